"""MI355X-native signature-kernel PDE solver (gfx950 HIP kernels behind a C ABI).

The public API lives in `sigpde_amd.torch` (SigPDE, RobustSigPDE, kernels), mirroring
stojl/SigPDE's `sigpde.torch`.
"""
from . import torch

__all__ = ["torch"]
