"""The reference's import name: `import sigpde.torch as sig` (stojl/SigPDE README.md:13-15,
sigpde/__init__.py:6, sigpde/torch/__init__.py:9-16) runs the MI355X build unchanged.

`sigpde.torch` is the module `sigpde_amd.torch` itself (SigPDE, RobustSigPDE, kernels), and the
reference's submodule paths `sigpde.torch.kernel` and `sigpde.torch.static_kernels` are the build's
modules of those names, so `from sigpde.torch.kernel import SigPDE` and
`from sigpde.torch.static_kernels import LinearKernel` also keep working. Every solve runs in the
HIP kernels of sigpde_amd/libsigpde_hip.so; there is no CPU path.
"""
import sys as _sys

import sigpde_amd.torch as _torch_api
from sigpde_amd.torch import kernel as _kernel
from sigpde_amd.torch import static_kernels as _static_kernels

torch = _torch_api
_sys.modules[__name__ + ".torch"] = _torch_api
_sys.modules[__name__ + ".torch.kernel"] = _kernel
_sys.modules[__name__ + ".torch.static_kernels"] = _static_kernels

__all__ = ["torch"]
