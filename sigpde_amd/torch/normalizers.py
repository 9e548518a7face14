"""Built-in normalisers of RobustSigPDE, evaluated on the device.

The reference takes any callable `normalizer(K(x,x)) -> target` and evaluates it with torch between
the K(x,x) solve and the root-find (sigpde/torch/kernel.py:180-195; default
`log_normalizer`, kernel.py:23-24). Any callable still works here that way. The objects below are
also plain callables (the same formula in torch, in the tensor's dtype), and RobustSigPDE
recognises them by `device_normalizer` and runs them inside the root-find kernel instead, so the
whole normalisation is one launch with no host round trip (sigpde_robust_norm_builtin):

  LogNormalizer()         2 - 1/(1 + log x)                              (the reference's default)
  TensorNormalizer(C, a)  x for x <= C, else C + C^(1+a) (C^-a - x^-a)/a  (C >= 1, a > 0): the
                          tensor normalisation of the robust signature (increasing, at most x,
                          bounded by C (1 + 1/a))
  ClipNormalizer(C)       min(x, C)                                       (C >= 1)
"""
from __future__ import annotations

import torch

from .. import _native


class LogNormalizer:
    device_normalizer = (_native.NORM_LOG, 0.0, 0.0)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return 2 - 1 / (1 + x.log())


class TensorNormalizer:
    def __init__(self, C: float = 2.0, a: float = 1.0):
        if not (C >= 1.0 and a > 0.0):
            raise ValueError(f"TensorNormalizer needs C >= 1 and a > 0, got C={C}, a={a}")
        self.C = float(C)
        self.a = float(a)

    @property
    def device_normalizer(self):
        return (_native.NORM_TENSOR, self.C, self.a)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        c = torch.tensor(self.C, dtype=x.dtype, device=x.device)
        a = torch.tensor(self.a, dtype=x.dtype, device=x.device)
        tail = c + c.pow(1 + a) * (c.pow(-a) - x.pow(-a)) / a
        return torch.where(x <= c, x, tail)


class ClipNormalizer:
    def __init__(self, C: float = 2.0):
        if not C >= 1.0:
            raise ValueError(f"ClipNormalizer needs C >= 1, got C={C}")
        self.C = float(C)

    @property
    def device_normalizer(self):
        return (_native.NORM_CLIP, self.C, 0.0)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return torch.clamp(x, max=self.C)


__all__ = ["LogNormalizer", "TensorNormalizer", "ClipNormalizer"]
