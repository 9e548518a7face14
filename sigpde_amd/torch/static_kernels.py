"""Static (point) kernels k(x_s, y_t) for the signature-kernel PDE.

Drop-ins for sigpde/torch/static_kernels.py:3-54 (stojl/SigPDE). Each object has
  .pairwise(X, Y) -> (A, M, N): k(X[a, s], Y[a, t])
  .gram(X, Y)     -> (A, B, M, N): k(X[a, s], Y[b, t])
evaluated with torch ops in the inputs' dtype, which is also the contract for user-defined
static kernels. The formulas are the reference's (they define the API's numbers):
  linear  scale * <x, y>
  RBF     exp(-(|x|^2 + |y|^2 - 2 <x, y>) / sigma)      (sigma not squared, no 1/2)
  L1-RBF  exp(-sum_k |x_k - y_k| / sigma)
LinearKernel additionally advertises `fused_linear_scale`, which lets SigPDE route it to the
fused increment-Gram kernels instead of materialising the point Gram.
"""
from __future__ import annotations

import torch


class LinearKernel:
    def __init__(self, scale: float = 1.0):
        self.scale = scale

    @property
    def fused_linear_scale(self) -> float:
        return float(self.scale)

    def pairwise(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        return self.scale * torch.bmm(X, Y.transpose(1, 2).contiguous())

    def gram(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        A, M, D = X.shape
        B, N, _ = Y.shape
        # (A*M, D) x (D, B*N) as one GEMM, then (A, M, B, N) -> (A, B, M, N)
        ip = torch.mm(X.reshape(A * M, D), Y.reshape(B * N, D).t())
        return self.scale * ip.view(A, M, B, N).permute(0, 2, 1, 3)


class RBFKernel:
    def __init__(self, sigma: float = 1.0):
        self.sigma = sigma

    def pairwise(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        A, M, _ = X.shape
        N = Y.shape[1]
        xx = (X ** 2).sum(dim=2).view(A, M, 1)
        yy = (Y ** 2).sum(dim=2).view(A, 1, N)
        dist = -2.0 * torch.bmm(X, Y.transpose(1, 2).contiguous())
        dist += xx + yy
        return torch.exp(-dist / self.sigma)

    def gram(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        A, M, D = X.shape
        B, N, _ = Y.shape
        xx = (X ** 2).sum(dim=2).view(A, 1, M, 1)
        yy = (Y ** 2).sum(dim=2).view(1, B, 1, N)
        ip = torch.mm(X.reshape(A * M, D), Y.reshape(B * N, D).t()).view(A, M, B, N)
        dist = -2.0 * ip.permute(0, 2, 1, 3)
        dist += xx + yy
        return torch.exp(-dist / self.sigma)


class L1RBFKernel:
    def __init__(self, sigma: float = 1.0):
        self.sigma = sigma

    def pairwise(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        dist = (X[:, :, None, :] - Y[:, None, :, :]).abs().sum(dim=-1)
        return torch.exp(-dist / self.sigma)

    def gram(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        dist = (X[:, None, :, None, :] - Y[None, :, None, :, :]).abs().sum(dim=-1)
        return torch.exp(-dist / self.sigma)
