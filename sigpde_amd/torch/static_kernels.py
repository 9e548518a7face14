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
fused increment-Gram kernels instead of materialising the point Gram. All three advertise
`fused_static`, which lets SigPDE.gram in precision "ref" form the point values and their second
difference inside the solve (sigpde_static_gram), with no increment tensor at all. All three also
advertise
`pairwise_increments` / `gram_increments`: the increments of the API's second difference
(sigpde/torch/kernel.py:13-21) computed in one HIP pass from the point inner products (linear and
RBF: the same float operations as the torch formulas below, so the same increments) or from the
paths (L1-RBF: without the (A, B, M, N, D) difference tensor). A subclass that overrides pairwise / gram gets the
generic torch route instead.
"""
from __future__ import annotations

import torch

from .. import _native


class LinearKernel:
    def __init__(self, scale: float = 1.0):
        self.scale = scale

    @property
    def fused_linear_scale(self) -> float:
        return float(self.scale)

    @property
    def fused_static(self):
        """(static kernel code, scale) for the fused reference-scheme Gram, or None when a subclass
        overrides gram."""
        if type(self).gram is not LinearKernel.gram:
            return None
        return _native.STATIC_LINEAR, float(self.scale)

    def pairwise(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        return self.scale * torch.bmm(X, Y.transpose(1, 2).contiguous())

    def gram(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        A, M, D = X.shape
        B, N, _ = Y.shape
        # (A*M, D) x (D, B*N) as one GEMM, then (A, M, B, N) -> (A, B, M, N)
        ip = torch.mm(X.reshape(A * M, D), Y.reshape(B * N, D).t())
        return self.scale * ip.view(A, M, B, N).permute(0, 2, 1, 3)

    def pairwise_increments(self, X: torch.Tensor, Y: torch.Tensor, dyadic_order: int):
        """(A, M-1, N-1) second differences of pairwise(X, Y) / 4^order in one pass (the same
        float operations), or None when a subclass overrides pairwise."""
        if type(self).pairwise is not LinearKernel.pairwise:
            return None
        A, M, _ = X.shape
        N = Y.shape[1]
        ip = torch.bmm(X, Y.transpose(1, 2).contiguous())
        out = torch.empty((A, M - 1, N - 1), dtype=X.dtype, device=X.device)
        _native.linear_increments(ip, (M * N, 0, N, 1), A, 1, M - 1, N - 1, self.scale,
                                  dyadic_order, out)
        return out

    def gram_increments(self, X: torch.Tensor, Y: torch.Tensor, dyadic_order: int):
        """(A, B, M-1, N-1) second differences of gram(X, Y) / 4^order in one pass (the same
        float operations), or None when a subclass overrides gram."""
        if type(self).gram is not LinearKernel.gram:
            return None
        A, M, D = X.shape
        B, N, _ = Y.shape
        ip = torch.mm(X.reshape(A * M, D), Y.reshape(B * N, D).t())  # element (a, i, b, j)
        out = torch.empty((A, B, M - 1, N - 1), dtype=X.dtype, device=X.device)
        _native.linear_increments(ip, (M * B * N, N, B * N, 1), A, B, M - 1, N - 1, self.scale,
                                  dyadic_order, out)
        return out


class RBFKernel:
    def __init__(self, sigma: float = 1.0):
        self.sigma = sigma

    @property
    def fused_static(self):
        """(static kernel code, sigma) for the fused reference-scheme Gram, or None when a
        subclass overrides gram."""
        if type(self).gram is not RBFKernel.gram:
            return None
        return _native.STATIC_RBF, float(self.sigma)

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

    def pairwise_increments(self, X: torch.Tensor, Y: torch.Tensor, dyadic_order: int):
        """(A, M-1, N-1) second differences of pairwise(X, Y) / 4^order in one pass, or None when
        a subclass overrides pairwise."""
        if type(self).pairwise is not RBFKernel.pairwise:
            return None
        A, M, _ = X.shape
        N = Y.shape[1]
        xx = (X ** 2).sum(dim=2)
        yy = (Y ** 2).sum(dim=2)
        ip = torch.bmm(X, Y.transpose(1, 2).contiguous())
        out = torch.empty((A, M - 1, N - 1), dtype=X.dtype, device=X.device)
        _native.rbf_increments(ip, (M * N, 0, N, 1), A, 1, M - 1, N - 1, xx, M, yy, N, 0,
                               self.sigma, dyadic_order, out)
        return out

    def gram_increments(self, X: torch.Tensor, Y: torch.Tensor, dyadic_order: int):
        """(A, B, M-1, N-1) second differences of gram(X, Y) / 4^order in one pass, or None when
        a subclass overrides gram."""
        if type(self).gram is not RBFKernel.gram:
            return None
        A, M, D = X.shape
        B, N, _ = Y.shape
        xx = (X ** 2).sum(dim=2)
        yy = (Y ** 2).sum(dim=2)
        ip = torch.mm(X.reshape(A * M, D), Y.reshape(B * N, D).t())  # element (a, i, b, j)
        out = torch.empty((A, B, M - 1, N - 1), dtype=X.dtype, device=X.device)
        _native.rbf_increments(ip, (M * B * N, N, B * N, 1), A, B, M - 1, N - 1, xx, M, yy, 0,
                               N, self.sigma, dyadic_order, out)
        return out


class L1RBFKernel:
    def __init__(self, sigma: float = 1.0):
        self.sigma = sigma

    @property
    def fused_static(self):
        """(static kernel code, sigma) for the fused reference-scheme Gram, or None when a
        subclass overrides gram."""
        if type(self).gram is not L1RBFKernel.gram:
            return None
        return _native.STATIC_L1RBF, float(self.sigma)

    def pairwise(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        dist = (X[:, :, None, :] - Y[:, None, :, :]).abs().sum(dim=-1)
        return torch.exp(-dist / self.sigma)

    def gram(self, X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        dist = (X[:, None, :, None, :] - Y[None, :, None, :, :]).abs().sum(dim=-1)
        return torch.exp(-dist / self.sigma)

    def pairwise_increments(self, X: torch.Tensor, Y: torch.Tensor, dyadic_order: int):
        """(A, M-1, N-1) second differences of pairwise(X, Y) / 4^order from the paths, or None
        when a subclass overrides pairwise."""
        if type(self).pairwise is not L1RBFKernel.pairwise:
            return None
        A, M, _ = X.shape
        N = Y.shape[1]
        out = torch.empty((A, M - 1, N - 1), dtype=X.dtype, device=X.device)
        _native.l1rbf_increments(X.contiguous(), Y.contiguous(), self.sigma, dyadic_order, out,
                                 pairwise=True)
        return out

    def gram_increments(self, X: torch.Tensor, Y: torch.Tensor, dyadic_order: int):
        """(A, B, M-1, N-1) second differences of gram(X, Y) / 4^order from the paths, or None
        when a subclass overrides gram."""
        if type(self).gram is not L1RBFKernel.gram:
            return None
        A, M, _ = X.shape
        B, N, _ = Y.shape
        out = torch.empty((A, B, M - 1, N - 1), dtype=X.dtype, device=X.device)
        _native.l1rbf_increments(X.contiguous(), Y.contiguous(), self.sigma, dyadic_order, out,
                                 pairwise=False)
        return out
