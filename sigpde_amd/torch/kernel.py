"""SigPDE / RobustSigPDE: drop-in for stojl/SigPDE's sigpde/torch/kernel.py on MI355X.

Same classes, methods, arguments, defaults and result semantics as the reference
(sigpde/torch/kernel.py:26-301); the PDE solves and the robust root-find run in the HIP kernels of
libsigpde_hip.so through sigpde_amd._native.

Semantics kept from the reference:
  * increments = 4-term second difference of the static kernel's point Gram divided by
    4**dyadic_order (sigpde/torch/kernel.py:13-21);
  * y=None means y=x and y_scale=x_scale; a missing scale is ones, a shape-(1,) scale is repeated
    (kernel.py:32-41, 80-84); gram(x) takes the symmetric path (kernel.py:70-71);
  * results are new tensors in x's dtype on x's device;
  * robust: K(x,x) -> NaN to +inf -> normalizer -> Chandrupatla -> lambda = 0 where K(x,x) < 1
    (kernel.py:180-195); default normalizer 2 - 1/(1 + log x) (kernel.py:23-24);
    tol=1e-8, maxit=100 defaults.
  * max_batch bounds the pairs whose increments are materialised at once (pairwise chunks of
    max_batch, Gram tiles of side ceil(sqrt(max_batch)), kernel.py:55, 68). max_threads is
    accepted for signature compatibility; the launch geometry is chosen by the library and does
    not change results.
Differences (documented, deliberate):
  * calls are stream-ordered on torch's current stream instead of ending in a device-wide
    synchronise; torch ops on the result are ordered after the solve.
  * inputs are validated (3-D, same dtype/device, length >= 2, float32/float64, GPU tensors).
  * RobustSigPDE.pairwise(x, y) normalises y on y's own length; the reference reuses x's refined
    length for y's solves (kernel.py:239-247 with PDESolvers.py:43-45), which only agrees when
    x and y have equal lengths.
  * symmetric Gram tiles lying entirely above the diagonal are skipped instead of computing and
    discarding their increments (kernel.py:143-150).
Tracing: with SIGPDE_ROCTX=1 the public calls and their phases (increments, root-find) run inside
named roctx ranges (sigpde_amd.trace).
"""
from __future__ import annotations

import math
from typing import Callable, Optional

import torch

from .. import _native
from ..trace import traced


@traced("increments.pairwise")
def pairwise_inner_product(x, y, static_kernel, dyadic_order):
    """Increments (B, Lx-1, Ly-1); sigpde/torch/kernel.py:13-16. Static kernels with a one-pass
    `pairwise_increments` (RBF, L1-RBF) use it."""
    hook = getattr(static_kernel, "pairwise_increments", None)
    inc = hook(x, y, dyadic_order) if hook is not None else None
    if inc is not None:
        return inc
    ip = static_kernel.pairwise(x, y)
    d = ip[:, 1:, 1:] + ip[:, :-1, :-1] - ip[:, 1:, :-1] - ip[:, :-1, 1:]
    return d / float(2 ** (2 * dyadic_order))


@traced("increments.gram")
def gram_inner_product(x, y, static_kernel, dyadic_order):
    """Increments (Bx, By, Lx-1, Ly-1); sigpde/torch/kernel.py:18-21. Static kernels with a one-pass
    `gram_increments` (RBF, L1-RBF) use it."""
    hook = getattr(static_kernel, "gram_increments", None)
    inc = hook(x, y, dyadic_order) if hook is not None else None
    if inc is not None:
        return inc
    ip = static_kernel.gram(x, y)
    d = ip[:, :, 1:, 1:] + ip[:, :, :-1, :-1] - ip[:, :, 1:, :-1] - ip[:, :, :-1, 1:]
    return d / float(2 ** (2 * dyadic_order))


def log_normalizer(x):
    """Default robust normaliser 2 - 1/(1 + log x); sigpde/torch/kernel.py:23-24."""
    return 2 - 1 / (1 + x.log())


def _chunks(total: int, size: int):
    """(start, stop) chunks of `size`, last one the remainder; sigpde/BatchIterator.py:3-18."""
    size = max(1, min(size, total))
    for start in range(0, total, size):
        yield start, min(start + size, total)


def _check_paths(x: torch.Tensor, name: str = "x") -> None:
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if x.dim() != 3:
        raise ValueError(f"{name} must have shape (batch, length, dim), got {tuple(x.shape)}")
    if x.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"{name} must be float32 or float64, got {x.dtype}")
    if not x.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor; the solver runs only on the device")
    if x.shape[1] < 2:
        raise ValueError(f"{name} needs length >= 2, got {x.shape[1]}")


def _check_pair(x, y):
    _check_paths(x, "x")
    _check_paths(y, "y")
    if y.dtype != x.dtype or y.device != x.device:
        raise ValueError("x and y must share dtype and device")
    if x.shape[2] != y.shape[2]:
        raise ValueError(f"x and y must have the same dim, got {x.shape[2]} and {y.shape[2]}")


def _scale_vector(s, batch: int, like: torch.Tensor) -> torch.Tensor:
    """Missing -> ones, shape (1,) -> repeated (sigpde/torch/kernel.py:36-41)."""
    if s is None:
        return torch.ones(batch, device=like.device, dtype=like.dtype)
    s = torch.as_tensor(s, device=like.device, dtype=like.dtype).reshape(-1)
    if s.shape[0] == 1:
        s = s.repeat(batch)
    if s.shape[0] != batch:
        raise ValueError(f"scale has {s.shape[0]} entries for a batch of {batch}")
    return s.contiguous()


PRECISIONS = ("ref", "fp16x2", "bf16x3", "fp32", "fp64", "fp16x2f")


def _fused_scale(static_kernel, precision: str):
    """kscale when the call takes the fused linear-kernel kernels, else None."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}, got {precision!r}")
    if precision == "ref":
        return None
    scale = getattr(static_kernel, "fused_linear_scale", None)
    if scale is None:
        raise ValueError(f"precision={precision!r} needs the linear static kernel "
                         "(kernels.LinearKernel); other static kernels run with precision='ref'")
    return float(scale)


class SigPDE:
    """Untruncated signature kernel by the Goursat PDE (stojl/SigPDE sigpde/torch/kernel.py:26-171).

    precision (an addition to the reference's signature; per instance, overridable per call):
      "ref"    -- the reference's arithmetic: increments from the static kernel's point Gram
                  (4-term second difference, in the inputs' dtype), float32 storage with double
                  per-cell arithmetic for float32 inputs, all-double for float64 inputs. Grams of
                  the built-in static kernels (linear, RBF, L1-RBF) within dyadic order <= 3 and
                  the dim / length limits of sigpde_static_gram form the point values and their
                  second difference inside the solve (no increment tensor; inner products and
                  distance sums in the kernel's summation order, equal to torch's when exact);
                  other Grams and every pairwise call run the reference's tiles of torch
                  increments;
      "fp16x2" -- linear kernel only: increments <dx, dy> formed from the paths and multiplied as
                  per-path-scaled fp16 hi/lo pairs on MFMA (fp32 accumulate), double wavefront;
                  relative error ~3e-7 against the float64 reference scheme;
      "bf16x3" -- linear kernel only: as "fp16x2" with the increments split into three bf16 parts
                  and six bf16 MFMA products (the terms of magnitude >= 2^-16 of hi*hi);
      "fp32"   -- linear kernel only, plain FP32 fused: increments rounded to float, <dx, dy> on
                  fp32 MFMA, the wavefront in the reference's float32 typing (no increment
                  tensor in HBM, as for "fp16x2");
      "fp64"   -- linear kernel only, dim <= 64: increments <dx, dy> in double on fp64 MFMA, the
                  wavefront in double (the reference scheme in float64 without the increment
                  tensor; results in the inputs' dtype);
      "fp16x2f" -- linear kernel only: the "fp16x2" increments with a float32 wavefront (float
                  coefficients and state); dyadic order <= 2 and up to 512 refined rows and
                  columns, where it stays within 1e-5 of float64 (4e-6 at 255 x 255 cells);
                  faster than "fp16x2".
    """

    def __init__(self, static_kernel, dyadic_order: int = 0, precision: str = "ref"):
        if int(dyadic_order) < 0:
            raise ValueError("dyadic_order must be >= 0")
        _fused_scale(static_kernel, precision)
        self.static_kernel = static_kernel
        self.dyadic_order = int(dyadic_order)
        self.precision = precision

    # ------------------------------------------------------------------ pairwise
    @traced("SigPDE.pairwise")
    def pairwise(self, x, y=None, x_scale=None, y_scale=None, max_batch=1000, max_threads=1024,
                 precision=None):
        y_scale = x_scale if y is None else y_scale
        y = x if y is None else y
        _check_pair(x, y)
        if y.shape[0] != x.shape[0]:
            raise ValueError("pairwise needs x and y with the same batch size")
        B = x.shape[0]
        scaled = x_scale is not None or y_scale is not None
        sx = _scale_vector(x_scale, B, x) if scaled else None
        sy = _scale_vector(y_scale, B, y) if scaled else None
        x, y = x.contiguous(), y.contiguous()
        out = torch.zeros(B, device=x.device, dtype=x.dtype)
        kscale = _fused_scale(self.static_kernel, precision or self.precision)
        if kscale is not None:
            _native.linear_pairwise(x, y, self.dyadic_order, kscale, out, precision or self.precision,
                                    sx, sy)
            return out
        for start, stop in _chunks(B, int(max_batch)):
            inc = pairwise_inner_product(x[start:stop], y[start:stop], self.static_kernel,
                                         self.dyadic_order).contiguous()
            _native.pde_pairwise(inc, self.dyadic_order, out[start:stop],
                                 None if sx is None else sx[start:stop],
                                 None if sy is None else sy[start:stop])
        return out

    # ------------------------------------------------------------------ gram
    @traced("SigPDE.gram")
    def gram(self, x, y=None, x_scale=None, y_scale=None, max_batch=1000, max_threads=1024,
             precision=None):
        precision = precision or self.precision
        kscale = _fused_scale(self.static_kernel, precision)
        if kscale is not None:
            return self._gram_fused(x, y, x_scale, y_scale, kscale, precision)
        spec = self._static_route(x, y)
        if spec is not None:
            return self._gram_static(spec, x, y, x_scale, y_scale)
        tile = math.ceil(math.sqrt(int(max_batch)))
        if y is None:
            return self._gram_sym(x, x_scale, tile)
        return self._gram(x, y, x_scale, y_scale, tile)

    def _static_route(self, x, y):
        """(static kernel code, parameter) when precision "ref" takes the fused static-kernel
        Gram (the point values and their second difference formed inside the solve), else None:
        built-in static kernels within the route's dyadic order, dim and length limits."""
        spec = getattr(self.static_kernel, "fused_static", None)
        if spec is None or not isinstance(x, torch.Tensor) or not x.is_cuda:
            return None
        if y is not None and (not isinstance(y, torch.Tensor) or y.dim() != 3 or x.dim() != 3
                              or y.shape[2] != x.shape[2]):
            return None
        if x.dim() != 3:
            return None
        Ly = x.shape[1] if y is None else y.shape[1]
        if not _native.static_gram_supported(spec[0], x.dtype, x.shape[1], Ly, x.shape[2],
                                             self.dyadic_order):
            return None
        return spec

    def _gram_static(self, spec, x, y, x_scale, y_scale):
        symmetric = y is None
        if symmetric:
            y, y_scale = x, x_scale
        _check_pair(x, y)
        Bx, By = x.shape[0], y.shape[0]
        scaled = x_scale is not None or y_scale is not None
        sx = _scale_vector(x_scale, Bx, x) if scaled else None
        sy = (sx if symmetric else _scale_vector(y_scale, By, y)) if scaled else None
        x = x.contiguous()
        y = x if symmetric else y.contiguous()
        out = torch.zeros((Bx, By), device=x.device, dtype=x.dtype)
        _native.static_gram(spec[0], x, y, self.dyadic_order, spec[1], out, sx, sy, symmetric)
        return out

    def _gram_fused(self, x, y, x_scale, y_scale, kscale, precision):
        symmetric = y is None
        if symmetric:
            y, y_scale = x, x_scale
        _check_pair(x, y)
        Bx, By = x.shape[0], y.shape[0]
        scaled = x_scale is not None or y_scale is not None
        sx = _scale_vector(x_scale, Bx, x) if scaled else None
        sy = (sx if symmetric else _scale_vector(y_scale, By, y)) if scaled else None
        x = x.contiguous()
        y = x if symmetric else y.contiguous()
        out = torch.zeros((Bx, By), device=x.device, dtype=x.dtype)
        _native.linear_gram(x, y, self.dyadic_order, kscale, out, precision, sx, sy, symmetric)
        return out

    def _gram(self, x, y, x_scale, y_scale, tile):
        _check_pair(x, y)
        Bx, By = x.shape[0], y.shape[0]
        scaled = x_scale is not None or y_scale is not None
        sx = _scale_vector(x_scale, Bx, x) if scaled else None
        sy = _scale_vector(y_scale, By, y) if scaled else None
        x, y = x.contiguous(), y.contiguous()
        out = torch.zeros((Bx, By), device=x.device, dtype=x.dtype)
        for ys, ye in _chunks(By, tile):
            for xs, xe in _chunks(Bx, tile):
                inc = gram_inner_product(x[xs:xe], y[ys:ye], self.static_kernel,
                                         self.dyadic_order).contiguous()
                _native.pde_gram(inc, self.dyadic_order, out, xs, ys,
                                 None if sx is None else sx[xs:xe],
                                 None if sy is None else sy[ys:ye])
        return out

    def _gram_sym(self, x, x_scale, tile):
        _check_paths(x)
        B = x.shape[0]
        s = _scale_vector(x_scale, B, x) if x_scale is not None else None
        x = x.contiguous()
        out = torch.zeros((B, B), device=x.device, dtype=x.dtype)
        for ys, ye in _chunks(B, tile):
            for xs, xe in _chunks(B, tile):
                if ys > xe - 1:  # tile entirely above the diagonal
                    continue
                inc = gram_inner_product(x[xs:xe], x[ys:ye], self.static_kernel,
                                         self.dyadic_order).contiguous()
                _native.pde_gram(inc, self.dyadic_order, out, xs, ys,
                                 None if s is None else s[xs:xe],
                                 None if s is None else s[ys:ye], symmetric=True)
        return out


class RobustSigPDE:
    """Robust (normalised) signature kernel (stojl/SigPDE sigpde/torch/kernel.py:174-301).

    precision (an addition to the reference's signature) selects how the final scaled solves of
    pairwise / gram run, with the same meaning as for SigPDE ("ref", "fp16x2", "bf16x3", "fp32",
    "fp64"; the fused modes need the linear static kernel). The normalisation root-find always
    runs the reference's arithmetic on the increments of the inputs' dtype.
    """

    def __init__(self, static_kernel, dyadic_order: int = 0, normalizer: Optional[Callable] = None,
                 precision: str = "ref"):
        if int(dyadic_order) < 0:
            raise ValueError("dyadic_order must be >= 0")
        _fused_scale(static_kernel, precision)
        self.static_kernel = static_kernel
        self.dyadic_order = int(dyadic_order)
        self.normalizer = log_normalizer if normalizer is None else normalizer
        self.precision = precision

    @traced("RobustSigPDE.root_find")
    def _norm_factors(self, inc, lam, knorm, normalizer, tol, maxit, iters=None):
        """sigpde/torch/kernel.py:180-195 on one chunk of self-increments (B, m, m). The default
        log normaliser and the built-in ones (normalizers.py: objects with `device_normalizer`)
        run inside the root-find kernel; any other callable between its two phases, on the host
        side, as in the reference."""
        if normalizer is log_normalizer:
            _native.robust_norm(inc, self.dyadic_order, lam, knorm, None, maxit, tol, iters)
            return
        builtin = getattr(normalizer, "device_normalizer", None)
        if builtin is not None:
            _native.robust_norm_builtin(inc, self.dyadic_order, lam, knorm, builtin[0], builtin[1],
                                        builtin[2], maxit, tol, iters)
            return
        _native.pde_pairwise(inc, self.dyadic_order, knorm)
        knorm[torch.isnan(knorm)] = torch.inf
        target = normalizer(knorm)
        target = torch.as_tensor(target, device=knorm.device, dtype=knorm.dtype).contiguous()
        _native.robust_norm(inc, self.dyadic_order, lam, knorm, target, maxit, tol, iters)

    def _normalization(self, x, normalizer, tol, maxit, max_batch, iters=None):
        B = x.shape[0]
        lam = torch.zeros(B, device=x.device, dtype=x.dtype)
        knorm = torch.zeros(B, device=x.device, dtype=x.dtype)
        for start, stop in _chunks(B, int(max_batch)):
            xi = x[start:stop]
            inc = pairwise_inner_product(xi, xi, self.static_kernel, self.dyadic_order).contiguous()
            self._norm_factors(inc, lam[start:stop], knorm[start:stop], normalizer, tol, maxit,
                               None if iters is None else iters[start:stop])
        return lam

    @traced("RobustSigPDE.normalization")
    def normalization(self, x, normalizer=None, tol=1e-8, maxit=100, max_batch=1000, max_threads=1024):
        normalizer = self.normalizer if normalizer is None else normalizer
        _check_paths(x)
        return self._normalization(x.contiguous(), normalizer, tol, maxit, max_batch)

    @staticmethod
    def _joint(x, y, normalizer):
        """x's and y's root-finds can share one launch: paths of one shape and dtype, and the
        default normaliser (elementwise; a custom one sees each batch as the reference passes it).
        Each path's root-find is independent of the other paths in its launch; one launch keeps
        more of the GPU busy when there are few paths."""
        return (normalizer is log_normalizer and x.shape[1:] == y.shape[1:]
                and x.dtype == y.dtype and x.device == y.device)

    @traced("RobustSigPDE.pairwise")
    def pairwise(self, x, y=None, normalizer=None, tol=1e-8, maxit=100, max_batch=1000, max_threads=1024):
        symmetric = y is None
        y = x if symmetric else y
        normalizer = self.normalizer if normalizer is None else normalizer
        _check_pair(x, y)
        if y.shape[0] != x.shape[0]:
            raise ValueError("pairwise needs x and y with the same batch size")
        x, y = x.contiguous(), y.contiguous()
        B = x.shape[0]
        out = torch.zeros(B, device=x.device, dtype=x.dtype)
        joint = not symmetric and self._joint(x, y, normalizer)
        for start, stop in _chunks(B, int(max_batch)):
            xi = x[start:stop]
            yi = xi if symmetric else y[start:stop]
            nb = stop - start
            if joint:
                xy = torch.cat([xi, yi])
                xy_inc = pairwise_inner_product(xy, xy, self.static_kernel,
                                                self.dyadic_order).contiguous()
                sxy = torch.zeros(2 * nb, device=x.device, dtype=x.dtype)
                kxy = torch.zeros(2 * nb, device=x.device, dtype=x.dtype)
                self._norm_factors(xy_inc, sxy, kxy, normalizer, tol, maxit)
                sx, sy = sxy[:nb], sxy[nb:]
                x_inc = None
            else:
                x_inc = pairwise_inner_product(xi, xi, self.static_kernel,
                                               self.dyadic_order).contiguous()
                sx = torch.zeros(nb, device=x.device, dtype=x.dtype)
                kx = torch.zeros(nb, device=x.device, dtype=x.dtype)
                self._norm_factors(x_inc, sx, kx, normalizer, tol, maxit)
            if symmetric:
                sy = sx
            elif not joint:
                y_inc = pairwise_inner_product(yi, yi, self.static_kernel, self.dyadic_order).contiguous()
                sy = torch.zeros(nb, device=x.device, dtype=x.dtype)
                ky = torch.zeros(nb, device=x.device, dtype=x.dtype)
                self._norm_factors(y_inc, sy, ky, normalizer, tol, maxit)
            if self.precision != "ref":
                out[start:stop] = SigPDE(self.static_kernel, self.dyadic_order, self.precision) \
                    .pairwise(xi, yi, sx, sy)
                continue
            inc = x_inc if symmetric else pairwise_inner_product(
                xi, yi, self.static_kernel, self.dyadic_order).contiguous()
            _native.pde_pairwise(inc, self.dyadic_order, out[start:stop], sx, sy)
        return out

    @traced("RobustSigPDE.gram")
    def gram(self, x, y=None, x_scale=None, y_scale=None, normalizer=None, tol=1e-8, maxit=100,
             max_batch=1000, max_threads=1024):
        nz = self.normalizer if normalizer is None else normalizer
        if y is not None and x_scale is None and y_scale is None:
            _check_paths(x)
            _check_paths(y)
            if self._joint(x, y, nz):
                lam = self._normalization(torch.cat([x, y]).contiguous(), nz, tol, maxit, max_batch)
                x_scale, y_scale = lam[:x.shape[0]], lam[x.shape[0]:]
        x_scales = self.normalization(x, normalizer, tol, maxit, max_batch, max_threads) \
            if x_scale is None else x_scale
        if y is None:
            y_scales = None
        else:
            y_scales = self.normalization(y, normalizer, tol, maxit, max_batch, max_threads) \
                if y_scale is None else y_scale
        return SigPDE(self.static_kernel, self.dyadic_order, self.precision).gram(
            x, y, x_scales, y_scales, max_batch, max_threads)
