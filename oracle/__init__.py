"""CPU oracle for the SigPDE signature-kernel scheme -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this
package, and only as the checker or the timed CPU baseline. The product (``sigpde_amd``) never
imports it and has no CPU fallback.

Contents:
  * ctypes bindings to ``liboracle.so`` (``sigpde_oracle.c``): the PDE solve in FP64 as written and
    in the reference's FP32 semantics, Gram / symmetric / scaled variants, Chandrupatla; pde_step
    as written (default) or with NVVM's FMA contraction (``contracted()``);
  * NumPy restatements of the static kernels (sigpde/torch/static_kernels.py:3-54), the increment
    second difference (sigpde/torch/kernel.py:13-21), the default normalizer
    (sigpde/torch/kernel.py:23-24) and the robust normalisation pipeline
    (sigpde/torch/kernel.py:180-195);
  * an independent NumPy anti-diagonal restatement of the wavefront (FP64), vectorised over each
    anti-diagonal like the reference's thread-per-row sweep (sigpde/cuda_pairwise_kernels.py:17-39),
    used to cross-check the C restatement bit for bit.

Parity pins for every function here live in ``tests/test_oracle.py`` (closed forms, invariants,
brute-force signatures, scipy root-finding, golden fixtures) and ``tests/test_oracle_typing.py``
(known answers of the float32 typing rules and of both contraction settings, computed in exact
rational arithmetic with explicit rounding).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Callable, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_ptr = ctypes.c_void_p


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("orc_pairwise_f64", "orc_pairwise_f32"):
            getattr(L, name).argtypes = [_ptr, _i64, _i64, _i64, _int, _ptr, _ptr, _ptr]
        L.orc_pairwise_f32_dscale.argtypes = [_ptr, _i64, _i64, _i64, _int, _ptr, _ptr]
        for name in ("orc_gram_f64", "orc_gram_f32"):
            getattr(L, name).argtypes = [_ptr, _i64, _i64, _i64, _i64, _int, _ptr, _ptr, _int, _ptr, _i64]
        for name in ("orc_linear_inc_f64", "orc_linear_inc_f32"):
            getattr(L, name).argtypes = [_ptr, _ptr, _i64, _i64, _i64, _dbl, _int, _ptr]
        for name in ("orc_linear_gram_f64", "orc_linear_gram_f32"):
            getattr(L, name).argtypes = [_ptr, _ptr, _i64, _i64, _i64, _i64, _i64, _dbl, _int,
                                         _ptr, _ptr, _ptr, _i64]
        for name in ("orc_chandrupatla_f64", "orc_chandrupatla_f32"):
            getattr(L, name).argtypes = [_ptr, _i64, _i64, _int, _ptr, _ptr, _int, _dbl, _ptr, _ptr,
                                         _ptr]
        L.orc_num_threads.restype = _int
        L.orc_set_num_threads.argtypes = [_int]
        L.orc_set_contract.argtypes = [_int]
        L.orc_get_contract.restype = _int
        _lib = L
    return _lib


def set_contract(on: bool) -> None:
    """Contraction setting of pde_step for every later solve: False (default) rounds every
    product and sum as written; True applies NVVM's FMA contraction of the expression
    (A = fma(c2, 1/12, fma(c, 0.5, 1)), B = fma(c2, -1/12, 1), k = fma(k01 + k10, A, -(k00 B))),
    the arithmetic of the reference on its NVIDIA target."""
    lib().orc_set_contract(int(bool(on)))


def get_contract() -> bool:
    return bool(lib().orc_get_contract())


class contracted:
    """Context manager: solves inside it use the contracted pde_step (set_contract(True))."""

    def __init__(self, on: bool = True):
        self.on = on

    def __enter__(self):
        self.prev = get_contract()
        set_contract(self.on)
        return self

    def __exit__(self, *exc):
        set_contract(self.prev)
        return False


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise TypeError(f"oracle supports float32/float64, got {dtype}")


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


# ----------------------------------------------------------------------------- PDE solves
def pde_pairwise(inc: np.ndarray, order: int, sx=None, sy=None) -> np.ndarray:
    """out[b] = K_b from increments (B, m, n); sigpde/cuda_pairwise_kernels.py:8-85."""
    dt = inc.dtype
    inc = _c(inc, dt)
    B, m, n = inc.shape
    sx, sy = _c(sx, dt), _c(sy, dt)
    out = np.zeros(B, dtype=dt)
    getattr(lib(), "orc_pairwise_" + _sfx(dt))(_p(inc), B, m, n, order, _p(sx), _p(sy), _p(out))
    return out


def pde_pairwise_f32_dscale(inc: np.ndarray, order: int, s: np.ndarray) -> np.ndarray:
    """The root-find's evaluation F for float32 paths: out[b] = K_b of c = inc[b] * s[b] with the
    double scale s = xt**2 (c and c**2 in double, cells stored as float32);
    sigpde/cuda_pairwise_kernels.py:114."""
    inc = _c(inc, np.float32)
    B, m, n = inc.shape
    s = _c(s, np.float64)
    out = np.zeros(B, dtype=np.float32)
    lib().orc_pairwise_f32_dscale(_p(inc), B, m, n, order, _p(s), _p(out))
    return out


def pde_gram(inc: np.ndarray, order: int, sx=None, sy=None, symmetric: bool = False) -> np.ndarray:
    """out[bx, by] from increments (Bx, By, m, n); sigpde/cuda_gram_kernels.py:8-165."""
    dt = inc.dtype
    inc = _c(inc, dt)
    Bx, By, m, n = inc.shape
    sx, sy = _c(sx, dt), _c(sy, dt)
    out = np.zeros((Bx, By), dtype=dt)
    getattr(lib(), "orc_gram_" + _sfx(dt))(_p(inc), Bx, By, m, n, order, _p(sx), _p(sy),
                                           int(symmetric), _p(out), By)
    return out


def linear_inc(x: np.ndarray, y: np.ndarray, order: int, kscale: float = 1.0) -> np.ndarray:
    """Increment matrix (Lx-1, Ly-1) of one pair for the linear static kernel, computed in the
    paths' dtype with a sequential feature-axis accumulation (sigpde/torch/kernel.py:13-16)."""
    dt = x.dtype
    x, y = _c(x, dt), _c(y, dt)
    Lx, d = x.shape
    Ly = y.shape[0]
    inc = np.zeros((Lx - 1, Ly - 1), dtype=dt)
    getattr(lib(), "orc_linear_inc_" + _sfx(dt))(_p(x), _p(y), Lx, Ly, d, float(kscale), order, _p(inc))
    return inc


def linear_gram(X: np.ndarray, Y: np.ndarray, order: int, kscale: float = 1.0, sx=None, sy=None) -> np.ndarray:
    """Full linear-kernel Gram from paths, pair by pair (no (Bx,By,m,n) tensor is formed)."""
    dt = X.dtype
    X, Y = _c(X, dt), _c(Y, dt)
    Bx, Lx, d = X.shape
    By, Ly, _ = Y.shape
    sx, sy = _c(sx, dt), _c(sy, dt)
    out = np.zeros((Bx, By), dtype=dt)
    getattr(lib(), "orc_linear_gram_" + _sfx(dt))(_p(X), _p(Y), Bx, By, Lx, Ly, d, float(kscale),
                                                  order, _p(sx), _p(sy), _p(out), By)
    return out


def chandrupatla(inc: np.ndarray, order: int, target: np.ndarray, knorm: np.ndarray,
                 maxit: int = 100, tol: float = 1e-8, fixed_points: bool = False):
    """Per-path lambda and evaluation count; sigpde/cuda_pairwise_kernels.py:88-166.
    fixed_points=True also returns, per path, the first evaluation count after which an iteration
    left the whole bracket state unchanged (-1 if none): the iteration repeats from there, so
    lambda is already final."""
    dt = inc.dtype
    inc = _c(inc, dt)
    B, m, m2 = inc.shape
    if m != m2:
        raise ValueError("chandrupatla takes (B, m, m) self-increments")
    target, knorm = _c(target, dt), _c(knorm, dt)
    out = np.zeros(B, dtype=dt)
    iters = np.zeros(B, dtype=np.int32)
    fixed = np.full(B, -1, dtype=np.int32)
    getattr(lib(), "orc_chandrupatla_" + _sfx(dt))(_p(inc), B, m, order, _p(target), _p(knorm),
                                                   int(maxit), float(tol), _p(out), _p(iters),
                                                   _p(fixed) if fixed_points else None)
    return (out, iters, fixed) if fixed_points else (out, iters)


# ----------------------------------------------------------------------------- static kernels
def static_pairwise(kind: str, X: np.ndarray, Y: np.ndarray, param: float = 1.0) -> np.ndarray:
    """Point Gram (A, M, N); sigpde/torch/static_kernels.py:3-54 (torch scalar semantics: the
    Python scalars do not promote float32 arrays)."""
    if kind == "linear":
        return param * np.matmul(X, np.swapaxes(Y, 1, 2))
    if kind == "rbf":
        A, M, _ = X.shape
        N = Y.shape[1]
        Xs = np.sum(X ** 2, axis=2)
        Ys = np.sum(Y ** 2, axis=2)
        dist = -2.0 * np.matmul(X, np.swapaxes(Y, 1, 2))
        dist += Xs.reshape(A, M, 1) + Ys.reshape(A, 1, N)
        return np.exp(-dist / param)
    if kind == "l1rbf":
        dist = np.abs(X[:, :, None, :] - Y[:, None, :, :]).sum(axis=-1)
        return np.exp(-dist / param)
    raise ValueError(kind)


def static_gram(kind: str, X: np.ndarray, Y: np.ndarray, param: float = 1.0) -> np.ndarray:
    """Point Gram (A, B, M, N); sigpde/torch/static_kernels.py:3-54."""
    if kind == "linear":
        return param * np.einsum("ipk,jqk->ijpq", X, Y)
    if kind == "rbf":
        A, M, _ = X.shape
        B, N, _ = Y.shape
        Xs = np.sum(X ** 2, axis=2)
        Ys = np.sum(Y ** 2, axis=2)
        dist = -2.0 * np.einsum("ipk,jqk->ijpq", X, Y)
        dist += Xs.reshape(A, 1, M, 1) + Ys.reshape(1, B, 1, N)
        return np.exp(-dist / param)
    if kind == "l1rbf":
        dist = np.abs(X[:, None, :, None, :] - Y[None, :, None, :, :]).sum(axis=-1)
        return np.exp(-dist / param)
    raise ValueError(kind)


def second_difference(ip: np.ndarray, order: int) -> np.ndarray:
    """sigpde/torch/kernel.py:15-16,20-21: ((ip[1:,1:] + ip[:-1,:-1]) - ip[1:,:-1]) - ip[:-1,1:],
    then / 2**(2*order), over the last two axes, in the array's dtype."""
    d = ip[..., 1:, 1:] + ip[..., :-1, :-1] - ip[..., 1:, :-1] - ip[..., :-1, 1:]
    return d / float(2 ** (2 * order))


def log_normalizer(x: np.ndarray) -> np.ndarray:
    """sigpde/torch/kernel.py:23-24: 2 - 1/(1 + log x)."""
    return 2 - 1 / (1 + np.log(x))


# ----------------------------------------------------------------------------- API mirrors
def sig_pairwise(X, Y, order, kind="linear", param=1.0, sx=None, sy=None):
    inc = second_difference(static_pairwise(kind, X, Y, param), order)
    return pde_pairwise(inc, order, sx, sy)


def sig_gram(X, Y, order, kind="linear", param=1.0, sx=None, sy=None, symmetric=False):
    inc = second_difference(static_gram(kind, X, Y, param), order)
    return pde_gram(inc, order, sx, sy, symmetric)


def norm_factors(inc: np.ndarray, order: int, normalizer: Optional[Callable] = None,
                 tol: float = 1e-8, maxit: int = 100, fixed_points: bool = False):
    """sigpde/torch/kernel.py:180-195 for increments (B, m, m) of each path with itself:
    K = PDE(inc); K[isnan] = inf; T = normalizer(K); lambda = Chandrupatla; lambda[K < 1] = 0.
    Returns (lambda, K, iterations), and the fixed points of chandrupatla() as a fourth value
    with fixed_points=True."""
    normalizer = log_normalizer if normalizer is None else normalizer
    knorm = pde_pairwise(inc, order)
    knorm[np.isnan(knorm)] = np.inf
    with np.errstate(all="ignore"):
        target = np.asarray(normalizer(knorm), dtype=inc.dtype)
    lam, iters, fixed = chandrupatla(inc, order, target, knorm, maxit, tol, fixed_points=True)
    lam[knorm < 1] = 0
    return (lam, knorm, iters, fixed) if fixed_points else (lam, knorm, iters)


def robust_normalization(X, order, kind="linear", param=1.0, normalizer=None, tol=1e-8, maxit=100):
    inc = second_difference(static_pairwise(kind, X, X, param), order)
    return norm_factors(inc, order, normalizer, tol, maxit)[0]


# ----------------------------------------------------------------------------- NumPy cross-check
def pde_pairwise_antidiag(inc: np.ndarray, order: int, scale: Optional[np.ndarray] = None,
                          f32ref: bool = False) -> np.ndarray:
    """Independent restatement: three live anti-diagonals indexed by row, updated one
    anti-diagonal p = i + j at a time with every row of the diagonal at once, as in the reference's
    thread-per-row sweep (sigpde/cuda_pairwise_kernels.py:17-39).
    f32ref=True applies the reference's float32 typing: c = float32(inc * scale), c*c in float32,
    the rest in float64, every stored cell rounded to float32."""
    inc = np.asarray(inc, dtype=np.float32 if f32ref else np.float64)
    B, m, n = inc.shape
    M, N = (m << order) + 1, (n << order) + 1  # refined lengths, sigpde/utils.py:6-7
    out = np.zeros(B, dtype=np.float32 if f32ref else np.float64)
    for b in range(B):
        st = np.float32 if f32ref else np.float64
        s = st(1.0) if scale is None else st(scale[b])
        d1 = np.ones(M, dtype=st)  # diagonal p-1, indexed by row i (k on the boundary is 1)
        d2 = np.ones(M, dtype=st)  # diagonal p-2
        for p in range(2, M + N - 1):
            i = np.arange(max(1, p - N + 1), min(M, p))
            j = p - i
            c = inc[b, (i - 1) >> order, (j - 1) >> order] * s  # in the storage dtype
            c2 = (c * c).astype(np.float64)
            c = c.astype(np.float64)
            k01 = np.where(i == 1, 1.0, d1[i - 1].astype(np.float64))
            k10 = np.where(j == 1, 1.0, d1[i].astype(np.float64))
            k00 = np.where((i == 1) | (j == 1), 1.0, d2[i - 1].astype(np.float64))
            val = (k01 + k10) * ((1.0 + 0.5 * c) + (1.0 / 12) * c2) - k00 * (1.0 - (1.0 / 12) * c2)
            d0 = np.ones(M, dtype=st)
            d0[i] = val.astype(st)
            d2, d1 = d1, d0
        out[b] = d1[M - 1]
    return out
