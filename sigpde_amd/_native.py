"""ctypes binding to libsigpde_hip.so (the C ABI declared in include/sigpde_hip.h).

The HIP library is the only compute backend: if it is missing or no GPU is visible, every entry
point raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsigpde_hip.so")

F32, F64 = 0, 1
EINVAL, EDEVICE, ENOMEM, ETIMEOUT = 1, 2, 3, 4
NORM_LOG, NORM_GIVEN, NORM_TENSOR, NORM_CLIP = 0, 1, 2, 3
PREC_FP16X2 = 1
PREC_FP32 = 2
PREC_BF16X3 = 3
PREC_FP64 = 4
PREC_FP16X2F = 5
PRECISIONS = {"fp16x2": PREC_FP16X2, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp64": PREC_FP64,
              "fp16x2f": PREC_FP16X2F}

_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double

# name -> argtypes; every entry point returns int status
_SIGNATURES = {
    "sigpde_pde_pairwise": [_int, _vp, _i64, _i64, _i64, _i64, _int, _vp, _vp, _vp, _vp],
    "sigpde_pde_gram": [_int, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _int, _vp, _vp, _int,
                        _i64, _i64, _vp, _i64, _vp],
    "sigpde_robust_norm": [_int, _vp, _i64, _i64, _i64, _int, _int, _vp, _vp, _int, _dbl, _vp,
                           _vp, _vp],
    "sigpde_linear_gram": [_int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _dbl, _int, _vp,
                           _vp, _int, _vp, _i64, _vp, ctypes.c_size_t, _vp],
    "sigpde_linear_pairwise": [_int, _int, _vp, _vp, _i64, _i64, _i64, _i64, _dbl, _int, _vp,
                               _vp, _vp, _vp, ctypes.c_size_t, _vp],
    "sigpde_linear_gram_batch": [_int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _dbl, _int,
                                 ctypes.POINTER(_i64), _i64, _vp, _vp, _vp, _vp, ctypes.c_size_t,
                                 _vp],
    "sigpde_rbf_increments": [_int, _vp, _i64, _i64, _i64, _i64, ctypes.POINTER(_i64), _vp, _i64,
                              _vp, _i64, _i64, _dbl, _int, _vp, _vp],
    "sigpde_l1rbf_increments": [_int, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _int, _dbl, _int,
                                _vp, _vp],
    "sigpde_linear_increments": [_int, _vp, _i64, _i64, _i64, _i64, ctypes.POINTER(_i64), _dbl,
                                 _int, _vp, _vp],
    "sigpde_static_gram": [_int, _int, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _dbl, _int, _vp,
                           _vp, _int, _vp, _i64, _vp, ctypes.c_size_t, _vp],
    "sigpde_static_gram_supported": [_int, _int, _i64, _i64, _i64, _int],
    "sigpde_robust_norm_builtin": [_int, _vp, _i64, _i64, _i64, _int, _int, _dbl, _dbl, _vp, _int,
                                   _dbl, _vp, _vp, _vp],
    "sigpde_set_spread": [_int, _dbl, _int],
}


class NativeError(RuntimeError):
    """Raised when a C ABI call returns a non-zero status (`status`: the SIGPDE_E* code)."""

    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status


def load(path: str | None = None) -> ctypes.CDLL:
    """Loads the HIP library (once). Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise ImportError(
                f"{p} is missing: build it with `python -m sigpde_amd.build` "
                "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        lib = ctypes.CDLL(p)
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = _int
        lib.sigpde_linear_workspace_bytes.restype = ctypes.c_size_t
        lib.sigpde_linear_workspace_bytes.argtypes = [_int, _i64, _i64, _i64, _i64, _i64, _int]
        lib.sigpde_static_workspace_bytes.restype = ctypes.c_size_t
        lib.sigpde_static_workspace_bytes.argtypes = [_int, _int, _i64, _i64, _i64, _i64, _i64,
                                                      _int]
        lib.sigpde_last_error.restype = ctypes.c_char_p
        lib.sigpde_last_error.argtypes = []
        lib.sigpde_version.restype = ctypes.c_char_p
        lib.sigpde_version.argtypes = []
        lib.sigpde_issue_rates.restype = _int
        lib.sigpde_issue_rates.argtypes = [ctypes.POINTER(_dbl), _int]
        _lib = lib
        return lib


def version() -> str:
    return load().sigpde_version().decode()


def issue_rates() -> dict:
    """Issue rates of the current GPU measured by sigpde_issue_rates (clock64 ticks per wave64
    instruction per SIMD, two waves per SIMD on every CU), with the vector and MFMA rates they
    imply at the reported clock, how far MFMA and FP32 VALU work of different waves on one SIMD
    overlap, the rate of an FP32 add with a DPP operand and two dependent-issue latencies."""
    buf = (_dbl * 13)()
    _check(load().sigpde_issue_rates(buf, 13), "sigpde_issue_rates")
    f32, f64, mfma, mix, hz, cus = buf[0], buf[1], buf[2], buf[3], buf[4], buf[5]
    simds = 4 * cus
    return {
        "ticks_per_fp32_fma": f32,
        "ticks_per_fp64_fma": f64,
        "ticks_per_mfma_f32_16x16x32_f16": mfma,
        "mfma_plus_valu_over_sum": mix,
        "split_1x1_over_sum": buf[7],
        "split_full_overlap_over_sum": buf[8],
        "split_2x2_over_sum": buf[9],
        "ticks_per_fp32_add_dpp": buf[10],
        "latency_fp32_fma_ticks": buf[11],
        "latency_fp32_add_dpp_ticks": buf[12],
        "clock_hz": hz,
        "cus": int(cus),
        "fp32_valu_tflops": 64 * 2 / f32 * simds * hz / 1e12,
        "fp64_valu_tflops": 64 * 2 / f64 * simds * hz / 1e12,
        "fp16_mfma_tflops": 16 * 16 * 32 * 2 / mfma * simds * hz / 1e12,
    }


def set_spread(groups: int = 0, wait_seconds: float = 2.0, test_hook: int = 0) -> None:
    """Settings of the route that spreads each of few long pairs over several workgroups
    (process-wide): groups 0 = automatic, 1 = never, 2..32 = that many whenever resident;
    wait_seconds bounds each wait between workgroups; test_hook 1 makes workgroup 0 of each pair
    publish no progress, so spread calls fail with ETIMEOUT (a test of the failure path)."""
    _check(load().sigpde_set_spread(int(groups), float(wait_seconds), int(test_hook)),
           "sigpde_set_spread")


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().sigpde_last_error().decode()
        raise NativeError(f"{what} failed (status {rc}): {msg}", rc)


def dtype_code(t: torch.dtype) -> int:
    if t == torch.float32:
        return F32
    if t == torch.float64:
        return F64
    raise TypeError(f"sigpde_amd supports float32 and float64 paths, got {t}")


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _require_gpu(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise ValueError("sigpde_amd kernels take GPU tensors (device='cuda'); got a CPU tensor")


def _require(ok: bool, msg: str) -> None:
    """Layout checks that stay on under `python -O`: a wrong stride or size would reach the
    kernels as out-of-bounds device accesses."""
    if not ok:
        raise ValueError(msg)


def _require_vec(t, n: int, what: str) -> None:
    """An optional per-path vector: None, or a contiguous 1-D tensor of n elements."""
    if t is not None:
        _require(t.dim() == 1 and t.shape[0] == n and (n <= 1 or t.stride(0) == 1),
                 f"{what} must be a contiguous vector of {n} elements, got {tuple(t.shape)}")


def pde_pairwise(inc: torch.Tensor, order: int, out: torch.Tensor, sx=None, sy=None) -> None:
    """out[b] = PDE solve of increments inc[b] (B, m, n)."""
    _require_gpu(inc, out, sx, sy)
    _require(inc.dim() == 3 and inc.stride(2) == 1 and inc.stride(1) == inc.shape[2],
             "pde_pairwise: increments must be (B, m, n) with unit-stride rows packed m x n")
    _require_vec(out, inc.shape[0], "pde_pairwise: out")
    _require_vec(sx, inc.shape[0], "pde_pairwise: sx")
    _require_vec(sy, inc.shape[0], "pde_pairwise: sy")
    B, m, n = inc.shape
    rc = load().sigpde_pde_pairwise(dtype_code(inc.dtype), _ptr(inc), B, m, n, inc.stride(0),
                                    int(order), _ptr(sx), _ptr(sy), _ptr(out), _stream(inc.device))
    _check(rc, "sigpde_pde_pairwise")


def pde_gram(inc: torch.Tensor, order: int, out: torch.Tensor, off_x: int = 0, off_y: int = 0,
             sx=None, sy=None, symmetric: bool = False) -> None:
    """Solves the (bx, by) pairs of an increment tile into out[off_x + i, off_y + j]."""
    _require_gpu(inc, out, sx, sy)
    _require(inc.dim() == 4 and inc.stride(3) == 1 and inc.stride(2) == inc.shape[3],
             "pde_gram: increments must be (bx, by, m, n) with unit-stride rows packed m x n")
    _require(out.dim() == 2 and out.stride(1) == 1, "pde_gram: out must be 2-D with unit column stride")
    _require(off_x >= 0 and off_y >= 0 and off_x + inc.shape[0] <= out.shape[0]
             and off_y + inc.shape[1] <= out.shape[1], "pde_gram: tile does not fit in out")
    if symmetric:
        _require(off_y + inc.shape[1] <= out.shape[0] and off_x + inc.shape[0] <= out.shape[1],
                 "pde_gram: the mirrored tile does not fit in out")
    _require_vec(sx, inc.shape[0], "pde_gram: sx")
    _require_vec(sy, inc.shape[1], "pde_gram: sy")
    bx, by, m, n = inc.shape
    rc = load().sigpde_pde_gram(dtype_code(inc.dtype), _ptr(inc), bx, by, m, n, inc.stride(0),
                                inc.stride(1), int(order), _ptr(sx), _ptr(sy), int(symmetric),
                                int(off_x), int(off_y), _ptr(out), out.stride(0),
                                _stream(inc.device))
    _check(rc, "sigpde_pde_gram")


def robust_norm(inc: torch.Tensor, order: int, lam: torch.Tensor, knorm: torch.Tensor,
                target=None, maxit: int = 100, tol: float = 1e-8, iters=None) -> None:
    """Chandrupatla normalisation per path. target=None fuses the log normaliser and writes
    K(x,x) into knorm; otherwise knorm and target are inputs."""
    _require_gpu(inc, lam, knorm, target, iters)
    _require(inc.dim() == 3 and inc.shape[1] == inc.shape[2] and inc.stride(2) == 1
             and inc.stride(1) == inc.shape[2],
             "robust_norm: increments must be (B, m, m) with unit-stride rows packed m x m")
    for t, name in ((lam, "lambda"), (knorm, "knorm"), (target, "target"), (iters, "iters")):
        _require_vec(t, inc.shape[0], f"robust_norm: {name}")
    B, m, _ = inc.shape
    mode = NORM_LOG if target is None else NORM_GIVEN
    rc = load().sigpde_robust_norm(dtype_code(inc.dtype), _ptr(inc), B, m, inc.stride(0),
                                   int(order), mode, _ptr(target), _ptr(knorm), int(maxit),
                                   float(tol), _ptr(lam), _ptr(iters), _stream(inc.device))
    _check(rc, "sigpde_robust_norm")


def robust_norm_builtin(inc: torch.Tensor, order: int, lam: torch.Tensor, knorm: torch.Tensor,
                        normalizer: int, p0: float = 0.0, p1: float = 0.0, maxit: int = 100,
                        tol: float = 1e-8, iters=None) -> None:
    """Chandrupatla normalisation per path with a built-in normaliser evaluated on the device
    (NORM_LOG, NORM_TENSOR with C = p0, a = p1, NORM_CLIP with C = p0); knorm receives K(x,x)."""
    _require_gpu(inc, lam, knorm, iters)
    _require(inc.dim() == 3 and inc.shape[1] == inc.shape[2] and inc.stride(2) == 1
             and inc.stride(1) == inc.shape[2],
             "robust_norm: increments must be (B, m, m) with unit-stride rows packed m x m")
    for t, name in ((lam, "lambda"), (knorm, "knorm"), (iters, "iters")):
        _require_vec(t, inc.shape[0], f"robust_norm: {name}")
    B, m, _ = inc.shape
    rc = load().sigpde_robust_norm_builtin(dtype_code(inc.dtype), _ptr(inc), B, m, inc.stride(0),
                                           int(order), int(normalizer), float(p0), float(p1),
                                           _ptr(knorm), int(maxit), float(tol), _ptr(lam),
                                           _ptr(iters), _stream(inc.device))
    _check(rc, "sigpde_robust_norm_builtin")


def linear_workspace_bytes(precision: str, Bx: int, Lx: int, By: int, Ly: int, d: int,
                           same: bool) -> int:
    """Device workspace of the fused linear-kernel entry points (host-side size computation)."""
    return int(load().sigpde_linear_workspace_bytes(PRECISIONS[precision], Bx, Lx, By, Ly, d,
                                                     int(same)))


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def linear_gram(X: torch.Tensor, Y: torch.Tensor, order: int, kscale: float, out: torch.Tensor,
                precision: str = "fp16x2", sx=None, sy=None, symmetric: bool = False) -> None:
    """Fused linear-kernel Gram: out[i, j] = K(X_i, Y_j), increments on MFMA."""
    _require_gpu(X, Y, out, sx, sy)
    _require(X.dim() == 3 and Y.dim() == 3 and X.is_contiguous() and Y.is_contiguous(),
             "linear_gram: X and Y must be contiguous (batch, length, dim) tensors")
    _require(out.dim() == 2 and out.stride(1) == 1 and tuple(out.shape) == (X.shape[0], Y.shape[0]),
             "linear_gram: out must be (Bx, By) with unit column stride")
    _require_vec(sx, X.shape[0], "linear_gram: sx")
    _require_vec(sy, Y.shape[0], "linear_gram: sy")
    Bx, Lx, d = X.shape
    By, Ly, _ = Y.shape
    same = X.data_ptr() == Y.data_ptr() and X.shape == Y.shape
    L = load()
    prec = PRECISIONS[precision]
    nbytes = L.sigpde_linear_workspace_bytes(prec, Bx, Lx, By, Ly, d, int(same))
    ws = _workspace(nbytes, X.device)
    rc = L.sigpde_linear_gram(dtype_code(X.dtype), prec, _ptr(X), Bx, Lx, _ptr(Y), By, Ly, d,
                              float(kscale), int(order), _ptr(sx), _ptr(sy), int(symmetric),
                              _ptr(out), out.stride(0), _ptr(ws), nbytes, _stream(X.device))
    _check(rc, "sigpde_linear_gram")


MAX_GRAM_BATCH = 64


def linear_gram_batch(X: torch.Tensor, Y: torch.Tensor, requests, order: int, kscale: float,
                      out: torch.Tensor, precision: str = "fp16x2", sx=None, sy=None) -> None:
    """Several fused linear-kernel Grams of one shape class in one launch. requests: up to 64
    (x0, bx, y0, by, symmetric) rows over the paths of X and Y; out: a flat tensor of
    sum(bx * by) elements receiving each request's row-major (bx, by) result in turn."""
    _require_gpu(X, Y, out, sx, sy)
    _require(X.dim() == 3 and Y.dim() == 3 and X.is_contiguous() and Y.is_contiguous()
             and X.shape[2] == Y.shape[2],
             "linear_gram_batch: X and Y must be contiguous (batch, length, dim) tensors")
    reqs = [tuple(int(v) for v in q) for q in requests]
    _require(len(reqs) <= MAX_GRAM_BATCH, f"linear_gram_batch: at most {MAX_GRAM_BATCH} requests")
    total = sum(q[1] * q[3] for q in reqs)
    _require(out.is_contiguous() and out.numel() == total,
             f"linear_gram_batch: out must be contiguous with {total} elements")
    _require_vec(sx, X.shape[0], "linear_gram_batch: sx")
    _require_vec(sy, Y.shape[0], "linear_gram_batch: sy")
    nx, Lx, d = X.shape
    ny, Ly, _ = Y.shape
    L = load()
    prec = PRECISIONS[precision]
    nbytes = L.sigpde_linear_workspace_bytes(prec, nx, Lx, ny, Ly, d, 0)
    ws = _workspace(nbytes, X.device)
    table = (_i64 * (5 * max(len(reqs), 1)))(*[v for q in reqs for v in q])
    rc = L.sigpde_linear_gram_batch(dtype_code(X.dtype), prec, _ptr(X), nx, Lx, _ptr(Y), ny, Ly, d,
                                    float(kscale), int(order), table, len(reqs), _ptr(sx), _ptr(sy),
                                    _ptr(out), _ptr(ws), nbytes, _stream(X.device))
    _check(rc, "sigpde_linear_gram_batch")


def linear_pairwise(X: torch.Tensor, Y: torch.Tensor, order: int, kscale: float,
                    out: torch.Tensor, precision: str = "fp16x2", sx=None, sy=None) -> None:
    """Fused linear-kernel pairwise: out[b] = K(X_b, Y_b)."""
    _require_gpu(X, Y, out, sx, sy)
    _require(X.dim() == 3 and Y.dim() == 3 and X.is_contiguous() and Y.is_contiguous()
             and X.shape[0] == Y.shape[0],
             "linear_pairwise: X and Y must be contiguous (batch, length, dim) with one batch")
    _require_vec(out, X.shape[0], "linear_pairwise: out")
    _require_vec(sx, X.shape[0], "linear_pairwise: sx")
    _require_vec(sy, X.shape[0], "linear_pairwise: sy")
    B, Lx, d = X.shape
    Ly = Y.shape[1]
    same = X.data_ptr() == Y.data_ptr() and X.shape == Y.shape
    L = load()
    prec = PRECISIONS[precision]
    nbytes = L.sigpde_linear_workspace_bytes(prec, B, Lx, B, Ly, d, int(same))
    ws = _workspace(nbytes, X.device)
    rc = L.sigpde_linear_pairwise(dtype_code(X.dtype), prec, _ptr(X), _ptr(Y), B, Lx, Ly, d,
                                  float(kscale), int(order), _ptr(sx), _ptr(sy), _ptr(out),
                                  _ptr(ws), nbytes, _stream(X.device))
    _check(rc, "sigpde_linear_pairwise")


STATIC_LINEAR = 0
STATIC_RBF = 1
STATIC_L1RBF = 2


def static_gram_supported(kind: int, dtype: torch.dtype, Lx: int, Ly: int, d: int,
                          order: int) -> bool:
    """True when sigpde_static_gram takes this shape (dyadic order, dim and column limits)."""
    if dtype not in (torch.float32, torch.float64):
        return False
    return bool(load().sigpde_static_gram_supported(dtype_code(dtype), int(kind), int(Lx),
                                                    int(Ly), int(d), int(order)))


def static_gram(kind: int, X: torch.Tensor, Y: torch.Tensor, order: int, param: float,
                out: torch.Tensor, sx=None, sy=None, symmetric: bool = False) -> None:
    """Reference-scheme Gram of a static kernel (STATIC_LINEAR: param = scale; STATIC_RBF,
    STATIC_L1RBF: param = sigma) with the increments formed inside the solve."""
    _require_gpu(X, Y, out, sx, sy)
    _require(X.dim() == 3 and Y.dim() == 3 and X.is_contiguous() and Y.is_contiguous()
             and X.shape[2] == Y.shape[2],
             "static_gram: X and Y must be contiguous (batch, length, dim) tensors")
    _require(out.dim() == 2 and out.stride(1) == 1 and tuple(out.shape) == (X.shape[0], Y.shape[0])
             and out.dtype == X.dtype,
             "static_gram: out must be (Bx, By) of X's dtype with unit column stride")
    _require_vec(sx, X.shape[0], "static_gram: sx")
    _require_vec(sy, Y.shape[0], "static_gram: sy")
    Bx, Lx, d = X.shape
    By, Ly, _ = Y.shape
    same = X.data_ptr() == Y.data_ptr() and X.shape == Y.shape
    L = load()
    code = dtype_code(X.dtype)
    nbytes = L.sigpde_static_workspace_bytes(code, int(kind), Bx, Lx, By, Ly, d, int(same))
    ws = _workspace(max(nbytes, 1), X.device)
    rc = L.sigpde_static_gram(code, int(kind), _ptr(X), Bx, Lx, _ptr(Y), By, Ly, d, float(param),
                              int(order), _ptr(sx), _ptr(sy), int(symmetric), _ptr(out),
                              out.stride(0), _ptr(ws), nbytes, _stream(X.device))
    _check(rc, "sigpde_static_gram")


def rbf_increments(ip: torch.Tensor, ip_strides, A: int, B: int, m: int, n: int,
                   xx: torch.Tensor, xx_stride_a: int, yy: torch.Tensor, yy_stride_a: int,
                   yy_stride_b: int, sigma: float, order: int, out: torch.Tensor) -> None:
    """RBF increments (A, B, m, n) from the point inner products ip (element (a, b, i, j) at the
    given element strides) and the row sums of squares xx (a, i), yy (a, b, j)."""
    _require_gpu(ip, xx, yy, out)
    _require(out.is_contiguous() and out.numel() == A * B * m * n,
             "rbf_increments: out must be contiguous with A * B * m * n elements")
    strides = (_i64 * 4)(*[int(v) for v in ip_strides])
    rc = load().sigpde_rbf_increments(dtype_code(ip.dtype), _ptr(ip), A, B, m, n, strides,
                                      _ptr(xx), int(xx_stride_a), _ptr(yy), int(yy_stride_a),
                                      int(yy_stride_b), float(sigma), int(order), _ptr(out),
                                      _stream(ip.device))
    _check(rc, "sigpde_rbf_increments")


def l1rbf_increments(X: torch.Tensor, Y: torch.Tensor, sigma: float, order: int,
                     out: torch.Tensor, pairwise: bool) -> None:
    """L1-RBF increments from the paths: (A, m, n) pairwise (X_a with Y_a) or (A, B, m, n)."""
    _require_gpu(X, Y, out)
    _require(X.dim() == 3 and Y.dim() == 3 and X.is_contiguous() and Y.is_contiguous()
             and out.is_contiguous(), "l1rbf_increments: contiguous (batch, length, dim) inputs")
    _require(out.numel() == (X.shape[0] * (X.shape[1] - 1) * (Y.shape[1] - 1) if pairwise else
                             X.shape[0] * Y.shape[0] * (X.shape[1] - 1) * (Y.shape[1] - 1)),
             "l1rbf_increments: out has the wrong number of elements")
    A, Lx, d = X.shape
    B, Ly, _ = Y.shape
    rc = load().sigpde_l1rbf_increments(dtype_code(X.dtype), _ptr(X), A, _ptr(Y), B, Lx - 1,
                                        Ly - 1, d, int(pairwise), float(sigma), int(order),
                                        _ptr(out), _stream(X.device))
    _check(rc, "sigpde_l1rbf_increments")


def linear_increments(ip: torch.Tensor, ip_strides, A: int, B: int, m: int, n: int,
                      kscale: float, order: int, out: torch.Tensor) -> None:
    """Linear-kernel increments (A, B, m, n) of kscale * ip (element (a, b, i, j) at the given
    element strides): the reference scheme's second difference in one pass."""
    _require_gpu(ip, out)
    _require(out.is_contiguous() and out.numel() == A * B * m * n,
             "linear_increments: out must be contiguous with A * B * m * n elements")
    strides = (_i64 * 4)(*[int(v) for v in ip_strides])
    rc = load().sigpde_linear_increments(dtype_code(ip.dtype), _ptr(ip), A, B, m, n, strides,
                                         float(kscale), int(order), _ptr(out), _stream(ip.device))
    _check(rc, "sigpde_linear_increments")
