"""GPU checks of the fused reference-scheme static-kernel Gram (sigpde_static_gram: the point values
of the linear / RBF / L1-RBF kernels and their 4-term second difference formed inside the solve,
sigpde_amd/csrc/point_producer.h), which SigPDE.gram takes in precision "ref":
  * against the unfused route (torch GEMM point inner products -> one-pass increments -> generic
    wavefront; every op of which is pinned by test_gpu_static_increments.py and the wavefront
    tests): equal bit for bit wherever the inner products and squared norms are exact in the
    dtype (quantized paths), since every later float operation is the same in both routes; for
    L1-RBF on any paths (both sum |x_k - y_k| over k in order);
  * against the float64 oracle (sigpde/torch/static_kernels.py, kernel.py:18-21 and the Gram
    solvers, restated in oracle/) at 1e-10 on Brownian float64 paths;
  * the reference's shape for the static kernels' float32 Gram, 128 x 128 paths of length 100,
    dim 8, dyadic order 2;
  * symmetric and scaled Grams, ragged batches, unequal lengths, and the route's limits (larger
    shapes fall back to the unfused route, the C ABI refuses them).
"""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from helpers import brownian_np, quantized_paths, rel_fro
from sigpde_amd import _native

pytestmark = pytest.mark.gpu
DEV = "cuda"

KERNELS = {
    "linear": (lambda p: sig.kernels.LinearKernel(p), 0.7),
    "rbf": (lambda p: sig.kernels.RBFKernel(p), 1.3),
    "l1rbf": (lambda p: sig.kernels.L1RBFKernel(p), 2.5),
}


def qpaths(B, L, d, seed, dtype):
    return torch.from_numpy(quantized_paths(B, L, d, seed, dtype=np.float64)).to(DEV, dtype)


def bpaths(B, L, d, seed, dtype):
    return torch.from_numpy(brownian_np(B, L, d, 1.0 / L, seed)).to(DEV, dtype)


def unfused(k, x, y, order, sx=None, sy=None, symmetric=False):
    """The tile route: torch increments (one pass) and the generic Gram wavefront."""
    S = sig.SigPDE(k, order)
    if symmetric:
        return S._gram_sym(x, sx, 1 << 20)
    return S._gram(x, y, sx, sy, 1 << 20)


def fused(k, x, y, order, sx=None, sy=None, symmetric=False):
    kind, param = k.fused_static
    if symmetric:
        y, sy = x, sx
    out = torch.full((x.shape[0], y.shape[0]), -7.0, dtype=x.dtype, device=DEV)
    _native.static_gram(kind, x, y, order, param, out, sx, sy, symmetric)
    return out


@pytest.mark.parametrize("name", ["linear", "rbf", "l1rbf"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_fused_equals_unfused_on_exact_inputs(name, dtype, order):
    make, p = KERNELS[name]
    k = make(p)
    dims = (3, 8) if name == "l1rbf" else (3, 8, 21)
    for d in dims:
        if not _native.static_gram_supported(k.fused_static[0], dtype, 37, 21, d, order):
            continue
        x, y = qpaths(7, 37, d, 10 + d, dtype), qpaths(5, 21, d, 20 + d, dtype)
        print(f"static fused {name} {dtype} order {order} d {d}", flush=True)
        a = fused(k, x, y, order)
        b = unfused(k, x, y, order)
        assert torch.equal(a, b), (name, d, float((a - b).abs().max()))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("order", [0, 2])
def test_fused_l1rbf_equals_unfused_on_brownian_paths(dtype, order):
    k = sig.kernels.L1RBFKernel(0.9)
    x, y = bpaths(9, 30, 5, 1, dtype), bpaths(6, 45, 5, 2, dtype)
    assert torch.equal(fused(k, x, y, order), unfused(k, x, y, order))


@pytest.mark.parametrize("name", ["linear", "rbf", "l1rbf"])
@pytest.mark.parametrize("order", [0, 1, 3])
def test_fused_float64_vs_oracle(name, order):
    make, p = KERNELS[name]
    k = make(p)
    x, y = bpaths(6, 26, 4, 3, torch.float64), bpaths(5, 33, 4, 4, torch.float64)
    got = fused(k, x, y, order).cpu().numpy()
    ref = O.sig_gram(x.cpu().numpy(), y.cpu().numpy(), order, name, p)
    assert rel_fro(got, ref) < 1e-10


@pytest.mark.parametrize("name", ["linear", "rbf", "l1rbf"])
def test_fused_float32_brownian_close_to_unfused_and_oracle(name):
    make, p = KERNELS[name]
    k = make(p)
    x, y = bpaths(8, 50, 6, 5, torch.float32), bpaths(7, 40, 6, 6, torch.float32)
    a = fused(k, x, y, 1)
    b = unfused(k, x, y, 1)
    assert rel_fro(a.cpu().numpy(), b.cpu().numpy()) < 1e-5
    ref = O.sig_gram(x.double().cpu().numpy(), y.double().cpu().numpy(), 1, name, p)
    assert rel_fro(a.cpu().numpy(), ref) < 1e-4


@pytest.mark.parametrize("name", ["linear", "rbf", "l1rbf"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_symmetric_and_scaled(name, dtype):
    make, p = KERNELS[name]
    k = make(p)
    x = qpaths(11, 29, 4, 7, dtype)
    s = torch.linspace(0.5, 1.25, 11, device=DEV, dtype=dtype)
    a = fused(k, x, None, 2, s, None, symmetric=True)
    b = unfused(k, x, None, 2, s, None, symmetric=True)
    assert torch.equal(a, b)
    assert torch.equal(a, a.T)
    y = qpaths(6, 17, 4, 8, dtype)
    sy = torch.linspace(1.5, 0.75, 6, device=DEV, dtype=dtype)
    assert torch.equal(fused(k, x, y, 1, s, sy), unfused(k, x, y, 1, s, sy))


@pytest.mark.parametrize("name", ["linear", "rbf", "l1rbf"])
def test_reference_shape_128x128_length100_dim8_order2(name):
    """The reference's float32 static-kernel Gram shape through SigPDE.gram (fused route)."""
    make, p = KERNELS[name]
    k = make(p)
    x, y = qpaths(128, 100, 8, 11, torch.float32), qpaths(128, 100, 8, 12, torch.float32)
    S = sig.SigPDE(k, 2)
    assert S._static_route(x, y) is not None
    a = S.gram(x, y)
    assert torch.equal(a, unfused(k, x, y, 2))
    # Brownian paths of the same shape against the float64 oracle on a corner
    xb, yb = bpaths(128, 100, 8, 13, torch.float32), bpaths(128, 100, 8, 14, torch.float32)
    g = S.gram(xb, yb)[:6, :5].double().cpu().numpy()
    ref = O.sig_gram(xb[:6].double().cpu().numpy(), yb[:5].double().cpu().numpy(), 2, name, p)
    assert rel_fro(g, ref) < 1e-4


def test_ragged_and_small_shapes():
    k = sig.kernels.RBFKernel(0.8)
    for (bx, by, lx, ly, d) in [(1, 1, 2, 2, 1), (1, 9, 2, 40, 2), (13, 3, 17, 2, 5),
                                (4, 4, 3, 3, 64), (5, 1, 64, 65, 9)]:
        x, y = qpaths(bx, lx, d, bx + lx, torch.float32), qpaths(by, ly, d, by + ly, torch.float32)
        assert torch.equal(fused(k, x, y, 1), unfused(k, x, y, 1)), (bx, by, lx, ly, d)


def test_api_routes_and_limits():
    x = qpaths(5, 20, 4, 1, torch.float32)
    lin = sig.SigPDE(sig.kernels.LinearKernel(), 0)
    assert lin._static_route(x, None) is not None
    # past the route (order 4, dim 80, a subclass with its own gram): the unfused route
    assert sig.SigPDE(sig.kernels.LinearKernel(), 4)._static_route(x, None) is None
    assert lin._static_route(qpaths(2, 20, 80, 2, torch.float32), None) is None

    class MyRBF(sig.kernels.RBFKernel):
        def gram(self, X, Y):
            return super().gram(X, Y)
    assert MyRBF(1.0).fused_static is None
    assert torch.equal(sig.SigPDE(MyRBF(1.0), 1).gram(x), unfused(sig.kernels.RBFKernel(1.0), x,
                                                                   None, 1, symmetric=True))
    # the C ABI refuses what the route does not take
    out = torch.empty((5, 5), device=DEV)
    with pytest.raises(_native.NativeError, match="dyadic order"):
        _native.static_gram(_native.STATIC_RBF, x, x, 4, 1.0, out)
    with pytest.raises(_native.NativeError, match="dim"):
        _native.static_gram(_native.STATIC_L1RBF, qpaths(5, 20, 17, 3, torch.float32),
                            qpaths(5, 20, 17, 3, torch.float32), 0, 1.0, out)
    with pytest.raises(_native.NativeError, match="sigma"):
        _native.static_gram(_native.STATIC_RBF, x, x, 0, 0.0, out)
