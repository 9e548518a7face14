"""GPU accuracy of the fused linear-kernel paths against the float64 oracle run on the same float32
inputs upcast:
  precision="fp16x2" (fp16 hi/lo increment Gram on MFMA, double wavefront): the headline bar,
    relative Frobenius error <= 1e-5 (measured ~3e-7), and 1e-4 on every entry whose magnitude is
    at least 1% of the largest;
  precision="bf16x3" (three bf16 parts, six bf16 MFMA products, double wavefront): same bars as
    fp16x2;
  precision="fp32" (plain FP32: fp32 increments on MFMA, the reference's float32 wavefront
    typing): float32 storage of every cell bounds it; bar 1e-4 Frobenius / 1e-3 per entry;
  precision="fp64" (double increments on fp64 MFMA, double wavefront, dim <= 64): the float32
    rounding of the result bounds it on float32 inputs (bar 2e-7); on float64 inputs it is the
    float64 reference scheme (bar 1e-12);
  precision="fp16x2f" (the fp16x2 increments, float32 wavefront, refined grids up to 512 x 512):
    the headline bar, 1e-5 Frobenius (measured 2e-6..9e-6 on the CPU model of its arithmetic).
"""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from helpers import brownian_np, rel_fro
from sigpde_amd import _native

pytestmark = pytest.mark.gpu
DEV = "cuda"
BARS = {"fp16x2": 1e-5, "bf16x3": 1e-5, "fp32": 1e-4, "fp64": 2e-7, "fp16x2f": 1e-5}
PRECS = list(BARS)


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def check(got, ref, prec="fp16x2", bar=None):
    bar = BARS[prec] if bar is None else bar
    assert rel_fro(got, ref) < bar, rel_fro(got, ref)
    big = np.abs(ref) >= 0.01 * np.abs(ref).max()
    assert np.max(np.abs(got - ref)[big] / np.abs(ref)[big]) < 10 * bar


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("L,d,order,var", [
    (256, 64, 0, 1 / 256),   # the 1024 x 1024 workload's path shape
    (100, 8, 2, 1 / 100),    # length 100, dim 8, order 2 (d padded to 32)
    (64, 16, 1, 9 / 64),     # length 64, dim 16, order 1, large K values
    (33, 4, 3, 1 / 33),      # order 3
    (20, 100, 0, 1 / 20),    # d = 100: four K blocks
    (513, 4, 0, 1 / 512),    # 512 increments: 32 bands of 16 rows
])
def test_fused_gram_vs_f64(L, d, order, var, prec):
    x = brownian_np(9, L, d, var, 5, np.float32)
    y = brownian_np(7, L, d, var, 6, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(), order, precision=prec)
    if prec == "fp64" and d > 64:
        with pytest.raises(_native.NativeError, match="dim <= 64"):
            k.gram(cuda(x), cuda(y))
        return
    if prec == "fp16x2f" and order > 2:
        with pytest.raises(_native.NativeError, match="dyadic order <= 2"):
            k.gram(cuda(x), cuda(y))
        return
    got = k.gram(cuda(x), cuda(y)).cpu().numpy().astype(np.float64)
    ref = O.linear_gram(x.astype(np.float64), y.astype(np.float64), order)
    check(got, ref, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_fused_gram_unequal_lengths_and_kscale(prec):
    x = brownian_np(6, 40, 8, 1 / 40, 1, np.float32)
    y = brownian_np(5, 71, 8, 1 / 70, 2, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(0.7), 1, precision=prec)
    got = k.gram(cuda(x), cuda(y)).cpu().numpy()
    ref = O.linear_gram(x.astype(np.float64), y.astype(np.float64), 1, kscale=0.7)
    check(got, ref, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_fused_gram_symmetric_and_scaled(prec):
    x = brownian_np(11, 50, 16, 1 / 50, 3, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(), 1, precision=prec)
    got = k.gram(cuda(x)).cpu().numpy()
    x64 = x.astype(np.float64)
    ref = O.linear_gram(x64, x64, 1)
    check(got, ref, prec)
    assert np.array_equal(got, got.T)
    s = np.linspace(0.3, 1.0, 11).astype(np.float32)
    got = k.gram(cuda(x), x_scale=cuda(s)).cpu().numpy()
    ref = O.linear_gram(x64, x64, 1, sx=s.astype(np.float64), sy=s.astype(np.float64))
    check(got, ref, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_fused_gram_float64_inputs(prec):
    x = brownian_np(6, 30, 4, 1 / 30, 8)
    y = brownian_np(6, 30, 4, 1 / 30, 9)
    k = sig.SigPDE(sig.kernels.LinearKernel(), 1, precision=prec)
    got = k.gram(cuda(x), cuda(y))
    assert got.dtype == torch.float64
    check(got.cpu().numpy(), O.linear_gram(x, y, 1), prec)
    if prec == "fp64":
        assert rel_fro(got.cpu().numpy(), O.linear_gram(x, y, 1)) < 1e-12


@pytest.mark.parametrize("prec", PRECS)
def test_fused_pairwise_vs_f64(prec):
    x = brownian_np(13, 90, 16, 1 / 90, 4, np.float32)
    y = brownian_np(13, 60, 16, 1 / 60, 5, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(), 2, precision=prec)
    got = k.pairwise(cuda(x), cuda(y)).cpu().numpy()
    ref = np.array([O.linear_gram(x[b:b + 1].astype(np.float64), y[b:b + 1].astype(np.float64), 2)[0, 0]
                    for b in range(13)])
    check(got, ref, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_fused_gram_many_pairs_matches_small_calls(prec):
    # results do not depend on how pairs are grouped into waves and streams
    x = brownian_np(70, 24, 4, 1 / 24, 10, np.float32)
    y = brownian_np(37, 24, 4, 1 / 24, 11, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(), 1, precision=prec)
    big = k.gram(cuda(x), cuda(y)).cpu().numpy()
    small = k.gram(cuda(x[5:9]), cuda(y[30:31])).cpu().numpy()
    assert np.array_equal(big[5:9, 30:31], small)


def test_fused_fp32_matches_fp32_reference_scheme_on_exact_increments():
    # Inputs on a dyadic grid make every increment product exact in fp32, so the fp32 fused path
    # and the reference-scheme path ("ref": point Gram, 4-term difference, float32 wavefront) see
    # the same increments; both then run the float32 wavefront typing and must agree to the
    # wavefront's own reassociation noise.
    rng = np.random.default_rng(12)
    x = (np.cumsum(rng.integers(-8, 9, size=(5, 33, 4)), axis=1) / 64.0).astype(np.float32)
    y = (np.cumsum(rng.integers(-8, 9, size=(4, 41, 4)), axis=1) / 64.0).astype(np.float32)
    x[:, 0] = 0
    y[:, 0] = 0
    fused = sig.SigPDE(sig.kernels.LinearKernel(), 1, precision="fp32").gram(cuda(x), cuda(y))
    ref = sig.SigPDE(sig.kernels.LinearKernel(), 1, precision="ref").gram(cuda(x), cuda(y))
    np.testing.assert_allclose(fused.cpu().numpy(), ref.cpu().numpy(), rtol=2e-6)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("L,order", [(2001, 0), (1001, 2), (301, 3)])
def test_fused_long_rows_of_columns(prec, L, order):
    """2000 / 4000 / 2400 refined columns: four LDS boundary lines that long would let fewer waves
    share a CU than the kernels' registers allow, so the launches keep their lines in global
    memory (the same arithmetic, the same results bit for bit as lines in LDS). Gram, symmetric
    Gram and pairwise against the oracle. fp16x2 and bf16x3 keep the headline bar of 1e-5
    (measured ~2e-7 at 2000 increments: their products of each magnitude are accumulated apart, so
    no rounding bias builds up along a path); fp32 gets 3e-4 (its float32 wavefront leaves ~2e-4
    absolute error on entries near 0.1 of paths this long); fp64 keeps 2e-7."""
    x = brownian_np(5, L, 4, 1 / L, 21, np.float32)
    y = brownian_np(3, L, 4, 1 / L, 22, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(), order, precision=prec)
    if prec == "fp16x2f":  # past its 512 refined columns / order 2
        with pytest.raises(_native.NativeError, match="fp16x2f"):
            k.gram(cuda(x), cuda(y))
        return
    bar = {"fp16x2": 1e-5, "bf16x3": 1e-5, "fp32": 3e-4, "fp64": 2e-7}[prec]
    x64, y64 = x.astype(np.float64), y.astype(np.float64)
    ref = O.linear_gram(x64, y64, order)
    check(k.gram(cuda(x), cuda(y)).cpu().numpy(), ref, prec, bar)
    check(k.pairwise(cuda(x[:3]), cuda(y)).cpu().numpy(), np.diag(ref[:3, :3]).copy(), prec, bar)
    got = k.gram(cuda(x)).cpu().numpy()
    check(got, O.linear_gram(x64, x64, order), prec, bar)
    assert np.array_equal(got, got.T)


@pytest.mark.parametrize("prec", PRECS)
def test_fused_rejects_nonlinear_kernel(prec):
    with pytest.raises(ValueError):
        sig.SigPDE(sig.kernels.RBFKernel(), 0, precision=prec)


@pytest.mark.parametrize("prec", ["fp16x2", "bf16x3", "fp64", "fp16x2f"])
def test_fused_full_1024x1024_gram_vs_gpu_float64(prec):
    """Every entry of the 1024 x 1024 Gram (length 256, dim 64, order 0) against the float64 GPU
    path (precision="ref" on the float64 inputs: the reference scheme in float64, pinned to the CPU
    oracle at 1e-10 by test_gpu_wavefront.py / test_gpu_api.py); the oracle itself is checked on a
    16 x 16 corner here as well."""
    from sigpde_amd.synthetic import brownian_paths
    X = brownian_paths(1024, 256, 64, 1 / 256, 5, torch.float32, DEV)
    Y = brownian_paths(1024, 256, 64, 1 / 256, 6, torch.float32, DEV)
    lin = sig.kernels.LinearKernel()
    K = sig.SigPDE(lin, 0, precision=prec).gram(X, Y)
    K64 = sig.SigPDE(lin, 0).gram(X.double(), Y.double(), max_batch=128 * 128)
    ref_corner = O.linear_gram(X[:16].double().cpu().numpy(), Y[:16].double().cpu().numpy(), 0)
    assert rel_fro(K64[:16, :16].cpu().numpy(), ref_corner) < 1e-12
    rel = float(torch.linalg.norm(K.double() - K64) / torch.linalg.norm(K64))
    assert rel < BARS[prec], rel


def test_fp16x2f_refuses_grids_past_512_refined_cells_a_side():
    x = brownian_np(2, 130, 4, 1 / 130, 3, np.float32)   # 129 << 2 = 516 refined rows
    k = sig.SigPDE(sig.kernels.LinearKernel(), 2, precision="fp16x2f")
    with pytest.raises(_native.NativeError, match="at most 512 refined"):
        k.gram(cuda(x), cuda(x))
    with pytest.raises(_native.NativeError, match="dyadic order <= 2"):
        sig.SigPDE(sig.kernels.LinearKernel(), 3, precision="fp16x2f").gram(cuda(x[:, :33]))
    # the same paths at order 1 (258 refined rows) run
    k1 = sig.SigPDE(sig.kernels.LinearKernel(), 1, precision="fp16x2f")
    got = k1.gram(cuda(x)).cpu().numpy().astype(np.float64)
    ref = O.linear_gram(x.astype(np.float64), x.astype(np.float64), 1)
    check(got, ref, "fp16x2f")


@pytest.mark.parametrize("order,L,d", [(0, 256, 64), (1, 129, 16), (2, 65, 8), (0, 40, 100)])
def test_fp16x2f_wave_pair_gram_equals_single_wave_pairwise(order, L, d):
    """The fp16x2f Gram (consumer / producer wave pairs sharing y operands) and the fp16x2f
    pairwise call (one wave per four pairs) form the same increment tiles in the same MFMA order
    and run the same float32 cell arithmetic, so the Gram's diagonal equals the pairwise results
    bit for bit."""
    x = brownian_np(21, L, d, 1 / L, 31, np.float32)
    y = brownian_np(21, L, d, 1 / L, 32, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(0.9), order, precision="fp16x2f")
    K = k.gram(cuda(x), cuda(y))
    p = k.pairwise(cuda(x), cuda(y))
    assert torch.equal(torch.diagonal(K), p)
