"""Bit-exact parity of the reference-scheme kernels with the oracle's contracted setting.

The reference-scheme arithmetic modes (float32 typing and float64) round every cell as the oracle
does under NVVM's contraction of pde_step (oracle.contracted(); the typing rules and both
contraction settings are pinned by tests/test_oracle_typing.py against exact rational
arithmetic), and the root-find's bracket arithmetic is uncontracted on both sides. Given the same
increments, every route of the library (64-lane groups, 16-lane groups for >= 2048 pairs, the
multi-wave route for few long pairs, boundary lines in global memory, Gram tiles, symmetric tiles,
scaled variants, the root-find) must therefore return the oracle's values bit for bit, in float32
and in float64. Through the public API, inputs whose point Gram is exact in float32 (quantized
paths) make the increments exact too, so the whole `precision="ref"` call on the 128 x 128 Gram of
length 100, dim 8, dyadic order 2 equals the oracle exactly.
"""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from helpers import evals_with_fixed_point_exit, quantized_paths, rel_fro
from sigpde_amd import _native

pytestmark = pytest.mark.gpu
DEV = "cuda"
NP = {torch.float64: np.float64, torch.float32: np.float32}
DTYPES = [torch.float64, torch.float32]


@pytest.fixture(autouse=True)
def contracted_oracle():
    with O.contracted(True):
        yield


def rand_inc(shape, seed, dtype, order=0):
    rng = np.random.default_rng(seed)
    m, n = shape[-2], shape[-1]
    return (rng.normal(size=shape) * (1.5 / np.sqrt(m * n)) / 4.0 ** order).astype(NP[dtype])


def same(got, ref):
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.dtype == ref.dtype and got.shape == ref.shape
    bad = ~((got == ref) | (np.isnan(got) & np.isnan(ref)))
    assert not bad.any(), (f"{int(bad.sum())} of {got.size} differ; first at "
                           f"{np.argwhere(bad)[0].tolist()}: {got[bad][:3]} vs {ref[bad][:3]}")


def gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("B,m,n,order", [
    (8, 31, 31, 1),      # one pair per wave (64-lane groups)
    (5, 1, 1, 0),        # single cell
    (3, 99, 99, 2),      # length 100, order 2
    (2, 127, 90, 3),     # order 3
    (2053, 17, 40, 0),   # >= 2048 pairs: 16-lane groups, two bands
    (2051, 33, 20, 1),   # 16-lane groups, refined
    (2050, 9, 7, 3),     # 16-lane groups, eight rows per lane
    (4, 1100, 40, 0),    # few long pairs: multi-wave workgroups
    (2, 600, 200, 1),    # multi-wave, two rows per lane
    (3, 2, 17696, 0),    # boundary lines in global memory
])
def test_pairwise_routes(dtype, B, m, n, order):
    inc = rand_inc((B, m, n), 7 * m + n + order + B, dtype, order)
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(gpu(inc), order, out)
    same(out.cpu().numpy(), O.pde_pairwise(inc, order))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("B,m,n,order", [(7, 20, 25, 1), (2049, 12, 9, 1), (3, 1030, 33, 1)])
def test_pairwise_scaled_routes(dtype, B, m, n, order):
    inc = rand_inc((B, m, n), 11 + B, dtype, order)
    rng = np.random.default_rng(3)
    sx = rng.uniform(0.2, 1.0, B).astype(NP[dtype])
    sy = rng.uniform(0.2, 1.0, B).astype(NP[dtype])
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(gpu(inc), order, out, gpu(sx), gpu(sy))
    same(out.cpu().numpy(), O.pde_pairwise(inc, order, sx, sy))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("bx,by,m,n,order,symmetric", [
    (6, 5, 17, 23, 1, False),
    (6, 6, 17, 17, 1, True),
    (48, 47, 12, 9, 0, False),     # 2256 pairs: 16-lane groups
    (50, 50, 10, 10, 2, True),     # symmetric, 16-lane groups
])
def test_gram_tile_routes(dtype, bx, by, m, n, order, symmetric):
    inc = rand_inc((bx, by, m, n), 5 + bx, dtype, order)
    rng = np.random.default_rng(9)
    sx = rng.uniform(0.3, 1.0, bx).astype(NP[dtype])
    sy = rng.uniform(0.3, 1.0, by).astype(NP[dtype])
    if symmetric:
        sy = sx
    out = torch.zeros((bx, by), dtype=dtype, device=DEV)
    _native.pde_gram(gpu(inc), order, out, 0, 0, gpu(sx), gpu(sy), symmetric)
    same(out.cpu().numpy(), O.pde_gram(inc, order, sx, sy, symmetric))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("B,m,order", [(16, 31, 1), (2050, 12, 1), (3, 300, 0), (2, 150, 2)])
def test_root_find_given_target(dtype, B, m, order):
    """K(x,x) equal bit for bit; with the same target (a custom normaliser computed on the host),
    every iterate and lambda equal the oracle's bit for bit, and the evaluation count is the
    oracle's, or its first fixed point of the iteration, where the device stops."""
    rng = np.random.default_rng(B + m)
    dx = rng.normal(size=(B, m, 5)) * (1.2 / np.sqrt(m))
    inc = (np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order).astype(NP[dtype])
    kn = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(gpu(inc), order, kn)
    k = kn.cpu().numpy()
    k_ref = O.pde_pairwise(inc, order)
    same(k, k_ref)
    with np.errstate(all="ignore"):
        target = (2 - 1 / np.power(np.maximum(k, 1), 0.25)).astype(NP[dtype])
    lam = torch.zeros(B, dtype=dtype, device=DEV)
    iters = torch.zeros(B, dtype=torch.int32, device=DEV)
    _native.robust_norm(gpu(inc), order, lam, kn, gpu(target), maxit=60, tol=1e-10, iters=iters)
    lam_ref, it_ref, fx_ref = O.chandrupatla(inc, order, target, k, 60, 1e-10, fixed_points=True)
    lam_ref[k < 1] = 0
    same(lam.cpu().numpy(), lam_ref)
    # the device stops at a fixed point of the iteration (lambda unchanged); the oracle runs on
    same(iters.cpu().numpy(), evals_with_fixed_point_exit(it_ref, fx_ref))
    if dtype == torch.float32:
        # float32 function values cannot reach tol 1e-10: paths end at a fixed point, and lambda
        # is still the oracle's 60-evaluation result bit for bit
        assert (fx_ref >= 0).any() and (fx_ref < 60).any()


@pytest.mark.parametrize("case", [
    ("-0x1.5aa548p+0",), ("-0x1.ec28e0p-1",), ("-0x1.5bd320p-1", "-0x1.4840acp-1")])
def test_typing_cells(case):
    """The float32 typing known answers (c**2 in float32, the rest in float64, float32 storage)
    through the GPU: each cell lands on the oracle's value, which equals the exact model."""
    inc = np.array([[[float.fromhex(v) for v in case]]], dtype=np.float32)
    out = torch.zeros(1, dtype=torch.float32, device=DEV)
    _native.pde_pairwise(gpu(inc), 0, out)
    same(out.cpu().numpy(), O.pde_pairwise(inc, 0))


def test_scaled_typing_cell():
    inc = np.array([[[float.fromhex("-0x1.d87866p-1")]]], dtype=np.float32)
    sx = np.array([float.fromhex("0x1.81db3ap-1")], dtype=np.float32)
    sy = np.array([float.fromhex("0x1.712314p-2")], dtype=np.float32)
    out = torch.zeros(1, dtype=torch.float32, device=DEV)
    _native.pde_pairwise(gpu(inc), 0, out, gpu(sx), gpu(sy))
    same(out.cpu().numpy(), O.pde_pairwise(inc, 0, sx, sy))


def test_api_gram_f32_128x128_length100_order2_quantized():
    """SigPDE(LinearKernel(), 2).gram(X, Y), float32, X and Y 128 paths of length 100, dim 8, on
    quantized paths (exact point Gram): equal to the oracle's float32 semantics bit for bit; the
    as-written contraction setting differs by at most a float32 rounding here and there."""
    X = quantized_paths(128, 100, 8, 2)
    Y = quantized_paths(128, 100, 8, 3)
    k = sig.SigPDE(sig.kernels.LinearKernel(), 2)
    got = k.gram(gpu(X), gpu(Y)).cpu().numpy()
    ref = O.linear_gram(X, Y, 2)
    same(got, ref)
    with O.contracted(False):
        ref_written = O.linear_gram(X, Y, 2)
    assert rel_fro(got, ref_written) < 1e-7
    # the fused plain-FP32 path forms the same exact increments on fp32 MFMA
    got32 = k.gram(gpu(X), gpu(Y), precision="fp32").cpu().numpy()
    same(got32, ref)
    # symmetric Gram of X with the default max_batch tiles (ceil(sqrt(1000)) = 32)
    gs = k.gram(gpu(X)).cpu().numpy()
    same(gs, O.linear_gram(X, X, 2))


def test_api_gram_f64_quantized_ref_and_fused_fp64():
    X = quantized_paths(40, 64, 16, 4, dtype=np.float64)
    Y = quantized_paths(36, 50, 16, 5, dtype=np.float64)
    k = sig.SigPDE(sig.kernels.LinearKernel(), 1)
    ref = O.linear_gram(X, Y, 1)
    same(k.gram(gpu(X), gpu(Y)).cpu().numpy(), ref)
    same(k.gram(gpu(X), gpu(Y), precision="fp64").cpu().numpy(), ref)
