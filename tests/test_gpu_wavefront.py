"""GPU parity: the HIP wavefront kernels (through the C ABI) against the CPU oracle.

Tolerances: float64 1e-10 relative (the oracle is the reference scheme in float64; the kernels may
contract multiply-adds into FMAs, as NVVM does for the reference); float32 storage 1e-5 relative
(per-cell double arithmetic with float storage: FMA contraction can flip the float rounding of a
cell now and then).
"""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import max_rel, rel_fro, evals_with_fixed_point_exit
from sigpde_amd import _native

pytestmark = pytest.mark.gpu

DEV = "cuda"
TOL = {torch.float64: 1e-10, torch.float32: 1e-5}
NP = {torch.float64: np.float64, torch.float32: np.float32}


def rand_inc(shape, scale, seed, dtype):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=shape) * scale).astype(NP[dtype])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("B,m,n,order", [
    (8, 31, 31, 1),      # 8 pairs of length 32, order 1
    (5, 1, 1, 0),        # single cell
    (3, 7, 13, 0),       # m != n, tiny
    (4, 13, 7, 2),
    (6, 63, 63, 1),      # length 64, order 1
    (3, 99, 99, 2),      # length 100, order 2: 396 rows, one band of 7 rows per lane
    (2, 255, 255, 0),    # length 256, order 0
    (2, 1100, 40, 0),    # two bands (rows > 64 * 16)
    (2, 127, 90, 3),     # order 3, 1016 rows
    (1, 511, 20, 3),     # 4088 rows: four bands, the largest refined side
    (3, 64, 3, 0),       # rows exactly one band, few columns (< 64: padded band length)
])
def test_pairwise_from_increments(dtype, B, m, n, order):
    # refined-cell increments of an (m, n) grid of original ones: c / 4^order with c ~ 1.5/sqrt(mn),
    # which keeps K of order 1..1e3 at every size (no float32 overflow)
    scale = 1.5 / np.sqrt(m * n) / 4.0 ** order
    inc = rand_inc((B, m, n), scale, seed=m * 1000 + n + order, dtype=dtype)
    ref = O.pde_pairwise(inc, order)
    t = torch.from_numpy(inc).to(DEV)
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(t, order, out)
    got = out.cpu().numpy()
    assert rel_fro(got, ref) < TOL[dtype], (got[:4], ref[:4])
    assert max_rel(got, ref, floor=1e-3) < 10 * TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_pairwise_scaled(dtype):
    B, m, n, order = 7, 20, 25, 1
    inc = rand_inc((B, m, n), 0.2, 11, dtype)
    rng = np.random.default_rng(3)
    sx = rng.uniform(0.2, 1.0, B).astype(NP[dtype])
    sy = rng.uniform(0.2, 1.0, B).astype(NP[dtype])
    ref = O.pde_pairwise(inc, order, sx, sy)
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(torch.from_numpy(inc).to(DEV), order, out,
                         torch.from_numpy(sx).to(DEV), torch.from_numpy(sy).to(DEV))
    assert rel_fro(out.cpu().numpy(), ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("symmetric", [False, True])
def test_gram_from_increments(dtype, symmetric):
    bx, by, m, order = 6, 5, 17, 1
    n = m if symmetric else 23
    if symmetric:
        by = bx
    inc = rand_inc((bx, by, m, n), 0.25, 5, dtype)
    rng = np.random.default_rng(9)
    sx = rng.uniform(0.3, 1.0, bx).astype(NP[dtype])
    sy = rng.uniform(0.3, 1.0, by).astype(NP[dtype])
    # place the tile at an offset inside a larger output
    off_x, off_y = (3, 3) if symmetric else (2, 4)
    out = torch.full((bx + 5, by + 7), -7.0, dtype=dtype, device=DEV)
    _native.pde_gram(torch.from_numpy(inc).to(DEV), order, out, off_x, off_y,
                     torch.from_numpy(sx).to(DEV), torch.from_numpy(sy).to(DEV), symmetric)
    got = out.cpu().numpy()
    ref = O.pde_gram(inc, order, sx, sy, symmetric)
    blk = got[off_x:off_x + bx, off_y:off_y + by]
    if symmetric:
        lo = np.tril_indices(bx)
        assert rel_fro(blk[lo], ref[lo]) < TOL[dtype]
        assert np.array_equal(blk, blk.T)  # exact mirror
    else:
        assert rel_fro(blk, ref) < TOL[dtype]
    # nothing outside the tile (and its mirror) was written
    mask = np.ones_like(got, dtype=bool)
    mask[off_x:off_x + bx, off_y:off_y + by] = False
    if symmetric:
        mask[off_y:off_y + by, off_x:off_x + bx] = False
    assert np.all(got[mask] == -7.0)


def test_gram_symmetric_offdiagonal_tile_skips_upper():
    # tile with off_y > off_x: every pair lies above the diagonal, nothing is solved
    inc = rand_inc((3, 3, 5, 5), 0.2, 1, torch.float64)
    out = torch.zeros((10, 10), dtype=torch.float64, device=DEV)
    _native.pde_gram(torch.from_numpy(inc).to(DEV), 0, out, 0, 5, symmetric=True)
    assert torch.count_nonzero(out).item() == 0


def test_zero_increments_give_one():
    inc = torch.zeros((4, 30, 40), dtype=torch.float64, device=DEV)
    out = torch.zeros(4, dtype=torch.float64, device=DEV)
    _native.pde_pairwise(inc, 2, out)
    assert torch.all(out == 1.0)


def test_empty_batch_is_noop():
    inc = torch.zeros((0, 5, 5), dtype=torch.float32, device=DEV)
    out = torch.zeros(0, dtype=torch.float32, device=DEV)
    _native.pde_pairwise(inc, 0, out)


def test_bad_arguments_raise():
    inc = torch.zeros((2, 5, 5), dtype=torch.float32, device=DEV)
    out = torch.zeros(2, dtype=torch.float32, device=DEV)
    with pytest.raises(_native.NativeError, match="refined columns"):
        big = torch.zeros((1, 1, (1 << 22) + 1), dtype=torch.float32, device=DEV)
        _native.pde_pairwise(big, 0, out[:1])  # 2^22 + 1 refined columns
    with pytest.raises(TypeError):
        _native.pde_pairwise(inc.half(), 0, out)
    # layout checks are ValueErrors (they stay on under python -O)
    with pytest.raises(ValueError, match="unit-stride"):
        _native.pde_pairwise(inc.transpose(1, 2), 0, out)
    with pytest.raises(ValueError, match="out must be a contiguous vector"):
        _native.pde_pairwise(inc, 0, out[:1])
    with pytest.raises(ValueError, match="does not fit"):
        _native.pde_gram(inc.reshape(1, 2, 5, 5), 0, torch.zeros((1, 1), device=DEV))
    X = torch.zeros((2, 6, 3), dtype=torch.float32, device=DEV)
    with pytest.raises(ValueError, match="contiguous"):
        _native.linear_gram(X.transpose(0, 1), X, 0, 1.0, torch.zeros((6, 2), device=DEV))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_robust_norm_log(dtype):
    B, m, order = 16, 31, 1
    rng = np.random.default_rng(21)
    dx = rng.normal(size=(B, m, 6)) * 0.3
    inc64 = np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order
    inc = inc64.astype(NP[dtype])
    lam_ref, k_ref, it_ref, fx_ref = O.norm_factors(inc, order, tol=1e-10, maxit=100, fixed_points=True)
    t = torch.from_numpy(inc).to(DEV)
    lam = torch.zeros(B, dtype=dtype, device=DEV)
    kn = torch.zeros(B, dtype=dtype, device=DEV)
    iters = torch.zeros(B, dtype=torch.int32, device=DEV)
    _native.robust_norm(t, order, lam, kn, None, maxit=100, tol=1e-10, iters=iters)
    assert rel_fro(kn.cpu().numpy(), k_ref) < TOL[dtype]
    tol_lam = 1e-9 if dtype == torch.float64 else 1e-5
    assert max_rel(lam.cpu().numpy(), lam_ref, floor=1e-3) < tol_lam
    if dtype == torch.float64:
        assert np.array_equal(iters.cpu().numpy(), evals_with_fixed_point_exit(it_ref, fx_ref))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_nonfinite_inputs_follow_the_reference(dtype):
    """NaN increments propagate to NaN solves; in the root-find K(x,x) = NaN becomes +inf
    (sigpde/torch/kernel.py:183, sigpde/cuda_pairwise_kernels.py:132), the target is then the
    normaliser's limit 2, and lambda follows the same iterate sequence as the oracle. Huge increments
    overflow float32 to +inf the same way."""
    B, m, order = 4, 9, 1
    rng = np.random.default_rng(31)
    dx = rng.normal(size=(B, m, 3)) * 0.5
    inc = (np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order).astype(NP[dtype])
    inc[1, 3, 4] = np.nan
    if dtype == torch.float32:
        inc[2] *= 1e4  # K(x,x) overflows float32
    t = torch.from_numpy(inc).to(DEV)
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(t, order, out)
    ref = O.pde_pairwise(inc, order)
    got = out.cpu().numpy()
    assert np.isnan(got[1]) and np.isnan(ref[1])
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    assert rel_fro(got[fin], ref[fin]) < TOL[dtype]
    lam_ref, k_ref, it_ref, fx_ref = O.norm_factors(inc, order, tol=1e-10, maxit=30, fixed_points=True)
    lam = torch.zeros(B, dtype=dtype, device=DEV)
    kn = torch.zeros(B, dtype=dtype, device=DEV)
    iters = torch.zeros(B, dtype=torch.int32, device=DEV)
    _native.robust_norm(t, order, lam, kn, None, maxit=30, tol=1e-10, iters=iters)
    kn, lam = kn.cpu().numpy(), lam.cpu().numpy()
    assert np.isinf(kn[1]) and np.isinf(k_ref[1])
    assert np.array_equal(np.isinf(kn), np.isinf(k_ref))
    tol_lam = 1e-9 if dtype == torch.float64 else 1e-5
    assert max_rel(lam, lam_ref, floor=1e-3) < tol_lam
    if dtype == torch.float64:
        assert np.array_equal(iters.cpu().numpy(), evals_with_fixed_point_exit(it_ref, fx_ref))


def test_robust_norm_given_target():
    B, m, order = 8, 20, 0
    rng = np.random.default_rng(5)
    dx = rng.normal(size=(B, m, 4)) * 0.4
    inc = np.einsum("bsk,btk->bst", dx, dx)
    k = O.pde_pairwise(inc, order)
    target = 2 - 1 / np.sqrt(k)  # custom normaliser, f(x) <= x for x >= 1
    lam_ref, _ = O.chandrupatla(inc, order, target, k, 100, 1e-10)
    lam_ref[k < 1] = 0
    lam = torch.zeros(B, dtype=torch.float64, device=DEV)
    kn = torch.from_numpy(k).to(DEV)
    _native.robust_norm(torch.from_numpy(inc).to(DEV), order, lam, kn,
                        torch.from_numpy(target).to(DEV), maxit=100, tol=1e-10)
    assert max_rel(lam.cpu().numpy(), lam_ref, floor=1e-3) < 1e-9
