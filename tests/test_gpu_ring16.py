"""GPU parity of the four-pairs-per-wave kernels (lane groups of 16), which the library picks for
batches of at least 2048 pairs or paths: pairwise, Gram (full, symmetric, offset tiles, scaled) and
the robust root-find, from increments through the C ABI, against the CPU oracle.

Small grids keep the oracle fast while covering the band / stream geometry: one band (M <= 16),
exactly one band (M = 16), several bands, a band length at its minimum (N + 1 <= 32), dyadic
refinement, m != n. Tolerances as in test_gpu_wavefront.py: float64 1e-10, float32 1e-5.
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
MANY = 2048  # the library's threshold for four pairs per wave


def rand_inc(shape, scale, seed, dtype):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=shape) * scale).astype(NP[dtype])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,n,order", [
    (1, 1, 0),     # single cell
    (5, 7, 0),     # one band, minimum band length
    (16, 16, 0),   # exactly one band of 16 rows
    (17, 40, 0),   # two bands, second of one row
    (15, 3, 1),    # refined 30 x 6
    (33, 20, 1),   # refined 66 x 40: five bands
    (70, 100, 0),  # wide
])
def test_pairwise_many(dtype, m, n, order):
    B = MANY + 5  # a partial last wave
    scale = 1.5 / np.sqrt(m * n) / 4.0 ** order
    inc = rand_inc((B, m, n), scale, seed=7 * m + n + order, dtype=dtype)
    ref = O.pde_pairwise(inc, order)
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(torch.from_numpy(inc).to(DEV), order, out)
    got = out.cpu().numpy()
    assert rel_fro(got, ref) < TOL[dtype]
    assert max_rel(got, ref, floor=1e-3) < 10 * TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_pairwise_many_scaled(dtype):
    B, m, n, order = MANY + 3, 12, 9, 1
    inc = rand_inc((B, m, n), 0.15, 31, dtype)
    rng = np.random.default_rng(4)
    sx = rng.uniform(0.2, 1.0, B).astype(NP[dtype])
    sy = rng.uniform(0.2, 1.0, B).astype(NP[dtype])
    ref = O.pde_pairwise(inc, order, sx, sy)
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(torch.from_numpy(inc).to(DEV), order, out,
                         torch.from_numpy(sx).to(DEV), torch.from_numpy(sy).to(DEV))
    assert rel_fro(out.cpu().numpy(), ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("symmetric", [False, True])
def test_gram_many(dtype, symmetric):
    bx = 66
    by = bx if symmetric else 35
    m = 11
    n = m if symmetric else 19
    order = 1
    inc = rand_inc((bx, by, m, n), 0.12, 8, dtype)
    rng = np.random.default_rng(12)
    sx = rng.uniform(0.3, 1.0, bx).astype(NP[dtype])
    sy = rng.uniform(0.3, 1.0, by).astype(NP[dtype])
    off_x, off_y = (2, 2) if symmetric else (1, 3)
    out = torch.full((bx + 4, by + 5), -7.0, dtype=dtype, device=DEV)
    _native.pde_gram(torch.from_numpy(inc).to(DEV), order, out, off_x, off_y,
                     torch.from_numpy(sx).to(DEV), torch.from_numpy(sy).to(DEV), symmetric)
    got = out.cpu().numpy()
    ref = O.pde_gram(inc, order, sx, sy, symmetric)
    blk = got[off_x:off_x + bx, off_y:off_y + by]
    if symmetric:
        lo = np.tril_indices(bx)
        assert rel_fro(blk[lo], ref[lo]) < TOL[dtype]
        assert np.array_equal(blk, blk.T)
    else:
        assert rel_fro(blk, ref) < TOL[dtype]
    mask = np.ones_like(got, dtype=bool)
    mask[off_x:off_x + bx, off_y:off_y + by] = False
    if symmetric:
        mask[off_y:off_y + by, off_x:off_x + bx] = False
    assert np.all(got[mask] == -7.0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_robust_norm_many(dtype):
    B, m, order = MANY + 1, 9, 1
    rng = np.random.default_rng(23)
    dx = rng.normal(size=(B, m, 3)) * 0.45
    inc64 = np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order
    inc = inc64.astype(NP[dtype])
    lam_ref, k_ref, it_ref, fx_ref = O.norm_factors(inc, order, tol=1e-10, maxit=100, fixed_points=True)
    lam = torch.zeros(B, dtype=dtype, device=DEV)
    kn = torch.zeros(B, dtype=dtype, device=DEV)
    iters = torch.zeros(B, dtype=torch.int32, device=DEV)
    _native.robust_norm(torch.from_numpy(inc).to(DEV), order, lam, kn, None, maxit=100,
                        tol=1e-10, iters=iters)
    assert rel_fro(kn.cpu().numpy(), k_ref) < TOL[dtype]
    tol_lam = 1e-9 if dtype == torch.float64 else 1e-5
    assert max_rel(lam.cpu().numpy(), lam_ref, floor=1e-3) < tol_lam
    if dtype == torch.float64:
        assert np.array_equal(iters.cpu().numpy(), evals_with_fixed_point_exit(it_ref, fx_ref))
