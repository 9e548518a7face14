"""GPU parity of the multi-wave route (sigpde_amd/csrc/mw_solver.h): pairs of at least 4 bands of
64 refined rows run one pair per workgroup of 8 waves (16 in float32 from 8 bands), the bands of each
pair pipelined across the waves, when the GPU holds all of the pairs' workgroups at once; from
dyadic order 1 a lane holds two refined rows (one row of increments) when that still fits. Checked
against the oracle, and bit for bit against the one-wave-per-pair route (the same pairs repeated
past the GPU's resident workgroups): the cell arithmetic is shared, so the route must not change a
result.
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


def rand_inc(shape, seed, dtype, order=0):
    rng = np.random.default_rng(seed)
    m, n = shape[-2], shape[-1]
    return (rng.normal(size=shape) * (1.5 / np.sqrt(m * n)) / 4.0 ** order).astype(NP[dtype])


def solve(inc, order, sx=None, sy=None):
    out = torch.zeros(inc.shape[0], dtype=inc.dtype, device=DEV)
    _native.pde_pairwise(inc, order, out, sx, sy)
    return out


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("B,m,n,order", [
    (4, 1100, 40, 0),    # 18 bands, short rows of columns
    (130, 1100, 9, 0),   # 130 pairs
    (6, 300, 50, 0),     # 5 bands: 8 waves in float32 too
    (300, 600, 5, 0),    # 300 pairs of 10 bands (float32: two workgroups per CU)
    (3, 300, 300, 2),    # 1200 x 1200 refined
    (2, 2000, 50, 0),    # 32 bands (16 waves)
    (5, 129, 2000, 3),   # 1032 x 16000 refined: the largest windows
    (1, 1024, 30, 0),    # exactly 16 bands, short bands (6 periods)
    (1, 1024, 7, 0),     # 5 periods a band: no overlap, the one-wave route
    (2, 600, 200, 1),    # order 1: 9 distinct increments per lane and period
    (2, 70, 40, 4),      # order 4: 16 columns span two increments
    (1, 40, 20, 5),      # order 5
])
def test_pairwise_multiwave(dtype, B, m, n, order):
    inc = rand_inc((B, m, n), 3 * m + n + order, dtype, order)
    ref = O.pde_pairwise(inc, order)
    got = solve(torch.from_numpy(inc).to(DEV), order).cpu().numpy()
    assert rel_fro(got, ref) < TOL[dtype]
    assert max_rel(got, ref, floor=1e-3) < 10 * TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_multiwave_route_equals_one_wave_route_bitwise(dtype):
    B, m, n, order = 3, 1030, 33, 1
    inc = torch.from_numpy(rand_inc((B, m, n), 5, dtype, order)).to(DEV)
    sx = torch.linspace(0.5, 1.0, B, dtype=dtype, device=DEV)
    sy = torch.linspace(0.9, 0.6, B, dtype=dtype, device=DEV)
    few = solve(inc, order, sx, sy)                        # multi-wave route
    # more pairs than the GPU holds workgroups of 8 waves (at most four per CU): one wave per pair
    nm = 4 * torch.cuda.get_device_properties(0).multi_processor_count + 1
    reps = (nm + B - 1) // B
    many_inc = inc.repeat(reps, 1, 1)[:nm].contiguous()
    many = solve(many_inc, order, sx.repeat(reps)[:nm].contiguous(),
                 sy.repeat(reps)[:nm].contiguous())
    assert torch.equal(few, many[:B])


@pytest.mark.parametrize("m,order", [(300, 0), (1100, 0), (2100, 0),
                                     (300, 1), (150, 2)])  # orders 1-2: two rows per lane
def test_robust_norm_multiwave(m, order):
    B = 2
    rng = np.random.default_rng(17)
    dx = rng.normal(size=(B, m, 3)) * (1.3 / np.sqrt(m))
    inc = np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order
    lam_ref, k_ref, it_ref, fx_ref = O.norm_factors(inc, order, tol=1e-10, maxit=100, fixed_points=True)
    lam = torch.zeros(B, dtype=torch.float64, device=DEV)
    kn = torch.zeros(B, dtype=torch.float64, device=DEV)
    iters = torch.zeros(B, dtype=torch.int32, device=DEV)
    t = torch.from_numpy(inc).to(DEV)
    _native.robust_norm(t, order, lam, kn, None, maxit=100, tol=1e-10, iters=iters)
    assert rel_fro(kn.cpu().numpy(), k_ref) < 1e-10
    assert max_rel(lam.cpu().numpy(), lam_ref, floor=1e-3) < 1e-9
    assert np.array_equal(iters.cpu().numpy(), evals_with_fixed_point_exit(it_ref, fx_ref))
    # bit for bit against the one-wave route: more paths than the GPU holds workgroups of the
    # float64 root-find (166-178 registers per lane: one workgroup of 8 waves per CU)
    nm = 2 * torch.cuda.get_device_properties(0).multi_processor_count + 1
    many = t.repeat((nm + B - 1) // B, 1, 1)[:nm].contiguous()
    lam2 = torch.zeros(nm, dtype=torch.float64, device=DEV)
    kn2 = torch.zeros(nm, dtype=torch.float64, device=DEV)
    _native.robust_norm(many, order, lam2, kn2, None, maxit=100, tol=1e-10)
    assert torch.equal(lam, lam2[:B]) and torch.equal(kn, kn2[:B])


def test_robust_api_on_long_paths():
    """RobustSigPDE end to end on few long float64 paths (the multi-wave route for both the
    root-find and the scaled pairwise solve) against the oracle."""
    import sigpde_amd.torch as sig
    from helpers import brownian_np
    x = brownian_np(3, 600, 3, 4.0 / 600, 31)
    y = brownian_np(3, 600, 3, 4.0 / 600, 32)
    rk = sig.RobustSigPDE(sig.kernels.LinearKernel(), 1)
    xc, yc = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    lx = rk.normalization(xc, tol=1e-10).cpu().numpy()
    lx_ref = O.robust_normalization(x, 1, tol=1e-10)
    assert np.max(np.abs(lx - lx_ref)) < 1e-9
    got = rk.pairwise(xc, yc, tol=1e-10).cpu().numpy()
    ly_ref = O.robust_normalization(y, 1, tol=1e-10)
    ref = O.sig_pairwise(x, y, 1, sx=lx_ref, sy=ly_ref)
    assert rel_fro(got, ref) < 1e-9
