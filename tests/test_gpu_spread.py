"""GPU parity of the spread route (sigpde_amd/csrc/mw_solver.h with XW, wavefront.hip
k_pde_pairwise_xw / k_robust_norm_xw): each of few long pairs runs on G workgroups whose waves
share its bands, handing band boundaries between workgroups through global memory with bounded
waits. The cell arithmetic is the other routes', so results must equal the single-workgroup
route bit for bit (sigpde_set_spread(1) turns the spread off) and the oracle within the usual
bars; sigpde_set_spread(G) forces G workgroups per pair on small grids. The failure path is
forced from the host: the test hook stops workgroup 0 from publishing its progress, so the waits
on it give up after the (short) wall-clock limit and the call reports SIGPDE_ETIMEOUT, with every
kernel still running to its end.
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


@pytest.fixture(autouse=True)
def _automatic_spread_after():
    yield
    _native.set_spread(0, 2.0, 0)


def rand_inc(shape, seed, dtype, order=0):
    rng = np.random.default_rng(seed)
    m, n = shape[-2], shape[-1]
    return (rng.normal(size=shape) * (1.5 / np.sqrt(m * n)) / 4.0 ** order).astype(NP[dtype])


def solve(inc, order, sx=None, sy=None):
    out = torch.zeros(inc.shape[0], dtype=inc.dtype, device=DEV)
    _native.pde_pairwise(inc, order, out, sx, sy)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("G", [2, 3, 5])
@pytest.mark.parametrize("B,m,n,order", [
    (2, 1100, 700, 0),   # 18 bands, 48 periods a band: bands overlap, windows wrap
    (3, 700, 300, 1),    # order 1: two rows per lane
    (1, 2100, 90, 0),    # 33 bands of short rows
    (2, 130, 400, 2),    # 520 x 1600 refined, 5 bands (some workgroups hold none)
])
def test_spread_pairwise_equals_single_workgroup_and_oracle(dtype, G, B, m, n, order):
    print(f"spread pairwise {dtype} G={G} B={B} m={m} n={n} order={order}", flush=True)
    inc = torch.from_numpy(rand_inc((B, m, n), 7 * m + n + order, dtype, order)).to(DEV)
    sx = torch.linspace(0.6, 1.0, B, dtype=dtype, device=DEV)
    sy = torch.linspace(1.1, 0.8, B, dtype=dtype, device=DEV)
    _native.set_spread(1)
    one = solve(inc, order, sx, sy)
    _native.set_spread(G)
    spread = solve(inc, order, sx, sy)
    assert torch.equal(one, spread)
    ref = O.pde_pairwise(inc.cpu().numpy(), order, sx=sx.cpu().numpy(), sy=sy.cpu().numpy())
    assert rel_fro(spread.cpu().numpy(), ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("G,m,order", [(2, 1100, 0), (4, 2100, 0), (3, 300, 2)])
def test_spread_root_find_equals_single_workgroup(dtype, G, m, order):
    print(f"spread root-find {dtype} G={G} m={m} order={order}", flush=True)
    B = 2
    rng = np.random.default_rng(29 + m)
    dx = rng.normal(size=(B, m, 3)) * (1.3 / np.sqrt(m))
    inc = torch.from_numpy((np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order).astype(NP[dtype]))
    inc = inc.to(DEV)

    def run():
        lam = torch.zeros(B, dtype=dtype, device=DEV)
        kn = torch.zeros(B, dtype=dtype, device=DEV)
        it = torch.zeros(B, dtype=torch.int32, device=DEV)
        _native.robust_norm(inc, order, lam, kn, None, maxit=100, tol=1e-10, iters=it)
        torch.cuda.synchronize()
        return lam, kn, it

    _native.set_spread(1)
    lam1, kn1, it1 = run()
    _native.set_spread(G)
    lam2, kn2, it2 = run()
    assert torch.equal(lam1, lam2) and torch.equal(kn1, kn2) and torch.equal(it1, it2)
    if dtype == torch.float64:
        lam_ref, k_ref, it_ref, fx_ref = O.norm_factors(inc.cpu().numpy(), order, tol=1e-10,
                                                        maxit=100, fixed_points=True)
        assert rel_fro(kn2.cpu().numpy(), k_ref) < 1e-10
        assert max_rel(lam2.cpu().numpy(), lam_ref, floor=1e-3) < 1e-9
        assert np.array_equal(it2.cpu().numpy(), evals_with_fixed_point_exit(it_ref, fx_ref))


def test_spread_given_normaliser_and_builtin():
    """The given-target root-find and a built-in device normaliser through the spread route."""
    B, m, order = 2, 900, 0
    rng = np.random.default_rng(5)
    dx = rng.normal(size=(B, m, 3)) * (1.3 / np.sqrt(m))
    inc = torch.from_numpy(np.einsum("bsk,btk->bst", dx, dx)).to(DEV)
    out = {}
    for G in (1, 3):
        _native.set_spread(G)
        lam = torch.zeros(B, dtype=torch.float64, device=DEV)
        kn = torch.zeros(B, dtype=torch.float64, device=DEV)
        _native.robust_norm_builtin(inc, order, lam, kn, _native.NORM_TENSOR, 4.0, 1.0,
                                    maxit=100, tol=1e-10)
        lam_g = torch.zeros(B, dtype=torch.float64, device=DEV)
        target = 1.0 + 0.5 * (kn - 1.0)
        _native.robust_norm(inc, order, lam_g, kn.clone(), target, maxit=100, tol=1e-10)
        torch.cuda.synchronize()
        out[G] = (lam, kn, lam_g)
    for a, b in zip(out[1], out[3]):
        assert torch.equal(a, b)


def test_automatic_route_on_two_long_float32_pairs():
    """2 x 4000^2 float32: the automatic choice spreads the pairs; same bits as one workgroup."""
    inc = torch.from_numpy(rand_inc((2, 4000, 4000), 41, torch.float32)).to(DEV)
    print("automatic spread 2 x 4000^2 f32", flush=True)
    _native.set_spread(1)
    one = solve(inc, 0)
    _native.set_spread(0)
    auto = solve(inc, 0)
    assert torch.equal(one, auto)
    ref = O.pde_pairwise(inc.cpu().numpy(), 0)
    assert rel_fro(auto.cpu().numpy(), ref) < 1e-5


def test_wait_that_gives_up_reports_timeout_and_the_next_call_works():
    inc = torch.from_numpy(rand_inc((2, 1100, 700), 3, torch.float32)).to(DEV)
    _native.set_spread(2, 0.05, 1)
    print("spread failure path (test hook, 50 ms waits)", flush=True)
    with pytest.raises(_native.NativeError) as ei:
        solve(inc, 0)
    assert ei.value.status == _native.ETIMEOUT
    assert "gave up" in str(ei.value)
    _native.set_spread(2, 2.0, 0)
    ok = solve(inc, 0)
    _native.set_spread(1)
    assert torch.equal(ok, solve(inc, 0))
