"""GPU parity for long paths. Refined column counts past what four boundary lines per wave fit in
LDS (float64 past 4559, float32 past 9679) run on lane groups of 64 (one line per wave in LDS);
past the longest line one wave holds in LDS (17695 float32 / 14927 float64 refined columns) on
lane groups of 16 with their lines in global memory, in launches of at most 4 GiB of lines. Also
long refined row counts (no LDS cost), the fused linear paths up to their 9679 columns, and the
errors past each limit. Oracle: the CPU restatement on the same increments / paths.
"""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from helpers import brownian_np, max_rel, rel_fro, evals_with_fixed_point_exit
from sigpde_amd import _native

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = {torch.float64: 1e-10, torch.float32: 1e-5}
NP = {torch.float64: np.float64, torch.float32: np.float32}


def rand_inc(shape, seed, dtype):
    rng = np.random.default_rng(seed)
    m, n = shape[-2], shape[-1]
    return (rng.normal(size=shape) * (1.5 / np.sqrt(m * n))).astype(NP[dtype])


@pytest.mark.parametrize("dtype,B,m,n,order", [
    (torch.float64, 4, 6, 5000, 0),       # one pair per wave; past the float64 four-line limit
    (torch.float64, 2049, 3, 5000, 0),    # many pairs, but four float64 lines do not fit: groups of 64
    (torch.float32, 2049, 3, 9000, 0),    # many pairs, four float32 lines still fit (groups of 16)
    (torch.float32, 2049, 2, 12000, 0),   # many pairs past the float32 four-line limit
    (torch.float32, 3, 2, 17695, 0),      # the longest float32 line
    (torch.float64, 3, 2, 14927, 0),      # the longest float64 line
    (torch.float64, 8, 2, 1200, 3),       # dyadic refinement to 9600 columns
    (torch.float64, 2, 9000, 3, 0),       # long rows: 9000 refined rows, no LDS cost
    (torch.float32, 3, 2, 17696, 0),      # one column past the float32 LDS line: global lines
    (torch.float64, 2049, 2, 14928, 0),   # many pairs, global float64 lines
    (torch.float32, 5, 3, 40000, 0),      # 40000 refined columns
    (torch.float64, 4, 2, 2500, 3),       # dyadic refinement to 20000 columns, 8 rows per lane
    (torch.float32, 6, 40, 18000, 0),     # several bands per pair with global lines
])
def test_pairwise_long(dtype, B, m, n, order):
    inc = rand_inc((B, m, n), 11 * m + n + order, dtype) / 4.0 ** order
    ref = O.pde_pairwise(inc, order)
    out = torch.zeros(B, dtype=dtype, device=DEV)
    _native.pde_pairwise(torch.from_numpy(inc).to(DEV), order, out)
    got = out.cpu().numpy()
    assert rel_fro(got, ref) < TOL[dtype]
    assert max_rel(got, ref, floor=1e-3) < 10 * TOL[dtype]


def test_pairwise_global_lines_in_several_launches():
    """70000 pairs of 17696 refined columns: 17500 waves of 277 KiB of float32 lines, more than the
    4 GiB one launch takes, so the batch runs as two launches reusing the lines. Every pair reads
    the same increments (batch stride 0), so every output must equal the one oracle value."""
    n, B = 17696, 70000
    inc1 = rand_inc((1, 1, n), 3, torch.float32)
    ref = O.pde_pairwise(inc1, 0)[0]
    inc = torch.from_numpy(inc1).to(DEV).expand(B, 1, n)
    out = torch.zeros(B, dtype=torch.float32, device=DEV)
    _native.pde_pairwise(inc, 0, out)
    got = out.cpu().numpy()
    assert np.all(got == got[0])
    assert abs(float(got[0]) - float(ref)) <= 1e-5 * abs(float(ref))


def test_gram_global_lines():
    bx, by, m, n = 3, 2, 2, 16000
    inc = rand_inc((bx, by, m, n), 17, torch.float64)
    ref = O.pde_gram(inc, 0)
    out = torch.zeros((bx, by), dtype=torch.float64, device=DEV)
    _native.pde_gram(torch.from_numpy(inc).to(DEV), 0, out)
    assert rel_fro(out.cpu().numpy(), ref) < 1e-10


def test_robust_norm_global_lines():
    """Root-find on a 15200 x 15200 refined grid (float64 lines in global memory), two iterations."""
    B, m, order = 1, 1900, 3
    rng = np.random.default_rng(13)
    dx = rng.normal(size=(B, m, 2)) * (1.3 / np.sqrt(m))
    inc = np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order
    lam_ref, k_ref, it_ref, fx_ref = O.norm_factors(inc, order, tol=1e-10, maxit=2, fixed_points=True)
    lam = torch.zeros(B, dtype=torch.float64, device=DEV)
    kn = torch.zeros(B, dtype=torch.float64, device=DEV)
    iters = torch.zeros(B, dtype=torch.int32, device=DEV)
    _native.robust_norm(torch.from_numpy(inc).to(DEV), order, lam, kn, None, maxit=2, tol=1e-10,
                        iters=iters)
    assert rel_fro(kn.cpu().numpy(), k_ref) < 1e-10
    assert max_rel(lam.cpu().numpy(), lam_ref, floor=1e-3) < 1e-9
    assert np.array_equal(iters.cpu().numpy(), evals_with_fixed_point_exit(it_ref, fx_ref))


def test_grid_too_large_raises():
    out = torch.zeros(1, dtype=torch.float32, device=DEV)
    with pytest.raises(_native.NativeError, match="refined columns"):
        inc = torch.zeros((1, 1, (1 << 22) + 1), dtype=torch.float32, device=DEV)
        _native.pde_pairwise(inc, 0, out)  # 2^22 + 1 refined columns
    with pytest.raises(_native.NativeError, match="too large"):
        inc = torch.zeros((1, 1 << 10, 1 << 10), dtype=torch.float32, device=DEV)
        _native.pde_pairwise(inc, 8, out)  # 2^18 x 2^18 refined grid


def test_symmetric_gram_long():
    B, m, order = 3, 5000, 0
    rng = np.random.default_rng(5)
    dx = rng.normal(size=(B, m, 2)) * (1.2 / np.sqrt(m))
    inc = np.einsum("isk,jtk->ijst", dx, dx)
    ref = O.pde_gram(inc, order, symmetric=True)
    out = torch.zeros((B, B), dtype=torch.float64, device=DEV)
    _native.pde_gram(torch.from_numpy(inc).to(DEV), order, out, 0, 0, None, None, True)
    got = out.cpu().numpy()
    lo = np.tril_indices(B)
    assert rel_fro(got[lo], ref[lo]) < 1e-10
    assert np.array_equal(got, got.T)


def test_robust_norm_long():
    B, m, order = 2, 5000, 0
    rng = np.random.default_rng(9)
    dx = rng.normal(size=(B, m, 3)) * (1.4 / np.sqrt(m))
    inc = np.einsum("bsk,btk->bst", dx, dx)
    lam_ref, k_ref, _ = O.norm_factors(inc, order, tol=1e-10, maxit=100)
    lam = torch.zeros(B, dtype=torch.float64, device=DEV)
    kn = torch.zeros(B, dtype=torch.float64, device=DEV)
    _native.robust_norm(torch.from_numpy(inc).to(DEV), order, lam, kn, None, maxit=100, tol=1e-10)
    assert rel_fro(kn.cpu().numpy(), k_ref) < 1e-10
    assert max_rel(lam.cpu().numpy(), lam_ref, floor=1e-3) < 1e-9


@pytest.mark.parametrize("prec", ["fp16x2", "bf16x3", "fp32", "fp64"])
@pytest.mark.parametrize("Lx,Ly", [(64, 9680), (12000, 40)])
def test_fused_gram_long(prec, Lx, Ly):
    """Fused linear Gram with 9679 refined columns (the longest float line; fp64, whose lines and
    ring are double, takes 4559 and must refuse it) and with 11999 refined rows."""
    X = brownian_np(3, Lx, 8, 1.0 / max(Lx, Ly), 21, np.float32)
    Y = brownian_np(2, Ly, 8, 1.0 / max(Lx, Ly), 22, np.float32)
    k = sig.SigPDE(sig.kernels.LinearKernel(), 0, precision=prec)
    if prec == "fp64" and Ly - 1 > 4559:
        with pytest.raises(_native.NativeError, match="refined columns"):
            k.gram(torch.from_numpy(X).to(DEV), torch.from_numpy(Y).to(DEV))
        Y = Y[:, :4560]
    ref = O.linear_gram(X.astype(np.float64), Y.astype(np.float64), 0)
    K = k.gram(torch.from_numpy(X).to(DEV), torch.from_numpy(np.ascontiguousarray(Y)).to(DEV))
    bar = {"fp32": 1e-4, "fp64": 2e-7}.get(prec, 1e-5)
    assert rel_fro(K.double().cpu().numpy(), ref) < bar


def test_fused_past_the_line_limit_raises():
    X = torch.zeros((1, 8, 4), dtype=torch.float32, device=DEV)
    Y = torch.zeros((1, 9681, 4), dtype=torch.float32, device=DEV)
    with pytest.raises(_native.NativeError, match="refined columns"):
        sig.SigPDE(sig.kernels.LinearKernel(), 0, precision="fp16x2").gram(X, Y)
