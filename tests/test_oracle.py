"""CPU pins for the oracle (no GPU). Each check is chosen so that a plausible restatement mistake
(a dropped term, a wrong index or sign, a transposed operand, wrong float32 typing) fails it:

  * two independent restatements (C row sweep, NumPy anti-diagonal sweep) agree bit for bit, in
    float64 and in the reference's float32 typing;
  * closed forms: one cell, zero increments, straight lines -> I0(2 sqrt c) / J0(2 sqrt |c|);
  * a brute-force truncated signature inner product (Chen's identity) that the refined PDE
    converges to;
  * invariants: transpose symmetry, dyadic order o == order 0 on 2^o-subdivided increments,
    scaling, symmetric Gram == lower triangle of the full Gram;
  * the increment second difference equals <dx_s, dy_t> / 4^o for the linear kernel;
  * static kernels against direct loops;
  * Chandrupatla roots against scipy's brentq and bisection, the first-iterate bookkeeping, and the
    normalisation special cases (NaN -> inf, K < 1 -> 0);
  * the golden fixtures in tests/golden (the oracle's outputs on seeded inputs of the smaller
    workloads, written by tests/golden/make_fixtures.py) are reproduced.
"""
import math

import numpy as np
import pytest
from scipy import optimize, special

import oracle as O
from helpers import brownian_np, rel_fro


def rand_inc(shape, scale, seed, dtype=np.float64):
    return (np.random.default_rng(seed).normal(size=shape) * scale).astype(dtype)


# ----------------------------------------------------------------------------- cross-restatement
@pytest.mark.parametrize("m,n,order", [(1, 1, 0), (5, 9, 0), (9, 5, 1), (7, 7, 2), (12, 4, 3)])
def test_c_and_numpy_restatements_bitwise_f64(m, n, order):
    inc = rand_inc((3, m, n), 0.4 / (1 << order), m * 31 + n + order)
    assert np.array_equal(O.pde_pairwise(inc, order), O.pde_pairwise_antidiag(inc, order))
    s = np.array([0.3, 0.9, 1.7])
    assert np.array_equal(O.pde_pairwise(inc, order, s, None),
                          O.pde_pairwise_antidiag(inc, order, s))


@pytest.mark.parametrize("m,n,order", [(1, 1, 0), (6, 11, 0), (10, 6, 2), (8, 8, 1)])
def test_c_and_numpy_restatements_bitwise_f32ref(m, n, order):
    inc = rand_inc((3, m, n), 0.4 / (1 << order), 7 + m + n, np.float32)
    assert np.array_equal(O.pde_pairwise(inc, order),
                          O.pde_pairwise_antidiag(inc, order, f32ref=True))
    s = np.float32([0.3, 0.9, 1.7])
    assert np.array_equal(O.pde_pairwise(inc, order, s, None),
                          O.pde_pairwise_antidiag(inc, order, s, f32ref=True))


# ----------------------------------------------------------------------------- closed forms
@pytest.mark.parametrize("c", [0.0, 0.37, -0.8, 2.5])
def test_single_cell_closed_form(c):
    # k[1][1] from three boundary ones: 2 * (1 + c/2 + c^2/12) - (1 - c^2/12) = 1 + c + c^2/4
    got = O.pde_pairwise(np.full((1, 1, 1), c), 0)[0]
    assert got == pytest.approx(1 + c + c * c / 4, rel=1e-15, abs=1e-15)


def test_zero_increments_are_exactly_one():
    for dt in (np.float64, np.float32):
        inc = np.zeros((2, 13, 7), dtype=dt)
        for order in range(4):
            assert np.all(O.pde_pairwise(inc, order) == 1)


@pytest.mark.parametrize("c", [0.7, 2.0, -0.9, -3.0])
def test_straight_lines_converge_to_bessel(c):
    # x(t) = t u, y(s) = s v on [0, 1]: K = sum_k c^k / (k!)^2 = I0(2 sqrt c) (c >= 0) or
    # J0(2 sqrt |c|) (c < 0), with c = <u, v>. With m x n segments every increment is c/(m n).
    exact = special.i0(2 * math.sqrt(c)) if c >= 0 else special.j0(2 * math.sqrt(-c))
    m, n = 3, 4
    errs = []
    for order in range(0, 8):
        inc = np.full((1, m, n), c / (m * n) / 4.0 ** order)
        errs.append(abs(O.pde_pairwise(inc, order)[0] - exact) / abs(exact))
    assert errs[-1] < 5e-6
    assert all(b < a for a, b in zip(errs, errs[1:]))  # monotone convergence in the order
    assert errs[-2] / errs[-1] > 3.5  # second-order scheme: ~4x per dyadic order


def _signature(path, level):
    """Truncated signature of a piecewise-linear path by Chen's identity: level-k tensors."""
    d = path.shape[1]
    sig = [np.ones(())] + [np.zeros((d,) * k) for k in range(1, level + 1)]
    for v in np.diff(path, axis=0):
        seg = [np.ones(())]
        for k in range(1, level + 1):
            seg.append(np.multiply.outer(seg[-1], v) / k)
        new = []
        for k in range(level + 1):
            acc = np.zeros((d,) * k)
            for i in range(k + 1):
                acc = acc + np.multiply.outer(sig[i], seg[k - i])
            new.append(acc)
        sig = new
    return sig


def test_pde_converges_to_truncated_signature_kernel():
    rng = np.random.default_rng(4)
    x = np.cumsum(rng.normal(size=(4, 2)) * 0.4, axis=0)
    y = np.cumsum(rng.normal(size=(5, 2)) * 0.4, axis=0)
    sx, sy = _signature(x, 12), _signature(y, 12)
    truth = sum(float(np.sum(a * b)) for a, b in zip(sx, sy))
    dx, dy = np.diff(x, axis=0), np.diff(y, axis=0)
    errs = []
    for order in (2, 4, 6, 8):
        inc = (dx @ dy.T)[None] / 4.0 ** order
        errs.append(abs(O.pde_pairwise(inc, order)[0] - truth) / abs(truth))
    assert errs[-1] < 1e-7 and errs[0] > errs[1] > errs[2] > errs[3]
    assert errs[-2] / errs[-1] > 10  # two dyadic orders: ~16x


# ----------------------------------------------------------------------------- invariants
def test_transpose_symmetry_bitwise():
    inc = rand_inc((4, 9, 14), 0.3, 3)
    assert np.array_equal(O.pde_pairwise(inc, 1), O.pde_pairwise(np.swapaxes(inc, 1, 2), 1))


@pytest.mark.parametrize("order", [1, 2, 3])
def test_dyadic_order_equals_subdivided_paths(order):
    inc = rand_inc((3, 6, 5), 0.5, 8) / 4.0 ** order
    R = 1 << order
    sub = np.repeat(np.repeat(inc, R, axis=1), R, axis=2)
    assert np.array_equal(O.pde_pairwise(inc, order), O.pde_pairwise(sub, 0))


def test_scaled_equals_prescaled_increments():
    inc = rand_inc((5, 8, 8), 0.3, 12)
    sx = np.linspace(0.1, 1.0, 5)
    sy = np.linspace(1.0, 0.4, 5)
    pre = inc * (sy * sx)[:, None, None]
    assert np.array_equal(O.pde_pairwise(inc, 1, sx, sy), O.pde_pairwise(pre, 1))


def test_symmetric_gram_is_lower_triangle_of_full_gram():
    x = brownian_np(7, 10, 3, 0.1, 1)
    full = O.sig_gram(x, x, 1)
    sym = O.sig_gram(x, x, 1, symmetric=True)
    lo = np.tril_indices(7)
    assert np.array_equal(sym[lo], full[lo])
    assert np.array_equal(sym, sym.T)


def test_gram_entries_equal_pairwise():
    inc = rand_inc((4, 3, 9, 11), 0.2, 5)
    sx, sy = np.linspace(0.5, 1, 4), np.linspace(1, 0.3, 3)
    g = O.pde_gram(inc, 2, sx, sy)
    p = O.pde_pairwise(inc.reshape(12, 9, 11), 2, np.repeat(sx, 3), np.tile(sy, 4))
    assert np.array_equal(g.reshape(-1), p)
    x = brownian_np(4, 9, 2, 0.2, 5)
    y = brownian_np(3, 11, 2, 0.2, 6)
    g = O.sig_gram(x, y, 2)
    for i in range(4):
        for j in range(3):
            assert g[i, j] == pytest.approx(O.sig_pairwise(x[i:i + 1], y[j:j + 1], 2)[0], rel=1e-13)


def test_f32ref_uses_double_cell_arithmetic():
    # float32 storage with double per-cell arithmetic stays within ~1e-5 of float64 on a 396 x 396
    # refined grid; all-float arithmetic would be ~1e-4 away (a typing mistake shows up here)
    x = brownian_np(6, 100, 8, 1 / 100, 2)
    y = brownian_np(6, 100, 8, 1 / 100, 3)
    inc = O.second_difference(O.static_pairwise("linear", x, y), 2)
    k64 = O.pde_pairwise(inc, 2)
    k32 = O.pde_pairwise(inc.astype(np.float32), 2)
    assert rel_fro(k32, k64) < 2e-5


# ----------------------------------------------------------------------------- increments, static kernels
def test_second_difference_is_increment_inner_product():
    x = brownian_np(3, 12, 4, 0.1, 7)
    y = brownian_np(3, 9, 4, 0.1, 8)
    for order in (0, 2):
        inc = O.second_difference(O.static_pairwise("linear", x, y, 1.5), order)
        want = 1.5 * np.einsum("bsk,btk->bst", np.diff(x, axis=1), np.diff(y, axis=1)) / 4.0 ** order
        assert np.allclose(inc, want, rtol=0, atol=1e-13)
        # the C path (sequential accumulation) agrees
        assert np.allclose(O.linear_inc(x[1], y[1], order, 1.5), want[1], rtol=0, atol=1e-13)


def test_static_kernels_against_loops():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(2, 3, 4))
    Y = rng.normal(size=(3, 5, 4))
    G = {k: O.static_gram(k, X, Y, 1.7) for k in ("linear", "rbf", "l1rbf")}
    for a in range(2):
        for b in range(3):
            for s in range(3):
                for t in range(5):
                    u, v = X[a, s], Y[b, t]
                    assert G["linear"][a, b, s, t] == pytest.approx(1.7 * u @ v)
                    assert G["rbf"][a, b, s, t] == pytest.approx(np.exp(-np.sum((u - v) ** 2) / 1.7))
                    assert G["l1rbf"][a, b, s, t] == pytest.approx(np.exp(-np.sum(np.abs(u - v)) / 1.7))
    P = O.static_pairwise("rbf", X, Y[:2], 1.7)
    assert np.allclose(P, np.stack([G["rbf"][a, a] for a in range(2)]))


def test_log_normalizer_closed_form():
    v = O.log_normalizer(np.array([1.0, math.e, 10.0, 1e6]))
    assert v[0] == 1.0 and v[1] == pytest.approx(1.5)
    assert np.all(v[1:] <= np.array([math.e, 10.0, 1e6]))


# ----------------------------------------------------------------------------- Chandrupatla
def _self_inc(B, m, d, sd, seed, order):
    rng = np.random.default_rng(seed)
    dx = rng.normal(size=(B, m, d)) * sd
    return np.einsum("bsk,btk->bst", dx, dx) / 4.0 ** order


@pytest.mark.parametrize("order", [0, 1])
def test_chandrupatla_roots_match_brentq_and_bisection(order):
    inc = _self_inc(6, 12, 3, 0.5, 9, order)
    lam, knorm, iters = O.norm_factors(inc, order, tol=1e-12, maxit=100)
    target = O.log_normalizer(knorm)
    for b in range(6):
        if knorm[b] < 1:
            continue

        def F(l):
            return O.pde_pairwise(inc[b:b + 1] * (l * l), order)[0] - target[b]

        root = optimize.brentq(F, 0.0, 1.0, xtol=1e-15, rtol=1e-15)
        lo, hi = 0.0, 1.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if F(mid) < 0 else (lo, mid)
        assert lam[b] == pytest.approx(root, abs=1e-10)
        assert lam[b] == pytest.approx(0.5 * (lo + hi), abs=1e-10)
        assert abs(F(lam[b])) < 1e-11
        assert iters[b] < 30  # superlinear: far fewer than bisection's ~40 steps


def test_chandrupatla_first_iterate_bookkeeping():
    # with maxit=1 the only evaluation is at 0.5 and the output is the better endpoint
    inc = _self_inc(4, 10, 3, 0.6, 1, 0)
    knorm = O.pde_pairwise(inc, 0)
    target = O.log_normalizer(knorm)
    lam, it = O.chandrupatla(inc, 0, target, knorm, maxit=1, tol=1e-14)
    f_half = O.pde_pairwise(inc * 0.25, 0) - target
    f_zero = 1.0 - target
    f_one = knorm - target
    # the bracket keeps the endpoint whose sign differs from F(0.5); the better of the two wins
    same = f_half * f_zero >= 0
    other_x = np.where(same, 1.0, 0.0)
    other_f = np.where(same, f_one, f_zero)
    want = np.where(np.abs(f_half) < np.abs(other_f), 0.5, other_x)
    assert np.array_equal(lam, want) and np.all(it == 1)
    lam0, it0 = O.chandrupatla(inc, 0, target, knorm, maxit=0, tol=1e-14)
    assert np.all(lam0 == 0) and np.all(it0 == 0)


def test_norm_factors_special_cases():
    # K < 1 -> lambda = 0; NaN increments -> K = inf
    inc = np.zeros((3, 1, 1))
    inc[0] = -0.1   # one cell: 1 - 0.1 + 0.0025 < 1
    inc[1] = np.nan
    inc[2] = 0.5
    lam, knorm, _ = O.norm_factors(inc, 0, tol=1e-12)
    assert knorm[0] < 1 and lam[0] == 0
    assert np.isinf(knorm[1])
    assert 0 < lam[2] < 1


def test_norm_factors_float32_close_to_float64():
    inc = _self_inc(5, 15, 3, 0.4, 3, 1)
    l64 = O.norm_factors(inc, 1, tol=1e-10)[0]
    l32 = O.norm_factors(inc.astype(np.float32), 1, tol=1e-10)[0]
    assert np.allclose(l32, l64, rtol=1e-5)


# ----------------------------------------------------------------------------- golden fixtures
def _golden(name):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
    return np.load(path, allow_pickle=False)


def test_oracle_reproduces_golden_fixtures():
    g = _golden("pairwise_f64_L32_d4_o1.npz")
    assert rel_fro(O.sig_pairwise(g["x"], g["y"], 1), g["K64"]) < 1e-14
    g = _golden("gram_f32_L100_d8_o2.npz")
    X, Y = g["X"], g["Y"]
    assert rel_fro(O.linear_gram(X.astype(np.float64), Y.astype(np.float64), 2), g["K64"]) < 1e-14
    assert rel_fro(O.linear_gram(X, Y, 2).astype(np.float64), g["K32"].astype(np.float64)) < 1e-7
    g = _golden("robust_f64_L64_d16_o1.npz")
    inc = O.second_difference(O.static_pairwise("linear", g["X"], g["X"]), 1)
    lam, knorm, iters = O.norm_factors(inc, 1, tol=1e-10, maxit=100)
    assert np.max(np.abs(lam - g["lam"])) < 1e-13
    assert rel_fro(knorm, g["knorm"]) < 1e-14
    assert np.array_equal(iters, g["iters"])
    assert rel_fro(O.linear_gram(g["X"], g["X"], 1, sx=lam, sy=lam), g["K"]) < 1e-13
