"""Known answers for the oracle's float32 typing and contraction settings (CPU, no GPU).

The expected values come from tests/exact_fp.py: exact rational arithmetic with an explicit
round-to-nearest-even, independent of numpy's and C's floating point. Each rule of the reference's
float32 arithmetic under Numba (sigpde/cuda_device_functions.py:5,
sigpde/cuda_pairwise_kernels.py:61,114, sigpde/torch/kernel.py:13-21) gets an input on which the
rule changes the stored float32 result; the test asserts both that the rule and its alternative
disagree there (so the case discriminates) and that the oracle lands on the rule's value, bit for
bit:
  * c**2 is float32 (float32 ** int literal stays float32), not float64;
  * every other operation is float64 (the float64 literals 0.5, 1./12, 1. and the ternaries with
    1.0 promote), not float32 (NEP-50 all-float32 arithmetic);
  * each cell is rounded to float32 when stored, before a neighbour reads it;
  * the scale s_y * s_x and c = inc * scale are float32 products;
  * the root-find's c = inc * xt**2 is float64 (xt is a float64), with c**2 in float64;
  * the two contraction settings of pde_step (as written; NVVM's FMA contraction) each match the
    exact model of their form (float64, where one cell shows the difference);
  * the linear increments: the point Gram entry scale * <x_s, y_t> and the 4-term difference
    ((ip11 + ip00) - ip10) - ip01 in float32, then / 4**order.
Random grids of up to 8 x 12 refined cells then check the whole scheme against the exact model
(float32 and float64, both contraction settings, dyadic orders 0-2, scaled), bit for bit.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle as O
from exact_fp import fr, r32, r64, solve, pde_step, to32

H = float.fromhex


def f32_bits(v) -> int:
    return int(np.float32(v).view(np.uint32))


def oracle_cell_f32(inc, order=0, sx=None, sy=None, contract=False):
    inc = np.asarray(inc, dtype=np.float32)[None]
    with O.contracted(contract):
        if sx is None:
            return O.pde_pairwise(inc, order)[0]
        return O.pde_pairwise(inc, order, np.float32([sx]), np.float32([sy]))[0]


RULE = dict(c2_float=True, rest_double=True)

# (contract, c): c**2 in float32 changes the stored cell
R1 = [(False, "-0x1.56aa5cp+0"), (True, "-0x1.5aa548p+0")]
# (contract, c): double arithmetic for the rest changes the stored cell
R2 = [(False, "-0x1.ad86f8p-1"), (True, "-0x1.ec28e0p-1")]
# (contract, c1, c2): the 1 x 2 grid's second cell reads the first cell rounded to float32
R3 = [(False, "-0x1.e4ce0ep-1", "-0x1.f1cefcp-1"), (True, "-0x1.5bd320p-1", "-0x1.4840acp-1")]
# (contract, inc, sx, sy): float32 scale product
R4 = [(False, "-0x1.7bf4ccp+0", "0x1.b6c670p-1", "0x1.ace130p-1"),
      (True, "-0x1.d87866p-1", "0x1.81db3ap-1", "0x1.712314p-2")]
# (contract, c1, c2, xt): the root-find's float64 c = inc * xt**2 on a 1 x 2 grid
R6 = [(False, "-0x1.6498f8p+0", "0x1.6de85cp-5", "0x1.0a0719bcbf191p-1"),
      (True, "-0x1.90c8d2p-2", "-0x1.7d21d2p+0", "0x1.b1af9d4bb629dp-1")]
# (contract, c): float64 single cell where the contraction changes the result
R5 = [(False, "-0x1.5061676100418p+0"), (True, "0x1.f247c45a19a50p-1")]


@pytest.mark.parametrize("contract,c", R1)
def test_c_squared_is_float32(contract, c):
    c = fr(np.float32(H(c)))
    want = r32(solve([[c]], contract=contract, **RULE))
    alt = r32(solve([[c]], contract=contract, c2_float=False, rest_double=True))
    assert want != alt
    got = oracle_cell_f32([[to32(c)]], contract=contract)
    assert f32_bits(got) == f32_bits(to32(want))


@pytest.mark.parametrize("contract,c", R2)
def test_rest_of_the_cell_is_float64(contract, c):
    c = fr(np.float32(H(c)))
    want = r32(solve([[c]], contract=contract, **RULE))
    alt = r32(solve([[c]], contract=contract, c2_float=True, rest_double=False))
    assert want != alt
    got = oracle_cell_f32([[to32(c)]], contract=contract)
    assert f32_bits(got) == f32_bits(to32(want))


@pytest.mark.parametrize("contract,c1,c2", R3)
def test_cells_are_stored_as_float32(contract, c1, c2):
    inc = [[fr(np.float32(H(c1))), fr(np.float32(H(c2)))]]
    want = r32(solve(inc, contract=contract, **RULE))
    alt = r32(solve(inc, contract=contract, store32=False, **RULE))
    assert want != alt
    got = oracle_cell_f32([[to32(v) for v in inc[0]]], contract=contract)
    assert f32_bits(got) == f32_bits(to32(want))


@pytest.mark.parametrize("contract,inc,sx,sy", R4)
def test_scale_product_is_float32(contract, inc, sx, sy):
    inc, sx, sy = (fr(np.float32(H(v))) for v in (inc, sx, sy))
    want = r32(solve([[r32(inc * r32(sy * sx))]], contract=contract, **RULE))
    alt_double_scale = r32(solve([[r32(inc * r64(sy * sx))]], contract=contract, **RULE))
    c64 = r64(inc * r64(sy * sx))
    alt_double_c = r32(solve([[c64]], contract=contract, c2_float=False, rest_double=True))
    assert want != alt_double_scale and want != alt_double_c
    got = oracle_cell_f32([[to32(inc)]], sx=to32(sx), sy=to32(sy), contract=contract)
    assert f32_bits(got) == f32_bits(to32(want))


@pytest.mark.parametrize("contract,c1,c2,xt", R6)
def test_root_find_evaluation_scales_in_float64(contract, c1, c2, xt):
    c1, c2 = fr(np.float32(H(c1))), fr(np.float32(H(c2)))
    xt = Fr(H(xt))
    s = r64(xt * xt)
    want = r32(solve([[r64(c1 * s), r64(c2 * s)]], contract=contract, c2_float=False,
                     rest_double=True))
    alt = r32(solve([[r32(c1 * s), r32(c2 * s)]], contract=contract, **RULE))
    assert want != alt
    inc = np.array([[[to32(c1), to32(c2)]]], dtype=np.float32)
    with O.contracted(contract):
        got = O.pde_pairwise_f32_dscale(inc, 0, np.array([float(s)]))[0]
    assert f32_bits(got) == f32_bits(to32(want))


@pytest.mark.parametrize("contract,c", R5)
def test_contraction_setting_float64(contract, c):
    c = Fr(H(c))
    want = solve([[c]], store32=False, contract=contract, c2_float=False, rest_double=True)
    alt = solve([[c]], store32=False, contract=not contract, c2_float=False, rest_double=True)
    assert want != alt
    with O.contracted(contract):
        got = O.pde_pairwise(np.array([[[float(c)]]]), 0)[0]
    assert Fr(float(got)) == want


def test_contracted_form_is_nvvm_fma_pattern():
    # fma(c, 0.5, 1) is 1 + 0.5 c exactly, so the contracted A differs from the written one only
    # in its last addition; B = fma(c2, -1/12, 1) and k = fma(k01 + k10, A, -(k00 B))
    c = Fr(H("0x1.f247c45a19a50p-1"))
    c2 = r64(c * c)
    k = pde_step(Fr(1), Fr(1), Fr(1), c, c2_float=False, rest_double=True, contract=True, c2=c2)
    tw = r64(Fr(1, 12))
    A = r64(c2 * tw + (1 + c / 2))
    B = r64(1 - c2 * tw)
    assert k == r64(2 * A - B)


# ----------------------------------------------------------------------------- linear increments
def test_linear_increments_float32_quantized_are_exact():
    # points on the grid 2^-6 with small steps: every point Gram entry and every 4-term difference
    # is exact in float32, so the oracle's increments equal the exact rational values
    from helpers import quantized_paths
    x = quantized_paths(1, 10, 8, 21)[0]
    y = quantized_paths(1, 13, 8, 22)[0]
    for order in (0, 2):
        got = O.linear_inc(x, y, order, 1.0)
        X = [[fr(v) for v in row] for row in x]
        Y = [[fr(v) for v in row] for row in y]
        ip = [[sum(a * b for a, b in zip(xs, yt)) for yt in Y] for xs in X]
        for s in range(len(X) - 1):
            for t in range(len(Y) - 1):
                exact = (ip[s + 1][t + 1] + ip[s][t] - ip[s + 1][t] - ip[s][t + 1]) / 4 ** order
                assert fr(got[s, t]) == exact


def test_linear_increments_float32_rounding_order():
    # dim 1: ip = float32(kscale_f32 * float32(x y)); D = ((ip11 + ip00) - ip10) - ip01 in float32,
    # then / 4^order (exact); the alternative grouping (ip11 - ip10) - (ip01 - ip00) and a float64
    # point Gram both change some entries
    rng = np.random.default_rng(5)
    x = rng.normal(size=(9, 1)).astype(np.float32)
    y = rng.normal(size=(11, 1)).astype(np.float32)
    ks = 0.7
    ks32 = r32(Fr(ks))
    ip = [[r32(ks32 * r32(fr(a[0]) * fr(b[0]))) for b in y] for a in x]
    ip64 = [[r64(Fr(ks) * fr(a[0]) * fr(b[0])) for b in y] for a in x]
    for order in (0, 1):
        got = O.linear_inc(x, y, order, ks)
        n_alt = n_64 = 0
        for s in range(8):
            for t in range(10):
                want = r32(r32(r32(ip[s + 1][t + 1] + ip[s][t]) - ip[s + 1][t]) - ip[s][t + 1])
                want = want / 4 ** order
                alt = r32(r32(ip[s + 1][t + 1] - ip[s + 1][t]) - r32(ip[s][t + 1] - ip[s][t]))
                a64 = r32(r32(r32(ip64[s + 1][t + 1] + ip64[s][t]) - ip64[s + 1][t]) -
                          ip64[s][t + 1]) / 4 ** order
                n_alt += alt / 4 ** order != want
                n_64 += a64 != want
                assert fr(got[s, t]) == want
        assert n_alt > 0 and n_64 > 0


# ----------------------------------------------------------------------------- whole grids
@pytest.mark.parametrize("contract", [False, True])
@pytest.mark.parametrize("m,n,order,scaled", [(2, 3, 0, False), (3, 2, 1, True), (2, 3, 2, False),
                                              (4, 6, 1, True)])
def test_small_grids_match_exact_model_float32(contract, m, n, order, scaled):
    rng = np.random.default_rng(100 * m + 10 * n + order)
    inc = (rng.normal(size=(1, m, n)) * 0.6 / (1 << order)).astype(np.float32)
    sx = np.float32([rng.uniform(0.3, 1.0)]) if scaled else None
    sy = np.float32([rng.uniform(0.3, 1.0)]) if scaled else None
    scale = r32(fr(sy[0]) * fr(sx[0])) if scaled else Fr(1)
    C = [[r32(fr(v) * scale) for v in row] for row in inc[0]]
    want = r32(solve(C, order=order, contract=contract, **RULE))
    with O.contracted(contract):
        got = O.pde_pairwise(inc, order, sx, sy)[0]
    assert f32_bits(got) == f32_bits(to32(want))


@pytest.mark.parametrize("contract", [False, True])
@pytest.mark.parametrize("m,n,order", [(2, 3, 0), (3, 3, 1), (2, 2, 2)])
def test_small_grids_match_exact_model_float64(contract, m, n, order):
    rng = np.random.default_rng(7 + m + n + order)
    inc = rng.normal(size=(1, m, n)) * 0.6 / (1 << order)
    C = [[fr(v) for v in row] for row in inc[0]]
    want = solve(C, order=order, store32=False, contract=contract, c2_float=False,
                 rest_double=True)
    with O.contracted(contract):
        got = O.pde_pairwise(inc, order)[0]
    assert Fr(float(got)) == want
