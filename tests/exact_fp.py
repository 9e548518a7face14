"""Exact rational model of the reference's per-cell arithmetic, for the oracle's known-answer tests.

Every value is a fractions.Fraction; rounding to binary32 / binary64 is done here, by an explicit
round-to-nearest-even on the significand (no numpy or C floating-point arithmetic is involved), so
the model is independent of the oracle it checks. Each typing choice of
sigpde/cuda_device_functions.py:5 under Numba is a switch, so a test can show
that the oracle lands on the value the rule gives and not on the value of the alternative.
"""
from __future__ import annotations

from fractions import Fraction as Fr

import numpy as np


def rnd(q: Fr, p: int) -> Fr:
    """Nearest binary floating-point value with a p-bit significand, ties to even (normal range)."""
    q = Fr(q)
    if q == 0:
        return Fr(0)
    neg = q < 0
    a = -q if neg else q
    e = a.numerator.bit_length() - a.denominator.bit_length()  # 2^e <= a < 2^(e+2)
    while Fr(2) ** e > a:
        e -= 1
    while Fr(2) ** (e + 1) <= a:
        e += 1
    ulp = Fr(2) ** (e - p + 1)
    n, r = divmod(a, ulp)  # a = n ulp + r, 0 <= r < ulp
    n = int(n)
    if r * 2 > ulp or (r * 2 == ulp and n % 2 == 1):
        n += 1
    v = n * ulp
    return -v if neg else v


def r32(q) -> Fr:
    return rnd(q, 24)


def r64(q) -> Fr:
    return rnd(q, 53)


def fr(x) -> Fr:
    """Exact value of a numpy / Python float."""
    return Fr(float(x)) if not isinstance(x, Fr) else x


def to32(q: Fr) -> np.float32:
    """A binary32 value as np.float32 (exact: q must already be a binary32 value)."""
    v = np.float32(float(q))
    assert Fr(float(v)) == q, "not a binary32 value"
    return v


def to64(q: Fr) -> float:
    v = float(q)
    assert Fr(v) == q, "not a binary64 value"
    return v


TWELFTH_64 = r64(Fr(1, 12))  # the literal 1./12 (a float64)
TWELFTH_32 = r32(Fr(1, 12))  # 1./12 cast to float32 (NEP-50 all-float arithmetic)


def pde_step(k00: Fr, k01: Fr, k10: Fr, c: Fr, *, c2_float: bool, rest_double: bool,
             contract: bool, c2: Fr | None = None) -> Fr:
    """(k01 + k10) * (1. + 0.5 * c + (1./12) * c**2) - k00 * (1. - (1./12) * c**2) with
    c**2 rounded to float32 (c2_float) or to the working precision, every other operation rounded
    to float64 (rest_double) or float32, and, with contract, NVVM's contraction
    A = fma(c2, 1/12, fma(c, 0.5, 1)), B = fma(c2, -1/12, 1), k = fma(k01 + k10, A, -(k00 B)).
    Returns the unrounded-to-storage result (a float64 or float32 value)."""
    R = r64 if rest_double else r32
    tw = TWELFTH_64 if rest_double else TWELFTH_32
    if c2 is None:
        c2 = r32(c * c) if c2_float else R(c * c)
    if contract:
        A = R(c2 * tw + R(c * Fr(1, 2) + 1))
        B = R(-c2 * tw + 1)
        return R((R(k01 + k10)) * A - R(k00 * B))
    A = R(R(1 + R(Fr(1, 2) * c)) + R(tw * c2))
    B = R(1 - R(tw * c2))
    return R(R(R(k01 + k10) * A) - R(k00 * B))


def solve(inc, *, store32: bool = True, c2_float: bool = True, rest_double: bool = True,
          contract: bool = False, order: int = 0) -> Fr:
    """k[M][N] of the Goursat scheme (sigpde/cuda_pairwise_kernels.py:8-39) on an m x n list of
    Fraction increments c (already scaled), boundary ones, row by row; each stored cell is rounded
    to float32 (store32) or kept in the working precision."""
    m, n = len(inc), len(inc[0])
    M, N = m << order, n << order
    prev = [Fr(1)] * (N + 1)
    for i in range(1, M + 1):
        cur = [Fr(1)] * (N + 1)
        for j in range(1, N + 1):
            c = inc[(i - 1) >> order][(j - 1) >> order]
            v = pde_step(prev[j - 1], prev[j], cur[j - 1], c, c2_float=c2_float,
                         rest_double=rest_double, contract=contract)
            cur[j] = r32(v) if store32 else v
        prev = cur
    return prev[N]
