"""CPU checks of the multi-wave schedule (sigpde_amd/csrc/mw_solver.h), restated here from its
stated rules (bands of 64 R refined rows, R per lane; band b on wave b mod W from period b D for
P = ceil((N + 63) / 16) periods,
D = max(5, ceil((P + 1) / (W - 1))), windows of 16 (D + 2) values, one barrier per period):
  * every window value is written in an earlier period than the one that reads it;
  * no window slot is overwritten before its reader has read it (neither by the same band's later
    columns nor by the band W later that reuses the window);
  * a wave's bands do not overlap in time;
and an emulation of the whole solve (periods, lanes, DPP shifts, tagged windows) on small grids
against the oracle.
"""
import numpy as np
import pytest

import oracle as O


def geom(M, N, W, R=1):
    nb = (M + 64 * R - 1) // (64 * R)
    P = (N + 63 + 15) // 16
    D = max(5, (P + 1 + W - 2) // (W - 1))
    return nb, P, D, 16 * (D + 2), (nb - 1) * D + P


@pytest.mark.parametrize("W", [4, 8, 16])
@pytest.mark.parametrize("M,N", [(1024, 1), (1024, 7), (1100, 40), (4088, 4088), (1032, 16000),
                                 (70000, 33), (2000, 600)])
def test_schedule_invariants(W, M, N):
    nb, P, D, win, T = geom(M, N, W)
    start = lambda b: b * D  # noqa: E731
    v = np.arange(1, N + 1)
    for b in range(1, min(nb, 3 * W)):
        wr = start(b - 1) + (v + 62) // 16       # lane 63 of band b-1 writes column v
        rd = start(b) + (v - 1) // 16            # lane 0 of band b reads it
        assert np.all(wr < rd)
        nxt = v + win                             # the same band's next write into the slot
        ok = nxt <= N
        assert np.all(start(b - 1) + (nxt[ok] + 62) // 16 > rd[ok])
        if b + W - 1 < nb:                        # band b+W-1 writes this window next
            assert start(b + W - 1) + 63 // 16 > rd[-1]
    for b in range(nb - W):
        assert start(b) + P <= start(b + W)      # one band at a time per wave
    assert T == (nb - 1) * D + P


def emulate(inc, o, W, R=1):
    """The solve as the schedule runs it: lane rho of band b holds refined rows
    64 R b + R rho + 1 .. + R (one row of increments), window values tagged with their band,
    column and period."""
    m, n = inc.shape
    M, N = m << o, n << o
    assert M % R == 0
    nb, P, D, win, T = geom(M, N, W, R)
    wins = [[None] * win for _ in range(W)]
    k = np.ones((W, 64, R))
    kprev = np.ones((W, 64))
    result = None
    for p in range(T):
        for wave in range(W):
            rel = p - wave * D
            if rel < 0:
                continue
            b = wave + W * (rel // (W * D))
            q = p - b * D
            if b >= nb or q >= P:
                continue
            if q == 0:
                k[wave][:] = 1
                kprev[wave][:] = 1
            for st in range(16):
                new = k[wave].copy()
                ups = np.empty(64)
                for lane in range(64):
                    v = 16 * q + st - lane + 1
                    wi = (v + 63) % win                 # lane 0's 16 columns: an aligned block
                    if lane == 0:
                        up = 1.0
                        if b > 0 and 1 <= v <= N:
                            tag = wins[wave][wi]
                            assert tag is not None and tag[:2] == (b, v) and tag[3] < p
                            up = tag[2]
                    else:
                        up = k[wave][lane - 1][R - 1]
                    ups[lane] = up
                    i0 = 64 * R * b + R * lane + 1
                    if 1 <= v <= N and i0 + R - 1 <= M:
                        c = inc[(i0 - 1) >> o][(v - 1) >> o]
                        A = (1 + 0.5 * c) + c * c / 12
                        B = 1 - c * c / 12
                        above, above_old = up, kprev[wave][lane]
                        for r in range(R):
                            old = k[wave][lane][r]
                            new[lane][r] = (above + old) * A - above_old * B
                            above_old, above = old, new[lane][r]
                        if lane == 63 and b + 1 < nb:
                            wins[(wave + 1) % W][wi] = (b + 1, v, new[lane][R - 1], p)
                        for r in range(R):
                            if i0 + r == M and v == N:
                                result = new[lane][r]
                k[wave] = new
                kprev[wave] = ups
    return result


@pytest.mark.parametrize("m,n,o,W,R", [(40, 9, 2, 4, 2), (70, 12, 1, 8, 2), (33, 5, 3, 4, 2),
                                       (17, 30, 2, 4, 4)])
def test_emulated_solve_rows_per_lane_matches_oracle(m, n, o, W, R):
    rng = np.random.default_rng(m + n + R)
    inc = rng.normal(size=(m, n)) * (1.2 / np.sqrt(m * n)) / 4 ** o
    ref = O.pde_pairwise(inc[None].copy(), o)[0]
    assert abs(emulate(inc, o, W, R) - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("m,n,o", [(130, 20, 0), (40, 9, 2)])
def test_emulated_solve_matches_oracle(m, n, o):
    rng = np.random.default_rng(m + n)
    inc = rng.normal(size=(m, n)) * (1.2 / np.sqrt(m * n)) / 4 ** o
    ref = O.pde_pairwise(inc[None].copy(), o)[0]
    assert abs(emulate(inc, o, 4) - ref) <= 1e-12 * abs(ref)
