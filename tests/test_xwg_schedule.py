"""CPU checks of the spread schedule (sigpde_amd/csrc/mw_solver.h with XW: one pair's bands over
G workgroups of W waves), restated from its stated rules:
  * WT = G W waves; band b on wave b mod WT (workgroup (b mod WT) // W), from period b D for
    P = ceil((N + 63) / 16) periods, D = max(Dmin, ceil((P + 1) / (WT - 1))), Dmin = 7;
  * LDS windows of 16 (D + 2) values inside a workgroup; between workgroups (wave W - 1 of
    workgroup g feeds wave 0 of workgroup g + 1 mod G) a global window of two buffers of
    16 (D + 2 + S) values, S = 8, buffer ((reader band) // WT) mod 2;
  * each workgroup runs its periods on its own; before period p it has published p completed
    periods; wave 0 (band b >= 1) waits for the feeding workgroup to have completed p - D + 5,
    wave W - 1 (band b with b + 1 < nb) for the fed workgroup to have completed p - 4 - S.
The emulation runs the workgroups in random interleavings that respect only those waits (and
runs of one workgroup as far as the waits let it), tags every window value with the band and
column it is meant for, asserts each read finds its own value (a read too early, or a slot
overwritten before it was read, fails), that the interleaving never deadlocks, and compares the
result with the oracle.
"""
import numpy as np
import pytest

import oracle as O

DMIN, SLACK = 7, 8


def geom(M, N, W, G, R=1):
    WT = W * G
    nb = (M + 64 * R - 1) // (64 * R)
    P = (N + 63 + 15) // 16
    D = max(DMIN, (P + 1 + WT - 2) // (WT - 1))
    return WT, nb, P, D, 16 * (D + 2), 16 * (D + 2 + SLACK), (nb - 1) * D + P


def spread_solve(inc, o, W, G, R, rng, runs=False, raw=5, war=4):
    m, n = inc.shape
    M, N = m << o, n << o
    WT, nb, P, D, win, wing, T = geom(M, N, W, G, R)
    lds = [[[None] * win for _ in range(W)] for _ in range(G)]
    glob = [[[None] * wing for _ in range(2)] for _ in range(G)]   # link l feeds workgroup l
    k = np.ones((WT, 64, R))
    kprev = np.ones((WT, 64))
    per = [0] * G                      # next period of each workgroup (= periods it completed)
    result = [None]

    def band_of(gw, p):
        rel = p - gw * D
        if rel < 0:
            return -1, 0
        b = gw + WT * (rel // (WT * D))
        q = p - b * D
        return (b, q) if b < nb and q < P else (-1, 0)

    def can_run(g):
        p = per[g]
        if p >= T:
            return False
        b0, _ = band_of(g * W, p)
        if b0 >= 1 and per[(g - 1) % G] < p - D + raw:
            return False
        bl, _ = band_of(g * W + W - 1, p)
        if bl >= 0 and bl + 1 < nb and per[(g + 1) % G] < p - war - SLACK:
            return False
        return True

    def run_period(g):
        p = per[g]
        for wave in range(W):
            gw = g * W + wave
            b, q = band_of(gw, p)
            if b < 0:
                continue
            if q == 0:
                k[gw][:] = 1
                kprev[gw][:] = 1
            if wave == 0 and b > 0:
                src, sz = glob[g][(b // WT) % 2], wing
            elif wave == 0:
                src, sz = None, win     # band 0: the boundary row of ones
            else:
                src, sz = lds[g][wave], win
            if wave == W - 1:
                dst, dsz = glob[(g + 1) % G][((b + 1) // WT) % 2], wing
            else:
                dst, dsz = lds[g][wave + 1], win
            i0 = 64 * R * b + R * np.arange(64) + 1
            for st in range(16):
                v = 16 * q + st - np.arange(64) + 1
                top = 1.0
                v0 = v[0]
                if src is not None and 1 <= v0 <= N:
                    tag = src[(v0 + 63) % sz]
                    assert tag is not None and tag[:2] == (b, v0), (g, wave, b, v0, tag, p)
                    top = tag[2]
                ups = np.concatenate(([top], k[gw][:-1, R - 1]))
                live = (v >= 1) & (v <= N) & (i0 + R - 1 <= M)
                c = np.zeros(64)
                rows = np.minimum((i0 - 1) >> o, m - 1)
                cols = np.clip((v - 1) >> o, 0, n - 1)
                c[live] = inc[rows[live], cols[live]]
                A = (1 + 0.5 * c) + c * c / 12
                B = 1 - c * c / 12
                above, above_old = ups.copy(), kprev[gw].copy()
                new = k[gw].copy()
                for r in range(R):
                    old = k[gw][:, r].copy()
                    val = (above + old) * A - above_old * B
                    new[live, r] = val[live]
                    above_old, above = old, new[:, r]
                if live[63] and b + 1 < nb:
                    dst[(v[63] + 63) % dsz] = (b + 1, int(v[63]), new[63, R - 1], p)
                for lane in np.nonzero(live & (v == N))[0]:
                    for r in range(R):
                        if i0[lane] + r == M:
                            result[0] = new[lane, r]
                k[gw] = new
                kprev[gw] = ups
        per[g] += 1

    while any(pp < T for pp in per):
        ready = [g for g in range(G) if can_run(g)]
        assert ready, f"deadlock at periods {per}"
        g = ready[rng.integers(len(ready))]
        if runs and rng.random() < 0.3:
            while can_run(g):            # one workgroup as far as its waits let it
                run_period(g)
        else:
            run_period(g)
    return result[0]


# small grids, and grids whose bands overlap (P well above D: WT - 1 >= (P + 1) / D) with
# windows that wrap (N above 16 (D + 2 + S))
@pytest.mark.parametrize("m,n,o,W,G,R", [(700, 40, 0, 2, 3, 1), (330, 24, 0, 1, 4, 1),
                                         (90, 10, 2, 2, 2, 2), (40, 9, 3, 1, 3, 2),
                                         (260, 3, 1, 2, 2, 2), (640, 700, 0, 2, 4, 1),
                                         (300, 330, 1, 2, 4, 2), (520, 400, 0, 1, 8, 1)])
@pytest.mark.parametrize("runs", [False, True])
def test_spread_solve_matches_oracle(m, n, o, W, G, R, runs):
    rng = np.random.default_rng(m * 7 + n + G + runs)
    inc = rng.normal(size=(m, n)) * (1.2 / np.sqrt(m * n)) / 4 ** o
    ref = O.pde_pairwise(inc[None].copy(), o)[0]
    got = spread_solve(inc, o, W, G, R, rng, runs)
    assert abs(got - ref) <= 1e-12 * abs(ref)


def test_wait_needs_stay_below_own_period():
    """The workgroup furthest behind never waits: both needs are below the waiting period."""
    for W, G in [(16, 2), (16, 4), (8, 8), (8, 32)]:
        for M, N in [(4000, 4000), (2048, 300), (16000, 16000), (640, 64)]:
            WT, nb, P, D, win, wing, T = geom(M, N, W, G)
            assert D >= 5 and DMIN - 5 >= 1
            for p in range(T):
                assert p - D + 5 <= p and p - 4 - SLACK < p
            # both buffers of a link: the band 2 WT later reuses one after its reader finished
            assert (2 * WT - 1) * D - P + 1 >= 5 + SLACK
            # within a band, a slot's next value comes >= 5 + S periods after its read
            v = np.arange(1, N + 1)
            rd = (v - 1) // 16 + D
            nxt = v + wing
            wr = (nxt + 62) // 16
            assert np.all(wr - rd >= 5 + SLACK)


@pytest.mark.parametrize("raw,war", [(4, 4), (5, 5)])
def test_weakened_wait_is_caught(raw, war):
    """A wait one period weaker than the kernel's lets some interleaving read a value before it
    is written (raw) or overwrite a slot before it is read (war): the emulation must see it."""
    rng = np.random.default_rng(11)
    inc = rng.normal(size=(640, 700)) * (1.2 / np.sqrt(640 * 700))
    caught = False
    for trial in range(6):
        try:
            spread_solve(inc, 0, 2, 4, 1, np.random.default_rng(trial), runs=True, raw=raw, war=war)
        except AssertionError:
            caught = True
            break
    assert caught
