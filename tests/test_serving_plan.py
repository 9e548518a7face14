"""Host logic of the serving loop (no GPU): length buckets, padding, batch classes, the HBM budget,
and the padding invariance the buckets rely on, checked with the oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from sigpde_amd import serving as S
from sigpde_amd.torch import kernels
from helpers import brownian_np


def req(kind, B, Lx, Ly=None, d=4, order=0, y=True, **kw):
    x = torch.zeros((B, Lx, d), dtype=torch.float32)
    yy = torch.zeros((B if kind != "gram" else 3, Ly or Lx, d), dtype=torch.float32) if y else None
    return S.Request(kind, x, yy, dyadic_order=order, **kw)


def test_bucket_length():
    b = S.DEFAULT_BUCKETS
    assert S.bucket_length(16, b) == 16
    assert S.bucket_length(17, b) == 24
    assert S.bucket_length(300, b) == 384
    assert S.bucket_length(5000, b) == 5000  # beyond every bucket: unchanged
    assert S.bucket_length(37, None) == 37


def test_pad_paths_repeats_last_point():
    x = torch.from_numpy(brownian_np(3, 10, 2, 0.1, 0))
    p = S.pad_paths(x, 16)
    assert p.shape == (3, 16, 2)
    assert torch.equal(p[:, :10], x)
    assert torch.equal(p[:, 10:], x[:, -1:].expand(3, 6, 2))
    assert S.pad_paths(x, 10) is x


@pytest.mark.parametrize("order", [0, 1, 2])
def test_padding_leaves_kernel_unchanged(order):
    # zero increments reduce the PDE step to k = k01 + k10 - k00: the corner value is unchanged
    x = brownian_np(4, 9, 3, 1 / 9, 1)
    y = brownian_np(4, 13, 3, 1 / 13, 2)
    xp = S.pad_paths(torch.from_numpy(x), 16).numpy()
    yp = S.pad_paths(torch.from_numpy(y), 24).numpy()
    a = O.sig_pairwise(x, y, order)
    b = O.sig_pairwise(xp, yp, order)
    np.testing.assert_allclose(b, a, rtol=1e-12)
    # a plausible bucketing mistake (padding with zeros instead of the last point) changes it
    xz = np.concatenate([x, np.zeros((4, 7, 3))], axis=1)
    assert np.max(np.abs(O.sig_pairwise(xz, y, order) - a) / np.abs(a)) > 1e-3


def test_batch_key_classes():
    b = S.DEFAULT_BUCKETS
    k1 = S.batch_key(req("pairwise", 2, 20), b)
    k2 = S.batch_key(req("pairwise", 5, 23), b)           # same bucket (24)
    k3 = S.batch_key(req("pairwise", 2, 20, order=1), b)  # other order
    k4 = S.batch_key(req("pairwise", 2, 20, d=16), b)     # other dim
    k5 = S.batch_key(req("normalization", 2, 20, y=False), b)
    k6 = S.batch_key(S.Request("pairwise", torch.zeros(2, 20, 4), torch.zeros(2, 20, 4),
                               static_kernel=kernels.RBFKernel(0.5)), b)
    assert k1 == k2
    assert len({k1, k3, k4, k5, k6}) == 5


def test_planner_groups_separable_requests_up_to_max_pairs():
    srv = S.Server("cpu", max_pairs=10)
    for B in (4, 4, 4, 3):
        srv.submit(req("pairwise", B, 20))
    srv.submit(req("gram", 2, 20))
    b1 = srv._take_batch()
    assert [r.x.shape[0] for r in b1.reqs] == [4, 4]  # a third would exceed max_pairs
    srv._inflight.append(b1)
    srv._inflight_bytes += b1.nbytes
    b2 = srv._take_batch()
    assert [r.x.shape[0] for r in b2.reqs] == [4, 3]
    srv._inflight.append(b2)
    srv._inflight_bytes += b2.nbytes
    b3 = srv._take_batch()
    assert b3.reqs[0].kind == "gram" and len(b3.reqs) == 1
    assert srv._take_batch() is None


def test_planner_budget():
    small = req("pairwise", 2, 20)
    need = S.estimate_bytes("pairwise", "ref", 2, 2, 24, 24, 4, 0, 4, 4096)  # bucket 24
    srv = S.Server("cpu", hbm_budget=need)
    srv.submit(small)
    srv.submit(req("pairwise", 2, 20, order=1))
    b = srv._take_batch()
    assert b is not None and b.nbytes <= srv.hbm_budget
    srv._inflight.append(b)
    srv._inflight_bytes += b.nbytes
    assert srv._take_batch() is None  # waits for the first to retire
    # a request that can never fit fails alone instead of blocking the queue
    srv2 = S.Server("cpu", hbm_budget=1000)
    big = srv2.submit(req("pairwise", 2, 20))
    assert srv2._take_batch() is None
    assert isinstance(big.error, MemoryError) and big in srv2.completed


def test_synthetic_mix_shapes():
    reqs = S.synthetic_requests(50, seed=7, device="cpu")
    kinds = {r.kind for r in reqs}
    assert kinds <= set(S.KINDS) and len(kinds) >= 3
    for r in reqs:
        assert 1 <= r.x.shape[0] <= 256 and 16 <= r.x.shape[1] <= 512
        assert ((max(r.x.shape[1], (r.y if r.y is not None else r.x).shape[1]) - 1)
                << r.dyadic_order) <= 4096


def test_fused_robust_estimate_counts_root_find_self_increments():
    # the root-find's self-increments exist whatever precision the final solves take
    for kind in ("robust_pairwise", "normalization"):
        ref = S.estimate_bytes(kind, "ref", 4096, 4096, 512, 512, 16, 0, 4, 4096)
        m = 511
        root_find = 2 * 4096 * (m * m + m * m) * 4
        assert ref >= root_find
    fused = S.estimate_bytes("robust_pairwise", "fp16x2", 4096, 4096, 512, 512, 16, 0, 4, 4096)
    assert fused >= root_find


def gram_req(B, L, By=None, d=4, order=0, precision="fp16x2", kind="gram", symmetric=False):
    x = torch.zeros((B, L, d), dtype=torch.float32)
    y = None if symmetric else torch.zeros((By or B, L, d), dtype=torch.float32)
    return S.Request(kind, x, y, dyadic_order=order, precision=precision)


def test_fused_gram_requests_coalesce_per_class():
    srv = S.Server("cpu", max_gram_pairs=1000)
    a = srv.submit(gram_req(10, 20, 12))
    b = srv.submit(gram_req(8, 23, 9))          # same class: lengths bucket to 24
    c = srv.submit(gram_req(5, 20, 7, d=16))    # another dim
    d = srv.submit(gram_req(6, 20, symmetric=True))   # symmetric: a class of its own
    e = srv.submit(gram_req(6, 20, 12, precision="ref"))  # reference scheme: alone
    f = srv.submit(gram_req(30, 20, 30))        # would pass max_gram_pairs with a and b
    assert S.gram_coalesced(a) and not S.gram_coalesced(e)
    got = []
    while True:
        bt = srv._take_batch()
        if bt is None:
            break
        got.append([r.id for r in bt.reqs])
    assert [a.id, b.id] in got and [f.id] in got and [c.id] in got and [d.id] in got
    assert [e.id] in got
    assert len(got) == 5


def test_gram_batch_caps_at_64_requests():
    srv = S.Server("cpu")
    for _ in range(70):
        srv.submit(gram_req(2, 20, 2))
    b1 = srv._take_batch()
    assert len(b1.reqs) == 64
    assert len(srv._take_batch().reqs) == 6


def test_cancel_and_timeout_of_queued_requests():
    srv = S.Server("cpu")
    a = srv.submit(req("pairwise", 2, 20))
    b = srv.submit(req("pairwise", 2, 20, order=1))
    c = srv.submit(req("pairwise", 2, 20, order=2, timeout=0.0))
    assert srv.cancel(a) is True
    assert isinstance(a.error, S.CancelledError) and a in srv.completed
    srv._expire()
    assert isinstance(c.error, TimeoutError) and c in srv.completed
    bt = srv._take_batch()
    assert bt.reqs == [b] and srv._take_batch() is None
