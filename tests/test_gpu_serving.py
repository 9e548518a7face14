"""Serving loop on the GPU: a request's result is the same whether it is served alone (a direct API
call) or inside a continuous-batching mix; sampled requests match the float64 oracle; the HBM
budget is respected; HIP-graph replays equal eager launches."""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from sigpde_amd import serving as S
from helpers import brownian_np

pytestmark = pytest.mark.gpu
DEV = "cuda"


def quantized(B, L, d, seed):
    # dyadic-grid paths: every point-Gram entry and increment is exact in float32, so results do
    # not depend on the summation order of any product (bitwise comparisons are meaningful)
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.integers(-4, 5, size=(B, L, d)), axis=1) / 32.0
    x[:, 0] = 0
    return torch.from_numpy(x.astype(np.float32)).to(DEV)


def alone(r):
    k = r.static_kernel
    if r.kind == "pairwise":
        return sig.SigPDE(k, r.dyadic_order, r.precision).pairwise(r.x, r.y)
    if r.kind == "gram":
        return sig.SigPDE(k, r.dyadic_order, r.precision).gram(r.x, r.y)
    rk = sig.RobustSigPDE(k, r.dyadic_order, r.normalizer, r.precision)
    if r.kind == "normalization":
        return rk.normalization(r.x, tol=r.tol, maxit=r.maxit)
    if r.kind == "robust_pairwise":
        return rk.pairwise(r.x, r.y, tol=r.tol, maxit=r.maxit)
    return rk.gram(r.x, r.y, tol=r.tol, maxit=r.maxit)


def mix(seed, n=24):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        kind = S.KINDS[i % len(S.KINDS)]
        B = int(rng.integers(1, 6))
        L = int(rng.choice([9, 12, 17]))     # few lengths: several requests share a class
        d = int(rng.choice([2, 4]))
        order = int(rng.integers(0, 2))
        x = quantized(B, L, d, 100 + i)
        if kind in ("gram", "robust_gram"):
            y = quantized(int(rng.integers(1, 5)), L, d, 200 + i) if i % 2 else None
        elif kind == "normalization":
            y = None
        else:
            y = quantized(B, L, d, 300 + i)
        # fused precision for some SigPDE requests and every robust Gram (launched alone); the other
        # robust kinds stay "ref" so that several requests share their shape classes
        prec = "fp32" if ((kind in ("pairwise", "gram") and i % 3 == 0)
                          or kind == "robust_gram") else "ref"
        reqs.append(S.Request(kind, x, y, dyadic_order=order, precision=prec,
                              tol=1e-6, maxit=60))
    return reqs


def test_alone_equals_in_mix_exact_lengths():
    reqs = mix(1)
    ref = [alone(r) for r in reqs]
    srv = S.Server(DEV, bucket_lengths=None, streams=3, max_pairs=12)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert srv.launches < len(reqs)  # requests were actually batched together
    for r, a in zip(reqs, ref):
        assert r.error is None, r.error
        assert torch.equal(r.result, a), (r.kind, r.id)


def test_alone_equals_in_mix_with_length_buckets():
    reqs = mix(2)
    ref = [alone(r) for r in reqs]
    srv = S.Server(DEV, streams=2)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    for r, a in zip(reqs, ref):
        assert r.error is None, r.error
        # padding adds zero-increment cells: equal up to float32 rounding of those cells
        # (root-find results move within the root-find's tolerance)
        tol = 1e-5 if r.kind in ("pairwise", "gram") else 2e-4
        torch.testing.assert_close(r.result, a, rtol=tol, atol=1e-6)


def test_sampled_requests_match_oracle():
    rng = np.random.default_rng(3)
    reqs = []
    for i in range(6):
        L = int(rng.choice([10, 14, 20]))
        x = brownian_np(3, L, 3, 1 / L, 10 + i)
        y = brownian_np(3, int(rng.choice([10, 14])), 3, 1 / L, 20 + i)
        kind = "pairwise" if i % 2 else "gram"
        reqs.append(S.Request(kind, torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV),
                              dyadic_order=i % 3))
    srv = S.Server(DEV)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    for r in reqs:
        x, y = r.x.cpu().numpy(), r.y.cpu().numpy()
        want = (O.sig_pairwise(x, y, r.dyadic_order) if r.kind == "pairwise"
                else O.sig_gram(x, y, r.dyadic_order))
        np.testing.assert_allclose(r.result.cpu().numpy(), want, rtol=1e-9)


def test_budget_serialises_and_rejects():
    reqs = [S.Request("pairwise", quantized(4, 17, 4, i), quantized(4, 17, 4, 50 + i),
                      dyadic_order=i % 2) for i in range(6)]
    one = max(S.estimate_bytes("pairwise", "ref", 4, 4, 24, 24, 4, 0, 4, 4096), 1)
    srv = S.Server(DEV, hbm_budget=one)  # room for one batch at a time
    peak = []
    orig = srv._launch

    def launch(b):
        orig(b)
        peak.append(srv._inflight_bytes)
    srv._launch = launch
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert all(r.error is None for r in reqs)
    assert max(peak) <= one
    big = S.Request("gram", quantized(64, 300, 64, 9), None)
    srv2 = S.Server(DEV, hbm_budget=1 << 20)
    srv2.run([big], realtime=False)
    assert isinstance(big.error, MemoryError)


@pytest.mark.parametrize("prec", ["fp16x2", "fp32", "fp64"])
def test_graph_replays_equal_eager(prec):
    k = sig.kernels.LinearKernel()
    srv = S.Server(DEV, use_graphs=True, streams=1)
    for rep in range(4):
        reqs = [S.Request("pairwise", quantized(3, 17, 4, 10 * rep + i),
                          quantized(3, 17, 4, 10 * rep + i + 5), dyadic_order=1, precision=prec)
                for i in range(3)]
        srv.run(reqs, realtime=False)
        torch.cuda.synchronize()
        # eager work between replays
        eager = [sig.SigPDE(k, 1, prec).pairwise(r.x, r.y) for r in reqs]
        torch.cuda.synchronize()
        for r, e in zip(reqs, eager):
            assert r.error is None, r.error
            assert torch.equal(r.result, e), (prec, rep)
    assert len(srv._graphs) == 1  # one shape class, one padded size: captured once


def test_graph_replays_equal_eager_long_rows_of_columns():
    """2000 refined columns: an eager fused launch keeps its boundary lines in global memory
    (allocated stream-ordered), a captured one keeps them in LDS so that the graph allocates
    nothing; the results are the same bit for bit."""
    k = sig.kernels.LinearKernel()
    srv = S.Server(DEV, use_graphs=True, streams=1)
    for rep in range(3):
        # scaled by 1/8 (exact): K of 1000-step paths stays moderate
        reqs = [S.Request("pairwise", quantized(2, 1001, 4, 40 + 10 * rep + i) / 8,
                          quantized(2, 1001, 4, 45 + 10 * rep + i) / 8, dyadic_order=1,
                          precision="fp16x2") for i in range(2)]
        srv.run(reqs, realtime=False)
        torch.cuda.synchronize()
        eager = [sig.SigPDE(k, 1, "fp16x2").pairwise(r.x, r.y) for r in reqs]
        torch.cuda.synchronize()
        for r, e in zip(reqs, eager):
            assert r.error is None, r.error
            assert torch.equal(r.result, e), rep


def test_summary_has_latency_percentiles():
    reqs = S.synthetic_requests(12, seed=5, device=DEV, max_batch_size=8, lengths=(16, 40),
                                orders=(0, 1))
    srv = S.Server(DEV)
    import time
    t0 = time.perf_counter()
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    s = S.summarize(reqs, time.perf_counter() - t0)
    assert s["failed"] == 0 and s["requests"] == 12
    assert 0 < s["latency_ms"]["p50"] <= s["latency_ms"]["p99"]


def test_warmup_serves_every_kind_and_forgets_it():
    """Server.warmup runs a tiny request of every kind (loading the device code) and leaves no
    trace in the server's counters; real requests are served after it as usual."""
    srv = S.Server(DEV, streams=2)
    srv.warmup(precisions=("ref", "fp16x2", "fp64"), dtypes=(torch.float32, torch.float64))
    assert srv.launches == 0 and srv.completed == [] and not srv.pending()
    reqs = S.synthetic_requests(4, seed=3, device=DEV, max_batch_size=4, lengths=(16, 32))
    srv.run(reqs, realtime=False)
    assert all(r.error is None for r in reqs) and len(srv.completed) == 4


def test_empty_requests_complete_with_empty_results():
    """Requests with a batch of no paths, of every kind, in a mix with ordinary ones: each completes
    with an empty result of the direct call's shape, and the ordinary ones are unaffected."""
    reqs = []
    for i, kind in enumerate(S.KINDS):
        e = quantized(0, 9, 3, 70 + i)
        y = None if kind in ("normalization", "gram", "robust_gram") else e
        prec = "fp16x2" if kind in ("pairwise", "gram") else "ref"
        reqs.append(S.Request(kind, e, y, dyadic_order=1, precision=prec))
        x = quantized(3, 9, 3, 80 + i)
        reqs.append(S.Request(kind, x, None if y is None else quantized(3, 9, 3, 90 + i),
                              dyadic_order=1, precision=prec))
    ref = [alone(r) for r in reqs]
    srv = S.Server(DEV, streams=2)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    for r, a in zip(reqs, ref):
        assert r.error is None, (r.kind, r.error)
        assert r.result.shape == a.shape, r.kind
        torch.testing.assert_close(r.result, a, rtol=2e-4, atol=1e-6)


@pytest.mark.parametrize("prec", ["fp16x2", "fp32"])
def test_graph_replays_on_several_streams_equal_eager(prec):
    """Several batches of one shape class in flight on three streams at once: each stream replays
    its own captured graph with its own static buffers, and every result equals the eager call."""
    k = sig.kernels.LinearKernel()
    srv = S.Server(DEV, use_graphs=True, streams=3, max_pairs=3)
    reqs = [S.Request("pairwise", quantized(3, 17, 4, 500 + i), quantized(3, 17, 4, 600 + i),
                      dyadic_order=1, precision=prec) for i in range(12)]
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert srv.launches == 12 and len(srv._graphs) == 3  # one graph per stream
    eager = [sig.SigPDE(k, 1, prec).pairwise(r.x, r.y) for r in reqs]
    torch.cuda.synchronize()
    for r, e in zip(reqs, eager):
        assert r.error is None, r.error
        assert torch.equal(r.result, e), r.id


def test_fused_gram_requests_share_a_launch_and_equal_alone():
    """Fused Gram and robust-Gram requests of one class are served by one batched launch each,
    and every result equals the direct API call bit for bit (exact lengths)."""
    reqs = []
    for i in range(6):
        x = quantized(3 + i, 17, 4, 700 + i)
        y = None if i % 2 else quantized(2 + i, 17, 4, 800 + i)
        kind = "robust_gram" if i >= 4 else "gram"
        reqs.append(S.Request(kind, x, y, dyadic_order=1, precision="fp16x2", tol=1e-6,
                              maxit=40))
    ref = [alone(r) for r in reqs]
    srv = S.Server(DEV, bucket_lengths=None, streams=2)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    # classes: gram non-symmetric (2), gram symmetric (2), robust_gram each (1 + 1)
    assert srv.launches == 4
    for r, a in zip(reqs, ref):
        assert r.error is None, r.error
        assert torch.equal(r.result, a), (r.kind, r.id)


def _serving_mix(n, seed, max_batch):
    return S.synthetic_requests(n, seed=seed, device=DEV, max_batch_size=max_batch,
                                precision="fp16x2")


def test_serving_mix_real_shapes_alone_equals_in_mix_and_oracle():
    """40 requests of the serving mix at its real shapes (70% SigPDE / 30% RobustSigPDE, pairwise
    or Gram, lengths 16-512, dims 4/16/64, dyadic orders 0-3, float32, final solves fp16x2), the
    batch size capped at 12 to keep the oracle cheap. Served with exact lengths, every result
    equals the direct API call bit for bit; with length buckets (zero-increment padding), within
    rounding of it (the robust kinds within the float32 root-find's resolution); sampled requests
    match the float64 oracle."""
    reqs = _serving_mix(40, 11, 12)
    kinds = {r.kind for r in reqs}
    assert {"pairwise", "gram", "robust_pairwise", "robust_gram"} <= kinds
    assert max(r.dyadic_order for r in reqs) == 3 and max(r.x.shape[1] for r in reqs) > 400
    ref = [alone(r) for r in reqs]
    torch.cuda.synchronize()
    srv = S.Server(DEV, bucket_lengths=None)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    for r, a in zip(reqs, ref):
        assert r.error is None, (r.kind, r.error)
        assert torch.equal(r.result, a), (r.kind, r.id, r.x.shape, r.dyadic_order)
    exact_launches = srv.launches
    reqs2 = _serving_mix(40, 11, 12)
    srv2 = S.Server(DEV)
    srv2.run(reqs2, realtime=False)
    torch.cuda.synchronize()
    # exact lengths rarely share a class in this mix; buckets do
    assert srv2.launches < len(reqs2) and srv2.launches <= exact_launches
    for r, a in zip(reqs2, ref):
        assert r.error is None, (r.kind, r.error)
        tol = 1e-5 if r.kind in ("pairwise", "gram") else 1e-4
        torch.testing.assert_close(r.result, a, rtol=tol, atol=1e-5 * float(a.abs().max()) + 1e-6)
    # sampled requests against the float64 oracle (float64 inputs equal to the float32 paths)
    checked = 0
    for r in reqs:
        cells = S.request_cells(r)
        if r.kind not in ("pairwise", "gram") or cells > 3e8:
            continue
        x = r.x.double().cpu().numpy()
        y = (r.x if r.y is None else r.y).double().cpu().numpy()
        want = (O.linear_gram(x, y, r.dyadic_order) if r.kind == "gram"
                else O.sig_pairwise(x, y, r.dyadic_order))
        got = r.result.double().cpu().numpy()
        assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want), (r.kind, r.id)
        checked += 1
    assert checked >= 5
