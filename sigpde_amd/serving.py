"""Serving loop for signature-kernel requests on one MI355X: continuous batching under an HBM
budget, length buckets, several HIP streams, optional HIP-graph replay, p50/p99 latency.

The reference has no serving layer (it is a library called once per batch); this module drives
the same entry points (sigpde_amd.torch.SigPDE / RobustSigPDE, i.e. the reference's
sigpde/torch/kernel.py:26-301 API) for a stream of independent requests:

  * Requests of one "shape class" (kind, static kernel, dyadic order, precision, dtype, dim,
    bucketed lengths, root-find settings) that are pair- or path-separable (pairwise,
    normalization, robust pairwise) are concatenated into one launch. Gram and robust-Gram
    requests with a fused precision (the linear kernel on MFMA) are coalesced too: up to 64 of
    one class run as one launch of the batched fused Gram (sigpde_linear_gram_batch), their
    robust root-finds as one normalisation launch over all their paths; reference-precision
    Grams (increment tiles in HBM) run as launches of their own. Every pair or path of a batch is
    solved independently, so a request's result does not depend on what it was batched with.
  * Length buckets: a path is padded to its bucket's length by repeating its last point. The
    appended increments are zero, where the PDE step reduces to k = k01 + k10 - k00 and the
    solution at the corner is unchanged (exactly in real arithmetic; to rounding in floating
    point). Disable with bucket_lengths=None.
  * HBM budget: a batch is admitted only while the estimated bytes of the batches in flight
    (inputs, outputs, increment tiles or fused workspace) stay within the budget; a request that
    cannot fit alone fails with MemoryError.
  * Streams: each batch goes to the one of `streams` HIP streams (default priority) with the least
    estimated work in flight (ties round-robin); completion is tracked with events. Results are
    recorded on the stream current when they are handed over (poll), so a caller's later use
    keeps their memory alive.
  * HIP graphs (use_graphs=True, the default): fused pairwise batches (precision "fp16x2",
    "bf16x3", "fp32" or "fp64") are captured once per (shape class, padded batch size, stream)
    and replayed with the inputs copied into that graph's static buffers (one set per stream:
    batches of one class on two streams never share them).
  * Deadlines and cancellation: a request with a `timeout` (seconds after submission) that is
    still queued when it expires fails with TimeoutError without being launched; cancel(r)
    removes a queued request (CancelledError), or drops the result of a launched one.

`python -m sigpde_amd.serving` runs a synthetic request mix and prints one JSON line.
"""
from __future__ import annotations

import argparse
import collections
import dataclasses
import json
import math
import time
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _native
from .torch import kernels as _kernels
from .torch.kernel import RobustSigPDE, SigPDE, log_normalizer

KINDS = ("pairwise", "gram", "normalization", "robust_pairwise", "robust_gram")
SEPARABLE = ("pairwise", "normalization", "robust_pairwise")
DEFAULT_BUCKETS = (16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072,
                   4096)


@dataclasses.dataclass(eq=False)
class Request:
    """One call of the drop-in API.

    kind: "pairwise" (SigPDE.pairwise), "gram" (SigPDE.gram), "normalization"
    (RobustSigPDE.normalization), "robust_pairwise" (RobustSigPDE.pairwise), "robust_gram"
    (RobustSigPDE.gram). y=None has the API's meaning (y = x; symmetric Gram)."""
    kind: str
    x: torch.Tensor
    y: Optional[torch.Tensor] = None
    static_kernel: object = None
    dyadic_order: int = 0
    precision: str = "ref"
    normalizer: Optional[Callable] = None
    tol: float = 1e-8
    maxit: int = 100
    arrival: float = 0.0          # seconds after the server's start (synthetic arrival time)
    timeout: float = math.inf     # seconds after submission; a request still queued then fails
    # filled by the server
    id: int = -1
    t_submit: float = math.nan
    t_done: float = math.nan
    result: Optional[torch.Tensor] = None
    error: Optional[BaseException] = None
    cancelled: bool = False

    def __post_init__(self):
        if self.kind not in KINDS:
            raise ValueError(f"kind must be one of {KINDS}")
        if self.static_kernel is None:
            self.static_kernel = _kernels.LinearKernel()
        if self.kind == "normalization" and self.precision != "ref":
            raise ValueError("normalization requests run with precision='ref' (the root-find "
                             "uses the reference's arithmetic)")

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    def pairs(self) -> int:
        """Output entries (the metric's "pairs")."""
        bx = self.x.shape[0]
        if self.kind in ("gram", "robust_gram"):
            return bx * (bx if self.y is None else self.y.shape[0])
        return bx


def _kernel_key(k) -> Tuple:
    return (type(k).__name__, tuple(sorted((a, float(v)) for a, v in vars(k).items()
                                           if isinstance(v, (int, float)))))


def bucket_length(L: int, buckets: Optional[Tuple[int, ...]]) -> int:
    """Smallest bucket >= L (L itself when bucketing is off or L exceeds every bucket)."""
    if not buckets:
        return L
    for b in buckets:
        if b >= L:
            return b
    return L


def pad_paths(x: torch.Tensor, L: int) -> torch.Tensor:
    """(B, L0, d) -> (B, L, d) by repeating the last point (zero increments)."""
    L0 = x.shape[1]
    if L0 == L:
        return x
    tail = x[:, -1:, :].expand(x.shape[0], L - L0, x.shape[2])
    return torch.cat([x, tail], dim=1)


class CancelledError(Exception):
    """The request was cancelled (Server.cancel) before its result was handed over."""


def gram_coalesced(r: Request) -> bool:
    """Gram kinds whose requests of one class share a launch: fused precisions (the linear static
    kernel on the MFMA producer)."""
    return (r.kind in ("gram", "robust_gram") and r.precision != "ref"
            and getattr(r.static_kernel, "fused_linear_scale", None) is not None)


def batch_key(r: Request, buckets) -> Tuple:
    """Requests with equal keys may share one launch."""
    Lx = bucket_length(r.x.shape[1], buckets)
    if r.kind == "normalization":
        Ly = Lx
    elif r.y is None:
        Ly = -1  # symmetric: y is x
    else:
        Ly = bucket_length(r.y.shape[1], buckets)
    norm = None if r.normalizer in (None, log_normalizer) else id(r.normalizer)
    return (r.kind, _kernel_key(r.static_kernel), r.dyadic_order, r.precision, r.x.dtype,
            r.x.shape[2], Lx, Ly, norm, float(r.tol), int(r.maxit))


def estimate_bytes(kind: str, precision: str, B: int, By: int, Lx: int, Ly: int, d: int,
                   order: int, itemsize: int, max_batch: int) -> int:
    """Upper estimate of the device memory one launch holds (inputs, outputs, increment tiles or
    fused workspace, robust scratch)."""
    m, n = Lx - 1, Ly - 1
    if kind in ("gram", "robust_gram"):
        inputs = (B * Lx + By * Ly) * d * itemsize
        out = B * By * itemsize
        if precision == "ref":
            t = math.ceil(math.sqrt(max_batch))
            tiles = min(B, t) * min(By, t) * m * n * itemsize * 2  # increments + point Gram
        else:
            tiles = _native.linear_workspace_bytes(precision, B, Lx, By, Ly, d, False)
        robust = (B + By) * (m * m + n * n) * itemsize if kind == "robust_gram" else 0
        return inputs + out + tiles + min(robust, 2 * max_batch * (m * m + n * n) * itemsize)
    inputs = B * (Lx + Ly) * d * itemsize
    out = 3 * B * itemsize
    chunk = min(B, max_batch)
    if precision == "ref":
        inc = chunk * m * n * itemsize * 2
    else:
        inc = _native.linear_workspace_bytes(precision, B, Lx, B, Ly, d, False)
    if kind != "pairwise":
        # the root-find's self-increments of x and y (or of x and y joined), whatever precision
        # the final solves take
        inc += 2 * chunk * (m * m + n * n) * itemsize
    return inputs + out + inc


class _Batch:
    def __init__(self, key, reqs: List[Request]):
        self.key = key
        self.reqs = reqs
        self.nbytes = 0
        self.cost = sum(request_cost(r) for r in reqs)
        self.event: Optional[torch.cuda.Event] = None
        self.outputs: List[torch.Tensor] = []
        self.stream = None


class _GraphEntry:
    def __init__(self):
        self.graph = None
        self.X = self.Y = self.out = None


class Server:
    """Continuous-batching server for one GPU.

    hbm_budget: bytes of device memory the in-flight launches may hold (default 64 GiB);
    max_pairs: pairs or paths per separable batch; max_gram_pairs: output pairs per coalesced
    Gram batch (at most 64 requests); streams: HIP streams (default priority); bucket_lengths:
    length buckets for separable and coalesced Gram batches (None: exact lengths only);
    max_batch: the API's max_batch (increment tile size of precision "ref"); use_graphs: replay
    fused pairwise batches through captured HIP graphs."""

    def __init__(self, device="cuda", hbm_budget: int = 64 << 30, max_pairs: int = 16384,
                 streams: int = 4, bucket_lengths=DEFAULT_BUCKETS, max_batch: int = 4096,
                 use_graphs: bool = True, max_gram_pairs: int = 1 << 22):
        self.device = torch.device(device)
        self.hbm_budget = int(hbm_budget)
        self.max_pairs = int(max_pairs)
        self.max_gram_pairs = int(max_gram_pairs)
        self.buckets = tuple(bucket_lengths) if bucket_lengths else None
        self.max_batch = int(max_batch)
        self.use_graphs = use_graphs
        self.nstreams = max(1, int(streams))
        self.streams: List = []  # created at the first launch (planning needs no GPU)
        self._next_stream = 0
        self._queue: "collections.OrderedDict[Tuple, List[Request]]" = collections.OrderedDict()
        self._inflight: List[_Batch] = []
        self._inflight_bytes = 0
        self._next_id = 0
        self._graphs: Dict[Tuple, _GraphEntry] = {}
        self.t0 = time.perf_counter()
        self.completed: List[Request] = []
        self.launches = 0

    # ------------------------------------------------------------------ submission
    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, r: Request) -> Request:
        r.id = self._next_id
        self._next_id += 1
        r.t_submit = self.now()
        key = batch_key(r, self.buckets if (r.kind in SEPARABLE or gram_coalesced(r)) else None)
        self._queue.setdefault(key, []).append(r)
        return r

    def cancel(self, r: Request) -> bool:
        """Cancels r: a queued request is removed and fails with CancelledError (True); a launched
        one still runs, and its result is dropped when its batch completes (False)."""
        r.cancelled = True
        for key, reqs in list(self._queue.items()):
            if any(q is r for q in reqs):
                self._remove(key, [r])
                self._fail(r, CancelledError(f"request {r.id} cancelled"))
                return True
        return False

    def _fail(self, r: Request, e: BaseException) -> None:
        r.error = e
        r.result = None
        r.t_done = self.now()
        self.completed.append(r)

    def _expire(self) -> None:
        """Fails the queued requests whose deadline has passed."""
        now = self.now()
        for key, reqs in list(self._queue.items()):
            late = [r for r in reqs if now - r.t_submit > r.timeout]
            if late:
                self._remove(key, late)
                for r in late:
                    self._fail(r, TimeoutError(f"request {r.id} still queued after its timeout "
                                               f"of {r.timeout} s"))

    # ------------------------------------------------------------------ planning
    def _request_bytes(self, key, reqs: List[Request]) -> int:
        kind, _, order, precision, dtype, d, Lx, Ly = key[:8]
        itemsize = torch.empty((), dtype=dtype).element_size()
        if kind in SEPARABLE:
            B = sum(r.x.shape[0] for r in reqs)
            Lyy = Lx if Ly < 0 else Ly
            return estimate_bytes(kind, precision, B, B, Lx, Lyy, d, order, itemsize,
                                  self.max_batch)
        if gram_coalesced(reqs[0]):
            # the joined paths of the batch: every request's x and y paths, one workspace
            Bx = sum(r.x.shape[0] for r in reqs)
            By = sum((r.x if r.y is None else r.y).shape[0] for r in reqs)
            Lyy = Lx if Ly < 0 else Ly
            n = estimate_bytes(kind, precision, Bx, By, Lx, Lyy, d, order, itemsize,
                               self.max_batch)
            # the outputs are per request, not Bx x By
            return n - Bx * By * itemsize + sum(r.pairs() for r in reqs) * itemsize
        r = reqs[0]
        By = r.x.shape[0] if r.y is None else r.y.shape[0]
        Lyy = r.x.shape[1] if r.y is None else r.y.shape[1]
        return estimate_bytes(kind, precision, r.x.shape[0], By, r.x.shape[1], Lyy, d, order,
                              itemsize, self.max_batch)

    def _take_batch(self) -> Optional[_Batch]:
        """Next batch that fits the budget, oldest shape class first."""
        for key in list(self._queue):
            reqs = self._queue[key]
            kind = key[0]
            if kind in SEPARABLE:
                take, pairs = [], 0
                for r in reqs:
                    b = r.x.shape[0]
                    if take and pairs + b > self.max_pairs:
                        break
                    take.append(r)
                    pairs += b
            elif gram_coalesced(reqs[0]):
                take, pairs = [], 0
                for r in reqs:
                    if take and (len(take) >= _native.MAX_GRAM_BATCH
                                 or pairs + r.pairs() > self.max_gram_pairs):
                        break
                    take.append(r)
                    pairs += r.pairs()
            else:
                take = reqs[:1]
            nbytes = self._request_bytes(key, take)
            while nbytes > self.hbm_budget and len(take) > 1:
                take = take[: len(take) // 2]
                nbytes = self._request_bytes(key, take)
            if nbytes > self.hbm_budget:
                r = take[0]
                r.error = MemoryError(f"request {r.id} needs ~{nbytes} bytes, over the HBM "
                                      f"budget of {self.hbm_budget}")
                r.t_done = self.now()
                self._remove(key, take)
                self.completed.append(r)
                return self._take_batch()
            if self._inflight and self._inflight_bytes + nbytes > self.hbm_budget:
                continue
            self._remove(key, take)
            b = _Batch(key, take)
            b.nbytes = nbytes
            return b
        return None

    def _remove(self, key, take):
        reqs = self._queue[key]
        ids = {id(r) for r in take}
        rest = [r for r in reqs if id(r) not in ids]
        if rest:
            self._queue[key] = rest
        else:
            del self._queue[key]

    # ------------------------------------------------------------------ execution
    def _launch(self, b: _Batch) -> None:
        if not self.streams:
            self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.nstreams)]
        # the stream with the least estimated work in flight (an idle one first), so that a small
        # request does not queue behind a large one while another stream has little to do; ties
        # go round-robin
        load = {id(st): 0.0 for st in self.streams}
        for f in self._inflight:
            load[id(f.stream)] += f.cost
        order = [(self._next_stream + i) % len(self.streams) for i in range(len(self.streams))]
        pick = min(order, key=lambda i: load[id(self.streams[i])])
        stream = self.streams[pick]
        self._next_stream = (pick + 1) % len(self.streams)
        b.stream = stream
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            try:
                b.outputs = self._run(b)
            except BaseException as e:  # the batch fails as a whole, reported per request
                for r in b.reqs:
                    r.error = e
                b.outputs = [None] * len(b.reqs)
            b.event = torch.cuda.Event()
            b.event.record(stream)
        self._inflight.append(b)
        self._inflight_bytes += b.nbytes
        self.launches += 1

    def _run(self, b: _Batch) -> List[torch.Tensor]:
        kind, _, order, precision, dtype, d, Lx, Ly = b.key[:8]
        r0 = b.reqs[0]
        k = r0.static_kernel
        if gram_coalesced(r0):
            return self._gram_batch(b)
        if kind == "gram":
            return [SigPDE(k, order, precision).gram(r0.x, r0.y, max_batch=self.max_batch)]
        if kind == "robust_gram":
            rk = RobustSigPDE(k, order, r0.normalizer, precision)
            return [rk.gram(r0.x, r0.y, tol=r0.tol, maxit=r0.maxit, max_batch=self.max_batch)]
        sizes = [r.x.shape[0] for r in b.reqs]
        X = torch.cat([pad_paths(r.x, Lx) for r in b.reqs]) if len(b.reqs) > 1 \
            else pad_paths(r0.x, Lx)
        if kind == "normalization":
            rk = RobustSigPDE(k, order, r0.normalizer)
            out = rk.normalization(X, tol=r0.tol, maxit=r0.maxit, max_batch=self.max_batch)
        else:
            if Ly < 0:
                Y = None
            else:
                Y = torch.cat([pad_paths(r.y, Ly) for r in b.reqs]) if len(b.reqs) > 1 \
                    else pad_paths(r0.y, Ly)
            if kind == "pairwise":
                if self.use_graphs and precision != "ref":
                    out = self._pairwise_graph(b.key, k, order, precision, X,
                                               X if Y is None else Y)
                else:
                    out = SigPDE(k, order, precision).pairwise(X, Y, max_batch=self.max_batch)
            else:
                rk = RobustSigPDE(k, order, r0.normalizer, precision)
                out = rk.pairwise(X, Y, tol=r0.tol, maxit=r0.maxit, max_batch=self.max_batch)
        return list(torch.split(out, sizes))

    def _gram_batch(self, b: _Batch) -> List[torch.Tensor]:
        """A coalesced batch of fused Gram / robust-Gram requests of one class: their paths
        joined (padded to the bucket lengths), the robust root-finds of all of them in one
        normalisation call, then one launch of the batched fused Gram."""
        kind, _, order, precision, dtype, d, Lx, Ly = b.key[:8]
        r0 = b.reqs[0]
        k = r0.static_kernel
        symmetric = Ly < 0
        X = torch.cat([pad_paths(r.x, Lx) for r in b.reqs])
        Y = X if symmetric else torch.cat([pad_paths(r.y, Ly) for r in b.reqs])
        rows, x0, y0 = [], 0, 0
        for r in b.reqs:
            bx = r.x.shape[0]
            by = bx if symmetric else r.y.shape[0]
            rows.append((x0, bx, x0 if symmetric else y0, by, int(symmetric)))
            x0 += bx
            y0 += by
        sx = sy = None
        if kind == "robust_gram":
            rk = RobustSigPDE(k, order, r0.normalizer)
            if symmetric:
                sx = sy = rk.normalization(X, tol=r0.tol, maxit=r0.maxit,
                                           max_batch=self.max_batch)
            elif Lx == Ly:
                lam = rk.normalization(torch.cat([X, Y]), tol=r0.tol, maxit=r0.maxit,
                                       max_batch=self.max_batch)
                sx, sy = lam[:X.shape[0]].contiguous(), lam[X.shape[0]:].contiguous()
            else:
                sx = rk.normalization(X, tol=r0.tol, maxit=r0.maxit, max_batch=self.max_batch)
                sy = rk.normalization(Y, tol=r0.tol, maxit=r0.maxit, max_batch=self.max_batch)
        out = torch.empty(sum(q[1] * q[3] for q in rows), device=X.device, dtype=dtype)
        _native.linear_gram_batch(X, Y, rows, order, float(k.fused_linear_scale), out, precision,
                                  sx, sy)
        res, off = [], 0
        for q in rows:
            res.append(out[off:off + q[1] * q[3]].view(q[1], q[3]))
            off += q[1] * q[3]
        return res

    def _pairwise_graph(self, key, k, order, precision, X, Y) -> torch.Tensor:
        """Fused pairwise through a captured HIP graph (one per shape class, padded size and
        stream: the static buffers of one graph are never shared by two streams)."""
        B = X.shape[0]
        Bp = max(4, 1 << (B - 1).bit_length())
        stream = torch.cuda.current_stream(self.device)
        gk = (key, Bp, stream.cuda_stream)
        g = self._graphs.get(gk)
        if g is None:
            g = _GraphEntry()
            g.X = torch.zeros((Bp,) + tuple(X.shape[1:]), device=X.device, dtype=X.dtype)
            g.Y = torch.zeros((Bp,) + tuple(Y.shape[1:]), device=Y.device, dtype=Y.dtype)
            model = SigPDE(k, order, precision)
            # warm up on a side stream (allocator pools, lazy init), then capture
            model.pairwise(g.X, g.Y)
            torch.cuda.synchronize(self.device)
            g.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g.graph):
                g.out = model.pairwise(g.X, g.Y)
            self._graphs[gk] = g
            stream.wait_stream(torch.cuda.current_stream(self.device))
        g.X[:B].copy_(X)
        g.Y[:B].copy_(Y)
        g.graph.replay()
        return g.out[:B].clone()

    def poll(self, block: bool = False) -> int:
        """Retires finished batches (all of them with block=True); returns how many."""
        done = 0
        keep = []
        for b in self._inflight:
            if block:
                b.event.synchronize()
            if block or b.event.query():
                t = self.now()
                cur = torch.cuda.current_stream(self.device)
                for r, o in zip(b.reqs, b.outputs):
                    if r.cancelled and r.error is None:
                        r.error = CancelledError(f"request {r.id} cancelled")
                        o = None
                    if o is not None:
                        # made on the batch's stream: keep its memory for the caller's stream
                        o.record_stream(cur)
                    r.result = o
                    r.t_done = t
                    self.completed.append(r)
                self._inflight_bytes -= b.nbytes
                done += 1
            else:
                keep.append(b)
        self._inflight = keep
        return done

    def step(self) -> int:
        """Launches every batch that fits now; returns how many."""
        self._expire()
        n = 0
        while self._queue:
            b = self._take_batch()
            if b is None:
                break
            self._launch(b)
            n += 1
        return n

    def pending(self) -> int:
        return sum(len(v) for v in self._queue.values()) + len(self._inflight)

    def warmup(self, precisions=("ref", "fp16x2"), dtypes=(torch.float32,)) -> None:
        """Serves one tiny request of every kind in each precision and dtype, then forgets them.
        The runtime loads the library's device code at the first launch of a kernel from each of
        its compiled units (0.1-0.3 s each); warming up at start keeps that, and the allocator's
        first blocks, off the first real requests' latency."""
        g = torch.Generator().manual_seed(0)
        reqs = []
        for dt in dtypes:
            x = (torch.randn(4, 9, 4, generator=g, dtype=torch.float64) * 0.3).cumsum(1)
            x = x.to(device=self.device, dtype=dt)
            for prec in precisions:
                for kind in KINDS:
                    if kind == "normalization" and prec != "ref":
                        continue
                    y = None if kind in ("gram", "robust_gram") else x
                    reqs.append(Request(kind, x, y, dyadic_order=1, precision=prec))
        launches, done = self.launches, len(self.completed)
        self.run(reqs, realtime=False)
        torch.cuda.synchronize(self.device)
        failed = [r for r in reqs if r.error is not None]
        if failed:
            raise RuntimeError(f"warmup request failed: {failed[0].error!r}")
        self.launches = launches
        del self.completed[done:]

    def run(self, requests: List[Request], realtime: bool = True,
            progress: Optional[Callable[[int, int, float], None]] = None,
            progress_every: float = 20.0) -> List[Request]:
        """Serves `requests` (sorted by arrival). realtime: submit each at its arrival time
        (seconds after start); otherwise all at once. progress(done, total, seconds) is called
        about every progress_every seconds while requests remain."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        done0 = len(self.completed)
        next_report = progress_every
        i = 0
        while i < len(reqs) or self.pending():
            if progress is not None and self.now() >= next_report:
                progress(len(self.completed) - done0, len(reqs), self.now())
                next_report = self.now() + progress_every
            now = self.now()
            while i < len(reqs) and (not realtime or reqs[i].arrival <= now):
                self.submit(reqs[i])
                i += 1
            self.step()
            if not self.poll():
                if self._inflight:
                    # wait for the oldest launch unless an arrival is due first
                    nxt = reqs[i].arrival if (realtime and i < len(reqs)) else math.inf
                    if nxt - self.now() > 1e-3:
                        self._inflight[0].event.synchronize()
                elif i < len(reqs):
                    time.sleep(max(0.0, min(reqs[i].arrival - self.now(), 0.01)))
        return reqs


# ---------------------------------------------------------------------- synthetic workload
def synthetic_requests(n: int, seed: int = 7, rate: float = math.inf, device="cuda",
                       dtype=torch.float32, max_batch_size: int = 256,
                       lengths=(16, 512), dims=(4, 16, 64), orders=(0, 1, 2, 3),
                       robust_frac: float = 0.3, gram_frac: float = 0.5,
                       precision: str = "ref") -> List[Request]:
    """Request mix: SigPDE vs RobustSigPDE (robust_frac), pairwise vs Gram (gram_frac), batch
    U{1..max_batch_size}, length U{lengths}, dim from dims, order from orders, increments
    N(0, 1/length); Poisson arrivals at `rate` per second (inf: all at t = 0). Orders are capped
    so that the refined grid side stays within 4096 (the largest of the request mix)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for _ in range(n):
        robust = rng.random() < robust_frac
        gram = rng.random() < gram_frac
        B = int(rng.integers(1, max_batch_size + 1))
        Lx = int(rng.integers(lengths[0], lengths[1] + 1))
        Ly = int(rng.integers(lengths[0], lengths[1] + 1))
        d = int(rng.choice(dims))
        order = int(rng.choice(orders))
        while ((max(Lx, Ly) - 1) << order) > 4096:
            order -= 1
        if robust:
            Ly = Lx  # robust pairwise normalises both sides with their own lengths
        x = _brownian(rng, B, Lx, d, dtype, device)
        if gram:
            By = int(rng.integers(1, max_batch_size + 1))
            y = _brownian(rng, By, Ly, d, dtype, device) if rng.random() < 0.5 else None
            kind = "robust_gram" if robust else "gram"
        else:
            y = _brownian(rng, B, Ly, d, dtype, device)
            kind = "robust_pairwise" if robust else "pairwise"
        if rate < math.inf:
            t += rng.exponential(1.0 / rate)
        out.append(Request(kind, x, y, dyadic_order=order, arrival=t, precision=precision))
    return out


def _brownian(rng, B, L, d, dtype, device):
    inc = rng.normal(0.0, math.sqrt(1.0 / L), size=(B, L - 1, d))
    x = np.concatenate([np.zeros((B, 1, d)), np.cumsum(inc, axis=1)], axis=1)
    return torch.from_numpy(x).to(device=device, dtype=dtype)


def request_cells(r: Request) -> int:
    R = 1 << r.dyadic_order
    m = r.x.shape[1] - 1
    if r.kind in ("gram", "robust_gram"):
        y = r.x if r.y is None else r.y
        return r.x.shape[0] * y.shape[0] * R * R * m * (y.shape[1] - 1)
    y = r.x if r.y is None else r.y
    return r.x.shape[0] * R * R * m * (y.shape[1] - 1)


def request_cost(r: Request) -> float:
    """Estimated device work of a request, in wavefront cells: its solves, plus for the robust
    kinds the root-find's evaluations of every path with itself (maxit of them for float32
    inputs, whose values never reach a tolerance below float resolution; about a dozen for
    float64)."""
    cost = float(request_cells(r))
    if r.kind in ("normalization", "robust_pairwise", "robust_gram"):
        R = 1 << r.dyadic_order
        iters = r.maxit if r.x.dtype == torch.float32 else min(r.maxit, 12)
        sides = [r.x] if r.y is None or r.kind == "normalization" else [r.x, r.y]
        for p in sides:
            m = (p.shape[1] - 1) * R
            cost += (iters + 1) * p.shape[0] * m * m
    return cost


def summarize(reqs: List[Request], wall: float) -> dict:
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.latency for r in ok]) * 1e3
    return {
        "requests": len(reqs),
        "failed": len(reqs) - len(ok),
        "wall_s": wall,
        "requests_per_s": len(ok) / wall,
        "pairs_per_s": sum(r.pairs() for r in ok) / wall,
        "cells_per_s": sum(request_cells(r) for r in ok) / wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50)) if len(lat) else None,
                       "p99": float(np.percentile(lat, 99)) if len(lat) else None,
                       "mean": float(lat.mean()) if len(lat) else None},
    }


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--requests", type=int, default=300)
    p.add_argument("--seed", type=int, default=7)
    p.add_argument("--rate", type=float, default=math.inf,
                   help="Poisson arrivals per second (default: all queued at t=0, saturation)")
    p.add_argument("--budget-gb", type=float, default=64.0)
    p.add_argument("--streams", type=int, default=4)
    p.add_argument("--precision", default="ref",
                   choices=["ref", "fp16x2", "bf16x3", "fp32", "fp64", "fp16x2f"])
    p.add_argument("--graphs", action="store_true")
    p.add_argument("--max-length", type=int, default=512)
    args = p.parse_args(argv)
    dev = torch.device("cuda:0")
    reqs = synthetic_requests(args.requests, args.seed, args.rate, dev,
                              lengths=(16, args.max_length), precision=args.precision)
    torch.cuda.synchronize()
    srv = Server(dev, hbm_budget=int(args.budget_gb * (1 << 30)), streams=args.streams,
                 use_graphs=args.graphs)
    srv.warmup(precisions=tuple(sorted({"ref", args.precision})))
    t0 = time.perf_counter()
    srv.run(reqs, realtime=args.rate < math.inf)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    line = {"metric": "signature-kernel request serving", "config": {
        "requests": args.requests, "seed": args.seed, "rate": args.rate,
        "budget_gb": args.budget_gb, "streams": args.streams, "precision": args.precision,
        "graphs": args.graphs, "lengths": [16, args.max_length]}, "launches": srv.launches}
    line.update(summarize(reqs, wall))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
