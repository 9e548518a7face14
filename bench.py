"""Benchmark: signature-kernel Gram pairs/sec on one MI355X.

Workload (default): Gram of X, Y each 1024 Brownian paths, length 256, dim 64, dyadic order 0,
linear static kernel (increments N(0, 1/256), seeds 5/6, float32 inputs resident in HBM).
One step = one full 1024 x 1024 Gram. value = pairs/sec = 1024^2 / step time of the
mixed-precision path (precision="fp16x2f": fp16 hi/lo increment Gram on MFMA, float32 wavefront;
--mode selects another). Timed beside it on the same inputs:
  mixed_fp16x2           precision="fp16x2": the same increments with a double wavefront (the
                         most accurate mixed mode)
  plain_fp32             precision="fp32": the same fused design in plain FP32 (fp32 increments on
                         fp32 MFMA, the reference's float32 wavefront typing)
  reference_scheme_fp32  precision="ref": the reference's algorithm in float32 (torch materialises
                         each increment tile from the point Gram, then the generic wavefront kernel)

Besides the JSON contract fields the line carries:
  rel_err_fp64   ||K - K64||_F / ||K64||_F on a sampled block, K64 = the float64 oracle on the same
                 float32 inputs upcast (computed in the CPU-baseline leg, which runs the oracle);
                 the same figure is given for the two comparison paths; per entry:
                 rel_err_fp64_max / _p99 (|K - K64| / |K64|), the |K64| of the worst entry and
                 the largest |K - K64| against the block's RMS entry (entries that cancel to near
                 zero carry large relative but small absolute errors)
  rel_err_fp64_full  the same norm over all entries, K64 = precision="ref" on the float64 inputs
                 (the reference scheme in float64 on the GPU)
  latency_ms     p50 / p99 of the per-step times
  roofline       dominant kernel: achieved vs peak, measured with HIP events on its stream
  cpu_baseline   the oracle (C restatement, float32 reference semantics) timed on this host's
                 cores on a sampled block, in pairs/sec (and on one thread: value_1_thread)
  workloads      the smaller workloads of the API, each with its float64-oracle error: float64
                 pairwise (8 pairs, length 32, dim 4, order 1: per-call latency p50 / p99), float32
                 Gram 128 x 128 (length 100, dim 8, order 2: every precision), float64
                 RobustSigPDE (256 paths, length 64, dim 16, order 1: normalization and Gram
                 times, root-find iteration histogram)
  serving        the continuous-batching request loop (sigpde_amd.serving) on a bounded sample of
                 the request mix (70% SigPDE / 30% RobustSigPDE, pairwise or Gram, batch 1-256,
                 length 16-512, dim 4/16/64, dyadic order 0-3, float32 Brownian paths, 64 GB HBM
                 budget; 300 requests by default): saturated throughput (all queued at once),
                 then the same requests with Poisson arrivals at 70% of it: p50 / p99 request
                 latency beside pairs/sec (sweep: at 50% and 90%)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import sigpde_amd.torch as sig  # noqa: E402
from sigpde_amd import _native  # noqa: E402
from sigpde_amd.synthetic import brownian_paths  # noqa: E402

MI355X_HBM_GBS = 8000.0
MI355X_FP64_VALU_TFLOPS = 78.6
MI355X_FP32_VALU_TFLOPS = 157.3
MI355X_FP16_MFMA_TFLOPS = 2500.0
MI355X_FP32_MFMA_TFLOPS = 157.3
MI355X_FP64_MFMA_TFLOPS = 78.6
FLOPS_PER_CELL = 10  # pde_step as written
# HBM bytes per launch of the fused fp16x2 and fp16x2f Grams at the default workload (1024 x 1024
# paths, length 256, dim 64, order 0), from rocprofv3 FETCH_SIZE + WRITE_SIZE passes scaled by a
# 256 MiB copy (profiles/gram_fp16x2*_1024x1024_L256_d64_o0_traffic.md); other workloads report
# null.
MEASURED_TRAFFIC = {("fp16x2", 1024, 256, 64, 0): 7.718e9,
                    ("fp16x2f", 1024, 256, 64, 0): 5.996e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--mode", default="fp16x2f",
                   choices=["fp16x2", "bf16x3", "fp32", "ref", "fp64", "fp16x2f"])
    p.add_argument("--ref-steps", type=int, default=1,
                   help="timed steps of the reference-scheme FP32 path")
    p.add_argument("--paths", type=int, default=1024)
    p.add_argument("--length", type=int, default=256)
    p.add_argument("--dim", type=int, default=64)
    p.add_argument("--order", type=int, default=0)
    p.add_argument("--max-batch", type=int, default=128 * 128)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline / accuracy leg")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--sample", type=int, default=32, help="side of the sampled accuracy block")
    p.add_argument("--no-compare", action="store_true",
                   help="time only --mode (skip the plain-FP32 and reference-scheme legs)")
    p.add_argument("--no-full-check", action="store_true",
                   help="skip the all-entries comparison with the float64 GPU path")
    p.add_argument("--no-workloads", action="store_true",
                   help="skip the smaller API workloads (pairwise, Gram, robust) beside the line")
    p.add_argument("--serving", type=int, default=300, metavar="N",
                   help="requests of the serving leg (0: skip it)")
    p.add_argument("--serving-sweep", type=int, default=300, metavar="N",
                   help="requests of the 50%% / 90%% load sweep (the first N of the same set; 0: "
                        "no sweep)")
    p.add_argument("--serving-rate", type=float, default=None,
                   help="Poisson arrival rate of the serving leg, requests per second (default: "
                        "70%% of the saturated rate measured on the same requests)")
    p.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                   help="only run a plain device-to-device copy of BYTES bytes 5 times (the "
                        "known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def gram_step(args, X, Y, mode=None):
    k = sig.SigPDE(sig.kernels.LinearKernel(), args.order, precision=mode or args.mode)
    return k.gram(X, Y, max_batch=args.max_batch)


def timed(args, X, Y, mode, steps, warmup):
    for _ in range(warmup):
        K = gram_step(args, X, Y, mode)
    torch.cuda.synchronize()
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        K = gram_step(args, X, Y, mode)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    total = time.perf_counter() - t0
    return K, total / steps * 1e3, step_ms


def _point_ne(d):
    """K elements per lane of the fused static-kernel Gram (float32): dpad = 4 NE >= d."""
    ne = 2
    while 4 * ne < d:
        ne *= 2
    return ne


def fused_roofline(args, X, Y, ms):
    """The fused kernel is one launch per step (plus a small split prepass); its time is the step
    time measured with HIP events on torch's current stream, the stream it is launched on."""
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = 2
    gram_step(args, X, Y, args.mode)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        gram_step(args, X, Y, args.mode)
    e1.record(st)
    torch.cuda.synchronize()
    dur_s = e0.elapsed_time(e1) / 1e3 / reps
    m = (X.shape[1] - 1) << args.order
    n = (Y.shape[1] - 1) << args.order
    pairs = X.shape[0] * Y.shape[0]
    cells = pairs * m * n
    flops_wave = cells * FLOPS_PER_CELL
    # MFMA work beside the wavefront: the increment Gram, or for "ref" the point Gram
    flops_inc = pairs * 2 * args.dim * (X.shape[1] - (args.mode != "ref")) * \
        (Y.shape[1] - (args.mode != "ref"))
    name = {"fp16x2": "k_linear_gram_fused<float, 2, 1, 1, false> (fp16x2 MFMA increments + "
                      "double wavefront, one launch)",
            "bf16x3": "k_linear_gram_fused<float, 2, 3, 1, false> (bf16x3 MFMA increments + "
                      "double wavefront, one launch)",
            "fp32": "k_linear_gram_fused<float, 2, 2, 1, false> (fp32 MFMA increments + "
                    "float32-typed wavefront, one launch)",
            "fp64": "k_linear_gram_fused<float, 2, 4, 1, false> (fp64 MFMA increments + double "
                    "wavefront, one launch)",
            "fp16x2f": "k_linear_gram_f16f<float, 2, 1, 4> (fp16x2 MFMA increments on producer "
                       "waves + float32 wavefront on consumer waves, four pairs of waves per "
                       "workgroup, one launch)",
            "ref": f"k_point_gram<float, 0, {_point_ne(args.dim)}, {min(1 << args.order, 8)}> "
                   "(fp32 MFMA point Gram, second difference and float32-typed wavefront, one "
                   "launch)"}[args.mode]
    traffic = MEASURED_TRAFFIC.get((args.mode, args.paths, args.length, args.dim, args.order))
    if args.mode == "ref":  # padded float points of X and Y, and the output
        alg_bytes = (X.shape[0] * X.shape[1] + Y.shape[0] * Y.shape[1]) * \
            4 * _point_ne(args.dim) * 4 + pairs * X.element_size()
    else:  # the split increments of X and Y, and the output
        alg_bytes = (X.shape[0] * (X.shape[1] - 1) + Y.shape[0] * (Y.shape[1] - 1)) * \
            ((args.dim + 31) // 32 * 32) * {"fp16x2": 4, "bf16x3": 6, "fp32": 4, "fp64": 8,
                                             "fp16x2f": 4}[args.mode] + \
            pairs * X.element_size()
    # the wavefront's VALU peak: FP32 cells (fp16x2f, fp32, ref) or FP64 cells
    f32_wave = args.mode in ("fp16x2f", "fp32", "ref")
    valu_peak = MI355X_FP32_VALU_TFLOPS if f32_wave else MI355X_FP64_VALU_TFLOPS
    mfma_peak = {"fp32": MI355X_FP32_MFMA_TFLOPS, "ref": MI355X_FP32_MFMA_TFLOPS,
                 "fp64": MI355X_FP64_MFMA_TFLOPS}.get(args.mode, MI355X_FP16_MFMA_TFLOPS)
    # MFMA work as issued: the split modes multiply 3 (fp16x2) or 6 (bf16x3) products over the
    # dimension padded to the MFMA K
    dpad = (args.dim + 31) // 32 * 32
    products = {"fp16x2": 3, "fp16x2f": 3, "bf16x3": 6}.get(args.mode, 1)
    flops_mfma_issued = flops_inc / args.dim * dpad * products
    # MFMA and VALU instructions share each SIMD: issued by the same wave they take the sum of
    # their times; issued by different waves they overlap in part (sigpde_issue_rates measures
    # both). The kernel cannot finish faster than the slower of the two at its peak
    # (overlap_bound_ms), and with no overlap not faster than their sum (serialized_bound_ms).
    t_mfma = flops_mfma_issued / (mfma_peak * 1e12)
    t_valu = flops_wave / (valu_peak * 1e12)
    t_serial = t_mfma + t_valu
    t_overlap = max(t_mfma, t_valu)
    return {
        "kernel": name,
        "bound": "valu",
        "achieved": flops_wave / dur_s / 1e12,
        "peak": valu_peak,
        "unit": "TFLOP/s",
        "frac": flops_wave / dur_s / 1e12 / valu_peak,
        "traffic": traffic,
        "algorithmic_bytes": alg_bytes,
        "hbm_gbs": None if traffic is None else traffic / dur_s / 1e9,
        "launch_ms": dur_s * 1e3,
        "cells_per_s": cells / dur_s,
        "increment_gram_tflops": flops_inc / dur_s / 1e12,
        "increment_gram_mfma_frac": flops_inc / dur_s / 1e12 / mfma_peak,
        "mfma_issued_tflops": flops_mfma_issued / dur_s / 1e12,
        "mfma_bound_ms": t_mfma * 1e3,
        "valu_bound_ms": t_valu * 1e3,
        "serialized_bound_ms": t_serial * 1e3,
        "frac_serialized": t_serial / dur_s,
        "overlap_bound_ms": t_overlap * 1e3,
        "frac_overlap": t_overlap / dur_s,
        "note": "bound: the wavefront's VALU (10 flops per refined cell, pde_step as written; "
                "float32 cells priced at the MI355X FP32 vector peak 157.3 TFLOP/s, which plain "
                "(unpacked) FP32 FMAs reach: issue_rates measures ~2 ticks per wave64 FMA; double "
                "cells at the FP64 vector peak 78.6). The increment Gram runs on MFMA on the same "
                "SIMDs: MFMA and VALU instructions of one wave take the sum of their times "
                "(issue_rates: mfma_plus_valu_over_sum ~1), those of different waves overlap in "
                "part (split_1x1_over_sum, split_2x2_over_sum against the full-overlap value "
                "split_full_overlap_over_sum), so the kernel's floor lies between "
                "overlap_bound_ms = max(issued MFMA flops at the MFMA peak, wavefront flops at the "
                "VALU peak) and serialized_bound_ms = their sum; frac_overlap and "
                "frac_serialized are the launch's fractions of them; traffic = measured HBM bytes "
                "per launch, algorithmic_bytes = operands + output",
    }


def cpu_leg(args, X, Y, K, others):
    """Runs the oracle: (1) the float64 reference on a sampled block for the accuracy figures,
    (2) the float32-semantics restatement timed on this host's cores as the CPU baseline."""
    import oracle

    s = args.sample
    Xs = X[:s].double().cpu().numpy()
    Ys = Y[:s].double().cpu().numpy()
    K64 = oracle.linear_gram(Xs, Ys, args.order)
    Kg = K[:s, :s].double().cpu().numpy()
    rel = float(np.linalg.norm(Kg - K64) / np.linalg.norm(K64))
    per = np.abs(Kg - K64) / np.abs(K64)
    rel_others = {}
    for name, Ko in others.items():
        Kr = Ko[:s, :s].double().cpu().numpy()
        rel_others[name] = float(np.linalg.norm(Kr - K64) / np.linalg.norm(K64))

    cores, host = _cpu_share()
    oracle.set_num_threads(cores)
    X32 = X.cpu().numpy()
    Y32 = Y.cpu().numpy()
    side = 8
    t0 = time.perf_counter()
    oracle.linear_gram(X32[:side], Y32[:side], args.order)
    dt = time.perf_counter() - t0
    # grow the block to about cpu_seconds of work
    pairs_target = side * side * args.cpu_seconds / max(dt, 1e-6)
    side2 = int(max(side, min(X.shape[0], np.sqrt(pairs_target))))
    t0 = time.perf_counter()
    oracle.linear_gram(X32[:side2], Y32[:side2], args.order)
    dt2 = time.perf_counter() - t0
    # one thread, on a block of about a tenth of the time
    oracle.set_num_threads(1)
    side1 = int(max(4, min(X.shape[0], np.sqrt(side2 * side2 * 0.1 / cores))))
    t0 = time.perf_counter()
    oracle.linear_gram(X32[:side1], Y32[:side1], args.order)
    dt1 = time.perf_counter() - t0
    oracle.set_num_threads(cores)
    return {
        "rel_err_fp64": rel,
        "rel_err_fp64_max": float(per.max()),
        "rel_err_fp64_p99": float(np.percentile(per, 99)),
        # the entry of largest relative error: |K64| there against the block's RMS entry
        "rel_err_fp64_max_at_abs_k": float(np.abs(K64).flat[int(np.argmax(per))]),
        "k64_rms": float(np.sqrt(np.mean(K64 * K64))),
        "abs_err_fp64_max_over_rms": float(np.abs(Kg - K64).max() / np.sqrt(np.mean(K64 * K64))),
        "rel_err_sample": f"{s}x{s} pairs",
    }, rel_others, {
        "value": side2 * side2 / dt2,
        "unit": "pairs/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{side2}x{side2} pairs of the same workload, float32 reference semantics "
                  f"(C restatement, OpenMP over pairs), {dt2:.1f} s",
        "value_1_thread": side1 * side1 / dt1,
        "sample_1_thread": f"{side1}x{side1} pairs on one thread, {dt1:.1f} s",
        **host,
    }


def _cpu_share():
    """Threads for the CPU baseline and a description of the host: the GPU box gives one GPU's
    job a share of 16 host threads (OMP_NUM_THREADS is set to it there), whatever nproc reports,
    so the baseline uses min(affinity, OMP_NUM_THREADS or 16) threads and records nproc, the
    affinity count and the CPU model beside it."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(omp) if omp.isdigit() and int(omp) > 0 else 16
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = max(1, min(affinity, cap))
    return cores, {"nproc": nproc, "affinity_cpus": affinity, "cpu_model": model,
                   "threads_cap": f"{cap} (OMP_NUM_THREADS)" if omp else f"{cap} (default)"}


def _rel(K, K64):
    K = np.asarray(K, dtype=np.float64)
    return float(np.linalg.norm(K - K64) / np.linalg.norm(K64))


def _wall_ms(fn, reps):
    """Per-call wall times (ms), each call followed by a device synchronise (the reference returns
    synchronised)."""
    out = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        out.append((time.perf_counter() - t0) * 1e3)
    return out


def _event_ms(fn, reps):
    """Mean device time (ms) of reps back-to-back calls, CUDA events around the whole run."""
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        out = fn()
    e1.record()
    torch.cuda.synchronize()
    return out, e0.elapsed_time(e1) / reps


def workloads_leg(args, dev):
    """The three smaller workloads of the reference's API, each through the public API and beside
    its float64-oracle error (the oracle runs on the same inputs, upcast for float32 ones):
      pairwise  SigPDE(LinearKernel(), 1).pairwise(x, y), float64, 8 pairs, length 32, dim 4
                (latency-bound: per-call wall time p50 / p99, synchronised on return)
      gram      SigPDE(LinearKernel(), 2).gram(X, Y), float32, 128 x 128, length 100, dim 8
                (reference scheme and the fused paths; full 128 x 128 float64 oracle); the same
                with RBFKernel / L1RBFKernel (one-pass increments vs the torch route)
      robust    RobustSigPDE(LinearKernel(), 1): .normalization(X) and .gram(X), float64, 256 paths,
                length 64, dim 16, increments N(0, 9/64), tol 1e-10 (root-find iteration histogram)
    """
    import oracle

    oracle.set_num_threads(_cpu_share()[0])
    lin = sig.kernels.LinearKernel()
    out = {}

    # pairwise, float64, order 1
    x = brownian_paths(8, 32, 4, 1.0 / 32, 0, torch.float64, dev)
    y = brownian_paths(8, 32, 4, 1.0 / 32, 1, torch.float64, dev)
    k1 = sig.SigPDE(lin, 1)
    for _ in range(5):
        k1.pairwise(x, y)
    torch.cuda.synchronize()
    lat = _wall_ms(lambda: k1.pairwise(x, y), 200)
    K = k1.pairwise(x, y).cpu().numpy()
    K64 = oracle.sig_pairwise(x.cpu().numpy(), y.cpu().numpy(), 1)
    out["pairwise_8_L32_d4_o1_f64"] = {
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "calls": len(lat), "rel_err_fp64": _rel(K, K64),
        "max_abs_err_fp64": float(np.abs(K - K64).max())}

    # Gram, float32, order 2
    X = brownian_paths(128, 100, 8, 1.0 / 100, 2, torch.float32, dev)
    Y = brownian_paths(128, 100, 8, 1.0 / 100, 3, torch.float32, dev)
    K64 = oracle.linear_gram(X.double().cpu().numpy(), Y.double().cpu().numpy(), 2)
    K32ref = oracle.linear_gram(X.cpu().numpy(), Y.cpu().numpy(), 2)  # float32 reference semantics
    g = {}
    for mode in ("ref", "fp32", "fp16x2", "bf16x3", "fp64", "fp16x2f"):
        k2 = sig.SigPDE(lin, 2, precision=mode)
        k2.gram(X, Y)
        torch.cuda.synchronize()
        Kg, ms = _event_ms(lambda: k2.gram(X, Y), 5)
        Kg = Kg.double().cpu().numpy()
        g[mode] = {"ms": ms, "pairs_per_s": 128 * 128 / (ms / 1e3), "rel_err_fp64": _rel(Kg, K64)}
        if mode == "ref":
            g[mode]["rel_err_vs_fp32_reference_semantics"] = _rel(Kg, K32ref.astype(np.float64))
            # the same reference scheme through tiles of one-pass increments (the route taken
            # past the fused static-kernel Gram's limits)
            tile = int(np.ceil(np.sqrt(1000)))
            k2._gram(X, Y, None, None, tile)
            torch.cuda.synchronize()
            _, ms_t = _event_ms(lambda: k2._gram(X, Y, None, None, tile), 3)
            g[mode]["route"] = "fused static-kernel Gram (sigpde_static_gram)"
            g[mode]["increment_tiles_ms"] = ms_t
    out["gram_128x128_L100_d8_o2_f32"] = g

    # the same Gram with the non-linear static kernels: the fused static-kernel Gram, tiles of
    # one-pass increments, and the torch route
    class _TorchRoute:
        """Hides the static kernel's one-pass increment hooks (the reference's torch route)."""
        def __init__(self, k):
            self.k = k

        def pairwise(self, X, Y):
            return self.k.pairwise(X, Y)

        def gram(self, X, Y):
            return self.k.gram(X, Y)

    Xs64, Ys64 = X[:16].double().cpu().numpy(), Y[:16].double().cpu().numpy()
    tile = int(np.ceil(np.sqrt(1000)))
    for kind, K in (("rbf", sig.kernels.RBFKernel(1.0)), ("l1rbf", sig.kernels.L1RBFKernel(1.0))):
        k1 = sig.SigPDE(K, 2)
        k2 = sig.SigPDE(_TorchRoute(K), 2)
        k1.gram(X, Y)
        k1._gram(X, Y, None, None, tile)
        k2.gram(X, Y)
        torch.cuda.synchronize()
        Kg, ms1 = _event_ms(lambda: k1.gram(X, Y), 3)
        _, ms_t = _event_ms(lambda: k1._gram(X, Y, None, None, tile), 3)
        _, ms2 = _event_ms(lambda: k2.gram(X, Y), 3)
        ref = oracle.sig_gram(Xs64, Ys64, 2, kind, 1.0)
        out[f"{kind}_gram_128x128_L100_d8_o2_f32"] = {
            "fused_ms": ms1, "pairs_per_s": 128 * 128 / (ms1 / 1e3),
            "route": "fused static-kernel Gram (sigpde_static_gram)",
            "one_pass_increment_tiles_ms": ms_t, "torch_route_ms": ms2,
            "rel_err_fp64_16x16": _rel(Kg[:16, :16].double().cpu().numpy(), ref)}

    # robust, float64, order 1
    Xr = brownian_paths(256, 64, 16, 9.0 / 64, 4, torch.float64, dev)
    k3 = sig.RobustSigPDE(lin, 1)
    k3.normalization(Xr, tol=1e-10)
    torch.cuda.synchronize()
    lam, ms_n = _event_ms(lambda: k3.normalization(Xr, tol=1e-10), 3)
    Kr, ms_g = _event_ms(lambda: k3.gram(Xr, tol=1e-10), 3)
    iters = torch.zeros(256, dtype=torch.int32, device=dev)
    k3._normalization(Xr, k3.normalizer, 1e-10, 100, 1000, iters=iters)
    it = iters.cpu().numpy()
    Xn = Xr.cpu().numpy()
    lam64 = oracle.robust_normalization(Xn, 1, tol=1e-10, maxit=100)
    Kr64 = oracle.linear_gram(Xn, Xn, 1, sx=lam64, sy=lam64)
    # the final scaled Gram solves on the fused fp16x2 path (root-find still float64)
    k3m = sig.RobustSigPDE(lin, 1, precision="fp16x2")
    k3m.gram(Xr, tol=1e-10)
    torch.cuda.synchronize()
    Krm, ms_gm = _event_ms(lambda: k3m.gram(Xr, tol=1e-10), 3)
    k3d = sig.RobustSigPDE(lin, 1, precision="fp64")
    k3d.gram(Xr, tol=1e-10)
    torch.cuda.synchronize()
    Krd, ms_gd = _event_ms(lambda: k3d.gram(Xr, tol=1e-10), 3)
    lam = lam.cpu().numpy()
    hist = np.bincount(it, minlength=int(it.max()) + 1)
    out["robust_256_L64_d16_o1_f64"] = {
        "normalization_ms": ms_n, "gram_ms": ms_g, "gram_pair_solves": 256 * 257 // 2,
        "gram_pairs_per_s": 256 * 256 / (ms_g / 1e3),
        "lambda_max_abs_err_fp64": float(np.abs(lam - lam64).max()),
        "gram_rel_err_fp64": _rel(Kr.cpu().numpy(), Kr64),
        "gram_fp16x2_ms": ms_gm, "gram_fp16x2_rel_err_fp64": _rel(Krm.cpu().numpy(), Kr64),
        "gram_fp64_ms": ms_gd, "gram_fp64_rel_err_fp64": _rel(Krd.cpu().numpy(), Kr64),
        "active_paths": int((lam64 < 1).sum()),
        "iterations_histogram": {str(i): int(c) for i, c in enumerate(hist) if c}}
    return out


def serving_leg(args, dev):
    """Serves a bounded sample of the request mix; every SigPDE and robust final solve runs in
    precision fp16x2 (the robust root-find in the reference's arithmetic on the float32
    increments). First the whole sample queued at once (saturated throughput), then the same
    requests with Poisson arrivals at 70% of that rate (or at --serving-rate): p50 / p99 latency;
    `sweep` repeats the Poisson run at 50% and 90% of the saturated rate."""
    from sigpde_amd import serving as S

    S.Server(dev, streams=4).warmup(precisions=("ref", "fp16x2"))
    torch.cuda.synchronize()

    # one request set (shapes and paths from seed 7) for every run; each rate draws its own
    # Poisson arrival times (seed 8)
    print(f"serving leg: generating {args.serving} requests", file=sys.stderr, flush=True)
    base = S.synthetic_requests(args.serving, seed=7, device=dev, precision="fp16x2")
    torch.cuda.synchronize()
    print("serving leg: requests generated", file=sys.stderr, flush=True)

    def report(done, total, t):
        print(f"serving leg: {done}/{total} requests served after {t:.0f} s", file=sys.stderr,
              flush=True)

    def serve(rate, n):
        sub = base[:n]
        if np.isfinite(rate):
            arrivals = np.cumsum(np.random.default_rng(8).exponential(1.0 / rate, len(sub)))
        else:
            arrivals = np.zeros(len(sub))
        reqs = [S.Request(r.kind, r.x, r.y, r.static_kernel, r.dyadic_order, r.precision,
                          r.normalizer, r.tol, r.maxit, arrival=float(a))
                for r, a in zip(sub, arrivals)]
        srv = S.Server(dev, hbm_budget=64 << 30, streams=4)
        t0 = time.perf_counter()
        srv.run(reqs, realtime=True, progress=report)
        torch.cuda.synchronize()
        out = S.summarize(reqs, time.perf_counter() - t0)
        print(f"serving leg: {len(reqs)} requests at rate {rate:.3g}/s served in "
              f"{out['wall_s']:.1f} s ({srv.launches} launches)", file=sys.stderr, flush=True)
        return out, srv.launches

    sat, sat_launches = serve(float("inf"), len(base))
    rate = args.serving_rate if args.serving_rate else 0.7 * sat["requests_per_s"]
    out, launches = serve(rate, len(base))
    sweep = []
    nsw = min(len(base), args.serving_sweep)  # 0: no sweep
    if nsw <= 0:
        sat_sw = None
    elif nsw == len(base):
        sat_sw = sat
    else:
        sat_sw, _ = serve(float("inf"), nsw)
    for frac in ((0.5, 0.9) if sat_sw else ()):
        r = frac * sat_sw["requests_per_s"]
        o, _ = serve(r, nsw)
        sweep.append({"load": frac, "requests": nsw, "rate_per_s": r,
                      "requests_per_s": o["requests_per_s"], "pairs_per_s": o["pairs_per_s"],
                      "latency_ms": o["latency_ms"]})
    out.update({"rate_per_s": rate, "launches": launches, "sweep": sweep,
                "saturated": {"requests_per_s": sat["requests_per_s"],
                              "pairs_per_s": sat["pairs_per_s"], "wall_s": sat["wall_s"],
                              "latency_ms": sat["latency_ms"], "launches": sat_launches},
                "mix": "70% SigPDE / 30% RobustSigPDE, pairwise or Gram 50/50, batch U{1..256}, "
                       "length U{16..512}, dim {4,16,64}, order U{0..3} (capped at a refined side "
                       "of 4096), float32, seed 7, precision fp16x2, 64 GB budget"})
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)
    dev = torch.device("cuda:0")
    if args.copy_probe:
        src = torch.ones(args.copy_probe // 4, device=dev, dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4}), flush=True)
        return
    X = brownian_paths(args.paths, args.length, args.dim, 1.0 / args.length, 5, torch.float32, dev)
    Y = brownian_paths(args.paths, args.length, args.dim, 1.0 / args.length, 6, torch.float32, dev)
    torch.cuda.synchronize()

    K, ms, step_ms = timed(args, X, Y, args.mode, args.steps, args.warmup)
    pairs = X.shape[0] * Y.shape[0]
    # comparison paths on the same workload (each leg only when it differs from --mode; none
    # with --no-compare)
    compare = not args.no_compare
    other16 = "fp16x2f" if args.mode != "fp16x2f" else "fp16x2"
    f32_wave_ok = args.order <= 2 and ((args.length - 1) << args.order) <= 512
    legs = [  # (line key, precision, timed steps, runs at this shape, path description)
        ("plain_fp32", "fp32", args.steps, True,
         "precision='fp32': fused, fp32 MFMA increments, float32 wavefront"),
        ("reference_scheme_fp32", "ref", args.ref_steps, True,
         "precision='ref': the reference scheme in float32, fused (point Gram on fp32 MFMA, "
         "4-term second difference and the float32-typed wavefront in one launch)"),
        ("mixed_bf16x3", "bf16x3", args.steps, True,
         "precision='bf16x3': three-part bf16 increments, six bf16 MFMA products, double "
         "wavefront"),
        ("fused_fp64", "fp64", 1, args.dim <= 64,
         "precision='fp64': double increments on fp64 MFMA, double wavefront (the reference "
         "scheme in float64, fused)"),
        ("mixed_" + other16, other16, args.steps, other16 == "fp16x2" or f32_wave_ok,
         {"fp16x2": "precision='fp16x2': fp16 hi/lo MFMA increments, double wavefront",
          "fp16x2f": "precision='fp16x2f': fp16 hi/lo MFMA increments, float32 wavefront"}[
             other16]),
    ]
    others = {}  # line key -> (K, ms, path)
    for key, prec, steps, ok, path in legs:
        if compare and ok and prec != args.mode:
            Ko, mso, _ = timed(args, X, Y, prec, steps, 1)
            others[key] = (Ko, mso, path)

    roof = fused_roofline(args, X, Y, ms)
    roof["issue_rates"] = _native.issue_rates()
    line = {
        "metric": "Gram pairs/sec vs plain FP32 at \u22641e-5 rel err (FP64 oracle); p50/p99 "
                  "latency",
        "value": pairs / (ms / 1e3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp16x2": "fp16x2+f64", "bf16x3": "bf16x3+f64", "fp32": "f32",
                  "ref": "f32", "fp64": "f64", "fp16x2f": "fp16x2+f32"}[args.mode],
        "data": "synthetic",
        "config": {"workload": f"SigPDE(LinearKernel(), {args.order}).gram(X, Y), X,Y "
                               f"{args.paths} Brownian paths, length {args.length}, dim "
                               f"{args.dim}, float32 inputs", "mode": args.mode,
                   "paths_x": args.paths, "paths_y": args.paths, "length": args.length,
                   "dim": args.dim, "dyadic_order": args.order},
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                       "p99": float(np.percentile(step_ms, 99))},
        "roofline": roof,
    }
    for key, (Ko, mso, path) in others.items():
        line[key] = {"value": pairs / (mso / 1e3), "unit": "pairs/s", "ms_per_step": mso,
                     "path": path}
    if "plain_fp32" in others:
        line["speedup_vs_plain_fp32"] = others["plain_fp32"][1] / ms
    if not args.no_cpu:
        acc, rel_others, cpu = cpu_leg(args, X, Y, K, {key: o[0] for key, o in others.items()
                                                        if key != "fused_fp64"})
        line.update(acc)
        for name, rel in rel_others.items():
            line[name]["rel_err_fp64"] = rel
        line["cpu_baseline"] = cpu
    if not args.no_full_check:
        # every entry of the Gram against the float64 GPU path (the reference scheme in float64 on
        # the same float32 inputs upcast; pinned to the CPU oracle by the GPU parity tests)
        t0 = time.perf_counter()
        K64 = sig.SigPDE(sig.kernels.LinearKernel(), args.order).gram(
            X.double(), Y.double(), max_batch=args.max_batch)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        n64 = torch.linalg.norm(K64)

        def rel_full(Kc):
            return float(torch.linalg.norm(Kc.double() - K64) / n64)

        line["rel_err_fp64_full"] = rel_full(K)
        line["rel_err_fp64_full_reference"] = (
            f"all {pairs} entries vs precision='ref' on float64 inputs (GPU, {t64:.1f} s)")
        for name, (Ko, _, _) in others.items():
            line[name]["rel_err_fp64_full"] = rel_full(Ko)
        del K64
    if not args.no_workloads:
        line["workloads"] = workloads_leg(args, dev)
    if args.serving > 0:
        line["serving"] = serving_leg(args, dev)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
