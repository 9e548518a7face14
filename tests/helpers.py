"""Shared test helpers: seeded inputs and error measures."""
from __future__ import annotations

import numpy as np


def rel_fro(a, ref) -> float:
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(a - ref) / (den if den > 0 else 1.0))


def max_rel(a, ref, floor: float = 0.0) -> float:
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(a - ref) / np.maximum(np.abs(ref), floor))) if a.size else 0.0


def brownian_np(batch, length, dim, var, seed, dtype=np.float64):
    from sigpde_amd.synthetic import brownian_paths
    import torch
    return brownian_paths(batch, length, dim, var, seed, torch.float64).numpy().astype(dtype)


def quantized_paths(batch, length, dim, seed, bits=6, dtype=np.float32):
    """Paths on the grid 2^-bits with small integer steps: every point Gram entry, and every
    4-term difference of it, is exact in float32 whatever the summation order, so float32
    results do not depend on the GEMM's reduction order."""
    rng = np.random.default_rng(seed)
    steps = rng.integers(-2, 3, size=(batch, length - 1, dim))
    x = np.zeros((batch, length, dim), dtype=np.float64)
    x[:, 1:, :] = np.cumsum(steps, axis=1) * 2.0 ** (-bits)
    return x.astype(dtype)


def evals_with_fixed_point_exit(iters, fixed):
    """Wavefront evaluations of a root-find that also stops at a fixed point of its state: the
    oracle's fixed point where it has one (the reference keeps iterating from there without
    changing lambda), else the oracle's own count."""
    iters = np.asarray(iters)
    fixed = np.asarray(fixed)
    return np.where(fixed >= 0, fixed, iters).astype(iters.dtype)
