"""Writes the golden fixtures in this folder: seeded inputs of the three smaller workloads of the
API and the CPU oracle's outputs on them (the reference has no golden vectors or known-answer
tests of its own, and it cannot run here, so these pin later oracle and GPU runs to the oracle's
present outputs). Run from the repository root: python tests/golden/make_fixtures.py

  pairwise_f64_L32_d4_o1.npz   SigPDE(LinearKernel(), 1).pairwise(x, y), 8 pairs, length 32,
                               dim 4, N(0, 1/32) increments, float64 (seeds 0 / 1)
  gram_f32_L100_d8_o2.npz      SigPDE(LinearKernel(), 2).gram(X, Y) on 12 x 12 paths, length 100,
                               dim 8, N(0, 1/100), float32 inputs (seeds 2 / 3): K64 on the inputs
                               upcast, K32 in the reference's float32 semantics
  robust_f64_L64_d16_o1.npz    RobustSigPDE(LinearKernel(), 1) on 16 paths, length 64, dim 16,
                               N(0, 9/64), float64 (seed 4), tol 1e-10, maxit 100: lambda, K(x,x),
                               root-find evaluations, and the normalised Gram
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from sigpde_amd.synthetic import brownian_paths  # noqa: E402


def paths(batch, length, dim, var, seed, dtype):
    import torch
    return brownian_paths(batch, length, dim, var, seed, torch.float64).numpy().astype(dtype)


def main():
    x = paths(8, 32, 4, 1 / 32, 0, np.float64)
    y = paths(8, 32, 4, 1 / 32, 1, np.float64)
    np.savez_compressed(os.path.join(HERE, "pairwise_f64_L32_d4_o1.npz"), x=x, y=y,
                        K64=oracle.sig_pairwise(x, y, 1))

    X = paths(12, 100, 8, 1 / 100, 2, np.float32)
    Y = paths(12, 100, 8, 1 / 100, 3, np.float32)
    np.savez_compressed(os.path.join(HERE, "gram_f32_L100_d8_o2.npz"), X=X, Y=Y,
                        K64=oracle.linear_gram(X.astype(np.float64), Y.astype(np.float64), 2),
                        K32=oracle.linear_gram(X, Y, 2))

    Xr = paths(16, 64, 16, 9 / 64, 4, np.float64)
    inc = oracle.second_difference(oracle.static_pairwise("linear", Xr, Xr), 1)
    lam, knorm, iters = oracle.norm_factors(inc, 1, tol=1e-10, maxit=100)
    np.savez_compressed(os.path.join(HERE, "robust_f64_L64_d16_o1.npz"), X=Xr, lam=lam,
                        knorm=knorm, iters=iters,
                        K=oracle.linear_gram(Xr, Xr, 1, sx=lam, sy=lam))


if __name__ == "__main__":
    main()
