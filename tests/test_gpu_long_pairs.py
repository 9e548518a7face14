"""Two pairs of 4000 x 4000 cells (paths of length 4001, float64, order 0): few pairs, many bands,
the multi-wave route of sigpde_pde_pairwise / sigpde_robust_norm (one pair per workgroup of 16
waves, bands pipelined over the waves). Against the float64 oracle: the solves at 1e-10, and the
robust normalisation (Chandrupatla over about ten such solves per path) at 1e-10 on lambda."""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from helpers import brownian_np

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_two_pairs_of_4000_squared_vs_oracle():
    L = 4001
    x = brownian_np(2, L, 3, 1.0 / L, 21)
    y = brownian_np(2, L, 3, 1.0 / L, 22)
    print("2 x 4000^2 pairwise", flush=True)
    got = sig.SigPDE(sig.kernels.LinearKernel(), 0).pairwise(
        torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)).cpu().numpy()
    ref = O.sig_pairwise(x, y, 0)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-10


def test_robust_normalization_of_two_4000_paths_vs_oracle():
    L = 4001
    x = brownian_np(2, L, 3, 9.0 / L, 23)
    print("2 x 4000^2 robust normalisation", flush=True)
    lam = sig.RobustSigPDE(sig.kernels.LinearKernel(), 0).normalization(
        torch.from_numpy(x).to(DEV), tol=1e-10).cpu().numpy()
    ref = O.robust_normalization(x, 0, tol=1e-10)
    assert (ref > 0).all()
    assert np.abs(lam - ref).max() < 1e-10
