"""GPU checks of the one-pass increments of the static kernels
(sigpde_amd/csrc/static_increments.hip) against the torch formulation of the reference
(point Gram by the static kernel's own gram / pairwise, then the 4-term second difference,
sigpde/torch/kernel.py:13-21):
  linear and RBF: the same float operations in the same order, so the increments must be equal
    bit for bit;
  L1-RBF: the distance is summed in another order than torch's reduction, so they agree to the
    operand type's rounding (relative to the point kernel values, which are at most 1);
and end to end through SigPDE against the float64 oracle, plus the fallback to the torch route
for a subclass that overrides the point kernel.
"""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from helpers import brownian_np, rel_fro

pytestmark = pytest.mark.gpu
DEV = "cuda"


def torch_pairwise_inc(k, x, y, order):
    ip = k.pairwise(x, y)
    return (ip[:, 1:, 1:] + ip[:, :-1, :-1] - ip[:, 1:, :-1] - ip[:, :-1, 1:]) / float(4 ** order)


def torch_gram_inc(k, x, y, order):
    ip = k.gram(x, y)
    d = ip[:, :, 1:, 1:] + ip[:, :, :-1, :-1] - ip[:, :, 1:, :-1] - ip[:, :, :-1, 1:]
    return d / float(4 ** order)


def paths(B, L, d, seed, dtype):
    return torch.from_numpy(brownian_np(B, L, d, 1.0 / L, seed)).to(DEV, dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("order", [0, 2])
def test_linear_increments_equal_torch_formula(dtype, order):
    k = sig.kernels.LinearKernel(0.7)
    x, y = paths(5, 37, 6, 1, dtype), paths(4, 21, 6, 2, dtype)
    assert torch.equal(k.gram_increments(x, y, order), torch_gram_inc(k, x, y, order))
    y2 = paths(5, 21, 6, 3, dtype)
    assert torch.equal(k.pairwise_increments(x, y2, order), torch_pairwise_inc(k, x, y2, order))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("order", [0, 2])
def test_rbf_increments_equal_torch_formula(dtype, order):
    k = sig.kernels.RBFKernel(0.7)
    x, y = paths(5, 37, 6, 1, dtype), paths(4, 21, 6, 2, dtype)
    g = k.gram_increments(x, y, order)
    assert torch.equal(g, torch_gram_inc(k, x, y, order))
    y2 = paths(5, 21, 6, 3, dtype)
    p = k.pairwise_increments(x, y2, order)
    assert torch.equal(p, torch_pairwise_inc(k, x, y2, order))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("d", [3, 40])  # one and two staged dimension chunks
def test_l1rbf_increments_match_torch_formula(dtype, d):
    k = sig.kernels.L1RBFKernel(1.3)
    x, y = paths(3, 19, d, 4, dtype), paths(2, 35, d, 5, dtype)
    atol = 2e-6 if dtype == torch.float32 else 1e-14
    g = k.gram_increments(x, y, 1)
    torch.testing.assert_close(g, torch_gram_inc(k, x, y, 1), rtol=0, atol=atol)
    y2 = paths(3, 35, d, 6, dtype)
    p = k.pairwise_increments(x, y2, 0)
    torch.testing.assert_close(p, torch_pairwise_inc(k, x, y2, 0), rtol=0, atol=atol)


@pytest.mark.parametrize("kind,param", [("rbf", 0.7), ("l1rbf", 1.3)])
def test_sigpde_with_nonlinear_kernels_vs_oracle(kind, param):
    x = brownian_np(6, 30, 4, 1 / 30, 7)
    y = brownian_np(5, 26, 4, 1 / 26, 8)
    K = {"rbf": sig.kernels.RBFKernel, "l1rbf": sig.kernels.L1RBFKernel}[kind](param)
    k = sig.SigPDE(K, 2)
    xc, yc = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    assert rel_fro(k.gram(xc, yc).cpu().numpy(), O.sig_gram(x, y, 2, kind, param)) < 1e-10
    assert rel_fro(k.gram(xc).cpu().numpy(), O.sig_gram(x, x, 2, kind, param)) < 1e-10
    assert rel_fro(k.pairwise(xc[:5], yc).cpu().numpy(),
                   O.sig_pairwise(x[:5], y, 2, kind, param)) < 1e-10


def test_subclass_overriding_the_point_kernel_takes_the_torch_route():
    class Scaled(sig.kernels.RBFKernel):
        def gram(self, X, Y):
            return 0.5 * super().gram(X, Y)

    k = Scaled(0.7)
    assert k.gram_increments(paths(2, 9, 3, 1, torch.float64), paths(2, 9, 3, 2, torch.float64),
                             0) is None
    x, y = paths(3, 12, 3, 1, torch.float64), paths(2, 15, 3, 2, torch.float64)
    got = sig.SigPDE(k, 1).gram(x, y)
    ref = sig.SigPDE(sig.kernels.RBFKernel(0.7), 1).gram(x, y)
    # K(0.5 k) != K(k) / 2 in general: check against the oracle on halved increments instead
    inc = 0.5 * O.second_difference(O.static_gram("rbf", x.cpu().numpy(), y.cpu().numpy(), 0.7), 1)
    assert rel_fro(got.cpu().numpy(), O.pde_gram(inc, 1)) < 1e-10
    assert not torch.allclose(got, ref)
