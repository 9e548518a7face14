"""Built-in device normalisers of RobustSigPDE (sigpde_amd/torch/normalizers.py,
sigpde_robust_norm_builtin): the target is evaluated inside the root-find kernel. Checked against
the same formula passed as a plain host callable (the reference's route for any normaliser,
sigpde/torch/kernel.py:180-195: K(x,x) solved, the callable applied with torch, the root-find on
the given targets), and against the float64 oracle with a NumPy restatement of each formula.
"""
import numpy as np
import pytest
import torch

import oracle as O
import sigpde_amd.torch as sig
from helpers import brownian_np
from sigpde_amd import _native

pytestmark = pytest.mark.gpu
DEV = "cuda"
N = sig.normalizers


def np_tensor(C, a):
    def f(x):
        x = np.asarray(x)
        with np.errstate(all="ignore"):
            tail = C + C ** (1 + a) * (C ** (-a) - x ** (-a)) / a
        return np.where(x <= C, x, tail)
    return f


def np_clip(C):
    return lambda x: np.minimum(np.asarray(x), C)


CASES = [
    ("log", N.LogNormalizer(), O.log_normalizer),
    ("tensor_2_1", N.TensorNormalizer(2.0, 1.0), np_tensor(2.0, 1.0)),
    ("tensor_1.5_0.5", N.TensorNormalizer(1.5, 0.5), np_tensor(1.5, 0.5)),
    ("clip_2", N.ClipNormalizer(2.0), np_clip(2.0)),
]


def paths(B, L, d, var, seed, dtype):
    return torch.from_numpy(brownian_np(B, L, d, var, seed)).to(DEV, dtype)


@pytest.mark.parametrize("name,obj,ref", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_builtin_equals_host_callable(name, obj, ref, dtype):
    x = paths(40, 24, 4, 9.0 / 24, 3, dtype)   # most K(x,x) > 1: normalisation active
    k = sig.RobustSigPDE(sig.kernels.LinearKernel(), 1)
    lam_dev = k.normalization(x, normalizer=obj, tol=1e-10)
    lam_host = k.normalization(x, normalizer=lambda kk: obj(kk), tol=1e-10)
    if name.startswith("clip") or name == "log":
        # min / the reference's own formula: the same targets to the bit
        tol = 0.0 if name.startswith("clip") else (1e-6 if dtype == torch.float32 else 1e-14)
    else:
        tol = 1e-5 if dtype == torch.float32 else 1e-12
    diff = (lam_dev.double() - lam_host.double()).abs().max().item()
    assert diff <= tol, (name, diff)
    assert (lam_dev > 0).sum() >= 30


@pytest.mark.parametrize("name,obj,ref", CASES, ids=[c[0] for c in CASES])
def test_builtin_float64_vs_oracle(name, obj, ref):
    x = paths(24, 20, 3, 9.0 / 20, 5, torch.float64)
    lam = sig.RobustSigPDE(sig.kernels.LinearKernel(), 1, normalizer=obj).normalization(
        x, tol=1e-10).cpu().numpy()
    xn = x.cpu().numpy()
    inc = O.second_difference(O.static_pairwise("linear", xn, xn), 1)
    lam64, knorm, _ = O.norm_factors(inc, 1, ref, tol=1e-10)
    assert np.abs(lam - lam64).max() < 1e-10
    assert (knorm > 1).sum() >= 18
    # the normalised kernels K(lambda x, lambda x) reach the normaliser's target
    tgt = ref(knorm)
    act = knorm > 1
    kk = sig.SigPDE(sig.kernels.LinearKernel(), 1).pairwise(
        x, x, x_scale=torch.from_numpy(lam).to(DEV), y_scale=torch.from_numpy(lam).to(DEV))
    assert np.abs(kk.cpu().numpy()[act] - tgt[act]).max() < 1e-8


def test_builtin_robust_gram_and_argument_checks():
    x = paths(12, 16, 3, 9.0 / 16, 7, torch.float64)
    obj = N.TensorNormalizer(2.0, 1.0)
    g_dev = sig.RobustSigPDE(sig.kernels.LinearKernel(), 0, normalizer=obj).gram(x, tol=1e-10)
    g_host = sig.RobustSigPDE(sig.kernels.LinearKernel(), 0,
                              normalizer=lambda kk: obj(kk)).gram(x, tol=1e-10)
    assert (g_dev - g_host).abs().max().item() < 1e-12
    with pytest.raises(ValueError):
        N.TensorNormalizer(0.5, 1.0)
    with pytest.raises(ValueError):
        N.ClipNormalizer(0.9)
    inc = torch.zeros((2, 3, 3), device=DEV, dtype=torch.float64)
    lam = torch.zeros(2, device=DEV, dtype=torch.float64)
    kn = torch.zeros(2, device=DEV, dtype=torch.float64)
    with pytest.raises(_native.NativeError, match="C >= 1"):
        _native.robust_norm_builtin(inc, 0, lam, kn, _native.NORM_TENSOR, 0.5, 1.0)
    with pytest.raises(_native.NativeError, match="unknown built-in"):
        _native.robust_norm_builtin(inc, 0, lam, kn, _native.NORM_GIVEN)
