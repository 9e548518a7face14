"""The reference's README usage (stojl/SigPDE README.md:13-57) through `import sigpde.torch as sig`,
with the README's own shapes and calls (a seed added), every result checked against the oracle.

The README samples points, not paths: torch.randn((30, 50, 30)) in float32 at dyadic order 2. Their
increments are large, so K(x, y) overflows float32 to NaN (inf - inf), in the oracle's float32
semantics as on the device; the robust kernel maps K(x, x) = NaN to +inf, normalises to target 2
and returns finite values, which are compared. The point Gram of such inputs is not exact in
float32, so float32 results carry the GEMM's summation-order rounding (~1e-6 here); the bars
below are 1e-4 relative (Frobenius) for kernel values and 1e-5 for the normalisation factors.
"""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import max_rel, rel_fro

pytestmark = pytest.mark.gpu


def test_reference_readme_usage():
    import sigpde.torch as sig

    torch.manual_seed(0)
    # Parameters
    batch_size = 30
    length = 50
    dim = 30
    dyadic_order = 2

    # Sample data on GPU
    x = torch.randn((batch_size, length, dim), device='cuda')
    y = torch.randn((batch_size, length, dim), device='cuda')

    # Define kernels
    static_kernel = sig.kernels.LinearKernel()
    kernel = sig.SigPDE(static_kernel, dyadic_order)
    robust_kernel = sig.RobustSigPDE(static_kernel, dyadic_order)

    out = {}
    out["pairwise_xy"] = kernel.pairwise(x, y)
    out["pairwise_x"] = kernel.pairwise(x)
    out["robust_pairwise_xy"] = robust_kernel.pairwise(x, y)
    out["robust_pairwise_x"] = robust_kernel.pairwise(x)
    out["gram_xy"] = kernel.gram(x, y)
    out["gram_x"] = kernel.gram(x)
    out["robust_gram_x"] = robust_kernel.gram(x)
    out["robust_gram_xy"] = robust_kernel.gram(x, y)
    out["normalization"] = robust_kernel.normalization(x)

    def my_normalizer(x):
        return 2 - 1 / x.pow(0.1)

    out["robust_pairwise_custom"] = robust_kernel.pairwise(x, normalizer=my_normalizer)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}

    X, Y = x.cpu().numpy(), y.cpu().numpy()
    for k, v in got.items():
        assert v.dtype == np.float32, k
    # plain kernel: float32 overflow to NaN, as in the oracle's float32 semantics
    for key, ref in (("pairwise_xy", O.sig_pairwise(X, Y, 2)), ("pairwise_x", O.sig_pairwise(X, X, 2)),
                     ("gram_xy", O.sig_gram(X, Y, 2)), ("gram_x", O.sig_gram(X, X, 2, symmetric=True))):
        g = got[key]
        assert g.shape == ref.shape, key
        assert np.array_equal(np.isnan(g), np.isnan(ref)), key
        fin = np.isfinite(ref)
        if fin.any():
            assert rel_fro(g[fin], ref[fin]) < 1e-4, key

    # robust normalisation factors of x, of y, and with the custom normaliser
    lam_x = O.robust_normalization(X, 2)
    lam_y = O.robust_normalization(Y, 2)
    assert got["normalization"].shape == (30,)
    assert max_rel(got["normalization"], lam_x, floor=1e-3) < 1e-5
    assert np.all((lam_x > 0) & (lam_x < 1))

    def np_normalizer(v):
        return 2 - 1 / np.power(v, 0.1)

    inc_xx = O.second_difference(O.static_pairwise("linear", X, X), 2)
    lam_c = O.norm_factors(inc_xx, 2, np_normalizer)[0]

    # robust kernel values: the plain kernel of the normalised paths (the oracle's own lambdas)
    for key, ref in (
            ("robust_pairwise_xy", O.sig_pairwise(X, Y, 2, sx=lam_x, sy=lam_y)),
            ("robust_pairwise_x", O.sig_pairwise(X, X, 2, sx=lam_x, sy=lam_x)),
            ("robust_gram_x", O.sig_gram(X, X, 2, sx=lam_x, sy=lam_x, symmetric=True)),
            ("robust_gram_xy", O.sig_gram(X, Y, 2, sx=lam_x, sy=lam_y)),
            ("robust_pairwise_custom", O.sig_pairwise(X, X, 2, sx=lam_c, sy=lam_c))):
        g = got[key]
        assert g.shape == ref.shape and np.all(np.isfinite(g)), key
        assert rel_fro(g, ref) < 1e-4, key
        ref64 = (O.sig_pairwise if g.ndim == 1 else O.sig_gram)(
            X.astype(np.float64), (X if key in ("robust_pairwise_x", "robust_gram_x",
                                                "robust_pairwise_custom") else Y).astype(np.float64),
            2, sx=(lam_c if key == "robust_pairwise_custom" else lam_x).astype(np.float64),
            sy=(lam_c if key == "robust_pairwise_custom" else
                lam_x if key in ("robust_pairwise_x", "robust_gram_x") else lam_y).astype(np.float64))
        assert rel_fro(g, ref64) < 1e-4, key
    assert np.array_equal(got["robust_gram_x"], got["robust_gram_x"].T)
