"""sigpde_linear_gram_batch (several fused Grams of one shape class in one launch, the serving
loop's coalesced Gram requests) against one sigpde_linear_gram call per request: equal bit for
bit, in every precision, with symmetric, scaled, empty and uneven requests."""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import rel_fro
from sigpde_amd import _native
from sigpde_amd.synthetic import brownian_paths

pytestmark = pytest.mark.gpu
DEV = "cuda"


def paths(B, L, d, seed, dtype=torch.float32):
    return brownian_paths(B, L, d, 1.0 / L, seed, dtype, DEV)


@pytest.mark.parametrize("precision,dtype", [("fp16x2", torch.float32), ("bf16x3", torch.float32),
                                             ("fp32", torch.float32), ("fp64", torch.float64),
                                             ("fp16x2f", torch.float32)])
@pytest.mark.parametrize("order", [0, 2])
def test_batch_equals_separate_grams(precision, dtype, order):
    L, d = 40, 16
    sizes = [(7, 5), (1, 1), (13, 30), (0, 4), (4, 0), (9, 9)]
    xs = [paths(bx, L, d, 10 + i, dtype) for i, (bx, _) in enumerate(sizes)]
    ys = [paths(by, L, d, 50 + i, dtype) for i, (_, by) in enumerate(sizes)]
    X, Y = torch.cat(xs), torch.cat(ys)
    rows, x0, y0 = [], 0, 0
    for bx, by in sizes:
        rows.append((x0, bx, y0, by, 0))
        x0 += bx
        y0 += by
    out = torch.full((sum(bx * by for bx, by in sizes),), -5.0, dtype=dtype, device=DEV)
    _native.linear_gram_batch(X, Y, rows, order, 1.3, out, precision)
    off = 0
    for (bx, by), x, y in zip(sizes, xs, ys):
        blk = out[off:off + bx * by].view(bx, by)
        off += bx * by
        if bx == 0 or by == 0:
            continue
        ref = torch.zeros((bx, by), dtype=dtype, device=DEV)
        _native.linear_gram(x.contiguous(), y.contiguous(), order, 1.3, ref, precision)
        assert torch.equal(blk, ref), (precision, bx, by)


@pytest.mark.parametrize("precision", ["fp16x2", "fp32", "fp16x2f"])
def test_batch_symmetric_and_scaled(precision):
    L, d, order = 33, 8, 1
    xs = [paths(b, L, d, 70 + b) for b in (6, 11, 1)]
    X = torch.cat(xs)
    rows, x0 = [], 0
    for x in xs:
        rows.append((x0, x.shape[0], x0, x.shape[0], 1))
        x0 += x.shape[0]
    s = torch.linspace(0.3, 1.0, X.shape[0], device=DEV)
    out = torch.empty(sum(x.shape[0] ** 2 for x in xs), device=DEV)
    _native.linear_gram_batch(X, X, rows, order, 1.0, out, precision, s, s)
    off, x0 = 0, 0
    for x in xs:
        b = x.shape[0]
        blk = out[off:off + b * b].view(b, b)
        ref = torch.zeros((b, b), device=DEV)
        sb = s[x0:x0 + b].contiguous()
        _native.linear_gram(x, x, order, 1.0, ref, precision, sb, sb, symmetric=True)
        assert torch.equal(blk, ref)
        assert torch.equal(blk, blk.T)
        # against the float64 oracle on the same float32 inputs
        k64 = O.linear_gram(x.double().cpu().numpy(), x.double().cpu().numpy(), order,
                            sx=sb.double().cpu().numpy(), sy=sb.double().cpu().numpy())
        assert rel_fro(blk.double().cpu().numpy(), k64) < (1e-4 if precision == "fp32" else 1e-5)
        off += b * b
        x0 += b


def test_batch_argument_checks():
    X = paths(4, 10, 3, 1)
    out = torch.empty(16, device=DEV)
    with pytest.raises(_native.NativeError, match="out of the"):
        _native.linear_gram_batch(X, X, [(2, 4, 0, 4, 0)], 0, 1.0, out)
    with pytest.raises(_native.NativeError, match="symmetric"):
        _native.linear_gram_batch(X, X, [(0, 4, 0, 2, 1)], 0, 1.0, torch.empty(8, device=DEV))
    with pytest.raises(ValueError, match="at most 64"):
        _native.linear_gram_batch(X, X, [(0, 1, 0, 1, 0)] * 65, 0, 1.0, torch.empty(65, device=DEV))
