"""CPU checks of the C ABI: the library loads, exports every entry point declared in
include/sigpde_hip.h, and rejects bad arguments before touching the device."""
import ctypes
import os
import re

import pytest

from sigpde_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sigpde_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sigpde_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_native.LIB_PATH):
        from sigpde_amd import build
        build.build()
    return ctypes.CDLL(_native.LIB_PATH)


def test_header_declares_entry_points():
    names = declared_functions()
    for want in ("sigpde_pde_pairwise", "sigpde_pde_gram", "sigpde_robust_norm",
                 "sigpde_last_error", "sigpde_version"):
        assert want in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string(lib):
    L = _native.load()
    assert "gfx950" in L.sigpde_version().decode()
    assert L.sigpde_last_error().decode() == ""


def test_argument_validation_without_device(lib):
    L = _native.load()
    # negative batch, null pointers, oversize grid, unknown dtype: rejected before any HIP call
    assert L.sigpde_pde_pairwise(0, None, -1, 5, 5, 25, 0, None, None, None, None) == 1
    assert "bad batch" in L.sigpde_last_error().decode()
    fake = ctypes.c_void_p(16)
    # 2^22 + 1 refined columns; 2^20 x 2^14 refined grid: more virtual columns than int holds
    assert L.sigpde_pde_pairwise(0, fake, 2, 5, (1 << 22) + 1, 1 << 23, 0, None, None, fake, None) == 1
    assert "refined columns" in L.sigpde_last_error().decode()
    assert L.sigpde_pde_pairwise(0, fake, 1, 1 << 20, 1 << 14, 1 << 34, 0, None, None, fake, None) == 1
    assert "too large" in L.sigpde_last_error().decode()
    assert L.sigpde_pde_pairwise(7, fake, 2, 5, 5, 25, 0, None, None, fake, None) == 1
    assert "dtype" in L.sigpde_last_error().decode()
    assert L.sigpde_pde_gram(1, fake, 2, 2, 5, 6, 60, 30, 0, None, None, 1, 0, 0, fake, 2, None) == 1
    assert "symmetric" in L.sigpde_last_error().decode()
    assert L.sigpde_robust_norm(1, fake, 2, 5, 25, 0, 9, None, fake, 10, 1e-8, fake, None, None) == 1
    assert "normalizer" in L.sigpde_last_error().decode()
    # empty work is a no-op success
    assert L.sigpde_pde_pairwise(0, None, 0, 5, 5, 25, 0, None, None, None, None) == 0


def test_spread_settings_without_device(lib):
    """sigpde_set_spread validates its settings and touches no device."""
    L = _native.load()
    assert L.sigpde_set_spread(-1, 2.0, 0) == 1 and "groups" in L.sigpde_last_error().decode()
    assert L.sigpde_set_spread(33, 2.0, 0) == 1
    assert L.sigpde_set_spread(0, 0.0, 0) == 1 and "wait_seconds" in L.sigpde_last_error().decode()
    assert L.sigpde_set_spread(0, 2.0, 2) == 1 and "test_hook" in L.sigpde_last_error().decode()
    with pytest.raises(_native.NativeError) as ei:
        _native.set_spread(1, -1.0)
    assert ei.value.status == _native.EINVAL
    for g in (1, 4, 32, 0):
        assert L.sigpde_set_spread(g, 2.0, 0) == 0


def test_fused_entry_points_empty_and_negative_batches_without_device(lib):
    """The fused linear entry points return at once for an empty result, with null pointers (what
    empty device tensors carry), and reject a negative batch; neither touches the device."""
    L = _native.load()
    for prec in (1, 2, 3, 4):
        # pairwise: dtype, precision, X, Y, B, Lx, Ly, d, kscale, order, sx, sy, out, ws, bytes, stream
        assert L.sigpde_linear_pairwise(0, prec, None, None, 0, 7, 7, 3, 1.0, 1, None, None, None,
                                        None, 0, None) == 0
        # gram: dtype, precision, X, Bx, Lx, Y, By, Ly, d, kscale, order, sx, sy, symmetric, out,
        #       ldo, ws, bytes, stream
        assert L.sigpde_linear_gram(1, prec, None, 0, 7, None, 4, 6, 3, 1.0, 1, None, None, 0,
                                    None, 4, None, 0, None) == 0
        assert L.sigpde_linear_gram(0, prec, None, 4, 7, None, 0, 6, 3, 1.0, 0, None, None, 0,
                                    None, 0, None, 0, None) == 0
        assert L.sigpde_linear_pairwise(0, prec, None, None, -1, 7, 7, 3, 1.0, 1, None, None, None,
                                        None, 0, None) == 1
        assert "negative batch" in L.sigpde_last_error().decode()
        assert L.sigpde_linear_gram(0, prec, None, 2, 7, None, 3, 6, 3, 1.0, 0, None, None, 0,
                                    None, 3, None, 0, None) == 1
        assert "null pointer" in L.sigpde_last_error().decode()
