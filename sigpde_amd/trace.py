"""roctx ranges around the API's calls and phases.

With SIGPDE_ROCTX=1 in the environment, every public call of sigpde_amd.torch (SigPDE.pairwise /
.gram, RobustSigPDE.normalization / .pairwise / .gram) and its phases (root-find, final solves,
increment tiles) push a named roctx range, so `rocprofv3 --marker-trace` (or any roctx consumer)
shows them on the host timeline beside the kernels they launch. Without the variable, or when
the roctx library cannot be loaded, the ranges cost one attribute test and nothing is called.
"""
from __future__ import annotations

import ctypes
import functools
import os

_LIBS = ("librocprofiler-sdk-roctx.so.1", "libroctx64.so.4", "libroctx64.so")


def _load():
    for name in _LIBS:
        for path in (os.path.join("/opt/rocm/lib", name), name):
            try:
                lib = ctypes.CDLL(path)
            except OSError:
                continue
            lib.roctxRangePushA.restype = ctypes.c_int
            lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
            lib.roctxRangePop.restype = ctypes.c_int
            lib.roctxRangePop.argtypes = []
            return lib
    return None


class _State:
    lib = None
    enabled = False


def enable(on: bool = True) -> bool:
    """Turns the ranges on (loading the roctx library) or off; returns whether they are on."""
    if on and _State.lib is None:
        _State.lib = _load()
    _State.enabled = bool(on and _State.lib is not None)
    return _State.enabled


def enabled() -> bool:
    return _State.enabled


class span:
    """Context manager: one roctx range named `name` while the block runs."""
    __slots__ = ("name",)

    def __init__(self, name: str):
        self.name = name

    def __enter__(self):
        if _State.enabled:
            _State.lib.roctxRangePushA(self.name.encode())
        return self

    def __exit__(self, *exc):
        if _State.enabled:
            _State.lib.roctxRangePop()
        return False


def traced(name: str):
    """Decorator: the call runs inside a roctx range `name`."""
    def wrap(fn):
        @functools.wraps(fn)
        def call(*args, **kwargs):
            if not _State.enabled:
                return fn(*args, **kwargs)
            _State.lib.roctxRangePushA(name.encode())
            try:
                return fn(*args, **kwargs)
            finally:
                _State.lib.roctxRangePop()
        return call
    return wrap


if os.environ.get("SIGPDE_ROCTX", "") == "1":
    enable(True)
