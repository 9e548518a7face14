"""roctx ranges of the API (sigpde_amd.trace): off by default, push / pop balanced around calls and
exceptions when on. Runs without a GPU (the roctx library records ranges with no tool attached)."""
import pytest

from sigpde_amd import trace


def test_ranges_off_by_default_and_balanced_when_on():
    calls = []

    @trace.traced("unit")
    def f(x):
        calls.append(x)
        return x + 1

    was = trace.enabled()
    try:
        trace.enable(False)
        assert f(1) == 2
        if not trace.enable(True):
            pytest.skip("roctx library not loadable")
        assert f(2) == 3
        with trace.span("outer"):
            assert f(3) == 4
        # a raising call still pops its range: the next push starts at depth 0
        @trace.traced("raises")
        def g():
            raise KeyError("x")
        with pytest.raises(KeyError):
            g()
        assert trace._State.lib.roctxRangePushA(b"probe") == 0
        trace._State.lib.roctxRangePop()
    finally:
        trace.enable(was)
    assert calls == [1, 2, 3]


def test_api_methods_are_traced():
    import sigpde_amd.torch as sig
    for name in ("pairwise", "gram"):
        assert hasattr(getattr(sig.SigPDE, name), "__wrapped__")
    for name in ("normalization", "pairwise", "gram"):
        assert hasattr(getattr(sig.RobustSigPDE, name), "__wrapped__")
