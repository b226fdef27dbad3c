"""Constructor keywords and object lifetime (CPU): user functions act as bound methods
without a reference cycle, so a dropped system is freed at once rather than by a later
cyclic collection (which may fall inside a graph capture); plans dropped during a capture
are destroyed after it."""
import gc
import weakref

import pytest

from accuratewigner_amd._params import KeywordParams


class _Sys(KeywordParams):
    def __init__(self, **kw):
        self._set_params(kw)


def test_function_keywords_bind_to_the_instance():
    s = _Sys(F=2.0, E=lambda self, t: self.F * t, V="0.5 * X * X")
    assert s.E(3.0) == 6.0
    assert s.V == "0.5 * X * X"
    assert hasattr(s, "E") and not hasattr(s, "missing")
    with pytest.raises(AttributeError):
        s.missing


def test_dropped_system_is_freed_without_the_cyclic_collector():
    was = gc.isenabled()
    gc.disable()
    try:
        s = _Sys(F=1.0, E=lambda self, t: self.F * t)
        s.E(1.0)
        r = weakref.ref(s)
        del s
        assert r() is None
    finally:
        if was:
            gc.enable()


def test_plans_dropped_during_a_capture_are_destroyed_after_it(monkeypatch):
    from accuratewigner_amd import engine

    destroyed = []

    class _Lib:
        def aw_plan_destroy(self, h):
            destroyed.append(h)
            return 0

    monkeypatch.setattr(engine._lib, "load", lambda: _Lib())
    capturing = [True]
    monkeypatch.setattr(engine, "_capturing", lambda: capturing[0])
    p = engine.Plan.__new__(engine.Plan)
    p.handle = 12345
    del p
    assert destroyed == [] and engine._DEFERRED_PLANS == [12345]
    engine.destroy_deferred_plans()            # still capturing: kept
    assert destroyed == []
    capturing[0] = False
    engine.destroy_deferred_plans()
    assert destroyed == [12345] and engine._DEFERRED_PLANS == []


def test_time_independence_rule_for_observables():
    from accuratewigner_amd.wigner_moyal_cuda_1d import time_independent
    assert time_independent("0.5 * P * P + 0.5 * X * X")
    assert time_independent("1.") and time_independent("X / pow(X * X + 1.37, 1.5)")
    assert not time_independent("X * E(t)")
    assert not time_independent("sin(omega * t) * X")
    assert time_independent("tanh(X) + t_0", "")              # identifiers that merely contain t
    # a helper macro that reads t makes every observable suspect
    assert not time_independent("X * Et", "#define Et (F * sin(omega * t))")
    assert not time_independent("X * Et", "#define Et \\\n    (F * sin(omega * t))")   # continued line
    assert time_independent("X * E0", "__device__ double E(double t) { return t; }")


def test_wm_time_dependence_detection():
    from accuratewigner_amd.wigner_moyal_cuda_1d import time_independent
    assert time_independent("0.5 * X * X + 0.1 * sqrt(X * X + 1.)")
    assert not time_independent("0.5 * X * X + X * E(t)")
    assert not time_independent("X * F", "#define F (0.1 * sin(t))\n")
    assert time_independent("X * E(1.)", "__device__ double E(double t) { return sin(t); }\n")
