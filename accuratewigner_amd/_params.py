"""Constructor keywords of the propagator classes.

The reference stores every keyword as an attribute and binds plain Python functions to the
instance with ``MethodType`` (wigner_moyal_cuda_1d.py:44-50, rho_vneumann_cuda_1d.py:24-30,
schrodinger_wigner_cuda_1d.py:30-36), so that ``self.E(t)`` works for a laser field given as
``E=lambda self, t: ...``.  Here the functions are bound on each access instead of being
stored bound: a stored bound method refers back to its instance, and that cycle would keep
a dropped system (its plan and its device memory) alive until Python's cyclic collector
runs, at an arbitrary later point, possibly while a hipGraph capture is open.
"""
from types import FunctionType, MethodType


class KeywordParams:
    def _set_params(self, kwargs):
        funcs = self.__dict__.setdefault("_user_funcs", {})
        for name, value in kwargs.items():
            if isinstance(value, FunctionType):
                funcs[name] = value
            else:
                setattr(self, name, value)

    def __getattr__(self, name):
        # only reached when normal lookup fails: user functions, bound afresh
        funcs = self.__dict__.get("_user_funcs")
        if funcs is not None and name in funcs:
            return MethodType(funcs[name], self)
        raise AttributeError("%r object has no attribute %r" % (type(self).__name__, name))
