"""Captured (hipGraph) Ehrenfest chunks against the same chunks run eagerly.

The graph replays the same kernels in the same order from the same buffers, so the results
are bit-identical.  Each propagate runs two full chunks (replays) and a shorter eager tail;
the second propagate replays the graph again after eager work and host synchronisation in
between, and every replay's observables are compared.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _wm():
    from accuratewigner_amd import WignerMoyalCUDA1D
    g = WignerMoyalCUDA1D(X_gridDIM=256, X_amplitude=10., P_gridDIM=256, P_amplitude=10., dt=0.01, t=0.,
                          K="0.5 * P * P", V="0.5 * X * X + 0.02 * X * sin(t)", diff_K="P",
                          diff_V="X + 0.02 * sin(t)")
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    return g


def _wm_merged():
    from accuratewigner_amd import WignerMoyalCUDA1D
    g = WignerMoyalCUDA1D(X_gridDIM=256, X_amplitude=10., P_gridDIM=256, P_amplitude=10., dt=0.01, t=0.,
                          K="0.5 * P * P", V="0.5 * X * X + 0.02 * X * sin(t)", precision="mixed")
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    return g


def _rv():
    from accuratewigner_amd import RhoVNeumannCUDA1D
    from accuratewigner_amd.strong_field import rv_sys_params
    p = rv_sys_params(X_gridDIM=256, X_amplitude=50., dt=0.02, omega=0.6, cycles=1)
    g = RhoVNeumannCUDA1D(**p)
    g.set_rho("exp(-X * X - X_prime * X_prime)")
    return g


def _sw():
    from accuratewigner_amd import SchrodingerWignerCUDA1D
    g = SchrodingerWignerCUDA1D(X_gridDIM=8192, X_amplitude=60., dt=0.01, t=0., K="0.5 * P * P",
                                V="-1. / sqrt(X * X + 1.37) + 0.05 * X * sin(0.06 * t)", diff_K="P",
                                diff_V="X / pow(X * X + 1.37, 1.5) + 0.05 * sin(0.06 * t)")
    g.set_wavefunction("exp(-X * X)")
    return g


@pytest.mark.parametrize("make,state", [(_wm, "wignerfunction"), (_rv, "rho"), (_wm_merged, "wignerfunction"),
                                        (_sw, "wavefunction")])
def test_captured_chunks_equal_eager(make, state, monkeypatch):
    import torch
    names = ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average")
    runs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("AWIG_GRAPHS", mode)
        g = make()
        out = []
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.propagate(600)          # two 256-step chunks and an 88-step tail
            torch.cuda.synchronize()
            print("graphs=%s propagate %d: %.1f ms" % (mode, rep, (time.perf_counter() - t0) * 1e3), flush=True)
            out.append(getattr(g, state).get())
        runs[mode] = (out, {n: np.array(getattr(g, n, [])) for n in names})
    (s0, l0), (s1, l1) = runs["0"], runs["1"]
    for a, b in zip(s0, s1):
        assert np.array_equal(a, b)
    if l0["X_average"].size:
        for n in names:
            assert l0[n].shape == (1200,)
            assert np.array_equal(l0[n], l1[n]), n
