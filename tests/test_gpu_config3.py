"""Strong-field von Neumann workload (rho(x, x') in a laser pulse, soft-core atom, absorbing
edge, dipole acceleration -<diff_V> every step).

* reduced grid, FP64: rho and every Ehrenfest list equal the oracle (1e-10);
* full size (1024^2, x in [-200, 200), dt = 0.02, 4-cycle pulse = 20944 steps): the mixed
  tier's dipole acceleration against the FP64 tier's, max error / max |a| below 1.5e-2
  (measured 6.1e-3: the acceleration, at most 2.3e-3, is a small difference of the force
  -<x/(x^2+1)^1.5> and the field E(t) of amplitude 0.05, so FP32 rounding in rho, about
  1e-4 after 20944 steps, is amplified about 20 times).  <x>, <p> and <H> are bounded at
  3e-2 of their maxima (measured 6.0e-3, 1.05e-2, 1.1e-3: the ionised tail reaches the
  absorbing edge at |x| ~ 180, where FP32 rounding in rho weighs most in <x> and <p>).
"""
import time

import numpy as np
import pytest

from oracle.awig_oracle import RhoVNeumannOracle

pytestmark = pytest.mark.gpu


def _oracle_for(p):
    F, om, T = p["F"], p["omega"], p["T_pulse"]
    Xa = p["X_amplitude"]
    E = lambda t: F * np.sin(np.pi * t / T) ** 2 * np.cos(om * t)    # noqa: E731
    return RhoVNeumannOracle(
        X_gridDIM=p["X_gridDIM"], X_amplitude=Xa, dt=p["dt"],
        V=lambda X, t: -1. / np.sqrt(X * X + 1.) + X * E(t), K=lambda P, t: 0.5 * P * P,
        diff_K=lambda X, P, t: P, diff_V=lambda X, P, t: X / (X * X + 1.) ** 1.5 + E(t),
        abs_boundary_x=lambda X: np.where(np.abs(X) < 0.9 * Xa, 1.,
                                          np.abs(np.cos(0.5 * np.pi * (np.abs(X) - 0.9 * Xa) / (0.1 * Xa)))
                                          ** 0.125))


def test_strong_field_rv_reduced_matches_oracle():
    from accuratewigner_amd import RhoVNeumannCUDA1D
    from accuratewigner_amd.strong_field import rv_ground_state_rho, rv_sys_params
    p = rv_sys_params(X_gridDIM=256, X_amplitude=50., dt=0.02, omega=0.6, cycles=1)
    rho0 = rv_ground_state_rho(p, nsteps=2000)
    g = RhoVNeumannCUDA1D(**p)
    g.set_rho(rho0)
    o = _oracle_for(p)
    o.set_rho(rho0)
    g.propagate(300)
    o.propagate(300)
    r = g.rho.get()
    assert np.linalg.norm(r - o.rho) / np.linalg.norm(o.rho) < 1e-10
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
        assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b))), name
    assert np.max(np.abs(g.P_average_RHS)) > 1e-3          # the field drives a dipole response


def test_strong_field_rv_full_mixed_vs_fp64():
    import torch
    from accuratewigner_amd import RhoVNeumannCUDA1D
    from accuratewigner_amd.strong_field import rv_ground_state_rho, rv_sys_params
    p = rv_sys_params()
    steps = int(round(p["T_pulse"] / p["dt"]))
    assert steps == 20944
    rho0 = rv_ground_state_rho(p, nsteps=5000)
    acc, obs, wall = {}, {}, {}
    for prec in ("fp64", "mixed"):
        g = RhoVNeumannCUDA1D(precision=prec, **p)
        g.set_rho(rho0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.propagate(steps)
        torch.cuda.synchronize()
        wall[prec] = time.perf_counter() - t0
        acc[prec] = np.array(g.P_average_RHS)
        obs[prec] = {n: np.array(getattr(g, n)) for n in ("X_average", "P_average", "hamiltonian_average")}
        print(prec, "steps", steps, "wall %.2f s" % wall[prec], flush=True)
        del g
    a64, a32 = acc["fp64"], acc["mixed"]
    assert a64.shape == (steps,) and np.all(np.isfinite(a32))
    err = np.max(np.abs(a32 - a64)) / np.max(np.abs(a64))
    print("dipole acceleration: max|a| %.3e, mixed vs fp64 max error / max|a| %.3e" % (np.max(np.abs(a64)), err),
          flush=True)
    assert err < 1.5e-2
    for n in ("X_average", "P_average", "hamiltonian_average"):
        e = np.max(np.abs(obs["mixed"][n] - obs["fp64"][n])) / np.max(np.abs(obs["fp64"][n]))
        print("%s: mixed vs fp64 max error / max |value| %.3e" % (n, e), flush=True)
        assert e < 3e-2, n


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
@pytest.mark.parametrize("N", [64, 256, 1024, 2048])
def test_rv_merged_ehrenfest_equals_per_step(prec, N):
    """Merged Ehrenfest steps (two passes a step, the trace and averages of each closed step
    stored by the merged column pass) against the per-step path (three passes, a reduction
    of the stored rho and a normalisation every step): the same lists and final rho, over
    two chunks and a second call, with a time-dependent field and an absorbing edge.  At 64
    points the column pass has no partials and the per-step path runs in both."""
    from accuratewigner_amd import RhoVNeumannCUDA1D
    from accuratewigner_amd.strong_field import rv_sys_params
    p = rv_sys_params(X_gridDIM=N, X_amplitude=50. * max(N, 256) / 256, dt=0.02, omega=0.6, cycles=1)
    init = "exp(-pow(X - 1., 2) / 2. + cuda_complex(0., 0.3) * X) * exp(-pow(X_prime - 1., 2) / 2. - cuda_complex(0., 0.3) * X_prime)"
    out = {}
    for merged in (True, False):
        g = RhoVNeumannCUDA1D(precision=prec, **p)
        g.merged_ehrenfest = merged
        g.set_rho(init)
        g.propagate(300)
        g.propagate(40)
        assert (g._ehr["nblk"] > 0) == (merged and N > 64)
        out[merged] = (g.rho.get().astype(np.complex128),
                       {n: np.array(getattr(g, n)) for n in ("X_average", "P_average", "X_average_RHS",
                                                              "P_average_RHS", "hamiltonian_average")})
        del g
    tol = 1e-11 if prec == "fp64" else 2e-5
    (r1, l1), (r0, l0) = out[True], out[False]
    assert np.linalg.norm(r1 - r0) / np.linalg.norm(r0) < tol
    for name in l0:
        assert l1[name].shape == (340,)
        assert np.max(np.abs(l1[name] - l0[name])) < tol * max(1.0, np.max(np.abs(l0[name]))), name
