"""Coherent strong-field driver on a reduced line (8192 points, two-level transform): the
stationary-state energies, frames (coordinate density, projected momentum density) and
Ehrenfest lists equal the oracle's run of the same flow (FP64)."""
import numpy as np
import pytest

from oracle.awig_oracle import SchrodingerWignerOracle

pytestmark = pytest.mark.gpu


def test_coherent_driver_matches_oracle(tmp_path):
    from accuratewigner_amd.coherent_strong_field import (CoherentRun, NpzWriter, coulomb_sys_params,
                                                          projected_momentum_density)
    p = coulomb_sys_params(X_gridDIM=8192, X_amplitude=500.)
    out = str(tmp_path / "c.npz")
    with NpzWriter(out) as w:
        run = CoherentRun(p, writer=w, nstates=3, nsteps=200, num_iteration=30)
        run.run(max_frames=2)
        run.save_ehrenfest()
    F, om, Xa = p["F"], p["omega"], p["X_amplitude"]
    E = lambda t: -F * np.sin(om * t) * np.sin(om * t / 16.) ** 2     # noqa: E731
    o = SchrodingerWignerOracle(
        X_gridDIM=8192, X_amplitude=Xa, dt=p["dt"],
        V=lambda X, t: -1. / np.sqrt(X * X + 1.37) + X * E(t), K=lambda P, t: 0.5 * P * P,
        diff_V=lambda X, P, t: X / (X * X + 1.37) ** 1.5 + E(t), diff_K=lambda X, P, t: P,
        abs_boundary_x=lambda X: (1. - np.exp(-50. * np.abs(X - Xa) / Xa)) * (1. - np.exp(-50. * np.abs(X + Xa) / Xa)))
    o.get_stationary_states(3, nsteps=200)
    o.wavefunction = o.stationary_states[-1].copy()
    top = o.get_energy()
    o.wavefunction = o.stationary_states[0].copy()
    Ip = -o.get_energy()
    assert abs(run.top_energy - top) < 1e-10 and abs(run.Ip - Ip) < 1e-10
    assert Ip > 0.0       # imaginary time has already lowered the energy below zero
    z = np.load(out, allow_pickle=False)
    for f in range(2):
        o.propagate(30)
        xd = np.abs(o.wavefunction.reshape(-1)) ** 2
        pd = projected_momentum_density(o.projectout_stationary_states(o.wavefunction.copy()))
        assert np.linalg.norm(z["frames/%d/x_density" % f] - xd) / np.linalg.norm(xd) < 1e-10
        assert np.max(np.abs(z["frames/%d/p_density" % f] - pd)) < 1e-8
        assert abs(float(z["frames/%d/t" % f]) - o.t) < 1e-12
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        a, b = z["ehrenfest/" + name], np.array(getattr(o, name))
        assert a.shape == (60,)
        assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b))), name
    h = np.array(o.hamiltonian_average)
    assert abs(run.energy_preservation_percent() - (1. - h.min() / h.max()) * 100) < 1e-8


def test_coherent_driver_full_line_14_states_matches_oracle():
    """The driver at the reference's own shape (coherent_strong_field_physics_cuda_1d.py:31,
    :171-199): 32768 points over [-2000, 2000), 14 stationary states by imaginary time with
    Gram-Schmidt against the states found so far (the multi-block projection; 40 imaginary
    steps each instead of 10000), then one 300-step frame with Ehrenfest lists."""
    from accuratewigner_amd.coherent_strong_field import CoherentRun, coulomb_sys_params, projected_momentum_density
    p = coulomb_sys_params()
    assert p["X_gridDIM"] == 32768 and p["X_amplitude"] == 2000.
    run = CoherentRun(p, nstates=14, nsteps=40, num_iteration=300)
    F, om, Xa = p["F"], p["omega"], p["X_amplitude"]
    E = lambda t: -F * np.sin(om * t) * np.sin(om * t / 16.) ** 2     # noqa: E731
    o = SchrodingerWignerOracle(
        X_gridDIM=32768, X_amplitude=Xa, dt=p["dt"],
        V=lambda X, t: -1. / np.sqrt(X * X + 1.37) + X * E(t), K=lambda P, t: 0.5 * P * P,
        diff_V=lambda X, P, t: X / (X * X + 1.37) ** 1.5 + E(t), diff_K=lambda X, P, t: P,
        abs_boundary_x=lambda X: (1. - np.exp(-50. * np.abs(X - Xa) / Xa)) * (1. - np.exp(-50. * np.abs(X + Xa) / Xa)))
    o.get_stationary_states(14, nsteps=40)
    sysm = run.coulomb_sys
    assert len(sysm.stationary_states) == 14
    for k, (a, b) in enumerate(zip(sysm.stationary_states, o.stationary_states)):
        a = a.get().reshape(-1)
        b = b.reshape(-1)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-9, k
    # the states are orthonormal (the projection's own invariant)
    S = np.array([[np.vdot(a.get().reshape(-1), b.get().reshape(-1)) for b in sysm.stationary_states]
                  for a in sysm.stationary_states])
    assert np.max(np.abs(S - np.eye(14))) < 1e-10
    o.wavefunction = o.stationary_states[-1].copy()
    top = o.get_energy()
    o.wavefunction = o.stationary_states[0].copy()
    Ip = -o.get_energy()
    assert abs(run.top_energy - top) < 1e-10 and abs(run.Ip - Ip) < 1e-10
    xd, pd = run.frame()
    o.propagate(300)
    xo = np.abs(o.wavefunction.reshape(-1)) ** 2
    assert np.linalg.norm(xd - xo) / np.linalg.norm(xo) < 1e-10
    assert np.max(np.abs(pd - projected_momentum_density(o.projectout_stationary_states(o.wavefunction.copy())))) < 1e-8
    for name in ("X_average", "P_average", "hamiltonian_average"):
        a, b = np.array(getattr(sysm, name)), np.array(getattr(o, name))
        assert a.shape == (300,)
        assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b))), name
