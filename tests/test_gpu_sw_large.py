"""Schrodinger lines longer than 4096 points (two-level transform over [N1][N2]) against the
oracle's plain length-N DFTs: propagation with a field, an absorbing edge and Ehrenfest
lists, merged propagation and energy, imaginary-time ground and excited states."""
import numpy as np
import pytest

from oracle.awig_oracle import SchrodingerWignerOracle

pytestmark = pytest.mark.gpu

XA = 100.
V_C = "-1. / sqrt(X * X + 1.37) + X * 0.05 * sin(0.06 * t)"
MASK_C = "(1. - exp(-50. * abs(X - X_amplitude) / X_amplitude)) * (1. - exp(-50. * abs(X + X_amplitude) / X_amplitude))"


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _pair(N, ehrenfest=True, dt=0.02):
    from accuratewigner_amd import SchrodingerWignerCUDA1D
    Vf = lambda X, t: -1. / np.sqrt(X * X + 1.37) + X * 0.05 * np.sin(0.06 * t)    # noqa: E731
    mask = lambda X: ((1. - np.exp(-50. * np.abs(X - XA) / XA))                     # noqa: E731
                      * (1. - np.exp(-50. * np.abs(X + XA) / XA)))
    extra, oextra = dict(abs_boundary_x=MASK_C), dict(abs_boundary_x=mask)
    if ehrenfest:
        extra.update(diff_V="X / pow(X * X + 1.37, 1.5) + 0.05 * sin(0.06 * t)", diff_K="P")
        oextra.update(diff_V=lambda X, P, t: X / (X * X + 1.37) ** 1.5 + 0.05 * np.sin(0.06 * t),
                      diff_K=lambda X, P, t: P)
    g = SchrodingerWignerCUDA1D(X_gridDIM=N, X_amplitude=XA, dt=dt, t=0., V=V_C, K="0.5 * P * P", **extra)
    o = SchrodingerWignerOracle(X_gridDIM=N, X_amplitude=XA, dt=dt, V=Vf, K=lambda P, t: 0.5 * P * P, **oextra)
    g.set_wavefunction("exp(-pow(X - 2., 2) / 2. + cuda_complex(0., 0.5) * X)")
    o.set_wavefunction(lambda X: np.exp(-(X - 2.) ** 2 / 2 + 0.5j * X))
    return g, o


@pytest.mark.parametrize("N", [8192, 32768])
def test_two_level_propagate_ehrenfest(N):
    g, o = _pair(N)
    assert g._two is not None
    g.propagate(150)
    o.propagate(150)
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
        assert a.shape == (150,)
        assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b))), name


@pytest.mark.parametrize("N", [8192, 32768])
def test_two_level_merged_propagate_and_energy(N):
    g, o = _pair(N, ehrenfest=False)
    g.propagate(300)
    o.propagate(300)
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    assert abs(g.get_energy() - o.get_energy()) < 1e-10
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    with pytest.raises(NotImplementedError):
        g.get_wigner()


def test_two_level_ground_and_stationary_states():
    g, o = _pair(8192, ehrenfest=False, dt=0.05)
    g.set_ground_state(nsteps=400)
    o.set_ground_state(nsteps=400)
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    g.get_stationary_states(2, nsteps=300)
    o.get_stationary_states(2, nsteps=300)
    for a, b in zip(g.stationary_states, o.stationary_states):
        assert rel(a.get(), b) < 1e-9
