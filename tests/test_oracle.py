"""Pins for the NumPy oracle (oracle/awig_oracle.py) against closed forms, invariants,
brute force and exact diagonalisation.  CPU only.

The reference ships no tests or golden data (its checks are the demo blocks, e.g. the
harmonic Gibbs state at wigner_bloch_cuda_1d.py:406-425 and the energy-conservation print
at wigner_moyal_cuda_1d.py:972-975); these re-derive those checks with fixed parameters.
"""
import numpy as np
import pytest

from oracle.awig_oracle import (DFT, RhoBlochOracle, RhoVNeumannOracle, SchrodingerWignerOracle,
                                WignerBlochOracle, WignerMoyalOracle, sign_matrix)

HARM_V = lambda X, t: 0.5 * X * X  # noqa: E731
HARM_K = lambda P, t: 0.5 * P * P  # noqa: E731


def brute_dft(a, axis, sign):
    n = a.shape[axis]
    k = np.arange(n)
    M = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return np.moveaxis(np.tensordot(M, np.moveaxis(a, axis, 0), axes=(1, 0)), 0, axis)


@pytest.mark.parametrize("workers", [None, 2])
def test_dft_conventions_match_brute_force(workers):
    rng = np.random.default_rng(0)
    a = rng.normal(size=(16, 32)) + 1j * rng.normal(size=(16, 32))
    d = DFT(workers)
    for axis in (0, 1):
        np.testing.assert_allclose(d.F(a, axis), brute_dft(a, axis, -1), rtol=0, atol=1e-12)
        np.testing.assert_allclose(d.I(a, axis), brute_dft(a, axis, +1), rtol=0, atol=1e-12)


def test_dft_brute_force_pins_sign():
    """A transposed sign (forward = +) must be caught by the brute-force pin."""
    a = np.exp(2j * np.pi * 3 * np.arange(8) / 8)   # pure frequency +3
    assert abs(DFT().F(a, 0)[3] - 8) < 1e-12         # forward maps e^{+2pi i 3j/8} to bin 3
    assert abs(DFT().I(a, 0)[5] - 8) < 1e-12         # inverse maps it to bin -3 = 5


def _coherent_wm(Nx=128, Np=128, Xa=10., Pa=10., dt=0.01, ehrenfest=True, **kw):
    o = WignerMoyalOracle(X_gridDIM=Nx, X_amplitude=Xa, P_gridDIM=Np, P_amplitude=Pa, dt=dt,
                          V=HARM_V, K=HARM_K,
                          diff_V=(lambda X, P, t: X) if ehrenfest else None,
                          diff_K=(lambda X, P, t: P) if ehrenfest else None, **kw)
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    return o


def test_wigner2rho_closed_form_coherent_state():
    """rho[i, j] = psi(XX_j) conj(psi(XX'_i)) / dP for a coherent state W."""
    for Nx, Np, Xa, Pa in [(256, 256, 10., 10.), (128, 64, 8., 6.)]:
        o = WignerMoyalOracle(X_gridDIM=Nx, X_amplitude=Xa, P_gridDIM=Np, P_amplitude=Pa,
                              dt=0.01, V=HARM_V, K=HARM_K)
        x0, p0 = 1.0, 0.5
        o.set_wignerfunction(lambda X, P: np.exp(-(X - x0) ** 2 - (P - p0) ** 2) / np.pi)
        psi = lambda x: np.pi ** -0.25 * np.exp(-(x - x0) ** 2 / 2 + 1j * p0 * x)  # noqa: E731
        R = psi(o.kXX)[None, :] * np.conj(psi(o.kXXp))[:, None] / o.dP
        assert np.linalg.norm(o.rho - R) / np.linalg.norm(R) < 1e-8


def test_wigner_round_trip_and_parseval():
    o = _coherent_wm(ehrenfest=False)
    W0 = o.wignerfunction.copy()
    o.rho2wigner_()
    # rho2wigner_ o wigner2rho = Np * identity (unnormalised transforms, cufft.py:233-238)
    assert np.linalg.norm(o.wignerfunction - o.P_gridDIM * W0) / np.linalg.norm(W0) < 1e-12
    o.normalize_rho_wigner()
    assert abs(np.sum(o.wignerfunction * o.wignerfunction).real
               - o.P_gridDIM * np.sum(np.abs(o.rho) ** 2)) < 1e-9 * np.sum(np.abs(o.rho) ** 2) * o.P_gridDIM


def test_wm_harmonic_rotation_and_conservation():
    o = _coherent_wm()
    o.propagate(200)
    t = o.dt * np.arange(1, 201)
    assert np.max(np.abs(np.array(o.X_average) - 2 * np.cos(t))) < 1e-4
    assert np.max(np.abs(np.array(o.P_average) + 2 * np.sin(t))) < 1e-4
    assert np.max(np.abs(np.array(o.hamiltonian_average) - 2.5)) < 1e-4
    assert abs(o.get_purity() - 1.) < 1e-10
    assert abs(o.wignerfunction.sum().real * o.dXdP - 1.) < 1e-12
    # first Ehrenfest theorem d<x>/dt = <p> to O(dt^2)
    dxdt = np.gradient(np.array(o.X_average), o.dt)
    assert np.max(np.abs(dxdt[1:-1] - np.array(o.X_average_RHS)[1:-1])) < 1e-3


def test_wm_planted_sign_error_fails_rotation():
    """Conjugating the expV phase (a plausible sign slip) must break the rotation pin."""
    class Bad(WignerMoyalOracle):
        def expV_factor(self, t):
            return np.conj(WignerMoyalOracle.expV_factor(self, t))
    o = Bad(X_gridDIM=128, X_amplitude=10., P_gridDIM=128, P_amplitude=10., dt=0.01,
            V=HARM_V, K=HARM_K, diff_V=lambda X, P, t: X, diff_K=lambda X, P, t: P)
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    o.propagate(100)
    t = o.dt * np.arange(1, 101)
    assert np.max(np.abs(np.array(o.P_average) + 2 * np.sin(t))) > 0.1


def test_wm_time_reversibility():
    o = _coherent_wm(ehrenfest=False, Nx=64, Np=64)
    o.V = lambda X, t: 0.5 * X * X + 0.05 * X ** 4
    r0 = o.rho.copy()
    for _ in range(20):
        o.single_step_propagation()
    o.dt = -o.dt
    for _ in range(20):
        o.single_step_propagation()
    assert np.linalg.norm(o.rho - r0) / np.linalg.norm(r0) < 1e-12


def test_wm_masks_absorb():
    """abs_boundary_x/p only ever multiply by <= 1: the trace may only decrease."""
    o = _coherent_wm(ehrenfest=False, Nx=64, Np=64,
                     abs_boundary_x=lambda XX, XXp: np.where(np.abs(XX) > 5., 0.9, 1.) * np.ones_like(XXp))
    n0 = np.sum(np.abs(o.rho) ** 2)
    for _ in range(50):
        o.single_step_propagation()
    assert np.sum(np.abs(o.rho) ** 2) < n0


def test_wb_harmonic_gibbs_exact():
    """wigner_bloch_cuda_1d.py:406-425 with fixed omega = 1, kT = 0.5."""
    omega, kT = 1.0, 0.5
    o = WignerBlochOracle(kT=kT, dbeta=0.01, X_gridDIM=128, X_amplitude=8., P_gridDIM=128,
                          P_amplitude=8., V=lambda X, t: 0.5 * omega ** 2 * X * X, K=HARM_K)
    W = o.get_gibbs_state().real
    H = 0.5 * (o.P ** 2 + omega ** 2 * o.X ** 2)
    th = np.tanh(0.5 * omega / kT)
    exact = th / np.pi * np.exp(-2. * th * H / omega)
    assert np.linalg.norm((exact - W).ravel(), np.inf) * o.dXdP < 1e-6
    Hav = o.get_average(lambda X, P, t: 0.5 * P * P + 0.5 * omega ** 2 * X * X)
    assert abs(Hav - 0.5 * omega / np.tanh(0.5 * omega / kT)) < 1e-5
    assert not o.warnings


def exact_thermal_energy(V, beta, L=10., N=256):
    """<H>(beta) and E0 by exact diagonalisation of the MUB Hamiltonian
    (oracle/band_structure.py, band_structure_1d.py:64-92)."""
    from oracle.band_structure import MUBBandStructure
    b = MUBBandStructure(N, L, lambda x: V(x, 0.), lambda p: 0.5 * p * p)
    E = np.linalg.eigvalsh(b.get_hamiltonian(0.0))
    return b.thermal_energy(beta), float(E[0])


def test_wb_anharmonic_energy_vs_diagonalisation():
    """Config-2 Hamiltonian (V = x^2/2 + 0.1 x^4) on a reduced 128^2 grid, beta 0 -> 2."""
    V = lambda X, t: 0.5 * X * X + 0.1 * X ** 4  # noqa: E731
    o = WignerBlochOracle(kT=0.5, dbeta=0.005, X_gridDIM=128, X_amplitude=8., P_gridDIM=128,
                          P_amplitude=8., V=V, K=HARM_K)
    o.get_gibbs_state()
    Hav = o.get_average(lambda X, P, t: 0.5 * P * P + V(X, t))
    Eex, _ = exact_thermal_energy(V, 2.0)
    assert abs(Hav - Eex) < 1e-4 * abs(Eex)


def test_wb_dbeta_rounding_py2_semantics():
    assert WignerBlochOracle.get_dbeta_num_beta_steps(0.5, 0.005) == (0.005, 400)
    db, n = WignerBlochOracle.get_dbeta_num_beta_steps(0.3, 0.01)   # 333.33 -> 333
    assert n == 333 and abs(db - 1. / (0.3 * 333)) < 1e-15
    db, n = WignerBlochOracle.get_dbeta_num_beta_steps(0.0, 0.01)
    assert n == np.inf


def _pure_pair(N=128, Xa=10., dt=0.01, Vf=None):
    Vf = Vf or (lambda X, t: 0.5 * X * X + 0.05 * X ** 4)
    x0, p0 = 1.5, -0.5
    psi0 = lambda X: np.exp(-(X - x0) ** 2 / 2 + 1j * p0 * X)  # noqa: E731
    sw = SchrodingerWignerOracle(X_gridDIM=N, X_amplitude=Xa, dt=dt, V=Vf, K=HARM_K)
    sw.set_wavefunction(psi0)
    rv = RhoVNeumannOracle(X_gridDIM=N, X_amplitude=Xa, dt=dt, V=Vf, K=HARM_K,
                           diff_V=lambda X, P, t: X + 0.2 * X ** 3, diff_K=lambda X, P, t: P)
    rv.set_rho(lambda X, Xp: psi0(X) * np.conj(psi0(Xp)))
    return sw, rv


def test_rv_pure_state_matches_schrodinger():
    """Cross-class invariant: von Neumann on |psi><psi| = Schrodinger psi(t) outer product
    (time-independent V)."""
    sw, rv = _pure_pair()
    sw.propagate(100)
    rv.propagate(100)
    psi = sw.wavefunction[0]
    R = psi[None, :] * np.conj(psi)[:, None]
    R /= np.trace(R).real * rv.dX
    assert np.linalg.norm(rv.rho - R) / np.linalg.norm(R) < 1e-10
    assert abs(rv.get_purity() - 1.) < 1e-10


def test_rv_planted_missing_sign_fails_crosscheck():
    """Dropping the (-1)^(i+j) factor of expV (a plausible slip) must break the pin.
    (An axis swap is NOT detectable with an even K(P): it maps P to -P.)"""
    class Bad(RhoVNeumannOracle):
        def expV_factor(self, t):
            return self.S * RhoVNeumannOracle.expV_factor(self, t)
    sw, rv = _pure_pair()
    bad = Bad(X_gridDIM=128, X_amplitude=10., dt=0.01, V=rv.V, K=HARM_K)
    bad.set_rho(rv.rho.copy())
    sw.propagate(50)
    bad.propagate(50)
    psi = sw.wavefunction[0]
    R = psi[None, :] * np.conj(psi)[:, None]
    R /= np.trace(R).real * bad.dX
    assert np.linalg.norm(bad.rho - R) / np.linalg.norm(R) > 1e-3


def test_rv_observables_gaussian_moments():
    """<x>, <p>, <p^2> of psi = exp(-(x-x0)^2/2 + i p0 x) / pi^(1/4): x0, p0, p0^2 + 1/2."""
    _, rv = _pure_pair()
    x0, p0 = 1.5, -0.5
    assert abs(rv.get_average((lambda X, P, t: X,)) - x0) < 1e-10
    assert abs(rv.get_average((None, lambda X, P, t: P)) - p0) < 1e-10
    assert abs(rv.get_average((None, lambda X, P, t: P * P)) - (p0 ** 2 + 0.5)) < 1e-10
    assert abs(rv.get_sigma_x_sigma_p() - 0.5) < 1e-9


def test_rv_ehrenfest_lists():
    _, rv = _pure_pair(dt=0.005)
    rv.propagate(200)
    dpdt = np.gradient(np.array(rv.P_average), rv.dt)
    assert np.max(np.abs(dpdt[1:-1] - np.array(rv.P_average_RHS)[1:-1])) < 1e-3
    h = np.array(rv.hamiltonian_average)
    assert (h.max() - h.min()) / abs(h.max()) < 1e-5


def test_rv_padded_wigner_closed_form():
    """Padded Wigner transform of a coherent state = exp(-(x-x0)^2 - (p-p0)^2)/pi."""
    _, rv = _pure_pair(N=128, Xa=10.)
    W = rv.get_wignerfunction().real
    x0, p0 = 1.5, -0.5
    exact = np.exp(-(rv.X_wigner - x0) ** 2 - (rv.P_wigner - p0) ** 2) / np.pi
    assert np.max(np.abs(W - exact)) < 1e-8


def test_rb_harmonic_gibbs():
    """<H> = (w/2) coth(w/2kT), purity = tanh(w/2kT) for the harmonic oscillator."""
    kT = 0.5
    rb = RhoBlochOracle(kT=kT, dbeta=0.005, X_gridDIM=128, X_amplitude=8., V=HARM_V, K=HARM_K)
    assert rb.K_min == 0.0
    rb.get_gibbs_state()
    E = (rb.get_average((None, lambda X, P, t: 0.5 * P * P))
         + rb.get_average((lambda X, P, t: 0.5 * X * X,)))
    assert abs(E - 0.5 / np.tanh(0.5 / kT)) < 1e-5
    assert abs(rb.get_purity() - np.tanh(0.5 / kT)) < 1e-5


def test_sw_stationary_states_harmonic():
    sw = SchrodingerWignerOracle(X_gridDIM=128, X_amplitude=10., dt=0.005, V=HARM_V, K=HARM_K)
    sw.get_stationary_states(3, nsteps=4000)
    for n, psi in enumerate(sw.stationary_states):
        sw.wavefunction = psi.copy()
        assert abs(sw.get_energy() - (n + 0.5)) < 1e-4


def test_sw_coherent_motion_and_wigner():
    sw = SchrodingerWignerOracle(X_gridDIM=128, X_amplitude=10., dt=0.01, V=HARM_V, K=HARM_K,
                                 diff_V=lambda X, P, t: X, diff_K=lambda X, P, t: P)
    sw.set_wavefunction(lambda X: np.exp(-(X - 2.) ** 2 / 2))
    sw.propagate(100)
    t = sw.dt * np.arange(1, 101)
    assert np.max(np.abs(np.array(sw.X_average) - 2 * np.cos(t))) < 1e-4
    W = sw.get_wigner().real
    assert abs(W.sum() * sw.wigner_dXdP - 1) < 1e-12


def test_sign_matrix():
    S = sign_matrix(3, 4)
    assert S[0, 0] == 1 and S[1, 0] == -1 and S[0, 1] == -1 and S[2, 3] == -1


def test_wm_precomputed_stepper_equals_reference_sequence():
    a = _coherent_wm(ehrenfest=False, Nx=64, Np=64)
    b = _coherent_wm(ehrenfest=False, Nx=64, Np=64)
    for _ in range(10):
        a.single_step_propagation()
    b.propagate_precomputed(10)
    assert np.linalg.norm(a.rho - b.rho) / np.linalg.norm(a.rho) < 1e-13
