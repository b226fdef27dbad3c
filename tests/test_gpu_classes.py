"""GPU parity of the Wigner-Bloch, von Neumann, rho-Bloch and Schrodinger-Wigner drop-ins
(and the cufft.py drop-in) against the NumPy oracle."""
import numpy as np
import pytest

from oracle.awig_oracle import (RhoBlochOracle, RhoVNeumannOracle, SchrodingerWignerOracle,
                                WignerBlochOracle)

pytestmark = pytest.mark.gpu

HARM_K = lambda P, t: 0.5 * P * P  # noqa: E731


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def test_cufft_dropin_matches_numpy():
    import torch  # noqa: F401
    from accuratewigner_amd import cufft
    from accuratewigner_amd.device import to_device
    rng = np.random.default_rng(3)
    a = rng.normal(size=(128, 256)) + 1j * rng.normal(size=(128, 256))
    x = to_device(a)
    p0 = cufft.Plan_Z2Z_2D_Axis0(a.shape)
    p1 = cufft.Plan_Z2Z_2D_Axis1(a.shape)
    cufft.fft_Z2Z(x, x, p0)
    assert rel(x.get(), np.fft.fft(a, axis=0)) < 1e-14
    cufft.ifft_Z2Z(x, x, p1)
    assert rel(x.get(), np.fft.ifft(np.fft.fft(a, axis=0), axis=1, norm="forward")) < 1e-14
    with pytest.raises(cufft.cufftInvalidValue):
        cufft.fft_Z2Z(to_device(a[:, :128].copy()), x, p0)


def test_wigner_bloch_gibbs_config2_reduced_matches_oracle():
    """Config 2 potential (V = x^2/2 + 0.1 x^4), x,p in [-8, 8), beta 0 -> 2 in dbeta 0.005,
    at 128^2 (the oracle's cost), renormalised every step; <H>(beta) recorded every step."""
    from accuratewigner_amd import WignerBlochCUDA1D
    V = "0.5 * X * X + 0.1 * X * X * X * X"
    Vf = lambda X, t: 0.5 * X * X + 0.1 * X ** 4  # noqa: E731
    kw = dict(X_gridDIM=128, X_amplitude=8., P_gridDIM=128, P_amplitude=8., kT=0.5, dbeta=0.005)
    g = WignerBlochCUDA1D(V=V, K="0.5 * P * P", t=0., **kw)
    W = g.get_gibbs_state(record=[g.hamiltonian]).get()
    o = WignerBlochOracle(V=Vf, K=HARM_K, **kw)
    Ho = []
    o.get_gibbs_state(callback=lambda k: Ho.append(o.get_average(lambda X, P, t: 0.5 * P * P + Vf(X, t))))
    assert rel(W, o.wignerfunction) < 1e-10
    assert g.gibbs_record.shape == (400, 1)
    assert np.max(np.abs(g.gibbs_record[:, 0] - np.array(Ho))) < 1e-10
    assert abs(g.V_min - o.V_min) < 1e-12 and abs(g.K_min - o.K_min) < 1e-12
    assert abs(g.get_purity() - o.get_purity()) < 1e-11


def test_wigner_bloch_mixed_gibbs_error():
    from accuratewigner_amd import WignerBlochCUDA1D
    kw = dict(X_gridDIM=128, X_amplitude=8., P_gridDIM=128, P_amplitude=8., kT=0.5, dbeta=0.005,
              V="0.5 * X * X + 0.1 * X * X * X * X", K="0.5 * P * P", t=0.)
    g = WignerBlochCUDA1D(precision="mixed", **kw)
    W = g.get_gibbs_state(record=[g.hamiltonian]).get()
    g64 = WignerBlochCUDA1D(**kw)
    W64 = g64.get_gibbs_state(record=[g64.hamiltonian]).get()
    assert rel(W.astype(np.complex128), W64) < 1e-5
    assert np.max(np.abs(g.gibbs_record - g64.gibbs_record) / np.abs(g64.gibbs_record)) < 5e-5


def test_wigner_bloch_ground_state_matches_oracle():
    from accuratewigner_amd import WignerBlochCUDA1D
    kw = dict(X_gridDIM=64, X_amplitude=6., P_gridDIM=64, P_amplitude=6., dbeta=0.05)
    g = WignerBlochCUDA1D(V="0.5 * X * X", K="0.5 * P * P", t=0., **kw)
    W = g.get_ground_state(abs_tol_purity=1e-6, max_steps=150).get()
    o = WignerBlochOracle(V=lambda X, t: 0.5 * X * X, K=HARM_K, **kw)
    Wo = o.get_ground_state(abs_tol_purity=1e-6, max_steps=150)
    assert g.ground_state_steps == o.ground_state_steps
    assert rel(W, Wo) < 1e-9


def _rv_pair(N=128, Xa=10., dt=0.01, masks=False, tdep=False, precision="fp64", ehrenfest=True):
    from accuratewigner_amd import RhoVNeumannCUDA1D
    V = "0.5 * X * X + 0.05 * X * X * X * X"
    Vf = lambda X, t: 0.5 * X * X + 0.05 * X ** 4  # noqa: E731
    if tdep:
        V = "-1. / sqrt(X * X + 1.) + X * 0.05 * pow(sin(M_PI * t / 4.), 2) * cos(0.06 * t)"
        Vf = lambda X, t: -1. / np.sqrt(X * X + 1.) + X * 0.05 * np.sin(np.pi * t / 4.) ** 2 * np.cos(0.06 * t)  # noqa: E731
    extra, oextra = {}, {}
    if ehrenfest:
        dV = "X + 0.2 * X * X * X" if not tdep else "X / pow(X * X + 1., 1.5) + 0.05 * pow(sin(M_PI * t / 4.), 2) * cos(0.06 * t)"
        dVf = (lambda X, P, t: X + 0.2 * X ** 3) if not tdep else \
            (lambda X, P, t: X / (X * X + 1.) ** 1.5 + 0.05 * np.sin(np.pi * t / 4.) ** 2 * np.cos(0.06 * t))
        extra = dict(diff_V=dV, diff_K="P")
        oextra = dict(diff_V=dVf, diff_K=lambda X, P, t: P)
    if masks:
        extra["abs_boundary_x"] = "pow(abs(cos(0.5 * M_PI * X / X_max)), 0.05)"
        Xm = Xa - 2 * Xa / N
        oextra["abs_boundary_x"] = lambda X: np.abs(np.cos(0.5 * np.pi * X / Xm)) ** 0.05
    g = RhoVNeumannCUDA1D(X_gridDIM=N, X_amplitude=Xa, dt=dt, t=0., V=V, K="0.5 * P * P",
                          precision=precision, **extra)
    o = RhoVNeumannOracle(X_gridDIM=N, X_amplitude=Xa, dt=dt, V=Vf, K=HARM_K, **oextra)
    init = "exp(-pow(X - 1.5, 2) / 2. + cuda_complex(0., -0.5) * X) * exp(-pow(X_prime - 1.5, 2) / 2. - cuda_complex(0., -0.5) * X_prime)"
    g.set_rho(init)
    psi0 = lambda X: np.exp(-(X - 1.5) ** 2 / 2 - 0.5j * X)  # noqa: E731
    o.set_rho(lambda X, Xp: psi0(X) * np.conj(psi0(Xp)))
    return g, o


def test_rv_propagate_and_ehrenfest_match_oracle():
    g, o = _rv_pair()
    assert rel(g.rho.get(), o.rho) < 1e-13
    g.propagate(150)
    o.propagate(150)
    assert rel(g.rho.get(), o.rho) < 1e-10
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
        assert a.shape == (150,)
        assert np.max(np.abs(a - b)) < 1e-10, name


def test_rv_time_dependent_masks_no_ehrenfest():
    g, o = _rv_pair(masks=True, tdep=True, ehrenfest=False)
    g.propagate(300)
    o.propagate(300)
    assert rel(g.rho.get(), o.rho) < 1e-10


def test_rv_observables_purity_wigner():
    g, o = _rv_pair(ehrenfest=False)
    g.propagate(40)
    o.propagate(40)
    for obs, oobs in [(("X",), (lambda X, P, t: X,)), ((None, "P * P"), (None, lambda X, P, t: P * P)),
                      (("X", "P", "X"), (lambda X, P, t: X, lambda X, P, t: P, lambda X, P, t: X))]:
        assert abs(g.get_average(obs) - o.get_average(oobs)) < 1e-10
    assert abs(g.get_purity() - o.get_purity()) < 1e-12
    assert abs(g.get_sigma_x_sigma_p() - o.get_sigma_x_sigma_p()) < 1e-10
    W = g.get_wignerfunction().get()
    assert rel(W, o.get_wignerfunction()) < 1e-10


def test_rho_bloch_gibbs_matches_oracle():
    from accuratewigner_amd import RhoBlochCUDA1D
    kw = dict(X_gridDIM=128, X_amplitude=8., kT=0.5, dbeta=0.005)
    g = RhoBlochCUDA1D(V="0.5 * X * X + 0.1 * X * X * X * X", K="0.5 * P * P", t=0., **kw)
    r = g.get_gibbs_state().get()
    o = RhoBlochOracle(V=lambda X, t: 0.5 * X * X + 0.1 * X ** 4, K=HARM_K, **kw)
    o.get_gibbs_state()
    assert g.K_min == 0.0 and abs(g.V_min - o.V_min) < 1e-12
    assert rel(r, o.rho) < 1e-10
    assert abs(g.get_purity() - o.get_purity()) < 1e-11


def test_rho_bloch_ground_state_matches_oracle():
    from accuratewigner_amd import RhoBlochCUDA1D
    kw = dict(X_gridDIM=64, X_amplitude=6., dbeta=0.05)
    g = RhoBlochCUDA1D(V="0.5 * X * X", K="0.5 * P * P", t=0., **kw)
    r = g.get_ground_state(abs_tol_purity=1e-6, max_steps=120).get()
    o = RhoBlochOracle(V=lambda X, t: 0.5 * X * X, K=HARM_K, **kw)
    o.get_ground_state(abs_tol_purity=1e-6, max_steps=120)
    assert g.ground_state_steps == o.ground_state_steps
    assert rel(r, o.rho) < 1e-9


def _sw_pair(N=256, ehrenfest=True, masks=False):
    from accuratewigner_amd import SchrodingerWignerCUDA1D
    V = "0.5 * X * X + 0.05 * X * X * X * X + 0.1 * X * sin(t)"
    Vf = lambda X, t: 0.5 * X * X + 0.05 * X ** 4 + 0.1 * X * np.sin(t)  # noqa: E731
    extra, oextra = {}, {}
    if ehrenfest:
        extra = dict(diff_V="X + 0.2 * X * X * X + 0.1 * sin(t)", diff_K="P")
        oextra = dict(diff_V=lambda X, P, t: X + 0.2 * X ** 3 + 0.1 * np.sin(t), diff_K=lambda X, P, t: P)
    if masks:
        extra["abs_boundary_x"] = "pow(abs(cos(0.5 * M_PI * X / 10.)), 0.01)"
        oextra["abs_boundary_x"] = lambda X: np.abs(np.cos(0.5 * np.pi * X / 10.)) ** 0.01
    g = SchrodingerWignerCUDA1D(X_gridDIM=N, X_amplitude=10., dt=0.01, t=0., V=V, K="0.5 * P * P", **extra)
    o = SchrodingerWignerOracle(X_gridDIM=N, X_amplitude=10., dt=0.01, V=Vf, K=HARM_K, **oextra)
    g.set_wavefunction("exp(-pow(X - 2., 2) / 2. + cuda_complex(0., 0.3) * X)")
    o.set_wavefunction(lambda X: np.exp(-(X - 2.) ** 2 / 2 + 0.3j * X))
    return g, o


def test_sw_propagate_ehrenfest_match_oracle():
    g, o = _sw_pair()
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-14
    g.propagate(200)
    o.propagate(200)
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
        assert np.max(np.abs(a - b)) < 1e-10, name


def test_sw_masks_energy_wigner():
    g, o = _sw_pair(ehrenfest=False, masks=True)
    g.propagate(300)
    o.propagate(300)
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    assert abs(g.get_energy() - o.get_energy()) < 1e-10
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    assert rel(g.get_wigner().get(), o.get_wigner()) < 1e-10


def test_sw_ground_and_stationary_states():
    from accuratewigner_amd import SchrodingerWignerCUDA1D
    g = SchrodingerWignerCUDA1D(X_gridDIM=128, X_amplitude=10., dt=0.005, t=0., V="0.5 * X * X", K="0.5 * P * P")
    o = SchrodingerWignerOracle(X_gridDIM=128, X_amplitude=10., dt=0.005, V=lambda X, t: 0.5 * X * X, K=HARM_K)
    g.set_ground_state(nsteps=1000)
    o.set_ground_state(nsteps=1000)
    assert rel(g.wavefunction.get(), o.wavefunction) < 1e-10
    g.get_stationary_states(3, nsteps=600)
    o.get_stationary_states(3, nsteps=600)
    for a, b in zip(g.stationary_states, o.stationary_states):
        assert rel(a.get(), b) < 1e-9


def test_sw_wigner_blackman_and_closed_form():
    """get_wigner against the coherent-state closed form exp(-(x-x0)^2 - (p-p0)^2) / (2 pi)
    on (X_wigner, P_wigner), and get_wigner_blackman (schrodinger_wigner_cuda_1d.py:627-663)
    against the oracle, FP64 and mixed."""
    from accuratewigner_amd import SchrodingerWignerCUDA1D
    x0, p0 = 1.3, -0.7
    for prec, tol in (("fp64", 1e-12), ("mixed", 2e-6), ("c64", 5e-7)):
        g = SchrodingerWignerCUDA1D(X_gridDIM=256, X_amplitude=12., dt=0.01, t=0., V="0.5 * X * X",
                                    K="0.5 * P * P", precision=prec)
        o = SchrodingerWignerOracle(X_gridDIM=256, X_amplitude=12., dt=0.01, V=lambda X, t: 0.5 * X * X, K=HARM_K)
        g.set_wavefunction("exp(-pow(X - 1.3, 2) / 2.) * cuda_complex(cos(-0.7 * X), sin(-0.7 * X))")
        W = g.get_wigner().get().astype(np.complex128)
        exact = np.exp(-(g.X_wigner[None, :] - x0) ** 2 - (g.P_wigner[:, None] - p0) ** 2) / (2 * np.pi)
        assert np.max(np.abs(W - exact)) < {"fp64": 1e-11, "mixed": 1e-6, "c64": 1e-7}[prec]
        rng = np.random.default_rng(9)
        psi = (rng.normal(size=256) + 1j * rng.normal(size=256)) * np.exp(-(g.X / 6.) ** 2)
        g.set_wavefunction(psi.reshape(1, -1).astype(g.state_dtype))
        o.set_wavefunction(psi.reshape(1, -1))
        Wb = g.get_wigner_blackman().get().astype(np.complex128)
        assert rel(Wb, o.get_wigner(blackman=True)) < tol
        assert rel(g.get_wigner().get().astype(np.complex128), o.get_wigner()) < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed", "c64"])
def test_wigner_bloch_gibbs_config2_full_size(precision):
    """Config 2 at its own size: V = x^2/2 + 0.1 x^4, 512^2, x, p in [-8, 8), beta 0 -> 2 in
    d beta = 0.005 (400 steps), <H>(beta) every step.  Here the merged chunks' FP64 weights
    are cut off where the opening Bloch factor C0 = exp(-d beta V / 4 ...) underflows below
    1e-150 (|XX'| reaches 71 on this grid).  FP64: W and <H>(beta) within 1e-10 of the
    oracle; mixed: within 1e-5 / 5e-5 of FP64."""
    import os
    from accuratewigner_amd import WignerBlochCUDA1D
    from oracle.awig_oracle import DFT
    V = "0.5 * X * X + 0.1 * X * X * X * X"
    Vf = lambda X, t: 0.5 * X * X + 0.1 * X ** 4  # noqa: E731
    kw = dict(X_gridDIM=512, X_amplitude=8., P_gridDIM=512, P_amplitude=8., kT=0.5, dbeta=0.005)
    g = WignerBlochCUDA1D(V=V, K="0.5 * P * P", t=0., precision=precision, **kw)
    W = g.get_gibbs_state(record=[g.hamiltonian]).get().astype(np.complex128)
    assert g.gibbs_record.shape == (400, 1)
    if precision == "fp64":
        o = WignerBlochOracle(V=Vf, K=HARM_K, dft=DFT(min(os.cpu_count() or 1, 16)), **kw)
        Ho = []
        o.get_gibbs_state(callback=lambda k: Ho.append(o.get_average(lambda X, P, t: 0.5 * P * P + Vf(X, t))))
        assert not o.warnings
        assert rel(W, o.wignerfunction) < 1e-10
        assert np.max(np.abs(g.gibbs_record[:, 0] - np.array(Ho))) < 1e-10
        assert abs(g.get_purity() - o.get_purity()) < 1e-11
    else:
        g64 = WignerBlochCUDA1D(V=V, K="0.5 * P * P", t=0., **kw)
        W64 = g64.get_gibbs_state(record=[g64.hamiltonian]).get()
        assert rel(W, W64) < 1e-5
        assert np.max(np.abs(g.gibbs_record - g64.gibbs_record) / np.abs(g64.gibbs_record)) < 5e-5
