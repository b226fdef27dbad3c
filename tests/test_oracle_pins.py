"""Pins for the oracle functions that the propagation pins do not reach: the Schrodinger
Wigner function (closed form of a coherent state), the ground states of the two Bloch
classes (harmonic closed forms, exact diagonalisation) and the Blackman-windowed Wigner
transforms (brute force with explicit DFT matrices).  Each pin is checked to fail on a
plausible planted mistake.  CPU only.
"""
import numpy as np
import pytest

from oracle.awig_oracle import (RhoBlochOracle, SchrodingerWignerOracle, WignerBlochOracle,
                                WignerMoyalOracle, sign_matrix)

HARM_V = lambda X, t: 0.5 * X * X  # noqa: E731
HARM_K = lambda P, t: 0.5 * P * P  # noqa: E731


# ---------------------------------------------------------------------------------------
# Schrodinger Wigner function (schrodinger_wigner_cuda_1d.py:593-663)
# ---------------------------------------------------------------------------------------
def _sw(N, Xa, cls=SchrodingerWignerOracle):
    return cls(X_gridDIM=N, X_amplitude=Xa, dt=0.01, V=HARM_V, K=HARM_K)


def _coherent_psi(o, x0, p0):
    return (np.pi ** -0.25 * np.exp(-(o.kX - x0) ** 2 / 2 + 1j * p0 * o.kX)).reshape(1, -1)


def _sw_closed_form_err(o, x0, p0):
    """max |W - exp(-(x - x0)^2 - (p - p0)^2) / (2 pi)| on (X_wigner[j], P_wigner[i]).  The
    factor 1/2 against the Wigner density is the reference's normalisation: it divides by
    sum W * dX dP (:622-623), and dX dP is twice the cell area dX_wigner dP_wigner."""
    o.set_wavefunction(_coherent_psi(o, x0, p0))
    W = o.get_wigner()
    exact = np.exp(-(o.X_wigner[None, :] - x0) ** 2 - (o.P_wigner[:, None] - p0) ** 2) / (2 * np.pi)
    return np.max(np.abs(W - exact))


@pytest.mark.parametrize("N,Xa,x0,p0", [(256, 12., 1.3, -0.7), (512, 20., -2., 1.5), (128, 10., 0.5, 2.)])
def test_sw_wigner_coherent_state_closed_form(N, Xa, x0, p0):
    o = _sw(N, Xa)
    assert _sw_closed_form_err(o, x0, p0) < 1e-11
    # the moments of W are those of the state: <x> = x0, <p> = p0, variances 1/2
    W = o.wignerfunction.real
    xg, pg = o.X_wigner[None, :], o.P_wigner[:, None]
    s = W.sum()
    assert abs((W * xg).sum() / s - x0) < 1e-10 and abs((W * pg).sum() / s - p0) < 1e-10
    assert abs((W * xg ** 2).sum() / s - x0 ** 2 - 0.5) < 1e-10


@pytest.mark.parametrize("plant", ["shear_x_sign", "shear_y_sign", "natural_order_p"])
def test_sw_wigner_planted_mistakes_fail_closed_form(plant):
    """A flipped shear sign or a P grid taken in natural instead of FFT order must break
    the closed-form pin."""
    class Bad(SchrodingerWignerOracle):
        def wigner_shear_factors(self):
            SX, SY = SchrodingerWignerOracle.wigner_shear_factors(self)
            if plant == "shear_x_sign":
                return np.conj(SX), SY
            if plant == "shear_y_sign":
                return SX, np.conj(SY)
            kP = self.kP
            self.kP = np.fft.fftshift(kP)   # natural order
            try:
                return SchrodingerWignerOracle.wigner_shear_factors(self)
            finally:
                self.kP = kP
    assert _sw_closed_form_err(_sw(256, 12., Bad), 1.3, -0.7) > 1e-2


def _dft_mats(n):
    k = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(k, k) / n)
    return F, np.conj(F)


def _sw_wigner_blackman_brute(o):
    """get_wigner_blackman (schrodinger_wigner_cuda_1d.py:627-663) with explicit DFT
    matrices: F1 A = A F^T, F0 A = F A (cufft.py:6-9 signs, unnormalised)."""
    N = o.X_gridDIM
    F, I = _dft_mats(N)
    psi = o.wavefunction[0]
    A = np.outer(np.conj(psi), psi)                  # psi2rho: rho[i, j] = psi_j conj(psi_i)
    c = np.arange(N) * 2.0 * np.pi / (N - 1.)
    win = 0.42 - 0.5 * np.cos(c) + 0.08 * np.cos(2.0 * c)
    P = o.kP                                         # FFT-ordered kernel momentum (:703, :719)
    X = o.kX
    SX = np.exp(-1j * np.tan(np.pi / 8.) * np.outer(X, P)) / N          # [i: X_prime, j: P]
    SY = np.exp(1j * np.sin(np.pi / 4.) * np.outer(P, X)) / N           # [i: P_prime, j: X]
    s = (1 - 2 * (np.arange(N) % 2)).astype(float)
    A = ((A * win[None, :]) @ F.T * SX) @ I.T        # blackmanX, F1, shearX, I1
    A = I @ (SY * (F @ (A * win[:, None])))          # blackmanY, F0, shearY, I0
    A = ((A * win[None, :]) @ F.T * SX) @ I.T        # blackmanX, F1, shearX, I1
    A = s[:, None] * (I @ (s[:, None] * (A * win[:, None])))   # blackmanY, sign, I0, sign
    return A / (A.sum().real * o.wigner_dXdP)


def test_sw_wigner_blackman_brute_force():
    o = _sw(16, 4.)
    rng = np.random.default_rng(5)
    o.set_wavefunction((rng.normal(size=16) + 1j * rng.normal(size=16)).reshape(1, -1))
    W = o.get_wigner(blackman=True)
    assert np.max(np.abs(W - _sw_wigner_blackman_brute(o))) < 1e-12 * np.max(np.abs(W))
    # the windows are numpy's Blackman window, and without them the plain transform results
    bx, by = o.blackman_factors()
    np.testing.assert_allclose(bx[0], np.blackman(16), rtol=0, atol=1e-15)
    np.testing.assert_allclose(by[:, 0], np.blackman(16), rtol=0, atol=1e-15)
    assert np.max(np.abs(W - o.get_wigner())) > 1e-3 * np.max(np.abs(W))


# ---------------------------------------------------------------------------------------
# Wigner-Moyal Blackman transform (wigner_moyal_cuda_1d.py:503-538, :678-698)
# ---------------------------------------------------------------------------------------
def _wm_blackman_brute(o, rho):
    """rho2wigner_blackman with explicit DFT matrices, transcribed from :503-538."""
    Np, Nx = rho.shape
    Fx, Ix = _dft_mats(Nx)
    Fp, Ip = _dft_mats(Np)
    S = sign_matrix(Np, Nx)
    cj = np.arange(Nx) * 2.0 * np.pi / (Nx - 1.)
    ci = np.arange(Np) * 2.0 * np.pi / (Np - 1.)
    bx = 0.42 - 0.5 * np.cos(cj) + 0.08 * np.cos(2.0 * cj)
    by = 0.42 - 0.5 * np.cos(ci) + 0.08 * np.cos(2.0 * ci)
    # shear phases from the kernel text (:639-670): X shear over (XX'_i, PP_j), Y over (PP'_i, XX_j)
    SX = np.exp(-1j * np.tan(np.pi / 8.) * np.outer(o.kXXp, o.kPP)) / Nx
    SY = np.exp(1j * np.sin(np.pi / 4.) * np.outer(o.kPPp, o.kXX)) / Np
    A = S * rho
    A = ((A * bx[None, :]) @ Fx.T * SX) @ Ix.T
    A = Ip @ (SY * (Fp @ (A * by[:, None])))
    A = ((A * bx[None, :]) @ Fx.T * SX) @ Ix.T
    A = Ip @ (A * by[:, None])
    return S * A


def _wm(Nx, Np, cls=WignerMoyalOracle):
    return cls(X_gridDIM=Nx, X_amplitude=5., P_gridDIM=Np, P_amplitude=4., dt=0.01, V=HARM_V, K=HARM_K)


@pytest.mark.parametrize("Nx,Np", [(16, 16), (16, 8), (8, 16)])
def test_wm_rho2wigner_blackman_brute_force(Nx, Np):
    o = _wm(Nx, Np)
    rng = np.random.default_rng(Nx + 3 * Np)
    rho = rng.normal(size=(Np, Nx)) + 1j * rng.normal(size=(Np, Nx))
    o.rho = rho.copy()
    W = o.rho2wigner_(blackman=True)
    B = _wm_blackman_brute(o, rho)
    assert np.max(np.abs(W - B)) < 1e-12 * np.max(np.abs(B))
    np.testing.assert_allclose(o.blackmanX_factor()[0], np.blackman(Nx), rtol=0, atol=1e-15)
    np.testing.assert_allclose(o.blackmanY_factor()[:, 0], np.blackman(Np), rtol=0, atol=1e-15)


@pytest.mark.parametrize("plant", ["periodic_window", "window_after_fft"])
def test_wm_blackman_planted_mistakes_fail(plant):
    """A periodic window (2 pi j / N instead of / (N - 1)), or windows applied after instead
    of before the forward FFT, must fail the brute-force pin."""
    class Bad(WignerMoyalOracle):
        def rho2wigner_(self, blackman=False):
            F, I = self.dft.F, self.dft.I
            SX, SY = self.shearX_factor(), self.shearY_factor()
            bx, by = self.blackmanX_factor(), self.blackmanY_factor()
            if plant == "periodic_window":
                cj = np.arange(self.X_gridDIM) * 2.0 * np.pi / self.X_gridDIM
                bx = (0.42 - 0.5 * np.cos(cj) + 0.08 * np.cos(2.0 * cj))[None, :]
            w = self.S * self.rho
            if plant == "window_after_fft":
                w = I(SX * bx * F(w, 1), 1)
            else:
                w = I(SX * F(bx * w, 1), 1)
            w = I(SY * F(by * w, 0), 0)
            w = I(SX * F(bx * w, 1), 1)
            w = I(by * w, 0)
            self.wignerfunction = self.S * w
            return self.wignerfunction
    o = _wm(16, 16, Bad)
    rng = np.random.default_rng(1)
    rho = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    o.rho = rho.copy()
    assert np.max(np.abs(o.rho2wigner_() - _wm_blackman_brute(o, rho))) > 1e-3


def test_wm_blackman_equals_plain_transform_without_windows():
    """With both windows replaced by ones the Blackman sequence is rho2wigner_ itself."""
    class Flat(WignerMoyalOracle):
        def blackmanX_factor(self):
            return np.ones((1, self.X_gridDIM))

        def blackmanY_factor(self):
            return np.ones((self.P_gridDIM, 1))
    o = _wm(32, 16, Flat)
    rng = np.random.default_rng(2)
    o.rho = rng.normal(size=(16, 32)) + 1j * rng.normal(size=(16, 32))
    a = o.rho2wigner_(blackman=True).copy()
    assert np.max(np.abs(a - o.rho2wigner_())) < 1e-12 * np.max(np.abs(a))


# ---------------------------------------------------------------------------------------
# Bloch ground states (wigner_bloch_cuda_1d.py:157-219, rho_bloch_cuda_1d.py:163-223)
# ---------------------------------------------------------------------------------------
def _wb_ground(dbeta, cls=WignerBlochOracle, V=HARM_V, N=64, L=6.):
    o = cls(X_gridDIM=N, X_amplitude=L, P_gridDIM=N, P_amplitude=L, dbeta=dbeta, V=V, K=HARM_K)
    W = o.get_ground_state()
    o.wignerfunction = W / (W.sum().real * o.dXdP)     # get_ground_state returns W unnormalised
    return o


def test_wb_ground_state_harmonic_closed_form():
    """E0 = 1/2, W -> exp(-x^2 - p^2) / pi and purity -> 1; the splitting error of W is
    O(dbeta^2) and that of E0 O(dbeta^4) (E is stationary at the ground state)."""
    errs = []
    for db in (0.05, 0.025):
        o = _wb_ground(db)
        E = o.get_average(lambda X, P, t: 0.5 * P * P + 0.5 * X * X)
        exact = np.exp(-o.kX[None, :] ** 2 - o.kP[:, None] ** 2) / np.pi
        errs.append((abs(E - 0.5), np.max(np.abs(o.wignerfunction - exact))))
        assert abs(o.get_purity() - 1.) < 1e-10
        assert o.get_sigma_x_sigma_p() >= 0.5 - 1e-9
    (e1, w1), (e2, w2) = errs
    assert e1 < 1e-8 and w1 < 2e-5
    assert 3. < w1 / w2 < 5.             # second order in dbeta
    assert e2 < e1 / 8.


def test_wb_ground_state_anharmonic_vs_diagonalisation():
    from oracle.band_structure import MUBBandStructure
    V = lambda X, t: 0.5 * X * X + 0.1 * X ** 4  # noqa: E731
    o = _wb_ground(0.02, V=V, N=64, L=6.)
    E = o.get_average(lambda X, P, t: 0.5 * P * P + V(X, t))
    b = MUBBandStructure(256, 10., lambda x: V(x, 0.), lambda p: 0.5 * p * p)
    E0 = float(np.linalg.eigvalsh(b.get_hamiltonian(0.0))[0])
    assert abs(E - E0) < 1e-6


def test_wb_ground_state_planted_mistake_fails():
    """Full instead of half potential kicks (a plausible slip in bloch_expV) converge to the
    ground state of a stiffer oscillator and fail the E0 = 1/2 pin."""
    class Bad(WignerBlochOracle):
        def bloch_expV_factor(self, dbeta):
            return WignerBlochOracle.bloch_expV_factor(self, dbeta) ** 2
    o = _wb_ground(0.05, cls=Bad)
    E = o.get_average(lambda X, P, t: 0.5 * P * P + 0.5 * X * X)
    assert abs(E - 0.5) > 1e-2


def _rb_ground(dbeta, cls=RhoBlochOracle, V=HARM_V):
    o = cls(X_gridDIM=64, X_amplitude=6., dbeta=dbeta, V=V, K=HARM_K)
    o.get_ground_state()
    return o


def test_rb_ground_state_harmonic_closed_form():
    """rho -> psi0 psi0^T with psi0 = pi^(-1/4) exp(-x^2/2), E0 = 1/2, purity -> 1."""
    errs = []
    for db in (0.05, 0.025):
        o = _rb_ground(db)
        x = o.X.ravel()
        psi0 = np.pi ** -0.25 * np.exp(-x ** 2 / 2)
        E = o.get_average((None, lambda X, P, t: 0.5 * P * P)) + o.get_average((lambda X, P, t: 0.5 * X * X,))
        errs.append((abs(E - 0.5), np.max(np.abs(o.rho - np.outer(psi0, psi0)))))
        assert abs(o.get_purity() - 1.) < 1e-10
        assert abs(np.trace(o.rho).real * o.dX - 1.) < 1e-12
    (e1, r1), (e2, r2) = errs
    assert e1 < 1e-8 and r1 < 5e-5
    assert 3. < r1 / r2 < 5.
    assert e2 < e1 / 8.


def test_rb_ground_state_planted_mistake_fails():
    class Bad(RhoBlochOracle):
        def bloch_expV_factor(self, dbeta):
            return RhoBlochOracle.bloch_expV_factor(self, dbeta) ** 2
    o = _rb_ground(0.05, cls=Bad)
    E = o.get_average((None, lambda X, P, t: 0.5 * P * P)) + o.get_average((lambda X, P, t: 0.5 * X * X,))
    assert abs(E - 0.5) > 1e-2
