"""NumPy complex128 restatement of the reference split-operator propagators (TEST ORACLE).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``.  Never imported by the product package.

Every operator sequence below follows the reference file:line cited in its docstring.
Conventions (cufft.py:6-7, 233-238, 274-345): ``F0/F1`` are unnormalised forward DFTs
(exp(-2 pi i jk/N)) along axis 0 / axis 1 of a C-order ``[rows, cols]`` array, ``I0/I1``
the unnormalised inverse DFTs (exp(+2 pi i jk/N)).  ``S[i, j] = (-1)**(i + j)``.

The physics that the reference passes as C-expression strings is passed here as NumPy
callables: ``V(X, t)``, ``K(P, t)``, observables ``f(X, P, t)``, Wigner-Moyal masks
``abs_boundary_x(XX, XX_prime)`` / ``abs_boundary_p(PP, PP_prime)`` and von Neumann /
Schrodinger masks ``abs_boundary_x(X)`` / ``abs_boundary_p(P)``.  Coordinates are the
ones the reference kernels recompute from their ``#define`` step sizes
(e.g. ``XX = dXX * (j - 0.5 * X_gridDIM)``, wigner_moyal_cuda_1d.py:769-770).

Pinning (tests/test_oracle.py):
  * FFT conventions: brute-force O(N^2) DFT at N <= 64.
  * WM step: exact harmonic phase-space rotation of a Gaussian (<x> = 2 cos t,
    <p> = -2 sin t, <H> = 2.5 + O(dt^2)); norm and purity conservation; time reversibility
    (step(dt) then step(-dt) = identity for time-independent H); the Ehrenfest identity
    d<x>/dt = <p> to O(dt^2); a planted sign error in expV fails the rotation test.
  * Wigner transforms: round trip rho2wigner_(wigner2rho(W)) = W; Gaussian closed forms
    of the x- and p-marginals; Parseval sum W^2 = Np * sum |rho|^2.
  * WB / RB: exact harmonic Gibbs state W = tanh(w/2kT)/pi exp(-2 tanh(w/2kT) H / w)
    (wigner_bloch_cuda_1d.py:406-409) and <H> = (w/2) coth(w/2kT); exact diagonalisation
    of the anharmonic Hamiltonian for <H>(beta) and E0 (band_structure_1d.py:64-92 pattern).
  * RV: unitary evolution of a pure state equals SchrodingerWigner's psi(t) outer product
    for time-independent V; trace/purity conservation; p-observables vs analytic Gaussian
    moments; padded Wigner transform marginal = diag(rho) closed form.
  * SW: harmonic ground state energy 0.5, excited states n + 0.5 (Gram-Schmidt), norm
    conservation, coherent-state <x>(t) = x0 cos t.
"""
import numpy as np

SQRT2 = np.sqrt(2.0)


def _py2_round(x):
    """Python 2 ``round()`` (half away from zero), used by get_dbeta_num_beta_steps
    (wigner_bloch_cuda_1d.py:107-110)."""
    return float(np.sign(x) * np.floor(abs(x) + 0.5))


class DFT:
    """Unnormalised batched DFT along one axis with cuFFT's sign convention
    (cufft.py:6-7 CUFFT_FORWARD = -1, CUFFT_INVERSE = +1; :233-238).

    ``workers=None`` uses numpy.fft (pocketfft, one core); an int uses scipy.fft with that
    many worker threads (the all-cores CPU baseline)."""

    def __init__(self, workers=None):
        self.workers = workers
        if workers is not None:
            import scipy.fft as sfft
            self._sfft = sfft

    def F(self, a, axis):
        if self.workers is None:
            return np.fft.fft(a, axis=axis)
        return self._sfft.fft(a, axis=axis, workers=self.workers)

    def I(self, a, axis):
        # norm="forward" puts 1/N on the forward transform only, so the inverse is unscaled
        if self.workers is None:
            return np.fft.ifft(a, axis=axis, norm="forward")
        return self._sfft.ifft(a, axis=axis, norm="forward", workers=self.workers)


def _as_grid(val, shape, dtype=np.float64):
    return np.broadcast_to(np.asarray(val, dtype=dtype), shape)


def kconst(x):
    """A float as a kernel sees it through ``#define name %.15e``
    (wigner_moyal_cuda_1d.py:174-176, rho_vneumann_cuda_1d.py and
    schrodinger_wigner_cuda_1d.py build their #define blocks the same way): 16 significant
    digits.  The kernels recompute coordinates from these rounded steps."""
    return float("%.15e" % float(x))


def sign_matrix(rows, cols):
    """(-1)**(i + j) (wigner_moyal_cuda_1d.py:707-715)."""
    si = 1 - 2 * (np.arange(rows) % 2)
    sj = 1 - 2 * (np.arange(cols) % 2)
    return np.outer(si, sj).astype(np.float64)


# ---------------------------------------------------------------------------------------
# Wigner-Moyal (wigner_moyal_cuda_1d.py)
# ---------------------------------------------------------------------------------------
class WignerMoyalOracle:
    """Restatement of WignerMoyalCUDA1D (wigner_moyal_cuda_1d.py:10-828)."""

    def __init__(self, X_gridDIM, X_amplitude, P_gridDIM, P_amplitude, dt, V, K, t=0.0,
                 diff_V=None, diff_K=None, abs_boundary_x=None, abs_boundary_p=None,
                 dft=None):
        # power-of-two checks, wigner_moyal_cuda_1d.py:58-67
        assert 2 ** int(np.log2(X_gridDIM)) == X_gridDIM
        assert 2 ** int(np.log2(P_gridDIM)) == P_gridDIM
        self.X_gridDIM, self.P_gridDIM = int(X_gridDIM), int(P_gridDIM)
        self.X_amplitude, self.P_amplitude = float(X_amplitude), float(P_amplitude)
        self.dt = float(dt)
        self.t = np.float64(t)
        self.V, self.K = V, K
        self.diff_V, self.diff_K = diff_V, diff_K
        self.abs_boundary_x, self.abs_boundary_p = abs_boundary_x, abs_boundary_p
        self.dft = dft if dft is not None else DFT()
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        Xa, Pa = self.X_amplitude, self.P_amplitude

        # grids, wigner_moyal_cuda_1d.py:110-134
        self.dX = 2. * Xa / Nx
        self.dP = 2. * Pa / Np
        self.dXdP = self.dX * self.dP
        X = np.linspace(-Xa, Xa - self.dX, Nx)
        P = np.linspace(-Pa, Pa - self.dP, Np)
        PP_prime = SQRT2 * P
        self.dPP_prime = PP_prime[1] - PP_prime[0]
        XX_prime = (np.arange(Np, dtype=np.float64) - Np // 2) * (np.pi / (SQRT2 * Pa))
        self.dXX_prime = XX_prime[1] - XX_prime[0]
        XX = SQRT2 * X
        self.dXX = XX[1] - XX[0]
        PP = (np.arange(Nx, dtype=np.float64) - Nx // 2) * (np.pi / (SQRT2 * Xa))
        self.dPP = PP[1] - PP[0]
        self.X = X[np.newaxis, :]
        self.P = P[:, np.newaxis]

        # coordinates as the kernels recompute them (:645-646, :662-663, :748-749, :769-770,
        # :797-798, :823-824)
        j = np.arange(Nx, dtype=np.float64)
        i = np.arange(Np, dtype=np.float64)
        self.kXX = kconst(self.dXX) * (j - 0.5 * Nx)
        self.kXXp = kconst(self.dXX_prime) * (i - 0.5 * Np)
        self.kPP = kconst(self.dPP) * (j - 0.5 * Nx)
        self.kPPp = kconst(self.dPP_prime) * (i - 0.5 * Np)
        self.kX = kconst(self.dX) * (j - 0.5 * Nx)
        self.kP = kconst(self.dP) * (i - 0.5 * Np)
        self.S = sign_matrix(Np, Nx)

        self.wignerfunction = np.zeros((Np, Nx), np.complex128)
        self.rho = np.zeros_like(self.wignerfunction)

        # Ehrenfest lists, :303-325
        self.isEhrenfest = diff_K is not None and diff_V is not None
        if self.isEhrenfest:
            self.X_average, self.P_average = [], []
            self.X_average_RHS, self.P_average_RHS = [], []
            self.hamiltonian_average = []

    # -- elementwise kernels ------------------------------------------------------------
    @staticmethod
    def _mask(f, a, b, t):
        """A mask callable f(XX, XX') or f(XX, XX', t): the reference inlines the mask's C
        expression into expV / expK, where the step time t is in scope (:753, :777)."""
        import inspect
        if len(inspect.signature(f).parameters) >= 3:
            return f(a, b, t)
        return f(a, b)

    def _absx(self, t=0.0):
        if self.abs_boundary_x is None:
            return 1.0
        return _as_grid(self._mask(self.abs_boundary_x, self.kXX[None, :], self.kXXp[:, None], t), self.rho.shape)

    def _absp(self, t=0.0):
        if self.abs_boundary_p is None:
            return 1.0
        return _as_grid(self._mask(self.abs_boundary_p, self.kPP[None, :], self.kPPp[:, None], t), self.rho.shape)

    def expV_factor(self, t):
        """expV kernel factor (wigner_moyal_cuda_1d.py:763-778)."""
        tt = t + 0.5 * self.dt
        Vj = _as_grid(self.V(self.kXX, tt), self.kXX.shape)
        Vi = _as_grid(self.V(self.kXXp, tt), self.kXXp.shape)
        phase = -0.5 * self.dt * (Vj[None, :] - Vi[:, None])
        return self.S * np.exp(1j * phase) * self._absx(t)

    def expK_factor(self, t):
        """expK kernel factor (wigner_moyal_cuda_1d.py:742-754)."""
        tt = t + 0.5 * self.dt
        Kj = _as_grid(self.K(self.kPP, tt), self.kPP.shape)
        Ki = _as_grid(self.K(self.kPPp, tt), self.kPPp.shape)
        phase = -self.dt * (Kj[None, :] - Ki[:, None])
        return np.exp(1j * phase) * self._absp(t)

    def shearX_factor(self):
        """phase_shearX (wigner_moyal_cuda_1d.py:639-654)."""
        a = np.tan(np.pi / 8.)
        phase = -a * self.kPP[None, :] * self.kXXp[:, None]
        return np.exp(1j * phase) / float(self.X_gridDIM)

    def shearY_factor(self):
        """phase_shearY (wigner_moyal_cuda_1d.py:656-670)."""
        b = -np.sin(np.pi / 4.)
        phase = -b * self.kPPp[:, None] * self.kXX[None, :]
        return np.exp(1j * phase) / float(self.P_gridDIM)

    def blackmanX_factor(self):
        """blackmanX (wigner_moyal_cuda_1d.py:678-687)."""
        Cj = np.arange(self.X_gridDIM) * 2.0 * np.pi / (self.X_gridDIM - 1.)
        return (0.42 - 0.5 * np.cos(Cj) + 0.08 * np.cos(2.0 * Cj))[None, :]

    def blackmanY_factor(self):
        """blackmanY (wigner_moyal_cuda_1d.py:689-698)."""
        Ci = np.arange(self.P_gridDIM) * 2.0 * np.pi / (self.P_gridDIM - 1.)
        return (0.42 - 0.5 * np.cos(Ci) + 0.08 * np.cos(2.0 * Ci))[:, None]

    # -- transforms ---------------------------------------------------------------------
    def wigner2rho(self):
        """wigner_moyal_cuda_1d.py:422-469."""
        F, I = self.dft.F, self.dft.I
        SX, SY = self.shearX_factor(), self.shearY_factor()
        r = self.S * self.wignerfunction
        r = F(r, 0)
        r = F(SX * I(r, 1), 1)
        r = F(SY * I(r, 0), 0)
        r = F(SX * I(r, 1), 1)
        self.rho = self.S * r
        return self.rho

    def rho2wigner_(self, blackman=False):
        """wigner_moyal_cuda_1d.py:471-501 (``blackman=True``: rho2wigner_blackman, :503-538)."""
        F, I = self.dft.F, self.dft.I
        SX, SY = self.shearX_factor(), self.shearY_factor()
        bx = self.blackmanX_factor() if blackman else 1.0
        by = self.blackmanY_factor() if blackman else 1.0
        w = self.S * self.rho
        w = I(SX * F(bx * w, 1), 1)
        w = I(SY * F(by * w, 0), 0)
        w = I(SX * F(bx * w, 1), 1)
        w = I(by * w, 0)
        self.wignerfunction = self.S * w
        return self.wignerfunction

    rho2wigner = rho2wigner_

    def normalize_rho_wigner(self):
        """wigner_moyal_cuda_1d.py:608-622."""
        self.rho2wigner()
        norm = self.wignerfunction.sum().real * self.dXdP
        self.wignerfunction /= norm
        self.rho /= norm
        return self

    # -- propagation --------------------------------------------------------------------
    def single_step_propagation(self):
        """wigner_moyal_cuda_1d.py:543-561."""
        F, I = self.dft.F, self.dft.I
        EV = self.expV_factor(self.t)
        r = EV * self.rho
        r = F(I(r, 0), 1)
        r = self.expK_factor(self.t) * r
        r = I(F(r, 0), 1)
        r /= r.shape[0] * r.shape[1]
        self.rho = EV * r
        return self.rho

    def propagate_precomputed(self, steps):
        """The single_step_propagation sequence (wigner_moyal_cuda_1d.py:543-561) with the
        expV / expK factors computed once (valid for a time-independent H).  Used as the CPU
        baseline timing loop; the result equals ``steps`` calls of single_step_propagation."""
        F, I = self.dft.F, self.dft.I
        EV = self.expV_factor(self.t) / np.sqrt(self.rho.shape[0] * self.rho.shape[1])
        EK = self.expK_factor(self.t)
        r = self.rho
        for _ in range(steps):
            r = F(I(EV * r, 0), 1)
            r = I(F(EK * r, 0), 1)
            r = EV * r
            self.t += self.dt
        self.rho = r
        return self.rho

    def propagate(self, steps=1):
        """wigner_moyal_cuda_1d.py:563-582 (t advanced before each step, :571)."""
        for _ in range(steps):
            self.t += self.dt
            self.single_step_propagation()
            self.get_Ehrenfest(self.t)
        if not self.isEhrenfest:
            self.normalize_rho_wigner()
        return self.wignerfunction

    def get_Ehrenfest(self, t):
        """wigner_moyal_cuda_1d.py:584-606."""
        if self.isEhrenfest:
            self.normalize_rho_wigner()
            self.X_average.append(self.get_average(lambda X, P, t: X))
            self.X_average_RHS.append(self.get_average(self.diff_K))
            self.P_average.append(self.get_average(lambda X, P, t: P))
            self.P_average_RHS.append(-self.get_average(self.diff_V))
            self.hamiltonian_average.append(self.get_average(
                lambda X, P, t: self.K(P, t) + self.V(X, t)))

    # -- observables --------------------------------------------------------------------
    def get_average(self, f):
        """Sum Re W * f(X, P, t) dXdP (wigner_moyal_cuda_1d.py:329-347, kernel :817-827)."""
        w = _as_grid(f(self.kX[None, :], self.kP[:, None], self.t), self.rho.shape)
        return float(np.sum(self.wignerfunction.real * w) * self.dXdP)

    def get_purity(self):
        """2 pi Re dot(W, W) dXdP, no conjugation (wigner_moyal_cuda_1d.py:349-354)."""
        return 2. * np.pi * np.sum(self.wignerfunction * self.wignerfunction).real * self.dXdP

    def get_sigma_x_sigma_p(self):
        """wigner_moyal_cuda_1d.py:356-366."""
        ga = self.get_average
        return np.sqrt((ga(lambda X, P, t: X * X) - ga(lambda X, P, t: X) ** 2)
                       * (ga(lambda X, P, t: P * P) - ga(lambda X, P, t: P) ** 2))

    def set_wignerfunction(self, new_wigner_func):
        """wigner_moyal_cuda_1d.py:379-420.  A callable is evaluated on the kernel grid
        f(X, P) like the C-string branch (:791-801); W /= (complex) sum W dXdP (:415)."""
        if isinstance(new_wigner_func, np.ndarray):
            assert new_wigner_func.shape == self.wignerfunction.shape
            self.wignerfunction = new_wigner_func.astype(np.complex128).copy()
        elif callable(new_wigner_func):
            self.wignerfunction = _as_grid(
                new_wigner_func(self.kX[None, :], self.kP[:, None]),
                self.rho.shape, np.complex128).copy()
        elif isinstance(new_wigner_func, (float, complex, int)):
            self.wignerfunction = np.full(self.rho.shape, np.complex128(new_wigner_func))
        else:
            raise NotImplementedError("new_wigner_func must be either function or numpy.array")
        self.wignerfunction /= self.wignerfunction.sum() * self.dXdP
        self.wigner2rho()
        return self


# ---------------------------------------------------------------------------------------
# Wigner-Bloch (wigner_bloch_cuda_1d.py)
# ---------------------------------------------------------------------------------------
class WignerBlochOracle(WignerMoyalOracle):
    """Restatement of WignerBlochCUDA1D (wigner_bloch_cuda_1d.py:9-366)."""

    def __init__(self, kT=0., dbeta=0.01, t_initial=0., dt=None, **kw):
        # :28-49
        self.kT = kT
        self.dbeta = dbeta
        self.t_initial = t_initial
        if dt is None:
            dt = dbeta
        WignerMoyalOracle.__init__(self, dt=dt, **kw)
        self.V_min = self.get_V_min()
        self.K_min = self.get_K_min()
        self.warnings = []

    def get_V_min(self):
        """fill_V + gpuarray.min (wigner_bloch_cuda_1d.py:70-80, :343-353)."""
        v = (_as_grid(self.V(self.kXX, self.t_initial), self.kXX.shape)[None, :]
             + _as_grid(self.V(self.kXXp, self.t_initial), self.kXXp.shape)[:, None])
        return float(v.min())

    def get_K_min(self):
        """fill_K + gpuarray.min (wigner_bloch_cuda_1d.py:82-92, :355-365)."""
        k = (_as_grid(self.K(self.kPP, self.t_initial), self.kPP.shape)[None, :]
             + _as_grid(self.K(self.kPPp, self.t_initial), self.kPPp.shape)[:, None])
        return float(k.min())

    def bloch_expK_factor(self, dbeta):
        """bloch_expK (wigner_bloch_cuda_1d.py:276-290); masks disabled (:289)."""
        Kj = _as_grid(self.K(self.kPP, self.t_initial), self.kPP.shape)
        Ki = _as_grid(self.K(self.kPPp, self.t_initial), self.kPPp.shape)
        return np.exp(-0.5 * dbeta * (Kj[None, :] + Ki[:, None] - self.K_min))

    def bloch_expV_factor(self, dbeta):
        """bloch_expV (wigner_bloch_cuda_1d.py:299-316); masks disabled (:315)."""
        Vj = _as_grid(self.V(self.kXX, self.t_initial), self.kXX.shape)
        Vi = _as_grid(self.V(self.kXXp, self.t_initial), self.kXXp.shape)
        return self.S * np.exp(-0.25 * dbeta * (Vj[None, :] + Vi[:, None] - self.V_min))

    def bloch_single_step_propagation(self, dbeta):
        """wigner_bloch_cuda_1d.py:221-240 (no 1/(Nx Np); normalised by the Wigner sum)."""
        F, I = self.dft.F, self.dft.I
        EV = self.bloch_expV_factor(dbeta)
        r = EV * self.rho
        r = F(I(r, 0), 1)
        r = self.bloch_expK_factor(dbeta) * r
        r = I(F(r, 0), 1)
        self.rho = EV * r
        self.normalize_rho_wigner()

    @classmethod
    def get_dbeta_num_beta_steps(cls, kT, dbeta):
        """wigner_bloch_cuda_1d.py:94-117 (kT = 0 gives an infinite step count)."""
        if kT > 0:
            num_beta_steps = 1. / (kT * dbeta)
            if _py2_round(num_beta_steps) != num_beta_steps:
                num_beta_steps = _py2_round(num_beta_steps)
                dbeta = 1. / (kT * num_beta_steps)
            num_beta_steps = int(num_beta_steps)
        else:
            num_beta_steps = np.inf
        return dbeta, num_beta_steps

    def get_gibbs_state(self, kT=None, dbeta=None, callback=None):
        """wigner_bloch_cuda_1d.py:119-155.  ``callback(k)`` (not in the reference) is
        called after every accepted step, e.g. to record <H>(beta)."""
        if dbeta is None:
            dbeta = self.dbeta
        if kT is None:
            kT = self.kT
        dbeta, num_beta_steps = self.get_dbeta_num_beta_steps(kT, dbeta)
        dbeta = np.float64(dbeta)
        self.set_wignerfunction(1. / np.prod(self.wignerfunction.shape))
        k = 0
        for k in range(num_beta_steps):
            self.bloch_single_step_propagation(dbeta)
            current_purity = self.get_purity()
            if not current_purity < 1.:
                self.warnings.append((k, current_purity))
                break
            if callback is not None:
                callback(k)
        return self.wignerfunction

    def get_ground_state(self, dbeta=None, abs_tol_purity=1e-12, max_steps=None):
        """wigner_bloch_cuda_1d.py:157-219 (returns the unnormalised rho2wigner, :217)."""
        previous_energy = current_energy = np.inf
        previous_purity = current_purity = 0.
        if dbeta is None:
            dbeta = self.dbeta
        dbeta = np.float64(dbeta)
        self.set_wignerfunction(1. / np.prod(self.wignerfunction.shape))
        previous_rho = self.rho.copy()
        nsteps = 0
        while current_purity < (1. - abs_tol_purity) and dbeta > 1e-11:
            if max_steps is not None and nsteps >= max_steps:
                break
            nsteps += 1
            self.bloch_single_step_propagation(dbeta)
            current_purity = self.get_purity()
            ok = current_purity <= 1.
            if ok:
                current_energy = self.get_average(lambda X, P, t: self.K(P, t) + self.V(X, t))
                ok = current_energy < previous_energy
            if ok:
                ok = self.get_sigma_x_sigma_p() >= 0.5
            if ok:
                previous_energy = current_energy
                previous_purity = current_purity
                previous_rho = self.rho.copy()
            else:
                self.rho = previous_rho.copy()
                dbeta *= 0.5
                current_energy = previous_energy
                current_purity = previous_purity
        self.ground_state_steps = nsteps
        self.rho2wigner()
        return self.wignerfunction


# ---------------------------------------------------------------------------------------
# von Neumann rho(x, x') (rho_vneumann_cuda_1d.py)
# ---------------------------------------------------------------------------------------
class RhoVNeumannOracle:
    """Restatement of RhoVNeumannCUDA1D (rho_vneumann_cuda_1d.py:15-822)."""

    def __init__(self, X_gridDIM, X_amplitude, dt, V, K, t=0.0, diff_V=None, diff_K=None,
                 abs_boundary_x=None, abs_boundary_p=None, dft=None):
        assert 2 ** int(np.log2(X_gridDIM)) == X_gridDIM
        self.X_gridDIM = N = int(X_gridDIM)
        self.X_amplitude = Xa = float(X_amplitude)
        self.dt = float(dt)
        self.t = np.float64(t)
        self.t_initial = float(t)  # rho_vneumann_cuda_1d.py:94
        self.V, self.K = V, K
        self.diff_V, self.diff_K = diff_V, diff_K
        self.abs_boundary_x, self.abs_boundary_p = abs_boundary_x, abs_boundary_p
        self.dft = dft if dft is not None else DFT()

        # grids, rho_vneumann_cuda_1d.py:105-145
        self.dX = 2. * Xa / N
        self.X = np.linspace(-Xa, Xa - self.dX, N)
        P = (np.arange(N, dtype=np.float64) - N // 2) * (2. * np.pi / (N * self.dX))
        self.dP = P[1] - P[0]
        self.P = P
        self.wigner_padding = N // 2
        self.XX_gridDIM = N + 2 * self.wigner_padding
        M = self.XX_gridDIM
        Xw = (np.arange(M, dtype=np.float64) - M // 2) * (self.dX / SQRT2)
        self.dX_wigner = Xw[1] - Xw[0]
        PPw = (np.arange(M, dtype=np.float64) - M // 2) * (2. * np.pi / (M * self.dX))
        self.dPP = PPw[1] - PPw[0]
        Pw = PPw / SQRT2
        self.dP_wigner = Pw[1] - Pw[0]
        self.dXdP_wigner = self.dX_wigner * self.dP_wigner
        self.X_wigner = Xw[np.newaxis, :]
        self.P_wigner = Pw[:, np.newaxis]

        j = np.arange(N, dtype=np.float64)
        self.kX = kconst(self.dX) * (j - 0.5 * N)   # :760-761
        self.kP = kconst(self.dP) * (j - 0.5 * N)   # :739-740
        self.S = sign_matrix(N, N)
        self.rho = np.zeros((N, N), np.complex128)
        self.wignerfunction = np.zeros((M, M), np.complex128)

        self.isEhrenfest = diff_K is not None and diff_V is not None
        if self.isEhrenfest:
            self.X_average, self.P_average = [], []
            self.X_average_RHS, self.P_average_RHS = [], []
            self.hamiltonian_average = []

    def _bx(self, x):
        if self.abs_boundary_x is None:
            return np.ones_like(x)
        return _as_grid(self.abs_boundary_x(x), x.shape)

    def _bp(self, p):
        if self.abs_boundary_p is None:
            return np.ones_like(p)
        return _as_grid(self.abs_boundary_p(p), p.shape)

    def expV_factor(self, t):
        """expV (rho_vneumann_cuda_1d.py:754-770)."""
        tt = t + 0.5 * self.dt
        v = _as_grid(self.V(self.kX, tt), self.kX.shape)
        phase = -0.5 * self.dt * (v[None, :] - v[:, None])
        b = self._bx(self.kX)
        return self.S * np.exp(1j * phase) * b[None, :] * b[:, None]

    def expK_factor(self, t):
        """expK (rho_vneumann_cuda_1d.py:733-745)."""
        tt = t + 0.5 * self.dt
        k = _as_grid(self.K(self.kP, tt), self.kP.shape)
        phase = -self.dt * (k[None, :] - k[:, None])
        b = self._bp(self.kP)
        return np.exp(1j * phase) * b[None, :] * b[:, None]

    def single_step_propagation(self):
        """rho_vneumann_cuda_1d.py:450-468 (1/N^2 commented out, :464)."""
        F, I = self.dft.F, self.dft.I
        EV = self.expV_factor(self.t)
        r = EV * self.rho
        r = F(I(r, 0), 1)
        r = self.expK_factor(self.t) * r
        r = I(F(r, 0), 1)
        self.rho = EV * r
        return self.rho

    def propagate(self, steps=1):
        """rho_vneumann_cuda_1d.py:429-448 (Re trace renormalisation every step, :443)."""
        for _ in range(steps):
            self.t += self.dt
            self.single_step_propagation()
            self.rho /= np.trace(self.rho).real * self.dX
            self.get_Ehrenfest(self.t)
        return self

    def get_Ehrenfest(self, t):
        """rho_vneumann_cuda_1d.py:470-501."""
        if self.isEhrenfest:
            self.X_average.append(self.get_average((lambda X, P, t: X,)))
            self.X_average_RHS.append(self.get_average((None, self.diff_K)))
            self.P_average.append(self.get_average((None, lambda X, P, t: P)))
            self.P_average_RHS.append(-self.get_average((self.diff_V,)))
            self.hamiltonian_average.append(
                self.get_average((None, lambda X, P, t: self.K(P, t)))
                + self.get_average((lambda X, P, t: self.V(X, t),)))

    def get_average(self, observable):
        """rho_vneumann_cuda_1d.py:520-563 with the observable kernel :811-821 (X and P
        both on the column index j, :817-818) and sign_flip_rho :676-684."""
        F, I = self.dft.F, self.dft.I
        N = self.X_gridDIM
        T = self.rho.copy()
        is_x = False
        for f in observable:
            is_x = not is_x
            if f is None:
                continue
            w = _as_grid(f(self.kX, self.kP, self.t), (N,))[None, :]
            if is_x:
                T = T * w
            else:
                T = F(self.S * T, 1) / N
                T = T * w
                T = self.S * I(T, 1)
        return float(np.trace(T).real * self.dX)

    def get_purity(self):
        """sum |rho|^2 dX^2 (rho_vneumann_cuda_1d.py:565-579)."""
        return float(np.sum(np.abs(self.rho) ** 2) * self.dX ** 2)

    def get_sigma_x_sigma_p(self):
        """rho_vneumann_cuda_1d.py:581-591."""
        ga = self.get_average
        return np.sqrt((ga((lambda X, P, t: X * X,)) - ga((lambda X, P, t: X,)) ** 2)
                       * (ga((None, lambda X, P, t: P * P)) - ga((None, lambda X, P, t: P)) ** 2))

    def set_rho(self, new_rho):
        """rho_vneumann_cuda_1d.py:348-385.  A callable f(X, X_prime) is evaluated on the
        kernel grid (rho[i, j] = f(X_j, X_i), :783-793).  rho /= Re tr(rho) dX (:383)."""
        N = self.X_gridDIM
        if isinstance(new_rho, np.ndarray):
            assert new_rho.shape == self.rho.shape
            self.rho = new_rho.astype(np.complex128).copy()
        elif callable(new_rho):
            self.rho = _as_grid(new_rho(self.kX[None, :], self.kX[:, None]),
                                (N, N), np.complex128).copy()
        elif isinstance(new_rho, (float, complex, int)):
            self.rho = np.full((N, N), np.complex128(new_rho))
        else:
            raise NotImplementedError("new_rho must be either function or numpy.array")
        self.rho /= np.trace(self.rho).real * self.dX
        return self

    def shear_factors(self):
        """phase_shearX / phase_shearY on the padded 2N grid (rho_vneumann_cuda_1d.py:608-639)."""
        M = self.XX_gridDIM
        k = np.arange(M, dtype=np.float64)
        PP = kconst(self.dPP) * (k - 0.5 * M)
        XX = kconst(self.dX) * (k - 0.5 * M)   # dXX = dX (:158)
        a = np.tan(np.pi / 8.)
        SX = np.exp(1j * (-a * PP[None, :] * XX[:, None])) / float(M)
        b = -np.sin(np.pi / 4.)
        SY = np.exp(1j * (-b * PP[:, None] * XX[None, :])) / float(M)
        return SX, SY

    def get_wignerfunction(self):
        """rho_vneumann_cuda_1d.py:387-427."""
        F, I = self.dft.F, self.dft.I
        N, M, pad = self.X_gridDIM, self.XX_gridDIM, self.wigner_padding
        SX, SY = self.shear_factors()
        Sw = sign_matrix(M, M)
        w = np.zeros((M, M), np.complex128)
        w[pad:pad + N, pad:pad + N] = self.rho
        w = Sw * w
        w = I(SX * F(w, 1), 1)
        w = I(SY * F(w, 0), 0)
        w = I(SX * F(w, 1), 1)
        w = Sw * I(w, 0)
        w /= w.sum().real * self.dXdP_wigner
        self.wignerfunction = w
        return self.wignerfunction


class RhoBlochOracle(RhoVNeumannOracle):
    """Restatement of RhoBlochCUDA1D (rho_bloch_cuda_1d.py:10-372)."""

    def __init__(self, kT=0., dbeta=0.01, dt=None, **kw):
        self.kT = kT
        self.dbeta = dbeta
        if dt is None:
            dt = dbeta
        RhoVNeumannOracle.__init__(self, dt=dt, **kw)
        # t_initial is overwritten by the parent with the initial t (rho_vneumann_cuda_1d.py:94)
        self.V_min = self.get_V_min()
        self.K_min = self.get_K_min()
        self.warnings = []

    def get_V_min(self):
        """rho_bloch_cuda_1d.py:70-83, fill_V :342-352."""
        v = _as_grid(self.V(self.kX, self.t_initial), self.kX.shape)
        return float((v[None, :] + v[:, None]).min())

    def get_K_min(self):
        """rho_bloch_cuda_1d.py:85-98: fill_K multiplies into a zeros array (:90, :370), so the
        minimum is 0 for any finite K."""
        k = _as_grid(self.K(self.kP, self.t_initial), self.kP.shape)
        return float((np.zeros((self.X_gridDIM,) * 2) * (k[None, :] + k[:, None])).min())

    def bloch_expK_factor(self, dbeta):
        """rho_bloch_cuda_1d.py:276-289."""
        k = _as_grid(self.K(self.kP, self.t_initial), self.kP.shape)
        return np.exp(-0.5 * dbeta * (k[None, :] + k[:, None] - self.K_min))

    def bloch_expV_factor(self, dbeta):
        """rho_bloch_cuda_1d.py:298-315."""
        v = _as_grid(self.V(self.kX, self.t_initial), self.kX.shape)
        return self.S * np.exp(-0.25 * dbeta * (v[None, :] + v[:, None] - self.V_min))

    def bloch_single_step_propagation(self, dbeta):
        """rho_bloch_cuda_1d.py:225-246 (complex trace normalisation, :244)."""
        F, I = self.dft.F, self.dft.I
        EV = self.bloch_expV_factor(dbeta)
        r = EV * self.rho
        r = F(I(r, 0), 1)
        r = self.bloch_expK_factor(dbeta) * r
        r = I(F(r, 0), 1)
        r = EV * r
        self.rho = r / (np.trace(r) * self.dX)
        return self.rho

    get_dbeta_num_beta_steps = WignerBlochOracle.get_dbeta_num_beta_steps

    def get_gibbs_state(self, kT=None, dbeta=None):
        """rho_bloch_cuda_1d.py:125-161 (identity start, :143)."""
        if dbeta is None:
            dbeta = self.dbeta
        if kT is None:
            kT = self.kT
        dbeta, num_beta_steps = self.get_dbeta_num_beta_steps(kT, dbeta)
        dbeta = np.float64(dbeta)
        self.set_rho(lambda X, Xp: np.asarray(X == Xp, np.float64))
        for k in range(num_beta_steps):
            self.bloch_single_step_propagation(dbeta)
            p = self.get_purity()
            if not p < 1.:
                self.warnings.append((k, p))
                break
        return self.rho

    def get_ground_state(self, dbeta=None, abs_tol_purity=1e-12, max_steps=None):
        """rho_bloch_cuda_1d.py:163-223 (uncertainty check commented out, :198)."""
        previous_energy = current_energy = np.inf
        previous_purity = current_purity = 0.
        if dbeta is None:
            dbeta = self.dbeta
        dbeta = np.float64(dbeta)
        self.set_rho(lambda X, Xp: np.asarray(X == Xp, np.float64))
        previous_rho = self.rho.copy()
        nsteps = 0
        while current_purity < (1. - abs_tol_purity) and dbeta > 1e-12:
            if max_steps is not None and nsteps >= max_steps:
                break
            nsteps += 1
            self.bloch_single_step_propagation(dbeta)
            current_purity = self.get_purity()
            ok = current_purity <= 1.
            if ok:
                current_energy = (self.get_average((None, lambda X, P, t: self.K(P, t)))
                                  + self.get_average((lambda X, P, t: self.V(X, t),)))
                ok = current_energy < previous_energy
            if ok:
                previous_energy = current_energy
                previous_purity = current_purity
                previous_rho = self.rho.copy()
            else:
                self.rho = previous_rho.copy()
                dbeta *= 0.5
                current_energy = previous_energy
                current_purity = previous_purity
        self.ground_state_steps = nsteps
        return self.rho


# ---------------------------------------------------------------------------------------
# Schrodinger psi(x) (schrodinger_wigner_cuda_1d.py)
# ---------------------------------------------------------------------------------------
class SchrodingerWignerOracle:
    """Restatement of SchrodingerWignerCUDA1D (schrodinger_wigner_cuda_1d.py:16-902)."""

    def __init__(self, X_gridDIM, X_amplitude, dt, V, K, t=0.0, diff_V=None, diff_K=None,
                 abs_boundary_x=None, abs_boundary_p=None, dft=None):
        assert 2 ** int(np.log2(X_gridDIM)) == X_gridDIM
        self.X_gridDIM = N = int(X_gridDIM)
        self.X_amplitude = Xa = float(X_amplitude)
        self.dt = float(dt)
        self.t = np.float64(t)
        self.t_initial = float(t)  # schrodinger_wigner_cuda_1d.py:96
        self.V, self.K = V, K
        self.diff_V, self.diff_K = diff_V, diff_K
        self.abs_boundary_x, self.abs_boundary_p = abs_boundary_x, abs_boundary_p
        self.dft = dft if dft is not None else DFT()
        # grids, :107-122
        self.dX = 2. * Xa / N
        self.X = np.linspace(-Xa, Xa - self.dX, N)
        self.P = np.fft.fftshift(np.fft.fftfreq(N, self.dX / (2. * np.pi)))
        self.dP = self.P[1] - self.P[0]
        self.X_wigner = self.X / SQRT2
        self.P_wigner = self.P / SQRT2
        self.wigner_dXdP = self.dX * self.dP
        j = np.arange(N)
        self.kX = kconst(self.dX) * (j - 0.5 * N)                          # :821
        self.kP = kconst(self.dP) * (((j + N // 2) % N) - 0.5 * N)          # :803 (FFT order)
        self.wavefunction = np.zeros((1, N), np.complex128)
        self.isEhrenfest = diff_K is not None and diff_V is not None
        if self.isEhrenfest:
            self.X_average, self.P_average = [], []
            self.X_average_RHS, self.P_average_RHS = [], []
            self.hamiltonian_average = []

    def _norm(self):
        return np.linalg.norm(self.wavefunction)

    def set_wavefunction(self, new_wave_func):
        """:316-354 (psi /= ||psi|| sqrt(dX)).  A callable f(X) is evaluated on the kernel grid."""
        N = self.X_gridDIM
        if isinstance(new_wave_func, np.ndarray):
            assert new_wave_func.shape == self.wavefunction.shape
            self.wavefunction = new_wave_func.astype(np.complex128).copy()
        elif callable(new_wave_func):
            self.wavefunction = _as_grid(new_wave_func(self.kX), (N,),
                                         np.complex128).reshape(1, N).copy()
        elif isinstance(new_wave_func, (float, complex, int)):
            self.wavefunction = np.full((1, N), np.complex128(new_wave_func))
        else:
            raise NotImplementedError("new_wave_func must be either function or numpy.array")
        self.wavefunction /= self._norm() * np.sqrt(self.dX)
        return self

    def expV_factor(self, t):
        """expV (:817-826): evaluated at t, not t + dt/2."""
        v = _as_grid(self.V(self.kX, t), self.kX.shape)
        b = 1.0 if self.abs_boundary_x is None else _as_grid(self.abs_boundary_x(self.kX), self.kX.shape)
        return np.exp(1j * (-0.5 * self.dt * v)) * b

    def expK_factor(self, t):
        """expK (:798-808)."""
        k = _as_grid(self.K(self.kP, t), self.kP.shape)
        b = 1.0 if self.abs_boundary_p is None else _as_grid(self.abs_boundary_p(self.kP), self.kP.shape)
        return np.exp(1j * (-self.dt * k)) * b

    def single_step_propagation(self):
        """:356-369 (no 1/N)."""
        F, I = self.dft.F, self.dft.I
        ev = self.expV_factor(self.t)
        psi = ev * self.wavefunction
        psi = I(self.expK_factor(self.t) * F(psi, 1), 1)
        self.wavefunction = ev * psi
        return self.wavefunction

    def propagate(self, steps=1):
        """:371-390."""
        for _ in range(steps):
            self.t += self.dt
            self.single_step_propagation()
            self.get_Ehrenfest(self.t)
            self.wavefunction /= self._norm() * np.sqrt(self.dX)
        return self

    def get_average(self, f):
        """sum |psi|^2 f(X, P, t) (:392-410, kernel :893-901)."""
        w = _as_grid(f(self.kX, self.kP, self.t), (self.X_gridDIM,))
        return float(np.sum(np.abs(self.wavefunction[0]) ** 2 * w))

    def get_Ehrenfest(self, t):
        """:412-459."""
        if self.isEhrenfest:
            F, I = self.dft.F, self.dft.I
            self.wavefunction /= self._norm()
            self.X_average.append(self.get_average(lambda X, P, t: X))
            self.P_average_RHS.append(-self.get_average(self.diff_V))
            self.hamiltonian_average.append(self.get_average(lambda X, P, t: self.V(X, t)))
            self.wavefunction = F(self.wavefunction, 1)
            self.wavefunction /= self._norm()
            self.X_average_RHS.append(self.get_average(self.diff_K))
            self.P_average.append(self.get_average(lambda X, P, t: P))
            self.hamiltonian_average[-1] += self.get_average(lambda X, P, t: self.K(P, t))
            self.wavefunction = I(self.wavefunction, 1)

    def get_energy(self):
        """:461-482 (leaves psi with 2-norm sqrt(N), like the reference)."""
        F, I = self.dft.F, self.dft.I
        self.wavefunction /= self._norm()
        av_V = self.get_average(lambda X, P, t: self.V(X, t))
        self.wavefunction = F(self.wavefunction, 1)
        self.wavefunction /= self._norm()
        av_K = self.get_average(lambda X, P, t: self.K(P, t))
        self.wavefunction = I(self.wavefunction, 1)
        return av_V + av_K

    def _bloch_step(self):
        """bloch_expV / bloch_expK (:834-854) in the :493-499 sequence."""
        F, I = self.dft.F, self.dft.I
        v = _as_grid(self.V(self.kX, self.t_initial), self.kX.shape)
        k = _as_grid(self.K(self.kP, self.t_initial), self.kP.shape)
        ev = np.exp(-0.5 * self.dt * v)
        ek = np.exp(-self.dt * k)
        psi = ev * self.wavefunction
        psi = I(ek * F(psi, 1), 1)
        self.wavefunction = ev * psi

    def set_ground_state(self, nsteps=10000):
        """:484-503."""
        self.set_wavefunction(1.)
        for _ in range(nsteps):
            self._bloch_step()
            self.wavefunction /= self._norm() * np.sqrt(self.dX)
        return self

    def projectout_stationary_states(self, wavefunction):
        """:505-550 (vdot = sum conj(bra) ket)."""
        wavefunction = wavefunction / np.linalg.norm(wavefunction)
        projs = [np.sum(np.conj(psi) * wavefunction) for psi in self.stationary_states]
        for psi, proj in zip(self.stationary_states, projs):
            wavefunction = wavefunction - proj * psi
        return wavefunction / np.linalg.norm(wavefunction)

    def get_stationary_states(self, nstates, nsteps=10000):
        """:552-591 (even states start from 1, odd from atan(X), :570-574)."""
        self.stationary_states = []
        even = True
        for _ in range(nstates):
            if even:
                self.set_wavefunction(1.)
            else:
                self.set_wavefunction(lambda X: np.arctan(X))
            even = not even
            for _ in range(nsteps):
                self._bloch_step()
                self.wavefunction = self.projectout_stationary_states(self.wavefunction)
            self.stationary_states.append(self.wavefunction.copy())
        return self

    def wigner_shear_factors(self):
        """phase_shearX / phase_shearY of the Schrodinger class (:695-726): kernel P in FFT
        order, X_prime / P_prime on the row index."""
        N = self.X_gridDIM
        a = np.tan(np.pi / 8.)
        b = -np.sin(np.pi / 4.)
        SX = np.exp(1j * (-a * self.kP[None, :] * self.kX[:, None])) / float(N)   # :695-710
        SY = np.exp(1j * (-b * self.kP[:, None] * self.kX[None, :])) / float(N)   # :712-726
        return SX, SY

    def blackman_factors(self):
        """blackmanX (along j) and blackmanY (along i) of the Schrodinger class (:728-750); both
        use X_gridDIM (the grid is square)."""
        c = np.arange(self.X_gridDIM) * 2.0 * np.pi / (self.X_gridDIM - 1.)
        w = 0.42 - 0.5 * np.cos(c) + 0.08 * np.cos(2.0 * c)
        return w[None, :], w[:, None]

    def get_wigner(self, blackman=False):
        """:593-625 (``blackman=True``: get_wigner_blackman, :627-663), with the missing array
        and plans created (the reference never allocates self.wignerfunction nor
        plan_Z2Z_ax0/1, :244, :255-256).  W[i, j] is the Wigner function at
        (P_wigner[i], X_wigner[j]), normalised so that sum W * wigner_dXdP = 1 with
        wigner_dXdP = dX dP, twice the cell area of that grid."""
        F, I = self.dft.F, self.dft.I
        N = self.X_gridDIM
        psi = self.wavefunction[0]
        w = psi[None, :] * np.conj(psi)[:, None]                  # psi2rho :680-687
        SX, SY = self.wigner_shear_factors()
        bx, by = self.blackman_factors() if blackman else (1.0, 1.0)
        w = I(SX * F(bx * w, 1), 1)
        w = I(SY * F(by * w, 0), 0)
        w = I(SX * F(bx * w, 1), 1)
        srow = (1 - 2 * (np.arange(N) % 2)).astype(np.float64)[:, None]           # :763-771
        w = srow * I(srow * (by * w), 0)
        w /= w.sum().real * self.wigner_dXdP
        self.wignerfunction = w
        return w
