"""RhoBlochCUDA1D: drop-in MI355X implementation of the reference class
(rho_bloch_cuda_1d.py:10-372): Gibbs and ground states of rho(x, x') by Bloch propagation.

Bloch step (rho_bloch_cuda_1d.py:225-246): EVb, I0, F1, EKb, F0, I1, EVb, then
rho /= tr(rho) dX with the complex trace (:244).  EVb = S exp(-dbeta/4 [V(X_j) + V(X_i) -
V_min]), EKb = exp(-dbeta/2 [K(P_j) + K(P_i) - K_min]) at t_initial, with K_min = 0 as the
reference computes it (fill_K multiplies into a zero array, :90, :370).
"""
import numpy as np

from .device import empty, to_device
from .engine import FWD, INV
from .rho_vneumann_cuda_1d import RhoVNeumannCUDA1D
from .wigner_bloch_cuda_1d import WignerBlochCUDA1D

CHECK_CHUNK = 32


class RhoBlochCUDA1D(RhoVNeumannCUDA1D):
    """Boltzmann-Gibbs state rho = exp(-H/kT) by the Bloch equation in the coordinate
    representation (reference: rho_bloch_cuda_1d.py:10-19)."""

    def __init__(self, **kwargs):
        """Keywords of RhoVNeumannCUDA1D plus kT, dbeta, t_initial (rho_bloch_cuda_1d.py:20-68).
        As in the reference, the parent overwrites t_initial with the initial t."""
        kwargs = dict(kwargs)
        if 't_initial' not in kwargs:
            kwargs.update(t_initial=0.)
            print("Warning: Initial time (t_initial) was not specified. So the default value was used t_initial = 0.")
        self.kT = kwargs.pop('kT', 0.)
        self.dbeta = kwargs.pop('dbeta', 0.01)
        if 'dt' not in kwargs:
            kwargs.update(dt=self.dbeta)
        RhoVNeumannCUDA1D.__init__(self, **kwargs)
        N = self.X_gridDIM
        out = empty((1, N), np.float64)
        ts = to_device(np.array([float(self.t_initial)]))
        self._eval_1d(0, self._dev_kX, N, ts, 1, 0.0, out)
        self._V0 = out
        outk = empty((1, N), np.float64)
        self._eval_1d(1, self._dev_kP, N, ts, 1, 0.0, outk)
        self._K0 = outk
        self._Vmin_1d = float(self._V0.get().min())
        self.V_min = 2.0 * self._Vmin_1d        # min_ij V(X_j) + V(X_i) (:70-83)
        self.K_min = 0.0                         # :85-98 with fill_K's '*=' into zeros
        self.cuda_consts += "#define V_min %.15e\n" % self.V_min
        self.cuda_consts += "#define K_min %.15e\n" % self.K_min
        self._bloch_cache = {}
        self.gibbs_stopped_at = None

    def get_V_min(self):
        return self.V_min

    def get_K_min(self):
        return self.K_min

    get_dbeta_num_beta_steps = WignerBlochCUDA1D.__dict__["get_dbeta_num_beta_steps"]

    def _bloch_factors(self, dbeta):
        key = float(dbeta)
        f = self._bloch_cache.get(key)
        if f is None:
            p, N = self.plan, self.X_gridDIM
            col = p.new_factors(N, 3)    # 0: opening/closing line+elem (same vector), 1: /N copy
            row = p.new_factors(N, 2)    # 0: line (/N), 1: elem
            q = -0.25 * dbeta
            p.factors(col, N, 1, A=self._V0.ptr, ca=q, c0=-q * self._Vmin_1d, alt_sign=True, mode=1)
            p.factors(col, N, 1, A=self._V0.ptr, ca=q, c0=-q * self._Vmin_1d, alt_sign=True, mode=1,
                      scale=1.0 / N, out_offset=1)
            h = -0.5 * dbeta
            # K_min = 0 split as 0 + 0 between the two factors
            p.factors(row, N, 1, A=self._K0.ptr, ca=h, mode=1, scale=1.0 / N)
            p.factors(row, N, 1, A=self._K0.ptr, ca=h, mode=1, out_offset=1)
            f = (col, row)
            self._bloch_cache[key] = f
        return f

    def _bloch_passes(self, dbeta):
        col, row = self._bloch_factors(dbeta)
        csz = self.plan.fsize
        N, p, rho = self.X_gridDIM, self.plan, self.rho
        p.run(rho, 0, 0, INV, line_fac=col.ptr, elem_fac=col.ptr,
              dscale=self._dscale if self._scale_pending else None)
        self._scale_pending = False
        p.run(rho, 1, FWD, INV, line_fac=row.ptr, elem_fac=row.ptr + N * csz)
        p.run(rho, 0, FWD, 0, line_fac=col.ptr, elem_fac=col.ptr + N * csz)

    def bloch_single_step_propagation(self, dbeta):
        """rho_bloch_cuda_1d.py:225-246 (complex trace normalisation)."""
        self._bloch_passes(float(dbeta))
        self._trace_normalize(1)
        return self.rho

    def _purity_from_stats(self, s):
        # purity of rho / (tr rho dX): sum |rho|^2 dX^2 / |tr rho dX|^2 (:565-579)
        tr2 = s[..., 5] ** 2 + s[..., 6] ** 2
        return s[..., 2] / tr2

    def get_gibbs_state(self, kT=None, dbeta=None):
        """rho_bloch_cuda_1d.py:125-161: identity start (set_rho("double(i == j)"), :143),
        n Bloch steps, purity < 1 checked after each (read once per chunk; a violation
        replays the chunk up to that step)."""
        if dbeta is None:
            dbeta = self.dbeta
        if kT is None:
            kT = self.kT
        dbeta, num_beta_steps = self.get_dbeta_num_beta_steps(kT, dbeta)
        if not np.isfinite(num_beta_steps):
            raise ValueError("kT = 0 gives an infinite number of beta steps (use get_ground_state)")
        dbeta = np.float64(dbeta)
        self.set_rho("double(i == j)")
        self.gibbs_stopped_at = None
        k = 0
        while k < num_beta_steps:
            n = min(CHECK_CHUNK, num_beta_steps - k)
            saved = self.rho.copy()
            saved_pending = self._scale_pending
            saved_ds = self._dscale.copy()
            stats = empty((n, 8), np.float64)
            for s in range(n):
                self._bloch_passes(dbeta)
                st = self.plan.stats(self.rho, out=stats[s])
                self.plan.scale_from_stats(st.ptr, 8, 5, 1, self.dX, self._dscale)
                self._scale_pending = True
            pur = self._purity_from_stats(stats.get())
            bad = np.nonzero(~(pur < 1.))[0]
            if len(bad):
                stop = int(bad[0])
                self.rho[:] = saved
                self._scale_pending = saved_pending
                self._dscale[:] = saved_ds
                for s in range(stop + 1):
                    self._bloch_passes(dbeta)
                    st = self.plan.stats(self.rho, out=stats[s])
                    self.plan.scale_from_stats(st.ptr, 8, 5, 1, self.dX, self._dscale)
                    self._scale_pending = True
                self._flush_scale()
                self.gibbs_stopped_at = k + stop
                print("Warning: Gibbs state calculations interupted because purity = %.10f > 1."
                      "Current kT = %.6f" % (pur[stop], 1 / (max(k + stop, 1) * dbeta)))
                break
            k += n
        self._flush_scale()
        return self.rho

    def get_ground_state(self, dbeta=None, abs_tol_purity=1e-12, max_steps=None):
        """rho_bloch_cuda_1d.py:163-223 (the uncertainty check is commented out there, :198)."""
        previous_energy = current_energy = np.inf
        previous_purity = current_purity = 0.
        if dbeta is None:
            dbeta = self.dbeta
        dbeta = np.float64(dbeta)
        self.set_rho("double(i == j)")
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
                current_energy = self.get_average((None, self.K)) + self.get_average((self.V,))
                ok = current_energy < previous_energy
            if ok:
                previous_energy = current_energy
                previous_purity = current_purity
                previous_rho[:] = self.rho
                print("Current energy: %.5f; purity: 1 - %.2e; dbeta: %.2e"
                      % (current_energy, 1 - current_purity, dbeta))
            else:
                self.rho[:] = previous_rho
                dbeta *= 0.5
                current_energy = previous_energy
                current_purity = previous_purity
        self.ground_state_steps = nsteps
        return self.rho
