"""WignerBlochCUDA1D: drop-in MI355X implementation of the reference class
(wigner_bloch_cuda_1d.py:9-366): the Boltzmann-Gibbs state and the ground state by
imaginary-time (Bloch) split-operator propagation in the Wigner-Moyal representation.

Bloch step (wigner_bloch_cuda_1d.py:221-240):  EVb, I0, F1, EKb, F0, I1, EVb, then
normalize_rho_wigner, with EVb = S exp(-dbeta/4 [V(XX_j) + V(XX'_i) - V_min]) and
EKb = exp(-dbeta/2 [K(PP_j) + K(PP'_i) - K_min]) at t_initial.  Both are separable, so
they run as FP64-built real factor vectors inside the same fused passes as the real-time
step: [EVb, I0] [F1, EKb, I1] [F0, EVb].
"""
import numpy as np

from . import graphs
from .device import empty, to_device
from .engine import FWD, INV, finish_sums
from .wigner_moyal_cuda_1d import WignerMoyalCUDA1D

CHECK_CHUNK = 32


def _py2_round(x):
    """Python 2 round() (half away from zero), as get_dbeta_num_beta_steps relies on."""
    return float(np.sign(x) * np.floor(abs(x) + 0.5))


class WignerBlochCUDA1D(WignerMoyalCUDA1D):
    """Gibbs state rho = exp(-H/kT) of H = K(p) + V(x) by Bloch propagation in phase space
    (reference: wigner_bloch_cuda_1d.py:9-18)."""

    # get_gibbs_state stops when the purity is not below this (the reference's check, :147)
    purity_limit = 1.

    def __init__(self, **kwargs):
        """Keywords of WignerMoyalCUDA1D plus kT, dbeta and t_initial
        (wigner_bloch_cuda_1d.py:19-68)."""
        kwargs = dict(kwargs)
        if 't_initial' not in kwargs:
            kwargs.update(t_initial=0.)
            print("Warning: Initial time (t_initial) was not specified. So the default value was used t_initial = 0.")
        self.kT = kwargs.pop('kT', 0.)
        self.dbeta = kwargs.pop('dbeta', 0.01)
        if 'dt' not in kwargs:
            kwargs.update(dt=self.dbeta)
        WignerMoyalCUDA1D.__init__(self, **kwargs)
        self.V_min_x, self.V_min_xp, self.K_min_p, self.K_min_pp = self._separable_minima()
        self.V_min = self.V_min_x + self.V_min_xp
        self.K_min = self.K_min_p + self.K_min_pp
        self.cuda_consts += "#define V_min %.15e \n" % self.V_min
        self.cuda_consts += "#define K_min %.15e \n" % self.K_min
        self._bloch_cache = {}
        self._gibbs_w = {}
        self._gibbs_mw = {}        # weights and factors of the merged Gibbs chunks (or None)
        self._gibbs_bufs = {}      # persistent per-chunk sums and captured chunks of the Gibbs loop
        self.gibbs_stopped_at = None

    def _eval_at_t0(self, which, xdev, n):
        out = empty((1, n), np.float64)
        ts = to_device(np.array([float(self.t_initial)]))
        self._eval_1d(which, xdev, n, ts, 1, 0.0, out)
        return out

    def _separable_minima(self):
        """fill_V / fill_K + gpuarray.min (wigner_bloch_cuda_1d.py:70-92, :343-365):
        min_ij [V(XX_j) + V(XX'_i)] = min_j V(XX_j) + min_i V(XX'_i)."""
        self._V0_x = self._eval_at_t0(0, self._dev_kXX, self.X_gridDIM)
        self._V0_xp = self._eval_at_t0(0, self._dev_kXXp, self.P_gridDIM)
        self._K0_p = self._eval_at_t0(1, self._dev_kPP, self.X_gridDIM)
        self._K0_pp = self._eval_at_t0(1, self._dev_kPPp, self.P_gridDIM)
        return (float(self._V0_x.get().min()), float(self._V0_xp.get().min()),
                float(self._K0_p.get().min()), float(self._K0_pp.get().min()))

    def get_V_min(self):
        return self.V_min

    def get_K_min(self):
        return self.K_min

    def _bloch_factors(self, dbeta):
        """Real factor vectors of one Bloch step: [start col, end col, row] pass vectors."""
        key = float(dbeta)
        f = self._bloch_cache.get(key)
        if f is not None:
            return f
        p = self.plan
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        colL = p.new_factors(Nx, 2)
        colE = p.new_factors(Np, 2)
        rowL = p.new_factors(Np, 1)
        rowE = p.new_factors(Nx, 1)
        q = -0.25 * dbeta
        for v, scale in ((0, 1.0), (1, 1.0 / Np)):
            p.factors(colL, Nx, 1, A=self._V0_x.ptr, ca=q, c0=-q * self.V_min_x, alt_sign=True,
                      mode=1, out_offset=v)
            p.factors(colE, Np, 1, A=self._V0_xp.ptr, ca=q, c0=-q * self.V_min_xp, alt_sign=True,
                      mode=1, scale=scale, out_offset=v)
        h = -0.5 * dbeta
        p.factors(rowL, Np, 1, A=self._K0_pp.ptr, ca=h, c0=-h * self.K_min_pp, mode=1,
                  scale=1.0 / Nx)
        p.factors(rowE, Nx, 1, A=self._K0_p.ptr, ca=h, c0=-h * self.K_min_p, mode=1)
        f = (colL, colE, rowL, rowE)
        self._bloch_cache[key] = f
        return f

    def _bloch_step_passes(self, dbeta):
        colL, colE, rowL, rowE = self._bloch_factors(dbeta)
        csz = self.plan.fsize
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        p, rho = self.plan, self.rho
        p.run(rho, 0, 0, INV, line_fac=colL.ptr, elem_fac=colE.ptr,
              dscale=self._dscale if self._scale_pending else None)
        self._scale_pending = False
        p.run(rho, 1, FWD, INV, line_fac=rowL.ptr, elem_fac=rowE.ptr)
        p.run(rho, 0, FWD, 0, line_fac=colL.ptr + Nx * csz, elem_fac=colE.ptr + Np * csz)

    def bloch_single_step_propagation(self, dbeta):
        """wigner_bloch_cuda_1d.py:221-240 (ends with normalize_rho_wigner)."""
        self._bloch_step_passes(float(dbeta))
        self.normalize_rho_wigner()

    def _normalized_step(self, dbeta, stats_out):
        """One Bloch step + Wigner transform; rho and W normalised lazily by a device-side
        scale.  stats_out receives the 8 statistics of the unnormalised W."""
        self._bloch_step_passes(dbeta)
        self.rho2wigner_()
        st = self.plan.stats(self.wignerfunction, out=stats_out)
        self.plan.scale_from_stats(st.ptr, 8, 0, 0, self.dXdP, self._dscale)
        self._scale_pending = True

    def _finish_normalization(self):
        if self._scale_pending:
            self.plan.scale(self.wignerfunction, self._dscale)
            self._flush_scale()

    @classmethod
    def get_dbeta_num_beta_steps(cls, kT, dbeta):
        """wigner_bloch_cuda_1d.py:94-117."""
        if kT > 0:
            num_beta_steps = 1. / (kT * dbeta)
            if _py2_round(num_beta_steps) != num_beta_steps:
                num_beta_steps = _py2_round(num_beta_steps)
                dbeta = 1. / (kT * num_beta_steps)
            num_beta_steps = int(num_beta_steps)
        else:
            num_beta_steps = np.inf
        return dbeta, num_beta_steps

    def _purities(self, stats_host, n, ncomp=8, comp=3, scale=1.0):
        # purity = 2 pi Re sum(W^2) dXdP with W normalised by n = Re sum W dXdP (:349-354);
        # stats rows hold Re sum W in component 0 and sum W^2 / scale in component comp
        s = stats_host.reshape(n, self.batch, ncomp)
        norm = s[:, :, 0] * self.dXdP
        return 2. * np.pi * scale * s[:, :, comp] * self.dXdP / norm ** 2

    def _gibbs_weights(self, record):
        """Weights of the Gibbs loop's per-step sums, FP64 [1 + m, Np, Nx]: L^H 1 (Re sum W,
        the norm) and L^H f for the m recorded expressions at the current t, with L the Wigner
        transform (the sums are linear in rho).  Cached per expression list and t."""
        key = (tuple(record), float(self.t))
        g = self._gibbs_w.get(key)
        if g is None:
            g = self._gibbs_w[key] = self._obs_weights(["1."] + list(record), t=float(self.t))
        return g

    def _normalized_step_sums(self, dbeta, g, nw, out_ptr):
        """One Bloch step, then one launch for Re sum W, the recorded sums and sum |rho|^2
        (sum W^2 = Np sum |rho|^2: the Wigner transform is Np times an isometry) into out
        (nw + 1 values per item), and the next normalisation scale.  W is not formed."""
        self._bloch_step_passes(dbeta)
        self.plan.weighted_sums(self.rho, g, nw, out_ptr, self.dXdP, self._dscale, norm2=True)
        self._scale_pending = True

    def _gibbs_merge(self, dbeta, record):
        """Weights and factors of merged Gibbs chunks, or None where they do not apply.

        Inside a chunk the closing column pass of step s, [F0, C1], and the opening one of
        step s + 1, [C0, I0], run as one pass [F0, C0 C1, I0]; its values right after the
        factor are v = C0 rho_s, with rho_s the state after step s (C0, C1 real).  So the sums
        of step s are Re <g / C0, v> and sum |v|^2 / C0^2, taken in that pass (aw_pass
        obs_mode 1).  The state is normalised once per chunk (by the last step's norm): the
        purity and the recorded averages are ratios of the sums, and the Bloch factors are at
        most 1, so the unnormalised state only decays within a chunk.  Where C0 is below
        1e-150 (FP64) or 1e-17 (mixed; the merged factor C0^2 / Np then stays a normal FP32
        number above it) the weights are 0: there |rho_s| <= C0 / Np |F0 Y| is below that
        fraction of the state (the potential on the X' grid grows far past V(X_amplitude),
        and exp(-dbeta V / 4) underflows to 0 at its edges)."""
        key = (float(dbeta), tuple(record), float(self.t))
        if key in self._gibbs_mw:
            return self._gibbs_mw[key]
        mw = None
        p = self.plan
        nblk = p.obs_blocks(0)
        if nblk:
            Nx, Np = self.X_gridDIM, self.P_gridDIM
            q = -0.25 * float(dbeta)
            vx = self._V0_x.get().reshape(-1)[:Nx]
            vxp = self._V0_xp.get().reshape(-1)[:Np]
            # C0 in FP64, as the opening factor vectors of _bloch_factors (v = 0)
            cl = np.exp(q * (vx - self.V_min_x)) * (1.0 - 2.0 * (np.arange(Nx) & 1))
            ce = np.exp(q * (vxp - self.V_min_xp)) * (1.0 - 2.0 * (np.arange(Np) & 1))
            c0 = np.outer(ce, cl)
            keep = np.abs(c0) > (1e-17 if self.precision == "mixed" else 1e-150)
            inv = np.where(keep, 1.0 / np.where(keep, c0, 1.0), 0.0)
            g = self._gibbs_weights(record)
            colLm, colEm = p.new_factors(Nx, 1), p.new_factors(Np, 1)
            p.factors(colLm, Nx, 1, A=self._V0_x.ptr, ca=2 * q, c0=-2 * q * self.V_min_x, mode=1)
            p.factors(colEm, Np, 1, A=self._V0_xp.ptr, ca=2 * q, c0=-2 * q * self.V_min_xp, mode=1,
                      scale=1.0 / Np)
            mw = dict(nblk=nblk, nc=len(record) + 2, g=g, colLm=colLm, colEm=colEm,
                      g_mid=to_device(g.get() * inv[None]), n2w=to_device(inv * inv))
        self._gibbs_mw[key] = mw
        return mw

    def _gibbs_merged_chunk(self, dbeta, n, mw, fb):
        """n Bloch steps in 2n + 1 passes; the per-step sums [Re sum W, recorded sums,
        sum |rho|^2] of the unnormalised states go to fb["stats"][:n], and the last step's
        normalisation scale to self._dscale (applied by the next pass that opens a step)."""
        colL, colE, rowL, rowE = self._bloch_factors(dbeta)
        csz = self.plan.fsize
        Nx, Np, B = self.X_gridDIM, self.P_gridDIM, self.batch
        p, rho = self.plan, self.rho
        nblk, nc = mw["nblk"], mw["nc"]
        ring = fb["ring"]
        stride = B * nblk * nc * 8
        p.run(rho, 0, 0, INV, line_fac=colL.ptr, elem_fac=colE.ptr, dscale=self._dscale)
        for s in range(n):
            p.run(rho, 1, FWD, INV, line_fac=rowL.ptr, elem_fac=rowE.ptr)
            part = ring.ptr + s * stride
            if s < n - 1:
                p.run(rho, 0, FWD, INV, line_fac=mw["colLm"].ptr, elem_fac=mw["colEm"].ptr,
                      obs_weights=mw["g_mid"], obs_nw=nc - 1, obs_partials=part, obs_mode=1, obs_norm2=1,
                      obs_norm2_weights=mw["n2w"])
            else:
                p.run(rho, 0, FWD, 0, line_fac=colL.ptr + Nx * csz, elem_fac=colE.ptr + Np * csz,
                      obs_weights=mw["g"], obs_nw=nc - 1, obs_partials=part, obs_mode=1, obs_norm2=1)
        finish_sums(ring, nblk, nc, n * B, fb["stats"])
        finish_sums(ring.ptr + (n - 1) * stride, nblk, nc, B, fb["last"], self.dXdP, self._dscale)
        self._scale_pending = True

    def get_gibbs_state(self, kT=None, dbeta=None, record=None):
        """wigner_bloch_cuda_1d.py:119-155: infinite-temperature start, n Bloch steps with a
        purity < 1 check after each (stop with a warning otherwise).  Steps run in chunks with
        the checks read once per chunk; a violation replays the chunk up to that step.
        ``record`` (not in the reference) may list C-expression observables whose averages
        are recorded after every step into self.gibbs_record (shape [steps, n_obs])."""
        if dbeta is None:
            dbeta = self.dbeta
        if kT is None:
            kT = self.kT
        dbeta, num_beta_steps = self.get_dbeta_num_beta_steps(kT, dbeta)
        if not np.isfinite(num_beta_steps):
            raise ValueError("kT = 0 gives an infinite number of beta steps (use get_ground_state)")
        dbeta = np.float64(dbeta)
        self.set_wignerfunction(1. / np.prod(self.wignerfunction.shape[-2:]))
        self.gibbs_stopped_at = None
        record = list(record) if record else []
        m = len(record)
        fast = self.obs_by_adjoint and m + 2 <= 8
        g = self._gibbs_weights(record) if fast else None
        nc = m + 2 if fast else 8         # per-step values per item in `stats`
        mw = None                         # merged chunks (fast path only)
        if fast:
            # every step applies the pending scale (1 at the start), so that the chunks are
            # alike and a full chunk can be captured once and replayed (fixed addresses)
            if not self._scale_pending:
                self._dscale.tensor[:, 0] = 1.0
                self._dscale.tensor[:, 1] = 0.0
                self._scale_pending = True
            fb = self._gibbs_bufs.get(nc)
            if fb is None:
                fb = self._gibbs_bufs[nc] = dict(stats=empty((CHECK_CHUNK, self.batch, nc), np.float64), graphs={})
                # the first call allocates the plan's scratch (never inside a capture)
                self.plan.weighted_sums(self.rho, g, m + 1, fb["stats"], norm2=True)
            self._bloch_factors(float(dbeta))      # cached factor vectors, made outside any capture
            mw = self._gibbs_merge(dbeta, record)  # likewise
            if mw is not None and "ring" not in fb:
                fb["ring"] = empty((CHECK_CHUNK, self.batch, mw["nblk"], nc), np.float64)
                fb["last"] = empty((self.batch, nc), np.float64)
            gkey = (float(dbeta), tuple(record), float(self.t), mw is not None)
        rec = []
        k = 0
        if fast:
            self._gibbs_fast_chunks(dbeta, num_beta_steps, fb, gkey, g, mw, m, nc, rec)
            if record:
                self.gibbs_record = np.concatenate(rec, axis=0) if rec else np.zeros((0, m))
            return self.wignerfunction
        # the loop that forms W every step (obs_by_adjoint off, or more than 6 recorded sums)
        while k < num_beta_steps:
            n = min(CHECK_CHUNK, num_beta_steps - k)
            saved = self.rho.copy()
            saved_scale = self._scale_pending
            saved_ds = self._dscale.copy()
            stats = empty((n, self.batch, nc), np.float64)
            ring = empty((n, self.batch, m), np.float64) if record else None

            def steps(n_):
                for s in range(n_):
                    self._normalized_step(dbeta, stats[s])
                    if record:
                        self._obs_launch(record, self.wignerfunction, float(self.t), ring.ptr + s * self.batch * m * 8)

            steps(n)
            st = stats.get()[:n]
            pur = self._purities(st, n)
            bad = np.nonzero(~(pur < self.purity_limit))[0]
            if len(bad):
                stop = int(bad[0])
                # replay up to and including the violating step (deterministic)
                self.rho[:] = saved
                self._scale_pending = saved_scale
                self._dscale[:] = saved_ds
                steps(stop + 1)
                self._finish_gibbs(False)
                self.gibbs_stopped_at = k + stop
                print("Warning: Gibbs state calculations interupted because purity = %.10f > 1."
                      "Current kT = %.6f" % (float(np.max(pur[stop])), 1 / (max(k + stop, 1) * dbeta)))
                if record:
                    rec.append(self._gibbs_record_values(False, st, ring, m, stop + 1))
                break
            if record:
                rec.append(self._gibbs_record_values(False, st, ring, m, n))
            k += n
        self._finish_gibbs(False)
        if record:
            self.gibbs_record = np.concatenate(rec, axis=0) if rec else np.zeros((0, m))
        return self.wignerfunction

    def _gibbs_fast_chunks(self, dbeta, num_beta_steps, fb, gkey, g, mw, m, nc, rec):
        """The Gibbs loop on the sums from rho, one chunk ahead: chunk c + 1 is queued
        before the host checks the purities of chunk c, whose sums are copied to pinned host
        memory behind it, so the GPU does not wait for the check.  A violation in chunk c
        restores the state saved before chunk c (the queued chunk is discarded) and replays
        chunk c up to the violating step, as the one-chunk-at-a-time loop does."""
        import torch
        B, Np = self.batch, float(self.P_gridDIM)

        def steps(n_):
            if mw is not None:
                self._gibbs_merged_chunk(dbeta, n_, mw, fb)
            else:
                for s in range(n_):
                    # [Re sum W, recorded sums, sum |rho|^2] and the scale, one launch
                    self._normalized_step_sums(dbeta, g, m + 1, fb["stats"].ptr + s * B * nc * 8)

        def launch(k0):
            n = min(CHECK_CHUNK, num_beta_steps - k0)
            saved, saved_ds = self.rho.copy(), self._dscale.copy()
            graphs.run_chunk(fb["graphs"], gkey, lambda n=n: steps(n), n == CHECK_CHUNK)
            host = torch.empty((n, B, nc), dtype=torch.float64, pin_memory=True)
            host.copy_(fb["stats"].tensor[:n], non_blocking=True)
            done = torch.cuda.Event()
            done.record()
            return k0, n, saved, saved_ds, host, done

        cur = launch(0)
        while cur is not None:
            k, n, saved, saved_ds, host, done = cur
            cur = launch(k + n) if k + n < num_beta_steps else None
            done.synchronize()
            st = host.numpy()
            pur = self._purities(st, n, nc, m + 1, Np)
            bad = np.nonzero(~(pur < self.purity_limit))[0]
            if len(bad):
                stop = int(bad[0])
                # replay up to and including the violating step (deterministic)
                self.rho[:] = saved
                self._scale_pending = True
                self._dscale[:] = saved_ds
                steps(stop + 1)
                self._finish_gibbs(True)
                self.gibbs_stopped_at = k + stop
                print("Warning: Gibbs state calculations interupted because purity = %.10f > 1."
                      "Current kT = %.6f" % (float(np.max(pur[stop])), 1 / (max(k + stop, 1) * dbeta)))
                if m:
                    rec.append(self._gibbs_record_values(True, st, None, m, stop + 1))
                return
            if m:
                rec.append(self._gibbs_record_values(True, st, None, m, n))
        self._finish_gibbs(True)

    def _finish_gibbs(self, fast):
        """Normalise rho and W by the last step's scale (W is formed here when the steps
        did not form it)."""
        if fast:
            self.rho2wigner_()           # applies the pending scale to rho first
        else:
            self._finish_normalization()

    def _gibbs_record_values(self, fast, st, ring, m, n):
        if fast:
            vals = st[:n, :, 1:1 + m] / st[:n, :, 0:1]
            return vals[:, 0, :] if self.batch == 1 else vals
        return self._record_values(ring.get()[:n], st[:n])

    def _record_values(self, ring, stats):
        # average = sum Re W f / Re sum W (the dXdP of W's normalisation cancels)
        vals = ring / stats[:, :, 0:1]
        return vals[:, 0, :] if self.batch == 1 else vals

    def get_ground_state(self, dbeta=None, abs_tol_purity=1e-12, max_steps=None):
        """wigner_bloch_cuda_1d.py:157-219: adaptive imaginary-time descent; a step that
        raises purity above 1, does not lower the energy or violates sigma_x sigma_p >= 1/2
        is rolled back and dbeta halved.  Returns the unnormalised rho2wigner (:217)."""
        if self.batch != 1:
            raise NotImplementedError("get_ground_state is defined for a single state")
        previous_energy = current_energy = np.inf
        previous_purity = current_purity = 0.
        if dbeta is None:
            dbeta = self.dbeta
        dbeta = np.float64(dbeta)
        self.set_wignerfunction(1. / np.prod(self.wignerfunction.shape))
        previous_rho = self.rho.copy()
        exprs = [self.hamiltonian, "X * X", "X", "P * P", "P"]
        stats = empty((1, 8), np.float64)
        sums = empty((1, len(exprs)), np.float64)
        nsteps = 0
        while current_purity < (1. - abs_tol_purity) and dbeta > 1e-11:
            if max_steps is not None and nsteps >= max_steps:
                break
            nsteps += 1
            self.bloch_single_step_propagation(dbeta)
            self.plan.stats(self.wignerfunction, out=stats)
            self._obs_launch(exprs, self.wignerfunction, float(self.t), sums)
            s = stats.get()[0]
            o = sums.get()[0] * self.dXdP
            current_purity = 2. * np.pi * s[3] * self.dXdP
            ok = current_purity <= 1.
            if ok:
                current_energy = o[0]
                ok = current_energy < previous_energy
            if ok:
                ok = np.sqrt((o[1] - o[2] ** 2) * (o[3] - o[4] ** 2)) >= 0.5
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
        self.rho2wigner()
        return self.wignerfunction
