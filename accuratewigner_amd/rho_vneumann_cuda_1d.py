"""RhoVNeumannCUDA1D: drop-in MI355X implementation of the reference class
(rho_vneumann_cuda_1d.py:15-822): split-operator propagation of the von Neumann equation
for rho(x, x') with absorbing boundaries, observables and the padded Wigner transform.

Step (rho_vneumann_cuda_1d.py:450-468): EV, I0, F1, EK, F0, I1, EV with
EV = S exp(-0.5 i dt [V(X_j) - V(X_i)]) bx(X_j) bx(X_i) and
EK = exp(-i dt [K(P_j) - K(P_i)]) bp(P_j) bp(P_i), both evaluated at t + dt/2, run as fused
passes [F0, EV EV', I0] [F1, EK, I1]; the trace renormalisation after every step (:443) is a
positive real scalar, applied lazily on the device.
"""
import ctypes
from types import FunctionType

import numpy as np

from . import graphs, rtc
from ._params import KeywordParams
from .device import DeviceArray, current_stream, empty, require_gpu, to_device
from .engine import FWD, INV, Plan, finish_sums, mem_info

SQRT2 = np.sqrt(2.0)
CHUNK_STEPS = 256


class RhoVNeumannCUDA1D(KeywordParams):
    """Second-order split-operator propagator of the von Neumann equation for rho(x, x', t)
    with H = K(p, t) + V(x, t) (reference: rho_vneumann_cuda_1d.py:15-21), on MI355X."""

    # Ehrenfest steps as merged passes whose column pass records the trace and averages of
    # the closed step (two passes a step); False keeps three passes and a reduction a step
    merged_ehrenfest = True

    def __init__(self, **kwargs):
        """Keywords as rho_vneumann_cuda_1d.py:22-47 (X_gridDIM, X_amplitude, V, K, dt, and
        optionally t, consts, functions, diff_V, diff_K, abs_boundary_x, abs_boundary_p,
        max_thread_block), plus ``precision`` ("fp64" | "mixed")."""
        kwargs = dict(kwargs)
        self.precision = kwargs.pop("precision", "fp64")
        self.batch = 1
        self._set_params(kwargs)
        if not hasattr(self, "X_gridDIM"):
            raise AttributeError("Coordinate grid size (X_gridDIM) was not specified")
        assert 2 ** int(np.log2(self.X_gridDIM)) == self.X_gridDIM, \
            "Coordinate grid size (X_gridDIM) must be a power of two"
        if not hasattr(self, "X_amplitude"):
            raise AttributeError("Coordinate grid range (X_amplitude) was not specified")
        if not hasattr(self, "V"):
            raise AttributeError("Potential energy (V) was not specified")
        if not hasattr(self, "K"):
            raise AttributeError("Momentum dependence (K) was not specified")
        if not hasattr(self, "dt"):
            raise AttributeError("Time-step (dt) was not specified")
        if hasattr(self, "t"):
            kwargs.pop("t", None)
        else:
            print("Warning: Initial time (t) was not specified, thus it is set to zero.")
            self.t = 0.
        # the initial time becomes t_initial (:94)
        kwargs.update(t_initial=self.t)
        self.t_initial = self.t
        self.t = np.float64(self.t)

        N = self.X_gridDIM
        # grids (:105-145)
        self.dX = 2. * self.X_amplitude / N
        self.X = np.linspace(-self.X_amplitude, self.X_amplitude - self.dX, N)
        self.P = (np.arange(N, dtype=np.float64) - N // 2) * (2. * np.pi / (N * self.dX))
        self.dP = self.P[1] - self.P[0]
        self.wigner_padding = N // 2
        self.XX_gridDIM = N + 2 * self.wigner_padding
        M = self.XX_gridDIM
        self.X_wigner = (np.arange(M, dtype=np.float64) - M // 2) * (self.dX / SQRT2)
        self.dX_wigner = self.X_wigner[1] - self.X_wigner[0]
        PP = (np.arange(M, dtype=np.float64) - M // 2) * (2. * np.pi / (M * self.dX))
        self.dPP = PP[1] - PP[0]
        self.P_wigner = PP / SQRT2
        self.dP_wigner = self.P_wigner[1] - self.P_wigner[0]
        self.dXdP_wigner = self.dX_wigner * self.dP_wigner
        self.X_wigner = self.X_wigner[np.newaxis, :]
        self.P_wigner = self.P_wigner[:, np.newaxis]
        j = np.arange(N, dtype=np.float64)
        # kernel coordinates from the #define'd steps (X = dX * (j - 0.5 * X_gridDIM))
        self._kX = rtc.kconst(self.dX) * (j - 0.5 * N)
        self._kP = rtc.kconst(self.dP) * (j - 0.5 * N)

        kwargs.update(dX=self.dX, dP=self.dP, XX_gridDIM=self.XX_gridDIM, dXX=self.dX,
                      dPP=self.dPP, P_max=self.P.max(), X_max=self.X.max())
        self.cuda_consts = rtc.cuda_consts(kwargs, getattr(self, "consts", None), None)
        if not hasattr(self, "abs_boundary_x"):
            self.abs_boundary_x = "1."
        if not hasattr(self, "abs_boundary_p"):
            self.abs_boundary_p = "1."
        if not hasattr(self, "max_thread_block"):
            self.max_thread_block = 512
        self.cuda_consts += getattr(self, "functions", "") or ""

        require_gpu()
        self.plan = Plan(N, N, 1, self.precision)
        self.state_dtype = self.plan.state_dtype
        self._wplan = None
        self._cplan = None      # 1 x N FP64 plan for momentum-diagonal transforms
        self._ehr = None        # persistent chunk buffers and captured graphs of the Ehrenfest loop
        self._mod = rtc.Module(self._physics_source(), "rho_vneumann_physics.hip")
        self._dev_kX = to_device(self._kX)
        self._dev_kP = to_device(self._kP)
        self.rho = DeviceArray((N, N), self.state_dtype)
        self.wignerfunction = None   # allocated on first get_wignerfunction ((2N)^2)
        self._tmp = None
        self._dscale = empty((1, 2), np.float64)
        self._scale_pending = False
        self._pass_events = None
        self._obs_cache = {}
        self._prepare_masks()

        self.isEhrenfest = hasattr(self, "diff_K") and hasattr(self, "diff_V")
        if self.isEhrenfest:
            self.X_average, self.P_average = [], []
            self.X_average_RHS, self.P_average_RHS = [], []
            self.hamiltonian_average = []
            self.wigner_time = []
        self.print_memory_info()

    # ------------------------------------------------------------------------------------
    def _state_t(self):
        return "double2" if self.state_dtype == np.complex128 else "float2"

    def _physics_source(self):
        return (rtc.PRELUDE + self.cuda_consts
                + "\n__device__ double K(double P, double t)\n{\n    return (%s);\n}\n" % self.K
                + "\n__device__ double V(double X, double t)\n{\n    return (%s);\n}\n" % self.V
                + "\n__device__ double abs_boundary_x(double X)\n{\n    return (%s);\n}\n" % self.abs_boundary_x
                + "\n__device__ double abs_boundary_p(double P)\n{\n    return (%s);\n}\n" % self.abs_boundary_p
                + rtc.EVAL_1D + r"""
extern "C" __global__ void aw_eval_masks(const double *X, const double *P, int n, double *bx, double *bp)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        bx[k] = abs_boundary_x(X[k]);
        bp[k] = abs_boundary_p(P[k]);
    }
}
""")

    def _eval_1d(self, which, xdev, n, ts_dev, nt, tshift, out):
        blocks, threads = rtc.launch_cfg(n * nt)
        self._mod.launch("aw_eval_1d", (blocks,), (threads,),
                         [ctypes.c_int(which), xdev.ptr, ctypes.c_int(n), ts_dev.ptr,
                          ctypes.c_int(nt), ctypes.c_double(tshift), out.ptr], current_stream())

    def _prepare_masks(self):
        """abs_boundary_x(X) and abs_boundary_p(P) are one-coordinate functions (:714-724), so
        their products bx(X_j) bx(X_i), bp(P_j) bp(P_i) are outer products of two vectors."""
        N = self.X_gridDIM
        self._bx = self._bp = None
        if self.abs_boundary_x.strip() in ("1.", "1", "1.0") and \
                self.abs_boundary_p.strip() in ("1.", "1", "1.0"):
            return
        bx = empty((N,), np.float64)
        bp = empty((N,), np.float64)
        blocks, threads = rtc.launch_cfg(N)
        self._mod.launch("aw_eval_masks", (blocks,), (threads,),
                         [self._dev_kX.ptr, self._dev_kP.ptr, ctypes.c_int(N), bx.ptr, bp.ptr],
                         current_stream())
        if self.abs_boundary_x.strip() not in ("1.", "1", "1.0"):
            self._bx = bx
        if self.abs_boundary_p.strip() not in ("1.", "1", "1.0"):
            self._bp = bp

    @classmethod
    def print_memory_info(cls):
        """rho_vneumann_cuda_1d.py:337-346."""
        free, total = mem_info()
        print("\n\n\t\tGPU memory Total %.2f GB\n\t\tGPU memory Free %.2f GB\n"
              % (total / 2. ** 30, free / 2. ** 30))

    # ------------------------------------------------------------------------------------
    # state
    # ------------------------------------------------------------------------------------
    def _trace_normalize(self, mode=0):
        """rho /= (Re) tr(rho) dX on the device (:383, :443; mode 1: complex trace)."""
        st = self.plan.stats(self.rho)
        self.plan.scale_from_stats(st.ptr, 8, 5, mode, self.dX, self._dscale)
        self.plan.scale(self.rho, self._dscale)
        self._scale_pending = False

    def _flush_scale(self):
        if self._scale_pending:
            self.plan.scale(self.rho, self._dscale)
            self._scale_pending = False

    def set_rho(self, new_rho):
        """rho_vneumann_cuda_1d.py:348-385: array, GPU array, Python function F(self, x, x'),
        C string (X, X_prime, i, j declared) or number; then rho /= Re tr(rho) dX."""
        N = self.X_gridDIM
        self._scale_pending = False
        if isinstance(new_rho, (np.ndarray, DeviceArray)):
            assert new_rho.shape == self.rho.shape, "The grid sizes does not match"
            self.rho[:] = new_rho
        elif isinstance(new_rho, FunctionType):
            self.rho[:] = np.asarray(new_rho(self, self.X[:, np.newaxis], self.X[np.newaxis, :]),
                                     dtype=np.complex128) * np.ones((N, N))
        elif isinstance(new_rho, str):
            src = rtc.PRELUDE + self.cuda_consts + r"""
extern "C" __global__ void aw_init(%(st)s *rho, long long n_item)
{
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n_item;
         idx += (long long)gridDim.x * blockDim.x) {
        const size_t i = idx / X_gridDIM, j = idx %% X_gridDIM;
        const double X = dX * (j - 0.5 * X_gridDIM);
        const double X_prime = dX * (i - 0.5 * X_gridDIM);
        aw_state_store<%(st)s>::put(rho + idx, cuda_complex(%(expr)s));
    }
}
""" % dict(st=self._state_t(), expr=new_rho)
            mod = rtc.Module(src, "rho_vneumann_init.hip")
            blocks, threads = rtc.launch_cfg(N * N)
            mod.launch("aw_init", (blocks,), (threads,), [self.rho.ptr, ctypes.c_longlong(N * N)],
                       current_stream())
        elif isinstance(new_rho, (float, complex, int)):
            self.rho.fill(np.complex128(new_rho))
        else:
            raise NotImplementedError("new_rho must be either function or numpy.array")
        self._trace_normalize(0)
        return self

    # ------------------------------------------------------------------------------------
    # propagation
    # ------------------------------------------------------------------------------------
    def _times(self, steps):
        ts = np.empty(steps, np.float64)
        t = self.t
        for k in range(steps):
            t = t + self.dt
            ts[k] = t
        return ts

    def _chunk_factors(self, ts, merged=True, bufs=None, both=False):
        """FP64 phase tables (hipRTC) and factor vectors for steps at times ts.  merged: the
        closing half-kick of a step and the opening one of the next share one column pass
        (n + 1 column vectors); otherwise every step has its own opening and closing
        vectors (2 n column vectors, open at index 2 s, close at 2 s + 1).  bufs: write the
        times, tables and (unmerged) factors into these preallocated arrays; both: also the
        merged column vectors, into bufs["facm"]."""
        n = len(ts)
        N, dt, p = self.X_gridDIM, self.dt, self.plan
        if bufs is None:
            ts_dev = to_device(ts)
            Vt = empty((n, N), np.float64)
            Kt = empty((n, N), np.float64)
        else:
            ts_dev, Vt, Kt = bufs["ts"], bufs["Vt"], bufs["Kt"]
            if isinstance(ts, np.ndarray):
                ts_dev[:n] = ts
            else:
                ts_dev.tensor[:n].copy_(ts)   # a device slice: ordered on the stream, no host wait
        self._eval_1d(0, self._dev_kX, N, ts_dev, n, 0.5 * dt, Vt)
        self._eval_1d(1, self._dev_kP, N, ts_dev, n, 0.5 * dt, Kt)
        bx = self._bx.ptr if self._bx is not None else None
        bp = self._bp.ptr if self._bp is not None else None
        if bufs is None:
            rowL = p.new_factors(N, n)
            rowE = p.new_factors(N, n)
        else:
            colL, colE, rowL, rowE = bufs["fac"]
        # expK (:733-745): line i -> exp(+i dt K(P_i)) bp_i / N, elem j -> exp(-i dt K(P_j)) bp_j
        p.factors(rowL, N, n, A=Kt.ptr, ca=dt, mag=bp, scale=1.0 / N)
        p.factors(rowE, N, n, A=Kt.ptr, ca=-dt, mag=bp)
        if not merged:
            if bufs is None:
                colL = p.new_factors(N, 2 * n)
                colE = p.new_factors(N, 2 * n)
            for v, scale in ((0, 1.0), (1, 1.0 / N)):
                # expV (:754-770) with S = (-1)^(i+j) split as (-1)^j (-1)^i
                p.factors(colL, N, n, A=Vt.ptr, ca=-0.5 * dt, mag=bx, alt_sign=True,
                          out_offset=v, out_vstride=2)
                p.factors(colE, N, n, A=Vt.ptr, ca=0.5 * dt, mag=bx, alt_sign=True, scale=scale,
                          out_offset=v, out_vstride=2)
            if not both:
                return colL, colE, rowL, rowE
            self._merged_cols(Vt, n, bx, *bufs["facm"])
            return colL, colE, rowL, rowE
        colL = p.new_factors(N, n + 1)
        colE = p.new_factors(N, n + 1)
        self._merged_cols(Vt, n, bx, colL, colE)
        return colL, colE, rowL, rowE

    def _merged_cols(self, Vt, n, bx, colL, colE):
        """Column vectors of n merged steps: 0 opens step 0, s (0 < s < n) closes step s - 1
        and opens step s, n closes step n - 1 (with the 1 / N of the inverse FFT)."""
        N, dt, p = self.X_gridDIM, self.dt, self.plan
        p.factors(colL, N, 1, A=Vt.ptr, ca=-0.5 * dt, mag=bx, alt_sign=True)
        p.factors(colE, N, 1, A=Vt.ptr, ca=0.5 * dt, mag=bx, alt_sign=True)
        if n > 1:
            p.factors(colL, N, n - 1, A=Vt.ptr, ca=-0.5 * dt, B=Vt.ptr + N * 8, cb=-0.5 * dt,
                      mag=bx, mag_pow=2, out_offset=1)
            p.factors(colE, N, n - 1, A=Vt.ptr, ca=0.5 * dt, B=Vt.ptr + N * 8, cb=0.5 * dt,
                      mag=bx, mag_pow=2, scale=1.0 / N, out_offset=1)
        last = Vt.ptr + (n - 1) * N * 8
        p.factors(colL, N, 1, A=last, ca=-0.5 * dt, mag=bx, alt_sign=True, out_offset=n)
        p.factors(colE, N, 1, A=last, ca=0.5 * dt, mag=bx, alt_sign=True, scale=1.0 / N, out_offset=n)

    def _run_chunk(self, ts, dscale=None):
        """Steps at times ts as 2n+1 fused passes."""
        n = len(ts)
        N = self.X_gridDIM
        csz = self.plan.fsize
        colL, colE, rowL, rowE = self._chunk_factors(ts)
        p, rho = self.plan, self.rho
        ev = self._pass_events
        if ev is not None:
            import torch
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        p.steps(rho, n, colL, colE, rowL, rowE, dscale=dscale)
        if ev is not None:
            e1.record()
            ev.append((e0, e1, 2 * n + 1))

    def _ehrenfest_chunk(self, n):
        """n steps [EV, I0] [F1, EK, I1] [F0, EV] from the chunk buffers (applying the pending
        normalisation), each followed by one native call for the trace, the x-observables
        from diag rho, the p-observables from the signed cyclic diagonal sums c_d (one read of
        rho for all: sum_j [S I1 (g/N) F1 S rho]_jj = Re sum_k g(P_k) F(c)_k / N
        = Re sum_d c_d F(g)_d / N, with the DFTs of the weights F(g) made once per chunk)
        and the next scale 1 / (Re tr(rho) dX).  Only launches, from fixed addresses."""
        bufs, p, rho = self._ehr, self.plan, self.rho
        N = self.X_gridDIM
        csz = self.plan.fsize
        colL, colE, rowL, rowE = bufs["fac"]
        stage, xw, pwc = bufs["stage"], bufs["xw"], bufs["pwc"]
        for k in range(n):
            p.steps(rho, 1, colL.ptr + 2 * k * N * csz, colE.ptr + 2 * k * N * csz,
                    rowL.ptr + k * N * csz, rowE.ptr + k * N * csz, dscale=self._dscale)
            self._scale_pending = False
            p.cyclic_diag_obs(rho, xw.ptr + k * 3 * N * 8, 3, pwc.ptr + k * 3 * N * 16, 3, stage.ptr + k * 64, 8,
                              self.dX, self._dscale)
            self._scale_pending = True

    def _ehrenfest_merged_chunk(self, n):
        """n steps as 2 n + 1 passes, [EV_0 -> I0] ([F1 -> EK_s -> I1] [F0 -> EV_s EV_s+1 -> I0])*
        ending [F0 -> EV_n-1]: each merged column pass also multiplies the values after its
        first FFT by step s's closing vectors, which is rho_s exactly as an unmerged step ends,
        and stores the block partials of its trace and Ehrenfest sums (obs_mode 2) in a ring;
        one launch per chunk adds them into the stage rows.  No per-step normalisation: the
        steps are linear and the averages are ratios to the recorded trace.  Only launches,
        from fixed addresses."""
        bufs, p, rho = self._ehr, self.plan, self.rho
        N, csz = self.X_gridDIM, self.plan.fsize
        colL, colE, rowL, rowE = bufs["fac"]
        mL, mE = bufs["facm"]
        xw, pwc, ring, nblk = bufs["xw"], bufs["pwc"], bufs["ring"], bufs["nblk"]
        nx, np_ = (len(e) for e in bufs["exprs"])
        nc = 1 + nx + np_
        p.run(rho, 0, 0, INV, line_fac=mL, elem_fac=mE)
        for k in range(n):
            last = k == n - 1
            p.run(rho, 1, FWD, INV, line_fac=rowL.ptr + k * N * csz, elem_fac=rowE.ptr + k * N * csz)
            p.run(rho, 0, FWD, 0 if last else INV, line_fac=mL.ptr + (k + 1) * N * csz,
                  elem_fac=mE.ptr + (k + 1) * N * csz, obs_mode=2,
                  obs_line_fac=colL.ptr + (2 * k + 1) * N * csz, obs_elem_fac=colE.ptr + (2 * k + 1) * N * csz,
                  obs_weights=pwc.ptr + k * np_ * N * 16, obs_nw=np_, obs_norm2_weights=xw.ptr + k * nx * N * 8,
                  obs_diag_nw=nx, obs_partials=ring.ptr + k * nblk * nc * 8)
        finish_sums(ring, nblk, nc, n, bufs["stage"])

    def single_step_propagation(self):
        """rho_vneumann_cuda_1d.py:450-468: unnormalised (no 1/N^2, :464 is commented out)."""
        self._flush_scale()
        self._run_chunk(np.array([self.t], np.float64))
        self.plan.scale(self.rho, None, float(self.X_gridDIM) ** 2)
        return self.rho

    def propagate(self, steps=1):
        """rho_vneumann_cuda_1d.py:429-448: t += dt, step, rho /= Re tr(rho) dX, Ehrenfest."""
        steps = int(steps)
        if steps <= 0:
            return self
        ts = self._times(steps)
        if not self.isEhrenfest:
            self._flush_scale()
            for c0 in range(0, steps, CHUNK_STEPS):
                self._run_chunk(ts[c0:c0 + CHUNK_STEPS])
                self._trace_normalize(0)
            self.t = ts[-1]
            return self
        self._flush_scale()
        N = self.X_gridDIM
        CH = CHUNK_STEPS
        wanted_x = ["X", self.diff_V, self.V]
        wanted_p = [self.diff_K, "P", self.K]
        if self._ehr is None:
            p = self.plan
            # merged steps when the column pass stores block partials at this size
            nblk = p.obs_blocks(0) if self.merged_ehrenfest else 0
            if nblk:
                # each distinct expression is summed once (the strong-field case has diff_K = "P")
                exprs_x = list(dict.fromkeys(e.strip() for e in wanted_x))
                exprs_p = list(dict.fromkeys(e.strip() for e in wanted_p))
            else:
                exprs_x, exprs_p = wanted_x, wanted_p
            ncol = 8 if nblk == 0 else 1 + len(exprs_x) + len(exprs_p)
            # DFTs of the chunk's momentum weights: CH x (up to 3) rows of N points, FP64
            self._cplan = Plan(1, N, CH * len(exprs_p), "fp64")
            self._ehr = dict(ts=empty((CH,), np.float64), Vt=empty((CH, N), np.float64),
                             Kt=empty((CH, N), np.float64),
                             fac=(p.new_factors(N, 2 * CH), p.new_factors(N, 2 * CH),
                                  p.new_factors(N, CH), p.new_factors(N, CH)),
                             xw=empty((CH, len(exprs_x), N), np.float64),
                             pwc=DeviceArray((CH, len(exprs_p), N), np.complex128),
                             stage=empty((CH, ncol), np.float64), graphs={}, nblk=nblk,
                             exprs=(exprs_x, exprs_p))
            if nblk:
                self._ehr["facm"] = (p.new_factors(N, CH + 1), p.new_factors(N, CH + 1))
                self._ehr["ring"] = empty((CH, nblk, ncol), np.float64)
            self._obs_module(exprs_x)
            self._obs_module(exprs_p)
            if not nblk:
                # the observable scratch is allocated by the first call, never in a capture
                self.plan.cyclic_diag_obs(self.rho, self._ehr["xw"], 3, self._ehr["pwc"], 3, self._ehr["stage"], 8)
        bufs = self._ehr
        merged = bufs["nblk"] > 0
        exprs_x, exprs_p = bufs["exprs"]
        ncol = bufs["stage"].shape[1]
        ring = empty((steps, ncol), np.float64)
        if not merged and not self._scale_pending:
            # every step applies the pending scale: start from 1 so that all chunks are alike
            self._dscale.tensor[:, 0] = 1.0
            self._dscale.tensor[:, 1] = 0.0
            self._scale_pending = True
        ts_all = to_device(ts)   # one upload; each chunk copies its slice on the device
        for c0 in range(0, steps, CH):
            n = min(CH, steps - c0)
            self._chunk_factors(ts_all.tensor[c0:c0 + n], merged=False, bufs=bufs, both=merged)
            # the observable weights of all the chunk's times at once
            self._weights_t(exprs_x, bufs["ts"], n, bufs["xw"])
            self._weights_t(exprs_p, bufs["ts"], n, bufs["pwc"], complex_out=True)
            self._cplan.fft(bufs["pwc"], 1, FWD)
            if merged:
                graphs.run_chunk(bufs["graphs"], "merged", lambda n=n: self._ehrenfest_merged_chunk(n), n == CH)
            else:
                graphs.run_chunk(bufs["graphs"], "ehrenfest", lambda n=n: self._ehrenfest_chunk(n), n == CH)
            ring.tensor[c0:c0 + n].copy_(bufs["stage"].tensor[:n])
        self.t = ts[-1]
        if merged:
            # rho /= Re tr(rho) dX with the last step's recorded trace (:443)
            self.plan.scale_from_stats(ring.ptr + (steps - 1) * ncol * 8, ncol, 0, 0, self.dX, self._dscale)
            self._scale_pending = True
        self._flush_scale()
        r = ring.get()
        tr = r[:, 0]
        # stage columns: trace, then the coordinate and the momentum sums of exprs_x / exprs_p
        cx = [1 + exprs_x.index(e.strip() if merged else e) for e in wanted_x]
        cp = [1 + len(exprs_x) + exprs_p.index(e.strip() if merged else e) for e in wanted_p]
        self.X_average.extend((r[:, cx[0]] / tr).tolist())
        self.P_average_RHS.extend((-r[:, cx[1]] / tr).tolist())
        self.X_average_RHS.extend((r[:, cp[0]] / tr).tolist())
        self.P_average.extend((r[:, cp[1]] / tr).tolist())
        self.hamiltonian_average.extend(((r[:, cp[2]] + r[:, cx[2]]) / tr).tolist())
        return self

    # ------------------------------------------------------------------------------------
    # observables
    # ------------------------------------------------------------------------------------
    def _obs_module(self, exprs):
        key = tuple(exprs)
        mod = self._obs_cache.get(key)
        if mod is None:
            body = "".join("        out[%d * n + k] = (%s);\n" % (q, e) for q, e in enumerate(exprs))
            cbody = "".join("        outc[%d * n + k] = make_double2((%s), 0.);\n" % (q, e)
                            for q, e in enumerate(exprs))
            src = (rtc.PRELUDE + self.cuda_consts + "\n#define NOBS_W %d\n" % len(exprs)
                   + "\n__device__ double K(double P, double t)\n{\n    return (%s);\n}\n" % self.K
                   + "\n__device__ double V(double X, double t)\n{\n    return (%s);\n}\n" % self.V
                   + r"""
extern "C" __global__ void aw_weights(const double *Xg, const double *Pg, int n, double t, double *out)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const double X = Xg[k], P = Pg[k];
%s    }
}
extern "C" __global__ void aw_weights_t(const double *Xg, const double *Pg, int n, const double *ts, int nt,
                                        double *out_all)
{
    // the same weights for nt times: out_all[s][q][k]
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * nt; idx += gridDim.x * blockDim.x) {
        const int s = idx / n, k = idx - s * n;
        const double X = Xg[k], P = Pg[k], t = ts[s];
        double *out = out_all + (long long)s * NOBS_W * n;
%s    }
}
extern "C" __global__ void aw_weights_tc(const double *Xg, const double *Pg, int n, const double *ts, int nt,
                                         double2 *outc_all)
{
    // complex (imaginary part 0) copy of aw_weights_t's output, ready for a DFT
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * nt; idx += gridDim.x * blockDim.x) {
        const int s = idx / n, k = idx - s * n;
        const double X = Xg[k], P = Pg[k], t = ts[s];
        double2 *outc = outc_all + (long long)s * NOBS_W * n;
%s    }
}
""" % (body, body, cbody))
            mod = rtc.Module(src, "rho_vneumann_obs.hip")
            self._obs_cache[key] = mod
        return mod

    def _weights(self, exprs, t, out):
        """out[q][k] = expr_q(X_k, P_k, t): the observable kernel puts X and P both on the
        column index (rho_vneumann_cuda_1d.py:811-821)."""
        N = self.X_gridDIM
        mod = self._obs_module(exprs)
        blocks, threads = rtc.launch_cfg(N)
        mod.launch("aw_weights", (blocks,), (threads,),
                   [self._dev_kX.ptr, self._dev_kP.ptr, ctypes.c_int(N), ctypes.c_double(t), out.ptr],
                   current_stream())

    def _weights_t(self, exprs, ts_dev, nt, out, complex_out=False):
        """out[s][q][k] = expr_q(X_k, P_k, ts[s]) for s < nt (one launch per chunk); as
        complex128 (imaginary part 0) with complex_out."""
        N = self.X_gridDIM
        mod = self._obs_module(exprs)
        blocks, threads = rtc.launch_cfg(N * nt)
        mod.launch("aw_weights_tc" if complex_out else "aw_weights_t", (blocks,), (threads,),
                   [self._dev_kX.ptr, self._dev_kP.ptr, ctypes.c_int(N), ts_dev.ptr, ctypes.c_int(nt), out.ptr],
                   current_stream())

    def _weight_vec(self, w, q, N):
        """Complex factor vector (plan precision) holding the real weights w[q]."""
        v = self.plan.new_factors(N, 1)
        self.plan.factors(v, N, 1, mag=w.ptr + q * N * 8)
        return v

    def get_observable(self, observable_str):
        """rho_vneumann_cuda_1d.py:503-518 (the compiled weight kernel for a string)."""
        return self._obs_module([observable_str])

    def get_average(self, observable):
        """rho_vneumann_cuda_1d.py:520-563: alternating x / p operators applied on the ket
        index to a copy of rho, then Re tr(.) dX."""
        N = self.X_gridDIM
        self._flush_scale()
        if self._tmp is None:
            self._tmp = DeviceArray((N, N), self.state_dtype)
        self._tmp[:] = self.rho
        is_x = False
        w = empty((1, N), np.float64)
        for obs in observable:
            is_x = not is_x
            if not obs:
                continue
            self._weights([obs], float(self.t), w)
            f = self._weight_vec(w, 0, N)
            if is_x:
                self.plan.run(self._tmp, 1, 0, 0, elem_fac=f)
            else:
                self.plan.run(self._tmp, 1, FWD, INV, sign_in=1, sign_out=1, elem_fac=f, scale=1.0 / N)
        st = self.plan.stats(self._tmp).get()[0]
        return float(st[5] * self.dX)

    def get_purity(self):
        """sum |rho|^2 dX^2 (rho_vneumann_cuda_1d.py:565-579)."""
        self._flush_scale()
        return float(self.plan.stats(self.rho).get()[0][2] * self.dX ** 2)

    def get_sigma_x_sigma_p(self):
        """rho_vneumann_cuda_1d.py:581-591."""
        ga = self.get_average
        return np.sqrt((ga(("X * X",)) - ga(("X",)) ** 2) * (ga((None, "P * P")) - ga((None, "P")) ** 2))

    def get_Ehrenfest(self, t):
        """rho_vneumann_cuda_1d.py:470-501."""
        if self.isEhrenfest:
            self.X_average.append(self.get_average(("X",)))
            self.X_average_RHS.append(self.get_average((None, self.diff_K)))
            self.P_average.append(self.get_average((None, "P")))
            self.P_average_RHS.append(-self.get_average((self.diff_V,)))
            self.hamiltonian_average.append(self.get_average((None, self.K)) + self.get_average((self.V,)))

    # ------------------------------------------------------------------------------------
    # Wigner function on the zero-padded 2N grid
    # ------------------------------------------------------------------------------------
    def get_wignerfunction(self):
        """rho_vneumann_cuda_1d.py:387-427: zero-pad rho into the centre of a (2N)^2 array,
        S'; [F1, SX', I1]; [F0, SY', I0]; [F1, SX', I1]; I0; S'; normalise by Re sum W dXdP."""
        self._flush_scale()
        N, M, pad = self.X_gridDIM, self.XX_gridDIM, self.wigner_padding
        if self._wplan is None:
            self._wplan = Plan(M, M, 1, self.precision)
            self.wignerfunction = DeviceArray((M, M), self.state_dtype)
            self._wscale = empty((1, 2), np.float64)
        W = self.wignerfunction
        W.fill(0.)
        W.tensor[pad:pad + N, pad:pad + N] = self.rho.tensor
        wp = self._wplan
        a, b = -np.tan(np.pi / 8.), np.sin(np.pi / 4.)
        kdx, kdpp = rtc.kconst(self.dX), rtc.kconst(self.dPP)
        shx = dict(chirp=a, lstep=kdx, lcenter=0.5 * M, estep=kdpp, ecenter=0.5 * M, scale=1.0 / M)
        shy = dict(chirp=b, lstep=kdx, lcenter=0.5 * M, estep=kdpp, ecenter=0.5 * M, scale=1.0 / M)
        wp.run(W, 1, FWD, INV, sign_in=1, **shx)
        wp.run(W, 0, FWD, INV, **shy)
        wp.run(W, 1, FWD, INV, **shx)
        wp.run(W, 0, INV, 0, sign_out=1)
        st = wp.stats(W)
        wp.scale_from_stats(st.ptr, 8, 0, 0, self.dXdP_wigner, self._wscale)
        wp.scale(W, self._wscale)
        return W
