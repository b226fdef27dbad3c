"""WignerMoyalCUDA1D: drop-in MI355X implementation of the reference class of the same name
(wigner_moyal_cuda_1d.py:10-828).

Same constructor keywords, attributes and methods.  Time stepping follows the reference's
operator sequence (wigner_moyal_cuda_1d.py:543-561):

    rho <- EV rho; I0; F1; EK; F0; I1; /(Nx Np); EV

but runs as fused line passes: a column pass [F0 -> EV(t_n) EV(t_{n+1}) -> I0] (the two
half-kicks of consecutive steps merged) and a row pass [F1 -> EK -> I1] per step, so each
step reads and writes the state twice.  Phase factors are outer products of 1D vectors
evaluated in FP64 from the user's C expressions (hipRTC), exactly the expV / expK phases
-0.5 dt [V(XX_j) - V(XX'_i)] and -dt [K(PP_j) - K(PP'_i)] at t + dt/2 (:751, :772).

Extra keywords (not in the reference): ``precision`` = "fp64" (default, complex128 state,
the reference's arithmetic) or "mixed" (complex64 state and FFT arithmetic, FP64 phase
arguments and FP64 reductions); ``batch`` = B > 1 propagates B independent states under the
same Hamiltonian in every launch (arrays become [B, Np, Nx], averages become length-B arrays).
"""
import ctypes
import re
from types import FunctionType

import numpy as np

from . import rtc
from ._params import KeywordParams
from .device import DeviceArray, current_stream, empty, require_gpu, to_device
from .engine import FWD, INV, Plan, finish_sums, mem_info, rank1_split, sum_partials
from . import graphs, wm_ops
from .wm_ops import WMGrid

CHUNK_STEPS = 256

_T_TOKEN = re.compile(r"\bt\b")


def time_independent(expr, functions=""):
    """True when the C expression cannot depend on t: it names no ``t`` and the user's
    helper code defines no macro that does (a helper function only sees t through its
    arguments).  Conservative: a false negative only keeps the per-step Wigner transform."""
    if _T_TOKEN.search(expr):
        return False
    for line in (functions or "").replace("\\\n", " ").splitlines():
        if line.strip().startswith("#define") and _T_TOKEN.search(line):
            return False
    return True


class WignerMoyalCUDA1D(KeywordParams):
    """The second-order split-operator propagator for the Moyal equation for the Wigner
    function W(x, p, t) with the time-dependent Hamiltonian H = K(p, t) + V(x, t)
    (reference: wigner_moyal_cuda_1d.py:10-16), on MI355X."""

    # record time-independent Ehrenfest observables as Re <L^H f, rho> (one reduction per step)
    # instead of forming W each step; False keeps the Wigner transform every step
    obs_by_adjoint = True

    def __init__(self, **kwargs):
        """Keywords as in the reference constructor (wigner_moyal_cuda_1d.py:17-41):
        X_gridDIM, X_amplitude, P_gridDIM, P_amplitude, V, K, dt, and optionally t, consts,
        functions, diff_V, diff_K, abs_boundary_x, abs_boundary_p, max_thread_block, plus
        ``precision`` ("fp64" | "mixed")."""
        kwargs = dict(kwargs)
        self.precision = kwargs.pop("precision", "fp64")
        self.batch = int(kwargs.pop("batch", 1))
        # save all attributes; Python functions become bound methods (:44-50)
        self._set_params(kwargs)

        # required-parameter checks (:53-92)
        if not hasattr(self, "X_gridDIM"):
            raise AttributeError("Coordinate grid size (X_gridDIM) was not specified")
        assert 2 ** int(np.log2(self.X_gridDIM)) == self.X_gridDIM, \
            "Coordinate grid size (X_gridDIM) must be a power of two"
        if not hasattr(self, "P_gridDIM"):
            raise AttributeError("Momentum grid size (P_gridDIM) was not specified")
        assert 2 ** int(np.log2(self.P_gridDIM)) == self.P_gridDIM, \
            "Momentum grid size (P_gridDIM) must be a power of two"
        if not hasattr(self, "X_amplitude"):
            raise AttributeError("Coordinate grid range (X_amplitude) was not specified")
        if not hasattr(self, "P_amplitude"):
            raise AttributeError("Momentum grid range (P_amplitude) was not specified")
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
        self.t = np.float64(self.t)

        self._make_grids()

        # #define block (:151-182)
        kwargs.update(
            dX=self.dX, dP=self.dP, dXX=self.dXX, dXX_prime=self.dXX_prime, dPP=self.dPP,
            dPP_prime=self.dPP_prime, XX_max=self.XX.max(), XX_prime_max=self.XX_prime.max(),
            PP_max=self.PP.max(), PP_prime_max=self.PP_prime.max())
        self._defines = kwargs
        self.cuda_consts = rtc.cuda_consts(kwargs, getattr(self, "consts", None), None)
        if not hasattr(self, "abs_boundary_x"):
            self.abs_boundary_x = "1."
        if not hasattr(self, "abs_boundary_p"):
            self.abs_boundary_p = "1."
        if not hasattr(self, "max_thread_block"):
            self.max_thread_block = 512
        self.cuda_consts += getattr(self, "functions", "") or ""

        require_gpu()
        self.plan = Plan(self.P_gridDIM, self.X_gridDIM, self.batch, self.precision)
        self.state_dtype = self.plan.state_dtype
        self._compile_physics()

        # state arrays (:268-271); W and rho in the chosen precision
        shape = (self.P_gridDIM, self.X_gridDIM)
        if self.batch > 1:
            shape = (self.batch,) + shape
        self.wignerfunction = DeviceArray(shape, self.state_dtype)
        self.rho = DeviceArray(shape, self.state_dtype)
        self._dscale = empty((self.batch, 2), np.float64)
        self._scale_pending = False
        self._pass_events = None   # list -> (start, end) torch events around each chunk's passes

        self.hamiltonian = self.K + ' + ' + self.V
        self.isEhrenfest = hasattr(self, "diff_K") and hasattr(self, "diff_V")
        if self.isEhrenfest:
            self.X_average, self.P_average = [], []
            self.X_average_RHS, self.P_average_RHS = [], []
            self.hamiltonian_average = []
        self._observable_kernels = {}
        self._obs_partials = {}
        self._g1 = None           # L^H 1: the norm weights of normalize_rho_wigner
        self._group_plans = {}    # plans over one MALL-resident group of items
        self._ehr = None          # persistent chunk buffers and captured graphs of the Ehrenfest loop
        self._prop = None         # the same for the merged-step loop
        self._prepare_masks()
        self._static = self._h_static()
        self.print_memory_info()

    # ------------------------------------------------------------------------------------
    # grids and compiled physics
    # ------------------------------------------------------------------------------------
    def _make_grids(self):
        """Grids exactly as wigner_moyal_cuda_1d.py:110-143 (1D vectors in WMGrid, exposed here
        with the reference's broadcast shapes)."""
        g = self._grid = WMGrid(self.X_gridDIM, self.X_amplitude, self.P_gridDIM, self.P_amplitude)
        self.dX, self.dP, self.dXdP = g.dX, g.dP, g.dXdP
        self.dPP_prime, self.dXX_prime, self.dXX, self.dPP = g.dPP_prime, g.dXX_prime, g.dXX, g.dPP
        self.P = g.P[:, np.newaxis]
        self.X = g.X[np.newaxis, :]
        self.XX = g.XX[np.newaxis, :]
        self.PP = g.PP[np.newaxis, :]
        self.XX_prime = g.XX_prime[:, np.newaxis]
        self.PP_prime = g.PP_prime[:, np.newaxis]
        self._kXX, self._kXXp, self._kPP, self._kPPp = g.kXX, g.kXXp, g.kPP, g.kPPp

    def _state_t(self):
        return "double2" if self.state_dtype == np.complex128 else "float2"

    def _physics_source(self):
        return (rtc.PRELUDE + self.cuda_consts
                + "\n__device__ double K(double P, double t)\n{\n    return (%s);\n}\n" % self.K
                + "\n__device__ double V(double X, double t)\n{\n    return (%s);\n}\n" % self.V
                + rtc.EVAL_1D
                + r"""
__device__ double aw_absx(double XX, double XX_prime, double t, size_t i, size_t j)
{
    return (%s);
}
__device__ double aw_absp(double PP, double PP_prime, double t, size_t i, size_t j)
{
    return (%s);
}
__device__ double aw_mask(int which, double xj, double xi, double t, size_t i, size_t j)
{
    return which == 0 ? aw_absx(xj, xi, t, i, j) : aw_absp(xj, xi, t, i, j);
}
// a separable mask M(t)[i][j] = v[i] u[j] at each of nt times: u[s][j] = M(t_s)[i0][j],
// v[s][i] = M(t_s)[i][j0] / M(t_s)[i0][j0]
extern "C" __global__ void aw_eval_mask_lines(int which, const double *xj, int nj, const double *xi, int ni,
                                              const double *ts, int nt, int i0, int j0, double *u, double *v)
{
    const long long total = (long long)nt * (nj + ni);
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(idx / (nj + ni)), k = (int)(idx %% (nj + ni));
        const double t = ts[s];
        if (k < nj) {
            u[(long long)s * nj + k] = aw_mask(which, xj[k], xi[i0], t, i0, k);
        } else {
            const int i = k - nj;
            v[(long long)s * ni + i] = aw_mask(which, xj[j0], xi[i], t, i, j0) / aw_mask(which, xj[j0], xi[i0], t, i0, j0);
        }
    }
}
// the mask table at the time *tp (read on the device, so that captured step loops can run it),
// in FP64 (out64) or FP32 (out32)
extern "C" __global__ void aw_eval_mask_at(int which, const double *xj, int nj, const double *xi, int ni,
                                           const double *tp, double *out64, float *out32)
{
    const long long total = (long long)ni * nj;
    const double t = *tp;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const size_t i = idx / nj, j = idx %% nj;
        const double m = which == 0 ? aw_absx(xj[j], xi[i], t, i, j) : aw_absp(xj[j], xi[i], t, i, j);
        if (out64) out64[idx] = m;
        else out32[idx] = (float)m;
    }
}
extern "C" __global__ void aw_eval_mask(int which, const double *xj, int nj, const double *xi,
                                        int ni, double t, double *out)
{
    const long long total = (long long)ni * nj;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const size_t i = idx / nj, j = idx %% nj;
        out[idx] = which == 0 ? aw_absx(xj[j], xi[i], t, i, j) : aw_absp(xj[j], xi[i], t, i, j);
    }
}
""" % (self.abs_boundary_x, self.abs_boundary_p))

    def _compile_physics(self):
        self._mod = rtc.Module(self._physics_source(), "wigner_moyal_physics.hip")
        self._dev_kXX = to_device(self._kXX)
        self._dev_kXXp = to_device(self._kXXp)
        self._dev_kPP = to_device(self._kPP)
        self._dev_kPPp = to_device(self._kPPp)

    def _eval_1d(self, which, xdev, n, ts_dev, nt, tshift, out):
        blocks, threads = rtc.launch_cfg(n * nt)
        self._mod.launch("aw_eval_1d", (blocks,), (threads,),
                         [ctypes.c_int(which), xdev.ptr, ctypes.c_int(n), ts_dev.ptr,
                          ctypes.c_int(nt), ctypes.c_double(tshift), out.ptr], current_stream())

    def _eval_mask(self, which, t):
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        out = empty((Np, Nx), np.float64)
        xj, xi = (self._dev_kXX, self._dev_kXXp) if which == 0 else (self._dev_kPP, self._dev_kPPp)
        blocks, threads = rtc.launch_cfg(Nx * Np)
        self._mod.launch("aw_eval_mask", (blocks,), (threads,),
                         [ctypes.c_int(which), xj.ptr, ctypes.c_int(Nx), xi.ptr, ctypes.c_int(Np),
                          ctypes.c_double(t), out.ptr], current_stream())
        return out.get()

    def _prepare_masks(self):
        """Evaluate abs_boundary_x(XX, XX', t) and abs_boundary_p(PP, PP', t), which expV /
        expK apply at the step time t (:753, :777).  A time-independent mask is evaluated once:
        two vectors when it is separable, else a 2D table.  A time-dependent mask that stays
        separable (rank 1 at every sampled time) has its line and element vectors evaluated on
        the device for every step of a chunk (``_mask_tables``), which multiply the factor
        vectors.  Any other time-dependent mask is evaluated as a 2D table at each step's time
        (``_ttable``) and its steps run unmerged, one step of three passes at a time."""
        self._masks = {}
        for which, expr in ((0, self.abs_boundary_x), (1, self.abs_boundary_p)):
            if expr.strip() in ("1.", "1", "1.0"):
                self._masks[which] = None
                continue
            M0 = self._eval_mask(which, float(self.t))
            M1 = self._eval_mask(which, float(self.t) + 1.2345)
            if not np.array_equal(M0, M1):
                M2 = self._eval_mask(which, float(self.t) + 2.7183)
                # pivot of the split: the largest entry, where the mask is nearest 1
                i0, j0 = np.unravel_index(np.argmax(np.abs(M0)), M0.shape)
                if any(rank1_split(M) is None for M in (M0, M1, M2)) or any(M[i0, j0] == 0.0 for M in (M1, M2)):
                    self._masks[which] = ("ttable",)
                else:
                    self._masks[which] = ("time", int(i0), int(j0))
                continue
            split = rank1_split(M0)
            if split is not None:
                v, u = split
                self._masks[which] = ("vec", to_device(v), to_device(u))
            else:
                rd = self.plan.real_dtype
                self._masks[which] = ("table", to_device(M0.astype(rd)),
                                      to_device((M0 * M0).astype(rd)))

    # ------------------------------------------------------------------------------------
    # observables
    # ------------------------------------------------------------------------------------
    def _observable_kernel(self, exprs):
        key = tuple(exprs)
        mod = self._observable_kernels.get(key)
        if mod is None:
            nobs = len(exprs)
            body = "".join("        acc[%d] += w * (%s);\n" % (k, e) for k, e in enumerate(exprs))
            grid_body = "".join("        out[%d * (long long)n_item + idx] = make_double2((%s), 0.);\n" % (k, e)
                                for k, e in enumerate(exprs))
            pt_body = "".join("        o[%d * (long long)npts] = (%s);\n" % (k, e) for k, e in enumerate(exprs))
            src = (rtc.PRELUDE + self.cuda_consts
                   + "\n__device__ double K(double P, double t)\n{\n    return (%s);\n}\n" % self.K
                   + "\n__device__ double V(double X, double t)\n{\n    return (%s);\n}\n" % self.V
                   + r"""
__device__ __forceinline__ void aw_obs_body(const %(st)s *W, int n_item, double t, double *partials)
{
    double acc[%(nobs)d];
    for (int c = 0; c < %(nobs)d; ++c) acc[c] = 0.;
    const %(st)s *Wb = W + (long long)blockIdx.y * n_item;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n_item; idx += gridDim.x * blockDim.x) {
        const int i = idx / X_gridDIM, j = idx - i * X_gridDIM;
        const double X = dX * (j - 0.5 * X_gridDIM);
        const double P = dP * (i - 0.5 * P_gridDIM);
        const double w = (double)Wb[idx].x;
%(body)s    }
""" % dict(st=self._state_t(), nobs=nobs, body=body) + rtc.REDUCE_TAIL % dict(nobs=nobs) + "}\n" + r"""
extern "C" __global__ void aw_obs(const %(st)s *W, int n_item, double t, double *partials)
{
    aw_obs_body(W, n_item, t, partials);
}
// the same with the time read from device memory (a captured step sequence replays with
// new times written into the buffer)
extern "C" __global__ void aw_obs_tp(const %(st)s *W, int n_item, const double *tp, double *partials)
{
    aw_obs_body(W, n_item, *tp, partials);
}
// the expressions themselves on the grid at time t: out[k][i][j] = (f_k(X_j, P_i, t), 0)
extern "C" __global__ void aw_obs_grid(double2 *out, int n_item, double t)
{
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n_item; idx += gridDim.x * blockDim.x) {
        const int i = idx / X_gridDIM, j = idx - i * X_gridDIM;
        const double X = dX * (j - 0.5 * X_gridDIM);
        const double P = dP * (i - 0.5 * P_gridDIM);
%(grid_body)s    }
}
// the expressions at npts grid points (flat indices i * X_gridDIM + j) and nt times:
// out[s][k][q] = f_k(X, P, ts[s]) at point pts[q]
extern "C" __global__ void aw_obs_points(const int *pts, int npts, const double *ts, int nt, double *out)
{
    const long long total = (long long)nt * npts;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(idx / npts), q = (int)(idx %% npts);
        const int g = pts[q], i = g / X_gridDIM, j = g - i * X_gridDIM;
        const double X = dX * (j - 0.5 * X_gridDIM);
        const double P = dP * (i - 0.5 * P_gridDIM);
        const double t = ts[s];
        double *o = out + (long long)s * %(nobs)d * npts + q;
%(pt_body)s    }
}
""" % dict(st=self._state_t(), grid_body=grid_body, nobs=nobs, pt_body=pt_body))
            mod = rtc.Module(src, "wigner_moyal_observables.hip")
            self._observable_kernels[key] = mod
        return mod

    def _obs_partials_for(self, nexprs):
        """(blocks per item, partial-sum buffer) of the observable kernels; the buffer is
        reused by every launch of this shape (stream order keeps the uses apart)."""
        n = self.X_gridDIM * self.P_gridDIM
        nblk = max(1, min(512, n // 2048))
        key = (nblk, nexprs)
        partials = self._obs_partials.get(key)
        if partials is None:
            partials = self._obs_partials[key] = empty((self.batch, nblk, nexprs), np.float64)
        return nblk, partials

    def _obs_launch(self, exprs, W, t, out):
        """Partial sums of Re W * f_k(X, P, t) for each expression, summed into out[k]."""
        mod = self._observable_kernel(exprs)
        n = self.X_gridDIM * self.P_gridDIM
        nblk, partials = self._obs_partials_for(len(exprs))
        mod.launch("aw_obs", (nblk, self.batch), (256,), [W.ptr, ctypes.c_int(n), ctypes.c_double(t),
                                                          partials.ptr], current_stream())
        sum_partials(partials, nblk, len(exprs), self.batch, out)

    def _obs_fin_launch(self, exprs, W, t_ptr, out_ptr, factor, dscale):
        """The sums of Re W * f_k(X, P, t) (t read from device memory at t_ptr) into
        out[item][k] and the normalisation scale 1 / (out[item][0] factor) into dscale: the
        block partials, then one launch that adds them and derives the scale."""
        mod = self._observable_kernel(exprs)
        n = self.X_gridDIM * self.P_gridDIM
        nblk, partials = self._obs_partials_for(len(exprs))
        mod.launch("aw_obs_tp", (nblk, self.batch), (256,), [W.ptr, ctypes.c_int(n), t_ptr, partials.ptr],
                   current_stream())
        finish_sums(partials, nblk, len(exprs), self.batch, out_ptr, factor, dscale)

    def _obs_weights(self, exprs, t=None):
        """g_k = L^H f_k for observables f_k(X, P) that do not depend on t, with L = rho2wigner_
        (the unnormalised transform of wigner_moyal_cuda_1d.py:471-501): then
        sum Re(W) f_k = Re <f_k, L rho> = Re <g_k, rho>, one weighted sum over rho per step
        instead of the Wigner transform and the observable pass.  complex128 [nobs, Np, Nx],
        or None when an expression may depend on t (or there are more than 8).  With t given,
        the expressions are taken at that fixed time and may depend on it."""
        if len(exprs) > 8:
            return None
        if t is None and not all(time_independent(e, getattr(self, "functions", "")) for e in exprs):
            return None
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        mod = self._observable_kernel(exprs)
        g = empty((len(exprs), Np, Nx), np.complex128)
        blocks, threads = rtc.launch_cfg(Nx * Np)
        mod.launch("aw_obs_grid", (blocks,), (threads,),
                   [g.ptr, ctypes.c_int(Nx * Np), ctypes.c_double(0.0 if t is None else float(t))], current_stream())
        self._obs_plan = Plan(Np, Nx, len(exprs), "fp64")   # lives with the system (its launches are queued)
        wm_ops.rho2wigner_adjoint(self._obs_plan, self._grid, g)
        return g

    def _obs_points(self, exprs, pts, ts):
        """f_k(X, P, t) at the flat grid indices pts for each time in ts: host [nt, nexprs, npts]."""
        mod = self._observable_kernel(exprs)
        pts_d = to_device(np.ascontiguousarray(pts, dtype=np.int32))
        ts_d = to_device(np.ascontiguousarray(ts, dtype=np.float64))
        out = empty((len(ts), len(exprs), len(pts)), np.float64)
        blocks, threads = rtc.launch_cfg(len(ts) * len(pts))
        mod.launch("aw_obs_points", (blocks,), (threads,),
                   [pts_d.ptr, ctypes.c_int(len(pts)), ts_d.ptr, ctypes.c_int(len(ts)), out.ptr], current_stream())
        return out.get()

    # sample times and tolerances of the time-separable observables
    _TSEP_TAUS = (0.3711, 1.9137, 3.1303, 7.7719, 12.4261, 29.3333)
    _TSEP_TOL = 1e-9

    def _tsep_setup(self, exprs):
        """Observables f_e(X, P, t) that depend on t only through a few functions of t, e.g.
        diff_V = X / (X^2 + a)^1.5 + E(t) or K + V with V = V0(X) + X E(t): at every t they lie
        in the span of r <= 8 time-independent grid functions, here actual evaluations
        b_m = f_e(., tau_m) at sample times.  Then sum W f_e(t) = sum_m c_em(t) Re <L^H b_m, rho>:
        the r weight arrays L^H b_m are formed once, and c(t) comes from the values of f_e(t)
        at r pivot grid points.  The span is found by a pivoted QR of the expressions sampled
        at 6 times on 4096 grid points and verified on 16384 more; each call checks it again
        at its own step times (``_tsep_coeffs``).  None when the observables do not fit."""
        import scipy.linalg as sla
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        ng = Nx * Np
        rng = np.random.default_rng(20231)
        nsamp = min(ng, 4096 + 16384)
        allp = rng.choice(ng, size=nsamp, replace=False).astype(np.int32)
        ns1 = min(4096, nsamp // 2)
        taus = float(self.t) + np.array(self._TSEP_TAUS)
        vals = self._obs_points(exprs, allp, taus)            # [ntau, ne, nsamp]
        ti = [time_independent(e, self.cuda_consts) for e in exprs]
        cols, keys = [], []
        for k in range(len(exprs)):
            for j in ([0] if ti[k] else range(len(taus))):
                cols.append(vals[j, k]); keys.append((k, j))
        M = np.stack(cols, axis=1)                             # [nsamp, ncols]
        scale = np.maximum(np.max(np.abs(M), axis=0), 1e-300)
        _, R, piv = sla.qr(M[:ns1] / scale, mode="economic", pivoting=True)
        d = np.abs(np.diag(R))
        r = int(np.sum(d > 1e-10 * d[0])) if d[0] > 0 else 0
        if r == 0 or r > 8:
            return None
        bcols = list(piv[:r])
        Mb = M[:, bcols]
        _, _, prow = sla.qr(Mb[:ns1].T, mode="economic", pivoting=True)
        prow = np.asarray(prow[:r])
        Bp = Mb[prow]
        C = np.linalg.solve(Bp, M[prow])
        err = np.max(np.abs(Mb @ C - M), axis=0) / scale
        if not np.all(err < self._TSEP_TOL):
            return None
        chk = np.setdiff1d(np.arange(ns1, nsamp), prow)[:64]
        # the basis functions on the whole grid, then their adjoint Wigner transforms
        G = empty((r, Np, Nx), np.complex128)
        mod = self._observable_kernel(exprs)
        blocks, threads = rtc.launch_cfg(ng)
        tmp = empty((len(exprs), Np, Nx), np.complex128)
        for m, c in enumerate(bcols):
            k, j = keys[c]
            mod.launch("aw_obs_grid", (blocks,), (threads,), [tmp.ptr, ctypes.c_int(ng), ctypes.c_double(taus[j])],
                       current_stream())
            G.tensor[m].copy_(tmp.tensor[k])
        del tmp
        self._tsep_plan = Plan(Np, Nx, r, "fp64")             # lives with the system (queued launches)
        wm_ops.rho2wigner_adjoint(self._tsep_plan, self._grid, G)
        return dict(nb=r, G=G, pts=np.concatenate([allp[prow], allp[chk]]).astype(np.int32),
                    Bp=Bp, Bchk=Mb[chk], tol=self._TSEP_TOL)

    def _tsep_coeffs(self, exprs, ts):
        """c[s, e, m] with f_e(., ts[s]) = sum_m c[s, e, m] b_m, from the pivot points; None
        when a step's values at the check points do not follow (the span does not hold)."""
        info = self._ehr["tsep"]
        r = info["nb"]
        v = self._obs_points(exprs, info["pts"], ts)           # [nt, ne, r + nchk]
        nt, ne = v.shape[0], v.shape[1]
        c = np.linalg.solve(info["Bp"], v[:, :, :r].reshape(nt * ne, r).T)     # [r, nt ne]
        rec = (info["Bchk"] @ c).T.reshape(nt, ne, -1)
        chk = v[:, :, r:]
        scale = np.maximum(np.max(np.abs(chk), axis=2, keepdims=True), 1e-300)
        if not np.all(np.abs(rec - chk) / scale < info["tol"]):
            return None
        return c.T.reshape(nt, ne, r)

    def _per_item(self, values):
        values = np.asarray(values, dtype=np.float64)
        return float(values[0]) if self.batch == 1 else values

    def get_average(self, observable):
        """Sum Re W * f(X, P, t) dXdP (wigner_moyal_cuda_1d.py:329-347)."""
        out = empty((self.batch, 1), np.float64)
        self._obs_launch([observable], self.wignerfunction, float(self.t), out)
        return self._per_item(out.get()[:, 0] * self.dXdP)

    def get_purity(self):
        """2 pi Re dot(W, W) dXdP without conjugation (wigner_moyal_cuda_1d.py:349-354)."""
        s = self.plan.stats(self.wignerfunction).get()
        return self._per_item(2. * np.pi * s[:, 3] * self.dXdP)

    def get_sigma_x_sigma_p(self):
        """wigner_moyal_cuda_1d.py:356-366."""
        return np.sqrt((self.get_average("X * X") - self.get_average("X") ** 2)
                       * (self.get_average("P * P") - self.get_average("P") ** 2))

    @classmethod
    def print_memory_info(cls):
        """wigner_moyal_cuda_1d.py:368-377."""
        free, total = mem_info()
        print("\n\n\t\tGPU memory Total %.2f GB\n\t\tGPU memory Free %.2f GB\n"
              % (total / 2. ** 30, free / 2. ** 30))

    # ------------------------------------------------------------------------------------
    # initial state and Wigner <-> rho transforms
    # ------------------------------------------------------------------------------------
    def _init_kernel_source(self, expr):
        return (rtc.PRELUDE + self.cuda_consts + r"""
extern "C" __global__ void aw_init(%(st)s *W, long long n_item, long long n_total)
{
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n_total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long loc = idx %% n_item;
        const size_t i = loc / X_gridDIM, j = loc %% X_gridDIM;
        const double X = dX * (j - 0.5 * X_gridDIM);
        const double P = dP * (i - 0.5 * P_gridDIM);
        aw_state_store<%(st)s>::put(W + idx, cuda_complex(%(expr)s));
    }
}
""" % dict(st=self._state_t(), expr=expr))

    def set_wignerfunction(self, new_wigner_func):
        """wigner_moyal_cuda_1d.py:379-420: array, GPU array, Python function F(self, X, P),
        C string (X, P declared) or number; then W /= (complex) sum W dXdP and wigner2rho."""
        if isinstance(new_wigner_func, (np.ndarray, DeviceArray)):
            assert new_wigner_func.shape in (self.wignerfunction.shape,
                                              self.wignerfunction.shape[-2:]), \
                "The grid sizes does not match with the Wigner function"
            self.wignerfunction[:] = new_wigner_func
        elif isinstance(new_wigner_func, FunctionType):
            self.wignerfunction[:] = np.asarray(new_wigner_func(self, self.X, self.P),
                                                dtype=np.complex128) * np.ones(self.wignerfunction.shape)
        elif isinstance(new_wigner_func, str):
            mod = rtc.Module(self._init_kernel_source(new_wigner_func), "wigner_moyal_init.hip")
            n = self.X_gridDIM * self.P_gridDIM
            blocks, threads = rtc.launch_cfg(n * self.batch)
            mod.launch("aw_init", (blocks,), (threads,), [self.wignerfunction.ptr, ctypes.c_longlong(n),
                                                          ctypes.c_longlong(n * self.batch)],
                       current_stream())
        elif isinstance(new_wigner_func, (float, complex, int)):
            self.wignerfunction.fill(np.complex128(new_wigner_func))
        else:
            raise NotImplementedError("new_wigner_func must be either function or numpy.array")
        # W /= sum(W) dXdP with the complex sum (:415)
        st = self.plan.stats(self.wignerfunction)
        self.plan.scale_from_stats(st.ptr, 8, 0, 1, self.dXdP, self._dscale)
        self.plan.scale(self.wignerfunction, self._dscale)
        self._scale_pending = False
        self.wigner2rho()
        return self

    def _shearX(self, data, dir_a, dir_b, **kw):
        wm_ops.shear_x(self.plan, self._grid, data, dir_a, dir_b, **kw)

    def _shearY(self, data, dir_a, dir_b, **kw):
        wm_ops.shear_y(self.plan, self._grid, data, dir_a, dir_b, **kw)

    def wigner2rho(self):
        """wigner_moyal_cuda_1d.py:422-469 as four passes:
        [S, F0] [I1, SX, F1] [I0, SY, F0] [I1, SX, F1, S]."""
        wm_ops.wigner2rho(self.plan, self._grid, self.wignerfunction, self.rho)
        return self.rho

    def rho2wigner_(self):
        """wigner_moyal_cuda_1d.py:471-501 as four passes:
        [S, F1, SX, I1] [F0, SY, I0] [F1, SX, I1] [I0, S]."""
        self._flush_scale()
        wm_ops.rho2wigner(self.plan, self._grid, self.rho, self.wignerfunction)
        return self.wignerfunction

    rho2wigner = rho2wigner_

    def rho2wigner_blackman(self):
        """wigner_moyal_cuda_1d.py:503-538 (Blackman windows before each forward FFT)."""
        self._flush_scale()
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        bx = self._blackman(Nx)
        by = self._blackman(Np)
        # [S, Bx] then [F1, SX, I1]: the window multiplies before the FFT, so apply it in a
        # plain pass first
        self.plan.run(self.rho, 1, 0, 0, sign_in=1, elem_fac=bx, out=self.wignerfunction)
        self._shearX(self.wignerfunction, FWD, INV)
        self.plan.run(self.wignerfunction, 0, 0, 0, elem_fac=by)
        self._shearY(self.wignerfunction, FWD, INV)
        self.plan.run(self.wignerfunction, 1, 0, 0, elem_fac=bx)
        self._shearX(self.wignerfunction, FWD, INV)
        self.plan.run(self.wignerfunction, 0, 0, INV, elem_fac=by, sign_out=1)
        return self.wignerfunction

    def _blackman(self, n):
        """blackmanX / blackmanY (:678-698) as a complex vector in the state precision."""
        c = np.arange(n) * 2.0 * np.pi / (n - 1.)
        return to_device((0.42 - 0.5 * np.cos(c) + 0.08 * np.cos(2.0 * c)).astype(self.plan.factor_dtype))

    def _flush_scale(self):
        """Apply a pending device-side normalisation to rho."""
        if self._scale_pending:
            self.plan.scale(self.rho, self._dscale)
            self._scale_pending = False

    def _norm_weights(self):
        """L^H 1 (FP64 [1, Np, Nx]) with L = rho2wigner_: Re <L^H 1, rho> = Re sum W."""
        if self._g1 is None:
            keep = getattr(self, "_obs_plan", None)
            self._g1 = self._obs_weights(["1."])
            self._g1_plan, self._obs_plan = self._obs_plan, keep   # the plan outlives its queued launches
            self._norm_out = empty((self.batch, 1), np.float64)
        return self._g1

    def normalize_rho_wigner(self):
        """wigner_moyal_cuda_1d.py:608-622: W = rho2wigner; n = Re sum W dXdP; W /= n;
        rho /= n (device-side, no host sync).  n is read from rho as Re <L^H 1, rho> dXdP (the
        transform L is linear), one launch pair; W is then formed already divided by n, and the
        transform's first pass, which reads rho, also writes it back divided by n: no read of W
        and no scale passes after the transform."""
        if not self.obs_by_adjoint:
            self.rho2wigner()
            st = self.plan.stats(self.wignerfunction)
            self.plan.scale_from_stats(st.ptr, 8, 0, 0, self.dXdP, self._dscale)
            self.plan.scale(self.wignerfunction, self._dscale)
            self.plan.scale(self.rho, self._dscale)
            return self
        self._flush_scale()
        g1 = self._norm_weights()
        self.plan.weighted_sums(self.rho, g1, 1, self._norm_out, self.dXdP, self._dscale)
        wm_ops.rho2wigner(self.plan, self._grid, self.rho, self.wignerfunction, dscale=self._dscale,
                          scale_rho=True)
        return self

    # ------------------------------------------------------------------------------------
    # propagation
    # ------------------------------------------------------------------------------------
    def _times(self, steps):
        """t after each ``t += dt`` (:571), accumulated in float64 like the reference."""
        ts = np.empty(steps, np.float64)
        t = self.t
        for k in range(steps):
            t = t + self.dt
            ts[k] = t
        return ts

    def _tdep_masks(self):
        return any(m is not None and m[0] in ("time", "ttable") for m in self._masks.values())

    def _ttables(self):
        """Masks evaluated as 2D tables at every step's time (time-dependent, not separable)."""
        return any(m is not None and m[0] == "ttable" for m in self._masks.values())

    def _ttable(self, which, t_ptr):
        """The table of mask ``which`` at the time stored at device address t_ptr, in the plan's
        real type, in a persistent buffer (one launch, capturable), or None for other masks."""
        m = self._masks.get(which)
        if m is None or m[0] != "ttable":
            return None
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        out = self._ttable_buf(which)
        xj, xi = (self._dev_kXX, self._dev_kXXp) if which == 0 else (self._dev_kPP, self._dev_kPPp)
        f64 = np.dtype(self.plan.real_dtype) == np.float64
        blocks, threads = rtc.launch_cfg(Nx * Np)
        self._mod.launch("aw_eval_mask_at", (blocks,), (threads,),
                         [ctypes.c_int(which), xj.ptr, ctypes.c_int(Nx), xi.ptr, ctypes.c_int(Np),
                          ctypes.c_void_p(t_ptr), ctypes.c_void_p(out.ptr if f64 else None),
                          ctypes.c_void_p(None if f64 else out.ptr)], current_stream())
        return out

    def _ttable_buf(self, which):
        """Persistent buffer of a per-step mask table (allocated before any capture uses it)."""
        tabs = self.__dict__.setdefault("_tt_bufs", {})
        if which not in tabs:
            tabs[which] = empty((self.P_gridDIM, self.X_gridDIM), self.plan.real_dtype)
        return tabs[which]

    def _alloc_ttables(self):
        for which, m in self._masks.items():
            if m is not None and m[0] == "ttable":
                self._ttable_buf(which)

    def _ttable_chunk(self, n):
        """n unmerged steps [EV, I0] [F1, EK, I1] [F0, EV] from the chunk buffers, each with its
        mask tables evaluated at the step's time (the time expV / expK see, :753, :777)."""
        colL, colE, rowL, rowE = self._prop["fac"]
        ts = self._prop["ts"]
        csz = self.plan.fsize
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        for k in range(n):
            tx, tp = self._step_tables(ts.ptr + 8 * k)
            self.plan.steps(self.rho, 1, colL.ptr + 2 * k * Nx * csz, colE.ptr + 2 * k * Np * csz,
                            rowL.ptr + k * Np * csz, rowE.ptr + k * Nx * csz, mask_open_close=tx, mask_row=tp)

    def _step_tables(self, t_ptr):
        """(coordinate mask, momentum mask) tables of one unmerged step at the time at t_ptr:
        the static table of a mask, its table at that time, or None."""
        tx = self._ttable(0, t_ptr)
        tp = self._ttable(1, t_ptr)
        return (tx if tx is not None else self._mask_args(0, 1)[2],
                tp if tp is not None else self._mask_args(1, 1)[2])

    def _mask_tables(self, ts_dev, n, out=None):
        """Per-step vectors of the time-dependent masks at the step times ts_dev[:n] (t after
        the increment, the time expV / expK see): {which: (line [n][rows], elem [n][cols])},
        written into ``out`` when given."""
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        tabs = {} if out is None else out
        for which, m in self._masks.items():
            if m is None or m[0] != "time":
                continue
            if which not in tabs or tabs[which][0].shape[0] < n:
                tabs[which] = (empty((n, Np), np.float64), empty((n, Nx), np.float64))
            v, u = tabs[which]
            xj, xi = (self._dev_kXX, self._dev_kXXp) if which == 0 else (self._dev_kPP, self._dev_kPPp)
            blocks, threads = rtc.launch_cfg(n * (Nx + Np))
            self._mod.launch("aw_eval_mask_lines", (blocks,), (threads,),
                             [ctypes.c_int(which), xj.ptr, ctypes.c_int(Nx), xi.ptr, ctypes.c_int(Np), ts_dev.ptr,
                              ctypes.c_int(n), ctypes.c_int(m[1]), ctypes.c_int(m[2]), u.ptr, v.ptr], current_stream())
        return tabs

    def _mask_args(self, which, power):
        m = self._masks.get(which)
        if m is None or m[0] in ("time", "ttable"):
            return None, None, None
        if m[0] == "vec":
            return m[1], m[2], None          # (v over lines-i, u over cols-j)
        return None, None, (m[1] if power == 1 else m[2])

    def _h_static(self):
        """True when V, K and the masks cannot depend on t: no token ``t`` in the V and K
        expressions nor in a macro of the constants and functions, and no time-dependent mask.
        The factor vectors of every step are then the same and are built once."""
        return (not self._tdep_masks() and time_independent(self.V, self.cuda_consts)
                and time_independent(self.K, self.cuda_consts))

    def _build_chunk_factors(self, ts, merged=True, bufs=None, both=False):
        """The factor vectors of one chunk (``_build_factors``); with a time-independent H and
        persistent buffers, the constant vectors of any chunk, built on first use: merged
        open / mid / close column vectors 0, 1, 2 and row vector 0, or one unmerged step's open
        and close column vectors 0, 1 and row vector 0 (with ``both``, merged into bufs["fac"]
        and unmerged into bufs["facu"])."""
        if not (self._static and bufs is not None):
            return self._build_factors(ts, merged, bufs, both)
        key = ("merged" if merged else "split") + ("+split" if both else "")
        if bufs.get("static") != key:
            n = 2 if merged else 1
            self._build_factors(np.full(n, float(self.t) + self.dt), merged, bufs, both)
            bufs["static"] = key
        if ts is not None:
            self._stage_times(bufs, ts)    # the step times, for observables that read t
        return bufs["fac"]

    @staticmethod
    def _stage_times(bufs, ts):
        ts_dev, n = bufs["ts"], len(ts)
        if isinstance(ts, np.ndarray):
            ts_dev[:n] = ts
        else:
            ts_dev.tensor[:n].copy_(ts)   # a device slice: ordered on the stream, no host wait
        return ts_dev

    def _build_factors(self, ts, merged=True, bufs=None, both=False):
        """Phase tables (FP64, hipRTC) and the per-pass complex factor vectors of one chunk.
        merged: the closing half-kick of a step and the opening one of the next share one
        column pass (n + 1 column vectors); otherwise each step has its own opening and
        closing vectors (2 n column vectors, open at 2 s, close at 2 s + 1).  bufs: write
        the times, tables and factors into these preallocated arrays (``_chunk_bufs``);
        both: also the unmerged vectors, into bufs["facu"]."""
        n = len(ts)
        Nx, Np, dt = self.X_gridDIM, self.P_gridDIM, self.dt
        if bufs is None:
            ts_dev = to_device(ts)
            Vj, Vi = empty((n, Nx), np.float64), empty((n, Np), np.float64)
            Kj, Ki = empty((n, Nx), np.float64), empty((n, Np), np.float64)
            dst = None
        else:
            ts_dev = self._stage_times(bufs, ts)
            Vj, Vi, Kj, Ki = bufs["Vj"], bufs["Vi"], bufs["Kj"], bufs["Ki"]
            dst = bufs["fac"]
        self._eval_1d(0, self._dev_kXX, Nx, ts_dev, n, 0.5 * dt, Vj)
        self._eval_1d(0, self._dev_kXXp, Np, ts_dev, n, 0.5 * dt, Vi)
        self._eval_1d(1, self._dev_kPP, Nx, ts_dev, n, 0.5 * dt, Kj)
        self._eval_1d(1, self._dev_kPPp, Np, ts_dev, n, 0.5 * dt, Ki)
        mask_t = None
        if self._tdep_masks():
            mask_t = self._mask_tables(ts_dev, n, None if bufs is None else bufs.setdefault("mask_t", {}))
        if both:
            self._factors_from_tables(n, Vj.ptr, Vi.ptr, Kj.ptr, Ki.ptr, dt, False, dst=bufs["facu"], mask_t=mask_t)
        return self._factors_from_tables(n, Vj.ptr, Vi.ptr, Kj.ptr, Ki.ptr, dt, merged, dst=dst, mask_t=mask_t)

    def _chunk_bufs(self, n):
        """Persistent per-chunk inputs of the per-step loops (fixed addresses for capture)."""
        Nx, Np, p = self.X_gridDIM, self.P_gridDIM, self.plan
        return dict(ts=empty((n,), np.float64),
                    Vj=empty((n, Nx), np.float64), Vi=empty((n, Np), np.float64),
                    Kj=empty((n, Nx), np.float64), Ki=empty((n, Np), np.float64),
                    fac=(p.new_factors(Nx, 2 * n), p.new_factors(Np, 2 * n),
                         p.new_factors(Np, n), p.new_factors(Nx, n)))

    def _factors_from_tables(self, n, Vj, Vi, Kj, Ki, dt, merged=True, k_stride=None, dst=None, mask_t=None):
        """Factor vectors of n steps from FP64 phase tables (device pointers) V(XX_j),
        V(XX'_i) [n, Nx] / [n, Np] and K(PP_j), K(PP'_i) (k_stride 0: one row for all
        steps).  A negative dt gives the conjugated factors: the adjoint step.  dst:
        preallocated (colL, colE, rowL, rowE) with room for the vectors.  mask_t: per-step
        vectors of time-dependent masks (``_mask_tables``)."""
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        p = self.plan
        mxp, mx, _ = self._mask_args(0, 1)
        mpp, mp, _ = self._mask_args(1, 1)
        mxp_p = mxp.ptr if mxp is not None else None
        mx_p = mx.ptr if mx is not None else None
        mt = mask_t or {}
        if 0 in mt or 1 in mt:
            return self._factors_masks_t(n, Vj, Vi, Kj, Ki, dt, merged, k_stride, dst, mt, mx_p, mxp_p,
                                         mpp.ptr if mpp is not None else None, mp.ptr if mp is not None else None)
        if dst is not None:
            colL, colE, rowL, rowE = dst
        else:
            rowL = p.new_factors(Np, n)
            rowE = p.new_factors(Nx, n)
        # row passes: exp(-i dt [K(PP_j) - K(PP'_i)]) absp / Nx
        p.factors(rowL, Np, n, A=Ki, ca=dt, mag=mpp.ptr if mpp is not None else None,
                  scale=1.0 / Nx, a_stride=k_stride)
        p.factors(rowE, Nx, n, A=Kj, ca=-dt, mag=mp.ptr if mp is not None else None, a_stride=k_stride)
        if not merged:
            if dst is None:
                colL = p.new_factors(Nx, 2 * n)
                colE = p.new_factors(Np, 2 * n)
            for v, scale in ((0, 1.0), (1, 1.0 / Np)):
                p.factors(colL, Nx, n, A=Vj, ca=-0.5 * dt, mag=mx_p, alt_sign=True,
                          out_offset=v, out_vstride=2)
                p.factors(colE, Np, n, A=Vi, ca=0.5 * dt, mag=mxp_p, alt_sign=True, scale=scale,
                          out_offset=v, out_vstride=2)
            return colL, colE, rowL, rowE
        if dst is None:
            colL = p.new_factors(Nx, n + 1)
            colE = p.new_factors(Np, n + 1)
        # column passes: v = 0 start EV(t_0) (with S), 1..n-1 merged EV EV', n end EV (with S)
        p.factors(colL, Nx, 1, A=Vj, ca=-0.5 * dt, mag=mx_p, alt_sign=True)
        p.factors(colE, Np, 1, A=Vi, ca=0.5 * dt, mag=mxp_p, alt_sign=True)
        if n > 1:
            p.factors(colL, Nx, n - 1, A=Vj, ca=-0.5 * dt, B=Vj + Nx * 8, cb=-0.5 * dt,
                      mag=mx_p, mag_pow=2, out_offset=1)
            p.factors(colE, Np, n - 1, A=Vi, ca=0.5 * dt, B=Vi + Np * 8, cb=0.5 * dt,
                      mag=mxp_p, mag_pow=2, scale=1.0 / Np, out_offset=1)
        p.factors(colL, Nx, 1, A=Vj + (n - 1) * Nx * 8, ca=-0.5 * dt, mag=mx_p, alt_sign=True,
                  out_offset=n)
        p.factors(colE, Np, 1, A=Vi + (n - 1) * Np * 8, ca=0.5 * dt, mag=mxp_p, alt_sign=True,
                  scale=1.0 / Np, out_offset=n)
        return colL, colE, rowL, rowE

    def _factors_masks_t(self, n, Vj, Vi, Kj, Ki, dt, merged, k_stride, dst, mt, mx_p, mxp_p, mpp_p, mp_p):
        """_factors_from_tables with at least one time-dependent mask: the step's mask vectors
        (mag, one per step, stride one vector) multiply its factor vectors; a merged column
        vector takes the closing step's mask times the opening step's."""
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        p = self.plan

        def mags(which, static_line, static_elem, nl, ne):
            """(line mag, line stride, elem mag, elem stride) of n per-step vectors."""
            if which in mt:
                v, u = mt[which]
                return v, nl, u, ne
            return static_line, 0, static_elem, 0

        pl, pls, pe, pes = mags(1, mpp_p, mp_p, Np, Nx)
        xl, xls, xe, xes = mags(0, mxp_p, mx_p, Np, Nx)
        ptr = lambda a: a.ptr if isinstance(a, DeviceArray) else a   # noqa: E731
        if dst is not None:
            colL, colE, rowL, rowE = dst
        else:
            rowL, rowE = p.new_factors(Np, n), p.new_factors(Nx, n)
        p.factors(rowL, Np, n, A=Ki, ca=dt, mag=ptr(pl), mag_stride=pls, scale=1.0 / Nx, a_stride=k_stride)
        p.factors(rowE, Nx, n, A=Kj, ca=-dt, mag=ptr(pe), mag_stride=pes, a_stride=k_stride)
        if not merged:
            if dst is None:
                colL, colE = p.new_factors(Nx, 2 * n), p.new_factors(Np, 2 * n)
            for v, scale in ((0, 1.0), (1, 1.0 / Np)):
                p.factors(colL, Nx, n, A=Vj, ca=-0.5 * dt, mag=ptr(xe), mag_stride=xes, alt_sign=True,
                          out_offset=v, out_vstride=2)
                p.factors(colE, Np, n, A=Vi, ca=0.5 * dt, mag=ptr(xl), mag_stride=xls, alt_sign=True, scale=scale,
                          out_offset=v, out_vstride=2)
            return colL, colE, rowL, rowE
        if dst is None:
            colL, colE = p.new_factors(Nx, n + 1), p.new_factors(Np, n + 1)

        def at(a, s, size):
            return None if a is None else ptr(a) + (s * size * 8 if isinstance(a, DeviceArray) else 0)

        p.factors(colL, Nx, 1, A=Vj, ca=-0.5 * dt, mag=at(xe, 0, Nx), alt_sign=True)
        p.factors(colE, Np, 1, A=Vi, ca=0.5 * dt, mag=at(xl, 0, Np), alt_sign=True)
        if n > 1:
            # merged kicks: the closing step's mask times the opening step's
            if isinstance(xe, DeviceArray):
                me = DeviceArray(None, np.float64, tensor=xe.tensor[:n - 1] * xe.tensor[1:n])
                ml = DeviceArray(None, np.float64, tensor=xl.tensor[:n - 1] * xl.tensor[1:n])
                kw_e, kw_l = dict(mag=me.ptr, mag_stride=Nx), dict(mag=ml.ptr, mag_stride=Np)
            else:
                kw_e, kw_l = dict(mag=xe, mag_pow=2), dict(mag=xl, mag_pow=2)
            p.factors(colL, Nx, n - 1, A=Vj, ca=-0.5 * dt, B=Vj + Nx * 8, cb=-0.5 * dt, out_offset=1, **kw_e)
            p.factors(colE, Np, n - 1, A=Vi, ca=0.5 * dt, B=Vi + Np * 8, cb=0.5 * dt, scale=1.0 / Np, out_offset=1,
                      **kw_l)
        p.factors(colL, Nx, 1, A=Vj + (n - 1) * Nx * 8, ca=-0.5 * dt, mag=at(xe, n - 1, Nx), alt_sign=True,
                  out_offset=n)
        p.factors(colE, Np, 1, A=Vi + (n - 1) * Np * 8, ca=0.5 * dt, mag=at(xl, n - 1, Np), alt_sign=True,
                  scale=1.0 / Np, out_offset=n)
        return colL, colE, rowL, rowE

    def _run_chunk(self, ts, dscale=None):
        """Steps with evaluation times ts (t after the increment) as 2 n + 1 fused passes."""
        n = len(ts)
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        colL, colE, rowL, rowE = self._build_chunk_factors(ts)
        csz = self.plan.fsize
        _, _, tab1 = self._mask_args(0, 1)
        _, _, tab2 = self._mask_args(0, 2)
        _, _, tabp = self._mask_args(1, 1)
        if self._ttables():
            # one step (single_step_propagation): the mask tables at its time
            if n != 1:
                raise ValueError("steps with tabulated time-dependent masks run one at a time")
            self._alloc_ttables()
            tdev = to_device(np.asarray(ts, np.float64))
            tab1, tabp = self._step_tables(tdev.ptr)
        p, rho = self.plan, self.rho
        ev = self._pass_events
        if ev is not None:
            import torch
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        p.steps(rho, n, colL, colE, rowL, rowE, mask_open_close=tab1, mask_mid=tab2, mask_row=tabp,
                dscale=dscale)
        if ev is not None:
            e1.record()
            ev.append((e0, e1, 2 * n + 1))

    def single_step_propagation(self):
        """wigner_moyal_cuda_1d.py:543-561 (unnormalised result in self.rho)."""
        self._flush_scale()
        self._run_chunk(np.array([self.t], np.float64))
        return self.rho

    def propagate(self, steps=1):
        """wigner_moyal_cuda_1d.py:563-582: t += dt before each step; Ehrenfest lists every
        step when diff_K and diff_V are given, otherwise one normalisation at the end."""
        steps = int(steps)
        if steps <= 0:
            if not self.isEhrenfest:
                self.normalize_rho_wigner()
            return self.wignerfunction
        ts = self._times(steps)
        # the step times of the whole call in one upload; each chunk copies its slice on the
        # device (a host upload per chunk would wait for the chunk before it)
        ts_all = None if (self._static and not self.isEhrenfest) else to_device(ts)
        if not self.isEhrenfest:
            self._flush_scale()
            CH = CHUNK_STEPS
            if self._prop is None:
                self._prop = dict(self._chunk_bufs(CH), graphs={})
            if self._ttables():
                # masks tabulated at every step's time: unmerged steps, three passes each
                self._alloc_ttables()
                for c0 in range(0, steps, CH):
                    n = min(CH, steps - c0)
                    self._build_chunk_factors(ts_all.tensor[c0:c0 + n], merged=False, bufs=self._prop)
                    self._timed_passes(lambda n=n: self._ttable_chunk(n), 3 * n)
                self.t = ts[-1]
                self.normalize_rho_wigner()
                return self.wignerfunction
            # latency-bound sizes (the small-pass kernels) replay a captured chunk
            small = self.batch * max(self.X_gridDIM, self.P_gridDIM) <= 2048
            normalized = False
            for c0 in range(0, steps, CH):
                n = min(CH, steps - c0)
                self._build_chunk_factors(None if ts_all is None else ts_all.tensor[c0:c0 + n], merged=True,
                                          bufs=self._prop)
                groups = self._item_groups(n) if (c0 + n == steps and self.obs_by_adjoint) else None
                if groups:
                    # the last chunk of a batch that outgrows the MALL: each group of items runs
                    # its steps, then at once its normalisation and Wigner transform, while it
                    # is still in the MALL
                    self._flush_scale()
                    for b0, nb in groups:
                        self._timed_passes(lambda b0=b0, nb=nb: self._merged_chunk(n, b0, nb),
                                           (2 * n + 1) * nb / self.batch)
                        self._normalize_items(b0, nb)
                    normalized = True
                    continue
                self._timed_passes(lambda n=n: graphs.run_chunk(self._prop["graphs"], "merged",
                                                                lambda: self._merged_chunk(n), n == CH and small),
                                   2 * n + 1)
            self.t = ts[-1]
            if not normalized:
                self.normalize_rho_wigner()
            return self.wignerfunction
        # Ehrenfest: per step normalise through the Wigner transform and record
        # <X>, <diff_K>, <P>, -<diff_V>, <K + V> (:584-606) into a device ring; one host copy
        wanted = ["1.", "X", self.diff_K, "P", self.diff_V, self.hamiltonian]
        # each distinct expression is summed once (config 1 has diff_K = "P" and diff_V = "X":
        # 4 weight arrays instead of 6)
        exprs = list(dict.fromkeys(e.strip() for e in wanted))
        col = [exprs.index(e.strip()) for e in wanted]
        nobs = len(exprs)
        ring = empty((steps, self.batch, nobs), np.float64)
        CH = CHUNK_STEPS
        if self._ehr is None:
            self._ehr = dict(self._chunk_bufs(CH), stage=empty((CH, self.batch, nobs), np.float64),
                             graphs={})
            self._observable_kernel(exprs)
            self._obs_partials_for(nobs)             # allocated here, never inside a capture
            self._alloc_ttables()
            # (time-dependent masks change the closing half-kick every step: no merged steps)
            g = self._obs_weights(exprs) if (self.obs_by_adjoint and not self._tdep_masks()) else None
            # time-independent observables imply a time-independent V (the Hamiltonian is one of
            # them): the steps are merged and the observables read from each row-pass output
            self._ehr["weights"] = self._closing_adjoint(g) if g is not None else None
            self._ehr["tsep"] = None
            if g is not None:
                # block partials of every step of a chunk, added by one launch per chunk
                nblk = self.plan.row_obs_blocks()
                self._ehr["wring"] = empty((CH, self.batch, nblk, nobs), np.float64)
                self._ehr["wnblk"] = nblk
            elif (self.obs_by_adjoint and self._mask_args(0, 1)[2] is None and not self._ttables()
                  and self.plan.obs_blocks(0) > 0):
                # observables that depend on t through a few functions of t: merged steps whose
                # column pass records the closed state's sums against time-independent weights
                tsep = self._tsep_setup(exprs)
                if tsep is not None:
                    nblk = self.plan.obs_blocks(0)
                    r = tsep["nb"]
                    self._ehr.update(tsep=tsep, tring=empty((CH, self.batch, nblk, r), np.float64), tnblk=nblk,
                                     tstage=empty((CH, self.batch, r), np.float64),
                                     facu=(self.plan.new_factors(self.X_gridDIM, 2 * CH),
                                           self.plan.new_factors(self.P_gridDIM, 2 * CH),
                                           self.plan.new_factors(self.P_gridDIM, CH),
                                           self.plan.new_factors(self.X_gridDIM, CH)))
        bufs = self._ehr
        coef = self._tsep_coeffs(exprs, ts) if bufs["tsep"] is not None else None
        if coef is not None:
            return self._propagate_tsep(exprs, col, steps, ts, ts_all, coef)
        merged = bufs["weights"] is not None
        if merged:
            # the steps are linear, so the reference's per-step normalisation (:584-622) is
            # applied once at the end; the recorded averages are ratios to the recorded norm
            self._flush_scale()
        elif not self._scale_pending:
            # every step applies the pending scale: start from 1 so that all chunks are alike
            self._dscale.tensor[:, 0] = 1.0
            self._dscale.tensor[:, 1] = 0.0
            self._scale_pending = True
        chunk = self._ehrenfest_merged_chunk if merged else self._ehrenfest_chunk
        for c0 in range(0, steps, CH):
            n = min(CH, steps - c0)
            self._build_chunk_factors(ts_all.tensor[c0:c0 + n], merged=merged, bufs=bufs)
            graphs.run_chunk(bufs["graphs"], (merged,) + tuple(exprs), lambda n=n: chunk(exprs, n), n == CH)
            ring.tensor[c0:c0 + n].copy_(bufs["stage"].tensor[:n])
        self.t = ts[-1]
        # final normalisation of rho and W by the last step's norm
        if merged:
            self.normalize_rho_wigner()  # W (not formed during the steps), then W and rho / norm
        else:
            self.plan.scale(self.wignerfunction, self._dscale)
            self._flush_scale()
        r = ring.get()
        s0 = r[:, :, 0]
        for lst, k, sgn in ((self.X_average, 1, 1.), (self.X_average_RHS, 2, 1.),
                            (self.P_average, 3, 1.), (self.P_average_RHS, 4, -1.),
                            (self.hamiltonian_average, 5, 1.)):
            vals = sgn * r[:, :, col[k]] / s0
            lst.extend(vals[:, 0].tolist() if self.batch == 1 else list(vals))
        return self.wignerfunction

    def _propagate_tsep(self, exprs, col, steps, ts, ts_all, coef):
        """The Ehrenfest loop with time-separable observables (``_tsep_setup``): merged steps,
        each merged column pass storing the closed state's sums against the basis weights; the
        averages are formed on the host from the coefficients coef[s, e, m]; rho and W are
        normalised once at the end (the steps are linear, the averages ratios)."""
        bufs, CH = self._ehr, CHUNK_STEPS
        r = bufs["tsep"]["nb"]
        ring = empty((steps, self.batch, r), np.float64)
        self._flush_scale()
        for c0 in range(0, steps, CH):
            n = min(CH, steps - c0)
            self._build_chunk_factors(ts_all.tensor[c0:c0 + n], merged=True, bufs=bufs, both=True)
            graphs.run_chunk(bufs["graphs"], ("tsep",) + tuple(exprs), lambda n=n: self._ehrenfest_tsep_chunk(n),
                             n == CH)
            ring.tensor[c0:c0 + n].copy_(bufs["tstage"].tensor[:n])
        self.t = ts[-1]
        self.normalize_rho_wigner()
        S = ring.get()                                           # [steps, batch, r]
        vals = np.einsum("sem,sbm->sbe", coef, S)                # [steps, batch, nexprs]
        s0 = vals[:, :, 0]
        for lst, k, sgn in ((self.X_average, 1, 1.), (self.X_average_RHS, 2, 1.),
                            (self.P_average, 3, 1.), (self.P_average_RHS, 4, -1.),
                            (self.hamiltonian_average, 5, 1.)):
            v = sgn * vals[:, :, col[k]] / s0
            lst.extend(v[:, 0].tolist() if self.batch == 1 else list(v))
        return self.wignerfunction

    def _ehrenfest_tsep_chunk(self, n):
        """n merged steps, 2 n + 1 passes, [EV_0 -> I0] ([F1 -> EK_s -> I1] [F0 -> EV_s EV_s+1 -> I0])*
        ending [F0 -> EV_n-1]: each merged column pass also multiplies its values after F0 by
        step s's closing vectors (rho_s exactly as an unmerged closing pass stores it) and
        stores the block partials of Re <L^H b_m, rho_s> (obs_mode 3) in a ring; one launch per
        chunk adds them.  Only launches, from fixed addresses."""
        bufs, p, rho = self._ehr, self.plan, self.rho
        colL, colE, rowL, rowE = bufs["fac"]
        uL, uE = bufs["facu"][0], bufs["facu"][1]
        _, _, tabp = self._mask_args(1, 1)
        csz = p.fsize
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        G, r = bufs["tsep"]["G"], bufs["tsep"]["nb"]
        ring, nblk = bufs["tring"], bufs["tnblk"]
        ring_row = self.batch * nblk * r * 8
        p.run(rho, 0, 0, INV, line_fac=colL, elem_fac=colE)
        for k in range(n):
            last = k == n - 1
            # vectors of step k: row k, merged column k + 1, unmerged closing 2 k + 1 (row 0,
            # merged 1 or 2 and closing 1 with a time-independent H)
            kr, kc, ku = (0, 2 if last else 1, 1) if self._static else (k, k + 1, 2 * k + 1)
            p.run(rho, 1, FWD, INV, line_fac=rowL.ptr + kr * Np * csz, elem_fac=rowE.ptr + kr * Nx * csz, mask=tabp)
            p.run(rho, 0, FWD, 0 if last else INV, line_fac=colL.ptr + kc * Nx * csz,
                  elem_fac=colE.ptr + kc * Np * csz, obs_mode=3, obs_weights=G, obs_nw=r,
                  obs_partials=ring.ptr + k * ring_row, obs_line_fac=uL.ptr + ku * Nx * csz,
                  obs_elem_fac=uE.ptr + ku * Np * csz)
        finish_sums(ring, nblk, r, n * self.batch, bufs["tstage"])

    def _merged_chunk(self, n, b0=0, nb=0):
        """n merged steps (2 n + 1 passes) from the chunk buffers, over the items b0 .. b0 + nb
        (nb = 0: all)."""
        colL, colE, rowL, rowE = self._prop["fac"]
        _, _, tab1 = self._mask_args(0, 1)
        _, _, tab2 = self._mask_args(0, 2)
        _, _, tabp = self._mask_args(1, 1)
        self.plan.steps(self.rho.ptr + b0 * self._item_bytes(), n, colL, colE, rowL, rowE, mask_open_close=tab1,
                        mask_mid=tab2, mask_row=tabp, batch=nb, const_steps=self._static)

    def _item_bytes(self):
        return self.X_gridDIM * self.P_gridDIM * np.dtype(self.state_dtype).itemsize

    def _timed_passes(self, launch, npass):
        """Run ``launch``; when a caller collects pass timings (_pass_events), bracket it with
        events and count it as npass passes over the whole batch."""
        ev = self._pass_events
        if ev is None:
            launch()
            return
        import torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        launch()
        e1.record()
        ev.append((e0, e1, npass))

    def _item_groups(self, n):
        """[(first item, items)] of the MALL-resident groups in which the step sequence runs
        a batch for n steps, or None when it runs the batch whole."""
        g = self.plan.step_group(n, self.batch)
        if g >= self.batch:
            return None
        return [(b0, min(g, self.batch - b0)) for b0 in range(0, self.batch, g)]

    def _normalize_items(self, b0, nb):
        """normalize_rho_wigner on the items b0 .. b0 + nb: the norm from rho, W formed already
        divided by it, and rho divided (a plan of nb items over the group's slices)."""
        gp = self._group_plans.get(nb)
        if gp is None:
            gp = self._group_plans[nb] = Plan(self.P_gridDIM, self.X_gridDIM, nb, self.precision)
        off = b0 * self._item_bytes()
        rho, W = self.rho.ptr + off, self.wignerfunction.ptr + off
        ds = self._dscale.ptr + 16 * b0
        gp.weighted_sums(rho, self._norm_weights(), 1, self._norm_out.ptr + 8 * b0, self.dXdP, ds)
        wm_ops.rho2wigner(gp, self._grid, rho, W, dscale=ds, scale_rho=True)

    def _ehrenfest_chunk(self, exprs, n):
        """n steps, each [EV, I0] [F1, EK, I1] [F0, EV] from the chunk buffers (applying the
        pending normalisation), then the Wigner transform, the observables into the stage
        ring and the next normalisation scale.  Only launches, from fixed addresses."""
        bufs, p, rho = self._ehr, self.plan, self.rho
        colL, colE, rowL, rowE = bufs["fac"]
        _, _, tab1 = self._mask_args(0, 1)
        _, _, tabp = self._mask_args(1, 1)
        csz = self.plan.fsize
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        nobs = len(exprs)
        row_bytes = self.batch * nobs * 8
        stage, ts = bufs["stage"], bufs["ts"]
        ttab = self._ttables()
        for k in range(n):
            # one step = three passes; its opening / closing vectors sit at 2 k and 2 k + 1
            # (at 0 and 1 for every step with a time-independent H)
            kc = 0 if self._static else k
            if ttab:
                # masks tabulated at the step's time
                tab1, tabp = self._step_tables(ts.ptr + 8 * k)
            p.steps(rho, 1, colL.ptr + 2 * kc * Nx * csz, colE.ptr + 2 * kc * Np * csz,
                    rowL.ptr + kc * Np * csz, rowE.ptr + kc * Nx * csz, mask_open_close=tab1,
                    mask_row=tabp, dscale=self._dscale)
            self._scale_pending = False
            self.rho2wigner_()
            # observables into the stage row and the next scale 1 / (norm dXdP), one launch
            self._obs_fin_launch(exprs, self.wignerfunction, ts.ptr + 8 * k, stage.ptr + k * row_bytes,
                                 self.dXdP, self._dscale)
            self._scale_pending = True

    def _ehrenfest_merged_chunk(self, exprs, n):
        """n steps with time-independent observables (and so a time-independent V), merged
        like the plain loop: [EV_0 -> I0] ([F1 -> EK_s -> I1] [sums] [F0 -> EV EV -> I0])*
        ending [F0 -> EV].  The step-complete rho_s = C1 F0(Y_s), with Y_s the row-pass output
        and C1 the closing half-kick, so its observables Re <g, rho_s> = Re <g', Y_s> with
        g' = I0(conj(C1) g) (``_closing_adjoint``).  Each step leaves block partials in a
        ring; one launch per chunk adds them into the stage rows.  No per-step
        normalisation: the steps are linear and the averages are ratios to the recorded norm.
        Two launches per step: the row pass stores the partials as it writes Y.  Only
        launches, from fixed addresses."""
        bufs, p, rho = self._ehr, self.plan, self.rho
        colL, colE, rowL, rowE = bufs["fac"]
        _, _, tab1 = self._mask_args(0, 1)
        _, _, tab2 = self._mask_args(0, 2)
        _, _, tabp = self._mask_args(1, 1)
        csz = self.plan.fsize
        Nx, Np = self.X_gridDIM, self.P_gridDIM
        nobs, nblk = len(exprs), bufs["wnblk"]
        gy, ring = bufs["weights"], bufs["wring"]
        ring_row = self.batch * nblk * nobs * 8
        p.run(rho, 0, 0, INV, line_fac=colL, elem_fac=colE, mask=tab1)
        for k in range(n):
            last = k == n - 1
            # vectors of step k: row k and column k + 1 (row 0 and column 1, or 2 when closing,
            # with a time-independent H)
            kr, kc = (0, 2 if last else 1) if self._static else (k, k + 1)
            # the row pass also leaves the block partials of Re <g', Y_k> in the ring
            p.run(rho, 1, FWD, INV, line_fac=rowL.ptr + kr * Np * csz, elem_fac=rowE.ptr + kr * Nx * csz,
                  mask=tabp, obs_weights=gy, obs_nw=nobs, obs_partials=ring.ptr + k * ring_row)
            p.run(rho, 0, FWD, 0 if last else INV, line_fac=colL.ptr + kc * Nx * csz,
                  elem_fac=colE.ptr + kc * Np * csz, mask=tab1 if last else tab2)
        finish_sums(ring, nblk, nobs, n * self.batch, bufs["stage"])

    def _closing_adjoint(self, g):
        """g' = I0(conj(C1) g) for each weight array g_q (FP64): C1 is the closing half-kick of
        a step, (-1)^(i+j) exp(-i dt/2 [V(XX_j) - V(XX'_i)]) / Np times the coordinate mask,
        as the merged sequence's last column vector builds it.  V must not depend on t."""
        Nx, Np, dt = self.X_gridDIM, self.P_gridDIM, self.dt
        p64 = self._obs_plan
        ts = to_device(np.array([float(self.t)]))
        Vj, Vi = empty((1, Nx), np.float64), empty((1, Np), np.float64)
        self._eval_1d(0, self._dev_kXX, Nx, ts, 1, 0.5 * dt, Vj)
        self._eval_1d(0, self._dev_kXXp, Np, ts, 1, 0.5 * dt, Vi)
        mxp, mx, _ = self._mask_args(0, 1)
        cl, ce = p64.new_factors(Nx, 1), p64.new_factors(Np, 1)
        # conjugates of the closing vectors: the phase coefficients change sign
        p64.factors(cl, Nx, 1, A=Vj.ptr, ca=0.5 * dt, mag=mx.ptr if mx is not None else None, alt_sign=True)
        p64.factors(ce, Np, 1, A=Vi.ptr, ca=-0.5 * dt, mag=mxp.ptr if mxp is not None else None,
                    alt_sign=True, scale=1.0 / Np)
        m = self._masks.get(0)
        table = to_device(self._eval_mask(0, float(self.t))) if m is not None and m[0] == "table" else None
        gy = g.copy()
        p64.run(gy, 0, 0, INV, line_fac=cl, elem_fac=ce, mask=table)
        return gy

    def get_Ehrenfest(self, t):
        """wigner_moyal_cuda_1d.py:584-606 (normalises, then appends the five averages)."""
        if self.isEhrenfest:
            self.normalize_rho_wigner()
            self.X_average.append(self.get_average("X"))
            self.X_average_RHS.append(self.get_average(self.diff_K))
            self.P_average.append(self.get_average("P"))
            self.P_average_RHS.append(-self.get_average(self.diff_V))
            self.hamiltonian_average.append(self.get_average(self.hamiltonian))
