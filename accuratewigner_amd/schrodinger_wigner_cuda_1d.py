"""SchrodingerWignerCUDA1D: drop-in MI355X implementation of the reference class
(schrodinger_wigner_cuda_1d.py:16-902): split-operator propagation of psi(x), imaginary-time
ground and excited states (Gram-Schmidt), Ehrenfest lists and the Wigner function of psi.

Step (schrodinger_wigner_cuda_1d.py:356-369): EV, F, EK, I, EV (no 1/N), with
EV = exp(-0.5 i dt V(X_j, t)) bx and EK = exp(-i dt K(P_j, t)) bp evaluated at t (not
t + dt/2, :805, :823) and P in FFT order (:803).  It runs as fused single-row passes
[EV EV' -> F] [EK -> I] (the closing and opening half-kicks of consecutive steps merged),
with the per-step normalisation psi /= ||psi|| sqrt(dX) (:388) applied lazily on the device.
get_wigner is implemented with the arrays and plans the reference forgot to create
(:244, :255-256).

Lines longer than 4096 points (e.g. the coherent strong-field driver's N = 32768,
coherent_strong_field_physics_cuda_1d.py:31) use a two-level transform on the same memory
viewed as [N1][N2] (n = n1 N2 + n2): column FFTs, twiddles exp(-2 pi i n2 k1 / N), row FFTs,
which leaves frequency k = k1 + N1 k2 at position k1 N2 + k2.  The kinetic factors are
evaluated on that permuted grid, so a step needs no transposes:
[EV -> col FFT -> twiddle] ([row FFT -> EK -> row IFFT -> twiddle*] [col IFFT -> EV EV' ->
col FFT -> twiddle])* ... [col IFFT -> EV].
"""
import ctypes
from types import FunctionType

import numpy as np

from . import _lib, graphs, rtc
from ._params import KeywordParams
from .device import DeviceArray, current_stream, empty, require_gpu, to_device
from .engine import FWD, INV, Plan, mem_info

SQRT2 = np.sqrt(2.0)
CHUNK_STEPS = 256


class SchrodingerWignerCUDA1D(KeywordParams):
    """Second-order split-operator propagator of the Schrodinger equation with
    H = K(p, t) + V(x, t) (reference: schrodinger_wigner_cuda_1d.py:16-25), on MI355X."""

    def __init__(self, **kwargs):
        """Keywords as schrodinger_wigner_cuda_1d.py:26-49, plus ``precision``."""
        kwargs = dict(kwargs)
        self.precision = kwargs.pop("precision", "fp64")
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
        kwargs.update(t_initial=self.t)
        self.t_initial = self.t
        self.t = np.float64(self.t)

        N = self.X_gridDIM
        # grids (:107-122)
        self.dX = 2. * self.X_amplitude / N
        self.X = np.linspace(-self.X_amplitude, self.X_amplitude - self.dX, N)
        self.P = np.fft.fftshift(np.fft.fftfreq(N, self.dX / (2. * np.pi)))
        self.dP = self.P[1] - self.P[0]
        self.X_wigner = self.X / SQRT2
        self.dXX = self.X_wigner[1] - self.X_wigner[0]
        self.P_wigner = self.P / SQRT2
        self.dPP = self.P_wigner[1] - self.P_wigner[0]
        self.wigner_dXdP = self.dX * self.dP
        j = np.arange(N)
        # kernel coordinates from the #define'd steps (:803, :821)
        self._kX = rtc.kconst(self.dX) * (j - 0.5 * N)
        self._kP = rtc.kconst(self.dP) * (((j + N // 2) % N) - 0.5 * N)
        # lines longer than 4096 points: two-level transform over [N1][N2]
        self._two = None
        self._kP_fac = self._kP
        if N > 4096:
            lg = int(np.log2(N))
            N1 = 2 ** ((lg + 1) // 2)
            N2 = N // N1
            if N2 < 64 or N1 > 4096 or N2 > 4096:
                raise ValueError("X_gridDIM = %d is outside the supported range" % N)
            self._two = (N1, N2)
            # momentum at the position where the two-level FFT leaves frequency k
            m = np.arange(N)
            self._kP_fac = self._kP[(m // N2) + N1 * (m % N2)]

        kwargs.update(dX=self.dX, dP=self.dP, P_max=self.P.max())
        self.cuda_consts = rtc.cuda_consts(kwargs, getattr(self, "consts", None), None)
        if not hasattr(self, "abs_boundary_x"):
            self.abs_boundary_x = "1."
        if not hasattr(self, "abs_boundary_p"):
            self.abs_boundary_p = "1."
        if not hasattr(self, "max_thread_block"):
            self.max_thread_block = 512
        self.cuda_consts += getattr(self, "functions", "") or ""

        require_gpu()
        self.plan = Plan(1, N, 1, self.precision) if self._two is None else Plan(self._two[0], self._two[1], 1,
                                                                                  self.precision)
        self.state_dtype = self.plan.state_dtype
        self._wplan = None
        self._blackman = None
        self._mod = rtc.Module(self._physics_source(), "schrodinger_physics.hip")
        self._dev_kX = to_device(self._kX)
        self._dev_kP = to_device(self._kP_fac)
        self.wavefunction = DeviceArray((1, N), self.state_dtype)
        self._tmp = DeviceArray((1, N), self.state_dtype)
        self._dscale = empty((1, 2), np.float64)
        self._norm_st = empty((1, 8), np.float64)
        self._scale_pending = False
        self._obs_cache = {}
        self._masks()
        self.isEhrenfest = hasattr(self, "diff_K") and hasattr(self, "diff_V")
        if self.isEhrenfest:
            self.X_average, self.P_average = [], []
            self.X_average_RHS, self.P_average_RHS = [], []
            self.hamiltonian_average = []
        self.print_memory_info()

    # ------------------------------------------------------------------------------------
    def _state_t(self):
        return "double2" if self.state_dtype == np.complex128 else "float2"

    def _physics_source(self):
        return (rtc.PRELUDE + self.cuda_consts
                + "\n__device__ double K(double P, double t)\n{\n    return (%s);\n}\n" % self.K
                + "\n__device__ double V(double X, double t)\n{\n    return (%s);\n}\n" % self.V
                + rtc.EVAL_1D + r"""
extern "C" __global__ void aw_eval_masks(const double *Xg, const double *Pg, int n, double *bx, double *bp)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        { const double X = Xg[k]; bx[k] = (%s); }
        { const double P = Pg[k]; bp[k] = (%s); }
    }
}
extern "C" __global__ void aw_outer(const %s *psi, %s *rho, int n)
{
    // psi2rho (schrodinger_wigner_cuda_1d.py:680-687): rho[i, j] = psi[j] conj(psi[i])
    const long long total = (long long)n * n;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long i = idx / n, j = idx %% n;
        const cuda_complex a((double)psi[j].x, (double)psi[j].y), b((double)psi[i].x, -(double)psi[i].y);
        aw_state_store<%s>::put(rho + idx, a * b);
    }
}
""" % (self.abs_boundary_x, self.abs_boundary_p, self._state_t(), self._state_t(), self._state_t()))

    def _masks(self):
        N = self.X_gridDIM
        self._bx = self._bp = None
        if self.abs_boundary_x.strip() in ("1.", "1", "1.0") and self.abs_boundary_p.strip() in ("1.", "1", "1.0"):
            return
        bx, bp = empty((N,), np.float64), empty((N,), np.float64)
        blocks, threads = rtc.launch_cfg(N)
        self._mod.launch("aw_eval_masks", (blocks,), (threads,),
                         [self._dev_kX.ptr, self._dev_kP.ptr, ctypes.c_int(N), bx.ptr, bp.ptr], current_stream())
        if self.abs_boundary_x.strip() not in ("1.", "1", "1.0"):
            self._bx = bx
        if self.abs_boundary_p.strip() not in ("1.", "1", "1.0"):
            self._bp = bp

    def _eval_1d(self, which, xdev, n, ts_dev, nt, tshift, out):
        blocks, threads = rtc.launch_cfg(n * nt)
        self._mod.launch("aw_eval_1d", (blocks,), (threads,),
                         [ctypes.c_int(which), xdev.ptr, ctypes.c_int(n), ts_dev.ptr,
                          ctypes.c_int(nt), ctypes.c_double(tshift), out.ptr], current_stream())

    @classmethod
    def print_memory_info(cls):
        """schrodinger_wigner_cuda_1d.py:305-314."""
        free, total = mem_info()
        print("\n\n\t\tGPU memory Total %.2f GB\n\t\tGPU memory Free %.2f GB\n"
              % (total / 2. ** 30, free / 2. ** 30))

    # ------------------------------------------------------------------------------------
    # norms and observables
    # ------------------------------------------------------------------------------------
    def _norm_scale(self, factor=1.0, target=None):
        """dscale = 1 / (||target|| * factor) on the device (the cu_linalg.norm divisions).
        The sums use the plan's own block partials and a persistent output row: nothing is
        allocated per call."""
        target = self.wavefunction if target is None else target
        st = self.plan.stats(target, out=self._norm_st)
        mod = self._obs_module([])
        mod.launch("aw_inv_norm", (1,), (1,), [st.ptr, ctypes.c_double(factor), self._dscale.ptr],
                   current_stream())
        return st

    def _normalize(self, factor=1.0, target=None):
        target = self.wavefunction if target is None else target
        self._flush_scale()
        self._norm_scale(factor, target)
        self.plan.scale(target, self._dscale)

    def _flush_scale(self):
        if self._scale_pending:
            self.plan.scale(self.wavefunction, self._dscale)
            self._scale_pending = False

    def _obs_module(self, exprs):
        key = tuple(exprs)
        mod = self._obs_cache.get(key)
        if mod is None:
            nobs = max(len(exprs), 1)
            body = "".join("        acc[%d] += w * (%s);\n" % (q, e) for q, e in enumerate(exprs))
            src = (rtc.PRELUDE + self.cuda_consts
                   + "\n__device__ double K(double P, double t)\n{\n    return (%s);\n}\n" % self.K
                   + "\n__device__ double V(double X, double t)\n{\n    return (%s);\n}\n" % self.V
                   + r"""
extern "C" __global__ void aw_weighted(const %(st)s *psi, int n, double t, int perm_n2, double *out)
{
    // out[q] = sum_j |psi_j|^2 f_q(X_j, P_j, t) (weighted_func_cuda_code, :893-901); one block.
    // perm_n2 > 0: psi is a two-level FFT output, position j holds frequency
    // (j / perm_n2) + (n / perm_n2) (j %% perm_n2), which sets P
    double acc[%(nobs)d];
    for (int q = 0; q < %(nobs)d; ++q) acc[q] = 0.;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int kp = perm_n2 > 0 ? (j / perm_n2) + (n / perm_n2) * (j %% perm_n2) : j;
        const double X = dX * (j - 0.5 * X_gridDIM);
        const double P = dP * ((kp + X_gridDIM / 2) %% X_gridDIM - 0.5 * X_gridDIM);
        const double w = (double)psi[j].x * psi[j].x + (double)psi[j].y * psi[j].y;
%(body)s    }
    __shared__ double red[4][%(nobs)d];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = 0; c < %(nobs)d; ++c) {
        const double s = aw_wave_sum(acc[c]);
        if (lane == 0) red[wid][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < %(nobs)d) {
        double s = 0.;
        for (int w = 0; w < 4; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}
// the same sums as block partials of a grid-stride loop, the time read from device memory
// (ts): partials[block][q]; aw_weighted_fin adds the blocks in order into out[q]
extern "C" __global__ void aw_weighted_mb(const %(st)s *psi, int n, const double *ts, int perm_n2,
                                          double *partials)
{
    const double t = *ts;
    double acc[%(nobs)d];
    for (int q = 0; q < %(nobs)d; ++q) acc[q] = 0.;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const int kp = perm_n2 > 0 ? (j / perm_n2) + (n / perm_n2) * (j %% perm_n2) : j;
        const double X = dX * (j - 0.5 * X_gridDIM);
        const double P = dP * ((kp + X_gridDIM / 2) %% X_gridDIM - 0.5 * X_gridDIM);
        const double w = (double)psi[j].x * psi[j].x + (double)psi[j].y * psi[j].y;
%(body)s    }
    __shared__ double red[4][%(nobs)d];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = 0; c < %(nobs)d; ++c) {
        const double s = aw_wave_sum(acc[c]);
        if (lane == 0) red[wid][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < %(nobs)d) {
        double s = 0.;
        for (int w = 0; w < 4; ++w) s += red[w][threadIdx.x];
        partials[blockIdx.x * %(nobs)d + threadIdx.x] = s;
    }
}
extern "C" __global__ void aw_weighted_fin(const double *partials, int nblk, double *out)
{
    if (threadIdx.x < %(nobs)d) {
        double s = 0.;
        for (int b = 0; b < nblk; ++b) s += partials[b * %(nobs)d + threadIdx.x];
        out[threadIdx.x] = s;
    }
}
extern "C" __global__ void aw_inv_norm(const double *stats, double factor, double *dscale)
{
    dscale[0] = 1.0 / (sqrt(stats[2]) * factor);
    dscale[1] = 0.0;
}
""" % dict(st=self._state_t(), nobs=nobs, body=body))
            mod = rtc.Module(src, "schrodinger_obs.hip")
            self._obs_cache[key] = mod
        return mod

    def _weighted(self, exprs, out_ptr, psi=None, t=None, p_rep=False):
        """p_rep: psi comes from ``_fft`` (permuted order for two-level lines)."""
        psi = self.wavefunction if psi is None else psi
        mod = self._obs_module(exprs)
        perm = self._two[1] if (p_rep and self._two is not None) else 0
        mod.launch("aw_weighted", (1,), (256,), [psi.ptr, ctypes.c_int(self.X_gridDIM),
                                                ctypes.c_double(float(self.t if t is None else t)),
                                                ctypes.c_int(perm), out_ptr],
                   current_stream())

    def _wblocks(self):
        return max(1, min(64, self.X_gridDIM // 1024))

    def _weighted_dev(self, exprs, out_ptr, t_ptr, partials, psi=None, p_rep=False):
        """``_weighted`` over several blocks (then one that adds them in order), with the time
        read from device memory: a fixed launch pair that a captured loop can replay."""
        psi = self.wavefunction if psi is None else psi
        mod = self._obs_module(exprs)
        perm = self._two[1] if (p_rep and self._two is not None) else 0
        nblk = self._wblocks()
        mod.launch("aw_weighted_mb", (nblk,), (256,), [psi.ptr, ctypes.c_int(self.X_gridDIM), ctypes.c_void_p(t_ptr),
                                                      ctypes.c_int(perm), partials.ptr], current_stream())
        mod.launch("aw_weighted_fin", (1,), (64,), [partials.ptr, ctypes.c_int(nblk), ctypes.c_void_p(out_ptr)],
                   current_stream())

    def _fft(self, data, direction, out=None):
        """Unnormalised DFT of the line (cuFFT convention).  Two-level lines: the forward
        output is in the permuted order (frequency k1 + N1 k2 at position k1 N2 + k2) and
        the inverse expects that order."""
        if self._two is None:
            self.plan.fft(data, 1, direction, out=out)
            return
        dst = data if out is None else out
        if direction == FWD:
            self._col(data, FWD, 0, twiddle=True, out=out)
            self._row(dst, FWD, 0)
        else:
            self._row(data, INV, 0, twiddle=True, out=out)
            self._col(dst, INV, 0)

    def _col(self, data, a, b, ctab=None, twiddle=False, dscale=None, out=None):
        """Two-level column pass; twiddle: exp(-2 pi i n2 k1 / N) after its FFTs."""
        self.plan.run(data, 0, a, b, ctab=ctab, chirp_out=(-2.0 * np.pi / self.X_gridDIM) if twiddle else 0.0,
                      lstep=1.0, lcenter=0.0, estep=1.0, ecenter=0.0, dscale=dscale, out=out)

    def _row(self, data, a, b, ctab=None, twiddle=False, out=None):
        """Two-level row pass; twiddle: exp(+2 pi i k1 n2 / N) after its FFTs."""
        self.plan.run(data, 1, a, b, ctab=ctab, chirp_out=(2.0 * np.pi / self.X_gridDIM) if twiddle else 0.0,
                      lstep=1.0, lcenter=0.0, estep=1.0, ecenter=0.0, out=out)

    def get_average(self, observable):
        """sum |psi_j|^2 f(X_j, P_j, t) with no normalisation (:392-410)."""
        self._flush_scale()
        out = empty((1,), np.float64)
        self._weighted([observable], out.ptr)
        return float(out.get()[0])

    def _p_average(self, observable):
        """get_average of psi in the momentum representation left by ``_fft``."""
        self._flush_scale()
        out = empty((1,), np.float64)
        self._weighted([observable], out.ptr, p_rep=True)
        return float(out.get()[0])

    # ------------------------------------------------------------------------------------
    # states
    # ------------------------------------------------------------------------------------
    def set_wavefunction(self, new_wave_func):
        """:316-354: array, GPU array, Python function F(self, X), C string (X declared) or
        number; then psi /= ||psi|| sqrt(dX)."""
        N = self.X_gridDIM
        self._scale_pending = False
        if isinstance(new_wave_func, (np.ndarray, DeviceArray)):
            assert new_wave_func.shape == self.wavefunction.shape, \
                "The grid sizes does not match with the wave function"
            self.wavefunction[:] = new_wave_func
        elif isinstance(new_wave_func, FunctionType):
            self.wavefunction[:] = np.asarray(new_wave_func(self, self.X), np.complex128).reshape(1, N)
        elif isinstance(new_wave_func, str):
            src = rtc.PRELUDE + self.cuda_consts + r"""
extern "C" __global__ void aw_init(%(st)s *psi, int n)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double X = dX * (j - 0.5 * X_gridDIM);
        aw_state_store<%(st)s>::put(psi + j, cuda_complex(%(expr)s));
    }
}
""" % dict(st=self._state_t(), expr=new_wave_func)
            mod = rtc.Module(src, "schrodinger_init.hip")
            blocks, threads = rtc.launch_cfg(N)
            mod.launch("aw_init", (blocks,), (threads,), [self.wavefunction.ptr, ctypes.c_int(N)],
                       current_stream())
        elif isinstance(new_wave_func, (float, complex, int)):
            self.wavefunction.fill(np.complex128(new_wave_func))
        else:
            raise NotImplementedError("new_wave_func must be either function or numpy.array")
        self._normalize(np.sqrt(self.dX))
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

    def _factors(self, ts, bloch=False, raw=False):
        """Per-step EV (with the merged pair for consecutive steps) and EK vectors.  Unless
        ``raw``, EK carries the 1/N of the F-I pair: the reference has no 1/N (:356-369) but
        renormalises every step (:388), so the scale cancels and the deferred normalisation
        of a chunk cannot overflow."""
        n = len(ts)
        N, dt, p = self.X_gridDIM, self.dt, self.plan
        ts_dev = to_device(ts)
        Vt = empty((n, N), np.float64)
        Kt = empty((n, N), np.float64)
        self._eval_1d(0, self._dev_kX, N, ts_dev, n, 0.0, Vt)
        self._eval_1d(1, self._dev_kP, N, ts_dev, n, 0.0, Kt)
        mode = 1 if bloch else 0
        bx = None if (bloch or self._bx is None) else self._bx.ptr
        bp = None if (bloch or self._bp is None) else self._bp.ptr
        ev = p.new_factors(N, n + 1)   # 0: first EV, s: EV_{s-1} EV_s, n: last EV
        ek = p.new_factors(N, n)
        cv = -0.5 * dt
        p.factors(ev, N, 1, A=Vt.ptr, ca=cv, mag=bx, mode=mode)
        if n > 1:
            p.factors(ev, N, n - 1, A=Vt.ptr, ca=cv, B=Vt.ptr + N * 8, cb=cv, mag=bx, mag_pow=2,
                      mode=mode, out_offset=1)
        p.factors(ev, N, 1, A=Vt.ptr + (n - 1) * N * 8, ca=cv, mag=bx, mode=mode, out_offset=n)
        p.factors(ek, N, n, A=Kt.ptr, ca=-dt, mag=bp, mode=mode, scale=1.0 if raw else 1.0 / N)
        return ev, ek

    def _run(self, ts, bloch=False, raw=False):
        """Steps at times ts with merged half-kicks (2 n + 1 passes)."""
        n = len(ts)
        N = self.X_gridDIM
        csz = self.plan.fsize
        p, psi = self.plan, self.wavefunction
        ev, ek = self._factors(ts, bloch, raw)
        if self._two is not None:
            # two-level lines: EV and EK are full-line tables applied inside the passes
            self._col(psi, 0, FWD, ctab=ev.ptr, twiddle=True,
                      dscale=self._dscale if self._scale_pending else None)
            self._scale_pending = False
            for s in range(n):
                self._row(psi, FWD, INV, ctab=ek.ptr + s * N * csz, twiddle=True)
                last = s == n - 1
                self._col(psi, INV, 0 if last else FWD, ctab=ev.ptr + (s + 1) * N * csz, twiddle=not last)
            return
        p.run(psi, 1, 0, FWD, elem_fac=ev.ptr, dscale=self._dscale if self._scale_pending else None)
        self._scale_pending = False
        for s in range(n):
            p.run(psi, 1, 0, INV, elem_fac=ek.ptr + s * N * csz)
            if s < n - 1:
                p.run(psi, 1, 0, FWD, elem_fac=ev.ptr + (s + 1) * N * csz)
        p.run(psi, 1, 0, 0, elem_fac=ev.ptr + n * N * csz)

    def _two_step(self, ev, ek):
        """One unmerged two-level step (three passes) from factor tables ev[0..1], ek[0]."""
        N, psi = self.X_gridDIM, self.wavefunction
        csz = self.plan.fsize
        self._col(psi, 0, FWD, ctab=ev.ptr, twiddle=True, dscale=self._dscale if self._scale_pending else None)
        self._scale_pending = False
        self._row(psi, FWD, INV, ctab=ek.ptr, twiddle=True)
        self._col(psi, INV, 0, ctab=ev.ptr + N * csz)

    def single_step_propagation(self):
        """:356-369 (unnormalised)."""
        self._flush_scale()
        self._run(np.array([self.t], np.float64), raw=True)
        return self.wavefunction

    def propagate(self, steps=1):
        """:371-390: t += dt; step; Ehrenfest; psi /= ||psi|| sqrt(dX)."""
        steps = int(steps)
        if steps <= 0:
            return self
        ts = self._times(steps)
        self._flush_scale()
        if not self.isEhrenfest:
            for c0 in range(0, steps, CHUNK_STEPS):
                self._run(ts[c0:c0 + CHUNK_STEPS])
                self._normalize(np.sqrt(self.dX))
            self.t = ts[-1]
            return self
        N, CH = self.X_gridDIM, CHUNK_STEPS
        ring = empty((steps, 8), np.float64)
        b = self._ehr_bufs()
        ts_all = to_device(ts)          # one upload; each chunk copies its slice on the device
        if not self._scale_pending:
            # every step applies the pending scale: start from 1 so that all chunks are alike
            self._dscale.tensor[:, 0] = 1.0
            self._dscale.tensor[:, 1] = 0.0
            self._scale_pending = True
        for c0 in range(0, steps, CH):
            n = min(CH, steps - c0)
            self._ehr_factors(ts_all.tensor[c0:c0 + n], n)
            graphs.run_chunk(b["graphs"], "ehrenfest", lambda n=n: self._ehr_chunk(n), n == CH)
            ring.tensor[c0:c0 + n].copy_(b["stage"].tensor[:n])
        self._flush_scale()
        self.t = ts[-1]
        r = ring.get()
        nx, npm = r[:, 0], r[:, 4]
        self.X_average.extend((r[:, 1] / nx).tolist())
        self.P_average_RHS.extend((-r[:, 2] / nx).tolist())
        self.X_average_RHS.extend((r[:, 5] / npm).tolist())
        self.P_average.extend((r[:, 6] / npm).tolist())
        self.hamiltonian_average.extend((r[:, 3] / nx + r[:, 7] / npm).tolist())
        return self

    def _ehr_bufs(self):
        """Persistent inputs and outputs of the Ehrenfest loop's chunks (fixed addresses, so
        that a chunk can be captured and replayed)."""
        b = getattr(self, "_ehr", None)
        if b is None:
            N, CH, p = self.X_gridDIM, CHUNK_STEPS, self.plan
            b = self._ehr = dict(ts=empty((CH,), np.float64), Vt=empty((CH, N), np.float64),
                                 Kt=empty((CH, N), np.float64), ev=p.new_factors(N, CH), ek=p.new_factors(N, CH),
                                 stage=empty((CH, 8), np.float64), part=empty((self._wblocks(), 4), np.float64),
                                 xs=["1.", "X", self.diff_V, self.V], ps=["1.", self.diff_K, "P", self.K],
                                 graphs={})
            # compile the observable modules and run the reductions once outside any capture
            self._obs_module(b["xs"])
            self._obs_module(b["ps"])
            self._obs_module([])
            self.plan.stats(self.wavefunction, out=self._norm_st)
        return b

    def _ehr_factors(self, ts_dev_slice, n):
        """EV and EK of the chunk's n steps (one vector each per step; a step applies its EV
        at both ends, :356-369) from the step times, which are copied on the device."""
        b, N, dt, p = self._ehr, self.X_gridDIM, self.dt, self.plan
        b["ts"].tensor[:n].copy_(ts_dev_slice)
        self._eval_1d(0, self._dev_kX, N, b["ts"], n, 0.0, b["Vt"])
        self._eval_1d(1, self._dev_kP, N, b["ts"], n, 0.0, b["Kt"])
        bx = None if self._bx is None else self._bx.ptr
        bp = None if self._bp is None else self._bp.ptr
        p.factors(b["ev"], N, n, A=b["Vt"].ptr, ca=-0.5 * dt, mag=bx)
        p.factors(b["ek"], N, n, A=b["Kt"].ptr, ca=-dt, mag=bp, scale=1.0 / N)

    def _ehr_chunk(self, n):
        """n steps EV F EK I EV (the first pass applies the pending normalisation), each
        followed by the Ehrenfest sums of the x- and p-representations into the stage row and
        the next normalisation scale 1 / (||psi|| sqrt(dX)).  Only launches, from fixed
        addresses."""
        b, N, p, psi = self._ehr, self.X_gridDIM, self.plan, self.wavefunction
        csz = p.fsize
        for s in range(n):
            ev, ek = b["ev"].ptr + s * N * csz, b["ek"].ptr + s * N * csz
            if self._two is None:
                p.run(psi, 1, 0, FWD, elem_fac=ev, dscale=self._dscale)
                p.run(psi, 1, 0, INV, elem_fac=ek)
                p.run(psi, 1, 0, 0, elem_fac=ev)
            else:
                self._col(psi, 0, FWD, ctab=ev, twiddle=True, dscale=self._dscale)
                self._row(psi, FWD, INV, ctab=ek, twiddle=True)
                self._col(psi, INV, 0, ctab=ev)
            t_ptr, row = b["ts"].ptr + 8 * s, b["stage"].ptr + 64 * s
            self._weighted_dev(b["xs"], row, t_ptr, b["part"])
            self._fft(psi, FWD, out=self._tmp)
            self._weighted_dev(b["ps"], row + 32, t_ptr, b["part"], psi=self._tmp, p_rep=True)
            self._norm_scale(np.sqrt(self.dX))

    def get_Ehrenfest(self, t):
        """:412-459 (single evaluation on the current state; normalises psi like the reference)."""
        if not self.isEhrenfest:
            return
        self._normalize(1.0)
        self.X_average.append(self.get_average("X"))
        self.P_average_RHS.append(-self.get_average(self.diff_V))
        self.hamiltonian_average.append(self.get_average(self.V))
        self._fft(self.wavefunction, FWD)
        self._normalize(1.0)
        self.X_average_RHS.append(self._p_average(self.diff_K))
        self.P_average.append(self._p_average("P"))
        self.hamiltonian_average[-1] += self._p_average(self.K)
        self._fft(self.wavefunction, INV)

    def get_energy(self):
        """:461-482 (leaves psi with 2-norm sqrt(N), like the reference)."""
        self._normalize(1.0)
        av_V = self.get_average(self.V)
        self._fft(self.wavefunction, FWD)
        self._normalize(1.0)
        av_K = self._p_average(self.K)
        self._fft(self.wavefunction, INV)
        return av_V + av_K

    def set_ground_state(self, nsteps=10000):
        """:484-503: imaginary-time steps exp(-dt V/2) F exp(-dt K) I exp(-dt V/2) from psi = 1,
        psi /= ||psi|| sqrt(dX) after each (applied per chunk: it only rescales)."""
        self.set_wavefunction(1.)
        ts = np.full(min(nsteps, CHUNK_STEPS), float(self.t_initial))
        done = 0
        while done < nsteps:
            n = min(CHUNK_STEPS, nsteps - done)
            self._run(ts[:n], bloch=True)
            self._normalize(np.sqrt(self.dX))
            done += n
        return self

    def projectout_stationary_states(self, wavefunction):
        """:505-550 (classical Gram-Schmidt against self.stationary_states, unit 2-norms)."""
        nphi = len(self.stationary_states)
        phis = self._phis if nphi else None
        _lib.call("aw_project_out", self.plan.prec, ctypes.c_void_p(wavefunction.ptr),
                  ctypes.c_void_p(phis.ptr if phis is not None else None), nphi, self.X_gridDIM,
                  ctypes.c_void_p(current_stream()))
        return wavefunction

    def get_stationary_states(self, nstates, nsteps=10000):
        """:552-591: even states start from 1, odd from atan(X); each imaginary-time step is
        followed by projecting out the states already found."""
        N = self.X_gridDIM
        self.stationary_states = []
        self._phis = DeviceArray((max(nstates, 1), N), self.state_dtype)
        even = True
        t0 = np.full(1, float(self.t_initial))
        ev, ek = self._factors(t0, bloch=True, raw=True)
        csz = self.plan.fsize
        p, psi = self.plan, self.wavefunction
        for n in range(nstates):
            if even:
                self.set_wavefunction(1.)
            else:
                self.set_wavefunction("atan(X)")
            even = not even
            for _ in range(nsteps):
                if self._two is None:
                    p.run(psi, 1, 0, FWD, elem_fac=ev.ptr)
                    p.run(psi, 1, 0, INV, elem_fac=ek.ptr)
                    p.run(psi, 1, 0, 0, elem_fac=ev.ptr + N * csz)
                else:
                    self._two_step(ev, ek)
                self.projectout_stationary_states(psi)
            self._phis.tensor[n].copy_(psi.tensor[0])
            self.stationary_states.append(psi.copy())
        return self

    # ------------------------------------------------------------------------------------
    # Wigner function of psi
    # ------------------------------------------------------------------------------------
    def get_wigner(self):
        """:593-625 with its missing N x N array and plans created: psi2rho; [F1, SX, I1];
        [F0, SY, I0]; [F1, SX, I1]; (-1)^i, I0, (-1)^i; W /= Re sum W dX dP.  W[i, j] is the
        Wigner function at (P_wigner[i], X_wigner[j]) divided by 2: the reference divides by
        sum W * dX dP, twice the cell area of that grid."""
        return self._wigner(blackman=False)

    def get_wigner_blackman(self):
        """:627-663: get_wigner with a Blackman window (0.42 - 0.5 cos c + 0.08 cos 2c,
        c = 2 pi k / (N - 1), :728-750) before each forward FFT: along j before the X shears,
        along i before the Y shear and the final inverse FFT.  A window that precedes a pass's
        first FFT is its own multiply pass; the last one is the factor of [(-1)^i, by, I0,
        (-1)^i]."""
        return self._wigner(blackman=True)

    def _wigner(self, blackman):
        self._flush_scale()
        N = self.X_gridDIM
        if self._two is not None:
            raise NotImplementedError("get_wigner needs an N x N grid; X_gridDIM = %d is above 4096" % N)
        if self._wplan is None:
            self._wplan = Plan(N, N, 1, self.precision)
            self.wignerfunction = DeviceArray((N, N), self.state_dtype)
            self._wscale = empty((1, 2), np.float64)
        W, wp = self.wignerfunction, self._wplan
        blocks, threads = rtc.launch_cfg(N * N)
        self._mod.launch("aw_outer", (blocks,), (threads,), [self.wavefunction.ptr, W.ptr, ctypes.c_int(N)],
                         current_stream())
        a, b = -np.tan(np.pi / 8.), np.sin(np.pi / 4.)
        kdx, kdp = rtc.kconst(self.dX), rtc.kconst(self.dP)
        shx = dict(chirp=a, lstep=kdx, lcenter=0.5 * N, estep=kdp, ecenter=0.5 * N, efft=1,
                   scale=1.0 / N)
        shy = dict(chirp=b, lstep=kdx, lcenter=0.5 * N, estep=kdp, ecenter=0.5 * N, efft=1,
                   scale=1.0 / N)
        if blackman:
            if self._blackman is None:
                c = np.arange(N) * 2.0 * np.pi / (N - 1.)
                self._blackman = to_device((0.42 - 0.5 * np.cos(c) + 0.08 * np.cos(2.0 * c)).astype(self.plan.factor_dtype))
            win = self._blackman
            wp.run(W, 1, 0, 0, elem_fac=win)
            wp.run(W, 1, FWD, INV, **shx)
            wp.run(W, 0, 0, 0, elem_fac=win)
            wp.run(W, 0, FWD, INV, **shy)
            wp.run(W, 1, 0, 0, elem_fac=win)
            wp.run(W, 1, FWD, INV, **shx)
            wp.run(W, 0, 0, INV, elem_fac=win, sign_in=2, sign_out=2)
        else:
            wp.run(W, 1, FWD, INV, **shx)
            wp.run(W, 0, FWD, INV, **shy)
            wp.run(W, 1, FWD, INV, **shx)
            wp.run(W, 0, INV, 0, sign_in=2, sign_out=2)
        st = wp.stats(W)
        wp.scale_from_stats(st.ptr, 8, 0, 0, self.wigner_dXdP, self._wscale)
        wp.scale(W, self._wscale)
        return W
