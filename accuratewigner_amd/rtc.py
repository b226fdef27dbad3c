"""Runtime compilation of the user's C-expression physics (V, K, masks, initial states,
observables) into gfx950 code objects with hipRTC, through libawig's aw_module_* ABI.

This is the counterpart of the reference's PyCUDA ``SourceModule`` JIT
(wigner_moyal_cuda_1d.py:241-258, :340-342, :403-405): user strings are substituted into
kernel templates next to a ``#define`` block built from the constructor keywords
(:169-182) and the user's ``consts`` / ``functions`` strings.  The templates here are this
package's own: they only *evaluate* the expressions (on 1D grids, 2D mask grids, initial
states and weighted reductions); the propagation itself is done by the fused FFT passes.
"""
import ctypes
import hashlib

import numpy as np

from . import _lib

# cuda_complex (the type name user strings use, e.g. rho_vneumann_cuda_1d.py:971-972) with the
# arithmetic and functions such strings need.
PRELUDE = r"""
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_SQRT2
#define M_SQRT2 1.41421356237309504880
#endif
struct cuda_complex {
    double re, im;
    __device__ cuda_complex(double r = 0., double i = 0.) : re(r), im(i) {}
    __device__ double real() const { return re; }
    __device__ double imag() const { return im; }
    __device__ cuda_complex &operator+=(const cuda_complex &o) { re += o.re; im += o.im; return *this; }
    __device__ cuda_complex &operator-=(const cuda_complex &o) { re -= o.re; im -= o.im; return *this; }
    __device__ cuda_complex &operator*=(const cuda_complex &o) {
        const double r = re * o.re - im * o.im; im = re * o.im + im * o.re; re = r; return *this; }
    __device__ cuda_complex &operator/=(const cuda_complex &o) {
        const double d = o.re * o.re + o.im * o.im;
        const double r = (re * o.re + im * o.im) / d; im = (im * o.re - re * o.im) / d; re = r;
        return *this; }
};
__device__ inline cuda_complex operator+(cuda_complex a, const cuda_complex &b) { return a += b; }
__device__ inline cuda_complex operator-(cuda_complex a, const cuda_complex &b) { return a -= b; }
__device__ inline cuda_complex operator*(cuda_complex a, const cuda_complex &b) { return a *= b; }
__device__ inline cuda_complex operator/(cuda_complex a, const cuda_complex &b) { return a /= b; }
__device__ inline cuda_complex operator+(cuda_complex a, double b) { return cuda_complex(a.re + b, a.im); }
__device__ inline cuda_complex operator+(double b, cuda_complex a) { return cuda_complex(a.re + b, a.im); }
__device__ inline cuda_complex operator-(cuda_complex a, double b) { return cuda_complex(a.re - b, a.im); }
__device__ inline cuda_complex operator-(double b, cuda_complex a) { return cuda_complex(b - a.re, -a.im); }
__device__ inline cuda_complex operator*(cuda_complex a, double b) { return cuda_complex(a.re * b, a.im * b); }
__device__ inline cuda_complex operator*(double b, cuda_complex a) { return cuda_complex(a.re * b, a.im * b); }
__device__ inline cuda_complex operator/(cuda_complex a, double b) { return cuda_complex(a.re / b, a.im / b); }
__device__ inline cuda_complex operator/(double b, cuda_complex a) { return cuda_complex(b, 0.) / a; }
__device__ inline cuda_complex operator-(const cuda_complex &a) { return cuda_complex(-a.re, -a.im); }
__device__ inline cuda_complex conj(const cuda_complex &a) { return cuda_complex(a.re, -a.im); }
__device__ inline double real(const cuda_complex &a) { return a.re; }
__device__ inline double imag(const cuda_complex &a) { return a.im; }
__device__ inline double abs(const cuda_complex &a) { return hypot(a.re, a.im); }
__device__ inline double norm(const cuda_complex &a) { return a.re * a.re + a.im * a.im; }
__device__ inline double arg(const cuda_complex &a) { return atan2(a.im, a.re); }
__device__ inline cuda_complex exp(const cuda_complex &a) {
    double s, c; sincos(a.im, &s, &c); const double m = exp(a.re); return cuda_complex(m * c, m * s); }
__device__ inline cuda_complex log(const cuda_complex &a) { return cuda_complex(log(abs(a)), arg(a)); }
__device__ inline cuda_complex sqrt(const cuda_complex &a) {
    const double m = sqrt(abs(a)), h = 0.5 * arg(a); double s, c; sincos(h, &s, &c);
    return cuda_complex(m * c, m * s); }
__device__ inline cuda_complex pow(const cuda_complex &a, double p) {
    const double m = pow(abs(a), p), h = p * arg(a); double s, c; sincos(h, &s, &c);
    return cuda_complex(m * c, m * s); }
template <typename T> struct aw_state_store;
template <> struct aw_state_store<double2> {
    static __device__ void put(double2 *p, const cuda_complex &v) { *p = make_double2(v.re, v.im); } };
template <> struct aw_state_store<float2> {
    static __device__ void put(float2 *p, const cuda_complex &v) { *p = make_float2((float)v.re, (float)v.im); } };
__device__ inline double aw_wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}
"""


def kconst(x):
    """The value a kernel sees for a float passed as ``#define name %.15e``
    (wigner_moyal_cuda_1d.py:174-176): 16 significant digits, so a computed step such as
    dXX can differ from the host double in its last bit.  Kernel coordinates are built from
    these values, as the reference kernels build them from their #defines."""
    return float("%.15e" % float(x))


def cuda_consts(kwargs, consts=None, functions=None):
    """``#define`` block of int / float keywords (wigner_moyal_cuda_1d.py:169-182), then
    the user's ``consts`` and ``functions`` strings verbatim (:179-182, :233-237; ``{{ }}``
    in ``functions`` reaches the compiler unchanged, strong_field_physics_cuda_1d.py:52-58)."""
    out = []
    for name, value in kwargs.items():
        if isinstance(value, (bool, int, np.integer)) and not isinstance(value, (float, np.floating)):
            out.append("#define %s %d\n" % (name, int(value)))
        elif isinstance(value, (float, np.floating)):
            out.append("#define %s %.15e\n" % (name, float(value)))
    s = "".join(out)
    if consts:
        s += consts
    if functions:
        s += functions
    return s


_CACHE = {}


class Module:
    """A hipRTC-compiled code object (cached per identical source within the process)."""

    def __init__(self, source, name="awig_rtc.hip"):
        self.source = source
        key = hashlib.sha1(source.encode()).hexdigest()
        self.handle = _CACHE.get(key)
        if self.handle is None:
            lib = _lib.load()
            h = ctypes.c_void_p()
            log = ctypes.create_string_buffer(1 << 16)
            _lib.check(lib.aw_module_compile(ctypes.byref(h), source.encode(), name.encode(),
                                             log, len(log)))
            self.handle = h
            _CACHE[key] = h

    def launch(self, kernel, grid, block, args, stream):
        """Launch ``kernel`` with ``args`` given as ctypes scalars / c_void_p values."""
        keep = [a if isinstance(a, ctypes._SimpleCData) else ctypes.c_void_p(a) for a in args]
        ptrs = (ctypes.c_void_p * len(keep))(*[ctypes.cast(ctypes.pointer(a), ctypes.c_void_p)
                                                for a in keep])
        gx, gy, gz = (tuple(grid) + (1, 1))[:3]
        bx, by, bz = (tuple(block) + (1, 1))[:3]
        _lib.check(_lib.load().aw_module_launch(self.handle, kernel.encode(), gx, gy, gz, bx, by,
                                                bz, 0, ctypes.c_void_p(stream), ptrs))


def check_source(source):
    """Compile without loading (no GPU needed).  Returns (ok, log)."""
    lib = _lib.load()
    log = ctypes.create_string_buffer(1 << 16)
    size = ctypes.c_size_t(0)
    rc = lib.aw_rtc_check(source.encode(), log, len(log), ctypes.byref(size))
    return rc == 0 and size.value > 0, log.value.decode(errors="replace")


# ---------------------------------------------------------------------------------------
# kernel templates (filled per propagator class)
# ---------------------------------------------------------------------------------------
EVAL_1D = r"""
extern "C" __global__ void aw_eval_1d(int which, const double *x, int n, const double *ts, int nt,
                                      double tshift, double *out)
{
    const long long total = (long long)n * nt;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long s = idx / n;
        const int k = (int)(idx - s * n);
        const double t = ts[s] + tshift;
        out[idx] = which == 0 ? V(x[k], t) : K(x[k], t);
    }
}
"""

REDUCE_TAIL = r"""
    __shared__ double aw_red[4][%(nobs)d];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = 0; c < %(nobs)d; ++c) {
        const double s = aw_wave_sum(acc[c]);
        if (lane == 0) aw_red[wid][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < %(nobs)d) {
        double s = 0.;
        for (int w = 0; w < 4; ++w) s += aw_red[w][threadIdx.x];
        partials[((long long)blockIdx.y * gridDim.x + blockIdx.x) * %(nobs)d + threadIdx.x] = s;
    }
"""

def launch_cfg(total, threads=256, max_blocks=65536):
    blocks = max(1, min(max_blocks, (total + threads - 1) // threads))
    return blocks, threads
