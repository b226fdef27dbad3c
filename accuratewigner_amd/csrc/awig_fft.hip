// libawig: fused FFT line passes, factor builders and reductions for gfx950 (MI355X).
//
// Every state array is C-order [batch][rows][cols] of interleaved complex numbers.  A
// "line" is one row (contiguous, axis-1 FFT) or one column (stride = cols, axis-0 FFT).
// One kernel, k_line_pass, does a whole fused pass over every line of a batch:
//
//   load -> [(-1)^(i+j)] -> [FFT a] -> * factor(line, elem) -> [FFT b] -> [(-1)^(i+j)]
//        -> store
//
// The FFT is a register-resident Stockham autosort FFT (each thread owns E elements at
// stride T = N/E, radix-E butterflies in registers, one LDS exchange between stages, none
// after the last stage), so a pass costs one HBM read and one HBM write of the state.
// Column lines are processed LPB at a time with lanes interleaved across the LPB adjacent
// columns, so every wave-wide load/store touches LPB*sizeof(complex)-byte contiguous
// segments of consecutive rows.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "awig_internal.h"

namespace awig {

thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// ---------------------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------------------
template <typename R> struct Cx;
template <> struct Cx<float> { using T = float2; };
template <> struct Cx<double> { using T = double2; };

template <typename T2> __device__ __forceinline__ T2 mk(double x, double y) {
    T2 r;
    r.x = static_cast<decltype(r.x)>(x);
    r.y = static_cast<decltype(r.y)>(y);
    return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) {
    T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r;
}
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) {
    T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r;
}
template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T2, typename S> __device__ __forceinline__ T2 cscale(T2 a, S s) {
    T2 r; r.x = a.x * s; r.y = a.y * s; return r;
}

// ---------------------------------------------------------------------------------------
// compile-time roots of unity: cos/sin(2 pi k / n) by exact octant reduction of k/n and a
// Taylor series on [0, pi/4] (evaluated by the compiler, never at run time)
// ---------------------------------------------------------------------------------------
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 16; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 16; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
struct CS { double c, s; };
// (cos, sin)(2 pi a / n) for 0 <= a <= n/8 (n divisible by 8) or any a when n < 8
constexpr CS root_small(long long a, long long n) {
    return CS{taylor_cos(2.0 * kPi * double(a) / double(n)),
              taylor_sin(2.0 * kPi * double(a) / double(n))};
}
constexpr CS root_cs(long long k, long long n) {
    long long a = ((k % n) + n) % n;
    bool neg_s = false;
    if (2 * a > n) { a = n - a; neg_s = true; }           // theta -> 2pi - theta
    bool neg_c = false;
    if (4 * a > n) { a = n / 2 - a; neg_c = true; }        // theta -> pi - theta
    CS r{1.0, 0.0};
    if (8 * a > n && n % 8 == 0) {                         // theta -> pi/2 - theta
        CS q = root_small(n / 4 - a, n);
        r = CS{q.s, q.c};
    } else {
        r = root_small(a, n);
    }
    if (neg_c) r.c = -r.c;
    if (neg_s) r.s = -r.s;
    return r;
}

// multiply by exp(DIR * 2 pi i K / LEN), all arguments compile-time
template <int DIR, int K, int LEN, typename T2>
__device__ __forceinline__ T2 rot(T2 d) {
    if constexpr (K == 0) {
        return d;
    } else if constexpr (4 * K == LEN) {
        T2 r;
        if constexpr (DIR < 0) { r.x = d.y; r.y = -d.x; }
        else { r.x = -d.y; r.y = d.x; }
        return r;
    } else {
        constexpr CS w = root_cs(K, LEN);
        using R = decltype(d.x);
        constexpr R c = static_cast<R>(w.c);
        constexpr R s = static_cast<R>(DIR * w.s);
        T2 r;
        r.x = d.x * c - d.y * s;
        r.y = d.x * s + d.y * c;
        return r;
    }
}

// in-register radix-2 decimation-in-frequency DFT of size N (natural order in and out)
template <int DIR, int N, int LEN, int K, typename T2>
__device__ __forceinline__ void dif_k(T2 *x) {
    if constexpr (K < LEN / 2) {
#pragma unroll
        for (int s = 0; s < N; s += LEN) {
            T2 a = x[s + K], b = x[s + K + LEN / 2];
            x[s + K] = cadd(a, b);
            x[s + K + LEN / 2] = rot<DIR, K, LEN>(csub(a, b));
        }
        dif_k<DIR, N, LEN, K + 1>(x);
    }
}
template <int DIR, int N, int LEN, typename T2>
__device__ __forceinline__ void dif_levels(T2 *x) {
    if constexpr (LEN >= 2) {
        dif_k<DIR, N, LEN, 0>(x);
        dif_levels<DIR, N, LEN / 2>(x);
    }
}
constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int bitrev(int i, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}
template <int DIR, int N, typename T2>
__device__ __forceinline__ void dft_reg(T2 *x) {
    if constexpr (N > 1) {
        dif_levels<DIR, N, N>(x);
        T2 y[N];
#pragma unroll
        for (int i = 0; i < N; ++i) y[i] = x[bitrev(i, ilog2(N))];
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = y[i];
    }
}

// LDS index with one padding slot every 32 elements (breaks power-of-two bank strides)
__device__ __forceinline__ int lds_idx(int p) { return p + (p >> 5); }
template <int N> constexpr int lds_region() { return N + (N >> 5) + 1; }

// Stockham stage with NS = product of the radices already applied.  Thread u owns the
// elements u + T*k (k < E).  Butterfly b = u + T*m reads positions b + r*N/RAD, which are
// the thread's own registers v[m + r*M]; outputs go to (b/NS)*NS*RAD + b%NS + r*NS.  In the
// last stage those are again the thread's own registers, so no exchange follows it.
template <typename T2, int N, int E, int NS, int DIR>
__device__ __forceinline__ void stockham(T2 *v, T2 *L, int u, const T2 *__restrict__ tw) {
    constexpr int T = N / E;
    constexpr int RAD = (N / NS < E) ? (N / NS) : E;
    constexpr int M = E / RAD;
    constexpr bool LAST = (NS * RAD == N);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int b = u + T * m;
        const int k = b & (NS - 1);
        T2 x[RAD];
#pragma unroll
        for (int r = 0; r < RAD; ++r) x[r] = v[m + r * M];
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                T2 w = tw[(r * k) * (N / (NS * RAD))];
                if constexpr (DIR > 0) w.y = -w.y;
                x[r] = cmul(x[r], w);
            }
        }
        dft_reg<DIR, RAD>(x);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[m + r * M] = x[r];
        } else {
            const int base = (b - k) * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) L[lds_idx(base + r * NS)] = x[r];
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < E; ++kk) v[kk] = L[lds_idx(u + T * kk)];
        __syncthreads();
        stockham<T2, N, E, NS * RAD, DIR>(v, L, u, tw);
    }
}

template <typename T2, int N, int E>
__device__ __forceinline__ void fft_line(int dir, T2 *v, T2 *L, int u, const T2 *tw) {
    if (dir < 0) stockham<T2, N, E, 1, -1>(v, L, u, tw);
    else stockham<T2, N, E, 1, +1>(v, L, u, tw);
}

// ---------------------------------------------------------------------------------------
// the fused line pass
// ---------------------------------------------------------------------------------------
template <typename R, int N, bool STRIDED> struct PassCfg {
    static constexpr int E = (N >= 256) ? 16 : 8;
    static constexpr int T = N / E;
    static constexpr int REGION = lds_region<N>();
    static constexpr int BYTES = sizeof(typename Cx<R>::T);
    static constexpr int WANT = STRIDED ? ((sizeof(R) == 8 ? 4 : 8) > 256 / T
                                               ? (sizeof(R) == 8 ? 4 : 8) : 256 / T)
                                        : (256 / T > 1 ? 256 / T : 1);
    static constexpr int FIT = (144 * 1024) / (REGION * BYTES);
    static constexpr int LPB = WANT < FIT ? WANT : (FIT >= 1 ? FIT : 1);
    static constexpr int THREADS = LPB * T;
};

template <typename R, int N, bool STRIDED>
__global__ __launch_bounds__((PassCfg<R, N, STRIDED>::THREADS))
void k_line_pass(PassArgs a) {
    using T2 = typename Cx<R>::T;
    using C = PassCfg<R, N, STRIDED>;
    constexpr int E = C::E, T = C::T, LPB = C::LPB;
    __shared__ T2 smem[LPB * C::REGION];

    const int tid = threadIdx.x;
    int l_loc, u;
    if constexpr (STRIDED) { l_loc = tid % LPB; u = tid / LPB; }
    else { l_loc = tid / T; u = tid % T; }
    const long long gline = (long long)blockIdx.x * LPB + l_loc;
    const bool active = gline < a.total_lines;
    const long long b = active ? gline / a.nlines : 0;
    const int line = active ? int(gline - b * a.nlines) : 0;
    T2 *L = smem + l_loc * C::REGION;
    const T2 *tw = reinterpret_cast<const T2 *>(a.tw);

    const T2 *src = reinterpret_cast<const T2 *>(a.in) + b * a.bstride + (long long)line * a.line_stride;
    T2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int e = u + T * k;
        v[k] = active ? src[(long long)e * a.elem_stride] : mk<T2>(0., 0.);
    }
    if (a.sign_in) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = u + T * k;
            const int row = STRIDED ? e : line;
            const int par = a.sign_in == 2 ? row : (line + e);
            if (par & 1) v[k] = cscale(v[k], R(-1));
        }
    }
    if (a.dir_a) fft_line<T2, N, E>(a.dir_a, v, L, u, tw);

    if (a.has_factor) {
        // line-constant part: scale * dscale[b] * line_fac[b][line]
        double2 lc = make_double2(a.scale, 0.0);
        if (a.dscale) {
            const double2 ds = reinterpret_cast<const double2 *>(a.dscale)[b];
            lc = make_double2(lc.x * ds.x - lc.y * ds.y, lc.x * ds.y + lc.y * ds.x);
        }
        T2 lf = mk<T2>(lc.x, lc.y);
        if (a.line_fac) {
            const T2 q = reinterpret_cast<const T2 *>(a.line_fac)[b * a.lf_bstride + line];
            lf = cmul(lf, q);
        }
        const T2 *ef = a.elem_fac ? reinterpret_cast<const T2 *>(a.elem_fac) + b * a.ef_bstride : nullptr;
        const R *mk2 = a.mask ? reinterpret_cast<const R *>(a.mask) + b * a.m_bstride : nullptr;
        double lcoord = 0.0;
        if (a.chirp != 0.0) {
            const int q = a.lfft ? ((line + a.nlines / 2) % a.nlines) : line;
            lcoord = a.lstep * (double(q) - a.lcenter);
        }
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = u + T * k;
            T2 f = lf;
            if (ef) f = cmul(f, ef[e]);
            if (mk2) {
                const long long row = STRIDED ? e : line, col = STRIDED ? line : e;
                f = cscale(f, mk2[row * a.m_rowstride + col]);
            }
            if (a.chirp != 0.0) {
                const int q = a.efft ? ((e + N / 2) % N) : e;
                const double ph = a.chirp * (lcoord * (a.estep * (double(q) - a.ecenter)));
                double s, c;
                sincos(ph, &s, &c);
                f = cmul(f, mk<T2>(c, s));
            }
            v[k] = cmul(v[k], f);
        }
    }

    if (a.dir_b) fft_line<T2, N, E>(a.dir_b, v, L, u, tw);
    if (a.sign_out) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = u + T * k;
            const int row = STRIDED ? e : line;
            const int par = a.sign_out == 2 ? row : (line + e);
            if (par & 1) v[k] = cscale(v[k], R(-1));
        }
    }
    if (!active) return;
    if (a.out_mode == 0) {
        T2 *dst = reinterpret_cast<T2 *>(a.out) + b * a.bstride + (long long)line * a.line_stride;
#pragma unroll
        for (int k = 0; k < E; ++k) dst[(long long)(u + T * k) * a.elem_stride] = v[k];
    } else {
        // keep only the element on the diagonal (e == line) of this line
        if ((line % T) == u && line < N) {
            const int kk = line / T;
            T2 val = v[0];
#pragma unroll
            for (int k = 0; k < E; ++k)
                if (k == kk) val = v[k];
            reinterpret_cast<T2 *>(a.out)[b * a.nlines + line] = val;
        }
    }
}

template <typename R, int N, bool STRIDED>
static int launch_pass_n(const PassArgs &a, hipStream_t st) {
    using C = PassCfg<R, N, STRIDED>;
    const long long blocks = (a.total_lines + C::LPB - 1) / C::LPB;
    if (blocks <= 0) return AW_OK;
    if (blocks > 0x7fffffffLL) return set_error(AW_ERR_INVALID_SIZE, "too many lines (%lld)", a.total_lines);
    hipLaunchKernelGGL((k_line_pass<R, N, STRIDED>), dim3((unsigned)blocks), dim3(C::THREADS), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_line_pass launch: %s", hipGetErrorString(e));
    return AW_OK;
}

template <typename R, bool STRIDED>
static int launch_pass_r(int n, const PassArgs &a, hipStream_t st) {
    switch (n) {
    case 64: return launch_pass_n<R, 64, STRIDED>(a, st);
    case 128: return launch_pass_n<R, 128, STRIDED>(a, st);
    case 256: return launch_pass_n<R, 256, STRIDED>(a, st);
    case 512: return launch_pass_n<R, 512, STRIDED>(a, st);
    case 1024: return launch_pass_n<R, 1024, STRIDED>(a, st);
    case 2048: return launch_pass_n<R, 2048, STRIDED>(a, st);
    case 4096: return launch_pass_n<R, 4096, STRIDED>(a, st);
    default: return set_error(AW_ERR_INVALID_SIZE, "unsupported line length %d (64..4096, power of two)", n);
    }
}

int launch_pass(int prec, int n, bool strided, const PassArgs &a, hipStream_t st) {
    if (prec == AW_PREC_F64)
        return strided ? launch_pass_r<double, true>(n, a, st) : launch_pass_r<double, false>(n, a, st);
    return strided ? launch_pass_r<float, true>(n, a, st) : launch_pass_r<float, false>(n, a, st);
}

// ---------------------------------------------------------------------------------------
// factor vectors: out[v][k] = scale * s(k) * mag^pow * E(c0 + ca*A + cb*B) in FP64
// ---------------------------------------------------------------------------------------
template <typename R>
__global__ void k_make_factors(FactorArgs f) {
    using T2 = typename Cx<R>::T;
    const long long total = (long long)f.n * f.nvec;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long vi = idx / f.n;
        const int k = int(idx - vi * f.n);
        double arg = f.c0;
        if (f.A) arg += f.ca * f.A[vi * f.a_stride + k];
        if (f.B) arg += f.cb * f.B[vi * f.b_stride + k];
        double mag = f.scale;
        if (f.mag) {
            const double m = f.mag[vi * f.mag_stride + k];
            for (int p = 0; p < f.mag_pow; ++p) mag *= m;
        }
        if (f.alt_sign && (k & 1)) mag = -mag;
        double re, im;
        if (f.mode == 0) {
            double s, c;
            sincos(arg, &s, &c);
            re = mag * c;
            im = mag * s;
        } else {
            re = mag * exp(arg);
            im = 0.0;
        }
        reinterpret_cast<T2 *>(f.out)[idx] = mk<T2>(re, im);
    }
}

int launch_make_factors(int prec, const FactorArgs &f, hipStream_t st) {
    const long long total = (long long)f.n * f.nvec;
    if (total <= 0) return AW_OK;
    const int threads = 256;
    long long blocks = (total + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_make_factors<double>, dim3((unsigned)blocks), dim3(threads), 0, st, f);
    else
        hipLaunchKernelGGL(k_make_factors<float>, dim3((unsigned)blocks), dim3(threads), 0, st, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_make_factors: %s", hipGetErrorString(e));
    return AW_OK;
}

// ---------------------------------------------------------------------------------------
// reductions: per-block partials (8 doubles) then a fixed-order final sum per batch item
// ---------------------------------------------------------------------------------------
constexpr int RED_THREADS = 256;
constexpr int NSTAT = 8;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}

template <int NV>
__device__ __forceinline__ void block_sum_store(double (&acc)[NV], double *dst) {
    __shared__ double red[RED_THREADS / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        const double s = wave_sum(acc[c]);
        if (lane == 0) red[wid][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < RED_THREADS / 64; ++w) s += red[w][threadIdx.x];
        dst[threadIdx.x] = s;
    }
}

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_stats(const void *data, long long n_item, int rows,
                                                         int cols, double *partials, int nblk) {
    using T2 = typename Cx<R>::T;
    const int b = blockIdx.y;
    const T2 *z = reinterpret_cast<const T2 *>(data) + (long long)b * n_item;
    double acc[NSTAT] = {0, 0, 0, 0, 0, 0, 0, 0};
    const long long stride = (long long)nblk * RED_THREADS;
    for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n_item; i += stride) {
        const T2 q = z[i];
        const double x = q.x, y = q.y;
        acc[0] += x;
        acc[1] += y;
        acc[2] += x * x + y * y;
        acc[3] += x * x - y * y;
        acc[4] += 2.0 * x * y;
    }
    const int nd = rows < cols ? rows : cols;
    for (long long k = (long long)blockIdx.x * RED_THREADS + threadIdx.x; k < nd; k += stride) {
        const T2 q = z[k * cols + k];
        acc[5] += q.x;
        acc[6] += q.y;
    }
    block_sum_store<NSTAT>(acc, partials + ((long long)b * nblk + blockIdx.x) * NSTAT);
}

__global__ void k_sum_partials(const double *partials, int nblk, int ncomp, double *out) {
    const int b = blockIdx.x;
    for (int c = threadIdx.x; c < ncomp; c += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < nblk; ++i) s += partials[((long long)b * nblk + i) * ncomp + c];
        out[(long long)b * ncomp + c] = s;
    }
}

int launch_stats(int prec, const void *data, long long n_item, int rows, int cols, int batch,
                 double *partials, int nblk, double *out, hipStream_t st) {
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_stats<double>, dim3(nblk, batch), dim3(RED_THREADS), 0, st, data, n_item,
                           rows, cols, partials, nblk);
    else
        hipLaunchKernelGGL(k_stats<float>, dim3(nblk, batch), dim3(RED_THREADS), 0, st, data, n_item,
                           rows, cols, partials, nblk);
    hipLaunchKernelGGL(k_sum_partials, dim3(batch), dim3(64), 0, st, partials, nblk, NSTAT, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_stats: %s", hipGetErrorString(e));
    return AW_OK;
}

int launch_sum_partials(const double *partials, int nblk, int ncomp, int batch, double *out,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partials, dim3(batch), dim3(64), 0, st, partials, nblk, ncomp, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_sum_partials: %s", hipGetErrorString(e));
    return AW_OK;
}

__global__ void k_scale_from_stats(const double *stats, int stride, int comp, int mode, double factor,
                                   double *dscale, int batch) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double re = stats[(long long)b * stride + comp] * factor;
    double im = mode == 1 ? stats[(long long)b * stride + comp + 1] * factor : 0.0;
    const double d = re * re + im * im;
    dscale[2 * b] = re / d;
    dscale[2 * b + 1] = -im / d;
}

int launch_scale_from_stats(const double *stats, int stride, int comp, int mode, double factor,
                            double *dscale, int batch, hipStream_t st) {
    hipLaunchKernelGGL(k_scale_from_stats, dim3((batch + 63) / 64), dim3(64), 0, st, stats, stride, comp,
                       mode, factor, dscale, batch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_scale_from_stats: %s", hipGetErrorString(e));
    return AW_OK;
}

template <typename R>
__global__ void k_scale(void *data, long long n, int batch, const double *dscale, double scale) {
    using T2 = typename Cx<R>::T;
    const long long total = n * batch;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long b = i / n;
        double sr = scale, si = 0.0;
        if (dscale) {
            sr = scale * dscale[2 * b];
            si = scale * dscale[2 * b + 1];
        }
        T2 *p = reinterpret_cast<T2 *>(data) + i;
        const double x = p->x, y = p->y;
        *p = mk<T2>(x * sr - y * si, x * si + y * sr);
    }
}

int launch_scale(int prec, void *data, long long n, int batch, const double *dscale, double scale,
                 hipStream_t st) {
    const long long total = n * batch;
    if (total <= 0) return AW_OK;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_scale<double>, dim3((unsigned)blocks), dim3(256), 0, st, data, n, batch, dscale, scale);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3((unsigned)blocks), dim3(256), 0, st, data, n, batch, dscale, scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_scale: %s", hipGetErrorString(e));
    return AW_OK;
}

template <typename RS, typename RD>
__global__ void k_convert(const void *src, void *dst, long long n) {
    using S2 = typename Cx<RS>::T;
    using D2 = typename Cx<RD>::T;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const S2 q = reinterpret_cast<const S2 *>(src)[i];
        reinterpret_cast<D2 *>(dst)[i] = mk<D2>(double(q.x), double(q.y));
    }
}

int launch_convert(const void *src, int sp, void *dst, int dp, long long n, hipStream_t st) {
    if (n <= 0) return AW_OK;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (sp == AW_PREC_F64 && dp == AW_PREC_F32)
        hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
    else if (sp == AW_PREC_F32 && dp == AW_PREC_F64)
        hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
    else if (sp == AW_PREC_F64)
        hipLaunchKernelGGL((k_convert<double, double>), dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
    else
        hipLaunchKernelGGL((k_convert<float, float>), dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_convert: %s", hipGetErrorString(e));
    return AW_OK;
}

int pass_lines_per_block(int prec, int n, bool strided) {
#define AW_LPB(R, NN) \
    (strided ? PassCfg<R, NN, true>::LPB : PassCfg<R, NN, false>::LPB)
    if (prec == AW_PREC_F64) {
        switch (n) {
        case 64: return AW_LPB(double, 64); case 128: return AW_LPB(double, 128);
        case 256: return AW_LPB(double, 256); case 512: return AW_LPB(double, 512);
        case 1024: return AW_LPB(double, 1024); case 2048: return AW_LPB(double, 2048);
        case 4096: return AW_LPB(double, 4096);
        }
    } else {
        switch (n) {
        case 64: return AW_LPB(float, 64); case 128: return AW_LPB(float, 128);
        case 256: return AW_LPB(float, 256); case 512: return AW_LPB(float, 512);
        case 1024: return AW_LPB(float, 1024); case 2048: return AW_LPB(float, 2048);
        case 4096: return AW_LPB(float, 4096);
        }
    }
#undef AW_LPB
    return 0;
}

}  // namespace awig
