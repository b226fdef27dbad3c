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
#include <mutex>
#include <string>
#include <vector>

#include "awig_internal.h"
#include "awig_pass_core.h"

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
// pass dispatch: plain factor passes run the specialised step kernels, the rest the general one
// ---------------------------------------------------------------------------------------
int launch_pass_f32(int n, bool strided, int kind, const PassArgs &a, hipStream_t st);
int launch_pass_f64(int n, bool strided, int kind, const PassArgs &a, hipStream_t st);
int launch_pass_c64(int n, bool strided, int kind, const PassArgs &a, hipStream_t st);
int obs_lpb_f32(int n, bool strided, long long total_lines);
int obs_lpb_f64(int n, bool strided, long long total_lines);

// the complex64-storage passes compute in FP64 with the FP64 tiling, so their blocks (and
// partial-sum layouts) are those of the FP64 passes
int obs_lpb(int prec, int n, bool strided, long long total_lines) {
    return prec == AW_PREC_F32 ? obs_lpb_f32(n, strided, total_lines) : obs_lpb_f64(n, strided, total_lines);
}

int launch_pass(int prec, int n, bool strided, const PassArgs &a, hipStream_t st) {
    int kind = 0;
    const bool plain = a.chirp == 0.0 && !a.sign_in && !a.sign_out && a.out_mode == 0 && !a.ctab &&
                       a.chirp_out == 0.0 && !a.aux_out;
    if (plain) {
        if (a.dir_a == 0 && a.dir_b == AW_INVERSE) kind = 1;
        else if (a.dir_a == AW_FORWARD && a.dir_b == AW_INVERSE) kind = 2;
        else if (a.dir_a == AW_FORWARD && a.dir_b == 0) kind = 3;
    } else if (a.out_mode == 0) {
        // general passes with the directions of the Wigner transform's passes fixed at compile
        // time (the kernel then holds one FFT body per slot)
        if (a.dir_a == AW_FORWARD && a.dir_b == AW_INVERSE) kind = 7;
        else if (a.dir_a == AW_INVERSE && a.dir_b == 0) kind = 8;
    }
    if (a.obs_mode == 2 || a.obs_mode == 3) {
        if (!strided || (kind != 2 && kind != 3))
            return set_error(AW_ERR_INVALID_VALUE,
                             "sums of the closed step need a plain [forward, factor(, inverse)] column pass");
        kind = a.obs_mode == 2 ? (kind == 2 ? 9 : 10) : (kind == 2 ? 11 : 12);
    } else if (a.obs_w) {
        if (a.obs_mode == 0) {
            if (kind != 2 || strided)
                return set_error(AW_ERR_INVALID_VALUE,
                                 "partials of the output need a plain [forward, factor, inverse] row pass");
            kind = 4;
        } else {
            if (!strided || (kind != 2 && kind != 3))
                return set_error(AW_ERR_INVALID_VALUE,
                                 "partials after the factor need a plain [forward, factor(, inverse)] column pass");
            kind = kind == 2 ? 5 : 6;
        }
    }
    if (prec == AW_PREC_C64) return launch_pass_c64(n, strided, kind, a, st);
    if (a.q) return set_error(AW_ERR_INVALID_VALUE, "line fixed point needs complex64 states with FP64 arithmetic");
    return prec == AW_PREC_F64 ? launch_pass_f64(n, strided, kind, a, st)
                               : launch_pass_f32(n, strided, kind, a, st);
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
        reinterpret_cast<T2 *>(f.out)[vi * f.out_stride + k] = mk<T2>(re, im);
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

__global__ void k_sum_partials(const double *partials, int nblk, int ncomp, double *out);
__global__ void k_finish_sums(const double *partials, int nblk, int ncomp, double *out, double factor,
                              double *dscale);

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

// weighted inner products of two states: per item, sum conj(a) b u[j] and sum conj(a) b v[i]
// (4 doubles: Re, Im of each); the adjoint's gradient of a dipole coupling
template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_pair_dot(const void *a, const void *bb, long long n_item,
                                                            int cols, const double *u, const double *v,
                                                            double *partials, int nblk) {
    using T2 = typename Cx<R>::T;
    const int b = blockIdx.y;
    const T2 *za = reinterpret_cast<const T2 *>(a) + (long long)b * n_item;
    const T2 *zb = reinterpret_cast<const T2 *>(bb) + (long long)b * n_item;
    double acc[4] = {0, 0, 0, 0};
    const long long stride = (long long)nblk * RED_THREADS;
    for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n_item; i += stride) {
        const T2 p = za[i], q = zb[i];
        const double re = double(p.x) * q.x + double(p.y) * q.y;   // conj(p) q
        const double im = double(p.x) * q.y - double(p.y) * q.x;
        const int ii = int(i), r = ii / cols;   // an item holds at most 4096^2 elements
        const double uj = u[ii - r * cols], vi = v[r];
        acc[0] += re * uj;
        acc[1] += im * uj;
        acc[2] += re * vi;
        acc[3] += im * vi;
    }
    block_sum_store<4>(acc, partials + ((long long)b * nblk + blockIdx.x) * 4);
}

int launch_pair_dot(int prec, const void *a, const void *b, long long n_item, int cols, const double *u,
                    const double *v, int batch, double *partials, int nblk, double *out, hipStream_t st) {
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_pair_dot<double>, dim3(nblk, batch), dim3(RED_THREADS), 0, st, a, b, n_item, cols,
                           u, v, partials, nblk);
    else
        hipLaunchKernelGGL(k_pair_dot<float>, dim3(nblk, batch), dim3(RED_THREADS), 0, st, a, b, n_item, cols,
                           u, v, partials, nblk);
    hipLaunchKernelGGL(k_finish_sums, dim3(batch), dim3(RED_THREADS), 0, st, partials, nblk, 4, out, 0.0, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_pair_dot: %s", hipGetErrorString(e));
    return AW_OK;
}

// Second stage of the block-partial reductions: one block per item adds the nblk partials of
// each of the ncomp components (thread t takes k = t, t + 256, ... in order, then a fixed-order
// block sum), writes them to out[b][c] and, when dscale is given, the normalisation scale
// 1 / (out[b][0] factor) (the formula of k_scale_from_stats, mode 0).  Deterministic.
// A one-launch "last block finishes" protocol was measured instead: its device-scope fence
// and ticket cost about 0.16 us per block, serialised (128 blocks: 25 us against 3 + 3 us).
__global__ __launch_bounds__(RED_THREADS) void k_finish_sums(const double *partials, int nblk, int ncomp,
                                                             double *out, double factor, double *dscale) {
    const int b = blockIdx.x;
    const double *pb = partials + (long long)b * nblk * ncomp;
    double tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < nblk; k += RED_THREADS) {
        for (int c = 0; c < ncomp && c < 8; ++c) tot[c] += pb[(long long)k * ncomp + c];
    }
    __shared__ double fin[8];
    block_sum_store<8>(tot, fin);
    __syncthreads();
    if (threadIdx.x < ncomp) {
        const double sum = fin[threadIdx.x];
        out[(long long)b * ncomp + threadIdx.x] = sum;
        if (threadIdx.x == 0 && dscale) {
            const double re = sum * factor, im = 0.0;
            const double d = re * re + im * im;
            dscale[2 * b] = re / d;
            dscale[2 * b + 1] = -im / d;
        }
    }
}

int launch_finish_sums(const double *partials, int nblk, int ncomp, int batch, double *out, double factor,
                       double *dscale, hipStream_t st) {
    hipLaunchKernelGGL(k_finish_sums, dim3(batch), dim3(RED_THREADS), 0, st, partials, nblk, ncomp, out, factor,
                       dscale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_finish_sums: %s", hipGetErrorString(e));
    return AW_OK;
}

// Real parts of weighted sums of each item: partials of out[b][q] = Re sum_i conj(w[q][i]) z[b][i]
// for q < NW, FP64 weights shared by the items, and with N2 also of sum_i |z[b][i]|^2.
template <typename R, int NW, int N2>
__global__ __launch_bounds__(RED_THREADS) void k_weighted_sums(const void *data, long long n_item,
                                                                 const double2 *w, double *partials) {
    using T2 = typename Cx<R>::T;
    constexpr int NC = NW + N2;
    constexpr int U = 4;   // independent element loads in flight per thread
    const int b = blockIdx.y;
    const int nblk = gridDim.x;
    const T2 *z = reinterpret_cast<const T2 *>(data) + (long long)b * n_item;
    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) acc[q] = 0.0;
    const long long stride = (long long)nblk * RED_THREADS;
    long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x;
    for (; i + (U - 1) * stride < n_item; i += U * stride) {
        T2 v[U];
        double2 c[U][NW];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v[u] = z[i + u * stride];
#pragma unroll
            for (int q = 0; q < NW; ++q) c[u][q] = w[q * n_item + i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double x = v[u].x, y = v[u].y;
#pragma unroll
            for (int q = 0; q < NW; ++q) acc[q] += c[u][q].x * x + c[u][q].y * y;
            if constexpr (N2 != 0) acc[NW] += x * x + y * y;
        }
    }
    for (; i < n_item; i += stride) {
        const T2 v = z[i];
        const double x = v.x, y = v.y;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const double2 c = w[q * n_item + i];
            acc[q] += c.x * x + c.y * y;
        }
        if constexpr (N2 != 0) acc[NW] += x * x + y * y;
    }
    block_sum_store<NC>(acc, partials + ((long long)b * nblk + blockIdx.x) * NC);
}

template <typename R, int NW>
static void launch_wsums_t(const void *data, long long n_item, const double *w, int norm2, int batch,
                           double *partials, int nblk, hipStream_t st) {
    if constexpr (NW < 8) {
        if (norm2) {
            hipLaunchKernelGGL((k_weighted_sums<R, NW, 1>), dim3(nblk, batch), dim3(RED_THREADS), 0, st, data,
                               n_item, reinterpret_cast<const double2 *>(w), partials);
            return;
        }
    }
    hipLaunchKernelGGL((k_weighted_sums<R, NW, 0>), dim3(nblk, batch), dim3(RED_THREADS), 0, st, data, n_item,
                       reinterpret_cast<const double2 *>(w), partials);
}

template <typename R>
static int launch_wsums_r(const void *data, long long n_item, const double *w, int nw, int norm2, int batch,
                          double *partials, int nblk, hipStream_t st) {
#define AW_WS(K) \
    case K: launch_wsums_t<R, K>(data, n_item, w, norm2, batch, partials, nblk, st); break;
    switch (nw) {
        AW_WS(1) AW_WS(2) AW_WS(3) AW_WS(4) AW_WS(5) AW_WS(6) AW_WS(7) AW_WS(8)
    default: return set_error(AW_ERR_INVALID_VALUE, "nw must be 1..8 (got %d)", nw);
    }
#undef AW_WS
    return AW_OK;
}

int launch_weighted_partials(int prec, const void *data, long long n_item, const double *w, int nw, int norm2,
                             int batch, double *partials, int nblk, hipStream_t st) {
    const int rc = prec == AW_PREC_F64
                       ? launch_wsums_r<double>(data, n_item, w, nw, norm2, batch, partials, nblk, st)
                       : launch_wsums_r<float>(data, n_item, w, nw, norm2, batch, partials, nblk, st);
    if (rc != AW_OK) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_weighted_sums: %s", hipGetErrorString(e));
    return AW_OK;
}

int launch_weighted_sums(int prec, const void *data, long long n_item, const double *w, int nw, int norm2,
                         int batch, double *partials, int nblk, double *out, double factor, double *dscale,
                         hipStream_t st) {
    const int rc = launch_weighted_partials(prec, data, n_item, w, nw, norm2, batch, partials, nblk, st);
    if (rc != AW_OK) return rc;
    return launch_finish_sums(partials, nblk, nw + norm2, batch, out, factor, dscale, st);
}

// signed cyclic diagonal sums of square items: out[b][d] = (-1)^d sum_i z[i][(i + d) mod n]
// (complex FP64).  For fixed i the threads d, d+1, ... read consecutive columns of row i.
template <typename R>
__global__ __launch_bounds__(256) void k_cyclic_diag(const void *data, int n, double *work) {
    using T2 = typename Cx<R>::T;
    const int d = blockIdx.x * 256 + threadIdx.x;
    const int r0 = blockIdx.y * kDiagRowsPerBlock;
    const int b = blockIdx.z;
    const T2 *z = reinterpret_cast<const T2 *>(data) + (long long)b * n * n;
    double re = 0.0, im = 0.0;
    if (d < n) {
#pragma unroll   // all row loads in flight at once (one 1024^2 step: -0.6 us)
        for (int i = r0; i < r0 + kDiagRowsPerBlock; ++i) {
            const T2 q = z[(long long)i * n + ((i + d) & (n - 1))];
            re += q.x;
            im += q.y;
        }
        const double s = (d & 1) ? -1.0 : 1.0;
        double *w = work + ((long long)b * gridDim.y + blockIdx.y) * 2 * n;
        w[2 * d] = s * re;
        w[2 * d + 1] = s * im;
    }
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
    if (!partials) {
        // one block per item writes its sums straight to out: no scratch, one launch
        if (prec == AW_PREC_F64)
            hipLaunchKernelGGL(k_stats<double>, dim3(1, batch), dim3(RED_THREADS), 0, st, data, n_item, rows,
                               cols, out, 1);
        else
            hipLaunchKernelGGL(k_stats<float>, dim3(1, batch), dim3(RED_THREADS), 0, st, data, n_item, rows,
                               cols, out, 1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_stats: %s", hipGetErrorString(e));
        return AW_OK;
    }
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_stats<double>, dim3(nblk, batch), dim3(RED_THREADS), 0, st, data, n_item,
                           rows, cols, partials, nblk);
    else
        hipLaunchKernelGGL(k_stats<float>, dim3(nblk, batch), dim3(RED_THREADS), 0, st, data, n_item,
                           rows, cols, partials, nblk);
    hipLaunchKernelGGL(k_finish_sums, dim3(batch), dim3(RED_THREADS), 0, st, partials, nblk, NSTAT, out, 0.0,
                       nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_stats: %s", hipGetErrorString(e));
    return AW_OK;
}

// second stage for many components: one thread per (item, component), fixed summation order
__global__ void k_sum_partials_wide(const double *partials, int nblk, int ncomp, double *out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (c >= ncomp) return;
    const double *p = partials + (long long)b * nblk * ncomp + c;
    double s = 0.0;
    for (int i = 0; i < nblk; ++i) s += p[(long long)i * ncomp];
    out[(long long)b * ncomp + c] = s;
}

int launch_cyclic_diag(int prec, const void *data, int n, int batch, double *work, double *out, hipStream_t st) {
    const dim3 grid((n + 255) / 256, n / kDiagRowsPerBlock, batch);
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_cyclic_diag<double>, grid, dim3(256), 0, st, data, n, work);
    else
        hipLaunchKernelGGL(k_cyclic_diag<float>, grid, dim3(256), 0, st, data, n, work);
    // second stage: out[b][c] = sum over the row blocks (fixed order), 2 n doubles per item
    hipLaunchKernelGGL(k_sum_partials_wide, dim3((2 * n + 255) / 256, batch), dim3(256), 0, st, work,
                       n / kDiagRowsPerBlock, 2 * n, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_cyclic_diag: %s", hipGetErrorString(e));
    return AW_OK;
}

// Per-step Ehrenfest sums of square items from the row-block partials of k_cyclic_diag.
// A block covers 64 diagonals d; its 4 wavefronts each add a quarter of the nrb row-block
// partials of every d (fixed order), and the quarters are added in order through LDS.  Then
// for each d: Re(c_d pw_hat[q][d]) for the momentum weights, and Re rho_dd (times xw[q][d])
// for the trace and the coordinate weights.  Block partials go to `partials`; the last
// block of the item (ticket counter, reset for the next launch; n/64 blocks, so the
// serialised fences stay short) adds them in block order into out[b][0 .. 1 + nx + np),
// divides the momentum terms by n, and writes the normalisation scale 1 / (out[b][0] factor)
// when dscale is given.
template <typename R>
__global__ __launch_bounds__(256) void k_cyclic_obs(const void *data, int n, const double *work, int nrb,
                                                    const double *xw, int nx, const double2 *pw_hat, int np,
                                                    double *partials, unsigned *counters, double *out,
                                                    long long out_bstride, double factor, double *dscale) {
    using T2 = typename Cx<R>::T;
    const int b = blockIdx.y;
    const int dl = threadIdx.x & 63, quarter = threadIdx.x >> 6;
    const int d = blockIdx.x * 64 + dl;
    __shared__ double2 csum[4][64];
    // the loads that do not depend on the row-block partials first, so that their latency
    // overlaps that of the partials
    double2 h[7];
    double xq[7], rd = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        h[q] = make_double2(0.0, 0.0);
        xq[q] = 0.0;
    }
    if (quarter == 0 && d < n) {
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            if (q < np) h[q] = pw_hat[(long long)q * n + d];
            if (q < nx) xq[q] = xw[(long long)q * n + d];
        }
        rd = reinterpret_cast<const T2 *>(data)[(long long)b * n * n + (long long)d * n + d].x;
    }
    {
        double cr = 0.0, ci = 0.0;
        if (d < n) {
            const double2 *w = reinterpret_cast<const double2 *>(work + (long long)b * nrb * 2 * n) + d;
            const int r0 = quarter * (nrb / 4), r1 = (quarter == 3) ? nrb : r0 + nrb / 4;
#pragma unroll 16
            for (int r = r0; r < r1; ++r) {
                const double2 v = w[(long long)r * n];
                cr += v.x;
                ci += v.y;
            }
        }
        csum[quarter][dl] = make_double2(cr, ci);
    }
    __syncthreads();
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (quarter == 0 && d < n) {
        double cr = 0.0, ci = 0.0;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
            cr += csum[qq][dl].x;
            ci += csum[qq][dl].y;
        }
#pragma unroll
        for (int q = 0; q < 7; ++q)
            if (q < np) acc[1 + nx + q] = cr * h[q].x - ci * h[q].y;
        acc[0] = rd;
#pragma unroll
        for (int q = 0; q < 7; ++q)
            if (q < nx) acc[1 + q] = rd * xq[q];
    }
    const int nblk = gridDim.x;
    double *pb = partials + (long long)b * nblk * 8;
    block_sum_store<8>(acc, pb + (long long)blockIdx.x * 8);
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(counters + b, 1u) == unsigned(nblk - 1);
    __syncthreads();
    if (!last) return;
    __threadfence();
    const int ncomp = 1 + nx + np;
    // all block partials at once (one load per thread, they were just written by blocks on
    // other XCDs), then the sums in block order from LDS
    __shared__ double fin[kCycObsMaxBlocks * 8];
    for (int i = threadIdx.x; i < nblk * 8; i += blockDim.x) fin[i] = pb[i];
    __syncthreads();
    if (threadIdx.x < ncomp) {
        double s = 0.0;
        for (int k = 0; k < nblk; ++k) s += fin[k * 8 + threadIdx.x];
        if (threadIdx.x > nx) s /= n;
        out[(long long)b * out_bstride + threadIdx.x] = s;
        if (threadIdx.x == 0 && dscale) {
            const double re = s * factor, im = 0.0;
            const double dd = re * re + im * im;
            dscale[2 * b] = re / dd;
            dscale[2 * b + 1] = -im / dd;
        }
    }
    if (threadIdx.x == 0) atomicExch(counters + b, 0u);
}

int launch_cyclic_obs(int prec, const void *data, int n, int batch, double *work, const double *xw, int nx,
                      const double *pw_hat, int np, double *partials, unsigned *counters, double *out,
                      long long out_bstride, double factor, double *dscale, hipStream_t st) {
    if ((n + 63) / 64 > kCycObsMaxBlocks)
        return set_error(AW_ERR_INVALID_SIZE, "cyclic observables: n = %d above %d", n, 64 * kCycObsMaxBlocks);
    const dim3 grid((n + 255) / 256, n / kDiagRowsPerBlock, batch);
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_cyclic_diag<double>, grid, dim3(256), 0, st, data, n, work);
    else
        hipLaunchKernelGGL(k_cyclic_diag<float>, grid, dim3(256), 0, st, data, n, work);
    const dim3 g2((n + 63) / 64, batch);
    const double2 *ph = reinterpret_cast<const double2 *>(pw_hat);
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_cyclic_obs<double>, g2, dim3(256), 0, st, data, n, work, n / kDiagRowsPerBlock, xw, nx,
                           ph, np, partials, counters, out, out_bstride, factor, dscale);
    else
        hipLaunchKernelGGL(k_cyclic_obs<float>, g2, dim3(256), 0, st, data, n, work, n / kDiagRowsPerBlock, xw, nx,
                           ph, np, partials, counters, out, out_bstride, factor, dscale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_cyclic_obs: %s", hipGetErrorString(e));
    return AW_OK;
}

int launch_sum_partials(const double *partials, int nblk, int ncomp, int batch, double *out,
                        hipStream_t st) {
    // a whole block per item adds up to 8 components; more components, one thread each
    if (ncomp <= 8)
        hipLaunchKernelGGL(k_finish_sums, dim3(batch), dim3(RED_THREADS), 0, st, partials, nblk, ncomp, out, 0.0,
                           nullptr);
    else
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

// ---------------------------------------------------------------------------------------
// real <-> complex transforms: the packing around a complex FFT of the full length
// ---------------------------------------------------------------------------------------
// The index helpers address a batch of lines of n points along one axis: axis 1, lines of n
// contiguous points (nl lines); axis 0, items of [n][cols] (nl items), the lines being their
// columns.  A half spectrum keeps h = n/2 + 1 points of each line.
//   mode 0: complex(re[i], 0) for every element (real input widened)
//   mode 1: out = the first h points of each line of the full spectrum in
//   mode 2: the full spectrum of each line from its half: k < h ? in[k] : conj(in[n - k])
//   mode 3: out = Re(in) for every element
template <typename R>
__global__ void k_real_complex(int mode, const void *in, void *out, long long nl, int n, int cols, int axis) {
    using T2 = typename Cx<R>::T;
    const int h = n / 2 + 1;
    const long long m = (mode == 1) ? (axis ? nl * h : nl * h * cols) : (axis ? nl * n : nl * n * cols);
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < m;
         idx += (long long)gridDim.x * blockDim.x) {
        if (mode == 0) {
            reinterpret_cast<T2 *>(out)[idx] = mk<T2>(double(reinterpret_cast<const R *>(in)[idx]), 0.0);
        } else if (mode == 3) {
            reinterpret_cast<R *>(out)[idx] = reinterpret_cast<const T2 *>(in)[idx].x;
        } else if (mode == 1) {
            long long src;
            if (axis) {
                const long long l = idx / h;
                src = l * n + (idx - l * h);
            } else {
                const long long t = idx / cols, c = idx - t * cols, i = t / h, k = t - i * h;
                src = (i * n + k) * cols + c;
            }
            reinterpret_cast<T2 *>(out)[idx] = reinterpret_cast<const T2 *>(in)[src];
        } else {
            long long l, k, c = 0;
            if (axis) {
                l = idx / n;
                k = idx - l * n;
            } else {
                const long long t = idx / cols;
                c = idx - t * cols;
                l = t / n;
                k = t - l * n;
            }
            const bool low = k < h;
            const long long kk = low ? k : n - k;
            const long long src = axis ? l * h + kk : (l * h + kk) * cols + c;
            T2 v = reinterpret_cast<const T2 *>(in)[src];
            if (!low) v.y = -v.y;
            reinterpret_cast<T2 *>(out)[idx] = v;
        }
    }
}

int launch_real_complex(int prec, int mode, const void *in, void *out, long long nl, int n, int cols, int axis,
                        hipStream_t st) {
    const long long m = (long long)(mode == 1 ? n / 2 + 1 : n) * nl * (axis ? 1 : cols);
    long long blocks = (m + 255) / 256;
    if (blocks < 1) return AW_OK;
    if (blocks > 65536) blocks = 65536;
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_real_complex<double>, dim3((unsigned)blocks), dim3(256), 0, st, mode, in, out, nl, n,
                           cols, axis);
    else
        hipLaunchKernelGGL(k_real_complex<float>, dim3((unsigned)blocks), dim3(256), 0, st, mode, in, out, nl, n,
                           cols, axis);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_real_complex: %s", hipGetErrorString(e));
    return AW_OK;
}

// ---------------------------------------------------------------------------------------
// Gram-Schmidt projection of one vector against nphi unit vectors (single workgroup)
// ---------------------------------------------------------------------------------------
// block-wide sum of two doubles (all threads get the result); uniform control flow only
__device__ __forceinline__ double2 block_sum2(double a, double b) {
    __shared__ double red[RED_THREADS / 64][2];
    __shared__ double bc[2];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) { red[wid][0] = a; red[wid][1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < RED_THREADS / 64; ++w) { sa += red[w][0]; sb += red[w][1]; }
        bc[0] = sa;
        bc[1] = sb;
    }
    __syncthreads();
    const double2 r = make_double2(bc[0], bc[1]);
    __syncthreads();
    return r;
}

template <typename T2>
__device__ __forceinline__ void block_normalize(T2 *psi, int n) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += RED_THREADS)
        acc += double(psi[k].x) * psi[k].x + double(psi[k].y) * psi[k].y;
    const double inv = 1.0 / sqrt(block_sum2(acc, 0.0).x);
    for (int k = threadIdx.x; k < n; k += RED_THREADS) psi[k] = mk<T2>(psi[k].x * inv, psi[k].y * inv);
    __syncthreads();
}

// psi /= ||psi||; c_q = <phi_q|psi> for all q (same psi); psi -= sum c_q phi_q; psi /= ||psi||
template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_project_out(void *psi_, const void *phis_, int nphi, int n) {
    using T2 = typename Cx<R>::T;
    extern __shared__ double coef[];
    T2 *psi = reinterpret_cast<T2 *>(psi_);
    const T2 *phis = reinterpret_cast<const T2 *>(phis_);
    block_normalize(psi, n);
    for (int q = 0; q < nphi; ++q) {
        const T2 *phi = phis + (long long)q * n;
        double cr = 0.0, ci = 0.0;
        for (int k = threadIdx.x; k < n; k += RED_THREADS) {
            const double ax = phi[k].x, ay = phi[k].y, bx = psi[k].x, by = psi[k].y;
            cr += ax * bx + ay * by;
            ci += ax * by - ay * bx;
        }
        const double2 c = block_sum2(cr, ci);
        if (threadIdx.x == 0) { coef[2 * q] = c.x; coef[2 * q + 1] = c.y; }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += RED_THREADS) {
        double x = psi[k].x, y = psi[k].y;
        for (int q = 0; q < nphi; ++q) {
            const T2 f = phis[(long long)q * n + k];
            const double c_r = coef[2 * q], c_i = coef[2 * q + 1];
            x -= c_r * f.x - c_i * f.y;
            y -= c_r * f.y + c_i * f.x;
        }
        psi[k] = mk<T2>(x, y);
    }
    __syncthreads();
    block_normalize(psi, n);
}

// Multi-block form for long lines (n >= kProjMultiMin): G blocks each take a slice of psi.
//  k_proj_partials: per block, sum |psi|^2 and sum conj(phi_q) psi over its slice
//  k_proj_update:   every block adds the G partials in the same fixed order (so all blocks
//                   hold the same norm and coefficients c_q = <phi_q|psi> / ||psi||), writes
//                   psi / ||psi|| - sum c_q phi_q over its slice and its partial of the new norm
//  k_proj_scale:    every block adds those partials and divides its slice by the norm
// (three launches of G blocks instead of one workgroup walking the whole line for every q).
constexpr int kProjMultiMin = 8192;
constexpr int kProjBlocks = 128;

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_proj_partials(const void *psi_, const void *phis_, int nphi, int n,
                                                                double *part) {
    using T2 = typename Cx<R>::T;
    const T2 *psi = reinterpret_cast<const T2 *>(psi_);
    const T2 *phis = reinterpret_cast<const T2 *>(phis_);
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int k0 = blockIdx.x * per, k1 = min(n, k0 + per);
    double acc = 0.0;
    for (int k = k0 + threadIdx.x; k < k1; k += RED_THREADS)
        acc += double(psi[k].x) * psi[k].x + double(psi[k].y) * psi[k].y;
    double2 r = block_sum2(acc, 0.0);
    double *pb = part + (long long)blockIdx.x * 2 * (nphi + 1);
    if (threadIdx.x == 0) { pb[0] = r.x; pb[1] = 0.0; }
    for (int q = 0; q < nphi; ++q) {
        const T2 *phi = phis + (long long)q * n;
        double cr = 0.0, ci = 0.0;
        for (int k = k0 + threadIdx.x; k < k1; k += RED_THREADS) {
            const double ax = phi[k].x, ay = phi[k].y, bx = psi[k].x, by = psi[k].y;
            cr += ax * bx + ay * by;
            ci += ax * by - ay * bx;
        }
        r = block_sum2(cr, ci);
        if (threadIdx.x == 0) { pb[2 * (q + 1)] = r.x; pb[2 * (q + 1) + 1] = r.y; }
    }
}

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_proj_update(void *psi_, const void *phis_, int nphi, int n,
                                                              const double *part, double *part2) {
    using T2 = typename Cx<R>::T;
    extern __shared__ double coef[];   // [0]: 1 / ||psi||, then (re, im) of c_q
    T2 *psi = reinterpret_cast<T2 *>(psi_);
    const T2 *phis = reinterpret_cast<const T2 *>(phis_);
    const int G = gridDim.x, stride = 2 * (nphi + 1);
    // every block adds the partials in block order (the same sums in every block)
    for (int c = threadIdx.x; c < stride; c += RED_THREADS) {
        double s = 0.0;
        for (int g = 0; g < G; ++g) s += part[(long long)g * stride + c];
        coef[c] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) coef[0] = 1.0 / sqrt(coef[0]);
    __syncthreads();
    const double inv = coef[0];
    const int per = (n + G - 1) / G;
    const int k0 = blockIdx.x * per, k1 = min(n, k0 + per);
    double acc = 0.0;
    for (int k = k0 + threadIdx.x; k < k1; k += RED_THREADS) {
        double x = psi[k].x * inv, y = psi[k].y * inv;
        for (int q = 0; q < nphi; ++q) {
            const T2 f = phis[(long long)q * n + k];
            const double c_r = coef[2 * (q + 1)] * inv, c_i = coef[2 * (q + 1) + 1] * inv;
            x -= c_r * f.x - c_i * f.y;
            y -= c_r * f.y + c_i * f.x;
        }
        const T2 o = mk<T2>(x, y);
        psi[k] = o;
        acc += double(o.x) * o.x + double(o.y) * o.y;
    }
    const double2 r = block_sum2(acc, 0.0);
    if (threadIdx.x == 0) part2[blockIdx.x] = r.x;
}

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_proj_scale(void *psi_, int n, const double *part2) {
    using T2 = typename Cx<R>::T;
    __shared__ double inv;
    T2 *psi = reinterpret_cast<T2 *>(psi_);
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int g = 0; g < (int)gridDim.x; ++g) s += part2[g];
        inv = 1.0 / sqrt(s);
    }
    __syncthreads();
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int k0 = blockIdx.x * per, k1 = min(n, k0 + per);
    for (int k = k0 + threadIdx.x; k < k1; k += RED_THREADS) psi[k] = mk<T2>(psi[k].x * inv, psi[k].y * inv);
}

// device scratch of the multi-block projection: kProjBlocks x (nphi + 1) complex partials and
// kProjBlocks norm partials, allocated on first use per device for the largest nphi (4096)
static double *g_proj_part[64];
static std::mutex g_proj_mutex;

template <typename R>
static int launch_project_multi(void *psi, const void *phis, int nphi, int n, hipStream_t st) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) return set_error(AW_ERR_HIP, "hipGetDevice failed");
    {
        std::lock_guard<std::mutex> lock(g_proj_mutex);
        if (!g_proj_part[dev]) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            e = hipStreamIsCapturing(st, &cs);
            if (e != hipSuccess || cs != hipStreamCaptureStatusNone)
                return set_error(AW_ERR_INVALID_VALUE, "the first multi-block aw_project_out call must not be captured");
            e = hipMalloc((void **)&g_proj_part[dev], sizeof(double) * (2ll * kProjBlocks * 4097 + kProjBlocks));
            if (e != hipSuccess) {
                g_proj_part[dev] = nullptr;
                return set_error(AW_ERR_ALLOC_FAILED, "projection scratch: %s", hipGetErrorString(e));
            }
        }
    }
    double *part = g_proj_part[dev], *part2 = part + 2ll * kProjBlocks * 4097;
    const int G = kProjBlocks;
    hipLaunchKernelGGL(k_proj_partials<R>, dim3(G), dim3(RED_THREADS), 0, st, psi, phis, nphi, n, part);
    hipLaunchKernelGGL(k_proj_update<R>, dim3(G), dim3(RED_THREADS), sizeof(double) * 2 * (nphi + 1), st, psi,
                       phis, nphi, n, part, part2);
    hipLaunchKernelGGL(k_proj_scale<R>, dim3(G), dim3(RED_THREADS), 0, st, psi, n, part2);
    e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "multi-block projection: %s", hipGetErrorString(e));
    return AW_OK;
}

int launch_project_out(int prec, void *psi, const void *phis, int nphi, int n, hipStream_t st) {
    if (n >= kProjMultiMin && nphi <= 2048)
        return prec == AW_PREC_F64 ? launch_project_multi<double>(psi, phis, nphi, n, st)
                                   : launch_project_multi<float>(psi, phis, nphi, n, st);
    const size_t shmem = sizeof(double) * 2 * (nphi > 0 ? nphi : 1);
    if (prec == AW_PREC_F64)
        hipLaunchKernelGGL(k_project_out<double>, dim3(1), dim3(RED_THREADS), shmem, st, psi, phis, nphi, n);
    else
        hipLaunchKernelGGL(k_project_out<float>, dim3(1), dim3(RED_THREADS), shmem, st, psi, phis, nphi, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_project_out: %s", hipGetErrorString(e));
    return AW_OK;
}

}  // namespace awig
