// Fused FFT line-pass kernel core (templates shared by libawig and its tuning probes).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "awig_internal.h"

namespace awig {

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
// a * conj(b)
template <typename T2> __device__ __forceinline__ T2 cmul_conj(T2 a, T2 b) {
    T2 r;
    r.x = a.x * b.x + a.y * b.y;
    r.y = a.y * b.x - a.x * b.y;
    return r;
}
template <typename T2, typename S> __device__ __forceinline__ T2 cscale(T2 a, S s) {
    T2 r; r.x = a.x * s; r.y = a.y * s; return r;
}
// complex value in another precision (the identity when the types agree)
template <typename D2, typename S2> __device__ __forceinline__ D2 cvt2(S2 a) {
    if constexpr (std::is_same<D2, S2>::value) {
        return a;
    } else {
        D2 r;
        r.x = static_cast<decltype(r.x)>(a.x);
        r.y = static_cast<decltype(r.y)>(a.y);
        return r;
    }
}

// ---------------------------------------------------------------------------------------
// Line fixed point: the FP64 values of a line stored as int32 pairs m with one power-of-two
// scale per line, x = m 2^e (the line's scale word is e).
//   decode  x = ldexp(double(m), e)                                       (exact)
//   encode  m = low word of (x + 1.5 2^(52+e))   (the add rounds x / 2^e to nearest, ties even:
//           x + 1.5 2^(52+e) lies in [2^(52+e), 2^(53+e)), whose spacing is 2^e)
// e follows from the line's largest |re|, |im| M (M < 2^(e+30), so |m| <= 2^30): a component
// is off by at most 2^-30 M, against 2^-24 |x| for complex64, so a line's L2 error is smaller
// unless its largest value exceeds about 2^5 times its root mean square.  The lines written
// here are FFT outputs, where one Inf or NaN input makes every output non-finite; such a line
// is stored as m = 1 with e = kBfpInf, which decodes to Inf and turns the next FFT to NaN.
// ---------------------------------------------------------------------------------------
constexpr int kBfpInf = 2000;
__device__ __forceinline__ double2 bfp_decode(int2 m, int e) {
    return make_double2(ldexp(double(m.x), e), ldexp(double(m.y), e));
}
// the scale word of a line whose largest |re|, |im| is mx: mx in [2^(ex-1023), 2^(ex-1022))
// gives e = ex - 1052 (ex: biased exponent; 0 for zero and subnormals, e = -1052)
__device__ __forceinline__ int bfp_scale(double mx) {
    const int ex = int((unsigned(__double2hiint(mx)) & 0x7fffffffu) >> 20);
    return ex > 2023 ? kBfpInf : ex - 1052;   // Inf, NaN or above 2^1001
}
__device__ __forceinline__ int2 bfp_encode(double2 v, double c) {
    return make_int2(__double2loint(v.x + c), __double2loint(v.y + c));
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

template <typename R> __device__ __forceinline__ R fmar(R a, R b, R c) {
    if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// In-register butterfly a <- a + w b, b <- a - w b with the compile-time root
// w = exp(DIR 2 pi i K / LEN).  Products by 1, -1 and +-i are exact; otherwise
// w b = c (b.x - t b.y, b.y + t b.x) with c = Re w and t = Im w / Re w, which makes the
// butterfly six fused multiply-adds instead of a complex product and four additions
// (Goedecker's rotation form; |t| <= tan(3 pi / 8) on the roots used here).
template <int DIR, int K, int LEN, typename T2>
__device__ __forceinline__ void bfly(T2 &a, T2 &b) {
    constexpr int Kn = ((K % LEN) + LEN) % LEN;
    using R = decltype(a.x);
    T2 p, m;
    if constexpr (Kn == 0) {
        p = cadd(a, b);
        m = csub(a, b);
    } else if constexpr (2 * Kn == LEN) {
        p = csub(a, b);
        m = cadd(a, b);
    } else if constexpr (4 * Kn == LEN || 4 * Kn == 3 * LEN) {
        // w = sg i, w b = sg (-b.y, b.x)
        constexpr int sg = (4 * Kn == LEN) ? DIR : -DIR;
        if constexpr (sg > 0) {
            p.x = a.x - b.y; p.y = a.y + b.x;
            m.x = a.x + b.y; m.y = a.y - b.x;
        } else {
            p.x = a.x + b.y; p.y = a.y - b.x;
            m.x = a.x - b.y; m.y = a.y + b.x;
        }
    } else {
        constexpr CS w = root_cs(Kn, LEN);
        constexpr R c = static_cast<R>(w.c);
        constexpr R t = static_cast<R>(DIR * w.s / w.c);
        const R u = fmar(-t, b.y, b.x);
        const R v = fmar(t, b.x, b.y);
        p.x = fmar(c, u, a.x); p.y = fmar(c, v, a.y);
        m.x = fmar(-c, u, a.x); m.y = fmar(-c, v, a.y);
    }
    a = p;
    b = m;
}

// X[k], X[k + N/2] = E[k] +- w^k O[k] for k = K .. N/2 - 1 (w = exp(DIR 2 pi i / N)), in place
template <int DIR, int N, int K, typename T2>
__device__ __forceinline__ void dit_combine(T2 *e, T2 *o) {
    if constexpr (K < N / 2) {
        bfly<DIR, K, N>(e[K], o[K]);
        dit_combine<DIR, N, K + 1>(e, o);
    }
}

// In-register DFT of size N (a power of two, natural order in and out), X[k] = sum_n x[n]
// exp(DIR 2 pi i n k / N), decimated in time so that every product by a root of unity is a
// butterfly above: N = 16 as 4 x 4 (n = 4 n1 + n2, k = k1 + 4 k2; the twiddles w^(n2 k1)
// are pushed into the second-level butterflies), other sizes by even / odd halves.
template <int DIR, int N, typename T2>
__device__ __forceinline__ void dft_reg(T2 *x) {
    if constexpr (N == 2) {
        bfly<DIR, 0, 2>(x[0], x[1]);
    } else if constexpr (N == 4) {
        bfly<DIR, 0, 4>(x[0], x[2]);      // E = (x0 + x2, x0 - x2)
        bfly<DIR, 0, 4>(x[1], x[3]);      // O = (x1 + x3, x1 - x3)
        bfly<DIR, 0, 4>(x[0], x[1]);      // X0, X2
        bfly<DIR, 1, 4>(x[2], x[3]);      // X1, X3
        const T2 t = x[1];
        x[1] = x[2];
        x[2] = t;
    } else if constexpr (N == 16) {
        T2 a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = x[4 * n1 + n2];
            dft_reg<DIR, 4>(a[n2]);
        }
        // for each k1: X[k1 + 4 k2] = DFT4 over n2 of w^(n2 k1) a[n2][k1], w = exp(DIR 2 pi i / 16):
        //   c0, c1 = a0 +- w^2k1 a2;  d0, d1 = a1 +- w^2k1 a3;
        //   X[k1], X[k1 + 8] = c0 +- w^k1 d0;  X[k1 + 4], X[k1 + 12] = c1 +- w^(k1 + 4) d1
        bfly<DIR, 0, 16>(a[0][0], a[2][0]); bfly<DIR, 0, 16>(a[1][0], a[3][0]);
        bfly<DIR, 0, 16>(a[0][0], a[1][0]); bfly<DIR, 4, 16>(a[2][0], a[3][0]);
        bfly<DIR, 2, 16>(a[0][1], a[2][1]); bfly<DIR, 2, 16>(a[1][1], a[3][1]);
        bfly<DIR, 1, 16>(a[0][1], a[1][1]); bfly<DIR, 5, 16>(a[2][1], a[3][1]);
        bfly<DIR, 4, 16>(a[0][2], a[2][2]); bfly<DIR, 4, 16>(a[1][2], a[3][2]);
        bfly<DIR, 2, 16>(a[0][2], a[1][2]); bfly<DIR, 6, 16>(a[2][2], a[3][2]);
        bfly<DIR, 6, 16>(a[0][3], a[2][3]); bfly<DIR, 6, 16>(a[1][3], a[3][3]);
        bfly<DIR, 3, 16>(a[0][3], a[1][3]); bfly<DIR, 7, 16>(a[2][3], a[3][3]);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            x[k1] = a[0][k1];
            x[k1 + 8] = a[1][k1];
            x[k1 + 4] = a[2][k1];
            x[k1 + 12] = a[3][k1];
        }
    } else if constexpr (N > 1) {
        T2 e[N / 2], o[N / 2];
#pragma unroll
        for (int n = 0; n < N / 2; ++n) {
            e[n] = x[2 * n];
            o[n] = x[2 * n + 1];
        }
        dft_reg<DIR, N / 2>(e);
        dft_reg<DIR, N / 2>(o);
        dit_combine<DIR, N, 0>(e, o);
#pragma unroll
        for (int k = 0; k < N / 2; ++k) {
            x[k] = e[k];
            x[k + N / 2] = o[k];
        }
    }
}

// LDS layouts of the line buffers.
//  rows (lanes of a wave walk one line): one region per line, one pad slot per 32 positions;
//  columns (lanes interleave LPB adjacent lines): position-major [p][l] with one pad row per
//  2^PADC positions (16 for complex exchanges, 32 for exchanges in real and imaginary
//  halves), so both the natural-order reads (u + T k) and the Stockham scatter writes
//  (16 u + r) of a wave spread over the banks.
template <bool STRIDED, int N, int LPB, int PADC = 4, int PADR = 5> struct LdsLayout {
    static constexpr int REGION = N + (N >> PADR) + 1;
    static constexpr int SIZE = STRIDED ? (N + (N >> PADC)) * LPB : LPB * REGION;
    static __device__ __forceinline__ int at(int p, int l) {
        if constexpr (STRIDED) return (p + (p >> PADC)) * LPB + l;
        else return l * REGION + p + (p >> PADR);
    }
    // at(p + q, l) - at(p, l) for a compile-time q, exact whenever the low PAD bits of p and
    // q add without a carry.  The exchanges only form such sums: a thread's reads p = u,
    // q = T k (u < T, both powers of two) and a butterfly's writes p = (b - k) RAD + k,
    // q = r NS (k < NS, (b - k) RAD a multiple of NS RAD).  A per-thread base plus this
    // offset lets the compiler encode the offsets in the LDS instructions instead of keeping
    // one address register per element.
    static constexpr int off(int q) {
        if constexpr (STRIDED) return (q + (q >> PADC)) * LPB;
        else return q + (q >> PADR);
    }
};

// Stockham stage with NS = product of the radices already applied.  Thread u owns the
// elements u + T*k (k < E).  Butterfly b = u + T*m reads positions b + r*N/RAD, which are
// the thread's own registers v[m + r*M]; outputs go to (b/NS)*NS*RAD + b%NS + r*NS.  In the
// last stage those are again the thread's own registers, so no exchange follows it.
// TWM 0/1: tw is the flat table exp(-2 pi i t / N) (global or LDS).  TWM 2: tw is the
// stage-major LDS table: for the stage of span NS and radix RAD, entry (NS - E) + k*(RAD-1)
// + (r-1) holds exp(-2 pi i r k / (NS RAD)), so a butterfly reads its RAD-1 twiddles at
// consecutive addresses from one base; TWM 3 reads that table through the cache.  In the flat
// table the lanes of a radix-16 stage read
// entries r k N / (NS RAD) (k = lane mod NS), which for FP64 fall into one or two LDS banks
// (8- to 16-way conflicts); the stage-major reads of 8 lanes cover distinct banks.
// XS 1: the exchange moves the real parts and then the imaginary parts through an LDS array
// of reals (half the LDS of a complex exchange, twice the barriers); the butterfly outputs
// wait in v meanwhile, so the register peak does not grow.
template <typename T2, int N, int E, int NS, int DIR, typename LAY, int TWM = 0, int XS = 0>
__device__ __forceinline__ void stockham(T2 *v, T2 *L, int l, int u, const T2 *__restrict__ tw) {
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
            const T2 *tws = (TWM >= 2) ? tw + (NS - E) + k * (RAD - 1) : tw;
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                const T2 w = (TWM >= 2) ? tws[r - 1] : tw[(r * k) * (N / (NS * RAD))];
                if constexpr (DIR > 0) x[r] = cmul_conj(x[r], w);
                else x[r] = cmul(x[r], w);
            }
        }
        dft_reg<DIR, RAD>(x);
        if constexpr (LAST || XS) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[m + r * M] = x[r];
        } else {
            const int wb = LAY::at((b - k) * RAD + k, l);
#pragma unroll
            for (int r = 0; r < RAD; ++r) L[wb + LAY::off(r * NS)] = x[r];
        }
    }
    if constexpr (!LAST && !XS) {
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < E; ++kk) v[kk] = L[LAY::at(u, l) + LAY::off(T * kk)];
        __syncthreads();
        stockham<T2, N, E, NS * RAD, DIR, LAY, TWM, XS>(v, L, l, u, tw);
    } else if constexpr (!LAST) {
        using R = decltype(v[0].x);
        R *Lr = reinterpret_cast<R *>(L);
        // real parts: outputs (b/NS)*NS*RAD + b%NS + r*NS of butterfly b = u + T*m sit in v[m + r*M]
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int b = u + T * m, k = b & (NS - 1), wb = LAY::at((b - k) * RAD + k, l);
#pragma unroll
            for (int r = 0; r < RAD; ++r) Lr[wb + LAY::off(r * NS)] = v[m + r * M].x;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < E; ++kk) v[kk].x = Lr[LAY::at(u, l) + LAY::off(T * kk)];
        __syncthreads();
        // imaginary parts (the .y of every v[m + r*M] is still the butterfly output)
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int b = u + T * m, k = b & (NS - 1), wb = LAY::at((b - k) * RAD + k, l);
#pragma unroll
            for (int r = 0; r < RAD; ++r) Lr[wb + LAY::off(r * NS)] = v[m + r * M].y;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < E; ++kk) v[kk].y = Lr[LAY::at(u, l) + LAY::off(T * kk)];
        __syncthreads();
        stockham<T2, N, E, NS * RAD, DIR, LAY, TWM, XS>(v, L, l, u, tw);
    }
}

template <typename T2, int N, int E, typename LAY, int TWM, int XS = 0>
__device__ __forceinline__ void fft_line(int dir, T2 *v, T2 *L, int l, int u, const T2 *tw) {
    if (dir < 0) stockham<T2, N, E, 1, -1, LAY, TWM, XS>(v, L, l, u, tw);
    else stockham<T2, N, E, 1, +1, LAY, TWM, XS>(v, L, l, u, tw);
}

// ---------------------------------------------------------------------------------------
// the fused line pass
// ---------------------------------------------------------------------------------------
// Compile-time configuration of one k_line_pass instantiation: precision R, line length N,
// elements per thread E (threads per line T = N / E), lines per block LPB, line kind,
// twiddle source (0: global table through the cache, 1: table copied into LDS) and the FFT
// directions: DA / DB in {-1, 0, +1} fixed at compile time, or 2 = chosen at run time.  A
// "general" pass (GEN = 1) also handles sign flips, chirps and diagonal output; the step
// passes (GEN = 0) only multiply by line x element factor vectors (and an optional mask),
// which keeps their code to two FFT bodies.
template <typename R_, int N_, int E_, int LPB_, bool STRIDED_, int TWM_, int DA_ = 2, int DB_ = 2,
          int GEN_ = 1, int MINW_ = 1, int PF_ = 0, int PADC_ = 4, int PADR_ = 5, int OBS_ = 0, int XS_ = 0,
          typename S_ = R_, int Q_ = 0>
struct KCfg {
    using R = R_;   // arithmetic, twiddles, factors and masks
    using S = S_;   // the state in HBM (converted on load and store)
    // line fixed-point storage (FP64 arithmetic on 8-byte states only): bit 0 the input,
    // bit 1 the output is int32 pairs scaled by a power of two per line (see bfp_decode)
    static constexpr int Q = Q_;
    static_assert(Q == 0 || (sizeof(R_) == 8 && sizeof(S_) == 4), "line fixed point: FP64 arithmetic, 8-byte states");
    static constexpr int N = N_, E = E_, LPB = LPB_, TWM = TWM_, DA = DA_, DB = DB_, GEN = GEN_;
    static constexpr int OBS = OBS_;     // weighted-sum block partials of the output (1) or of the
                                         // values right after the factor multiply (2); von Neumann
                                         // observables (3) or dense weighted sums (4) of the closed
                                         // step (columns)
    static constexpr int XS = XS_;       // 1: LDS exchange in real and imaginary halves
    static constexpr int MINW = MINW_;   // minimum waves per SIMD the register budget must allow
    static constexpr int PF = PF_;       // 1: load the factor vectors with the data, before the first FFT;
                                         // 2: the mask as well
    static constexpr bool STRIDED = STRIDED_;
    static constexpr int T = N / E;
    static constexpr int THREADS = LPB * T;
    using LAY = LdsLayout<STRIDED, N, LPB, PADC_, PADR_>;
    static constexpr int XCH_ELEMS = XS ? (LAY::SIZE + 1) / 2 : LAY::SIZE;   // in complex units
    // OBS 3: elements whose weight gathers are in flight together (register budget)
    static constexpr int OBS3_KB = E < 4 ? E : (E >= 16 ? 1 : 4);
    static constexpr int BASE_ELEMS = XCH_ELEMS + ((TWM == 1 || TWM == 2) ? N : 0);
    // OBS 3: the cyclic weights (kVnMaxNw rows of N complex FP64) and the closing element
    // vector (N complex) staged in LDS after the exchange and twiddle regions, where they fit
    static constexpr int OBS3_WOFF = (BASE_ELEMS + 1) & ~1;                 // 16-byte aligned
    static constexpr int OBS3_ELEMS =
        kVnMaxNw * N * int(sizeof(double2) / sizeof(typename Cx<R>::T)) + N;
    static constexpr bool OBS3_LDS =
        OBS == 3 && (OBS3_WOFF + OBS3_ELEMS) * int(sizeof(typename Cx<R>::T)) <= 160 * 1024;
    static constexpr int LDS_ELEMS = OBS3_LDS ? OBS3_WOFF + OBS3_ELEMS : BASE_ELEMS;
};

// Block partials of the weighted sums of a pass's line values v (all lines of the block
// belong to one item):  Re sum conj(w_q) v for q < obs_nw, then (obs_norm2) sum |v|^2, with
// real weights obs_n2w when given, into obs_partials[item][block][component].  Each thread
// adds its elements, then the waves and then the waves of the block are added in order.
// Step passes of up to 4 elements per thread (the single-state latency configuration) load
// the first kObsPre weight arrays (and the real weights of the norm2 component) with the data,
// so that their latency overlaps the FFTs instead of following them.
constexpr int kObsPre = 4;
template <typename C> constexpr int obs_prefetch() { return ((C::OBS == 1 || C::OBS == 2) && C::E <= 4) ? kObsPre : 0; }

// Adds each of the nc (uniform) components of acc over the block, in a fixed order (the
// lanes of each wave, then the waves), into obs_partials[item][block][component].
template <typename C>
__device__ __forceinline__ void store_block_partials(const double (&acc)[8], int nc, const PassArgs &a, long long blk,
                                                     int nl, int tid) {
    constexpr int LPB = C::LPB;
    static_assert(C::THREADS % 64 == 0, "block partials of whole waves");
    __shared__ double ored[C::THREADS / 64][8];
    const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (q < nc) {   // nc is uniform: only the components in use are reduced
            double s = acc[q];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
            if (lane == 0) ored[wid][q] = s;
        }
    }
    __syncthreads();
    if (tid < nc) {
        double s = 0.0;
        for (int q = 0; q < C::THREADS / 64; ++q) s += ored[q][tid];
        const long long g0 = blk * LPB;
        const long long bb = g0 >> a.nl_shift;
        const long long bi = (g0 & (nl - 1)) / LPB;
        a.obs_partials[(bb * (nl / LPB) + bi) * nc + tid] = s;
    }
}

template <typename C, typename T2>
__device__ __forceinline__ void obs_block_partials(const PassArgs &a, const T2 *v, bool active, int off0,
                                                   int kstep, long long blk, int nl, int tid,
                                                   const double2 (*wpre)[C::E] = nullptr,
                                                   const double *n2pre = nullptr) {
    constexpr int N = C::N, E = C::E;
    constexpr int NPRE = obs_prefetch<C>();
    static_assert(C::THREADS % 64 == 0, "weighted-sum partials of whole waves");
    const int nw = a.obs_nw, nc = a.obs_nw + a.obs_norm2;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (active) {
        const long long nit = (long long)N * nl;
        const double2 *w = a.obs_w + off0;   // the weights are shared by the items
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const double x = v[k].x, y = v[k].y;
            if constexpr (NPRE > 0) {
#pragma unroll
                for (int q = 0; q < NPRE; ++q) {
                    if (q < nw) acc[q] += wpre[q][k].x * x + wpre[q][k].y * y;
                }
            }
            for (int q = NPRE; q < nw; ++q) {
                const double2 c = w[q * nit + (long long)k * kstep];
                acc[q] += c.x * x + c.y * y;
            }
            if (a.obs_norm2) {
                double m2 = x * x + y * y;
                if (a.obs_n2w) m2 *= (NPRE > 0) ? n2pre[k] : a.obs_n2w[off0 + (long long)k * kstep];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (q == nw) acc[q] += m2;
            }
        }
    }
    store_block_partials<C>(acc, nc, a, blk, nl, tid);
}

// The von Neumann observables of a square item's state between two steps, from a column
// pass (line j = column j, element e = row i): the values right after the first FFT times
// the step's closing vectors, z(i, j) = v * (obs_lf[j] * obs_ef[i]) (the same product as the
// closing pass of an unmerged step), give
//   [0] Re z(j, j), [1 + q] Re z(j, j) obs_n2w[q][j]  (on the diagonal element of the line),
//   [1 + nd + q] Re (-1)^(j-i) z(i, j) obs_w[q][(j - i) mod N] / N,
// i.e. the trace, coordinate averages and, through the signed cyclic diagonal sums, the
// momentum averages (the weights are DFTs of the momentum functions).
// (The thread's sums; the block sums are stored at the end of the pass, so that the block
// barrier does not hold the factor registers of the merged pass.)
template <typename C, typename T2>
__device__ __forceinline__ void cyc_partials(const PassArgs &a, const T2 *v, bool active, int line, int u,
                                             long long b, double (&acc)[8], const T2 *smem) {
    constexpr int N = C::N, E = C::E, T = C::T;
    constexpr int NW = kVnMaxNw, ND = kVnMaxNd;
    const int nd = a.obs_nd, nw = a.obs_nw;
    double acc0 = 0.0, accx[ND], accp[NW];
#pragma unroll
    for (int q = 0; q < ND; ++q) accx[q] = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) accp[q] = 0.0;
    if (active) {
        // rows past nw read row 0 and count zero: no branch between the gathers, so that
        // they are all in flight at once
        // weights and closing element vector from LDS where the pass staged them
        const double2 *w0 = C::OBS3_LDS ? reinterpret_cast<const double2 *>(smem + C::OBS3_WOFF) : a.obs_w;
        const double2 *wq[NW];
        double mq[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            wq[q] = w0 + (q < nw ? q : 0) * N;
            mq[q] = q < nw ? 1.0 : 0.0;
        }
        T2 lf = mk<T2>(1.0, 0.0);
        if (a.obs_lf) lf = reinterpret_cast<const T2 *>(a.obs_lf)[b * a.lf_bstride + line];
        const T2 *ef = C::OBS3_LDS ? smem + C::OBS3_WOFF + kVnMaxNw * N * int(sizeof(double2) / sizeof(T2)) + u
                       : (a.obs_ef ? reinterpret_cast<const T2 *>(a.obs_ef) + b * a.ef_bstride + u : nullptr);
        // the loads of KB elements at a time in flight (registers: KB (NW + 1) complex)
        constexpr int KB = C::OBS3_KB;
#pragma unroll
        for (int k0 = 0; k0 < E; k0 += KB) {
            // keep the scheduler from hoisting every batch's loads to the top (that takes
            // (NW + 1) E complex registers at once and spills)
            __builtin_amdgcn_sched_barrier(0);
            double2 h[KB][NW];
            T2 efv[KB];
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const int d = (line - (u + T * (k0 + k))) & (N - 1);
#pragma unroll
                for (int q = 0; q < NW; ++q) h[k][q] = wq[q][d];
                efv[k] = ef ? ef[T * (k0 + k)] : mk<T2>(1.0, 0.0);
            }
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const int e = u + T * (k0 + k);
                const T2 z = cmul(v[k0 + k], cmul(lf, efv[k]));
                const double zx = double(z.x), zy = double(z.y);
                const int d = (line - e) & (N - 1);
                const double sx = (d & 1) ? -zx : zx, sy = (d & 1) ? -zy : zy;
#pragma unroll
                for (int q = 0; q < NW; ++q) accp[q] += mq[q] * (sx * h[k][q].x - sy * h[k][q].y);
                if (e == line) {
                    acc0 += zx;
#pragma unroll
                    for (int q = 0; q < ND; ++q)
                        if (q < nd) accx[q] += zx * a.obs_n2w[q * N + line];
                }
            }
        }
    }
    // components [trace, diagonal sums (nd), cyclic sums / N (nw)] at compile-time slots
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        double x = c == 0 ? acc0 : 0.0;
#pragma unroll
        for (int q = 0; q < ND; ++q)
            if (c == 1 + q && q < nd) x = accx[q];
#pragma unroll
        for (int q = 0; q < NW; ++q)
            if (c == 1 + nd + q && q < nw) x = accp[q] * (1.0 / N);
        acc[c] = x;
    }
}

// Weighted sums of the state between two steps, from a column pass (obs_mode 3): z = the
// values right after the first FFT times the step's closing vectors (as above), and
// Re sum conj(w_q) z over the item for dense complex weights w [nw][rows][cols] shared by the
// items.  The weight rows are taken one at a time, the E loads of a row issued together.
template <typename C, typename T2>
__device__ __forceinline__ void closed_partials(const PassArgs &a, const T2 *v, bool active, int line, int u,
                                                long long b, int off0, int kstep, int nl, double (&acc)[8]) {
    constexpr int N = C::N, E = C::E, T = C::T;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
    if (!active) return;
    T2 lf = mk<T2>(1.0, 0.0);
    if (a.obs_lf) lf = reinterpret_cast<const T2 *>(a.obs_lf)[b * a.lf_bstride + line];
    T2 z[E];
    if (a.obs_ef) {
        const T2 *ef = reinterpret_cast<const T2 *>(a.obs_ef) + b * a.ef_bstride + u;
#pragma unroll
        for (int k = 0; k < E; ++k) z[k] = cmul(v[k], cmul(lf, ef[T * k]));
    } else {
#pragma unroll
        for (int k = 0; k < E; ++k) z[k] = cmul(v[k], lf);
    }
    const long long nit = (long long)N * nl;
    const double2 *w = a.obs_w + off0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (q < a.obs_nw) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < E; ++k) {
                const double2 c = w[q * nit + (long long)k * kstep];
                s += c.x * double(z[k].x) + c.y * double(z[k].y);
            }
            acc[q] = s;
        }
    }
}

// Multiply the thread's elements v[k] (element e = u + T k of its line) by the chirp
// exp(i c lcoord ecoord(e)), ecoord(e) = estep (q - ecenter) with q = e, or its FFT-ordered
// index (e + N/2) % N.  The phase is linear in k between the wrap of the FFT order, which
// falls at k = E/2, so it is evaluated in FP64 with one sincos every R elements and rotated
// by exp(i c lcoord estep T) in between (FP64 complex products: a relative error of a few
// 1e-16 after R - 1 rotations), instead of a sincos per element.
template <typename C, typename T2>
__device__ __forceinline__ void apply_chirp(T2 *v, double c, double lcoord, const PassArgs &a, int u) {
    constexpr int N = C::N, E = C::E, T = C::T;
    constexpr int R = (E >= 16) ? 8 : (E >= 2 ? E / 2 : 1);
    const double cl = c * lcoord;
    double sd, cd;
    sincos(cl * (a.estep * double(T)), &sd, &cd);
#pragma unroll
    for (int k0 = 0; k0 < E; k0 += R) {
        const int e = u + T * k0;
        const int q = a.efft ? ((e + N / 2) % N) : e;
        double sn, cs;
        sincos(c * (lcoord * (a.estep * (double(q) - a.ecenter))), &sn, &cs);
#pragma unroll
        for (int k = k0; k < k0 + R; ++k) {
            v[k] = cmul(v[k], mk<T2>(cs, sn));
            if (k + 1 < k0 + R) {
                const double c2 = cs * cd - sn * sd;
                sn = cs * sd + sn * cd;
                cs = c2;
            }
        }
    }
}

template <typename C>
__global__ __launch_bounds__(C::THREADS) __attribute__((amdgpu_waves_per_eu(C::MINW)))
void k_line_pass(PassArgs a) {
    using R = typename C::R;
    using T2 = typename Cx<R>::T;
    using LAY = typename C::LAY;
    constexpr int N = C::N, E = C::E, T = C::T, LPB = C::LPB;
    constexpr bool STRIDED = C::STRIDED;
    __shared__ T2 smem[C::LDS_ELEMS];
    // line maxima (high words) of a fixed-point store whose lines span several waves
    __shared__ unsigned q_lmax[(C::Q & 2) && (C::STRIDED || C::T > 64) ? LPB : 1];
    static_assert(!(C::Q & 2) || C::DA != 0 || C::DB != 0, "the FFT's barriers order q_lmax's reset");

    const int tid = threadIdx.x;
    int l_loc, u;
    if constexpr (STRIDED) { l_loc = tid % LPB; u = tid / LPB; }
    else { l_loc = tid / T; u = tid % T; }
    // Column tiles that share 128-byte row segments run on the same XCD (blocks are dealt
    // to the 8 XCDs round-robin), so the second half of each line hits that XCD's L2.
    long long blk = blockIdx.x;
    if constexpr (STRIDED) {
        const long long nb = gridDim.x;
        if ((nb & 7) == 0) blk = (blk & 7) * (nb >> 3) + (blk >> 3);
    }
    long long gline = blk * LPB + l_loc;
    const bool active = gline < a.total_lines;
    if (!active) gline = a.total_lines - 1;   // compute on a valid line, store nothing
    const int nl = a.nlines;
    const long long b = gline >> a.nl_shift;
    const int line = int(gline & (nl - 1));
    T2 *L = smem;
    const T2 *tw = reinterpret_cast<const T2 *>(a.tw);
    if constexpr ((C::Q & 2) && (STRIDED || T > 64)) {
        if (tid < LPB) q_lmax[tid] = 0u;
    }
    if constexpr (C::TWM == 3) {
        // the stage-major table of this E read through the cache (see TWM 2 below)
        static_assert(E == 4 || E == 8 || E == 16 || E == 32, "stage-major twiddles: E = 4..32");
        tw += (E == 4 ? 1 : (E == 8 ? 2 : (E == 16 ? 3 : 4))) * N;
    } else if constexpr (C::TWM != 0) {
        // twiddle table into LDS; the first exchange barrier of any FFT precedes its first use
        T2 *tl = smem + C::XCH_ELEMS;
        if ((C::DA == 2 ? a.dir_a != 0 : C::DA != 0) || (C::DB == 2 ? a.dir_b != 0 : C::DB != 0)) {
            if constexpr (C::TWM == 1) {
                for (int q = tid; q < N; q += C::THREADS) tl[q] = tw[q];
            } else {
                // the stage-major table of this E, which the plan stores after the flat one
                // (tables for E = 4, 8, 16, 32 at N, 2N, 3N, 4N)
                static_assert(E == 4 || E == 8 || E == 16 || E == 32, "stage-major twiddles: E = 4..32");
                constexpr int EI = E == 4 ? 1 : (E == 8 ? 2 : (E == 16 ? 3 : 4));
                for (int q = tid; q < N - E; q += C::THREADS) tl[q] = tw[EI * N + q];
            }
        }
        tw = tl;
    }
    if constexpr (C::OBS3_LDS) {
        // the cyclic weights and the closing element vector into LDS (read after the first
        // FFT, whose barriers order these writes before them)
        double2 *wl = reinterpret_cast<double2 *>(smem + C::OBS3_WOFF);
#pragma unroll
        for (int r = 0; r < kVnMaxNw; ++r) {
            if (r < a.obs_nw) {
#pragma unroll
                for (int q0 = 0; q0 < N; q0 += C::THREADS)
                    if (q0 + tid < N) wl[r * N + q0 + tid] = a.obs_w[r * N + q0 + tid];
            }
        }
        T2 *el = smem + C::OBS3_WOFF + kVnMaxNw * N * int(sizeof(double2) / sizeof(T2));
        const long long bi = (blk * LPB) >> a.nl_shift;   // the block's item
        const T2 *ef = reinterpret_cast<const T2 *>(a.obs_ef) + bi * a.ef_bstride;
#pragma unroll
        for (int q0 = 0; q0 < N; q0 += C::THREADS)
            if (q0 + tid < N) el[q0 + tid] = a.obs_ef ? ef[q0 + tid] : mk<T2>(1.0, 0.0);
    }

    // Addressing: an item is N x nl (columns) or nl x N (rows) elements; offsets inside an
    // item fit 32 bits (<= 4096^2).  Element e = u + T k of the line sits at
    //   rows: line * N + e          columns: e * nl + line
    const long long item = b * (long long)N * nl;
    const int off0 = STRIDED ? (u * nl + line) : (line * N + u);
    const int kstep = STRIDED ? T * nl : T;
    using S2 = typename Cx<typename C::S>::T;
    const S2 *src = reinterpret_cast<const S2 *>(a.in) + item + off0;
    T2 v[E];
    if constexpr (C::Q & 1) {
        // element e of this line lies on line e of the previous pass (the other axis), whose
        // scale word that pass stored
        const int2 *qsrc = reinterpret_cast<const int2 *>(a.in) + item + off0;
        const int *qh = a.q_in + b * N + u;
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = bfp_decode(qsrc[k * kstep], qh[T * k]);
    } else {
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = cvt2<T2>(src[k * kstep]);
    }
    constexpr int NPRE = obs_prefetch<C>();
    double2 wpre[NPRE > 0 ? NPRE : 1][E];
    double n2pre[NPRE > 0 ? E : 1];
    if constexpr (NPRE > 0) {
        if (active) {
            const long long nit = (long long)N * nl;
#pragma unroll
            for (int q = 0; q < NPRE; ++q) {
                if (q < a.obs_nw) {
#pragma unroll
                    for (int k = 0; k < E; ++k) wpre[q][k] = a.obs_w[q * nit + off0 + (long long)k * kstep];
                }
            }
            if (a.obs_norm2 && a.obs_n2w) {
#pragma unroll
                for (int k = 0; k < E; ++k) n2pre[k] = a.obs_n2w[off0 + (long long)k * kstep];
            }
        }
    }
    if (C::GEN && a.aux_out && active) {
        // the input times dscale[b], in FP64 and rounded once (as a separate scale pass would)
        const double2 ds = reinterpret_cast<const double2 *>(a.dscale)[b];
        S2 *aux = reinterpret_cast<S2 *>(a.aux_out) + item + off0;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const double x = v[k].x, y = v[k].y;
            aux[k * kstep] = mk<S2>(x * ds.x - y * ds.y, x * ds.y + y * ds.x);
        }
    }
    // optional prefetch of the element factors: their load latency then overlaps the data
    // loads instead of stalling between the two FFTs (costs 2E registers)
    T2 efp[C::PF ? E : 1];
    R mkp[C::PF >= 2 ? E : 1];
    T2 lfp = mk<T2>(1.0, 0.0);
    if constexpr (C::PF) {
        if (a.has_factor && a.elem_fac) {
            const T2 *ef = reinterpret_cast<const T2 *>(a.elem_fac) + b * a.ef_bstride + u;
#pragma unroll
            for (int k = 0; k < E; ++k) efp[k] = ef[T * k];
        }
        if (a.has_factor && a.line_fac) lfp = reinterpret_cast<const T2 *>(a.line_fac)[b * a.lf_bstride + line];
        if constexpr (C::PF >= 2) {
            if (a.has_factor && a.mask) {
                const R *mk2 = reinterpret_cast<const R *>(a.mask) + b * a.m_bstride + off0;
#pragma unroll
                for (int k = 0; k < E; ++k) mkp[k] = mk2[k * kstep];
            }
        }
    }

    if (C::GEN && a.sign_in) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = u + T * k;
            const int row = STRIDED ? e : line;
            const int par = a.sign_in == 2 ? row : (line + e);
            if (par & 1) v[k] = cscale(v[k], R(-1));
        }
    }
    if constexpr (C::DA == 2) {
        if (a.dir_a) fft_line<T2, N, E, LAY, C::TWM, C::XS>(a.dir_a, v, L, l_loc, u, tw);
    } else if constexpr (C::DA != 0) {
        stockham<T2, N, E, 1, C::DA, LAY, C::TWM, C::XS>(v, L, l_loc, u, tw);
    }

    double cacc[(C::OBS == 3 || C::OBS == 4) ? 8 : 1];
    if constexpr (C::OBS == 3) cyc_partials<C>(a, v, active, line, u, b, cacc, smem);
    if constexpr (C::OBS == 4) closed_partials<C>(a, v, active, line, u, b, off0, kstep, nl, cacc);

    if (a.has_factor) {
        // line-constant part: scale * dscale[b] * line_fac[b][line]
        double2 lc = make_double2(a.scale, 0.0);
        if (a.dscale) {
            const double2 ds = reinterpret_cast<const double2 *>(a.dscale)[b];
            lc = make_double2(lc.x * ds.x - lc.y * ds.y, lc.x * ds.y + lc.y * ds.x);
        }
        T2 lf = mk<T2>(lc.x, lc.y);
        if (a.line_fac) {
            const T2 q = C::PF ? lfp : reinterpret_cast<const T2 *>(a.line_fac)[b * a.lf_bstride + line];
            lf = cmul(lf, q);
        }
        if (a.elem_fac) {
            const T2 *ef = reinterpret_cast<const T2 *>(a.elem_fac) + b * a.ef_bstride + u;
#pragma unroll
            for (int k = 0; k < E; ++k) v[k] = cmul(v[k], cmul(lf, C::PF ? efp[k] : ef[T * k]));
        } else {
#pragma unroll
            for (int k = 0; k < E; ++k) v[k] = cmul(v[k], lf);
        }
        if (a.mask) {
            const R *mk2 = reinterpret_cast<const R *>(a.mask) + b * a.m_bstride + off0;
#pragma unroll
            for (int k = 0; k < E; ++k) v[k] = cscale(v[k], C::PF >= 2 ? mkp[k] : mk2[k * kstep]);
        }
        if (C::GEN && a.ctab) {
            const T2 *ct = reinterpret_cast<const T2 *>(a.ctab) + b * a.c_bstride + off0;
#pragma unroll
            for (int k = 0; k < E; ++k) v[k] = cmul(v[k], ct[k * kstep]);
        }
        if (C::GEN && a.chirp != 0.0) {
            const int ql = a.lfft ? ((line + nl / 2) % nl) : line;
            apply_chirp<C>(v, a.chirp, a.lstep * (double(ql) - a.lcenter), a, u);
        }
    }

    if constexpr (C::OBS == 2) obs_block_partials<C>(a, v, active, off0, kstep, blk, nl, tid, wpre, n2pre);
    // The second FFT would reuse the first one's LDS addresses (exchange reads and writes
    // and twiddle reads: a VGPR each per element, ~80 at E = 32) and keep them live through
    // the factor multiply.  Hiding the lane's line and position from the compiler here makes
    // it recompute them, a few integer instructions per FFT.  FP32 1024-point columns go
    // from 234 to 156 VGPRs (2 -> 3 waves per SIMD, 1.46e11 -> 1.51e11 grid-point steps/s
    // on 1024^2 x 96).  FP64 passes are held at 2 waves by their LDS exchange and only pay
    // the recomputation (1.8 % on 1024^2 x 48), so they keep the addresses.
    int l_b = l_loc, u_b = u;
    if constexpr (C::DA != 0 && C::DB != 0 && sizeof(R) == 4) __asm__ volatile("" : "+v"(l_b), "+v"(u_b));
    if constexpr (C::DB == 2) {
        if (a.dir_b) fft_line<T2, N, E, LAY, C::TWM, C::XS>(a.dir_b, v, L, l_b, u_b, tw);
    } else if constexpr (C::DB != 0) {
        stockham<T2, N, E, 1, C::DB, LAY, C::TWM, C::XS>(v, L, l_b, u_b, tw);
    }
    if (C::GEN && a.chirp_out != 0.0) {
        // a second chirp after the second FFT (the twiddles of a two-level transform)
        const int ql = a.lfft ? ((line + nl / 2) % nl) : line;
        apply_chirp<C>(v, a.chirp_out, a.lstep * (double(ql) - a.lcenter), a, u);
    }
    if (C::GEN && a.sign_out) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = u + T * k;
            const int row = STRIDED ? e : line;
            const int par = a.sign_out == 2 ? row : (line + e);
            if (par & 1) v[k] = cscale(v[k], R(-1));
        }
    }
    if constexpr (C::OBS == 1) obs_block_partials<C>(a, v, active, off0, kstep, blk, nl, tid, wpre, n2pre);
    if constexpr (C::OBS == 3) store_block_partials<C>(cacc, 1 + a.obs_nd + a.obs_nw, a, blk, nl, tid);
    if constexpr (C::OBS == 4) store_block_partials<C>(cacc, a.obs_nw, a, blk, nl, tid);
    int q_e = 0;
    if constexpr (C::Q & 2) {
        // the line's largest |re|, |im| (a NaN line is NaN in every element), as the high word
        // of its FP64 pattern, which orders non-negative values
        double mx = fabs(v[0].x);
#pragma unroll
        for (int k = 0; k < E; ++k) mx = fmax(fmax(mx, fabs(v[k].x)), fabs(v[k].y));
        unsigned key = unsigned(__double2hiint(mx)) & 0x7fffffffu;
        if constexpr (!STRIDED && T <= 64) {
            // the line's T lanes are adjacent in one wave
#pragma unroll
            for (int o = T / 2; o >= 1; o >>= 1) key = max(key, unsigned(__shfl_xor(int(key), o)));
        } else {
            atomicMax(&q_lmax[l_loc], key);
            __syncthreads();
            key = q_lmax[l_loc];
        }
        q_e = bfp_scale(__hiloint2double(int(key), 0));
        if (active && u == 0) a.q_out[b * nl + line] = q_e;
    }
    if (!active) return;
    if constexpr (C::Q & 2) {
        int2 *dst = reinterpret_cast<int2 *>(a.out) + item + off0;
        if (q_e != kBfpInf) {
            const double c = ldexp(1.5, 52 + q_e);
#pragma unroll
            for (int k = 0; k < E; ++k) dst[k * kstep] = bfp_encode(v[k], c);
        } else {
#pragma unroll
            for (int k = 0; k < E; ++k) dst[k * kstep] = make_int2(1, 1);
        }
    } else if (!C::GEN || a.out_mode == 0) {
        S2 *dst = reinterpret_cast<S2 *>(a.out) + item + off0;
#pragma unroll
        for (int k = 0; k < E; ++k) dst[k * kstep] = cvt2<S2>(v[k]);
    } else {
        // keep only the element on the diagonal (e == line) of this line
        if ((line % T) == u && line < N) {
            const int kk = line / T;
            T2 val = v[0];
#pragma unroll
            for (int k = 0; k < E; ++k)
                if (k == kk) val = v[k];
            reinterpret_cast<S2 *>(a.out)[b * nl + line] = cvt2<S2>(val);
        }
    }
}

}  // namespace awig
