// Tuned k_line_pass configurations and their launchers (instantiated per precision in
// awig_pass_f32.hip / awig_pass_f64.hip, so the two compile in parallel).
#pragma once
#include "awig_pass_core.h"

namespace awig {

constexpr int kLdsBudget = 160 * 1024;

// Per (precision, line length, line kind):
//  * E = 16 elements per thread (8 below 256), so a 1024-point line is 64 lanes, radix 16-16-4;
//    FP32 columns of 512 and 1024 points use E = 32 (radix 32-16 / 32-32: one LDS exchange,
//    8 columns per 128 / 256-thread block; 14% faster at 1024^2 x 96 than E = 16, and still
//    8 % faster than E = 16 with the split exchange and 32-position pads), and
//    exchange the real and the imaginary parts in turn (43 instead of 76 KB of LDS; 6% faster
//    on a MALL-resident 24-state group of 1024^2, 2% slower on 96 states from HBM);
//  * FP32 rows of E = 16 exchange the same way: a 1024-point row block then takes 25 instead
//    of 42 KB of LDS and 4 instead of 3 waves run per SIMD (110 VGPRs; 1.2 % faster on
//    1024^2 x 96, neutral at 512^2).  FP64 passes keep the complex exchange: with split
//    halves they reach 3 waves only by spilling, and measured no faster;
//  * rows: 4 lines per 256-thread block in FP32, 2 lines per 128-thread block in FP64; FP32
//    rows of 256 and of 2048+ points load their factor vectors with the data (PF, 4-7 %
//    faster on MALL-resident batches when measured; at 512 points it costs a wave per SIMD
//    and 1.5 %, at 1024 points now 3 %; FP64 rows go from 244 to 282 VGPRs and lose 40 %);
//  * columns: 8 adjacent columns per block (64-byte row segments in FP32, 128 in FP64; FP64
//    columns of 512 and 1024 points: 4, see C64);
//  * twiddles in LDS (FP32 always, FP64 columns when they fit) or read through the cache;
//  * the largest power-of-two LPB that fits the 160 KB LDS.
// Measured at 1024^2 x 48 (FP32) / x 24 (FP64) states on MI355X.
template <typename R, int N, bool STRIDED> struct PassTune {
    static constexpr bool F32 = sizeof(R) == 4;
    static constexpr bool WIDE = F32 && STRIDED && (N == 512 || N == 1024);
    static constexpr int E = WIDE ? 32 : ((N >= 256) ? 16 : 8);
    static constexpr int T = N / E;
    static constexpr int WANT_ROWS = F32 ? (256 / T > 1 ? 256 / T : 1) : (128 / T > 1 ? 128 / T : 1);
    static constexpr int WANT_COLS = WIDE ? 8 : ((8 > 256 / T) ? 8 : 256 / T);
    static constexpr int WANT0 = STRIDED ? WANT_COLS : WANT_ROWS;
    // at most 512 threads per block (a 1024-thread block caps a thread at 128 VGPRs)
    static constexpr int WANT = (WANT0 * T > 512) ? (512 / T > 1 ? 512 / T : 1) : WANT0;
    // FP64 rows of 16 elements per thread exchange in halves with the twiddles in LDS (below)
    static constexpr bool R64X = !F32 && !STRIDED && E == 16;
    static constexpr int TW_PREF = (F32 || STRIDED || R64X) ? 1 : 0;
    static constexpr int fit(int lpb, int tw) {
        return lpb <= 1 ? 1 : (lds_bytes_dyn(lpb, tw) <= kLdsBudget ? lpb : fit(lpb / 2, tw));
    }
    // row pad: one slot per 32 positions in FP32; per 16 with FP64 arithmetic (FP64 1024-point
    // rows, 12-state group launches: 80.3 -> 78.4 us; complex64 states with FP64 arithmetic,
    // 24 states: 141.3 -> 136.5 us; per 8 or 64: slower)
    static constexpr int PADR = F32 ? 5 : 4;
    static constexpr int lds_bytes_dyn(int lpb, int tw) {
        return ((STRIDED ? (N + (N >> 4)) * lpb : lpb * (N + (N >> PADR) + 1)) + (tw ? N : 0)) *
               int(sizeof(typename Cx<R>::T));
    }
    // FP64 columns of 512 and 1024 points go 4 to a 256-thread block with the twiddles read
    // through the cache (70 instead of 155 KB of LDS at 1024 points; TWM 3), so that two blocks
    // share a CU and one block's loads and stores overlap the other's FFTs: 6.7e10 -> 7.35e10
    // grid-point steps/s on 1024^2 x 48 (8 columns with the split exchange: 7.2e10; 2 columns:
    // 7.1e10; 4 columns with both: no faster); at 512 points +2 % on 512^2 x 192, equal on x 48
    static constexpr bool C64 = !F32 && STRIDED && E == 16 && N >= 512 && N <= 1024;
    // keep LDS twiddles only if they do not cost lines per block, read stage-major (TWM 2:
    // the flat table's radix-16 reads conflict in the LDS banks): FP64 1024-point rows
    // 78.4 -> 75.0 us per 12-state launch, complex64 states with FP64 arithmetic 144.5 ->
    // 128.2 us per 24 states; FP32 rows 76.3 -> 76.0 us, FP32 1024^2 x 96 steps +0.2 %.
    // The FP64 columns below read the stage-major table through the cache (TWM 3: a
    // butterfly's 15 twiddles in two cache lines instead of 15): 84.3 -> 80.4 us per
    // 12-state 1024-point launch, complex64 states 138 -> 136 us per 24
    static constexpr int TWM = C64 ? 3 : ((TW_PREF && fit(WANT, 1) == fit(WANT, 0)) ? 2 : 0);
    static constexpr int LPB = C64 ? 4 : fit(WANT, TWM);
    // FP32 columns of 2048 points exchange in halves as well (86 -> 51 KB of LDS, 2 -> 3
    // blocks per CU): 1.00e11 -> 1.17e11 grid-point steps/s on 2048^2 x 24.  At 4096 points
    // that is 6 % slower, and FP64 columns of 2048 / 4096 points with half the columns per
    // block and the twiddles through the cache lose 15 / 32 %.
    // FP64 rows (E = 16, 2 rows per 128 threads) with the split exchange and LDS twiddles:
    // 94.3 -> 87.6 us per 12-state 1024^2 launch (complex64 states with FP64 arithmetic:
    // 161.8 -> 160.6 us per 24 states); either change alone gains 1-3 %
    static constexpr int XS = (WIDE || (F32 && STRIDED && N == 2048) || (F32 && !STRIDED && E == 16) || R64X) ? 1 : 0;
    // (FP32 rows of 1024 points without the prefetch: 85 VGPRs and 5 waves per SIMD, 74.5-75.5
    // -> 72.2-73.0 us per 24-state launch, since the exchange addressing and the butterflies
    // were slimmed; with it 108 VGPRs and 4 waves)
    static constexpr int PF = (F32 && !STRIDED && (N == 256 || N > 1024)) ? 1 : 0;
};

// Small passes (lines up to 2048 points and few lines in the whole batch: one 256^2 state,
// a few 128^2 states, one 1024^2 state) are bound by the latency of one block, not by HBM:
// fewer elements per thread (E = N/128, 4..16) and 1-2 lines per block spread the lines over
// many more CUs and shorten each thread's chain.  Twiddles come from LDS, stage-major (config
// 1 solve 10.84 -> 10.59 ms, config 3 FP64 step 42.7 -> 42.1 us, FP32 unchanged; the same
// table read through the cache instead: config 3 -1.8 %, 512^2 Gibbs solves +2-3 %).  Measured on
// MI355X: one 256^2 FP64 state 6.8 -> 4.7 us per pass (the back-to-back launch floor is
// 2.5 us).  The crossover to the batch configuration (back-to-back passes, us):
//   1024 lines: small wins or ties everywhere (1024^2 FP32 row 7.5 vs 8.6, column 10.4 vs 11.3);
//   2048 lines of 256 points: small still wins (FP32 row 4.8 vs 6.7, FP64 column 7.8 vs 7.9);
//   2048 lines of 512 points: FP64 loses (column 13.2 vs 10.8), FP32 ties (column 9.9 vs 10.0);
//   2048 lines of 1024 points: small loses (FP64 row 24.2 vs 16.1, FP32 column 16.5 vs 12.6).
constexpr long long kSmallLines = 1024;
template <typename R, int N> constexpr long long small_lines_max() {
    return (N <= 256 || (sizeof(R) == 4 && N <= 512)) ? 2 * kSmallLines : kSmallLines;
}

// The step kinds of up to 8 elements per thread load their factor vectors and mask with the
// data (PF = 2), so that their latency overlaps the data loads instead of following the first
// FFT; not the kinds with von Neumann observables (OBS 3), whose epilogue after the first FFT
// would then hold those registers too (FP64 1024-point columns: 256 VGPRs and spills).
// Columns of 1024 points go 4 to a block (32 / 64-byte row segments): the von Neumann step
// on one 1024^2 state (three passes) takes 64.3 instead of 68.1 us in FP64 and 45.8 instead
// of 49.6 us in mixed precision.  At 256 and 512 points, and with 1 or 8 columns at 1024,
// it is no faster or slower; E = 4 or 16 at 1024 points measured no faster either.
template <typename R, int N, bool STRIDED> struct SmallTune {
    static constexpr int E = (N / 128 < 4) ? 4 : ((N / 128 > 16) ? 16 : N / 128);
    static constexpr int T = N / E;
    static constexpr int LPB = STRIDED ? (N == 1024 ? 4 : 2) : (T >= 64 ? 1 : 64 / T);
    // (E <= 8: with 16 elements FP64 passes of 2048 points would go to 240-256 VGPRs and one
    // wave per SIMD)
    static constexpr int PF = E <= 8 ? 2 : 0;
};

// Column exchanges in real and imaginary halves pad one row per 32 positions (PADC 5)
// instead of 16: +2 % on 1024^2 x 96 mixed (8: +0.7 %, 64 and 128: slower); the row pad
// stays at 32 positions (16 is 3 % slower, 64 no faster).
// The step passes of complex64 states with FP64 arithmetic over 1024-point rows go 4 rows to
// a 256-thread block under the register budget of 3 waves per SIMD (152 instead of 174
// VGPRs, no spill; 51 KB of LDS, 3 blocks per CU): 129.4 -> 122.7 us per 24-state launch.
// FP64 states lose with it (75.4 -> 86.5 us per 12 states), and the kinds with block
// partials keep the tiling that obs_lpb reports.
template <typename R, int N, bool STRIDED, int GEN, int OBS, typename S> constexpr bool c64_row3() {
    return sizeof(R) == 8 && sizeof(S) == 4 && !STRIDED && N == 1024 && GEN == 0 && OBS == 0;
}

template <typename R, int N, bool STRIDED, int DA, int DB, int GEN, int OBS = 0, typename S = R, int Q = 0>
using PassK = KCfg<R, N, PassTune<R, N, STRIDED>::E,
                   c64_row3<R, N, STRIDED, GEN, OBS, S>() ? 4 : PassTune<R, N, STRIDED>::LPB, STRIDED,
                   PassTune<R, N, STRIDED>::TWM, DA, DB, GEN, c64_row3<R, N, STRIDED, GEN, OBS, S>() ? 3 : 1,
                   PassTune<R, N, STRIDED>::PF,
                   (PassTune<R, N, STRIDED>::XS && STRIDED) ? 5 : 4, PassTune<R, N, STRIDED>::PADR, OBS,
                   PassTune<R, N, STRIDED>::XS, S, Q>;

template <typename R, int N, bool STRIDED, int DA, int DB, int GEN, int OBS = 0, typename S = R, int Q = 0>
using PassKS = KCfg<R, N, SmallTune<R, N, STRIDED>::E, SmallTune<R, N, STRIDED>::LPB, STRIDED, 2, DA,
                    DB, GEN, 1, (GEN || OBS >= 3) ? 0 : SmallTune<R, N, STRIDED>::PF, 4, 5, OBS, 0, S, Q>;

template <typename C>
int launch_cfg(const PassArgs &a, hipStream_t st) {
    static_assert(C::LDS_ELEMS * int(sizeof(typename Cx<typename C::R>::T)) <= kLdsBudget, "LDS budget");
    const long long blocks = (a.total_lines + C::LPB - 1) / C::LPB;
    if (blocks <= 0) return AW_OK;
    if (blocks > 0x7fffffffLL) return set_error(AW_ERR_INVALID_SIZE, "too many lines (%lld)", a.total_lines);
    hipLaunchKernelGGL((k_line_pass<C>), dim3((unsigned)blocks), dim3(C::THREADS), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AW_ERR_EXEC_FAILED, "k_line_pass launch: %s", hipGetErrorString(e));
    return AW_OK;
}

// kind: 0 general (run-time directions, signs, chirps, diagonal output);
//       1 step [factor, inverse FFT]; 2 step [forward, factor, inverse]; 3 step [forward, factor];
//       4 row step [forward, factor, inverse] that also stores weighted-sum block partials of
//         its output;
//       5 / 6 column step [forward, factor, inverse] / [forward, factor] that stores the block
//         partials of the values right after the factor multiply;
//       9 / 10 column step [forward, factor, inverse] / [forward, factor] that stores the
//         block partials of the von Neumann observables of the values after the first FFT
//         times the closing vectors of a step (the state between two steps);
//       11 / 12 the same with dense weighted sums of that state instead;
//       7 / 8 general pass with the directions [forward, inverse] / [inverse, none] fixed (the
//         passes of the Wigner transform: one FFT body per slot instead of both directions,
//         104 -> 92 us for an FP32 1024-point row pass over 24 1024^2 states); the latency
//         configuration runs them as kind 0
// (the partial kinds need blocks of whole waves: obs_lpb_n reports 0 where they are not)
template <typename R, int N, bool STRIDED, typename S = R>
int launch_kind_n(int kind, const PassArgs &a, hipStream_t st) {
    if constexpr (N <= 2048) {
        if (a.total_lines <= small_lines_max<R, N>()) {
            switch (kind) {
            case 4:
                if constexpr (!STRIDED) return launch_cfg<PassKS<R, N, STRIDED, -1, 1, 0, 1, S>>(a, st);
                return set_error(AW_ERR_INVALID_VALUE, "weighted-sum partials need a row pass");
            case 5:
            case 6:
                if constexpr (STRIDED && PassKS<R, N, STRIDED, -1, 1, 0>::THREADS % 64 == 0) {
                    if (kind == 5) return launch_cfg<PassKS<R, N, STRIDED, -1, 1, 0, 2, S>>(a, st);
                    return launch_cfg<PassKS<R, N, STRIDED, -1, 0, 0, 2, S>>(a, st);
                }
                return set_error(AW_ERR_INVALID_VALUE,
                                 "partials after the factor: column passes of whole waves only");
            case 9:
            case 10:
                if constexpr (STRIDED && PassKS<R, N, STRIDED, -1, 1, 0>::THREADS % 64 == 0) {
                    if (kind == 9) return launch_cfg<PassKS<R, N, STRIDED, -1, 1, 0, 3, S>>(a, st);
                    return launch_cfg<PassKS<R, N, STRIDED, -1, 0, 0, 3, S>>(a, st);
                }
                return set_error(AW_ERR_INVALID_VALUE, "von Neumann observables: column passes of whole waves only");
            case 11:
            case 12:
                if constexpr (STRIDED && PassKS<R, N, STRIDED, -1, 1, 0>::THREADS % 64 == 0) {
                    if (kind == 11) return launch_cfg<PassKS<R, N, STRIDED, -1, 1, 0, 4, S>>(a, st);
                    return launch_cfg<PassKS<R, N, STRIDED, -1, 0, 0, 4, S>>(a, st);
                }
                return set_error(AW_ERR_INVALID_VALUE, "sums of the closed step: column passes of whole waves only");
            case 7:
            case 8:
            case 1: if (kind == 1) return launch_cfg<PassKS<R, N, STRIDED, 0, 1, 0, 0, S>>(a, st);
                    return launch_cfg<PassKS<R, N, STRIDED, 2, 2, 1, 0, S>>(a, st);
            case 2: return launch_cfg<PassKS<R, N, STRIDED, -1, 1, 0, 0, S>>(a, st);
            case 3: return launch_cfg<PassKS<R, N, STRIDED, -1, 0, 0, 0, S>>(a, st);
            default: return launch_cfg<PassKS<R, N, STRIDED, 2, 2, 1, 0, S>>(a, st);
            }
        }
    }
    switch (kind) {
    case 4:
        if constexpr (!STRIDED) return launch_cfg<PassK<R, N, STRIDED, -1, 1, 0, 1, S>>(a, st);
        return set_error(AW_ERR_INVALID_VALUE, "weighted-sum partials need a row pass");
    case 5:
    case 6:
        if constexpr (STRIDED && PassK<R, N, STRIDED, -1, 1, 0>::THREADS % 64 == 0) {
            if (kind == 5) return launch_cfg<PassK<R, N, STRIDED, -1, 1, 0, 2, S>>(a, st);
            return launch_cfg<PassK<R, N, STRIDED, -1, 0, 0, 2, S>>(a, st);
        }
        return set_error(AW_ERR_INVALID_VALUE, "partials after the factor: column passes of whole waves only");
    case 9:
    case 10:
        if constexpr (STRIDED && PassK<R, N, STRIDED, -1, 1, 0>::THREADS % 64 == 0) {
            if (kind == 9) return launch_cfg<PassK<R, N, STRIDED, -1, 1, 0, 3, S>>(a, st);
            return launch_cfg<PassK<R, N, STRIDED, -1, 0, 0, 3, S>>(a, st);
        }
        return set_error(AW_ERR_INVALID_VALUE, "von Neumann observables: column passes of whole waves only");
    case 11:
    case 12:
        if constexpr (STRIDED && PassK<R, N, STRIDED, -1, 1, 0>::THREADS % 64 == 0) {
            if (kind == 11) return launch_cfg<PassK<R, N, STRIDED, -1, 1, 0, 4, S>>(a, st);
            return launch_cfg<PassK<R, N, STRIDED, -1, 0, 0, 4, S>>(a, st);
        }
        return set_error(AW_ERR_INVALID_VALUE, "sums of the closed step: column passes of whole waves only");
    case 7: return launch_cfg<PassK<R, N, STRIDED, -1, 1, 1, 0, S>>(a, st);
    case 8: return launch_cfg<PassK<R, N, STRIDED, 1, 0, 1, 0, S>>(a, st);
    case 1: return launch_cfg<PassK<R, N, STRIDED, 0, 1, 0, 0, S>>(a, st);
    case 2: return launch_cfg<PassK<R, N, STRIDED, -1, 1, 0, 0, S>>(a, st);
    case 3: return launch_cfg<PassK<R, N, STRIDED, -1, 0, 0, 0, S>>(a, st);
    default: return launch_cfg<PassK<R, N, STRIDED, 2, 2, 1, 0, S>>(a, st);
    }
}

// lines per block of the pass that launch_kind_n picks for total_lines lines of n points,
// or 0 where its blocks are not whole waves (no weighted-sum partials there)
template <typename R, int N, bool STRIDED> int obs_lpb_n(long long total_lines) {
    if constexpr (N <= 2048) {
        if (total_lines <= small_lines_max<R, N>()) {
            using C = PassKS<R, N, STRIDED, -1, 1, 0>;
            return C::THREADS % 64 == 0 ? C::LPB : 0;
        }
    }
    using C = PassK<R, N, STRIDED, -1, 1, 0>;
    return C::THREADS % 64 == 0 ? C::LPB : 0;
}

template <typename R> int obs_lpb(int n, bool strided, long long total_lines) {
#define AW_CASE(NN) \
    case NN: return strided ? obs_lpb_n<R, NN, true>(total_lines) : obs_lpb_n<R, NN, false>(total_lines);
    switch (n) {
        AW_CASE(64) AW_CASE(128) AW_CASE(256) AW_CASE(512) AW_CASE(1024) AW_CASE(2048) AW_CASE(4096)
    default: return 0;
    }
#undef AW_CASE
}

// The step passes of a sequence that keeps its state in line fixed point (complex64 states
// with FP64 arithmetic, a.q): the opening column pass [factor, inverse] writes it (q = 2),
// the passes [forward, factor, inverse] read and write it (q = 3) and the closing column pass
// [forward, factor] reads it (q = 1).  Same tiling as the complex64 step passes.
template <typename R, int N, bool STRIDED, typename S>
int launch_step_q_n(int kind, const PassArgs &a, hipStream_t st) {
    const int want = kind == 1 ? 2 : (kind == 2 ? 3 : (kind == 3 ? 1 : -1));
    if (a.q != want || (!STRIDED && kind != 2) || !a.q_in != !(a.q & 1) || !a.q_out != !(a.q & 2))
        return set_error(AW_ERR_INVALID_VALUE, "line fixed point: no step pass of kind %d with q = %d here", kind, a.q);
    if constexpr (N <= 2048) {
        if (a.total_lines <= small_lines_max<R, N>()) {
            if constexpr (STRIDED) {
                if (kind == 1) return launch_cfg<PassKS<R, N, true, 0, 1, 0, 0, S, 2>>(a, st);
                if (kind == 3) return launch_cfg<PassKS<R, N, true, -1, 0, 0, 0, S, 1>>(a, st);
            }
            return launch_cfg<PassKS<R, N, STRIDED, -1, 1, 0, 0, S, 3>>(a, st);
        }
    }
    if constexpr (STRIDED) {
        if (kind == 1) return launch_cfg<PassK<R, N, true, 0, 1, 0, 0, S, 2>>(a, st);
        if (kind == 3) return launch_cfg<PassK<R, N, true, -1, 0, 0, 0, S, 1>>(a, st);
    }
    return launch_cfg<PassK<R, N, STRIDED, -1, 1, 0, 0, S, 3>>(a, st);
}

template <typename R, typename S>
int launch_step_q(int n, bool strided, int kind, const PassArgs &a, hipStream_t st) {
#define AW_CASE(NN) \
    case NN: return strided ? launch_step_q_n<R, NN, true, S>(kind, a, st) : launch_step_q_n<R, NN, false, S>(kind, a, st);
    switch (n) {
        AW_CASE(64) AW_CASE(128) AW_CASE(256) AW_CASE(512) AW_CASE(1024) AW_CASE(2048) AW_CASE(4096)
    default: return set_error(AW_ERR_INVALID_SIZE, "unsupported line length %d (64..4096, power of two)", n);
    }
#undef AW_CASE
}

template <typename R, typename S = R>
int launch_pass_prec(int n, bool strided, int kind, const PassArgs &a, hipStream_t st) {
#define AW_CASE(NN) \
    case NN: return strided ? launch_kind_n<R, NN, true, S>(kind, a, st) : launch_kind_n<R, NN, false, S>(kind, a, st);
    switch (n) {
        AW_CASE(64) AW_CASE(128) AW_CASE(256) AW_CASE(512) AW_CASE(1024) AW_CASE(2048) AW_CASE(4096)
    default: return set_error(AW_ERR_INVALID_SIZE, "unsupported line length %d (64..4096, power of two)", n);
    }
#undef AW_CASE
}

}  // namespace awig
