/*
 * libawig: MI355X (gfx950) split-operator engine for phase-space propagators.
 *
 * C ABI: plain pointers and sizes, no torch types.  Every function returns an int status
 * (AW_OK = 0); the message of the last failure on the calling thread is returned by
 * aw_last_error().  Device pointers are HIP device addresses owned by the caller; the
 * library owns plans (twiddle tables, partial-sum scratch) and compiled expression modules.
 * A plan is used from one stream at a time (not re-entrant per plan, safe across plans).
 *
 * Arrays are C-order [batch][rows][cols] of complex (interleaved re, im), complex128
 * (prec AW_PREC_F64) or complex64 (AW_PREC_F32 and AW_PREC_C64).  Factor vectors, complex
 * tables and real masks are in the plan's arithmetic precision: FP32 for AW_PREC_F32, FP64
 * otherwise.  "axis 0" FFTs run along rows index i
 * (strided, one FFT per column), "axis 1" FFTs along the contiguous index j, exactly the
 * two plan kinds of the reference FFT wrapper.
 */
#ifndef AWIG_H
#define AWIG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the nonzero codes mirror the exception classes of cufft.py:31-104 */
#define AW_OK 0
#define AW_ERR_INVALID_PLAN 1
#define AW_ERR_ALLOC_FAILED 2
#define AW_ERR_INVALID_TYPE 3
#define AW_ERR_INVALID_VALUE 4
#define AW_ERR_INTERNAL 5
#define AW_ERR_EXEC_FAILED 6
#define AW_ERR_SETUP_FAILED 7
#define AW_ERR_INVALID_SIZE 8
#define AW_ERR_RTC 9
#define AW_ERR_HIP 10

#define AW_PREC_F64 0
#define AW_PREC_F32 1
#define AW_PREC_C64 2   /* complex64 states in HBM; twiddles, factor vectors, masks and every
                           FFT and factor operation in FP64 (converted on load and store) */

#define AW_FORWARD (-1) /* exp(-2 pi i jk/N), unnormalised (cufft.py:6) */
#define AW_INVERSE 1    /* exp(+2 pi i jk/N), unnormalised (cufft.py:7) */

typedef struct aw_plan aw_plan;
typedef struct aw_module aw_module;
typedef void *aw_stream; /* hipStream_t; NULL = default stream */

/* Library version (major*10000 + minor*100 + patch) and last error text. */
int aw_version(void);
const char *aw_last_error(void);

/*
 * Plan for a batch of [rows, cols] grids.  Replaces cufftPlanMany as called by
 * Plan_Z2Z_2D_Axis0 / Plan_Z2Z_2D_Axis1 (cufft.py:274-345; the plan call itself at
 * cufft.py:107-116): one plan serves both axes.  rows, cols: powers of two, 64..4096
 * (cols may also be a single row of length 64..4096 with rows = 1).
 */
int aw_plan_create(aw_plan **plan, int rows, int cols, int batch, int prec);
int aw_plan_destroy(aw_plan *plan); /* cufftDestroy, cufft.py:119-127 */

/*
 * In-place (or out-of-place when out != data) unnormalised batched FFT along one axis
 * (axis 0 or 1), dir = AW_FORWARD / AW_INVERSE.  Replaces cufftExecZ2Z through
 * fft_Z2Z / ifft_Z2Z (cufft.py:226-238).
 */
int aw_fft_axis(aw_plan *plan, void *data, void *out, int axis, int dir, aw_stream stream);

/*
 * Real-to-complex forward FFT along one axis (unnormalised): real input with the plan's
 * shape (double for AW_PREC_F64, float for AW_PREC_F32), complex output keeping the n/2 + 1
 * non-negative frequencies of each line: [batch][rows][cols/2 + 1] along axis 1,
 * [batch][rows/2 + 1][cols] along axis 0.  Replaces cufftExecD2Z / cufftExecR2C through
 * cu_fft_D2Z / cu_ifft_D2Z (cufft.py:129-138).  aw_fft_c2r is the complex-to-real inverse
 * (the half spectrum in, the n reals of each line out; the imaginary parts of the zero and
 * Nyquist frequencies are ignored), replacing cufftExecZ2D / cufftExecC2R through
 * cu_fft_Z2D / cu_ifft_Z2D (cufft.py:146-157).  Out of place.  The first call on a plan
 * allocates a complex workspace of the plan's size and must not be captured.
 */
int aw_fft_r2c(aw_plan *plan, const void *in, void *out, int axis, aw_stream stream);
int aw_fft_c2r(aw_plan *plan, const void *in, void *out, int axis, aw_stream stream);

/*
 * One fused line pass over every line of the chosen axis of every batch item:
 *     load -> [sign_in] -> [FFT dir_a] -> multiply by factor -> [FFT dir_b] -> [sign_out]
 *          -> store (or store the diagonal element only)
 * with factor(line l, element e) =
 *     scale * dscale[b] * line_fac[b][l] * elem_fac[b][e] * mask[b][row, col]
 *     * ctab[b][row, col] * exp(i * chirp * lcoord(l) * ecoord(e)),
 * and, after the second FFT, a multiply by exp(i * chirp_out * lcoord(l) * ecoord(e)) (the
 * inter-level twiddles of a line longer than 4096 points split over rows and columns).
 * lcoord(l) = lstep * (lpos(l) - lcenter), ecoord likewise, where lpos is l itself or its
 * FFT-ordered index ((l + n/2) % n - n/2 + lcenter) when lfftorder is set.
 * It replaces the reference's expV / expK / bloch_expV / bloch_expK / phase_shearX /
 * phase_shearY / sign_flip / observable kernels together with the cuFFT calls around
 * them, e.g. the sequence of single_step_propagation (wigner_moyal_cuda_1d.py:543-561).
 * Vectors and ctab are complex in the plan's precision; mask is real in the plan's precision.
 * NULL pointers mean "all ones".  out = NULL or data means in place.
 */
typedef struct aw_pass {
    int axis;          /* 0: lines are columns (FFT along rows index i); 1: lines are rows */
    int dir_a, dir_b;  /* 0 (none), AW_FORWARD or AW_INVERSE */
    int sign_in, sign_out; /* 1: multiply by (-1)^(i+j) (0: no, 2: (-1)^i only) */
    int out_mode;      /* 0: store the line; 1: store only element e == l into out[b][l] */
    const void *line_fac; long long line_fac_bstride; /* elements between batch items */
    const void *elem_fac; long long elem_fac_bstride;
    const void *mask; long long mask_bstride;         /* real [rows][cols] */
    double chirp, lstep, lcenter, estep, ecenter;
    int lfftorder, efftorder;
    double scale;
    const double *dscale; /* complex (re, im) per batch item, or NULL */
    void *out;
    const void *ctab; long long ctab_bstride;         /* complex [rows][cols] factor table */
    double chirp_out;
    /* Block partials of the weighted sums Re sum conj(obs_weights[q]) v (q < obs_nw <= 8;
     * complex128 [obs_nw][rows][cols], shared by the items) into obs_partials[b][k][c],
     * k < aw_pass_obs_blocks(), c < obs_nw + obs_norm2, to be added by aw_finish_sums.
     *   obs_mode 0: v is the pass output; plain [forward, factor, inverse] row passes.
     *   obs_mode 1: v is the line right after the factor multiply; plain [forward, factor,
     *               inverse] or [forward, factor] column passes.  With the factor of a
     *               merged pass F = C0 * C1 (C1 closing one step, C0 opening the next) and
     *               weights w / conj(C0), these are the sums of the state between the steps
     *               (wigner_bloch_cuda_1d.py:139-147 takes them after every step).
     * obs_norm2 = 1 adds the component sum |v|^2 (times obs_norm2_weights, real
     * [rows][cols], when given).  NULL weights: none. */
    const void *obs_weights; int obs_nw;
    double *obs_partials;
    int obs_mode, obs_norm2;
    const double *obs_norm2_weights;
    /* With dscale: also store each loaded element times dscale[b] (before the sign flip and
     * the first FFT) into in_scaled_out (the input's layout; may be the input itself), e.g. the
     * division of rho by its norm riding on the first pass of its Wigner transform
     * (normalize_rho_wigner, wigner_moyal_cuda_1d.py:608-622).  NULL: none. */
    void *in_scaled_out;
    /* obs_mode 2 (column step passes of square items, [forward, factor(, inverse)]): block
     * partials of the von Neumann observables of the state between two steps,
     *   z(i, j) = (line j right after the first FFT)[i] * obs_line_fac[b][j] * obs_elem_fac[b][i]
     * (the closing factor vectors of the step, complex in the plan's precision, with the
     * batch strides of line_fac / elem_fac; NULL: ones), into obs_partials[b][k][c]:
     *   c = 0:                 Re sum_j z(j, j)                              (the trace)
     *   c = 1 .. nd:           Re sum_j z(j, j) obs_norm2_weights[c - 1][j]  (real [nd][n])
     *   c = 1 + nd .. nd + nw: Re sum_ij (-1)^(j-i) z(i, j) obs_weights[q][(j - i) mod n] / n
     *                          (complex128 [nw][n]: the DFTs of momentum weights, so that the
     *                          sum is Re sum_k w(P_k) rho~(k, k) of the momentum representation)
     * with nd = obs_diag_nw (0..4), nw = obs_nw (1..3) and obs_norm2 = 0.  This is the
     * trace and the Ehrenfest averages of rho_vneumann_cuda_1d.py:441-448 and :596-644 taken
     * inside the merged column pass that closes one step and opens the next. */
    /* obs_mode 3 (column step passes, [forward, factor(, inverse)]): the weighted sums
     * Re sum conj(obs_weights[q]) z (q < obs_nw <= 8, complex128 [obs_nw][rows][cols] shared by
     * the items, obs_norm2 = 0) of the same state z between two steps, into obs_partials[b][k][q]:
     * the Wigner-Moyal Ehrenfest averages of a step (wigner_moyal_cuda_1d.py:584-606) as
     * Re <L^H f, rho> when the phases change every step (a time-dependent V). */
    const void *obs_line_fac, *obs_elem_fac;
    int obs_diag_nw;
} aw_pass;

int aw_pass_run(aw_plan *plan, void *data, const aw_pass *pass, aw_stream stream);

/* Blocks per item (k range) of the obs_partials of a pass along axis (0: columns, 1: rows)
 * over the plan's batch; AW_ERR_INVALID_SIZE where that pass has no partials (blocks of
 * fewer than 64 threads: columns of 64 points in single-state passes). */
int aw_pass_obs_blocks(const aw_plan *plan, int axis, int *nblk);
/* The same for rows (axis 1). */
int aw_row_obs_blocks(const aw_plan *plan, int *nblk);

/*
 * n merged Strang steps of a 2D split-operator propagator in 2n + 1 fused passes:
 *   [col: factor c_0 -> I0]  then for s < n:  [row: F1 -> factor r_s -> I1]
 *   [col: F0 -> factor c_{s+1} -> (I0 unless s == n-1)]
 * which is the sequence EV I0 F1 EK F0 I1 EV of single_step_propagation
 * (wigner_moyal_cuda_1d.py:543-561, rho_vneumann_cuda_1d.py:450-468, and the Bloch steps
 * wigner_bloch_cuda_1d.py:221-240, rho_bloch_cuda_1d.py:225-246) repeated, with the closing
 * and opening half-kicks of consecutive steps merged into one column factor.
 * Column factor c_v = col_line[b][v][j] * col_elem[b][v][i] (v = 0..n), row factor
 * r_s = row_line[b][s][i] * row_elem[b][s][j]; *_bstride are elements between batch items
 * (0 shares one table across the batch).  Optional real masks (rows x cols per item):
 * mask_open_close for columns 0 and n, mask_mid for the merged ones, mask_row for rows.
 * dscale (complex per item) multiplies the opening pass.  Launches only; no host sync.
 * const_steps = 1 (time-independent H): the column table holds 3 vectors per item
 * (0 opening, 1 merged, 2 closing) and the row table 1, reused for every step.
 * batch > 0 runs only the first `batch` items of the plan (a slot pool's active prefix).
 * A batch larger than 192 MiB with n >= 2 is stepped in groups of items of at most that
 * size, each group through all 2n + 1 passes, so that it stays in the MALL between passes;
 * the results are bit-identical to stepping the whole batch per pass.
 * AW_PREC_C64 plans hold the state between the opening and the closing pass as int32 pairs
 * with one power-of-two scale per line (rounding to 2^-30 of the line's largest component
 * instead of 2^-24 of each value); `data` is complex64 on entry and on return.
 */
typedef struct aw_steps {
    int nsteps;
    int const_steps;
    int batch;
    const void *col_line, *col_elem, *row_line, *row_elem;
    long long col_line_bstride, col_elem_bstride, row_line_bstride, row_elem_bstride;
    const void *mask_open_close, *mask_mid, *mask_row;
    long long mask_bstride;
    const double *dscale;
} aw_steps;

int aw_step_sequence(aw_plan *plan, void *data, const aw_steps *steps, aw_stream stream);

/*
 * Items per group that aw_step_sequence uses for `nsteps` steps over the first `batch`
 * items (0: the plan's batch); equal to the batch when it is not split.  No reference
 * counterpart (the reference propagates one state): lets a caller relate kernel launches
 * to passes over its batch.
 */
int aw_step_group(const aw_plan *plan, int nsteps, int batch, int *group);

/*
 * Build complex factor vectors in the plan's precision from FP64 inputs:
 *   out[v * out_stride + k] = scale * s(k) * mag[v][k]^mag_pow * E(c0 + ca * A[v][k] + cb * B[v][k])
 * with E(x) = exp(i x) (mode 0) or exp(x) (mode 1), s(k) = (-1)^k if alt_sign else 1.
 * nvec vectors of length n, out_stride elements apart (0 means n, i.e. packed; a larger
 * stride interleaves several tables); A, B, mag may be NULL (zero / zero / one) and have strides
 * a_stride, b_stride, mag_stride (0 broadcasts one vector to all nvec).  Phase arguments
 * are summed and exponentiated in FP64 (the reference evaluates every phase in double,
 * e.g. wigner_moyal_cuda_1d.py:751, 772); only the result is rounded to the plan precision.
 */
int aw_make_factors(aw_plan *plan, void *out, long long out_stride, int n, int nvec, int mode,
                    const double *A, long long a_stride, double ca,
                    const double *B, long long b_stride, double cb,
                    double c0, const double *mag, long long mag_stride, int mag_pow,
                    double scale, int alt_sign, aw_stream stream);

/*
 * Reductions over each batch item of a [rows, cols] state (FP64 accumulation,
 * deterministic order).  out[b][0..7] = { Re sum z, Im sum z, sum |z|^2, Re sum z^2,
 * Im sum z^2, Re trace, Im trace, 0 } (trace = sum of z[k, k], square grids only).
 * Replaces gpuarray.sum (wigner_moyal_cuda_1d.py:415, 617), gpuarray.dot without
 * conjugation (:354), cu_linalg.trace (rho_vneumann_cuda_1d.py:383, 443) and the purity
 * ReductionKernel (:574-579).  out is device memory (8 doubles per item).
 */
int aw_reduce_stats(aw_plan *plan, const void *data, double *out, aw_stream stream);

/*
 * Weighted inner products of two states of the plan's shape, per batch item (FP64):
 *   out[b][0..3] = { Re, Im of sum_ij conj(a_ij) b_ij u_j,  Re, Im of sum_ij conj(a_ij) b_ij v_i }
 * u has cols entries, v has rows entries (device FP64).  No reference counterpart: the
 * gradient of the adjoint with respect to a dipole field (the derivative of the expV
 * factor, wigner_moyal_cuda_1d.py:763-778, is -i dt/2 (XX_j - XX'_i) times that factor).
 */
int aw_pair_dot(aw_plan *plan, const void *a, const void *b, const double *u, const double *v,
                double *out, aw_stream stream);

/*
 * Real parts of weighted sums of each batch item (FP64 accumulation, fixed order):
 *   out[b][q] = Re sum_ij conj(w[q][i][j]) data[b][i][j],   q < nw,
 * and with norm2 != 0 also out[b][nw] = sum_ij |data[b][i][j]|^2 (nw + norm2 <= 8);
 * weights complex128 [nw][rows][cols], shared by every item; out [batch][nw + norm2] FP64.
 * With dscale != NULL also dscale[b] = 1 / (out[b][0] * factor) (complex, imaginary part 0),
 * the normalisation of the next step.  Two launches (block partials, then one block per
 * item adds them).  The propagators record the reference's per-step Ehrenfest averages
 * sum Re W f_q dXdP (wigner_moyal_cuda_1d.py:584-606, W from rho2wigner_ :471-501) as
 * Re <g_q, rho> with g_q = rho2wigner^H f_q when f_q does not depend on t; the Gibbs loop's
 * purity check 2 pi Re sum W^2 dXdP < 1 (wigner_bloch_cuda_1d.py:139-147 with get_purity,
 * wigner_moyal_cuda_1d.py:349-354) from
 * sum W^2 = Np sum |rho|^2 (the transform is Np times an isometry).  The first call
 * on a plan allocates its scratch and must not be made under stream capture; calls on one
 * plan must not overlap.
 */
int aw_weighted_sums(aw_plan *plan, const void *data, const double *weights, int nw, int norm2, double *out,
                     double factor, double *dscale, aw_stream stream);

/*
 * aw_weighted_sums without a plan and without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The block partials live in
 * a fixed device array of 512 x 8 x 256 doubles (8 MiB, held by each GPU the library runs on
 * until the process ends), so nblk * (nw + norm2) * batch must fit in it (nblk = rows * cols /
 * 2048, clamped to 1..512), else AW_ERR_INVALID_SIZE.  The results are bit-identical to
 * aw_weighted_sums on a plan of the same shape and precision.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (different streams, or threads) race.
 */
int aw_weighted_sums_noalloc(int prec, int rows, int cols, int batch, const void *data, const double *weights,
                             int nw, int norm2, double *out, double factor, double *dscale, aw_stream stream);

/*
 * The first stage of aw_weighted_sums alone: block partials
 * partials[b][k][c] (c < nw + norm2, k < *nblk) into a caller buffer, one launch, no scratch
 * (so any call may be captured).  *nblk receives the block count (partials = NULL: only that).
 * aw_finish_sums over several such buffers laid end to end adds many steps' sums in one launch:
 * the propagators record a chunk of steps this way (the sums are ratios, so the per-step
 * normalisation of the reference, wigner_moyal_cuda_1d.py:608-622, commutes with the linear
 * steps and is applied once at the end).
 */
int aw_weighted_partials(aw_plan *plan, const void *data, const double *weights, int nw, int norm2,
                         double *partials, int *nblk, aw_stream stream);

/*
 * Per-step Ehrenfest sums of square density matrices (n >= 64), FP64, two launches:
 *   out[b][0]          = Re tr rho[b],
 *   out[b][1 + q]      = sum_i Re rho[b][i][i] xw[q][i]                  (q < nx),
 *   out[b][1 + nx + q] = Re sum_d c_d pw_hat[q][d] / n                   (q < np),
 * with c_d = (-1)^d sum_i rho[b][i][(i + d) mod n] and pw_hat[q] the forward DFT of a real
 * momentum weight pw[q] (complex128 [np][n]), so that the last group equals
 * sum_k pw[q][k] Re FFT(c)_k / n: the averages that the reference's get_Ehrenfest
 * (rho_vneumann_cuda_1d.py:470-501) takes through get_average (:520-563), i.e. weighted
 * traces after the ket index is moved to momentum space.  xw: FP64 [nx][n].  1 + nx + np <= 8; rows of out are
 * out_bstride doubles apart.  With dscale != NULL also dscale[b] = 1 / (out[b][0] * factor).
 * The first call on a plan allocates its scratch and must not be made under stream capture.
 */
int aw_cyclic_diag_obs(aw_plan *plan, const void *data, const double *xw, int nx, const double *pw_hat,
                       int np, double *out, long long out_bstride, double factor, double *dscale,
                       aw_stream stream);

/*
 * Signed cyclic diagonal sums of square items (n >= 64), FP64 complex:
 *   out[b][d] = (-1)^d sum_i data[b][i][(i + d) mod n],  d = 0..n-1   (2 n doubles per item).
 * The forward DFT of this vector is the momentum-space diagonal of rho scaled by n, so a
 * momentum observable sum_j [S I1 (g(P_j)/n) F1 S rho]_jj (rho_vneumann_cuda_1d.py:551-560,
 * :676-684, :811-821) is Re sum_k g(P_k) F(out)_k / n: one read of rho for any number of
 * g.  The plan allocates its scratch (2 n^2 / 16 doubles per item) on first use.
 */
int aw_cyclic_diag_sums(aw_plan *plan, const void *data, double *out, aw_stream stream);

/*
 * out[b][c] = sum_i partials[b][i][c] for i < nblk, c < ncomp, summed in fixed order (the
 * second stage of every block-partial reduction, including the hipRTC observable kernels).
 */
int aw_sum_partials(const double *partials, int nblk, int ncomp, int batch, double *out,
                    aw_stream stream);

/*
 * The same sums (ncomp <= 8) added by a whole block per item in a fixed order, and with
 * dscale != NULL the normalisation scale dscale[b] = 1 / (out[b][0] * factor) in the same
 * launch: sum_partials followed by scale_from_stats (mode 0) of the reference's per-step
 * normalisation, e.g. wigner_moyal_cuda_1d.py:608-622 after the Ehrenfest sums (:584-606).
 */
int aw_finish_sums(const double *partials, int nblk, int ncomp, int batch, double *out, double factor,
                   double *dscale, aw_stream stream);

/* Same statistics for an arbitrary complex vector of length n per item (for 1D states). */
int aw_reduce_vec(int prec, const void *data, long long n, int batch, double *out,
                  aw_stream stream);

/*
 * Scale factors from reductions (device-side, no host sync):
 *   dscale[b] = 1 / ((mode 0: Re stats[b][comp]) or (mode 1: complex stats[b][comp..comp+1])
 *               * factor)
 * stats rows are stats_stride doubles apart.  Used for normalize_rho_wigner
 * (wigner_moyal_cuda_1d.py:617-620) and trace renormalisation (rho_vneumann_cuda_1d.py:443,
 * rho_bloch_cuda_1d.py:244).
 */
int aw_scale_from_stats(const double *stats, int stats_stride, int comp, int mode,
                        double factor, double *dscale, int batch, aw_stream stream);

/* data[b][k] *= dscale[b] (complex) * scale for n complex elements per item. */
int aw_scale(int prec, void *data, long long n, int batch, const double *dscale,
             double scale, aw_stream stream);

/* Copy with precision conversion between complex128 and complex64 (n complex elements). */
int aw_convert(const void *src, int src_prec, void *dst, int dst_prec, long long n,
               aw_stream stream);

/*
 * Runtime compilation of generated HIP C source (the analogue of PyCUDA SourceModule,
 * wigner_moyal_cuda_1d.py:241-258) for gfx950 with hipRTC, and generic launch of its
 * extern "C" kernels.  log receives the compiler log (may be NULL).
 */
int aw_module_compile(aw_module **mod, const char *source, const char *name,
                      char *log, size_t log_len);
int aw_module_destroy(aw_module *mod);
/* Compile only (no device needed): status, compiler log and code-object size. */
int aw_rtc_check(const char *source, char *log, size_t log_len, size_t *code_size);
int aw_module_launch(aw_module *mod, const char *kernel, unsigned gx, unsigned gy, unsigned gz,
                     unsigned bx, unsigned by, unsigned bz, unsigned shmem, aw_stream stream,
                     void **args);

/*
 * Gram-Schmidt step of projectout_stationary_states (schrodinger_wigner_cuda_1d.py:505-550)
 * on one vector of n complex elements (FP64 sums):
 *   psi /= ||psi||;  c_k = sum conj(phi_k) psi;  psi -= sum_k c_k phi_k;  psi /= ||psi||.
 * phis holds nphi vectors of length n contiguously (unit 2-norm, like the reference's
 * stationary_states); n <= 65536, nphi <= 4096.  Lines below 8192 points run in one
 * workgroup; longer ones (nphi <= 2048) in three launches of 128 workgroups over slices of
 * the line, whose partial sums live in a device buffer the library allocates on the first
 * such call on a device (not under stream capture) and shares between calls: such calls on
 * one device must not overlap.
 */
int aw_project_out(int prec, void *psi, const void *phis, int nphi, int n, aw_stream stream);

/* Free / total device memory in bytes (print_memory_info, wigner_moyal_cuda_1d.py:368-377). */
int aw_mem_info(size_t *free_bytes, size_t *total_bytes);

/* hipDeviceSynchronize / hipStreamSynchronize with error capture. */
int aw_sync(aw_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* AWIG_H */
