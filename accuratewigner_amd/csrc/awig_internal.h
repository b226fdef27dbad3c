// Internal declarations shared by the libawig translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/awig.h"

namespace awig {

extern thread_local std::string g_err;
int set_error(int code, const char *fmt, ...);

// kernel-side arguments of one fused line pass (see k_line_pass)
struct PassArgs {
    const void *in;
    void *out;
    long long bstride;      // elements between batch items
    long long line_stride;  // elements between consecutive lines
    long long elem_stride;  // elements between consecutive elements of a line
    long long total_lines;  // nlines * batch
    int nlines;             // lines per batch item (a power of two, like every plan axis)
    int nl_shift;           // log2(nlines)
    int dir_a, dir_b;
    int sign_in, sign_out, out_mode;
    int has_factor;
    const void *line_fac; long long lf_bstride;
    const void *elem_fac; long long ef_bstride;
    const void *mask; long long m_bstride; long long m_rowstride;
    double chirp, lstep, lcenter, estep, ecenter;
    int lfft, efft;
    double scale;
    const double *dscale;
    const void *tw;         // N roots exp(-2 pi i t / N), plan precision
    const void *ctab; long long c_bstride;   // complex per-element factor table (general passes)
    double chirp_out;                        // exp(i chirp_out lcoord ecoord) after the second FFT
    const double2 *obs_w;   // complex128 weights [obs_nw][rows][cols] shared by the items
    int obs_nw;             // 1..8
    double *obs_partials;   // [batch][nlines / LPB][obs_nw + obs_norm2]: block partials
    int obs_norm2;          // 1: one more component, sum |v|^2 (times obs_n2w when given)
    const double *obs_n2w;  // real weights [rows][cols] of that component, or NULL
    int obs_mode;           // 0: partials of the pass output (row passes); 1: of the values
                            // right after the factor multiply (column passes); 2: von Neumann
                            // observables of the state after the first FFT and the closing
                            // vectors obs_lf / obs_ef (column passes of square items); 3: dense
                            // weighted sums of that state
    const void *obs_lf, *obs_ef;
    int obs_nd;             // obs_mode 2: diagonal weights (obs_n2w rows)
    void *aux_out;          // general passes: the loaded input times dscale[b], stored here
    int q;                  // step passes of complex64 states with FP64 arithmetic: bit 0 the
                            // input, bit 1 the output in line fixed point (int32 pairs, one
                            // power-of-two scale word per line)
    const int *q_in;        // [batch][N] scale words of the input (lines of the other axis)
    int *q_out;             // [batch][nlines] scale words of the output
};

// obs_mode 2: at most this many cyclic (momentum) and diagonal (coordinate) weight rows
constexpr int kVnMaxNw = 3;
constexpr int kVnMaxNd = 4;

struct FactorArgs {
    void *out; long long out_stride;
    int n, nvec, mode;
    const double *A; long long a_stride; double ca;
    const double *B; long long b_stride; double cb;
    double c0;
    const double *mag; long long mag_stride; int mag_pow;
    double scale;
    int alt_sign;
};

int launch_pass(int prec, int n, bool strided, const PassArgs &a, hipStream_t st);
// lines per block of the row step pass with weighted-sum partials (axis 1, line length n)
int obs_lpb(int prec, int n, bool strided, long long total_lines);
int launch_make_factors(int prec, const FactorArgs &f, hipStream_t st);
int launch_stats(int prec, const void *data, long long n_item, int rows, int cols, int batch,
                 double *partials, int nblk, double *out, hipStream_t st);
int launch_pair_dot(int prec, const void *a, const void *b, long long n_item, int cols, const double *u,
                    const double *v, int batch, double *partials, int nblk, double *out, hipStream_t st);
constexpr int kDiagRowsPerBlock = 16;
// blocks of k_cyclic_obs per item (64 diagonals each): lines up to 4096 points
constexpr int kCycObsMaxBlocks = 64;
int launch_cyclic_diag(int prec, const void *data, int n, int batch, double *work, double *out, hipStream_t st);
constexpr int kWsumMaxBlocks = 512;
int launch_weighted_sums(int prec, const void *data, long long n_item, const double *w, int nw, int norm2,
                         int batch, double *partials, int nblk, double *out, double factor, double *dscale,
                         hipStream_t st);
int launch_finish_sums(const double *partials, int nblk, int ncomp, int batch, double *out, double factor,
                       double *dscale, hipStream_t st);
int launch_weighted_partials(int prec, const void *data, long long n_item, const double *w, int nw, int norm2,
                             int batch, double *partials, int nblk, hipStream_t st);
int launch_cyclic_obs(int prec, const void *data, int n, int batch, double *work, const double *xw, int nx,
                      const double *pw_hat, int np, double *partials, unsigned *counters, double *out,
                      long long out_bstride, double factor, double *dscale, hipStream_t st);
int launch_sum_partials(const double *partials, int nblk, int ncomp, int batch, double *out,
                        hipStream_t st);
int launch_scale_from_stats(const double *stats, int stride, int comp, int mode, double factor,
                            double *dscale, int batch, hipStream_t st);
int launch_scale(int prec, void *data, long long n, int batch, const double *dscale, double scale,
                 hipStream_t st);
int launch_project_out(int prec, void *psi, const void *phis, int nphi, int n, hipStream_t st);
int launch_convert(const void *src, int sp, void *dst, int dp, long long n, hipStream_t st);
// the packing of real <-> complex transforms (modes: see k_real_complex)
int launch_real_complex(int prec, int mode, const void *in, void *out, long long nl, int n, int cols, int axis,
                        hipStream_t st);

}  // namespace awig
