// libawig C ABI: plans, fused passes, factor builders, reductions and hipRTC modules.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "awig_internal.h"

using namespace awig;

struct aw_plan {
    int rows, cols, batch, prec;
    int csize;                 // bytes per complex element
    void *tw_rows = nullptr;   // roots for FFTs of length rows (axis 0)
    void *tw_cols = nullptr;   // roots for FFTs of length cols (axis 1)
    double *partials = nullptr;
    int red_blocks = 0;
    double *diag_work = nullptr;   // cyclic-diagonal partial sums, allocated on first use
    double *wsum_partials = nullptr;    // weighted-sum block partials, allocated on first use
    unsigned *wsum_counters = nullptr;  // their per-item tickets (zero between launches)
    void *rc_work = nullptr;            // full complex spectrum of the real transforms, on first use
    int *q_scales = nullptr;            // complex64 states with FP64 arithmetic: the scale words of
                                        // the line fixed point of step sequences, [batch][rows]
                                        // (row passes) then [batch][cols] (column passes)
};

struct aw_module {
    hipModule_t mod = nullptr;
    std::string name;
};

#define AW_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return set_error(AW_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// precision of a plan's twiddles, factor vectors, tables and masks (its arithmetic)
static int arith_prec(int prec) { return prec == AW_PREC_F32 ? AW_PREC_F32 : AW_PREC_F64; }
// precision of the states it reads and writes
static int state_prec(int prec) { return prec == AW_PREC_F64 ? AW_PREC_F64 : AW_PREC_F32; }

// exp(-2 pi i t / n) in FP64 with exact octant reduction of t/n before calling libm
static void host_root(long long t, long long n, double &c, double &s) {
    long long a = ((t % n) + n) % n;
    bool neg_s = false, neg_c = false, swap = false;
    if (2 * a > n) { a = n - a; neg_s = true; }
    if (4 * a > n && n % 2 == 0) { a = n / 2 - a; neg_c = true; }
    if (8 * a > n && n % 8 == 0) { a = n / 4 - a; swap = true; }
    const double th = 2.0 * 3.141592653589793238462643383279502884 * double(a) / double(n);
    double cc = std::cos(th), ss = std::sin(th);
    if (swap) std::swap(cc, ss);
    if (neg_c) cc = -cc;
    if (neg_s) ss = -ss;
    c = cc;
    s = -ss;  // forward root exp(-i theta)
}

// The flat table exp(-2 pi i t / n), t < n, followed by the stage-major tables of the
// Stockham FFTs with E = 4, 8, 16 and 32 elements per thread (n entries each): after the
// first stage (radix E), the stage of span NS and radix RAD = min(E, n / NS) owns entries
// [NS - E, NS RAD - E), entry NS - E + k (RAD - 1) + r - 1 = exp(-2 pi i r k / (NS RAD)).
static int make_twiddles(int n, int prec, void **out) {
    std::vector<double> h(2 * (size_t)n * 5, 0.0);
    for (int t = 0; t < n; ++t) host_root(t, n, h[2 * t], h[2 * t + 1]);
    for (int ei = 1; ei <= 4; ++ei) {
        const int E = 2 << ei;
        if (E >= n) continue;
        double *tab = h.data() + 2 * (size_t)n * ei;
        for (int ns = E; ns < n;) {
            const int rad = (n / ns < E) ? n / ns : E;
            for (int k = 0; k < ns; ++k)
                for (int r = 1; r < rad; ++r) {
                    const size_t q = (size_t)(ns - E + k * (rad - 1) + r - 1);
                    const size_t t = (size_t)(r * k) * (n / (ns * rad));
                    tab[2 * q] = h[2 * t];
                    tab[2 * q + 1] = h[2 * t + 1];
                }
            ns *= rad;
        }
    }
    size_t bytes;
    std::vector<float> hf;
    const void *src = h.data();
    if (prec == AW_PREC_F32) {
        hf.resize(h.size());
        for (size_t i = 0; i < hf.size(); ++i) hf[i] = (float)h[i];
        src = hf.data();
        bytes = hf.size() * sizeof(float);
    } else {
        bytes = h.size() * sizeof(double);
    }
    AW_HIP(hipMalloc(out, bytes));
    AW_HIP(hipMemcpy(*out, src, bytes, hipMemcpyHostToDevice));
    return AW_OK;
}

extern "C" {

int aw_version(void) { return 100; }

const char *aw_last_error(void) { return g_err.c_str(); }

int aw_plan_create(aw_plan **plan, int rows, int cols, int batch, int prec) {
    if (!plan) return set_error(AW_ERR_INVALID_VALUE, "plan pointer is NULL");
    *plan = nullptr;
    if (prec != AW_PREC_F64 && prec != AW_PREC_F32 && prec != AW_PREC_C64)
        return set_error(AW_ERR_INVALID_TYPE, "unsupported precision %d", prec);
    if (batch < 1) return set_error(AW_ERR_INVALID_SIZE, "batch must be >= 1 (got %d)", batch);
    if (!is_pow2(cols) || cols < 64 || cols > 4096)
        return set_error(AW_ERR_INVALID_SIZE, "cols = %d: must be a power of two in 64..4096", cols);
    if (!(rows == 1 || (is_pow2(rows) && rows >= 64 && rows <= 4096)))
        return set_error(AW_ERR_INVALID_SIZE, "rows = %d: must be 1 or a power of two in 64..4096", rows);
    aw_plan *p = new aw_plan();
    p->rows = rows;
    p->cols = cols;
    p->batch = batch;
    p->prec = prec;
    p->csize = prec == AW_PREC_F64 ? 16 : 8;
    int rc = make_twiddles(cols, arith_prec(prec), &p->tw_cols);
    if (rc == AW_OK && rows > 1) rc = make_twiddles(rows, arith_prec(prec), &p->tw_rows);
    if (rc == AW_OK) {
        const long long n = (long long)rows * cols;
        long long nb = (n + 256 * 8 - 1) / (256 * 8);
        if (nb > 1024) nb = 1024;
        if (nb < 1) nb = 1;
        p->red_blocks = (int)nb;
        hipError_t e = hipMalloc((void **)&p->partials, sizeof(double) * 8 * (size_t)nb * batch);
        if (e != hipSuccess) rc = set_error(AW_ERR_ALLOC_FAILED, "partials: %s", hipGetErrorString(e));
    }
    if (rc == AW_OK && prec == AW_PREC_C64 && rows > 1) {
        hipError_t e = hipMalloc((void **)&p->q_scales, sizeof(int) * (size_t)(rows + cols) * batch);
        if (e != hipSuccess) rc = set_error(AW_ERR_ALLOC_FAILED, "scale words: %s", hipGetErrorString(e));
    }
    if (rc != AW_OK) {
        aw_plan_destroy(p);
        return rc;
    }
    *plan = p;
    return AW_OK;
}

int aw_plan_destroy(aw_plan *p) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (p->tw_rows) (void)hipFree(p->tw_rows);
    if (p->tw_cols) (void)hipFree(p->tw_cols);
    if (p->partials) (void)hipFree(p->partials);
    if (p->diag_work) (void)hipFree(p->diag_work);
    if (p->wsum_partials) (void)hipFree(p->wsum_partials);
    if (p->wsum_counters) (void)hipFree(p->wsum_counters);
    if (p->rc_work) (void)hipFree(p->rc_work);
    if (p->q_scales) (void)hipFree(p->q_scales);
    delete p;
    return AW_OK;
}

// The line fixed point of a step pass (step sequences of complex64 states with FP64
// arithmetic): q bit 0 reads it with the scale words q_in, bit 1 writes it and its scale
// words to q_out.
struct StepQ {
    int q = 0;
    const int *q_in = nullptr;
    int *q_out = nullptr;
};

// One pass over the first `nbatch` items at `data` (the plan's batch for aw_pass_run; a group
// of items for the step sequences).
static int pass_run_n(aw_plan *p, void *data, const aw_pass *ps, int nbatch, aw_stream stream,
                      const StepQ &sq = StepQ()) {
    if (!data || !ps) return set_error(AW_ERR_INVALID_VALUE, "NULL data or pass");
    if (ps->axis != 0 && ps->axis != 1) return set_error(AW_ERR_INVALID_VALUE, "axis must be 0 or 1");
    if (ps->axis == 0 && p->rows == 1)
        return set_error(AW_ERR_INVALID_VALUE, "axis-0 pass on a single-row plan");
    for (int d : {ps->dir_a, ps->dir_b})
        if (d != 0 && d != AW_FORWARD && d != AW_INVERSE)
            return set_error(AW_ERR_INVALID_VALUE, "FFT direction must be 0, -1 or +1 (got %d)", d);
    if (ps->out_mode == 1 && p->rows != p->cols)
        return set_error(AW_ERR_INVALID_VALUE, "diagonal output needs a square grid");
    PassArgs a{};
    a.in = data;
    a.out = ps->out ? ps->out : data;
    a.bstride = (long long)p->rows * p->cols;
    const bool strided = ps->axis == 0;
    const int n = strided ? p->rows : p->cols;
    a.nlines = strided ? p->cols : p->rows;
    a.nl_shift = 0;
    while ((1 << a.nl_shift) < a.nlines) ++a.nl_shift;
    a.line_stride = strided ? 1 : p->cols;
    a.elem_stride = strided ? p->cols : 1;
    a.total_lines = (long long)a.nlines * nbatch;
    a.dir_a = ps->dir_a;
    a.dir_b = ps->dir_b;
    a.sign_in = ps->sign_in;
    a.sign_out = ps->sign_out;
    a.out_mode = ps->out_mode;
    a.line_fac = ps->line_fac;
    a.lf_bstride = ps->line_fac_bstride;
    a.elem_fac = ps->elem_fac;
    a.ef_bstride = ps->elem_fac_bstride;
    a.mask = ps->mask;
    a.m_bstride = ps->mask_bstride;
    a.m_rowstride = p->cols;
    a.chirp = ps->chirp;
    a.lstep = ps->lstep;
    a.lcenter = ps->lcenter;
    a.estep = ps->estep;
    a.ecenter = ps->ecenter;
    a.lfft = ps->lfftorder;
    a.efft = ps->efftorder;
    a.scale = ps->scale;
    a.dscale = ps->dscale;
    a.ctab = ps->ctab;
    a.c_bstride = ps->ctab_bstride;
    a.chirp_out = ps->chirp_out;
    if (ps->obs_mode == 2) {
        // von Neumann observables of the state between two steps, from a column pass
        if (!ps->obs_partials) return set_error(AW_ERR_INVALID_VALUE, "obs_mode 2 needs an obs_partials buffer");
        if (!strided || p->rows != p->cols)
            return set_error(AW_ERR_INVALID_VALUE, "obs_mode 2 needs a column pass over square items");
        if (ps->obs_nw < 1 || ps->obs_nw > kVnMaxNw || ps->obs_diag_nw < 0 || ps->obs_diag_nw > kVnMaxNd)
            return set_error(AW_ERR_INVALID_VALUE, "obs_mode 2: obs_nw must be 1..%d and obs_diag_nw 0..%d",
                             kVnMaxNw, kVnMaxNd);
        if (!ps->obs_weights || (ps->obs_diag_nw > 0 && !ps->obs_norm2_weights))
            return set_error(AW_ERR_INVALID_VALUE, "obs_mode 2: NULL weights");
        if (ps->obs_norm2) return set_error(AW_ERR_INVALID_VALUE, "obs_mode 2 takes no obs_norm2");
        a.obs_w = static_cast<const double2 *>(ps->obs_weights);
        a.obs_nw = ps->obs_nw;
        a.obs_partials = ps->obs_partials;
        a.obs_mode = 2;
        a.obs_n2w = ps->obs_norm2_weights;
        a.obs_nd = ps->obs_diag_nw;
        a.obs_lf = ps->obs_line_fac;
        a.obs_ef = ps->obs_elem_fac;
    } else if (ps->obs_mode == 3) {
        // dense weighted sums of the state between two steps, from a column pass
        if (!ps->obs_partials || !ps->obs_weights)
            return set_error(AW_ERR_INVALID_VALUE, "obs_mode 3 needs weights and an obs_partials buffer");
        if (!strided) return set_error(AW_ERR_INVALID_VALUE, "obs_mode 3 needs a column pass");
        if (ps->obs_nw < 1 || ps->obs_nw > 8) return set_error(AW_ERR_INVALID_VALUE, "obs_mode 3: obs_nw must be 1..8");
        if (ps->obs_norm2) return set_error(AW_ERR_INVALID_VALUE, "obs_mode 3 takes no obs_norm2");
        a.obs_w = static_cast<const double2 *>(ps->obs_weights);
        a.obs_nw = ps->obs_nw;
        a.obs_partials = ps->obs_partials;
        a.obs_mode = 3;
        a.obs_lf = ps->obs_line_fac;
        a.obs_ef = ps->obs_elem_fac;
    } else if (ps->obs_weights) {
        if (ps->obs_nw < 1 || ps->obs_nw > 8 || !ps->obs_partials)
            return set_error(AW_ERR_INVALID_VALUE, "obs_nw must be 1..8 with an obs_partials buffer");
        if (ps->obs_mode != 0 && ps->obs_mode != 1)
            return set_error(AW_ERR_INVALID_VALUE, "obs_mode must be 0, 1 or 2 (got %d)", ps->obs_mode);
        if (ps->obs_norm2 != 0 && ps->obs_norm2 != 1)
            return set_error(AW_ERR_INVALID_VALUE, "obs_norm2 must be 0 or 1 (got %d)", ps->obs_norm2);
        if (ps->obs_nw + ps->obs_norm2 > 8)
            return set_error(AW_ERR_INVALID_VALUE, "obs_nw + obs_norm2 must be at most 8");
        if (ps->obs_norm2_weights && !ps->obs_norm2)
            return set_error(AW_ERR_INVALID_VALUE, "obs_norm2_weights without obs_norm2");
        a.obs_w = static_cast<const double2 *>(ps->obs_weights);
        a.obs_nw = ps->obs_nw;
        a.obs_partials = ps->obs_partials;
        a.obs_mode = ps->obs_mode;
        a.obs_norm2 = ps->obs_norm2;
        a.obs_n2w = ps->obs_norm2_weights;
    }
    if (ps->in_scaled_out && !ps->dscale)
        return set_error(AW_ERR_INVALID_VALUE, "in_scaled_out needs dscale");
    a.aux_out = ps->in_scaled_out;
    a.has_factor = (ps->line_fac || ps->elem_fac || ps->mask || ps->ctab || ps->chirp != 0.0 ||
                    ps->scale != 1.0 || ps->dscale) ? 1 : 0;
    a.tw = strided ? p->tw_rows : p->tw_cols;
    a.q = sq.q;
    a.q_in = sq.q_in;
    a.q_out = sq.q_out;
    return launch_pass(p->prec, n, strided, a, (hipStream_t)stream);
}

int aw_row_obs_blocks(const aw_plan *p, int *nblk) { return aw_pass_obs_blocks(p, 1, nblk); }

int aw_pass_obs_blocks(const aw_plan *p, int axis, int *nblk) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!nblk) return set_error(AW_ERR_INVALID_VALUE, "NULL nblk");
    if (axis != 0 && axis != 1) return set_error(AW_ERR_INVALID_VALUE, "axis must be 0 or 1");
    const bool strided = axis == 0;
    const int n = strided ? p->rows : p->cols, nl = strided ? p->cols : p->rows;
    const int lpb = obs_lpb(p->prec, n, strided, (long long)nl * p->batch);
    if (lpb < 1 || nl % lpb)
        return set_error(AW_ERR_INVALID_SIZE, "no weighted-sum partials for %s of %d points over %d lines",
                         strided ? "columns" : "rows", n, nl);
    *nblk = nl / lpb;
    return AW_OK;
}

int aw_pass_run(aw_plan *p, void *data, const aw_pass *ps, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    return pass_run_n(p, data, ps, p->batch, stream);
}

// Bytes of state a step sequence keeps in flight per group of items.  Every pass of a
// sequence reads and writes the whole state; once the batch outgrows the 256 MB MALL (the
// memory-side Infinity Cache) each pass streams it from HBM.  Stepping the batch in groups
// of at most this many bytes through all the passes of the sequence keeps a group resident
// in the MALL between passes.  Items are independent and every line computes the same
// arithmetic in any grouping, so results are bit-identical to one launch per pass.
// Measured at 1024^2 x 96 complex64, 129 passes: 24-item groups (192 MiB) 1.45e11
// gp-steps/s against 1.28e11 for the whole batch (row pass 4.64 -> 4.98 TB/s of
// algorithmic traffic, column pass 3.54 -> 4.26 TB/s); 8-item groups are slower again
// (too few lines per launch).
constexpr long long kGroupBytes = 192ll << 20;
constexpr long long kGroupMinLines = 8192;   // keep >= 1024 blocks per launch

// items per group: the whole batch unless it outgrows kGroupBytes (only worth it for several
// steps, and only while a group still fills the GPU); equal groups otherwise
static int step_group(const aw_plan *p, int nsteps, int active) {
    const long long fit = kGroupBytes / ((long long)p->rows * p->cols * p->csize);
    if (nsteps < 2 || fit < 1 || active <= fit || fit * std::min(p->rows, p->cols) < kGroupMinLines)
        return active;
    const long long ngroups = (active + fit - 1) / fit;
    return int((active + ngroups - 1) / ngroups);
}

int aw_step_group(const aw_plan *p, int nsteps, int batch, int *group) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!group) return set_error(AW_ERR_INVALID_VALUE, "NULL group");
    if (nsteps < 1) return set_error(AW_ERR_INVALID_VALUE, "nsteps must be >= 1");
    if (batch < 0 || batch > p->batch)
        return set_error(AW_ERR_INVALID_VALUE, "batch %d outside 0..%d", batch, p->batch);
    *group = step_group(p, nsteps, batch > 0 ? batch : p->batch);
    return AW_OK;
}

int aw_step_sequence(aw_plan *p, void *data, const aw_steps *st, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!data || !st) return set_error(AW_ERR_INVALID_VALUE, "NULL data or steps");
    if (st->nsteps < 1) return set_error(AW_ERR_INVALID_VALUE, "nsteps must be >= 1");
    if (p->rows < 2) return set_error(AW_ERR_INVALID_VALUE, "step sequences need a 2D plan");
    if (st->batch < 0 || st->batch > p->batch)
        return set_error(AW_ERR_INVALID_VALUE, "batch %d outside 0..%d", st->batch, p->batch);
    const long long esz = p->csize;                                   // state elements
    const long long fsz = arith_prec(p->prec) == AW_PREC_F64 ? 16 : 8;  // factor-vector elements
    const char *cl = static_cast<const char *>(st->col_line), *ce = static_cast<const char *>(st->col_elem);
    const char *rl = static_cast<const char *>(st->row_line), *re = static_cast<const char *>(st->row_elem);
    if (!cl || !ce || !rl || !re) return set_error(AW_ERR_INVALID_VALUE, "NULL factor table");
    const int n = st->nsteps;
    const int active = st->batch > 0 ? st->batch : p->batch;   // a slot pool runs its active prefix
    const long long item_elems = (long long)p->rows * p->cols;
    const int group = step_group(p, n, active);
    // column vector v of the sequence and row vector s (constant tables reuse 0/1/2 and 0)
    auto cvec = [&](int v) { return st->const_steps ? (v == 0 ? 0 : (v == n ? 2 : 1)) : v; };
    auto rvec = [&](int s) { return st->const_steps ? 0 : s; };
    const long long rsz = fsz / 2;   // masks are real, in the arithmetic precision
    int rc = AW_OK;
    for (int b0 = 0; rc == AW_OK && b0 < active; b0 += group) {
        const int nb = std::min(group, active - b0);
        // per-item tables advance by their batch strides (0: shared by all items)
        char *d = static_cast<char *>(data) + b0 * item_elems * esz;
        const char *gcl = cl + b0 * st->col_line_bstride * fsz, *gce = ce + b0 * st->col_elem_bstride * fsz;
        const char *grl = rl + b0 * st->row_line_bstride * fsz, *gre = re + b0 * st->row_elem_bstride * fsz;
        auto moff = [&](const void *m) -> const void * {
            return m ? static_cast<const char *>(m) + b0 * st->mask_bstride * rsz : nullptr;
        };
        const void *m_oc = moff(st->mask_open_close), *m_mid = moff(st->mask_mid), *m_row = moff(st->mask_row);
        // complex64 states with FP64 arithmetic: between the opening and the closing pass the
        // state is held in line fixed point (int32 pairs, a power-of-two scale word per line:
        // 2^-30 of the line's largest value instead of 2^-24 of each value), so the sequence
        // rounds to complex64 once, at its end, instead of on every pass
        int *qr = p->q_scales ? p->q_scales + (long long)b0 * p->rows : nullptr;
        int *qc = p->q_scales ? p->q_scales + (long long)p->batch * p->rows + (long long)b0 * p->cols : nullptr;
        const StepQ q_open = qr ? StepQ{2, nullptr, qc} : StepQ(), q_row = qr ? StepQ{3, qc, qr} : StepQ();
        const StepQ q_mid = qr ? StepQ{3, qr, qc} : StepQ(), q_close = qr ? StepQ{1, qr, nullptr} : StepQ();
        aw_pass ps{};
        ps.scale = 1.0;
        ps.mask_bstride = st->mask_bstride;
        // opening column pass: factor c_0 then inverse FFT along axis 0
        ps.axis = 0; ps.dir_a = 0; ps.dir_b = AW_INVERSE;
        ps.line_fac = gcl; ps.line_fac_bstride = st->col_line_bstride;
        ps.elem_fac = gce; ps.elem_fac_bstride = st->col_elem_bstride;
        ps.mask = m_oc;
        ps.dscale = st->dscale ? st->dscale + 2 * b0 : nullptr;
        rc = pass_run_n(p, d, &ps, nb, stream, q_open);
        ps.dscale = nullptr;
        for (int s = 0; rc == AW_OK && s < n; ++s) {
            ps.axis = 1; ps.dir_a = AW_FORWARD; ps.dir_b = AW_INVERSE;
            ps.line_fac = grl + rvec(s) * p->rows * fsz; ps.line_fac_bstride = st->row_line_bstride;
            ps.elem_fac = gre + rvec(s) * p->cols * fsz; ps.elem_fac_bstride = st->row_elem_bstride;
            ps.mask = m_row;
            rc = pass_run_n(p, d, &ps, nb, stream, q_row);
            if (rc != AW_OK) break;
            const bool last = s == n - 1;
            ps.axis = 0; ps.dir_a = AW_FORWARD; ps.dir_b = last ? 0 : AW_INVERSE;
            ps.line_fac = gcl + cvec(s + 1) * p->cols * fsz; ps.line_fac_bstride = st->col_line_bstride;
            ps.elem_fac = gce + cvec(s + 1) * p->rows * fsz; ps.elem_fac_bstride = st->col_elem_bstride;
            ps.mask = last ? m_oc : m_mid;
            rc = pass_run_n(p, d, &ps, nb, stream, last ? q_close : q_mid);
        }
    }
    return rc;
}

int aw_fft_axis(aw_plan *p, void *data, void *out, int axis, int dir, aw_stream stream) {
    if (dir != AW_FORWARD && dir != AW_INVERSE)
        return set_error(AW_ERR_INVALID_VALUE, "direction must be -1 or +1 (got %d)", dir);
    aw_pass ps{};
    ps.axis = axis;
    ps.dir_a = dir;
    ps.scale = 1.0;
    ps.out = out;
    return aw_pass_run(p, data, &ps, stream);
}

// Real transforms run the complex FFT of the full length in a plan-owned workspace (one
// complex array of the plan's size, allocated by the first call, which must not be captured).
static int real_transform(aw_plan *p, const void *in, void *out, int axis, int fwd, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!in || !out) return set_error(AW_ERR_INVALID_VALUE, "NULL data");
    if (axis != 0 && axis != 1) return set_error(AW_ERR_INVALID_VALUE, "axis must be 0 or 1");
    if (axis == 0 && p->rows == 1) return set_error(AW_ERR_INVALID_VALUE, "axis-0 transform on a single-row plan");
    if (p->prec != AW_PREC_F64 && p->prec != AW_PREC_F32)
        return set_error(AW_ERR_INVALID_TYPE, "real transforms need an F64 or F32 plan");
    if (!p->rc_work) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        AW_HIP(hipStreamIsCapturing((hipStream_t)stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return set_error(AW_ERR_INVALID_VALUE, "the first real transform on a plan must not be captured");
        AW_HIP(hipMalloc(&p->rc_work, (size_t)p->csize * p->rows * p->cols * p->batch));
    }
    const int n = axis ? p->cols : p->rows;
    const long long nl = axis ? (long long)p->rows * p->batch : p->batch;
    const hipStream_t st = (hipStream_t)stream;
    int rc = launch_real_complex(p->prec, fwd ? 0 : 2, in, p->rc_work, nl, n, p->cols, axis, st);
    if (rc == AW_OK) rc = aw_fft_axis(p, p->rc_work, nullptr, axis, fwd ? AW_FORWARD : AW_INVERSE, stream);
    if (rc == AW_OK) rc = launch_real_complex(p->prec, fwd ? 1 : 3, p->rc_work, out, nl, n, p->cols, axis, st);
    return rc;
}

int aw_fft_r2c(aw_plan *p, const void *in, void *out, int axis, aw_stream stream) {
    return real_transform(p, in, out, axis, 1, stream);
}

int aw_fft_c2r(aw_plan *p, const void *in, void *out, int axis, aw_stream stream) {
    return real_transform(p, in, out, axis, 0, stream);
}

int aw_make_factors(aw_plan *p, void *out, long long out_stride, int n, int nvec, int mode, const double *A,
                    long long a_stride, double ca, const double *B, long long b_stride, double cb,
                    double c0, const double *mag, long long mag_stride, int mag_pow, double scale,
                    int alt_sign, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!out || n < 1 || nvec < 1) return set_error(AW_ERR_INVALID_VALUE, "bad factor arguments");
    if (mode != 0 && mode != 1) return set_error(AW_ERR_INVALID_VALUE, "mode must be 0 or 1");
    if (mag_pow < 0 || mag_pow > 8) return set_error(AW_ERR_INVALID_VALUE, "mag_pow out of range");
    if (out_stride != 0 && out_stride < n) return set_error(AW_ERR_INVALID_VALUE, "out_stride < n");
    FactorArgs f{};
    f.out = out;
    f.out_stride = out_stride ? out_stride : n;
    f.n = n;
    f.nvec = nvec;
    f.mode = mode;
    f.A = A; f.a_stride = a_stride; f.ca = ca;
    f.B = B; f.b_stride = b_stride; f.cb = cb;
    f.c0 = c0;
    f.mag = mag; f.mag_stride = mag_stride; f.mag_pow = mag_pow;
    f.scale = scale;
    f.alt_sign = alt_sign;
    return launch_make_factors(arith_prec(p->prec), f, (hipStream_t)stream);
}

int aw_reduce_stats(aw_plan *p, const void *data, double *out, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!data || !out) return set_error(AW_ERR_INVALID_VALUE, "NULL data or out");
    return launch_stats(p->prec, data, (long long)p->rows * p->cols, p->rows, p->cols, p->batch,
                        p->partials, p->red_blocks, out, (hipStream_t)stream);
}

int aw_pair_dot(aw_plan *p, const void *a, const void *b, const double *u, const double *v, double *out,
                aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!a || !b || !u || !v || !out) return set_error(AW_ERR_INVALID_VALUE, "NULL pair-dot argument");
    return launch_pair_dot(p->prec, a, b, (long long)p->rows * p->cols, p->cols, u, v, p->batch, p->partials,
                           p->red_blocks, out, (hipStream_t)stream);
}

int aw_cyclic_diag_sums(aw_plan *p, const void *data, double *out, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!data || !out) return set_error(AW_ERR_INVALID_VALUE, "NULL data or out");
    if (p->rows != p->cols || p->rows < 64)
        return set_error(AW_ERR_INVALID_SIZE, "cyclic diagonal sums need a square plan with n >= 64");
    const int n = p->rows;
    const int nblk = n / kDiagRowsPerBlock;
    if (!p->diag_work) {
        AW_HIP(hipMalloc((void **)&p->diag_work, sizeof(double) * 2 * (size_t)n * nblk * p->batch));
    }
    return launch_cyclic_diag(p->prec, data, n, p->batch, p->diag_work, out, (hipStream_t)stream);
}

// Scratch of the one-launch reductions (block partials and per-item ticket counters, which
// start at zero and are reset by each launch's last block), allocated by the first call on a
// plan.  Allocation and the counter reset cannot be captured into a graph, so that first call
// must run eagerly.
static int ensure_finish_scratch(aw_plan *p, aw_stream stream, const char *who) {
    if (p->wsum_partials) return AW_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    AW_HIP(hipStreamIsCapturing((hipStream_t)stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return set_error(AW_ERR_INVALID_VALUE, "the first %s call on a plan must not be captured", who);
    AW_HIP(hipMalloc((void **)&p->wsum_counters, sizeof(unsigned) * (size_t)p->batch));
    AW_HIP(hipMemsetAsync(p->wsum_counters, 0, sizeof(unsigned) * (size_t)p->batch, (hipStream_t)stream));
    AW_HIP(hipMalloc((void **)&p->wsum_partials, sizeof(double) * 8 * (size_t)kWsumMaxBlocks * p->batch));
    return AW_OK;
}

// blocks per item of the weighted sums: 8 elements per thread (4 independent loads in
// flight), at most kWsumMaxBlocks
static int wsum_blocks(const aw_plan *p) {
    const long long nb = (long long)p->rows * p->cols / 2048;
    return int(nb < 1 ? 1 : (nb > kWsumMaxBlocks ? kWsumMaxBlocks : nb));
}

int aw_weighted_partials(aw_plan *p, const void *data, const double *weights, int nw, int norm2,
                         double *partials, int *nblk, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (nw < 1 || nw + (norm2 ? 1 : 0) > 8)
        return set_error(AW_ERR_INVALID_VALUE, "nw must be 1..%d (got %d)", norm2 ? 7 : 8, nw);
    const int nb = wsum_blocks(p);
    if (nblk) *nblk = nb;
    if (!partials) return AW_OK;   // a query of the block count
    if (!data || !weights) return set_error(AW_ERR_INVALID_VALUE, "NULL data or weights");
    return launch_weighted_partials(p->prec, data, (long long)p->rows * p->cols, weights, nw, norm2 ? 1 : 0,
                                    p->batch, partials, nb, (hipStream_t)stream);
}

int aw_weighted_sums(aw_plan *p, const void *data, const double *weights, int nw, int norm2, double *out,
                     double factor, double *dscale, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!data || !weights || !out) return set_error(AW_ERR_INVALID_VALUE, "NULL data, weights or out");
    if (nw < 1 || nw + (norm2 ? 1 : 0) > 8)
        return set_error(AW_ERR_INVALID_VALUE, "nw must be 1..%d (got %d)", norm2 ? 7 : 8, nw);
    const int rc = ensure_finish_scratch(p, stream, "aw_weighted_sums");
    if (rc != AW_OK) return rc;
    const long long n_item = (long long)p->rows * p->cols;
    const int nblk = wsum_blocks(p);
    return launch_weighted_sums(p->prec, data, n_item, weights, nw, norm2 ? 1 : 0, p->batch, p->wsum_partials,
                                nblk, out, factor, dscale, (hipStream_t)stream);
}

int aw_cyclic_diag_obs(aw_plan *p, const void *data, const double *xw, int nx, const double *pw_hat, int np,
                       double *out, long long out_bstride, double factor, double *dscale, aw_stream stream) {
    if (!p) return set_error(AW_ERR_INVALID_PLAN, "NULL plan");
    if (!data || !out || (nx > 0 && !xw) || (np > 0 && !pw_hat))
        return set_error(AW_ERR_INVALID_VALUE, "NULL data, weights or out");
    if (nx < 0 || np < 0 || 1 + nx + np > 8)
        return set_error(AW_ERR_INVALID_VALUE, "1 + nx + np must be at most 8 (nx %d, np %d)", nx, np);
    if (p->rows != p->cols || p->rows < 64)
        return set_error(AW_ERR_INVALID_SIZE, "cyclic diagonal sums need a square plan with n >= 64");
    if (out_bstride < 1 + nx + np) return set_error(AW_ERR_INVALID_VALUE, "out_bstride below 1 + nx + np");
    const int n = p->rows;
    if (!p->diag_work) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        AW_HIP(hipStreamIsCapturing((hipStream_t)stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return set_error(AW_ERR_INVALID_VALUE, "the first aw_cyclic_diag_obs call on a plan must not be captured");
        AW_HIP(hipMalloc((void **)&p->diag_work,
                         sizeof(double) * 2 * (size_t)n * (n / kDiagRowsPerBlock) * p->batch));
    }
    const int rc = ensure_finish_scratch(p, stream, "aw_cyclic_diag_obs");
    if (rc != AW_OK) return rc;
    return launch_cyclic_obs(p->prec, data, n, p->batch, p->diag_work, xw, nx, pw_hat, np, p->wsum_partials,
                             p->wsum_counters, out, out_bstride, factor, dscale, (hipStream_t)stream);
}

int aw_finish_sums(const double *partials, int nblk, int ncomp, int batch, double *out, double factor,
                   double *dscale, aw_stream stream) {
    if (!partials || !out || nblk < 1 || ncomp < 1 || ncomp > 8 || batch < 1)
        return set_error(AW_ERR_INVALID_VALUE, "bad partial-sum arguments");
    return launch_finish_sums(partials, nblk, ncomp, batch, out, factor, dscale, (hipStream_t)stream);
}

int aw_sum_partials(const double *partials, int nblk, int ncomp, int batch, double *out,
                    aw_stream stream) {
    if (!partials || !out || nblk < 1 || ncomp < 1 || batch < 1)
        return set_error(AW_ERR_INVALID_VALUE, "bad partial-sum arguments");
    return launch_sum_partials(partials, nblk, ncomp, batch, out, (hipStream_t)stream);
}

int aw_reduce_vec(int prec, const void *data, long long n, int batch, double *out, aw_stream stream) {
    if (!data || !out || n < 1 || batch < 1) return set_error(AW_ERR_INVALID_VALUE, "bad vector reduction");
    if (prec != AW_PREC_F64 && prec != AW_PREC_F32 && prec != AW_PREC_C64)
        return set_error(AW_ERR_INVALID_TYPE, "bad precision");
    // one block per item in a fixed order, no scratch (nothing is allocated; a plan's
    // aw_reduce_stats splits large items over many blocks)
    return launch_stats(prec, data, n, 1, (int)(n > 0x7fffffff ? 0x7fffffff : n), batch, nullptr, 1, out,
                        (hipStream_t)stream);
}

int aw_scale_from_stats(const double *stats, int stats_stride, int comp, int mode, double factor,
                        double *dscale, int batch, aw_stream stream) {
    if (!stats || !dscale || batch < 1) return set_error(AW_ERR_INVALID_VALUE, "bad scale arguments");
    return launch_scale_from_stats(stats, stats_stride, comp, mode, factor, dscale, batch,
                                   (hipStream_t)stream);
}

int aw_scale(int prec, void *data, long long n, int batch, const double *dscale, double scale,
             aw_stream stream) {
    if (!data) return set_error(AW_ERR_INVALID_VALUE, "NULL data");
    return launch_scale(prec, data, n, batch, dscale, scale, (hipStream_t)stream);
}

int aw_convert(const void *src, int src_prec, void *dst, int dst_prec, long long n, aw_stream stream) {
    if (!src || !dst) return set_error(AW_ERR_INVALID_VALUE, "NULL pointer");
    for (int q : {src_prec, dst_prec})
        if (q != AW_PREC_F64 && q != AW_PREC_F32 && q != AW_PREC_C64)
            return set_error(AW_ERR_INVALID_TYPE, "bad precision %d", q);
    return launch_convert(src, state_prec(src_prec), dst, state_prec(dst_prec), n, (hipStream_t)stream);
}

int aw_project_out(int prec, void *psi, const void *phis, int nphi, int n, aw_stream stream) {
    if (!psi || n < 1 || n > 65536 || nphi < 0 || (nphi > 0 && !phis))
        return set_error(AW_ERR_INVALID_VALUE, "bad projection arguments");
    if (nphi > 4096) return set_error(AW_ERR_INVALID_VALUE, "too many states to project out");
    return launch_project_out(prec, psi, phis, nphi, n, (hipStream_t)stream);
}

int aw_mem_info(size_t *free_bytes, size_t *total_bytes) {
    AW_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return AW_OK;
}

int aw_sync(aw_stream stream) {
    if (stream) AW_HIP(hipStreamSynchronize((hipStream_t)stream));
    else AW_HIP(hipDeviceSynchronize());
    return AW_OK;
}

// ---------------------------------------------------------------------------------------
// hipRTC
// ---------------------------------------------------------------------------------------
static int rtc_compile(const char *source, const char *name, char *log, size_t log_len,
                       std::vector<char> &code) {
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, source, name ? name : "awig_rtc.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) return set_error(AW_ERR_RTC, "hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast"};
    r = hiprtcCompileProgram(prog, 4, opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string lg(ls + 1, '\0');
    if (ls) hiprtcGetProgramLog(prog, &lg[0]);
    if (log && log_len) {
        std::strncpy(log, lg.c_str(), log_len - 1);
        log[log_len - 1] = '\0';
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return set_error(AW_ERR_RTC, "hipRTC compilation failed: %s\n%s", hiprtcGetErrorString(r), lg.c_str());
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return AW_OK;
}

int aw_rtc_check(const char *source, char *log, size_t log_len, size_t *code_size) {
    std::vector<char> code;
    int rc = rtc_compile(source, "awig_check.hip", log, log_len, code);
    if (code_size) *code_size = code.size();
    return rc;
}

int aw_module_compile(aw_module **mod, const char *source, const char *name, char *log, size_t log_len) {
    if (!mod || !source) return set_error(AW_ERR_INVALID_VALUE, "NULL module or source");
    *mod = nullptr;
    std::vector<char> code;
    int rc = rtc_compile(source, name, log, log_len, code);
    if (rc != AW_OK) return rc;
    aw_module *m = new aw_module();
    hipError_t e = hipModuleLoadData(&m->mod, code.data());
    if (e != hipSuccess) {
        delete m;
        return set_error(AW_ERR_HIP, "hipModuleLoadData: %s", hipGetErrorString(e));
    }
    m->name = name ? name : "";
    *mod = m;
    return AW_OK;
}

int aw_module_destroy(aw_module *m) {
    if (!m) return set_error(AW_ERR_INVALID_VALUE, "NULL module");
    if (m->mod) (void)hipModuleUnload(m->mod);
    delete m;
    return AW_OK;
}

int aw_module_launch(aw_module *m, const char *kernel, unsigned gx, unsigned gy, unsigned gz,
                     unsigned bx, unsigned by, unsigned bz, unsigned shmem, aw_stream stream,
                     void **args) {
    if (!m || !kernel) return set_error(AW_ERR_INVALID_VALUE, "NULL module or kernel name");
    hipFunction_t f;
    AW_HIP(hipModuleGetFunction(&f, m->mod, kernel));
    AW_HIP(hipModuleLaunchKernel(f, gx, gy, gz, bx, by, bz, shmem, (hipStream_t)stream, args, nullptr));
    return AW_OK;
}

}  // extern "C"
