// FP64 (complex128) instantiations of the fused line pass.
#include "awig_pass_cfg.h"

namespace awig {
int launch_pass_f64(int n, bool strided, int kind, const PassArgs &a, hipStream_t st) {
    return launch_pass_prec<double>(n, strided, kind, a, st);
}
int row_obs_lpb_f64(int n, long long total_lines) { return row_lpb<double>(n, total_lines); }
}  // namespace awig
