// FP64 (complex128) instantiations of the fused line pass.
#include "awig_pass_cfg.h"

namespace awig {
int launch_pass_f64(int n, bool strided, int kind, const PassArgs &a, hipStream_t st) {
    return launch_pass_prec<double>(n, strided, kind, a, st);
}
int obs_lpb_f64(int n, bool strided, long long total_lines) { return obs_lpb<double>(n, strided, total_lines); }
}  // namespace awig
