// FP32 (complex64) instantiations of the fused line pass.
#include "awig_pass_cfg.h"

namespace awig {
int launch_pass_f32(int n, bool strided, int kind, const PassArgs &a, hipStream_t st) {
    return launch_pass_prec<float>(n, strided, kind, a, st);
}
int obs_lpb_f32(int n, bool strided, long long total_lines) { return obs_lpb<float>(n, strided, total_lines); }
}  // namespace awig
