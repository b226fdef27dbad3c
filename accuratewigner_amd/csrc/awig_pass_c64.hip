// Instantiations of the fused line pass for complex64 states with FP64 arithmetic: the FP64
// tiling, twiddles, factors and masks, with the state converted on load and store.
#include "awig_pass_cfg.h"

namespace awig {
int launch_pass_c64(int n, bool strided, int kind, const PassArgs &a, hipStream_t st) {
    if (a.q) return launch_step_q<double, float>(n, strided, kind, a, st);
    return launch_pass_prec<double, float>(n, strided, kind, a, st);
}
}  // namespace awig
