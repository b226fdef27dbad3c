// aw_weighted_sums_noalloc: the weighted sums of aw_weighted_sums without a plan and without
// any allocation.  The block partials live in a file-scope device array of fixed size; each
// GPU holds its own copy at its own address, looked up once per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "awig_internal.h"

namespace awig {

// room for the block partials: 512 blocks x 8 components x 256 items (8 MiB per device)
constexpr long long kScratchDoubles = 512ll * 8 * 256;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static std::once_flag g_scratch_once[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];

// device address of `scratch` on the current device (hipGetSymbolAddress, once per device,
// with that device current; the host-side name of the array is not a device address)
static int scratch_ptr(double **ptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(AW_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_error(AW_ERR_INVALID_VALUE, "device %d above %d", dev, kMaxDevices);
    std::call_once(g_scratch_once[dev], [dev] {
        void *p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    });
    if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
        return set_error(AW_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(g_scratch_err[dev]));
    *ptr = g_scratch_ptr[dev];
    return AW_OK;
}

}  // namespace awig

using namespace awig;

extern "C" int aw_weighted_sums_noalloc(int prec, int rows, int cols, int batch, const void *data,
                                        const double *weights, int nw, int norm2, double *out, double factor,
                                        double *dscale, aw_stream stream) {
    if (prec != AW_PREC_F64 && prec != AW_PREC_F32 && prec != AW_PREC_C64)
        return set_error(AW_ERR_INVALID_TYPE, "unsupported precision %d", prec);
    if (rows < 1 || cols < 1 || batch < 1) return set_error(AW_ERR_INVALID_SIZE, "bad shape %d x %d x %d", batch, rows, cols);
    if (nw < 1 || nw + (norm2 ? 1 : 0) > 8)
        return set_error(AW_ERR_INVALID_VALUE, "nw must be 1..%d (got %d)", norm2 ? 7 : 8, nw);
    const long long n_item = (long long)rows * cols;
    // the block count of aw_weighted_sums (8 elements per thread, at most kWsumMaxBlocks), so
    // that both entry points add in the same order
    long long nb = n_item / 2048;
    nb = nb < 1 ? 1 : (nb > kWsumMaxBlocks ? kWsumMaxBlocks : nb);
    const int nc = nw + (norm2 ? 1 : 0);
    if (nb * nc * (long long)batch > kScratchDoubles)
        return set_error(AW_ERR_INVALID_SIZE,
                         "weighted sums of %d items x %d components need %lld scratch doubles, above the %lld "
                         "of aw_weighted_sums_noalloc (use aw_weighted_sums)",
                         batch, nc, nb * nc * (long long)batch, kScratchDoubles);
    if (!data || !weights || !out) return set_error(AW_ERR_INVALID_VALUE, "NULL data, weights or out");
    double *partials = nullptr;
    const int rc = scratch_ptr(&partials);
    if (rc != AW_OK) return rc;
    const int sp = prec == AW_PREC_F64 ? AW_PREC_F64 : AW_PREC_F32;   // the states' precision
    return launch_weighted_sums(sp, data, n_item, weights, nw, norm2 ? 1 : 0, batch, partials, (int)nb, out, factor,
                                dscale, (hipStream_t)stream);
}
