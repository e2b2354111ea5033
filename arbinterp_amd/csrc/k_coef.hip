// Sentinel row nc of the coefficient table, the reference's alpha*[:, nc]:
// zero for a scalar field (ARBInterp.py:31; 4D :650) and for an explicit
// mode='vector' (whose NaN column nc+1 no query reaches, :42-45; 4D :661-664),
// NaN in the other modes (:57, :70, :85, :99; 4D :676, :689, :704, :718).
// Rows that reduce to the sentinel cell are evaluated from it.  The
// coefficient builds themselves are in k_coef_mfma.hip.
#include "arb_internal.h"

namespace {

template <typename TO>
__global__ void k_sentinel(TO* table, int64_t nc, int32_t width, TO v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < width) table[nc * width + i] = v;
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t arb_launch_sentinel(void* table, int32_t f32, int64_t nc, int32_t ninterp, int32_t K,
                               int32_t nan, hipStream_t s) {
    const int width = ninterp * K;
    if (f32)
        hipLaunchKernelGGL((k_sentinel<float>), dim3(grid1(width, 256)), dim3(256), 0, s,
                           (float*)table, nc, width, nan ? (float)NAN : 0.0f);
    else
        hipLaunchKernelGGL((k_sentinel<double>), dim3(grid1(width, 256)), dim3(256), 0, s,
                           (double*)table, nc, width, nan ? (double)NAN : 0.0);
    return hipGetLastError();
}
