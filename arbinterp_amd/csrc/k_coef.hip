// Coefficient table build: alpha[cell][interp][i] = sum_j A_ij F[nbr_j(cell)].
//
// Replaces calcCoefficients1/2/3 + neighbourInd + allCoeffs
// (ARBInterp.py:523-525, 570-621; 4D :1260-1262, 1322-1381): the reference
// runs one 64x64 (256x256) FP64 GEMV per touched cell from a Python map.
// Here every cell of the grid is built in one launch.
//
// This file holds the reference-form kernels: one thread per (cell, interp,
// coefficient) with A = inv(B).D read from global memory (L2 resident) and
// the 4^d neighbourhood gathered from the canonical node arrays.
#include "arb_internal.h"

namespace {

template <int D, typename TO>
__global__ void k_coef_ref(ArbGeom g, const double* __restrict__ vals, int32_t ninterp,
                           const double* __restrict__ A, TO* __restrict__ table) {
    constexpr int K = (D == 3) ? 64 : 256;
    const int64_t total = g.nc * ninterp * K;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int i = (int)(t % K);
    const int64_t ck = t / K;
    const int k = (int)(ck % ninterp);
    const int64_t c = ck / ninterp;
    // base of the 4^D neighbourhood: node (cx, cy, cz[, ct]) = cell coordinates
    int64_t base = 0, rem = c;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        base += (rem % g.npos[a]) * g.nstride[a];
        rem /= g.npos[a];
    }
    const double* F = vals + (int64_t)k * g.nnodes + base;
    const double* Ai = A + (int64_t)i * K;
    double acc = 0.0;
    for (int j = 0; j < K; ++j) {
        int64_t off = (j & 3) + ((j >> 2) & 3) * g.nstride[1] + ((j >> 4) & 3) * g.nstride[2];
        if (D == 4) off += ((j >> 6) & 3) * g.nstride[3];
        acc = fma(Ai[j], F[off], acc);
    }
    table[t] = (TO)acc;
}

template <typename TO>
__global__ void k_sentinel(TO* table, int64_t nc, int32_t width, TO v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < width) table[nc * width + i] = v;
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t arb_launch_coef_fp64(const ArbGeom& g, const double* vals, int32_t ninterp,
                                const double* A, double* table, hipStream_t s) {
    const int K = g.ndim == 3 ? 64 : 256;
    const int64_t total = g.nc * ninterp * K;
    if (total == 0) return hipSuccess;
    if (g.ndim == 3)
        hipLaunchKernelGGL((k_coef_ref<3, double>), dim3(grid1(total, 256)), dim3(256), 0, s, g,
                           vals, ninterp, A, table);
    else
        hipLaunchKernelGGL((k_coef_ref<4, double>), dim3(grid1(total, 256)), dim3(256), 0, s, g,
                           vals, ninterp, A, table);
    return hipGetLastError();
}

hipError_t arb_launch_coef_mixed(const ArbGeom& g, const double* vals, int32_t ninterp,
                                 const double* A, float* table, hipStream_t s) {
    const int K = g.ndim == 3 ? 64 : 256;
    const int64_t total = g.nc * ninterp * K;
    if (total == 0) return hipSuccess;
    if (g.ndim == 3)
        hipLaunchKernelGGL((k_coef_ref<3, float>), dim3(grid1(total, 256)), dim3(256), 0, s, g,
                           vals, ninterp, A, table);
    else
        hipLaunchKernelGGL((k_coef_ref<4, float>), dim3(grid1(total, 256)), dim3(256), 0, s, g,
                           vals, ninterp, A, table);
    return hipGetLastError();
}

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
