// Tricubic coefficients of a tile of cells by separable passes.
//
// The reference's 3D operator A = inv(B).D (ARBInterp.py:107-175) is the
// Kronecker product M (x) M (x) M of one 4x4 matrix: M maps the four node
// values at offsets -1, 0, 1, 2 along an axis to the cubic coefficients of
// the cell [0, 1] on that axis (Hermite interpolation of the values at 0 and
// 1 with central-difference slopes):
//     M = [[0, 1, 0, 0], [-1/2, 0, 1/2, 0], [1, -5/2, 2, -1/2], [-1/2, 3/2, -3/2, 1/2]]
// (row = power of u, column = node offset + 1), with coefficient index
// ex + 4 ey + 16 ez and neighbourhood index ox + 4 oy + 16 oz.  So
// alpha = A . F is three 4-tap passes, x then y then z, and neighbouring
// cells share the earlier passes: about 625 FMAs per cell instead of 4096.
// Every coefficient is formed by the same fma sequence whatever the tile
// shape, so the table build and the fused query kernel agree bitwise.
#pragma once
#include "arb_internal.h"

namespace arbs3 {

// c[e] = sum_o M[e][o] n_o, as a fixed fma chain per e
__device__ __forceinline__ void cub4(double n0, double n1, double n2, double n3, double* c) {
    c[0] = n1;
    c[1] = fma(0.5, n2, -0.5 * n0);
    c[2] = fma(-0.5, n3, fma(2.0, n2, fma(-2.5, n1, n0)));
    c[3] = fma(0.5, n3, fma(-1.5, n2, fma(1.5, n1, -0.5 * n0)));
}

// LDS doubles the passes need for a tile of bx x by x bz cells: node tile,
// x-pass and y-pass results.
__host__ __device__ inline int sep3_nodes(int bx, int by, int bz) {
    return (bx + 3) * (by + 3) * (bz + 3);
}
__host__ __device__ inline int sep3_t1(int bx, int by, int bz) { return (by + 3) * (bz + 3) * bx * 4; }
__host__ __device__ inline int sep3_t2(int bx, int by, int bz) { return (bz + 3) * by * bx * 16; }

// Block-wide (every thread must call it; it synchronises between passes):
// nodes[(z * NY + y) * NX + x] (NX = bx+3, NY = by+3, NZ = bz+3) holds the
// node tile whose (0,0,0) is the neighbourhood origin of cell (0,0,0).  For
// every cell (cx, cy, cz) of the tile and coefficient i = ex + 4 ey + 16 ez,
// out(cx, cy, cz, i, alpha_i) is called once; sixteen consecutive threads
// handle the sixteen (ex, ey) of one cell and ez, so a row-major store of
// alpha is coalesced.  t1, t2: scratch of sep3_t1 / sep3_t2 doubles.
// The caller synchronises before it reuses nodes, t1 or t2.
template <typename OUT>
__device__ __forceinline__ void sep3_passes(const double* nodes, double* t1, double* t2, int bx,
                                            int by, int bz, const OUT& out) {
    const int NX = bx + 3, NY = by + 3, NZ = bz + 3;
    // x: t1[((z NY + y) bx + cx) 4 + ex]
    for (int t = threadIdx.x; t < bx * NY * NZ; t += blockDim.x) {
        const int cx = t % bx, yz = t / bx;
        const double* n = nodes + yz * NX + cx;
        double c[4];
        cub4(n[0], n[1], n[2], n[3], c);
        double* o = t1 + (yz * bx + cx) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = c[e];
    }
    __syncthreads();
    // y: t2[(((z by + cy) bx + cx) 4 + ex) 4 + ey]
    for (int t = threadIdx.x; t < bx * by * NZ * 4; t += blockDim.x) {
        const int ex = t & 3;
        int r = t >> 2;
        const int cx = r % bx;
        r /= bx;
        const int cy = r % by;
        const int z = r / by;
        const double* n = t1 + ((z * NY + cy) * bx + cx) * 4 + ex;
        const int st = bx * 4;                      // next y
        double c[4];
        cub4(n[0], n[st], n[2 * st], n[3 * st], c);
        double* o = t2 + (((z * by + cy) * bx + cx) * 4 + ex) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = c[e];
    }
    __syncthreads();
    // z: alpha_i, i = ex + 4 ey + 16 ez
    for (int t = threadIdx.x; t < bx * by * bz * 16; t += blockDim.x) {
        const int exy = t & 15;
        int r = t >> 4;
        const int cx = r % bx;
        r /= bx;
        const int cy = r % by;
        const int cz = r / by;
        const int ex = exy & 3, ey = exy >> 2;
        const double* n = t2 + ((cz * by + cy) * bx + cx) * 16 + ex * 4 + ey;
        const int st = by * bx * 16;                // next z
        double c[4];
        cub4(n[0], n[st], n[2 * st], n[3 * st], c);
#pragma unroll
        for (int e = 0; e < 4; ++e) out(cx, cy, cz, exy + 16 * e, c[e]);
    }
}

}  // namespace arbs3
