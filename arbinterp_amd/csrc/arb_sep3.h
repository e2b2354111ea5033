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

// Non-finite node values.  The reference forms a cell's coefficients as
// alpha = A . F[stencil] (np.dot, ARBInterp.py:599-602, :1351-1354); A has a
// zero in every column and 0 * NaN = 0 * Inf = NaN, so a NaN or +-Inf value
// anywhere in a cell's 4^d stencil makes every coefficient of the cell NaN
// (or +-Inf), and with it every output of the cell; no other cell changes.
// cub4 skips M's zero entries, which would leave some of those coefficients
// finite, so the builds mark such cells and store NaN for all of their
// coefficients.  (For +-Inf values the reference has NaN or +-Inf per
// coefficient; the outputs are NaN either way.)
//
// cell_bad3: per cell of a BX x BY x BZ tile (node tile laid out as in
// sep3_passes), whether its 4 x 4 x 4 stencil holds a non-finite value.
// Block-wide; the callers take this path only for a tile that holds one
// (64 LDS reads per cell).
template <int BX, int BY, int BZ>
__device__ void cell_bad3(const double* nodes, unsigned char* bad) {
    constexpr int NX = BX + 3, NY = BY + 3;
    for (int c = threadIdx.x; c < BX * BY * BZ; c += blockDim.x) {
        const int cx = c % BX, cy = (c / BX) % BY, cz = c / (BX * BY);
        bool b = false;
        for (int z = 0; z < 4; ++z)
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x)
                    b |= !isfinite(nodes[((cz + z) * NY + cy + y) * NX + cx + x]);
        bad[c] = b ? 1 : 0;
    }
}

// LDS doubles the passes need for a tile of bx x by x bz cells: node tile,
// x-pass and y-pass results.
__host__ __device__ inline int sep3_nodes(int bx, int by, int bz) {
    return (bx + 3) * (by + 3) * (bz + 3);
}
__host__ __device__ inline int sep3_t1(int bx, int by, int bz) { return (by + 3) * (bz + 3) * bx * 4; }
__host__ __device__ inline int sep3_t2(int bx, int by, int bz) { return (bz + 3) * by * bx * 16; }

// Block-wide (every thread of a SEP_THREADS-thread workgroup must call it; it
// synchronises between passes):
// nodes[(z * NY + y) * NX + x] (NX = BX+3, NY = BY+3, NZ = BZ+3) holds the
// node tile whose (0,0,0) is the neighbourhood origin of cell (0,0,0).  For
// every cell (cx, cy, cz) of the tile and coefficient i = ex + 4 ey + 16 ez,
// out(cx, cy, cz, i, alpha_i) is called once.
//   CELL_FAST = false: sixteen consecutive threads of the z pass take the
//     sixteen (ex, ey) of one cell, so a row-major store of alpha is
//     coalesced; the pass buffers keep the coefficient index innermost,
//     t1[(yz BX + cx) 4 + ex], t2[((z BY + cy) BX + cx) 16 + 4 ex + ey], and
//     each thread writes its four results as two 16-byte stores;
//   CELL_FAST = true: consecutive threads take consecutive cells
//     (cx + BX (cy + BY cz)); the pass buffers keep the cell x index
//     innermost, t1[((ex NZ + z) NY + y) BX + cx],
//     t2[(((4 ex + ey) NZ + z) BY + cy) BX + cx], so the lanes of a wave
//     touch consecutive doubles.
// Either way each alpha_i comes from the same three cub4 chains on the same
// values, so it does not depend on the tile shape, the layout or the thread
// order.  t1, t2: scratch of sep3_t1 / sep3_t2 doubles (16-byte aligned).
// The caller synchronises before it reuses nodes, t1 or t2.
// sep3_passes4 hands each z-pass thread's four results over at once:
// out4(cx, cy, cz, exy, c) with c[ez] = alpha_{exy + 16 ez}.  The workgroup
// has SEP_THREADS threads (a compile-time stride lets the loops unroll).
constexpr int SEP_THREADS = 256;

template <int BX, int BY, int BZ, bool CELL_FAST, typename OUT4>
__device__ __forceinline__ void sep3_passes4(const double* nodes, double* t1, double* t2,
                                             const OUT4& out4) {
    constexpr int NX = BX + 3, NY = BY + 3, NZ = BZ + 3;
    // x
    for (int t = threadIdx.x; t < BX * NY * NZ; t += SEP_THREADS) {
        const int cx = t % BX, yz = t / BX;
        const double* n = nodes + yz * NX + cx;
        double c[4];
        cub4(n[0], n[1], n[2], n[3], c);
        if constexpr (CELL_FAST) {
#pragma unroll
            for (int e = 0; e < 4; ++e) t1[(e * NY * NZ + yz) * BX + cx] = c[e];
        } else {
            double2* o = reinterpret_cast<double2*>(t1 + (yz * BX + cx) * 4);
            o[0] = make_double2(c[0], c[1]);
            o[1] = make_double2(c[2], c[3]);
        }
    }
    __syncthreads();
    // y
    for (int t = threadIdx.x; t < BX * BY * NZ * 4; t += SEP_THREADS) {
        double c[4];
        if constexpr (CELL_FAST) {
            const int cx = t % BX;
            int r = t / BX;
            const int cy = r % BY;
            r /= BY;
            const int z = r % NZ;
            const int ex = r / NZ;
            const double* n = t1 + ((ex * NZ + z) * NY + cy) * BX + cx;
            cub4(n[0], n[BX], n[2 * BX], n[3 * BX], c);
#pragma unroll
            for (int e = 0; e < 4; ++e) t2[(((ex * 4 + e) * NZ + z) * BY + cy) * BX + cx] = c[e];
        } else {
            const int ex = t & 3;
            int r = t >> 2;
            const int cx = r % BX;
            r /= BX;
            const int cy = r % BY;
            const int z = r / BY;
            const double* n = t1 + ((z * NY + cy) * BX + cx) * 4 + ex;
            constexpr int st = BX * 4;                  // next y
            cub4(n[0], n[st], n[2 * st], n[3 * st], c);
            double2* o = reinterpret_cast<double2*>(t2 + (((z * BY + cy) * BX + cx) * 4 + ex) * 4);
            o[0] = make_double2(c[0], c[1]);
            o[1] = make_double2(c[2], c[3]);
        }
    }
    __syncthreads();
    // z: alpha_i, i = ex + 4 ey + 16 ez
    constexpr int NC = BX * BY * BZ;
    for (int t = threadIdx.x; t < NC * 16; t += SEP_THREADS) {
        const int exy = CELL_FAST ? t / NC : (t & 15);
        const int cell = CELL_FAST ? t % NC : (t >> 4);
        const int cx = cell % BX, cy = (cell / BX) % BY, cz = cell / (BX * BY);
        const int ex = exy & 3, ey = exy >> 2;
        double c[4];
        if constexpr (CELL_FAST) {
            const double* n = t2 + (((ex * 4 + ey) * NZ + cz) * BY + cy) * BX + cx;
            constexpr int st = BY * BX;                 // next z
            cub4(n[0], n[st], n[2 * st], n[3 * st], c);
        } else {
            const double* n = t2 + ((cz * BY + cy) * BX + cx) * 16 + ex * 4 + ey;
            constexpr int st = BY * BX * 16;            // next z
            cub4(n[0], n[st], n[2 * st], n[3 * st], c);
        }
        out4(cx, cy, cz, exy, c);
    }
}

template <int BX, int BY, int BZ, bool CELL_FAST, typename OUT>
__device__ __forceinline__ void sep3_passes(const double* nodes, double* t1, double* t2,
                                            const OUT& out) {
    sep3_passes4<BX, BY, BZ, CELL_FAST>(nodes, t1, t2,
                                        [&](int cx, int cy, int cz, int exy, const double* c) {
#pragma unroll
        for (int e = 0; e < 4; ++e) out(cx, cy, cz, exy + 16 * e, c[e]);
    });
}

}  // namespace arbs3
