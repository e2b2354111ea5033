// Coefficient table: the build (calcCoefficients*, allCoeffs:
// ARBInterp.py:523-525, 570-603; 4D :1260-1262, 1322-1355) and the sentinel
// row.
//
// alpha = A . F for every cell, in FP64, stored as FP64 or rounded once to
// FP32 (mixed precision), by separable passes:
//   3D: A = M (x) M (x) M (arb_sep3.h), about 625 FMAs per cell;
//   4D: A = M (x) M (x) M (x) M + a rank-16 correction from the reference's
//       row shift (arb_sep4.h).
//
// Sentinel row nc, the reference's alpha*[:, nc]: zero for a scalar field
// (ARBInterp.py:31; 4D :650) and for an explicit mode='vector' (whose NaN
// column nc+1 no query reaches, :42-45; 4D :661-664), NaN in the other modes
// (:57, :70, :85, :99; 4D :676, :689, :704, :718).  Rows that reduce to the
// sentinel cell are evaluated from it.
#include "arb_internal.h"
#include "arb_sep4.h"

namespace {


template <typename TO>
__global__ void k_sentinel(TO* table, int64_t nc, int32_t width, TO v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < width) table[nc * width + i] = v;
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// 3D, both precisions: alpha = (M (x) M (x) M) . F by separable passes
// (arbs3::sep3_passes) over an 8 x 4 x 4 tile of cells per workgroup, in
// FP64, stored as TO (FP64 table, or FP32 rounded once for mixed).  The tile
// reads 11 x 7 x 7 nodes per interpolant for 128 cells (4.2 per cell); the
// loads are unconditional (index 0 for nodes past the grid edge, which only
// cells outside the grid use) so that they issue back to back.
constexpr int SX = 8, SY = 4, SZ3 = 4;

template <typename TO>
__global__ void __launch_bounds__(256)
k_coef_sep3(ArbGeom g, const double* __restrict__ vals, int32_t ninterp, TO* __restrict__ table) {
    constexpr int NX = SX + 3, NY = SY + 3, NZ = SZ3 + 3, NT = NX * NY * NZ;
    constexpr int PER = (NT + 255) / 256;
    __shared__ double nodes[NT];
    __shared__ __align__(16) double t1[NY * NZ * SX * 4];
    __shared__ __align__(16) double t2[NZ * SY * SX * 16];
    __shared__ unsigned char cbad[SX * SY * SZ3];       // cells with a non-finite stencil
    const int64_t cx0 = (int64_t)blockIdx.x * SX, cy0 = (int64_t)blockIdx.y * SY;
    const int64_t cz0 = (int64_t)blockIdx.z * SZ3;
    for (int k = 0; k < ninterp; ++k) {
        const double* F = vals + (int64_t)k * g.nnodes;
        double v[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = threadIdx.x + 256 * j;
            const int x = i % NX, y = (i / NX) % NY, z = i / (NX * NY);
            const int64_t nx = cx0 + x, ny = cy0 + y, nz = cz0 + z;
            const bool in = i < NT && nx < g.n[0] && ny < g.n[1] && nz < g.n[2];
            const double t = F[in ? nx + g.nstride[1] * ny + g.nstride[2] * nz : 0];
            v[j] = in ? t : 0.0;
        }
        __syncthreads();                            // the previous interpolant is done
        int nonfinite = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = threadIdx.x + 256 * j;
            if (i < NT) {
                nodes[i] = v[j];
                nonfinite |= !isfinite(v[j]);
            }
        }
        const bool anybad = __syncthreads_or(nonfinite) != 0;
        if (anybad) {
            arbs3::cell_bad3<SX, SY, SZ3>(nodes, cbad);
            __syncthreads();
        }
        arbs3::sep3_passes<SX, SY, SZ3, false>(nodes, t1, t2,
                                              [&](int cx, int cy, int cz, int i, double a) {
            const int64_t x = cx0 + cx, y = cy0 + cy, z = cz0 + cz;
            if (x < g.npos[0] && y < g.npos[1] && z < g.npos[2]) {
                const int64_t cell = x + g.npos[0] * (y + g.npos[1] * z);
                if (anybad && cbad[cx + SX * (cy + SY * cz)]) a = NAN;
                table[(cell * ninterp + k) * 64 + i] = (TO)a;
            }
        });
    }
}

// 4D, both precisions: alpha = the separable M passes + the row-shift
// correction (arb_sep4.h), in FP64, stored as TO.  Tile: 8 x 4 x SZ4 cells
// in (x, y, z) at one t.  Per interpolant: the 11 x 7 x (SZ4+3) x 4 node
// tile, the corner differences G and the per-cell c_k from it, then the node
// tile is replaced in place by its four powers of t (cub4 along t; each
// thread rewrites the nodes it read), and each power's 3D slice goes through
// the separable passes, whose z-pass threads add the correction for their
// (ex, ey) and store four coefficients.
constexpr int SZ4 = 2;

template <typename TO>
__global__ void __launch_bounds__(256)
k_coef_sep4(ArbGeom g, const double* __restrict__ vals, int32_t ninterp, TO* __restrict__ table) {
    constexpr int NX = SX + 3, NY = SY + 3, NZ = SZ4 + 3;
    constexpr int GX = SX + 1, GY = SY + 1, GZ = SZ4 + 1;     // corner nodes, GT = 2
    constexpr int NC = SX * SY * SZ4;
    __shared__ __align__(16) double n4[4][NZ][NY][NX];
    __shared__ __align__(16) double t1[NY * NZ * SX * 4];
    __shared__ __align__(16) double t2[NZ * SY * SX * 16];
    __shared__ double gc[2][GZ][GY][GX];
    __shared__ __align__(16) double cc[NC][16];
    __shared__ unsigned char cbad[NC];                  // cells with a non-finite stencil
    // 32-bit node offsets and cell indices: the launcher admits grids of
    // fewer than 2^31 nodes (a table of such a grid holds 2^39 bytes)
    const int32_t n0 = (int32_t)g.n[0], n1 = (int32_t)g.n[1], n2 = (int32_t)g.n[2];
    const int32_t s1 = (int32_t)g.nstride[1], s2 = (int32_t)g.nstride[2], s3 = (int32_t)g.nstride[3];
    const int32_t p0 = (int32_t)g.npos[0], p1 = (int32_t)g.npos[1], p2 = (int32_t)g.npos[2];
    const int32_t nzt = (p2 + SZ4 - 1) / SZ4;
    const int32_t cx0 = blockIdx.x * SX, cy0 = blockIdx.y * SY;
    const int32_t cz0 = (blockIdx.z % nzt) * SZ4, ct = blockIdx.z / nzt;
    for (int k = 0; k < ninterp; ++k) {
        const double* F = vals + (int64_t)k * g.nnodes + (cx0 + s1 * cy0 + s2 * cz0 + s3 * ct);
        __syncthreads();                            // the previous interpolant is done
        // node tile; nodes past the grid read 0 (only cells outside it use them)
        int nonfinite = 0;
        for (int i = threadIdx.x; i < 4 * NZ * NY * NX; i += blockDim.x) {
            const int x = i % NX, y = (i / NX) % NY, z = (i / (NX * NY)) % NZ;
            const int t = i / (NX * NY * NZ);
            double v = 0.0;
            if (cx0 + x < n0 && cy0 + y < n1 && cz0 + z < n2) v = F[x + s1 * y + s2 * z + s3 * t];
            (&n4[0][0][0][0])[i] = v;
            nonfinite |= !isfinite(v);
        }
        const bool anybad = __syncthreads_or(nonfinite) != 0;
        if (anybad) {
            // per cell: a non-finite value among its 4^4 stencil nodes
            for (int c = threadIdx.x; c < NC; c += blockDim.x) {
                const int cx = c % SX, cy = (c / SX) % SY, cz = c / (SX * SY);
                bool b = false;
                for (int t = 0; t < 4; ++t)
                    for (int z = 0; z < 4; ++z)
                        for (int y = 0; y < 4; ++y)
                            for (int x = 0; x < 4; ++x) b |= !isfinite(n4[t][cz + z][cy + y][cx + x]);
                cbad[c] = b ? 1 : 0;
            }
        }
        for (int i = threadIdx.x; i < 2 * GZ * GY * GX; i += blockDim.x) {
            const int gx = i % GX, gy = (i / GX) % GY, gz = (i / (GX * GY)) % GZ;
            const int gt = i / (GX * GY * GZ);
            (&gc[0][0][0][0])[i] =
                arbs4::corner_diff<NX, NX * NY, NX * NY * NZ>(&n4[gt + 1][gz + 1][gy + 1][gx + 1]);
        }
        __syncthreads();                            // n4 read, gc written
        // powers of t in place, and c_k = G(corner k-1) - G(corner k) per cell
        for (int i = threadIdx.x; i < NZ * NY * NX; i += blockDim.x) {
            double* p = &n4[0][0][0][0] + i;
            constexpr int st = NZ * NY * NX;
            double c[4];
            arbs3::cub4(p[0], p[st], p[2 * st], p[3 * st], c);
#pragma unroll
            for (int e = 0; e < 4; ++e) p[e * st] = c[e];
        }
        for (int i = threadIdx.x; i < NC * 16; i += blockDim.x) {
            const int kk = i & 15, cell = i >> 4;
            const int cx = cell % SX, cy = (cell / SX) % SY, cz = cell / (SX * SY);
            auto G = [&](int c) {
                return gc[(c >> 3) & 1][cz + ((c >> 2) & 1)][cy + ((c >> 1) & 1)][cx + (c & 1)];
            };
            cc[cell][kk] = (kk ? G(kk - 1) : 0.0) - G(kk);
        }
        __syncthreads();
        for (int et = 0; et < 4; ++et) {
            arbs3::sep3_passes4<SX, SY, SZ4, false>(
                &n4[et][0][0][0], t1, t2, [&](int cx, int cy, int cz, int exy, const double* c0) {
                    const int cell = cx + SX * (cy + SY * cz);
                    double c[4] = {c0[0], c0[1], c0[2], c0[3]};
                    arbs4::shift_rows(cc[cell], exy & 3, exy >> 2, et, c);
                    const int32_t x = cx0 + cx, y = cy0 + cy, z = cz0 + cz;
                    if (x < p0 && y < p1 && z < p2) {
                        const uint32_t cl = (uint32_t)(x + p0 * (y + p1 * (z + p2 * ct)));
                        const bool bad = anybad && cbad[cell];
                        TO* o = table + ((uint64_t)cl * (uint32_t)ninterp + k) * 256 + exy + 64 * et;
#pragma unroll
                        for (int ez = 0; ez < 4; ++ez) o[16 * ez] = bad ? (TO)NAN : (TO)c[ez];
                    }
                });
        }
    }
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// alpha for every cell into table ([nc+1][ninterp][K]; FP64 for ARB_PREC_FP64,
// FP32 for ARB_PREC_MIXED).
hipError_t arb_launch_coef(const ArbGeom& g, int32_t prec, const double* vals, int32_t ninterp,
                           void* table, hipStream_t s) {
    if (g.nc == 0) return hipSuccess;
    const dim3 block(256);
    if (g.ndim == 3) {
        const dim3 grid(cdiv(g.npos[0], SX), cdiv(g.npos[1], SY), cdiv(g.npos[2], SZ3));
        if (prec == ARB_PREC_FP64)
            hipLaunchKernelGGL(k_coef_sep3<double>, grid, block, 0, s, g, vals, ninterp, (double*)table);
        else
            hipLaunchKernelGGL(k_coef_sep3<float>, grid, block, 0, s, g, vals, ninterp, (float*)table);
    } else {
        if (g.nnodes > INT32_MAX) return hipErrorInvalidValue;   // 32-bit offsets
        const dim3 grid(cdiv(g.npos[0], SX), cdiv(g.npos[1], SY),
                        cdiv(g.npos[2], SZ4) * (unsigned)g.npos[3]);
        if (prec == ARB_PREC_FP64)
            hipLaunchKernelGGL(k_coef_sep4<double>, grid, block, 0, s, g, vals, ninterp, (double*)table);
        else
            hipLaunchKernelGGL(k_coef_sep4<float>, grid, block, 0, s, g, vals, ninterp, (float*)table);
    }
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
