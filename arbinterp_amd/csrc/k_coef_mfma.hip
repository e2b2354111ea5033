// Coefficient table builds (calcCoefficients*, allCoeffs: ARBInterp.py:523-525,
// 570-603; 4D :1260-1262, 1322-1355): alpha = A . F for every cell.
//
// 3D, both precisions (k_coef_sep3): A = M (x) M (x) M, three separable 4-tap
// passes in FP64 over an LDS node tile (arb_sep3.h); stored as FP64, or
// rounded once to FP32 for mixed precision.
//
// 4D, FP64 (k_coef_f64mfma4): dense alpha = A . F on v_mfma_f64_16x16x4f64.
//
// 4D, mixed (k_coef_bf16x3<4>): the reference's row shift in the d4/dxdydzdt
// block (ARBInterp.py:860) breaks the Kronecker form, so alpha = inv(B) . b
// with b the FD b-vector (16 blocks x 16 corners, ARBInterp.py:738-757),
// formed in FP64 from an LDS tile of node values, one dense contraction per
// tile of 16 cells, (256 x 256) x (256 x 16), on v_mfma_f32_16x16x32_bf16
// with FP32 accumulation, made exact:
//   * the cell's base value f(corner 0) is subtracted from its 16 corner
//     values (inv(B) maps constant data to e_0, so alpha_0 takes it back);
//   * the residual b' is scaled by a power of two 2^-E so |b'| < 1 and split
//     on a common fixed-point grid into four 8-bit integer parts,
//     b' 2^-E = p0/2^8 + p1/2^16 + p2/2^24 + p3/2^32 (+ < 2^-33),
//     |p_k| <= 256, exact in BF16;
//   * inv(B) is integer (|.| <= 81), exact in BF16.
// Every product is then an integer below 2^15 and every partial sum an
// integer below 2^23 (256 * 81 * 256), so each part's MFMA chain is exact in
// FP32 whatever its internal order.  The four accumulators are combined in
// FP64 (exact), scaled back by 2^E and rounded once to FP32: the stored
// alpha is the FP64 alpha of the quantised b' (quantisation < 2^-33 of the
// cell's largest residual) correctly rounded to FP32.
// Work split: one workgroup = 4 waves = a 16(x) x 4(y) tile of cells at one
// (z, t); wave w owns cell row y0+w, lane l the cell x0 + (l & 15).  The node
// tile (19 x 7 x 4 x 4) is loaded once per interpolant into LDS with
// row-contiguous reads.
#include "arb_internal.h"
#include "arb_sep3.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int TXC = 16, TYC = 4;            // cells per tile
constexpr int TX = TXC + 3, TY = TYC + 3;   // nodes per tile

// b-vector block -> axis bitmask (bit a = derivative along axis a), in the
// reference's block order (ARBInterp.py:118-125, :738-757).
__constant__ unsigned char kBlockMask3[8] = {0, 1, 2, 4, 3, 5, 6, 7};
__constant__ unsigned char kBlockMask4[16] = {0, 1, 2, 4, 8, 3, 5, 9, 6, 10, 12, 7, 11, 13, 14, 15};

template <int D>
struct TileGeom {
    static constexpr int TZ = 4;
    static constexpr int TT = (D == 4) ? 4 : 1;
    static constexpr int NODES = TX * TY * TZ * TT;
};

// Finite-difference entry e of the b-vector (block e / nco, corner e % nco)
// for the cell whose 4^D neighbourhood starts at tile offset (lx, ly, 0, 0).
template <int D>
__device__ __forceinline__ double fd_entry(const double* __restrict__ tile, int lx, int ly, int e) {
    constexpr int nco = 1 << D;
    const int blk = e / nco;
    int corner = e % nco;
    const int bm = (D == 3) ? kBlockMask3[blk] : kBlockMask4[blk];
    if (D == 4 && blk == 15) {
        // ARBInterp.py:860: row 240 is zero, row 240+k uses corner k-1
        if (corner == 0) return 0.0;
        corner -= 1;
    }
    const int nb = __builtin_popcount(bm);
    const double scale = (nb == 0) ? 1.0 : (nb == 1) ? 0.5 : (nb == 2) ? 0.25 : (nb == 3) ? 0.125 : 0.0625;
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < (1 << D); ++t) {
        if (t & ~bm) continue;
        int o[4] = {0, 0, 0, 0};
        int neg = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const int c = 1 + ((corner >> a) & 1);
            if (bm >> a & 1) {
                const int s = (t >> a & 1) ? 1 : -1;
                neg += (s < 0);
                o[a] = c + s;
            } else {
                o[a] = c;
            }
        }
        const int idx = (((o[3] * TileGeom<D>::TZ + o[2]) * TY + (ly + o[1])) * TX + (lx + o[0]));
        const double v = tile[idx];
        acc += (neg & 1) ? -v : v;
    }
    return acc * scale;
}

template <int D>
__global__ void __launch_bounds__(256)
k_coef_bf16x3(ArbGeom g, const double* __restrict__ vals, int32_t ninterp,
              const unsigned short* __restrict__ Ab, float* __restrict__ table) {
    constexpr int K = (D == 3) ? 64 : 256;
    constexpr int KS = K / 32;      // MFMA k-steps
    constexpr int MT = K / 16;      // MFMA row tiles
    constexpr int nco = 1 << D;
    __shared__ double tile[TileGeom<D>::NODES];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int grp = lane >> 4;      // 16-lane group: k-slice of the operands
    const int lx = lane & 15;
    const int64_t cx0 = (int64_t)blockIdx.x * TXC;
    const int64_t cy0 = (int64_t)blockIdx.y * TYC;
    int64_t cz = blockIdx.z, ct = 0;
    if (D == 4) { ct = cz / g.npos[2]; cz = cz % g.npos[2]; }
    const int64_t cx = cx0 + lx, cy = cy0 + w;
    const bool valid = (cx < g.npos[0]) && (cy < g.npos[1]);
    const int64_t cell = cx + g.npos[0] * (cy + g.npos[1] * (cz + (D == 4 ? g.npos[2] * ct : 0)));

    for (int k = 0; k < ninterp; ++k) {
        const double* F = vals + (int64_t)k * g.nnodes;
        __syncthreads();
        for (int i = threadIdx.x; i < TileGeom<D>::NODES; i += 256) {
            const int x = i % TX;
            const int y = (i / TX) % TY;
            const int z = (i / (TX * TY)) % TileGeom<D>::TZ;
            const int t = i / (TX * TY * TileGeom<D>::TZ);
            const int64_t nx = cx0 + x, ny = cy0 + y, nz = cz + z, nt = ct + t;
            double v = 0.0;
            if (nx < g.n[0] && ny < g.n[1])
                v = F[nx + g.nstride[1] * ny + g.nstride[2] * nz + (D == 4 ? g.nstride[3] * nt : 0)];
            tile[i] = v;
        }
        __syncthreads();

        // b-vector slice of this lane: entries ks*32 + 8*grp + j, j = 0..7
        const int lxc = valid ? lx : 0;
        const int lyc = w;
        const double f0 = fd_entry<D>(tile, lxc, lyc, 0);
        double bv[KS][8];
        double mx = 0.0;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int e = ks * 32 + 8 * grp + j;
                double b = fd_entry<D>(tile, lxc, lyc, e);
                if (e < nco) b -= f0;                   // base value carried by alpha_0
                bv[ks][j] = b;
                mx = fmax(mx, fabs(b));
                if (b != b) mx = b;                      // NaN data: NaN coefficients
            }
        }
        // the cell's entries live in lanes lx, lx+16, lx+32, lx+48
        double m1 = __shfl_xor(mx, 16);
        mx = (m1 != m1) ? m1 : fmax(mx, m1);
        m1 = __shfl_xor(mx, 32);
        mx = (m1 != m1) ? m1 : fmax(mx, m1);
        int E = 0;
        if (mx > 0.0 && mx == mx && mx < INFINITY) (void)frexp(mx, &E);   // mx < 2^E
        us8 bp[4][KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                double x = ldexp(bv[ks][j], 8 - E);      // |x| < 2^8
#pragma unroll
                for (int pp = 0; pp < 4; ++pp) {
                    const double pk = rint(x);           // integer, |pk| <= 256
                    x = (x - pk) * 256.0;                // exact
                    bp[pp][ks][j] = (unsigned short)(__float_as_uint((float)pk) >> 16);
                }
            }
        }
        float* out = table + (cell * ninterp + k) * K;
#pragma unroll 1
        for (int mt = 0; mt < MT; ++mt) {
            f4 acc[4];
#pragma unroll
            for (int pp = 0; pp < 4; ++pp) {
                acc[pp] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const us8 a = *reinterpret_cast<const us8*>(Ab + (((mt * KS + ks) * 64 + lane) * 8));
                    acc[pp] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                                      __builtin_bit_cast(bf8, bp[pp][ks]),
                                                                      acc[pp], 0, 0, 0);
                }
            }
            // lane holds alpha[mt*16 + 4*grp + r] of cell lx
            f4 res;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double v = (double)acc[0][r] * 0x1p-8 + (double)acc[1][r] * 0x1p-16 +
                           (double)acc[2][r] * 0x1p-24 + (double)acc[3][r] * 0x1p-32;
                v = ldexp(v, E);
                if (mt == 0 && grp == 0 && r == 0) v += f0;
                res[r] = (float)v;
            }
            if (valid) *reinterpret_cast<f4*>(out + mt * 16 + 4 * grp) = res;
        }
    }
}

// 3D, both precisions: alpha = (M (x) M (x) M) . F by separable passes
// (arbs3::sep3_passes) over an 8 x 4 x 1 tile of cells per workgroup, in
// FP64, stored as TO (FP64 table, or FP32 rounded once for mixed).
constexpr int SX = 8, SY = 4;

template <typename TO>
__global__ void __launch_bounds__(256)
k_coef_sep3(ArbGeom g, const double* __restrict__ vals, int32_t ninterp, TO* __restrict__ table) {
    constexpr int NX = SX + 3, NY = SY + 3, NZ = 4;
    __shared__ double nodes[NX * NY * NZ];
    __shared__ double t1[NY * NZ * SX * 4];
    __shared__ double t2[NZ * SY * SX * 16];
    const int64_t cx0 = (int64_t)blockIdx.x * SX, cy0 = (int64_t)blockIdx.y * SY;
    const int64_t cz = blockIdx.z;
    for (int k = 0; k < ninterp; ++k) {
        const double* F = vals + (int64_t)k * g.nnodes;
        __syncthreads();                            // the previous interpolant is done
        for (int i = threadIdx.x; i < NX * NY * NZ; i += blockDim.x) {
            const int x = i % NX, y = (i / NX) % NY, z = i / (NX * NY);
            const int64_t nx = cx0 + x, ny = cy0 + y, nz = cz + z;
            double v = 0.0;
            if (nx < g.n[0] && ny < g.n[1]) v = F[nx + g.nstride[1] * ny + g.nstride[2] * nz];
            nodes[i] = v;
        }
        __syncthreads();
        arbs3::sep3_passes(nodes, t1, t2, SX, SY, 1, [&](int cx, int cy, int, int i, double a) {
            const int64_t x = cx0 + cx, y = cy0 + cy;
            if (x < g.npos[0] && y < g.npos[1]) {
                const int64_t cell = x + g.npos[0] * (y + g.npos[1] * cz);
                table[(cell * ninterp + k) * 64 + i] = (TO)a;
            }
        });
    }
}

// Plain FP64, 4D: alpha = A . F, K = 256.  One workgroup = 16 cells along x
// (one row at fixed y, z, t); its 4 waves split the 16 row tiles (4 each)
// and share the 19 x 4 x 4 x 4 node tile in LDS.  A operands stream from
// L2 (512 KB, shared by every workgroup).
__global__ void __launch_bounds__(256)
k_coef_f64mfma4(ArbGeom g, const double* __restrict__ vals, int32_t ninterp,
                const double* __restrict__ A, double* __restrict__ table) {
    constexpr int K = 256, TXN = TXC + 3;
    constexpr int NODES = TXN * 4 * 4 * 4;
    __shared__ double tile[NODES];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int lx = lane & 15, kq = lane >> 4;
    const int64_t cx0 = (int64_t)blockIdx.x * TXC;
    const int64_t cy = blockIdx.y;
    const int64_t cz = blockIdx.z % g.npos[2], ct = blockIdx.z / g.npos[2];
    const bool colvalid = (cx0 + lx) < g.npos[0];
    const int lxc = colvalid ? lx : 0;
    const int64_t cell = (cx0 + lx) + g.npos[0] * (cy + g.npos[1] * (cz + g.npos[2] * ct));
    for (int k = 0; k < ninterp; ++k) {
        const double* F = vals + (int64_t)k * g.nnodes;
        __syncthreads();
        for (int i = threadIdx.x; i < NODES; i += 256) {
            const int x = i % TXN;
            const int y = (i / TXN) & 3;
            const int z = (i / (TXN * 4)) & 3;
            const int tt = i / (TXN * 16);
            const int64_t nx = cx0 + x;
            double v = 0.0;
            if (nx < g.n[0])
                v = F[nx + g.nstride[1] * (cy + y) + g.nstride[2] * (cz + z) + g.nstride[3] * (ct + tt)];
            tile[i] = v;
        }
        __syncthreads();
        double* out = table + (cell * ninterp + k) * K;
        for (int mq = 0; mq < 4; ++mq) {
            const int mt = w * 4 + mq;
            d4 acc = {0.0, 0.0, 0.0, 0.0};
            const double* Arow = A + (int64_t)(mt * 16 + lx) * K + kq;
#pragma unroll 8
            for (int ks = 0; ks < 64; ++ks) {
                const int j = ks * 4 + kq;
                const int ox = j & 3, oy = (j >> 2) & 3, oz = (j >> 4) & 3, ot = j >> 6;
                const double b = tile[((ot * 4 + oz) * 4 + oy) * TXN + (lxc + ox)];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Arow[ks * 4], b, acc, 0, 0, 0);
            }
            if (colvalid) {
#pragma unroll
                for (int r = 0; r < 4; ++r) out[mt * 16 + kq + 4 * r] = acc[r];
            }
        }
    }
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

hipError_t arb_launch_coef_f64mfma(const ArbGeom& g, const double* vals, int32_t ninterp,
                                   const double* A, double* table, hipStream_t s) {
    if (g.nc == 0) return hipSuccess;
    if (g.ndim == 3) {
        dim3 grid(cdiv(g.npos[0], SX), cdiv(g.npos[1], SY), (unsigned)g.npos[2]), block(256);
        hipLaunchKernelGGL(k_coef_sep3<double>, grid, block, 0, s, g, vals, ninterp, table);
    } else {
        dim3 grid(cdiv(g.npos[0], TXC), (unsigned)g.npos[1], (unsigned)(g.npos[2] * g.npos[3]));
        hipLaunchKernelGGL(k_coef_f64mfma4, grid, dim3(256), 0, s, g, vals, ninterp, A, table);
    }
    return hipGetLastError();
}

// 3D: separable passes in FP64, rounded to FP32.  4D: Ab = inv(B) as BF16 in
// operand order [mt][ks][lane][8] (host-prepared), k_coef_bf16x3<4>.
hipError_t arb_launch_coef_bf16x3(const ArbGeom& g, const double* vals, int32_t ninterp,
                                  const unsigned short* Ab, float* table, hipStream_t s) {
    if (g.nc == 0) return hipSuccess;
    dim3 block(256);
    if (g.ndim == 3) {
        dim3 grid(cdiv(g.npos[0], SX), cdiv(g.npos[1], SY), (unsigned)g.npos[2]);
        hipLaunchKernelGGL(k_coef_sep3<float>, grid, block, 0, s, g, vals, ninterp, table);
    } else {
        dim3 grid(cdiv(g.npos[0], TXC), cdiv(g.npos[1], TYC), (unsigned)(g.npos[2] * g.npos[3]));
        hipLaunchKernelGGL(k_coef_bf16x3<4>, grid, block, 0, s, g, vals, ninterp, Ab, table);
    }
    return hipGetLastError();
}
