// Quadcubic coefficients: the pieces the 4D table build (k_coef.hip) and the
// fused 4D query kernel (k_query.hip) share, so that the two agree bitwise.
//
// The reference's 4D operator (ARBInterp.py:726-878) is
//   A4 = M (x) M (x) M (x) M + P . dD,
//   P  = Hd (x) Hd (x) Hd (x) Hd, Hd = the derivative columns of the 1D
//        Hermite map, [[0,0],[1,0],[-2,-1],[1,1]] (row = power, column = end);
//   dD row k = S(k-1) - S(k), S(k) the xyzt central-difference stencil at
//        corner k (S(-1) = 0): the reference's row shift in the d4/dxdydzdt
//        block (ARBInterp.py:860).
// So alpha = the separable M passes (per power et of t: the t-combined 3D
// node tile through arbs3::sep3_passes4) + P . c, with c_k = G(k-1) - G(k),
// G the xyzt central difference at the cell's corner nodes.  P vanishes
// unless every power is >= 1 (81 of 256 rows), and row 85 (u v w t) of A4 is
// identically zero: alpha_85 is set to 0 exactly.
//
// P . c is separable as well: with the powers (ex, ey, et) of one z-pass
// thread fixed,
//   X(my, mz, mt) = Hd[ex][0] c(0, my, mz, mt) + Hd[ex][1] c(1, my, mz, mt)
//   Y(mz, mt)     = Hd[ey][0] X(0, mz, mt)     + Hd[ey][1] X(1, mz, mt)
//   S(mz)         = Hd[et][0] Y(mz, 0)         + Hd[et][1] Y(mz, 1)
//   corr(ez)      = Hd[ez][0] S(0)             + Hd[ez][1] S(1)
// (c index mx + 2 my + 4 mz + 8 mt).  Hd's entries are 0, +-1 and -2, so
// every product is exact and each line is one rounded sum, formed as
// fma(h1, b, h0 * a): about 30 FP64 operations per thread for its three
// corrected coefficients.
#pragma once
#include "arb_sep3.h"

namespace arbs4 {

__device__ __forceinline__ double hd(int e, int k) {
    return e == 0 ? 0.0 : e == 1 ? (k ? 0.0 : 1.0) : e == 2 ? (k ? -1.0 : -2.0) : 1.0;
}

// xyzt central difference at the node p points to, in a node tile with
// strides 1, SY_, SZ_, ST_ (doubles) along x, y, z, t: the sixteen
// neighbours at +-1 per axis, summed in the order of m = bit x + 2 bit y +
// 4 bit z + 8 bit t (bit set: +1), each with the sign of its parity
template <int SY_, int SZ_, int ST_>
__device__ __forceinline__ double corner_diff(const double* p) {
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int off = ((m & 1) ? 1 : -1) + ((m & 2) ? SY_ : -SY_) + ((m & 4) ? SZ_ : -SZ_) +
                        ((m & 8) ? ST_ : -ST_);
        const double v = p[off];
        acc += (__builtin_popcount(~m & 15) & 1) ? -v : v;
    }
    return acc * 0.0625;
}

// The row-shift correction of one z-pass thread: c[ez] (the separable
// alpha_{ex + 4 ey + 16 ez + 64 et}) becomes the 4D coefficient.  cv: the
// cell's sixteen c_k.
__device__ __forceinline__ void shift_rows(const double* cv, int ex, int ey, int et, double* c) {
    if (!(ex && ey && et)) return;                  // P's rows with a zero power vanish
    const double hx0 = hd(ex, 0), hx1 = hd(ex, 1), hy0 = hd(ey, 0), hy1 = hd(ey, 1);
    const double ht0 = hd(et, 0), ht1 = hd(et, 1);
    double S[2];
#pragma unroll
    for (int mz = 0; mz < 2; ++mz) {
        double Y[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const double* w = cv + 4 * mz + 8 * mt;
            const double x0 = fma(hx1, w[1], hx0 * w[0]);              // my = 0
            const double x1 = fma(hx1, w[3], hx0 * w[2]);              // my = 1
            Y[mt] = fma(hy1, x1, hy0 * x0);
        }
        S[mz] = fma(ht1, Y[1], ht0 * Y[0]);
    }
#pragma unroll
    for (int ez = 1; ez < 4; ++ez) c[ez] += fma(hd(ez, 1), S[1], hd(ez, 0) * S[0]);
    if (ex == 1 && ey == 1 && et == 1) c[1] = 0.0;  // row 85: the reference's row is zero
}

}  // namespace arbs4
