// Mixed-precision tricubic coefficients of one cell on the BF16 matrix cores
// (gfx950 v_mfma_f32_16x16x32_bf16), from the cell's 4x4x4 node
// neighbourhood F:
//     alpha = A . F,   A = inv(B) . D   (ARBInterp.py:107-175)
// A is exact dyadic with 16 A integer, |16 A_ij| <= 250: exact in BF16.  The
// contraction is made exact:
//   * F's corner-0 value f0 (neighbourhood index 21) is subtracted; A maps a
//     constant to e_0, so alpha_0 takes f0 back at the end;
//   * the residual r = F - f0 is scaled by 2^(8-E), E = the exponent of the
//     cell's largest |r|, so x = r 2^(8-E) lies in (-256, 256), and x is cut
//     into integer parts on a 2^-24 grid (truncation below it):
//         x = p0 + p1 2^-8 + p2 2^-16 + p3 2^-24,
//     p0 = rint(x) (|p0| <= 256), m = trunc((x - p0) 2^24), p1 = m >> 16
//     (|p1| <= 128), p2, p3 = the low bytes of m (0..255): all exact in BF16;
//   * every product (16 A_ik) p_k is an integer below 2^16 and every partial
//     sum over the 64 k an integer below 2^22, so each part's MFMA chain is
//     exact in FP32 whatever its order;
//   * the four accumulators combine into one integer below 2^45 in FP64
//     (exact), which is scaled by 2^(E-36) (the 2^-24 grid, 2^(8-E) and the
//     16 of 16 A), f0 is added to alpha_0, and the result is rounded once to
//     FP32.
// The stored alpha is therefore the FP64 alpha of F quantised to 2^(E-32)
// (a node-value error the interpolation weights carry with gain ~1), correctly
// rounded to FP32.  The table build and the fused query kernel both call this
// function, so every query path sees bitwise the same coefficients.
#pragma once
#include "arb_internal.h"

namespace arbc3 {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

// BF16 bits of a small integer (|v| <= 256: exact)
__device__ __forceinline__ unsigned short bf16_int(int v) {
    return (unsigned short)(__float_as_uint((float)v) >> 16);
}

// Lane layout (one wave, one column tile of 16 cells): lane l works on cell
// l & 15 and holds the k-slice ks*32 + 8*(l >> 4) + j (ks = 0, 1; j = 0..7)
// of the B operand.  nb(kk) returns F at neighbourhood index
// kk = ox + 4 oy + 16 oz of this lane's cell (it is called twice per entry,
// for the maximum and for the split, instead of holding 16 FP64 registers).
// Af is 16 A as BF16 in operand order [mt][ks][lane][8] (row mt*16 + (l & 15),
// column ks*32 + 8*(l >> 4) + j).  out(mt, v) receives
// v[r] = alpha[mt*16 + 4*(l >> 4) + r] of the lane's cell, mt = 0..3.
// Every lane of the wave must take part (the MFMAs and shuffles span it).
template <typename NB, typename OUT>
__device__ __forceinline__ void coef3_mixed(const NB& nb, const unsigned short* __restrict__ Af,
                                            const OUT& out) {
    const int lane = threadIdx.x & 63;
    const int grp = lane >> 4;
    const double f0 = nb(21);
    double mx = 0.0;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double v = nb(ks * 32 + 8 * grp + j) - f0;
            mx = fmax(mx, fabs(v));
            if (v != v) mx = v;                          // NaN data: NaN coefficients
        }
    }
    // the cell's entries live in lanes l, l^16, l^32, l^48
    double m1 = __shfl_xor(mx, 16);
    mx = (m1 != m1) ? m1 : fmax(mx, m1);
    m1 = __shfl_xor(mx, 32);
    mx = (m1 != m1) ? m1 : fmax(mx, m1);
    int E = 0;
    if (mx > 0.0 && mx < INFINITY) (void)frexp(mx, &E);   // mx < 2^E
    us8 bp[4][2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double x = ldexp(nb(ks * 32 + 8 * grp + j) - f0, 8 - E);   // |x| < 2^8
            const bool fin = x == x && fabs(x) < 512.0;
            const double p0 = fin ? rint(x) : 0.0;
            const int m = fin ? (int)((x - p0) * 0x1p24) : 0;   // exact, |.| <= 2^23; truncation
            bp[0][ks][j] = fin ? bf16_int((int)p0) : (unsigned short)0x7fc0;   // NaN/Inf -> NaN
            bp[1][ks][j] = bf16_int(m >> 16);
            bp[2][ks][j] = bf16_int((m >> 8) & 255);
            bp[3][ks][j] = bf16_int(m & 255);
        }
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f4 acc[4];
#pragma unroll
        for (int pp = 0; pp < 4; ++pp) acc[pp] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const us8 a = *reinterpret_cast<const us8*>(Af + (((mt * 2 + ks) * 64 + lane) * 8));
            const bf8 ab = __builtin_bit_cast(bf8, a);
#pragma unroll
            for (int pp = 0; pp < 4; ++pp)
                acc[pp] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    ab, __builtin_bit_cast(bf8, bp[pp][ks]), acc[pp], 0, 0, 0);
        }
        f4 v4;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // ((acc0 2^8 + acc1) 2^8 + acc2) 2^8 + acc3: an integer below 2^45
            double v = fma((double)acc[0][r], 256.0, (double)acc[1][r]);
            v = fma(v, 256.0, (double)acc[2][r]);
            v = fma(v, 256.0, (double)acc[3][r]);
            v = ldexp(v, E - 36);
            if (mt == 0 && grp == 0 && r == 0) v += f0;
            v4[r] = (float)v;
        }
        out(mt, v4);
    }
}

}  // namespace arbc3
