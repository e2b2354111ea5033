// Backward pass of the scalar interpolant (value + gradient outputs).
//
// No reference counterpart: the reference has no derivatives with respect to
// its inputs.  For upstream gradients gv (of U) and gg (of grad U):
//   d/dq   = gv grad U + Hess U . gg                  (per query, analytic)
//   d/df_n = sum_i sum_j c_ij A_{j, l(n, cell_i)}    (U is linear in f)
//            c_i = gv_i m(x_i) + sum_k gg_ik dm_k(x_i)/h_k
// The node adjoint is deterministic: queries are grouped per cell in
// query-index order (a stable radix sort of the cell keys, k_sort.hip), and
// the per-cell 64 (256)-vector alpha_bar is summed in that order.  3D: the
// per-query records are written by row (table path) or, from the node
// tiles, at the queries' grouped positions (a cell's records consecutive).  3D: blocks of 4^3
// cells apply the transpose of the separable coefficient passes to their
// alpha_bar, giving partial gradients on the block's 7^3 node tile, and
// every node sums the tiles holding it in a fixed block order.  4D:
// b_bar = A^T alpha_bar per cell, stored slot-major bbar[l][cell], and every
// node gathers b_bar from its <= 4^4 cells in a fixed slot order.  No
// floating-point atomics.
#include "arb_internal.h"

namespace {

template <int D>
struct Basis1 {
    double b0[D][4], b1[D][4], b2[D][4];
};

// t[e] for a runtime e in 0..3 by selects (a dynamic index into a register
// array would place the array in scratch memory)
__device__ __forceinline__ double sel4(const double* t, int e) {
    return e == 0 ? t[0] : e == 1 ? t[1] : e == 2 ? t[2] : t[3];
}

template <int D>
__device__ __forceinline__ void make_basis(const double* loc, Basis1<D>& B) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const double u = loc[a];
        B.b0[a][0] = 1.0; B.b0[a][1] = u; B.b0[a][2] = u * u; B.b0[a][3] = u * u * u;
        B.b1[a][0] = 0.0; B.b1[a][1] = 1.0; B.b1[a][2] = 2.0 * u; B.b1[a][3] = 3.0 * u * u;
        B.b2[a][0] = 0.0; B.b2[a][1] = 0.0; B.b2[a][2] = 2.0; B.b2[a][3] = 6.0 * u;
    }
}

// p(u), p'(u) and p''(u)/2 of c0 + c1 u + c2 u^2 + c3 u^3 (synthetic division).
__device__ __forceinline__ void h4dd(double c0, double c1, double c2, double c3, double u,
                                     double& p, double& d, double& hh) {
    p = c3;
    d = 0.0;
    hh = 0.0;
    hh = fma(hh, u, d); d = fma(d, u, p); p = fma(p, u, c2);
    hh = fma(hh, u, d); d = fma(d, u, p); p = fma(p, u, c1);
    hh = fma(hh, u, d); d = fma(d, u, p); p = fma(p, u, c0);
}

// Value, gradient and Hessian (cuboid units) of sum_i a_i u^ex v^ey w^ez by
// tensor Horner: rows along u (p, p', p''), then v, then w, keeping the
// derivative combinations of total order <= 2.
template <typename TC>
__device__ __forceinline__ void hess3(const TC* __restrict__ a, double u, double v, double w,
                                      double& val, double* g, double (*H)[3]) {
    double Q[4], Qv[4], Qvv[4], R[4], Rv[4], S[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double P[4], Pu[4], Puu[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const TC* c = a + 16 * k + 4 * j;
            double hh;
            h4dd((double)c[0], (double)c[1], (double)c[2], (double)c[3], u, P[j], Pu[j], hh);
            Puu[j] = 2.0 * hh;
        }
        double hq;
        h4dd(P[0], P[1], P[2], P[3], v, Q[k], Qv[k], hq);
        Qvv[k] = 2.0 * hq;
        h4d<double>(Pu[0], Pu[1], Pu[2], Pu[3], v, R[k], Rv[k]);
        S[k] = h4<double>(Puu[0], Puu[1], Puu[2], Puu[3], v);
    }
    double hw;
    h4dd(Q[0], Q[1], Q[2], Q[3], w, val, g[2], hw);
    H[2][2] = 2.0 * hw;
    h4d<double>(R[0], R[1], R[2], R[3], w, g[0], H[0][2]);
    h4d<double>(Qv[0], Qv[1], Qv[2], Qv[3], w, g[1], H[1][2]);
    H[0][0] = h4<double>(S[0], S[1], S[2], S[3], w);
    H[1][1] = h4<double>(Qvv[0], Qvv[1], Qvv[2], Qvv[3], w);
    H[0][1] = h4<double>(Rv[0], Rv[1], Rv[2], Rv[3], w);
    H[1][0] = H[0][1];
    H[2][0] = H[0][2];
    H[2][1] = H[1][2];
}

// 4D: the 3D quantities of the four t-slices, combined along t.
template <typename TC>
__device__ __forceinline__ void hess4(const TC* __restrict__ a, double u, double v, double w,
                                      double t, double& val, double* g, double (*H)[4]) {
    double r[4], gs[3][4], hs[3][3][4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        double gl[3], Hl[3][3];
        hess3<TC>(a + 64 * l, u, v, w, r[l], gl, Hl);
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            gs[x][l] = gl[x];
#pragma unroll
            for (int y = 0; y < 3; ++y) hs[x][y][l] = Hl[x][y];
        }
    }
    double ht;
    h4dd(r[0], r[1], r[2], r[3], t, val, g[3], ht);
    H[3][3] = 2.0 * ht;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        h4d<double>(gs[x][0], gs[x][1], gs[x][2], gs[x][3], t, g[x], H[x][3]);
        H[3][x] = H[x][3];
#pragma unroll
        for (int y = 0; y < 3; ++y)
            H[x][y] = h4<double>(hs[x][y][0], hs[x][y][1], hs[x][y][2], hs[x][y][3], t);
    }
}

// Cell key of each query for the node adjoint's grouping: its flat cell, or
// nc for rows outside the box (the group after the cells).
template <int D>
__global__ void __launch_bounds__(256)
k_bwd_keys(ArbGeom g, const double* __restrict__ q, int64_t n, int64_t ldq,
           int32_t* __restrict__ key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double loc[D];
    bool oob;
    const int64_t flat = arb_reduce<D>(q + i * ldq, g, loc, &oob);
    key[i] = flat < g.nc ? (int32_t)flat : (int32_t)g.nc;
}

// d/dq per query, and, for the node adjoint, the query's cell key and (3D)
// its record at row i.  Without perm, thread i takes query i: 3D reads a
// cell's 64 table coefficients as one contiguous 512-byte run, so query
// order costs nothing and the queries' own arrays are read and written in
// order.  With perm (the queries grouped by cell; 4D, whose 256
// coefficients of a cell are 2 KB), thread s takes query perm[s], so
// neighbouring threads read the same cell's coefficients.
template <int D, typename TC>
__global__ void __launch_bounds__(256)
k_bwd_points(ArbGeom g, const TC* __restrict__ table, int32_t ninterp, int32_t kn,
             const double* __restrict__ q, int64_t n, int64_t ldq, const double* __restrict__ gv,
             const double* __restrict__ gg, double* __restrict__ dq,
             const int32_t* __restrict__ perm, int32_t* __restrict__ key,
             double* __restrict__ rec) {
    constexpr int K = 1 << (2 * D);
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t i = perm ? (int64_t)perm[s] : s;
    if (i < 0 || i >= n) return;                    // unreachable: perm holds row indices
    double loc[D];
    bool oob;
    int64_t flat = arb_reduce<D>(q + i * ldq, g, loc, &oob);
    const bool ok = flat < g.nc;
    if (key) key[i] = ok ? (int32_t)flat : (int32_t)g.nc;  // rows outside: after the cells
    if (rec && ok) {
        // the node adjoint's per-query record: local coordinates, gv, gg / h
        double4* r4 = reinterpret_cast<double4*>(rec + i * 8);
        r4[0] = make_double4(loc[0], loc[1], loc[2], gv[i]);
        r4[1] = make_double4(gg[i * 3] / g.h[0], gg[i * 3 + 1] / g.h[1], gg[i * 3 + 2] / g.h[2], 0.0);
    }
    if (!dq) return;
    if (!ok) {
#pragma unroll
        for (int x = 0; x < D; ++x) dq[i * D + x] = NAN;
        return;
    }
    double val, gr[D], H[D][D];
    const TC* a = table + (flat * ninterp + kn) * K;
    if constexpr (D == 3) hess3<TC>(a, loc[0], loc[1], loc[2], val, gr, H);
    else hess4<TC>(a, loc[0], loc[1], loc[2], loc[3], val, gr, H);
    const double v = gv[i];
#pragma unroll
    for (int x = 0; x < D; ++x) {
        double acc = v * (gr[x] / g.h[x]);
#pragma unroll
        for (int y = 0; y < D; ++y) acc += (H[x][y] / (g.h[x] * g.h[y])) * gg[i * D + y];
        dq[i * D + x] = acc;
    }
}

// alpha_bar[j] of one 4D cell for the entries j = j0 + 64 r (r < 4; j0 a
// lane): the sum over the cell's queries, in query-index order, of
// gv m_j + sum_x gk_x dm_j/dx.  With j = e0 + 4 e1 + 16 e2 + 64 e3, lane j0
// fixes e0, e1, e2 and r is e3, so m_j = P b0_t[r] with P = u^e0 v^e1 w^e2
// shared by the lane's four entries:
//   t_j = (gv P + gk_u P_u + gk_v P_v + gk_w P_w) b0_t[r] + gk_t P b1_t[r].
template <int D, int R>
__device__ __forceinline__ void cell_abar(const ArbGeom& g, const double* __restrict__ q,
                                          int64_t ldq, const double* __restrict__ gv,
                                          const double* __restrict__ gg, const int32_t* perm,
                                          int32_t s0, int32_t s1, int j0, int stride,
                                          double* acc) {
    static_assert(D == 4 && R == 4, "4D: lane j0 holds the entries j0 + 64 r");
    (void)stride;
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    const int e0 = j0 & 3, e1 = (j0 >> 2) & 3, e2 = (j0 >> 4) & 3;
    for (int32_t s = s0; s < s1; ++s) {
        const int64_t i = perm[s];
        double loc[D];
        bool oob;
        (void)arb_reduce<D>(q + i * ldq, g, loc, &oob);
        Basis1<D> B;
        make_basis<D>(loc, B);
        const double v = gv[i];
        double gk[D];
#pragma unroll
        for (int x = 0; x < D; ++x) gk[x] = gg[i * D + x] / g.h[x];
        const double a0 = sel4(B.b0[0], e0), a1 = sel4(B.b0[1], e1), a2 = sel4(B.b0[2], e2);
        const double d0 = sel4(B.b1[0], e0), d1 = sel4(B.b1[1], e1), d2 = sel4(B.b1[2], e2);
        const double a01 = a0 * a1;
        const double P = a01 * a2;
        const double Pu = d0 * a1 * a2, Pv = a0 * d1 * a2, Pw = a01 * d2;
        const double c = v * P + gk[0] * Pu + gk[1] * Pv + gk[2] * Pw;
        const double e = gk[3] * P;
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += c * B.b0[3][r] + e * B.b1[3][r];
    }
}

// 3D: a workgroup takes a block of BB^3 cells.  Four threads per cell form
// alpha_bar into LDS from the queries' records (u, v, w, gv,
// gg / h): thread g of a cell sums, over the cell's queries in index order,
// the entries ex + 4 ey + 16 ez with ex + 4 ey in 4g..4g+3,
//   gv u^ex v^ey w^ez + gk_u (u^ex)' v^ey w^ez + gk_v u^ex (v^ey)' w^ez
//                     + gk_w u^ex v^ey (w^ez)'      (gk = gg / h).
// A = M (x) M (x) M (z slowest; M the 4x4 map of arb_sep3.h), so the node
// adjoint of the block, A^T alpha_bar spread over the cells' neighbourhoods,
// is the transpose of the separable passes: z, then y, then x, each a gather
//   out(n) = sum over (c, o) with c + o = n of  sum_e M[e][o] in(c, e)
// over the cells c of the block and offsets o = 0..3.  The result is the
// block's node tile, (BB+3)^3 partial node gradients, written to
// tiles[block][node].  Every sum runs in a fixed order (no atomics).
__device__ __forceinline__ void pow3(double x, double* p, double* dp) {
    p[0] = 1.0; p[1] = x; p[2] = x * x; p[3] = p[2] * x;
    dp[0] = 0.0; dp[1] = 1.0; dp[2] = 2.0 * x; dp[3] = 3.0 * p[2];
}

// sum_e M[e][o] a_e (column o of M applied to a; M as in arb_sep3.h)
__device__ __forceinline__ double mcol(int o, double a0, double a1, double a2, double a3) {
    switch (o) {
        case 0: return fma(-0.5, a3, fma(-0.5, a1, a2));
        case 1: return fma(1.5, a3, fma(-2.5, a2, a0));
        case 2: return fma(-1.5, a3, fma(0.5, a1, 2.0 * a2));
        default: return 0.5 * (a3 - a2);
    }
}

constexpr int BB = 4;                              // block edge (cells)
constexpr int BN = BB + 3;                         // node tile edge
constexpr int BCELLS = BB * BB * BB;
constexpr int BTILE = BN * BN * BN;

// A cell holding more than BWD_CHUNK queries is summed in pieces: the
// queries at sorted positions [o C, (o + 1) C) of such a cell (C =
// BWD_CHUNK, o C inside the cell) are summed by a chunk kernel into
// part[o], in index order, and the cell's own sum takes the queries before
// its first multiple of C in index order, then part[o] for each of its
// chunks in order: a fixed order, so the result still repeats bit for bit,
// and a cell of many queries (particles clustered in a few cells) is
// spread over many threads instead of four.
constexpr int32_t BWD_CHUNK = 1024;

// The chunk starting at sorted position o C, if it belongs to a cell of
// more than BWD_CHUNK queries: [s0, s1) and that cell; q, ldq: the query
// positions (the cell of the query at o C, by the same reduction as the
// keys).
template <int D>
__device__ __forceinline__ bool bwd_chunk(const ArbGeom& g, const double* __restrict__ q,
                                          int64_t ldq, const int32_t* __restrict__ offs,
                                          const int32_t* __restrict__ perm, int64_t o,
                                          int32_t* s0, int32_t* s1) {
    const int64_t p0 = o * BWD_CHUNK;
    if (p0 >= offs[g.nc]) return false;              // the rows outside the box follow
    const int64_t i = perm[p0];
    double loc[D];
    bool oob;
    const int64_t c = arb_reduce<D>(q + i * ldq, g, loc, &oob);
    if (c < 0 || c >= g.nc) return false;            // cannot happen: p0 is a cell's row
    const int32_t e = offs[c + 1];
    if (e - offs[c] <= BWD_CHUNK) return false;      // summed whole by the cell kernel
    *s0 = (int32_t)p0;
    *s1 = p0 + BWD_CHUNK < e ? (int32_t)(p0 + BWD_CHUNK) : e;
    return true;
}

// 3D: thread gq (of four) of a cell: its 16 entries of alpha_bar, summed over
// the queries at sorted positions [s0, s1) in order, into acc[p][ez]
// (entries 4 gq + p + 16 ez).
__device__ __forceinline__ void abar_row3(const double4& r0, const double4& r1, int gq,
                                          double (*acc)[4]);

template <int NB>
__device__ __forceinline__ void abar_rows3(const double* __restrict__ rec,
                                           const int32_t* __restrict__ rperm, int32_t s0,
                                           int32_t s1, int gq, double (*acc)[4]) {
    // the record of sorted position s: rec[s] (written at the grouped
    // positions: a cell's records are consecutive) or rec[rperm[s]] (written
    // by row); NB (1 or 2) loaded together, then added in query order
    int32_t s = s0;
    if constexpr (NB == 2) {
        for (; s + 2 <= s1; s += 2) {
            const int64_t ra = rperm ? (int64_t)rperm[s] : s;
            const int64_t rb = rperm ? (int64_t)rperm[s + 1] : s + 1;
            const double4* r4a = reinterpret_cast<const double4*>(rec + ra * 8);
            const double4* r4b = reinterpret_cast<const double4*>(rec + rb * 8);
            const double4 a0 = r4a[0], a1 = r4a[1], b0 = r4b[0], b1 = r4b[1];
            abar_row3(a0, a1, gq, acc);
            abar_row3(b0, b1, gq, acc);
        }
    }
    for (; s < s1; ++s) {
        const int64_t ra = rperm ? (int64_t)rperm[s] : s;
        const double4* r4 = reinterpret_cast<const double4*>(rec + ra * 8);
        abar_row3(r4[0], r4[1], gq, acc);
    }
}

// One query's contribution (record r0 = u, v, w, gv; r1 = gg / h) to the 16
// entries of thread gq.
__device__ __forceinline__ void abar_row3(const double4& r0, const double4& r1, int gq,
                                          double (*acc)[4]) {
    {
        double pu[4], du[4], pv[4], dv[4], pw[4], dw[4];
        pow3(r0.x, pu, du);
        pow3(r0.y, pv, dv);
        pow3(r0.z, pw, dw);
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int exy = 4 * gq + p, ex = exy & 3, ey = exy >> 2;
            const double a = sel4(pu, ex) * sel4(pv, ey);
            const double au = sel4(du, ex) * sel4(pv, ey);
            const double av = sel4(pu, ex) * sel4(dv, ey);
            const double base = fma(r1.y, av, fma(r1.x, au, r0.w * a));
            const double gwa = r1.z * a;
#pragma unroll
            for (int ez = 0; ez < 4; ++ez) acc[p][ez] += fma(gwa, dw[ez], base * pw[ez]);
        }
    }
}

// 3D chunk sums: four threads per chunk, part[o][entry] (64 entries).
__global__ void __launch_bounds__(256)
k_bwd_chunk3(ArbGeom g, const double* __restrict__ q, int64_t ldq, const double* __restrict__ rec,
             const int32_t* __restrict__ offs, const int32_t* __restrict__ perm,
             const int32_t* __restrict__ rperm, int64_t nchunk, double* __restrict__ part) {
    const int64_t o = (int64_t)blockIdx.x * 64 + (threadIdx.x >> 2);
    const int gq = threadIdx.x & 3;
    int32_t s0, s1;
    if (o >= nchunk || !bwd_chunk<3>(g, q, ldq, offs, perm, o, &s0, &s1)) return;
    double acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int ez = 0; ez < 4; ++ez) acc[p][ez] = 0.0;
    abar_rows3<2>(rec, rperm, s0, s1, gq, acc);
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int ez = 0; ez < 4; ++ez) part[o * 64 + 4 * gq + p + 16 * ez] = acc[p][ez];
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_bwd_block3(ArbGeom g, const double* __restrict__ rec, const int32_t* __restrict__ offs,
             const int32_t* __restrict__ rperm, int64_t nbx, int64_t nby,
             const double* __restrict__ part, double* __restrict__ tiles) {
    // the z pass's output, then the y pass's (20.6 KB in all)
    __shared__ double t2[BN][BB][BB][16];           // [z][cy][cx][ex + 4 ey]
    __shared__ double t1[BN][BN][BB][4];            // [z][y][cx][ex]
    for (int t = threadIdx.x; t < BN * BB * BB * 16; t += 256) (&t2[0][0][0][0])[t] = 0.0;
    const int64_t blk = blockIdx.x;
    const int64_t bx = blk % nbx, by = (blk / nbx) % nby, bz = blk / (nbx * nby);
    {
        const int cl = threadIdx.x >> 2, gq = threadIdx.x & 3;
        const int cx = cl & 3, cy = (cl >> 2) & 3, cz = cl >> 4;
        const int64_t x = bx * BB + cx, y = by * BB + cy, z = bz * BB + cz;
        double acc[4][4];                          // [pair][ez]: alpha_bar[4 gq + p + 16 ez]
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int ez = 0; ez < 4; ++ez) acc[p][ez] = 0.0;
        if (x < g.npos[0] && y < g.npos[1] && z < g.npos[2]) {
            const int64_t c = x + g.npos[0] * (y + g.npos[1] * z);
            const int32_t s0 = offs[c], s1 = offs[c + 1];
            if (s1 - s0 <= BWD_CHUNK) {
                abar_rows3<1>(rec, rperm, s0, s1, gq, acc);
            } else {
                // the queries before the cell's first chunk, then its chunks
                const int64_t ofirst = ((int64_t)s0 + BWD_CHUNK - 1) / BWD_CHUNK;
                const int32_t h1 = ofirst * BWD_CHUNK < s1 ? (int32_t)(ofirst * BWD_CHUNK) : s1;
                abar_rows3<1>(rec, rperm, s0, h1, gq, acc);
                // then the chunks' sums, in chunk order
                for (int64_t o = ofirst; o * BWD_CHUNK < s1; ++o) {
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int ez = 0; ez < 4; ++ez)
                            acc[p][ez] += part[o * 64 + 4 * gq + p + 16 * ez];
                }
            }
        }
        // z: t2[z][cy][cx][exy] = sum_{cz + oz = z} sum_ez M[ez][oz] alpha_bar[cell][exy + 16 ez],
        // from the registers: one wave (one cz) at a time, cz descending, so
        // every entry adds its terms in ascending oz from 0.0
        __syncthreads();                           // t2 cleared
        for (int ph = BB - 1; ph >= 0; --ph) {
            if (cz == ph) {
#pragma unroll
                for (int p = 0; p < 4; ++p)
#pragma unroll
                    for (int oz = 0; oz < 4; ++oz)
                        t2[cz + oz][cy][cx][4 * gq + p] +=
                            mcol(oz, acc[p][0], acc[p][1], acc[p][2], acc[p][3]);
            }
            __syncthreads();
        }
    }
    // y: t1[z][y][cx][ex] = sum_{cy + oy = y} sum_ey M[ey][oy] t2[z][cy][cx][ex + 4 ey]
    for (int t = threadIdx.x; t < BN * BN * BB * 4; t += 256) {
        const int ex = t & 3, cx = (t >> 2) & 3, r = t >> 4, y = r % BN, z = r / BN;
        double sum = 0.0;
#pragma unroll
        for (int oy = 0; oy < 4; ++oy) {
            const int cy = y - oy;
            if (cy >= 0 && cy < BB) {
                const double* a = &t2[z][cy][cx][ex];
                sum += mcol(oy, a[0], a[4], a[8], a[12]);
            }
        }
        t1[z][y][cx][ex] = sum;
    }
    __syncthreads();
    // x: tile[z][y][x] = sum_{cx + ox = x} sum_ex M[ex][ox] t1[z][y][cx][ex]
    double* out = tiles + blk * BTILE;
    for (int t = threadIdx.x; t < BTILE; t += 256) {
        const int x = t % BN, r = t / BN, y = r % BN, z = r / BN;
        double sum = 0.0;
#pragma unroll
        for (int ox = 0; ox < 4; ++ox) {
            const int cx = x - ox;
            if (cx >= 0 && cx < BB) {
                const double* a = &t1[z][y][cx][0];
                sum += mcol(ox, a[0], a[1], a[2], a[3]);
            }
        }
        out[t] = sum;
    }
}

// Node gradient: the sum, over the blocks whose tile holds the node (at most
// two per axis), of the tile entries, blocks in ascending z, y, x order.
__global__ void __launch_bounds__(256)
k_bwd_nodes3(ArbGeom g, const double* __restrict__ tiles, int64_t nbx, int64_t nby, int64_t nbz,
             double* __restrict__ dval) {
    const int64_t nidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nidx >= g.nnodes) return;
    const int64_t i = nidx % g.n[0], j = (nidx / g.n[0]) % g.n[1], k = nidx / (g.n[0] * g.n[1]);
    // block b's tile holds nodes BB*b .. BB*b + BN - 1
    auto lo = [](int64_t v) { const int64_t b = v - (BN - 1); return b <= 0 ? (int64_t)0 : (b + BB - 1) / BB; };
    const int64_t x0 = lo(i), x1 = min(i / BB, nbx - 1);
    const int64_t y0 = lo(j), y1 = min(j / BB, nby - 1);
    const int64_t z0 = lo(k), z1 = min(k / BB, nbz - 1);
    double sum = 0.0;
    for (int64_t bz = z0; bz <= z1; ++bz)
        for (int64_t by = y0; by <= y1; ++by)
            for (int64_t bx = x0; bx <= x1; ++bx) {
                const int64_t blk = bx + nbx * (by + nby * bz);
                const int64_t t = (i - BB * bx) + BN * ((j - BB * by) + BN * (k - BB * bz));
                sum += tiles[blk * BTILE + t];
            }
    dval[nidx] = sum;
}

// 4D: a workgroup per 16 cells: one wave per cell at a time sums alpha_bar
// (256) over the cell's queries in index order, then b_bar = A^T alpha_bar
// for the 16 cells together (each element of A read once per workgroup),
// written slot-major bbar[l][cell].
// 4D chunk sums (cells of more than BWD_CHUNK queries): one wave per chunk,
// part[o][entry] (256 entries).
template <int D>
__global__ void __launch_bounds__(256)
k_bwd_chunk(ArbGeom g, const double* __restrict__ q, int64_t ldq, const double* __restrict__ gv,
            const double* __restrict__ gg, const int32_t* __restrict__ offs,
            const int32_t* __restrict__ perm, int64_t nchunk, double* __restrict__ part) {
    constexpr int K = 1 << (2 * D);
    constexpr int PER = K / 64;
    const int lane = threadIdx.x & 63;
    const int64_t o = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int32_t s0, s1;
    if (o >= nchunk || !bwd_chunk<D>(g, q, ldq, offs, perm, o, &s0, &s1)) return;
    double acc[PER];
    cell_abar<D, PER>(g, q, ldq, gv, gg, perm, s0, s1, lane, 64, acc);
#pragma unroll
    for (int r = 0; r < PER; ++r) part[o * K + lane + 64 * r] = acc[r];
}

constexpr int BWD_CELLS4 = 16;                    // cells per workgroup, 4D

template <int D>
__global__ void __launch_bounds__(256)
k_bwd_cells(ArbGeom g, const double* __restrict__ A, const double* __restrict__ q, int64_t ldq,
            const double* __restrict__ gv, const double* __restrict__ gg,
            const int32_t* __restrict__ offs, const int32_t* __restrict__ perm,
            const double* __restrict__ part, double* __restrict__ bbar) {
    constexpr int K = 1 << (2 * D);
    constexpr int PER = K / 64;
    constexpr int CB = BWD_CELLS4;
    static_assert(K == 256, "one thread per entry l of b_bar");
    __shared__ double abar[CB][K];
    __shared__ int32_t any;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t c0 = (int64_t)blockIdx.x * CB;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    // alpha_bar of the workgroup's cells: wave w forms cells w, w + 4, ...
    for (int k = w; k < CB; k += 4) {
        const int64_t c = c0 + k;
        const bool live = c < g.nc;
        const int32_t s0 = live ? offs[c] : 0, s1 = live ? offs[c + 1] : 0;
        double acc[PER];
        if (s1 - s0 <= BWD_CHUNK) {
            cell_abar<D, PER>(g, q, ldq, gv, gg, perm, s0, s1, lane, 64, acc);
        } else {
            // the queries before the cell's first chunk, then its chunks
            const int64_t ofirst = ((int64_t)s0 + BWD_CHUNK - 1) / BWD_CHUNK;
            const int32_t h1 = ofirst * BWD_CHUNK < s1 ? (int32_t)(ofirst * BWD_CHUNK) : s1;
            cell_abar<D, PER>(g, q, ldq, gv, gg, perm, s0, h1, lane, 64, acc);
            for (int64_t o = ofirst; o * BWD_CHUNK < s1; ++o) {
#pragma unroll
                for (int r = 0; r < PER; ++r) acc[r] += part[o * K + lane + 64 * r];
            }
        }
#pragma unroll
        for (int r = 0; r < PER; ++r) abar[k][lane + 64 * r] = acc[r];
        if (s0 != s1 && lane == 0) any = 1;
    }
    __syncthreads();
    // b_bar[l][c] = sum_j A[j][l] alpha_bar[c][j], j ascending from 0, for
    // the CB cells at once: thread l reads each A[j][l] once (a row of A is
    // one coalesced read) and alpha_bar[c][j] is an LDS broadcast.  A cell
    // without queries gets 0.
    const int l = threadIdx.x;
    double b[CB];
#pragma unroll
    for (int k = 0; k < CB; ++k) b[k] = 0.0;
    if (any) {
        for (int j = 0; j < K; ++j) {
            const double a = A[(int64_t)j * K + l];
#pragma unroll
            for (int k = 0; k < CB; ++k) b[k] += a * abar[k][j];
        }
    }
#pragma unroll
    for (int k = 0; k < CB; ++k) {
        const int64_t c = c0 + k;
        if (c < g.nc) {
            const int32_t s0 = offs[c], s1 = offs[c + 1];
            bbar[(int64_t)l * g.nc + c] = s0 != s1 ? b[k] : 0.0;
        }
    }
}

// Node adjoint: gather b_bar from the cells whose neighbourhood holds the node.
template <int D>
__global__ void __launch_bounds__(256)
k_bwd_nodes(ArbGeom g, const double* __restrict__ bbar, double* __restrict__ dval) {
    constexpr int K = 1 << (2 * D);
    const int64_t nidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nidx >= g.nnodes) return;
    int64_t x[D];
    int64_t rem = nidx;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        x[a] = rem % g.n[a];
        rem /= g.n[a];
    }
    double sum = 0.0;
    for (int l = 0; l < K; ++l) {          // neighbour slot l: offsets o_a = (l >> 2a) & 3
        int64_t cell = 0, mul = 1;
        bool ok = true;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const int64_t ca = x[a] - ((l >> (2 * a)) & 3);
            ok &= (ca >= 0) && (ca < g.npos[a]);
            cell += ca * mul;
            mul *= g.npos[a];
        }
        if (ok) sum += bbar[(int64_t)l * g.nc + cell];
    }
    dval[nidx] = sum;
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t arb_launch_bwd_keys(const ArbGeom& g, const double* q, int64_t n, int64_t ldq,
                               int32_t* key, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (g.ndim == 3)
        hipLaunchKernelGGL(k_bwd_keys<3>, dim3(grid1(n, 256)), dim3(256), 0, s, g, q, n, ldq, key);
    else
        hipLaunchKernelGGL(k_bwd_keys<4>, dim3(grid1(n, 256)), dim3(256), 0, s, g, q, n, ldq, key);
    return hipGetLastError();
}

hipError_t arb_launch_bwd_points(const ArbGeom& g, int32_t prec, const void* table,
                                 int32_t ninterp, int32_t kn, const double* q, int64_t n,
                                 int64_t ldq, const double* gv, const double* gg, double* dq,
                                 const int32_t* perm, int32_t* key, double* rec,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 grid(grid1(n, 256)), block(256);
    if (g.ndim == 3) {
        if (prec == ARB_PREC_FP64)
            hipLaunchKernelGGL((k_bwd_points<3, double>), grid, block, 0, s, g, (const double*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, perm, key, rec);
        else
            hipLaunchKernelGGL((k_bwd_points<3, float>), grid, block, 0, s, g, (const float*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, perm, key, rec);
    } else {
        if (prec == ARB_PREC_FP64)
            hipLaunchKernelGGL((k_bwd_points<4, double>), grid, block, 0, s, g, (const double*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, perm, key, rec);
        else
            hipLaunchKernelGGL((k_bwd_points<4, float>), grid, block, 0, s, g, (const float*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, perm, key, rec);
    }
    return hipGetLastError();
}

// Blocks of the 3D node adjoint per axis, and the scratch it needs (doubles).
void arb_bwd_blocks3(const ArbGeom& g, int64_t* nb3) {
    for (int a = 0; a < 3; ++a) nb3[a] = (g.npos[a] + BB - 1) / BB;
}

int64_t arb_bwd_scratch(const ArbGeom& g) {
    if (g.ndim == 3) {
        int64_t nb3[3];
        arb_bwd_blocks3(g, nb3);
        return nb3[0] * nb3[1] * nb3[2] * BTILE;
    }
    return g.nc * 256;
}

int64_t arb_bwd_part_elems(const ArbGeom& g, int64_t n) {
    const int64_t nchunk = (n + BWD_CHUNK - 1) / BWD_CHUNK;
    return (nchunk ? nchunk : 1) * (g.ndim == 3 ? 64 : 256);
}

hipError_t arb_launch_bwd_cells(const ArbGeom& g, const double* A, const double* q, int64_t ldq,
                                const double* gv, const double* gg, const double* rec,
                                int32_t rec_grouped, const int32_t* offs, const int32_t* perm,
                                int64_t n, double* part, double* bbar, hipStream_t s) {
    const int32_t* rperm = rec_grouped ? nullptr : perm;
    const int64_t nchunk = (n + BWD_CHUNK - 1) / BWD_CHUNK;
    if (g.ndim == 3) {
        if (nchunk > 0)
            hipLaunchKernelGGL(k_bwd_chunk3, dim3(grid1(nchunk, 64)), dim3(256), 0, s, g, q, ldq,
                               rec, offs, perm, rperm, nchunk, part);
        int64_t nb3[3];
        arb_bwd_blocks3(g, nb3);
        hipLaunchKernelGGL(k_bwd_block3, dim3((unsigned)(nb3[0] * nb3[1] * nb3[2])), dim3(256), 0,
                           s, g, rec, offs, rperm, nb3[0], nb3[1], (const double*)part, bbar);
    } else {
        if (nchunk > 0)
            hipLaunchKernelGGL(k_bwd_chunk<4>, dim3(grid1(nchunk, 4)), dim3(256), 0, s, g, q, ldq,
                               gv, gg, offs, perm, nchunk, part);
        hipLaunchKernelGGL(k_bwd_cells<4>, dim3(grid1(g.nc, BWD_CELLS4)), dim3(256), 0, s, g, A,
                           q, ldq, gv, gg, offs, perm, (const double*)part, bbar);
    }
    return hipGetLastError();
}

hipError_t arb_launch_bwd_nodes(const ArbGeom& g, const double* bbar, double* dval, hipStream_t s) {
    dim3 grid(grid1(g.nnodes, 256)), block(256);
    if (g.ndim == 3) {
        int64_t nb3[3];
        arb_bwd_blocks3(g, nb3);
        hipLaunchKernelGGL(k_bwd_nodes3, grid, block, 0, s, g, bbar, nb3[0], nb3[1], nb3[2], dval);
    } else {
        hipLaunchKernelGGL(k_bwd_nodes<4>, grid, block, 0, s, g, bbar, dval);
    }
    return hipGetLastError();
}
