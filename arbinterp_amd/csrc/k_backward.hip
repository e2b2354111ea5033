// Backward pass of the scalar interpolant (value + gradient outputs).
//
// No reference counterpart: the reference has no derivatives with respect to
// its inputs.  For upstream gradients gv (of U) and gg (of grad U):
//   d/dq   = gv grad U + Hess U . gg                  (per query, analytic)
//   d/df_n = sum_i sum_j c_ij A_{j, l(n, cell_i)}    (U is linear in f)
//            c_i = gv_i m(x_i) + sum_k gg_ik dm_k(x_i)/h_k
// The node adjoint is deterministic: queries are grouped per cell (counting
// sort), each cell's segment is put in query-index order, the per-cell
// 64 (256)-vector alpha_bar is summed in that order, b_bar = A^T alpha_bar,
// and every node gathers b_bar from its <= 4^d cells in a fixed order.  No
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

// Value, gradient and Hessian (cuboid units) of sum_i a_i m_i.
template <int D, typename TC>
__device__ void val_grad_hess(const TC* __restrict__ a, const Basis1<D>& B, double& val,
                              double* g, double (*H)[D]) {
    constexpr int K = 1 << (2 * D);
    val = 0.0;
    for (int x = 0; x < D; ++x) {
        g[x] = 0.0;
        for (int y = 0; y < D; ++y) H[x][y] = 0.0;
    }
    for (int i = 0; i < K; ++i) {
        const double c = (double)a[i];
        int e[D];
#pragma unroll
        for (int x = 0; x < D; ++x) e[x] = (i >> (2 * x)) & 3;
        double p0 = c;
#pragma unroll
        for (int x = 0; x < D; ++x) p0 *= sel4(B.b0[x], e[x]);
        val += p0;
#pragma unroll
        for (int x = 0; x < D; ++x) {
            double gx = c * sel4(B.b1[x], e[x]);
            double hxx = c * sel4(B.b2[x], e[x]);
#pragma unroll
            for (int y = 0; y < D; ++y)
                if (y != x) { gx *= sel4(B.b0[y], e[y]); hxx *= sel4(B.b0[y], e[y]); }
            g[x] += gx;
            H[x][x] += hxx;
#pragma unroll
            for (int y = x + 1; y < D; ++y) {
                double hxy = c * sel4(B.b1[x], e[x]) * sel4(B.b1[y], e[y]);
#pragma unroll
                for (int z = 0; z < D; ++z)
                    if (z != x && z != y) hxy *= sel4(B.b0[z], e[z]);
                H[x][y] += hxy;
            }
        }
    }
#pragma unroll
    for (int x = 0; x < D; ++x)
#pragma unroll
        for (int y = 0; y < x; ++y) H[x][y] = H[y][x];
}

// d/dq per query, and the cell key for the node adjoint.
template <int D, typename TC>
__global__ void __launch_bounds__(256)
k_bwd_points(ArbGeom g, const TC* __restrict__ table, int32_t ninterp, int32_t kn,
             const double* __restrict__ q, int64_t n, int64_t ldq, const double* __restrict__ gv,
             const double* __restrict__ gg, double* __restrict__ dq, int32_t* __restrict__ key,
             int32_t* __restrict__ count) {
    constexpr int K = 1 << (2 * D);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double loc[D];
    bool oob;
    int64_t flat = arb_reduce<D>(q + i * ldq, g, loc, &oob);
    const bool ok = flat < g.nc;
    if (key) {
        key[i] = ok ? (int32_t)flat : -1;
        if (ok) atomicAdd(&count[flat], 1);
    }
    if (!dq) return;
    if (!ok) {
#pragma unroll
        for (int x = 0; x < D; ++x) dq[i * D + x] = NAN;
        return;
    }
    Basis1<D> B;
    make_basis<D>(loc, B);
    double val, gr[D], H[D][D];
    val_grad_hess<D, TC>(table + (flat * ninterp + kn) * K, B, val, gr, H);
    const double v = gv[i];
#pragma unroll
    for (int x = 0; x < D; ++x) {
        double acc = v * (gr[x] / g.h[x]);
#pragma unroll
        for (int y = 0; y < D; ++y) acc += (H[x][y] / (g.h[x] * g.h[y])) * gg[i * D + y];
        dq[i * D + x] = acc;
    }
}

__global__ void k_seg_scatter(const int32_t* __restrict__ key, const int32_t* __restrict__ offs,
                              int32_t* __restrict__ cursor, int32_t* __restrict__ perm, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = key[i];
    if (c >= 0) perm[offs[c] + atomicAdd(&cursor[c], 1)] = (int32_t)i;
}

// Put every cell's segment in query-index order (segments are short).
__global__ void k_seg_sort(const int32_t* __restrict__ offs, int32_t* __restrict__ perm,
                           int64_t nc) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int32_t s0 = offs[c], s1 = offs[c + 1];
    for (int32_t s = s0 + 1; s < s1; ++s) {
        const int32_t v = perm[s];
        int32_t t = s - 1;
        while (t >= s0 && perm[t] > v) {
            perm[t + 1] = perm[t];
            --t;
        }
        perm[t + 1] = v;
    }
}

// One wave per cell: alpha_bar (K) summed over the cell's queries in index
// order, then b_bar = A^T alpha_bar, written to bbar[cell][K].
template <int D>
__global__ void __launch_bounds__(256)
k_bwd_cells(ArbGeom g, const double* __restrict__ A, const double* __restrict__ q, int64_t ldq,
            const double* __restrict__ gv, const double* __restrict__ gg,
            const int32_t* __restrict__ offs, const int32_t* __restrict__ perm,
            double* __restrict__ bbar) {
    constexpr int K = 1 << (2 * D);
    constexpr int PER = K / 64;
    __shared__ double abar[4][K];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t c = (int64_t)blockIdx.x * 4 + w;
    const bool live = c < g.nc;
    const int32_t s0 = live ? offs[c] : 0, s1 = live ? offs[c + 1] : 0;
    double acc[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) acc[r] = 0.0;
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
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int j = lane + 64 * r;
            int e[D];
#pragma unroll
            for (int x = 0; x < D; ++x) e[x] = (j >> (2 * x)) & 3;
            double m = 1.0;
#pragma unroll
            for (int x = 0; x < D; ++x) m *= sel4(B.b0[x], e[x]);
            double t = v * m;
#pragma unroll
            for (int x = 0; x < D; ++x) {
                double dm = sel4(B.b1[x], e[x]);
#pragma unroll
                for (int y = 0; y < D; ++y)
                    if (y != x) dm *= sel4(B.b0[y], e[y]);
                t += gk[x] * dm;
            }
            acc[r] += t;
        }
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) abar[w][lane + 64 * r] = acc[r];
    __syncthreads();
    if (!live) return;
    double* out = bbar + c * K;
    if (s0 == s1) {
#pragma unroll
        for (int r = 0; r < PER; ++r) out[lane + 64 * r] = 0.0;
        return;
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int l = lane + 64 * r;
        double b = 0.0;
        for (int j = 0; j < K; ++j) b += A[(int64_t)j * K + l] * abar[w][j];
        out[l] = b;
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
        if (ok) sum += bbar[cell * K + l];
    }
    dval[nidx] = sum;
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t arb_launch_bwd_points(const ArbGeom& g, int32_t prec, const void* table,
                                 int32_t ninterp, int32_t kn, const double* q, int64_t n,
                                 int64_t ldq, const double* gv, const double* gg, double* dq,
                                 int32_t* key, int32_t* count, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 grid(grid1(n, 256)), block(256);
    if (g.ndim == 3) {
        if (prec == ARB_PREC_FP64)
            hipLaunchKernelGGL((k_bwd_points<3, double>), grid, block, 0, s, g, (const double*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, key, count);
        else
            hipLaunchKernelGGL((k_bwd_points<3, float>), grid, block, 0, s, g, (const float*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, key, count);
    } else {
        if (prec == ARB_PREC_FP64)
            hipLaunchKernelGGL((k_bwd_points<4, double>), grid, block, 0, s, g, (const double*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, key, count);
        else
            hipLaunchKernelGGL((k_bwd_points<4, float>), grid, block, 0, s, g, (const float*)table,
                               ninterp, kn, q, n, ldq, gv, gg, dq, key, count);
    }
    return hipGetLastError();
}

hipError_t arb_launch_seg_scatter(const int32_t* key, const int32_t* offs, int32_t* cursor,
                                  int32_t* perm, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_seg_scatter, dim3(grid1(n, 256)), dim3(256), 0, s, key, offs, cursor, perm, n);
    return hipGetLastError();
}

hipError_t arb_launch_seg_sort(const int32_t* offs, int32_t* perm, int64_t nc, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_sort, dim3(grid1(nc, 256)), dim3(256), 0, s, offs, perm, nc);
    return hipGetLastError();
}

hipError_t arb_launch_bwd_cells(const ArbGeom& g, const double* A, const double* q, int64_t ldq,
                                const double* gv, const double* gg, const int32_t* offs,
                                const int32_t* perm, double* bbar, hipStream_t s) {
    dim3 grid(grid1(g.nc, 4)), block(256);
    if (g.ndim == 3)
        hipLaunchKernelGGL(k_bwd_cells<3>, grid, block, 0, s, g, A, q, ldq, gv, gg, offs, perm, bbar);
    else
        hipLaunchKernelGGL(k_bwd_cells<4>, grid, block, 0, s, g, A, q, ldq, gv, gg, offs, perm, bbar);
    return hipGetLastError();
}

hipError_t arb_launch_bwd_nodes(const ArbGeom& g, const double* bbar, double* dval, hipStream_t s) {
    dim3 grid(grid1(g.nnodes, 256)), block(256);
    if (g.ndim == 3) hipLaunchKernelGGL(k_bwd_nodes<3>, grid, block, 0, s, g, bbar, dval);
    else hipLaunchKernelGGL(k_bwd_nodes<4>, grid, block, 0, s, g, bbar, dval);
    return hipGetLastError();
}
