// Many fields in one launch, and the RK4 particle push.
//
// No reference counterpart: the reference evaluates one field per Query call
// and dropped its particle integrator in v1.8 (CHANGELOG.md:8, "Created for
// simulating particle motion in magnetic fields", CHANGELOG.md:22).
//
// Batched query: one persistent grid; each workgroup pulls work items
// (field, chunk of `chunk` queries) from a queue counter until the queue is
// empty, finds the field by binary search over the per-field item offsets,
// and evaluates its chunk lane-per-query from that field's table.  Every wave
// exits once the counter passes the item count, so the grid always drains.
//
// RK4 push: each lane advances one particle (x, v) by `nsteps` classical RK4
// steps under a = -grad U; the current cell's coefficients stay in registers
// and are reloaded only when a stage lands in another cell.  Positions that
// leave the interpolation box turn NaN (as the reference's batch query does)
// and the NaN then propagates through position and velocity.
#include "arb_internal.h"

namespace {

struct BatchField {
    ArbGeom g;
    const void* table;
    int64_t q0;          // first query row of this field
    int64_t nq;          // its number of queries
    int64_t item0;       // first work item of this field
};

template <typename TO, typename T>
__device__ __forceinline__ void put(void* p, int64_t idx, T v) {
    reinterpret_cast<TO*>(p)[idx] = (TO)v;
}

template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(256)
k_batch_query(const BatchField* __restrict__ fields, int32_t nfields, int64_t nitems, int32_t chunk,
              int32_t ninterp, int32_t has_comps, int32_t has_norm, const double* __restrict__ q,
              int64_t ldq, void* comps, void* norms, void* grads, int32_t* flag,
              unsigned int* counter) {
    constexpr int K = (D == 3) ? 64 : 256;
    __shared__ long long item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = (long long)atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t item = item_s;
        __syncthreads();
        if (item >= nitems) return;                 // uniform: every wave leaves
        int lo = 0, hi = nfields - 1;
        while (lo < hi) {                           // last field with item0 <= item
            const int mid = (lo + hi + 1) >> 1;
            if (fields[mid].item0 <= item) lo = mid; else hi = mid - 1;
        }
        const BatchField& F = fields[lo];
        const int64_t base = (item - F.item0) * chunk;
        for (int r = threadIdx.x; r < chunk; r += blockDim.x) {
            const int64_t j = base + r;
            if (j >= F.nq) break;
            const int64_t i = F.q0 + j;
            double loc[D];
            bool oob;
            int64_t flat = arb_reduce<D>(q + i * ldq, F.g, loc, &oob);
            if (flat > F.g.nc) { flag[0] = 1; flat = F.g.nc; }
            const TC* a = reinterpret_cast<const TC*>(F.table) + flat * (int64_t)ninterp * K;
            T x[D];
#pragma unroll
            for (int d = 0; d < D; ++d) x[d] = (T)loc[d];
            int k = 0;
            if (has_comps) {
                for (int c = 0; c < 3; ++c) {
                    T v;
                    if constexpr (D == 3) v = arb_val3<T, TC>(a + c * K, x[0], x[1], x[2]);
                    else v = arb_val4<T, TC>(a + c * K, x[0], x[1], x[2], x[3]);
                    put<TO>(comps, i * 3 + c, v);
                }
                k = 3;
            }
            if (has_norm) {
                T v, gr[D];
                if constexpr (D == 3) arb_valgrad3<T, TC>(a + k * K, x[0], x[1], x[2], v, gr);
                else arb_valgrad4<T, TC>(a + k * K, x[0], x[1], x[2], x[3], v, gr);
                put<TO>(norms, i, v);
#pragma unroll
                for (int d = 0; d < D; ++d) put<TO>(grads, i * D + d, gr[d] / (T)F.g.h[d]);
            }
        }
    }
}

__global__ void k_zero_u32(unsigned int* p) { p[0] = 0u; }

// -grad U at x (physical units), with the lane's cell cache.
template <typename T, typename TC>
__device__ __forceinline__ void accel3(const ArbGeom& g, const TC* __restrict__ table,
                                       int32_t ninterp, int32_t kn, const double* x,
                                       int64_t& cached, T* c, double* acc) {
    double loc[3];
    bool oob;
    int64_t flat = arb_reduce<3>(x, g, loc, &oob);
    if (flat > g.nc) flat = g.nc;
    if (flat != cached) {
        const TC* a = table + (flat * ninterp + kn) * 64;
#pragma unroll
        for (int i = 0; i < 64; ++i) c[i] = (T)a[i];
        cached = flat;
    }
    T v, gr[3];
    arb_valgrad3<T, T>(c, (T)loc[0], (T)loc[1], (T)loc[2], v, gr);
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] = -(double)(gr[d] / (T)g.h[d]);
}

template <typename T, typename TC>
__global__ void __launch_bounds__(256)
k_rk4(ArbGeom g, const TC* __restrict__ table, int32_t ninterp, int32_t kn,
      double* __restrict__ X, double* __restrict__ V, int64_t n, double dt, int32_t nsteps) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double x[3], v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) { x[d] = X[p * 3 + d]; v[d] = V[p * 3 + d]; }
    int64_t cached = -1;
    T c[64];
    const double h2 = 0.5 * dt, h6 = dt / 6.0;
    for (int s = 0; s < nsteps; ++s) {
        double a1[3], a2[3], a3[3], a4[3], xs[3], v2[3], v3[3], v4[3];
        accel3<T, TC>(g, table, ninterp, kn, x, cached, c, a1);
#pragma unroll
        for (int d = 0; d < 3; ++d) { xs[d] = x[d] + h2 * v[d]; v2[d] = v[d] + h2 * a1[d]; }
        accel3<T, TC>(g, table, ninterp, kn, xs, cached, c, a2);
#pragma unroll
        for (int d = 0; d < 3; ++d) { xs[d] = x[d] + h2 * v2[d]; v3[d] = v[d] + h2 * a2[d]; }
        accel3<T, TC>(g, table, ninterp, kn, xs, cached, c, a3);
#pragma unroll
        for (int d = 0; d < 3; ++d) { xs[d] = x[d] + dt * v3[d]; v4[d] = v[d] + dt * a3[d]; }
        accel3<T, TC>(g, table, ninterp, kn, xs, cached, c, a4);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[d] += h6 * (v[d] + 2.0 * v2[d] + 2.0 * v3[d] + v4[d]);
            v[d] += h6 * (a1[d] + 2.0 * a2[d] + 2.0 * a3[d] + a4[d]);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) { X[p * 3 + d] = x[d]; V[p * 3 + d] = v[d]; }
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

size_t arb_batch_field_desc_size() { return sizeof(BatchField); }

void arb_batch_fill_desc(void* dst, const ArbGeom& g, const void* table, int64_t q0, int64_t nq,
                         int64_t item0) {
    BatchField* b = static_cast<BatchField*>(dst);
    b->g = g;
    b->table = table;
    b->q0 = q0;
    b->nq = nq;
    b->item0 = item0;
}

hipError_t arb_launch_batch_query(int32_t ndim, int32_t prec, int32_t out_f64, const void* fields,
                                  int32_t nfields, int64_t nitems, int32_t chunk, int32_t threads,
                                  int32_t nblocks, int32_t ninterp, int32_t has_comps,
                                  int32_t has_norm, const double* q, int64_t ldq, void* comps,
                                  void* norms, void* grads, int32_t* flag, unsigned int* counter,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, counter);
    const BatchField* F = static_cast<const BatchField*>(fields);
    dim3 grid(nblocks), block(threads);
#define L_BATCH(DD, T, TC, TO)                                                                 \
    hipLaunchKernelGGL((k_batch_query<DD, T, TC, TO>), grid, block, 0, s, F, nfields, nitems,     \
                       chunk, ninterp, has_comps, has_norm, q, ldq, comps, norms, grads, flag,  \
                       counter)
    if (ndim == 3) {
        if (prec == ARB_PREC_FP64) L_BATCH(3, double, double, double);
        else if (out_f64) L_BATCH(3, float, float, double);
        else L_BATCH(3, float, float, float);
    } else {
        if (prec == ARB_PREC_FP64) L_BATCH(4, double, double, double);
        else if (out_f64) L_BATCH(4, float, float, double);
        else L_BATCH(4, float, float, float);
    }
#undef L_BATCH
    return hipGetLastError();
}

hipError_t arb_launch_rk4(const ArbGeom& g, int32_t prec, const void* table, int32_t ninterp,
                          int32_t kn, double* X, double* V, int64_t n, double dt, int32_t nsteps,
                          hipStream_t s) {
    if (n <= 0 || nsteps <= 0) return hipSuccess;
    dim3 grid(grid1(n, 256)), block(256);
    if (prec == ARB_PREC_FP64)
        hipLaunchKernelGGL((k_rk4<double, double>), grid, block, 0, s, g, (const double*)table,
                           ninterp, kn, X, V, n, dt, nsteps);
    else
        hipLaunchKernelGGL((k_rk4<float, float>), grid, block, 0, s, g, (const float*)table,
                           ninterp, kn, X, V, n, dt, nsteps);
    return hipGetLastError();
}
