// Query evaluation kernels.
//
// Replaces rQuery1/2/3 (ARBInterp.py:344-521; 4D :1064-1258): bounds mask,
// cell reduction, coefficient gather and the power-basis sums for values and
// gradients.  The reference builds ~40 N x 64 temporaries per call; here each
// query is one lane.
#include "arb_internal.h"
#include "arb_sep3.h"

namespace {

template <typename TO, typename T>
__device__ __forceinline__ void put(void* p, int64_t idx, T v) {
    reinterpret_cast<TO*>(p)[idx] = (TO)v;
}

// Evaluate the interpolants of one point from its cell's coefficients `a`
// ([interp][K], in global memory or LDS) and write row i of the outputs.
template <int D, typename T, typename TC, typename TO>
__device__ __forceinline__ void eval_point(const TC* a, int32_t has_comps, int32_t has_norm,
                                           const double* loc, const ArbGeom& g, const ArbOut& o,
                                           int64_t i) {
    constexpr int K = (D == 3) ? 64 : 256;
    T x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = (T)loc[d];
    TO rec[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) rec[r] = (TO)0;
    int k = 0, w = 0;
    if (has_comps) {
#pragma unroll 1
        for (int c = 0; c < 3; ++c) {
            T v;
            if constexpr (D == 3) v = arb_val3<T, TC>(a + c * K, x[0], x[1], x[2]);
            else v = arb_val4<T, TC>(a + c * K, x[0], x[1], x[2], x[3]);
            rec[c] = (TO)v;
        }
        k = 3;
        w = 3;
    }
    if (has_norm) {
        T v, gr[D];
        if constexpr (D == 3) arb_valgrad3<T, TC>(a + k * K, x[0], x[1], x[2], v, gr);
        else arb_valgrad4<T, TC>(a + k * K, x[0], x[1], x[2], x[3], v, gr);
        if (has_comps) {
            rec[3] = (TO)v;
#pragma unroll
            for (int d = 0; d < D; ++d) rec[4 + d] = (TO)(gr[d] / (T)g.h[d]);
        } else {
            rec[0] = (TO)v;
#pragma unroll
            for (int d = 0; d < D; ++d) rec[1 + d] = (TO)(gr[d] / (T)g.h[d]);
        }
        w += 1 + D;
    }
    if (o.packed) {
        // one aligned record per row, written as full 16-byte vectors
        TO* dst = reinterpret_cast<TO*>(o.comps) + i * o.packed;
        if constexpr (sizeof(TO) == 4) {
            float4* d4 = reinterpret_cast<float4*>(dst);
            d4[0] = make_float4(rec[0], rec[1], rec[2], rec[3]);
            if (o.packed == 8) d4[1] = make_float4(rec[4], rec[5], rec[6], rec[7]);
        } else {
            double2* d2 = reinterpret_cast<double2*>(dst);
            d2[0] = make_double2(rec[0], rec[1]);
            d2[1] = make_double2(rec[2], rec[3]);
            if (o.packed == 8) {
                d2[2] = make_double2(rec[4], rec[5]);
                d2[3] = make_double2(rec[6], rec[7]);
            }
        }
        return;
    }
    if (has_comps) {
#pragma unroll
        for (int c = 0; c < 3; ++c) put<TO>(o.comps, i * 3 + c, rec[c]);
    }
    if (has_norm) {
        const int b = has_comps ? 3 : 0;
        put<TO>(o.norms, i, rec[b]);
#pragma unroll
        for (int d = 0; d < D; ++d) put<TO>(o.grads, i * D + d, rec[b + 1 + d]);
    }
    (void)w;
}

// Cell reduction of row i with the side outputs every path writes: queryInds,
// the past-the-sentinel flag and the NaN row fill.  Returns the cell used for
// the coefficients (clamped to the sentinel nc).
template <int D>
__device__ __forceinline__ int64_t reduce_row(const double* __restrict__ q, int64_t ldq,
                                              const ArbGeom& g, const ArbOut& o, int64_t i,
                                              double* loc) {
    bool oob;
    int64_t flat = arb_reduce<D>(q + i * ldq, g, loc, &oob);
    if (o.cells) o.cells[i] = flat;
    if (flat > g.nc) {
        o.flag[0] = 1;
        flat = g.nc;
    }
    if (oob && o.qmut) {
        for (int64_t c = 0; c < o.ldq; ++c) o.qmut[i * o.ldq + c] = NAN;
    }
    return flat;
}

// Lane-per-query kernel: coefficients read straight from the table.
template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(256)
k_query_direct(ArbGeom g, const TC* __restrict__ table, int32_t ninterp, int32_t has_comps,
               int32_t has_norm, const double* __restrict__ q, int64_t n, int64_t ldq, ArbOut o) {
    constexpr int K = (D == 3) ? 64 : 256;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double loc[D];
    const int64_t flat = reduce_row<D>(q, ldq, g, o, i, loc);
    eval_point<D, T, TC, TO>(table + flat * (int64_t)ninterp * K, has_comps, has_norm, loc, g, o, i);
}

// ---------------------------------------------------------------------------
// Cell-bucketed path.  Queries are grouped by the block of cells they fall in
// (k_keys + the radix sort of k_sort.hip by default; with ARB_Q_GROUP_ATOMIC
// a counting sort: atomic rank per bucket, one scan, one scatter of the query
// index).  A workgroup then stages a block's coefficients in LDS with
// contiguous 16-byte loads -- each cell's coefficients leave HBM once per
// call -- and evaluates the bucket's queries from LDS.  Output rows are
// written at the caller's row index, so no unsort pass is needed.  Results do
// not depend on the order inside a bucket (each row is computed alone), so
// every grouping gives bitwise identical outputs.
// ---------------------------------------------------------------------------

// Bucket of a cell and the cell's index inside its block.  The per-axis cell
// coordinates come from the reduction; a query at IntMax can alias into the
// next row/slab (floor(iu) == nPos on an axis while the flat index stays
// below nc, as in the reference), and only then is the flat index
// decomposed.  Block edges are powers of two: shifts and masks only.
template <int D>
__device__ __forceinline__ int64_t bucket_of(int64_t flat, const int64_t* ia, const ArbGeom& g,
                                             const ArbBuckets& B, int* lc) {
    int64_t c[D];
    bool alias = false;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        c[a] = ia[a];
        alias |= ia[a] >= g.npos[a];
    }
    if (alias) {
        int64_t rem = flat;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            c[a] = rem % g.npos[a];
            rem /= g.npos[a];
        }
    }
    int64_t bidx = 0;
    int l = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        bidx += (c[a] >> B.bsh[a]) * B.nbs[a];
        l |= (int)(c[a] & (B.b[a] - 1)) << B.lsh[a];
    }
    *lc = l;
    return bidx;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Reduction + bucket rank.  Rows that reduce to the sentinel (out of range,
// NaN, or past the sentinel) are evaluated here from the sentinel row.
template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(256)
k_bucket(ArbGeom g, ArbBuckets B, const TC* __restrict__ table, int32_t ninterp, int32_t has_comps,
         int32_t has_norm, const double* __restrict__ q, int64_t n, int64_t ldq, ArbOut o,
         int32_t* __restrict__ bkt, int32_t* __restrict__ rank, int32_t* __restrict__ count) {
    constexpr int K = (D == 3) ? 64 : 256;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double loc[D];
    int64_t ia[D];
    bool oob;
    int64_t flat = arb_reduce_axes<D>(q + i * ldq, g, loc, ia, &oob);
    if (o.cells) o.cells[i] = flat;
    if (oob && o.qmut) {
        for (int64_t c = 0; c < o.ldq; ++c) o.qmut[i * o.ldq + c] = NAN;
    }
    if (flat >= g.nc) {
        if (flat > g.nc) o.flag[0] = 1;
        eval_point<D, T, TC, TO>(table + g.nc * (int64_t)ninterp * K, has_comps, has_norm, loc, g,
                                 o, i);
        bkt[i] = -1;
        return;
    }
    int lc;
    const int64_t b = bucket_of<D>(flat, ia, g, B, &lc);
    bkt[i] = (int32_t)b;
    rank[i] = atomicAdd(&count[b], 1);
}

// Reduction + bucket key for the radix-sorted path (no atomics): key = bucket,
// or nbt for rows evaluated here from the sentinel row (the sort's value is
// the row index, implicit).
// One workgroup of KEYS_THREADS threads per radix tile (ARB_RS_THREADS x
// ARB_RS_ITEMS rows): it also counts the low 8 bits of its keys in LDS and
// writes the tile's histogram, which is the first radix pass's upsweep.
// Positions are loaded KEYS_BATCH rows at a time so their loads are in flight
// together.
constexpr int KEYS_THREADS = 256;
constexpr int KEYS_ITEMS = ARB_RS_THREADS * ARB_RS_ITEMS / KEYS_THREADS;
constexpr int KEYS_BATCH = 4;
static_assert(KEYS_ITEMS % KEYS_BATCH == 0, "rows per thread must be a multiple of the batch");

template <int D, typename T, typename TC, typename TO>
__device__ __forceinline__ uint32_t key_of_row(const ArbGeom& g, const ArbBuckets& B,
                                               const TC* __restrict__ table, int32_t ninterp,
                                               int32_t has_comps, int32_t has_norm,
                                               const ArbOut& o, int64_t i, const double* qp,
                                               int32_t cellbits) {
    constexpr int K = (D == 3) ? 64 : 256;
    double loc[D];
    int64_t ia[D];
    bool oob;
    const int64_t flat = arb_reduce_axes<D>(qp, g, loc, ia, &oob);
    if (o.cells) o.cells[i] = flat;
    if (oob && o.qmut) {
        for (int64_t c = 0; c < o.ldq; ++c) o.qmut[i * o.ldq + c] = NAN;
    }
    if (flat >= g.nc) {
        if (flat > g.nc) o.flag[0] = 1;
        eval_point<D, T, TC, TO>(table + g.nc * (int64_t)ninterp * K, has_comps, has_norm, loc, g,
                                 o, i);
        return (uint32_t)B.nbt << cellbits;
    }
    int lc;
    const uint32_t b = (uint32_t)bucket_of<D>(flat, ia, g, B, &lc);
    return (b << cellbits) | (cellbits ? (uint32_t)lc : 0u);
}

template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(KEYS_THREADS)
k_keys(ArbGeom g, ArbBuckets B, const TC* __restrict__ table, int32_t ninterp, int32_t has_comps,
       int32_t has_norm, const double* __restrict__ q, int64_t n, int64_t ldq, ArbOut o,
       uint32_t* __restrict__ key, int32_t cellbits, int32_t* __restrict__ hist,
       int64_t ntiles) {
    __shared__ int32_t cnt[256];
    cnt[threadIdx.x] = 0;                           // KEYS_THREADS == 256 == digits
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * (KEYS_THREADS * KEYS_ITEMS) + threadIdx.x;
    for (int j0 = 0; j0 < KEYS_ITEMS; j0 += KEYS_BATCH) {
        double qv[KEYS_BATCH][D];
#pragma unroll
        for (int u = 0; u < KEYS_BATCH; ++u) {
            const int64_t i = base + (int64_t)(j0 + u) * KEYS_THREADS;
            const int64_t r = i < n ? i : 0;        // row 0 exists: n > 0
#pragma unroll
            for (int a = 0; a < D; ++a) qv[u][a] = q[r * ldq + a];
        }
#pragma unroll
        for (int u = 0; u < KEYS_BATCH; ++u) {
            const int64_t i = base + (int64_t)(j0 + u) * KEYS_THREADS;
            if (i < n) {
                const uint32_t k = key_of_row<D, T, TC, TO>(g, B, table, ninterp, has_comps,
                                                            has_norm, o, i, qv[u], cellbits);
                key[i] = k;
                atomicAdd(&cnt[k & 255u], 1);
            }
        }
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of nb counts into offs[0..nb] with one workgroup of 1024
// threads: each thread keeps 64 consecutive counts in registers per pass, so
// a pass costs one round of independent loads and one of stores.
__global__ void __launch_bounds__(1024) k_scan(const int32_t* __restrict__ count,
                                               int32_t* __restrict__ offs, int64_t nb) {
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    int32_t carry = 0;
    for (int64_t base = 0; base < nb; base += 1024 * 64) {
        const int64_t i0 = base + (int64_t)t * 64;
        int32_t v[64];
        int32_t sum = 0;
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            v[k] = (i0 + k < nb) ? count[i0 + k] : 0;
            sum += v[k];
        }
        part[t] = sum;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int32_t x = (t >= off) ? part[t - off] : 0;
            __syncthreads();
            part[t] += x;
            __syncthreads();
        }
        int32_t run = carry + part[t] - sum;
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            if (i0 + k < nb) offs[i0 + k] = run;
            run += v[k];
        }
        carry += part[1023];
        __syncthreads();
    }
    if (t == 0) offs[nb] = carry;
}

// Scatter: query index into bucket order.
__global__ void k_scatter(const int32_t* __restrict__ bkt, const int32_t* __restrict__ rank,
                          const int32_t* __restrict__ offs, int32_t* __restrict__ perm, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bkt[i];
    if (b < 0) return;
    perm[offs[b] + rank[i]] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// Staging map.  Slot v (0 <= v < ncell * cellvec) of a bucket holds 16-byte
// vector j of local cell lc (v = lc * cellvec + j); in LDS it sits at
// lc * cellstride + j.  Its table source is the bucket's first cell plus
// sum_a la_a * cs_a cells, la_a the local coordinates of lc; cs_a, the cell
// stride of axis a, fits 32 bits (fewer than 2^32 cells, checked when the
// field is created), so the per-slot arithmetic is 32-bit multiply-adds.
// ---------------------------------------------------------------------------
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

template <int D>
struct StageBase {
    uint64_t vbase;          // first cell of the bucket, in 16-byte vectors
    int32_t ext[D];          // cells of the bucket inside the grid, per axis
    bool edge;               // some axis has ext < block edge
};

template <int D>
__device__ __forceinline__ void stage_base(int64_t b, const ArbGeom& g, const ArbBuckets& B,
                                           StageBase<D>& sb) {
    int64_t rem = b;
    uint64_t cell = 0, st = 1;
    bool edge = false;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const int64_t bi = rem % B.nb[a];
        rem /= B.nb[a];
        const int64_t c0 = bi << B.bsh[a];
        const int64_t left = g.npos[a] - c0;
        sb.ext[a] = (int32_t)(left < B.b[a] ? left : B.b[a]);
        edge |= sb.ext[a] < B.b[a];
        cell += (uint64_t)c0 * st;
        st *= (uint64_t)g.npos[a];
    }
    sb.edge = edge;
    sb.vbase = cell * (uint64_t)B.cellvec;
}

template <int D>
__device__ __forceinline__ void cell_strides32(const ArbGeom& g, uint32_t* cs) {
    uint32_t st = 1;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        cs[a] = st;
        st *= (uint32_t)g.npos[a];
    }
}

// Table source of slot v (vector index; 0, a valid dummy, for a slot past the
// block or of a cell outside the grid) and its LDS slot (the padding vector
// of cell 0 for a slot past the block).
template <int D>
__device__ __forceinline__ uint64_t stage_slot(int v, int total, const ArbBuckets& B,
                                               const uint32_t* cs, const StageBase<D>& sb,
                                               int* dst) {
    const int lc = B.cvsh >= 0 ? (v >> B.cvsh) : v / B.cellvec;
    const int j = v - lc * B.cellvec;
    bool inside = v < total;
    uint64_t off = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const uint32_t la = (uint32_t)(lc >> B.lsh[a]) & (uint32_t)(B.b[a] - 1);
        if (sb.edge) inside &= (int32_t)la < sb.ext[a];
        off += (uint64_t)la * cs[a];
    }
    *dst = v < total ? lc * B.cellstride + j : B.cellvec;
    return inside ? sb.vbase + off * (uint32_t)B.cellvec + (uint32_t)j : 0;
}

// Evaluate row qi (position qp) of bucket b from the block staged in LDS.
template <int D, typename T, typename TC, typename TO>
__device__ __forceinline__ void eval_row_lds(const ArbGeom& g, const ArbBuckets& B,
                                             const uint4* smem, int32_t has_comps,
                                             int32_t has_norm, const ArbOut& o, int64_t b,
                                             int64_t qi, const double* qp) {
    double loc[D];
    int64_t ia[D];
    bool oob;
    const int64_t flat = arb_reduce_axes<D>(qp, g, loc, ia, &oob);
    int lc;
    const int64_t bb = bucket_of<D>(flat, ia, g, B, &lc);
    if (bb != b || flat >= g.nc) return;            // cannot happen: same reduction as k_keys
    const TC* a = reinterpret_cast<const TC*>(smem + lc * B.cellstride);
    eval_point<D, T, TC, TO>(a, has_comps, has_norm, loc, g, o, qi);
}

// One workgroup per bucket: stage the block's coefficients in LDS, then
// evaluate the bucket's queries from LDS.
template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(256)
k_eval_bucket(ArbGeom g, ArbBuckets B, const TC* __restrict__ table, int32_t has_comps,
              int32_t has_norm, const double* __restrict__ q, int64_t n, int64_t ldq, ArbOut o,
              const int32_t* __restrict__ offs, const int32_t* __restrict__ perm) {
    extern __shared__ uint4 smem[];
    const int64_t b = blockIdx.x;
    const int32_t start = offs[b], end = offs[b + 1];
    if (start == end) return;                       // uniform over the workgroup
    StageBase<D> sb;
    stage_base<D>(b, g, B, sb);
    uint32_t cs[D];
    cell_strides32<D>(g, cs);
    const v4u* __restrict__ tv = reinterpret_cast<const v4u*>(table);
    v4u* lds = reinterpret_cast<v4u*>(smem);
    const int total = B.ncell * B.cellvec;
    // this thread's first query: issue its loads before the staging so the
    // latencies overlap
    int32_t s = start + threadIdx.x;
    int64_t qi = -1;
    double qp[D];
    if (s < end) {
        const int64_t pi = perm[s];
        qi = (pi >= 0 && pi < n) ? pi : -1;         // always in range: perm holds row indices
        const int64_t qq = qi >= 0 ? qi : 0;
#pragma unroll
        for (int a = 0; a < D; ++a) qp[a] = q[qq * ldq + a];
    }
    // stage the block: 8 independent 16-byte loads in flight per thread.  Loads
    // and LDS stores are unconditional so that they issue back to back.  The
    // table is read once per call, so the loads are nontemporal: the stream
    // does not evict the positions, which the queries read back at random
    // (the 128^3 mixed eval: 0.89 -> 0.80 ms).
    constexpr int BATCH = 8;
    for (int v0 = threadIdx.x; v0 < total; v0 += BATCH * blockDim.x) {
        v4u r[BATCH];
        int dst[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u)
            r[u] = __builtin_nontemporal_load(
                &tv[stage_slot<D>(v0 + u * blockDim.x, total, B, cs, sb, &dst[u])]);
#pragma unroll
        for (int u = 0; u < BATCH; ++u) lds[dst[u]] = r[u];
    }
    __syncthreads();
    if (qi >= 0) eval_row_lds<D, T, TC, TO>(g, B, smem, has_comps, has_norm, o, b, qi, qp);
    for (s += blockDim.x; s < end; s += blockDim.x) {
        const int64_t pi = perm[s];
        if (pi < 0 || pi >= n) continue;            // unreachable: perm holds row indices
#pragma unroll
        for (int a = 0; a < D; ++a) qp[a] = q[pi * ldq + a];
        eval_row_lds<D, T, TC, TO>(g, B, smem, has_comps, has_norm, o, b, pi, qp);
    }
}

// 3D without the coefficient table: one workgroup per bucket forms its
// cells' coefficients from node values (per interpolant: node tile, cells + 3
// per axis, into LDS, then the separable passes of arb_sep3.h, the routine
// the table build uses, so the results are bitwise those of the table),
// writes them into the same LDS block layout the staged kernel uses, and
// evaluates the bucket's queries.  The node tiles of all interpolants are
// loaded into registers first (FUSED_REGS per thread, all loads in flight at
// once), so the passes never wait on global memory.  Dynamic LDS: ncell *
// cellstride * 16 bytes of coefficients, then one node tile and the pass
// buffers (doubles).
constexpr int FUSED_REGS = 8;

template <typename T, typename TO, int BX, int BY, int BZ>
__global__ void __launch_bounds__(256)
k_eval_fused3(ArbGeom g, ArbBuckets B, const double* __restrict__ vals, int32_t ninterp,
              int32_t has_comps, int32_t has_norm, const double* __restrict__ q, int64_t n,
              int64_t ldq, ArbOut o, const int32_t* __restrict__ offs,
              const int32_t* __restrict__ perm) {
    extern __shared__ uint4 smem[];
    const int64_t b = blockIdx.x;
    const int32_t start = offs[b], end = offs[b + 1];
    if (start == end) return;                       // uniform over the workgroup
    int64_t c0[3];
    {
        int64_t rem = b;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            c0[a] = (rem % B.nb[a]) << B.bsh[a];
            rem /= B.nb[a];
        }
    }
    constexpr int NX = BX + 3, NY = BY + 3, NT = NX * NY * (BZ + 3);
    double* nodes = reinterpret_cast<double*>(smem + B.ncell * B.cellstride);
    double* t1 = nodes + NT;
    double* t2 = t1 + arbs3::sep3_t1(BX, BY, BZ);
    T* cf = reinterpret_cast<T*>(smem);
    const int cstr = B.cellstride * (16 / (int)sizeof(T));   // elements per LDS cell
    // this thread's first query, loaded before the coefficient work
    int32_t s = start + threadIdx.x;
    int64_t qi = -1;
    double qp[3];
    if (s < end) {
        const int64_t pi = perm[s];
        qi = (pi >= 0 && pi < n) ? pi : -1;         // always in range: perm holds row indices
        const int64_t qq = qi >= 0 ? qi : 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) qp[a] = q[qq * ldq + a];
    }
    // node tiles of every interpolant, x fastest: element e = k * NT + r of
    // the flattened [interp][tile] space sits in register j of thread
    // e - j * blockDim.x (the launcher checks ninterp * NT <= FUSED_REGS *
    // blockDim.x).  Nodes past the grid edge read as 0 (only cells outside
    // the grid use them); the loads are unconditional (index 0 when unused).
    double nv[FUSED_REGS];
    const int total = ninterp * NT;
#pragma unroll
    for (int j = 0; j < FUSED_REGS; ++j) {
        const int e = threadIdx.x + j * blockDim.x;
        const int k = e / NT, r = e - k * NT;
        const int x = r % NX, y = (r / NX) % NY, z = r / (NX * NY);
        const int64_t nx = c0[0] + x, ny = c0[1] + y, nz = c0[2] + z;
        const bool in = e < total && nx < g.n[0] && ny < g.n[1] && nz < g.n[2];
        const int64_t src = in ? (int64_t)k * g.nnodes + nx + g.nstride[1] * ny + g.nstride[2] * nz
                               : 0;
        const double v = vals[src];
        nv[j] = in ? v : 0.0;
    }
    for (int k = 0; k < ninterp; ++k) {
        // tile k into LDS: the x pass of interpolant k - 1, the last reader
        // of `nodes`, ended at that interpolant's first pass barrier
#pragma unroll
        for (int j = 0; j < FUSED_REGS; ++j) {
            const int e = threadIdx.x + j * blockDim.x - k * NT;
            if (e >= 0 && e < NT) nodes[e] = nv[j];
        }
        __syncthreads();
        // consecutive threads, consecutive cells: local cell cx + BX (cy + BY cz)
        arbs3::sep3_passes<BX, BY, BZ, true>(nodes, t1, t2,
                                             [&](int cx, int cy, int cz, int i, double a) {
            cf[(cx + BX * (cy + BY * cz)) * cstr + k * 64 + i] = (T)a;
        });
    }
    __syncthreads();
    if (qi >= 0) eval_row_lds<3, T, T, TO>(g, B, smem, has_comps, has_norm, o, b, qi, qp);
    for (s += blockDim.x; s < end; s += blockDim.x) {
        const int64_t pi = perm[s];
        if (pi < 0 || pi >= n) continue;            // unreachable: perm holds row indices
#pragma unroll
        for (int a = 0; a < 3; ++a) qp[a] = q[pi * ldq + a];
        eval_row_lds<3, T, T, TO>(g, B, smem, has_comps, has_norm, o, b, pi, qp);
    }
}

// Sorted order, coefficients read from the table in global memory.
template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(256)
k_eval_perm(ArbGeom g, const TC* __restrict__ table, int32_t ninterp, int32_t has_comps,
            int32_t has_norm, const double* __restrict__ q, int64_t n, int64_t ldq, ArbOut o,
            const int32_t* __restrict__ total, const int32_t* __restrict__ perm) {
    constexpr int K = (D == 3) ? 64 : 256;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || s >= total[0]) return;
    const int64_t i = perm[s];
    if (i < 0 || i >= n) return;                    // unreachable: perm holds row indices
    double loc[D];
    bool oob;
    const int64_t flat = arb_reduce<D>(q + i * ldq, g, loc, &oob);
    if (flat >= g.nc) return;
    eval_point<D, T, TC, TO>(table + flat * (int64_t)ninterp * K, has_comps, has_norm, loc, g, o, i);
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Dispatch over (ndim, precision, output type).
#define ARB_DISPATCH(D, PREC, OUTF64, LAUNCH)                                   \
    do {                                                                        \
        if (PREC == ARB_PREC_FP64) { LAUNCH(D, double, double, double); }       \
        else if (OUTF64) { LAUNCH(D, float, float, double); }                   \
        else { LAUNCH(D, float, float, float); }                                \
    } while (0)

template <int D>
hipError_t launch_direct(const ArbGeom& g, int32_t prec, int32_t out_f64, const void* table,
                         int32_t ninterp, int32_t has_comps, int32_t has_norm, const double* q,
                         int64_t n, int64_t ldq, const ArbOut& o, hipStream_t s) {
    dim3 grid(grid1(n, 256)), block(256);
    if (prec == ARB_PREC_FP64) {
        hipLaunchKernelGGL((k_query_direct<D, double, double, double>), grid, block, 0, s, g,
                           (const double*)table, ninterp, has_comps, has_norm, q, n, ldq, o);
    } else if (out_f64) {
        hipLaunchKernelGGL((k_query_direct<D, float, float, double>), grid, block, 0, s, g,
                           (const float*)table, ninterp, has_comps, has_norm, q, n, ldq, o);
    } else {
        hipLaunchKernelGGL((k_query_direct<D, float, float, float>), grid, block, 0, s, g,
                           (const float*)table, ninterp, has_comps, has_norm, q, n, ldq, o);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t arb_launch_query_direct(const ArbGeom& g, int32_t prec, int32_t out_f64,
                                   const void* table, int32_t ninterp, int32_t has_comps,
                                   int32_t has_norm, const double* q, int64_t n, int64_t ldq,
                                   const ArbOut& o, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (g.ndim == 3)
        return launch_direct<3>(g, prec, out_f64, table, ninterp, has_comps, has_norm, q, n, ldq,
                                o, s);
    return launch_direct<4>(g, prec, out_f64, table, ninterp, has_comps, has_norm, q, n, ldq, o,
                            s);
}

hipError_t arb_launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid1(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t arb_launch_bucket(const ArbGeom& g, const ArbBuckets& B, int32_t prec, int32_t out_f64,
                             const void* table, int32_t ninterp, int32_t has_comps,
                             int32_t has_norm, const double* q, int64_t n, int64_t ldq,
                             const ArbOut& o, int32_t* bkt, int32_t* rank, int32_t* count,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define L_BUCKET(DD, T, TC, TO)                                                              \
    hipLaunchKernelGGL((k_bucket<DD, T, TC, TO>), dim3(grid1(n, 256)), dim3(256), 0, s, g, B,   \
                       (const TC*)table, ninterp, has_comps, has_norm, q, n, ldq, o, bkt, rank, \
                       count)
    if (g.ndim == 3) ARB_DISPATCH(3, prec, out_f64, L_BUCKET);
    else ARB_DISPATCH(4, prec, out_f64, L_BUCKET);
#undef L_BUCKET
    return hipGetLastError();
}

hipError_t arb_launch_scan(const int32_t* count, int32_t* offs, int64_t nb, hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, count, offs, nb);
    return hipGetLastError();
}

hipError_t arb_launch_scatter(const int32_t* bkt, const int32_t* rank, const int32_t* offs,
                              int32_t* perm, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3(grid1(n, 256)), dim3(256), 0, s, bkt, rank, offs, perm, n);
    return hipGetLastError();
}

hipError_t arb_launch_eval_bucket(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                  int32_t out_f64, const void* table, int32_t ninterp,
                                  int32_t has_comps, int32_t has_norm, const double* q,
                                  int64_t n, int64_t ldq, const ArbOut& o, const int32_t* offs,
                                  const int32_t* perm, int32_t block_threads, hipStream_t s) {
    (void)ninterp;
    const size_t lds = (size_t)B.ncell * B.cellstride * 16;
    dim3 grid((unsigned)B.nbt), block(block_threads);
#define L_EVALB(DD, T, TC, TO)                                                              \
    hipLaunchKernelGGL((k_eval_bucket<DD, T, TC, TO>), grid, block, lds, s, g, B,             \
                       (const TC*)table, has_comps, has_norm, q, n, ldq, o, offs, perm)
    if (g.ndim == 3) ARB_DISPATCH(3, prec, out_f64, L_EVALB);
    else ARB_DISPATCH(4, prec, out_f64, L_EVALB);
#undef L_EVALB
    return hipGetLastError();
}

// Cell blocks k_eval_fused3 is instantiated for (the ladder the bucket
// geometry is chosen from).
#define ARB_FUSED_SHAPES(X) X(4, 4, 4) X(4, 4, 2) X(4, 2, 2) X(2, 2, 2) X(2, 2, 1) X(2, 1, 1) X(1, 1, 1)

// LDS bytes of k_eval_fused3 for the bucket geometry B (0 when the path does
// not apply: 3D only, an instantiated block shape, within 160 KiB, node tiles
// of all interpolants within the registers of `threads` threads).
size_t arb_fused_lds(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp, int32_t threads) {
    if (g.ndim != 3) return 0;
    const int bx = B.b[0], by = B.b[1], bz = B.b[2];
    bool shape = false;
#define ARB_FUSED_IS(X, Y, Z) shape |= (bx == X && by == Y && bz == Z);
    ARB_FUSED_SHAPES(ARB_FUSED_IS)
#undef ARB_FUSED_IS
    if (!shape) return 0;
    // the local cell index must be cx + bx (cy + by cz), as the kernel writes it
    if (B.lsh[0] != 0 || (1 << B.lsh[1]) != bx || (1 << B.lsh[2]) != bx * by) return 0;
    if ((int64_t)ninterp * arbs3::sep3_nodes(bx, by, bz) > (int64_t)FUSED_REGS * threads) return 0;
    const size_t lds = (size_t)B.ncell * B.cellstride * 16 +
                       sizeof(double) * (size_t)(arbs3::sep3_nodes(bx, by, bz) +
                                                 arbs3::sep3_t1(bx, by, bz) +
                                                 arbs3::sep3_t2(bx, by, bz));
    return lds <= 160 * 1024 ? lds : 0;            // gfx950: 160 KiB of LDS per workgroup
}

hipError_t arb_launch_eval_fused(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                 int32_t out_f64, const double* vals, int32_t ninterp,
                                 int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                 int64_t ldq, const ArbOut& o, const int32_t* offs,
                                 const int32_t* perm, int32_t block_threads, hipStream_t s) {
    const size_t lds = arb_fused_lds(g, B, ninterp, block_threads);
    if (lds == 0) return hipErrorInvalidValue;
    dim3 grid((unsigned)B.nbt), block(block_threads);
#define L_FUSED(T, TO, X, Y, Z)                                                               \
    hipLaunchKernelGGL((k_eval_fused3<T, TO, X, Y, Z>), grid, block, lds, s, g, B, vals,      \
                       ninterp, has_comps, has_norm, q, n, ldq, o, offs, perm)
#define ARB_FUSED_CASE(X, Y, Z)                                                               \
    if (B.b[0] == X && B.b[1] == Y && B.b[2] == Z) {                                          \
        if (prec == ARB_PREC_FP64) L_FUSED(double, double, X, Y, Z);                          \
        else if (out_f64) L_FUSED(float, double, X, Y, Z);                                    \
        else L_FUSED(float, float, X, Y, Z);                                                  \
        return hipGetLastError();                                                             \
    }
    ARB_FUSED_SHAPES(ARB_FUSED_CASE)
#undef ARB_FUSED_CASE
#undef L_FUSED
    return hipErrorInvalidValue;
}

hipError_t arb_launch_eval_perm(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                int32_t out_f64, const void* table, int32_t ninterp,
                                int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                int64_t ldq, const ArbOut& o, const int32_t* offs,
                                const int32_t* perm, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int32_t* total = offs + B.nbt;
#define L_EVALP(DD, T, TC, TO)                                                              \
    hipLaunchKernelGGL((k_eval_perm<DD, T, TC, TO>), dim3(grid1(n, 256)), dim3(256), 0, s, g,  \
                       (const TC*)table, ninterp, has_comps, has_norm, q, n, ldq, o, total, perm)
    if (g.ndim == 3) ARB_DISPATCH(3, prec, out_f64, L_EVALP);
    else ARB_DISPATCH(4, prec, out_f64, L_EVALP);
#undef L_EVALP
    return hipGetLastError();
}

hipError_t arb_launch_keys(const ArbGeom& g, const ArbBuckets& B, int32_t prec, int32_t out_f64,
                           const void* table, int32_t ninterp, int32_t has_comps,
                           int32_t has_norm, const double* q, int64_t n, int64_t ldq,
                           const ArbOut& o, uint32_t* key, int32_t cellbits, int32_t* hist,
                           hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t ntiles = arb_rs_ntiles(n);
#define L_KEYS(DD, T, TC, TO)                                                                 \
    hipLaunchKernelGGL((k_keys<DD, T, TC, TO>), dim3((unsigned)ntiles), dim3(KEYS_THREADS), 0,   \
                       s, g, B, (const TC*)table, ninterp, has_comps, has_norm, q, n, ldq, o,    \
                       key, cellbits, hist, ntiles)
    if (g.ndim == 3) ARB_DISPATCH(3, prec, out_f64, L_KEYS);
    else ARB_DISPATCH(4, prec, out_f64, L_KEYS);
#undef L_KEYS
    return hipGetLastError();
}
