// Query evaluation kernels.
//
// Replaces rQuery1/2/3 (ARBInterp.py:344-521; 4D :1064-1258): bounds mask,
// cell reduction, coefficient gather and the power-basis sums for values and
// gradients.  The reference builds ~40 N x 64 temporaries per call; here each
// query is one lane.
#include "arb_internal.h"
#include "arb_sep4.h"

namespace {

template <typename TO, typename T>
__device__ __forceinline__ void put(void* p, int64_t idx, T v) {
    reinterpret_cast<TO*>(p)[idx] = (TO)v;
}

// Write row i of the outputs from its record [comps(3)] [value, grad(D)]
// (rec[0] = value when there are no components).
template <int D, typename TO>
__device__ __forceinline__ void store_rec(const TO* rec, int32_t has_comps, int32_t has_norm,
                                          const ArbOut& o, int64_t i) {
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
    int k = 0;
    if (has_comps) {
#pragma unroll 1
        for (int c = 0; c < 3; ++c) {
            T v;
            if constexpr (D == 3) v = arb_val3<T, TC>(a + c * K, x[0], x[1], x[2]);
            else v = arb_val4<T, TC>(a + c * K, x[0], x[1], x[2], x[3]);
            rec[c] = (TO)v;
        }
        k = 3;
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
    }
    store_rec<D, TO>(rec, has_comps, has_norm, o, i);
}

// A row whose local coordinates hold a NaN (outside the box, or a NaN
// position) evaluates to NaN in every output whatever the coefficients:
// Horner's rule in a NaN coordinate gives NaN for the value and every
// derivative.  Written here without evaluating; the record's padding stays
// zero, as eval_point leaves it.
template <int D, typename TO>
__device__ __forceinline__ void store_nan_row(int32_t has_comps, int32_t has_norm,
                                              const ArbOut& o, int64_t i) {
    const int used = (has_comps ? 3 : 0) + (has_norm ? 1 + D : 0);
    TO rec[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) rec[r] = r < used ? (TO)NAN : (TO)0;
    store_rec<D, TO>(rec, has_comps, has_norm, o, i);
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
// Float32 fields queried with float32 positions.  numpy then reduces in
// float32 (arb_reduce with g.red32) and forms the power basis from the
// float32 local coordinate with float32 powers before the FP64 sums
// (ARBInterp.py:493-503: queryCoords**2, **3 and 3*queryCoords**2 are
// float32 arrays): X = (1, u, fl32(u^2), fl32(u^3)), dX = (0, 1, fl32(2u),
// fl32(3 fl32(u^2))).  The cube is the float32 nearest to u^3 (numpy's
// vectorised float32 power is itself within an ulp of it, not always equal).
// Values are then sum_ijk alpha_ijk X_i Y_j Z_k in FP64, contracted x first.
// ---------------------------------------------------------------------------
__device__ float cube_f32_nearest(float u) {
    const double u2 = (double)u * (double)u;    // exact: 48 significant bits
    const double c = u2 * (double)u;            // u^3 rounded to FP64
    const double e = fma(u2, (double)u, -c);    // u^3 - c exactly
    float r = (float)c;
    const double dr = (double)r;
    if (e != 0.0 && c != dr) {
        // c lies exactly between r and its neighbour: the remainder decides
        const float nb = c > dr ? nextafterf(r, INFINITY) : nextafterf(r, -INFINITY);
        if (c - dr == (double)nb - c) r = ((e > 0.0) == (nb > r)) ? nb : r;
    }
    return r;
}

__device__ __forceinline__ void basis_f32(double loc, double* X, double* dX) {
    const float u = (float)loc;                 // loc holds a float32 value
    const float u2 = u * u;
    X[0] = 1.0;
    X[1] = (double)u;
    X[2] = (double)u2;
    X[3] = (double)cube_f32_nearest(u);
    dX[0] = 0.0;
    dX[1] = 1.0;
    dX[2] = (double)(2.0f * u);
    dX[3] = (double)(3.0f * u2);
}

// sum over the K coefficients a of a_i * prod_axis b[axis][e_axis]
template <int D, typename TC>
__device__ __forceinline__ double contract_basis(const TC* a, const double* const* b) {
    double out = 0.0;
#pragma unroll 1
    for (int l = 0; l < (D == 4 ? 4 : 1); ++l) {
        double z = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double y = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const TC* c = a + 64 * l + 16 * k + 4 * j;
                double x = 0.0;
#pragma unroll
                for (int i = 0; i < 4; ++i) x = fma((double)c[i], b[0][i], x);
                y = fma(x, b[1][j], y);
            }
            z = fma(y, b[2][k], z);
        }
        out = D == 4 ? fma(z, b[3][l], out) : z;
    }
    return out;
}

template <int D, typename TC, typename TO>
__global__ void __launch_bounds__(256)
k_query_direct_red32(ArbGeom g, const TC* __restrict__ table, int32_t ninterp,
                     int32_t has_comps, int32_t has_norm, const double* __restrict__ q,
                     int64_t n, int64_t ldq, ArbOut o) {
    constexpr int K = (D == 3) ? 64 : 256;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double loc[D];
    const int64_t flat = reduce_row<D>(q, ldq, g, o, i, loc);
    const TC* a = table + flat * (int64_t)ninterp * K;
    double X[D][4], dX[D][4];
#pragma unroll
    for (int d = 0; d < D; ++d) basis_f32(loc[d], X[d], dX[d]);
    const double* b[4] = {X[0], X[1], X[2], D == 4 ? X[D - 1] : X[0]};
    TO rec[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) rec[r] = (TO)0;
    int k = 0;
    if (has_comps) {
        for (int c = 0; c < 3; ++c) rec[c] = (TO)contract_basis<D, TC>(a + c * K, b);
        k = 3;
    }
    if (has_norm) {
        const int r0 = has_comps ? 3 : 0;
        rec[r0] = (TO)contract_basis<D, TC>(a + k * K, b);
        for (int d = 0; d < D; ++d) {
            const double* bd[4] = {b[0], b[1], b[2], b[3]};
            bd[d] = dX[d];
            rec[r0 + 1 + d] = (TO)(contract_basis<D, TC>(a + k * K, bd) / g.h[d]);
        }
    }
    store_rec<D, TO>(rec, has_comps, has_norm, o, i);
}

// ---------------------------------------------------------------------------
// Cell-bucketed path.  Queries are grouped by the block of cells they fall in
// (k_keys + the grouping of k_sort.hip by default: high key digit, then low
// digit inside each segment, no order inside a bucket; its LSD radix sort
// past 2^24 rows or 16 key bits, or with the cell in the key; with
// ARB_Q_GROUP_ATOMIC a counting sort: atomic rank per bucket, one scan, one
// scatter of the query index).  A workgroup then stages a block's coefficients in LDS with
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

// Cell-local coordinates of the node-tile paths, as every one of them (and
// the per-row records of the bucketed path) carries them: the FP32 value
// per axis, which the Hermite weights use, plus an 8-bit correction per
// axis in units of 2^-32, so that the FP64 model sees the FP64 coordinate
// to within 2^-32 (the FP32 rounding alone would move the point by up to
// 2^-25 of a cell, too much near a zero of a fast-varying scalar).  The
// corrections share a 32-bit word with the local cell index (8 bits: up to
// 256 cells per bucket).
__device__ __forceinline__ uint32_t loc_encode(const double* loc, float* hi) {
    uint32_t packed = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        hi[a] = (float)loc[a];
        int k = (int)rint((loc[a] - (double)hi[a]) * 4294967296.0);
        k = k < -128 ? -128 : (k > 127 ? 127 : k);
        packed |= (uint32_t)(k & 255) << (8 * a + 8);
    }
    return packed;
}

__device__ __forceinline__ void loc_decode(const float* hi, uint32_t packed, double* loc) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
        loc[a] = (double)hi[a] + (double)(int8_t)((packed >> (8 * a + 8)) & 255u) * 0x1p-32;
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

// Reduction + bucket key for the sorted path (the grouping or the radix sort;
// no atomics): key = bucket,
// or nbt for rows evaluated here from the sentinel row (the sort's value is
// the row index, implicit).
// One workgroup of KEYS_THREADS threads per radix tile (ARB_RS_THREADS x
// ARB_RS_ITEMS rows): it also counts the 8-bit digit (key >> hshift) & 255
// of its keys in LDS and writes the tile's histogram, which is the first
// sort pass's upsweep.
// Positions are loaded KEYS_BATCH rows at a time so their loads are in flight
// together.
constexpr int KEYS_THREADS = 256;
constexpr int KEYS_ITEMS = ARB_RS_THREADS * ARB_RS_ITEMS / KEYS_THREADS;
constexpr int KEYS_BATCH = 4;
static_assert(KEYS_ITEMS * KEYS_THREADS == ARB_RS_THREADS * ARB_RS_ITEMS, "a tile per workgroup");
static_assert(KEYS_THREADS == 256, "a thread per digit");
static_assert(KEYS_ITEMS % KEYS_BATCH == 0 && KEYS_ITEMS <= 32,
              "rows per thread: a multiple of the batch, one bit each in a 32-bit mask");

// The row's record for the node-tile evaluation (3D): its local coordinates
// (FP32 and corrections) and its local cell in the bucket.  *has = false for
// a row that reduces to the sentinel (no record); *sent_row = true for such a
// row, which the caller evaluates from the sentinel row afterwards, and
// *nan_loc = true when its local coordinates hold a NaN (its outputs are NaN).
template <int D>
__device__ __forceinline__ uint32_t key_of_row(const ArbGeom& g, const ArbBuckets& B,
                                               const ArbOut& o, int64_t i, const double* qp,
                                               int32_t cellbits, float4* rec, bool* has,
                                               bool* sent_row, bool* nan_loc) {
    double loc[D];
    int64_t ia[D];
    bool oob;
    const int64_t flat = arb_reduce_axes<D>(qp, g, loc, ia, &oob);
    *has = false;
    *sent_row = false;
    *nan_loc = false;
    if (o.cells) o.cells[i] = flat;
    if (oob && o.qmut) {
        for (int64_t c = 0; c < o.ldq; ++c) o.qmut[i * o.ldq + c] = NAN;
    }
    if (flat >= g.nc) {
        if (flat > g.nc) o.flag[0] = 1;
        *sent_row = true;
        bool nl = false;
#pragma unroll
        for (int a = 0; a < D; ++a) nl |= isnan(loc[a]);
        *nan_loc = nl;
        return (uint32_t)B.nbt << cellbits;
    }
    int lc;
    const uint32_t b = (uint32_t)bucket_of<D>(flat, ia, g, B, &lc);
    if constexpr (D == 3) {
        float lf[3];
        const uint32_t lcq = loc_encode(loc, lf) | (uint32_t)lc;
        *rec = make_float4(lf[0], lf[1], lf[2], __uint_as_float(lcq));
        *has = true;
    }
    return (b << cellbits) | (cellbits ? (uint32_t)lc : 0u);
}

// Store the records of a wave's 64 consecutive rows (row0 .. row0 + 63; lane
// l holds row row0 + l).  16-byte slots (qstride 1): one store per lane.
// 32-byte slots (qstride 2): each slot is written whole, its second half
// zero, by two adjacent lanes, so that every store instruction covers 1 KiB
// of consecutive bytes (where a lane writing both halves of its own slot
// would leave every line half-written by each instruction).  Every lane of
// the wave must call it.
__device__ __forceinline__ void store_recs(float4* __restrict__ qrec, int32_t qstride, int64_t row0,
                                           int64_t n, float4 r, bool has) {
    const int lane = threadIdx.x & 63;
    if (qstride == 1) {
        if (has) qrec[row0 + lane] = r;
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int src = (lane >> 1) + 32 * h;
        const float x = __shfl(r.x, src, 64), y = __shfl(r.y, src, 64);
        const float z = __shfl(r.z, src, 64), w = __shfl(r.w, src, 64);
        const int hs = __shfl((int)has, src, 64);
        const int64_t row = row0 + src;
        if (hs && row < n)
            qrec[row * 2 + (lane & 1)] =
                (lane & 1) ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(x, y, z, w);
    }
}

template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(KEYS_THREADS)
k_keys(ArbGeom g, ArbBuckets B, const TC* __restrict__ sent, int32_t has_comps,
       int32_t has_norm, const double* __restrict__ q, int64_t n, int64_t ldq, ArbOut o,
       uint32_t* __restrict__ key, int32_t cellbits, int32_t* __restrict__ hist, int32_t hshift,
       int64_t ntiles, float4* __restrict__ qrec, int32_t qstride) {
    __shared__ int32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * (KEYS_THREADS * KEYS_ITEMS) + threadIdx.x;
    uint32_t sent_rows = 0;     // bit j: row j reduces to the sentinel, finite local coordinates
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
            float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
            bool has = false, sr = false, nl = false;
            if (i < n) {
                const uint32_t k =
                    key_of_row<D>(g, B, o, i, qv[u], cellbits, &r, &has, &sr, &nl);
                key[i] = k;
                atomicAdd(&cnt[(k >> hshift) & 255u], 1);
                if (nl) store_nan_row<D, TO>(has_comps, has_norm, o, i);
            }
            sent_rows |= (sr && !nl ? 1u : 0u) << (j0 + u);
            if (qrec) store_recs(qrec, qstride, i - (threadIdx.x & 63), n, r, has);
        }
    }
    // rows that reduce to the sentinel with finite local coordinates (rare:
    // the rows outside the box or NaN got their NaN outputs above): evaluated
    // from the sentinel row one at a time, their positions loaded again
    // (unchanged: only rows outside the box are NaN-filled)
#pragma unroll 1
    for (int u = 0; u < KEYS_ITEMS; ++u) {
        if (!((sent_rows >> u) & 1u)) continue;
        const int64_t i = base + (int64_t)u * KEYS_THREADS;
        double loc[D];
        bool oob;
        arb_reduce<D>(q + i * ldq, g, loc, &oob);
        eval_point<D, T, TC, TO>(sent, has_comps, has_norm, loc, g, o, i);
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

// Split packed records (one aligned record of `width` elements per row, as
// the bucketed kernels write them) into the caller's separate comps / norms /
// grads arrays.  Row i is thread i, so both the record reads and the array
// writes are contiguous over a wave; the bucketed kernels' random writes then
// cost one full record per row instead of three partial-line writes.
template <typename TO, int D, int W>
__global__ void __launch_bounds__(256)
k_unpack(const TO* __restrict__ rec, int32_t has_comps, int32_t has_norm,
         TO* __restrict__ comps, TO* __restrict__ norms, TO* __restrict__ grads, int64_t n) {
    static_assert(W == 4 || W == 8, "records are 4 or 8 elements");
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the record is 16-byte aligned (W * sizeof(TO) is 16, 32 or 64 bytes)
    TO r[W];
    using V = typename std::conditional<sizeof(TO) == 4, float4, double2>::type;
    constexpr int PER = 16 / sizeof(TO);
    const V* src = reinterpret_cast<const V*>(rec + i * W);
#pragma unroll
    for (int v = 0; v < W / PER; ++v) {
        const V x = src[v];
        r[v * PER] = x.x;
        r[v * PER + 1] = x.y;
        if constexpr (PER == 4) {
            r[v * PER + 2] = x.z;
            r[v * PER + 3] = x.w;
        }
    }
    // constant indices in each branch keep r in registers; an index past the
    // record (possible only in combinations the width rules out) reads 0
    auto at = [&](int k) { return k < W ? r[k < W ? k : 0] : TO(0); };
    if (has_comps) {
#pragma unroll
        for (int c = 0; c < 3; ++c) comps[i * 3 + c] = at(c);
    }
    if (has_norm) {
        if (has_comps) {
            norms[i] = at(3);
#pragma unroll
            for (int d = 0; d < D; ++d) grads[i * D + d] = at(4 + d);
        } else {
            norms[i] = at(0);
#pragma unroll
            for (int d = 0; d < D; ++d) grads[i * D + d] = at(1 + d);
        }
    }
}

// The same split for 3D records when the three arrays are 16-byte aligned:
// a workgroup reads its 256 rows' records (one per thread) into LDS, then
// writes each array's part for those rows (768 comps, 256 norms, 768 grads
// values, contiguous) as consecutive 16-byte vectors, so every store
// instruction of a wave covers whole cache lines.
constexpr int UNPACK_ROWS = 256;

template <typename TO, int W>
__global__ void __launch_bounds__(UNPACK_ROWS)
k_unpack_lds(const TO* __restrict__ rec, int32_t has_comps, TO* __restrict__ comps,
             TO* __restrict__ norms, TO* __restrict__ grads, int64_t n) {
    static_assert(W == 4 || W == 8, "records are 4 or 8 elements");
    using V = typename std::conditional<sizeof(TO) == 4, float4, double2>::type;
    constexpr int PER = 16 / sizeof(TO);
    __shared__ __align__(16) TO sc[UNPACK_ROWS * 3];
    __shared__ __align__(16) TO sn[UNPACK_ROWS];
    __shared__ __align__(16) TO sg[UNPACK_ROWS * 3];
    const int64_t r0 = (int64_t)blockIdx.x * UNPACK_ROWS;
    const int64_t i = r0 + threadIdx.x;
    if (i < n) {
        TO r[W];
        const V* src = reinterpret_cast<const V*>(rec + i * W);
#pragma unroll
        for (int v = 0; v < W / PER; ++v) {
            const V x = src[v];
            r[v * PER] = x.x;
            r[v * PER + 1] = x.y;
            if constexpr (PER == 4) {
                r[v * PER + 2] = x.z;
                r[v * PER + 3] = x.w;
            }
        }
        // constant indices in each branch keep r in registers
        if (has_comps) {
#pragma unroll
            for (int c = 0; c < 3; ++c) sc[threadIdx.x * 3 + c] = r[c < W ? c : 0];
            sn[threadIdx.x] = r[3 < W ? 3 : 0];
#pragma unroll
            for (int d = 0; d < 3; ++d) sg[threadIdx.x * 3 + d] = r[4 + d < W ? 4 + d : 0];
        } else {
            sn[threadIdx.x] = r[0];
#pragma unroll
            for (int d = 0; d < 3; ++d) sg[threadIdx.x * 3 + d] = r[1 + d];
        }
    }
    __syncthreads();
    const int64_t rows = (n - r0) < UNPACK_ROWS ? (n - r0) : UNPACK_ROWS;
    if (rows == UNPACK_ROWS) {
        // r0 is a multiple of 256, so each part starts 16-byte aligned
        auto put = [&](TO* dst, const TO* lds, int m) {
            V* d = reinterpret_cast<V*>(dst);
            const V* l = reinterpret_cast<const V*>(lds);
            for (int v = threadIdx.x; v < m / PER; v += UNPACK_ROWS) d[v] = l[v];
        };
        if (has_comps) put(comps + r0 * 3, sc, UNPACK_ROWS * 3);
        put(norms + r0, sn, UNPACK_ROWS);
        put(grads + r0 * 3, sg, UNPACK_ROWS * 3);
    } else {
        for (int e = threadIdx.x; e < rows * 3; e += UNPACK_ROWS) {
            if (has_comps) comps[r0 * 3 + e] = sc[e];
            grads[r0 * 3 + e] = sg[e];
        }
        for (int e = threadIdx.x; e < rows; e += UNPACK_ROWS) norms[r0 + e] = sn[e];
    }
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

// Rows of the bucketed evaluation taken by this workgroup (uniform over it).
// Workgroup w < nbt takes the first `chunk` rows of bucket w.  Workgroup
// nbt + o takes, among the sorted positions [o chunk, (o + 1) chunk), the
// rows past the first chunk of their bucket: those of one bucket at most,
// the one holding position o chunk (a bucket that starts later inside the
// range has its first chunk of rows there).  So every row of a bucket below
// nbt is taken exactly once, and a bucket of many rows (queries clustered in
// a few cells) is shared by as many workgroups.  hot[0] == 0 (no bucket
// holds more than a chunk) ends the workgroups past nbt at once.
__device__ __forceinline__ bool bucket_rows(const int32_t* __restrict__ offs, int64_t nbt,
                                            const int32_t* __restrict__ hot, int32_t chunk,
                                            int64_t* b, int32_t* start, int32_t* end) {
    const int64_t w = blockIdx.x;
    if (w < nbt) {
        const int32_t s0 = offs[w], e0 = offs[w + 1];
        *b = w;
        *start = s0;
        *end = e0 - s0 > chunk ? s0 + chunk : e0;
        return s0 < e0;
    }
    if (hot[0] == 0) return false;
    const int64_t p0 = (w - nbt) * (int64_t)chunk;
    if (p0 >= offs[nbt]) return false;              // the rows outside the grid follow
    int64_t lo = 0, hi = nbt;                       // offs[lo] <= p0 < offs[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (offs[mid] <= p0) lo = mid;
        else hi = mid;
    }
    const int64_t s0 = (int64_t)offs[lo] + chunk, e0 = offs[lo + 1];
    const int64_t s1 = p0 > s0 ? p0 : s0;
    const int64_t e1 = p0 + chunk < e0 ? p0 + chunk : e0;
    *b = lo;
    *start = (int32_t)s1;
    *end = (int32_t)e1;
    return s1 < e1;
}

// One workgroup per bucket: stage the block's coefficients in LDS, then
// evaluate the bucket's queries from LDS.
template <int D, typename T, typename TC, typename TO>
__global__ void __launch_bounds__(256)
k_eval_bucket(ArbGeom g, ArbBuckets B, const TC* __restrict__ table, int32_t has_comps,
              int32_t has_norm, const double* __restrict__ q, int64_t n, int64_t ldq, ArbOut o,
              const int32_t* __restrict__ offs, const int32_t* __restrict__ perm,
              const int32_t* __restrict__ hot) {
    extern __shared__ uint4 smem[];
    int64_t b;
    int32_t start, end;
    if (!bucket_rows(offs, B.nbt, hot, ARB_EVAL_CHUNK, &b, &start, &end)) return;
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
              const int32_t* __restrict__ perm, const int32_t* __restrict__ hot) {
    extern __shared__ uint4 smem[];
    int64_t b;
    int32_t start, end;
    if (!bucket_rows(offs, B.nbt, hot, ARB_EVAL_CHUNK, &b, &start, &end)) return;
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
    __shared__ unsigned char cbad[BX * BY * BZ];        // cells with a non-finite stencil
    for (int k = 0; k < ninterp; ++k) {
        // tile k into LDS: the x pass of interpolant k - 1, the last reader
        // of `nodes`, ended at that interpolant's first pass barrier (so did
        // its last reader of cbad, the z pass's stores)
        int nonfinite = 0;
#pragma unroll
        for (int j = 0; j < FUSED_REGS; ++j) {
            const int e = threadIdx.x + j * blockDim.x - k * NT;
            if (e >= 0 && e < NT) {
                nodes[e] = nv[j];
                nonfinite |= !isfinite(nv[j]);
            }
        }
        const bool anybad = __syncthreads_or(nonfinite) != 0;
        if (anybad) {
            arbs3::cell_bad3<BX, BY, BZ>(nodes, cbad);
            __syncthreads();
        }
        // consecutive threads, consecutive cells: local cell cx + BX (cy + BY cz)
        arbs3::sep3_passes<BX, BY, BZ, true>(nodes, t1, t2,
                                             [&](int cx, int cy, int cz, int i, double a) {
            const int lc = cx + BX * (cy + BY * cz);
            if (anybad && cbad[lc]) a = NAN;
            cf[lc * cstr + k * 64 + i] = (T)a;
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

// ---------------------------------------------------------------------------
// 4D without the coefficient table (the table does not fit, or a table size
// limit is set): one workgroup per bucket forms its cells' coefficients in
// LDS with the arithmetic of the 4D table build (arb_sep4.h: per power of t
// the t-combined 3D node tile through the separable passes, then the
// row-shift correction P . c with c_k = G(k-1) - G(k), row 85 zero; and NaN
// for every coefficient of a cell whose stencil holds a non-finite value),
// so the results are bitwise those of the table, and evaluates the bucket's
// queries as the staged kernel does.  Dynamic LDS: the coefficient block
// (ncell * cellstride * 16 bytes), then doubles: the block's 4D node tile,
// one t-combined 3D tile, the pass buffers, the corner differences G and
// the per-cell c.
// ---------------------------------------------------------------------------
template <int BX, int BY, int BZ, int BT>
struct Fused4Shape {
    static constexpr int NX = BX + 3, NY = BY + 3, NZ = BZ + 3, NTT = BT + 3;
    static constexpr int NT4 = NX * NY * NZ * NTT;                 // node tile
    static constexpr int NT3 = NX * NY * NZ;                       // t-combined tile
    static constexpr int GX = BX + 1, GY = BY + 1, GZ = BZ + 1, GT = BT + 1;
    static constexpr int NG = GX * GY * GZ * GT;                   // corner differences
    static constexpr int NC = BX * BY * BZ * BT;
    static constexpr int DOUBLES = NT4 + NT3 + (NY * NZ * BX * 4) + (NZ * BY * BX * 16) + NG + NC * 16;
};

template <typename T, typename TO, int BX, int BY, int BZ, int BT>
__global__ void __launch_bounds__(256)
k_eval_fused4(ArbGeom g, ArbBuckets B, const double* __restrict__ vals, int32_t ninterp,
              int32_t has_comps, int32_t has_norm, const double* __restrict__ q, int64_t n,
              int64_t ldq, ArbOut o, const int32_t* __restrict__ offs,
              const int32_t* __restrict__ perm, const int32_t* __restrict__ hot) {
    using S = Fused4Shape<BX, BY, BZ, BT>;
    extern __shared__ uint4 smem[];
    __shared__ unsigned char cbad[S::NC];           // cells with a non-finite stencil
    int64_t b;
    int32_t start, end;
    if (!bucket_rows(offs, B.nbt, hot, ARB_EVAL_CHUNK, &b, &start, &end)) return;
    int64_t c0[4];
    {
        int64_t rem = b;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            c0[a] = (rem % B.nb[a]) << B.bsh[a];
            rem /= B.nb[a];
        }
    }
    double* nodes = reinterpret_cast<double*>(smem + B.ncell * B.cellstride);
    double* t3 = nodes + S::NT4;
    double* t1 = t3 + S::NT3;
    double* t2 = t1 + S::NY * S::NZ * BX * 4;
    double* gcorn = t2 + S::NZ * BY * BX * 16;
    double* cc = gcorn + S::NG;
    T* cf = reinterpret_cast<T*>(smem);
    const int cstr = B.cellstride * (16 / (int)sizeof(T));   // elements per LDS cell
    // this thread's first query, loaded before the coefficient work
    int32_t s = start + threadIdx.x;
    int64_t qi = -1;
    double qp[4];
    if (s < end) {
        const int64_t pi = perm[s];
        qi = (pi >= 0 && pi < n) ? pi : -1;         // always in range: perm holds row indices
        const int64_t qq = qi >= 0 ? qi : 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) qp[a] = q[qq * ldq + a];
    }
    for (int k = 0; k < ninterp; ++k) {
        const double* F = vals + (int64_t)k * g.nnodes;
        __syncthreads();                            // the previous interpolant is done
        // node tile, x fastest then y, z, t; nodes past the grid read 0 (only
        // cells outside the grid use them)
        int nonfinite = 0;
        for (int e = threadIdx.x; e < S::NT4; e += blockDim.x) {
            const int x = e % S::NX, y = (e / S::NX) % S::NY, z = (e / (S::NX * S::NY)) % S::NZ;
            const int t = e / (S::NX * S::NY * S::NZ);
            const int64_t nx = c0[0] + x, ny = c0[1] + y, nz = c0[2] + z, nt = c0[3] + t;
            double v = 0.0;
            if (nx < g.n[0] && ny < g.n[1] && nz < g.n[2] && nt < g.n[3])
                v = F[nx + g.nstride[1] * ny + g.nstride[2] * nz + g.nstride[3] * nt];
            nodes[e] = v;
            nonfinite |= !isfinite(v);
        }
        const bool anybad = __syncthreads_or(nonfinite) != 0;
        auto N4 = [&](int x, int y, int z, int t) {
            return nodes[((t * S::NZ + z) * S::NY + y) * S::NX + x];
        };
        if (anybad) {
            for (int c = threadIdx.x; c < S::NC; c += blockDim.x) {
                const int cx = c % BX, cy = (c / BX) % BY, cz = (c / (BX * BY)) % BZ;
                const int ct = c / (BX * BY * BZ);
                bool bad = false;
                for (int t = 0; t < 4; ++t)
                    for (int z = 0; z < 4; ++z)
                        for (int y = 0; y < 4; ++y)
                            for (int x = 0; x < 4; ++x)
                                bad |= !isfinite(N4(cx + x, cy + y, cz + z, ct + t));
                cbad[c] = bad ? 1 : 0;
            }
        }
        // xyzt central differences at the cells' corner nodes (tile node + 1),
        // in the table build's order of terms
        for (int i = threadIdx.x; i < S::NG; i += blockDim.x) {
            const int gx = i % S::GX, gy = (i / S::GX) % S::GY, gz = (i / (S::GX * S::GY)) % S::GZ;
            const int gt = i / (S::GX * S::GY * S::GZ);
            gcorn[i] = arbs4::corner_diff<S::NX, S::NX * S::NY, S::NX * S::NY * S::NZ>(
                nodes + (((gt + 1) * S::NZ + gz + 1) * S::NY + gy + 1) * S::NX + gx + 1);
        }
        __syncthreads();
        // c_k = G(corner k-1) - G(corner k) per cell
        for (int i = threadIdx.x; i < S::NC * 16; i += blockDim.x) {
            const int kk = i & 15, c = i >> 4;
            const int cx = c % BX, cy = (c / BX) % BY, cz = (c / (BX * BY)) % BZ;
            const int ct = c / (BX * BY * BZ);
            auto G = [&](int cn) {
                const int gx = cx + (cn & 1), gy = cy + ((cn >> 1) & 1);
                const int gz = cz + ((cn >> 2) & 1), gt = ct + ((cn >> 3) & 1);
                return gcorn[((gt * S::GZ + gz) * S::GY + gy) * S::GX + gx];
            };
            cc[c * 16 + kk] = (kk ? G(kk - 1) : 0.0) - G(kk);
        }
        for (int ct = 0; ct < BT; ++ct) {
            for (int et = 0; et < 4; ++et) {
                // t-combined 3D node tile for power et of cell row ct
                for (int i = threadIdx.x; i < S::NT3; i += blockDim.x) {
                    const int x = i % S::NX, y = (i / S::NX) % S::NY, z = i / (S::NX * S::NY);
                    double c[4];
                    arbs3::cub4(N4(x, y, z, ct), N4(x, y, z, ct + 1), N4(x, y, z, ct + 2),
                                N4(x, y, z, ct + 3), c);
                    t3[i] = c[et];
                }
                __syncthreads();
                arbs3::sep3_passes4<BX, BY, BZ, false>(
                    t3, t1, t2, [&](int cx, int cy, int cz, int exy, const double* c0) {
                        const int c = cx + BX * (cy + BY * (cz + BZ * ct));
                        double a[4] = {c0[0], c0[1], c0[2], c0[3]};
                        arbs4::shift_rows(cc + c * 16, exy & 3, exy >> 2, et, a);
                        const bool bad = anybad && cbad[c];
                        T* o = cf + c * cstr + k * 256 + exy + 64 * et;
#pragma unroll
                        for (int ez = 0; ez < 4; ++ez) o[16 * ez] = bad ? (T)NAN : (T)a[ez];
                    });
            }
        }
    }
    __syncthreads();
    if (qi >= 0) eval_row_lds<4, T, T, TO>(g, B, smem, has_comps, has_norm, o, b, qi, qp);
    for (s += blockDim.x; s < end; s += blockDim.x) {
        const int64_t pi = perm[s];
        if (pi < 0 || pi >= n) continue;            // unreachable: perm holds row indices
#pragma unroll
        for (int a = 0; a < 4; ++a) qp[a] = q[pi * ldq + a];
        eval_row_lds<4, T, T, TO>(g, B, smem, has_comps, has_norm, o, b, pi, qp);
    }
}

// ---------------------------------------------------------------------------
// Node form (mixed precision, 3D).  Since the coefficients are (M x M x M) F,
// the interpolant of a cell is
//   p(u, v, w) = sum_ijk wx_i(u) wy_j(v) wz_k(w) F(cx + i, cy + j, cz + k),
// w_o(u) = sum_e M[e][o] u^e (the cubic Hermite weights with central-
// difference slopes), so a query can be evaluated from its 4 x 4 x 4 nodes.
//
// Node tiles.  Per bucket, the block's node tile (cells + 3 per axis, x
// fastest) is kept as FP32 differences from a polynomial model of each
// interpolant around the tile's middle node m, with p = node - m:
//   L(p) = F(m) + G . p for all but the last interpolant,
//   Q(p) = L(p) + p^T H p / 2 for the last one (|B| in mode 'both', the last
//   component in 'vector'), and C(p) = Q(p) + T[p, p, p] / 6 for the one
//   interpolant of scalar fields and mode 'norm' (whose values cross zero:
//   the per-value error bound needs the smaller differences),
// G, H and T the central first, second and third differences at m (FP64),
// together with the models (TileHead).  Central differences are exact for
// quadratics, so the weights reproduce every polynomial of degree <= 2 per
// axis (sum_o w_o = 1, sum_o w_o (o - 1) = u, sum_o w_o (o - 1)^2 = u^2,
// and the same for w' with 0, 1, 2u); of the cubic terms only x^3, y^3, z^3
// are not, and their interpolant on a unit cell is x^3 + phi(u) with
// phi(u) = u (u - 1) (2u - 1) (slopes 3x^2 + 1 from central differences).
// So
//   p = C(query) + sum_a (T_aaa / 6) phi(u_a) + sum w (F - C),
//   grad p = grad C + (T_aaa / 6) phi'(u_a) + sum w' (F - C):
// the model is evaluated in FP64 and only the differences, whose size is set
// by the field's fourth derivatives across the tile (third for Q, second
// for L), carry FP32 rounding.  (The cubic terms cost 'both' 50 registers
// and 0.09 ms per 10^7 queries, for no gain its bound needs.)  (Rounding the node values themselves to
// FP32 would leave an error of order 2^-24 |F| in the gradient, since the
// gradient weights sum to zero.)  A tile record holds the NI interpolants of
// one node: 16 bytes for NI = 4 ('both') and NI = 3 ('vector', fourth lane
// 0), 4 bytes for NI = 1 (scalar fields, 'norm').
// ---------------------------------------------------------------------------
struct TileHead {
    double f[4];             // F(m) per interpolant
    double gx[4], gy[4], gz[4];   // central differences at m, per node step
    double hxx, hyy, hzz, hxy, hxz, hyz;   // second differences, last interpolant
    double txxx, tyyy, tzzz, txxy, txxz, txyy, tyyz, txzz, tyzz, txyz;   // third
    int32_t m[3];            // m in tile coordinates
    int32_t pad;
};

// Whether the last interpolant's model has the cubic terms.
template <int NI>
constexpr bool tile_cubic() { return NI == 1; }

// The model of interpolant k (of NI) at p (tile coordinates relative to m).
template <int NI>
__device__ __forceinline__ double tile_model(const TileHead& h, int k, double px, double py,
                                            double pz) {
    double v = fma(h.gz[k], pz, fma(h.gy[k], py, fma(h.gx[k], px, h.f[k])));
    if (k == NI - 1) {
        const double q = fma(0.5 * h.hzz, pz * pz,
                             fma(0.5 * h.hyy, py * py, 0.5 * h.hxx * (px * px)));
        v += fma(h.hyz, py * pz, fma(h.hxz, px * pz, fma(h.hxy, px * py, q)));
    }
    if (tile_cubic<NI>() && k == NI - 1) {
        const double px2 = px * px, py2 = py * py, pz2 = pz * pz;
        double c = (h.txxx * px2 * px + h.tyyy * py2 * py + h.tzzz * pz2 * pz) * (1.0 / 6.0);
        c = fma(0.5, h.txxy * px2 * py + h.txxz * px2 * pz + h.txyy * px * py2 +
                         h.tyyz * py2 * pz + h.txzz * px * pz2 + h.tyzz * py * pz2, c);
        v += fma(h.txyz, px * py * pz, c);
    }
    return v;
}

// Gradient (per node step) of the last interpolant's model at p.
template <int NI>
__device__ __forceinline__ void tile_model_grad(const TileHead& h, int k, double px, double py,
                                                double pz, double* gm) {
    gm[0] = fma(h.hxz, pz, fma(h.hxy, py, fma(h.hxx, px, h.gx[k])));
    gm[1] = fma(h.hyz, pz, fma(h.hyy, py, fma(h.hxy, px, h.gy[k])));
    gm[2] = fma(h.hzz, pz, fma(h.hyz, py, fma(h.hxz, px, h.gz[k])));
    if constexpr (tile_cubic<NI>()) {
        const double px2 = px * px, py2 = py * py, pz2 = pz * pz;
        gm[0] += 0.5 * (h.txxx * px2 + h.txyy * py2 + h.txzz * pz2) + h.txxy * px * py +
                 h.txxz * px * pz + h.txyz * py * pz;
        gm[1] += 0.5 * (h.tyyy * py2 + h.txxy * px2 + h.tyzz * pz2) + h.txyy * px * py +
                 h.tyyz * py * pz + h.txyz * px * pz;
        gm[2] += 0.5 * (h.tzzz * pz2 + h.txxz * px2 + h.tyyz * py2) + h.txzz * px * pz +
                 h.tyzz * py * pz + h.txyz * px * py;
    }
}

template <int NI>
struct TileRec {
    using type = float4;
    using bits = v4u;        // same bytes, for nontemporal loads
};
template <>
struct TileRec<1> {
    using type = float;
    using bits = uint32_t;
};

template <int NI, int BX, int BY, int BZ>
struct TileShape {
    static constexpr int NX = BX + 3, NY = BY + 3, NZ = BZ + 3;
    static constexpr int NT = NX * NY * NZ;
    static constexpr int NTP = (NT + 15) / 16 * 16;      // records per tile in HBM
    // LDS strides in records.  A lane reads record
    // (lx + i) + YS (ly + j) + ZS (lz + k); LDS serves a wave's reads from
    // 64 banks of 4 bytes, i.e. 16 groups for 16-byte records and 64 for
    // 4-byte ones, so the strides are padded (by < 16 records each) to
    // spread the block's cells evenly over the groups: the first (YS, ZS)
    // with the smallest largest group.
    static constexpr int G = NI == 1 ? 64 : 16;
    static constexpr int load(int ys, int zs) {
        int cnt[64] = {};
        int mx = 0;
        for (int z = 0; z < BZ; ++z)
            for (int y = 0; y < BY; ++y)
                for (int x = 0; x < BX; ++x) {
                    const int gi = (x + ys * y + zs * z) & (G - 1);
                    ++cnt[gi];
                    mx = cnt[gi] > mx ? cnt[gi] : mx;
                }
        return mx;
    }
    static constexpr int pick(bool want_z) {
        int by = NX, bz = NX * NY, best = 1 << 30;
        for (int py = 0; py < 16; ++py)
            for (int pz = 0; pz < 16; ++pz) {
                const int ys = NX + py, zs = NY * ys + pz;
                const int l = load(ys, zs);
                if (l < best) {
                    best = l;
                    by = ys;
                    bz = zs;
                }
            }
        return want_z ? bz : by;
    }
    static constexpr int YS = pick(false), ZS = pick(true);
};

// Tile origin (first cell = first node) of bucket b, per axis.
__device__ __forceinline__ void bucket_origin3(int64_t b, const ArbBuckets& B, int64_t* c0) {
    int64_t rem = b;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c0[a] = (rem % B.nb[a]) << B.bsh[a];
        rem /= B.nb[a];
    }
}

// The model of interpolant k (values v = that interpolant's node values)
// around grid node mi: F(mi) and the central first differences, and for the
// last interpolant the second (and, with the cubic model, third) differences.
// Returns whether every number of it is finite.
template <int NI>
__device__ bool tile_head_at(const ArbGeom& g, const double* __restrict__ v, int64_t mi, int k,
                             TileHead& h) {
    const int64_t sx = 1, sy = g.nstride[1], sz = g.nstride[2];
    const double f = v[mi];
    h.f[k] = f;
    h.gx[k] = 0.5 * (v[mi + sx] - v[mi - sx]);
    h.gy[k] = 0.5 * (v[mi + sy] - v[mi - sy]);
    h.gz[k] = 0.5 * (v[mi + sz] - v[mi - sz]);
    bool ok = isfinite(f) && isfinite(h.gx[k]) && isfinite(h.gy[k]) && isfinite(h.gz[k]);
    if (k != NI - 1) return ok;
    h.hxx = v[mi + sx] - 2.0 * f + v[mi - sx];
    h.hyy = v[mi + sy] - 2.0 * f + v[mi - sy];
    h.hzz = v[mi + sz] - 2.0 * f + v[mi - sz];
    h.hxy = 0.25 * ((v[mi + sx + sy] - v[mi + sx - sy]) - (v[mi - sx + sy] - v[mi - sx - sy]));
    h.hxz = 0.25 * ((v[mi + sx + sz] - v[mi + sx - sz]) - (v[mi - sx + sz] - v[mi - sx - sz]));
    h.hyz = 0.25 * ((v[mi + sy + sz] - v[mi + sy - sz]) - (v[mi - sy + sz] - v[mi - sy - sz]));
    ok = ok && isfinite(h.hxx) && isfinite(h.hyy) && isfinite(h.hzz) && isfinite(h.hxy) &&
         isfinite(h.hxz) && isfinite(h.hyz);
    if constexpr (tile_cubic<NI>()) {
        // third differences: along one axis (5 nodes; 0 where the axis has
        // only 4), and the central differences of the second ones
        auto d3 = [&](int64_t st, int a) {
            return g.n[a] >= 5 ? 0.5 * ((v[mi + 2 * st] - v[mi - 2 * st]) -
                                        2.0 * (v[mi + st] - v[mi - st]))
                               : 0.0;
        };
        auto d2d1 = [&](int64_t s2, int64_t s1) {   // d/d(s1) of d2/d(s2)^2
            return 0.5 * ((v[mi + s2 + s1] - 2.0 * v[mi + s1] + v[mi - s2 + s1]) -
                          (v[mi + s2 - s1] - 2.0 * v[mi - s1] + v[mi - s2 - s1]));
        };
        h.txxx = d3(sx, 0);
        h.tyyy = d3(sy, 1);
        h.tzzz = d3(sz, 2);
        h.txxy = d2d1(sx, sy);
        h.txxz = d2d1(sx, sz);
        h.txyy = d2d1(sy, sx);
        h.tyyz = d2d1(sy, sz);
        h.txzz = d2d1(sz, sx);
        h.tyzz = d2d1(sz, sy);
        h.txyz = 0.125 * ((v[mi + sx + sy + sz] - v[mi + sx + sy - sz]) -
                          (v[mi + sx - sy + sz] - v[mi + sx - sy - sz]) -
                          (v[mi - sx + sy + sz] - v[mi - sx + sy - sz]) +
                          (v[mi - sx - sy + sz] - v[mi - sx - sy - sz]));
        ok = ok && isfinite(h.txxx) && isfinite(h.tyyy) && isfinite(h.tzzz) &&
             isfinite(h.txxy) && isfinite(h.txxz) && isfinite(h.txyy) && isfinite(h.tyyz) &&
             isfinite(h.txzz) && isfinite(h.tyzz) && isfinite(h.txyz);
    } else {
        h.txxx = h.tyyy = h.tzzz = h.txxy = h.txxz = 0.0;
        h.txyy = h.tyyz = h.txzz = h.tyzz = h.txyz = 0.0;
    }
    return ok;
}

// A zero model for interpolant k (its records then hold the node values).
template <int NI>
__device__ void tile_head_zero(TileHead& h, int k) {
    h.f[k] = h.gx[k] = h.gy[k] = h.gz[k] = 0.0;
    if (k == NI - 1) {
        h.hxx = h.hyy = h.hzz = h.hxy = h.hxz = h.hyz = 0.0;
        h.txxx = h.tyyy = h.tzzz = h.txxy = h.txxz = 0.0;
        h.txyy = h.tyyz = h.txzz = h.tyzz = h.txyz = 0.0;
    }
}

// One workgroup per bucket: the head, then every record of the tile.
//
// Non-finite node values.  In the reference a NaN or +-Inf node value makes
// every coefficient of every cell whose 4x4x4 stencil holds it NaN (alpha =
// A . F, ARBInterp.py:599-602, and A has a zero in every column), and no
// other cell.  Here such a node's record is NaN, so exactly the cells whose
// 64 records include it evaluate to NaN (every record enters every output of
// its interpolant).  The model must stay finite, or it would spread the NaN
// to the whole bucket: when the model around the middle node m is not
// finite, the tile node nearest to m (then lowest index) whose model is
// finite for every interpolant becomes the centre; if there is none, each
// interpolant whose model at m is not finite gets a zero model.
template <int NI, int BX, int BY, int BZ>
__global__ void __launch_bounds__(256)
k_build_tiles3(ArbGeom g, ArbBuckets B, const double* __restrict__ vals,
               typename TileRec<NI>::type* __restrict__ tiles, TileHead* __restrict__ heads) {
    using S = TileShape<NI, BX, BY, BZ>;
    __shared__ TileHead h;
    __shared__ int best;
    const int64_t b = blockIdx.x;
    int64_t c0[3];
    bucket_origin3(b, B, c0);
    // the centre range per axis: one node inside the grid so that the central
    // differences exist (two where the axis has 5 nodes or more, for the third
    // differences along it); the default centre is the tile's middle node
    int64_t lo[3], hi[3], m[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int64_t in = g.n[a] >= 5 ? 2 : 1;
        lo[a] = in;
        hi[a] = g.n[a] - 1 - in;
        const int64_t nd = (a == 0 ? S::NX : a == 1 ? S::NY : S::NZ) / 2;
        m[a] = c0[a] + nd;
        m[a] = m[a] < lo[a] ? lo[a] : (m[a] > hi[a] ? hi[a] : m[a]);
    }
    const int k = threadIdx.x;
    bool bad = false;
    if (k < 4) {
        if (k < NI) {
            const int64_t mi = m[0] + g.nstride[1] * m[1] + g.nstride[2] * m[2];
            bad = !tile_head_at<NI>(g, vals + k * g.nnodes, mi, k, h);
        } else {
            h.f[k] = h.gx[k] = h.gy[k] = h.gz[k] = 0.0;
        }
    }
    if (threadIdx.x == 0) best = INT32_MAX;
    if (__syncthreads_or(bad)) {
        // search the tile for the nearest centre with a finite model
        for (int e = threadIdx.x; e < S::NT; e += blockDim.x) {
            const int x = e % S::NX, y = (e / S::NX) % S::NY, z = e / (S::NX * S::NY);
            const int64_t cx = c0[0] + x, cy = c0[1] + y, cz = c0[2] + z;
            if (cx < lo[0] || cx > hi[0] || cy < lo[1] || cy > hi[1] || cz < lo[2] || cz > hi[2])
                continue;
            const int64_t ci = cx + g.nstride[1] * cy + g.nstride[2] * cz;
            TileHead t;
            bool ok = true;
            for (int kk = 0; kk < NI && ok; ++kk)
                ok = tile_head_at<NI>(g, vals + kk * g.nnodes, ci, kk, t);
            if (ok) {
                const int64_t dx = cx - m[0], dy = cy - m[1], dz = cz - m[2];
                atomicMin(&best, (int)((dx * dx + dy * dy + dz * dz) * 1024 + e));
            }
        }
        __syncthreads();
        const int e = best;
        if (e != INT32_MAX) {
            const int r = e & 1023;
            m[0] = c0[0] + r % S::NX;
            m[1] = c0[1] + (r / S::NX) % S::NY;
            m[2] = c0[2] + r / (S::NX * S::NY);
        }
        if (k < NI) {
            const int64_t mi = m[0] + g.nstride[1] * m[1] + g.nstride[2] * m[2];
            if (!tile_head_at<NI>(g, vals + k * g.nnodes, mi, k, h)) tile_head_zero<NI>(h, k);
        }
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) h.m[a] = (int32_t)(m[a] - c0[a]);
        h.pad = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) heads[b] = h;
    for (int e = threadIdx.x; e < S::NTP; e += blockDim.x) {
        const int x = e % S::NX, y = (e / S::NX) % S::NY, z = e / (S::NX * S::NY);
        const int64_t nx = c0[0] + x, ny = c0[1] + y, nz = c0[2] + z;
        // records past the tile or nodes past the grid (used only by cells
        // outside the grid) are 0; non-finite node values are NaN
        const bool in = e < S::NT && nx < g.n[0] && ny < g.n[1] && nz < g.n[2];
        const int64_t src = in ? nx + g.nstride[1] * ny + g.nstride[2] * nz : 0;
        const double px = x - h.m[0], py = y - h.m[1], pz = z - h.m[2];
        float d[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < NI; ++kk) {
            const double f = vals[kk * g.nnodes + src];
            d[kk] = !in ? 0.f
                        : isfinite(f) ? (float)(f - tile_model<NI>(h, kk, px, py, pz))
                                      : __builtin_nanf("");
        }
        if constexpr (NI == 1) tiles[b * S::NTP + e] = d[0];
        else tiles[b * S::NTP + e] = make_float4(d[0], d[1], d[2], d[3]);
    }
}

__device__ __forceinline__ void hermite_weights(float u, float* w, float* dw) {
    w[0] = u * fmaf(u, fmaf(-0.5f, u, 1.0f), -0.5f);
    w[1] = fmaf(u * u, fmaf(1.5f, u, -2.5f), 1.0f);
    w[2] = u * fmaf(u, fmaf(-1.5f, u, 2.0f), 0.5f);
    w[3] = (u * u) * fmaf(0.5f, u, -0.5f);
    dw[0] = fmaf(u, fmaf(-1.5f, u, 2.0f), -0.5f);
    dw[1] = u * fmaf(4.5f, u, -5.0f);
    dw[2] = fmaf(u, fmaf(-4.5f, u, 4.0f), 0.5f);
    dw[3] = u * fmaf(1.5f, u, -1.0f);
}

__device__ __forceinline__ float4 fma4(float a, float4 x, float4 y) {
    return make_float4(fmaf(a, x.x, y.x), fmaf(a, x.y, y.y), fmaf(a, x.z, y.z), fmaf(a, x.w, y.w));
}
__device__ __forceinline__ float fma4(float a, float x, float y) { return fmaf(a, x, y); }
__device__ __forceinline__ float last_of(float4 x) { return x.w; }
__device__ __forceinline__ float last_of(float x) { return x; }

// Evaluate row qi at local cell lc of a block and local coordinates loc
// from the block's node tile (records with row stride YS and plane stride
// ZS, in LDS or in global memory) and its head.  The gradient is of the last
// interpolant (GRAD: modes with a value and gradient output).  The same
// operations in the same order wherever the tile lives, so every node-tile
// path gives the same bits.
template <typename TO, int NI, int BX, int BY, int BZ, int YS, int ZS>
__device__ __forceinline__ void tile_eval(const ArbGeom& g, const ArbBuckets& B,
                                          const typename TileRec<NI>::type* tile,
                                          const TileHead& h, const ArbOut& o,
                                          const float* lf, uint32_t lcq, int64_t qi) {
    using R = typename TileRec<NI>::type;
    constexpr bool GRAD = NI != 3;
    const int lc = (int)(lcq & 255u);
    const int lx = lc & (BX - 1), ly = (lc >> B.lsh[1]) & (BY - 1), lz = (lc >> B.lsh[2]) & (BZ - 1);
    // compiler barriers around the contraction: they keep the tile head's
    // LDS loads (32 doubles) from being hoisted out of the query loop and
    // held in registers through it (121 -> 76 VGPRs, 4 -> 6 waves per SIMD)
    asm volatile("" ::: "memory");
    float wx[4], dx[4], wy[4], dy[4], wz[4], dz[4];
    hermite_weights(lf[0], wx, dx);
    hermite_weights(lf[1], wy, dy);
    hermite_weights(lf[2], wz, dz);
    const R* base = tile + lz * ZS + ly * YS + lx;
    R val{};
    float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        R sy{};
        float sdx = 0.f, sdy = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const R* r = base + k * ZS + j * YS;
            const R r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
            R sx{};
            sx = fma4(wx[0], r0, sx);
            sx = fma4(wx[1], r1, sx);
            sx = fma4(wx[2], r2, sx);
            sx = fma4(wx[3], r3, sx);
            sy = fma4(wy[j], sx, sy);
            if constexpr (GRAD) {
                float ddx = dx[0] * last_of(r0);
                ddx = fmaf(dx[1], last_of(r1), ddx);
                ddx = fmaf(dx[2], last_of(r2), ddx);
                ddx = fmaf(dx[3], last_of(r3), ddx);
                sdx = fmaf(wy[j], ddx, sdx);
                sdy = fmaf(dy[j], last_of(sx), sdy);
            }
        }
        val = fma4(wz[k], sy, val);
        if constexpr (GRAD) {
            gx = fmaf(wz[k], sdx, gx);
            gy = fmaf(wz[k], sdy, gy);
            gz = fmaf(dz[k], last_of(sy), gz);
        }
    }
    asm volatile("" ::: "memory");
    // the models at the query, in tile coordinates (cell lx spans nodes
    // lx + 1 and lx + 2)
    double loc[3];
    loc_decode(lf, lcq, loc);
    const double px = (double)(lx + 1 - h.m[0]) + loc[0];
    const double py = (double)(ly + 1 - h.m[1]) + loc[1];
    const double pz = (double)(lz + 1 - h.m[2]) + loc[2];
    TO rec[8];
    float vs[4];
    if constexpr (NI == 1) {
        vs[0] = val;
    } else {
        vs[0] = val.x;
        vs[1] = val.y;
        vs[2] = val.z;
        vs[3] = val.w;
    }
    // the interpolant of the last model's x^3, y^3, z^3 terms: phi(u) more
    // than the terms themselves, phi(u) = u (u - 1) (2u - 1)
    constexpr int L = NI - 1;
    double corr = 0.0;
    if constexpr (tile_cubic<NI>()) {
        const double phix = loc[0] * (loc[0] - 1.0) * (2.0 * loc[0] - 1.0);
        const double phiy = loc[1] * (loc[1] - 1.0) * (2.0 * loc[1] - 1.0);
        const double phiz = loc[2] * (loc[2] - 1.0) * (2.0 * loc[2] - 1.0);
        corr = (h.txxx * phix + h.tyyy * phiy + h.tzzz * phiz) * (1.0 / 6.0);
    }
#pragma unroll
    for (int k = 0; k < NI; ++k)
        rec[k] = (TO)(tile_model<NI>(h, k, px, py, pz) + (k == L ? corr : 0.0) + (double)vs[k]);
    if constexpr (GRAD) {
        double gm[3];
        tile_model_grad<NI>(h, L, px, py, pz, gm);
        if constexpr (tile_cubic<NI>()) {
            gm[0] += h.txxx * (1.0 / 6.0) * fma(6.0 * loc[0], loc[0] - 1.0, 1.0);
            gm[1] += h.tyyy * (1.0 / 6.0) * fma(6.0 * loc[1], loc[1] - 1.0, 1.0);
            gm[2] += h.tzzz * (1.0 / 6.0) * fma(6.0 * loc[2], loc[2] - 1.0, 1.0);
        }
        rec[NI] = (TO)((gm[0] + (double)gx) / g.h[0]);
        rec[NI + 1] = (TO)((gm[1] + (double)gy) / g.h[1]);
        rec[NI + 2] = (TO)((gm[2] + (double)gz) / g.h[2]);
    }
#pragma unroll
    for (int r = GRAD ? NI + 3 : NI; r < 8; ++r) rec[r] = (TO)0;
    // record layout [comps(3)] [value, grad(3)] (value first without comps)
    store_rec<3, TO>(rec, NI != 1, GRAD, o, qi);
}


// Lane-per-query node-tile kernel (small batches and single points): each
// lane reads its cell's 64 records and its block's head from the tiles in
// global memory.  Rows outside the box, or past the sentinel, are evaluated
// from the sentinel row `sent`, as everywhere.
template <typename TO, int NI, int BX, int BY, int BZ>
__global__ void __launch_bounds__(256)
k_query_tiles3(ArbGeom g, ArbBuckets B, const typename TileRec<NI>::type* __restrict__ tiles,
               const TileHead* __restrict__ heads, const float* __restrict__ sent,
               int32_t has_comps, int32_t has_norm, const double* __restrict__ q, int64_t n,
               int64_t ldq, ArbOut o) {
    using S = TileShape<NI, BX, BY, BZ>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double loc[3];
    int64_t ia[3];
    bool oob;
    const int64_t flat = arb_reduce_axes<3>(q + i * ldq, g, loc, ia, &oob);
    if (o.cells) o.cells[i] = flat;
    if (oob && o.qmut) {
        for (int64_t c = 0; c < o.ldq; ++c) o.qmut[i * o.ldq + c] = NAN;
    }
    if (flat >= g.nc) {
        if (flat > g.nc) o.flag[0] = 1;
        eval_point<3, float, float, TO>(sent, has_comps, has_norm, loc, g, o, i);
        return;
    }
    int lc;
    const int64_t b = bucket_of<3>(flat, ia, g, B, &lc);
    float lf[3];
    const uint32_t lcq = loc_encode(loc, lf) | (uint32_t)lc;
    tile_eval<TO, NI, BX, BY, BZ, S::NX, S::NX * S::NY>(g, B, tiles + b * S::NTP, heads[b], o,
                                                        lf, lcq, i);
}

// Second derivatives of the Hermite weights (w_o'' for hermite_weights).
__device__ __forceinline__ void hermite_weights2(float u, float* ddw) {
    ddw[0] = fmaf(-3.0f, u, 2.0f);
    ddw[1] = fmaf(9.0f, u, -5.0f);
    ddw[2] = fmaf(-9.0f, u, 4.0f);
    ddw[3] = fmaf(3.0f, u, -1.0f);
}

// Backward of the last interpolant's (value, gradient) outputs from the node
// tiles (mixed precision, 3D).  Per query: d/dq = gv grad U + Hess U . gg,
// with grad U and Hess U formed the way the forward forms U from the tiles:
// the FP64 model's derivatives at the query (plus the interpolant of its
// x^3, y^3, z^3 terms, phi''(u) = 12u - 6) and the FP32 contractions of the
// residual records with the Hermite weights w, w' and w''.  The model
// carries the large, cancelling part of the Hessian in FP64, so only the
// residuals' rounding reaches it.  With perm (the queries grouped by cell),
// thread s takes the query perm[s] and writes the node adjoint's record of
// that query at position s (as k_bwd_points does); without, query s.
template <int NI, int BX, int BY, int BZ>
__global__ void __launch_bounds__(256)
k_bwd_tiles3(ArbGeom g, ArbBuckets B, const typename TileRec<NI>::type* __restrict__ tiles,
             const TileHead* __restrict__ heads, const double* __restrict__ q, int64_t n,
             int64_t ldq, const double* __restrict__ gv, const double* __restrict__ gg,
             double* __restrict__ dq, const int32_t* __restrict__ perm,
             double* __restrict__ rec) {
    using S = TileShape<NI, BX, BY, BZ>;
    using R = typename TileRec<NI>::type;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t i = perm ? (int64_t)perm[s] : s;
    if (i < 0 || i >= n) return;                    // unreachable: perm holds row indices
    double loc[3];
    int64_t ia[3];
    bool oob;
    const int64_t flat = arb_reduce_axes<3>(q + i * ldq, g, loc, ia, &oob);
    const bool ok = flat < g.nc;
    if (rec && ok) {
        double4* r4 = reinterpret_cast<double4*>(rec + s * 8);
        r4[0] = make_double4(loc[0], loc[1], loc[2], gv[i]);
        r4[1] = make_double4(gg[i * 3] / g.h[0], gg[i * 3 + 1] / g.h[1], gg[i * 3 + 2] / g.h[2], 0.0);
    }
    if (!dq) return;
    if (!ok) {
#pragma unroll
        for (int x = 0; x < 3; ++x) dq[i * 3 + x] = NAN;
        return;
    }
    int lc;
    const int64_t b = bucket_of<3>(flat, ia, g, B, &lc);
    const int lx = lc & (BX - 1), ly = (lc >> B.lsh[1]) & (BY - 1), lz = (lc >> B.lsh[2]) & (BZ - 1);
    // the forward's local coordinates (FP32 for the weights, decoded for
    // the model)
    {
        float lf[3];
        const uint32_t lcq = loc_encode(loc, lf);
        loc_decode(lf, lcq, loc);
    }
    float wx[4], dx[4], ddx[4], wy[4], dy[4], ddy[4], wz[4], dz[4], ddz[4];
    hermite_weights((float)loc[0], wx, dx);
    hermite_weights((float)loc[1], wy, dy);
    hermite_weights((float)loc[2], wz, dz);
    hermite_weights2((float)loc[0], ddx);
    hermite_weights2((float)loc[1], ddy);
    hermite_weights2((float)loc[2], ddz);
    const R* base = tiles + b * S::NTP + (lz * S::NY + ly) * S::NX + lx;
    // residual part: value, gradient and Hessian (node units), FP32
    float gr[3] = {0.f, 0.f, 0.f}, hr[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // xx yy zz xy xz yz
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float b00 = 0.f, b10 = 0.f, b01 = 0.f, b20 = 0.f, b11 = 0.f, b02 = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const R* r = base + (k * S::NY + j) * S::NX;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const float v = last_of(r[x]);
                a0 = fmaf(wx[x], v, a0);
                a1 = fmaf(dx[x], v, a1);
                a2 = fmaf(ddx[x], v, a2);
            }
            b00 = fmaf(wy[j], a0, b00);
            b10 = fmaf(wy[j], a1, b10);
            b01 = fmaf(dy[j], a0, b01);
            b20 = fmaf(wy[j], a2, b20);
            b11 = fmaf(dy[j], a1, b11);
            b02 = fmaf(ddy[j], a0, b02);
        }
        gr[0] = fmaf(wz[k], b10, gr[0]);
        gr[1] = fmaf(wz[k], b01, gr[1]);
        gr[2] = fmaf(dz[k], b00, gr[2]);
        hr[0] = fmaf(wz[k], b20, hr[0]);
        hr[1] = fmaf(wz[k], b02, hr[1]);
        hr[2] = fmaf(ddz[k], b00, hr[2]);
        hr[3] = fmaf(wz[k], b11, hr[3]);
        hr[4] = fmaf(dz[k], b10, hr[4]);
        hr[5] = fmaf(dz[k], b01, hr[5]);
    }
    // model part (FP64) at the query, in tile coordinates
    const TileHead& h = heads[b];
    constexpr int L = NI - 1;
    const double px = (double)(lx + 1 - h.m[0]) + loc[0];
    const double py = (double)(ly + 1 - h.m[1]) + loc[1];
    const double pz = (double)(lz + 1 - h.m[2]) + loc[2];
    double gm[3];
    tile_model_grad<NI>(h, L, px, py, pz, gm);
    double H[3][3];
    H[0][0] = h.hxx;
    H[1][1] = h.hyy;
    H[2][2] = h.hzz;
    H[0][1] = h.hxy;
    H[0][2] = h.hxz;
    H[1][2] = h.hyz;
    if constexpr (tile_cubic<NI>()) {
        gm[0] += h.txxx * (1.0 / 6.0) * fma(6.0 * loc[0], loc[0] - 1.0, 1.0);
        gm[1] += h.tyyy * (1.0 / 6.0) * fma(6.0 * loc[1], loc[1] - 1.0, 1.0);
        gm[2] += h.tzzz * (1.0 / 6.0) * fma(6.0 * loc[2], loc[2] - 1.0, 1.0);
        H[0][0] += h.txxx * px + h.txxy * py + h.txxz * pz +
                   h.txxx * (1.0 / 6.0) * fma(12.0, loc[0], -6.0);
        H[1][1] += h.tyyy * py + h.txyy * px + h.tyyz * pz +
                   h.tyyy * (1.0 / 6.0) * fma(12.0, loc[1], -6.0);
        H[2][2] += h.tzzz * pz + h.txzz * px + h.tyzz * py +
                   h.tzzz * (1.0 / 6.0) * fma(12.0, loc[2], -6.0);
        H[0][1] += h.txxy * px + h.txyy * py + h.txyz * pz;
        H[0][2] += h.txxz * px + h.txzz * pz + h.txyz * py;
        H[1][2] += h.tyyz * py + h.tyzz * pz + h.txyz * px;
    }
    H[0][0] += (double)hr[0];
    H[1][1] += (double)hr[1];
    H[2][2] += (double)hr[2];
    H[0][1] += (double)hr[3];
    H[0][2] += (double)hr[4];
    H[1][2] += (double)hr[5];
    H[1][0] = H[0][1];
    H[2][0] = H[0][2];
    H[2][1] = H[1][2];
    const double v = gv[i];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        double acc = v * ((gm[x] + (double)gr[x]) / g.h[x]);
#pragma unroll
        for (int y = 0; y < 3; ++y) acc += (H[x][y] / (g.h[x] * g.h[y])) * gg[i * 3 + y];
        dq[i * 3 + x] = acc;
    }
}

template <typename TO, int NI, int BX, int BY, int BZ>
__global__ void __launch_bounds__(256)
k_eval_tiles3(ArbGeom g, ArbBuckets B, const typename TileRec<NI>::type* __restrict__ tiles,
              const TileHead* __restrict__ heads, const float4* qrec, int32_t qstride, int64_t n,
              ArbOut o, const int32_t* __restrict__ offs, const int32_t* __restrict__ perm,
              const int32_t* __restrict__ hot) {
    using S = TileShape<NI, BX, BY, BZ>;
    using RB = typename TileRec<NI>::bits;
    __shared__ typename TileRec<NI>::type tile[S::NZ * S::ZS];
    __shared__ TileHead h;
    int64_t b;
    int32_t start, end;
    if (!bucket_rows(offs, B.nbt, hot, ARB_EVAL_CHUNK, &b, &start, &end)) return;
    int32_t s = start + threadIdx.x;
    int64_t qi = -1;
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    if (s < end) {
        const int64_t pi = perm[s];
        qi = (pi >= 0 && pi < n) ? pi : -1;         // always in range: perm holds row indices
        r = qrec[(qi >= 0 ? qi : 0) * qstride];
    }
    // stage the tile (read once per call: nontemporal) and its head
    const RB* src = reinterpret_cast<const RB*>(tiles + b * S::NTP);
    RB* dst = reinterpret_cast<RB*>(tile);
    for (int e = threadIdx.x; e < S::NT; e += blockDim.x) {
        const int x = e % S::NX, y = (e / S::NX) % S::NY, z = e / (S::NX * S::NY);
        dst[z * S::ZS + y * S::YS + x] = __builtin_nontemporal_load(&src[e]);
    }
    if (threadIdx.x < (int)(sizeof(TileHead) / 8))
        reinterpret_cast<double*>(&h)[threadIdx.x] =
            reinterpret_cast<const double*>(heads + b)[threadIdx.x];
    __syncthreads();
    // each row's record, written by k_keys: FP32 local coordinates and the
    // local cell inside the bucket (the same reduction as every path's)
    if (qi >= 0) {
        const float lf[3] = {r.x, r.y, r.z};
        tile_eval<TO, NI, BX, BY, BZ, S::YS, S::ZS>(g, B, tile, h, o, lf, __float_as_uint(r.w), qi);
    }
    for (s += blockDim.x; s < end; s += blockDim.x) {
        const int64_t pi = perm[s];
        if (pi < 0 || pi >= n) continue;            // unreachable: perm holds row indices
        const float4 rr = qrec[pi * qstride];
        const float lf[3] = {rr.x, rr.y, rr.z};
        tile_eval<TO, NI, BX, BY, BZ, S::YS, S::ZS>(g, B, tile, h, o, lf, __float_as_uint(rr.w),
                                                    pi);
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

template <int D>
hipError_t launch_direct_red32(const ArbGeom& g, int32_t prec, int32_t out_f64,
                               const void* table, int32_t ninterp, int32_t has_comps,
                               int32_t has_norm, const double* q, int64_t n, int64_t ldq,
                               const ArbOut& o, hipStream_t s) {
    dim3 grid(grid1(n, 256)), block(256);
    if (prec == ARB_PREC_FP64)
        hipLaunchKernelGGL((k_query_direct_red32<D, double, double>), grid, block, 0, s, g,
                           (const double*)table, ninterp, has_comps, has_norm, q, n, ldq, o);
    else if (out_f64)
        hipLaunchKernelGGL((k_query_direct_red32<D, float, double>), grid, block, 0, s, g,
                           (const float*)table, ninterp, has_comps, has_norm, q, n, ldq, o);
    else
        hipLaunchKernelGGL((k_query_direct_red32<D, float, float>), grid, block, 0, s, g,
                           (const float*)table, ninterp, has_comps, has_norm, q, n, ldq, o);
    return hipGetLastError();
}

}  // namespace

hipError_t arb_launch_query_direct_red32(const ArbGeom& g, int32_t prec, int32_t out_f64,
                                         const void* table, int32_t ninterp, int32_t has_comps,
                                         int32_t has_norm, const double* q, int64_t n,
                                         int64_t ldq, const ArbOut& o, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (g.ndim == 3)
        return launch_direct_red32<3>(g, prec, out_f64, table, ninterp, has_comps, has_norm, q,
                                      n, ldq, o, s);
    return launch_direct_red32<4>(g, prec, out_f64, table, ninterp, has_comps, has_norm, q, n,
                                  ldq, o, s);
}

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

namespace {
__global__ void __launch_bounds__(256)
k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
}  // namespace

hipError_t arb_launch_copy16(void* dst, const void* src, int64_t n16, hipStream_t s) {
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy16, dim3(grid1(n16, 256)), dim3(256), 0, s, (const uint4*)src,
                       (uint4*)dst, n16);
    return hipGetLastError();
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

hipError_t arb_launch_unpack(int32_t ndim, int32_t out_f64, int32_t has_comps, int32_t has_norm,
                             const void* rec, void* comps, void* norms, void* grads, int64_t n,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int32_t w = arb_packed_width(has_comps, has_norm, ndim);
    const bool aligned = ((uintptr_t)comps | (uintptr_t)norms | (uintptr_t)grads) % 16 == 0;
    if (ndim == 3 && has_norm && aligned) {
        const dim3 grid(grid1(n, UNPACK_ROWS));
#define L_UNPACK4(TO, WW)                                                                      \
    hipLaunchKernelGGL((k_unpack_lds<TO, WW>), grid, dim3(UNPACK_ROWS), 0, s, (const TO*)rec,     \
                       has_comps, (TO*)comps, (TO*)norms, (TO*)grads, n)
        if (out_f64) {
            if (w == 4) L_UNPACK4(double, 4);
            else L_UNPACK4(double, 8);
        } else {
            if (w == 4) L_UNPACK4(float, 4);
            else L_UNPACK4(float, 8);
        }
#undef L_UNPACK4
        return hipGetLastError();
    }
#define L_UNPACK(TO, DD, WW)                                                                   \
    hipLaunchKernelGGL((k_unpack<TO, DD, WW>), dim3(grid1(n, 256)), dim3(256), 0, s,             \
                       (const TO*)rec, has_comps, has_norm, (TO*)comps, (TO*)norms,             \
                       (TO*)grads, n)
#define L_UNPACK_W(TO, DD)                                                                     \
    if (w == 4) L_UNPACK(TO, DD, 4);                                                           \
    else L_UNPACK(TO, DD, 8)
    if (out_f64) {
        if (ndim == 3) L_UNPACK_W(double, 3);
        else L_UNPACK_W(double, 4);
    } else {
        if (ndim == 3) L_UNPACK_W(float, 3);
        else L_UNPACK_W(float, 4);
    }
#undef L_UNPACK_W
#undef L_UNPACK
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

// Workgroups of a bucketed evaluation: one per bucket, then one per chunk of
// sorted positions for the rows of buckets past their first chunk.
inline unsigned eval_grid(const ArbBuckets& B, int64_t n) {
    return (unsigned)(B.nbt + (n + ARB_EVAL_CHUNK - 1) / ARB_EVAL_CHUNK);
}

hipError_t arb_launch_eval_bucket(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                  int32_t out_f64, const void* table, int32_t ninterp,
                                  int32_t has_comps, int32_t has_norm, const double* q,
                                  int64_t n, int64_t ldq, const ArbOut& o, const int32_t* offs,
                                  const int32_t* perm, const int32_t* hot, int32_t block_threads,
                                  hipStream_t s) {
    (void)ninterp;
    const size_t lds = (size_t)B.ncell * B.cellstride * 16;
    dim3 grid(eval_grid(B, n)), block(block_threads);
#define L_EVALB(DD, T, TC, TO)                                                              \
    hipLaunchKernelGGL((k_eval_bucket<DD, T, TC, TO>), grid, block, lds, s, g, B,             \
                       (const TC*)table, has_comps, has_norm, q, n, ldq, o, offs, perm, hot)
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
// of all interpolants within the registers of the workgroup).  The fused
// kernels run arbs3::SEP_THREADS threads per workgroup whatever the tuning's
// thread count (the separable passes stride by it).
size_t arb_fused_lds(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp, int32_t threads) {
    (void)threads;
    if (g.ndim != 3) return 0;
    const int bx = B.b[0], by = B.b[1], bz = B.b[2];
    bool shape = false;
#define ARB_FUSED_IS(X, Y, Z) shape |= (bx == X && by == Y && bz == Z);
    ARB_FUSED_SHAPES(ARB_FUSED_IS)
#undef ARB_FUSED_IS
    if (!shape) return 0;
    // the local cell index must be cx + bx (cy + by cz), as the kernel writes it
    if (B.lsh[0] != 0 || (1 << B.lsh[1]) != bx || (1 << B.lsh[2]) != bx * by) return 0;
    if ((int64_t)ninterp * arbs3::sep3_nodes(bx, by, bz) > (int64_t)FUSED_REGS * arbs3::SEP_THREADS)
        return 0;
    const size_t lds = (size_t)B.ncell * B.cellstride * 16 +
                       sizeof(double) * (size_t)(arbs3::sep3_nodes(bx, by, bz) +
                                                 arbs3::sep3_t1(bx, by, bz) +
                                                 arbs3::sep3_t2(bx, by, bz));
    // gfx950: 160 KiB of LDS per workgroup, with room for the kernel's static
    // LDS (one byte per cell)
    return lds + 256 <= 160 * 1024 ? lds : 0;
}

hipError_t arb_launch_eval_fused(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                 int32_t out_f64, const double* vals, int32_t ninterp,
                                 int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                 int64_t ldq, const ArbOut& o, const int32_t* offs,
                                 const int32_t* perm, const int32_t* hot, int32_t block_threads,
                                 hipStream_t s) {
    (void)block_threads;
    const size_t lds = arb_fused_lds(g, B, ninterp, arbs3::SEP_THREADS);
    if (lds == 0) return hipErrorInvalidValue;
    dim3 grid(eval_grid(B, n)), block(arbs3::SEP_THREADS);
#define L_FUSED(T, TO, X, Y, Z)                                                               \
    hipLaunchKernelGGL((k_eval_fused3<T, TO, X, Y, Z>), grid, block, lds, s, g, B, vals,      \
                       ninterp, has_comps, has_norm, q, n, ldq, o, offs, perm, hot)
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

// Cell blocks of the 4D fused kernel (the 4D bucket ladder).
#define ARB_FUSED4_SHAPES(X)                                                                  \
    X(4, 4, 2, 2) X(4, 2, 2, 2) X(2, 2, 2, 2) X(2, 2, 2, 1) X(2, 2, 1, 1) X(2, 1, 1, 1)      \
    X(1, 1, 1, 1)

// LDS bytes of k_eval_fused4 for the bucket geometry B (0 when it does not
// apply: 4D, an instantiated block shape, the local cell index laid out as
// the kernel writes it, within 160 KiB with its static LDS).
size_t arb_fused4_lds(const ArbGeom& g, const ArbBuckets& B) {
    if (g.ndim != 4) return 0;
    const int bx = B.b[0], by = B.b[1], bz = B.b[2], bt = B.b[3];
    size_t dbl = 0;
#define ARB_F4_IS(X, Y, Z, T)                                                                 \
    if (bx == X && by == Y && bz == Z && bt == T) dbl = Fused4Shape<X, Y, Z, T>::DOUBLES;
    ARB_FUSED4_SHAPES(ARB_F4_IS)
#undef ARB_F4_IS
    if (dbl == 0) return 0;
    if (B.lsh[0] != 0 || (1 << B.lsh[1]) != bx || (1 << B.lsh[2]) != bx * by ||
        (1 << B.lsh[3]) != bx * by * bz)
        return 0;
    const size_t lds = (size_t)B.ncell * B.cellstride * 16 + sizeof(double) * dbl;
    return lds + 256 <= 160 * 1024 ? lds : 0;
}

hipError_t arb_launch_eval_fused4(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                  int32_t out_f64, const double* vals, int32_t ninterp,
                                  int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                  int64_t ldq, const ArbOut& o, const int32_t* offs,
                                  const int32_t* perm, const int32_t* hot, int32_t block_threads,
                                  hipStream_t s) {
    (void)block_threads;
    const size_t lds = arb_fused4_lds(g, B);
    if (lds == 0) return hipErrorInvalidValue;
    dim3 grid(eval_grid(B, n)), block(arbs3::SEP_THREADS);
#define L_FUSED4(T, TO, X, Y, Z, TT)                                                          \
    hipLaunchKernelGGL((k_eval_fused4<T, TO, X, Y, Z, TT>), grid, block, lds, s, g, B, vals,  \
                       ninterp, has_comps, has_norm, q, n, ldq, o, offs, perm, hot)
#define ARB_FUSED4_CASE(X, Y, Z, TT)                                                          \
    if (B.b[0] == X && B.b[1] == Y && B.b[2] == Z && B.b[3] == TT) {                          \
        if (prec == ARB_PREC_FP64) L_FUSED4(double, double, X, Y, Z, TT);                     \
        else if (out_f64) L_FUSED4(float, double, X, Y, Z, TT);                               \
        else L_FUSED4(float, float, X, Y, Z, TT);                                             \
        return hipGetLastError();                                                             \
    }
    ARB_FUSED4_SHAPES(ARB_FUSED4_CASE)
#undef ARB_FUSED4_CASE
#undef L_FUSED4
    return hipErrorInvalidValue;
}

// Cell blocks the node-tile kernels are instantiated for.// Cell blocks the node-tile kernels are instantiated for.
#define ARB_NODE_SHAPES(X)                                                                    \
    X(8, 8, 4) X(8, 4, 4) X(4, 4, 4) X(8, 4, 2) X(4, 4, 2) X(4, 2, 2) X(2, 2, 2) X(2, 2, 1)    \
    X(2, 1, 1) X(1, 1, 1)

// Whether the node-tile path applies: 3D, mode 'both' (four interpolants,
// components and a gradient), 'vector' (three, components) or scalar /
// 'norm' (one, value and gradient); an instantiated block shape with the
// local cell index cx + bx (cy + by cz).
bool arb_nodes_ok(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp, int32_t has_comps,
                  int32_t has_norm) {
    if (g.ndim != 3) return false;
    const bool mode_ok = (ninterp == 4 && has_comps && has_norm) ||
                         (ninterp == 3 && has_comps && !has_norm) ||
                         (ninterp == 1 && !has_comps && has_norm);
    if (!mode_ok) return false;
    for (int a = 0; a < 3; ++a)
        if (g.n[a] < 4) return false;
    const int bx = B.b[0], by = B.b[1], bz = B.b[2];
    bool shape = false;
#define ARB_NODE_IS(X, Y, Z) shape |= (bx == X && by == Y && bz == Z);
    ARB_NODE_SHAPES(ARB_NODE_IS)
#undef ARB_NODE_IS
    return shape && B.lsh[0] == 0 && (1 << B.lsh[1]) == bx && (1 << B.lsh[2]) == bx * by;
}

// Bytes of the node tiles of B for ninterp interpolants: records, then the
// heads (16-byte aligned).
int64_t arb_tiles_bytes(const ArbBuckets& B, int32_t ninterp, int64_t* head_offset) {
    const int64_t ntp = ((int64_t)(B.b[0] + 3) * (B.b[1] + 3) * (B.b[2] + 3) + 15) / 16 * 16;
    const int64_t rec = B.nbt * ntp * (ninterp == 1 ? 4 : 16);
    if (head_offset) *head_offset = rec;
    return rec + B.nbt * (int64_t)sizeof(TileHead);
}

hipError_t arb_launch_build_tiles(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                  const double* vals, void* tiles, hipStream_t s) {
    int64_t hoff = 0;
    arb_tiles_bytes(B, ninterp, &hoff);
    TileHead* heads = (TileHead*)((char*)tiles + hoff);
#define L_TILES(NI, X, Y, Z)                                                                  \
    hipLaunchKernelGGL((k_build_tiles3<NI, X, Y, Z>), dim3((unsigned)B.nbt), dim3(256), 0, s, g, \
                       B, vals, (typename TileRec<NI>::type*)tiles, heads)
#define ARB_TILES_CASE(X, Y, Z)                                                               \
    if (B.b[0] == X && B.b[1] == Y && B.b[2] == Z) {                                          \
        if (ninterp == 4) L_TILES(4, X, Y, Z);                                                \
        else if (ninterp == 3) L_TILES(3, X, Y, Z);                                           \
        else L_TILES(1, X, Y, Z);                                                             \
        return hipGetLastError();                                                             \
    }
    ARB_NODE_SHAPES(ARB_TILES_CASE)
#undef ARB_TILES_CASE
#undef L_TILES
    return hipErrorInvalidValue;
}

hipError_t arb_launch_bwd_tiles(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                const void* tiles, const double* q, int64_t n, int64_t ldq,
                                const double* gv, const double* gg, double* dq,
                                const int32_t* perm, double* rec, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (ninterp != 4 && ninterp != 1) return hipErrorInvalidValue;   // modes with a gradient
    int64_t hoff = 0;
    arb_tiles_bytes(B, ninterp, &hoff);
    const TileHead* heads = (const TileHead*)((const char*)tiles + hoff);
    dim3 grid(grid1(n, 256)), block(256);
#define L_BWD(NI, X, Y, Z)                                                                    \
    hipLaunchKernelGGL((k_bwd_tiles3<NI, X, Y, Z>), grid, block, 0, s, g, B,                  \
                       (const typename TileRec<NI>::type*)tiles, heads, q, n, ldq, gv, gg, dq, \
                       perm, rec)
#define ARB_BWD_CASE(X, Y, Z)                                                                 \
    if (B.b[0] == X && B.b[1] == Y && B.b[2] == Z) {                                          \
        if (ninterp == 4) L_BWD(4, X, Y, Z);                                                  \
        else L_BWD(1, X, Y, Z);                                                               \
        return hipGetLastError();                                                             \
    }
    ARB_NODE_SHAPES(ARB_BWD_CASE)
#undef ARB_BWD_CASE
#undef L_BWD
    return hipErrorInvalidValue;
}

hipError_t arb_launch_eval_nodes(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                 int32_t out_f64, const void* tiles, const void* qrec,
                                 int32_t qstride, int64_t n,
                                 const ArbOut& o, const int32_t* offs, const int32_t* perm,
                                 const int32_t* hot, int32_t block_threads, hipStream_t s) {
    int64_t hoff = 0;
    arb_tiles_bytes(B, ninterp, &hoff);
    const TileHead* heads = (const TileHead*)((const char*)tiles + hoff);
    dim3 grid(eval_grid(B, n)), block(block_threads);
#define L_NODES(TO, NI, X, Y, Z)                                                              \
    hipLaunchKernelGGL((k_eval_tiles3<TO, NI, X, Y, Z>), grid, block, 0, s, g, B,             \
                       (const typename TileRec<NI>::type*)tiles, heads, (const float4*)qrec,  \
                       qstride, n, o, offs, perm, hot)
#define L_NODES_NI(TO, X, Y, Z)                                                               \
    if (ninterp == 4) L_NODES(TO, 4, X, Y, Z);                                                \
    else if (ninterp == 3) L_NODES(TO, 3, X, Y, Z);                                           \
    else L_NODES(TO, 1, X, Y, Z)
#define ARB_NODES_CASE(X, Y, Z)                                                               \
    if (B.b[0] == X && B.b[1] == Y && B.b[2] == Z) {                                          \
        if (out_f64) L_NODES_NI(double, X, Y, Z);                                             \
        else L_NODES_NI(float, X, Y, Z);                                                      \
        return hipGetLastError();                                                             \
    }
    ARB_NODE_SHAPES(ARB_NODES_CASE)
#undef ARB_NODES_CASE
#undef L_NODES_NI
#undef L_NODES
    return hipErrorInvalidValue;
}

hipError_t arb_launch_query_tiles(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                  int32_t out_f64, const void* tiles, const void* sent,
                                  int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                  int64_t ldq, const ArbOut& o, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t hoff = 0;
    arb_tiles_bytes(B, ninterp, &hoff);
    const TileHead* heads = (const TileHead*)((const char*)tiles + hoff);
    dim3 grid(grid1(n, 256)), block(256);
#define L_QT(TO, NI, X, Y, Z)                                                                 \
    hipLaunchKernelGGL((k_query_tiles3<TO, NI, X, Y, Z>), grid, block, 0, s, g, B,            \
                       (const typename TileRec<NI>::type*)tiles, heads, (const float*)sent,   \
                       has_comps, has_norm, q, n, ldq, o)
#define L_QT_NI(TO, X, Y, Z)                                                                  \
    if (ninterp == 4) L_QT(TO, 4, X, Y, Z);                                                   \
    else if (ninterp == 3) L_QT(TO, 3, X, Y, Z);                                              \
    else L_QT(TO, 1, X, Y, Z)
#define ARB_QT_CASE(X, Y, Z)                                                                  \
    if (B.b[0] == X && B.b[1] == Y && B.b[2] == Z) {                                          \
        if (out_f64) L_QT_NI(double, X, Y, Z);                                                \
        else L_QT_NI(float, X, Y, Z);                                                         \
        return hipGetLastError();                                                             \
    }
    ARB_NODE_SHAPES(ARB_QT_CASE)
#undef ARB_QT_CASE
#undef L_QT_NI
#undef L_QT
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
                           const void* sent, int32_t has_comps, int32_t has_norm,
                           const double* q, int64_t n, int64_t ldq, const ArbOut& o,
                           uint32_t* key, int32_t cellbits, int32_t* hist, int32_t hshift,
                           void* qrec, int32_t qstride, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t ntiles = arb_rs_ntiles(n);
#define L_KEYS(DD, T, TC, TO)                                                                 \
    hipLaunchKernelGGL((k_keys<DD, T, TC, TO>), dim3((unsigned)ntiles), dim3(KEYS_THREADS), 0,   \
                       s, g, B, (const TC*)sent, has_comps, has_norm, q, n, ldq, o, key,         \
                       cellbits, hist, hshift, ntiles, (float4*)qrec, qstride)
    if (g.ndim == 3) ARB_DISPATCH(3, prec, out_f64, L_KEYS);
    else ARB_DISPATCH(4, prec, out_f64, L_KEYS);
#undef L_KEYS
    return hipGetLastError();
}
