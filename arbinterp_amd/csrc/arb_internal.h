// Internal declarations shared by the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "../../include/arbinterp.h"

#define ARB_MAXD 4

// Grid constants as the kernels see them (by value).
struct ArbGeom {
    double lo[ARB_MAXD];      // IntMin per axis
    double hi[ARB_MAXD];      // IntMax per axis
    double h[ARB_MAXD];       // spacing |node0 - node1|
    int64_t n[ARB_MAXD];      // nodes per axis
    int64_t npos[ARB_MAXD];   // cells per axis = n - 3
    int64_t nstride[ARB_MAXD];// node strides (x fastest)
    int64_t nc;               // number of cells
    int64_t nnodes;
    int32_t ndim;
    int32_t red32;            // reduce in float32 (float32 field and queries)
};

// Outputs of one query launch.
struct ArbOut {
    void* comps;      // n x 3
    void* norms;      // n
    void* grads;      // n x d
    int64_t* cells;   // n (queryInds), nullable
    double* qmut;     // positions to NaN-fill for out-of-range rows (nullable)
    int64_t ldq;      // row stride of the positions (also the fill width)
    int32_t* flag;    // set to 1 when a row reduced past the sentinel
    int32_t packed;   // outputs are records of `packed` elements in `comps`
};

// Width (elements) of a packed output record: [comps(3)] [value, grad(d)],
// padded to 4 or 8 so every record is one aligned 16/32/64-byte block.
__host__ __device__ inline int arb_packed_width(int has_comps, int has_norm, int d) {
    const int w = (has_comps ? 3 : 0) + (has_norm ? 1 + d : 0);
    return w <= 4 ? 4 : 8;
}

// Cell blocks used to bucket queries: a bucket is a block of b[0] x b[1] x
// b[2] [x b[3]] cells; its coefficients are staged in LDS together.
struct ArbBuckets {
    int32_t b[ARB_MAXD];     // block edge (cells) per axis, a power of two
    int32_t bsh[ARB_MAXD];   // log2 of b
    int32_t lsh[ARB_MAXD];   // shift of axis a inside the local cell index
    int64_t nb[ARB_MAXD];    // blocks per axis
    int64_t nbs[ARB_MAXD];   // bucket-index stride of axis a
    int64_t nbt;             // number of buckets
    int32_t cellvec;         // 16-byte vectors per cell of coefficients
    int32_t cvsh;            // log2(cellvec) when cellvec is a power of two, else -1
    int32_t cellstride;      // LDS stride per cell in 16-byte vectors
    int32_t ncell;           // cells per block
};

// ---------------------------------------------------------------------------
// Cell reduction (ARBInterp.py:350-373, 1069-1092): strict bounds test, then
// iu = (q - IntMin)/h in IEEE FP64, ix = floor(iu), local = iu - ix, flat cell
// index ix + iy*npx + iz*npx*npy [+ it*npx*npy*npz].  The reference forms the
// flat index in FP64; every term is an integer below 2^53, so the int64 form
// is identical.  Rows outside the box (or with a NaN coordinate) map to the
// sentinel cell nc with NaN local coordinates.
// Returns the unclamped flat index; *oob is set for rows strictly outside.
//
// With g.red32 (a float32 field queried with float32 positions) the reduction
// runs in float32, as numpy does for those dtypes (ARBInterp.py:358-373 with
// float32 IntMin and h, :547-556): iu = fl32(fl32(q - IntMin) / h32), h32 the
// float32 spacing, ix = floor(iu), local = iu - ix in float32; the flat index
// is formed from the integer ix exactly as before.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void arb_reduce_axis(double x, const ArbGeom& g, int a, bool out,
                                                double* iu, double* fl, double* loc) {
    if (g.red32) {
        const float uf = out ? NAN : ((float)x - (float)g.lo[a]) / (float)g.h[a];
        const float ff = floorf(uf);
        *iu = (double)uf;
        *fl = (double)ff;
        *loc = (double)(uf - ff);
    } else {
        *iu = out ? NAN : (x - g.lo[a]) / g.h[a];
        *fl = floor(*iu);
        *loc = *iu - *fl;
    }
}

template <int D>
__device__ __forceinline__ int64_t arb_reduce(const double* __restrict__ q, const ArbGeom& g,
                                              double* __restrict__ loc, bool* oob) {
    bool out = false;
    double x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        x[a] = q[a];
        out |= (x[a] < g.lo[a]) | (x[a] > g.hi[a]);
    }
    *oob = out;
    int64_t flat = 0, mul = 1;
    bool bad = out;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double iu, fl;
        arb_reduce_axis(x[a], g, a, out, &iu, &fl, &loc[a]);
        bad |= isnan(iu);
        int64_t ia = bad ? 0 : (int64_t)(uint32_t)fl;   // 0 <= fl <= npos < 2^32
        flat += ia * mul;
        mul *= g.npos[a];
    }
    return bad ? g.nc : flat;
}

// Same reduction, also returning the per-axis cell coordinates floor(iu)
// (valid only when the returned flat index is below nc).
template <int D>
__device__ __forceinline__ int64_t arb_reduce_axes(const double* __restrict__ q, const ArbGeom& g,
                                                   double* __restrict__ loc, int64_t* ia,
                                                   bool* oob) {
    bool out = false;
    double x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        x[a] = q[a];
        out |= (x[a] < g.lo[a]) | (x[a] > g.hi[a]);
    }
    *oob = out;
    int64_t flat = 0, mul = 1;
    bool bad = out;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double iu, fl;
        arb_reduce_axis(x[a], g, a, out, &iu, &fl, &loc[a]);
        bad |= isnan(iu);
        ia[a] = bad ? 0 : (int64_t)(uint32_t)fl;        // 0 <= fl <= npos < 2^32
        flat += ia[a] * mul;
        mul *= g.npos[a];
    }
    return bad ? g.nc : flat;
}

// ---------------------------------------------------------------------------
// Tensor Horner evaluation of p(u,v,w[,t]) = sum_i a_i u^ex v^ey w^ez [t^et],
// i = ex + 4ey + 16ez [+ 64et].  Mathematically equal to the reference's
// sum over the tile/repeat power basis (ARBInterp.py:380-397, 493-521,
// 1097-1127, 1223-1258); the derivative of the basis along each axis is
// carried alongside (Horner's rule for p and p').  All arithmetic is explicit
// fma() so the result does not depend on compiler contraction.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T h4(T c0, T c1, T c2, T c3, T u) {
    return fma(fma(fma(c3, u, c2), u, c1), u, c0);
}
// value and d/du of c0 + c1 u + c2 u^2 + c3 u^3
template <typename T>
__device__ __forceinline__ void h4d(T c0, T c1, T c2, T c3, T u, T& p, T& dp) {
    T b = fma(c3, u, c2);
    T d = fma(c3, u, b);
    b = fma(b, u, c1);
    d = fma(d, u, b);
    p = fma(b, u, c0);
    dp = d;
}

// value only, 3D; a: 64 coefficients
template <typename T, typename C>
__device__ __forceinline__ T arb_val3(const C* a, T u, T v, T w) {
    T q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T p[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const C* c = a + 16 * k + 4 * j;
            p[j] = h4<T>((T)c[0], (T)c[1], (T)c[2], (T)c[3], u);
        }
        q[k] = h4<T>(p[0], p[1], p[2], p[3], v);
    }
    return h4<T>(q[0], q[1], q[2], q[3], w);
}

// value and gradient in cuboid units, 3D
template <typename T, typename C>
__device__ __forceinline__ void arb_valgrad3(const C* a, T u, T v, T w, T& val, T* g) {
    T q[4], qu[4], qv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T p[4], pu[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const C* c = a + 16 * k + 4 * j;
            h4d<T>((T)c[0], (T)c[1], (T)c[2], (T)c[3], u, p[j], pu[j]);
        }
        h4d<T>(p[0], p[1], p[2], p[3], v, q[k], qv[k]);
        qu[k] = h4<T>(pu[0], pu[1], pu[2], pu[3], v);
    }
    T dw;
    h4d<T>(q[0], q[1], q[2], q[3], w, val, dw);
    g[0] = h4<T>(qu[0], qu[1], qu[2], qu[3], w);
    g[1] = h4<T>(qv[0], qv[1], qv[2], qv[3], w);
    g[2] = dw;
}

// 4D: a has 256 coefficients; reduce over x, y, z (3D blocks), then t.
template <typename T, typename C>
__device__ __forceinline__ T arb_val4(const C* a, T u, T v, T w, T t) {
    T r[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) r[l] = arb_val3<T, C>(a + 64 * l, u, v, w);
    return h4<T>(r[0], r[1], r[2], r[3], t);
}

template <typename T, typename C>
__device__ __forceinline__ void arb_valgrad4(const C* a, T u, T v, T w, T t, T& val, T* g) {
    T r[4], ru[4], rv[4], rw[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        T gg[3];
        arb_valgrad3<T, C>(a + 64 * l, u, v, w, r[l], gg);
        ru[l] = gg[0];
        rv[l] = gg[1];
        rw[l] = gg[2];
    }
    T dt;
    h4d<T>(r[0], r[1], r[2], r[3], t, val, dt);
    g[0] = h4<T>(ru[0], ru[1], ru[2], ru[3], t);
    g[1] = h4<T>(rv[0], rv[1], rv[2], rv[3], t);
    g[2] = h4<T>(rw[0], rw[1], rw[2], rw[3], t);
    g[3] = dt;
}

// ---------------------------------------------------------------------------
// Host-side launchers (defined in the .hip kernel files).
// ---------------------------------------------------------------------------
struct ArbField;

hipError_t arb_launch_finite_check(const double* rows, int64_t nrows, int32_t ncols, int64_t ld,
                                   int32_t ndim, int32_t* flag, hipStream_t s);
hipError_t arb_launch_hash_insert(const double* rows, int64_t nrows, int64_t ld, int32_t axis,
                                  unsigned long long* tab, int64_t cap, hipStream_t s);
hipError_t arb_launch_hash_compact(const unsigned long long* tab, int64_t cap, double* out,
                                   unsigned int* count, hipStream_t s);
hipError_t arb_launch_rank_scatter(const double* rows, int64_t nrows, int64_t ld,
                                   const double* const* axes_dev, const ArbGeom& g,
                                   int32_t mode, int32_t vcol0, double* vals, int32_t ninterp,
                                   int32_t* seen, hipStream_t s);
hipError_t arb_launch_check_seen(const int32_t* seen, int64_t nnodes, int32_t* flag, hipStream_t s);
// Extend the node values src ([ninterp][go.nnodes]) to the grid gn (one ghost
// layer per face) in dst ([ninterp][gn.nnodes]) with boundary rule bc.
hipError_t arb_launch_pad_boundary(const double* src, const ArbGeom& go, double* dst,
                                   const ArbGeom& gn, int32_t ninterp, int32_t bc,
                                   hipStream_t s);

hipError_t arb_launch_coef(const ArbGeom& g, int32_t prec, const double* vals, int32_t ninterp,
                           void* table, hipStream_t s);
hipError_t arb_launch_sentinel(void* table, int32_t f32, int64_t nc, int32_t ninterp, int32_t K,
                               int32_t nan, hipStream_t s);

hipError_t arb_launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
hipError_t arb_launch_copy16(void* dst, const void* src, int64_t n16, hipStream_t s);
hipError_t arb_launch_bucket(const ArbGeom& g, const ArbBuckets& B, int32_t prec, int32_t out_f64,
                             const void* table, int32_t ninterp, int32_t has_comps,
                             int32_t has_norm, const double* q, int64_t n, int64_t ldq,
                             const ArbOut& o, int32_t* bkt, int32_t* rank, int32_t* count,
                             hipStream_t s);
hipError_t arb_launch_scan(const int32_t* count, int32_t* offs, int64_t nb, hipStream_t s);
hipError_t arb_launch_unpack(int32_t ndim, int32_t out_f64, int32_t has_comps, int32_t has_norm,
                             const void* rec, void* comps, void* norms, void* grads, int64_t n,
                             hipStream_t s);
// Rows of one bucket a workgroup of the bucketed evaluation takes at most.
// A bucket holding more (queries clustered in a few cells) is shared by the
// workgroups nbt + o that cover its sorted positions past the first chunk,
// o = position / chunk; the launchers size their grids nbt + ceil(n /
// chunk).  hot[0] (set with the bucket bounds) is nonzero when some bucket
// holds more than a chunk; when it is zero those workgroups end at once.
constexpr int32_t ARB_EVAL_CHUNK = 512;
// Radix-sort tiles: ARB_RS_THREADS threads x ARB_RS_ITEMS pairs (8192; the
// longer digit runs of the larger tile make the scattered writes wider).
constexpr int ARB_RS_THREADS = 512;
constexpr int ARB_RS_ITEMS = 16;
// Keys (bucket, or bucket and cell) of n queries, and the tile histograms of
// the keys' 8-bit digit (key >> hshift) & 255 (hist[digit * ntiles + tile]:
// hshift 0 for the radix sort, bits - 8 for the grouping sort).  The value of key i is
// the row index i, implicit (the radix sort's index_vals).  Rows outside the
// box are evaluated here from the sentinel row `sent` (ninterp x K values).
// With qrec (3D node-tile path) every row i inside the box also gets its
// 16-byte record, at qrec + i * qstride (in 16-byte units): FP32 local
// coordinates and the local cell in its bucket.
hipError_t arb_launch_keys(const ArbGeom& g, const ArbBuckets& B, int32_t prec, int32_t out_f64,
                           const void* sent, int32_t has_comps, int32_t has_norm,
                           const double* q, int64_t n, int64_t ldq, const ArbOut& o,
                           uint32_t* key, int32_t cellbits, int32_t* hist, int32_t hshift,
                           void* qrec, int32_t qstride, hipStream_t s);
int64_t arb_scan_tmp_elems(int64_t len);
hipError_t arb_launch_scan_long(const int32_t* in, int32_t* out, int64_t len, int32_t* tmp,
                                int32_t* tmp_scan, hipStream_t s);
int64_t arb_rs_ntiles(int64_t n);
// scratch ints of arb_launch_radix_sort and arb_launch_group_sort; the last
// one is the hot flag of the bucketed evaluation (ARB_EVAL_CHUNK)
int64_t arb_rs_first_elems(int64_t ngroups);
hipError_t arb_launch_radix_sort(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int64_t n,
                                 int bits, int32_t* hist, int32_t* hscan, int hist0_ready,
                                 int index_vals, int32_t* offs, int64_t ngroups, int32_t gshift,
                                 int32_t* first, int* in_first, int stable, hipStream_t s);
// Grouping of the rows by key for the bucketed evaluation (k_sort.hip): keys
// below 2^bits with 9 <= bits <= 16 and n < 2^24 (arb_group_sort_ok); no
// order inside a key.  kcnt: arb_rs_first_elems(nbt) ints of scratch, whose
// last one becomes the hot flag (a key below nbt holds more than
// ARB_EVAL_CHUNK rows).
int arb_group_sort_ok(int64_t n, int bits);
hipError_t arb_launch_group_sort(const uint32_t* keys, uint32_t* ent, int32_t* perm, int64_t n,
                                 int bits, const int32_t* hist, int32_t* hscan, int32_t* offs,
                                 int64_t nbt, int32_t* kcnt, hipStream_t s);
hipError_t arb_launch_scatter(const int32_t* bkt, const int32_t* rank, const int32_t* offs,
                              int32_t* perm, int64_t n, hipStream_t s);
size_t arb_fused_lds(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp, int32_t threads);
hipError_t arb_launch_eval_fused(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                 int32_t out_f64, const double* vals, int32_t ninterp,
                                 int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                 int64_t ldq, const ArbOut& o, const int32_t* offs,
                                 const int32_t* perm, const int32_t* hot, int32_t block_threads,
                                 hipStream_t s);
// 4D bucketed query without the table (coefficients formed per bucket with
// the 4D table build's arithmetic); arb_fused4_lds is 0 where it does not
// apply.
size_t arb_fused4_lds(const ArbGeom& g, const ArbBuckets& B);
hipError_t arb_launch_eval_fused4(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                  int32_t out_f64, const double* vals, int32_t ninterp,
                                  int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                  int64_t ldq, const ArbOut& o, const int32_t* offs,
                                  const int32_t* perm, const int32_t* hot, int32_t block_threads,
                                  hipStream_t s);
bool arb_nodes_ok(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp, int32_t has_comps,
                  int32_t has_norm);
int64_t arb_tiles_bytes(const ArbBuckets& B, int32_t ninterp, int64_t* head_offset);
hipError_t arb_launch_build_tiles(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                  const double* vals, void* tiles, hipStream_t s);
hipError_t arb_launch_query_tiles(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                  int32_t out_f64, const void* tiles, const void* sent,
                                  int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                  int64_t ldq, const ArbOut& o, hipStream_t s);
// Backward of the last interpolant from the node tiles (mixed 3D; ninterp 4
// or 1): d/dq, and the node adjoint's per-query record.  perm (or nullptr):
// the queries in grouped order; thread s then takes query perm[s] and writes
// its record at position s.
hipError_t arb_launch_bwd_tiles(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                const void* tiles, const double* q, int64_t n, int64_t ldq,
                                const double* gv, const double* gg, double* dq,
                                const int32_t* perm, double* rec, hipStream_t s);
// Node-tile evaluation of the bucketed path; qrec: the per-row records k_keys
// wrote (16 bytes each at a stride of qstride x 16 bytes: FP32 local
// coordinates, local cell in the bucket).  The records may live in the
// output records themselves (row i's output overwrites its own record).
hipError_t arb_launch_eval_nodes(const ArbGeom& g, const ArbBuckets& B, int32_t ninterp,
                                 int32_t out_f64, const void* tiles, const void* qrec,
                                 int32_t qstride, int64_t n,
                                 const ArbOut& o, const int32_t* offs, const int32_t* perm,
                                 const int32_t* hot, int32_t block_threads, hipStream_t s);
hipError_t arb_launch_eval_bucket(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                  int32_t out_f64, const void* table, int32_t ninterp,
                                  int32_t has_comps, int32_t has_norm, const double* q,
                                  int64_t n, int64_t ldq, const ArbOut& o, const int32_t* offs,
                                  const int32_t* perm, const int32_t* hot, int32_t block_threads,
                                  hipStream_t s);
hipError_t arb_launch_eval_perm(const ArbGeom& g, const ArbBuckets& B, int32_t prec,
                                int32_t out_f64, const void* table, int32_t ninterp,
                                int32_t has_comps, int32_t has_norm, const double* q, int64_t n,
                                int64_t ldq, const ArbOut& o, const int32_t* offs,
                                const int32_t* perm, hipStream_t s);
// The node adjoint's cell key per query (flat cell, nc outside the box).
hipError_t arb_launch_bwd_keys(const ArbGeom& g, const double* q, int64_t n, int64_t ldq,
                               int32_t* key, hipStream_t s);
// Backward of interpolant kn from the table: d/dq, and for the node adjoint
// (key, rec non-null) the query's cell key and (3D) its record at row i.
// perm (or nullptr): the queries in grouped order; thread s then takes
// query perm[s].
hipError_t arb_launch_bwd_points(const ArbGeom& g, int32_t prec, const void* table,
                                 int32_t ninterp, int32_t kn, const double* q, int64_t n,
                                 int64_t ldq, const double* gv, const double* gg, double* dq,
                                 const int32_t* perm, int32_t* key, double* rec,
                                 hipStream_t s);
// part: arb_bwd_part_elems(g, n) doubles of scratch (the sums of chunks of
// cells holding many queries).
int64_t arb_bwd_part_elems(const ArbGeom& g, int64_t n);
// rec_grouped: the 3D records sit at the queries' grouped positions (else by
// row, reached through perm).
hipError_t arb_launch_bwd_cells(const ArbGeom& g, const double* A, const double* q, int64_t ldq,
                                const double* gv, const double* gg, const double* rec,
                                int32_t rec_grouped, const int32_t* offs, const int32_t* perm,
                                int64_t n, double* part, double* bbar, hipStream_t s);
hipError_t arb_launch_bwd_nodes(const ArbGeom& g, const double* bbar, double* dval, hipStream_t s);
// Doubles of node-adjoint scratch (bbar of arb_launch_bwd_cells / _nodes).
int64_t arb_bwd_scratch(const ArbGeom& g);
size_t arb_batch_field_desc_size();
void arb_batch_fill_desc(void* dst, const ArbGeom& g, const void* table, int64_t q0, int64_t nq,
                         int64_t item0);
hipError_t arb_launch_batch_query(int32_t ndim, int32_t prec, int32_t out_f64, const void* fields,
                                  int32_t nfields, int64_t nitems, int32_t chunk, int32_t threads,
                                  int32_t nblocks, int32_t ninterp, int32_t has_comps,
                                  int32_t has_norm, const double* q, int64_t ldq, void* comps,
                                  void* norms, void* grads, int32_t* flag, unsigned int* counter,
                                  hipStream_t s);
hipError_t arb_launch_rk4(const ArbGeom& g, int32_t prec, const void* table, int32_t ninterp,
                          int32_t kn, double* X, double* V, int64_t n, double dt, int32_t nsteps,
                          hipStream_t s);
// Direct kernel for float32 fields queried with float32 positions (g.red32):
// float32 reduction, and the power basis with numpy's float32 powers.
hipError_t arb_launch_query_direct_red32(const ArbGeom& g, int32_t prec, int32_t out_f64,
                                         const void* table, int32_t ninterp, int32_t has_comps,
                                         int32_t has_norm, const double* q, int64_t n,
                                         int64_t ldq, const ArbOut& o, hipStream_t s);
hipError_t arb_launch_query_direct(const ArbGeom& g, int32_t prec, int32_t out_f64,
                                   const void* table, int32_t ninterp, int32_t has_comps,
                                   int32_t has_norm, const double* q, int64_t n, int64_t ldq,
                                   const ArbOut& o, hipStream_t s);
