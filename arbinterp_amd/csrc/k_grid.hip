// Grid detection: per-axis unique coordinates, rank of every row, scatter of
// the value columns to canonical x-fastest node order.
//
// Replaces getFieldParams (ARBInterp.py:528-568, 4D :1264-1320), which sorts
// the rows lexicographically with three (four) argsorts and counts the axis
// lengths along the first node's axis lines.  For a complete grid the sorted
// row order is unique, so ranking each coordinate in its axis' sorted unique
// values gives the same canonical order without a full-row sort.
//
// Unique values are found with an open-addressing hash set per axis (64-bit
// atomicCAS on the coordinate's bit pattern, -0.0 folded into +0.0 as the
// reference's `==` does); the unique sets are small and are sorted on the host.
#include "arb_internal.h"

namespace {

constexpr unsigned long long kEmpty = 0xFFFFFFFFFFFFFFFFull;  // a NaN pattern

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

__global__ void k_finite_check(const double* __restrict__ rows, int64_t nrows, int32_t ncols,
                               int64_t ld, int32_t ndim, int32_t* flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    bool bad = false;
    for (int a = 0; a < ndim; ++a) bad |= !isfinite(rows[i * ld + a]);
    if (bad) flag[0] = 1;
}

__global__ void k_hash_insert(const double* __restrict__ rows, int64_t nrows, int64_t ld,
                              int32_t axis, unsigned long long* tab, int64_t cap) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    double x = rows[i * ld + axis] + 0.0;   // -0.0 -> +0.0
    unsigned long long key = (unsigned long long)__double_as_longlong(x);
    unsigned long long mask = (unsigned long long)cap - 1;
    unsigned long long slot = mix64(key) & mask;
    // The table holds at most nrows distinct keys in cap >= 2*nrows slots, so
    // the probe always reaches an empty slot or the key itself.  A grid has
    // few distinct coordinates per axis and every row repeats one of them: a
    // plain read first finds the key already present without an atomic (a
    // filled slot never changes; a stale empty read falls through to the CAS).
    for (int64_t probe = 0; probe < cap; ++probe) {
        const unsigned long long cur = __hip_atomic_load(&tab[slot], __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
        if (cur == key) return;
        if (cur == kEmpty) {
            const unsigned long long old = atomicCAS(&tab[slot], kEmpty, key);
            if (old == kEmpty || old == key) return;
        }
        slot = (slot + 1) & mask;
    }
}

__global__ void k_hash_compact(const unsigned long long* __restrict__ tab, int64_t cap,
                               double* out, unsigned int* count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    unsigned long long k = tab[i];
    if (k != kEmpty) {
        unsigned int p = atomicAdd(count, 1u);
        out[p] = __longlong_as_double((long long)k);
    }
}

// Index of x in the ascending array ax[0..n-1] (exact match expected).
__device__ __forceinline__ int64_t rank_of(const double* __restrict__ ax, int64_t n, double x) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (ax[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

struct AxisPtrs { const double* p[ARB_MAXD]; };

// mode: ARB_MODE_*; vcol0 = first value column (= ndim).
// Interpolant order in `vals`: scalar -> [f]; vector -> [Bx,By,Bz];
// norm -> [|B|]; both -> [Bx,By,Bz,|B|].
__global__ void k_rank_scatter(const double* __restrict__ rows, int64_t nrows, int64_t ld,
                               AxisPtrs axes, ArbGeom g, int32_t mode, int32_t vcol0,
                               double* __restrict__ vals, int32_t* __restrict__ seen) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const double* r = rows + i * ld;
    int64_t node = 0;
    for (int a = 0; a < g.ndim; ++a) {
        int64_t k = rank_of(axes.p[a], g.n[a], r[a] + 0.0);
        node += k * g.nstride[a];
    }
    atomicAdd(&seen[node], 1);
    const int64_t nn = g.nnodes;
    const bool f32_norm = (mode & ARB_MODE_F32_NORM) != 0;
    mode &= ~ARB_MODE_F32_NORM;
    if (mode == ARB_MODE_SCALAR) {
        vals[node] = r[vcol0];
        return;
    }
    double bx = r[vcol0], by = r[vcol0 + 1], bz = r[vcol0 + 2];
    // |B| as numpy.linalg.norm(axis=1) forms it: sqrt((x*x + y*y) + z*z);
    // compiled without FP contraction, so each product is rounded.  A float32
    // source array is normed in float32 (products, sums and the square root
    // each rounded to float32); the float32 square root is the FP64 one
    // rounded once more, which is exact for float32 operands (53 >= 2*24 + 2).
    double bn;
    if (f32_norm) {
        const float x = (float)bx, y = (float)by, z = (float)bz;
        const float s2 = (x * x + y * y) + z * z;
        bn = (double)(float)sqrt((double)s2);
    } else {
        bn = sqrt((bx * bx + by * by) + bz * bz);
    }
    if (mode == ARB_MODE_NORM) {
        vals[node] = bn;
        return;
    }
    vals[node] = bx;
    vals[nn + node] = by;
    vals[2 * nn + node] = bz;
    if (mode == ARB_MODE_BOTH) vals[3 * nn + node] = bn;
}

__global__ void k_check_seen(const int32_t* __restrict__ seen, int64_t nnodes, int32_t* flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    if (seen[i] != 1) flag[0] = 1;
}

// ---------------------------------------------------------------------------
// Boundary extension (opt-in; no reference counterpart: README.md:35 plans
// boundary conditions of the Neumann or Dirichlet type).  One ghost node
// layer per face: the extended grid has n + 2 nodes per axis, the original
// node i at index i + 1.  Axes are extended one after another (x, y, z[, t]);
// the faces of axis a span the already extended axes b < a and the original
// nodes of the axes b > a, so edge and corner ghosts follow from the same
// per-axis rule (indices of the extended axis, m = n + 2 nodes):
//   Neumann   v(0) = v(2)               v(m-1) = v(m-3)
//   Dirichlet v(0) = 2 v(1) - v(2)      v(m-1) = 2 v(m-2) - v(m-3)
// i.e. an even (zero normal slope at the edge node) or odd (the edge node's
// value held, one-sided slope) reflection about the edge node.
// ---------------------------------------------------------------------------
__global__ void k_pad_copy(const double* __restrict__ src, ArbGeom go, double* __restrict__ dst,
                           ArbGeom gn, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t k = i / go.nnodes;
    int64_t r = i - k * go.nnodes;
    int64_t o = k * gn.nnodes;
    for (int a = 0; a < go.ndim; ++a) {
        const int64_t c = r % go.n[a];
        r /= go.n[a];
        o += (c + 1) * gn.nstride[a];
    }
    dst[o] = src[i];
}

__global__ void k_pad_axis(double* __restrict__ v, ArbGeom gn, int32_t axis, int32_t bc,
                           int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t r = i >> 1;
    const bool hi = (i & 1) != 0;
    int64_t off = 0;
    for (int b = 0; b < gn.ndim; ++b) {
        if (b == axis) continue;
        const int64_t ext = b < axis ? gn.n[b] : gn.n[b] - 2;
        const int64_t c = r % ext;
        r /= ext;
        off += (b < axis ? c : c + 1) * gn.nstride[b];
    }
    off += r * gn.nnodes;                           // r: the interpolant
    const int64_t st = gn.nstride[axis], m = gn.n[axis];
    const int64_t ghost = off + (hi ? m - 1 : 0) * st;
    const int64_t edge = off + (hi ? m - 2 : 1) * st;
    const int64_t next = off + (hi ? m - 3 : 2) * st;
    const double e = v[edge], nx = v[next];
    v[ghost] = bc == ARB_BC_NEUMANN ? nx : 2.0 * e - nx;
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t arb_launch_finite_check(const double* rows, int64_t nrows, int32_t ncols, int64_t ld,
                                   int32_t ndim, int32_t* flag, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finite_check, dim3(grid1(nrows, 256)), dim3(256), 0, s, rows, nrows,
                       ncols, ld, ndim, flag);
    return hipGetLastError();
}

hipError_t arb_launch_hash_insert(const double* rows, int64_t nrows, int64_t ld, int32_t axis,
                                  unsigned long long* tab, int64_t cap, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hash_insert, dim3(grid1(nrows, 256)), dim3(256), 0, s, rows, nrows, ld,
                       axis, tab, cap);
    return hipGetLastError();
}

hipError_t arb_launch_hash_compact(const unsigned long long* tab, int64_t cap, double* out,
                                   unsigned int* count, hipStream_t s) {
    hipLaunchKernelGGL(k_hash_compact, dim3(grid1(cap, 256)), dim3(256), 0, s, tab, cap, out,
                       count);
    return hipGetLastError();
}

hipError_t arb_launch_rank_scatter(const double* rows, int64_t nrows, int64_t ld,
                                   const double* const* axes_dev, const ArbGeom& g,
                                   int32_t mode, int32_t vcol0, double* vals, int32_t ninterp,
                                   int32_t* seen, hipStream_t s) {
    (void)ninterp;
    AxisPtrs ap{};
    for (int a = 0; a < g.ndim; ++a) ap.p[a] = axes_dev[a];
    hipLaunchKernelGGL(k_rank_scatter, dim3(grid1(nrows, 256)), dim3(256), 0, s, rows, nrows, ld,
                       ap, g, mode, vcol0, vals, seen);
    return hipGetLastError();
}

hipError_t arb_launch_check_seen(const int32_t* seen, int64_t nnodes, int32_t* flag,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_check_seen, dim3(grid1(nnodes, 256)), dim3(256), 0, s, seen, nnodes,
                       flag);
    return hipGetLastError();
}

hipError_t arb_launch_pad_boundary(const double* src, const ArbGeom& go, double* dst,
                                   const ArbGeom& gn, int32_t ninterp, int32_t bc,
                                   hipStream_t s) {
    const int64_t total = go.nnodes * ninterp;
    hipLaunchKernelGGL(k_pad_copy, dim3(grid1(total, 256)), dim3(256), 0, s, src, go, dst, gn,
                       total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int a = 0; a < gn.ndim; ++a) {
        int64_t count = 2 * (int64_t)ninterp;
        for (int b = 0; b < gn.ndim; ++b)
            if (b != a) count *= b < a ? gn.n[b] : gn.n[b] - 2;
        hipLaunchKernelGGL(k_pad_axis, dim3(grid1(count, 256)), dim3(256), 0, s, dst, gn, a, bc,
                           count);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
