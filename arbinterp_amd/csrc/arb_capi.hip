// Host runtime and C ABI of libarbinterp.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "arb_internal.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                 \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess)                                                     \
            return fail(e_ == hipErrorOutOfMemory ? ARB_ENOMEM : ARB_EHIP,         \
                        "%s failed: %s", #x, hipGetErrorString(e_));              \
    } while (0)

// Device buffer freed on scope exit unless released.
struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
    void* release() { void* r = p; p = nullptr; return r; }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// Operator matrices, built exactly on the host (makeAMatrix,
// ARBInterp.py:107-175 / :726-878).
//
// B: rows (block, corner) -> monomial columns; inv(B) is the tensor product of
// the 1D cubic Hermite map (p0, p1, p0', p1') -> (a0, a1, a2, a3), so its
// entries are integers (|.| <= 27 in 3D, 81 in 4D) and are formed exactly.
// D: tensor products of the central difference (-1/2, 0, 1/2) in cuboid units;
// in 4D the d4/dxdydzdt block carries the reference's row shift
// (ARBInterp.py:860: rows 241..255 take the stencils of corners 0..14, row 240
// stays zero).  A = inv(B).D is a sum of integer x dyadic products, exact in
// FP64, equal to np.matmul(np.linalg.inv(B), D).
// ---------------------------------------------------------------------------
namespace {

// b-vector blocks: subsets of axes in order of size, lexicographic.
std::vector<int> deriv_blocks(int d) {
    std::vector<int> out;
    for (int order = 0; order <= d; ++order)
        for (int m = 0; m < (1 << d); ++m)
            if (__builtin_popcount(m) == order) out.push_back(m);
    // within one order the reference lists x<y<z<t combinations
    // lexicographically: (x,y),(x,z),(x,t),(y,z),(y,t),(z,t); sort by the
    // tuple of axis ids.
    auto key = [&](int m) {
        std::vector<int> ax;
        for (int a = 0; a < d; ++a) if (m >> a & 1) ax.push_back(a);
        return ax;
    };
    std::stable_sort(out.begin(), out.end(), [&](int a, int b) {
        int pa = __builtin_popcount(a), pb = __builtin_popcount(b);
        if (pa != pb) return pa < pb;
        return key(a) < key(b);
    });
    return out;
}

void make_B(int d, std::vector<double>& B) {
    const int K = 1 << (2 * d), nco = 1 << d;
    B.assign((size_t)K * K, 0.0);
    auto blocks = deriv_blocks(d);
    for (int i = 0; i < K; ++i)
        for (size_t b = 0; b < blocks.size(); ++b)
            for (int k = 0; k < nco; ++k) {
                double v = 1.0;
                for (int a = 0; a < d; ++a) {
                    int e = (i >> (2 * a)) & 3;
                    double x = (k >> a) & 1;
                    if (blocks[b] >> a & 1) v *= (e == 0) ? 0.0 : e * std::pow(x, e - 1);
                    else v *= std::pow(x, e);
                }
                B[(size_t)(b * nco + k) * K + i] = v;
            }
}

// 1D Hermite: coefficient e from (p0, p1, d0, d1)
const double kH1[4][4] = {{1, 0, 0, 0}, {0, 0, 1, 0}, {-3, 3, -2, -1}, {2, -2, 1, 1}};

void make_invB(int d, std::vector<double>& Bi) {
    const int K = 1 << (2 * d), nco = 1 << d;
    Bi.assign((size_t)K * K, 0.0);
    auto blocks = deriv_blocks(d);
    for (int i = 0; i < K; ++i)
        for (size_t b = 0; b < blocks.size(); ++b)
            for (int k = 0; k < nco; ++k) {
                double v = 1.0;
                for (int a = 0; a < d; ++a) {
                    int e = (i >> (2 * a)) & 3;
                    int col = ((blocks[b] >> a & 1) ? 2 : 0) + ((k >> a) & 1);
                    v *= kH1[e][col];
                }
                Bi[(size_t)i * K + (b * nco + k)] = v;
            }
}

void make_D(int d, std::vector<double>& D) {
    const int K = 1 << (2 * d), nco = 1 << d;
    D.assign((size_t)K * K, 0.0);
    auto blocks = deriv_blocks(d);
    std::vector<int> C(nco);
    for (int k = 0; k < nco; ++k) {
        int c = 0, p = 1;
        for (int a = 0; a < d; ++a, p *= 4) c += (1 + ((k >> a) & 1)) * p;
        C[k] = c;
    }
    for (size_t b = 0; b < blocks.size(); ++b) {
        std::vector<std::pair<int, double>> taps{{0, 1.0}};
        for (int a = 0; a < d; ++a) {
            if (!(blocks[b] >> a & 1)) continue;
            std::vector<std::pair<int, double>> nt;
            int st = 1 << (2 * a);
            for (auto& t : taps)
                for (int s = -1; s <= 1; s += 2) nt.push_back({t.first + s * st, t.second * 0.5 * s});
            taps.swap(nt);
        }
        int r0 = (int)b * nco, nrow = nco, cshift = 0;
        if (d == 4 && __builtin_popcount(blocks[b]) == 4) { r0 += 1; nrow -= 1; cshift = 0; }
        for (int r = 0; r < nrow; ++r)
            for (auto& t : taps) D[(size_t)(r0 + r) * K + (C[r + cshift] + t.first)] = t.second;
    }
}

void matmul(int K, const std::vector<double>& X, const std::vector<double>& Y,
            std::vector<double>& Z) {
    Z.assign((size_t)K * K, 0.0);
    for (int i = 0; i < K; ++i)
        for (int r = 0; r < K; ++r) {
            double x = X[(size_t)i * K + r];
            if (x == 0.0) continue;
            for (int j = 0; j < K; ++j) Z[(size_t)i * K + j] += x * Y[(size_t)r * K + j];
        }
}

}  // namespace

// ---------------------------------------------------------------------------
// field handle
// ---------------------------------------------------------------------------
struct arb_field {
    ArbGeom g{};
    int32_t mode = 0, precision = 0, ninterp = 0, has_comps = 0, has_norm = 0;
    int32_t table_built = 0;
    int64_t table_limit = 0;    // largest table the caller allows (0: no limit)
    int32_t otf = 0;            // the table does not fit: coefficients formed on the fly
    int32_t sentinel_set = 0;   // sentinel row `sent` written
    int32_t table_loaded = 0;   // rows came from arb_table_load
    int32_t bc = ARB_BC_NONE;   // boundary extension (arb_field_set_boundary)
    double edge_lo[ARB_MAXD] = {0, 0, 0, 0};   // first / last node coordinate per axis
    double edge_hi[ARB_MAXD] = {0, 0, 0, 0};
    int device = 0;
    double* vals = nullptr;     // [ninterp][nnodes] FP64, canonical node order
    void* tiles = nullptr;      // node tiles of the buckets of B (node-form path)
    int64_t tiles_bytes = 0;
    int32_t tiles_ok = 0;       // tiles hold the current node values
    void* table = nullptr;      // [nc+1][ninterp][K]: FP64 or FP32, allocated on first build
    int64_t table_bytes = 0;
    void* sent = nullptr;       // [ninterp][K] the sentinel row, for paths without the table
    double* A = nullptr;        // K x K FP64
    int32_t* flag = nullptr;    // device status word
    // bucketed query path: block geometry of the table kernels (B) and of the
    // fused kernel (BF, without the larger blocks that only save a sort pass)
    ArbBuckets B{};
    ArbBuckets BF{};
    int32_t user_block[ARB_MAXD] = {0, 0, 0, 0};
    int32_t eval_threads = 256;
    int32_t* ws_bkt = nullptr;  // per query: bucket (-1 = handled in place)
    int32_t* ws_rank = nullptr; // per query: rank inside its bucket
    int32_t* ws_perm = nullptr; // bucket-sorted query indices
    void* ws_rec = nullptr;     // packed output records of a bucketed call (split afterwards)
    int64_t ws_rec_bytes = 0;
    void* ws_qrec = nullptr;    // per-row records of the node-tile path (16 bytes per row)
    int64_t ws_qrec_n = 0;
    int64_t ws_n = 0;
    int32_t* ws_count = nullptr;
    int32_t* ws_offs = nullptr;
    int64_t ws_nb = 0;
    // radix-sort workspace: keys/values x2, digit histograms, scan temporaries
    uint32_t* rs_k[2] = {nullptr, nullptr};
    int32_t* rs_v[2] = {nullptr, nullptr};
    int32_t* rs_hist = nullptr;    // tile digit histograms
    int32_t* rs_hscan = nullptr;   // their per-digit scans, then the digit totals
    int64_t rs_n = 0;
    // backward workspace
    int32_t* bw_key = nullptr;     // per query: cell (nc outside)
    int32_t* bw_perm = nullptr;    // queries grouped per cell, index order
    uint32_t* bw_k1 = nullptr;     // the stable radix sort's second key and
    int32_t* bw_v1 = nullptr;      // value buffers (n each),
    int32_t* bw_hist = nullptr;    // its tile histograms (256 per tile)
    int32_t* bw_hscan = nullptr;   // and their scans (+ 256 digit totals)
    double* bw_part = nullptr;     // sums of chunks of cells holding many queries
    double* bw_rec = nullptr;      // 3D: per query u, v, w, gv, gg / h (8 doubles)
    int64_t bw_n = 0;
    int32_t* bw_count = nullptr;   // the radix sort's first positions
    int32_t* bw_offs = nullptr;    // nc + 1
    double* bw_bbar = nullptr;     // node-adjoint scratch: 3D block tiles, 4D K x nc
    // particle push graph
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t push_exec = nullptr;
    double push_key_d = 0.0;
    int64_t push_key[5] = {0, 0, 0, 0, 0};
    hipEvent_t ev[6] = {};
    int32_t ev_ready = 0;
    float last_ms[5] = {0, 0, 0, 0, 0};
    // single-point mailbox: pinned host record the point kernel reads and
    // writes in place (host address, and the same memory as the GPU sees it)
    double* pin = nullptr;
    double* pin_dev = nullptr;
    // batches holding this field (destroying the field invalidates them)
    std::vector<arb_batch*> batches;
};

// Choose the cell block of one bucket: the largest of a fixed ladder whose
// coefficients (padded by 16 bytes per cell) fit in 48 KiB of LDS, unless the
// caller set one.  With `upgrade`, when that leaves 2^16 buckets or more (a
// third 8-bit radix pass), a larger block within 80 KiB that brings the
// count below 2^16 is taken instead: on the 128^3 FP64 field this trades
// 0.06 ms of evaluation (fewer workgroups per CU) for a 0.08 ms sort pass.
static int64_t bucket_count(const arb_field* f, const int* dims) {
    int64_t nb = 1;
    for (int a = 0; a < f->g.ndim; ++a) nb *= (f->g.npos[a] + dims[a] - 1) / dims[a];
    return nb;
}

static void choose_buckets(arb_field* f, ArbBuckets& B, bool upgrade) {
    const int d = f->g.ndim;
    const int K = 1 << (2 * d);
    const int esz = f->precision == ARB_PREC_FP64 ? 8 : 4;
    B.cellvec = f->ninterp * K * esz / 16;
    B.cellstride = B.cellvec + 1;
    static const int ladder3[][3] = {{4, 4, 4}, {4, 4, 2}, {4, 2, 2}, {2, 2, 2}, {2, 2, 1}, {2, 1, 1}, {1, 1, 1}};
    static const int ladder4[][4] = {{4, 4, 2, 2}, {4, 2, 2, 2}, {2, 2, 2, 2}, {2, 2, 2, 1},
                                     {2, 2, 1, 1}, {2, 1, 1, 1}, {1, 1, 1, 1}};
    const int max_cells = (48 * 1024) / (B.cellstride * 16);
    int dims[4] = {1, 1, 1, 1};
    if (f->user_block[0] > 0) {
        for (int a = 0; a < d; ++a) dims[a] = f->user_block[a];
    } else {
        const int big_cells = (80 * 1024) / (B.cellstride * 16);
        const int nl = d == 3 ? (int)(sizeof ladder3 / sizeof ladder3[0])
                              : (int)(sizeof ladder4 / sizeof ladder4[0]);
        auto entry = [&](int i, int* e) {
            for (int a = 0; a < 4; ++a) e[a] = a < d ? (d == 3 ? ladder3[i][a] : ladder4[i][a]) : 1;
        };
        auto cells = [&](int i) {
            int e[4];
            entry(i, e);
            return e[0] * e[1] * e[2] * e[3];
        };
        auto count = [&](int i) {
            int e[4];
            entry(i, e);
            return bucket_count(f, e);
        };
        int pick = -1;
        for (int i = 0; i < nl && pick < 0; ++i)
            if (cells(i) <= max_cells) pick = i;
        if (pick < 0) pick = nl - 1;
        if (upgrade && count(pick) >= ((int64_t)1 << 16))
            for (int i = pick - 1; i >= 0; --i)
                if (cells(i) <= big_cells && count(i) < ((int64_t)1 << 16)) {
                    pick = i;
                    break;
                }
        entry(pick, dims);
    }
    B.ncell = 1;
    B.nbt = 1;
    B.cvsh = -1;
    for (int sh = 0; sh < 16; ++sh)
        if ((1 << sh) == B.cellvec) B.cvsh = sh;
    int lsh = 0;
    for (int a = 0; a < ARB_MAXD; ++a) {
        B.b[a] = a < d ? dims[a] : 1;
        B.bsh[a] = 0;
        while ((1 << B.bsh[a]) < B.b[a]) ++B.bsh[a];
        B.lsh[a] = lsh;
        lsh += B.bsh[a];
        B.nb[a] = a < d ? (f->g.npos[a] + B.b[a] - 1) / B.b[a] : 1;
        B.nbs[a] = B.nbt;
        B.ncell *= B.b[a];
        B.nbt *= B.nb[a];
    }
}

static int ensure_workspace(arb_field* f, int64_t n, int64_t nbt) {
    if (n > f->ws_n) {
        if (f->ws_bkt) (void)hipFree(f->ws_bkt);
        if (f->ws_rank) (void)hipFree(f->ws_rank);
        if (f->ws_perm) (void)hipFree(f->ws_perm);
        f->ws_bkt = f->ws_rank = f->ws_perm = nullptr;
        f->ws_n = 0;
        HIPCHK(hipMalloc(&f->ws_bkt, sizeof(int32_t) * n));
        HIPCHK(hipMalloc(&f->ws_rank, sizeof(int32_t) * n));
        HIPCHK(hipMalloc(&f->ws_perm, sizeof(int32_t) * n));
        f->ws_n = n;
    }
    if (n > f->rs_n) {
        for (int k = 0; k < 2; ++k) {
            if (f->rs_k[k]) (void)hipFree(f->rs_k[k]);
            if (f->rs_v[k]) (void)hipFree(f->rs_v[k]);
            f->rs_k[k] = nullptr;
            f->rs_v[k] = nullptr;
        }
        if (f->rs_hist) (void)hipFree(f->rs_hist);
        if (f->rs_hscan) (void)hipFree(f->rs_hscan);
        f->rs_hist = f->rs_hscan = nullptr;
        f->rs_n = 0;
        const int64_t hl = 256 * arb_rs_ntiles(n);
        for (int k = 0; k < 2; ++k) {
            HIPCHK(hipMalloc(&f->rs_k[k], sizeof(uint32_t) * n));
            HIPCHK(hipMalloc(&f->rs_v[k], sizeof(int32_t) * n));
        }
        HIPCHK(hipMalloc(&f->rs_hist, sizeof(int32_t) * hl));
        HIPCHK(hipMalloc(&f->rs_hscan, sizeof(int32_t) * (hl + 256)));   // + digit totals
        f->rs_n = n;
    }
    if (nbt + 1 > f->ws_nb) {
        if (f->ws_count) (void)hipFree(f->ws_count);
        if (f->ws_offs) (void)hipFree(f->ws_offs);
        f->ws_count = f->ws_offs = nullptr;
        f->ws_nb = 0;
        HIPCHK(hipMalloc(&f->ws_count, sizeof(int32_t) * arb_rs_first_elems(nbt)));
        HIPCHK(hipMalloc(&f->ws_offs, sizeof(int32_t) * (nbt + 1)));
        f->ws_nb = nbt + 1;
    }
    return ARB_OK;
}

// Per-row records of the node-tile bucketed path (n x 16 bytes), made on
// first use.
static int ensure_qrec(arb_field* f, int64_t n) {
    if (n > f->ws_qrec_n) {
        if (f->ws_qrec) (void)hipFree(f->ws_qrec);
        f->ws_qrec = nullptr;
        f->ws_qrec_n = 0;
        HIPCHK(hipMalloc(&f->ws_qrec, (size_t)n * 16));
        f->ws_qrec_n = n;
    }
    return ARB_OK;
}

// Room for `bytes` of packed output records, made on first use.
static int ensure_records(arb_field* f, int64_t bytes) {
    if (bytes > f->ws_rec_bytes) {
        if (f->ws_rec) (void)hipFree(f->ws_rec);
        f->ws_rec = nullptr;
        f->ws_rec_bytes = 0;
        HIPCHK(hipMalloc(&f->ws_rec, (size_t)bytes));
        f->ws_rec_bytes = bytes;
    }
    return ARB_OK;
}

static int interp_count(int mode, int* comps, int* norm) {
    switch (mode) {
        case ARB_MODE_SCALAR: *comps = 0; *norm = 1; return 1;
        case ARB_MODE_VECTOR:
        case ARB_MODE_VECTOR_NAN: *comps = 1; *norm = 0; return 3;
        case ARB_MODE_NORM: *comps = 0; *norm = 1; return 1;
        case ARB_MODE_BOTH: *comps = 1; *norm = 1; return 4;
        default: return 0;
    }
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free query (arb_query_noalloc): one fixed array
// per GPU, resident from the library's load until the process ends.  Host
// code never uses the name `scratch` as an address (on the host it is only a
// stand-in for the device symbol): the device address of each GPU's copy is
// looked up once per device with hipGetSymbolAddress and kept in
// g_scratch_dev.  Every call on one GPU shares that GPU's array, so calls
// that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[ARB_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[kMaxDevices] = {};

// Device address of the current device's scratch (looked up on the first
// call on that device).
static int scratch_address(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices)
        return fail(ARB_EINVAL, "device %d is beyond the %d devices the scratch table holds", dev,
                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ARB_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_dev[dev];
    return ARB_OK;
}

extern "C" {

const char* arb_last_error(void) { return g_err.c_str(); }

const char* arb_version(void) { return "arbinterp-mi355x 0.1 (gfx950)"; }

int arb_operator_matrix(int32_t ndim, int32_t which, double* out) {
    if ((ndim != 3 && ndim != 4) || !out || which < 0 || which > 3)
        return fail(ARB_EINVAL, "arb_operator_matrix: bad argument");
    const int K = 1 << (2 * ndim);
    std::vector<double> M, Bi, D;
    if (which == 0) make_B(ndim, M);
    else if (which == 1) make_invB(ndim, M);
    else if (which == 2) make_D(ndim, M);
    else { make_invB(ndim, Bi); make_D(ndim, D); matmul(K, Bi, D, M); }
    std::memcpy(out, M.data(), sizeof(double) * (size_t)K * K);
    return ARB_OK;
}

int arb_field_create(const double* rows_dev, int64_t nrows, int32_t ncols, int64_t ld,
                     int32_t ndim, int32_t mode, int32_t precision, void* stream,
                     arb_field** out) {
    if (!out) return fail(ARB_EINVAL, "arb_field_create: out is NULL");
    *out = nullptr;
    if (ndim != 3 && ndim != 4) return fail(ARB_EINVAL, "ndim must be 3 or 4 (got %d)", ndim);
    if (ncols != ndim + 1 && ncols != ndim + 3)
        return fail(ARB_ESHAPE, "--- Input not shaped as expected - should be N x %d or N x %d ---",
                    ndim + 1, ndim + 3);
    if (!rows_dev || nrows <= 0 || ld < ncols)
        return fail(ARB_EINVAL, "arb_field_create: bad rows/nrows/ld");
    if (precision != ARB_PREC_FP64 && precision != ARB_PREC_MIXED)
        return fail(ARB_EINVAL, "unknown precision %d", precision);
    const int32_t f32_norm = mode & ARB_MODE_F32_NORM;
    mode &= ~ARB_MODE_F32_NORM;
    if (ncols == ndim + 1) mode = ARB_MODE_SCALAR;
    else if (mode == ARB_MODE_SCALAR || mode < 0 || mode > ARB_MODE_VECTOR_NAN)
        return fail(ARB_EINVAL, "vector field needs mode vector/norm/both (got %d)", mode);
    hipStream_t s = (hipStream_t)stream;

    arb_field* f = new (std::nothrow) arb_field();
    if (!f) return fail(ARB_ENOMEM, "host allocation failed");
    struct Guard { arb_field* f; ~Guard() { if (f) arb_field_destroy(f); } } guard{f};
    HIPCHK(hipGetDevice(&f->device));
    // set-up of the allocation-free query on this device: the first lookup of
    // the scratch array's address makes the runtime load the code object that
    // holds it (its storage is committed then), so it happens here, in a call
    // that allocates anyway, and never inside arb_query_noalloc
    {
        unsigned char* sp = nullptr;
        int rc = scratch_address(f->device, &sp);
        if (rc) return rc;
    }
    f->mode = mode;
    f->precision = precision;
    f->ninterp = interp_count(mode, &f->has_comps, &f->has_norm);
    HIPCHK(hipMalloc(&f->flag, sizeof(int32_t) * 4));
    HIPCHK(hipMemsetAsync(f->flag, 0, sizeof(int32_t) * 4, s));

    // 1. coordinates must be finite
    HIPCHK(arb_launch_finite_check(rows_dev, nrows, ncols, ld, ndim, f->flag, s));
    int32_t hflag = 0;
    HIPCHK(hipMemcpyAsync(&hflag, f->flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hflag) return fail(ARB_EGRID, "field has a non-finite coordinate");

    // 2. unique values per axis
    int64_t cap = 1024;
    while (cap < 2 * nrows) cap <<= 1;
    DevBuf tab, uniq, cnt;
    HIPCHK(tab.alloc(sizeof(unsigned long long) * cap));
    HIPCHK(uniq.alloc(sizeof(double) * cap));
    HIPCHK(cnt.alloc(sizeof(unsigned int)));
    std::vector<std::vector<double>> axes(ndim);
    for (int a = 0; a < ndim; ++a) {
        HIPCHK(hipMemsetAsync(tab.p, 0xFF, sizeof(unsigned long long) * cap, s));
        HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned int), s));
        HIPCHK(arb_launch_hash_insert(rows_dev, nrows, ld, a, tab.as<unsigned long long>(), cap, s));
        HIPCHK(arb_launch_hash_compact(tab.as<unsigned long long>(), cap, uniq.as<double>(),
                                       cnt.as<unsigned int>(), s));
        unsigned int nu = 0;
        HIPCHK(hipMemcpyAsync(&nu, cnt.p, sizeof nu, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        axes[a].resize(nu);
        HIPCHK(hipMemcpyAsync(axes[a].data(), uniq.p, sizeof(double) * nu, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::sort(axes[a].begin(), axes[a].end());
    }
    ArbGeom& g = f->g;
    g.ndim = ndim;
    int64_t prod = 1;
    for (int a = 0; a < ndim; ++a) {
        g.n[a] = (int64_t)axes[a].size();
        if (g.n[a] < 4)
            return fail(ARB_EGRID, "axis %d has %lld distinct coordinates; at least 4 are needed",
                        a, (long long)g.n[a]);
        if (prod > (int64_t)1 << 40) return fail(ARB_EGRID, "grid too large");
        prod *= g.n[a];
    }
    if (prod != nrows)
        return fail(ARB_EGRID, "rows do not form a complete grid: %lld rows, axis lengths give %lld",
                    (long long)nrows, (long long)prod);
    int64_t st = 1;
    g.nc = 1;
    for (int a = 0; a < ndim; ++a) {
        g.nstride[a] = st;
        st *= g.n[a];
        g.npos[a] = g.n[a] - 3;
        g.nc *= g.npos[a];
        g.h[a] = std::fabs(axes[a][0] - axes[a][1]);
        g.lo[a] = axes[a][1];
        g.hi[a] = axes[a][g.n[a] - 2];
        f->edge_lo[a] = axes[a][0];
        f->edge_hi[a] = axes[a][g.n[a] - 1];
    }
    g.nnodes = prod;
    // cell strides are 32-bit in the staging arithmetic of the query kernels
    if (g.nc >= ((int64_t)1 << 32))
        return fail(ARB_EGRID, "grid too large: %lld cells (limit 2^32)", (long long)g.nc);

    // 3. rank rows, scatter values into canonical order, check completeness
    DevBuf axes_dev[ARB_MAXD], seen;
    const double* axp[ARB_MAXD] = {nullptr, nullptr, nullptr, nullptr};
    for (int a = 0; a < ndim; ++a) {
        HIPCHK(axes_dev[a].alloc(sizeof(double) * g.n[a]));
        HIPCHK(hipMemcpyAsync(axes_dev[a].p, axes[a].data(), sizeof(double) * g.n[a],
                              hipMemcpyHostToDevice, s));
        axp[a] = axes_dev[a].as<double>();
    }
    HIPCHK(seen.alloc(sizeof(int32_t) * g.nnodes));
    HIPCHK(hipMemsetAsync(seen.p, 0, sizeof(int32_t) * g.nnodes, s));
    HIPCHK(hipMalloc(&f->vals, sizeof(double) * g.nnodes * f->ninterp));
    HIPCHK(arb_launch_rank_scatter(rows_dev, nrows, ld, axp, g, mode | f32_norm, ndim, f->vals,
                                   f->ninterp, seen.as<int32_t>(), s));
    HIPCHK(arb_launch_check_seen(seen.as<int32_t>(), g.nnodes, f->flag, s));
    HIPCHK(hipMemcpyAsync(&hflag, f->flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hflag) return fail(ARB_EGRID, "rows do not form a complete grid (duplicate points)");

    // 4. operator matrix and table storage
    const int K = 1 << (2 * ndim);
    std::vector<double> Bi, D, A;
    make_invB(ndim, Bi);
    make_D(ndim, D);
    matmul(K, Bi, D, A);
    HIPCHK(hipMalloc(&f->A, sizeof(double) * K * K));
    HIPCHK(hipMemcpyAsync(f->A, A.data(), sizeof(double) * K * K, hipMemcpyHostToDevice, s));
    const size_t esz = precision == ARB_PREC_FP64 ? sizeof(double) : sizeof(float);
    // the table itself is allocated by its first build (node-tile queries
    // never read it)
    f->table_bytes = (int64_t)((g.nc + 1) * f->ninterp * K * esz);
    HIPCHK(hipMalloc(&f->sent, esz * f->ninterp * K));
    HIPCHK(hipStreamSynchronize(s));
    guard.f = nullptr;
    *out = f;
    return ARB_OK;
}

int arb_field_info_get(const arb_field* f, arb_field_info* out) {
    if (!f || !out) return fail(ARB_EINVAL, "arb_field_info_get: NULL argument");
    std::memset(out, 0, sizeof *out);
    out->ndim = f->g.ndim;
    out->mode = f->mode;
    out->precision = f->precision;
    out->ninterp = f->ninterp;
    out->has_comps = f->has_comps;
    out->has_norm = f->has_norm;
    out->table_built = f->table_built;
    for (int a = 0; a < f->g.ndim; ++a) {
        out->n[a] = f->g.n[a];
        out->npos[a] = f->g.npos[a];
        out->h[a] = f->g.h[a];
        out->int_min[a] = f->g.lo[a];
        out->int_max[a] = f->g.hi[a];
    }
    out->nc = f->g.nc;
    out->nnodes = f->g.nnodes;
    out->table_bytes = f->table_bytes;
    out->packed_width = arb_packed_width(f->has_comps, f->has_norm, f->g.ndim);
    out->boundary = f->bc;
    for (int a = 0; a < f->g.ndim; ++a) out->block[a] = f->B.nbt ? f->B.b[a] : 0;
    out->tiles_bytes = f->tiles_bytes;
    out->on_the_fly = f->otf;
    return ARB_OK;
}

int arb_field_set_boundary(arb_field* f, int32_t bc, void* stream) {
    if (!f) return fail(ARB_EINVAL, "arb_field_set_boundary: NULL handle");
    if (bc != ARB_BC_NONE && bc != ARB_BC_NEUMANN && bc != ARB_BC_DIRICHLET)
        return fail(ARB_EINVAL, "unknown boundary rule %d", bc);
    if (bc == f->bc) return ARB_OK;
    if (f->bc != ARB_BC_NONE) return fail(ARB_EINVAL, "the field's boundary is already extended");
    if (f->table_built || f->sentinel_set || f->table_loaded || f->B.nbt != 0 || f->bw_n != 0 ||
        f->push_exec)
        return fail(ARB_EINVAL, "set the boundary before the first table build or query");
    hipStream_t s = (hipStream_t)stream;
    const ArbGeom& go = f->g;
    ArbGeom gn = go;
    int64_t st = 1;
    gn.nc = 1;
    for (int a = 0; a < go.ndim; ++a) {
        gn.n[a] = go.n[a] + 2;
        gn.npos[a] = go.n[a] - 1;
        gn.nstride[a] = st;
        st *= gn.n[a];
        gn.nc *= gn.npos[a];
        gn.lo[a] = f->edge_lo[a];
        gn.hi[a] = f->edge_hi[a];
    }
    gn.nnodes = st;
    if (gn.nc >= ((int64_t)1 << 32))
        return fail(ARB_EGRID, "grid too large: %lld cells (limit 2^32)", (long long)gn.nc);
    const int K = 1 << (2 * go.ndim);
    const size_t esz = f->precision == ARB_PREC_FP64 ? sizeof(double) : sizeof(float);
    const int64_t tbytes = (int64_t)((gn.nc + 1) * f->ninterp * K * esz);
    DevBuf nv;
    HIPCHK(nv.alloc(sizeof(double) * gn.nnodes * f->ninterp));
    HIPCHK(arb_launch_pad_boundary(f->vals, go, nv.as<double>(), gn, f->ninterp, bc, s));
    HIPCHK(hipStreamSynchronize(s));
    (void)hipFree(f->vals);
    if (f->table) (void)hipFree(f->table);    // not built yet: allocated again by the build
    f->vals = nv.as<double>();
    f->table = nullptr;
    nv.p = nullptr;
    f->table_bytes = tbytes;
    f->g = gn;
    f->bc = bc;
    return ARB_OK;
}

// The sentinel row, which rows outside the box are evaluated from: zeros in
// scalar and vector mode, NaN otherwise (ARBInterp.py:31,42-45,57,70,85,99).
// Written to `sent` (paths without the table) and, with the table, to its
// row nc.
static int sentinel_nan(const arb_field* f) {
    return (f->mode == ARB_MODE_SCALAR || f->mode == ARB_MODE_VECTOR) ? 0 : 1;
}

static int ensure_sentinel(arb_field* f, hipStream_t s) {
    if (f->sentinel_set) return ARB_OK;
    const int K = 1 << (2 * f->g.ndim);
    HIPCHK(arb_launch_sentinel(f->sent, f->precision != ARB_PREC_FP64, 0, f->ninterp, K,
                               sentinel_nan(f), s));
    f->sentinel_set = 1;
    return ARB_OK;
}

// The table's storage, allocated by the first build.  A table larger than
// the caller's limit (arb_set_table_limit), or one the device cannot hold,
// switches the field to forming coefficients on the fly (otf): queries then
// take the fused bucketed kernels, which give the table's bits, and the
// calls that need the table itself return ARB_ENOMEM.
static int ensure_table_alloc(arb_field* f) {
    if (f->table) return ARB_OK;
    if (!f->otf && f->table_limit > 0 && f->table_bytes > f->table_limit) f->otf = 1;
    if (!f->otf) {
        const hipError_t e = hipMalloc(&f->table, (size_t)f->table_bytes);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            f->table = nullptr;
            f->otf = 1;
        } else if (e != hipSuccess) {
            f->table = nullptr;
            return fail(ARB_EHIP, "hipMalloc of the coefficient table failed: %s",
                        hipGetErrorString(e));
        }
    }
    if (f->otf)
        return fail(ARB_ENOMEM, "the coefficient table (%lld bytes) does not fit%s; queries form "
                    "coefficients on the fly", (long long)f->table_bytes,
                    f->table_limit > 0 ? " the table size limit" : " in device memory");
    return ARB_OK;
}

int arb_build_coeffs(arb_field* f, void* stream) {
    if (!f) return fail(ARB_EINVAL, "arb_build_coeffs: NULL handle");
    if (f->table_built) return ARB_OK;
    hipStream_t s = (hipStream_t)stream;
    int rc = ensure_table_alloc(f);
    if (rc) return rc;
    HIPCHK(arb_launch_coef(f->g, f->precision, f->vals, f->ninterp, f->table, s));
    const int K = 1 << (2 * f->g.ndim);
    HIPCHK(arb_launch_sentinel(f->table, f->precision != ARB_PREC_FP64, f->g.nc, f->ninterp, K,
                               sentinel_nan(f), s));
    rc = ensure_sentinel(f, s);
    if (rc) return rc;
    f->table_built = 1;
    return ARB_OK;
}

// Whether queries of the field are evaluated from node tiles: mixed 3D,
// unless the table is asked for or has been loaded.  Not on Neumann-extended
// fields: the table's FP64-formed slopes make the normal slope on their faces
// exactly zero, node tiles only to 2^-24 of the tile differences.  Chooses
// the bucket geometry if needed.
static bool node_path(arb_field* f, int32_t flags) {
    if (f->precision != ARB_PREC_MIXED || f->g.ndim != 3 || f->table_loaded ||
        f->bc == ARB_BC_NEUMANN || (flags & ARB_Q_TABLE))
        return false;
    if (f->B.nbt == 0) choose_buckets(f, f->B, true);
    return arb_nodes_ok(f->g, f->B, f->ninterp, f->has_comps, f->has_norm);
}

// The node tiles of the current node values and bucket geometry.
static int ensure_tiles(arb_field* f, hipStream_t s) {
    if (f->tiles_ok) return ARB_OK;
    const int64_t tb = arb_tiles_bytes(f->B, f->ninterp, nullptr);
    if (tb != f->tiles_bytes) {
        if (f->tiles) (void)hipFree(f->tiles);
        f->tiles = nullptr;
        f->tiles_bytes = 0;
        HIPCHK(hipMalloc(&f->tiles, (size_t)tb));
        f->tiles_bytes = tb;
    }
    HIPCHK(arb_launch_build_tiles(f->g, f->B, f->ninterp, f->vals, f->tiles, s));
    f->tiles_ok = 1;
    return ARB_OK;
}

// ---------------------------------------------------------------------------
// Per-call workspace of the bucketed query path: the sort's key and
// row-index buffers (two each), its tile histograms and their scans, and the
// bucket bounds (`first` positions, then `offs`).  Carved from one base
// address, each piece 256-byte aligned; with base == nullptr only the size
// is computed.
// ---------------------------------------------------------------------------
struct SortWs {
    uint32_t* k[2] = {nullptr, nullptr};
    int32_t* v[2] = {nullptr, nullptr};
    int32_t* hist = nullptr;
    int32_t* hscan = nullptr;
    int32_t* first = nullptr;
    int32_t* offs = nullptr;
    void* qrec = nullptr;       // node-tile path: 16-byte record per row
};

static int64_t carve_sort_ws(unsigned char* base, int64_t n, int64_t nbt, bool qrec, SortWs* w) {
    int64_t off = 0;
    auto take = [&](int64_t bytes) -> void* {
        void* p = base ? (void*)(base + off) : nullptr;
        off += (bytes + 255) / 256 * 256;
        return p;
    };
    const int64_t hl = 256 * arb_rs_ntiles(n);
    for (int k = 0; k < 2; ++k) {
        w->k[k] = static_cast<uint32_t*>(take(sizeof(uint32_t) * n));
        w->v[k] = static_cast<int32_t*>(take(sizeof(int32_t) * n));
    }
    w->hist = static_cast<int32_t*>(take(sizeof(int32_t) * hl));
    w->hscan = static_cast<int32_t*>(take(sizeof(int32_t) * (hl + 256)));
    w->first = static_cast<int32_t*>(take(sizeof(int32_t) * arb_rs_first_elems(nbt)));
    w->offs = static_cast<int32_t*>(take(sizeof(int32_t) * (nbt + 1)));
    if (qrec) w->qrec = take(16 * n);
    return off;
}

// arb_query and arb_query_noalloc.  With noalloc the call allocates nothing:
// the table or node tiles must already be allocated (they are rebuilt in
// place when the node values changed), the bucketed path takes its sort
// workspace from `scratch` and writes the output arrays directly (no packed
// record workspace), and flags that need other workspaces are refused.
static int query_impl(arb_field* f, double* q_dev, int64_t n, int64_t ldq, void* comps_dev,
                      void* norms_dev, void* grads_dev, int64_t* cells_dev, int32_t flags,
                      void* stream, bool noalloc) {
    const char* fn = noalloc ? "arb_query_noalloc" : "arb_query";
    if (!f) return fail(ARB_EINVAL, "%s: NULL handle", fn);
    if (n < 0 || (n > 0 && (!q_dev || ldq < f->g.ndim)))
        return fail(ARB_EINVAL, "%s: bad positions (n=%lld ldq=%lld)", fn, (long long)n,
                    (long long)ldq);
    const bool packed = (flags & ARB_Q_PACKED) != 0;
    if (n > 0 && (f->has_comps || packed) && !comps_dev)
        return fail(ARB_EINVAL, "%s: comps is NULL", fn);
    if (n > 0 && !packed && f->has_norm && (!norms_dev || !grads_dev))
        return fail(ARB_EINVAL, "%s: norms/grads is NULL", fn);
    if (noalloc) {
        if (flags & (ARB_Q_TIMING | ARB_Q_GROUP_ATOMIC))
            return fail(ARB_EINVAL, "%s: ARB_Q_TIMING and ARB_Q_GROUP_ATOMIC need workspaces "
                        "of their own and are not available here", fn);
        if (n > ARB_NOALLOC_MAX_ROWS)
            return fail(ARB_EINVAL, "%s: %lld rows exceed the limit of %d per call", fn,
                        (long long)n, ARB_NOALLOC_MAX_ROWS);
        int dev = -1;
        HIPCHK(hipGetDevice(&dev));
        if (dev != f->device)
            return fail(ARB_EINVAL, "%s: the field lives on device %d, device %d is current", fn,
                        f->device, dev);
    }
    hipStream_t s = (hipStream_t)stream;
    // float32 field and positions: float32 reduction on the direct table kernel
    const bool red32 = (flags & ARB_Q_REDUCE_F32) != 0;
    if (red32) flags = (flags | ARB_Q_DIRECT | ARB_Q_TABLE) & ~(ARB_Q_SORTED | ARB_Q_SORTED_GLOBAL);
    ArbGeom gq = f->g;
    gq.red32 = red32 ? 1 : 0;
    bool sorted = (flags & (ARB_Q_SORTED | ARB_Q_SORTED_GLOBAL)) != 0;
    // automatic choice: the bucketed path from about one query per four cells
    // with node tiles (it stages ~128 bytes per cell), from one per two cells
    // with the table (1 KB per cell in mixed 'both'); measured crossovers on
    // the 128^3 field: 0.5 M and 1-2 M queries
    if (!(flags & ARB_Q_DIRECT) && !sorted)
        sorted = n >= 65536 && (node_path(f, flags) ? 4 * n : 2 * n) >= f->g.nc;
    if (n > INT32_MAX - 1) sorted = false;
    if (sorted && f->B.nbt == 0) choose_buckets(f, f->B, true);
    // 3D, bucketed, on request: coefficients formed in the query kernel from
    // node values, no table (ignored once a table has been loaded, or when
    // the block does not fit the kernel's LDS)
    bool fused = false;
    if ((flags & ARB_Q_FUSED) && sorted && f->g.ndim == 3 && !f->table_loaded &&
        !(flags & (ARB_Q_SORTED_GLOBAL | ARB_Q_GROUP_ATOMIC))) {
        if (f->BF.nbt == 0) choose_buckets(f, f->BF, false);
        fused = arb_fused_lds(f->g, f->BF, f->ninterp, f->eval_threads) > 0;
    }
    // node tiles (mixed 3D): the direct path, and the bucketed path unless
    // another grouping or the fused kernel is asked for; no table
    bool nodes = !fused && node_path(f, flags) &&
                 !(flags & (ARB_Q_SORTED_GLOBAL | ARB_Q_GROUP_ATOMIC));
    // a table path whose table does not fit: coefficients on the fly, in the
    // fused bucketed kernels (3D and 4D), whatever path was asked for
    bool fused4 = false;
    if (!fused && !nodes && !f->table && !noalloc) {
        const int rc0 = ensure_table_alloc(f);
        if (rc0 && !f->otf) return rc0;
    }
    if (!fused && !nodes && f->otf) {
        if (red32)
            return fail(ARB_ENOMEM, "%s: the float32 reduction reads the coefficient table, which "
                        "does not fit", fn);
        if (n > INT32_MAX - 1)
            return fail(ARB_EINVAL, "%s: too many rows for the on-the-fly path", fn);
        flags &= ~(ARB_Q_DIRECT | ARB_Q_SORTED_GLOBAL | ARB_Q_GROUP_ATOMIC);
        sorted = true;
        if (f->BF.nbt == 0) choose_buckets(f, f->BF, false);
        if (f->g.ndim == 3)
            fused = arb_fused_lds(f->g, f->BF, f->ninterp, f->eval_threads) > 0;
        else
            fused4 = arb_fused4_lds(f->g, f->BF) > 0;
        if (!fused && !fused4)
            return fail(ARB_ENOMEM, "%s: the coefficient table does not fit and the bucket of "
                        "this field is too large to form its coefficients on the fly", fn);
    }
    const bool nodeform = nodes && sorted;
    const ArbBuckets& BB = (fused || fused4) ? f->BF : f->B;
    if (noalloc) {
        if (nodes && !(f->tiles && f->tiles_bytes == arb_tiles_bytes(f->B, f->ninterp, nullptr)))
            return fail(ARB_EINVAL, "%s: the field's node tiles are not allocated; prepare the "
                        "field with one arb_query call with the same flags (n may be 0)", fn);
        if (!fused && !fused4 && !nodes && !f->table)
            return fail(ARB_EINVAL, "%s: the field's coefficient table is not allocated; "
                        "prepare the field with arb_build_coeffs", fn);
    }
    int rc = (fused || fused4 || nodes) ? ensure_sentinel(f, s) : arb_build_coeffs(f, stream);
    if (rc) return rc;
    if (nodes) {
        rc = ensure_tiles(f, s);
        if (rc) return rc;
    }

    if (n == 0) return ARB_OK;
    if (flags & ARB_Q_CHECK) HIPCHK(arb_launch_fill_i32(f->flag + 1, 1, 0, s));
    ArbOut o{};
    o.comps = comps_dev;
    o.norms = norms_dev;
    o.grads = grads_dev;
    o.cells = cells_dev;
    o.qmut = (flags & ARB_Q_NAN_ROWS) ? q_dev : nullptr;
    o.ldq = ldq;
    o.flag = f->flag + 1;
    o.packed = packed ? arb_packed_width(f->has_comps, f->has_norm, f->g.ndim) : 0;
    const int out64 = (flags & ARB_Q_OUT_F64) ? 1 : 0;
    const bool timing = (flags & ARB_Q_TIMING) != 0;
    if (timing && !f->ev_ready) {
        for (auto& e : f->ev) HIPCHK(hipEventCreate(&e));
        f->ev_ready = 1;
    }
    for (float& m : f->last_ms) m = 0.f;
    if (!sorted) {
        if (timing) HIPCHK(hipEventRecord(f->ev[0], s));
        if (red32)
            HIPCHK(arb_launch_query_direct_red32(gq, f->precision, out64, f->table, f->ninterp,
                                                 f->has_comps, f->has_norm, q_dev, n, ldq, o,
                                                 s));
        else if (nodes)
            HIPCHK(arb_launch_query_tiles(f->g, f->B, f->ninterp, out64, f->tiles, f->sent,
                                          f->has_comps, f->has_norm, q_dev, n, ldq, o, s));
        else
            HIPCHK(arb_launch_query_direct(f->g, f->precision, out64, f->table, f->ninterp,
                                           f->has_comps, f->has_norm, q_dev, n, ldq, o, s));
        if (timing) {
            HIPCHK(hipEventRecord(f->ev[4], s));
            HIPCHK(hipEventSynchronize(f->ev[4]));
            HIPCHK(hipEventElapsedTime(&f->last_ms[3], f->ev[0], f->ev[4]));
        }
    } else {
        SortWs w;
        if (noalloc) {
            unsigned char* base = nullptr;
            rc = scratch_address(f->device, &base);
            if (rc) return rc;
            const int64_t need = carve_sort_ws(nullptr, n, BB.nbt, nodeform, &w);
            if (need > (int64_t)ARB_NOALLOC_SCRATCH_BYTES)
                return fail(ARB_EINVAL, "%s: %lld rows over %lld buckets need %lld bytes of "
                            "scratch; the limit is %lld", fn, (long long)n, (long long)BB.nbt,
                            (long long)need, (long long)ARB_NOALLOC_SCRATCH_BYTES);
            carve_sort_ws(base, n, BB.nbt, nodeform, &w);
        } else {
            rc = ensure_workspace(f, n, BB.nbt);
            if (rc) return rc;
            for (int k = 0; k < 2; ++k) {
                w.k[k] = f->rs_k[k];
                w.v[k] = f->rs_v[k];
            }
            w.hist = f->rs_hist;
            w.hscan = f->rs_hscan;
            w.first = f->ws_count;
            w.offs = f->ws_offs;
        }
        // separate output arrays of a 3D field: the bucketed kernels write
        // each row as one packed record into the workspace, and k_unpack
        // splits the records into the arrays in row order (random
        // partial-line writes to several arrays per row cost more than the
        // extra ordered pass).  Measured on 10^7 rows of the 128^3 field:
        // 'both' 0.93 -> 0.68 ms; FP64 'norm' / scalar rows (norms and
        // grads) 1.24 ms as arrays against 0.96 ms as records.  FP32 norms
        // and grads, comps alone, and 4D (whose evaluation hides the writes:
        // 2e6 rows, eval time unchanged, +33 us of split) gain nothing.
        const bool wide = f->precision == ARB_PREC_FP64 || out64;
        const bool split =
            !noalloc && !packed && f->g.ndim == 3 && f->has_norm && (f->has_comps || wide);
        if (split) {
            const int32_t pw = arb_packed_width(f->has_comps, f->has_norm, f->g.ndim);
            const int64_t esz = wide ? (int64_t)sizeof(double) : (int64_t)sizeof(float);
            rc = ensure_records(f, n * pw * esz);
            if (rc) return rc;
            o.comps = f->ws_rec;
            o.norms = nullptr;
            o.grads = nullptr;
            o.packed = pw;
        }
        // the node-tile rows' 16-byte records: inside the output records when
        // those are 16 or 32 bytes (k_keys writes each slot whole; each row's
        // output then overwrites its own record, so the evaluation reads and
        // writes the same line of a row instead of one line at random for
        // each: 10^7 rows of the 128^3 field, FP32 'both', eval 0.50 ->
        // 0.40 ms, k_keys 0.10 -> 0.14 ms, step 0.89 -> 0.82 ms), else in a
        // workspace of their own (64-byte slots written whole would double
        // k_keys' writes for less than they save)
        int32_t qstride = 1;
        if (nodeform) {
            const int64_t slot = (int64_t)o.packed * (wide ? 8 : 4);
            if ((slot == 16 || slot == 32) && (uintptr_t)o.comps % 16 == 0) {
                w.qrec = o.comps;
                qstride = (int32_t)(slot / 16);
            } else if (!noalloc) {
                rc = ensure_qrec(f, n);
                if (rc) return rc;
                w.qrec = f->ws_qrec;
            }
        }
        const int32_t* perm = f->ws_perm;
        // the evaluation's hot flag (a bucket holds more rows than one
        // workgroup takes): the last int of w.first, set by the grouping
        // sort, and set to 1 (look) by the other groupings
        int32_t* hot = w.first + arb_rs_first_elems(BB.nbt) - 1;
        if (timing) HIPCHK(hipEventRecord(f->ev[0], s));
        if (flags & ARB_Q_GROUP_ATOMIC) {
            HIPCHK(arb_launch_fill_i32(f->ws_count, f->B.nbt, 0, s));
            HIPCHK(arb_launch_bucket(f->g, f->B, f->precision, out64, f->table, f->ninterp,
                                     f->has_comps, f->has_norm, q_dev, n, ldq, o, f->ws_bkt,
                                     f->ws_rank, f->ws_count, s));
            if (timing) HIPCHK(hipEventRecord(f->ev[1], s));
            HIPCHK(arb_launch_scan(f->ws_count, f->ws_offs, f->B.nbt, s));
            if (timing) HIPCHK(hipEventRecord(f->ev[2], s));
            HIPCHK(arb_launch_scatter(f->ws_bkt, f->ws_rank, f->ws_offs, f->ws_perm, n, s));
            HIPCHK(arb_launch_fill_i32(hot, 1, 1, s));
        } else {
            // key = bucket (then the cell inside it with ARB_Q_SORT_CELL)
            int cellbits = 0;
            if (flags & ARB_Q_SORT_CELL)
                while ((1 << cellbits) < BB.ncell) ++cellbits;
            int bits = 1;
            while (((int64_t)1 << bits) <= BB.nbt) ++bits;        // buckets are 0..nbt
            bits += cellbits;
            if (bits > 32) {          // no room for the cell: order by bucket only
                bits -= cellbits;
                cellbits = 0;
            }
            // rows outside the box: the table's row nc on the table paths
            // (a loaded table brings its own), the sentinel row otherwise
            const size_t esz = f->precision == ARB_PREC_FP64 ? sizeof(double) : sizeof(float);
            const void* sentrow =
                (fused || fused4 || nodes) ? f->sent
                                 : (const void*)((const char*)f->table +
                                                 esz * (size_t)f->g.nc * f->ninterp *
                                                     ((size_t)1 << (2 * f->g.ndim)));
            // rows grouped by bucket in any order inside it (the evaluation
            // of a row does not depend on it): the grouping sort, whose
            // second digit needs no stable ranking; with the cell in the key
            // (rows ordered by cell inside a bucket) or keys past its range,
            // the LSD radix sort
            const bool grp = cellbits == 0 && arb_group_sort_ok(n, bits);
            HIPCHK(arb_launch_keys(f->g, BB, f->precision, out64, sentrow, f->has_comps,
                                   f->has_norm, q_dev, n, ldq, o, w.k[0], cellbits, w.hist,
                                   grp ? bits - 8 : 0, nodeform ? w.qrec : nullptr, qstride, s));
            if (timing) HIPCHK(hipEventRecord(f->ev[1], s));
            if (grp) {
                // each bucket's first position goes to w.offs (w.first: the
                // per-bucket counters)
                HIPCHK(arb_launch_group_sort(w.k[0], w.k[1], w.v[0], n, bits, w.hist, w.hscan,
                                             w.offs, BB.nbt, w.first, s));
                perm = w.v[0];
            } else {
                int in_first = 1;
                // the sort also leaves each bucket's first sorted position in
                // w.offs (w.first as scratch)
                HIPCHK(arb_launch_radix_sort(w.k[0], w.v[0], w.k[1], w.v[1], n, bits, w.hist,
                                             w.hscan, 1, 1, w.offs, BB.nbt, cellbits, w.first,
                                             &in_first, 0, s));
                HIPCHK(arb_launch_fill_i32(hot, 1, 1, s));
                perm = w.v[in_first ? 0 : 1];
            }
            if (timing) HIPCHK(hipEventRecord(f->ev[2], s));
        }
        const int32_t* offs = w.offs;
        if (timing) HIPCHK(hipEventRecord(f->ev[3], s));
        if (flags & ARB_Q_SORTED_GLOBAL)
            HIPCHK(arb_launch_eval_perm(f->g, f->B, f->precision, out64, f->table, f->ninterp,
                                        f->has_comps, f->has_norm, q_dev, n, ldq, o, offs,
                                        perm, s));
        else if (nodeform)
            HIPCHK(arb_launch_eval_nodes(f->g, BB, f->ninterp, out64, f->tiles, w.qrec, qstride,
                                         n, o, offs, perm, hot, f->eval_threads, s));
        else if (fused)
            HIPCHK(arb_launch_eval_fused(f->g, BB, f->precision, out64, f->vals, f->ninterp,
                                         f->has_comps, f->has_norm, q_dev, n, ldq, o, offs,
                                         perm, hot, f->eval_threads, s));
        else if (fused4)
            HIPCHK(arb_launch_eval_fused4(f->g, BB, f->precision, out64, f->vals, f->ninterp,
                                          f->has_comps, f->has_norm, q_dev, n, ldq, o, offs,
                                          perm, hot, f->eval_threads, s));
        else
            HIPCHK(arb_launch_eval_bucket(f->g, f->B, f->precision, out64, f->table, f->ninterp,
                                          f->has_comps, f->has_norm, q_dev, n, ldq, o,
                                          offs, perm, hot, f->eval_threads, s));
        if (timing) HIPCHK(hipEventRecord(f->ev[4], s));
        if (split)
            HIPCHK(arb_launch_unpack(f->g.ndim, f->precision == ARB_PREC_FP64 || out64,
                                     f->has_comps, f->has_norm, f->ws_rec, comps_dev, norms_dev,
                                     grads_dev, n, s));
        if (timing) {
            HIPCHK(hipEventRecord(f->ev[5], s));
            HIPCHK(hipEventSynchronize(f->ev[5]));
            for (int k = 0; k < 5; ++k)
                HIPCHK(hipEventElapsedTime(&f->last_ms[k], f->ev[k], f->ev[k + 1]));
        }
    }
    if (flags & ARB_Q_CHECK) {
        int32_t h = 0;
        HIPCHK(hipMemcpyAsync(&h, f->flag + 1, sizeof h, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (h) return fail(ARB_EINDEX, "a query reduced to a cell index past the sentinel");
    }
    return ARB_OK;
}

int arb_query(arb_field* f, double* q_dev, int64_t n, int64_t ldq, void* comps_dev,
              void* norms_dev, void* grads_dev, int64_t* cells_dev, int32_t flags,
              void* stream) {
    return query_impl(f, q_dev, n, ldq, comps_dev, norms_dev, grads_dev, cells_dev, flags, stream,
                      false);
}

int arb_query_noalloc(arb_field* f, double* q_dev, int64_t n, int64_t ldq, void* comps_dev,
                      void* norms_dev, void* grads_dev, int64_t* cells_dev, int32_t flags,
                      void* stream) {
    return query_impl(f, q_dev, n, ldq, comps_dev, norms_dev, grads_dev, cells_dev, flags, stream,
                      true);
}

// Layout of the single-point record, in doubles: point, components, norm,
// gradient, cell (int64) and status word (int32).
enum { PIN_Q = 0, PIN_COMPS = 4, PIN_NORM = 7, PIN_GRAD = 8, PIN_CELL = 12, PIN_FLAG = 13,
       PIN_DOUBLES = 16 };

int arb_query_point(arb_field* f, const double* pt, double* comps, double* norm, double* grad,
                    int64_t* cell, void* stream) {
    if (!f || !pt) return fail(ARB_EINVAL, "arb_query_point: NULL argument");
    if (f->has_comps && !comps) return fail(ARB_EINVAL, "arb_query_point: comps is NULL");
    if (f->has_norm && (!norm || !grad))
        return fail(ARB_EINVAL, "arb_query_point: norm/grad is NULL");
    hipStream_t s = (hipStream_t)stream;
    const bool nodes = node_path(f, 0);
    int rc = nodes ? ensure_sentinel(f, s) : arb_build_coeffs(f, stream);
    // no table (it does not fit): the point goes through the bucketed
    // on-the-fly path as a one-row batch, with the same results
    const bool otf_point = rc == ARB_ENOMEM && f->otf;
    if (rc && !otf_point) return rc;
    if (nodes) {
        rc = ensure_tiles(f, s);
        if (rc) return rc;
    }
    if (!f->pin) {
        void* h = nullptr;
        HIPCHK(hipHostMalloc(&h, sizeof(double) * PIN_DOUBLES, hipHostMallocDefault));
        f->pin = static_cast<double*>(h);
        void* dptr = nullptr;
        HIPCHK(hipHostGetDevicePointer(&dptr, h, 0));
        f->pin_dev = static_cast<double*>(dptr);
    }
    const int d = f->g.ndim;
    volatile double* h = f->pin;
    for (int a = 0; a < d; ++a) h[PIN_Q + a] = pt[a];
    volatile int32_t* hflag = reinterpret_cast<volatile int32_t*>(f->pin + PIN_FLAG);
    *hflag = 0;
    ArbOut o{};
    o.comps = f->has_comps ? f->pin_dev + PIN_COMPS : nullptr;
    o.norms = f->has_norm ? f->pin_dev + PIN_NORM : nullptr;
    o.grads = f->has_norm ? f->pin_dev + PIN_GRAD : nullptr;
    o.cells = reinterpret_cast<int64_t*>(f->pin_dev + PIN_CELL);
    o.ldq = d;
    o.flag = reinterpret_cast<int32_t*>(f->pin_dev + PIN_FLAG);
    bool past = false;           // the point reduced past the sentinel cell
    if (otf_point) {
        // with ARB_Q_CHECK the call synchronises and reports a row past the
        // sentinel after the cell has been written
        rc = query_impl(f, f->pin_dev + PIN_Q, 1, d, o.comps, o.norms, o.grads, o.cells,
                        ARB_Q_SORTED | ARB_Q_OUT_F64 | ARB_Q_CHECK, stream, false);
        past = rc == ARB_EINDEX;
        if (rc && !past) return rc;
    } else if (nodes)
        HIPCHK(arb_launch_query_tiles(f->g, f->B, f->ninterp, 1, f->tiles, f->sent, f->has_comps,
                                      f->has_norm, f->pin_dev + PIN_Q, 1, d, o, s));
    else
        HIPCHK(arb_launch_query_direct(f->g, f->precision, 1, f->table, f->ninterp, f->has_comps,
                                       f->has_norm, f->pin_dev + PIN_Q, 1, d, o, s));
    HIPCHK(hipStreamSynchronize(s));
    if (f->has_comps)
        for (int c = 0; c < 3; ++c) comps[c] = h[PIN_COMPS + c];
    if (f->has_norm) {
        *norm = h[PIN_NORM];
        for (int a = 0; a < d; ++a) grad[a] = h[PIN_GRAD + a];
    }
    const int64_t flat = *reinterpret_cast<volatile int64_t*>(f->pin + PIN_CELL);
    if (cell) *cell = flat;
    if (*hflag || past) return fail(ARB_EINDEX, "the point reduced to a cell index past the sentinel");
    return ARB_OK;
}

int arb_field_values(const arb_field* f, double* dst_dev, void* stream) {
    if (!f || !dst_dev) return fail(ARB_EINVAL, "arb_field_values: NULL argument");
    if (f->bc) return fail(ARB_EINVAL, "node values are not available on a boundary-extended field");
    HIPCHK(hipMemcpyAsync(dst_dev, f->vals, sizeof(double) * f->g.nnodes * f->ninterp,
                          hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return ARB_OK;
}

int arb_field_set_values(arb_field* f, const double* src_dev, void* stream) {
    if (!f || !src_dev) return fail(ARB_EINVAL, "arb_field_set_values: NULL argument");
    if (f->bc) return fail(ARB_EINVAL, "node values are not settable on a boundary-extended field");
    HIPCHK(hipMemcpyAsync(f->vals, src_dev, sizeof(double) * f->g.nnodes * f->ninterp,
                          hipMemcpyDeviceToDevice, (hipStream_t)stream));
    f->table_built = 0;
    f->tiles_ok = 0;
    return ARB_OK;
}

int arb_query_backward(arb_field* f, const double* q_dev, int64_t n, int64_t ldq,
                       const double* gv_dev, const double* gg_dev, double* dq_dev,
                       double* dval_dev, void* stream) {
    if (!f) return fail(ARB_EINVAL, "arb_query_backward: NULL handle");
    if (f->bc && dval_dev)
        return fail(ARB_EINVAL, "node gradients are not available on a boundary-extended field");
    if (!f->has_norm) return fail(ARB_EINVAL, "backward needs a scalar interpolant (scalar, norm or both)");
    if (n < 0 || (n > 0 && (!q_dev || !gv_dev || !gg_dev || ldq < f->g.ndim)))
        return fail(ARB_EINVAL, "arb_query_backward: bad arguments");
    if (n > INT32_MAX - 1 || f->g.nc > INT32_MAX - 1)
        return fail(ARB_EINVAL, "arb_query_backward: too many queries or cells for int32 indices");
    hipStream_t s = (hipStream_t)stream;
    // mixed 3D fields take U's derivatives from the node tiles (FP64 model
    // plus FP32 residuals), as their forward queries do; others from the table
    const bool nodes = node_path(f, 0);
    int rc = nodes ? ensure_sentinel(f, s) : arb_build_coeffs(f, stream);
    if (rc) return rc;
    if (nodes) {
        rc = ensure_tiles(f, s);
        if (rc) return rc;
    }
    const int64_t nc = f->g.nc;
    if (dval_dev) {
        if (n > f->bw_n) {
            if (f->bw_key) (void)hipFree(f->bw_key);
            if (f->bw_perm) (void)hipFree(f->bw_perm);
            if (f->bw_rec) (void)hipFree(f->bw_rec);
            if (f->bw_k1) (void)hipFree(f->bw_k1);
            if (f->bw_v1) (void)hipFree(f->bw_v1);
            if (f->bw_hist) (void)hipFree(f->bw_hist);
            if (f->bw_hscan) (void)hipFree(f->bw_hscan);
            if (f->bw_part) (void)hipFree(f->bw_part);
            f->bw_key = f->bw_perm = f->bw_v1 = f->bw_hist = f->bw_hscan = nullptr;
            f->bw_k1 = nullptr;
            f->bw_part = nullptr;
            f->bw_rec = nullptr;
            f->bw_n = 0;
            const int64_t m = n ? n : 1;
            const int64_t hl = 256 * arb_rs_ntiles(m);
            HIPCHK(hipMalloc(&f->bw_key, sizeof(int32_t) * m));
            HIPCHK(hipMalloc(&f->bw_perm, sizeof(int32_t) * m));
            HIPCHK(hipMalloc(&f->bw_k1, sizeof(uint32_t) * m));
            HIPCHK(hipMalloc(&f->bw_v1, sizeof(int32_t) * m));
            HIPCHK(hipMalloc(&f->bw_hist, sizeof(int32_t) * hl));
            HIPCHK(hipMalloc(&f->bw_hscan, sizeof(int32_t) * (hl + 256)));
            HIPCHK(hipMalloc(&f->bw_part, sizeof(double) * arb_bwd_part_elems(f->g, m)));
            if (f->g.ndim == 3) HIPCHK(hipMalloc(&f->bw_rec, sizeof(double) * 8 * m));
            f->bw_n = n;
        }
        if (!f->bw_bbar) {
            // the radix sort's first positions, then the cell bounds
            HIPCHK(hipMalloc(&f->bw_count, sizeof(int32_t) * arb_rs_first_elems(nc)));
            HIPCHK(hipMalloc(&f->bw_offs, sizeof(int32_t) * (nc + 1)));
            HIPCHK(hipMalloc(&f->bw_bbar, sizeof(double) * arb_bwd_scratch(f->g)));
        }
    }
    // The node adjoint groups the queries per cell in query-index order (so
    // every per-cell sum runs in a fixed order): a stable radix sort of the
    // cell keys, O(n) whatever the distribution (a per-cell sort of the
    // cell's queries was quadratic in their number).  From the node tiles,
    // d/dq runs in that grouped order (neighbouring threads then read the
    // same tiles) and writes each query's record at its grouped position.
    // From the 3D table (a cell's 64 coefficients one contiguous 512-byte
    // read), in query order, with the keys and records by row.  From the
    // 4D table (2 KB per cell), in the grouped order again; 4D has no
    // records.
    const bool grouped = dval_dev && nodes;
    const bool grouped4 = dval_dev && !nodes && f->g.ndim == 4;
    if ((grouped || grouped4) && n > 0)
        HIPCHK(arb_launch_bwd_keys(f->g, q_dev, n, ldq, f->bw_key, s));
    if (n > 0 && !nodes && !grouped4)
        HIPCHK(arb_launch_bwd_points(f->g, f->precision, f->table, f->ninterp, f->ninterp - 1,
                                     q_dev, n, ldq, gv_dev, gg_dev, dq_dev, nullptr,
                                     dval_dev ? f->bw_key : nullptr,
                                     dval_dev ? f->bw_rec : nullptr, s));
    const int32_t* perm = nullptr;
    if (dval_dev) {
        if (n == 0) {
            HIPCHK(arb_launch_fill_i32(f->bw_offs, nc + 1, 0, s));
        } else {
            int bits = 1;
            while (((int64_t)1 << bits) <= nc) ++bits;      // keys 0..nc
            int in_first = 1;
            HIPCHK(arb_launch_radix_sort(reinterpret_cast<uint32_t*>(f->bw_key), f->bw_perm,
                                         f->bw_k1, f->bw_v1, n, bits, f->bw_hist, f->bw_hscan, 0,
                                         1, f->bw_offs, nc, 0, f->bw_count, &in_first, 1, s));
            perm = in_first ? f->bw_perm : f->bw_v1;
        }
    }
    if (n > 0 && nodes)
        HIPCHK(arb_launch_bwd_tiles(f->g, f->B, f->ninterp, f->tiles, q_dev, n, ldq, gv_dev,
                                    gg_dev, dq_dev, perm, dval_dev ? f->bw_rec : nullptr, s));
    if (n > 0 && grouped4)
        HIPCHK(arb_launch_bwd_points(f->g, f->precision, f->table, f->ninterp, f->ninterp - 1,
                                     q_dev, n, ldq, gv_dev, gg_dev, dq_dev, perm, nullptr,
                                     nullptr, s));
    if (dval_dev) {
        HIPCHK(arb_launch_bwd_cells(f->g, f->A, q_dev, ldq, gv_dev, gg_dev, f->bw_rec,
                                    grouped ? 1 : 0, f->bw_offs, perm, n, f->bw_part, f->bw_bbar,
                                    s));
        HIPCHK(arb_launch_bwd_nodes(f->g, f->bw_bbar, dval_dev, s));
    }
    return ARB_OK;
}

int arb_push_rk4(arb_field* f, double* x_dev, double* v_dev, int64_t n, double dt,
                 int64_t nsteps, int32_t steps_per_launch, int32_t use_graph, void* stream) {
    if (!f) return fail(ARB_EINVAL, "arb_push_rk4: NULL handle");
    if (f->g.ndim != 3 || !f->has_norm)
        return fail(ARB_EINVAL, "the particle push needs a 3D field with a scalar interpolant");
    if (n < 0 || nsteps < 0 || (n > 0 && (!x_dev || !v_dev)) || steps_per_launch < 1)
        return fail(ARB_EINVAL, "arb_push_rk4: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    int rc = arb_build_coeffs(f, stream);
    if (rc) return rc;
    if (n == 0 || nsteps == 0) return ARB_OK;
    const int64_t nlaunch = (nsteps + steps_per_launch - 1) / steps_per_launch;
    auto issue = [&](hipStream_t st) -> hipError_t {
        int64_t left = nsteps;
        for (int64_t k = 0; k < nlaunch; ++k) {
            const int32_t m = (int32_t)(left < steps_per_launch ? left : steps_per_launch);
            hipError_t e = arb_launch_rk4(f->g, f->precision, f->table, f->ninterp,
                                          f->ninterp - 1, x_dev, v_dev, n, dt, m, st);
            if (e != hipSuccess) return e;
            left -= m;
        }
        return hipSuccess;
    };
    if (!use_graph) {
        HIPCHK(issue(s));
        return ARB_OK;
    }
    const int64_t key[5] = {(int64_t)(uintptr_t)x_dev, (int64_t)(uintptr_t)v_dev, n, nsteps,
                            steps_per_launch};
    bool same = f->push_exec && f->push_key_d == dt;
    for (int k = 0; k < 5 && same; ++k) same = f->push_key[k] == key[k];
    if (!same) {
        if (f->push_exec) {
            HIPCHK(hipGraphExecDestroy(f->push_exec));
            f->push_exec = nullptr;
        }
        if (!f->cap_stream) HIPCHK(hipStreamCreateWithFlags(&f->cap_stream, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(f->cap_stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = issue(f->cap_stream);
        hipError_t e2 = hipStreamEndCapture(f->cap_stream, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(ARB_EHIP, "graph capture failed: %s",
                        hipGetErrorString(e != hipSuccess ? e : e2));
        }
        e = hipGraphInstantiate(&f->push_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            f->push_exec = nullptr;
            return fail(ARB_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        for (int k = 0; k < 5; ++k) f->push_key[k] = key[k];
        f->push_key_d = dt;
    }
    HIPCHK(hipGraphLaunch(f->push_exec, s));
    return ARB_OK;
}

// ---------------------------------------------------------------------------
// batches of fields
// ---------------------------------------------------------------------------
struct arb_batch {
    std::vector<arb_field*> fields;
    std::vector<const void*> tables;   // each field's table as the descriptors hold it
    int32_t dead = 0;                  // a field of the batch was destroyed
    int32_t ndim = 3, precision = 0, ninterp = 0, has_comps = 0, has_norm = 0;
    int32_t chunk = 256, threads = 256, blocks_per_cu = 4, ncu = 256;
    std::vector<int64_t> last_counts;
    std::vector<unsigned char> desc_host;
    void* desc_dev = nullptr;
    int64_t nitems = 0;
    unsigned int* counter = nullptr;
    int32_t* flag = nullptr;
};

int arb_batch_create(arb_field* const* fields, int32_t nfields, void* stream, arb_batch** out) {
    if (!out || !fields || nfields < 1) return fail(ARB_EINVAL, "arb_batch_create: bad arguments");
    *out = nullptr;
    arb_batch* b = new (std::nothrow) arb_batch();
    if (!b) return fail(ARB_ENOMEM, "host allocation failed");
    struct Guard { arb_batch* b; ~Guard() { if (b) arb_batch_destroy(b); } } guard{b};
    for (int k = 0; k < nfields; ++k) {
        arb_field* f = fields[k];
        if (!f) return fail(ARB_EINVAL, "field %d is NULL", k);
        if (k == 0) {
            b->ndim = f->g.ndim;
            b->precision = f->precision;
            b->ninterp = f->ninterp;
            b->has_comps = f->has_comps;
            b->has_norm = f->has_norm;
        } else if (f->g.ndim != b->ndim || f->precision != b->precision ||
                   f->ninterp != b->ninterp || f->has_comps != b->has_comps ||
                   f->has_norm != b->has_norm) {
            return fail(ARB_EINVAL, "field %d differs in ndim/mode/precision from field 0", k);
        }
        int rc = arb_build_coeffs(f, stream);
        if (rc) return rc;
        b->fields.push_back(f);
        b->tables.push_back(f->table);
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t dsz = arb_batch_field_desc_size();
    b->desc_host.assign(dsz * nfields, 0);
    HIPCHK(hipMalloc(&b->desc_dev, dsz * nfields));
    HIPCHK(hipMalloc(&b->counter, sizeof(unsigned int)));
    HIPCHK(hipMalloc(&b->flag, sizeof(int32_t)));
    HIPCHK(hipMemsetAsync(b->flag, 0, sizeof(int32_t), (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    for (arb_field* f : b->fields)
        if (std::find(f->batches.begin(), f->batches.end(), b) == f->batches.end())
            f->batches.push_back(b);
    guard.b = nullptr;
    *out = b;
    return ARB_OK;
}

int arb_batch_set_tuning(arb_batch* b, int32_t chunk, int32_t threads, int32_t blocks_per_cu) {
    if (!b) return fail(ARB_EINVAL, "arb_batch_set_tuning: NULL handle");
    if (chunk && (chunk < 64 || chunk % 64 || chunk > 65536))
        return fail(ARB_EINVAL, "chunk must be a multiple of 64 in [64, 65536]");
    if (threads && (threads < 64 || threads > 1024 || threads % 64))
        return fail(ARB_EINVAL, "threads must be a multiple of 64 in [64, 1024]");
    if (blocks_per_cu && (blocks_per_cu < 1 || blocks_per_cu > 64))
        return fail(ARB_EINVAL, "blocks_per_cu must be 1..64");
    if (chunk) b->chunk = chunk;
    if (threads) b->threads = threads;
    if (blocks_per_cu) b->blocks_per_cu = blocks_per_cu;
    b->last_counts.clear();          // item layout depends on chunk
    return ARB_OK;
}

int arb_batch_query(arb_batch* b, double* q_dev, const int64_t* counts, int64_t ldq,
                    void* comps_dev, void* norms_dev, void* grads_dev, int32_t flags,
                    void* stream) {
    if (!b || !counts) return fail(ARB_EINVAL, "arb_batch_query: NULL argument");
    if (b->dead) return fail(ARB_EINVAL, "arb_batch_query: a field of the batch was destroyed");
    const int nf = (int)b->fields.size();
    hipStream_t s = (hipStream_t)stream;
    // tables made stale by arb_field_set_values are rebuilt first; a field
    // whose table moved gets its descriptor written again
    bool moved = false;
    for (int k = 0; k < nf; ++k) {
        int rc = arb_build_coeffs(b->fields[k], stream);
        if (rc) return rc;
        if (b->fields[k]->table != b->tables[k]) {
            b->tables[k] = b->fields[k]->table;
            moved = true;
        }
    }
    if (moved) b->last_counts.clear();
    std::vector<int64_t> cnt(counts, counts + nf);
    int64_t total = 0;
    for (int64_t c : cnt) {
        if (c < 0) return fail(ARB_EINVAL, "negative query count");
        total += c;
    }
    if (total == 0) return ARB_OK;
    if (!q_dev || ldq < b->ndim) return fail(ARB_EINVAL, "arb_batch_query: bad positions");
    if ((b->has_comps && !comps_dev) || (b->has_norm && (!norms_dev || !grads_dev)))
        return fail(ARB_EINVAL, "arb_batch_query: missing output buffer");
    if (cnt != b->last_counts) {
        const size_t dsz = arb_batch_field_desc_size();
        int64_t q0 = 0, item = 0;
        for (int k = 0; k < nf; ++k) {
            arb_batch_fill_desc(b->desc_host.data() + dsz * k, b->fields[k]->g, b->fields[k]->table,
                                q0, cnt[k], item);
            q0 += cnt[k];
            item += (cnt[k] + b->chunk - 1) / b->chunk;
        }
        if (item >= (int64_t)UINT32_MAX - (int64_t)b->ncu * 64)
            return fail(ARB_EINVAL, "too many work items");
        b->nitems = item;
        HIPCHK(hipMemcpyAsync(b->desc_dev, b->desc_host.data(), dsz * nf, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        b->last_counts = cnt;
    }
    if (flags & ARB_Q_CHECK) HIPCHK(hipMemsetAsync(b->flag, 0, sizeof(int32_t), s));
    int64_t nblocks = (int64_t)b->ncu * b->blocks_per_cu;
    if (nblocks > b->nitems) nblocks = b->nitems;
    HIPCHK(arb_launch_batch_query(b->ndim, b->precision, (flags & ARB_Q_OUT_F64) ? 1 : 0,
                                  b->desc_dev, nf, b->nitems, b->chunk, b->threads, (int32_t)nblocks,
                                  b->ninterp, b->has_comps, b->has_norm, q_dev, ldq, comps_dev,
                                  norms_dev, grads_dev, b->flag, b->counter, s));
    if (flags & ARB_Q_CHECK) {
        int32_t h = 0;
        HIPCHK(hipMemcpyAsync(&h, b->flag, sizeof h, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (h) return fail(ARB_EINDEX, "a query reduced to a cell index past the sentinel");
    }
    return ARB_OK;
}

void arb_batch_destroy(arb_batch* b) {
    if (!b) return;
    for (arb_field* f : b->fields)
        if (f) f->batches.erase(std::remove(f->batches.begin(), f->batches.end(), b),
                                f->batches.end());
    if (b->desc_dev) (void)hipFree(b->desc_dev);
    if (b->counter) (void)hipFree(b->counter);
    if (b->flag) (void)hipFree(b->flag);
    delete b;
}

int arb_set_table_limit(arb_field* f, int64_t max_bytes) {
    if (!f || max_bytes < 0) return fail(ARB_EINVAL, "arb_set_table_limit: bad argument");
    f->table_limit = max_bytes;
    if (max_bytes > 0 && f->table && f->table_bytes > max_bytes) {
        if (f->table_loaded)
            return fail(ARB_EINVAL, "the loaded coefficient table (%lld bytes) exceeds the limit",
                        (long long)f->table_bytes);
        HIPCHK(hipDeviceSynchronize());     // no queued work may still read the table
        (void)hipFree(f->table);
        f->table = nullptr;
        f->table_built = 0;
    }
    // the table is tried again under the new limit at its next build
    f->otf = (max_bytes > 0 && f->table_bytes > max_bytes) ? 1 : 0;
    return ARB_OK;
}

int arb_last_timings(const arb_field* f, float* ms5) {
    if (!f || !ms5) return fail(ARB_EINVAL, "arb_last_timings: NULL argument");
    for (int k = 0; k < 5; ++k) ms5[k] = f->last_ms[k];
    return ARB_OK;
}

int arb_device_copy(void* dst_dev, const void* src_dev, int64_t bytes, void* stream) {
    if (!dst_dev || !src_dev || bytes < 0 || bytes % 16 ||
        ((uintptr_t)dst_dev | (uintptr_t)src_dev) % 16)
        return fail(ARB_EINVAL, "arb_device_copy: pointers and size must be 16-byte multiples");
    HIPCHK(arb_launch_copy16(dst_dev, src_dev, bytes / 16, (hipStream_t)stream));
    return ARB_OK;
}

int arb_set_tuning(arb_field* f, const int32_t* block_cells, int32_t threads) {
    if (!f) return fail(ARB_EINVAL, "arb_set_tuning: NULL handle");
    if (threads != 0 && (threads < 64 || threads > 256 || threads % 64))
        return fail(ARB_EINVAL, "threads must be 64, 128, 192 or 256");
    const int d = f->g.ndim;
    if (block_cells) {
        int64_t cells = 1;
        for (int a = 0; a < d; ++a) {
            if (block_cells[a] < 1 || block_cells[a] > 64 || (block_cells[a] & (block_cells[a] - 1)))
                return fail(ARB_EINVAL, "block edge must be a power of two in 1..64 cells");
            cells *= block_cells[a];
        }
        const int K = 1 << (2 * d);
        const int esz = f->precision == ARB_PREC_FP64 ? 8 : 4;
        const int64_t lds = cells * (f->ninterp * K * esz / 16 + 1) * 16;
        if (lds > 160 * 1024)
            return fail(ARB_EINVAL, "block needs %lld bytes of LDS (max 163840)", (long long)lds);
        for (int a = 0; a < d; ++a) f->user_block[a] = block_cells[a];
    } else {
        for (int a = 0; a < ARB_MAXD; ++a) f->user_block[a] = 0;
    }
    f->eval_threads = threads ? threads : 256;
    choose_buckets(f, f->B, true);
    f->BF.nbt = 0;                                  // chosen again at the next fused query
    f->tiles_ok = 0;                                // node tiles follow B
    return ARB_OK;
}

int arb_table_copy(arb_field* f, int64_t cell0, int64_t ncells, void* dst_dev, void* stream) {
    if (!f || !dst_dev || cell0 < 0 || ncells < 0 || cell0 + ncells > f->g.nc + 1)
        return fail(ARB_EINVAL, "arb_table_copy: bad range");
    int rc = arb_build_coeffs(f, stream);
    if (rc) return rc;
    const int K = 1 << (2 * f->g.ndim);
    const size_t esz = f->precision == ARB_PREC_FP64 ? sizeof(double) : sizeof(float);
    const size_t row = esz * (size_t)f->ninterp * K;
    HIPCHK(hipMemcpyAsync(dst_dev, (char*)f->table + row * cell0, row * ncells,
                          hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return ARB_OK;
}

int arb_table_load(arb_field* f, int64_t cell0, int64_t ncells, const void* src_dev,
                   void* stream) {
    if (!f || !src_dev || cell0 < 0 || ncells < 0 || cell0 + ncells > f->g.nc + 1)
        return fail(ARB_EINVAL, "arb_table_load: bad range");
    const int K = 1 << (2 * f->g.ndim);
    const size_t esz = f->precision == ARB_PREC_FP64 ? sizeof(double) : sizeof(float);
    const size_t row = esz * (size_t)f->ninterp * K;
    // rows not loaded keep the built coefficients; a partial load overlays them
    int rc = ensure_table_alloc(f);
    if (rc) return rc;
    rc = ensure_sentinel(f, (hipStream_t)stream);
    if (rc) return rc;
    if (!(cell0 == 0 && ncells == f->g.nc + 1)) {
        rc = arb_build_coeffs(f, stream);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync((char*)f->table + row * cell0, src_dev, row * ncells,
                          hipMemcpyDeviceToDevice, (hipStream_t)stream));
    f->table_built = 1;
    f->table_loaded = 1;      // the table is authoritative: queries read it
    return ARB_OK;
}

void arb_field_destroy(arb_field* f) {
    if (!f) return;
    for (arb_batch* b : f->batches) {
        b->dead = 1;
        for (auto& bf : b->fields)
            if (bf == f) bf = nullptr;
    }
    if (f->vals) (void)hipFree(f->vals);
    if (f->tiles) (void)hipFree(f->tiles);
    if (f->table) (void)hipFree(f->table);
    if (f->sent) (void)hipFree(f->sent);
    if (f->A) (void)hipFree(f->A);
    if (f->flag) (void)hipFree(f->flag);
    if (f->ws_bkt) (void)hipFree(f->ws_bkt);
    if (f->ws_rank) (void)hipFree(f->ws_rank);
    if (f->ws_perm) (void)hipFree(f->ws_perm);
    if (f->ws_rec) (void)hipFree(f->ws_rec);
    if (f->ws_qrec) (void)hipFree(f->ws_qrec);
    if (f->ws_count) (void)hipFree(f->ws_count);
    if (f->ws_offs) (void)hipFree(f->ws_offs);
    for (int k = 0; k < 2; ++k) {
        if (f->rs_k[k]) (void)hipFree(f->rs_k[k]);
        if (f->rs_v[k]) (void)hipFree(f->rs_v[k]);
    }
    if (f->rs_hist) (void)hipFree(f->rs_hist);
    if (f->rs_hscan) (void)hipFree(f->rs_hscan);
    if (f->bw_key) (void)hipFree(f->bw_key);
    if (f->bw_perm) (void)hipFree(f->bw_perm);
    if (f->bw_k1) (void)hipFree(f->bw_k1);
    if (f->bw_v1) (void)hipFree(f->bw_v1);
    if (f->bw_hist) (void)hipFree(f->bw_hist);
    if (f->bw_hscan) (void)hipFree(f->bw_hscan);
    if (f->bw_part) (void)hipFree(f->bw_part);
    if (f->bw_rec) (void)hipFree(f->bw_rec);
    if (f->bw_count) (void)hipFree(f->bw_count);
    if (f->bw_offs) (void)hipFree(f->bw_offs);
    if (f->bw_bbar) (void)hipFree(f->bw_bbar);
    if (f->ev_ready)
        for (auto& e : f->ev) (void)hipEventDestroy(e);
    if (f->push_exec) (void)hipGraphExecDestroy(f->push_exec);
    if (f->pin) (void)hipHostFree(f->pin);
    if (f->cap_stream) (void)hipStreamDestroy(f->cap_stream);
    delete f;
}

}  // extern "C"
