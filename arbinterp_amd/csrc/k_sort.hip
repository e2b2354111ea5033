// Stable LSD radix sort of (key, value) pairs with 8-bit digits, and the
// helpers the bucketed query path needs around it.  Used to group queries by
// bucket (block of cells) without device-scope atomics.
//
// One pass over digit d:
//   histogram  each tile of 8192 pairs counts its digits in LDS and writes
//              hist[digit * ntiles + tile] (k_rs_upsweep; for the first pass
//              the query kernel that makes the keys writes it);
//   digit scan one workgroup per digit scans that digit's tile counts:
//              hscan[digit * ntiles + tile] = pairs of the digit in earlier
//              tiles, tot[digit] = the digit's total;
//   downsweep  each tile adds the exclusive scan of tot (the start of every
//              digit's run) to its hscan entries, ranks its pairs stably
//              (wave ballots match equal digits, per-wave LDS counters keep
//              element order) and writes them to that start + rank.
// Tiles are ARB_RS_THREADS threads x 16 items; items are striped inside each
// wave (item j of lane l of wave w is element w*1024 + j*64 + l), which keeps
// the ranking stable in element order.  Threads below 256 also own a digit
// each in the per-digit steps.
#include "arb_internal.h"

namespace {

constexpr int RS_THREADS = ARB_RS_THREADS;
constexpr int RS_ITEMS = ARB_RS_ITEMS;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;   // 8192
constexpr int RS_WAVES = RS_THREADS / 64;

// Tile of a scatter pass for workgroup b of ntiles.  Workgroups are handed
// to the eight XCDs in turn (b -> XCD b % 8), each XCD with its own L2.  A
// digit's run from tile t ends where the same digit's run from tile t + 1
// begins, so the line they share is written by both: giving each XCD a
// contiguous range of tiles, taken in order, has both halves of such a line
// written through the same L2, close together in time, instead of two
// partial lines from two L2s.
constexpr int RS_XCDS = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t ntiles) {
    const int64_t x = b % RS_XCDS, k = b / RS_XCDS;
    const int64_t q = ntiles / RS_XCDS, r = ntiles % RS_XCDS;
    return x * q + (x < r ? x : r) + k;        // XCD x holds tiles [x q + min(x, r), +q + (x < r))
}

__global__ void __launch_bounds__(RS_THREADS)
k_rs_upsweep(const uint32_t* __restrict__ keys, int64_t n, int shift, int32_t* __restrict__ hist,
             int64_t ntiles) {
    __shared__ int32_t cnt[256];
    const bool dt = threadIdx.x < 256;              // this thread owns a digit
    if (dt) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t e = base + j * RS_THREADS + threadIdx.x;
        if (e < n) atomicAdd(&cnt[(__builtin_nontemporal_load(&keys[e]) >> shift) & 255u], 1);
    }
    __syncthreads();
    if (dt) hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// One workgroup per digit: hscan[d * ntiles + t] = sum of hist[d * ntiles + u]
// over u < t, tot[d] = the sum over all tiles.  Chunks of RS_TILE counts,
// RS_ITEMS consecutive ones per thread.
// zero[0 .. zlen): cleared on the side (the grouping sort's per-key counts).
__global__ void __launch_bounds__(RS_THREADS)
k_rs_digit_scan(const int32_t* __restrict__ hist, int64_t ntiles, int32_t* __restrict__ hscan,
                int32_t* __restrict__ tot, int32_t* __restrict__ zero, int64_t zlen) {
    __shared__ int32_t part[RS_THREADS];
    for (int64_t i = (int64_t)blockIdx.x * RS_THREADS + threadIdx.x; i < zlen;
         i += (int64_t)gridDim.x * RS_THREADS)
        zero[i] = 0;
    const int64_t base = (int64_t)blockIdx.x * ntiles;
    int32_t carry = 0;
    for (int64_t c0 = 0; c0 < ntiles; c0 += RS_TILE) {
        const int64_t t0 = c0 + (int64_t)threadIdx.x * RS_ITEMS;
        int32_t v[RS_ITEMS];
        int32_t sum = 0;
#pragma unroll
        for (int j = 0; j < RS_ITEMS; ++j) {
            v[j] = (t0 + j < ntiles) ? hist[base + t0 + j] : 0;
            sum += v[j];
        }
        part[threadIdx.x] = sum;
        __syncthreads();
        for (int off = 1; off < RS_THREADS; off <<= 1) {
            const int32_t x = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += x;
            __syncthreads();
        }
        int32_t run = carry + part[threadIdx.x] - sum;
#pragma unroll
        for (int j = 0; j < RS_ITEMS; ++j) {
            if (t0 + j < ntiles) hscan[base + t0 + j] = run;
            run += v[j];
        }
        carry += part[RS_THREADS - 1];
        __syncthreads();                            // part is reused by the next chunk
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Exclusive scan over the 256 digits of one value per thread (threads 0..255,
// the first four waves) and the inclusive total, with wave shuffles and one
// barrier; every thread of the workgroup must call it.
__device__ __forceinline__ int32_t digits_exscan(int32_t v, int32_t* wsum, int32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t x = threadIdx.x < 256 ? v : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (w < 4 && lane == 63) wsum[w] = x;
    __syncthreads();
    int32_t before = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) before += (k < w) ? wsum[k] : 0;
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    return before + x - (threadIdx.x < 256 ? v : 0);
}

// First pass of the LSD sort (the lowest digit, shift 0; the values are the
// element indices).  An LSD sort needs only its later passes to be stable, so
// each pair's place inside its digit comes from an LDS atomic instead of the
// stable wave-ballot ranking of k_rs_downsweep.  The tile's digit counts are
// the first pass's histogram, made with the keys.
// PACK (the grouping sort's first pass, digit = the high 8 bits, shift =
// bits - 8): the output is one word per pair, the key's low `shift` bits
// above the 24-bit element index, in kout; vout is not used.
// The tile in LDS, in digit order: keys and values, or (PACK) the packed
// words and their digits (40 KB instead of 64 KB: more tiles per CU).
template <bool PACK>
struct Pass1Tile {
    uint32_t skey[RS_TILE];
    int32_t sval[RS_TILE];
};
template <>
struct Pass1Tile<true> {
    uint32_t sword[RS_TILE];
    uint8_t sdig[RS_TILE];
};

template <bool PACK>
__global__ void __launch_bounds__(RS_THREADS)
k_rs_pass1(const uint32_t* __restrict__ kin, uint32_t* __restrict__ kout,
           int32_t* __restrict__ vout, int64_t n, const int32_t* __restrict__ hist,
           const int32_t* __restrict__ hscan, const int32_t* __restrict__ tot,
           int64_t ntiles, int shift) {
    __shared__ int32_t cnt[256];                // next free place per digit in the tile
    __shared__ int32_t gbase[256];              // global start of the tile's part of each run
    __shared__ int32_t toff[256];               // tile-local start of each digit's run
    __shared__ int32_t wsum[2][4];
    __shared__ Pass1Tile<PACK> L;
    const bool dt = threadIdx.x < 256;
    const int dg0 = dt ? threadIdx.x : 0;
    const int64_t tile = xcd_tile(blockIdx.x, ntiles);
    const int32_t tc = dt ? hist[(int64_t)dg0 * ntiles + tile] : 0;
    const int32_t gt = dt ? tot[dg0] : 0;
    const int32_t hs = dt ? hscan[(int64_t)dg0 * ntiles + tile] : 0;
    int32_t ttot, gtot;
    const int32_t tex = digits_exscan(tc, wsum[0], &ttot);
    const int32_t gex = digits_exscan(gt, wsum[1], &gtot);
    if (dt) {
        toff[dg0] = tex;
        cnt[dg0] = tex;
        gbase[dg0] = gex + hs;
    }
    __syncthreads();
    const int64_t base = tile * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t e = base + (int64_t)j * RS_THREADS + threadIdx.x;
        if (e < n) {
            const uint32_t k = __builtin_nontemporal_load(&kin[e]);
            const uint32_t dg = (k >> shift) & 255u;
            const int lp = atomicAdd(&cnt[dg], 1);
            // lp < RS_TILE: the tile holds tc[d] pairs of digit d, its run
            // starts at toff[d] and the runs tile [0, RS_TILE)
            if constexpr (PACK) {
                L.sword[lp] = ((k & ((1u << shift) - 1u)) << 24) | (uint32_t)e;
                L.sdig[lp] = (uint8_t)dg;
            } else {
                L.skey[lp] = k;
                L.sval[lp] = (int32_t)e;
            }
        }
    }
    __syncthreads();
    const int cntt = (int)((n - base) < RS_TILE ? (n - base) : RS_TILE);
    for (int i = threadIdx.x; i < cntt; i += RS_THREADS) {
        uint32_t dg;
        if constexpr (PACK) dg = L.sdig[i];
        else dg = (L.skey[i] >> shift) & 255u;
        const int64_t pos = (int64_t)gbase[dg] + (i - toff[dg]);
        if (pos < 0 || pos >= n) continue;    // unreachable: a digit's run stays inside [0, n)
        if constexpr (PACK) {
            kout[pos] = L.sword[i];
        } else {
            vout[pos] = L.sval[i];
            kout[pos] = L.skey[i];
        }
    }
}

__global__ void __launch_bounds__(RS_THREADS)
k_rs_downsweep(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
               uint32_t* __restrict__ kout, int32_t* __restrict__ vout, int64_t n, int shift,
               const int32_t* __restrict__ hscan, const int32_t* __restrict__ tot,
               int64_t ntiles, int32_t* __restrict__ first, int32_t fshift) {
    __shared__ int32_t wcnt[RS_WAVES][256];     // running digit counts per wave
    __shared__ int32_t gbase[256];              // global start of each digit's run
    __shared__ int32_t toff[256];               // tile-local start of each digit's run
    __shared__ uint32_t skey[RS_TILE];          // the tile's pairs in digit order
    __shared__ int32_t sval[RS_TILE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool dt = threadIdx.x < 256;              // this thread owns digit threadIdx.x
    const int dg0 = dt ? threadIdx.x : 0;
    for (int d = threadIdx.x; d < RS_WAVES * 256; d += RS_THREADS) (&wcnt[0][0])[d] = 0;
    const int32_t dtot = tot[dg0];
    if (dt) gbase[dg0] = dtot;
    __syncthreads();
    const int64_t tile = xcd_tile(blockIdx.x, ntiles);
    const int64_t base = tile * RS_TILE + (int64_t)w * (64 * RS_ITEMS);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t key[RS_ITEMS];
    int32_t val[RS_ITEMS];
    int32_t rnk[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t e = base + j * 64 + lane;
        const bool ok = e < n;
        // pairs are read once: nontemporal, to keep the query positions cached
        key[j] = ok ? __builtin_nontemporal_load(&kin[e]) : 0xFFFFFFFFu;
        val[j] = ok ? (vin ? __builtin_nontemporal_load(&vin[e]) : (int32_t)e) : 0;   // no vin: the index
        const uint32_t dg = (key[j] >> shift) & 255u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t on = __ballot(ok && ((dg >> b) & 1u));
            peers &= ((dg >> b) & 1u) ? on : ~on;
        }
        // peers: lanes of this item slot holding the same digit (all valid)
        const int before = __popcll(peers & lt);
        const int run = wcnt[w][dg];
        rnk[j] = run + before;
        if (ok && before == 0) wcnt[w][dg] = run + __popcll(peers);   // group leader
    }
    __syncthreads();
    // exclusive prefix over the waves of this tile, per digit
    int32_t acc = 0;
    if (dt) {
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            const int32_t c = wcnt[ww][dg0];
            wcnt[ww][dg0] = acc;
            acc += c;
        }
        toff[dg0] = acc;
    }
    // exclusive scans over the 256 digits, side by side: the tile's digit
    // totals (tile-local run starts) and the global digit totals (global run
    // starts, to which the pairs of the digit in earlier tiles are added)
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const bool on = dt && dg0 >= off;
        const int32_t x = on ? toff[dg0 - off] : 0;
        const int32_t y = on ? gbase[dg0 - off] : 0;
        __syncthreads();
        if (dt) {
            toff[dg0] += x;
            gbase[dg0] += y;
        }
        __syncthreads();
    }
    if (dt) {
        toff[dg0] -= acc;
        gbase[dg0] += hscan[(int64_t)dg0 * ntiles + tile] - dtot;
    }
    __syncthreads();
    // local reorder: digit-major, element order inside each digit
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t e = base + j * 64 + lane;
        if (e >= n) continue;
        const uint32_t dg = (key[j] >> shift) & 255u;
        const int lp = toff[dg] + wcnt[w][dg] + rnk[j];
        skey[lp] = key[j];
        sval[lp] = val[j];
    }
    __syncthreads();
    // write out: consecutive slots of one digit go to consecutive addresses.
    // Last pass with `first`: the keys are not written; instead every slot
    // that starts a run of equal (key >> fshift) inside its digit's run
    // offers its position to first[key >> fshift] (the sorted order makes
    // the minimum over tiles the group's first position).
    const int64_t tbase = tile * RS_TILE;
    const int cnt = (int)((n - tbase) < RS_TILE ? (n - tbase) : RS_TILE);
    for (int i = threadIdx.x; i < cnt; i += RS_THREADS) {
        const uint32_t k = skey[i];
        const uint32_t dg = (k >> shift) & 255u;
        const int64_t pos = (int64_t)gbase[dg] + (i - toff[dg]);
        if (pos < 0 || pos >= n) continue;    // unreachable: a digit's run stays inside [0, n)
        vout[pos] = sval[i];
        if (first) {
            if (i == toff[dg] || (skey[i - 1] >> fshift) != (k >> fshift))
                atomicMin(&first[k >> fshift], (int32_t)pos);
        } else {
            kout[pos] = k;
        }
    }
}

// ---------------------------------------------------------------------------
// Grouping sort: keys below 2^bits (9 <= bits <= 16), n < 2^24 rows.  The
// evaluation needs the rows grouped by key, in any order inside a group, so
// the second digit need not keep the first pass's order: the first pass
// (k_rs_pass1<true>) groups the rows by the key's high 8 bits into 256
// segments, and each segment is then grouped by the remaining low bits on its
// own, with no stable ranking:
//   k_grp_count  each tile of a segment counts its low digits in LDS and adds
//                them to the segment's per-key counts (global atomics, at most
//                256 per tile);
//   k_grp_starts per segment, the exclusive scan of its keys' counts from the
//                segment start: each key's first position, written to offs
//                (the bucket bounds) and to the counters;
//   k_grp_place  each tile ranks its rows per low digit in LDS, takes the
//                places of each digit's rows with one atomic on the key's
//                counter, and writes the row indices there in digit order.
// A segment of c rows has ceil(c / RS_TILE) tiles; the grids hold one
// workgroup per tile of any segment (at most ntiles + 256), and each
// workgroup finds its segment and tile from the 256 segment sizes.
// ---------------------------------------------------------------------------
struct SegTile {
    int64_t lo, hi;     // the tile's entries [lo, hi)
    int d;              // its segment (high digit)
};

// Block-wide; false (for every thread) when blockIdx.x is past the last tile.
__device__ __forceinline__ bool seg_tile(const int32_t* __restrict__ tot, int32_t* sS, int32_t* sT,
                                         int32_t (*wsum)[4], int* sd, SegTile* st) {
    const bool dt = threadIdx.x < 256;
    const int32_t c = dt ? tot[threadIdx.x] : 0;
    const int32_t nt = (c + RS_TILE - 1) / RS_TILE;
    int32_t ctot, ttot;
    const int32_t cex = digits_exscan(c, wsum[0], &ctot);
    const int32_t tex = digits_exscan(nt, wsum[1], &ttot);
    if (dt) {
        sS[threadIdx.x] = cex;
        sT[threadIdx.x] = tex;
    }
    if (threadIdx.x == 0) {
        sS[256] = ctot;
        sT[256] = ttot;
    }
    __syncthreads();
    const int32_t b = (int32_t)blockIdx.x;
    if (b >= sT[256]) return false;
    // exactly one segment holds tile b: sT[d] <= b < sT[d + 1] (empty
    // segments have sT[d] == sT[d + 1])
    if (dt && sT[threadIdx.x] <= b && b < sT[threadIdx.x + 1]) *sd = threadIdx.x;
    __syncthreads();
    const int d = *sd;
    st->d = d;
    st->lo = (int64_t)sS[d] + (int64_t)(b - sT[d]) * RS_TILE;
    const int64_t end = sS[d + 1];
    st->hi = st->lo + RS_TILE < end ? st->lo + RS_TILE : end;
    return true;
}

__global__ void __launch_bounds__(RS_THREADS)
k_grp_count(const uint32_t* __restrict__ ent, const int32_t* __restrict__ tot, int hs,
            int32_t* __restrict__ kcnt, int64_t nbt) {
    __shared__ int32_t sS[257], sT[257], wsum[2][4], cnt[256];
    __shared__ int sd;
    if (threadIdx.x < 256) cnt[threadIdx.x] = 0;
    SegTile st;
    if (!seg_tile(tot, sS, sT, wsum, &sd, &st)) return;
    for (int64_t e = st.lo + threadIdx.x; e < st.hi; e += RS_THREADS)
        atomicAdd(&cnt[ent[e] >> 24], 1);
    __syncthreads();
    if (threadIdx.x < (1 << hs)) {
        const int64_t key = ((int64_t)st.d << hs) | threadIdx.x;
        const int32_t c = cnt[threadIdx.x];
        if (c > 0 && key <= nbt) atomicAdd(&kcnt[key], c);
    }
}

// One workgroup of 256 threads per segment.  hot[0] = 1 when a key below
// nbt (a bucket; nbt is the rows outside the grid) holds more than chunk rows.
__global__ void __launch_bounds__(256)
k_grp_starts(const int32_t* __restrict__ tot, int hs, int32_t* __restrict__ kcnt,
             int32_t* __restrict__ offs, int64_t nbt, int32_t* __restrict__ hot, int32_t chunk) {
    __shared__ int32_t wsum[2][4];
    __shared__ int32_t sbase;
    int32_t total;
    const int32_t sex = digits_exscan(tot[threadIdx.x], wsum[0], &total);
    if (threadIdx.x == blockIdx.x) sbase = sex;  // the segment's start
    const int64_t key = ((int64_t)blockIdx.x << hs) + threadIdx.x;
    const bool in = threadIdx.x < (1 << hs) && key <= nbt;
    const int32_t c = in ? kcnt[key] : 0;
    int32_t ctot;
    const int32_t lex = digits_exscan(c, wsum[1], &ctot);   // its barrier publishes sbase
    if (in) {
        const int32_t first = sbase + lex;
        kcnt[key] = first;
        offs[key] = first;
        if (key < nbt && c > chunk) hot[0] = 1;
    }
}

__global__ void __launch_bounds__(RS_THREADS)
k_grp_place(const uint32_t* __restrict__ ent, const int32_t* __restrict__ tot, int hs,
            int32_t* __restrict__ kcur, int32_t* __restrict__ perm, int64_t n, int64_t nbt) {
    __shared__ int32_t sS[257], sT[257], wsum[3][4], cnt[256], toff[256], gbase[256];
    __shared__ uint32_t sorted[RS_TILE];        // the tile's entries in low-digit order
    __shared__ int sd;
    if (threadIdx.x < 256) cnt[threadIdx.x] = 0;
    SegTile st;
    if (!seg_tile(tot, sS, sT, wsum, &sd, &st)) return;
    uint32_t v[RS_ITEMS];
    int32_t r[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t e = st.lo + (int64_t)j * RS_THREADS + threadIdx.x;
        v[j] = 0;
        r[j] = 0;
        if (e < st.hi) {
            v[j] = __builtin_nontemporal_load(&ent[e]);
            r[j] = atomicAdd(&cnt[v[j] >> 24], 1);
        }
    }
    __syncthreads();
    const int32_t c = threadIdx.x < 256 ? cnt[threadIdx.x] : 0;
    int32_t ttot;
    const int32_t ex = digits_exscan(c, wsum[2], &ttot);
    if (threadIdx.x < 256) {
        toff[threadIdx.x] = ex;
        const int64_t key = ((int64_t)st.d << hs) | threadIdx.x;
        gbase[threadIdx.x] = (c > 0 && key <= nbt) ? atomicAdd(&kcur[key], c) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t e = st.lo + (int64_t)j * RS_THREADS + threadIdx.x;
        if (e < st.hi) sorted[toff[v[j] >> 24] + r[j]] = v[j];
    }
    __syncthreads();
    const int cntt = (int)(st.hi - st.lo);
    for (int i = threadIdx.x; i < cntt; i += RS_THREADS) {
        const uint32_t sv = sorted[i];
        const uint32_t l = sv >> 24;
        const int64_t pos = (int64_t)gbase[l] + (i - toff[l]);
        if (pos < 0 || pos >= n) continue;    // unreachable: a key's places stay inside [0, n)
        perm[pos] = (int32_t)(sv & 0xFFFFFFu);
    }
}

// offs[b] = min(first[b .. len-1]) for b in [0, len): the first sorted
// position whose group is >= b (first[b] = n for an absent group).  Two
// launches over blocks of SF_BLOCK entries (SF_ITEMS consecutive ones per
// thread): the minimum of every block, then per block the suffix minimum,
// seeded with the minimum of the blocks after it.
constexpr int SF_THREADS = 256, SF_ITEMS = 16, SF_BLOCK = SF_THREADS * SF_ITEMS;

__global__ void __launch_bounds__(SF_THREADS)
k_block_min(const int32_t* __restrict__ first, int64_t len, int32_t* __restrict__ bmin) {
    __shared__ int32_t part[SF_THREADS];
    const int64_t b0 = (int64_t)blockIdx.x * SF_BLOCK + (int64_t)threadIdx.x * SF_ITEMS;
    int32_t m = INT32_MAX;
#pragma unroll
    for (int j = 0; j < SF_ITEMS; ++j)
        if (b0 + j < len) m = min(m, first[b0 + j]);
    part[threadIdx.x] = m;
    __syncthreads();
    for (int off = SF_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) part[threadIdx.x] = min(part[threadIdx.x], part[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) bmin[blockIdx.x] = part[0];
}

__global__ void __launch_bounds__(SF_THREADS)
k_suffix_min(const int32_t* __restrict__ first, int64_t len, const int32_t* __restrict__ bmin,
             int32_t* __restrict__ offs) {
    __shared__ int32_t part[SF_THREADS];
    // minimum over the later blocks
    int32_t carry = INT32_MAX;
    for (int64_t k = blockIdx.x + 1 + threadIdx.x; k < gridDim.x; k += SF_THREADS)
        carry = min(carry, bmin[k]);
    part[threadIdx.x] = carry;
    __syncthreads();
    for (int off = SF_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) part[threadIdx.x] = min(part[threadIdx.x], part[threadIdx.x + off]);
        __syncthreads();
    }
    carry = part[0];
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * SF_BLOCK + (int64_t)threadIdx.x * SF_ITEMS;
    int32_t v[SF_ITEMS];
    int32_t m = INT32_MAX;
#pragma unroll
    for (int j = 0; j < SF_ITEMS; ++j) {
        v[j] = (b0 + j < len) ? first[b0 + j] : INT32_MAX;
        m = min(m, v[j]);
    }
    // inclusive suffix minimum over the threads of the block
    part[threadIdx.x] = m;
    __syncthreads();
    for (int off = 1; off < SF_THREADS; off <<= 1) {
        const int32_t x = threadIdx.x + off < SF_THREADS ? part[threadIdx.x + off] : INT32_MAX;
        __syncthreads();
        part[threadIdx.x] = min(part[threadIdx.x], x);
        __syncthreads();
    }
    int32_t run = min(carry, threadIdx.x + 1 < SF_THREADS ? part[threadIdx.x + 1] : INT32_MAX);
#pragma unroll
    for (int j = SF_ITEMS - 1; j >= 0; --j) {
        run = min(run, v[j]);
        if (b0 + j < len) offs[b0 + j] = run;
    }
}

// Exclusive scan of len int32 values (in -> out, out[len] = total) in three
// launches: block sums, scan of the block sums, block-local scans.
constexpr int SC_THREADS = 256, SC_ITEMS = 16, SC_BLOCK = SC_THREADS * SC_ITEMS;

__global__ void __launch_bounds__(SC_THREADS)
k_sc_sums(const int32_t* __restrict__ in, int64_t len, int32_t* __restrict__ sums) {
    __shared__ int32_t red[SC_THREADS];
    const int64_t base = (int64_t)blockIdx.x * SC_BLOCK;
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
        const int64_t e = base + j * SC_THREADS + threadIdx.x;
        if (e < len) s += in[e];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = SC_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(SC_THREADS)
k_sc_local(const int32_t* __restrict__ in, int64_t len, const int32_t* __restrict__ bscan,
           int32_t* __restrict__ out) {
    __shared__ int32_t part[SC_THREADS];
    const int64_t base = (int64_t)blockIdx.x * SC_BLOCK + (int64_t)threadIdx.x * SC_ITEMS;
    int32_t v[SC_ITEMS];
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
        v[j] = (base + j < len) ? in[base + j] : 0;
        s += v[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < SC_THREADS; off <<= 1) {
        const int32_t x = (threadIdx.x >= off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    int32_t run = bscan[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
        if (base + j < len) out[base + j] = run;
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[len] = bscan[gridDim.x];   // total
}

__global__ void k_fill_first(int32_t* __restrict__ p, int64_t len, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) p[i] = v;
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

int64_t arb_scan_tmp_elems(int64_t len) { return (int64_t)cdiv(len, SC_BLOCK) + 1; }

// out[0..len] (len + 1 entries, out[len] = the total) exclusive scan of
// in[0..len); tmp and tmp_scan hold arb_scan_tmp_elems(len) ints each (block
// sums and their scan).
hipError_t arb_launch_scan_long(const int32_t* in, int32_t* out, int64_t len, int32_t* tmp,
                                int32_t* tmp_scan, hipStream_t s) {
    const unsigned nblk = cdiv(len, SC_BLOCK);
    if (len <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sc_sums, dim3(nblk), dim3(SC_THREADS), 0, s, in, len, tmp);
    hipError_t e = arb_launch_scan(tmp, tmp_scan, nblk, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sc_local, dim3(nblk), dim3(SC_THREADS), 0, s, in, len, tmp_scan, out);
    return hipGetLastError();
}

int64_t arb_rs_ntiles(int64_t n) { return (int64_t)cdiv(n, RS_TILE); }

int64_t arb_rs_first_elems(int64_t ngroups) {
    return ngroups + 1 + (int64_t)cdiv(ngroups + 1, SF_BLOCK) + 1;
}

// Sort (k0, v0) by the low `bits` bits of the keys.  Buffers: k1/v1 the same
// size as k0/v0; hist 256 * ntiles ints; hscan 256 * ntiles + 256 ints (the
// per-tile digit prefixes, then the digit totals).  hist0_ready: hist already
// holds the first pass's tile histograms (written with the keys).
// index_vals: v0 is not read; the value of pair i is i.  stable: every pass
// keeps the order of equal keys (with index_vals, the pairs of one key end in
// index order); else the first pass ranks with LDS atomics.  The sorted values
// end in v0 when the number of passes is even, else in v1; *in_first reports
// which.  The sorted keys are not kept: instead offs[g] (g in [0, ngroups])
// becomes the first sorted position whose key >> gshift is >= g, keys being
// below (ngroups + 1) << gshift; first holds arb_rs_first_elems(ngroups)
// ints of scratch.
hipError_t arb_launch_radix_sort(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int64_t n,
                                 int bits, int32_t* hist, int32_t* hscan, int hist0_ready,
                                 int index_vals, int32_t* offs, int64_t ngroups, int32_t gshift,
                                 int32_t* first, int* in_first, int stable, hipStream_t s) {
    const int64_t ntiles = arb_rs_ntiles(n);
    int32_t* tot = hscan + 256 * ntiles;
    uint32_t* ka = k0;
    int32_t* va = v0;
    uint32_t* kb = k1;
    int32_t* vb = v1;
    int passes = 0;
    hipLaunchKernelGGL(k_fill_first, dim3(cdiv(ngroups + 1, 256)), dim3(256), 0, s, first,
                       ngroups + 1, (int32_t)n);
    for (int shift = 0; shift < bits; shift += 8, ++passes) {
        const bool last = shift + 8 >= bits;
        if (shift > 0 || !hist0_ready)
            hipLaunchKernelGGL(k_rs_upsweep, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, ka, n,
                               shift, hist, ntiles);
        hipLaunchKernelGGL(k_rs_digit_scan, dim3(256), dim3(RS_THREADS), 0, s, hist, ntiles, hscan,
                           tot, (int32_t*)nullptr, (int64_t)0);
        const int32_t* vin = (shift == 0 && index_vals) ? nullptr : va;
        if (shift == 0 && index_vals && !last && !stable)
            hipLaunchKernelGGL(k_rs_pass1<false>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s,
                               ka, kb, vb, n, hist, hscan, tot, ntiles, 0);
        else
            hipLaunchKernelGGL(k_rs_downsweep, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, ka,
                               vin, kb, vb, n, shift, hscan, tot, ntiles, last ? first : nullptr,
                               gshift);
        uint32_t* tk = ka; ka = kb; kb = tk;
        int32_t* tv = va; va = vb; vb = tv;
    }
    const unsigned nsb = cdiv(ngroups + 1, SF_BLOCK);
    int32_t* bmin = first + ngroups + 1;
    hipLaunchKernelGGL(k_block_min, dim3(nsb), dim3(SF_THREADS), 0, s, first, ngroups + 1, bmin);
    hipLaunchKernelGGL(k_suffix_min, dim3(nsb), dim3(SF_THREADS), 0, s, first, ngroups + 1, bmin,
                       offs);
    *in_first = (passes % 2 == 0);
    return hipGetLastError();
}

int arb_group_sort_ok(int64_t n, int bits) {
    return n > 0 && n < ((int64_t)1 << 24) && bits >= 9 && bits <= 16;
}

// Group the rows 0..n-1 by their keys (keys < 2^bits, key <= nbt; see
// arb_group_sort_ok): perm lists the rows of key 0, then of key 1, ..., in
// no set order inside a key, and offs[b] (b in [0, nbt]) is the first
// position of key b.  hist: the tile histograms of the keys' high 8 bits
// (key >> (bits - 8)), written with the keys; hscan as for the radix sort;
// ent: n words; kcnt: arb_rs_first_elems(nbt) ints (the last: the hot flag).
hipError_t arb_launch_group_sort(const uint32_t* keys, uint32_t* ent, int32_t* perm, int64_t n,
                                 int bits, const int32_t* hist, int32_t* hscan, int32_t* offs,
                                 int64_t nbt, int32_t* kcnt, hipStream_t s) {
    if (!arb_group_sort_ok(n, bits) || nbt >= ((int64_t)1 << bits)) return hipErrorInvalidValue;
    const int64_t ntiles = arb_rs_ntiles(n);
    int32_t* tot = hscan + 256 * ntiles;
    const int hs = bits - 8;
    const int64_t klen = arb_rs_first_elems(nbt);
    int32_t* hot = kcnt + klen - 1;
    hipLaunchKernelGGL(k_rs_digit_scan, dim3(256), dim3(RS_THREADS), 0, s, hist, ntiles, hscan,
                       tot, kcnt, klen);
    hipLaunchKernelGGL(k_rs_pass1<true>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, keys,
                       ent, (int32_t*)nullptr, n, hist, hscan, tot, ntiles, hs);
    const unsigned gt = (unsigned)(ntiles + 256);    // tiles of all segments: at most this
    hipLaunchKernelGGL(k_grp_count, dim3(gt), dim3(RS_THREADS), 0, s, ent, tot, hs, kcnt, nbt);
    hipLaunchKernelGGL(k_grp_starts, dim3(256), dim3(256), 0, s, tot, hs, kcnt, offs, nbt, hot,
                       ARB_EVAL_CHUNK);
    hipLaunchKernelGGL(k_grp_place, dim3(gt), dim3(RS_THREADS), 0, s, ent, tot, hs, kcnt, perm, n,
                       nbt);
    return hipGetLastError();
}
