// xss_hybrid.hpp -- MSD-first sort of one large keys-only array.
//
// An LSD sort of 4-byte keys streams the array through HBM four times (each
// pass: a histogram read, then a read and a write).  For large arrays this
// path does two passes and then finishes every bucket on chip:
//
//   pass A   LSD pass on the top digit over the whole array (keys -> scratch)
//   table    the 256 top-digit buckets are cut into ranges (built on the
//            device from pass A's digit totals)
//   pass B   segmented pass on the second digit inside every top bucket
//            (scratch -> keys); its per-bucket digit totals are the exact
//            sizes of the 65536 buckets of the top two digits
//   buckets  bucket starts; buckets larger than one workgroup's LDS are listed
//   local    one workgroup per bucket sorts it by the remaining low digits in
//            LDS (an 8-bit LSD radix per digit, stable) and writes it back
//   large    listed buckets (skewed or duplicate-heavy data) get segmented
//            LSD passes on the low digits (keys -> scratch -> keys)
//
// Everything is enqueued on the stream without a host round trip (the range
// tables are built by kernels and passes over them are launched at the
// tables' capacity, surplus workgroups exiting), so the call stays
// asynchronous and capturable.  Keys-only sorts need no stability, but every
// step is stable anyway.
#pragma once
#include <atomic>
#include <type_traits>
#include <vector>

#include "xss_count.hpp"
#include "xss_sort.hpp"

namespace xss {

#ifndef XSS_HYB_BLOCK
#define XSS_HYB_BLOCK 512
#endif
#ifndef XSS_HYB_ITEMS
#define XSS_HYB_ITEMS 12
#endif
#ifndef XSS_HYB_WAVES
#define XSS_HYB_WAVES 8
#endif
// waves per SIMD the local sort is compiled for (its SGPRs alone would cap it at 7)
#define XSS_HYB_OCC __attribute__((amdgpu_waves_per_eu(XSS_HYB_WAVES)))
constexpr int HYB_BLOCK = XSS_HYB_BLOCK;  // local sort workgroup (build-time knobs)
constexpr int HYB_ITEMS = XSS_HYB_ITEMS;
constexpr uint32_t HYB_CAP = HYB_BLOCK * HYB_ITEMS;  // largest bucket sorted in LDS
constexpr uint32_t HYB_BUCKETS = 65536;              // buckets of the top two digits
constexpr size_t HYB_MIN_N = size_t(1) << 25;        // smaller arrays take the LSD path

struct HybTable {        // a device-built range table of segmented passes
    uint64_t *begin;
    uint64_t *end;
    uint64_t *sbase;     // first key of the range's segment
    uint32_t *bseg;      // segment index
    uint32_t *head;      // 1 at a segment's first range
    uint32_t *nr;        // ranges built (device)
    uint32_t cap;        // capacity = grid of the passes over the table
};

struct HybArgs {
    const uint32_t *top_total;  // [256] key counts per top digit (pass A)
    uint64_t *seg_off;          // [257] top-bucket starts
    const uint32_t *segtot;     // [256][256] second-digit counts per top bucket (pass B)
    uint32_t *bstart;           // [65536] bucket starts
    uint32_t *bsize;            // [65536] bucket sizes
    uint32_t *nlarge;           // buckets larger than HYB_CAP (device)
    uint32_t *lstart;           // [lcap] the large buckets
    uint32_t *lsize;
    uint32_t lcap;
    uint32_t cap;               // largest bucket the LDS sort takes
    HybTable tb;                // ranges of pass B (segments: top buckets)
    HybTable td;                // ranges of the low passes (segments: large buckets)
    uint64_t rk_b;              // keys per range
    uint64_t rk_d;
};

// ranges of pass B from pass A's digit totals (one workgroup)
template <int = 0>
__global__ __launch_bounds__(256) void hyb_top_ranges_kernel(HybArgs a)
{
    __shared__ uint32_t w1[4], w2[4];
    const uint32_t s = threadIdx.x, lane = s & 63, w = s >> 6;
    const uint32_t size = a.top_total[s];
    const uint32_t nrs = uint32_t((size + a.rk_b - 1) / a.rk_b);
    uint32_t i1 = size, i2 = nrs;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t1 = __shfl_up(i1, off), t2 = __shfl_up(i2, off);
        if (lane >= uint32_t(off)) {
            i1 += t1;
            i2 += t2;
        }
    }
    if (lane == 63) {
        w1[w] = i1;
        w2[w] = i2;
    }
    __syncthreads();
    uint32_t p1 = 0, p2 = 0;
    for (uint32_t k = 0; k < w; ++k) {
        p1 += w1[k];
        p2 += w2[k];
    }
    const uint64_t start = p1 + i1 - size;
    const uint32_t r0 = p2 + i2 - nrs;
    a.seg_off[s] = start;
    if (s == 255) {
        a.seg_off[256] = start + size;
        *a.tb.nr = r0 + nrs;
        *a.nlarge = 0;
    }
    for (uint32_t j = 0; j < nrs; ++j) {
        const uint32_t r = r0 + j;
        if (r >= a.tb.cap) break;  // cannot happen: cap = ceil(n / rk) + 256
        const uint64_t b = start + uint64_t(j) * a.rk_b;
        a.tb.begin[r] = b;
        a.tb.end[r] = min(start + size, b + a.rk_b);
        a.tb.sbase[r] = start;
        a.tb.bseg[r] = s;
        a.tb.head[r] = j == 0 ? 1u : 0u;
    }
}

// bucket starts / sizes of the top two digits; list the large ones
// (workgroup = top bucket, thread = second digit)
template <int = 0>
__global__ __launch_bounds__(256) void hyb_buckets_kernel(HybArgs a)
{
    __shared__ uint32_t wsum[4];
    const uint32_t s = blockIdx.x, d = threadIdx.x, lane = d & 63, w = d >> 6;
    const uint64_t s0 = a.seg_off[s], s1 = a.seg_off[s + 1];
    // an empty top bucket has no ranges in pass B, so no totals were written for it
    const uint32_t size = s1 > s0 ? a.segtot[size_t(s) * 256 + d] : 0u;
    uint32_t incl = size;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t k = 0; k < w; ++k) pre += wsum[k];
    const uint32_t start = uint32_t(s0) + pre + incl - size;
    const uint32_t b = s * 256 + d;
    a.bstart[b] = start;
    a.bsize[b] = size;
    if (size > a.cap) {
        const uint32_t i = atomicAdd(a.nlarge, 1u);
        if (i < a.lcap) {  // lcap >= n / (cap + 1): always true
            a.lstart[i] = start;
            a.lsize[i] = size;
        }
    }
}

// ranges of the low passes over the large buckets (one workgroup)
template <int = 0>
__global__ __launch_bounds__(1024) void hyb_large_ranges_kernel(HybArgs a)
{
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t nl = min(*a.nlarge, a.lcap);
    const uint32_t per = (nl + 1023) / 1024;
    const uint32_t c0 = min(nl, t * per), c1 = min(nl, c0 + per);
    uint32_t cnt = 0;
    for (uint32_t i = c0; i < c1; ++i) cnt += uint32_t((a.lsize[i] + a.rk_d - 1) / a.rk_d);
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t x = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += x;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t r = incl - cnt;
    for (uint32_t k = 0; k < w; ++k) r += wsum[k];
    if (t == 1023) *a.td.nr = min(a.td.cap, r + cnt);
    for (uint32_t i = c0; i < c1; ++i) {
        const uint64_t st = a.lstart[i], sz = a.lsize[i];
        for (uint64_t x = st; x < st + sz; x += a.rk_d, ++r) {
            if (r >= a.td.cap) return;  // cannot happen: cap = ceil(n / rk) + lcap
            a.td.begin[r] = x;
            a.td.end[r] = min(st + sz, x + a.rk_d);
            a.td.sbase[r] = st;
            a.td.bseg[r] = i;
            a.td.head[r] = x == st ? 1u : 0u;
        }
    }
}

// Shape of the per-bucket LDS sort: 512 threads, up to ITEMS keys (and
// payloads) each.  4-byte keys only: compiled for 8 waves per SIMD (its
// SGPRs alone would cap it at 7).
#ifndef XSS_HYB_ITEMS_W8
#define XSS_HYB_ITEMS_W8 12  // items per thread for 5..8-byte items (build-time knob)
#endif
// SMALL (items wider than 4 bytes): 6 items per thread (5 for items wider
// than 8 bytes), a 3072- (2560-) key bucket capacity, for arrays whose
// buckets average at most that capacity less HYB_SMALL_MARGIN keys; the
// smaller LDS footprint fits more workgroups per CU (64-bit keys at 2^27:
// 3.52 vs 4.21 ms per sort).
// MID (items of at most 8 bytes): 9 items per thread, a 4608-key capacity
// for buckets averaging at most that less HYB_SMALL_MARGIN (argsort of
// 32-bit keys at 2^28: 4.61 vs 4.83 ms).  SIZE: 0 = the default shape,
// 1 = MID, 2 = SMALL.
template <class KD, int PL, int SIZE = 0>
struct HybShape {
    static constexpr int BYTES = ScatterClass<KD, PL>::bytes;
    static constexpr int ITEMS = SIZE == 2   ? (BYTES > 8 ? 5 : 6)
                               : SIZE == 1   ? 9
                               : BYTES > 8   ? 8
                               : BYTES > 4   ? XSS_HYB_ITEMS_W8
                                             : HYB_ITEMS;
    static constexpr uint32_t CAP = uint32_t(HYB_BLOCK) * ITEMS;
    static constexpr int WAVES = BYTES == 4 ? XSS_HYB_WAVES : 1;
    static constexpr bool HAS_SMALL = BYTES > 4;
    static constexpr bool HAS_MID = BYTES <= 8;
};
constexpr uint32_t HYB_LARGE_GRID = 1024;  // workgroups of each large-bucket pass

// Tile shape of the segmented passes of this sort: the batch form's
// (SegShape), except 4-byte keys, which take 1024 x 16 tiles as the
// one-array pass does.  Pass B ranks without a peer match, and the longer
// runs per digit then pay: 0.68 vs 0.75 ms per pass at 2^28 f32 (8K-key
// tiles were the faster shape while every pass ranked stably).
// 5..8-byte items (64-bit keys, 32-bit argsort with its index) take
// 1024 x 12 tiles, as the one-array pass does: argsort f32 2^28 4.60 ->
// 4.46 ms, qsort u64 2^27 3.60 -> 3.57 ms (the batch form's 512 x 16
// before).  Build-time knobs XSS_HYB_SEG_BLOCK_C2 / XSS_HYB_SEG_ITEMS_C2.
#ifndef XSS_HYB_SEG_BLOCK_C1
#define XSS_HYB_SEG_BLOCK_C1 1024
#endif
#ifndef XSS_HYB_SEG_BLOCK_C2
#define XSS_HYB_SEG_BLOCK_C2 1024
#endif
#ifndef XSS_HYB_SEG_ITEMS_C2
#define XSS_HYB_SEG_ITEMS_C2 12
#endif
template <class KD, int PL>
struct HybSegShape {
    static constexpr int BYTES = SegShape<KD, PL>::BYTES;
    static constexpr int BLOCK = BYTES == 4   ? XSS_HYB_SEG_BLOCK_C1
                               : BYTES <= 8 ? XSS_HYB_SEG_BLOCK_C2
                                            : SegShape<KD, PL>::BLOCK;
    static constexpr int ITEMS = (BYTES > 4 && BYTES <= 8) ? XSS_HYB_SEG_ITEMS_C2 : SegShape<KD, PL>::ITEMS;
};
// average bucket size below the SMALL capacity (about 8 standard deviations
// of a uniform bucket's size at ~2000 keys)
constexpr uint32_t HYB_SMALL_MARGIN = 384;

// One workgroup per bucket of at most CAP keys: LSD radix by the low P - 2
// digits in LDS (the top two are the bucket's), written back in place; the
// payload (values, or argsort's 32-bit indices) moves with its key.
// PL_IDX: the sorted indices go to idx_out (64-bit) at the bucket's place.
template <class KD, int PL, bool NL, int SIZE>
__global__ __launch_bounds__(HYB_BLOCK) __attribute__((amdgpu_waves_per_eu(HybShape<KD, PL, SIZE>::WAVES)))
void hyb_local_kernel(const uint32_t *bstart, const uint32_t *bsize, void *keys_v, void *vals_v, uint64_t *idx_out,
                      uint64_t flip64, uint32_t canon)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_PL = PL != PL_NONE;
    constexpr int P = KD::PASSES;
    constexpr int WAVES = HYB_BLOCK / 64;
    constexpr int ITEMS = HybShape<KD, PL, SIZE>::ITEMS;
    constexpr uint32_t CAP = HybShape<KD, PL, SIZE>::CAP;
    __shared__ U s_keys[CAP];
    __shared__ V s_vals[HAS_PL ? CAP : 1];
    __shared__ uint32_t s_off[WAVES * 256];
    __shared__ uint32_t s_wsum[WAVES];
    const uint32_t b = blockIdx.x;
    const uint32_t m = bsize[b];
    if (m > CAP) return;  // a large bucket: the low segmented passes sort it
    const uint32_t st = bstart[b];
    U *keys = static_cast<U *>(keys_v) + st;
    V *vals = HAS_PL ? static_cast<V *>(vals_v) + st : nullptr;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto finish = [&](U x) {
        if constexpr (KD::KIND == Kind::Float)
            if (canon && KD::is_nan(x)) x = KD::QNAN;
        return x;
    };
    auto put = [&](uint32_t idx, U x, V v) {  // final key (and payload) at bucket slot idx
        if constexpr (PL == PL_IDX) {
            idx_out[st + idx] = uint64_t(v);
        } else {
            keys[idx] = finish(x);
            if constexpr (HAS_PL) vals[idx] = v;
        }
    };
    if (m < 2) {
        if (m == 1 && tid == 0) put(0, keys[0], HAS_PL ? vals[0] : V(0));
        return;
    }
    const U flip = U(flip64);
    // the bucket is spread over all waves: `it` items per thread (uniform),
    // wave w holding keys [w * 64 * it, (w + 1) * 64 * it)
    const uint32_t it = (m + HYB_BLOCK - 1) / HYB_BLOCK;
    const uint32_t lbase = wave * 64 * it + lane;
    U key[ITEMS];
    V val[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t idx = lbase + i * 64;
        const bool ok = uint32_t(i) < it && idx < m;
        key[i] = ok ? keys[idx] : U(0);
        if constexpr (HAS_PL) val[i] = ok ? vals[idx] : V(0);
    }
    // one LSD pass over the digit at `shift`; stable unless `unst` (the
    // first pass, whose input order does not matter; never for argsort)
    auto pass = [&](const int shift, const bool unst) {
        for (uint32_t i = tid; i < WAVES * 256; i += HYB_BLOCK) s_off[i] = 0;
        __syncthreads();  // counters cleared; the previous pass's reads of s_keys are done
        // rank: stable within the bucket (wave-striped order), per-wave
        // counters; the lanes holding one digit reserve their slots with one
        // LDS atomic by the lowest of them, so successive items do not wait
        // on each other's counter update.  Padding takes digit 255 and comes
        // after every key.
        uint32_t rd[ITEMS];
        uint32_t *wc = s_off + wave * 256;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                const uint32_t idx = lbase + i * 64;
                const uint32_t d = idx < m ? KD::digit(key[i], flip, NL, shift) : 255u;
                if (PL != PL_IDX && unst) {
                    // each key reserves its own slot with one returning LDS
                    // add (a wave of one digit: 64 slots by lane 0, bit 24);
                    // padding takes no slot (bit 25), so the keys fill
                    // [0, m) and the next pass sees padding at idx >= m
                    const bool real = idx < m;
                    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
                    const bool uni = __ballot(!real || d != d0) == 0;
                    uint32_t old = 0;
                    if (real && (!uni || lane == 0)) old = atomicAdd(&wc[d], uni ? 64u : 1u);
                    rd[i] = old | (d << 16) | (uni ? (1u << 24) : 0u) | (real ? 0u : (1u << 25));
                    continue;
                }
                const uint64_t peers = match_any8(d);
                const uint32_t below = lanes_below(peers);
                uint32_t old = 0;
                if (below == 0) old = atomicAdd(&wc[d], uint32_t(__popcll(peers)));
                old = __shfl(old, int(__builtin_ctzll(peers)));
                rd[i] = (old + below) | (d << 16);
            }
        }
        __syncthreads();
        // digit counts, their exclusive scan, and the slot base of every (wave, digit)
        uint32_t tcount = 0;
        if (tid < 256) {
#pragma unroll
            for (int w = 0; w < WAVES; ++w) tcount += s_off[w * 256 + tid];
        }
        uint32_t incl = tcount;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off);
            if (lane >= uint32_t(off)) incl += t;
        }
        if (lane == 63) s_wsum[wave] = incl;
        __syncthreads();
        if (tid < 256) {
            uint32_t sw = incl - tcount;
            for (uint32_t w = 0; w < wave; ++w) sw += s_wsum[w];
#pragma unroll
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t c = s_off[w * 256 + tid];
                s_off[w * 256 + tid] = sw;
                sw += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                uint32_t r = rd[i];
                if (PL != PL_IDX && unst) {
                    if (r >> 25) continue;
                    if (r >> 24) r = (__builtin_amdgcn_readfirstlane(r) + lane) & 0xffffffu;
                }
                const uint32_t d = r >> 16;
                const uint32_t slot = wc[d] + (r & 0xffffu);
                s_keys[slot] = key[i];
                if constexpr (HAS_PL) s_vals[slot] = val[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                key[i] = s_keys[lbase + i * 64];
                if constexpr (HAS_PL) val[i] = s_vals[lbase + i * 64];
            }
        }
    };
    if constexpr (P == 4) {
        // 4-byte keys: both passes inline (measured 8% faster than the loop)
        pass(0, true);
        pass(8, false);
    } else {
#pragma unroll 1
        for (int p = 0; p < P - 2; ++p) pass(8 * p, p == 0);
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t idx = lbase + i * 64;
        if (uint32_t(i) < it && idx < m) put(idx, key[i], HAS_PL ? val[i] : V(0));
    }
}

template <class KD, int PL>
struct HybPlan {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr bool HAS_PL = PL != PL_NONE;
    static constexpr uint32_t CAP = HybShape<KD, PL>::CAP;
    // the shape (and bucket capacity) of the local sort for n keys: the
    // smallest whose capacity exceeds the average bucket by HYB_SMALL_MARGIN
    static int shape(size_t n)
    {
        const size_t need = n / HYB_BUCKETS + HYB_SMALL_MARGIN;
        if (HybShape<KD, PL>::HAS_SMALL && need <= HybShape<KD, PL, 2>::CAP) return 2;
        if (HybShape<KD, PL>::HAS_MID && need <= HybShape<KD, PL, 1>::CAP) return 1;
        return 0;
    }
    static uint32_t cap(size_t n)
    {
        const int sh = shape(n);
        return sh == 2 ? HybShape<KD, PL, 2>::CAP : sh == 1 ? HybShape<KD, PL, 1>::CAP : CAP;
    }
    // tile shape of the segmented passes (16K-key tiles for 4-byte keys,
    // see HybSegShape)
    using SS = HybSegShape<KD, PL>;
    static constexpr uint64_t SEG_TILE = uint64_t(SS::BLOCK) * SS::ITEMS;
    // pass B's table must fit the workspace's [256][MAX_RANGES] range tables
    static uint64_t rk_b(size_t n)
    {
        const uint64_t per = uint64_t(MAX_RANGES - 256) * SEG_TILE;
        return SEG_TILE * std::max<uint64_t>(1, (n + per - 1) / per);
    }
    static uint32_t cap_b(size_t n) { return uint32_t((n + rk_b(n) - 1) / rk_b(n) + 256); }
    static uint64_t rk_d(size_t) { return 8 * SEG_TILE; }
    static uint32_t lcap(size_t n) { return uint32_t(std::min<size_t>(HYB_BUCKETS, n / (cap(n) + 1) + 1)); }
    static uint32_t cap_d(size_t n) { return uint32_t((n + rk_d(n) - 1) / rk_d(n) + lcap(n)); }
    static size_t table_bytes(uint32_t cap) { return 3 * align256(size_t(cap) * 8) + 2 * align256(size_t(cap) * 4); }
    // key / payload buffers: keys-only and key-value sorts ping-pong with one
    // scratch pair; argsort (read-only keys) runs between two scratch pairs
    static size_t data_bytes(size_t n)
    {
        const size_t pair = align256(n * sizeof(U)) + (HAS_PL ? align256(n * sizeof(V)) : 0);
        return PL == PL_IDX ? 2 * pair : pair;
    }
    static size_t scratch_bytes(size_t n)
    {
        return data_bytes(n) + align256(256 * 256 * 4) + align256(257 * 8) + 2 * align256(HYB_BUCKETS * 4) +
               align256(16) + table_bytes(cap_b(n)) + 2 * align256(size_t(lcap(n)) * 4) + table_bytes(cap_d(n)) +
               2 * align256(size_t(256) * cap_d(n) * 4) + align256(size_t(lcap(n)) * 256 * 4);
    }
};

// MSD-first sort.  PL_NONE / PL_V32 / PL_V64: keys (and vals) sorted in
// place; PL_IDX: keys read-only, the sorted order's indices to idx_out.
template <class KD, int PL>
int enqueue_hybrid(Workspace *ws, char *scratch, void *keys, void *vals, uint64_t *idx_out, size_t n, bool hasnan,
                   bool desc, bool canon, hipStream_t s)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    using HP = HybPlan<KD, PL>;
    using SS = typename HP::SS;
    constexpr bool HAS_PL = PL != PL_NONE;
    constexpr int P = KD::PASSES;
    static_assert(P >= 4, "two top digits and at least two low digits");
    const uint64_t flip = desc ? ~0ull : 0ull;
    const bool nan_last = (KD::KIND == Kind::Float) && hasnan;
    const bool canon_nan = canon && nan_last && PL == PL_NONE;
    const uint32_t nn = uint32_t(n);
    const uint32_t cap_b = HP::cap_b(n), cap_d = HP::cap_d(n), lcap = HP::lcap(n);

    // scratch layout (HybPlan::scratch_bytes)
    char *p = scratch;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align256(bytes);
        return q;
    };
    // X: where the second-digit pass writes and the buckets are sorted; Y: the other buffer
    void *xk = keys, *xv = vals;
    void *yk = take(n * sizeof(U));
    void *yv = HAS_PL ? static_cast<void *>(take(n * sizeof(V))) : nullptr;
    if (PL == PL_IDX) {
        xk = take(n * sizeof(U));
        xv = take(n * sizeof(V));
    }
    uint32_t *segtot = reinterpret_cast<uint32_t *>(take(256 * 256 * 4));
    HybArgs h{};
    h.seg_off = reinterpret_cast<uint64_t *>(take(257 * 8));
    h.bstart = reinterpret_cast<uint32_t *>(take(HYB_BUCKETS * 4));
    h.bsize = reinterpret_cast<uint32_t *>(take(HYB_BUCKETS * 4));
    uint32_t *ctr = reinterpret_cast<uint32_t *>(take(16));
    auto table = [&](HybTable &t, uint32_t cap, uint32_t *nr) {
        t.begin = reinterpret_cast<uint64_t *>(take(size_t(cap) * 8));
        t.end = reinterpret_cast<uint64_t *>(take(size_t(cap) * 8));
        t.sbase = reinterpret_cast<uint64_t *>(take(size_t(cap) * 8));
        t.bseg = reinterpret_cast<uint32_t *>(take(size_t(cap) * 4));
        t.head = reinterpret_cast<uint32_t *>(take(size_t(cap) * 4));
        t.nr = nr;
        t.cap = cap;
    };
    table(h.tb, cap_b, ctr + 1);
    h.lstart = reinterpret_cast<uint32_t *>(take(size_t(lcap) * 4));
    h.lsize = reinterpret_cast<uint32_t *>(take(size_t(lcap) * 4));
    table(h.td, cap_d, ctr + 2);
    uint32_t *rsum_d = reinterpret_cast<uint32_t *>(take(size_t(256) * cap_d * 4));
    uint32_t *rpre_d = reinterpret_cast<uint32_t *>(take(size_t(256) * cap_d * 4));
    uint32_t *segtot_d = reinterpret_cast<uint32_t *>(take(size_t(lcap) * 256 * 4));
    h.top_total = ws->total();
    h.segtot = segtot;
    h.nlarge = ctr;
    h.lcap = lcap;
    h.cap = HP::cap(n);
    h.rk_b = HP::rk_b(n);
    h.rk_d = HP::rk_d(n);

    // pass A: top digit over the whole array (keys -> Y; argsort: indices generated)
    {
        const int rc = enqueue_lsd_pass<KD, PL>(ws, keys, vals, yk, yv, nullptr, n, 8 * (P - 1), flip, nan_last,
                                                true, false, false, s, nullptr, 0, false, nullptr, true);
        if (rc) return rc;
    }
    {
        KernelTimer t("hyb_tables", s);
        hipLaunchKernelGGL(hyb_top_ranges_kernel<0>, dim3(1), dim3(256), 0, s, h);
    }
    XSS_HIP_CHECK(hipGetLastError());
    // one segmented pass over a device-built table (src -> dst)
    auto seg_pass = [&](const char *const names[3], const HybTable &t, uint32_t *rsum, uint32_t *rpre,
                        uint32_t *stot, int shift, const void *sk, const void *sv, void *dk, void *dv,
                        bool last, bool unstable, uint32_t max_grid) -> int {
        // the kernels walk the table's ranges with a bounded grid: a table
        // that turns out (nearly) empty costs few workgroups
        const uint32_t grid = std::min<uint32_t>(t.cap, max_grid);
        UpArgs u{};
        u.keys = sk; u.n = nn; u.shift = uint32_t(shift); u.nan_last = nan_last; u.flip = flip;
        u.rsum = rsum; u.ranges = t.cap; u.range_begin = t.begin; u.range_end = t.end; u.nr_dev = t.nr;
        {
            KernelTimer kt(names[0], s);
            hipLaunchKernelGGL((upsweep_kernel<KD>), dim3(grid), dim3(UP_BLOCK), 0, s, u);
        }
        XSS_HIP_CHECK(hipGetLastError());
        SegScanArgs sc{rsum, rpre, stot, t.head, t.bseg, 0, t.cap, t.nr};
        {
            KernelTimer kt(names[1], s);
            hipLaunchKernelGGL(seg_scan_kernel<0>, dim3(256), dim3(256), 0, s, sc);
        }
        XSS_HIP_CHECK(hipGetLastError());
        PassArgs a{};
        a.keys_in = sk; a.vals_in = sv; a.keys_out = dk; a.vals_out = dv; a.idx_out = idx_out; a.n = nn;
        a.shift = uint32_t(shift); a.flip = flip; a.nan_last = nan_last; a.first = false; a.last = last;
        a.canon_nan = last && canon_nan;
        SegRanges r{t.begin, t.end, t.sbase, t.bseg, rpre, stot, 0, t.cap, t.nr};
        KernelTimer kt(names[2], s);
        // pass B's input order does not matter (it is an MSD pass inside
        // each top bucket), so it ranks without keeping it; argsort keeps
        // every pass stable (ties in input order, the identity for sorted
        // input)
        constexpr bool UN = PL != PL_IDX;
        if (unstable && nan_last)
            hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, SS::BLOCK, SS::ITEMS, true, UN>), dim3(grid),
                               dim3(SS::BLOCK), 0, s, a, r);
        else if (unstable)
            hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, SS::BLOCK, SS::ITEMS, false, UN>), dim3(grid),
                               dim3(SS::BLOCK), 0, s, a, r);
        else if (nan_last)
            hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, SS::BLOCK, SS::ITEMS, true>), dim3(grid),
                               dim3(SS::BLOCK), 0, s, a, r);
        else
            hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, SS::BLOCK, SS::ITEMS, false>), dim3(grid),
                               dim3(SS::BLOCK), 0, s, a, r);
        XSS_HIP_CHECK(hipGetLastError());
        return XSS_OK;
    };
    // pass B: second digit inside every top bucket (Y -> X)
    static const char *const msd_names[3] = {"hyb_msd_upsweep", "hyb_msd_scan", "hyb_msd_downsweep"};
    static const char *const low_names[3] = {"hyb_large_upsweep", "hyb_large_scan", "hyb_large_downsweep"};
    int rc = seg_pass(msd_names, h.tb, ws->rsum(), ws->rpre(), segtot, 8 * (P - 2), yk, yv, xk, xv, false, true,
                      MAX_RANGES);
    if (rc) return rc;
    {
        KernelTimer t("hyb_tables", s);
        hipLaunchKernelGGL(hyb_buckets_kernel<0>, dim3(256), dim3(256), 0, s, h);
    }
    XSS_HIP_CHECK(hipGetLastError());
    {
        KernelTimer t("hyb_local", s);
        auto launch = [&](auto size_tag) {
            constexpr int SZ = decltype(size_tag)::value;
            if (nan_last)
                hipLaunchKernelGGL((hyb_local_kernel<KD, PL, true, SZ>), dim3(HYB_BUCKETS), dim3(HYB_BLOCK), 0, s,
                                   h.bstart, h.bsize, xk, xv, idx_out, flip, uint32_t(canon_nan));
            else
                hipLaunchKernelGGL((hyb_local_kernel<KD, PL, false, SZ>), dim3(HYB_BUCKETS), dim3(HYB_BLOCK), 0, s,
                                   h.bstart, h.bsize, xk, xv, idx_out, flip, 0u);
        };
        const int sh = HP::shape(n);
        if constexpr (HybShape<KD, PL>::HAS_SMALL) {
            if (sh == 2) launch(std::integral_constant<int, 2>{});
        }
        if constexpr (HybShape<KD, PL>::HAS_MID) {
            if (sh == 1) launch(std::integral_constant<int, 1>{});
        }
        if (sh == 0) launch(std::integral_constant<int, 0>{});
    }
    XSS_HIP_CHECK(hipGetLastError());
    {
        KernelTimer t("hyb_tables", s);
        hipLaunchKernelGGL(hyb_large_ranges_kernel<0>, dim3(1), dim3(1024), 0, s, h);
    }
    XSS_HIP_CHECK(hipGetLastError());
    // low digits of the large buckets (X -> Y -> X ...; P - 2 is even; the
    // last pass of argsort writes the indices).  Their table is sized for
    // the worst case (every key in a large bucket) but is usually empty
    // (spread data), so the passes take HYB_LARGE_GRID persistent
    // workgroups: 4 per CU still fill the GPU when the table is full, and
    // an empty table costs a quarter of a 4096-workgroup launch.
    void *sk = xk, *sv = xv, *dk = yk, *dv = yv;
    for (int q = 0; q < P - 2; ++q) {
        rc = seg_pass(low_names, h.td, rsum_d, rpre_d, segtot_d, 8 * q, sk, sv, dk, dv, q == P - 3, q == 0,
                      HYB_LARGE_GRID);
        if (rc) return rc;
        std::swap(sk, dk);
        std::swap(sv, dv);
    }
    return XSS_OK;
}

}  // namespace xss

namespace xss {

// arrays of at least this many keys take the MSD-first path, and whether the
// sampled spread check runs first (xss_gpu_set_tuning)
std::atomic<uint64_t> &hybrid_min_n();
std::atomic<uint64_t> &hybrid_check();
std::atomic<uint64_t> &count_min_n();  // 16-bit qsort: counting sort from this size

constexpr uint32_t HYB_SAMPLES = 4096;

// top 16 bits of the order keys of HYB_SAMPLES evenly spaced keys
template <class KD>
__global__ __launch_bounds__(256) void hyb_sample_kernel(const void *keys_v, uint64_t n, uint64_t flip64,
                                                         uint32_t nan_last, uint16_t *out)
{
    using U = typename KD::U;
    const U *keys = static_cast<const U *>(keys_v);
    for (uint32_t i = threadIdx.x; i < HYB_SAMPLES; i += 256) {
        const uint64_t pos = (uint64_t(i) * n) / HYB_SAMPLES;
        const U o = KD::ord(keys[pos], U(flip64), nan_last != 0);
        out[i] = uint16_t(o >> (8 * sizeof(U) - 16));
    }
}

// Whether the top 16 bits spread the keys over buckets small enough for
// the LDS sort: in a sample of HYB_SAMPLES keys, samples that share their
// top 16 bits with another sample point at buckets of >= ~2n/HYB_SAMPLES
// keys (far above an LDS bucket at these sizes).  Skewed or
// duplicate-heavy data -- most samples in such buckets -- sorts faster on
// the LSD path.  Costs one small kernel and one wait for the stream; a
// stream being captured into a graph cannot wait, and takes the MSD-first
// path unconditionally.
template <class KD>
bool hybrid_spread_ok(Workspace *ws, const void *keys, size_t n, bool nan_last, uint64_t flip, hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return true;
    static thread_local uint16_t *host = nullptr;
    if (!host && hipHostMalloc(reinterpret_cast<void **>(&host), HYB_SAMPLES * 2) != hipSuccess) {
        host = nullptr;
        return true;
    }
    uint16_t *dev = reinterpret_cast<uint16_t *>(ws->rpre());  // free until the sort's first scan
    hipLaunchKernelGGL(hyb_sample_kernel<KD>, dim3(1), dim3(256), 0, s, keys, uint64_t(n), flip,
                       uint32_t(nan_last), dev);
    if (hipGetLastError() != hipSuccess) return true;
    if (hipMemcpyAsync(host, dev, HYB_SAMPLES * 2, hipMemcpyDeviceToHost, s) != hipSuccess) return true;
    if (hipStreamSynchronize(s) != hipSuccess) return true;
    static thread_local std::vector<uint8_t> seen(65536, 0);
    uint32_t shared = 0;
    for (uint32_t i = 0; i < HYB_SAMPLES; ++i) {
        uint8_t &c = seen[host[i]];
        shared += c == 1 ? 2u : c > 1 ? 1u : 0u;  // the first repeat counts both samples
        c = uint8_t(c < 2 ? c + 1 : 2);
    }
    for (uint32_t i = 0; i < HYB_SAMPLES; ++i) seen[host[i]] = 0;
    return shared * 5 <= HYB_SAMPLES;  // at most 20% of the samples in shared buckets
}

// qsort / keyvalue_qsort / argsort on the device: MSD-first for large arrays
// of 4- and 8-byte keys, LSD otherwise (same arguments as run_sort)
template <class KD, int PL>
int run_sort_auto(const void *keys_in, void *keys_out, const void *vals_in, void *vals_out, uint64_t *idx_out,
                  size_t n, bool hasnan, bool desc, bool canon, hipStream_t s)
{
    if constexpr (sizeof(typename KD::U) == 2 && PL == PL_NONE) {
        // 16-bit keys only: a counting sort (one read, one write).  Every
        // NaN leaves it as the canonical NaN, so it serves qsort (canon).
        if (keys_in && keys_in == keys_out && canon && (reinterpret_cast<uintptr_t>(keys_out) & 1u) == 0 &&
            n >= std::max<uint64_t>(count_min_n().load(), 2) && n <= MAX_N) {
            const uint32_t rows = cs_rows();
            Workspace *ws = nullptr;
            std::unique_lock<std::mutex> lock;
            int rc = acquire_workspace(s, cs_scratch_bytes(rows), &ws, &lock);
            if (rc) return rc;
            return enqueue_count_sort<KD>(ws->data, keys_out, n, hasnan, desc, canon, rows, s);
        }
    }
    if constexpr (sizeof(typename KD::U) >= 4) {
        const bool in_place = PL == PL_IDX || (keys_in == keys_out && vals_in == vals_out);
        const bool args_ok = keys_in && (PL == PL_IDX ? idx_out != nullptr : keys_out != nullptr) &&
                             (PL == PL_NONE || PL == PL_IDX || (vals_in && vals_out));
        if (in_place && args_ok && n >= std::max<uint64_t>(hybrid_min_n().load(), 2) && n <= MAX_N) {
            Workspace *ws = nullptr;
            std::unique_lock<std::mutex> lock;
            int rc = acquire_workspace(s, HybPlan<KD, PL>::scratch_bytes(n), &ws, &lock);
            if (rc) return rc;
            const bool nan_last = (KD::KIND == Kind::Float) && hasnan;
            const uint64_t flip = desc ? ~0ull : 0ull;
            if (hybrid_check() == 0 || hybrid_spread_ok<KD>(ws, keys_in, n, nan_last, flip, s))
                return enqueue_hybrid<KD, PL>(ws, ws->data, const_cast<void *>(keys_in),
                                              const_cast<void *>(vals_in), idx_out, n, hasnan, desc, canon, s);
            // skewed data: the LSD path (its scratch fits in the workspace just grown)
            return enqueue_sort<KD, PL>(ws, ws->data, keys_in, keys_out, vals_in, vals_out, idx_out, n, hasnan,
                                        desc, canon, s);
        }
    }
    return run_sort<KD, PL>(keys_in, keys_out, vals_in, vals_out, idx_out, n, hasnan, desc, canon, s);
}

}  // namespace xss
