// xss_sample.hpp -- segmented sample sort of 4-byte keys: the batch form's
// segments too long for one workgroup.
//
// A radix MSD split of such a segment depends on how its keys spread over
// the order-key space: uniform floats in [-1, 1], for instance, put half of
// a segment into two top-digit buckets (their exponents), and those buckets
// would need further passes.  Splitters taken from a sample of the segment
// cut it into buckets of about equal size whatever the distribution, so
// every key is moved once through HBM and then sorted on chip:
//
//   ss_sample   per segment: S stratified samples (S = a power of two, about
//               `oversample` per bucket; in groups of 16 consecutive keys,
//               one line each), bitonic-sorted in LDS by order key; nb - 1
//               evenly spaced ones are the splitters
//   ss_hist     per range (ranges never cross a segment): the bucket of each
//               key (binary search over the segment's splitters in LDS),
//               counted into 256 bins -- the digit histogram of a pass
//   seg_scan    (xss_kernels.hpp) per-segment bucket totals and range bases
//   ss_scatter  the scatter of a segmented pass with the bucket as the digit
//               (keys -> scratch; argsort generates its indices here)
//   ss_buckets  bucket starts and sizes; buckets too large for the LDS sort
//               are listed (select: only the bucket holding rank len / 2,
//               whose slot is kept per segment)
//   ss_local    one workgroup per bucket sorts it by its whole order key in
//               LDS: the bits where its keys differ (AND / OR of the bucket),
//               a 12-bit pass whose order inside a digit is free, then
//               stable 4-bit passes; argsort: stable 4-bit passes only
//   ss_big      listed buckets: one workgroup runs 8-bit LSD passes over
//               the bucket in HBM (digits its keys share are skipped)
//   select      the histogram read also copies the keys to scratch;
//               ss_part3 writes them back in three regions (below, inside,
//               above the bucket holding rank len / 2) and only that bucket
//               is sorted, in place
//
// Moved bytes per key: a histogram read, one scatter (read + write) and the
// LDS sort's read + write: 20 bytes for 4-byte keys, against 36 for four LSD
// passes; select: 16 (read, copy, partition read + write).  Everything is planned on the host (segment lengths) or on the
// device (bucket sizes), with no host wait: a plan's run is capturable.
#pragma once
#include "xss_sort.hpp"

namespace xss {

constexpr int SS_SPLIT = 256;         // buckets per segment at most (one scatter digit)
constexpr int SS_SAMPLE_BLOCK = 1024;
constexpr uint32_t SS_SAMPLE_MAX = 4096;
constexpr int SS_LOCAL_BLOCK = 512;
constexpr int SS_BIG_BLOCK = 1024;
constexpr int SS_BIG_ITEMS = 8;
// the scatter's wide tile: 1024 threads x 16 items of 4 bytes, 12 of 5..8
// bytes (a 4-byte key with argsort's index, an 8-byte key), 8 of 12 bytes
constexpr int ss_wide_items(int item_bytes) { return item_bytes <= 4 ? 16 : item_bytes <= 8 ? 12 : 8; }
template <class KD, int PL>
struct SsWideShape {
    static constexpr int BLOCK = 1024;
    static constexpr int ITEMS = ss_wide_items(int(sizeof(typename KD::U)) + (PL == PL_IDX ? 4 : 0));
};
// keys per thread of the smaller LDS bucket sort of 8-byte keys (capacity
// 4096; the default is 12 items, argsort 11)
constexpr int SS_LOCAL_ITEMS_W8 = 8;

// a wave-uniform value moved to scalar registers (64-bit values in halves)
template <class U>
__device__ __forceinline__ U uniform(U x)
{
    if constexpr (sizeof(U) == 8) {
        const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
        const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(uint64_t(x) >> 32));
        return U((uint64_t(hi) << 32) | lo);
    } else {
        return U(__builtin_amdgcn_readfirstlane(uint32_t(x)));
    }
}

struct SsArgs {
    // host-built plan
    const uint64_t *seg_off;  // [nseg + 1]
    const uint32_t *lseg;     // [nlarge] segment id of each large segment
    const uint32_t *lnb;      // [nlarge] its buckets (2..256)
    const uint32_t *lbb;      // [nlarge] its first bucket slot
    const uint32_t *bk_seg;   // [nbk] large segment of each bucket slot
    uint32_t oversample;      // samples per bucket
    // built on the device
    void *split;              // [nlarge][256] splitters (order keys; 255 used)
    uint32_t *bstart;         // [nbk] bucket starts (positions in the batch)
    uint32_t *bsize;          // [nbk]
    void *brange;             // [nbk][2] of the key type: lowest order key and
                              // key range of the bucket from its splitters
                              // ([1] = 0: unknown)
    uint32_t *ctl;            // [0]: buckets listed in `big`
    uint32_t *big;            // [nbk] slots of the buckets larger than lcap
    uint32_t *ksel;           // [nlarge] select: the slot of the bucket holding rank len / 2
    uint32_t lcap;            // largest bucket ss_local sorts
    uint32_t select;          // list only the buckets that hold rank len / 2
    uint64_t flip;
    uint32_t nan_last;
};

// Bucket of order key o among the splitters sp[0..254] (sorted; the entries
// past the segment's nb - 1 splitters are all-ones): the number of
// splitters <= o, capped at nb - 1.  Two levels: every 16th splitter
// (sup[i] = sp[16 i + 15], wave-uniform, in scalar registers) is compared
// with o in ALU operations, which gives the group of 16; then a 4-step
// binary search inside the group in LDS (half the LDS lookups of a search
// over all 255).
template <class U>
struct SplitSearch {
    U sup[15];
    __device__ __forceinline__ void load(const U *sp)
    {
#pragma unroll
        for (int i = 0; i < 15; ++i) sup[i] = uniform(sp[16 * i + 15]);
    }
    __device__ __forceinline__ uint32_t operator()(U o, const U *sp, uint32_t nbm1) const
    {
        uint32_t g = 0;
#pragma unroll
        for (int i = 0; i < 15; ++i) g += sup[i] <= o ? 1u : 0u;
        uint32_t p = 16 * g;
#pragma unroll
        for (uint32_t b = 8; b; b >>= 1)
            if (sp[p + b - 1] <= o) p += b;
        return min(p, nbm1);
    }
};

template <class KD, bool NL>
struct SplitDigit {
    using U = typename KD::U;
    const U *sp;
    SplitSearch<U> ss;
    U flip;
    uint32_t nbm1;
    __device__ __forceinline__ uint32_t operator()(U x) const { return ss(KD::ord(x, flip, NL), sp, nbm1); }
};

__device__ __forceinline__ uint32_t ss_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// ---------------------------------------------------------------- splitters

// One workgroup per large segment.  Workgroup 0 also clears the list
// counters of this run (read by kernels later on the stream).
template <class KD>
__global__ __launch_bounds__(SS_SAMPLE_BLOCK) void ss_sample_kernel(SsArgs a, const void *keys_v)
{
    using U = typename KD::U;
    __shared__ U s[SS_SAMPLE_MAX];
    const uint32_t j = blockIdx.x, tid = threadIdx.x;
    if (j == 0 && tid < 4) a.ctl[tid] = 0;
    const uint32_t seg = a.lseg[j], nb = a.lnb[j];
    const uint64_t b0 = a.seg_off[seg], len = a.seg_off[seg + 1] - b0;
    uint32_t S = 1;
    while (S < nb * a.oversample && S < SS_SAMPLE_MAX) S <<= 1;
    while (S > len) S >>= 1;
    const U *keys = static_cast<const U *>(keys_v) + b0;
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    // S / G groups of G consecutive keys (one 64-byte line), one group per
    // stratum [g len / (S / G), (g + 1) len / (S / G)) at a hashed offset:
    // single-key samples would fetch a line each
    constexpr uint32_t G = 64 / sizeof(U);
    const uint32_t ng = S >= G ? S / G : 1, gl = S >= G ? G : S;
    for (uint32_t i = tid; i < S; i += SS_SAMPLE_BLOCK) {
        const uint32_t gi = i / gl, e = i % gl;
        const uint64_t lo = uint64_t(gi) * len / ng, hi = uint64_t(gi + 1) * len / ng;  // hi - lo >= gl
        const uint64_t pos = lo + ss_hash(gi * 0x9e3779b9u ^ (j + 1) * 0x85ebca6bu) % (hi - lo - gl + 1) + e;
        s[i] = KD::ord(keys[pos], flip, nl);
    }
    __syncthreads();
    for (uint32_t k = 2; k <= S; k <<= 1) {
        for (uint32_t h = k >> 1; h > 0; h >>= 1) {
            for (uint32_t i = tid; i < S; i += SS_SAMPLE_BLOCK) {
                const uint32_t q = i ^ h;
                if (q > i) {
                    const U x = s[i], y = s[q];
                    if ((x > y) == ((i & k) == 0)) {
                        s[i] = y;
                        s[q] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    U *out = static_cast<U *>(a.split) + size_t(j) * SS_SPLIT;
    if (tid < SS_SPLIT - 1) out[tid] = tid + 1 < nb ? s[size_t(tid + 1) * S / nb] : U(KD::ONES);
}

// ---------------------------------------------------------------- histogram

struct SsRanges {
    const uint64_t *begin;  // [nr]
    const uint64_t *end;
    const uint32_t *lidx;   // [nr] large segment of each range
};

// One workgroup per range: the bucket counts of its keys (rsum[d][range]).
template <class KD>
__global__ __launch_bounds__(256) void ss_hist_kernel(SsArgs a, SsRanges r, const void *keys_v, uint32_t *rsum,
                                                      uint32_t stride, void *copy_v)
{
    using U = typename KD::U;
    __shared__ U sp[SS_SPLIT];
    __shared__ uint32_t h[256];
    const uint32_t rg = blockIdx.x, tid = threadIdx.x;
    const uint32_t j = r.lidx[rg];
    sp[tid] = tid < SS_SPLIT - 1 ? static_cast<const U *>(a.split)[size_t(j) * SS_SPLIT + tid] : U(KD::ONES);
    h[tid] = 0;
    const uint32_t nbm1 = a.lnb[j] - 1;
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    __syncthreads();
    SplitSearch<U> search;
    search.load(sp);
    const U *keys = static_cast<const U *>(keys_v);
    const uint64_t r0 = r.begin[rg], r1 = r.end[rg];
    auto count = [&](U x) {
        const uint32_t d = search(KD::ord(x, flip, nl), sp, nbm1);
        const uint64_t act = __ballot(1);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        if (__ballot(d != d0) == 0) {
            if (lanes_below(act) == 0) atomicAdd(&h[d0], uint32_t(__popcll(act)));
        } else {
            atomicAdd(&h[d], 1u);
        }
    };
    constexpr int VEC = 16 / sizeof(U);
    const uint64_t mis = (reinterpret_cast<uintptr_t>(keys + r0) & 15u) / sizeof(U);
    const uint64_t head = min(r1 - r0, uint64_t((VEC - mis) % VEC));
    const uint64_t v0 = r0 + head;
    const uint64_t nvec = (r1 - v0) / VEC;
    const uint4 *kv = reinterpret_cast<const uint4 *>(keys + v0);
    // copy_v (select): the keys are also copied to the same positions there
    U *copy = static_cast<U *>(copy_v);
    uint4 *cv = copy ? reinterpret_cast<uint4 *>(copy + v0) : nullptr;
    constexpr uint32_t UNR = 2;
    uint64_t v = tid;
    for (; v + (UNR - 1) * 256 < nvec; v += UNR * 256) {
        uint4 q[UNR];
#pragma unroll
        for (uint32_t u = 0; u < UNR; ++u) q[u] = kv[v + u * 256];
#pragma unroll
        for (uint32_t u = 0; u < UNR; ++u) {
            if (cv) cv[v + u * 256] = q[u];
            const U *e = reinterpret_cast<const U *>(&q[u]);
#pragma unroll
            for (int k = 0; k < VEC; ++k) count(e[k]);
        }
    }
    for (; v < nvec; v += 256) {
        const uint4 q = kv[v];
        if (cv) cv[v] = q;
        const U *e = reinterpret_cast<const U *>(&q);
#pragma unroll
        for (int k = 0; k < VEC; ++k) count(e[k]);
    }
    const uint64_t t0 = v0 + nvec * VEC;
    const uint64_t rest = head + (r1 - t0);
    for (uint64_t i = tid; i < rest; i += 256) {
        const uint64_t p = i < head ? r0 + i : t0 + (i - head);
        const U x = keys[p];
        if (copy) copy[p] = x;
        count(x);
    }
    __syncthreads();
    rsum[size_t(tid) * stride + rg] = h[tid];
}

// ---------------------------------------------------------------- scatter

// One workgroup per range: the tiles of the range scattered to their
// buckets (positions from seg_scan: segment start + bucket start + the
// range's base inside the bucket).  UNST: ranked without keeping the order
// inside a bucket (keys-only sorts: the LDS sort orders every bucket).
template <class KD, int PL, int BLOCK, int ITEMS, bool NL, bool UNST>
__global__ __launch_bounds__(BLOCK) void ss_scatter_kernel(PassArgs p, SegRanges r, SsArgs a)
{
    using U = typename KD::U;
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ DownsweepSmem<KD, PL, BLOCK, ITEMS, true> sm;  // bucket of each slot kept in LDS
    __shared__ U sp[SS_SPLIT];
    const uint32_t tid = threadIdx.x, b = blockIdx.x;
    const uint64_t rb = r.begin[b], re = r.end[b], sbase = r.sbase[b];
    const uint32_t j = r.rbseg[b];
    if (tid < SS_SPLIT)
        sp[tid] = tid < SS_SPLIT - 1 ? static_cast<const U *>(a.split)[size_t(j) * SS_SPLIT + tid] : U(KD::ONES);
    // bucket starts inside the segment (its barrier also publishes sp)
    const uint32_t dstart = digit_starts(r.segtot + size_t(j) * 256, sm.wsum);
    uint64_t run = tid < 256 ? sbase + dstart + r.rpre[size_t(tid) * r.stride + b] : 0u;
    SplitDigit<KD, NL> dig;
    dig.sp = sp;
    dig.ss.load(sp);
    dig.flip = U(p.flip);
    dig.nbm1 = a.lnb[j] - 1;
    uint64_t t = rb;
    for (; t + TILE <= re; t += TILE)
        downsweep_tile_g<KD, PL, BLOCK, ITEMS, true, NL, UNST>(p, sm, t, TILE, sbase, run, dig);
    if (t < re) downsweep_tile_g<KD, PL, BLOCK, ITEMS, false, NL, UNST>(p, sm, t, uint32_t(re - t), sbase, run, dig);
}

// ---------------------------------------------------------------- bucket table

// One workgroup per large segment, a thread per bucket.
template <class U>
__global__ __launch_bounds__(256) void ss_buckets_kernel(SsArgs a, const uint32_t *segtot)
{
    __shared__ uint32_t wsum[4];
    const uint32_t j = blockIdx.x, d = threadIdx.x, lane = d & 63, w = d >> 6;
    const uint32_t nb = a.lnb[j];
    const uint64_t b0 = a.seg_off[a.lseg[j]], len = a.seg_off[a.lseg[j] + 1] - b0;
    const uint32_t size = d < nb ? segtot[size_t(j) * 256 + d] : 0u;
    uint32_t incl = size;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = incl - size;
    for (uint32_t k = 0; k < w; ++k) pre += wsum[k];
    if (d >= nb) return;
    const uint32_t slot = a.lbb[j] + d;
    a.bstart[slot] = uint32_t(b0) + pre;
    a.bsize[slot] = size;
    // a middle bucket holds order keys in [sp[d - 1], sp[d]); the first and
    // last are open-ended (their LDS sort measures min and max)
    const U *sp = static_cast<const U *>(a.split) + size_t(j) * SS_SPLIT;
    const bool mid = d > 0 && d + 1 < nb && sp[d] > sp[d - 1];
    U *br = static_cast<U *>(a.brange) + 2 * size_t(slot);
    br[0] = mid ? sp[d - 1] : U(0);
    br[1] = mid ? U(sp[d] - 1 - sp[d - 1]) : U(0);
    const uint64_t kth = len / 2;
    const bool holds_k = pre <= kth && kth < uint64_t(pre) + size;
    if (a.select && holds_k) a.ksel[j] = slot;
    if ((!a.select || holds_k) && size > a.lcap) a.big[atomicAdd(a.ctl, 1u)] = slot;
}

// ---------------------------------------------------------------- LDS sort

struct SsLocalArgs {
    const uint32_t *slots;    // bucket mode: the slots to sort (null: slot = workgroup)
    const uint32_t *bstart;
    const uint32_t *bsize;
    const void *brange;       // [nbk][2] key range from the splitters ([1] = 0: unknown)
    const uint32_t *bk_seg;
    const uint32_t *lseg;
    const uint64_t *seg_off;
    const uint32_t *med;      // segment mode: segment ids (whole segments, sorted in place)
    const void *keys_in;      // bucket mode: the scatter's output; segment mode: the keys
    const void *vals_in;      // bucket mode, argsort: the indices the scatter generated
    void *keys_out;
    void *vals_out;           // key-value sort: where the payloads go (with keys_out)
    uint32_t *idx32_out;      // argsort: segment-local indices
    uint64_t *idx64_out;      // argsort of one array (non-null): 64-bit indices, no segment base
    uint64_t flip;
    uint32_t canon;           // qsort(hasnan): canonical NaN bits
    uint32_t cap;             // bucket mode: buckets above this (0: CAP) are sorted elsewhere
    uint32_t n1;              // one array of n1 keys at keys_in (non-zero: no tables; keys only)
};

// The key whose order key is o (the inverse of KD::ord; the NaN order key
// of hasnan gives a NaN pattern).
template <class KD>
__device__ __forceinline__ typename KD::U key_of_ord(typename KD::U o, typename KD::U flip)
{
    using U = typename KD::U;
    const U k = U(o ^ flip);
    if constexpr (KD::KIND == Kind::Unsigned) return k;
    else if constexpr (KD::KIND == Kind::Signed) return U(k ^ KD::SIGN);
    else return (k & KD::SIGN) ? U(k ^ KD::SIGN) : U(~k);
}

// Peer mask of the lanes holding the same 4-bit digit (4 ballots; see
// match_any8).
__device__ __forceinline__ uint64_t match_any4(uint32_t d)
{
    uint32_t lo = ~0u, hi = ~0u;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint32_t x = uint32_t(__builtin_amdgcn_sbfe(int32_t(d), b, 1));
        const uint64_t bal = __builtin_amdgcn_uicmp(x, 0u, 33 /* ne */);
        lo = __builtin_amdgcn_bitop3_b32(lo, uint32_t(bal), x, 0x90);
        hi = __builtin_amdgcn_bitop3_b32(hi, uint32_t(bal >> 32), x, 0x90);
    }
    return (uint64_t(hi) << 32) | lo;
}

// One workgroup sorts a bucket (or a whole medium segment) of at most
// 512 * ITEMS keys (4-byte keys: 12 or 15 items, 8-byte keys: 8, 11 or 12)
// by its order key in LDS, and writes it out.  Keys are
// ranked by r = ord - min (min and max of the bucket): only the bits of
// max - min are sorted.
//
// MSD split, then a rank by comparison: r >> shift, with the shift that
// makes max - min span 4096 values, splits the keys into 4096 sub-buckets (each key takes its slot with one returning LDS
// add on 16-bit counters, two to a word; the order inside a sub-bucket is
// free), which hold about one key each for spread data.  Every key's final
// place is then its sub-bucket's start plus the number of keys of the
// sub-bucket that order before it (by order key, ties by slot -- argsort: by
// original index, so equal keys keep the input order).  Keys only: an
// inverse map in LDS (the slot that ends at each place) then lets the
// workgroup write the bucket out in place order, so the stores are
// contiguous; argsort stores each index at its place directly.
//
// A sub-bucket of more than SS_SUB_MAX keys (duplicates, or keys crowded in
// a small part of the range) would make the comparisons quadratic; the
// bucket is then sorted LSD by r instead: keys only, the lowest 12 bits in one pass whose
// order inside a digit is free, then stable 4-bit passes ranked per wave (a
// 4-bit peer match, per-wave counters); argsort: stable 4-bit passes only.
constexpr uint32_t SS_SUB_MAX = 32;
// waves per SIMD the keys-only LDS sort is compiled for (build-time knob)
#ifndef XSS_SS_LOCAL_WAVES
#define XSS_SS_LOCAL_WAVES 6
#endif
template <class KD, int PL, bool NL, int ITEMS>
__global__ __launch_bounds__(SS_LOCAL_BLOCK) __attribute__((amdgpu_waves_per_eu(PL != PL_NONE ? 4 : XSS_SS_LOCAL_WAVES)))
void ss_local_kernel(SsLocalArgs a)
{
    using U = typename KD::U;
    // PL_IDX: 32-bit indices (ties by index, output indices); PL_V32 /
    // PL_V64: payloads that move with their keys (ties in any order)
    using V = typename PayloadT<PL>::type;
    constexpr int BLOCK = SS_LOCAL_BLOCK;
    constexpr int WAVES = BLOCK / 64;
    constexpr uint32_t CAP = uint32_t(BLOCK) * ITEMS;
    constexpr bool IDX = PL == PL_IDX;
    constexpr bool HAS_V = PL != PL_NONE;
    static_assert(CAP <= 8192, "12-bit ranks fit 13 bits, counts 16 bits");
    static_assert(WAVES * 16 <= 128, "the fallback's counters are scanned by two waves");
    __shared__ U s_keys[CAP];
    __shared__ V s_vals[HAS_V ? CAP : 1];
    __shared__ uint32_t s_cnt[2048];  // 4096 16-bit counters / sub-bucket starts
    // rank path, keys only: the slot whose key ends at each position (argsort
    // stores directly: its key, index and this map would not leave room for
    // two workgroups per CU)
    __shared__ uint16_t s_inv[IDX ? 1 : CAP];
    __shared__ uint32_t s_wc[WAVES * 16];
    __shared__ uint32_t s_wsum[WAVES];
    __shared__ U s_mm[2];          // measured min, max of the order keys
    __shared__ uint32_t s_mx;      // largest sub-bucket
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t st, segb;
    uint32_t m;
    U known_lo = 0, known_range = 0;
    if (a.n1) {
        st = 0;
        m = a.n1;
        segb = 0;
        if (m > CAP) return;  // the caller sends longer arrays elsewhere
    } else if (a.med) {
        const uint32_t seg = a.med[blockIdx.x];
        st = a.seg_off[seg];
        m = uint32_t(a.seg_off[seg + 1] - st);
        segb = st;
        if (m > CAP) return;  // the plan sends longer segments elsewhere
    } else {
        const uint32_t slot = a.slots ? a.slots[blockIdx.x] : blockIdx.x;
        st = a.bstart[slot];
        m = a.bsize[slot];
        if (a.brange) {
            known_lo = static_cast<const U *>(a.brange)[2 * size_t(slot)];
            known_range = static_cast<const U *>(a.brange)[2 * size_t(slot) + 1];
        }
        if (m > CAP || (a.cap && m > a.cap)) return;  // listed for the large-bucket passes
        segb = IDX && !a.idx64_out ? a.seg_off[a.lseg[a.bk_seg[slot]]] : 0;
    }
    if (m == 0) return;
    const U flip = U(a.flip);
    const U *kin = static_cast<const U *>(a.keys_in) + st;
    const V *vin = HAS_V && !a.med && a.vals_in ? static_cast<const V *>(a.vals_in) + st : nullptr;
    U *kout = static_cast<U *>(a.keys_out) + st;
    V *vout = HAS_V && !IDX ? static_cast<V *>(a.vals_out) + st : nullptr;
    uint32_t *iout = a.idx32_out + st;
    // the key (and index) that ends at bucket position q
    auto put = [&](uint32_t q, U x, V v) {
        if constexpr (IDX) {
            if (a.idx64_out) a.idx64_out[st + q] = uint64_t(v);
            else iout[q] = v - uint32_t(segb);
        } else {
            if constexpr (KD::KIND == Kind::Float && NL)
                if (a.canon && KD::is_nan(x)) x = KD::QNAN;
            kout[q] = x;
            if constexpr (HAS_V) vout[q] = v;
        }
    };
    // the bucket's key range: from its splitters, or measured (min / max)
    const bool measure = known_range == 0;
    if (tid == 0) {
        s_mm[0] = U(~U(0));
        s_mm[1] = U(0);
        s_mx = 0u;
    }
    // keys into registers, striped: wave w holds positions [w 64 it, (w + 1) 64 it)
    const uint32_t it = (m + BLOCK - 1) / BLOCK;
    const uint32_t lbase = wave * 64 * it + lane;
    U k[ITEMS];
    V v[ITEMS];
    U t_min = U(~U(0)), t_max = U(0);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t idx = lbase + i * 64;
        const bool real = uint32_t(i) < it && idx < m;
        k[i] = real ? kin[idx] : U(0);
        if constexpr (IDX) v[i] = real ? (vin ? vin[idx] : uint32_t(st + idx)) : V(0);
        else if constexpr (HAS_V) v[i] = real ? vin[idx] : V(0);
        if (measure && real) {
            const U o = KD::ord(k[i], flip, NL);
            t_min = min(t_min, o);
            t_max = max(t_max, o);
        }
    }
    if (measure) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            t_min = min(t_min, U(__shfl_xor(t_min, off)));
            t_max = max(t_max, U(__shfl_xor(t_max, off)));
        }
        __syncthreads();  // s_mm initialised
        if (lane == 0) {
            if constexpr (sizeof(U) == 8) {
                atomicMin(reinterpret_cast<unsigned long long *>(&s_mm[0]), (unsigned long long)t_min);
                atomicMax(reinterpret_cast<unsigned long long *>(&s_mm[1]), (unsigned long long)t_max);
            } else {
                atomicMin(&s_mm[0], t_min);
                atomicMax(&s_mm[1], t_max);
            }
        }
        __syncthreads();
    }
    // keys are sorted by r = ord - min, whose range fits `bits` bits (a
    // range from the splitters may include values no key has: then the
    // bucket is sorted as if keys were spread over all of it)
    const U omin = measure ? uniform(s_mm[0]) : known_lo;
    const U range = measure ? U(uniform(s_mm[1]) - omin) : known_range;
    auto rel = [&](U x) { return U(KD::ord(x, flip, NL) - omin); };
    if (range == 0) {  // all keys equal: in input order
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t idx = lbase + i * 64;
            if (uint32_t(i) < it && idx < m) put(idx, k[i], HAS_V ? v[i] : V(0));
        }
        return;
    }
    auto start_of = [&](uint32_t d) { return (s_cnt[d >> 1] >> ((d & 1u) * 16u)) & 0xffffu; };
    // 12-bit split at `sh` of the keys in registers (order inside a digit
    // free): counts, their scan (largest count to s_mx); the scatter is
    // separate
    uint32_t rd[ITEMS];
    auto rank12 = [&](int sh) {
        for (uint32_t i = tid; i < 2048; i += BLOCK) s_cnt[i] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                const uint32_t idx = lbase + i * 64;
                const bool real = idx < m;
                const uint32_t d = uint32_t(rel(k[i]) >> sh) & 0xfffu;
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
                const bool uni = __ballot(!real || d != d0) == 0;
                uint32_t old = 0;
                if (real && (!uni || lane == 0)) {
                    const uint32_t hs = (d & 1u) * 16u;
                    old = (atomicAdd(&s_cnt[d >> 1], (uni ? 64u : 1u) << hs) >> hs) & 0xffffu;
                }
                rd[i] = old | (d << 13) | (uni ? (1u << 25) : 0u) | (real ? 0u : (1u << 26));
            }
        }
        __syncthreads();
        // exclusive scan: 8 counters (4 words) per thread
        uint32_t c[8], sum = 0, mx = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t w = s_cnt[tid * 4 + j];
            c[2 * j] = w & 0xffffu;
            c[2 * j + 1] = w >> 16;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            sum += c[j];
            mx = max(mx, c[j]);
        }
        uint32_t incl = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off);
            if (lane >= uint32_t(off)) incl += t;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = max(mx, uint32_t(__shfl_xor(mx, off)));
        if (lane == 63) s_wsum[wave] = incl;
        if (lane == 0) atomicMax(&s_mx, mx);
        __syncthreads();
        uint32_t run = incl - sum;
        for (uint32_t w = 0; w < wave; ++w) run += s_wsum[w];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t lo = run;
            run += c[2 * j];
            s_cnt[tid * 4 + j] = lo | (run << 16);
            run += c[2 * j + 1];
        }
        __syncthreads();
    };
    // as_rel: store r = ord - min instead of the key (the rank compares
    // order keys; the write-out rebuilds the key)
    auto scatter12 = [&](bool as_rel) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                uint32_t r = rd[i];
                if (r >> 26) continue;
                if (r >> 25) r = (__builtin_amdgcn_readfirstlane(r) + lane) & 0x1ffffffu;
                const uint32_t slot = start_of(r >> 13) + (r & 0x1fffu);
                s_keys[slot] = as_rel ? rel(k[i]) : k[i];
                if constexpr (HAS_V) s_vals[slot] = v[i];
            }
        }
        __syncthreads();
    };
    const int bits = sizeof(U) == 8 ? 64 - __builtin_clzll(uint64_t(range)) : 32 - __builtin_clz(uint32_t(range));
    const int msh = bits > 12 ? bits - 12 : 0;  // the digit (ord - min) >> msh spans [0, 4096)
    rank12(msh);
    if (__builtin_amdgcn_readfirstlane(s_mx) <= SS_SUB_MAX) {
        // the rank compares r = ord - min, kept in LDS in place of the key
        // when the key can be rebuilt from it (not for NaN payloads that
        // a select or key-value sort with hasnan must keep).  4-byte keys
        // only: the 8-byte forms needed more registers (147 VGPRs for the
        // 12-item u64 sort, 3 waves per SIMD) and ran slower
        const bool as_rel = sizeof(U) == 4 && (IDX || !NL || a.canon);
        scatter12(as_rel);
        // one item at a time (no register arrays here): keeps the loads of
        // different items from being hoisted together
#pragma unroll 1
        for (uint32_t i = 0; i < it; ++i) {
            const uint32_t p = lbase + i * 64;
            if (p < m) {
                const U x = s_keys[p];
                const V vx = HAS_V ? s_vals[p] : V(0);
                const U ox = as_rel ? x : U(KD::ord(x, flip, NL) - omin);
                const uint32_t d = uint32_t(ox >> msh);
                const uint32_t lo = start_of(d), hi = d < 4095 ? start_of(d + 1) : m;
                uint32_t r = 0;
                for (uint32_t q = lo; q < hi; ++q) {
                    const U oy = as_rel ? s_keys[q] : U(KD::ord(s_keys[q], flip, NL) - omin);
                    bool before = oy < ox;
                    if (oy == ox) {
                        if constexpr (IDX) before = s_vals[q] < vx;
                        else before = q < p;
                    }
                    r += before ? 1u : 0u;
                }
                if constexpr (IDX) put(lo + r, x, vx);
                else s_inv[lo + r] = uint16_t(p);
            }
        }
        if constexpr (!IDX) {
            __syncthreads();
            // written in position order: consecutive threads, consecutive places
            for (uint32_t q = tid; q < m; q += BLOCK) {
                const uint32_t p = s_inv[q];
                const U y = s_keys[p];
                put(q, as_rel ? key_of_ord<KD>(U(y + omin), flip) : y, HAS_V ? s_vals[p] : V(0));
            }
        }
        return;
    }
    // LSD fallback.  4 bits at `sh`, stable: items ranked in position order
    // (wave, item, lane) with a 4-bit peer match and per-wave counters; the
    // keys come from registers (the first pass) or from LDS.
    auto pass4 = [&](int sh, bool from_lds) {
        if (tid < WAVES * 16) s_wc[tid] = 0;
        __syncthreads();  // also: the previous pass's scatter is complete
        uint32_t *wc = s_wc + wave * 16;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                const uint32_t idx = lbase + i * 64;
                const bool real = idx < m;
                if (from_lds && real) {
                    k[i] = s_keys[idx];
                    if constexpr (HAS_V) v[i] = s_vals[idx];
                }
                const uint32_t d = uint32_t(rel(k[i]) >> sh) & 0xfu;
                const uint64_t peers = match_any4(d) & __ballot(real);
                uint32_t r = 0;
                if (real) {
                    const uint32_t below = lanes_below(peers), tot = uint32_t(__popcll(peers));
                    const uint32_t c = wc[d];
                    r = c + below;
                    if (below + 1 == tot) wc[d] = c + tot;
                }
                rd[i] = r | (d << 16);
            }
        }
        __syncthreads();
        // slot base of every (wave, digit), digit-major then wave: a scan of
        // the 128 counts held by threads 0..127 (waves 0 and 1)
        uint32_t c = 0;
        if (tid < WAVES * 16) c = s_wc[(tid % WAVES) * 16 + tid / WAVES];
        uint32_t incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off);
            if (lane >= uint32_t(off)) incl += t;
        }
        if (tid == 63) s_wsum[0] = incl;
        __syncthreads();
        if (tid < WAVES * 16) s_wc[(tid % WAVES) * 16 + tid / WAVES] = (tid >= 64 ? s_wsum[0] : 0u) + incl - c;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it && lbase + i * 64 < m) {
                const uint32_t slot = wc[rd[i] >> 16] + (rd[i] & 0xffffu);
                s_keys[slot] = k[i];
                if constexpr (HAS_V) s_vals[slot] = v[i];
            }
        }
        __syncthreads();  // the scatter and its reads of s_wc are done
    };
    int sh = 0;
    bool from_lds = false;
    if constexpr (!IDX) {
        rank12(sh);
        scatter12(false);
        sh += 12;
        from_lds = true;
    }
    for (; sh < bits; sh += 4) {
        pass4(sh, from_lds);
        from_lds = true;
    }
    for (uint32_t p = tid; p < m; p += BLOCK) put(p, s_keys[p], HAS_V ? s_vals[p] : V(0));
}

// ---------------------------------------------------------------- large buckets

struct SsBigArgs {
    const uint32_t *ctl;       // [0]: listed buckets
    const uint32_t *big;
    const uint32_t *bstart;
    const uint32_t *bsize;
    const uint32_t *bk_seg;
    const uint32_t *lseg;
    const uint64_t *seg_off;
    void *y_keys;              // the scatter's output (where each listed bucket is)
    void *y_vals;              // argsort: its indices
    void *z_keys;              // the other buffer of the ping-pong: the caller's keys
    void *z_vals;              // (argsort: a second scratch pair)
    uint32_t *idx32_out;       // argsort
    uint64_t flip;
    uint32_t canon;
    uint32_t end_in_y;         // the keys end in y (select: y is the caller's array)
};

// One workgroup per listed bucket: 8-bit LSD passes through HBM over the
// digits on which the bucket's keys differ, stable, tile by tile as a
// downsweep does (each pass: a histogram loop, a scan, the tiles).  Keys end
// in the caller's array (argsort: the last pass writes the indices).
template <class KD, int PL, bool NL>
__global__ __launch_bounds__(SS_BIG_BLOCK) void ss_big_kernel(SsBigArgs g)
{
    using U = typename KD::U;
    using V = uint32_t;
    constexpr int BLOCK = SS_BIG_BLOCK, ITEMS = SS_BIG_ITEMS, TILE = BLOCK * ITEMS;
    constexpr bool IDX = PL == PL_IDX;
    __shared__ DownsweepSmem<KD, PL, BLOCK, ITEMS> sm;
    __shared__ uint32_t h[256];
    __shared__ U s_ao[2];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t nlist = g.ctl[0];
    const U flip = U(g.flip);
    for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
        const uint32_t slot = g.big[li];
        const uint64_t st = g.bstart[slot], m = g.bsize[slot];
        const uint64_t segb = g.seg_off[g.lseg[g.bk_seg[slot]]];
        __syncthreads();  // the previous bucket's LDS readers are done
        if (tid == 0) {
            s_ao[0] = U(~U(0));
            s_ao[1] = U(0);
        }
        __syncthreads();
        const U *yk = static_cast<const U *>(g.y_keys);
        U t_and = U(~U(0)), t_or = U(0);
        for (uint64_t p = st + tid; p < st + m; p += BLOCK) {
            const U o = KD::ord(yk[p], flip, NL);
            t_and &= o;
            t_or |= o;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            t_and &= __shfl_xor(t_and, off);
            t_or |= __shfl_xor(t_or, off);
        }
        if (lane == 0) {
            if constexpr (sizeof(U) == 8) {
                atomicAnd(reinterpret_cast<unsigned long long *>(&s_ao[0]), (unsigned long long)t_and);
                atomicOr(reinterpret_cast<unsigned long long *>(&s_ao[1]), (unsigned long long)t_or);
            } else {
                atomicAnd(&s_ao[0], uint32_t(t_and));
                atomicOr(&s_ao[1], uint32_t(t_or));
            }
        }
        __syncthreads();
        const U diff = uniform(U(s_ao[0] ^ s_ao[1]));
        uint32_t dmask = 0;
        for (int q = 0; q < KD::PASSES; ++q)
            if (uint32_t(diff >> (8 * q)) & 0xffu) dmask |= 1u << q;
        void *sk = g.y_keys, *sv = g.y_vals, *dk = g.z_keys, *dv = g.z_vals;
        bool in_y = true;
        for (int q = 0; q < KD::PASSES; ++q) {
            if (!((dmask >> q) & 1u)) continue;
            const bool last = (dmask >> (q + 1)) == 0;
            // histogram of the digit over the bucket
            for (uint32_t i = tid; i < 256; i += BLOCK) h[i] = 0;
            __syncthreads();
            const U *kin = static_cast<const U *>(sk);
            for (uint64_t p = st + tid; p < st + m; p += BLOCK) atomicAdd(&h[KD::digit(kin[p], flip, NL, 8 * q)], 1u);
            __syncthreads();
            const uint32_t dstart = digit_starts(h, sm.wsum);
            uint64_t run = tid < 256 ? st + dstart : 0u;
            PassArgs a{};
            a.keys_in = sk; a.vals_in = sv; a.keys_out = dk; a.vals_out = dv;
            a.idx32_out = g.idx32_out; a.shift = uint32_t(8 * q); a.flip = g.flip; a.nan_last = NL;
            a.first = 0; a.last = last; a.canon_nan = last ? g.canon : 0u;
            uint64_t t = st;
            for (; t + TILE <= st + m; t += TILE) downsweep_tile<KD, PL, BLOCK, ITEMS, true, NL>(a, sm, t, TILE, segb, run);
            if (t < st + m) downsweep_tile<KD, PL, BLOCK, ITEMS, false, NL>(a, sm, t, uint32_t(st + m - t), segb, run);
            __syncthreads();
            void *tk = sk, *tv = sv;
            sk = dk;
            sv = dv;
            dk = tk;
            dv = tv;
            in_y = !in_y;
        }
        // keys left in the other buffer than the one they end in (by the
        // parity of the passes): copied (argsort: the indices written)
        if constexpr (IDX) {
            if (dmask == 0) {
                const V *yv = static_cast<const V *>(g.y_vals);
                for (uint64_t p = st + tid; p < st + m; p += BLOCK) g.idx32_out[p] = yv[p] - uint32_t(segb);
            }
        } else {
            if (in_y != (g.end_in_y != 0)) {
                const U *src = static_cast<const U *>(in_y ? g.y_keys : g.z_keys);
                U *dst = static_cast<U *>(in_y ? g.z_keys : g.y_keys);
                for (uint64_t p = st + tid; p < st + m; p += BLOCK) {
                    U x = src[p];
                    if constexpr (KD::KIND == Kind::Float && NL)
                        if (g.canon && KD::is_nan(x)) x = KD::QNAN;
                    dst[p] = x;
                }
            } else if constexpr (KD::KIND == Kind::Float && NL) {
                // no pass ran (every order key equal, e.g. all NaN): the
                // NaNs still become the canonical NaN where they are
                if (dmask == 0 && g.canon) {
                    U *k = static_cast<U *>(in_y ? g.y_keys : g.z_keys);
                    for (uint64_t p = st + tid; p < st + m; p += BLOCK)
                        if (KD::is_nan(k[p])) k[p] = KD::QNAN;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- select

// Select's partition: every key of a large segment goes to one of three
// regions of the caller's array -- below, inside or above the bucket that
// holds rank len / 2 (order-key bounds from its splitters) -- from the copy
// the histogram read made.  One workgroup per range (4 waves); a tile of
// 16 keys per thread is classified, each key ranked inside its wave and
// class by ballots (no atomics), the waves' class counts scanned in LDS,
// the tile laid out in LDS by class, and written out as three contiguous
// runs at the regions' running bases.  The order inside a region is free.  Range bases come from the bucket histogram's scan: the
// region below starts at the segment's start plus the range's counts of
// the lower buckets in earlier ranges, and so on.
constexpr int SS_P3_BLOCK = 256;
constexpr int SS_P3_ITEMS = 16;
template <class KD, bool NL>
__global__ __launch_bounds__(SS_P3_BLOCK) void ss_part3_kernel(SsArgs a, SegRanges r, const void *src_v,
                                                                void *dst_v)
{
    using U = typename KD::U;
    constexpr int WAVES = SS_P3_BLOCK / 64;
    constexpr uint32_t TILE = SS_P3_BLOCK * SS_P3_ITEMS;
    __shared__ uint32_t s_red[3][256];
    __shared__ uint32_t s_wc[WAVES][3];
    __shared__ U s_tile[TILE];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b = blockIdx.x;
    const uint32_t j = r.rbseg[b];
    const uint32_t kd = a.ksel[j] - a.lbb[j], nb = a.lnb[j];
    const U *sp = static_cast<const U *>(a.split) + size_t(j) * SS_SPLIT;
    const U lo = kd > 0 ? sp[kd - 1] : U(0);
    const U hi = kd + 1 < nb ? U(sp[kd] - 1) : U(KD::ONES);  // inclusive
    // region bases of this range: segment totals and the range's prefix
    // (digit-major rpre) summed below / at / above kd
    const uint32_t tot = r.segtot[size_t(j) * 256 + tid];
    const uint32_t pre = r.rpre[size_t(tid) * r.stride + b];
    s_red[0][tid] = tid < kd ? tot : 0u;
    s_red[1][tid] = tid < kd ? pre : 0u;
    s_red[2][tid] = tid > kd ? pre : 0u;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s_red[0][tid] += s_red[0][tid + h];
            s_red[1][tid] += s_red[1][tid + h];
            s_red[2][tid] += s_red[2][tid + h];
        }
        __syncthreads();
    }
    const uint64_t sbase = r.sbase[b];
    const uint32_t l_lt = s_red[0][0], l_eq = r.segtot[size_t(j) * 256 + kd];
    uint64_t base[3] = {sbase + s_red[1][0], sbase + l_lt + r.rpre[size_t(kd) * r.stride + b],
                        sbase + l_lt + l_eq + s_red[2][0]};
    const U *src = static_cast<const U *>(src_v);
    U *dst = static_cast<U *>(dst_v);
    const U flip = U(a.flip);
    const uint64_t rb = r.begin[b], re = r.end[b];
    for (uint64_t t = rb; t < re; t += TILE) {
        U key[SS_P3_ITEMS];
        uint32_t off[SS_P3_ITEMS];  // class (bits 30..31), place in the wave's class run of this tile
        uint32_t wc0 = 0, wc1 = 0, wc2 = 0;
#pragma unroll
        for (int i = 0; i < SS_P3_ITEMS; ++i) {
            const uint64_t p = t + uint64_t(i) * SS_P3_BLOCK + tid;
            const bool real = p < re;
            key[i] = real ? src[p] : U(0);
            const U o = KD::ord(key[i], flip, NL);
            const uint32_t c = o < lo ? 0u : (o <= hi ? 1u : 2u);
            const uint64_t m0 = __ballot(real && c == 0), m1 = __ballot(real && c == 1), m2 = __ballot(real && c == 2);
            const uint64_t mine = c == 0 ? m0 : c == 1 ? m1 : m2;
            const uint32_t run = c == 0 ? wc0 : c == 1 ? wc1 : wc2;
            off[i] = real ? ((c << 30) | (run + lanes_below(mine))) : ~0u;
            wc0 += uint32_t(__popcll(m0));
            wc1 += uint32_t(__popcll(m1));
            wc2 += uint32_t(__popcll(m2));
        }
        if (lane == 0) {
            s_wc[wave][0] = wc0;
            s_wc[wave][1] = wc1;
            s_wc[wave][2] = wc2;
        }
        __syncthreads();
        // tile slots: class c starts at cs[c]; wave w's keys of class c
        // after those of the waves before it
        uint32_t wslot[3], total[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            uint32_t before = 0, all = 0;
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t x = s_wc[w][c];
                before += w < int(wave) ? x : 0u;
                all += x;
            }
            wslot[c] = before;
            total[c] = all;
        }
        const uint32_t cs1 = total[0], cs2 = total[0] + total[1], n_tile = cs2 + total[2];
        wslot[1] += cs1;
        wslot[2] += cs2;
#pragma unroll
        for (int i = 0; i < SS_P3_ITEMS; ++i) {
            if (off[i] != ~0u) {
                const uint32_t c = off[i] >> 30;
                s_tile[wslot[c] + (off[i] & 0x3fffffffu)] = key[i];
            }
        }
        __syncthreads();
        // write-out: consecutive threads take consecutive slots, so each
        // class is one contiguous run in HBM
        for (uint32_t q = tid; q < n_tile; q += SS_P3_BLOCK) {
            const uint64_t g = q < cs1 ? base[0] + q : q < cs2 ? base[1] + (q - cs1) : base[2] + (q - cs2);
            dst[g] = s_tile[q];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) base[c] += total[c];
        __syncthreads();  // s_wc and s_tile are read before the next tile writes them
    }
}

}  // namespace xss
