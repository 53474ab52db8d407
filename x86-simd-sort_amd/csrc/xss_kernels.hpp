// xss_kernels.hpp -- CDNA4 (gfx950) kernels of the LSD onesweep radix sort.
//
// The reference sorts with a recursive AVX-512 quicksort (partition by
// compress-store, src/xss-common-qsort.h:199-605, pivots from
// src/xss-pivot-selection.hpp).  None of that maps to a GPU; this file
// replaces it with an 8-bit-digit least-significant-digit radix sort in the
// onesweep form (one digit-histogram read of the keys, then one scatter pass
// per digit whose tiles chain their bucket offsets by decoupled look-back):
//
//   hist_kernel    one read of all keys, all P digit histograms at once
//                  (LDS atomics), one partial histogram row per workgroup
//   scan_kernel    sums the partial rows, exclusive scan per digit position
//                  -> the 256 global bucket offsets of each pass
//   pass_kernel    per tile of BLOCK*ITEMS keys: wave64 ballot ranking
//                  (8 ballots + mbcnt per item), per-wave LDS digit counters,
//                  tile-local sort in LDS, look-back over predecessor tiles
//                  for the per-digit global base, coalesced write-out
//   block_sort     one workgroup bitonic-sorts a whole (small) array or a
//                  segment in LDS
//
// Keys travel as their original bits; ord() of xss_keys.hpp is recomputed
// where a digit is needed.  Payloads (kv values, argsort indices) move with
// their keys.  LSD needs a stable pass: items are ranked in input order
// (wave-striped layout, items in order, lanes in order).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xss_keys.hpp"

namespace xss {

// payload modes
enum : int {
    PL_NONE = 0,  // keys only (qsort)
    PL_V32 = 1,   // 32-bit values carried (keyvalue with 32-bit values)
    PL_V64 = 2,   // 64-bit values carried (keyvalue with 64-bit values)
    PL_IDX = 3,   // 32-bit index generated in the first pass; the last pass
                  // writes it widened to size_t and drops the keys (argsort)
};
template <int PL>
struct PayloadT { using type = uint32_t; };
template <>
struct PayloadT<PL_V64> { using type = uint64_t; };

constexpr int HIST_BLOCK = 256;
constexpr int HIST_MAX_BLOCKS = 1024;
constexpr uint32_t FLAG_AGG = 1u, FLAG_PRE = 2u;
constexpr uint32_t COUNT_MASK = 0x3fffffffu;  // 30-bit counts: n < 2^30 per call

// control block in device memory (one per workspace)
struct Ctrl {
    uint32_t epoch;           // look-back epoch base of the current sort
    uint32_t pad0;
    uint32_t tile_counter[8]; // dynamic tile ids, one counter per pass
    uint32_t misc[6];
};

__device__ __forceinline__ uint64_t match_any8(uint32_t d)
{
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}
__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int = 0>
__global__ void fill_u64_kernel(uint64_t *p, uint64_t v, uint64_t n)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------- histogram

struct HistArgs {
    const void *keys;
    uint32_t n;
    uint32_t nan_last;
    uint64_t flip;
    uint32_t *partial;   // [gridDim.x][P][256]
    Ctrl *ctrl;
    uint64_t *status;    // zeroed here when the epoch counter would wrap
    uint64_t status_words;
};

template <class KD>
__global__ __launch_bounds__(HIST_BLOCK) void hist_kernel(HistArgs a)
{
    using U = typename KD::U;
    constexpr int P = KD::PASSES;
    __shared__ uint32_t h[P * 256];
    for (int i = threadIdx.x; i < P * 256; i += HIST_BLOCK) h[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 8) a.ctrl->tile_counter[threadIdx.x] = 0;
    // epoch wrap guard: the scan kernel restarts epochs at 1 when they would
    // overflow; stale status words must not carry a reusable epoch then.
    if (a.ctrl->epoch > 0xffffff00u) {
        for (uint64_t i = uint64_t(blockIdx.x) * HIST_BLOCK + threadIdx.x; i < a.status_words;
             i += uint64_t(gridDim.x) * HIST_BLOCK)
            a.status[i] = 0;
    }
    __syncthreads();
    const U *keys = static_cast<const U *>(a.keys);
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    constexpr int VEC = 16 / sizeof(U);
    // keys before the first 16-byte boundary are read one by one
    const uint32_t mis = uint32_t(reinterpret_cast<uintptr_t>(keys) & 15u);
    const uint32_t head = min(a.n, uint32_t(((16u - mis) & 15u) / sizeof(U)));
    const uint32_t nvec = (a.n - head) / VEC;
    const uint4 *kv = reinterpret_cast<const uint4 *>(keys + head);
    const uint32_t gtid = blockIdx.x * HIST_BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * HIST_BLOCK;
    for (uint32_t v = gtid; v < nvec; v += stride) {
        uint4 q = kv[v];
        const U *e = reinterpret_cast<const U *>(&q);
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
            const U o = KD::ord(e[j], flip, nl);
#pragma unroll
            for (int p = 0; p < P; ++p) atomicAdd(&h[p * 256 + ((o >> (8 * p)) & 0xff)], 1u);
        }
    }
    // unaligned head and the tail that does not fill a vector
    const uint32_t tail0 = head + nvec * VEC;
    const uint32_t rest = head + (a.n - tail0);
    for (uint32_t r = gtid; r < rest; r += stride) {
        const uint32_t i = r < head ? r : tail0 + (r - head);
        const U o = KD::ord(keys[i], flip, nl);
#pragma unroll
        for (int p = 0; p < P; ++p) atomicAdd(&h[p * 256 + ((o >> (8 * p)) & 0xff)], 1u);
    }
    __syncthreads();
    uint32_t *dst = a.partial + size_t(blockIdx.x) * P * 256;
    for (int i = threadIdx.x; i < P * 256; i += HIST_BLOCK) dst[i] = h[i];
}

// ---------------------------------------------------------------- scan

struct ScanArgs {
    const uint32_t *partial;
    uint32_t nblocks;
    uint32_t passes;
    uint32_t *offsets;   // [P][256]
    Ctrl *ctrl;
};

// one workgroup per digit position: sum the partial rows, exclusive scan
template <int = 0>
__global__ __launch_bounds__(256) void scan_kernel(ScanArgs a)
{
    __shared__ uint32_t wsum[4];
    const uint32_t p = blockIdx.x, d = threadIdx.x, lane = d & 63, w = d >> 6;
    uint32_t tot = 0;
    for (uint32_t b = 0; b < a.nblocks; ++b) tot += a.partial[(size_t(b) * a.passes + p) * 256 + d];
    uint32_t incl = tot;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t i = 0; i < w; ++i) pre += wsum[i];
    a.offsets[p * 256 + d] = pre + incl - tot;
    if (p == 0 && d == 0) {
        uint32_t e = a.ctrl->epoch;
        a.ctrl->epoch = (e > 0xffffff00u) ? 1u : e + 8u;
    }
}

// ---------------------------------------------------------------- scatter pass

struct PassArgs {
    const void *keys_in;
    void *keys_out;
    const void *vals_in;
    void *vals_out;
    uint64_t *idx_out;       // PL_IDX, last pass
    uint32_t n;
    uint32_t shift;
    uint64_t flip;
    uint32_t nan_last;
    uint32_t first;          // PL_IDX: generate indices instead of loading them
    uint32_t last;           // PL_IDX: write size_t indices only
    uint32_t canon_nan;      // qsort(hasnan): write canonical NaN bits
    const uint32_t *offsets; // 256 global bucket offsets of this pass
    uint64_t *status;        // [tiles][256] look-back words
    Ctrl *ctrl;
    uint32_t pass;
};

template <class KD, int PL, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void pass_kernel(PassArgs a)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_PL = PL != PL_NONE;
    constexpr int WAVES = BLOCK / 64;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int WTILE = 64 * ITEMS;
    static_assert(BLOCK >= 256 && BLOCK % 64 == 0, "one thread per digit needed");
    static_assert(WTILE <= 65535, "rank packing");

    __shared__ uint32_t s_cnt[WAVES * 256];
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_wsum[WAVES];
    __shared__ uint32_t s_tile;
    __shared__ U s_keys[TILE];
    __shared__ V s_vals[HAS_PL ? TILE : 1];

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t i = tid; i < WAVES * 256; i += BLOCK) s_cnt[i] = 0;
    if (tid == 0)
        s_tile = __hip_atomic_fetch_add(&a.ctrl->tile_counter[a.pass], 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t tbase = tile * TILE;
    const uint32_t valid = min(uint32_t(TILE), a.n - tbase);
    const uint32_t epoch = a.ctrl->epoch + a.pass;
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const int shift = int(a.shift);

    const U *kin = static_cast<const U *>(a.keys_in);
    const V *vin = static_cast<const V *>(a.vals_in);
    U key[ITEMS];
    V val[ITEMS];
    const uint32_t lbase = wave * WTILE + lane;
    if (valid == TILE) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t lp = lbase + i * 64;
            key[i] = kin[tbase + lp];
            if constexpr (PL == PL_IDX) val[i] = a.first ? (tbase + lp) : vin[tbase + lp];
            else if constexpr (HAS_PL) val[i] = vin[tbase + lp];
        }
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t lp = lbase + i * 64;
            key[i] = U(0);
            if constexpr (HAS_PL) val[i] = V(0);
            if (lp < valid) {
                key[i] = kin[tbase + lp];
                if constexpr (PL == PL_IDX) val[i] = a.first ? (tbase + lp) : vin[tbase + lp];
                else if constexpr (HAS_PL) val[i] = vin[tbase + lp];
            }
        }
    }

    // --- rank: stable within the tile (wave-striped order), per-wave counters
    uint32_t rd[ITEMS];  // rank in low 16 bits, digit in bits 16..23
    uint32_t *wc = s_cnt + wave * 256;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t lp = lbase + i * 64;
        const uint32_t d = (lp < valid) ? KD::digit(key[i], flip, nl, shift) : 255u;
        const uint64_t peers = match_any8(d);
        const uint32_t below = lanes_below(peers);
        const uint32_t tot = uint32_t(__popcll(peers));
        const uint32_t c = wc[d];
        rd[i] = (c + below) | (d << 16);
        if (below + 1 == tot) wc[d] = c + tot;
    }
    __syncthreads();

    // --- per-digit tile counts, wave prefixes, publish the tile aggregate
    uint32_t tcount = 0, pub = 0;
    uint64_t *my_status = a.status + size_t(tile) * 256 + tid;
    if (tid < 256) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_cnt[w * 256 + tid];
            s_cnt[w * 256 + tid] = s;
            s += c;
        }
        tcount = s;
        pub = tcount - (tid == 255 ? (TILE - valid) : 0u);
        st_relaxed(my_status, (uint64_t(epoch) << 32) | ((tile == 0 ? FLAG_PRE : FLAG_AGG) << 30) | pub);
    }
    // exclusive scan of tcount over the 256 digits (tile-local bucket starts)
    uint32_t incl = tcount;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += t;
    }
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    if (tid < 256) {
        uint32_t pre = 0;
        for (uint32_t w = 0; w < wave; ++w) pre += s_wsum[w];
        s_start[tid] = pre + incl - tcount;
    }
    __syncthreads();

    // --- tile-local sort: scatter into LDS at bucket start + wave prefix + rank
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t d = rd[i] >> 16;
        const uint32_t lpos = s_start[d] + wc[d] + (rd[i] & 0xffffu);
        s_keys[lpos] = key[i];
        if constexpr (HAS_PL) s_vals[lpos] = val[i];
    }

    // --- decoupled look-back: exclusive prefix of this digit over earlier tiles
    if (tid < 256) {
        uint32_t excl = 0;
        if (tile > 0) {
            int32_t j = int32_t(tile) - 1;
            while (true) {
                const uint64_t s = ld_relaxed(a.status + size_t(j) * 256 + tid);
                if (uint32_t(s >> 32) != epoch) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const uint32_t lo = uint32_t(s);
                excl += lo & COUNT_MASK;
                if ((lo >> 30) == FLAG_PRE) break;
                --j;
            }
            st_relaxed(my_status, (uint64_t(epoch) << 32) | (FLAG_PRE << 30) | (excl + pub));
        }
        s_base[tid] = a.offsets[tid] + excl - s_start[tid];
    }
    __syncthreads();

    // --- write-out: consecutive threads take consecutive sorted positions
    U *kout = static_cast<U *>(a.keys_out);
    V *vout = static_cast<V *>(a.vals_out);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t lpos = i * BLOCK + tid;
        if (lpos < valid) {
            const U x = s_keys[lpos];
            const uint32_t d = KD::digit(x, flip, nl, shift);
            const uint32_t g = s_base[d] + lpos;
            if constexpr (PL == PL_IDX) {
                if (a.last) {
                    a.idx_out[g] = uint64_t(s_vals[lpos]);
                    continue;
                }
            }
            if constexpr (KD::KIND == Kind::Float) {
                kout[g] = (a.canon_nan && KD::is_nan(x)) ? KD::QNAN : x;
            } else {
                kout[g] = x;
            }
            if constexpr (HAS_PL) vout[g] = s_vals[lpos];
        }
    }
}

// ---------------------------------------------------------------- block sort

struct BlockSortArgs {
    const void *keys_in;
    void *keys_out;
    const void *vals_in;
    void *vals_out;
    uint64_t *idx_out;         // PL_IDX
    const uint64_t *seg_off;   // segment s = [seg_off[s], seg_off[s+1]); null: one array [0, n)
    const uint32_t *seg_list;  // segment ids handled by this launch (null: blockIdx.x)
    uint32_t n;
    uint32_t nan_last;
    uint64_t flip;
    uint32_t canon_nan;
    uint32_t idx_local;        // PL_IDX: write segment-local indices
};

// MAXN-element LDS bitonic sort by (ord(key), original position): the
// position tie-break makes the network's result unique (and stable).
template <class KD, int PL, int BLOCK, int MAXN>
__global__ __launch_bounds__(BLOCK) void block_sort_kernel(BlockSortArgs a)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_VAL = PL == PL_V32 || PL == PL_V64;
    using IdxT = typename std::conditional<(MAXN <= 65536), uint16_t, uint32_t>::type;
    __shared__ U s_ord[MAXN];
    __shared__ IdxT s_idx[MAXN];
    __shared__ U s_raw[MAXN];
    __shared__ V s_val[HAS_VAL ? MAXN : 1];

    const uint32_t seg = a.seg_list ? a.seg_list[blockIdx.x] : blockIdx.x;
    const uint64_t base = a.seg_off ? a.seg_off[seg] : 0;
    const uint32_t n = a.seg_off ? uint32_t(a.seg_off[seg + 1] - base) : a.n;
    if (n < 2) {
        if (n == 1) {
            const U *kin = static_cast<const U *>(a.keys_in);
            if (threadIdx.x == 0) {
                if constexpr (PL == PL_IDX) {
                    a.idx_out[base] = a.idx_local ? 0 : base;
                } else {
                    U x = kin[base];
                    if constexpr (KD::KIND == Kind::Float)
                        if (a.canon_nan && KD::is_nan(x)) x = KD::QNAN;
                    static_cast<U *>(a.keys_out)[base] = x;
                    if constexpr (HAS_VAL)
                        static_cast<V *>(a.vals_out)[base] = static_cast<const V *>(a.vals_in)[base];
                }
            }
        }
        return;
    }
    uint32_t npow = 2;
    while (npow < n) npow <<= 1;
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const U *kin = static_cast<const U *>(a.keys_in) + base;
    for (uint32_t i = threadIdx.x; i < npow; i += BLOCK) {
        if (i < n) {
            const U x = kin[i];
            s_raw[i] = x;
            s_ord[i] = KD::ord(x, flip, nl);
            if constexpr (HAS_VAL) s_val[i] = static_cast<const V *>(a.vals_in)[base + i];
        } else {
            s_ord[i] = KD::ONES;
        }
        s_idx[i] = IdxT(i);
    }
    __syncthreads();
    for (uint32_t k = 2; k <= npow; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < npow / 2; t += BLOCK) {
                const uint32_t i = 2 * t - (t & (j - 1));
                const uint32_t l = i + j;
                const U oi = s_ord[i], ol = s_ord[l];
                const IdxT ii = s_idx[i], il = s_idx[l];
                const bool gt = (oi > ol) || (oi == ol && ii > il);
                const bool up = (i & k) == 0;
                if (gt == up) {
                    s_ord[i] = ol; s_ord[l] = oi;
                    s_idx[i] = il; s_idx[l] = ii;
                }
            }
            __syncthreads();
        }
    }
    U *kout = static_cast<U *>(a.keys_out) + base;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        const uint32_t src = s_idx[i];
        if constexpr (PL == PL_IDX) {
            a.idx_out[base + i] = a.idx_local ? uint64_t(src) : base + src;
        } else {
            U x = s_raw[src];
            if constexpr (KD::KIND == Kind::Float)
                if (a.canon_nan && KD::is_nan(x)) x = KD::QNAN;
            kout[i] = x;
            if constexpr (HAS_VAL) static_cast<V *>(a.vals_out)[base + i] = s_val[src];
        }
    }
}

}  // namespace xss
