// xss_select.hpp -- radix select: qselect, keyvalue_select, argselect and
// the partial sorts built on them.
//
// The reference selects with a quickselect that partitions around sampled
// pivots and recurses into one side (qselect_ src/xss-common-qsort.h:607-649,
// kvselect_ src/xss-common-keyvaluesort.hpp:519-562, argselect_
// src/xss-common-argsort.h:551-585).  Here the k-th key is found by digit
// histograms, most significant digit first, entirely on the device (no host
// round trip between rounds):
//
//   round 0   sel_hist     histogram of the top digit of every key (and a
//                          copy of keys/values into scratch for the in-place
//                          forms); sel_choose picks the bucket holding rank k
//   round 1   sel_filter   keys of that bucket are compacted into per-block
//                          candidate lists; histogram of the next digit
//   round r   sel_cand     histogram of digit r over the candidates still
//                          matching the known high digits
//   final     sel_partition  3-way split of all keys around K = ord of the
//                          k-th key: [< K | == K | > K] at [0, cl) [cl, cl+ce)
//                          [cl+ce, n); every tile reserves its slots with one
//                          atomic per non-empty class
//
// The arrangement inside the three classes is unspecified, as in the
// reference (its tests check only the partition property,
// tests/test-qsort-common.h:52-93).  NaNs (hasnan) rank above +inf, so they
// end up at the back (ascending) or front (descending) with their bits kept,
// as move_nans_to_end/start_of_array do (src/xss-common-qsort.h:137-180).
#pragma once
#include "xss_sort.hpp"

namespace xss {

constexpr int SEL_BLOCK = 256;
constexpr int SEL_TILE = SEL_BLOCK * 16;
constexpr int SEL_MAX_BLOCKS = 1024;  // chunked select kernels; partial rows live in the rsum area

struct SelState {
    uint64_t prefix;    // ord bits of K fixed so far
    uint64_t mask;      // which ord bits are fixed
    uint32_t k_rem;     // rank of K among the current candidates
    uint32_t cl;        // keys known to rank below K
    uint32_t ce;        // keys in the chosen bucket (== K after the last round)
    uint32_t cursor[3]; // partition write cursors
};
static_assert(sizeof(SelState) <= 128, "SelState must fit the control block");

inline SelState *sel_state(Workspace *ws) { return reinterpret_cast<SelState *>(ws->meta + 128); }

struct SelArgs {
    const void *src_k;     // keys read by the kernel
    const void *src_v;
    void *dst_k;           // copy (sel_hist) or partition (sel_partition) destination
    void *dst_v;
    uint64_t *idx_out;     // argselect
    void *cand;            // candidate ord keys, per-block regions of `chunk`
    uint32_t *ccount;      // per-block candidate counts
    uint32_t *partial;     // [blocks][256] digit histograms
    SelState *st;
    uint32_t n;
    uint32_t chunk;        // keys per block in the chunked kernels
    uint32_t nblocks;
    uint32_t shift;        // digit of this round
    uint32_t k;
    uint32_t nan_last;
    uint64_t flip;
    uint32_t copy;         // sel_hist: also copy keys (and values) to dst
};

__device__ __forceinline__ void block_hist_flush(uint32_t *h, uint32_t *dst)
{
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += blockDim.x) dst[i] = h[i];
}

// round 0: top-digit histogram of every key (+ optional copy to scratch)
template <class KD, int PL>
__global__ __launch_bounds__(SEL_BLOCK) void sel_hist_kernel(SelArgs a)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_VAL = PL == PL_V32 || PL == PL_V64;
    __shared__ uint32_t h[256];
    for (int i = threadIdx.x; i < 256; i += SEL_BLOCK) h[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.st->prefix = 0;
        a.st->mask = 0;
        a.st->k_rem = a.k;
        a.st->cl = 0;
        a.st->ce = a.n;
    }
    __syncthreads();
    const U *src = static_cast<const U *>(a.src_k);
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const uint32_t b0 = blockIdx.x * a.chunk, b1 = min(a.n, b0 + a.chunk);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += SEL_BLOCK) {
        const U x = src[i];
        if (a.copy) {
            static_cast<U *>(a.dst_k)[i] = x;
            if constexpr (HAS_VAL) static_cast<V *>(a.dst_v)[i] = static_cast<const V *>(a.src_v)[i];
        }
        atomicAdd(&h[KD::digit(x, flip, nl, a.shift)], 1u);
    }
    block_hist_flush(h, a.partial + size_t(blockIdx.x) * 256);
}

// pick the bucket of rank k_rem in the summed histogram; narrow the state
template <int = 0>
__global__ __launch_bounds__(256) void sel_choose_kernel(SelArgs a)
{
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t s_kr;
    const uint32_t d = threadIdx.x, lane = d & 63, w = d >> 6;
    if (d == 0) s_kr = a.st->k_rem;  // read before the chosen thread rewrites it
    uint32_t tot = 0;
    for (uint32_t b = 0; b < a.nblocks; ++b) tot += a.partial[size_t(b) * 256 + d];
    uint32_t incl = tot;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t i = 0; i < w; ++i) pre += wsum[i];
    const uint32_t excl = pre + incl - tot;
    const uint32_t kr = s_kr;
    if (tot > 0 && excl <= kr && kr < excl + tot) {
        SelState *st = a.st;
        st->prefix |= uint64_t(d) << a.shift;
        st->mask |= uint64_t(0xff) << a.shift;
        st->k_rem = kr - excl;
        st->cl += excl;
        st->ce = tot;
        if (a.shift == 0) {
            st->cursor[0] = 0;
            st->cursor[1] = st->cl;
            st->cursor[2] = st->cl + tot;
        }
    }
}

// round 1: compact the keys of the chosen top bucket, histogram the next digit
template <class KD>
__global__ __launch_bounds__(SEL_BLOCK) void sel_filter_kernel(SelArgs a)
{
    using U = typename KD::U;
    __shared__ uint32_t h[256];
    __shared__ uint32_t cnt;
    for (int i = threadIdx.x; i < 256; i += SEL_BLOCK) h[i] = 0;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const U *src = static_cast<const U *>(a.src_k);
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const U mask = U(a.st->mask), prefix = U(a.st->prefix);
    U *cand = static_cast<U *>(a.cand) + size_t(blockIdx.x) * a.chunk;
    const uint32_t b0 = blockIdx.x * a.chunk, b1 = min(a.n, b0 + a.chunk);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += SEL_BLOCK) {
        const U o = KD::ord(src[i], flip, nl);
        if (U(o & mask) == prefix) {
            cand[atomicAdd(&cnt, 1u)] = o;
            atomicAdd(&h[uint32_t(o >> a.shift) & 0xffu], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) a.ccount[blockIdx.x] = cnt;
    block_hist_flush(h, a.partial + size_t(blockIdx.x) * 256);
}

// rounds >= 2: histogram of the next digit over the matching candidates
template <class KD>
__global__ __launch_bounds__(SEL_BLOCK) void sel_cand_kernel(SelArgs a)
{
    using U = typename KD::U;
    __shared__ uint32_t h[256];
    for (int i = threadIdx.x; i < 256; i += SEL_BLOCK) h[i] = 0;
    __syncthreads();
    const U mask = U(a.st->mask), prefix = U(a.st->prefix);
    const U *cand = static_cast<const U *>(a.cand) + size_t(blockIdx.x) * a.chunk;
    const uint32_t c = a.ccount[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < c; i += SEL_BLOCK) {
        const U o = cand[i];
        if (U(o & mask) == prefix) atomicAdd(&h[uint32_t(o >> a.shift) & 0xffu], 1u);
    }
    block_hist_flush(h, a.partial + size_t(blockIdx.x) * 256);
}

// 3-way split around K; MODE 0: keys (+values) src -> dst; MODE 1: indices
template <class KD, int PL, int MODE>
__global__ __launch_bounds__(SEL_BLOCK) void sel_partition_kernel(SelArgs a)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_VAL = PL == PL_V32 || PL == PL_V64;
    constexpr int ITEMS = SEL_TILE / SEL_BLOCK;
    constexpr int WAVES = SEL_BLOCK / 64;
    __shared__ uint32_t s_wc[WAVES][3];
    __shared__ uint32_t s_base[3];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t tbase = blockIdx.x * SEL_TILE;
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const U K = U(a.st->prefix);
    const U *src = static_cast<const U *>(a.src_k);
    U key[ITEMS];
    V val[ITEMS];
    uint32_t cr[ITEMS];  // class in bits 16.., rank in wave in low bits
    uint32_t wc0 = 0, wc1 = 0, wc2 = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t g = tbase + wave * (64 * ITEMS) + i * 64 + lane;
        uint32_t c = 3;
        if (g < a.n) {
            key[i] = src[g];
            if constexpr (HAS_VAL) val[i] = static_cast<const V *>(a.src_v)[g];
            const U o = KD::ord(key[i], flip, nl);
            c = o < K ? 0u : (o == K ? 1u : 2u);
        }
        const uint64_t b0 = __ballot(c == 0), b1 = __ballot(c == 1), b2 = __ballot(c == 2);
        const uint32_t r = c == 0 ? wc0 + lanes_below(b0) : c == 1 ? wc1 + lanes_below(b1) : wc2 + lanes_below(b2);
        cr[i] = (c << 16) | r;
        wc0 += uint32_t(__popcll(b0));
        wc1 += uint32_t(__popcll(b1));
        wc2 += uint32_t(__popcll(b2));
    }
    if (lane == 0) {
        s_wc[wave][0] = wc0;
        s_wc[wave][1] = wc1;
        s_wc[wave][2] = wc2;
    }
    __syncthreads();
    if (tid < 3) {
        uint32_t s = 0;
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_wc[w][tid];
            s_wc[w][tid] = s;
            s += c;
        }
        s_base[tid] = s ? __hip_atomic_fetch_add(&a.st->cursor[tid], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                        : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t c = cr[i] >> 16;
        if (c < 3) {
            const uint32_t pos = s_base[c] + s_wc[wave][c] + (cr[i] & 0xffffu);
            if constexpr (MODE == 1) {
                a.idx_out[pos] = tbase + wave * (64 * ITEMS) + i * 64 + lane;
            } else {
                static_cast<U *>(a.dst_k)[pos] = key[i];
                if constexpr (HAS_VAL) static_cast<V *>(a.dst_v)[pos] = val[i];
            }
        }
    }
}

template <class KD, int PL>
struct SelPlan {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr bool HAS_VAL = PL == PL_V32 || PL == PL_V64;
    static uint32_t nblocks(size_t n) { return uint32_t(std::min<size_t>(SEL_MAX_BLOCKS, (n + 4095) / 4096)); }
    static uint32_t chunk(size_t n)
    {
        const uint32_t nb = nblocks(n);
        return uint32_t(((n + nb - 1) / nb + 255) / 256 * 256);
    }
    // scratch: copy of keys (+values) for in-place forms, candidates, counts
    static size_t scratch_bytes(size_t n, bool in_place)
    {
        size_t s = align256(size_t(nblocks(n)) * chunk(n) * sizeof(U)) + align256(SEL_MAX_BLOCKS * 4);
        if (in_place) s += align256(n * sizeof(U)) + (HAS_VAL ? align256(n * sizeof(V)) : 0);
        return s;
    }
};

// Enqueue a radix select of rank k.  In place (keys/vals) unless idx_out is
// given (argselect: keys read-only, indices written).
template <class KD, int PL>
int enqueue_select(Workspace *ws, char *scratch, void *keys, void *vals, uint64_t *idx_out, size_t n, size_t k,
                   bool hasnan, bool desc, hipStream_t s)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    using SP = SelPlan<KD, PL>;
    constexpr int P = KD::PASSES;
    const bool in_place = idx_out == nullptr;
    SelArgs a{};
    a.n = uint32_t(n);
    a.k = uint32_t(k);
    a.nblocks = SP::nblocks(n);
    a.chunk = SP::chunk(n);
    a.nan_last = (KD::KIND == Kind::Float) && hasnan;
    a.flip = desc ? ~0ull : 0ull;
    a.st = sel_state(ws);
    a.partial = ws->rsum();
    char *p = scratch;
    a.cand = p;
    p += align256(size_t(a.nblocks) * a.chunk * sizeof(U));
    a.ccount = reinterpret_cast<uint32_t *>(p);
    p += align256(SEL_MAX_BLOCKS * 4);
    char *tmp_k = nullptr, *tmp_v = nullptr;
    if (in_place) {
        tmp_k = p;
        p += align256(n * sizeof(U));
        if (SP::HAS_VAL) tmp_v = p;
    }
    // round 0
    a.src_k = keys;
    a.src_v = vals;
    a.dst_k = tmp_k;
    a.dst_v = tmp_v;
    a.copy = in_place;
    a.shift = 8 * (P - 1);
    hipLaunchKernelGGL((sel_hist_kernel<KD, PL>), dim3(a.nblocks), dim3(SEL_BLOCK), 0, s, a);
    XSS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(sel_choose_kernel<0>, dim3(1), dim3(256), 0, s, a);
    XSS_HIP_CHECK(hipGetLastError());
    const void *src = in_place ? static_cast<const void *>(tmp_k) : keys;
    for (int r = 1; r < P; ++r) {
        a.shift = 8 * (P - 1 - r);
        a.src_k = src;
        if (r == 1) hipLaunchKernelGGL((sel_filter_kernel<KD>), dim3(a.nblocks), dim3(SEL_BLOCK), 0, s, a);
        else hipLaunchKernelGGL((sel_cand_kernel<KD>), dim3(a.nblocks), dim3(SEL_BLOCK), 0, s, a);
        XSS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(sel_choose_kernel<0>, dim3(1), dim3(256), 0, s, a);
        XSS_HIP_CHECK(hipGetLastError());
    }
    // partition
    const uint32_t tiles = uint32_t((n + SEL_TILE - 1) / SEL_TILE);
    a.src_k = src;
    a.src_v = in_place ? tmp_v : nullptr;
    a.dst_k = keys;
    a.dst_v = vals;
    a.idx_out = idx_out;
    if (in_place) hipLaunchKernelGGL((sel_partition_kernel<KD, PL, 0>), dim3(tiles), dim3(SEL_BLOCK), 0, s, a);
    else hipLaunchKernelGGL((sel_partition_kernel<KD, PL, 1>), dim3(tiles), dim3(SEL_BLOCK), 0, s, a);
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

// kind: 0 select, 1 partial sort.  PL: PL_NONE (keys), PL_V32/PL_V64 (kv), PL_IDX (argselect)
template <class KD, int PL>
int run_select(int kind, void *keys, void *vals, uint64_t *idx_out, size_t n, size_t k, bool hasnan, bool desc,
               hipStream_t s)
{
    constexpr int SPL = PL == PL_IDX ? PL_NONE : PL;
    if (n > MAX_N) return set_error(XSS_ERR_INVALID, "n = %zu exceeds the per-call limit %u", n, MAX_N);
    if (kind == 1 && k == 0) return XSS_OK;  // partial sort of 0 elements (src/xss-common-qsort.h:757)
    if (n <= 1) {
        if (PL == PL_IDX && n == 1) return run_sort<KD, PL_IDX>(keys, nullptr, nullptr, nullptr, idx_out, 1, hasnan,
                                                                 desc, false, s);
        return XSS_OK;
    }
    const size_t kk = kind == 1 ? k - 1 : k;
    if (kk >= n) return set_error(XSS_ERR_INVALID, "k = %zu out of range for n = %zu", k, n);
    if (!keys || (PL == PL_IDX && !idx_out) || ((PL == PL_V32 || PL == PL_V64) && !vals))
        return set_error(XSS_ERR_INVALID, "null pointer argument");
    // small arrays: one LDS sort is a valid selection (NaN bits kept)
    if (SortPlan<KD, PL>::small(n) && kind == 0) {
        if (PL == PL_IDX) return run_sort<KD, PL_IDX>(keys, nullptr, nullptr, nullptr, idx_out, n, hasnan, desc, false, s);
        return run_sort<KD, PL>(keys, keys, vals, vals, nullptr, n, hasnan, desc, false, s);
    }
    const bool in_place = PL != PL_IDX;
    size_t sb = SelPlan<KD, SPL>::scratch_bytes(n, in_place);
    if (kind == 1) sb = std::max(sb, SortPlan<KD, SPL>::scratch_bytes(kk));
    Workspace *ws = nullptr;
    std::unique_lock<std::mutex> lock;
    int rc = acquire_workspace(s, sb, &ws, &lock);
    if (rc) return rc;
    if (SortPlan<KD, PL>::small(n)) {
        // partial sort of a small array: sort it whole with qsort semantics
        rc = enqueue_sort<KD, SPL>(ws, ws->data, keys, keys, vals, vals, nullptr, n, hasnan, desc, false, s);
        if (rc) return rc;
    } else {
        rc = enqueue_select<KD, SPL>(ws, ws->data, keys, vals, PL == PL_IDX ? idx_out : nullptr, n, kk, hasnan,
                                     desc, s);
        if (rc) return rc;
    }
    if (kind == 1 && kk > 1) {
        // xss_partial_qsort: qsort of the first k-1 (NaNs there become canonical)
        rc = enqueue_sort<KD, SPL>(ws, ws->data, keys, keys, vals, vals, nullptr, kk, hasnan, desc, PL == PL_NONE, s);
    }
    return rc;
}

}  // namespace xss
