// xss_select.hpp -- radix select: qselect, keyvalue_select, argselect and
// the partial sorts built on them.
//
// The reference selects with a quickselect that partitions around sampled
// pivots and recurses into one side (qselect_ src/xss-common-qsort.h:607-649,
// kvselect_ src/xss-common-keyvaluesort.hpp:519-562, argselect_
// src/xss-common-argsort.h:551-585).  Here the k-th key K is found by digit
// histograms, most significant digit first, entirely on the device (no host
// round trip).  The array is cut into B contiguous blocks (one workgroup
// each, the same cut in every kernel):
//
//   sel_init      clear the state and the per-block counters
//   round 0       histogram of the top digit of every order key, per block
//                 and summed; AND / OR of all order keys (digits that every
//                 key shares need no round); copy of keys / values to
//                 scratch for the in-place forms
//   sel_choose    pick the digit bucket holding rank k; fix that digit of K,
//                 and every following digit all keys share; decide whether
//                 the next round compacts the keys still matching K's known
//                 digits into per-block candidate lists (once they are at
//                 most n/16 of the keys)
//   round r       histogram of the next digit over the matching keys (whole
//                 blocks or candidate lists); each block first folds the
//                 previous round's counts below the chosen digit into its
//                 "keys below K" total
//   sel_bases     per block: keys below K, equal to K, above K, and their
//                 exclusive scans over blocks
//   sel_partition every block writes its keys, in input order, to
//                 [< K | == K | > K] at its own bases: no atomics, so the
//                 result is deterministic
//
// The arrangement inside the three classes is unspecified by the reference
// (its tests check only the partition property, tests/test-qsort-common.h:
// 52-93).  NaNs (hasnan) rank above +inf, so they end up at the back
// (ascending) or front (descending) with their bits kept, as
// move_nans_to_end/start_of_array do (src/xss-common-qsort.h:137-180).
#pragma once
#include "xss_sort.hpp"

namespace xss {

constexpr int SEL_BLOCK = 512;
constexpr int SEL_ITEMS = 8;
constexpr int SEL_TILE = SEL_BLOCK * SEL_ITEMS;  // partition tile
constexpr int SEL_MAX_BLOCKS = 2048;            // [blocks][256] histograms fit the 4 MiB range tables
constexpr int SEL_BASES_BLOCK = 1024;

struct SelState {
    uint64_t prefix;     // order-key bits of K fixed so far
    uint64_t mask;       // which bits are fixed
    uint64_t all_and;    // AND / OR of every order key (round 0)
    uint64_t all_or;
    uint32_t k_rem;      // rank of K among the keys matching prefix
    uint32_t cl;         // keys below K found so far (all blocks)
    uint32_t ce;         // keys matching prefix (all == K once done)
    int32_t shift;       // digit of the next round; < 0: K is known
    uint32_t filtered;   // candidate lists hold the keys matching prefix
    uint32_t filter_now; // the next round compacts the matching keys
    int32_t pend;        // histogram buffer of the last round
    uint32_t pend_d;     // its chosen digit
};
static_assert(sizeof(SelState) <= 128, "SelState must fit the control block");

inline SelState *sel_state(Workspace *ws) { return reinterpret_cast<SelState *>(ws->meta + 128); }

struct SelArgs {
    const void *src_k;     // caller's keys (read by every round, and by the partition of argselect)
    const void *src_v;     // caller's values (round 0 copy)
    void *copy_k;          // in-place forms: scratch copy of keys / values (partition source)
    void *copy_v;
    void *dst_k;           // partition destination
    void *dst_v;
    uint64_t *idx_out;     // argselect
    void *cand;            // candidate order keys, block b at cand + b * chunk
    uint32_t *ccount;      // [blocks] candidates per block
    uint32_t *part0;       // [blocks][256] per-block digit counts, rounds 0, 2, ...
    uint32_t *part1;       //                                       rounds 1, 3, ...
    uint32_t *tot;         // [2][256] summed digit counts
    uint32_t *below;       // [blocks] keys below K per block
    uint32_t *bases;       // [3][blocks] output base of each class per block
    SelState *st;
    uint32_t n;
    uint32_t chunk;        // keys per block
    uint32_t nblocks;
    uint32_t k;
    uint32_t nan_last;
    uint32_t round;
    uint64_t flip;
    uint32_t copy;         // in-place forms: round 0 copies keys (and values)
    uint32_t first_shift;  // top digit
};

template <int = 0>
__global__ __launch_bounds__(256) void sel_init_kernel(SelArgs a)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        SelState *st = a.st;
        st->prefix = 0;
        st->mask = 0;
        st->all_and = ~0ull;
        st->all_or = 0;
        st->k_rem = a.k;
        st->cl = 0;
        st->ce = a.n;
        st->shift = int32_t(a.first_shift);
        st->filtered = 0;
        st->filter_now = 0;
        st->pend = -1;
        st->pend_d = 0;
    }
    if (i < 512) a.tot[i] = 0;
    if (i < a.nblocks) a.below[i] = 0;
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *red)
{
    constexpr int WAVES = SEL_BLOCK / 64;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) s += red[w];
    return s;
}

// One order key into the block histogram: lanes of the wave holding the same
// digit add once (wave64 match-any), so heavy duplicates do not serialise
// the LDS atomics.  Inactive or non-matching lanes pass ok = false.
__device__ __forceinline__ void sel_count(uint32_t *h, uint32_t d, bool ok)
{
    const uint64_t okm = __ballot(ok);
    if (okm == 0) return;
    const uint64_t peers = match_any8(d) & okm;
    if (ok && lanes_below(peers) == 0) atomicAdd(&h[d], uint32_t(__popcll(peers)));
}

// Rounds: histogram of digit `shift` of the order keys matching prefix/mask.
// R0: first round (no prefix; AND/OR; copy).  The block reads its own chunk
// of the caller's keys, or its candidate list once the keys were compacted.
template <class KD, int PL, bool R0>
__global__ __launch_bounds__(SEL_BLOCK) void sel_round_kernel(SelArgs a)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_VAL = PL == PL_V32 || PL == PL_V64;
    constexpr int WAVES = SEL_BLOCK / 64;
    __shared__ uint32_t h[256];
    __shared__ uint32_t red[WAVES];
    __shared__ U s_and[WAVES], s_or[WAVES];
    __shared__ uint32_t s_cnt;
    SelState *st = a.st;
    const int shift = R0 ? int(a.first_shift) : st->shift;
    if (shift < 0) return;  // K already known
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const uint32_t buf = a.round & 1;
    uint32_t *part = buf ? a.part1 : a.part0;
    for (uint32_t i = tid; i < 256; i += SEL_BLOCK) h[i] = 0;
    if (tid == 0) s_cnt = 0;
    if (!R0) {
        // fold the previous round: keys of this block below its chosen digit
        const uint32_t *prev = (st->pend ? a.part1 : a.part0) + size_t(b) * 256;
        const uint32_t pd = st->pend_d;
        const uint32_t c = (tid < pd) ? prev[tid] : 0u;
        const uint32_t s = block_sum_u32(c, red);
        if (tid == 0) a.below[b] += s;
    }
    __syncthreads();
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const U mask = R0 ? U(0) : U(st->mask), prefix = R0 ? U(0) : U(st->prefix);
    const bool from_cand = !R0 && st->filtered;
    const bool compact = !R0 && !from_cand && st->filter_now;
    const uint32_t b0 = min(a.n, b * a.chunk), b1 = min(a.n, b0 + a.chunk);
    U *cand = static_cast<U *>(a.cand) + size_t(b) * a.chunk;
    U v_and = U(~U(0)), v_or = U(0);

    auto visit = [&](U o, bool valid) {
        const bool ok = valid && U(o & mask) == prefix;
        sel_count(h, uint32_t(o >> shift) & 0xffu, ok);
        if (R0 && valid) {
            v_and &= o;
            v_or |= o;
        }
        if (compact) {
            const uint64_t m = __ballot(ok);
            if (m) {
                const uint64_t act = __ballot(1);
                uint32_t base = 0;
                if (lanes_below(act) == 0) base = atomicAdd(&s_cnt, uint32_t(__popcll(m)));
                base = __builtin_amdgcn_readfirstlane(base);
                const uint32_t slot = base + lanes_below(m);
                if (ok && slot < a.chunk) cand[slot] = o;  // at most b1 - b0 <= chunk keys match
            }
        }
    };

    if (from_cand) {
        const uint32_t c = a.ccount[b];
        for (uint32_t i = tid; i < c + SEL_BLOCK - 1; i += SEL_BLOCK) {
            if (i - tid >= c) break;  // uniform exit
            const bool valid = i < c;
            visit(valid ? cand[i] : U(0), valid);
        }
    } else {
        const U *keys = static_cast<const U *>(a.src_k);
        constexpr int VEC = 16 / sizeof(U);
        // keys before the first 16-byte boundary are read one by one
        const uint32_t mis = uint32_t(reinterpret_cast<uintptr_t>(keys + b0) & 15u);
        const uint32_t head = min(b1 - b0, uint32_t(((16u - mis) & 15u) / sizeof(U)));
        const uint32_t v0 = b0 + head;
        const uint32_t nvec = (b1 - v0) / VEC;
        const uint4 *kv = reinterpret_cast<const uint4 *>(keys + v0);
        U *ck = static_cast<U *>(a.copy_k);
        // the copy is stored as 16-byte vectors when it shares the keys' alignment
        const bool cvec = ((reinterpret_cast<uintptr_t>(ck) ^ reinterpret_cast<uintptr_t>(keys)) & 15u) == 0;
        auto do_vec = [&](const uint4 &q, uint32_t vi) {
            const U *e = reinterpret_cast<const U *>(&q);
            if (R0 && a.copy) {
                if (cvec) {
                    reinterpret_cast<uint4 *>(ck + v0)[vi] = q;
                } else {
#pragma unroll
                    for (int j = 0; j < VEC; ++j) ck[v0 + vi * VEC + j] = e[j];
                }
            }
#pragma unroll
            for (int j = 0; j < VEC; ++j) visit(KD::ord(e[j], flip, nl), true);
        };
        // whole blocks of UNR x SEL_BLOCK vectors, UNR loads in flight per thread
        constexpr uint32_t UNR = 4;
        const uint32_t nsuper = nvec / (UNR * SEL_BLOCK);
        for (uint32_t it = 0; it < nsuper; ++it) {
            const uint32_t vb = it * UNR * SEL_BLOCK + tid;
            uint4 q[UNR];
#pragma unroll
            for (uint32_t u = 0; u < UNR; ++u) q[u] = kv[vb + u * SEL_BLOCK];
#pragma unroll
            for (uint32_t u = 0; u < UNR; ++u) do_vec(q[u], vb + u * SEL_BLOCK);
        }
        // remaining vectors, then the unaligned head and the tail; every lane
        // of the block stays in these loops until the last row (the counting
        // uses wave ballots)
        const uint32_t vdone = nsuper * UNR * SEL_BLOCK;
        const uint32_t nv_left = nvec - vdone;
        for (uint32_t i = tid; i < nv_left + SEL_BLOCK - 1; i += SEL_BLOCK) {
            if (i - tid >= nv_left) break;
            if (i < nv_left) {
                do_vec(kv[vdone + i], vdone + i);
            } else {
#pragma unroll
                for (int j = 0; j < VEC; ++j) visit(U(0), false);
            }
        }
        const uint32_t t0 = v0 + nvec * VEC;
        const uint32_t rest = head + (b1 - t0);
        for (uint32_t r = tid; r < rest + SEL_BLOCK - 1; r += SEL_BLOCK) {
            if (r - tid >= rest) break;
            const bool valid = r < rest;
            const uint32_t i = r < head ? b0 + r : t0 + (r - head);
            U x = U(0);
            if (valid) {
                x = keys[i];
                if (R0 && a.copy) ck[i] = x;
            }
            visit(valid ? KD::ord(x, flip, nl) : U(0), valid);
        }
        if constexpr (R0 && HAS_VAL) {
            if (a.copy) {
                const V *sv = static_cast<const V *>(a.src_v);
                V *cv = static_cast<V *>(a.copy_v);
                for (uint32_t i = b0 + tid; i < b1; i += SEL_BLOCK) cv[i] = sv[i];
            }
        }
    }
    if constexpr (R0) {
        // AND / OR of the block's order keys
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            v_and &= U(__shfl_xor(v_and, off));
            v_or |= U(__shfl_xor(v_or, off));
        }
        if ((tid & 63) == 0) {
            s_and[tid >> 6] = v_and;
            s_or[tid >> 6] = v_or;
        }
    }
    __syncthreads();
    if constexpr (R0) {
        if (tid == 0) {
            U x = U(~U(0)), y = U(0);
            for (int w = 0; w < WAVES; ++w) {
                x &= s_and[w];
                y |= s_or[w];
            }
            // bits above the key width stay set in all_and and clear in all_or
            unsigned long long xa = (unsigned long long)x;
            if constexpr (sizeof(U) < 8) xa |= ~0ull << (8 * sizeof(U));
            atomicAnd(reinterpret_cast<unsigned long long *>(&st->all_and), xa);
            atomicOr(reinterpret_cast<unsigned long long *>(&st->all_or), (unsigned long long)y);
        }
    }
    for (uint32_t i = tid; i < 256; i += SEL_BLOCK) {
        const uint32_t c = h[i];
        part[size_t(b) * 256 + i] = c;
        if (c) atomicAdd(&a.tot[buf * 256 + i], c);
    }
    if (compact && tid == 0) a.ccount[b] = s_cnt;
}

// pick the bucket of rank k_rem; fix the digit and every following digit all
// keys share; decide the next round's source
template <class KD>
__global__ __launch_bounds__(256) void sel_choose_kernel(SelArgs a)
{
    using U = typename KD::U;
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t s_d, s_excl, s_tot;
    SelState *st = a.st;
    const int shift = st->shift;
    if (shift < 0) return;
    const uint32_t d = threadIdx.x, lane = d & 63, w = d >> 6;
    const uint32_t buf = a.round & 1;
    const uint32_t kr = st->k_rem;
    const uint32_t tot = a.tot[buf * 256 + d];
    a.tot[buf * 256 + d] = 0;  // this buffer is summed into again two rounds later
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
    if (d == 0) {
        s_d = 0;
        s_excl = 0;
        s_tot = 0;
    }
    __syncthreads();
    if (tot > 0 && excl <= kr && kr < excl + tot) {  // exactly one digit: k_rem < keys matching prefix
        s_d = d;
        s_excl = excl;
        s_tot = tot;
    }
    __syncthreads();
    if (d == 0) {
        const uint32_t D = s_d;
        uint64_t prefix = st->prefix | (uint64_t(D) << shift);
        uint64_t mask = st->mask | (uint64_t(0xff) << shift);
        const uint32_t ce = s_tot;
        st->k_rem = kr - s_excl;
        st->cl += s_excl;
        st->ce = ce;
        st->pend = int32_t(buf);
        st->pend_d = D;
        const uint32_t filtered = st->filtered | st->filter_now;
        st->filtered = filtered;
        // digits below that every key shares are K's digits too
        const uint64_t diff = st->all_and ^ st->all_or;
        int s = shift - 8;
        while (s >= 0 && ((diff >> s) & 0xffu) == 0) {
            prefix |= st->all_and & (uint64_t(0xff) << s);
            mask |= uint64_t(0xff) << s;
            s -= 8;
        }
        st->prefix = prefix;
        st->mask = mask;
        st->shift = s;
        st->filter_now = (!filtered && s >= 0 && ce <= a.n / 16) ? 1u : 0u;
    }
}

// per-block class counts (below / equal / above K) and their exclusive scans
template <int = 0>
__global__ __launch_bounds__(SEL_BASES_BLOCK) void sel_bases_kernel(SelArgs a)
{
    __shared__ uint32_t s_c[3][SEL_MAX_BLOCKS];
    __shared__ uint32_t s_w[3][SEL_BASES_BLOCK / 64];
    const SelState *st = a.st;
    const uint32_t *last = st->pend ? a.part1 : a.part0;
    const uint32_t pd = st->pend_d;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t b = tid; b < a.nblocks; b += SEL_BASES_BLOCK) {
        const uint4 *row = reinterpret_cast<const uint4 *>(last + size_t(b) * 256);
        uint32_t lo = 0;
        for (uint32_t q = 0; q < pd / 4; ++q) {
            const uint4 r = row[q];
            lo += r.x + r.y + r.z + r.w;
        }
        for (uint32_t dd = pd & ~3u; dd < pd; ++dd) lo += last[size_t(b) * 256 + dd];
        const uint32_t eq = last[size_t(b) * 256 + pd];
        const uint32_t below = a.below[b] + lo;
        const uint32_t b0 = min(a.n, b * a.chunk), b1 = min(a.n, b0 + a.chunk);
        s_c[0][b] = below;
        s_c[1][b] = eq;
        s_c[2][b] = (b1 - b0) - below - eq;
    }
    __syncthreads();
    // exclusive scan of each class over blocks: every thread owns a run of
    // consecutive blocks
    constexpr uint32_t PER = SEL_MAX_BLOCKS / SEL_BASES_BLOCK;
    const uint32_t j0 = tid * PER;
    uint32_t t[3] = {0, 0, 0};
    for (uint32_t j = j0; j < j0 + PER && j < a.nblocks; ++j)
        for (int c = 0; c < 3; ++c) t[c] += s_c[c][j];
    uint32_t incl[3];
    for (int c = 0; c < 3; ++c) {
        incl[c] = t[c];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t x = __shfl_up(incl[c], off);
            if (lane >= uint32_t(off)) incl[c] += x;
        }
        if (lane == 63) s_w[c][wave] = incl[c];
    }
    __syncthreads();
    // class bases: [0, cl) below, [cl, cl + ce) equal, the rest above
    const uint32_t cbase[3] = {0u, st->cl, st->cl + st->ce};
    for (int c = 0; c < 3; ++c) {
        uint32_t run = cbase[c] + incl[c] - t[c];
        for (uint32_t w = 0; w < wave; ++w) run += s_w[c][w];
        for (uint32_t j = j0; j < j0 + PER && j < a.nblocks; ++j) {
            a.bases[c * a.nblocks + j] = run;
            run += s_c[c][j];
        }
    }
}

// 3-way split around K, block by block in input order.  MODE 0: keys
// (+values) copy -> caller's arrays; MODE 1: indices (argselect).
template <class KD, int PL, int MODE>
__global__ __launch_bounds__(SEL_BLOCK) void sel_partition_kernel(SelArgs a)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_VAL = PL == PL_V32 || PL == PL_V64;
    constexpr int WAVES = SEL_BLOCK / 64;
    __shared__ uint32_t s_wc[WAVES][3];
    __shared__ uint32_t s_run[3];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t b = blockIdx.x;
    const uint32_t b0 = min(a.n, b * a.chunk), b1 = min(a.n, b0 + a.chunk);
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const U K = U(a.st->prefix);
    const U *src = static_cast<const U *>(MODE == 1 ? a.src_k : a.copy_k);
    const V *srcv = static_cast<const V *>(a.copy_v);
    if (tid < 3) s_run[tid] = a.bases[tid * a.nblocks + b];
    for (uint32_t t0 = b0; t0 < b1; t0 += SEL_TILE) {
        U key[SEL_ITEMS];
        V val[SEL_ITEMS];
        uint32_t cr[SEL_ITEMS];  // class << 16 | rank among the wave's keys of that class
        uint32_t wc0 = 0, wc1 = 0, wc2 = 0;
#pragma unroll
        for (int i = 0; i < SEL_ITEMS; ++i) {
            const uint32_t g = t0 + wave * (64 * SEL_ITEMS) + i * 64 + lane;
            uint32_t c = 3;
            if (g < b1) {
                key[i] = src[g];
                if constexpr (HAS_VAL) val[i] = srcv[g];
                const U o = KD::ord(key[i], flip, nl);
                c = o < K ? 0u : (o == K ? 1u : 2u);
            }
            const uint64_t m0 = __ballot(c == 0), m1 = __ballot(c == 1), m2 = __ballot(c == 2);
            const uint32_t r = c == 0 ? wc0 + lanes_below(m0) : c == 1 ? wc1 + lanes_below(m1)
                                                               : wc2 + lanes_below(m2);
            cr[i] = (c << 16) | r;
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
        if (tid < 3) {
            uint32_t s = s_run[tid];
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t c = s_wc[w][tid];
                s_wc[w][tid] = s;
                s += c;
            }
            s_run[tid] = s;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SEL_ITEMS; ++i) {
            const uint32_t c = cr[i] >> 16;
            const uint32_t pos = c < 3 ? s_wc[wave][c] + (cr[i] & 0xffffu) : a.n;
            if (pos < a.n) {  // the class bases partition [0, n)
                if constexpr (MODE == 1) {
                    a.idx_out[pos] = t0 + wave * (64 * SEL_ITEMS) + i * 64 + lane;
                } else {
                    static_cast<U *>(a.dst_k)[pos] = key[i];
                    if constexpr (HAS_VAL) static_cast<V *>(a.dst_v)[pos] = val[i];
                }
            }
        }
        __syncthreads();  // s_wc is rewritten by the next tile
    }
}

template <class KD, int PL>
struct SelPlan {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr bool HAS_VAL = PL == PL_V32 || PL == PL_V64;
    // B blocks of `chunk` keys (a multiple of 256); every block holds at least one key
    static uint32_t chunk(size_t n)
    {
        const size_t nb = std::min<size_t>(SEL_MAX_BLOCKS, std::max<size_t>(1, (n + SEL_TILE - 1) / SEL_TILE));
        return uint32_t(((n + nb - 1) / nb + 255) / 256 * 256);
    }
    static uint32_t nblocks(size_t n) { return uint32_t((n + chunk(n) - 1) / chunk(n)); }
    // scratch: candidates, counters, bases, copy of keys (+values) for in-place forms
    static size_t scratch_bytes(size_t n, bool in_place)
    {
        const size_t nb = nblocks(n);
        size_t s = align256(nb * chunk(n) * sizeof(U)) + align256(nb * 4) + align256(512 * 4) +
                   align256(nb * 4) + align256(3 * nb * 4);
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
    a.part0 = ws->rsum();
    a.part1 = ws->rpre();
    a.first_shift = 8 * (P - 1);
    a.copy = in_place;
    char *p = scratch;
    a.cand = p;
    p += align256(size_t(a.nblocks) * a.chunk * sizeof(U));
    a.ccount = reinterpret_cast<uint32_t *>(p);
    p += align256(size_t(a.nblocks) * 4);
    a.tot = reinterpret_cast<uint32_t *>(p);
    p += align256(512 * 4);
    a.below = reinterpret_cast<uint32_t *>(p);
    p += align256(size_t(a.nblocks) * 4);
    a.bases = reinterpret_cast<uint32_t *>(p);
    p += align256(size_t(3) * a.nblocks * 4);
    if (in_place) {
        a.copy_k = p;
        p += align256(n * sizeof(U));
        if (SP::HAS_VAL) a.copy_v = p;
    }
    a.src_k = keys;
    a.src_v = vals;
    a.dst_k = keys;
    a.dst_v = vals;
    a.idx_out = idx_out;
    KernelTimer t("select", s);
    hipLaunchKernelGGL(sel_init_kernel<0>, dim3((std::max<uint32_t>(a.nblocks, 512) + 255) / 256), dim3(256), 0, s,
                       a);
    XSS_HIP_CHECK(hipGetLastError());
    for (int r = 0; r < P; ++r) {
        a.round = uint32_t(r);
        if (r == 0) hipLaunchKernelGGL((sel_round_kernel<KD, PL, true>), dim3(a.nblocks), dim3(SEL_BLOCK), 0, s, a);
        else hipLaunchKernelGGL((sel_round_kernel<KD, PL, false>), dim3(a.nblocks), dim3(SEL_BLOCK), 0, s, a);
        XSS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(sel_choose_kernel<KD>, dim3(1), dim3(256), 0, s, a);
        XSS_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(sel_bases_kernel<0>, dim3(1), dim3(SEL_BASES_BLOCK), 0, s, a);
    XSS_HIP_CHECK(hipGetLastError());
    if (in_place) hipLaunchKernelGGL((sel_partition_kernel<KD, PL, 0>), dim3(a.nblocks), dim3(SEL_BLOCK), 0, s, a);
    else hipLaunchKernelGGL((sel_partition_kernel<KD, PL, 1>), dim3(a.nblocks), dim3(SEL_BLOCK), 0, s, a);
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
