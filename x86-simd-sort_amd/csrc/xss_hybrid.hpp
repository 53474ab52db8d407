// xss_hybrid.hpp -- MSD-first sort of one large array (keys only, key-value,
// or argsort).
//
// An LSD sort of 4-byte keys streams the array through HBM four times (each
// pass: a histogram read, then a read and a write).  For large arrays this
// path does two passes and then finishes every bucket on chip:
//
//   pass A   LSD pass on the top digit over the whole array (keys -> scratch);
//            its histogram read also takes the AND / OR of all keys, and a
//            plan made on the device moves passes A and B down to the
//            highest digits on which keys differ (keys below 2^32 in a
//            64-bit type sort like 32-bit keys)
//   table    the 256 top-digit buckets are cut into ranges (built on the
//            device from pass A's digit totals)
//   pass B   segmented pass on the next digit inside every top bucket
//            (scratch -> keys); its per-bucket digit totals are the exact
//            sizes of the 65536 buckets of the two MSD digits
//   buckets  bucket starts; buckets larger than one workgroup's LDS are listed
//   local    one workgroup per bucket sorts it by the remaining low digits in
//            LDS (an 8-bit LSD radix per digit, stable) and writes it back
//   large    listed buckets (skewed or duplicate-heavy data) get segmented
//            LSD passes on the low digits (keys -> scratch -> keys); a
//            bucket whose keys are all equal is left in place
//
// Everything is enqueued on the stream without a host round trip: the range
// tables and the plan are built by kernels, passes over the tables hand
// their ranges out from counters to a bounded grid, and passes the plan
// does not need exit at once.  The call stays asynchronous and capturable.
// Passes whose input order does not matter rank without keeping it, except
// for argsort, whose every pass is stable.
#pragma once
#include <atomic>
#include <type_traits>

#include "xss_count.hpp"
#include "xss_sample.hpp"
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
// 4-byte keys sorted in LDS as 12 unstable + 4 stable bits (1) or 8 + 8 (0)
#ifndef XSS_HYB_SPLIT
#define XSS_HYB_SPLIT 1
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
    uint32_t *wctr;             // work-stealing counters of the segmented passes (cleared by top_ranges)
    uint32_t nwctr;
    uint64_t *seg_andor;        // [lcap][2] AND / OR of each large bucket's raw key bits
    const uint32_t *plan;       // the device plan (hyb_plan_kernel)
    uint32_t *lrfirst;          // [lcap + 1] first range of each large bucket (large_ranges)
    HybTable tl;                // ranges of the low passes after the first: large buckets not constant
    uint32_t *lrfirst2;         // [lcap + 1] first range of each large bucket in tl
};

// The sort's plan, made on the device from the AND / OR of all order keys
// (pass A's histogram read): digits above the highest one on which keys
// differ are skipped in pairs, so that an even count of low digits remains
// and the keys end in the caller's buffer.  plan[0] = shift of pass A's
// digit dA, plan[1] = shift of pass B's digit dA - 1, plan[2] = the low
// digits (dA - 1) sorted in LDS or by the large-bucket passes, plan[3] = 1
// if dA is not the top digit (pass A's histogram is then taken again),
// plan[4] = 1 if no low digit is left (dA = 1).  Keys that differ only in
// digit 0 (or not at all) take dA = 1.
template <int P>
__global__ __launch_bounds__(256) void hyb_plan_kernel(const uint64_t *slots, uint32_t nslots, uint64_t *andor,
                                                       uint32_t *plan)
{
    fold_andor(slots, nslots, andor);  // pass A's per-workgroup AND / OR
    if (threadIdx.x != 0) return;
    const uint64_t diff = andor[0] ^ andor[1];
    int t = -1;
    for (int d = P - 1; d >= 0; --d)
        if ((diff >> (8 * d)) & 0xffu) {
            t = d;
            break;
        }
    int dA = t < 1 ? 1 : ((P - 1 - t) % 2 == 0 ? t : t + 1);
    plan[0] = uint32_t(8 * dA);
    plan[1] = uint32_t(8 * (dA - 1));
    plan[2] = uint32_t(dA - 1);
    plan[3] = dA != P - 1 ? 1u : 0u;
    plan[4] = dA == 1 ? 1u : 0u;
}

// A plan without low digits (plan[4]): after passes A and B every bucket
// holds equal keys, so the sort is done except for what the last pass
// writes: argsort's indices (from the payload, in order) and qsort(hasnan)'s
// canonical NaNs.  Grid-stride over the array; exits unless plan[4] is set.
template <class KD, int PL>
__global__ __launch_bounds__(256) void hyb_finish_kernel(const uint32_t *plan, void *keys_v, const void *vals_v,
                                                         uint64_t *idx_out, uint64_t n, uint32_t canon)
{
    using U = typename KD::U;
    if (gate_closed(plan + 4, 0)) return;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
        if constexpr (PL == PL_IDX) {
            idx_out[i] = uint64_t(static_cast<const uint32_t *>(vals_v)[i]);
        } else if constexpr (KD::KIND == Kind::Float) {
            U *k = static_cast<U *>(keys_v);
            if (canon && KD::is_nan(k[i])) k[i] = KD::QNAN;
        }
    }
}

// ---------------------------------------------------------------- two-digit histogram of pass A
//
// Pass A's histogram read counts the keys of its range by both MSD digits
// at once (65536 bins of 16 bits in LDS), so that pass B, whose segments are
// the top buckets, needs no histogram read of its own: the keys of pass-A
// range r with top digit d1 land in one contiguous run of Y, and a pass-B
// range made of consecutive such runs has, as its digit counts, the sum of
// their rows.  Pass A uses H2_RANGES ranges (one workgroup per CU).
#ifndef XSS_H2_RANGES
#define XSS_H2_RANGES 256  // pass-A ranges (build-time knob)
#endif
constexpr uint32_t H2_RANGES = XSS_H2_RANGES;
constexpr uint32_t H2_BLOCK = 1024;
constexpr uint32_t H2_SPILL = 0x4000;  // a 16-bit bin reaching this moves it to the spill log

struct H2Args {
    const void *keys;
    uint32_t n;
    uint32_t range_keys;
    uint32_t ranges;
    uint32_t shift_hi;          // shift of the high digit (null shift_dev: the top digit)
    const uint32_t *shift_dev;  // the plan's pass-A digit
    const uint32_t *gate;       // runs only if *gate > 0 (null: always)
    uint64_t flip;
    uint32_t nan_last;
    uint64_t *andor_slots;      // first launch: per-workgroup AND / OR of the order keys
    uint32_t *res;              // [ranges][32768]: two 16-bit bins per word, bin = (hi << 8) | lo
    uint32_t *spill_cnt;        // [ranges]
    uint32_t *spill_bin;        // [ranges][spill_cap]: bins that moved H2_SPILL keys out of LDS
    uint32_t spill_cap;
    uint32_t *rsum;             // [256][ranges] high-digit counts (pass A's range scan)
};

template <class KD>
__global__ __launch_bounds__(H2_BLOCK) void hyb_h2_kernel(H2Args a)
{
    using U = typename KD::U;
    __shared__ uint32_t h[32768];
    __shared__ unsigned long long s_ao[2];
    if (a.gate && __builtin_amdgcn_readfirstlane(*a.gate) == 0) return;
    const uint32_t shift_hi = dev_shift(a.shift_dev, a.shift_hi);
    const int shift = int(shift_hi) - 8;  // the 16-bit bin: high digit and the one below it
    const uint32_t r = blockIdx.x, tid = threadIdx.x;
    for (uint32_t i = tid; i < 32768; i += H2_BLOCK) h[i] = 0;
    if (tid == 0) {
        __hip_atomic_store(a.spill_cnt + r, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ao[0] = ~0ull;
        s_ao[1] = 0ull;
    }
    __syncthreads();
    const bool track = a.andor_slots != nullptr;
    const U flip = U(a.flip);
    const bool nl = a.nan_last != 0;
    const U *keys = static_cast<const U *>(a.keys);
    const uint32_t r0 = min(a.n, r * a.range_keys), r1 = uint32_t(min<uint64_t>(a.n, uint64_t(r0) + a.range_keys));
    U t_and = U(~U(0)), t_or = U(0);
    // an add of c to bin b found `word`: crossing H2_SPILL moves H2_SPILL
    // keys out of the bin into the spill log.  One move per bin is pending
    // at a time and other waves add at most a few thousand meanwhile, so a
    // bin never carries into its neighbour.
    auto add = [&](uint32_t b, uint32_t c) {
        const uint32_t sh = (b & 1u) * 16u;
        const uint32_t w = atomicAdd(&h[b >> 1], c << sh);
        const uint32_t old = (w >> sh) & 0xffffu;
        if (old < H2_SPILL && old + c >= H2_SPILL) {
            atomicSub(&h[b >> 1], H2_SPILL << sh);
            const uint32_t i = atomicAdd(&a.spill_cnt[r], 1u);
            if (i < a.spill_cap)
                __hip_atomic_store(a.spill_bin + size_t(r) * a.spill_cap + i, b, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    auto count = [&](U x) {
        const U o = KD::ord(x, flip, nl);
        if (track) {
            t_and &= o;
            t_or |= o;
        }
        const uint32_t b = uint32_t(o >> shift) & 0xffffu;
        const uint64_t act = __ballot(1);
        const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
        if (__ballot(b != b0) == 0) {  // a wave of one bin (sorted or constant runs): one add
            if (lanes_below(act) == 0) add(b0, uint32_t(__popcll(act)));
        } else {
            add(b, 1u);
        }
    };
    constexpr int VEC = 16 / sizeof(U);
    const uint32_t mis = uint32_t(reinterpret_cast<uintptr_t>(keys + r0) & 15u);
    const uint32_t head = min(r1 - r0, uint32_t(((16u - mis) & 15u) / sizeof(U)));
    const uint32_t v0 = r0 + head;
    const uint32_t nvec = (r1 - v0) / VEC;
    const uint4 *kv = reinterpret_cast<const uint4 *>(keys + v0);
    constexpr uint32_t UNR = 4;
    uint32_t v = tid;
    for (; v + (UNR - 1) * H2_BLOCK < nvec; v += UNR * H2_BLOCK) {
        uint4 q[UNR];
#pragma unroll
        for (uint32_t u = 0; u < UNR; ++u) q[u] = kv[v + u * H2_BLOCK];
#pragma unroll
        for (uint32_t u = 0; u < UNR; ++u) {
            const U *e = reinterpret_cast<const U *>(&q[u]);
#pragma unroll
            for (int j = 0; j < VEC; ++j) count(e[j]);
        }
    }
    for (; v < nvec; v += H2_BLOCK) {
        const uint4 q = kv[v];
        const U *e = reinterpret_cast<const U *>(&q);
#pragma unroll
        for (int j = 0; j < VEC; ++j) count(e[j]);
    }
    const uint32_t t0 = v0 + nvec * VEC;
    const uint32_t rest = head + (r1 - t0);
    for (uint32_t i = tid; i < rest; i += H2_BLOCK) count(keys[i < head ? r0 + i : t0 + (i - head)]);
    if (track) {
        atomicAnd(&s_ao[0], (unsigned long long)t_and | ~(unsigned long long)U(~U(0)));
        atomicOr(&s_ao[1], (unsigned long long)t_or);
    }
    __syncthreads();
    if (track && tid == 0) {
        a.andor_slots[2 * r] = s_ao[0];
        a.andor_slots[2 * r + 1] = s_ao[1];
    }
    uint32_t *out = a.res + size_t(r) * 32768;
    for (uint32_t i = tid; i < 32768; i += H2_BLOCK) out[i] = h[i];
    // high-digit counts: four threads per digit (64 bins each), plus spills
    const uint32_t d1 = tid >> 2, qtr = tid & 3u;
    uint32_t c = 0;
    for (uint32_t i = 0; i < 32; ++i) {
        const uint32_t w = h[d1 * 128 + qtr * 32 + i];
        c += (w & 0xffffu) + (w >> 16);
    }
    c += __shfl_xor(c, 1);
    c += __shfl_xor(c, 2);
    if (qtr == 0) {
        // the log was written by other threads' device atomics and stores
        // in this kernel: read past the CU's cache
        const uint32_t ns = min(__hip_atomic_load(a.spill_cnt + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                a.spill_cap);
        for (uint32_t i = 0; i < ns; ++i)
            if ((__hip_atomic_load(a.spill_bin + size_t(r) * a.spill_cap + i, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) >> 8) == d1)
                c += H2_SPILL;
        a.rsum[size_t(d1) * a.ranges + r] = c;
    }
}

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
    __shared__ uint32_t s_r0[257];
    __shared__ uint64_t s_start[256];
    a.seg_off[s] = start;
    s_r0[s] = r0;
    s_start[s] = start;
    if (s == 255) {
        a.seg_off[256] = start + size;
        *a.tb.nr = r0 + nrs;
        *a.nlarge = 0;
        s_r0[256] = r0 + nrs;
    }
    if (s < a.nwctr) a.wctr[s] = 0;
    __syncthreads();
    // every range by its own thread (one bucket may hold nearly every key)
    const uint32_t total = min(s_r0[256], a.tb.cap);  // cap = ceil(n / rk) + 256: never below the count
    for (uint32_t r = s; r < total; r += 256) {
        uint32_t lo = 0, hi = 256;  // the top bucket with s_r0[lo] <= r < s_r0[lo + 1]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (s_r0[mid] <= r) lo = mid;
            else hi = mid;
        }
        const uint64_t st = s_start[lo], end = lo == 255 ? st + a.top_total[255] : s_start[lo + 1];
        const uint64_t b = st + uint64_t(r - s_r0[lo]) * a.rk_b;
        a.tb.begin[r] = b;
        a.tb.end[r] = min(end, b + a.rk_b);
        a.tb.sbase[r] = st;
        a.tb.bseg[r] = lo;
        a.tb.head[r] = b == st ? 1u : 0u;
    }
}

// Pass B's per-range digit counts without a histogram read of Y.  The keys
// of pass-A range r with top digit d1 lie in one run of Y inside top bucket
// d1, at offset rpre_a[d1][r] (pass A's scan); the two-digit histogram of
// pass A already counts each run's keys by the next digit.  A pass-B range
// [a, b) of bucket d1 takes the rows of the runs it covers whole and counts
// the keys of the (at most two) runs it covers in part directly from Y --
// about a run's length per range.  A workgroup per range.
constexpr uint32_t GRP_MAX_RUNS = 4096;
template <class KD>
__global__ __launch_bounds__(1024) void hyb_msd_hist_kernel(HybTable tb, const void *ykeys_v, uint64_t flip64,
                                                            uint32_t nan_last, const uint32_t *shift_dev,
                                                            const uint32_t *rpre_a, const uint32_t *top_total,
                                                            uint32_t ranges_a, const uint32_t *res,
                                                            const uint32_t *spill_cnt, const uint32_t *spill_bin,
                                                            uint32_t spill_cap, uint32_t *rsum_b)
{
    using U = typename KD::U;
    __shared__ uint32_t s_off[GRP_MAX_RUNS + 1];
    __shared__ uint32_t hist[4][256];  // one copy per 4 waves: fewer waves meet on a counter
    __shared__ uint32_t part[4][256];
    const uint32_t nr = *tb.nr;
    const uint32_t tid = threadIdx.x, d2 = tid & 255u, q = tid >> 8;
    uint32_t *const hw = hist[(tid >> 6) & 3u];
    const int shift = int(dev_shift(shift_dev, 0));
    const U flip = U(flip64);
    const bool nl = nan_last != 0;
    const U *ykeys = static_cast<const U *>(ykeys_v);
    const uint16_t *res16 = reinterpret_cast<const uint16_t *>(res);
    uint32_t loaded = ~0u;  // the top digit whose run offsets s_off holds
    for (uint32_t g = blockIdx.x; g < nr; g += gridDim.x) {
        const uint32_t d1 = tb.bseg[g];
        const uint64_t sb = tb.sbase[g];
        const uint32_t a = uint32_t(tb.begin[g] - sb), b = uint32_t(tb.end[g] - sb);
        __syncthreads();  // the previous range's readers of s_off / hist / part are done
        if (d1 != loaded) {
            for (uint32_t r = tid; r < ranges_a; r += 1024) s_off[r] = rpre_a[size_t(d1) * ranges_a + r];
            if (tid == 0) s_off[ranges_a] = top_total[d1];
            loaded = d1;
        }
        hist[q][d2] = 0;
        __syncthreads();
        // the runs holding positions a and b - 1: the last run starting at or before them
        auto run_of = [&](uint32_t x) {
            uint32_t lo = 0, hi = ranges_a;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (s_off[mid] <= x) lo = mid;
                else hi = mid;
            }
            return lo;
        };
        const uint32_t ra = run_of(a), rb = run_of(b - 1);
        // a wave whose keys share one digit (sorted runs) adds once, since
        // same-address adds serialise
        auto add = [&](U key) {
            const uint32_t d = KD::digit(key, flip, nl, shift);
            const uint64_t act = __ballot(1);
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            if (__ballot(d != d0) == 0) {
                if (lanes_below(act) == 0) atomicAdd(&hw[d0], uint32_t(__popcll(act)));
            } else {
                atomicAdd(&hw[d], 1u);
            }
        };
        auto count = [&](uint32_t lo, uint32_t hi) {
            uint32_t x = lo + tid;
            for (; x + 3 * 1024 < hi; x += 4 * 1024) {  // four loads in flight
                U k[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) k[j] = ykeys[sb + x + j * 1024];
#pragma unroll
                for (int j = 0; j < 4; ++j) add(k[j]);
            }
            for (; x < hi; x += 1024) add(ykeys[sb + x]);
        };
        if (ra == rb) {
            count(a, b);
        } else {
            count(a, s_off[ra + 1]);
            count(s_off[rb], b);
        }
        // the runs covered whole: their rows of pass A's histogram
        const uint32_t bin = (d1 << 8) | d2;
        uint32_t c = 0;
        for (uint32_t r = ra + 1 + q; r < rb; r += 4) {
            c += res16[size_t(r) * 65536 + bin];
            const uint32_t ns = min(spill_cnt[r], spill_cap);
            for (uint32_t i = 0; i < ns; ++i)
                if (spill_bin[size_t(r) * spill_cap + i] == bin) c += H2_SPILL;
        }
        part[q][d2] = c;
        __syncthreads();
        if (q == 0)
            rsum_b[size_t(d2) * tb.cap + g] = hist[0][d2] + hist[1][d2] + hist[2][d2] + hist[3][d2] + part[0][d2] +
                                              part[1][d2] + part[2][d2] + part[3][d2];
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
    // a plan without low digits (plan[4]) leaves every bucket as it is:
    // marked larger than any LDS capacity, and listed nowhere
    if (a.plan[4]) {
        a.bsize[b] = ~0u;
        return;
    }
    a.bsize[b] = size;
    if (size > a.cap) {
        const uint32_t i = atomicAdd(a.nlarge, 1u);
        if (i < a.lcap) {  // lcap >= n / (cap + 1): always true
            a.lstart[i] = start;
            a.lsize[i] = size;
        }
    }
}

// ranges of the low passes over the large buckets (one workgroup): the
// first range of every bucket by a scan, then every range written by its own
// thread (a bucket of millions of keys has thousands of ranges)
template <int = 0>
__global__ __launch_bounds__(1024) void hyb_large_ranges_kernel(HybArgs a)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t s_total;
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
    for (uint32_t i = c0; i < c1; ++i) {
        a.seg_andor[2 * size_t(i)] = ~0ull;
        a.seg_andor[2 * size_t(i) + 1] = 0ull;
        a.lrfirst[i] = r;
        r += uint32_t((a.lsize[i] + a.rk_d - 1) / a.rk_d);
    }
    if (t == 1023) {
        const uint32_t total = min(a.td.cap, r);  // cap = ceil(n / rk) + lcap: never below r
        *a.td.nr = total;
        a.lrfirst[nl] = r;
        s_total = total;
    }
    __syncthreads();
    const uint32_t total = s_total;
    for (uint32_t q = t; q < total; q += 1024) {
        uint32_t lo = 0, hi = nl;  // the bucket i with lrfirst[i] <= q < lrfirst[i + 1]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (a.lrfirst[mid] <= q) lo = mid;
            else hi = mid;
        }
        const uint64_t st = a.lstart[lo], sz = a.lsize[lo];
        const uint64_t x = st + uint64_t(q - a.lrfirst[lo]) * a.rk_d;
        a.td.begin[q] = x;
        a.td.end[q] = min(st + sz, x + a.rk_d);
        a.td.sbase[q] = st;
        a.td.bseg[q] = lo;
        a.td.head[q] = x == st ? 1u : 0u;
    }
}

// The ranges of the low passes after the first (one workgroup): those of
// the large buckets whose keys are not all equal, by the AND / OR the first
// pass's histogram read took (a bucket of equal keys stays where it is, so
// the later passes need not walk its ranges; duplicate-heavy data has
// thousands of them).  Segment indices stay the large buckets' own.
template <class KD>
__global__ __launch_bounds__(1024) void hyb_live_ranges_kernel(HybArgs a, uint32_t canon)
{
    using U = typename KD::U;
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t s_total;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t nl = min(*a.nlarge, a.lcap);
    const uint32_t per = (nl + 1023) / 1024;
    const uint32_t c0 = min(nl, t * per), c1 = min(nl, c0 + per);
    auto ranges_of = [&](uint32_t i) {
        const U x = U(a.seg_andor[2 * size_t(i)]), o = U(a.seg_andor[2 * size_t(i) + 1]);
        bool c = x == o;
        if constexpr (KD::KIND == Kind::Float) {
            if (c && canon && KD::is_nan(x) && x != KD::QNAN) c = false;
        }
        return c ? 0u : uint32_t((a.lsize[i] + a.rk_d - 1) / a.rk_d);
    };
    uint32_t cnt = 0;
    for (uint32_t i = c0; i < c1; ++i) cnt += ranges_of(i);
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
    for (uint32_t i = c0; i < c1; ++i) {
        a.lrfirst2[i] = r;
        r += ranges_of(i);
    }
    if (t == 1023) {
        const uint32_t total = min(a.tl.cap, r);
        *a.tl.nr = total;
        a.lrfirst2[nl] = r;
        s_total = total;
    }
    __syncthreads();
    const uint32_t total = s_total;
    for (uint32_t q = t; q < total; q += 1024) {
        // the bucket i with lrfirst2[i] <= q < lrfirst2[i + 1] (the last of
        // equal starts: buckets without ranges share their successor's)
        uint32_t lo = 0, hi = nl;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (a.lrfirst2[mid] <= q) lo = mid;
            else hi = mid;
        }
        const uint64_t st = a.lstart[lo], sz = a.lsize[lo];
        const uint64_t x = st + uint64_t(q - a.lrfirst2[lo]) * a.rk_d;
        a.tl.begin[q] = x;
        a.tl.end[q] = min(st + sz, x + a.rk_d);
        a.tl.sbase[q] = st;
        a.tl.bseg[q] = lo;
        a.tl.head[q] = x == st ? 1u : 0u;
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
// Grids of the segmented passes, whose ranges are handed out by a counter
// (work stealing): a scatter workgroup holds a CU (1024 threads), so one per
// CU; histogram workgroups (256 threads) four per CU.  A pass with little
// or no work (the large-bucket passes of spread data) then costs little.
#ifndef XSS_HYB_DOWN_GRID
#define XSS_HYB_DOWN_GRID 256
#endif
#ifndef XSS_HYB_UP_GRID
#define XSS_HYB_UP_GRID 1024
#endif
// control words of the sort: plan[8], AND / OR [2 x u64], work counters
constexpr size_t HYB_CTL_BYTES = 32 + 16 + 64 * 4;
constexpr uint32_t HYB_NWCTR = 64;

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
                      uint64_t flip64, uint32_t canon, const uint32_t *plan)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_PL = PL != PL_NONE;
    constexpr int P = KD::PASSES;
    constexpr int WAVES = HYB_BLOCK / 64;
    constexpr int ITEMS = HybShape<KD, PL, SIZE>::ITEMS;
    constexpr uint32_t CAP = HybShape<KD, PL, SIZE>::CAP;
    // 4-byte keys (not argsort): the 16 low bits sort as one unstable pass on
    // 12 bits and one stable pass on 4 (SPLIT), so that only 4 bits pay the
    // stable peer match
    constexpr bool SPLIT = XSS_HYB_SPLIT && P == 4 && PL != PL_IDX;
    __shared__ U s_keys[CAP];
    __shared__ V s_vals[HAS_PL ? CAP : 1];
    __shared__ uint32_t s_off[SPLIT ? 4096 : WAVES * 256];
    __shared__ uint32_t s_wsum[WAVES];
    const uint32_t b = blockIdx.x;
    const uint32_t m = bsize[b];
    if (m > CAP) return;  // a large bucket: the low segmented passes sort it (or a plan with no low digit)
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
    // SPLIT: pass 1 on the low 12 bits, unstable: every key takes its slot
    // in its digit with one returning LDS add on a counter shared by the
    // workgroup (4096 digits: few collisions), the counters are scanned, and
    // the keys scattered.  Pass 2 on bits 12..15, stable, with per-thread
    // counters (below).  Padding takes no slot in pass 1 and digit 15 (last)
    // in pass 2.
    auto split_12_4 = [&]() {
        uint32_t *cnt = s_off;
        for (uint32_t i = tid; i < 4096; i += HYB_BLOCK) cnt[i] = 0;
        __syncthreads();
        uint32_t rd[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                const uint32_t idx = lbase + i * 64;
                const bool real = idx < m;
                const uint32_t d = uint32_t(KD::ord(key[i], flip, NL)) & 0xfffu;
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
                const bool uni = __ballot(!real || d != d0) == 0;
                uint32_t old = 0;
                if (real && (!uni || lane == 0)) old = atomicAdd(&cnt[d], uni ? 64u : 1u);
                rd[i] = old | (d << 13) | (uni ? (1u << 25) : 0u) | (real ? 0u : (1u << 26));
            }
        }
        __syncthreads();
        // exclusive scan of the 4096 counters: 8 per thread, then across the workgroup
        {
            constexpr uint32_t PER = 4096 / HYB_BLOCK;
            uint32_t v[PER], sum = 0;
#pragma unroll
            for (uint32_t j = 0; j < PER; ++j) {
                v[j] = cnt[tid * PER + j];
                sum += v[j];
            }
            uint32_t incl = sum;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(incl, off);
                if (lane >= uint32_t(off)) incl += t;
            }
            if (lane == 63) s_wsum[wave] = incl;
            __syncthreads();
            uint32_t run = incl - sum;
            for (uint32_t w = 0; w < wave; ++w) run += s_wsum[w];
#pragma unroll
            for (uint32_t j = 0; j < PER; ++j) {
                cnt[tid * PER + j] = run;
                run += v[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                uint32_t r = rd[i];
                if (r >> 26) continue;  // padding
                if (r >> 25) r = (__builtin_amdgcn_readfirstlane(r) + lane) & 0x1ffffffu;
                const uint32_t d = r >> 13;
                const uint32_t slot = cnt[d] + (r & 0x1fffu);
                s_keys[slot] = key[i];
                if constexpr (HAS_PL) s_vals[slot] = val[i];
            }
        }
        // pass 2: bits 12..15, stable.  Each thread takes `it` consecutive
        // positions (blocked) and counts its keys per digit in registers
        // (16 counters of 4 bits: at most 12 keys each); the per-thread
        // counts, digit-major, are scanned in LDS; a key's slot is its
        // (digit, thread) base plus its rank among the thread's keys.
        uint16_t *cnt16 = reinterpret_cast<uint16_t *>(s_off);  // [16][HYB_BLOCK]
        static_assert(16 * HYB_BLOCK * 2 <= 4096 * 4, "counters fit");
        static_assert(ITEMS < 16, "4-bit counters");
        U k2[ITEMS];
        V v2[ITEMS];
        uint32_t lr[ITEMS];
        uint32_t c_lo = 0, c_hi = 0;
        __syncthreads();  // pass 1's scatter is complete
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                const uint32_t pos = tid * it + i;
                const bool real = pos < m;
                k2[i] = real ? s_keys[pos] : U(0);
                if constexpr (HAS_PL) v2[i] = real ? s_vals[pos] : V(0);
                const uint32_t d = real ? (uint32_t(KD::ord(k2[i], flip, NL)) >> 12) & 0xfu : 15u;
                const uint32_t sh = (d & 7u) * 4u;
                const uint32_t r = ((d < 8 ? c_lo : c_hi) >> sh) & 0xfu;
                if (d < 8) c_lo += 1u << sh;
                else c_hi += 1u << sh;
                lr[i] = r | (d << 8);
            }
        }
#pragma unroll
        for (uint32_t d = 0; d < 16; ++d)
            cnt16[d * HYB_BLOCK + tid] = uint16_t(((d < 8 ? c_lo : c_hi) >> ((d & 7u) * 4u)) & 0xfu);
        __syncthreads();
        {
            constexpr uint32_t PER = 16;  // 16 * HYB_BLOCK counters, HYB_BLOCK threads
            uint32_t v[PER], sum = 0;
#pragma unroll
            for (uint32_t j = 0; j < PER; ++j) {
                v[j] = cnt16[tid * PER + j];
                sum += v[j];
            }
            uint32_t incl = sum;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(incl, off);
                if (lane >= uint32_t(off)) incl += t;
            }
            if (lane == 63) s_wsum[wave] = incl;
            __syncthreads();
            uint32_t run = incl - sum;
            for (uint32_t w = 0; w < wave; ++w) run += s_wsum[w];
#pragma unroll
            for (uint32_t j = 0; j < PER; ++j) {
                cnt16[tid * PER + j] = uint16_t(run);
                run += v[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (uint32_t(i) < it) {
                const uint32_t d = lr[i] >> 8;
                const uint32_t slot = cnt16[d * HYB_BLOCK + tid] + (lr[i] & 0xffu);
                s_keys[slot] = k2[i];
                if constexpr (HAS_PL) s_vals[slot] = v2[i];
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
    if constexpr (SPLIT) {
        split_12_4();
    } else if constexpr (P == 4) {
        // 4-byte keys: both passes inline (measured 8% faster than the loop)
        pass(0, true);
        pass(8, false);
    } else {
        // low digits to sort (plan[2]): P - 2, or fewer when the top digits
        // are the same in every key
        const int npass = int(__builtin_amdgcn_readfirstlane(plan[2]));
#pragma unroll 1
        for (int p = 0; p < npass; ++p) pass(8 * p, p == 0);
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t idx = lbase + i * 64;
        if (uint32_t(i) < it && idx < m) put(idx, key[i], HAS_PL ? val[i] : V(0));
    }
}

// 8-byte keys (qsort, argsort, key-value sort): the buckets are sorted by the batch
// form's LDS sort (ss_local_kernel: an MSD split of the bucket's measured key
// range into 4096 sub-buckets, then a rank by comparison), which costs the
// same whatever the keys' width, in place of six 8-bit LSD passes in LDS
// (build-time knob; 0: hyb_local_kernel)
#ifndef XSS_HYB_SS_LOCAL
#define XSS_HYB_SS_LOCAL 1
#endif
template <class KD, int PL>
constexpr bool hyb_ss_local()
{
    return XSS_HYB_SS_LOCAL && sizeof(typename KD::U) == 8;
}
constexpr uint32_t HYB_SS_CAP = uint32_t(SS_LOCAL_BLOCK) * SS_LOCAL_ITEMS_W8;
// a smaller shape (6 keys per thread, 3072-key capacity) for arrays whose
// buckets average at most that less HYB_SMALL_MARGIN: less LDS per
// workgroup, more workgroups per CU (build-time knob; 0: always 8 items)
#ifndef XSS_HYB_SS_SMALL
#define XSS_HYB_SS_SMALL 1
#endif
constexpr int HYB_SS_ITEMS_SMALL = 6;
constexpr uint32_t HYB_SS_CAP_SMALL = uint32_t(SS_LOCAL_BLOCK) * HYB_SS_ITEMS_SMALL;
inline bool hyb_ss_small(size_t n)
{
    return XSS_HYB_SS_SMALL && n / 65536 + 384 <= HYB_SS_CAP_SMALL;
}
// and a larger one (12 per thread, 6144 keys) when the average bucket is
// above 3712 keys (arrays above about 2^27.8 8-byte keys)
constexpr int HYB_SS_ITEMS_LARGE = 12;
constexpr uint32_t HYB_SS_CAP_LARGE = uint32_t(SS_LOCAL_BLOCK) * HYB_SS_ITEMS_LARGE;
inline bool hyb_ss_large(size_t n)
{
    return n / 65536 + 384 > HYB_SS_CAP;
}
inline uint32_t hyb_ss_cap(size_t n)
{
    return hyb_ss_small(n) ? HYB_SS_CAP_SMALL : hyb_ss_large(n) ? HYB_SS_CAP_LARGE : HYB_SS_CAP;
}

template <class KD, int PL>
struct HybPlan {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr bool HAS_PL = PL != PL_NONE;
    static constexpr uint32_t CAP = hyb_ss_local<KD, PL>() ? HYB_SS_CAP_LARGE : HybShape<KD, PL>::CAP;
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
        if constexpr (hyb_ss_local<KD, PL>()) return hyb_ss_cap(n);
        const int sh = shape(n);
        return sh == 2 ? HybShape<KD, PL, 2>::CAP : sh == 1 ? HybShape<KD, PL, 1>::CAP : CAP;
    }
    // tile shape of the segmented passes (16K-key tiles for 4-byte keys,
    // see HybSegShape)
    using SS = HybSegShape<KD, PL>;
    static constexpr uint64_t SEG_TILE = uint64_t(SS::BLOCK) * SS::ITEMS;
    // pass B's table must fit the workspace's [256][MAX_RANGES] range tables
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
    // pass A: H2_RANGES ranges of whole tiles (its histogram read counts both MSD digits)
    static uint32_t range_tiles_a(size_t n)
    {
        const size_t tiles = SortPlan<KD, PL>::tiles(n);
        return uint32_t((tiles + H2_RANGES - 1) / H2_RANGES);
    }
    static uint32_t ranges_a(size_t n)
    {
        return uint32_t((SortPlan<KD, PL>::tiles(n) + range_tiles_a(n) - 1) / range_tiles_a(n));
    }
    static uint32_t spill_cap(size_t n)
    {
        return uint32_t(size_t(range_tiles_a(n)) * SortPlan<KD, PL>::tile(n) / H2_SPILL + 1);
    }
    static size_t h2_bytes(size_t n)
    {
        const size_t ra = ranges_a(n);
        return align256(ra * 32768 * 4) + align256(ra * 4) + align256(ra * size_t(spill_cap(n)) * 4);
    }
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
               2 * align256(size_t(256) * cap_d(n) * 4) + align256(size_t(lcap(n)) * 256 * 4) +
               align256(HYB_CTL_BYTES) + align256(size_t(lcap(n)) * 16) + align256((size_t(lcap(n)) + 1) * 4) +
               table_bytes(cap_d(n)) + align256((size_t(lcap(n)) + 1) * 4) + h2_bytes(n);
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
    char *ctl = take(HYB_CTL_BYTES);
    uint32_t *plan = reinterpret_cast<uint32_t *>(ctl);
    uint64_t *andor = reinterpret_cast<uint64_t *>(ctl + 32);
    h.wctr = reinterpret_cast<uint32_t *>(ctl + 48);
    h.nwctr = HYB_NWCTR;
    h.seg_andor = reinterpret_cast<uint64_t *>(take(size_t(lcap) * 16));
    h.lrfirst = reinterpret_cast<uint32_t *>(take((size_t(lcap) + 1) * 4));
    table(h.tl, cap_d, ctr + 3);
    h.lrfirst2 = reinterpret_cast<uint32_t *>(take((size_t(lcap) + 1) * 4));
    h.plan = plan;
    const uint32_t ranges_a = HP::ranges_a(n), rtiles_a = HP::range_tiles_a(n), spill_cap = HP::spill_cap(n);
    uint32_t *h2_res = reinterpret_cast<uint32_t *>(take(size_t(ranges_a) * 32768 * 4));
    uint32_t *h2_scnt = reinterpret_cast<uint32_t *>(take(size_t(ranges_a) * 4));
    uint32_t *h2_sbin = reinterpret_cast<uint32_t *>(take(size_t(ranges_a) * spill_cap * 4));
    if (size_t(p - scratch) > HP::scratch_bytes(n))
        return set_error(XSS_ERR_NOMEM, "MSD-first sort: scratch layout (%zu B) exceeds its size (%zu B)",
                         size_t(p - scratch), HP::scratch_bytes(n));

    h.top_total = ws->total();
    h.segtot = segtot;
    h.nlarge = ctr;
    h.lcap = lcap;
    h.cap = HP::cap(n);
    h.rk_b = HP::rk_b(n);
    h.rk_d = HP::rk_d(n);

    // pass A: top digit over the whole array (keys -> Y; argsort: indices
    // generated).  Its histogram read counts the keys of each of its ranges
    // by both MSD digits (pass B then needs no histogram read) and takes the
    // AND / OR of every order key, from which the plan kernel picks the
    // digit pass A really sorts by (the top one unless all keys share it);
    // the histograms are taken again only in that case.
    {
        using Plan = SortPlan<KD, PL>;
        H2Args u{};
        u.keys = keys; u.n = nn; u.range_keys = uint32_t(rtiles_a * Plan::tile(n)); u.ranges = ranges_a;
        u.shift_hi = uint32_t(8 * (P - 1)); u.flip = flip; u.nan_last = nan_last;
        u.andor_slots = ws->andor_slots(); u.res = h2_res; u.spill_cnt = h2_scnt; u.spill_bin = h2_sbin;
        u.spill_cap = spill_cap; u.rsum = ws->rsum();
        {
            KernelTimer t("upsweep", s);
            hipLaunchKernelGGL((hyb_h2_kernel<KD>), dim3(ranges_a), dim3(H2_BLOCK), 0, s, u);
        }
        XSS_HIP_CHECK(hipGetLastError());
        {
            KernelTimer t("hyb_tables", s);
            hipLaunchKernelGGL(hyb_plan_kernel<P>, dim3(1), dim3(256), 0, s, ws->andor_slots(), ranges_a, andor,
                               plan);
        }
        XSS_HIP_CHECK(hipGetLastError());
        H2Args u2 = u;  // the histograms of plan[0]'s digits, if they are not the top ones
        u2.andor_slots = nullptr; u2.shift_dev = plan; u2.gate = plan + 3;
        {
            KernelTimer t("upsweep", s);
            hipLaunchKernelGGL((hyb_h2_kernel<KD>), dim3(ranges_a), dim3(H2_BLOCK), 0, s, u2);
        }
        XSS_HIP_CHECK(hipGetLastError());
        ScanArgs sc{};
        sc.rsum = ws->rsum(); sc.rpre = ws->rpre(); sc.total = ws->total(); sc.ranges = ranges_a;
        {
            KernelTimer t("range_scan", s);
            hipLaunchKernelGGL(range_scan_kernel<0>, dim3(256), dim3(SCAN_BLOCK), 0, s, sc);
        }
        XSS_HIP_CHECK(hipGetLastError());
        PassArgs a{};
        a.keys_in = keys; a.vals_in = vals; a.keys_out = yk; a.vals_out = yv; a.idx_out = nullptr;
        a.n = nn; a.shift = uint32_t(8 * (P - 1)); a.shift_dev = plan; a.flip = flip; a.nan_last = nan_last;
        a.range_tiles = rtiles_a; a.ranges = ranges_a; a.rpre = ws->rpre(); a.total = ws->total();
        a.first = true; a.last = false; a.canon_nan = false;
        const int rc = Plan::small_tiles(n)
                           ? launch_downsweep<KD, PL, SmallPassShape<KD, PL>>(a, ranges_a, nan_last, true, s)
                           : launch_downsweep<KD, PL, PassShape<KD, PL>>(a, ranges_a, nan_last, true, s);
        if (rc) return rc;
    }
    {
        KernelTimer t("hyb_tables", s);
        hipLaunchKernelGGL(hyb_top_ranges_kernel<0>, dim3(1), dim3(256), 0, s, h);
    }
    XSS_HIP_CHECK(hipGetLastError());
    {
        // pass B's digit counts: pass A's rows for the runs a range covers
        // whole, the keys of the rest counted from Y
        static_assert(H2_RANGES <= GRP_MAX_RUNS, "run offsets of a top bucket fit in LDS");
        KernelTimer t("hyb_msd_upsweep", s);
        hipLaunchKernelGGL((hyb_msd_hist_kernel<KD>), dim3(std::min<uint32_t>(cap_b, 2048)), dim3(1024), 0, s, h.tb,
                           yk, flip, uint32_t(nan_last), plan + 1, ws->rpre(), ws->total(), ranges_a, h2_res, h2_scnt,
                           h2_sbin, spill_cap, ws->rsum());
    }
    XSS_HIP_CHECK(hipGetLastError());
    uint32_t wnext = 0;  // next free work-stealing counter
    // one segmented pass over a device-built table (src -> dst)
    auto seg_pass = [&](const char *const names[3], const HybTable &t, uint32_t *rsum, uint32_t *rpre,
                        uint32_t *stot, int shift, const uint32_t *shift_dev, const uint32_t *gate,
                        uint32_t gate_min, const void *sk, const void *sv, void *dk, void *dv, bool last,
                        bool unstable, bool large, bool first_large) -> int {
        // The large-bucket passes (usually near empty) take their ranges from
        // counters with small grids, so that a table that turns out (nearly)
        // empty costs few workgroups; pass B (always full) takes one range
        // per workgroup, and its digit counts come from pass A's two-digit
        // histograms (hyb_group_hist_kernel), so it has no histogram read.
        uint32_t *up_ctr = large ? h.wctr + wnext++ : nullptr;
        uint32_t *down_ctr = large ? h.wctr + wnext++ : nullptr;
        if (large) {
            UpArgs u{};
            u.keys = sk; u.n = nn; u.shift = uint32_t(shift); u.shift_dev = shift_dev; u.nan_last = nan_last;
            u.flip = flip; u.rsum = rsum; u.ranges = t.cap; u.range_begin = t.begin; u.range_end = t.end;
            u.nr_dev = t.nr; u.gate = gate; u.gate_min = gate_min; u.wctr = up_ctr; u.rbseg = t.bseg;
            if (first_large) u.seg_andor_out = h.seg_andor;
            else u.seg_andor_in = h.seg_andor;
            u.seg_canon = canon_nan;
            KernelTimer kt(names[0], s);
            hipLaunchKernelGGL((upsweep_kernel<KD, true>), dim3(std::min<uint32_t>(t.cap, XSS_HYB_UP_GRID)),
                               dim3(UP_BLOCK), 0, s, u);
        }
        XSS_HIP_CHECK(hipGetLastError());
        SegScanArgs sc{rsum, rpre, stot, t.head, t.bseg, 0, t.cap, t.nr, gate, gate_min, down_ctr};
        {
            KernelTimer kt(names[1], s);
            hipLaunchKernelGGL(seg_scan_kernel<0>, dim3(256), dim3(256), 0, s, sc);
        }
        XSS_HIP_CHECK(hipGetLastError());
        PassArgs a{};
        a.keys_in = sk; a.vals_in = sv; a.keys_out = dk; a.vals_out = dv; a.idx_out = idx_out; a.n = nn;
        a.shift = uint32_t(shift); a.shift_dev = shift_dev; a.gate = gate; a.gate_min = gate_min; a.flip = flip;
        a.nan_last = nan_last; a.first = false; a.last = last;
        a.canon_nan = last && canon_nan;
        SegRanges r{t.begin, t.end, t.sbase, t.bseg, rpre, stot, 0, t.cap, t.nr, down_ctr};
        if (large) {
            r.seg_andor = h.seg_andor;
            r.seg_canon = canon_nan;
            r.home_vals = xv;
            a.last_gate = plan + 2;
            a.last_q = gate_min;
            a.canon_if_last = canon_nan;
        }
        const uint32_t grid = std::min<uint32_t>(t.cap, large ? XSS_HYB_DOWN_GRID * (1024 / SS::BLOCK) : MAX_RANGES);
        KernelTimer kt(names[2], s);
        // pass B's input order does not matter (it is an MSD pass inside
        // each top bucket), so it ranks without keeping it; argsort keeps
        // every pass stable (ties in input order, the identity for sorted
        // input)
        constexpr bool UN = PL != PL_IDX;
        auto go = [&](auto nl_tag, auto un_tag, auto seg_tag) {
            hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, SS::BLOCK, SS::ITEMS, decltype(nl_tag)::value,
                                                     decltype(un_tag)::value, decltype(seg_tag)::value>),
                               dim3(grid), dim3(SS::BLOCK), 0, s, a, r);
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        using U_ = std::integral_constant<bool, UN>;
        if (large) {
            if (unstable && nan_last) go(T_{}, U_{}, T_{});
            else if (unstable) go(F_{}, U_{}, T_{});
            else if (nan_last) go(T_{}, F_{}, T_{});
            else go(F_{}, F_{}, T_{});
        } else {
            if (unstable && nan_last) go(T_{}, U_{}, F_{});
            else if (unstable) go(F_{}, U_{}, F_{});
            else if (nan_last) go(T_{}, F_{}, F_{});
            else go(F_{}, F_{}, F_{});
        }
        XSS_HIP_CHECK(hipGetLastError());
        return XSS_OK;
    };
    // pass B: the digit below pass A's inside every top bucket (Y -> X)
    static const char *const msd_names[3] = {"hyb_msd_upsweep", "hyb_msd_scan", "hyb_msd_downsweep"};
    static const char *const low_names[3] = {"hyb_large_upsweep", "hyb_large_scan", "hyb_large_downsweep"};
    int rc = seg_pass(msd_names, h.tb, ws->rsum(), ws->rpre(), segtot, 8 * (P - 2), plan + 1, nullptr, 0, yk, yv,
                      xk, xv, false, true, false, false);
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
                                   h.bstart, h.bsize, xk, xv, idx_out, flip, uint32_t(canon_nan), plan);
            else
                hipLaunchKernelGGL((hyb_local_kernel<KD, PL, false, SZ>), dim3(HYB_BUCKETS), dim3(HYB_BLOCK), 0, s,
                                   h.bstart, h.bsize, xk, xv, idx_out, flip, 0u, plan);
        };
        if constexpr (hyb_ss_local<KD, PL>()) {
            SsLocalArgs la{};
            la.bstart = h.bstart; la.bsize = h.bsize; la.keys_in = xk; la.keys_out = xk; la.vals_in = xv;
            la.vals_out = xv;
            la.idx64_out = idx_out; la.flip = flip; la.canon = nan_last ? uint32_t(canon_nan) : 0u;
            la.cap = HP::cap(n);
            auto go = [&](auto nl_tag, auto items_tag) {
                hipLaunchKernelGGL((ss_local_kernel<KD, PL, decltype(nl_tag)::value, decltype(items_tag)::value>),
                                   dim3(HYB_BUCKETS), dim3(SS_LOCAL_BLOCK), 0, s, la);
            };
            using SMALL = std::integral_constant<int, HYB_SS_ITEMS_SMALL>;
            using MID = std::integral_constant<int, SS_LOCAL_ITEMS_W8>;
            using LARGE = std::integral_constant<int, HYB_SS_ITEMS_LARGE>;
            const int shape = hyb_ss_small(n) ? 0 : hyb_ss_large(n) ? 2 : 1;
            if (nan_last) {
                if (shape == 0) go(std::true_type{}, SMALL{});
                else if (shape == 1) go(std::true_type{}, MID{});
                else go(std::true_type{}, LARGE{});
            } else {
                if (shape == 0) go(std::false_type{}, SMALL{});
                else if (shape == 1) go(std::false_type{}, MID{});
                else go(std::false_type{}, LARGE{});
            }
        } else {
            const int sh = HP::shape(n);
            if constexpr (HybShape<KD, PL>::HAS_SMALL) {
                if (sh == 2) launch(std::integral_constant<int, 2>{});
            }
            if constexpr (HybShape<KD, PL>::HAS_MID) {
                if (sh == 1) launch(std::integral_constant<int, 1>{});
            }
            if (sh == 0) launch(std::integral_constant<int, 0>{});
        }
    }
    XSS_HIP_CHECK(hipGetLastError());
    {
        KernelTimer t("hyb_tables", s);
        hipLaunchKernelGGL(hyb_large_ranges_kernel<0>, dim3(1), dim3(1024), 0, s, h);
    }
    XSS_HIP_CHECK(hipGetLastError());
    // low digits of the large buckets (X -> Y -> X ...; the plan's count of
    // low digits is even; the last pass of argsort writes the indices).
    // Pass q runs only if the plan has more than q low digits.  The first
    // pass also takes each large bucket's AND / OR: a bucket of equal keys
    // (duplicate-heavy data) is then left where it is by every pass.  The
    // table is sized for the worst case (every key in a large bucket) but
    // is usually empty (spread data).
    void *sk = xk, *sv = xv, *dk = yk, *dv = yv;
    for (int q = 0; q < P - 2; ++q) {
        // the last pass (the plan's last low digit) writes argsort's
        // indices and canonical NaNs; which one that is is read on the device.
        // qsort / keyvalue_qsort: passes after the first walk only the
        // ranges of large buckets that are not constant (argsort's last pass
        // writes the indices of every bucket, so it keeps the whole table).
        const bool live_only = q > 0 && PL != PL_IDX;
        rc = seg_pass(low_names, live_only ? h.tl : h.td, rsum_d, rpre_d, segtot_d, 8 * q, nullptr, plan + 2,
                      uint32_t(q), sk, sv, dk, dv, false, q == 0, true, q == 0);
        if (rc) return rc;
        if (q == 0 && PL != PL_IDX && P > 3) {
            KernelTimer t("hyb_tables", s);
            hipLaunchKernelGGL(hyb_live_ranges_kernel<KD>, dim3(1), dim3(1024), 0, s, h, uint32_t(canon_nan));
            XSS_HIP_CHECK(hipGetLastError());
        }
        std::swap(sk, dk);
        std::swap(sv, dv);
    }
    if (PL == PL_IDX || canon_nan) {
        KernelTimer t("hyb_tables", s);
        hipLaunchKernelGGL((hyb_finish_kernel<KD, PL>), dim3(1024), dim3(256), 0, s, plan, xk, xv, idx_out,
                           uint64_t(n), uint32_t(canon_nan));
    }
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

// ---------------------------------------------------------------- one array in one workgroup
//
// qsort (in place) and argsort of an array of 4- or 8-byte keys above the
// LDS bitonic size and up to the capacity of the batch form's LDS sort
// (512 x 12 keys): one workgroup of that sort (ss_local_kernel) in one
// launch, against the LSD plan's 14 launches.  Argsort ranks equal keys by
// index, so its indices are the stable sort's.
constexpr int L1_ITEMS = 12;
constexpr uint32_t L1_CAP = uint32_t(SS_LOCAL_BLOCK) * L1_ITEMS;

template <class KD, int PL, bool NL>
int enqueue_local1(const void *keys, void *keys_out, uint64_t *idx_out, size_t n, bool desc, bool canon,
                   hipStream_t s)
{
    SsLocalArgs la{};
    la.n1 = uint32_t(n);
    la.keys_in = keys;
    la.keys_out = keys_out;
    la.idx64_out = idx_out;  // argsort: 64-bit indices (the keys are not written)
    la.flip = desc ? ~0ull : 0ull;
    la.canon = (canon && NL && PL == PL_NONE) ? 1u : 0u;
    KernelTimer t("local1", s);
    hipLaunchKernelGGL((ss_local_kernel<KD, PL, NL, L1_ITEMS>), dim3(1), dim3(SS_LOCAL_BLOCK), 0, s, la);
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

// qsort (in place) of an array of 4- or 8-byte keys a little above the LDS
// sort's capacity: one 1024-thread workgroup runs the LSD passes over the
// array in tiles through the cache (ss_big_kernel, the batch form's pass
// for large buckets, with the array as its one bucket; passes over digits
// every key shares are skipped), in two launches against the LSD plan's 14.
template <int = 0>
__global__ __launch_bounds__(64) void big1_init_kernel(uint32_t n, uint32_t *t)
{
    if (threadIdx.x == 0) {
        t[0] = 1;                                  // ctl: one listed bucket
        t[1] = 0;                                  // big[0]: slot 0
        t[2] = 0;                                  // bstart[0]
        t[3] = n;                                  // bsize[0]
        t[4] = 0;                                  // bk_seg[0]
        t[5] = 0;                                  // lseg[0]
        reinterpret_cast<uint64_t *>(t + 6)[0] = 0;  // seg_off[0]
    }
}

inline size_t big1_scratch_bytes(size_t n, size_t key_bytes)
{
    return align256(n * key_bytes) + 256;
}

template <class KD, bool NL>
int enqueue_big1(char *scratch, void *keys, size_t n, bool desc, bool canon, hipStream_t s)
{
    using U = typename KD::U;
    uint32_t *t = reinterpret_cast<uint32_t *>(scratch + align256(n * sizeof(U)));
    {
        KernelTimer kt("big1_init", s);
        hipLaunchKernelGGL(big1_init_kernel<0>, dim3(1), dim3(64), 0, s, uint32_t(n), t);
    }
    XSS_HIP_CHECK(hipGetLastError());
    SsBigArgs g{};
    g.ctl = t; g.big = t + 1; g.bstart = t + 2; g.bsize = t + 3; g.bk_seg = t + 4; g.lseg = t + 5;
    g.seg_off = reinterpret_cast<const uint64_t *>(t + 6);
    g.y_keys = keys; g.z_keys = scratch; g.flip = desc ? ~0ull : 0ull; g.canon = (canon && NL) ? 1u : 0u;
    g.end_in_y = 1;
    {
        KernelTimer kt("big1", s);
        hipLaunchKernelGGL((ss_big_kernel<KD, PL_NONE, NL>), dim3(1), dim3(SS_BIG_BLOCK), 0, s, g);
    }
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

}  // namespace xss

namespace xss {

// qsort of 4- and 8-byte keys above L1_CAP and up to this many keys takes
// the one-workgroup LSD (xss_gpu_set_tuning "big1_max_n"; 0: off)
std::atomic<uint64_t> &big1_max_n();

// qsort of 4- and 8-byte keys up to this many keys (above the LDS bitonic
// size, at most L1_CAP) is one LDS workgroup (xss_gpu_set_tuning
// "local1_max_n"; 0: off)
std::atomic<uint64_t> &local1_max_n();

// arrays of at least this many keys take the MSD-first path (xss_gpu_set_tuning)
std::atomic<uint64_t> &hybrid_min_n();
std::atomic<uint64_t> &count_min_n();  // 16-bit qsort: counting sort from this size

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
    if constexpr (sizeof(typename KD::U) >= 4 && (PL == PL_NONE || PL == PL_IDX)) {
        // arrays up to one LDS workgroup's capacity: qsort in place, argsort
        const bool args_ok = keys_in && (PL == PL_IDX ? idx_out != nullptr : keys_in == keys_out);
        if (args_ok && !SortPlan<KD, PL>::small(n) && n <= std::min<uint64_t>(local1_max_n().load(), L1_CAP)) {
            if ((KD::KIND == Kind::Float) && hasnan)
                return enqueue_local1<KD, PL, true>(keys_in, keys_out, idx_out, n, desc, canon, s);
            return enqueue_local1<KD, PL, false>(keys_in, keys_out, idx_out, n, desc, canon, s);
        }
        if (PL == PL_NONE && keys_in && keys_in == keys_out && n > L1_CAP && n <= big1_max_n().load()) {
            Workspace *ws = nullptr;
            std::unique_lock<std::mutex> lock;
            int rc = acquire_workspace(s, big1_scratch_bytes(n, sizeof(typename KD::U)), &ws, &lock);
            if (rc) return rc;
            if ((KD::KIND == Kind::Float) && hasnan) return enqueue_big1<KD, true>(ws->data, keys_out, n, desc, canon, s);
            return enqueue_big1<KD, false>(ws->data, keys_out, n, desc, canon, s);
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
            return enqueue_hybrid<KD, PL>(ws, ws->data, const_cast<void *>(keys_in), const_cast<void *>(vals_in),
                                          idx_out, n, hasnan, desc, canon, s);
        }
    }
    return run_sort<KD, PL>(keys_in, keys_out, vals_in, vals_out, idx_out, n, hasnan, desc, canon, s);
}

}  // namespace xss
