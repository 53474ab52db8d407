// xss_kernels.hpp -- CDNA4 (gfx950) kernels of the LSD radix sort.
//
// The reference sorts with a recursive AVX-512 quicksort (partition by
// compress-store, src/xss-common-qsort.h:199-605, pivots from
// src/xss-pivot-selection.hpp).  None of that maps to a GPU; this file
// replaces it with an 8-bit-digit least-significant-digit radix sort.  Each
// of the P = key bits / 8 passes is a stable counting scatter, organised as
// reduce-then-scan over persistent "range" workgroups:
//
//   upsweep_kernel     workgroup b reads its contiguous key range once and
//                      writes the 256-bin histogram of this pass's digit
//   range_scan_kernel  one workgroup turns the [ranges][256] histograms into
//                      the global start of every (range, digit) bucket slice
//   downsweep_kernel   workgroup b walks its range tile by tile (BLOCK*ITEMS
//                      keys): wave64 ballot ranking (8 ballots, v_bitop3 peer
//                      masks, mbcnt), per-wave LDS digit counters, tile-local
//                      sort in LDS, coalesced write-out at running per-digit
//                      bases it keeps itself -- no cross-workgroup waiting
//   block_sort_kernel  one workgroup bitonic-sorts a whole small array (or a
//                      segment) in LDS
//
// A decoupled look-back (onesweep) variant was measured on MI355X and
// rejected: a device-coherent hand-off between workgroups on different XCDs
// costs ~0.55 us, so look-back chains over tens of thousands of tiles were
// latency-bound (4 ms per pass for 2^28 keys against 0.5 ms without them).
//
// Keys travel as their original bits; ord() of xss_keys.hpp is recomputed
// where a digit is needed.  Payloads (kv values, argsort indices) move with
// their keys.  LSD needs a stable pass: items are ranked in input order
// (wave-striped layout, items in order, lanes in order) and a workgroup's
// tiles are written in range order.
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

constexpr int MAX_RANGES = 4096;   // persistent workgroups per pass
constexpr int UP_BLOCK = 256;
constexpr int SCAN_BLOCK = 256;

// Peer mask of the lanes holding the same 8-bit digit: per bit b, the ballot
// B_b of bit b, and every lane keeps the lanes that agree with it,
// peers &= ~(B_b ^ x_b) with x_b = all-ones if its bit is set.  v_bitop3 does
// the and-not-xor of each 32-bit half in one instruction (table 0x90).
__device__ __forceinline__ uint64_t match_any8(uint32_t d)
{
    uint32_t lo = ~0u, hi = ~0u;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        // x: bit b replicated (v_bfe_i32); the ballot compares x itself so the
        // compiler does not re-derive the bit with a second shift
        const uint32_t x = uint32_t(__builtin_amdgcn_sbfe(int32_t(d), b, 1));
        const uint64_t bal = __builtin_amdgcn_uicmp(x, 0u, 33 /* ne */);
        lo = __builtin_amdgcn_bitop3_b32(lo, uint32_t(bal), x, 0x90);
        hi = __builtin_amdgcn_bitop3_b32(hi, uint32_t(bal >> 32), x, 0x90);
    }
    return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

template <int = 0>
__global__ void fill_u64_kernel(uint64_t *p, uint64_t v, uint64_t n)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------- upsweep

struct UpArgs {
    const void *keys;
    uint32_t n;
    uint32_t range_keys;   // keys per range (multiple of the downsweep tile)
    uint32_t shift;
    uint32_t nan_last;
    uint64_t flip;
    uint32_t *rsum;        // [256][ranges] (digit-major)
    uint32_t ranges;
    const uint64_t *range_begin;  // segmented form: explicit key range of
    const uint64_t *range_end;    // each workgroup (null: blockIdx.x * range_keys)
    const uint32_t *nr_dev;       // device-built range tables: the range count
    // passes that skip a constant digit: `ran` [pass] = 1 if that pass moved
    // the keys; with an odd count of moves before this pass they are in keys_alt
    const void *keys_alt;
    const uint32_t *ran;
    uint32_t pass;
    // andor (with ran): AND and OR of the order keys, accumulated by pass 0's
    // upsweep; a later pass whose digit they show constant skips its read
    // (skip_ok: the pass may skip)
    uint64_t *andor;
    uint32_t skip_ok;
    uint64_t *andor_slots;      // pass 0 (andor set): per-workgroup AND / OR, [gridDim.x][2]
    // passes whose plan is made on the device (the MSD-first sort):
    const uint32_t *shift_dev;  // the digit's shift (null: `shift`)
    const uint32_t *gate;       // the pass runs only if *gate > gate_min (null: always)
    uint32_t gate_min;
    uint32_t nr_static;         // ranges walked by a bounded grid (0: one per workgroup)
    uint32_t *wctr;             // ranges handed out by this counter (work stealing), or null
    // per-segment AND / OR of the raw key bits (segmented passes over large
    // buckets): accumulated here (seg_andor_out), or read to skip segments
    // whose keys are all equal (seg_andor_in with seg_canon)
    const uint32_t *rbseg;
    uint64_t *seg_andor_out;
    const uint64_t *seg_andor_in;
    uint32_t seg_canon;
};

// Number of the earlier passes (of a sort that skips constant digits) that
// moved the keys: its parity says which of the two buffers holds them.  The
// flags are written by one workgroup of a downsweep while workgroups on
// other XCDs read the same line, so every access is an agent-scope atomic
// (coherent across the XCDs' L2 caches).
__device__ __forceinline__ uint32_t moves_before(const uint32_t *ran, uint32_t pass)
{
    uint32_t c = 0;
    for (uint32_t p = 0; p < pass; ++p)
        c += __hip_atomic_load(const_cast<uint32_t *>(ran) + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return c;
}
// The digit at `shift` is the same in every key: AND and OR of the order
// keys agree on it (andor written by pass 0's upsweep; wave-uniform).
__device__ __forceinline__ bool digit_constant(const uint64_t *andor, int shift)
{
    const uint64_t x = __hip_atomic_load(const_cast<uint64_t *>(andor), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ^
                       __hip_atomic_load(const_cast<uint64_t *>(andor) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_amdgcn_readfirstlane(uint32_t((x >> shift) & 0xffu) == 0u ? 1u : 0u) != 0u;
}
// 1 if the keys are in the alternate buffer.  Wave-uniform by construction
// (readfirstlane): a pointer chosen from a per-lane load would make the
// compiler treat the pass's loops as divergent, which breaks the ballot-based
// wave-uniform paths of the histogram and ranking code.
__device__ __forceinline__ uint32_t moves_odd(const uint32_t *ran, uint32_t pass)
{
    return __builtin_amdgcn_readfirstlane(moves_before(ran, pass) & 1u);
}

// A pass of a plan made on the device runs only if *gate > gate_min.  The
// plan words are written by an earlier kernel on the stream, so plain
// (cacheable, wave-uniform) loads see them.
__device__ __forceinline__ bool gate_closed(const uint32_t *gate, uint32_t gate_min)
{
    return gate && __builtin_amdgcn_readfirstlane(*gate) <= gate_min;
}
__device__ __forceinline__ uint32_t dev_shift(const uint32_t *shift_dev, uint32_t shift)
{
    return shift_dev ? __builtin_amdgcn_readfirstlane(*shift_dev) : shift;
}
// A segment whose keys are all the same bits needs no pass at all: AND and
// OR of its raw bits agree.  A NaN-canonicalising sort still rewrites a
// segment of equal non-canonical NaNs.
template <class KD>
__device__ __forceinline__ bool segment_constant(const uint64_t *seg_andor, uint32_t seg, bool canon)
{
    if (!seg_andor) return false;
    using U = typename KD::U;
    // written by an earlier kernel (the first large-bucket histogram pass)
    const U a = U(seg_andor[2 * size_t(seg)]), o = U(seg_andor[2 * size_t(seg) + 1]);
    bool c = a == o;
    if constexpr (KD::KIND == Kind::Float) {
        if (c && canon && KD::is_nan(a) && a != KD::QNAN) c = false;
    }
    return __builtin_amdgcn_readfirstlane(c ? 1u : 0u) != 0u;
}
// The next range of a work-stealing walk (all threads call it): one device
// atomic per range by thread 0, broadcast through LDS.
__device__ __forceinline__ uint32_t next_range(uint32_t *wctr, uint32_t *slot)
{
    __syncthreads();  // the previous range's readers of *slot are done
    if (threadIdx.x == 0) *slot = atomicAdd(wctr, 1u);
    __syncthreads();
    return *slot;
}

// Histogram copies of an upsweep workgroup (its waves share them round-robin;
// more copies, fewer same-bank collisions of the LDS adds) and 16-byte loads
// in flight per thread (build-time knobs).
#ifndef XSS_UP_HISTS
#define XSS_UP_HISTS 1
#endif
#ifndef XSS_UP_UNR
#define XSS_UP_UNR 4
#endif
// SEG: the features of the large-bucket passes (per-segment AND / OR,
// skipped segments, ranges from a counter), compiled out of the others.
template <class KD, bool SEG = false>
__global__ __launch_bounds__(UP_BLOCK) void upsweep_kernel(UpArgs a)
{
    using U = typename KD::U;
    constexpr int NH = XSS_UP_HISTS;
    __shared__ uint32_t hh[NH * 256];
    uint32_t *h = hh + ((threadIdx.x >> 6) % NH) * 256;
    // the buffer holding the keys: a.keys, or (a sort that skips constant
    // digits, after an odd number of moves) a.keys_alt.  Chosen by a load
    // from a 2-entry LDS table rather than a select on the pointers: the
    // select of two kernel-argument pointers on a readfirstlane'd flag was
    // compiled to take keys_alt unconditionally (hipcc, ROCm 7.2).
    __shared__ const void *s_kbuf[2];
    __shared__ uint64_t s_andor[2];
    // a device-built table that came out empty: nothing to read (checked
    // first, so that such a launch ends after one load)
    if (a.nr_dev && *a.nr_dev == 0) return;
    if (gate_closed(a.gate, a.gate_min)) return;
    a.shift = dev_shift(a.shift_dev, a.shift);
    // a later pass over a digit every key shares reads nothing
    if (a.andor && a.pass > 0 && a.skip_ok && digit_constant(a.andor, int(a.shift))) return;
    const bool track = a.andor && a.pass == 0;  // pass 0: AND / OR of all order keys
    const bool seg_track = SEG && a.seg_andor_out != nullptr;  // AND / OR of raw bits per segment
    __shared__ uint64_t s_segao[2];
    __shared__ uint32_t s_rg;
    if (threadIdx.x == 0) {
        s_kbuf[0] = a.keys;
        s_kbuf[1] = a.ran ? a.keys_alt : a.keys;
        s_andor[0] = ~0ull;
        s_andor[1] = 0ull;
    }
    const uint32_t odd = a.ran ? moves_odd(a.ran, a.pass) : 0u;
    __syncthreads();
    const U *keys = static_cast<const U *>(s_kbuf[odd]);
    U t_and = U(~U(0)), t_or = U(0);
    // one range per workgroup, or (device-built tables, launched with a
    // bounded grid) every gridDim.x-th range
    const uint32_t nr = a.nr_dev ? *a.nr_dev : a.nr_static ? a.nr_static : gridDim.x;
    if (nr == 0) return;  // an empty table: no counter traffic
    const bool steal = SEG && a.wctr;
    for (uint32_t rg = steal ? next_range(a.wctr, &s_rg) : blockIdx.x; rg < nr;
         rg = steal ? next_range(a.wctr, &s_rg) : rg + gridDim.x) {
        __syncthreads();  // the previous range's histogram was written out
        // a segment of equal keys takes no pass (its digit counts go unused)
        if (SEG && a.seg_andor_in && segment_constant<KD>(a.seg_andor_in, a.rbseg[rg], a.seg_canon != 0)) continue;
        for (int i = threadIdx.x; i < NH * 256; i += UP_BLOCK) hh[i] = 0;
        if (seg_track && threadIdx.x == 0) {
            s_segao[0] = ~0ull;
            s_segao[1] = 0ull;
        }
        U r_and = U(~U(0)), r_or = U(0);
        __syncthreads();
        const U flip = U(a.flip);
        const bool nl = a.nan_last != 0;
        const int shift = int(a.shift);
        const uint32_t r0 = a.range_begin ? uint32_t(a.range_begin[rg]) : rg * a.range_keys;
        const uint32_t r1 = a.range_begin ? uint32_t(a.range_end[rg]) : min(a.n, r0 + a.range_keys);
        constexpr int VEC = 16 / sizeof(U);
        // keys before the first 16-byte boundary are read one by one
        const uint32_t mis = uint32_t(reinterpret_cast<uintptr_t>(keys + r0) & 15u);
        const uint32_t head = min(r1 - r0, uint32_t(((16u - mis) & 15u) / sizeof(U)));
        const uint32_t v0 = r0 + head;
        const uint32_t nvec = (r1 - v0) / VEC;
        const uint4 *kv = reinterpret_cast<const uint4 *>(keys + v0);
        // count one 16-byte vector of keys; whole-wave runs of one digit (sorted
        // or constant data) add once
        auto count_vec = [&](const uint4 &q) {
            const U *e = reinterpret_cast<const U *>(&q);
#pragma unroll
            for (int j = 0; j < VEC; ++j) {
                if (track) {
                    const U o = KD::ord(e[j], flip, nl);
                    t_and &= o;
                    t_or |= o;
                }
                if (seg_track) {
                    r_and &= e[j];
                    r_or |= e[j];
                }
                const uint32_t d = KD::digit(e[j], flip, nl, shift);
                const uint64_t act = __ballot(1);
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
                if (__ballot(d != d0) == 0) {
                    if (lanes_below(act) == 0) atomicAdd(&h[d0], uint32_t(__popcll(act)));
                } else {
                    atomicAdd(&h[d], 1u);
                }
            }
        };
        // XSS_UP_UNR loads in flight per thread before any is consumed
        constexpr uint32_t UNR = XSS_UP_UNR;
        uint32_t v = threadIdx.x;
        for (; v + (UNR - 1) * UP_BLOCK < nvec; v += UNR * UP_BLOCK) {
            uint4 q[UNR];
#pragma unroll
            for (uint32_t u = 0; u < UNR; ++u) q[u] = kv[v + u * UP_BLOCK];
#pragma unroll
            for (uint32_t u = 0; u < UNR; ++u) count_vec(q[u]);
        }
        for (; v < nvec; v += UP_BLOCK) count_vec(kv[v]);
        const uint32_t t0 = v0 + nvec * VEC;
        const uint32_t rest = head + (r1 - t0);
        for (uint32_t r = threadIdx.x; r < rest; r += UP_BLOCK) {
            const uint32_t i = r < head ? r0 + r : t0 + (r - head);
            if (track) {
                const U o = KD::ord(keys[i], flip, nl);
                t_and &= o;
                t_or |= o;
            }
            if (seg_track) {
                r_and &= keys[i];
                r_or |= keys[i];
            }
            atomicAdd(&h[KD::digit(keys[i], flip, nl, shift)], 1u);
        }
        if (seg_track) {
            atomicAnd(reinterpret_cast<unsigned long long *>(&s_segao[0]),
                      (unsigned long long)r_and | ~(unsigned long long)U(~U(0)));
            atomicOr(reinterpret_cast<unsigned long long *>(&s_segao[1]), (unsigned long long)r_or);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < 256; i += UP_BLOCK) {
            uint32_t c = 0;
#pragma unroll
            for (int j = 0; j < NH; ++j) c += hh[j * 256 + i];
            a.rsum[size_t(i) * a.ranges + rg] = c;
        }
        if (seg_track && threadIdx.x == 0) {
            uint64_t *so = a.seg_andor_out + 2 * size_t(a.rbseg[rg]);
            atomicAnd(reinterpret_cast<unsigned long long *>(so), (unsigned long long)s_segao[0]);
            atomicOr(reinterpret_cast<unsigned long long *>(so + 1), (unsigned long long)s_segao[1]);
        }
    }
    if (track) {
        atomicAnd(reinterpret_cast<unsigned long long *>(&s_andor[0]), (unsigned long long)t_and | ~(unsigned long long)U(~U(0)));
        atomicOr(reinterpret_cast<unsigned long long *>(&s_andor[1]), (unsigned long long)t_or);
        __syncthreads();
        // one slot per workgroup (a later kernel folds them): thousands of
        // device atomics on one address would serialise
        if (threadIdx.x == 0) {
            a.andor_slots[2 * blockIdx.x] = s_andor[0];
            a.andor_slots[2 * blockIdx.x + 1] = s_andor[1];
        }
    }
}

// AND / OR of the order keys from the upsweep's per-workgroup slots (all
// threads of the calling workgroup; the result in out[0], out[1])
__device__ __forceinline__ void fold_andor(const uint64_t *slots, uint32_t nslots, uint64_t *out)
{
    __shared__ unsigned long long s_ao[2];
    if (threadIdx.x == 0) {
        s_ao[0] = ~0ull;
        s_ao[1] = 0ull;
    }
    __syncthreads();
    unsigned long long x = ~0ull, y = 0ull;
    for (uint32_t i = threadIdx.x; i < nslots; i += blockDim.x) {
        x &= slots[2 * i];
        y |= slots[2 * i + 1];
    }
    atomicAnd(&s_ao[0], x);
    atomicOr(&s_ao[1], y);
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = s_ao[0];
        out[1] = s_ao[1];
    }
}

// AND = all ones, OR = 0 before a sort that skips constant digits
template <int = 0>
__global__ void andor_init_kernel(uint64_t *andor)
{
    andor[0] = ~0ull;
    andor[1] = 0ull;
}

// ---------------------------------------------------------------- range scan

struct ScanArgs {
    const uint32_t *rsum;  // [256][ranges]
    uint32_t *rpre;        // [256][ranges] exclusive prefix over ranges, per digit
    uint32_t *total;       // [256] keys per digit
    uint32_t ranges;
    const uint64_t *andor_slots;  // pass 0 of a sort that skips constant digits: fold the
    uint64_t *andor;              // upsweep's per-workgroup AND / OR into andor (workgroup 0)
};

// workgroup d scans digit d's column (ranges <= 16 * SCAN_BLOCK)
template <int = 0>
__global__ __launch_bounds__(SCAN_BLOCK) void range_scan_kernel(ScanArgs a)
{
    constexpr int PER = MAX_RANGES / SCAN_BLOCK;
    __shared__ uint32_t wsum[SCAN_BLOCK / 64];
    const uint32_t d = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t *col = a.rsum + size_t(d) * a.ranges;
    uint32_t *out = a.rpre + size_t(d) * a.ranges;
    uint32_t v[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const uint32_t b = t * PER + i;
        v[i] = b < a.ranges ? col[b] : 0u;
        sum += v[i];
    }
    uint32_t incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t x = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += x;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (uint32_t i = 0; i < w; ++i) run += wsum[i];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const uint32_t b = t * PER + i;
        if (b < a.ranges) out[b] = run;
        run += v[i];
    }
    if (t == SCAN_BLOCK - 1) a.total[d] = run;
    if (a.andor_slots && d == 0) fold_andor(a.andor_slots, a.ranges, a.andor);
}

// exclusive scan of the 256 digit totals by the first 256 threads of a
// workgroup (all threads must call it): returns the global start of digit tid
__device__ __forceinline__ uint32_t digit_starts(const uint32_t *total, uint32_t *wsum4)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t tot = tid < 256 ? total[tid] : 0u;
    uint32_t incl = tot;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t x = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += x;
    }
    if (lane == 63 && w < 4) wsum4[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t i = 0; i < w && i < 4; ++i) pre += wsum4[i];
    return pre + incl - tot;
}

// ---------------------------------------------------------------- downsweep

// Scatter ranking: 0 = each digit's run base read from the wave's counter and
// the last peer writes it back (default); 1 = the lowest peer reserves the
// run with a returning LDS add and shuffles the base (as the bucket sort
// does), measured 4% slower on the 16K-key tiles (build-time knob).
#ifndef XSS_RANK_ATOMIC
#define XSS_RANK_ATOMIC 0
#endif

struct PassArgs {
    const void *keys_in;
    void *keys_out;
    const void *vals_in;
    void *vals_out;
    uint64_t *idx_out;       // PL_IDX, last pass (size_t indices)
    uint32_t *idx32_out;     // PL_IDX, last pass: segment-local 32-bit indices instead
    uint32_t n;
    uint32_t shift;
    uint64_t flip;
    uint32_t nan_last;
    uint32_t first;          // PL_IDX: generate indices instead of loading them
    uint32_t last;           // PL_IDX: write size_t indices only
    uint32_t canon_nan;      // qsort(hasnan): write canonical NaN bits
    uint32_t range_tiles;    // tiles per workgroup
    uint32_t ranges;
    const uint32_t *rpre;    // [256][ranges]
    const uint32_t *total;   // [256]
    // a sort that skips constant digits (keys and values ping-pong between
    // the *_in and *_out buffers by the parity of the earlier moves, see
    // UpArgs): `ran` [pass] records whether this pass moved the keys; it
    // skips when one digit holds all n keys and skip_ok is set
    uint32_t *ran;
    uint32_t pass;
    uint32_t skip_ok;
    const uint64_t *andor;  // AND / OR of the order keys (pass 0's upsweep), or null
    const uint32_t *shift_dev;  // plans made on the device: the digit's shift (null: `shift`)
    const uint32_t *gate;       // and the pass runs only if *gate > gate_min (null: always)
    uint32_t gate_min;
    // a pass that is the last only if the plan says so: last = (*last_gate
    // == last_q + 1), and it then writes canonical NaNs if canon_if_last
    const uint32_t *last_gate;
    uint32_t last_q;
    uint32_t canon_if_last;
};

// 1: the scatter keeps each sorted slot's digit in an LDS byte array; 0:
// the write-out recomputes it from the key (one LDS write and read fewer
// per key; build-time knob)
#ifndef XSS_DIG_LDS
#define XSS_DIG_LDS 0
#endif
// DIG: keep each sorted slot's digit in an LDS byte array (the default for
// radix digits is XSS_DIG_LDS; digits that cost more than a shift, such as
// a sample sort's bucket, are always kept)
template <class KD, int PL, int BLOCK, int ITEMS, bool DIG = XSS_DIG_LDS != 0>
struct DownsweepSmem {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr int WAVES = BLOCK / 64;
    static constexpr int TILE = BLOCK * ITEMS;
    uint32_t off[WAVES * 256];   // per-wave digit counters, then tile slot bases
    uint64_t base[256];          // global position of tile slot 0 of each digit
    uint32_t wsum[WAVES < 4 ? 4 : WAVES];
    U keys[TILE];                // the tile, sorted by digit
    uint8_t dig[DIG ? TILE : 1];  // digit of each sorted slot (DIG)
    V vals[PL == PL_NONE ? 1 : TILE];
};

// The digit of a radix pass: 8 bits of the order key at `shift`.
template <class KD, bool NL>
struct RadixDigit {
    using U = typename KD::U;
    U flip;
    int shift;
    __device__ __forceinline__ uint32_t operator()(U x) const { return KD::digit(x, flip, NL, shift); }
};

// One tile of a downsweep: load, rank, local sort, write-out.  FULL: all
// TILE keys valid (no per-item bounds checks); NL: floats with hasnan.
// The tile holds keys [tbase, tbase + valid) (valid == TILE when FULL).
// seg_base: start of the segment the tile belongs to (segment-local argsort
// indices are position - seg_base).  dig maps a key to its digit (0..255):
// the radix digit of the pass, or (the sample sort) its bucket among a
// segment's splitters.
template <class KD, int PL, int BLOCK, int ITEMS, bool FULL, bool NL, bool UNST, class DG, bool DIGC>
__device__ __forceinline__ void downsweep_tile_g(const PassArgs &a, DownsweepSmem<KD, PL, BLOCK, ITEMS, DIGC> &sm,
                                                 uint64_t tbase, uint32_t valid_in, uint64_t seg_base, uint64_t &run,
                                                 const DG &dig)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    constexpr bool HAS_PL = PL != PL_NONE;
    constexpr int WAVES = BLOCK / 64;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int WTILE = 64 * ITEMS;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const U *kin = static_cast<const U *>(a.keys_in);
    const V *vin = static_cast<const V *>(a.vals_in);

    for (uint32_t i = tid; i < WAVES * 256; i += BLOCK) sm.off[i] = 0;
    const uint32_t valid = FULL ? uint32_t(TILE) : valid_in;
    U key[ITEMS];
    V val[ITEMS];
    const uint32_t lbase = wave * WTILE + lane;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t lp = lbase + i * 64;
        if (FULL || lp < valid) {
            key[i] = kin[tbase + lp];
            if constexpr (PL == PL_IDX) val[i] = a.first ? uint32_t(tbase + lp) : vin[tbase + lp];
            else if constexpr (HAS_PL) val[i] = vin[tbase + lp];
        } else {
            key[i] = U(0);
            if constexpr (HAS_PL) val[i] = V(0);
        }
    }
    __syncthreads();  // counters cleared; the previous tile's LDS reads are done

    // --- rank: stable within the tile (wave-striped order), per-wave counters
    uint32_t rd[ITEMS];  // rank in the wave's digit run (low 16 bits), digit (bits 16..23)
    uint32_t *wc = sm.off + wave * 256;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t lp = lbase + i * 64;
        const bool real = FULL || lp < valid;
        const uint32_t d = real ? dig(key[i]) : 255u;
        if constexpr (UNST && PL != PL_IDX) {
            // keys only, or key-value (whose payloads inside a run of equal
            // keys may come in any order): each lane reserves its own slot
            // with one returning LDS add, no peer match.  A wave whose lanes share one digit (sorted
            // runs) reserves its 64 slots with one add by lane 0, since
            // same-address adds serialise; bit 24 marks it, and each lane's
            // slot is then lane 0's base + lane (resolved after the barrier).
            // The padding of a partial tile takes no slot (bit 25).
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            const bool uni = __ballot(!real || d != d0) == 0;
            uint32_t old = 0;
            if (real && (!uni || lane == 0)) old = atomicAdd(&wc[d], uni ? 64u : 1u);
            rd[i] = old | (d << 16) | (uni ? (1u << 24) : 0u) | (real ? 0u : (1u << 25));
            continue;
        }
        const uint64_t peers = match_any8(d);
        const uint32_t below = lanes_below(peers);
#if XSS_RANK_ATOMIC
        // the lowest lane of each digit reserves the run with one returning
        // LDS add; items do not wait on each other's counter update
        uint32_t old = 0;
        if (below == 0) old = atomicAdd(&wc[d], uint32_t(__popcll(peers)));
        old = __shfl(old, int(__builtin_ctzll(peers)));
        rd[i] = (old + below) | (d << 16);
#else
        const uint32_t tot = uint32_t(__popcll(peers));
        const uint32_t c = wc[d];
        rd[i] = (c + below) | (d << 16);
        if (below + 1 == tot) wc[d] = c + tot;
#endif
    }
    __syncthreads();

    // --- digit counts of the tile, their exclusive scan, and the slot base of
    // every (wave, digit): off[w][d] = start of digit d in the tile + the
    // counts of d in waves < w
    uint32_t tcount = 0;
    if (tid < 256) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) tcount += sm.off[w * 256 + tid];
    }
    uint32_t incl = tcount;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= uint32_t(off)) incl += t;
    }
    if (lane == 63) sm.wsum[wave] = incl;
    __syncthreads();
    if (tid < 256) {
        uint32_t st = incl - tcount;
        for (uint32_t w = 0; w < wave; ++w) st += sm.wsum[w];
        sm.base[tid] = run - st;
        // padding of a partial last tile was ranked as digit 255 (last),
        // except by the unstable ranking, where it takes no slot
        constexpr bool PAD_SLOTS = !(UNST && PL != PL_IDX);
        run += tcount - ((!FULL && PAD_SLOTS && tid == 255) ? (TILE - valid) : 0u);
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = sm.off[w * 256 + tid];
            sm.off[w * 256 + tid] = st;
            st += c;
        }
    }
    __syncthreads();

    // --- tile-local sort: scatter into LDS at slot base + rank
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        uint32_t r = rd[i];
        if constexpr (UNST && PL != PL_IDX) {
            if (r >> 25) continue;  // padding
            if (r >> 24) r = (__builtin_amdgcn_readfirstlane(r) + lane) & 0xffffffu;
        }
        const uint32_t d = r >> 16;
        const uint32_t lpos = wc[d] + (r & 0xffffu);
        sm.keys[lpos] = key[i];
        if constexpr (DIGC) sm.dig[lpos] = uint8_t(d);
        if constexpr (HAS_PL) sm.vals[lpos] = val[i];
    }
    __syncthreads();

    // --- write-out: consecutive threads take consecutive sorted slots
    U *kout = static_cast<U *>(a.keys_out);
    V *vout = static_cast<V *>(a.vals_out);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t lpos = i * BLOCK + tid;
        if (FULL || lpos < valid) {
            const U x = sm.keys[lpos];
            // the slot's digit: kept in LDS, or recomputed from its key
            const uint64_t g = sm.base[DIGC ? uint32_t(sm.dig[lpos]) : dig(x)] + lpos;
            if constexpr (PL == PL_IDX) {
                if (a.last) {
                    if (a.idx32_out) a.idx32_out[g] = uint32_t(sm.vals[lpos] - uint32_t(seg_base));
                    else a.idx_out[g] = uint64_t(sm.vals[lpos]);
                    continue;
                }
            }
            if constexpr (KD::KIND == Kind::Float && NL) {
                kout[g] = (a.canon_nan && KD::is_nan(x)) ? KD::QNAN : x;
            } else {
                kout[g] = x;
            }
            if constexpr (HAS_PL) vout[g] = sm.vals[lpos];
        }
    }
}

template <class KD, int PL, int BLOCK, int ITEMS, bool FULL, bool NL, bool UNST = false, bool DIGC>
__device__ __forceinline__ void downsweep_tile(const PassArgs &a, DownsweepSmem<KD, PL, BLOCK, ITEMS, DIGC> &sm,
                                               uint64_t tbase, uint32_t valid_in, uint64_t seg_base, uint64_t &run)
{
    const RadixDigit<KD, NL> dig{typename KD::U(a.flip), int(a.shift)};
    downsweep_tile_g<KD, PL, BLOCK, ITEMS, FULL, NL, UNST>(a, sm, tbase, valid_in, seg_base, run, dig);
}

// Tile shape and register budget of the scatter passes, per class of item
// width (key + payload bytes): C0 <= 2 (16-bit keys), C1 = 4 (32-bit keys),
// C2 = 5..8 (64-bit keys, or a 16/32-bit key with a 32-bit payload or
// index), C3 > 8.  Build-time knobs; the defaults are the measured best on
// MI355X.  A workgroup's waves must fit the register file: 1024 threads
// leave 128 VGPRs per lane, 512 threads at most 256 (XSS_SCATTER_WAVES_Cn
// caps the budget at 512 / waves VGPRs; 1 = the compiler's choice).  Larger
// tiles write longer runs per digit, so fewer partially written lines.
#ifndef XSS_PASS_BLOCK_C0
#define XSS_PASS_BLOCK_C0 512
#endif
#ifndef XSS_PASS_ITEMS_C0
#define XSS_PASS_ITEMS_C0 16
#endif
#ifndef XSS_SCATTER_WAVES_C0
#define XSS_SCATTER_WAVES_C0 1
#endif
#ifndef XSS_PASS_BLOCK_C1
#define XSS_PASS_BLOCK_C1 1024
#endif
#ifndef XSS_PASS_ITEMS_C1
#define XSS_PASS_ITEMS_C1 16
#endif
#ifndef XSS_SCATTER_WAVES_C1
#define XSS_SCATTER_WAVES_C1 1
#endif
#ifndef XSS_PASS_BLOCK_C2
#define XSS_PASS_BLOCK_C2 1024
#endif
#ifndef XSS_PASS_ITEMS_C2
#define XSS_PASS_ITEMS_C2 12
#endif
#ifndef XSS_SCATTER_WAVES_C2
#define XSS_SCATTER_WAVES_C2 1
#endif
#ifndef XSS_PASS_BLOCK_C3
#define XSS_PASS_BLOCK_C3 1024
#endif
#ifndef XSS_PASS_ITEMS_C3
#define XSS_PASS_ITEMS_C3 8
#endif
#ifndef XSS_SCATTER_WAVES_C3
#define XSS_SCATTER_WAVES_C3 1
#endif
template <class KD, int PL>
struct ScatterClass {
    static constexpr int bytes =
        int(sizeof(typename KD::U)) + (PL == PL_NONE ? 0 : int(sizeof(typename PayloadT<PL>::type)));
    static constexpr int value = bytes <= 2 ? 0 : bytes <= 4 ? 1 : bytes <= 8 ? 2 : 3;
};
// 512-thread tiles of 4-byte keys (the shape for smaller arrays) get 128
// VGPRs = two workgroups per CU.
template <class KD, int PL, int BLOCK>
struct ScatterWaves {
    static constexpr int c = ScatterClass<KD, PL>::value;
    static constexpr int value = (BLOCK <= 512 && ScatterClass<KD, PL>::bytes == 4) ? 4
                               : c == 0 ? XSS_SCATTER_WAVES_C0 : c == 1 ? XSS_SCATTER_WAVES_C1
                               : c == 2 ? XSS_SCATTER_WAVES_C2 : XSS_SCATTER_WAVES_C3;
};
#define XSS_SCATTER_OCCUPANCY(KD, PL, BLOCK) \
    __attribute__((amdgpu_waves_per_eu(ScatterWaves<KD, PL, BLOCK>::value)))

// UNST (not argsort): full tiles rank without keeping the order inside a
// digit; for a pass whose input order does not matter (the first LSD pass,
// the MSD passes of the MSD-first sort).
template <class KD, int PL, int BLOCK, int ITEMS, bool NL, bool UNST = false>
__global__ __launch_bounds__(BLOCK) XSS_SCATTER_OCCUPANCY(KD, PL, BLOCK) void downsweep_kernel(PassArgs a)
{
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(BLOCK >= 256 && BLOCK % 64 == 0, "one thread per digit needed");
    static_assert(64 * ITEMS <= 65535, "rank packing");
    __shared__ DownsweepSmem<KD, PL, BLOCK, ITEMS> sm;
    const uint32_t tid = threadIdx.x;
    const uint32_t ntiles = (a.n + TILE - 1) / TILE;
    const uint32_t tile0 = blockIdx.x * a.range_tiles;
    const uint32_t tile1 = min(ntiles, tile0 + a.range_tiles);
    if (gate_closed(a.gate, a.gate_min)) return;
    PassArgs b = a;
    b.shift = dev_shift(a.shift_dev, a.shift);
    if (a.ran) {
        // all n keys in one digit: a stable pass leaves them in place
        const bool constant = a.andor ? digit_constant(a.andor, int(a.shift))
                                      : __syncthreads_or(tid < 256 && a.total[tid] == a.n) != 0;
        const bool skip = constant && a.skip_ok;
        if (blockIdx.x == 0 && tid == 0)
            __hip_atomic_store(a.ran + a.pass, skip ? 0u : 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (skip) return;
        if (moves_odd(a.ran, a.pass)) {  // the keys are in the *_out buffers
            b.keys_in = a.keys_out;
            b.keys_out = const_cast<void *>(a.keys_in);
            b.vals_in = a.vals_out;
            b.vals_out = const_cast<void *>(a.vals_in);
        }
    }
    // running global position of each digit of this range (thread d owns digit d)
    const uint32_t dstart = digit_starts(b.total, sm.wsum);
    uint64_t run = tid < 256 ? dstart + b.rpre[size_t(tid) * b.ranges + blockIdx.x] : 0u;
    const uint32_t nfull = min(tile1, b.n / TILE);
    uint32_t tile = tile0;
    for (; tile < nfull; ++tile)
        downsweep_tile<KD, PL, BLOCK, ITEMS, true, NL, UNST>(b, sm, uint64_t(tile) * TILE, TILE, 0, run);
    for (; tile < tile1; ++tile)
        downsweep_tile<KD, PL, BLOCK, ITEMS, false, NL, UNST>(b, sm, uint64_t(tile) * TILE, b.n - tile * TILE, 0, run);
}

// After a sort that skips constant digits: an odd number of moves left the
// keys (and values) in the scratch buffers; copy them back.
template <int = 0>
__global__ __launch_bounds__(256) void skip_finish_kernel(void *keys, const void *keys_tmp, void *vals,
                                                          const void *vals_tmp, uint64_t key_bytes,
                                                          uint64_t val_bytes, const uint32_t *ran, uint32_t passes)
{
    if (!moves_odd(ran, passes)) return;
    const uint64_t stride = uint64_t(gridDim.x) * 256u;
    auto copy = [&](void *dst, const void *src, uint64_t bytes) {
        // both buffers are 256-byte aligned allocations or the caller's array:
        // 16-byte vectors where both ends allow it, bytes otherwise
        const bool vec = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) == 0;
        if (vec) {
            const uint64_t nv = bytes / 16;
            uint4 *d = static_cast<uint4 *>(dst);
            const uint4 *q = static_cast<const uint4 *>(src);
            for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < nv; i += stride) d[i] = q[i];
            for (uint64_t i = nv * 16 + uint64_t(blockIdx.x) * 256u + threadIdx.x; i < bytes; i += stride)
                static_cast<uint8_t *>(dst)[i] = static_cast<const uint8_t *>(src)[i];
        } else {
            for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < bytes; i += stride)
                static_cast<uint8_t *>(dst)[i] = static_cast<const uint8_t *>(src)[i];
        }
    };
    copy(keys, keys_tmp, key_bytes);
    if (vals) copy(vals, vals_tmp, val_bytes);
}

// ---------------------------------------------------------------- segmented passes
//
// A segmented LSD pass runs over a table of ranges: every range lies inside
// one segment, and a segment's ranges are consecutive in the table.  Used by
// the batch form (host-built tables) and by the MSD-first one-array sort
// (tables built on the device; their count is then read from nr_dev and
// grids are launched at the table's capacity, surplus workgroups exiting).

struct SegScanArgs {
    const uint32_t *rsum;    // [256][stride]
    uint32_t *rpre;          // [256][stride] exclusive prefix within the segment
    uint32_t *segtot;        // [segments][256] digit totals per segment
    const uint32_t *head;    // [nr] 1 at the first range of a segment
    const uint32_t *rbseg;   // [nr] segment index of each range
    uint32_t nr;             // ranges (when nr_dev is null)
    uint32_t stride;         // row length of rsum / rpre
    const uint32_t *nr_dev;  // device-built tables: the range count
    const uint32_t *gate;    // the pass runs only if *gate > gate_min (null: always)
    uint32_t gate_min;
    uint32_t *wctr_reset;    // work-stealing counters of the pass's downsweep, cleared here
};

// workgroup d: segmented exclusive scan of column d over the ranges
template <int = 0>
__global__ __launch_bounds__(256) void seg_scan_kernel(SegScanArgs a)
{
    __shared__ uint32_t s_sum[256];
    __shared__ uint32_t s_flag[256];
    if (gate_closed(a.gate, a.gate_min)) return;
    if (a.wctr_reset && blockIdx.x == 0 && threadIdx.x == 0) *a.wctr_reset = 0;
    const uint32_t d = blockIdx.x, t = threadIdx.x;
    const uint32_t nr = a.nr_dev ? *a.nr_dev : a.nr;
    const uint32_t per = (nr + 255) / 256;
    const uint32_t c0 = min(nr, t * per), c1 = min(nr, c0 + per);
    const uint32_t *col = a.rsum + size_t(d) * a.stride;
    // chunk summary: sum since the chunk's last head, and whether it has one
    uint32_t sum = 0, flag = 0;
    for (uint32_t r = c0; r < c1; ++r) {
        if (a.head[r]) {
            sum = 0;
            flag = 1;
        }
        sum += col[r];
    }
    s_sum[t] = sum;
    s_flag[t] = flag;
    __syncthreads();
    // inclusive segmented scan over the 256 chunk summaries (Hillis-Steele)
    for (uint32_t off = 1; off < 256; off <<= 1) {
        uint32_t ps = 0, pf = 0;
        if (t >= off) {
            ps = s_sum[t - off];
            pf = s_flag[t - off];
        }
        __syncthreads();
        if (t >= off && !flag) {
            sum += ps;
            flag |= pf;
        }
        s_sum[t] = sum;
        s_flag[t] = flag;
        __syncthreads();
    }
    uint32_t carry = t > 0 ? s_sum[t - 1] : 0u;  // running sum entering this chunk
    uint32_t *out = a.rpre + size_t(d) * a.stride;
    for (uint32_t r = c0; r < c1; ++r) {
        if (a.head[r]) carry = 0;
        out[r] = carry;
        carry += col[r];
        if (r + 1 == nr || a.head[r + 1]) a.segtot[size_t(a.rbseg[r]) * 256 + d] = carry;
    }
}

struct SegRanges {
    const uint64_t *begin;   // [nr] first key of each range
    const uint64_t *end;     // [nr]
    const uint64_t *sbase;   // [nr] first key of the range's segment
    const uint32_t *rbseg;   // [nr] segment index
    const uint32_t *rpre;    // [256][stride]
    const uint32_t *segtot;  // [segments][256]
    uint32_t nr;
    uint32_t stride;
    const uint32_t *nr_dev;  // device-built tables: the range count
    uint32_t *wctr;          // ranges handed out by this counter (work stealing), or null
    // skip segments whose keys are all equal (segment_constant); the last
    // pass of an argsort then writes such a segment's indices from home_vals,
    // where every pass left them
    const uint64_t *seg_andor;
    uint32_t seg_canon;
    const void *home_vals;
};

// Tile shape of the segmented passes: 512 threads x 16 items (8 for items
// wider than 8 bytes), smaller than the one-array pass for 4-byte keys
// because a segment of a few tiles wastes less of a smaller last tile.  The
// host-side plan (shape_of in xss_segmented.hip) cuts ranges with the same
// tile size.
// Keys-only 4-byte tiles get a 128-VGPR budget (two workgroups per CU).
#ifndef XSS_SEG_BLOCK_C1
#define XSS_SEG_BLOCK_C1 512   // segmented-pass tile of 4-byte items (build-time knobs)
#endif
#ifndef XSS_SEG_ITEMS_C1
#define XSS_SEG_ITEMS_C1 16
#endif
template <class KD, int PL>
struct SegShape {
    static constexpr int BYTES = ScatterClass<KD, PL>::bytes;
    static constexpr int BLOCK = BYTES == 4 ? XSS_SEG_BLOCK_C1 : 512;
    static constexpr int ITEMS = BYTES > 8 ? 8 : BYTES == 4 ? XSS_SEG_ITEMS_C1 : 16;
    static constexpr int WAVES = BYTES == 4 ? 4 : 1;
};

// SEG: the features of the large-bucket passes (plan gates, the last pass
// chosen on the device, skipped segments, ranges from a counter).
template <class KD, int PL, int BLOCK, int ITEMS, bool NL, bool UNST = false, bool SEG = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(SegShape<KD, PL>::WAVES)))
void seg_downsweep_kernel(PassArgs a, SegRanges r)
{
    constexpr int TILE = BLOCK * ITEMS;
    __shared__ DownsweepSmem<KD, PL, BLOCK, ITEMS> sm;
    __shared__ uint32_t s_b;
    const uint32_t tid = threadIdx.x;
    const uint32_t nr = r.nr_dev ? *r.nr_dev : r.nr;
    if (nr == 0) return;  // an empty table
    if constexpr (SEG) {
        if (gate_closed(a.gate, a.gate_min)) return;
        if (a.last_gate) {
            a.last = dev_shift(a.last_gate, 0) == a.last_q + 1;
            a.canon_nan = a.last ? a.canon_if_last : 0u;
        }
    }
    a.shift = dev_shift(a.shift_dev, a.shift);
    // one range per workgroup, or (device-built tables, launched with a
    // bounded grid) every gridDim.x-th range, or ranges taken from a counter
    const bool steal = SEG && r.wctr;
    for (uint32_t b = steal ? next_range(r.wctr, &s_b) : blockIdx.x; b < nr;
         b = steal ? next_range(r.wctr, &s_b) : b + gridDim.x) {
        __syncthreads();  // the previous range's LDS reads are done
        const uint64_t rb = r.begin[b], re = r.end[b], sbase = r.sbase[b];
        if (SEG && r.seg_andor && segment_constant<KD>(r.seg_andor, r.rbseg[b], r.seg_canon != 0)) {
            // all keys equal: the segment stays where it is; an argsort's
            // last pass writes its indices (in order: every pass is stable)
            if constexpr (PL == PL_IDX) {
                if (a.last) {
                    const uint32_t *hv = static_cast<const uint32_t *>(r.home_vals);
                    for (uint64_t i = rb + tid; i < re; i += BLOCK) {
                        if (a.idx32_out) a.idx32_out[i] = hv[i] - uint32_t(sbase);
                        else a.idx_out[i] = uint64_t(hv[i]);
                    }
                }
            }
            continue;
        }
        // digit starts inside this range's segment
        const uint32_t dstart = digit_starts(r.segtot + size_t(r.rbseg[b]) * 256, sm.wsum);
        uint64_t run = tid < 256 ? sbase + dstart + r.rpre[size_t(tid) * r.stride + b] : 0u;
        uint64_t t = rb;
        for (; t + TILE <= re; t += TILE)
            downsweep_tile<KD, PL, BLOCK, ITEMS, true, NL, UNST>(a, sm, t, TILE, sbase, run);
        if (t < re) downsweep_tile<KD, PL, BLOCK, ITEMS, false, NL, UNST>(a, sm, t, uint32_t(re - t), sbase, run);
    }
}


// ---------------------------------------------------------------- block sort

struct BlockSortArgs {
    const void *keys_in;
    void *keys_out;
    const void *vals_in;
    void *vals_out;
    uint64_t *idx_out;         // PL_IDX
    uint32_t *idx32_out;       // PL_IDX: segment-local 32-bit indices instead
    const uint64_t *arr_keys;  // pointer-list form: array s at arr_keys[s] (values at
    const uint64_t *arr_vals;  // arr_vals[s]) with arr_len[s] keys, sorted in place;
    const uint32_t *arr_len;   // null: the segment forms below
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
    if (a.arr_keys) {
        // pointer-list form: rebase this workgroup onto array `seg`
        a.keys_in = a.keys_out = reinterpret_cast<void *>(a.arr_keys[seg]);
        if (a.arr_vals) a.vals_in = a.vals_out = reinterpret_cast<void *>(a.arr_vals[seg]);
    }
    const uint64_t base = a.seg_off ? a.seg_off[seg] : 0;
    const uint32_t n = a.arr_keys ? a.arr_len[seg] : a.seg_off ? uint32_t(a.seg_off[seg + 1] - base) : a.n;
    if (n < 2) {
        if (n == 1) {
            const U *kin = static_cast<const U *>(a.keys_in);
            if (threadIdx.x == 0) {
                if constexpr (PL == PL_IDX) {
                    if (a.idx32_out) a.idx32_out[base] = 0;
                    else a.idx_out[base] = a.idx_local ? 0 : base;
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
            if (a.idx32_out) a.idx32_out[base + i] = src;
            else a.idx_out[base + i] = a.idx_local ? uint64_t(src) : base + src;
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
