// xss_count.hpp -- counting sort of large arrays of 16-bit keys (qsort of
// uint16_t, int16_t and _Float16).
//
// A keys-only sort of 16-bit keys is fixed by how many keys carry each of
// the 65536 order-key values, so it is one read and one write of the array:
//
//   cs_zero      clear the overflow counters
//   cs_hist      one workgroup per CU counts its contiguous share of the keys
//                into 65536 16-bit LDS bins (128 KiB; two bins per 32-bit
//                word; a bin reaching CS_SPILL moves CS_SPILL to a global
//                overflow counter), then writes its packed bins to its row
//                of a [workgroups][32768] word table
//   cs_bins      per bin: the column sum over the rows (+ overflow), exclusive
//                scan inside blocks of 256 bins
//   cs_starts    start of every bin in the output (256 workgroups, each adds
//                the total of the blocks before its own)
//   cs_expand    each wave writes a span of output positions as runs of the
//                bins that overlap it; it reads the ends of 64 bins at once
//                (one load per 64 bins, empty bins skipped), writes the short
//                runs lane by lane and the long ones with the whole wave, and
//                finds its first bin (or the next after a gap of more than 64
//                empty bins) with a three-level 64-way search
//
// Keys are written back from their order key, so the run of a bin holds
// one bit pattern: every NaN of a hasnan sort is the bin above +inf and is
// written as the canonical NaN (src/xss-custom-float.h:81-85), as qsort
// does; -0.0 and +0.0 are different order keys, so both keep their bits.
#pragma once
#include "xss_sort.hpp"

#ifndef XSS_CS_WAVE_SPAN
#define XSS_CS_WAVE_SPAN 8192
#endif
#ifndef XSS_CS_UNR
#define XSS_CS_UNR 4
#endif

namespace xss {

constexpr int CS_BLOCK = 1024;
constexpr uint32_t CS_BINS = 65536;
constexpr uint32_t CS_WAVE_SPAN = XSS_CS_WAVE_SPAN;  // output positions per expand wave (multiple of 8)
constexpr uint32_t CS_SPILL = 0x4000;                // LDS bin level moved to the global overflow counter
constexpr uint32_t CS_UNR = XSS_CS_UNR;              // 16-byte loads per thread in flight while counting
constexpr uint32_t CS_SHORT = 32;                    // runs up to this long are written by one lane
constexpr size_t CS_MIN_N = size_t(1) << 23;         // smaller arrays take the LSD path
static_assert(CS_WAVE_SPAN % 8 == 0, "wave spans start 16-byte aligned when the array is");

// bits of the key whose order key is o (the inverse of KeyDesc::ord without
// the NaN mapping)
template <class KD>
__device__ __forceinline__ typename KD::U unord(typename KD::U o, typename KD::U flip)
{
    using U = typename KD::U;
    const U k = U(o ^ flip);
    if constexpr (KD::KIND == Kind::Unsigned) return k;
    else if constexpr (KD::KIND == Kind::Signed) return U(k ^ KD::SIGN);
    else return (k & KD::SIGN) ? U(k ^ KD::SIGN) : U(~k);
}

template <int = 0>
__global__ __launch_bounds__(256) void cs_zero_kernel(uint32_t *ovf)
{
    ovf[blockIdx.x * 256 + threadIdx.x] = 0;
}

template <class KD>
__global__ __launch_bounds__(CS_BLOCK) void cs_hist_kernel(const void *keys_v, uint32_t n, uint32_t per,
                                                           uint64_t flip64, uint32_t nan_last, uint32_t *part,
                                                           uint32_t *ovf)
{
    using U = typename KD::U;
    static_assert(sizeof(U) == 2, "16-bit keys");
    __shared__ uint32_t h[CS_BINS / 2];
    for (uint32_t i = threadIdx.x; i < CS_BINS / 2; i += CS_BLOCK) h[i] = 0;
    __syncthreads();
    const U *keys = static_cast<const U *>(keys_v);
    const U flip = U(flip64);
    const bool nl = nan_last != 0;
    const uint32_t lo = min(n, blockIdx.x * per), hi = min(n, lo + per);
    // An add of c (<= 512) to bin o returned the word `word` it found.  A
    // bin never wraps: the add that takes it across CS_SPILL moves CS_SPILL
    // to the bin's global overflow counter.  One such move is pending at a
    // time (the bin stays >= CS_SPILL until it lands), and meanwhile each
    // other wave can land at most a few batches of adds on the bin (a wave
    // checks its batch of 8 adds once all 8 have returned), at most 512
    // keys per batch: far fewer than the 0xc000 left below the wrap, so no
    // carry ever reaches the neighbouring bin.
    auto spill = [&](uint32_t o, uint32_t word, uint32_t c) {
        const uint32_t sh = (o & 1u) * 16u;
        const uint32_t old = (word >> sh) & 0xffffu;
        if (old < CS_SPILL && old + c >= CS_SPILL) {
            atomicSub(&h[o >> 1], CS_SPILL << sh);
            atomicAdd(&ovf[o], CS_SPILL);
        }
    };
    auto add = [&](uint32_t o, uint32_t c) { spill(o, atomicAdd(&h[o >> 1], c << ((o & 1u) * 16u)), c); };
    // the 8 keys of a 16-byte load
    auto count8 = [&](const uint4 &q) {
        const U *e = reinterpret_cast<const U *>(&q);
        const bool same = q.x == q.y && q.y == q.z && q.z == q.w && (q.x >> 16) == (q.x & 0xffffu);
        if (__ballot(!same) == 0) {
            // every lane holds 8 equal keys (runs of sorted or constant
            // data): one add per lane, or one per wave of a single key
            const uint32_t o = uint32_t(KD::ord(e[0], flip, nl));
            const uint64_t act = __ballot(1);
            const uint32_t o0 = __builtin_amdgcn_readfirstlane(o);
            if (__ballot(o != o0) == 0) {
                if (lanes_below(act) == 0) add(o0, 8u * uint32_t(__popcll(act)));
            } else {
                add(o, 8u);
            }
        } else {
            // 8 adds in flight, then one test for a spill among them (an
            // add of 1 spills when it found CS_SPILL - 1)
            uint32_t o[8], w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) o[j] = uint32_t(KD::ord(e[j], flip, nl));
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = atomicAdd(&h[o[j] >> 1], 1u << ((o[j] & 1u) * 16u));
            uint32_t hit = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                hit |= uint32_t(((w[j] >> ((o[j] & 1u) * 16u)) & 0xffffu) == CS_SPILL - 1u) << j;
            if (__ballot(hit != 0) != 0) {
#pragma unroll
                for (int j = 0; j < 8; ++j) spill(o[j], w[j], 1u);
            }
        }
    };
    // keys before the first 16-byte boundary one by one, then 8 keys per
    // load: tiles of UNR loads per thread with the next tile's loads issued
    // before the current tile is counted, then the last vectors, then the
    // keys after the last 16-byte boundary
    const uint32_t mis = uint32_t(reinterpret_cast<uintptr_t>(keys + lo) & 15u);
    const uint32_t head = min(hi - lo, uint32_t(((16u - mis) & 15u) / 2u));
    const uint32_t v0 = lo + head;
    const uint32_t nvec = (hi - v0) / 8;
    const uint4 *kv = reinterpret_cast<const uint4 *>(keys + v0);
    constexpr uint32_t UNR = CS_UNR;
    const uint32_t tiles = nvec / (UNR * CS_BLOCK);
    uint4 cur[UNR];
    if (tiles > 0) {
#pragma unroll
        for (uint32_t u = 0; u < UNR; ++u) cur[u] = kv[u * CS_BLOCK + threadIdx.x];
    }
    for (uint32_t t = 0; t < tiles; ++t) {
        uint4 nxt[UNR];
        const bool more = t + 1 < tiles;
        if (more) {
#pragma unroll
            for (uint32_t u = 0; u < UNR; ++u) nxt[u] = kv[((t + 1) * UNR + u) * CS_BLOCK + threadIdx.x];
        }
#pragma unroll
        for (uint32_t u = 0; u < UNR; ++u) count8(cur[u]);
        if (more) {
#pragma unroll
            for (uint32_t u = 0; u < UNR; ++u) cur[u] = nxt[u];
        }
    }
    for (uint32_t v = tiles * UNR * CS_BLOCK + threadIdx.x; v < nvec; v += CS_BLOCK) count8(kv[v]);
    const uint32_t t0 = v0 + nvec * 8;
    const uint32_t rest = head + (hi - t0);
    for (uint32_t r = threadIdx.x; r < rest; r += CS_BLOCK) {
        const uint32_t i = r < head ? lo + r : t0 + (r - head);
        add(uint32_t(KD::ord(keys[i], flip, nl)), 1u);
    }
    __syncthreads();
    // every bin is below 0x10000 here (no spill is pending): the packed words
    uint32_t *row = part + size_t(blockIdx.x) * (CS_BINS / 2);
    for (uint32_t i = threadIdx.x; i < CS_BINS / 2; i += CS_BLOCK) row[i] = h[i];
}

// workgroup j = bins [256 j, 256 j + 256) = words [128 j, 128 j + 128) of
// every row: 8 groups of 128 threads sum every 8th row, then 256 threads
// add the groups and the overflow and scan the block; bsum[j] = its total
template <int = 0>
__global__ __launch_bounds__(1024) void cs_bins_kernel(const uint32_t *part, uint32_t rows, const uint32_t *ovf,
                                                       uint32_t *loc, uint32_t *bsum)
{
    __shared__ uint32_t gsum[8][256];
    __shared__ uint32_t wsum[4];
    const uint32_t t = threadIdx.x, j = blockIdx.x;
    {
        const uint32_t wd = t & 127u, g = t >> 7;
        const uint32_t *col = part + size_t(j) * 128 + wd;
        uint32_t lo = 0, hi = 0;
#pragma unroll 4
        for (uint32_t r = g; r < rows; r += 8) {
            const uint32_t x = col[size_t(r) * (CS_BINS / 2)];
            lo += x & 0xffffu;
            hi += x >> 16;
        }
        gsum[g][2 * wd] = lo;
        gsum[g][2 * wd + 1] = hi;
    }
    __syncthreads();
    // waves 0-3 (threads 0-255, one per bin) add the groups and scan; every
    // thread reaches both barriers
    const bool scan = t < 256;
    const uint32_t lane = t & 63, w = t >> 6;
    const uint32_t o = j * 256 + (t & 255u);
    uint32_t c = 0, incl = 0;
    if (scan) {
        c = ovf[o];
#pragma unroll
        for (int g = 0; g < 8; ++g) c += gsum[g][t];
        incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t x = __shfl_up(incl, off);
            if (lane >= uint32_t(off)) incl += x;
        }
        if (lane == 63) wsum[w] = incl;
    }
    __syncthreads();
    if (scan) {
        uint32_t pre = 0;
        for (uint32_t k = 0; k < w; ++k) pre += wsum[k];
        loc[o] = pre + incl - c;
        if (t == 255) bsum[j] = pre + incl;
    }
}

// workgroup j writes starts[256 j .. 256 j + 255] (the total of blocks
// before j plus the in-block offsets); starts[65536] = n
template <int = 0>
__global__ __launch_bounds__(256) void cs_starts_kernel(const uint32_t *loc, const uint32_t *bsum, uint32_t n,
                                                        uint32_t *starts)
{
    __shared__ uint32_t wsum[4];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6, j = blockIdx.x;
    uint32_t x = t < j ? bsum[t] : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0) wsum[w] = x;
    __syncthreads();
    const uint32_t boff = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    starts[j * 256 + t] = boff + loc[j * 256 + t];
    if (j == 255 && t == 0) starts[CS_BINS] = n;
}

// the bin holding output position pos (pos < n = starts[65536]): the number
// of b in [1, 65535] with starts[b] <= pos, counted over a wave in three
// rounds of 63-way (1024-, 16-, 1-bin strides) samples
__device__ __forceinline__ uint32_t cs_find_bin(const uint32_t *starts, uint32_t pos, uint32_t lane)
{
    // base + stride * 63 <= 65535 in every round, so b stays inside starts[]
    auto round = [&](uint32_t base, uint32_t stride, uint32_t lanes) {
        const uint32_t b = min(base + stride * (lane + 1), CS_BINS);
        const bool in = lane < lanes && starts[b] <= pos;
        return base + stride * uint32_t(__popcll(__ballot(in)));
    };
    uint32_t base = round(0, 1024, 63);  // [base, base + 1024)
    base = round(base, 16, 63);          // [base, base + 16)
    return round(base, 1, 15);
}

template <class KD>
__global__ __launch_bounds__(256) void cs_expand_kernel(const uint32_t *starts, void *out_v, uint32_t n,
                                                        uint64_t flip64, uint32_t nan_last, uint32_t canon)
{
    using U = typename KD::U;
    U *out = static_cast<U *>(out_v);
    const U flip = U(flip64);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (wave >= (n + CS_WAVE_SPAN - 1) / CS_WAVE_SPAN) return;
    uint32_t pos = wave * CS_WAVE_SPAN;
    const uint32_t end = min(n, pos + CS_WAVE_SPAN);
    // positions p with (p + m8) % 8 == 0 are 16-byte aligned in memory
    const uint32_t m8 = uint32_t(reinterpret_cast<uintptr_t>(out) >> 1) & 7u;
    uint32_t o = cs_find_bin(starts, pos, lane);
    while (pos < end) {
        // window: bins o + lane, their ends (bins past the last end at n)
        const uint32_t bl = o + lane;
        const uint32_t te = bl < CS_BINS ? starts[bl + 1] : n;
        const uint32_t tp = __shfl_up(te, 1);
        // this lane's bin holds keys at pos or later
        const bool live = te > pos && (lane == 0 || te != tp);
        if (__ballot(live) == 0) {  // a gap of empty bins: search again
            o = cs_find_bin(starts, pos, lane);
            continue;
        }
        U xl = unord<KD>(U(bl), flip);
        if constexpr (KD::KIND == Kind::Float)
            if (nan_last && canon && U(bl) == U(KD::ONES ^ flip)) xl = KD::QNAN;  // the NaN bin of a hasnan sort
        // the bin's run inside [pos, end): short runs are written by their
        // lane, long ones by the whole wave one after another
        const uint32_t rs = lane == 0 ? pos : max(pos, tp);
        const uint32_t re = min(end, te);
        const bool run = live && rs < re;
        if (run && re - rs <= CS_SHORT)
            for (uint32_t p = rs; p < re; ++p) out[p] = xl;
        uint64_t big = __ballot(run && re - rs > CS_SHORT);
        while (big != 0) {
            const int r = int(__builtin_ctzll(big));
            big &= big - 1;
            const uint32_t s0 = uint32_t(__builtin_amdgcn_readlane(int(rs), r));
            const uint32_t e = uint32_t(__builtin_amdgcn_readlane(int(re), r));
            const U x = U(__builtin_amdgcn_readlane(int(xl), r));
            // [s0, e) <- x: 16-byte stores over its aligned middle [a, b)
            const uint32_t a = min(e + m8, (s0 + m8 + 7u) & ~7u) - m8;
            const uint32_t b = max(a + m8, (e + m8) & ~7u) - m8;
            if (lane < a - s0) out[s0 + lane] = x;
            if (lane < e - b) out[b + lane] = x;
            const uint32_t w = uint32_t(x) | (uint32_t(x) << 16);
            const uint4 q = make_uint4(w, w, w, w);
            uint4 *ov = reinterpret_cast<uint4 *>(out + a);
            for (uint32_t i = lane; i < (b - a) / 8; i += 64) ov[i] = q;
        }
        // every bin of the window ends by te(63): go on from there
        pos = min(end, uint32_t(__builtin_amdgcn_readlane(int(te), 63)));
        o += 64;
    }
}

// rows of the histogram table: one per CU
inline uint32_t cs_rows()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return uint32_t(cus);
}
inline size_t cs_scratch_bytes(uint32_t rows)
{
    return align256(size_t(rows) * CS_BINS * 2) + 2 * align256(CS_BINS * 4) + align256((CS_BINS + 1) * 4) +
           align256(256 * 4);
}

// Enqueue a counting sort of n (< 2^30) 2-byte aligned 16-bit keys in place,
// NaNs of a hasnan sort written as the canonical NaN.  scratch:
// cs_scratch_bytes(rows).
template <class KD>
int enqueue_count_sort(char *scratch, void *keys, size_t n, bool hasnan, bool desc, bool canon, uint32_t rows,
                       hipStream_t s)
{
    char *p = scratch;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align256(bytes);
        return reinterpret_cast<uint32_t *>(q);
    };
    uint32_t *part = take(size_t(rows) * CS_BINS * 2);
    uint32_t *ovf = take(CS_BINS * 4);
    uint32_t *loc = take(CS_BINS * 4);
    uint32_t *starts = take((CS_BINS + 1) * 4);
    uint32_t *bsum = take(256 * 4);
    const uint64_t flip = desc ? ~0ull : 0ull;
    const uint32_t nl = (KD::KIND == Kind::Float) && hasnan;
    const uint32_t nn = uint32_t(n);
    const uint32_t per = uint32_t(((n + rows - 1) / rows + 7) / 8 * 8);
    const uint32_t waves = (nn + CS_WAVE_SPAN - 1) / CS_WAVE_SPAN;
    KernelTimer t("count_sort", s);
    hipLaunchKernelGGL(cs_zero_kernel<0>, dim3(CS_BINS / 256), dim3(256), 0, s, ovf);
    XSS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((cs_hist_kernel<KD>), dim3(rows), dim3(CS_BLOCK), 0, s, keys, nn, per, flip, nl, part, ovf);
    XSS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cs_bins_kernel<0>, dim3(CS_BINS / 256), dim3(1024), 0, s, part, rows, ovf, loc, bsum);
    XSS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cs_starts_kernel<0>, dim3(CS_BINS / 256), dim3(256), 0, s, loc, bsum, nn, starts);
    XSS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((cs_expand_kernel<KD>), dim3((waves + 3) / 4), dim3(256), 0, s, starts, keys, nn, flip, nl,
                       uint32_t(canon));
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

}  // namespace xss
