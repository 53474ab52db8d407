// xss_sort.hpp -- host-side pipeline of one radix sort on a stream.
//
// Replaces the driver layer of the reference (xss_qsort
// src/xss-common-qsort.h:652-710, xss_qsort_kv
// src/xss-common-keyvaluesort.hpp:564-653, xss_argsort
// src/xss-common-argsort.h:587-669): instead of NaN pre/post passes, an
// OpenMP team and a recursion, a sort is either one LDS block-sort launch
// (small arrays) or hist -> scan -> P onesweep passes (P = key bits / 8).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/xss_gpu.h"
#include "xss_kernels.hpp"
#include "xss_runtime.hpp"

namespace xss {

// Largest array of one call: positions, ranks and digit counts are 32-bit
// on the device, and tile / range arithmetic needs 2^24 of headroom below
// 2^32 (4,278,190,080 keys: 16 GiB of 4-byte keys, 32 GiB of 8-byte ones).
constexpr uint32_t MAX_N = 0xFF000000u;

template <class T>
struct TypeKey;
template <> struct TypeKey<uint16_t> { using KD = KeyDesc<16, Kind::Unsigned>; };
template <> struct TypeKey<int16_t> { using KD = KeyDesc<16, Kind::Signed>; };
struct f16_tag {};
template <> struct TypeKey<f16_tag> { using KD = KeyDesc<16, Kind::Float>; };
template <> struct TypeKey<uint32_t> { using KD = KeyDesc<32, Kind::Unsigned>; };
template <> struct TypeKey<int32_t> { using KD = KeyDesc<32, Kind::Signed>; };
template <> struct TypeKey<float> { using KD = KeyDesc<32, Kind::Float>; };
template <> struct TypeKey<uint64_t> { using KD = KeyDesc<64, Kind::Unsigned>; };
template <> struct TypeKey<int64_t> { using KD = KeyDesc<64, Kind::Signed>; };
template <> struct TypeKey<double> { using KD = KeyDesc<64, Kind::Float>; };

// payload mode of a value type: its width is all that matters
template <class V>
constexpr int payload_of() { return sizeof(V) == 4 ? PL_V32 : PL_V64; }

// tile shape of a downsweep: BLOCK threads x ITEMS keys
template <class KD, int PL>
struct PassShape {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr int CLS = ScatterClass<KD, PL>::value;
    static constexpr int BLOCK = CLS == 0 ? XSS_PASS_BLOCK_C0 : CLS == 1 ? XSS_PASS_BLOCK_C1
                               : CLS == 2 ? XSS_PASS_BLOCK_C2 : XSS_PASS_BLOCK_C3;
    static constexpr int ITEMS = CLS == 0 ? XSS_PASS_ITEMS_C0 : CLS == 1 ? XSS_PASS_ITEMS_C1
                               : CLS == 2 ? XSS_PASS_ITEMS_C2 : XSS_PASS_ITEMS_C3;
    static constexpr int TILE = BLOCK * ITEMS;
};

// Arrays below SMALL_PASS_N keys use 512-thread tiles: the large tiles of
// PassShape would leave too few workgroups to fill the GPU.
constexpr size_t SMALL_PASS_N = size_t(1) << 24;
template <class KD, int PL>
struct SmallPassShape {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr int BLOCK = 512;
    static constexpr int ITEMS = (sizeof(U) + (PL == PL_NONE ? 0 : sizeof(V))) > 8 ? 8 : 16;
    static constexpr int TILE = BLOCK * ITEMS;
};

// largest array one workgroup sorts in LDS (<= 64 KiB of LDS)
template <class KD, int PL>
struct BlockShape {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr int BPE = 2 * sizeof(U) + 2 + ((PL == PL_V32 || PL == PL_V64) ? sizeof(V) : 0);
    static constexpr int MAXN = (65536 / BPE) >= 4096 ? 4096 : (65536 / BPE) >= 2048 ? 2048 : 1024;
    static constexpr int BLOCK = 512;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

template <class KD, int PL>
struct SortPlan {
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    static constexpr int P = KD::PASSES;
    static bool small(size_t n) { return n <= size_t(BlockShape<KD, PL>::MAXN); }
    static bool small_tiles(size_t n) { return n < SMALL_PASS_N; }
    static size_t tile(size_t n)
    {
        return small_tiles(n) ? size_t(SmallPassShape<KD, PL>::TILE) : size_t(PassShape<KD, PL>::TILE);
    }
    static size_t tiles(size_t n) { return (n + tile(n) - 1) / tile(n); }
    // persistent workgroups: contiguous ranges of whole tiles
    static uint32_t range_tiles(size_t n) { return uint32_t((tiles(n) + MAX_RANGES - 1) / MAX_RANGES); }
    static uint32_t ranges(size_t n) { return uint32_t((tiles(n) + range_tiles(n) - 1) / range_tiles(n)); }
    static size_t scratch_bytes(size_t n)
    {
        if (small(n)) return 0;
        if (PL == PL_IDX) return (P > 2 ? 2 : 1) * (align256(n * sizeof(U)) + align256(n * 4));
        if (PL == PL_NONE) return align256(n * sizeof(U));
        return align256(n * sizeof(U)) + align256(n * sizeof(V));
    }
};

// Enqueue one 8-bit LSD pass over n keys (digit at `shift`): upsweep ->
// range_scan -> downsweep, src -> dst, with the tile shape SortPlan picks
// for n.  PL_IDX: `first` generates the indices, `last` writes them to
// idx_out.
template <class KD, int PL, class PS>
int launch_downsweep(const PassArgs &a, uint32_t ranges, bool nan_last, bool unstable, hipStream_t s)
{
    KernelTimer t("downsweep", s);
    if constexpr (PL != PL_IDX) {
        if (unstable) {
            if (nan_last)
                hipLaunchKernelGGL((downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, true, true>), dim3(ranges),
                                   dim3(PS::BLOCK), 0, s, a);
            else
                hipLaunchKernelGGL((downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, false, true>), dim3(ranges),
                                   dim3(PS::BLOCK), 0, s, a);
            XSS_HIP_CHECK(hipGetLastError());
            return XSS_OK;
        }
    }
    if (nan_last)
        hipLaunchKernelGGL((downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, true>), dim3(ranges), dim3(PS::BLOCK), 0,
                           s, a);
    else
        hipLaunchKernelGGL((downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, false>), dim3(ranges), dim3(PS::BLOCK),
                           0, s, a);
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

// ran != null: a pass of a sort that skips constant digits (src/dst are
// then the two buffers the keys ping-pong between, chosen on the device).
// unstable: a pass whose input order does not matter may give up the order
// inside a digit (cheaper ranking); argsort passes always keep it.
template <class KD, int PL>
int enqueue_lsd_pass(Workspace *ws, const void *src_k, const void *src_v, void *dst_k, void *dst_v,
                     uint64_t *idx_out, size_t n, int shift, uint64_t flip, bool nan_last, bool first, bool last,
                     bool canon_nan, hipStream_t s, uint32_t *ran = nullptr, uint32_t pass = 0, bool skip_ok = false,
                     uint64_t *andor = nullptr, bool unstable = false)
{
    using Plan = SortPlan<KD, PL>;
    const uint32_t nn = uint32_t(n);
    const uint32_t rtiles = Plan::range_tiles(n), ranges = Plan::ranges(n);
    UpArgs u{};
    u.keys = src_k; u.n = nn; u.range_keys = uint32_t(rtiles * Plan::tile(n)); u.shift = uint32_t(shift);
    u.nan_last = nan_last; u.flip = flip; u.rsum = ws->rsum(); u.ranges = ranges;
    u.keys_alt = dst_k; u.ran = ran; u.pass = pass; u.andor = andor; u.skip_ok = skip_ok;
    u.andor_slots = ws->andor_slots();
    {
        KernelTimer t("upsweep", s);
        hipLaunchKernelGGL((upsweep_kernel<KD>), dim3(ranges), dim3(UP_BLOCK), 0, s, u);
    }
    XSS_HIP_CHECK(hipGetLastError());
    ScanArgs sc{};
    sc.rsum = ws->rsum(); sc.rpre = ws->rpre(); sc.total = ws->total(); sc.ranges = ranges;
    if (andor && pass == 0) {
        sc.andor_slots = ws->andor_slots();
        sc.andor = andor;
    }
    {
        KernelTimer t("range_scan", s);
        hipLaunchKernelGGL(range_scan_kernel<0>, dim3(256), dim3(SCAN_BLOCK), 0, s, sc);
    }
    XSS_HIP_CHECK(hipGetLastError());
    PassArgs a{};
    a.keys_in = src_k; a.vals_in = src_v; a.keys_out = dst_k; a.vals_out = dst_v; a.idx_out = idx_out;
    a.n = nn; a.shift = uint32_t(shift); a.flip = flip; a.nan_last = nan_last; a.range_tiles = rtiles;
    a.ranges = ranges; a.rpre = ws->rpre(); a.total = ws->total();
    a.first = first; a.last = last; a.canon_nan = canon_nan;
    a.ran = ran; a.pass = pass; a.skip_ok = skip_ok; a.andor = andor;
    if (Plan::small_tiles(n))
        return launch_downsweep<KD, PL, SmallPassShape<KD, PL>>(a, ranges, nan_last, unstable, s);
    return launch_downsweep<KD, PL, PassShape<KD, PL>>(a, ranges, nan_last, unstable, s);
}

// Enqueue one sort of n keys (in place when keys_in == keys_out).
// `scratch` must hold SortPlan::scratch_bytes(n).
template <class KD, int PL>
int enqueue_sort(Workspace *ws, char *scratch, const void *keys_in, void *keys_out, const void *vals_in,
                 void *vals_out, uint64_t *idx_out, size_t n, bool hasnan, bool desc, bool canon,
                 hipStream_t s)
{
    using U = typename KD::U;
    using Plan = SortPlan<KD, PL>;
    const uint64_t flip = desc ? ~0ull : 0ull;
    const bool nan_last = (KD::KIND == Kind::Float) && hasnan;
    if (n < 2) {
        // n <= 1 is a no-op in the reference (src/xss-common-qsort.h:660,
        // src/xss-common-keyvaluesort.hpp:585); argsort still returns [0].
        if (n == 1 && PL == PL_IDX) {
            hipLaunchKernelGGL(fill_u64_kernel<0>, dim3(1), dim3(64), 0, s, idx_out, uint64_t(0), uint64_t(1));
            XSS_HIP_CHECK(hipGetLastError());
        }
        return XSS_OK;
    }
    if (Plan::small(n)) {
        using BS = BlockShape<KD, PL>;
        BlockSortArgs b{};
        b.keys_in = keys_in; b.keys_out = keys_out; b.vals_in = vals_in; b.vals_out = vals_out;
        b.idx_out = idx_out; b.n = uint32_t(n); b.nan_last = nan_last; b.flip = flip;
        b.canon_nan = canon && nan_last;
        {
            KernelTimer t("block_sort", s);
            hipLaunchKernelGGL((block_sort_kernel<KD, PL, BS::BLOCK, BS::MAXN>), dim3(1), dim3(BS::BLOCK), 0, s, b);
        }
        XSS_HIP_CHECK(hipGetLastError());
        return XSS_OK;
    }
    constexpr int P = KD::PASSES;
    // ping-pong buffers
    char *k0 = scratch, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
    if (PL == PL_IDX) {
        v0 = k0 + align256(n * sizeof(U));
        if (P > 2) {
            k1 = v0 + align256(n * 4);
            v1 = k1 + align256(n * sizeof(U));
        }
    } else if (PL != PL_NONE) {
        v0 = k0 + align256(n * sizeof(U));
    }
    if (PL != PL_IDX && keys_in == keys_out && vals_in == vals_out) {
        // In place: a pass whose digit is the same for all n keys (keys of a
        // narrow range, constant arrays) is skipped on the device; the keys
        // ping-pong between the array and the scratch buffer by the parity of
        // the passes that ran.  The last pass of a NaN-canonicalising qsort
        // always runs (it writes the canonical NaN).
        uint32_t *ran = ws->ran();
        uint64_t *andor = ws->andor();
        hipLaunchKernelGGL(andor_init_kernel<0>, dim3(1), dim3(1), 0, s, andor);
        XSS_HIP_CHECK(hipGetLastError());
        for (int p = 0; p < P; ++p) {
            const bool last = p == P - 1;
            const bool canon_pass = last && canon && nan_last;
            const int rc = enqueue_lsd_pass<KD, PL>(ws, keys_out, vals_out, k0, v0, idx_out, n, 8 * p, flip, nan_last,
                                                    p == 0, last, canon_pass, s, ran, uint32_t(p), !canon_pass, andor,
                                                    p == 0);
            if (rc) return rc;
        }
        using V = typename PayloadT<PL>::type;
        const uint64_t kb = uint64_t(n) * sizeof(U), vb = PL == PL_NONE ? 0 : uint64_t(n) * sizeof(V);
        const uint32_t blocks = uint32_t(std::min<uint64_t>(4096, (kb / 16 + 255) / 256 + 1));
        hipLaunchKernelGGL(skip_finish_kernel<0>, dim3(blocks), dim3(256), 0, s, keys_out, k0,
                           PL == PL_NONE ? nullptr : vals_out, v0, kb, vb, ran, uint32_t(P));
        XSS_HIP_CHECK(hipGetLastError());
        return XSS_OK;
    }
    const void *src_k = keys_in, *src_v = vals_in;
    for (int p = 0; p < P; ++p) {
        void *dk, *dv;
        if (PL == PL_IDX) {
            dk = p == P - 1 ? nullptr : (p % 2 == 0) ? k0 : k1;
            dv = p == P - 1 ? nullptr : (p % 2 == 0) ? v0 : v1;
        } else {
            dk = (p % 2 == 0) ? static_cast<void *>(k0) : keys_out;
            dv = (p % 2 == 0) ? static_cast<void *>(v0) : vals_out;
        }
        const int rc = enqueue_lsd_pass<KD, PL>(ws, src_k, src_v, dk, dv, idx_out, n, 8 * p, flip, nan_last, p == 0,
                                                p == P - 1, (p == P - 1) && canon && nan_last, s, nullptr, 0, false,
                                                nullptr, p == 0);
        if (rc) return rc;
        src_k = dk;
        src_v = dv;
    }
    return XSS_OK;
}

// Whole-call wrappers: validate, acquire the stream's workspace, enqueue.
template <class KD, int PL>
int run_sort(const void *keys_in, void *keys_out, const void *vals_in, void *vals_out, uint64_t *idx_out,
             size_t n, bool hasnan, bool desc, bool canon, hipStream_t s)
{
    if (n > MAX_N) return set_error(XSS_ERR_INVALID, "n = %zu exceeds the per-call limit %u", n, MAX_N);
    // n <= 1 is a no-op (src/xss-common-qsort.h:660); argsort of one key still writes index 0
    if (n == 0 || (n == 1 && PL != PL_IDX)) return XSS_OK;
    if (!keys_in || (PL == PL_IDX ? !idx_out : !keys_out) || ((PL == PL_V32 || PL == PL_V64) && (!vals_in || !vals_out)))
        return set_error(XSS_ERR_INVALID, "null pointer argument");
    using Plan = SortPlan<KD, PL>;
    Workspace *ws = nullptr;
    std::unique_lock<std::mutex> lock;
    int rc = acquire_workspace(s, Plan::scratch_bytes(n), &ws, &lock);
    if (rc) return rc;
    return enqueue_sort<KD, PL>(ws, ws->data, keys_in, keys_out, vals_in, vals_out, idx_out, n, hasnan, desc,
                                canon, s);
}

}  // namespace xss
