// xss_server.hip -- a continuous-batching request loop for sort / select
// requests on one GPU.  No counterpart in the reference (a synchronous
// library); it is the serving form built on the same kernels.
//
// Clients submit requests (operation, key/value types, device pointers,
// length, k, flags) at any time.  One worker thread owns S HIP streams
// (default priority) and, in a loop:
//   * retires finished launch groups (their HIP event has completed) and
//     releases their HBM footprint and stream;
//   * admits queued requests in arrival order while a stream is free and the
//     admitted footprint (keys + values + scratch) stays within the budget;
//   * coalesces queued qsort and keyvalue_qsort requests of at most
//     coalesce_max keys with the same key type, value width and flags into
//     one launch group: those short enough for one LDS workgroup are sorted
//     in place by one launch of the pointer-list block sort; the others are
//     gathered into the stream's arena as segments of one buffer, sorted by
//     one set of segmented LSD passes (upsweep -> per-segment scan ->
//     downsweep per digit, over a range table built on the host and uploaded
//     with one copy) and scattered back;
//   * runs every other request through its C-ABI entry point on its stream,
//     so up to S requests overlap on the device while new ones keep arriving
//     and are admitted as streams and memory free up.
// Each stream keeps the library's pooled workspace, reserved up front, and a
// fixed arena for coalesced groups, so steady-state requests allocate
// nothing.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "xss_sort.hpp"

using namespace xss;
using Clock = std::chrono::steady_clock;

constexpr uint32_t MAX_GROUP = 1024;  // requests per coalesced launch group

namespace {

// Coalesced groups: medium requests (longer than one LDS workgroup sorts)
// share segmented passes over up to GROUP_KEYS keys gathered into the
// stream's arena, cut into ranges of 1..MAX_RANGE_TILES downsweep tiles
// (tiles are at least 4096 keys, so GROUP_RANGES bounds the range table).
constexpr uint64_t GROUP_KEYS = uint64_t(1) << 24;
constexpr uint32_t MAX_RANGE_TILES = 8;
constexpr uint32_t GROUP_RANGES = uint32_t(GROUP_KEYS / 4096u) + MAX_GROUP;

enum Op : int { OP_QSORT = 0, OP_KV_QSORT = 1, OP_QSELECT = 2, OP_PARTIAL = 3 };
// key/value type codes: index into this list
const char *const TAG[] = {"u16", "i16", "f16", "u32", "i32", "f32", "u64", "i64", "f64"};
constexpr int NTAGS = 9;
size_t tag_bytes(int t) { return t < 3 ? 2 : t < 6 ? 4 : 8; }

struct Req {
    xss_request *r;        // caller's record (latency written back)
    Clock::time_point submitted;
    size_t footprint = 0;
    bool done = false;
};

struct Group {
    std::vector<Req *> reqs;
    int stream = -1;
    hipEvent_t ev = nullptr;
    size_t footprint = 0;
};

size_t footprint_of(const xss_request &r)
{
    const size_t kb = tag_bytes(r.key_type), vb = r.op == OP_KV_QSORT ? tag_bytes(r.val_type) : 0;
    // data + ping-pong scratch (+ candidates for selects) + fixed metadata
    return r.n * (kb + vb) * (r.op == OP_QSELECT || r.op == OP_PARTIAL ? 4 : 2) + (1u << 20);
}

// launch one request through its C-ABI entry point
int launch_one(const xss_request &r, hipStream_t s)
{
    const int kt = r.key_type;
    if (kt < 0 || kt >= NTAGS) return set_error(XSS_ERR_INVALID, "bad key type %d", kt);
    switch (r.op) {
    case OP_QSORT:
        switch (kt) {
        case 0: return xss_gpu_qsort_u16((uint16_t *)r.keys, r.n, r.hasnan, r.descending, s);
        case 1: return xss_gpu_qsort_i16((int16_t *)r.keys, r.n, r.hasnan, r.descending, s);
        case 2: return xss_gpu_qsort_f16((xss_f16_bits *)r.keys, r.n, r.hasnan, r.descending, s);
        case 3: return xss_gpu_qsort_u32((uint32_t *)r.keys, r.n, r.hasnan, r.descending, s);
        case 4: return xss_gpu_qsort_i32((int32_t *)r.keys, r.n, r.hasnan, r.descending, s);
        case 5: return xss_gpu_qsort_f32((float *)r.keys, r.n, r.hasnan, r.descending, s);
        case 6: return xss_gpu_qsort_u64((uint64_t *)r.keys, r.n, r.hasnan, r.descending, s);
        case 7: return xss_gpu_qsort_i64((int64_t *)r.keys, r.n, r.hasnan, r.descending, s);
        default: return xss_gpu_qsort_f64((double *)r.keys, r.n, r.hasnan, r.descending, s);
        }
    case OP_QSELECT:
    case OP_PARTIAL: {
        const bool sel = r.op == OP_QSELECT;
#define XSS_SEL(tag, T)                                                                                     \
    return sel ? xss_gpu_qselect_##tag((T *)r.keys, r.k, r.n, r.hasnan, r.descending, s)                   \
               : xss_gpu_partial_qsort_##tag((T *)r.keys, r.k, r.n, r.hasnan, r.descending, s)
        switch (kt) {
        case 0: XSS_SEL(u16, uint16_t);
        case 1: XSS_SEL(i16, int16_t);
        case 2: XSS_SEL(f16, xss_f16_bits);
        case 3: XSS_SEL(u32, uint32_t);
        case 4: XSS_SEL(i32, int32_t);
        case 5: XSS_SEL(f32, float);
        case 6: XSS_SEL(u64, uint64_t);
        case 7: XSS_SEL(i64, int64_t);
        default: XSS_SEL(f64, double);
        }
#undef XSS_SEL
    }
    case OP_KV_QSORT: {
        // value types that share a width share the kernel; key types need their own order
        const bool v64 = tag_bytes(r.val_type) == 8;
#define XSS_KV(tag, K)                                                                                      \
    return v64 ? xss_gpu_keyvalue_qsort_##tag##_u64((K *)r.keys, (uint64_t *)r.vals, r.n, r.hasnan, r.descending, s) \
               : xss_gpu_keyvalue_qsort_##tag##_u32((K *)r.keys, (uint32_t *)r.vals, r.n, r.hasnan, r.descending, s)
        if (r.val_type < 3 || r.val_type >= NTAGS)
            return set_error(XSS_ERR_INVALID, "keyvalue_qsort has no %s values", TAG[r.val_type < 0 ? 0 : r.val_type % NTAGS]);
        switch (kt) {
        case 0: XSS_KV(u16, uint16_t);
        case 1: XSS_KV(i16, int16_t);
        case 2: XSS_KV(f16, xss_f16_bits);
        case 3: XSS_KV(u32, uint32_t);
        case 4: XSS_KV(i32, int32_t);
        case 5: XSS_KV(f32, float);
        case 6: XSS_KV(u64, uint64_t);
        case 7: XSS_KV(i64, int64_t);
        default: XSS_KV(f64, double);
        }
#undef XSS_KV
    }
    }
    return set_error(XSS_ERR_INVALID, "bad op %d", r.op);
}

// ---------------------------------------------------------------- groups

// Device addresses of one group's tables, packed by the host into the
// stream's arena (the same offsets in its pinned mirror).
struct GroupDev {
    const uint64_t *sm_k, *sm_v;   // [nsmall] small requests: keys / values pointers
    const uint32_t *sm_len;        // [nsmall]
    const uint64_t *md_k, *md_v;   // [nmed] medium requests: keys / values pointers
    const uint64_t *r_begin, *r_end, *r_sbase;  // [nr] ranges of the staging buffer
    const uint32_t *r_bseg, *r_head;            // [nr] medium index, first range of one
    uint32_t *rsum, *rpre, *segtot;             // pass counts
    char *kA, *kB, *vA, *vB;                    // staging: keys and values, ping-pong
    uint32_t nsmall, nmed, nr;
    uint64_t nkeys;                             // medium keys in the staging buffer
};

// Medium requests <-> the staging buffer: one workgroup per range (a range
// lies inside one request).  back = 0 gathers, 1 scatters.
template <class T>
__global__ __launch_bounds__(256) void group_copy_kernel(const uint64_t *r_begin, const uint64_t *r_end,
                                                         const uint64_t *r_sbase, const uint32_t *r_bseg,
                                                         const uint64_t *ptrs, T *stage, int back)
{
    const uint32_t r = blockIdx.x;
    const uint64_t b = r_begin[r], e = r_end[r], sb = r_sbase[r];
    T *user = reinterpret_cast<T *>(ptrs[r_bseg[r]]) - sb;
    if (back)
        for (uint64_t x = b + threadIdx.x; x < e; x += 256) user[x] = stage[x];
    else
        for (uint64_t x = b + threadIdx.x; x < e; x += 256) stage[x] = user[x];
}

template <class T>
int group_copy(const GroupDev &g, const uint64_t *ptrs, char *stage, int back, hipStream_t s)
{
    KernelTimer t(back ? "serve_scatter" : "serve_gather", s);
    hipLaunchKernelGGL(group_copy_kernel<T>, dim3(g.nr), dim3(256), 0, s, g.r_begin, g.r_end, g.r_sbase, g.r_bseg,
                       ptrs, reinterpret_cast<T *>(stage), back);
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

// One coalesced group of keys-only (PL_NONE) or key-value (PL_V32 / PL_V64)
// sorts: the pointer-list block sort for the small requests, then the
// medium ones gathered, sorted by P segmented LSD passes (P is even, so the
// result is back in kA / vA) and scattered.  qsort(hasnan) writes the
// canonical NaN, as the plain call does; key-value sorts keep NaN bits.
template <class KD, int PL>
int run_group(const GroupDev &g, bool hasnan, bool desc, hipStream_t s)
{
    using U = typename KD::U;
    using V = typename PayloadT<PL>::type;
    using BS = BlockShape<KD, PL>;
    using PS = SegShape<KD, PL>;
    constexpr int P = KD::PASSES;
    static_assert(P % 2 == 0, "an even pass count ends in the gathered buffer");
    constexpr bool HAS_VAL = PL != PL_NONE;
    const bool nan_last = (KD::KIND == Kind::Float) && hasnan;
    const bool canon = nan_last && !HAS_VAL;
    const uint64_t flip = desc ? ~0ull : 0ull;
    if (g.nsmall) {
        BlockSortArgs b{};
        b.arr_keys = g.sm_k;
        b.arr_vals = HAS_VAL ? g.sm_v : nullptr;
        b.arr_len = g.sm_len;
        b.nan_last = nan_last;
        b.flip = flip;
        b.canon_nan = canon;
        KernelTimer t("serve_small", s);
        hipLaunchKernelGGL((block_sort_kernel<KD, PL, BS::BLOCK, BS::MAXN>), dim3(g.nsmall), dim3(BS::BLOCK), 0, s, b);
        XSS_HIP_CHECK(hipGetLastError());
    }
    if (!g.nr) return XSS_OK;
    int rc = group_copy<U>(g, g.md_k, g.kA, 0, s);
    if (!rc && HAS_VAL) rc = group_copy<V>(g, g.md_v, g.vA, 0, s);
    if (rc) return rc;
    const void *src_k = g.kA, *src_v = g.vA;
    for (int pass = 0; pass < P; ++pass) {
        UpArgs u{};
        u.keys = src_k;
        u.n = uint32_t(g.nkeys);
        u.shift = 8 * pass;
        u.nan_last = nan_last;
        u.flip = flip;
        u.rsum = g.rsum;
        u.ranges = g.nr;
        u.range_begin = g.r_begin;
        u.range_end = g.r_end;
        {
            KernelTimer t("serve_upsweep", s);
            hipLaunchKernelGGL((upsweep_kernel<KD>), dim3(g.nr), dim3(UP_BLOCK), 0, s, u);
        }
        XSS_HIP_CHECK(hipGetLastError());
        SegScanArgs sc{g.rsum, g.rpre, g.segtot, g.r_head, g.r_bseg, g.nr, g.nr, nullptr};
        {
            KernelTimer t("serve_scan", s);
            hipLaunchKernelGGL(seg_scan_kernel<0>, dim3(256), dim3(256), 0, s, sc);
        }
        XSS_HIP_CHECK(hipGetLastError());
        PassArgs a{};
        a.keys_in = src_k;
        a.vals_in = src_v;
        a.keys_out = (pass % 2 == 0) ? g.kB : g.kA;
        a.vals_out = (pass % 2 == 0) ? g.vB : g.vA;
        a.n = uint32_t(g.nkeys);
        a.shift = 8 * pass;
        a.flip = flip;
        a.nan_last = nan_last;
        a.first = pass == 0;
        a.last = pass == P - 1;
        a.canon_nan = (pass == P - 1) && canon;
        SegRanges r{g.r_begin, g.r_end, g.r_sbase, g.r_bseg, g.rpre, g.segtot, g.nr, g.nr, nullptr};
        {
            KernelTimer t("serve_downsweep", s);
            if (nan_last)
                hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, true>), dim3(g.nr),
                                   dim3(PS::BLOCK), 0, s, a, r);
            else
                hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, false>), dim3(g.nr),
                                   dim3(PS::BLOCK), 0, s, a, r);
        }
        XSS_HIP_CHECK(hipGetLastError());
        src_k = a.keys_out;
        src_v = a.vals_out;
    }
    rc = group_copy<U>(g, g.md_k, g.kA, 1, s);
    if (!rc && HAS_VAL) rc = group_copy<V>(g, g.md_v, g.vA, 1, s);
    return rc;
}

template <class KD>
int run_group_pl(int val_bytes, const GroupDev &g, bool hasnan, bool desc, hipStream_t s)
{
    if (val_bytes == 0) return run_group<KD, PL_NONE>(g, hasnan, desc, s);
    if (val_bytes == 4) return run_group<KD, PL_V32>(g, hasnan, desc, s);
    return run_group<KD, PL_V64>(g, hasnan, desc, s);
}

int run_group_typed(int kt, int val_bytes, const GroupDev &g, bool hasnan, bool desc, hipStream_t s)
{
    switch (kt) {
    case 0: return run_group_pl<TypeKey<uint16_t>::KD>(val_bytes, g, hasnan, desc, s);
    case 1: return run_group_pl<TypeKey<int16_t>::KD>(val_bytes, g, hasnan, desc, s);
    case 2: return run_group_pl<TypeKey<f16_tag>::KD>(val_bytes, g, hasnan, desc, s);
    case 3: return run_group_pl<TypeKey<uint32_t>::KD>(val_bytes, g, hasnan, desc, s);
    case 4: return run_group_pl<TypeKey<int32_t>::KD>(val_bytes, g, hasnan, desc, s);
    case 5: return run_group_pl<TypeKey<float>::KD>(val_bytes, g, hasnan, desc, s);
    case 6: return run_group_pl<TypeKey<uint64_t>::KD>(val_bytes, g, hasnan, desc, s);
    case 7: return run_group_pl<TypeKey<int64_t>::KD>(val_bytes, g, hasnan, desc, s);
    default: return run_group_pl<TypeKey<double>::KD>(val_bytes, g, hasnan, desc, s);
    }
}

// largest request the pointer-list block sort takes (BlockShape's MAXN)
template <class KD>
uint32_t small_max_pl(int val_bytes)
{
    if (val_bytes == 0) return BlockShape<KD, PL_NONE>::MAXN;
    if (val_bytes == 4) return BlockShape<KD, PL_V32>::MAXN;
    return BlockShape<KD, PL_V64>::MAXN;
}

uint32_t small_limit(int kt, int val_bytes)
{
    switch (tag_bytes(kt)) {
    case 2: return small_max_pl<TypeKey<uint16_t>::KD>(val_bytes);
    case 4: return small_max_pl<TypeKey<uint32_t>::KD>(val_bytes);
    default: return small_max_pl<TypeKey<uint64_t>::KD>(val_bytes);
    }
}

// keys per range of the segmented passes: whole tiles of SegShape, as many
// per range as keep about 1024 workgroups per pass (at most
// MAX_RANGE_TILES), so a group of a few requests still fills the GPU
uint32_t range_keys(int kt, int val_bytes, uint64_t med_keys)
{
    const int bytes = int(tag_bytes(kt)) + val_bytes;
    const uint32_t tile = bytes == 4 ? uint32_t(XSS_SEG_BLOCK_C1) * XSS_SEG_ITEMS_C1 : 512u * (bytes > 8 ? 8 : 16);
    const uint64_t want = (med_keys + uint64_t(tile) * 1024 - 1) / (uint64_t(tile) * 1024);
    return tile * uint32_t(std::min<uint64_t>(std::max<uint64_t>(want, 1), MAX_RANGE_TILES));
}

// value bytes of a request that may join a group (0: keys only), or -1
int group_val_bytes(const xss_request &r)
{
    if (r.key_type < 0 || r.key_type >= NTAGS) return -1;
    if (r.op == OP_QSORT) return 0;
    if (r.op == OP_KV_QSORT && r.val_type >= 3 && r.val_type < NTAGS) return int(tag_bytes(r.val_type));
    return -1;
}

// Per-stream memory of coalesced groups: tables (device + pinned mirror),
// pass counts and the staging buffers.
struct Arena {
    char *h_tab = nullptr, *d_tab = nullptr;
    uint32_t *rsum = nullptr, *rpre = nullptr, *segtot = nullptr;
    char *stage = nullptr;   // kA | kB | vA | vB, GROUP_KEYS * 8 bytes each
    static size_t table_bytes()
    {
        return size_t(MAX_GROUP) * (8 + 8 + 4 + 8 + 8) + size_t(GROUP_RANGES) * (8 * 3 + 4 * 2) + 16 * 256;
    }
};

}  // namespace

struct xss_server {
    size_t budget = 0, used = 0;
    uint32_t coalesce_max = 0;
    bool one_at_a_time = false;
    int device = 0;
    std::vector<hipStream_t> streams;
    std::vector<bool> busy;
    std::vector<Arena> arenas;        // per stream (when coalescing)
    std::deque<Req *> queue;
    std::vector<Group *> inflight;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    int error = XSS_OK;
    std::string error_text;
    uint64_t completed = 0, coalesced = 0, groups = 0;

    void loop();
    bool retire();
    bool admit();
    int launch_group(int s, int val_bytes, Group *g);
};

// Take the queued requests that can share the head's group (same op, key
// type, value width and flags, at most coalesce_max keys, medium keys within
// GROUP_KEYS, footprint within the budget) off the queue, pack their tables
// into the stream's pinned mirror, upload them with one copy and launch.
int xss_server::launch_group(int s, int vb, Group *g)
{
    const xss_request &fr = *queue.front()->r;
    const uint32_t smax = small_limit(fr.key_type, vb);
    const uint64_t cap = std::min<uint64_t>(coalesce_max, GROUP_KEYS);
    std::vector<Req *> sm, md;
    uint64_t med_keys = 0;
    for (auto it = queue.begin(); it != queue.end() && sm.size() + md.size() < MAX_GROUP;) {
        const xss_request &q = *(*it)->r;
        const bool match = q.op == fr.op && q.key_type == fr.key_type && group_val_bytes(q) == vb &&
                           q.hasnan == fr.hasnan && q.descending == fr.descending && q.n <= cap;
        const bool small = q.n <= smax;
        // the head is admitted whatever its size (the caller checked the budget)
        const bool fits = g->reqs.empty() || used + g->footprint + (*it)->footprint <= budget;
        if (match && fits && (small || med_keys + q.n <= GROUP_KEYS)) {
            (small ? sm : md).push_back(*it);
            if (!small) med_keys += q.n;
            g->reqs.push_back(*it);
            g->footprint += (*it)->footprint;
            it = queue.erase(it);
        } else {
            ++it;
        }
    }
    if (g->reqs.size() == 1) {  // nothing to share a launch with: the plain call, without the copies
        g->footprint = g->reqs[0]->footprint;
        return launch_one(fr, streams[s]);
    }
    const uint32_t rk = range_keys(fr.key_type, vb, med_keys);
    size_t nr = 0;
    for (Req *r : md) nr += (r->r->n + rk - 1) / rk;
    if (nr > GROUP_RANGES) return set_error(XSS_ERR_INVALID, "group range table overflow (%zu ranges)", nr);
    // pack the tables
    Arena &A = arenas[s];
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 15) & ~size_t(15);
        return o;
    };
    const size_t nsm = sm.size(), nmd = md.size();
    const size_t o_smk = take(nsm * 8), o_smv = take(nsm * 8), o_sml = take(nsm * 4);
    const size_t o_mdk = take(nmd * 8), o_mdv = take(nmd * 8);
    const size_t o_rb = take(nr * 8), o_re = take(nr * 8), o_rs = take(nr * 8);
    const size_t o_rbs = take(nr * 4), o_rh = take(nr * 4);
    if (off > Arena::table_bytes()) return set_error(XSS_ERR_INVALID, "group tables exceed the arena");
    char *h = A.h_tab;
    for (size_t i = 0; i < nsm; ++i) {
        const xss_request &q = *sm[i]->r;
        reinterpret_cast<uint64_t *>(h + o_smk)[i] = reinterpret_cast<uint64_t>(q.keys);
        reinterpret_cast<uint64_t *>(h + o_smv)[i] = reinterpret_cast<uint64_t>(q.vals);
        reinterpret_cast<uint32_t *>(h + o_sml)[i] = uint32_t(q.n);
    }
    uint64_t base = 0;
    size_t r = 0;
    for (size_t i = 0; i < nmd; ++i) {
        const xss_request &q = *md[i]->r;
        reinterpret_cast<uint64_t *>(h + o_mdk)[i] = reinterpret_cast<uint64_t>(q.keys);
        reinterpret_cast<uint64_t *>(h + o_mdv)[i] = reinterpret_cast<uint64_t>(q.vals);
        for (uint64_t x = base; x < base + q.n; x += rk, ++r) {
            reinterpret_cast<uint64_t *>(h + o_rb)[r] = x;
            reinterpret_cast<uint64_t *>(h + o_re)[r] = std::min<uint64_t>(base + q.n, x + rk);
            reinterpret_cast<uint64_t *>(h + o_rs)[r] = base;
            reinterpret_cast<uint32_t *>(h + o_rbs)[r] = uint32_t(i);
            reinterpret_cast<uint32_t *>(h + o_rh)[r] = x == base ? 1u : 0u;
        }
        base += q.n;
    }
    if (off) {
        const hipError_t e = hipMemcpyAsync(A.d_tab, h, off, hipMemcpyHostToDevice, streams[s]);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
    }
    char *d = A.d_tab;
    GroupDev gd{};
    gd.sm_k = reinterpret_cast<const uint64_t *>(d + o_smk);
    gd.sm_v = reinterpret_cast<const uint64_t *>(d + o_smv);
    gd.sm_len = reinterpret_cast<const uint32_t *>(d + o_sml);
    gd.md_k = reinterpret_cast<const uint64_t *>(d + o_mdk);
    gd.md_v = reinterpret_cast<const uint64_t *>(d + o_mdv);
    gd.r_begin = reinterpret_cast<const uint64_t *>(d + o_rb);
    gd.r_end = reinterpret_cast<const uint64_t *>(d + o_re);
    gd.r_sbase = reinterpret_cast<const uint64_t *>(d + o_rs);
    gd.r_bseg = reinterpret_cast<const uint32_t *>(d + o_rbs);
    gd.r_head = reinterpret_cast<const uint32_t *>(d + o_rh);
    gd.rsum = A.rsum;
    gd.rpre = A.rpre;
    gd.segtot = A.segtot;
    const size_t sb = size_t(GROUP_KEYS) * 8;
    gd.kA = A.stage;
    gd.kB = A.stage + sb;
    gd.vA = A.stage + 2 * sb;
    gd.vB = A.stage + 3 * sb;
    gd.nsmall = uint32_t(nsm);
    gd.nmed = uint32_t(nmd);
    gd.nr = uint32_t(nr);
    gd.nkeys = base;
    coalesced += g->reqs.size() > 1 ? g->reqs.size() : 0;
    return run_group_typed(fr.key_type, vb, gd, fr.hasnan != 0, fr.descending != 0, streams[s]);
}

bool xss_server::retire()
{
    bool any = false;
    for (size_t i = 0; i < inflight.size();) {
        Group *g = inflight[i];
        if (hipEventQuery(g->ev) == hipSuccess) {
            const auto now = Clock::now();
            for (Req *r : g->reqs) {
                r->r->latency_us = std::chrono::duration<double, std::micro>(now - r->submitted).count();
                r->done = true;
            }
            used -= g->footprint;
            busy[g->stream] = false;
            completed += g->reqs.size();
            (void)hipEventDestroy(g->ev);
            delete g;
            inflight[i] = inflight.back();
            inflight.pop_back();
            any = true;
        } else {
            ++i;
        }
    }
    if (any) cv_done.notify_all();
    return any;
}

bool xss_server::admit()
{
    bool any = false;
    while (!queue.empty()) {
        if (one_at_a_time && !inflight.empty()) break;
        int s = -1;
        for (size_t i = 0; i < streams.size(); ++i)
            if (!busy[i]) {
                s = int(i);
                break;
            }
        if (s < 0) break;
        Req *front = queue.front();
        if (used + front->footprint > budget && !inflight.empty()) break;  // wait for memory
        Group *g = new Group();
        g->stream = s;
        const xss_request &fr = *front->r;
        int rc = XSS_OK;
        const int vb = group_val_bytes(fr);
        if (coalesce_max && vb >= 0 && fr.n <= std::min<uint64_t>(coalesce_max, GROUP_KEYS)) {
            rc = launch_group(s, vb, g);
        } else {
            queue.pop_front();
            g->reqs.push_back(front);
            g->footprint = front->footprint;
            rc = launch_one(fr, streams[s]);
        }
        if (!rc) {
            hipError_t e = hipEventCreateWithFlags(&g->ev, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventRecord(g->ev, streams[s]);
            if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord");
        }
        if (rc) {
            error = rc;
            error_text = xss_gpu_last_error();
            for (Req *r : g->reqs) {
                r->r->latency_us = -1;
                r->done = true;
            }
            if (g->ev) (void)hipEventDestroy(g->ev);
            delete g;
            cv_done.notify_all();
            continue;
        }
        used += g->footprint;
        busy[s] = true;
        ++groups;
        inflight.push_back(g);
        any = true;
    }
    return any;
}

void xss_server::loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> l(mu);
    while (!stop || !queue.empty() || !inflight.empty()) {
        const bool r = retire();
        const bool a = admit();
        if (!r && !a) {
            if (inflight.empty()) cv_work.wait_for(l, std::chrono::microseconds(200));
            else {
                // GPU work outstanding: poll its events again (a timed sleep
                // would add the OS timer slack, ~50 us, to every latency)
                l.unlock();
                std::this_thread::yield();
                l.lock();
            }
        }
    }
}

extern "C" {

int xss_server_create(xss_server_t *out, size_t budget_bytes, int nstreams, uint32_t coalesce_max,
                      size_t reserve_bytes, int one_at_a_time)
{
    if (!out || nstreams < 1 || nstreams > 16) return set_error(XSS_ERR_INVALID, "bad server arguments");
    std::unique_ptr<xss_server> sv(new xss_server());
    sv->budget = budget_bytes;
    sv->coalesce_max = coalesce_max;
    sv->one_at_a_time = one_at_a_time != 0;
    XSS_HIP_CHECK(hipGetDevice(&sv->device));
    for (int i = 0; i < nstreams; ++i) {
        hipStream_t s;
        XSS_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        sv->streams.push_back(s);
        sv->busy.push_back(false);
        Arena A;
        if (coalesce_max) {
            const size_t tb = Arena::table_bytes();
            XSS_HIP_CHECK(hipHostMalloc(&A.h_tab, tb));
            XSS_HIP_CHECK(hipMalloc(&A.d_tab, tb));
            XSS_HIP_CHECK(hipMalloc(&A.rsum, size_t(GROUP_RANGES) * 256 * 4));
            XSS_HIP_CHECK(hipMalloc(&A.rpre, size_t(GROUP_RANGES) * 256 * 4));
            XSS_HIP_CHECK(hipMalloc(&A.segtot, size_t(MAX_GROUP) * 256 * 4));
            XSS_HIP_CHECK(hipMalloc(&A.stage, size_t(GROUP_KEYS) * 8 * 4));
        }
        sv->arenas.push_back(A);
        if (reserve_bytes) {
            // pool: grow this stream's workspace now, not on the first large request
            Workspace *ws = nullptr;
            std::unique_lock<std::mutex> lock;
            int rc = acquire_workspace(s, reserve_bytes, &ws, &lock);
            if (rc) return rc;
        }
    }
    XSS_HIP_CHECK(hipDeviceSynchronize());
    xss_server *p = sv.get();
    p->worker = std::thread([p] { p->loop(); });
    *out = sv.release();
    return XSS_OK;
}

int xss_server_submit(xss_server_t sv, xss_request *req)
{
    if (!sv || !req) return set_error(XSS_ERR_INVALID, "null argument");
    Req *r = new Req();
    r->r = req;
    r->submitted = Clock::now();
    r->footprint = footprint_of(*req);
    req->latency_us = 0;
    req->internal = r;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->queue.push_back(r);
    }
    sv->cv_work.notify_one();
    return XSS_OK;
}

int xss_server_wait(xss_server_t sv, xss_request *req)
{
    if (!sv || !req || !req->internal) return set_error(XSS_ERR_INVALID, "request was not submitted");
    Req *r = static_cast<Req *>(req->internal);
    std::unique_lock<std::mutex> l(sv->mu);
    sv->cv_done.wait(l, [r] { return r->done; });
    delete r;
    req->internal = nullptr;
    if (req->latency_us < 0) return set_error(sv->error, "request failed: %s", sv->error_text.c_str());
    return XSS_OK;
}

int xss_server_run_trace(xss_server_t sv, xss_request *reqs, size_t count)
{
    if (!sv || (!reqs && count)) return set_error(XSS_ERR_INVALID, "null argument");
    const auto t0 = Clock::now();
    for (size_t i = 0; i < count; ++i) {
        const auto due = t0 + std::chrono::nanoseconds(int64_t(reqs[i].arrival_us * 1000.0));
        while (Clock::now() < due) {
            const auto left = due - Clock::now();
            if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left - std::chrono::microseconds(100));
        }
        int rc = xss_server_submit(sv, &reqs[i]);
        if (rc) return rc;
    }
    int first = XSS_OK;
    for (size_t i = 0; i < count; ++i) {
        const int rc = xss_server_wait(sv, &reqs[i]);
        if (rc && !first) first = rc;
    }
    return first;
}

int xss_server_stats(xss_server_t sv, uint64_t *completed, uint64_t *launch_groups, uint64_t *coalesced)
{
    if (!sv) return set_error(XSS_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> g(sv->mu);
    if (completed) *completed = sv->completed;
    if (launch_groups) *launch_groups = sv->groups;
    if (coalesced) *coalesced = sv->coalesced;
    return XSS_OK;
}

int xss_server_destroy(xss_server_t sv)
{
    if (!sv) return XSS_OK;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    for (size_t i = 0; i < sv->streams.size(); ++i) {
        (void)hipStreamSynchronize(sv->streams[i]);
        Arena &A = sv->arenas[i];
        if (A.h_tab) (void)hipHostFree(A.h_tab);
        for (void *m : {(void *)A.d_tab, (void *)A.rsum, (void *)A.rpre, (void *)A.segtot, (void *)A.stage})
            if (m) (void)hipFree(m);
        (void)hipStreamDestroy(sv->streams[i]);
    }
    delete sv;
    return XSS_OK;
}

}  // extern "C"
