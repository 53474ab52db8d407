// xss_segmented.hip -- segmented sort / argsort / select of many independent
// arrays (segments of one concatenated buffer) in one pipeline.
//
// The reference sorts one array per call; a batch of thousands of uneven
// arrays costs a call per array.  Here a host-built plan buckets the
// segments by size once:
//   * small segments (len <= small_max): one LDS bitonic workgroup each
//     (block_sort_kernel over a segment list) -- one launch for all of them;
//   * larger segments: cut into ranges of range_tiles downsweep tiles; every
//     LSD pass is upsweep (one workgroup per range) -> seg_scan_kernel (per
//     digit, an exclusive scan over the ranges of each segment, restarting at
//     every segment) -> seg_downsweep_kernel (one workgroup per range, running
//     per-digit positions inside its segment).  A 1M-key segment is spread
//     over many workgroups; 4K-key segments take one range each.
// The plan owns all device memory it needs, so xss_gpu_segplan_run allocates
// nothing and its launches can be captured in a hipGraph and replayed.
#include <memory>
#include <string>
#include <vector>

#include "xss_sort.hpp"

using namespace xss;

namespace {

struct Tuning {
    uint32_t small_max;    // segments up to this length: LDS bitonic
    uint32_t range_tiles;  // tiles per workgroup range in larger segments
};

}  // namespace

// ------------------------------------------------------------------ plan

struct xss_segplan {
    int op = 0;                 // 0 qsort, 1 argsort, 2 qselect
    std::string tag;            // key type tag
    size_t nseg = 0, n = 0;
    Tuning tune{};
    uint32_t nsmall = 0, nr = 0, nbig = 0;
    int device = 0;
    char *mem = nullptr;        // one device allocation for everything below
    uint64_t *seg_off = nullptr;
    uint32_t *small_list = nullptr;
    uint64_t *r_begin = nullptr, *r_end = nullptr, *r_sbase = nullptr;
    uint32_t *r_bseg = nullptr, *r_head = nullptr;
    uint32_t *rsum = nullptr, *rpre = nullptr, *segtot = nullptr;
    char *scratch = nullptr;    // ping-pong keys (+ indices)
};

namespace {

int key_bytes(const std::string &tag)
{
    if (tag == "u16" || tag == "i16" || tag == "f16") return 2;
    if (tag == "u32" || tag == "i32" || tag == "f32") return 4;
    if (tag == "u64" || tag == "i64" || tag == "f64") return 8;
    return 0;
}

template <class KD, int PL>
int seg_run(xss_segplan *p, void *keys, uint32_t *idx32, bool hasnan, bool desc, hipStream_t s)
{
    using U = typename KD::U;
    using PS = SegShape<KD, PL>;
    using BS = BlockShape<KD, PL>;
    constexpr int P = KD::PASSES;
    const uint64_t flip = desc ? ~0ull : 0ull;
    const bool nan_last = (KD::KIND == Kind::Float) && hasnan;
    const bool canon = p->op == 0;
    if (p->nsmall) {
        BlockSortArgs b{};
        b.keys_in = keys; b.keys_out = keys; b.idx32_out = idx32; b.seg_off = p->seg_off;
        b.seg_list = p->small_list; b.nan_last = nan_last; b.flip = flip; b.canon_nan = canon && nan_last;
        KernelTimer t("seg_small", s);
        hipLaunchKernelGGL((block_sort_kernel<KD, PL, BS::BLOCK, BS::MAXN>), dim3(p->nsmall), dim3(BS::BLOCK), 0, s, b);
        XSS_HIP_CHECK(hipGetLastError());
    }
    if (!p->nr) return XSS_OK;
    char *k0 = p->scratch, *i0 = nullptr, *k1 = nullptr, *i1 = nullptr;
    if (PL == PL_IDX) {
        i0 = k0 + align256(p->n * sizeof(U));
        k1 = i0 + align256(p->n * 4);
        i1 = k1 + align256(p->n * sizeof(U));
    }
    const void *src_k = keys, *src_v = nullptr;
    for (int pass = 0; pass < P; ++pass) {
        UpArgs u{};
        u.keys = src_k; u.n = uint32_t(p->n); u.shift = 8 * pass; u.nan_last = nan_last; u.flip = flip;
        u.rsum = p->rsum; u.ranges = p->nr; u.range_begin = p->r_begin; u.range_end = p->r_end;
        {
            KernelTimer t("seg_upsweep", s);
            hipLaunchKernelGGL((upsweep_kernel<KD>), dim3(p->nr), dim3(UP_BLOCK), 0, s, u);
        }
        XSS_HIP_CHECK(hipGetLastError());
        SegScanArgs sc{p->rsum, p->rpre, p->segtot, p->r_head, p->r_bseg, p->nr, p->nr, nullptr};
        {
            KernelTimer t("seg_scan", s);
            hipLaunchKernelGGL(seg_scan_kernel<0>, dim3(256), dim3(256), 0, s, sc);
        }
        XSS_HIP_CHECK(hipGetLastError());
        PassArgs a{};
        a.keys_in = src_k; a.vals_in = src_v; a.n = uint32_t(p->n); a.shift = 8 * pass; a.flip = flip;
        a.nan_last = nan_last; a.first = pass == 0; a.last = pass == P - 1;
        a.canon_nan = (pass == P - 1) && canon && nan_last;
        a.idx32_out = idx32;
        if (PL == PL_IDX) {
            if (pass < P - 1) {
                a.keys_out = (pass % 2 == 0) ? k0 : k1;
                a.vals_out = (pass % 2 == 0) ? i0 : i1;
            }
        } else {
            a.keys_out = (pass % 2 == 0) ? static_cast<void *>(k0) : keys;
        }
        SegRanges r{p->r_begin, p->r_end, p->r_sbase, p->r_bseg, p->rpre, p->segtot, p->nr, p->nr, nullptr};
        {
            KernelTimer t("seg_downsweep", s);
            if (nan_last)
                hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, true>), dim3(p->nr),
                                   dim3(PS::BLOCK), 0, s, a, r);
            else
                hipLaunchKernelGGL((seg_downsweep_kernel<KD, PL, PS::BLOCK, PS::ITEMS, false>), dim3(p->nr),
                                   dim3(PS::BLOCK), 0, s, a, r);
        }
        XSS_HIP_CHECK(hipGetLastError());
        src_k = a.keys_out;
        src_v = a.vals_out;
    }
    return XSS_OK;
}

template <class KD>
int seg_dispatch_op(xss_segplan *p, void *keys, uint32_t *idx32, bool hasnan, bool desc, hipStream_t s)
{
    if (p->op == 1) return seg_run<KD, PL_IDX>(p, keys, idx32, hasnan, desc, s);
    return seg_run<KD, PL_NONE>(p, keys, nullptr, hasnan, desc, s);
}

int seg_dispatch(xss_segplan *p, void *keys, uint32_t *idx32, bool hasnan, bool desc, hipStream_t s)
{
    const std::string &t = p->tag;
    if (t == "u16") return seg_dispatch_op<TypeKey<uint16_t>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "i16") return seg_dispatch_op<TypeKey<int16_t>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "f16") return seg_dispatch_op<TypeKey<f16_tag>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "u32") return seg_dispatch_op<TypeKey<uint32_t>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "i32") return seg_dispatch_op<TypeKey<int32_t>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "f32") return seg_dispatch_op<TypeKey<float>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "u64") return seg_dispatch_op<TypeKey<uint64_t>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "i64") return seg_dispatch_op<TypeKey<int64_t>::KD>(p, keys, idx32, hasnan, desc, s);
    if (t == "f64") return seg_dispatch_op<TypeKey<double>::KD>(p, keys, idx32, hasnan, desc, s);
    return set_error(XSS_ERR_INVALID, "unknown key type '%s'", t.c_str());
}

// largest LDS bitonic size and downsweep tile for a type / op
void shape_of(const std::string &tag, int op, uint32_t *maxn, uint32_t *tile)
{
    const int kb = key_bytes(tag);
    const bool idx = op == 1;
    // BlockShape of xss_sort.hpp and SegShape above, for keys-only and index payloads
    const int bpe = 2 * kb + 2;
    *maxn = (65536 / bpe) >= 4096 ? 4096 : (65536 / bpe) >= 2048 ? 2048 : 1024;
    const int bytes = kb + (idx ? 4 : 0);
    *tile = bytes == 4 ? uint32_t(XSS_SEG_BLOCK_C1) * XSS_SEG_ITEMS_C1 : 512u * (bytes > 8 ? 8 : 16);
}

}  // namespace

extern "C" {

int xss_gpu_segplan_create(xss_segplan_t *out, const char *op, const char *key_type, const uint64_t *offsets,
                           size_t nseg, uint32_t small_max, uint32_t range_tiles)
{
    if (!out || !op || !key_type || (!offsets && nseg)) return set_error(XSS_ERR_INVALID, "null pointer argument");
    *out = nullptr;
    std::unique_ptr<xss_segplan> p(new xss_segplan());
    const std::string o(op);
    p->op = o == "qsort" ? 0 : o == "argsort" ? 1 : o == "qselect" ? 2 : -1;
    if (p->op < 0) return set_error(XSS_ERR_INVALID, "unknown segmented op '%s'", op);
    p->tag = key_type;
    const int kb = key_bytes(p->tag);
    if (!kb) return set_error(XSS_ERR_INVALID, "unknown key type '%s'", key_type);
    p->nseg = nseg;
    p->n = nseg ? offsets[nseg] : 0;
    if (p->n > MAX_N) return set_error(XSS_ERR_INVALID, "batch of %zu keys exceeds the limit %u", p->n, MAX_N);
    for (size_t i = 0; i < nseg; ++i)
        if (offsets[i + 1] < offsets[i]) return set_error(XSS_ERR_INVALID, "segment offsets must not decrease");
    uint32_t maxn, tile;
    shape_of(p->tag, p->op, &maxn, &tile);
    p->tune.small_max = small_max ? std::min(small_max, maxn) : maxn;
    p->tune.range_tiles = range_tiles ? range_tiles : 8;
    const uint64_t rkeys = uint64_t(p->tune.range_tiles) * tile;
    // bucket the segments
    std::vector<uint32_t> small;
    std::vector<uint64_t> rb, re, rs;
    std::vector<uint32_t> rbs, rh;
    uint32_t nbig = 0;
    for (size_t i = 0; i < nseg; ++i) {
        const uint64_t b = offsets[i], len = offsets[i + 1] - b;
        if (len < 2) {
            if (len == 1 && p->op == 1) small.push_back(uint32_t(i));  // argsort still writes [0]
            continue;
        }
        if (len <= p->tune.small_max) {
            small.push_back(uint32_t(i));
            continue;
        }
        for (uint64_t x = b; x < b + len; x += rkeys) {
            rb.push_back(x);
            re.push_back(std::min(b + len, x + rkeys));
            rs.push_back(b);
            rbs.push_back(nbig);
            rh.push_back(x == b ? 1u : 0u);
        }
        ++nbig;
    }
    p->nsmall = uint32_t(small.size());
    p->nr = uint32_t(rb.size());
    p->nbig = nbig;
    // device memory: metadata + ping-pong scratch
    const size_t nr = p->nr;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o2 = off;
        off += align256(bytes);
        return o2;
    };
    const size_t o_seg = take((nseg + 1) * 8), o_small = take(small.size() * 4 + 4);
    const size_t o_rb = take(nr * 8 + 8), o_re = take(nr * 8 + 8), o_rs = take(nr * 8 + 8);
    const size_t o_rbs = take(nr * 4 + 4), o_rh = take(nr * 4 + 4);
    const size_t o_rsum = take(nr * 256 * 4 + 4), o_rpre = take(nr * 256 * 4 + 4), o_tot = take(size_t(nbig) * 1024 + 4);
    const size_t scratch = p->nr ? (p->op == 1 ? 2 * (align256(p->n * kb) + align256(p->n * 4)) : align256(p->n * kb)) : 0;
    const size_t o_scr = take(scratch + 256);
    XSS_HIP_CHECK(hipGetDevice(&p->device));
    XSS_HIP_CHECK(hipMalloc(&p->mem, off));
    char *m = p->mem;
    p->seg_off = reinterpret_cast<uint64_t *>(m + o_seg);
    p->small_list = reinterpret_cast<uint32_t *>(m + o_small);
    p->r_begin = reinterpret_cast<uint64_t *>(m + o_rb);
    p->r_end = reinterpret_cast<uint64_t *>(m + o_re);
    p->r_sbase = reinterpret_cast<uint64_t *>(m + o_rs);
    p->r_bseg = reinterpret_cast<uint32_t *>(m + o_rbs);
    p->r_head = reinterpret_cast<uint32_t *>(m + o_rh);
    p->rsum = reinterpret_cast<uint32_t *>(m + o_rsum);
    p->rpre = reinterpret_cast<uint32_t *>(m + o_rpre);
    p->segtot = reinterpret_cast<uint32_t *>(m + o_tot);
    p->scratch = m + o_scr;
    if (nseg) XSS_HIP_CHECK(hipMemcpy(p->seg_off, offsets, (nseg + 1) * 8, hipMemcpyHostToDevice));
    if (!small.empty()) XSS_HIP_CHECK(hipMemcpy(p->small_list, small.data(), small.size() * 4, hipMemcpyHostToDevice));
    if (nr) {
        XSS_HIP_CHECK(hipMemcpy(p->r_begin, rb.data(), nr * 8, hipMemcpyHostToDevice));
        XSS_HIP_CHECK(hipMemcpy(p->r_end, re.data(), nr * 8, hipMemcpyHostToDevice));
        XSS_HIP_CHECK(hipMemcpy(p->r_sbase, rs.data(), nr * 8, hipMemcpyHostToDevice));
        XSS_HIP_CHECK(hipMemcpy(p->r_bseg, rbs.data(), nr * 4, hipMemcpyHostToDevice));
        XSS_HIP_CHECK(hipMemcpy(p->r_head, rh.data(), nr * 4, hipMemcpyHostToDevice));
    }
    *out = p.release();
    return XSS_OK;
}

int xss_gpu_segplan_run(xss_segplan_t p, void *keys, uint32_t *idx_out, int hasnan, int descending, xss_stream_t s)
{
    if (!p) return set_error(XSS_ERR_INVALID, "null plan");
    if (p->n == 0) return XSS_OK;
    if (!keys || (p->op == 1 && !idx_out)) return set_error(XSS_ERR_INVALID, "null pointer argument");
    return seg_dispatch(p, keys, idx_out, hasnan != 0, descending != 0, s);
}

int xss_gpu_segplan_info(xss_segplan_t p, uint32_t *nsmall, uint32_t *nranges, uint32_t *nlarge)
{
    if (!p) return set_error(XSS_ERR_INVALID, "null plan");
    if (nsmall) *nsmall = p->nsmall;
    if (nranges) *nranges = p->nr;
    if (nlarge) *nlarge = p->nbig;
    return XSS_OK;
}

int xss_gpu_segplan_destroy(xss_segplan_t p)
{
    if (!p) return XSS_OK;
    if (p->mem) XSS_HIP_CHECK(hipFree(p->mem));
    delete p;
    return XSS_OK;
}

// The baseline the segmented form is measured against: one plain call per
// segment (xss_gpu_qsort_f32 on each sub-array), enqueued from C++.
int xss_gpu_loop_qsort_f32(float *keys, const uint64_t *offsets, size_t nseg, int hasnan, int descending,
                           xss_stream_t s)
{
    for (size_t i = 0; i < nseg; ++i) {
        const int rc = xss_gpu_qsort_f32(keys + offsets[i], offsets[i + 1] - offsets[i], hasnan, descending, s);
        if (rc) return rc;
    }
    return XSS_OK;
}

int xss_gpu_loop_argsort_i32(const int32_t *keys, uint64_t *idx, const uint64_t *offsets, size_t nseg,
                             xss_stream_t s)
{
    for (size_t i = 0; i < nseg; ++i) {
        const int rc = xss_gpu_argsort_i32(keys + offsets[i], idx + offsets[i], offsets[i + 1] - offsets[i], 0, 0, s);
        if (rc) return rc;
    }
    return XSS_OK;
}

int xss_gpu_loop_qselect_f32(float *keys, const uint64_t *offsets, size_t nseg, int hasnan, xss_stream_t s)
{
    for (size_t i = 0; i < nseg; ++i) {
        const size_t len = offsets[i + 1] - offsets[i];
        if (len < 2) continue;
        const int rc = xss_gpu_qselect_f32(keys + offsets[i], len / 2, len, hasnan, 0, s);
        if (rc) return rc;
    }
    return XSS_OK;
}

}  // extern "C"
