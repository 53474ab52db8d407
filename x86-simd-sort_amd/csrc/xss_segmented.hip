// xss_segmented.hip -- segmented sort / argsort / select of many independent
// arrays (segments of one concatenated buffer) in one pipeline.
//
// The reference sorts one array per call; a batch of thousands of uneven
// arrays costs a call per array.  Here a host-built plan buckets the
// segments by size once:
//   * small segments (len <= small_max): one LDS bitonic workgroup each
//     (block_sort_kernel over a segment list) -- one launch for all of them;
//   * 4- and 8-byte keys (u32 / i32 / f32 / u64 / i64 / f64):
//       - medium segments (up to the LDS radix sort's capacity): one
//         workgroup each sorts the segment by its order key in LDS
//         (ss_local_kernel, xss_sample.hpp);
//       - large segments: a sample sort (xss_sample.hpp) -- splitters from a
//         sorted sample, one histogram read and one scatter into buckets of
//         about `bucket_keys` keys, then every bucket sorted in LDS; qselect
//         sorts only the bucket holding rank len / 2;
//   * 2-byte keys, larger segments: cut into ranges of range_tiles
//     downsweep tiles; every LSD pass is upsweep (one workgroup per range) ->
//     seg_scan_kernel (per digit, an exclusive scan over the ranges of each
//     segment, restarting at every segment) -> seg_downsweep_kernel (one
//     workgroup per range, running per-digit positions inside its segment).
// The plan owns all device memory it needs, so xss_gpu_segplan_run allocates
// nothing and its launches can be captured in a hipGraph and replayed.
#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "xss_sample.hpp"

using namespace xss;

namespace {

struct Tuning {
    uint32_t small_max;    // segments up to this length: LDS bitonic
    uint32_t range_tiles;  // tiles per workgroup range in larger segments
    uint32_t bucket_keys;  // sample sort: keys per bucket (average)
    uint32_t oversample;   // sample sort: samples per bucket
    uint32_t local_items;  // LDS sort: keys per thread, 12 or 15 (capacity 512 x items)
    uint32_t wide_tiles;   // sample sort's scatter: 1024-thread tiles (else 512)
};

}  // namespace

namespace xss {

// Plan knobs of the segmented form's sample sort (xss_gpu_set_tuning,
// read when a plan is created); 0 = the default.
std::atomic<uint64_t> &seg_knob(int which)
{
    static std::atomic<uint64_t> v[5];
    return v[which];
}
enum { SEG_BUCKET_KEYS = 0, SEG_OVERSAMPLE, SEG_LOCAL_ITEMS, SEG_WIDE_TILES, SEG_SAMPLE_OFF };

bool seg_tuning_set(const std::string &name, uint64_t value)
{
    static const char *const names[5] = {"seg_bucket_keys", "seg_oversample", "seg_local_items", "seg_wide_tiles",
                                         "seg_sample_off"};
    for (int i = 0; i < 5; ++i)
        if (name == names[i]) {
            seg_knob(i).store(value);
            return true;
        }
    return false;
}

}  // namespace xss

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
    // sample sort (4- and 8-byte keys)
    bool sample = false;
    uint32_t nmed = 0, nbk = 0, lcap = 0;
    uint32_t *med_list = nullptr, *lseg = nullptr, *lnb = nullptr, *lbb = nullptr, *bk_seg = nullptr;
    void *split = nullptr;
    uint32_t *bstart = nullptr, *bsize = nullptr, *ctl = nullptr, *big = nullptr, *ksel = nullptr;
    void *brange = nullptr;
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

// sample-sort plan of 4- and 8-byte keys: small segments (bitonic), medium (LDS
// radix, in place), large (sample sort: keys -> scratch -> keys)
template <class KD, int PL, bool NL, int ITEMS>
int ss_run(xss_segplan *p, void *keys, uint32_t *idx32, bool desc, hipStream_t s)
{
    using U = typename KD::U;
    using BS = BlockShape<KD, PL>;
    constexpr bool IDX = PL == PL_IDX;
    const uint64_t flip = desc ? ~0ull : 0ull;
    const uint32_t canon = (p->op == 0 && NL) ? 1u : 0u;
    if (p->nsmall) {
        BlockSortArgs b{};
        b.keys_in = keys; b.keys_out = keys; b.idx32_out = idx32; b.seg_off = p->seg_off;
        b.seg_list = p->small_list; b.nan_last = NL; b.flip = flip; b.canon_nan = canon;
        KernelTimer t("seg_small", s);
        hipLaunchKernelGGL((block_sort_kernel<KD, PL, BS::BLOCK, BS::MAXN>), dim3(p->nsmall), dim3(BS::BLOCK), 0, s, b);
        XSS_HIP_CHECK(hipGetLastError());
    }
    // scratch: the scatter's output (keys, argsort indices) and, for
    // argsort, the other buffer pair of the large-bucket passes
    char *yk = p->scratch, *yv = nullptr, *zk = static_cast<char *>(keys), *zv = nullptr;
    if (IDX) {
        yv = yk + align256(p->n * sizeof(U));
        zk = yv + align256(p->n * 4);
        zv = zk + align256(p->n * sizeof(U));
    }
    SsLocalArgs la{};
    la.bstart = p->bstart; la.bsize = p->bsize; la.brange = p->brange; la.bk_seg = p->bk_seg; la.lseg = p->lseg; la.seg_off = p->seg_off;
    la.idx32_out = idx32; la.flip = flip; la.canon = canon; la.keys_out = keys;
    if (p->nmed) {
        SsLocalArgs m = la;
        m.med = p->med_list; m.keys_in = keys;
        KernelTimer t("seg_medium", s);
        hipLaunchKernelGGL((ss_local_kernel<KD, PL, NL, ITEMS>), dim3(p->nmed), dim3(SS_LOCAL_BLOCK), 0, s, m);
        XSS_HIP_CHECK(hipGetLastError());
    }
    if (!p->nbig) return XSS_OK;
    SsArgs a{};
    a.seg_off = p->seg_off; a.lseg = p->lseg; a.lnb = p->lnb; a.lbb = p->lbb; a.bk_seg = p->bk_seg;
    a.oversample = p->tune.oversample; a.split = p->split; a.bstart = p->bstart; a.bsize = p->bsize;
    a.brange = p->brange;
    a.ctl = p->ctl; a.big = p->big; a.ksel = p->ksel; a.lcap = p->lcap; a.select = p->op == 2;
    a.flip = flip; a.nan_last = NL;
    {
        KernelTimer t("ss_sample", s);
        hipLaunchKernelGGL((ss_sample_kernel<KD>), dim3(p->nbig), dim3(SS_SAMPLE_BLOCK), 0, s, a,
                           static_cast<const void *>(keys));
    }
    XSS_HIP_CHECK(hipGetLastError());
    {
        KernelTimer t("ss_hist", s);
        // select: the histogram read also copies the keys to scratch, from
        // where the partition writes them back
        hipLaunchKernelGGL((ss_hist_kernel<KD>), dim3(p->nr), dim3(256), 0, s, a, SsRanges{p->r_begin, p->r_end, p->r_bseg},
                           static_cast<const void *>(keys), p->rsum, p->nr,
                           p->op == 2 ? static_cast<void *>(yk) : nullptr);
    }
    XSS_HIP_CHECK(hipGetLastError());
    SegScanArgs sc{p->rsum, p->rpre, p->segtot, p->r_head, p->r_bseg, p->nr, p->nr, nullptr};
    {
        KernelTimer t("seg_scan", s);
        hipLaunchKernelGGL(seg_scan_kernel<0>, dim3(256), dim3(256), 0, s, sc);
    }
    XSS_HIP_CHECK(hipGetLastError());
    if (p->op != 2) {
        PassArgs pa{};
        pa.keys_in = keys; pa.keys_out = yk; pa.vals_out = yv; pa.n = uint32_t(p->n); pa.flip = flip;
        pa.nan_last = NL; pa.first = 1; pa.last = 0;
        SegRanges r{p->r_begin, p->r_end, p->r_sbase, p->r_bseg, p->rpre, p->segtot, p->nr, p->nr, nullptr};
        constexpr bool UN = !IDX;  // keys only: the LDS sort orders each bucket
        KernelTimer t("ss_scatter", s);
        if (p->tune.wide_tiles) {
            using SS = SsWideShape<KD, PL>;
            hipLaunchKernelGGL((ss_scatter_kernel<KD, PL, SS::BLOCK, SS::ITEMS, NL, UN>), dim3(p->nr), dim3(SS::BLOCK), 0,
                               s, pa, r, a);
        } else {
            using SS = SegShape<KD, PL>;
            hipLaunchKernelGGL((ss_scatter_kernel<KD, PL, SS::BLOCK, SS::ITEMS, NL, UN>), dim3(p->nr), dim3(SS::BLOCK), 0,
                               s, pa, r, a);
        }
    }
    XSS_HIP_CHECK(hipGetLastError());
    {
        KernelTimer t("ss_buckets", s);
        hipLaunchKernelGGL(ss_buckets_kernel<U>, dim3(p->nbig), dim3(256), 0, s, a, static_cast<const uint32_t *>(p->segtot));
    }
    XSS_HIP_CHECK(hipGetLastError());
    la.keys_in = yk; la.vals_in = yv;
    if (p->op == 2) {
        // select: the copy partitioned back into the caller's keys (below /
        // inside / above the bucket holding rank len / 2), then that bucket
        // of every large segment sorted in place
        SegRanges r{p->r_begin, p->r_end, p->r_sbase, p->r_bseg, p->rpre, p->segtot, p->nr, p->nr, nullptr};
        KernelTimer t("ss_part3", s);
        hipLaunchKernelGGL((ss_part3_kernel<KD, NL>), dim3(p->nr), dim3(SS_P3_BLOCK), 0, s, a, r,
                           static_cast<const void *>(yk), keys);
        XSS_HIP_CHECK(hipGetLastError());
        la.slots = p->ksel;
        la.keys_in = keys;
    }
    {
        KernelTimer t("ss_local", s);
        hipLaunchKernelGGL((ss_local_kernel<KD, PL, NL, ITEMS>), dim3(p->op == 2 ? p->nbig : p->nbk),
                           dim3(SS_LOCAL_BLOCK), 0, s, la);
    }
    XSS_HIP_CHECK(hipGetLastError());
    {
        SsBigArgs g{};
        g.ctl = p->ctl; g.big = p->big; g.bstart = p->bstart; g.bsize = p->bsize; g.bk_seg = p->bk_seg; g.lseg = p->lseg;
        g.seg_off = p->seg_off; g.y_keys = yk; g.y_vals = yv; g.z_keys = zk; g.z_vals = zv; g.idx32_out = idx32;
        g.flip = flip; g.canon = canon;
        if (p->op == 2) {  // select: the bucket is in the caller's keys and ends there
            g.y_keys = keys;
            g.z_keys = yk;
            g.end_in_y = 1;
        }
        KernelTimer t("ss_big", s);
        hipLaunchKernelGGL((ss_big_kernel<KD, PL, NL>), dim3(std::min<uint32_t>(p->nbk, 256)), dim3(SS_BIG_BLOCK), 0, s,
                           g);
    }
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

template <class KD, int PL>
int ss_dispatch(xss_segplan *p, void *keys, uint32_t *idx32, bool hasnan, bool desc, hipStream_t s)
{
    const bool nl = (KD::KIND == Kind::Float) && hasnan;
    if constexpr (sizeof(typename KD::U) == 8) {
        // a larger capacity for 8-byte keys: 12 items (keys only) or 11
        // (argsort: keys and indices of two workgroups fit one CU's LDS)
        constexpr int WIDE = PL == PL_IDX ? 11 : 12;
        if (p->tune.local_items == WIDE) {
            if (nl) return ss_run<KD, PL, true, WIDE>(p, keys, idx32, desc, s);
            return ss_run<KD, PL, false, WIDE>(p, keys, idx32, desc, s);
        }
        if (nl) return ss_run<KD, PL, true, SS_LOCAL_ITEMS_W8>(p, keys, idx32, desc, s);
        return ss_run<KD, PL, false, SS_LOCAL_ITEMS_W8>(p, keys, idx32, desc, s);
    } else {
        if (p->tune.local_items == 15) {
            if (nl) return ss_run<KD, PL, true, 15>(p, keys, idx32, desc, s);
            return ss_run<KD, PL, false, 15>(p, keys, idx32, desc, s);
        }
        if (nl) return ss_run<KD, PL, true, 12>(p, keys, idx32, desc, s);
        return ss_run<KD, PL, false, 12>(p, keys, idx32, desc, s);
    }
}

template <class KD>
int seg_dispatch_op(xss_segplan *p, void *keys, uint32_t *idx32, bool hasnan, bool desc, hipStream_t s)
{
    if constexpr (sizeof(typename KD::U) >= 4) {
        if (p->sample) {
            if (p->op == 1) return ss_dispatch<KD, PL_IDX>(p, keys, idx32, hasnan, desc, s);
            return ss_dispatch<KD, PL_NONE>(p, keys, nullptr, hasnan, desc, s);
        }
    }
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
    // sample sort of 4- and 8-byte keys (unless switched off): plan knobs
    p->sample = kb >= 4 && seg_knob(SEG_SAMPLE_OFF).load() == 0;
    if (p->sample) {
        const uint64_t items = seg_knob(SEG_LOCAL_ITEMS).load();
        if (kb == 8) {
            const uint32_t wide = p->op == 1 ? 11 : 12;
            p->tune.local_items = items == 8 ? SS_LOCAL_ITEMS_W8 : (items == 0 || items == wide) ? wide
                                                                                                   : SS_LOCAL_ITEMS_W8;
        } else {
            p->tune.local_items = items == 15 ? 15 : 12;
        }
        p->lcap = uint32_t(SS_LOCAL_BLOCK) * p->tune.local_items;
        const uint64_t bk = seg_knob(SEG_BUCKET_KEYS).load();
        // measured on 8192 arrays of log-uniform length: 4096-key buckets,
        // 32 samples per bucket (capped at SS_SAMPLE_MAX per segment)
        p->tune.bucket_keys = bk ? uint32_t(std::min<uint64_t>(bk, 1u << 30)) : p->lcap * 2 / 3;
        const uint64_t ov = seg_knob(SEG_OVERSAMPLE).load();
        p->tune.oversample = ov ? uint32_t(std::min<uint64_t>(ov, SS_SAMPLE_MAX)) : 32;
        p->tune.wide_tiles = seg_knob(SEG_WIDE_TILES).load() ? 1 : 0;
        // the scatter's tile: SsWideShape with wide tiles, else SegShape
        if (p->tune.wide_tiles) tile = 1024u * uint32_t(ss_wide_items(kb + (p->op == 1 ? 4 : 0)));
    }
    const uint64_t rkeys = uint64_t(p->tune.range_tiles) * tile;
    // bucket the segments
    std::vector<uint32_t> small, med, lseg, lnb, lbb, bkseg;
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
        if (p->sample && len <= p->lcap) {
            med.push_back(uint32_t(i));
            continue;
        }
        if (p->sample) {
            const uint64_t nb = std::min<uint64_t>(SS_SPLIT, std::max<uint64_t>(2, (len + p->tune.bucket_keys - 1) /
                                                                                     p->tune.bucket_keys));
            lseg.push_back(uint32_t(i));
            lnb.push_back(uint32_t(nb));
            lbb.push_back(uint32_t(bkseg.size()));
            bkseg.insert(bkseg.end(), nb, nbig);
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
    p->nmed = uint32_t(med.size());
    p->nr = uint32_t(rb.size());
    p->nbig = nbig;
    p->nbk = uint32_t(bkseg.size());
    // device memory: metadata + scratch
    const size_t nr = p->nr, nbk = p->nbk;
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
    size_t scratch = 0;
    if (p->nr) {
        // LSD: keys ping-pong (argsort: two key + index pairs); sample sort:
        // the scatter's output (argsort: two key + index pairs as well)
        scratch = p->op == 1 ? 2 * (align256(p->n * kb) + align256(p->n * 4)) : align256(p->n * kb);
    }
    const size_t o_scr = take(scratch + 256);
    const size_t o_med = take(med.size() * 4 + 4), o_lseg = take(size_t(nbig) * 4 + 4), o_lnb = take(size_t(nbig) * 4 + 4);
    const size_t o_lbb = take(size_t(nbig) * 4 + 4), o_bks = take(nbk * 4 + 4), o_split = take(size_t(nbig) * SS_SPLIT * kb + 4);
    const size_t o_bst = take(nbk * 4 + 4), o_bsz = take(nbk * 4 + 4), o_big = take(nbk * 4 + 4);
    const size_t o_brg = take(nbk * 2 * kb + 8);
    const size_t o_ksel = take(size_t(nbig) * 4 + 4), o_ctl = take(16);
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
    p->med_list = reinterpret_cast<uint32_t *>(m + o_med);
    p->lseg = reinterpret_cast<uint32_t *>(m + o_lseg);
    p->lnb = reinterpret_cast<uint32_t *>(m + o_lnb);
    p->lbb = reinterpret_cast<uint32_t *>(m + o_lbb);
    p->bk_seg = reinterpret_cast<uint32_t *>(m + o_bks);
    p->split = m + o_split;
    p->bstart = reinterpret_cast<uint32_t *>(m + o_bst);
    p->bsize = reinterpret_cast<uint32_t *>(m + o_bsz);
    p->brange = m + o_brg;
    p->big = reinterpret_cast<uint32_t *>(m + o_big);
    p->ksel = reinterpret_cast<uint32_t *>(m + o_ksel);
    p->ctl = reinterpret_cast<uint32_t *>(m + o_ctl);
    auto up = [](void *dst, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    XSS_HIP_CHECK(up(p->seg_off, offsets, nseg ? (nseg + 1) * 8 : 0));
    XSS_HIP_CHECK(up(p->small_list, small.data(), small.size() * 4));
    XSS_HIP_CHECK(up(p->r_begin, rb.data(), nr * 8));
    XSS_HIP_CHECK(up(p->r_end, re.data(), nr * 8));
    XSS_HIP_CHECK(up(p->r_sbase, rs.data(), nr * 8));
    XSS_HIP_CHECK(up(p->r_bseg, rbs.data(), nr * 4));
    XSS_HIP_CHECK(up(p->r_head, rh.data(), nr * 4));
    XSS_HIP_CHECK(up(p->med_list, med.data(), med.size() * 4));
    XSS_HIP_CHECK(up(p->lseg, lseg.data(), lseg.size() * 4));
    XSS_HIP_CHECK(up(p->lnb, lnb.data(), lnb.size() * 4));
    XSS_HIP_CHECK(up(p->lbb, lbb.data(), lbb.size() * 4));
    XSS_HIP_CHECK(up(p->bk_seg, bkseg.data(), bkseg.size() * 4));
    // select's slot per segment starts valid (slot 0) before any run writes it
    XSS_HIP_CHECK(hipMemset(p->ksel, 0, size_t(nbig) * 4 + 4));
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
