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
//   * coalesces queued qsort requests short enough for one LDS workgroup
//     (same key type and flags) into one launch of the pointer-list block
//     sort; every other request runs through its C-ABI entry point on its
//     stream, so up to S requests overlap on the device while new ones keep
//     arriving and are admitted as streams and memory free up.
// Each stream keeps the library's pooled workspace, reserved up front, so
// steady-state requests allocate nothing.
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

namespace {

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
        switch (kt) {
        case 3: XSS_KV(u32, uint32_t);
        case 4: XSS_KV(i32, int32_t);
        case 5: XSS_KV(f32, float);
        case 6: XSS_KV(u64, uint64_t);
        case 7: XSS_KV(i64, int64_t);
        case 8: XSS_KV(f64, double);
        default: return set_error(XSS_ERR_INVALID, "keyvalue_qsort has no %s keys", TAG[kt]);
        }
#undef XSS_KV
    }
    }
    return set_error(XSS_ERR_INVALID, "bad op %d", r.op);
}

// one launch sorting many short arrays (pointer-list block sort)
template <class KD>
int launch_small_group(const uint64_t *d_ptrs, const uint32_t *d_len, uint32_t count, bool hasnan, bool desc,
                       hipStream_t s)
{
    using BS = BlockShape<KD, PL_NONE>;
    BlockSortArgs b{};
    b.arr_keys = d_ptrs;
    b.arr_len = d_len;
    b.nan_last = (KD::KIND == Kind::Float) && hasnan;
    b.flip = desc ? ~0ull : 0ull;
    b.canon_nan = b.nan_last;
    KernelTimer t("serve_small", s);
    hipLaunchKernelGGL((block_sort_kernel<KD, PL_NONE, BS::BLOCK, BS::MAXN>), dim3(count), dim3(BS::BLOCK), 0, s, b);
    XSS_HIP_CHECK(hipGetLastError());
    return XSS_OK;
}

int small_group(int kt, const uint64_t *d_ptrs, const uint32_t *d_len, uint32_t count, bool hasnan, bool desc,
                hipStream_t s)
{
    switch (kt) {
    case 0: return launch_small_group<TypeKey<uint16_t>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    case 1: return launch_small_group<TypeKey<int16_t>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    case 2: return launch_small_group<TypeKey<f16_tag>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    case 3: return launch_small_group<TypeKey<uint32_t>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    case 4: return launch_small_group<TypeKey<int32_t>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    case 5: return launch_small_group<TypeKey<float>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    case 6: return launch_small_group<TypeKey<uint64_t>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    case 7: return launch_small_group<TypeKey<int64_t>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    default: return launch_small_group<TypeKey<double>::KD>(d_ptrs, d_len, count, hasnan, desc, s);
    }
}

// the largest request one LDS workgroup sorts, per key type (keys only)
uint32_t small_limit(int kt) { return tag_bytes(kt) == 8 ? 2048 : 4096; }

}  // namespace

constexpr uint32_t MAX_GROUP = 1024;  // requests per coalesced launch

struct xss_server {
    size_t budget = 0, used = 0;
    uint32_t coalesce_max = 0;
    bool one_at_a_time = false;
    int device = 0;
    std::vector<hipStream_t> streams;
    std::vector<bool> busy;
    std::vector<uint64_t *> h_ptrs;   // pinned staging of the pointer lists, per stream
    std::vector<uint32_t *> h_len;
    std::vector<uint64_t *> d_ptrs;   // device copies, per stream
    std::vector<uint32_t *> d_len;
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
};

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
        if (coalesce_max && fr.op == OP_QSORT && fr.n <= std::min<uint64_t>(coalesce_max, small_limit(fr.key_type))) {
            // gather queued short qsorts of the same key type and flags
            uint32_t cnt = 0;
            for (auto it = queue.begin(); it != queue.end() && cnt < MAX_GROUP;) {
                const xss_request &q = *(*it)->r;
                const bool match = q.op == OP_QSORT && q.key_type == fr.key_type && q.hasnan == fr.hasnan &&
                                   q.descending == fr.descending &&
                                   q.n <= std::min<uint64_t>(coalesce_max, small_limit(q.key_type));
                if (match && used + g->footprint + (*it)->footprint <= budget) {
                    h_ptrs[s][cnt] = reinterpret_cast<uint64_t>(q.keys);
                    h_len[s][cnt] = uint32_t(q.n);
                    g->reqs.push_back(*it);
                    g->footprint += (*it)->footprint;
                    ++cnt;
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
            if (cnt == 0) {  // the head alone exceeds the budget: run it by itself
                queue.erase(std::find(queue.begin(), queue.end(), front));
                h_ptrs[s][0] = reinterpret_cast<uint64_t>(fr.keys);
                h_len[s][0] = uint32_t(fr.n);
                g->reqs.push_back(front);
                g->footprint = front->footprint;
                cnt = 1;
            }
            hipError_t e = hipMemcpyAsync(d_ptrs[s], h_ptrs[s], cnt * 8, hipMemcpyHostToDevice, streams[s]);
            if (e == hipSuccess) e = hipMemcpyAsync(d_len[s], h_len[s], cnt * 4, hipMemcpyHostToDevice, streams[s]);
            if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync");
            if (!rc) rc = small_group(fr.key_type, d_ptrs[s], d_len[s], cnt, fr.hasnan, fr.descending, streams[s]);
            coalesced += cnt > 1 ? cnt : 0;
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
        uint64_t *hp;
        uint32_t *hl;
        uint64_t *dp;
        uint32_t *dl;
        XSS_HIP_CHECK(hipHostMalloc(&hp, MAX_GROUP * 8));
        XSS_HIP_CHECK(hipHostMalloc(&hl, MAX_GROUP * 4));
        XSS_HIP_CHECK(hipMalloc(&dp, MAX_GROUP * 8));
        XSS_HIP_CHECK(hipMalloc(&dl, MAX_GROUP * 4));
        sv->h_ptrs.push_back(hp);
        sv->h_len.push_back(hl);
        sv->d_ptrs.push_back(dp);
        sv->d_len.push_back(dl);
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
        (void)hipHostFree(sv->h_ptrs[i]);
        (void)hipHostFree(sv->h_len[i]);
        (void)hipFree(sv->d_ptrs[i]);
        (void)hipFree(sv->d_len[i]);
        (void)hipStreamDestroy(sv->streams[i]);
    }
    delete sv;
    return XSS_OK;
}

}  // extern "C"
