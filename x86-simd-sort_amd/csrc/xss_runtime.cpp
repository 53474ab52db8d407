// xss_runtime.cpp -- workspaces, errors, library-level C entry points.
#include "xss_runtime.hpp"

#include <stdarg.h>
#include <stdio.h>

#include <atomic>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/xss_gpu.h"

namespace xss {

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return set_error(e == hipErrorOutOfMemory ? XSS_ERR_NOMEM : XSS_ERR_HIP, "%s failed: %s", what,
                     hipGetErrorString(e));
}

std::atomic<uint64_t> &workspace_limit()
{
    static std::atomic<uint64_t> v{0};
    return v;
}

namespace {
std::mutex g_map_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Workspace>> g_ws;

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
}  // namespace

int acquire_workspace(hipStream_t stream, size_t data_bytes, Workspace **out, std::unique_lock<std::mutex> *lock)
{
    int dev = 0;
    XSS_HIP_CHECK(hipGetDevice(&dev));
    Workspace *ws;
    {
        std::lock_guard<std::mutex> g(g_map_mu);
        auto &slot = g_ws[{dev, stream}];
        if (!slot) {
            slot.reset(new Workspace());
            slot->device = dev;
            slot->stream = stream;
        }
        ws = slot.get();
    }
    std::unique_lock<std::mutex> l(ws->mu);
    const bool need_meta = ws->meta == nullptr;
    const bool grow_data = data_bytes > ws->data_cap;
    if (need_meta || grow_data) {
        const uint64_t lim = workspace_limit().load();
        if (lim && META_BYTES + data_bytes > lim)
            return set_error(XSS_ERR_NOMEM, "workspace of %zu bytes exceeds the limit of %llu bytes",
                             META_BYTES + data_bytes, (unsigned long long)lim);
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        XSS_HIP_CHECK(hipStreamIsCapturing(stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return set_error(XSS_ERR_CAPTURE,
                             "workspace must grow to %zu scratch bytes during stream capture; "
                             "run the call once before capturing",
                             data_bytes);
        if (need_meta) XSS_HIP_CHECK(hipMalloc(&ws->meta, META_BYTES));
        if (grow_data) {
            // stream-ordered: the old buffer is released once earlier work on
            // this stream is done, with no host wait
            if (ws->data) XSS_HIP_CHECK(hipFreeAsync(ws->data, stream));
            ws->data = nullptr;
            ws->data_cap = 0;
            size_t cap = round_up(data_bytes + data_bytes / 4, 1 << 20);
            XSS_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&ws->data), cap, stream));
            ws->data_cap = cap;
        }
    }
    *out = ws;
    *lock = std::move(l);
    return XSS_OK;
}

int acquire_aux(hipStream_t stream, size_t bytes, char **out, std::unique_lock<std::mutex> *lock)
{
    int dev = 0;
    XSS_HIP_CHECK(hipGetDevice(&dev));
    Workspace *ws;
    {
        std::lock_guard<std::mutex> g(g_map_mu);
        auto &slot = g_ws[{dev, stream}];
        if (!slot) {
            slot.reset(new Workspace());
            slot->device = dev;
            slot->stream = stream;
        }
        ws = slot.get();
    }
    std::unique_lock<std::mutex> l(ws->aux_mu);
    if (bytes > ws->aux_cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        XSS_HIP_CHECK(hipStreamIsCapturing(stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return set_error(XSS_ERR_CAPTURE,
                             "staging buffer must grow to %zu bytes during stream capture; "
                             "run the call once before capturing",
                             bytes);
        if (ws->aux) XSS_HIP_CHECK(hipFreeAsync(ws->aux, stream));
        ws->aux = nullptr;
        ws->aux_cap = 0;
        const size_t cap = round_up(bytes + bytes / 4, 1 << 20);
        XSS_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&ws->aux), cap, stream));
        ws->aux_cap = cap;
    }
    *out = ws->aux;
    *lock = std::move(l);
    return XSS_OK;
}

namespace {
struct TimedLaunch {
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_t_mu;
bool g_timing = false;
std::vector<TimedLaunch> g_launches;
std::vector<hipEvent_t> g_event_pool;
hipEvent_t pooled_event()
{
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}
}  // namespace

thread_local int g_timing_quiet = 0;
bool timing_on() { return g_timing && g_timing_quiet == 0; }
TimingSuppressed::TimingSuppressed() { ++g_timing_quiet; }
TimingSuppressed::~TimingSuppressed() { --g_timing_quiet; }

void timing_begin(const char *name, hipStream_t s)
{
    std::lock_guard<std::mutex> g(g_t_mu);
    TimedLaunch t{name, pooled_event(), pooled_event()};
    (void)hipEventRecord(t.a, s);
    g_launches.push_back(t);
}

void timing_end(hipStream_t s)
{
    std::lock_guard<std::mutex> g(g_t_mu);
    if (!g_launches.empty()) (void)hipEventRecord(g_launches.back().b, s);
}

}  // namespace xss

extern "C" {

int xss_gpu_timing_enable(int on)
{
    std::lock_guard<std::mutex> g(xss::g_t_mu);
    xss::g_timing = on != 0;
    return XSS_OK;
}

int xss_gpu_timing_collect(int max_names, const char **names, double *total_ms, uint64_t *counts)
{
    std::lock_guard<std::mutex> g(xss::g_t_mu);
    static std::vector<std::string> keep;
    std::map<std::string, std::pair<double, uint64_t>> agg;
    for (auto &t : xss::g_launches) {
        XSS_HIP_CHECK(hipEventSynchronize(t.b));
        float ms = 0.f;
        XSS_HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
        auto &slot = agg[t.name];
        slot.first += ms;
        slot.second += 1;
        xss::g_event_pool.push_back(t.a);
        xss::g_event_pool.push_back(t.b);
    }
    xss::g_launches.clear();
    keep.clear();
    int i = 0;
    for (auto &kv : agg) {
        if (i >= max_names) break;
        keep.push_back(kv.first);
        names[i] = nullptr;
        total_ms[i] = kv.second.first;
        counts[i] = kv.second.second;
        ++i;
    }
    for (int j = 0; j < i; ++j) names[j] = keep[j].c_str();
    return i;
}

const char *xss_gpu_last_error(void) { return xss::g_err.c_str(); }

const char *xss_gpu_version(void) { return "xss-amd 0.1 (gfx950 onesweep radix)"; }

int xss_gpu_release_workspaces(void)
{
    std::lock_guard<std::mutex> g(xss::g_map_mu);
    for (auto &kv : xss::g_ws) {
        xss::Workspace *ws = kv.second.get();
        std::lock_guard<std::mutex> la(ws->aux_mu);
        std::lock_guard<std::mutex> l(ws->mu);
        int cur = 0;
        XSS_HIP_CHECK(hipGetDevice(&cur));
        XSS_HIP_CHECK(hipSetDevice(ws->device));
        if (ws->data) XSS_HIP_CHECK(hipFreeAsync(ws->data, ws->stream));
        if (ws->aux) XSS_HIP_CHECK(hipFreeAsync(ws->aux, ws->stream));
        XSS_HIP_CHECK(hipStreamSynchronize(ws->stream));
        if (ws->meta) XSS_HIP_CHECK(hipFree(ws->meta));
        ws->meta = ws->data = ws->aux = nullptr;
        ws->data_cap = ws->aux_cap = 0;
        XSS_HIP_CHECK(hipSetDevice(cur));
    }
    xss::g_ws.clear();
    return XSS_OK;
}

size_t xss_gpu_workspace_bytes(void)
{
    std::lock_guard<std::mutex> g(xss::g_map_mu);
    size_t s = 0;
    for (auto &kv : xss::g_ws)
        s += (kv.second->meta ? xss::META_BYTES : 0) + kv.second->data_cap + kv.second->aux_cap;
    return s;
}

}  // extern "C"
