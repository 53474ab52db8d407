// xss_runtime.cpp -- workspaces, errors, library-level C entry points.
#include "xss_runtime.hpp"

#include <stdarg.h>
#include <stdio.h>

#include <map>
#include <memory>
#include <string>
#include <utility>

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

namespace {
std::mutex g_map_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Workspace>> g_ws;

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
}  // namespace

int acquire_workspace(hipStream_t stream, size_t status_words, size_t data_bytes, Workspace **out,
                      std::unique_lock<std::mutex> *lock)
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
    const size_t meta_need = META_FIXED + status_words * sizeof(uint64_t);
    const bool grow_meta = meta_need > ws->meta_cap;
    const bool grow_data = data_bytes > ws->data_cap;
    if (grow_meta || grow_data) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        XSS_HIP_CHECK(hipStreamIsCapturing(stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return set_error(XSS_ERR_CAPTURE,
                             "workspace must grow (meta %zu, data %zu bytes) during stream capture; "
                             "run the call once before capturing",
                             meta_need, data_bytes);
        // earlier work on this stream may still use the old buffers
        XSS_HIP_CHECK(hipStreamSynchronize(stream));
        if (grow_meta) {
            if (ws->meta) XSS_HIP_CHECK(hipFree(ws->meta));
            ws->meta = nullptr;
            ws->meta_cap = 0;
            size_t cap = round_up(META_FIXED + (status_words + status_words / 4) * sizeof(uint64_t), 1 << 16);
            XSS_HIP_CHECK(hipMalloc(&ws->meta, cap));
            ws->meta_cap = cap;
            // epoch 0 and all-zero look-back words: no stale word matches a live epoch
            XSS_HIP_CHECK(hipMemsetAsync(ws->meta, 0, cap, stream));
        }
        if (grow_data) {
            if (ws->data) XSS_HIP_CHECK(hipFree(ws->data));
            ws->data = nullptr;
            ws->data_cap = 0;
            size_t cap = round_up(data_bytes + data_bytes / 4, 1 << 20);
            XSS_HIP_CHECK(hipMalloc(&ws->data, cap));
            ws->data_cap = cap;
        }
    } else if (!ws->meta) {
        size_t cap = round_up(META_FIXED + 4096 * sizeof(uint64_t), 1 << 16);
        XSS_HIP_CHECK(hipMalloc(&ws->meta, cap));
        ws->meta_cap = cap;
        XSS_HIP_CHECK(hipMemsetAsync(ws->meta, 0, cap, stream));
    }
    *out = ws;
    *lock = std::move(l);
    return XSS_OK;
}

}  // namespace xss

extern "C" {

const char *xss_gpu_last_error(void) { return xss::g_err.c_str(); }

const char *xss_gpu_version(void) { return "xss-amd 0.1 (gfx950 onesweep radix)"; }

int xss_gpu_release_workspaces(void)
{
    std::lock_guard<std::mutex> g(xss::g_map_mu);
    for (auto &kv : xss::g_ws) {
        xss::Workspace *ws = kv.second.get();
        std::lock_guard<std::mutex> l(ws->mu);
        int cur = 0;
        XSS_HIP_CHECK(hipGetDevice(&cur));
        XSS_HIP_CHECK(hipSetDevice(ws->device));
        XSS_HIP_CHECK(hipStreamSynchronize(ws->stream));
        if (ws->meta) XSS_HIP_CHECK(hipFree(ws->meta));
        if (ws->data) XSS_HIP_CHECK(hipFree(ws->data));
        ws->meta = ws->data = nullptr;
        ws->meta_cap = ws->data_cap = 0;
        XSS_HIP_CHECK(hipSetDevice(cur));
    }
    xss::g_ws.clear();
    return XSS_OK;
}

size_t xss_gpu_workspace_bytes(void)
{
    std::lock_guard<std::mutex> g(xss::g_map_mu);
    size_t s = 0;
    for (auto &kv : xss::g_ws) s += kv.second->meta_cap + kv.second->data_cap;
    return s;
}

}  // extern "C"
