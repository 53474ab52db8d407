// xss_api.cpp -- the host-pointer C++ API of include/x86simdsort.h.
//
// Explicit specialisations of the x86simdsort:: templates for every type the
// reference exports (lib/x86simdsort.cpp:60-109 for the single-array
// functions, :207-293 for the 36 key-value pairs), with its signatures, so
// the symbols are the reference's.  Each call stages the array(s) into a
// per-thread device buffer on a per-thread stream, runs the C-ABI device
// entry point and copies the result back before returning.  GPU failures
// throw std::runtime_error.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/x86simdsort.h"

namespace {

struct HostCtx {
    hipStream_t s = nullptr;
    char *buf = nullptr;
    size_t cap = 0;
    ~HostCtx()
    {
        // process teardown: release quietly
        if (buf) (void)hipFree(buf);
        if (s) (void)hipStreamDestroy(s);
    }
};
thread_local HostCtx g_ctx;

[[noreturn]] void fail(const std::string &what) { throw std::runtime_error("x86simdsort (MI355X): " + what); }

void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

void xss_ok(int rc, const char *what)
{
    if (rc != XSS_OK) fail(std::string(what) + ": " + xss_gpu_last_error());
}

hipStream_t stream()
{
    if (!g_ctx.s) hip_ok(hipStreamCreateWithFlags(&g_ctx.s, hipStreamNonBlocking), "hipStreamCreate");
    return g_ctx.s;
}

char *device_buffer(size_t bytes)
{
    if (bytes > g_ctx.cap) {
        if (g_ctx.buf) {
            hip_ok(hipStreamSynchronize(stream()), "hipStreamSynchronize");
            hip_ok(hipFree(g_ctx.buf), "hipFree");
            g_ctx.buf = nullptr;
            g_ctx.cap = 0;
        }
        const size_t cap = (bytes + bytes / 4 + 4095) & ~size_t(4095);
        hip_ok(hipMalloc(&g_ctx.buf, cap), "hipMalloc");
        g_ctx.cap = cap;
    }
    return g_ctx.buf;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

void h2d(void *d, const void *h, size_t bytes)
{
    hip_ok(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream()), "hipMemcpyAsync H2D");
}
void d2h(void *h, const void *d, size_t bytes)
{
    hip_ok(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, stream()), "hipMemcpyAsync D2H");
}
void sync() { hip_ok(hipStreamSynchronize(stream()), "hipStreamSynchronize"); }

}  // namespace

namespace xss {
int set_error(int code, const char *fmt, ...);
}

// Host-pointer NumPy-style arg forms: `arr` and `arg` staged to the device,
// the xss_gpu_ form run, `arg` copied back.  Status codes as the C ABI.
#define XSS_HOST_ARGFORM(tag, T)                                                                        \
    extern "C" int xss_host_argsort_arg_##tag(const T *arr, size_t *arg, size_t n, int hasnan,          \
                                              int descending)                                           \
    {                                                                                                   \
        try {                                                                                           \
            if (n <= 1) return XSS_OK;                                                                  \
            if (!arr || !arg) return xss::set_error(XSS_ERR_INVALID, "null pointer argument");         \
            const size_t b = n * sizeof(T);                                                             \
            char *d = device_buffer(align256(b) + n * 8);                                               \
            uint64_t *di = reinterpret_cast<uint64_t *>(d + align256(b));                               \
            h2d(d, arr, b);                                                                             \
            h2d(di, arg, n * 8);                                                                        \
            const int rc = xss_gpu_argsort_arg_##tag(reinterpret_cast<const T *>(d), di, n, hasnan,     \
                                                     descending, stream());                             \
            if (rc != XSS_OK) return rc;                                                                \
            d2h(arg, di, n * 8);                                                                        \
            sync();                                                                                     \
            return XSS_OK;                                                                              \
        } catch (const std::exception &e) {                                                             \
            return xss::set_error(XSS_ERR_HIP, "%s", e.what());                                         \
        }                                                                                               \
    }                                                                                                   \
    extern "C" int xss_host_argselect_arg_##tag(const T *arr, size_t *arg, size_t k, size_t n, int hasnan) \
    {                                                                                                   \
        try {                                                                                           \
            if (n <= 1) return XSS_OK;                                                                  \
            if (!arr || !arg) return xss::set_error(XSS_ERR_INVALID, "null pointer argument");         \
            const size_t b = n * sizeof(T);                                                             \
            char *d = device_buffer(align256(b) + n * 8);                                               \
            uint64_t *di = reinterpret_cast<uint64_t *>(d + align256(b));                               \
            h2d(d, arr, b);                                                                             \
            h2d(di, arg, n * 8);                                                                        \
            const int rc = xss_gpu_argselect_arg_##tag(reinterpret_cast<const T *>(d), di, k, n, hasnan, \
                                                       stream());                                       \
            if (rc != XSS_OK) return rc;                                                                \
            d2h(arg, di, n * 8);                                                                        \
            sync();                                                                                     \
            return XSS_OK;                                                                              \
        } catch (const std::exception &e) {                                                             \
            return xss::set_error(XSS_ERR_HIP, "%s", e.what());                                         \
        }                                                                                               \
    }

XSS_HOST_ARGFORM(u16, uint16_t)
XSS_HOST_ARGFORM(i16, int16_t)
XSS_HOST_ARGFORM(f16, xss_f16_bits)
XSS_HOST_ARGFORM(u32, uint32_t)
XSS_HOST_ARGFORM(i32, int32_t)
XSS_HOST_ARGFORM(f32, float)
XSS_HOST_ARGFORM(u64, uint64_t)
XSS_HOST_ARGFORM(i64, int64_t)
XSS_HOST_ARGFORM(f64, double)

namespace x86simdsort {

#define XSS_HOST_T(tag, T, CT)                                                                          \
    template <>                                                                                         \
    void qsort<T>(T * arr, size_t n, bool hasnan, bool descending)                                      \
    {                                                                                                   \
        if (n <= 1) return;                                                                             \
        const size_t b = n * sizeof(T);                                                                 \
        char *d = device_buffer(b);                                                                     \
        h2d(d, arr, b);                                                                                 \
        xss_ok(xss_gpu_qsort_##tag(reinterpret_cast<CT *>(d), n, hasnan, descending, stream()), "qsort"); \
        d2h(arr, d, b);                                                                                 \
        sync();                                                                                         \
    }                                                                                                   \
    template <>                                                                                         \
    void qselect<T>(T * arr, size_t k, size_t n, bool hasnan, bool descending)                          \
    {                                                                                                   \
        if (n <= 1) return;                                                                             \
        const size_t b = n * sizeof(T);                                                                 \
        char *d = device_buffer(b);                                                                     \
        h2d(d, arr, b);                                                                                 \
        xss_ok(xss_gpu_qselect_##tag(reinterpret_cast<CT *>(d), k, n, hasnan, descending, stream()),    \
               "qselect");                                                                              \
        d2h(arr, d, b);                                                                                 \
        sync();                                                                                         \
    }                                                                                                   \
    template <>                                                                                         \
    void partial_qsort<T>(T * arr, size_t k, size_t n, bool hasnan, bool descending)                    \
    {                                                                                                   \
        if (n <= 1 || k == 0) return;                                                                   \
        const size_t b = n * sizeof(T);                                                                 \
        char *d = device_buffer(b);                                                                     \
        h2d(d, arr, b);                                                                                 \
        xss_ok(xss_gpu_partial_qsort_##tag(reinterpret_cast<CT *>(d), k, n, hasnan, descending, stream()), \
               "partial_qsort");                                                                        \
        d2h(arr, d, b);                                                                                 \
        sync();                                                                                         \
    }                                                                                                   \
    template <>                                                                                         \
    std::vector<size_t> argsort<T>(T * arr, size_t n, bool hasnan, bool descending)                     \
    {                                                                                                   \
        std::vector<size_t> out(n);                                                                     \
        if (n <= 1) {                                                                                   \
            std::iota(out.begin(), out.end(), size_t(0));                                               \
            return out;                                                                                 \
        }                                                                                               \
        const size_t b = n * sizeof(T);                                                                 \
        char *d = device_buffer(align256(b) + n * 8);                                                   \
        uint64_t *di = reinterpret_cast<uint64_t *>(d + align256(b));                                   \
        h2d(d, arr, b);                                                                                 \
        xss_ok(xss_gpu_argsort_##tag(reinterpret_cast<const CT *>(d), di, n, hasnan, descending, stream()), \
               "argsort");                                                                              \
        d2h(out.data(), di, n * 8);                                                                     \
        sync();                                                                                         \
        return out;                                                                                     \
    }                                                                                                   \
    template <>                                                                                         \
    std::vector<size_t> argselect<T>(T * arr, size_t k, size_t n, bool hasnan)                          \
    {                                                                                                   \
        std::vector<size_t> out(n);                                                                     \
        if (n <= 1) {                                                                                   \
            std::iota(out.begin(), out.end(), size_t(0));                                               \
            return out;                                                                                 \
        }                                                                                               \
        const size_t b = n * sizeof(T);                                                                 \
        char *d = device_buffer(align256(b) + n * 8);                                                   \
        uint64_t *di = reinterpret_cast<uint64_t *>(d + align256(b));                                   \
        h2d(d, arr, b);                                                                                 \
        xss_ok(xss_gpu_argselect_##tag(reinterpret_cast<const CT *>(d), di, k, n, hasnan, stream()),    \
               "argselect");                                                                            \
        d2h(out.data(), di, n * 8);                                                                     \
        sync();                                                                                         \
        return out;                                                                                     \
    }

XSS_HOST_T(u16, uint16_t, uint16_t)
XSS_HOST_T(i16, int16_t, int16_t)
XSS_HOST_T(f16, _Float16, xss_f16_bits)
XSS_HOST_T(u32, uint32_t, uint32_t)
XSS_HOST_T(i32, int32_t, int32_t)
XSS_HOST_T(f32, float, float)
XSS_HOST_T(u64, uint64_t, uint64_t)
XSS_HOST_T(i64, int64_t, int64_t)
XSS_HOST_T(f64, double, double)

#define XSS_HOST_KV(ktag, K, CK, vtag, V)                                                                \
    template <>                                                                                         \
    void keyvalue_qsort<K, V>(K * key, V * val, size_t n, bool hasnan, bool descending)                 \
    {                                                                                                   \
        if (n <= 1) return;                                                                             \
        const size_t kb = n * sizeof(K), vb = n * sizeof(V);                                            \
        char *d = device_buffer(align256(kb) + vb);                                                     \
        char *dv = d + align256(kb);                                                                    \
        h2d(d, key, kb);                                                                                \
        h2d(dv, val, vb);                                                                               \
        xss_ok(xss_gpu_keyvalue_qsort_##ktag##_##vtag(reinterpret_cast<CK *>(d), reinterpret_cast<V *>(dv), n, \
                                                      hasnan, descending, stream()),                    \
               "keyvalue_qsort");                                                                       \
        d2h(key, d, kb);                                                                                \
        d2h(val, dv, vb);                                                                               \
        sync();                                                                                         \
    }                                                                                                   \
    template <>                                                                                         \
    void keyvalue_select<K, V>(K * key, V * val, size_t k, size_t n, bool hasnan, bool descending)      \
    {                                                                                                   \
        if (n <= 1) return;                                                                             \
        const size_t kb = n * sizeof(K), vb = n * sizeof(V);                                            \
        char *d = device_buffer(align256(kb) + vb);                                                     \
        char *dv = d + align256(kb);                                                                    \
        h2d(d, key, kb);                                                                                \
        h2d(dv, val, vb);                                                                               \
        xss_ok(xss_gpu_keyvalue_select_##ktag##_##vtag(reinterpret_cast<CK *>(d), reinterpret_cast<V *>(dv), k, \
                                                       n, hasnan, descending, stream()),                \
               "keyvalue_select");                                                                      \
        d2h(key, d, kb);                                                                                \
        d2h(val, dv, vb);                                                                               \
        sync();                                                                                         \
    }                                                                                                   \
    template <>                                                                                         \
    void keyvalue_partial_sort<K, V>(K * key, V * val, size_t k, size_t n, bool hasnan, bool descending) \
    {                                                                                                   \
        if (n <= 1 || k == 0) return;                                                                   \
        const size_t kb = n * sizeof(K), vb = n * sizeof(V);                                            \
        char *d = device_buffer(align256(kb) + vb);                                                     \
        char *dv = d + align256(kb);                                                                    \
        h2d(d, key, kb);                                                                                \
        h2d(dv, val, vb);                                                                               \
        xss_ok(xss_gpu_keyvalue_partial_sort_##ktag##_##vtag(reinterpret_cast<CK *>(d),                  \
                                                             reinterpret_cast<V *>(dv), k, n, hasnan,   \
                                                             descending, stream()),                     \
               "keyvalue_partial_sort");                                                                \
        d2h(key, d, kb);                                                                                \
        d2h(val, dv, vb);                                                                               \
        sync();                                                                                         \
    }

#define XSS_HOST_KV_ALLV(ktag, K, CK)       \
    XSS_HOST_KV(ktag, K, CK, u32, uint32_t) \
    XSS_HOST_KV(ktag, K, CK, i32, int32_t)  \
    XSS_HOST_KV(ktag, K, CK, f32, float)    \
    XSS_HOST_KV(ktag, K, CK, u64, uint64_t) \
    XSS_HOST_KV(ktag, K, CK, i64, int64_t)  \
    XSS_HOST_KV(ktag, K, CK, f64, double)

XSS_HOST_KV_ALLV(u16, uint16_t, uint16_t)
XSS_HOST_KV_ALLV(i16, int16_t, int16_t)
XSS_HOST_KV_ALLV(f16, _Float16, xss_f16_bits)
XSS_HOST_KV_ALLV(u32, uint32_t, uint32_t)
XSS_HOST_KV_ALLV(i32, int32_t, int32_t)
XSS_HOST_KV_ALLV(f32, float, float)
XSS_HOST_KV_ALLV(u64, uint64_t, uint64_t)
XSS_HOST_KV_ALLV(i64, int64_t, int64_t)
XSS_HOST_KV_ALLV(f64, double, double)

namespace detail {
#define XSS_HOST_OBJ(tag, K)                                                                            \
    template <>                                                                                         \
    void object_sort_by_keys<K>(void *objs, size_t obj_bytes, const K *keys, size_t n)                  \
    {                                                                                                   \
        if (n <= 1) return;                                                                             \
        const size_t ob = n * obj_bytes, kb = n * sizeof(K);                                            \
        char *d = device_buffer(align256(ob) + kb);                                                     \
        char *dk = d + align256(ob);                                                                    \
        h2d(d, objs, ob);                                                                               \
        h2d(dk, keys, kb);                                                                              \
        xss_ok(xss_gpu_object_sort_by_keys_##tag(d, obj_bytes, reinterpret_cast<const K *>(dk), n, stream()), \
               "object_qsort");                                                                         \
        d2h(objs, d, ob);                                                                               \
        sync();                                                                                         \
    }
XSS_HOST_OBJ(f64, double)
XSS_HOST_OBJ(f32, float)
XSS_HOST_OBJ(u64, uint64_t)
XSS_HOST_OBJ(i64, int64_t)
XSS_HOST_OBJ(u32, uint32_t)
XSS_HOST_OBJ(i32, int32_t)
}  // namespace detail

}  // namespace x86simdsort

// The reference's C symbols (lib/x86simdsort.cpp:298-339): ascending,
// hasnan = true.  A C caller cannot catch exceptions, so a GPU failure is
// reported and the process stops instead of returning unsorted data.
#define XSS_C_KV(name, K, V)                                                      \
    extern "C" void name(K *key, V *val, size_t size)                             \
    {                                                                             \
        try {                                                                     \
            x86simdsort::keyvalue_qsort(key, val, size, true);                    \
        } catch (const std::exception &e) {                                       \
            std::fprintf(stderr, "%s: %s\n", #name, e.what());                    \
            std::abort();                                                         \
        }                                                                         \
    }
XSS_C_KV(keyvalue_qsort_float_uint32, float, uint32_t)
XSS_C_KV(keyvalue_qsort_float_uint64, float, uint64_t)
XSS_C_KV(keyvalue_qsort_uint64_uint32, uint64_t, uint32_t)
XSS_C_KV(keyvalue_qsort_uint64_uint64, uint64_t, uint64_t)
XSS_C_KV(keyvalue_qsort_int32_uint32, int32_t, uint32_t)
XSS_C_KV(keyvalue_qsort_int32_uint64, int32_t, uint64_t)
XSS_C_KV(keyvalue_qsort_uint32_uint32, uint32_t, uint32_t)
XSS_C_KV(keyvalue_qsort_uint32_uint64, uint32_t, uint64_t)
