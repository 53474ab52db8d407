// xss_argform.hip -- C-ABI entry points of the NumPy-style arg forms:
//   xss_gpu_argsort_arg_<T>    x86simdsortStatic::argsort(T *, size_t *arg, size, hasnan, descending)
//   xss_gpu_argselect_arg_<T>  x86simdsortStatic::argselect(T *, size_t *arg, k, size, hasnan)
//     (src/x86simdsort-static-incl.h:31-36,43-46,92-97,108-113)
//
// `arg` holds indices into `arr` on entry (NumPy passes 0..n-1) and is
// reordered in place so that arr[arg[i]] is sorted, or partitioned around
// rank k, with the reference's NaN and descending rules
// (src/xss-common-argsort.h:609-662,693-731): the keys arr[arg[i]] are
// gathered into a staging buffer, then that (key, index) array is sorted or
// selected in place as a key-value array with 64-bit payloads.  The keys in
// `arr` are not modified.
#include "xss_hybrid.hpp"
#include "xss_select.hpp"

using namespace xss;

namespace {

// out[i] = arr[arg[i]]; an index >= n (undefined input for the reference) is
// read as 0 so that no load leaves the array
template <class U>
__global__ __launch_bounds__(256) void argform_gather_kernel(const U *arr, const uint64_t *arg, U *out, uint64_t n)
{
    const uint64_t stride = uint64_t(gridDim.x) * 256u;
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) {
        const uint64_t j = arg[i];
        out[i] = arr[j < n ? j : 0];
    }
}

// Stage arr[arg[i]] in the stream's aux buffer; on success *keys points at it
// and `lock` holds the buffer until the caller's sort or select is enqueued.
template <class U>
int gather_keys(const void *arr, const uint64_t *arg, size_t n, hipStream_t s, U **keys,
                std::unique_lock<std::mutex> *lock)
{
    if (n > MAX_N) return set_error(XSS_ERR_INVALID, "n = %zu exceeds the per-call limit %u", n, MAX_N);
    if (!arr || !arg) return set_error(XSS_ERR_INVALID, "null pointer argument");
    char *buf = nullptr;
    int rc = acquire_aux(s, n * sizeof(U), &buf, lock);
    if (rc) return rc;
    U *g = reinterpret_cast<U *>(buf);
    const uint32_t blocks = uint32_t(std::min<size_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(argform_gather_kernel<U>, dim3(blocks), dim3(256), 0, s, static_cast<const U *>(arr), arg, g,
                       uint64_t(n));
    XSS_HIP_CHECK(hipGetLastError());
    *keys = g;
    return XSS_OK;
}

template <class CT>
int run_argsort_arg(const void *arr, uint64_t *arg, size_t n, bool hasnan, bool desc, hipStream_t s)
{
    using KD = typename TypeKey<CT>::KD;
    using U = typename KD::U;
    if (n <= 1) return XSS_OK;
    U *g = nullptr;
    std::unique_lock<std::mutex> lock;
    int rc = gather_keys<U>(arr, arg, n, s, &g, &lock);
    if (rc) return rc;
    return run_sort_auto<KD, PL_V64>(g, g, arg, arg, nullptr, n, hasnan, desc, false, s);
}

template <class CT>
int run_argselect_arg(const void *arr, uint64_t *arg, size_t k, size_t n, bool hasnan, hipStream_t s)
{
    using KD = typename TypeKey<CT>::KD;
    using U = typename KD::U;
    if (n <= 1) return XSS_OK;
    if (k >= n) return set_error(XSS_ERR_INVALID, "k = %zu out of range for n = %zu", k, n);
    U *g = nullptr;
    std::unique_lock<std::mutex> lock;
    int rc = gather_keys<U>(arr, arg, n, s, &g, &lock);
    if (rc) return rc;
    return run_select<KD, PL_V64>(0, g, arg, nullptr, n, k, hasnan, false, s);
}

}  // namespace

#define XSS_DEF_ARGFORM(tag, T, CT)                                                                  \
    extern "C" int xss_gpu_argsort_arg_##tag(const T *arr, uint64_t *arg, size_t n, int hasnan,      \
                                             int descending, xss_stream_t s)                         \
    {                                                                                                \
        return run_argsort_arg<CT>(arr, arg, n, hasnan != 0, descending != 0, s);                    \
    }                                                                                                \
    extern "C" int xss_gpu_argselect_arg_##tag(const T *arr, uint64_t *arg, size_t k, size_t n,      \
                                               int hasnan, xss_stream_t s)                           \
    {                                                                                                \
        return run_argselect_arg<CT>(arr, arg, k, n, hasnan != 0, s);                                \
    }

XSS_DEF_ARGFORM(u16, uint16_t, uint16_t)
XSS_DEF_ARGFORM(i16, int16_t, int16_t)
XSS_DEF_ARGFORM(f16, xss_f16_bits, f16_tag)
XSS_DEF_ARGFORM(u32, uint32_t, uint32_t)
XSS_DEF_ARGFORM(i32, int32_t, int32_t)
XSS_DEF_ARGFORM(f32, float, float)
XSS_DEF_ARGFORM(u64, uint64_t, uint64_t)
XSS_DEF_ARGFORM(i64, int64_t, int64_t)
XSS_DEF_ARGFORM(f64, double, double)
