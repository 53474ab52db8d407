/*
 * x86simdsort-gpu.h -- header-only device forms of x86simdsort::object_qsort
 * (lib/x86simdsort.h:72-106) for objects that already live in HBM.
 *
 * The reference evaluates key_func on the host for every object, argsorts
 * the keys with keyvalue_qsort(keys, iota) and permutes the objects in place
 * along cycles.  Here the caller's translation unit (compiled by hipcc)
 * instantiates a kernel that evaluates the key functor on the device, one
 * object per lane, into a temporary key array; the library then argsorts
 * the keys (ascending, hasnan = false, as the reference) and gathers the
 * objects into that order in place (xss_gpu_object_sort_by_keys_<K>).
 *
 *     struct P { double x, y, z; };
 *     P *d_pts = ...;                                   // device memory
 *     int rc = x86simdsort::gpu::object_qsort(
 *         d_pts, n, [] __device__(const P &p) { return p.x * p.x + p.y * p.y; }, stream);
 *
 * key_fn is any callable usable in device code (a __device__ lambda or a
 * functor with a __device__ operator()); it takes the object (by value or
 * const reference) and returns a 32- or 64-bit arithmetic key (float,
 * double, or a signed / unsigned 32- or 64-bit integer), the key types the
 * reference accepts.  Returns an xss_status: the call is stream-ordered and
 * asynchronous (the key array is allocated and freed stream-ordered; no
 * host wait).  Objects must be trivially copyable.  Equal keys may leave
 * their objects in any order, as in the reference (its key-value sort is
 * not stable); this form keeps them in input order.
 *
 * No counterpart in the reference's headers: the device functor is this
 * build's addition for callers whose objects are on the GPU.
 */
#ifndef XSS_AMD_X86SIMDSORT_GPU_H
#define XSS_AMD_X86SIMDSORT_GPU_H

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <type_traits>
#include <utility>

#include "xss_gpu.h"

namespace x86simdsort {
namespace gpu {
namespace detail {

template <class T, class Fn, class K>
__global__ __launch_bounds__(256) void object_keys_kernel(const T *arr, uint64_t n, Fn fn, K *keys)
{
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) keys[i] = K(fn(arr[i]));
}

// the key functor's result type; the functor may be device-only, so the
// call is named in a host-device context, where its device overload is
// viable
template <class Fn, class T>
struct key_result {
    __host__ __device__ static auto probe(Fn f, const T &t) -> decltype(f(t));
    using type = typename std::decay<decltype(probe(std::declval<Fn>(), std::declval<const T &>()))>::type;
};

// the library's key type of a 32/64-bit arithmetic key
template <class R>
struct lib_key {
    static_assert(std::is_arithmetic<R>::value && (sizeof(R) == 4 || sizeof(R) == 8),
                  "key_fn must return a 32- or 64-bit arithmetic key");
    using type = typename std::conditional<
        std::is_floating_point<R>::value, R,
        typename std::conditional<
            sizeof(R) == 4,
            typename std::conditional<std::is_signed<R>::value, int32_t, uint32_t>::type,
            typename std::conditional<std::is_signed<R>::value, int64_t, uint64_t>::type>::type>::type;
};

inline int sort_by_keys(void *objs, size_t ob, const double *k, size_t n, hipStream_t s)
{
    return xss_gpu_object_sort_by_keys_f64(objs, ob, k, n, s);
}
inline int sort_by_keys(void *objs, size_t ob, const float *k, size_t n, hipStream_t s)
{
    return xss_gpu_object_sort_by_keys_f32(objs, ob, k, n, s);
}
inline int sort_by_keys(void *objs, size_t ob, const uint64_t *k, size_t n, hipStream_t s)
{
    return xss_gpu_object_sort_by_keys_u64(objs, ob, k, n, s);
}
inline int sort_by_keys(void *objs, size_t ob, const int64_t *k, size_t n, hipStream_t s)
{
    return xss_gpu_object_sort_by_keys_i64(objs, ob, k, n, s);
}
inline int sort_by_keys(void *objs, size_t ob, const uint32_t *k, size_t n, hipStream_t s)
{
    return xss_gpu_object_sort_by_keys_u32(objs, ob, k, n, s);
}
inline int sort_by_keys(void *objs, size_t ob, const int32_t *k, size_t n, hipStream_t s)
{
    return xss_gpu_object_sort_by_keys_i32(objs, ob, k, n, s);
}

}  // namespace detail

template <typename T, typename U, typename Fn>
int object_qsort(T *d_arr, U arrsize, Fn key_fn, hipStream_t stream = nullptr)
{
    static_assert(std::is_integral<U>::value, "arrsize must be an integral type");
    static_assert(sizeof(U) == sizeof(int32_t) || sizeof(U) == sizeof(int64_t), "arrsize must be 32 or 64 bits");
    static_assert(std::is_trivially_copyable<T>::value, "objects are moved as bytes on the GPU");
    using R = typename detail::key_result<Fn, T>::type;
    using K = typename detail::lib_key<R>::type;
    if constexpr (std::is_signed<U>::value)
        if (arrsize < 0) return XSS_ERR_INVALID;
    const size_t n = size_t(arrsize);
    if (n < 2) return XSS_OK;
    if (!d_arr) return XSS_ERR_INVALID;
    K *keys = nullptr;
    if (hipMallocAsync(reinterpret_cast<void **>(&keys), n * sizeof(K), stream) != hipSuccess)
        return XSS_ERR_NOMEM;
    const uint64_t blocks = (uint64_t(n) + 255) / 256;
    hipLaunchKernelGGL((detail::object_keys_kernel<T, Fn, K>), dim3(unsigned(blocks < 8192 ? blocks : 8192)),
                       dim3(256), 0, stream, d_arr, uint64_t(n), key_fn, keys);
    int rc = hipGetLastError() == hipSuccess ? int(XSS_OK) : int(XSS_ERR_HIP);
    if (rc == XSS_OK) rc = detail::sort_by_keys(static_cast<void *>(d_arr), sizeof(T), keys, n, stream);
    if (hipFreeAsync(keys, stream) != hipSuccess && rc == XSS_OK) rc = XSS_ERR_HIP;
    return rc;
}

}  // namespace gpu
}  // namespace x86simdsort

#endif
