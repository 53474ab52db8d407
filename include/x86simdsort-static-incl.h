/*
 * x86simdsort-static-incl.h -- the x86simdsortStatic:: API of the reference
 * (src/x86simdsort-static-incl.h:7-71, the header NumPy includes) on the
 * MI355X library, with the same names, argument order and defaults.
 *
 * The reference defines these header-only over its SIMD kernels.  Here each
 * one forwards to the library: the value forms to the x86simdsort::
 * templates of x86simdsort.h (host memory, staged through the GPU), the
 * NumPy arg forms argsort(T *, size_t *arg, ...) and argselect(T *, size_t
 * *arg, k, ...) to xss_host_argsort_arg_<T> / xss_host_argselect_arg_<T> of
 * xss_gpu.h.  `arg` holds indices into `arr` on entry (NumPy passes
 * 0..size-1) and is reordered in place; `arr` is left unchanged.  GPU
 * failures throw std::runtime_error (the reference cannot fail).
 *
 * Link with -lxss_amd, as for x86simdsort.h.
 */
#ifndef XSS_AMD_X86SIMDSORT_STATIC_INCL_H
#define XSS_AMD_X86SIMDSORT_STATIC_INCL_H

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "x86simdsort.h"

namespace x86simdsortStatic {

namespace detail {

template <typename T>
struct argform;

#define XSS_STATIC_ARGFORM(T, tag, CT)                                                                 \
    template <>                                                                                        \
    struct argform<T> {                                                                                \
        static int sort(T *arr, size_t *arg, size_t n, bool hasnan, bool desc)                        \
        {                                                                                              \
            return xss_host_argsort_arg_##tag(reinterpret_cast<const CT *>(arr), arg, n, hasnan, desc); \
        }                                                                                              \
        static int select(T *arr, size_t *arg, size_t k, size_t n, bool hasnan)                        \
        {                                                                                              \
            return xss_host_argselect_arg_##tag(reinterpret_cast<const CT *>(arr), arg, k, n, hasnan); \
        }                                                                                              \
    };

XSS_STATIC_ARGFORM(uint16_t, u16, uint16_t)
XSS_STATIC_ARGFORM(int16_t, i16, int16_t)
#if defined(__FLT16_MAX__)
XSS_STATIC_ARGFORM(_Float16, f16, xss_f16_bits)
#endif
XSS_STATIC_ARGFORM(uint32_t, u32, uint32_t)
XSS_STATIC_ARGFORM(int32_t, i32, int32_t)
XSS_STATIC_ARGFORM(float, f32, float)
XSS_STATIC_ARGFORM(uint64_t, u64, uint64_t)
XSS_STATIC_ARGFORM(int64_t, i64, int64_t)
XSS_STATIC_ARGFORM(double, f64, double)
#undef XSS_STATIC_ARGFORM

inline void check(int rc, const char *what)
{
    if (rc != XSS_OK) throw std::runtime_error(std::string("x86simdsortStatic::") + what + " (MI355X): " +
                                               xss_gpu_last_error());
}

}  // namespace detail

template <typename T>
inline void qsort(T *arr, size_t size, bool hasnan = false, bool descending = false)
{
    x86simdsort::qsort(arr, size, hasnan, descending);
}

template <typename T>
inline void qselect(T *arr, size_t k, size_t size, bool hasnan = false, bool descending = false)
{
    x86simdsort::qselect(arr, k, size, hasnan, descending);
}

template <typename T>
inline void partial_qsort(T *arr, size_t k, size_t size, bool hasnan = false, bool descending = false)
{
    x86simdsort::partial_qsort(arr, k, size, hasnan, descending);
}

template <typename T>
inline std::vector<size_t> argsort(T *arr, size_t size, bool hasnan = false, bool descending = false)
{
    return x86simdsort::argsort(arr, size, hasnan, descending);
}

/* argsort API required by NumPy */
template <typename T>
inline void argsort(T *arr, size_t *arg, size_t size, bool hasnan = false, bool descending = false)
{
    detail::check(detail::argform<T>::sort(arr, arg, size, hasnan, descending), "argsort");
}

template <typename T>
inline std::vector<size_t> argselect(T *arr, size_t k, size_t size, bool hasnan = false)
{
    return x86simdsort::argselect(arr, k, size, hasnan);
}

/* argselect API required by NumPy */
template <typename T>
inline void argselect(T *arr, size_t *arg, size_t k, size_t size, bool hasnan = false)
{
    detail::check(detail::argform<T>::select(arr, arg, k, size, hasnan), "argselect");
}

template <typename T1, typename T2>
inline void keyvalue_qsort(T1 *key, T2 *val, size_t size, bool hasnan = false, bool descending = false)
{
    x86simdsort::keyvalue_qsort(key, val, size, hasnan, descending);
}

template <typename T1, typename T2>
inline void keyvalue_select(T1 *key, T2 *val, size_t k, size_t size, bool hasnan = false, bool descending = false)
{
    x86simdsort::keyvalue_select(key, val, k, size, hasnan, descending);
}

template <typename T1, typename T2>
inline void keyvalue_partial_sort(T1 *key, T2 *val, size_t k, size_t size, bool hasnan = false,
                                  bool descending = false)
{
    x86simdsort::keyvalue_partial_sort(key, val, k, size, hasnan, descending);
}

}  // namespace x86simdsortStatic

#endif /* XSS_AMD_X86SIMDSORT_STATIC_INCL_H */
