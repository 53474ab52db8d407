// xss_select.hip -- C-ABI entry points of the select family:
//   xss_gpu_qselect_<T>, xss_gpu_partial_qsort_<T>, xss_gpu_argselect_<T>
//     (x86simdsort::qselect / partial_qsort / argselect, lib/x86simdsort.h:20-44)
//   xss_gpu_keyvalue_select_<K>_<V>, xss_gpu_keyvalue_partial_sort_<K>_<V>
//     (x86simdsort::keyvalue_select / keyvalue_partial_sort, lib/x86simdsort.h:54-70;
//     16-bit keys beyond the reference's 36 pairs as in keyvalue_qsort)
#include "xss_select.hpp"

using namespace xss;

#define XSS_DEF_SELECT(tag, T, CT)                                                                   \
    extern "C" int xss_gpu_qselect_##tag(T *arr, size_t k, size_t n, int hasnan, int descending,     \
                                         xss_stream_t s)                                             \
    {                                                                                                \
        return run_select<TypeKey<CT>::KD, PL_NONE>(0, arr, nullptr, nullptr, n, k, hasnan != 0,     \
                                                    descending != 0, s);                             \
    }                                                                                                \
    extern "C" int xss_gpu_partial_qsort_##tag(T *arr, size_t k, size_t n, int hasnan,               \
                                               int descending, xss_stream_t s)                       \
    {                                                                                                \
        return run_select<TypeKey<CT>::KD, PL_NONE>(1, arr, nullptr, nullptr, n, k, hasnan != 0,     \
                                                    descending != 0, s);                             \
    }                                                                                                \
    extern "C" int xss_gpu_argselect_##tag(const T *arr, uint64_t *arg, size_t k, size_t n,          \
                                           int hasnan, xss_stream_t s)                               \
    {                                                                                                \
        return run_select<TypeKey<CT>::KD, PL_IDX>(0, const_cast<T *>(arr), nullptr, arg, n, k,      \
                                                   hasnan != 0, false, s);                           \
    }

XSS_DEF_SELECT(u16, uint16_t, uint16_t)
XSS_DEF_SELECT(i16, int16_t, int16_t)
XSS_DEF_SELECT(f16, xss_f16_bits, f16_tag)
XSS_DEF_SELECT(u32, uint32_t, uint32_t)
XSS_DEF_SELECT(i32, int32_t, int32_t)
XSS_DEF_SELECT(f32, float, float)
XSS_DEF_SELECT(u64, uint64_t, uint64_t)
XSS_DEF_SELECT(i64, int64_t, int64_t)
XSS_DEF_SELECT(f64, double, double)

#define XSS_DEF_KVSEL(ktag, K, CK, vtag, V)                                                            \
    extern "C" int xss_gpu_keyvalue_select_##ktag##_##vtag(K *key, V *val, size_t k, size_t n,       \
                                                           int hasnan, int descending,               \
                                                           xss_stream_t s)                           \
    {                                                                                                \
        return run_select<TypeKey<CK>::KD, payload_of<V>()>(0, key, val, nullptr, n, k, hasnan != 0,  \
                                                           descending != 0, s);                      \
    }                                                                                                \
    extern "C" int xss_gpu_keyvalue_partial_sort_##ktag##_##vtag(K *key, V *val, size_t k, size_t n, \
                                                                 int hasnan, int descending,         \
                                                                 xss_stream_t s)                     \
    {                                                                                                \
        return run_select<TypeKey<CK>::KD, payload_of<V>()>(1, key, val, nullptr, n, k, hasnan != 0,  \
                                                           descending != 0, s);                      \
    }

#define XSS_DEF_KVSEL_ALLV(ktag, K, CK)       \
    XSS_DEF_KVSEL(ktag, K, CK, u32, uint32_t) \
    XSS_DEF_KVSEL(ktag, K, CK, i32, int32_t)  \
    XSS_DEF_KVSEL(ktag, K, CK, f32, float)    \
    XSS_DEF_KVSEL(ktag, K, CK, u64, uint64_t) \
    XSS_DEF_KVSEL(ktag, K, CK, i64, int64_t)  \
    XSS_DEF_KVSEL(ktag, K, CK, f64, double)

XSS_DEF_KVSEL_ALLV(u16, uint16_t, uint16_t)
XSS_DEF_KVSEL_ALLV(i16, int16_t, int16_t)
XSS_DEF_KVSEL_ALLV(f16, xss_f16_bits, f16_tag)
XSS_DEF_KVSEL_ALLV(u32, uint32_t, uint32_t)
XSS_DEF_KVSEL_ALLV(i32, int32_t, int32_t)
XSS_DEF_KVSEL_ALLV(f32, float, float)
XSS_DEF_KVSEL_ALLV(u64, uint64_t, uint64_t)
XSS_DEF_KVSEL_ALLV(i64, int64_t, int64_t)
XSS_DEF_KVSEL_ALLV(f64, double, double)
