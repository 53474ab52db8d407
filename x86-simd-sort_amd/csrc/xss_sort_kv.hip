// xss_sort_kv.hip -- C-ABI entry points xss_gpu_keyvalue_qsort_<K>_<V>
// (x86simdsort::keyvalue_qsort, lib/x86simdsort.h:46-52; the 36 key/value
// pairs of lib/x86simdsort.cpp:280-293, plus 16-bit keys u16 / i16 / f16
// with the same six value types, which the reference does not sort).
// Values move as 32- or 64-bit payloads; the value type matters only
// through its width.
#include "xss_hybrid.hpp"

using namespace xss;

#define XSS_DEF_KV(ktag, K, CK, vtag, V)                                                          \
    extern "C" int xss_gpu_keyvalue_qsort_##ktag##_##vtag(K *key, V *val, size_t n, int hasnan,   \
                                                          int descending, xss_stream_t s)         \
    {                                                                                             \
        return run_sort_auto<TypeKey<CK>::KD, payload_of<V>()>(key, key, val, val, nullptr, n,    \
                                                         hasnan != 0, descending != 0, false, s); \
    }

#define XSS_DEF_KV_ALLV(ktag, K, CK)       \
    XSS_DEF_KV(ktag, K, CK, u32, uint32_t) \
    XSS_DEF_KV(ktag, K, CK, i32, int32_t)  \
    XSS_DEF_KV(ktag, K, CK, f32, float)    \
    XSS_DEF_KV(ktag, K, CK, u64, uint64_t) \
    XSS_DEF_KV(ktag, K, CK, i64, int64_t)  \
    XSS_DEF_KV(ktag, K, CK, f64, double)

XSS_DEF_KV_ALLV(u16, uint16_t, uint16_t)
XSS_DEF_KV_ALLV(i16, int16_t, int16_t)
XSS_DEF_KV_ALLV(f16, xss_f16_bits, f16_tag)
XSS_DEF_KV_ALLV(u32, uint32_t, uint32_t)
XSS_DEF_KV_ALLV(i32, int32_t, int32_t)
XSS_DEF_KV_ALLV(f32, float, float)
XSS_DEF_KV_ALLV(u64, uint64_t, uint64_t)
XSS_DEF_KV_ALLV(i64, int64_t, int64_t)
XSS_DEF_KV_ALLV(f64, double, double)
