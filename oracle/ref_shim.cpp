// ref_shim.cpp -- C-ABI binding around the reference library, so that tests
// and the CPU-baseline leg of bench.py can call it through ctypes.
//
// TEST INFRASTRUCTURE ONLY.  It is compiled (by oracle/Makefile) together
// with the reference's own sources, read where they lie, into
// oracle/_ref/libxss_ref.so.  No reference source is copied into this
// repository; this file only includes the reference's public header
// lib/x86simdsort.h and calls the exported templates declared there
// (lib/x86simdsort.h:15-106).
#include "x86simdsort.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SHIM extern "C" __attribute__((visibility("default")))

#define SHIM_TYPE(tag, T)                                                              \
    SHIM void xssref_qsort_##tag(T *a, size_t n, int hasnan, int desc)                  \
    {                                                                                   \
        x86simdsort::qsort(a, n, hasnan != 0, desc != 0);                               \
    }                                                                                   \
    SHIM void xssref_qselect_##tag(T *a, size_t k, size_t n, int hasnan, int desc)      \
    {                                                                                   \
        x86simdsort::qselect(a, k, n, hasnan != 0, desc != 0);                          \
    }                                                                                   \
    SHIM void xssref_partial_qsort_##tag(T *a, size_t k, size_t n, int hasnan,          \
                                         int desc)                                      \
    {                                                                                   \
        x86simdsort::partial_qsort(a, k, n, hasnan != 0, desc != 0);                    \
    }                                                                                   \
    SHIM void xssref_argsort_##tag(T *a, size_t n, int hasnan, int desc, uint64_t *out) \
    {                                                                                   \
        std::vector<size_t> r = x86simdsort::argsort(a, n, hasnan != 0, desc != 0);     \
        std::memcpy(out, r.data(), n * sizeof(uint64_t));                               \
    }                                                                                   \
    SHIM void xssref_argselect_##tag(T *a, size_t k, size_t n, int hasnan,              \
                                     uint64_t *out)                                     \
    {                                                                                   \
        std::vector<size_t> r = x86simdsort::argselect(a, k, n, hasnan != 0);           \
        std::memcpy(out, r.data(), n * sizeof(uint64_t));                               \
    }

SHIM_TYPE(u16, uint16_t)
SHIM_TYPE(i16, int16_t)
SHIM_TYPE(f16, _Float16)
SHIM_TYPE(u32, uint32_t)
SHIM_TYPE(i32, int32_t)
SHIM_TYPE(f32, float)
SHIM_TYPE(u64, uint64_t)
SHIM_TYPE(i64, int64_t)
SHIM_TYPE(f64, double)

#define SHIM_KV(ktag, K, vtag, V)                                                       \
    SHIM void xssref_keyvalue_qsort_##ktag##_##vtag(K *k, V *v, size_t n, int hasnan,   \
                                                    int desc)                           \
    {                                                                                   \
        x86simdsort::keyvalue_qsort(k, v, n, hasnan != 0, desc != 0);                   \
    }                                                                                   \
    SHIM void xssref_keyvalue_select_##ktag##_##vtag(K *k, V *v, size_t kk, size_t n,   \
                                                     int hasnan, int desc)              \
    {                                                                                   \
        x86simdsort::keyvalue_select(k, v, kk, n, hasnan != 0, desc != 0);              \
    }                                                                                   \
    SHIM void xssref_keyvalue_partial_sort_##ktag##_##vtag(K *k, V *v, size_t kk,       \
                                                           size_t n, int hasnan,        \
                                                           int desc)                    \
    {                                                                                   \
        x86simdsort::keyvalue_partial_sort(k, v, kk, n, hasnan != 0, desc != 0);        \
    }

#define SHIM_KV_ALLV(ktag, K)            \
    SHIM_KV(ktag, K, u32, uint32_t)      \
    SHIM_KV(ktag, K, i32, int32_t)       \
    SHIM_KV(ktag, K, f32, float)         \
    SHIM_KV(ktag, K, u64, uint64_t)      \
    SHIM_KV(ktag, K, i64, int64_t)       \
    SHIM_KV(ktag, K, f64, double)

SHIM_KV_ALLV(u32, uint32_t)
SHIM_KV_ALLV(i32, int32_t)
SHIM_KV_ALLV(f32, float)
SHIM_KV_ALLV(u64, uint64_t)
SHIM_KV_ALLV(i64, int64_t)
SHIM_KV_ALLV(f64, double)

// object_qsort over {double x, y, z} keyed by the Euclidean norm, the
// workload of benchmarks/bench-objsort.hpp (lib/x86simdsort.h:72-106).
struct Point3 {
    double x, y, z;
};
SHIM void xssref_object_qsort_xyz(double *xyz, size_t n)
{
    x86simdsort::object_qsort(reinterpret_cast<Point3 *>(xyz), n, [](Point3 p) -> double {
        return std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    });
}

// Threads the reference's OpenMP team may use (it caps itself at 16,
// src/xss-common-includes.h:91-94).
SHIM int xssref_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
SHIM void xssref_set_threads(int t)
{
#ifdef _OPENMP
    omp_set_num_threads(t);
#else
    (void)t;
#endif
}

// The ISA path the reference's dispatch selects for 32/64-bit, 16-bit and
// _Float16 keys, restating check_cpu_feature_support / find_preferred_cpu
// (lib/x86simdsort.cpp:8-43).
SHIM const char *xssref_isa(void)
{
    static char buf[160];
    __builtin_cpu_init();
    bool skx = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq") &&
               __builtin_cpu_supports("avx512vl");
    bool icl = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vbmi2") &&
               __builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("avx512vl");
    bool spr = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512fp16") &&
               __builtin_cpu_supports("avx512vbmi2");
    bool avx2 = __builtin_cpu_supports("avx2");
    if (std::getenv("XSS_DISABLE_AVX512")) skx = icl = spr = false;
    std::snprintf(buf, sizeof buf, "32/64-bit:%s 16-bit:%s _Float16:%s",
                  skx ? "avx512_skx" : (avx2 ? "avx2" : "scalar"), icl ? "avx512_icl" : "scalar",
                  spr ? "avx512_spr" : (icl ? "avx512_icl(emulated)" : "scalar"));
    return buf;
}
