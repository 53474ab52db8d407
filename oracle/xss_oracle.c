/*
 * xss_oracle.c -- CPU oracle for the x86simdsort:: sort/select family.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py may use it, and only as the checker.
 *
 * It is a plain-C restatement of the reference's *observable semantics*,
 * written as a textbook stable merge sort driven by the reference's
 * NaN-aware comparator.  Each function cites the reference file:line whose
 * behaviour it restates (paths relative to the intel/x86-simd-sort tree):
 *
 *   comparator  compare<T,less|greater>      utils/custom-compare.h:11-34
 *   qsort       xss_qsort + NaN canonicalise src/xss-common-qsort.h:652-710,
 *               replace_nan_with_inf/replace_inf_with_nan :54-79,107-135,
 *               canonical NaN constants      src/xss-custom-float.h:34-38,81-85
 *               (_Float16 0x7c01: src/avx512fp16-16bit-qsort.hpp:178-199,
 *                emulated fp16: src/avx512-16bit-qsort.hpp:522-543,582-608)
 *   qselect     xss_qselect                  src/xss-common-qsort.h:712-752
 *   partial     xss_partial_qsort            src/xss-common-qsort.h:754-762
 *   argsort     xss_argsort                  src/xss-common-argsort.h:587-669
 *   argselect   xss_argselect                src/xss-common-argsort.h:693-731
 *   kv sort     xss_qsort_kv                 src/xss-common-keyvaluesort.hpp:564-653
 *   kv select   xss_select_kv                src/xss-common-keyvaluesort.hpp:655-718
 *   kv partial  xss_partial_sort_kv          src/xss-common-keyvaluesort.hpp:720-738
 *   object      object_qsort                 lib/x86simdsort.h:72-106
 *
 * Where the reference leaves an arrangement unspecified (order inside a run
 * of equal keys, the non-k elements after a select), this oracle returns one
 * valid arrangement -- the fully sorted one -- and the checkers in tests/
 * compare those parts as multisets or through the partition property, as the
 * reference's own tests do (tests/test-qsort-common.h:29-104,
 * tests/test-keyvalue.cpp:44-171).
 *
 * _Float16 is handled as its uint16_t bit pattern (gcc 11 has no _Float16 in
 * C); values are widened exactly to double for comparison.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- helpers */

static double half_to_double(uint16_t h)
{
    int sign = h >> 15, exp = (h >> 10) & 0x1f, mant = h & 0x3ff;
    double v;
    if (exp == 0) v = ldexp((double)mant, -24);            /* subnormal / 0 */
    else if (exp == 31) v = mant ? NAN : INFINITY;
    else v = ldexp((double)(mant | 0x400), exp - 25);
    return sign ? -v : v;
}
static int half_isnan(uint16_t h) { return (h & 0x7c00u) == 0x7c00u && (h & 0x3ffu); }

/* Canonical NaN bit patterns written by qsort(hasnan=true):
 * std::numeric_limits<T>::quiet_NaN() for float/double, 0x7c01 for _Float16
 * (src/xss-custom-float.h:34-38,81-85). */
static const uint32_t QNAN_F32 = 0x7fc00000u;
static const uint64_t QNAN_F64 = 0x7ff8000000000000ull;
static const uint16_t QNAN_F16 = 0x7c01u;

/* compare<T, Op> of utils/custom-compare.h:11-34, for a floating type whose
 * values are given as doubles plus explicit NaN flags.  op: 0 = less, 1 = greater. */
static int fp_compare(double a, int na, double b, int nb, int op)
{
#define OP(x, y) (op ? ((x) > (y)) : ((x) < (y)))
    const double inf = INFINITY, one = 1.0;
    if (!na && !nb) return OP(a, b);
    if (na && !nb) return b == inf ? OP(inf, one) : OP(inf, b);
    if (!na && nb) return a == inf ? OP(one, inf) : OP(a, inf);
    return OP(one, one);
#undef OP
}

/* ------------------------------------------------------ per-type plumbing */
/* X(tag, C type, is_float, to_double, isnan, set_canonical_nan) */
#define F32_D(x) ((double)(x))
#define F32_NAN(x) isnan(x)
#define F64_D(x) (x)
#define F64_NAN(x) isnan(x)
#define F16_D(x) half_to_double(x)
#define F16_NAN(x) half_isnan(x)
#define INT_D(x) 0.0
#define INT_NAN(x) 0

static void canon_f32(float *p) { memcpy(p, &QNAN_F32, 4); }
static void canon_f64(double *p) { memcpy(p, &QNAN_F64, 8); }
static void canon_f16(uint16_t *p) { *p = QNAN_F16; }
#define canon_int(p) ((void)(p))

#define ALL_TYPES(X)                                              \
    X(u16, uint16_t, 0, INT_D, INT_NAN, canon_int)                \
    X(i16, int16_t, 0, INT_D, INT_NAN, canon_int)                 \
    X(f16, uint16_t, 1, F16_D, F16_NAN, canon_f16)                \
    X(u32, uint32_t, 0, INT_D, INT_NAN, canon_int)                \
    X(i32, int32_t, 0, INT_D, INT_NAN, canon_int)                 \
    X(f32, float, 1, F32_D, F32_NAN, canon_f32)                   \
    X(u64, uint64_t, 0, INT_D, INT_NAN, canon_int)                \
    X(i64, int64_t, 0, INT_D, INT_NAN, canon_int)                 \
    X(f64, double, 1, F64_D, F64_NAN, canon_f64)

/*
 * For each type: before(a, b, desc) is compare<T, less> (desc = 0) or
 * compare<T, greater> (desc = 1); a stable bottom-up merge sort of values and
 * one of indices; and the public entry points.
 */
#define DEFINE_TYPE(tag, T, ISF, TOD, ISNAN, CANON)                                \
    static inline int tag##_before(T a, T b, int desc)                            \
    {                                                                              \
        if (ISF) return fp_compare(TOD(a), ISNAN(a), TOD(b), ISNAN(b), desc);     \
        return desc ? (a > b) : (a < b);                                           \
    }                                                                              \
    static void tag##_msort(T *a, size_t n, int desc)                              \
    {                                                                              \
        if (n < 2) return;                                                         \
        T *tmp = (T *)malloc(n * sizeof(T));                                       \
        T *src = a, *dst = tmp;                                                    \
        for (size_t w = 1; w < n; w *= 2) {                                        \
            for (size_t lo = 0; lo < n; lo += 2 * w) {                             \
                size_t mid = lo + w < n ? lo + w : n;                              \
                size_t hi = lo + 2 * w < n ? lo + 2 * w : n;                       \
                size_t i = lo, j = mid, o = lo;                                    \
                while (i < mid && j < hi)                                          \
                    dst[o++] = tag##_before(src[j], src[i], desc) ? src[j++]       \
                                                                   : src[i++];     \
                while (i < mid) dst[o++] = src[i++];                               \
                while (j < hi) dst[o++] = src[j++];                                \
            }                                                                      \
            T *t = src; src = dst; dst = t;                                        \
        }                                                                          \
        if (src != a) memcpy(a, src, n * sizeof(T));                               \
        free(tmp);                                                                 \
    }                                                                              \
    /* stable argsort of keys a[0..n) (ascending when desc = 0) */                 \
    static void tag##_argmsort(const T *a, uint64_t *idx, size_t n, int desc)      \
    {                                                                              \
        for (size_t i = 0; i < n; i++) idx[i] = i;                                 \
        if (n < 2) return;                                                         \
        uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));                  \
        uint64_t *src = idx, *dst = tmp;                                           \
        for (size_t w = 1; w < n; w *= 2) {                                        \
            for (size_t lo = 0; lo < n; lo += 2 * w) {                             \
                size_t mid = lo + w < n ? lo + w : n;                              \
                size_t hi = lo + 2 * w < n ? lo + 2 * w : n;                       \
                size_t i = lo, j = mid, o = lo;                                    \
                while (i < mid && j < hi)                                          \
                    dst[o++] = tag##_before(a[src[j]], a[src[i]], desc) ? src[j++] \
                                                                        : src[i++]; \
                while (i < mid) dst[o++] = src[i++];                               \
                while (j < hi) dst[o++] = src[j++];                                \
            }                                                                      \
            uint64_t *t = src; src = dst; dst = t;                                 \
        }                                                                          \
        if (src != idx) memcpy(idx, src, n * sizeof(uint64_t));                    \
        free(tmp);                                                                 \
    }                                                                              \
    /* xss_qsort (src/xss-common-qsort.h:652-710): NaN-aware order; with         \
     * hasnan the nan_count slots at the end (asc) / start (desc) hold the       \
     * canonical quiet NaN (replace_inf_with_nan, :107-135). */                    \
    EXPORT void xss_oracle_qsort_##tag(T *a, size_t n, int hasnan, int desc)       \
    {                                                                              \
        if (n < 2) return;                                                         \
        size_t nan_count = 0;                                                      \
        if (ISF && hasnan)                                                         \
            for (size_t i = 0; i < n; i++) nan_count += ISNAN(a[i]) ? 1 : 0;       \
        tag##_msort(a, n, desc);                                                   \
        for (size_t c = 0; c < nan_count; c++) CANON(&a[desc ? c : n - 1 - c]);    \
    }                                                                              \
    /* xss_qselect (src/xss-common-qsort.h:712-752): NaN bits are preserved and  \
     * moved to the end (asc) / start (desc); the fully sorted arrangement is    \
     * one valid result. */                                                        \
    EXPORT void xss_oracle_qselect_##tag(T *a, size_t k, size_t n, int hasnan,     \
                                         int desc)                                 \
    {                                                                              \
        (void)k; (void)hasnan;                                                     \
        tag##_msort(a, n, desc);                                                   \
    }                                                                              \
    /* xss_partial_qsort (src/xss-common-qsort.h:754-762):                       \
     * k == 0 -> no-op; else qselect(k-1) then qsort of the first k-1. */          \
    EXPORT void xss_oracle_partial_qsort_##tag(T *a, size_t k, size_t n,           \
                                               int hasnan, int desc)               \
    {                                                                              \
        if (k == 0) return;                                                        \
        xss_oracle_qselect_##tag(a, k - 1, n, hasnan, desc);                       \
        xss_oracle_qsort_##tag(a, k - 1, hasnan, desc);                            \
    }                                                                              \
    /* xss_argsort (src/xss-common-argsort.h:587-669): ascending NaN-last order, \
     * reversed for descending. */                                                 \
    EXPORT void xss_oracle_argsort_##tag(const T *a, size_t n, int hasnan,         \
                                         int desc, uint64_t *out)                  \
    {                                                                              \
        (void)hasnan;                                                              \
        tag##_argmsort(a, out, n, 0);                                              \
        if (desc)                                                                  \
            for (size_t i = 0, j = n ? n - 1 : 0; i < j; i++, j--) {               \
                uint64_t t = out[i]; out[i] = out[j]; out[j] = t;                  \
            }                                                                      \
    }                                                                              \
    /* xss_argselect (src/xss-common-argsort.h:693-731): ascending only. */        \
    EXPORT void xss_oracle_argselect_##tag(const T *a, size_t k, size_t n,         \
                                           int hasnan, uint64_t *out)              \
    {                                                                              \
        (void)k;                                                                   \
        xss_oracle_argsort_##tag(a, n, hasnan, 0, out);                            \
    }                                                                              \
    /* xss_qsort_kv (src/xss-common-keyvaluesort.hpp:564-653): ascending sort    \
     * with NaN keys (bits kept) last, then both arrays reversed if desc.        \
     * The value type only matters through its width. */                           \
    static void tag##_kvsort(T *key, void *val, size_t vsz, size_t n, int desc)    \
    {                                                                              \
        if (n < 2) return;                                                         \
        uint64_t *idx = (uint64_t *)malloc(n * sizeof(uint64_t));                  \
        tag##_argmsort(key, idx, n, 0);                                            \
        T *k2 = (T *)malloc(n * sizeof(T));                                        \
        unsigned char *v2 = (unsigned char *)malloc(n * vsz);                      \
        for (size_t i = 0; i < n; i++) {                                           \
            size_t s = desc ? idx[n - 1 - i] : idx[i];                             \
            k2[i] = key[s];                                                        \
            memcpy(v2 + i * vsz, (unsigned char *)val + s * vsz, vsz);             \
        }                                                                          \
        memcpy(key, k2, n * sizeof(T));                                            \
        memcpy(val, v2, n * vsz);                                                  \
        free(idx); free(k2); free(v2);                                             \
    }                                                                              \
    EXPORT void xss_oracle_keyvalue_qsort_##tag(T *key, void *val, size_t vsz,     \
                                                size_t n, int hasnan, int desc)    \
    {                                                                              \
        (void)hasnan;                                                              \
        tag##_kvsort(key, val, vsz, n, desc);                                      \
    }                                                                              \
    /* xss_select_kv (:655-718): the full kv sort is one valid result (it is     \
     * what xss::scalar::keyvalue_select does, lib/x86simdsort-scalar.h:110-121). */ \
    EXPORT void xss_oracle_keyvalue_select_##tag(T *key, void *val, size_t vsz,    \
                                                 size_t k, size_t n, int hasnan,   \
                                                 int desc)                         \
    {                                                                              \
        (void)k; (void)hasnan;                                                     \
        tag##_kvsort(key, val, vsz, n, desc);                                      \
    }                                                                              \
    /* xss_partial_sort_kv (:720-738): select(k-1) + kv sort of the first k-1. */  \
    EXPORT void xss_oracle_keyvalue_partial_sort_##tag(T *key, void *val,          \
                                                       size_t vsz, size_t k,       \
                                                       size_t n, int hasnan,       \
                                                       int desc)                   \
    {                                                                              \
        (void)hasnan;                                                              \
        if (k == 0) return;                                                        \
        tag##_kvsort(key, val, vsz, n, desc);                                      \
    }

ALL_TYPES(DEFINE_TYPE)

/*
 * object_qsort (lib/x86simdsort.h:72-106) for objects {double x, y, z} keyed
 * by sqrt(x*x + y*y + z*z) (the Euclidean-distance benchmark of
 * benchmarks/bench-objsort.hpp).  Keys are computed in ISO C arithmetic
 * (built with -ffp-contract=off), sorted ascending (hasnan = false), and the
 * objects permuted by the resulting order.  Also returns the keys used.
 */
EXPORT void xss_oracle_object_qsort_xyz(double *xyz, size_t n, double *keys_out)
{
    if (n < 2) {
        for (size_t i = 0; i < n; i++) {
            double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
            keys_out[i] = sqrt(x * x + y * y + z * z);
        }
        return;
    }
    double *keys = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) {
        double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        keys[i] = sqrt(x * x + y * y + z * z);
    }
    uint64_t *idx = (uint64_t *)malloc(n * sizeof(uint64_t));
    f64_argmsort(keys, idx, n, 0);
    double *tmp = (double *)malloc(3 * n * sizeof(double));
    for (size_t i = 0; i < n; i++) {
        memcpy(tmp + 3 * i, xyz + 3 * idx[i], 3 * sizeof(double));
        keys_out[i] = keys[idx[i]];
    }
    memcpy(xyz, tmp, 3 * n * sizeof(double));
    free(keys); free(idx); free(tmp);
}

/* compare<T, Op> for the comparator known-answer test of
 * tests/test-qsort.cpp:261-292: op 0 less, 1 greater, 2 less_equal,
 * 3 greater_equal, 4 equal_to, applied to f64 values (NaN allowed). */
EXPORT int xss_oracle_compare_f64(double a, double b, int op)
{
    int na = isnan(a), nb = isnan(b);
    const double inf = INFINITY, one = 1.0;
#define APPLY(x, y) (op == 0 ? (x) < (y) : op == 1 ? (x) > (y) : op == 2 ? (x) <= (y) \
                     : op == 3 ? (x) >= (y) : (x) == (y))
    if (!na && !nb) return APPLY(a, b);
    if (na && !nb) return b == inf ? APPLY(inf, one) : APPLY(inf, b);
    if (!na && nb) return a == inf ? APPLY(one, inf) : APPLY(a, inf);
    return APPLY(one, one);
#undef APPLY
}

EXPORT const char *xss_oracle_version(void) { return "xss-oracle 1 (stable merge sort restatement)"; }
