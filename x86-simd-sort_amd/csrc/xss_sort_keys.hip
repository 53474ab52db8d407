// xss_sort_keys.hip -- C-ABI entry points xss_gpu_qsort_<T> and
// xss_gpu_argsort_<T> (x86simdsort::qsort / argsort, lib/x86simdsort.h:15-18,36-39),
// and the library's tuning switches.
#include <string>

#include "xss_hybrid.hpp"

using namespace xss;

#define XSS_DEF_SORT(tag, T, CT)                                                                      \
    extern "C" int xss_gpu_qsort_##tag(T *arr, size_t n, int hasnan, int descending, xss_stream_t s)  \
    {                                                                                                 \
        return run_sort_auto<TypeKey<CT>::KD, PL_NONE>(arr, arr, nullptr, nullptr, nullptr, n, hasnan != 0,  \
                                                        descending != 0, true, s);                    \
    }                                                                                                 \
    extern "C" int xss_gpu_argsort_##tag(const T *arr, uint64_t *arg, size_t n, int hasnan,           \
                                         int descending, xss_stream_t s)                              \
    {                                                                                                 \
        return run_sort_auto<TypeKey<CT>::KD, PL_IDX>(arr, nullptr, nullptr, nullptr, arg, n, hasnan != 0, \
                                                       descending != 0, false, s);                    \
    }

XSS_DEF_SORT(u16, uint16_t, uint16_t)
XSS_DEF_SORT(i16, int16_t, int16_t)
XSS_DEF_SORT(f16, xss_f16_bits, f16_tag)
XSS_DEF_SORT(u32, uint32_t, uint32_t)
XSS_DEF_SORT(i32, int32_t, int32_t)
XSS_DEF_SORT(f32, float, float)
XSS_DEF_SORT(u64, uint64_t, uint64_t)
XSS_DEF_SORT(i64, int64_t, int64_t)
XSS_DEF_SORT(f64, double, double)

namespace xss {
std::atomic<uint64_t> &hybrid_min_n()
{
    static std::atomic<uint64_t> v{HYB_MIN_N};
    return v;
}
std::atomic<uint64_t> &count_min_n()
{
    static std::atomic<uint64_t> v{CS_MIN_N};
    return v;
}
std::atomic<uint64_t> &local1_max_n()
{
    static std::atomic<uint64_t> v{L1_CAP};
    return v;
}
std::atomic<uint64_t> &big1_max_n()
{
    // measured: f32 12288 keys 0.063 vs 0.086 ms on the LSD plan, 16384
    // 0.068 vs 0.084, 24576 0.095 vs 0.085; u64 16384 0.135 vs 0.162
    static std::atomic<uint64_t> v{16384};
    return v;
}
// plan knobs of the segmented form (xss_segmented.hip)
bool seg_tuning_set(const std::string &name, uint64_t value);
}  // namespace xss

extern "C" int xss_gpu_set_tuning(const char *name, uint64_t value)
{
    if (!name) return set_error(XSS_ERR_INVALID, "null tuning name");
    if (seg_tuning_set(name, value)) return XSS_OK;
    if (std::string(name) == "hybrid_min_n") {
        hybrid_min_n().store(value);
        return XSS_OK;
    }
    if (std::string(name) == "workspace_limit") {
        workspace_limit().store(value);
        return XSS_OK;
    }
    if (std::string(name) == "big1_max_n") {
        big1_max_n().store(value);
        return XSS_OK;
    }
    if (std::string(name) == "local1_max_n") {
        local1_max_n().store(value);
        return XSS_OK;
    }
    if (std::string(name) == "count_min_n") {
        count_min_n().store(value);
        return XSS_OK;
    }
    return set_error(XSS_ERR_INVALID, "unknown tuning '%s'", name);
}
