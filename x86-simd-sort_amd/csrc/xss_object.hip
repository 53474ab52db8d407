// xss_object.hip -- object_qsort on the GPU (lib/x86simdsort.h:72-106).
//
// The reference argsorts the host-computed keys with keyvalue_qsort(keys,
// iota) and then permutes the objects in place by walking the permutation's
// cycles with swaps (lib/x86simdsort.h:93-105) -- random-access traffic
// that dominates its run time at 16M objects.  Here the keys are argsorted
// by the radix sort (index payload generated in the first pass) and the
// objects are moved once by an out-of-place gather: word w of output object
// i is word w of input object idx[i]; consecutive threads write consecutive
// words, so stores are coalesced and each object's words are read from one
// place.  The gathered copy is then written back over the input.
#include "xss_sort.hpp"

using namespace xss;

namespace {

// word-wise gather of n objects of `words` WordT each
template <class WordT>
__global__ __launch_bounds__(256) void gather_objects_kernel(const WordT *src, WordT *dst, const uint64_t *idx,
                                                             uint64_t n, uint32_t words)
{
    const uint64_t total = n * words;
    for (uint64_t w = uint64_t(blockIdx.x) * 256 + threadIdx.x; w < total; w += uint64_t(gridDim.x) * 256) {
        const uint64_t i = w / words;
        const uint32_t j = uint32_t(w - i * words);
        dst[w] = src[idx[i] * words + j];
    }
}

// sqrt(x*x + y*y + z*z) in IEEE double without contraction (this file is
// built with -ffp-contract=off); sqrt is the correctly rounded one
__global__ __launch_bounds__(256) void xyz_norm_kernel(const double *xyz, double *keys, uint64_t n)
{
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const double xx = x * x, yy = y * y, zz = z * z;
        keys[i] = __builtin_sqrt((xx + yy) + zz);
    }
}

__global__ __launch_bounds__(256) void gather_keys_kernel(const double *keys, const uint64_t *idx, double *out,
                                                          uint64_t n)
{
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256)
        out[i] = keys[idx[i]];
}

uint32_t grid_for(uint64_t work) { return uint32_t(std::min<uint64_t>(8192, (work + 255) / 256)); }

int enqueue_gather(void *objs, size_t obj_bytes, const uint64_t *idx, char *tmp, size_t n, hipStream_t s)
{
    KernelTimer t("gather_objects", s);
    if (obj_bytes % 8 == 0) {
        const uint32_t words = uint32_t(obj_bytes / 8);
        hipLaunchKernelGGL(gather_objects_kernel<uint64_t>, dim3(grid_for(uint64_t(n) * words)), dim3(256), 0, s,
                           static_cast<const uint64_t *>(objs), reinterpret_cast<uint64_t *>(tmp), idx, uint64_t(n),
                           words);
    } else if (obj_bytes % 4 == 0) {
        const uint32_t words = uint32_t(obj_bytes / 4);
        hipLaunchKernelGGL(gather_objects_kernel<uint32_t>, dim3(grid_for(uint64_t(n) * words)), dim3(256), 0, s,
                           static_cast<const uint32_t *>(objs), reinterpret_cast<uint32_t *>(tmp), idx, uint64_t(n),
                           words);
    } else {
        const uint32_t words = uint32_t(obj_bytes);
        hipLaunchKernelGGL(gather_objects_kernel<uint8_t>, dim3(grid_for(uint64_t(n) * words)), dim3(256), 0, s,
                           static_cast<const uint8_t *>(objs), reinterpret_cast<uint8_t *>(tmp), idx, uint64_t(n),
                           words);
    }
    XSS_HIP_CHECK(hipGetLastError());
    XSS_HIP_CHECK(hipMemcpyAsync(objs, tmp, n * obj_bytes, hipMemcpyDeviceToDevice, s));
    return XSS_OK;
}

template <class K>
int run_object_sort(void *objs, size_t obj_bytes, const K *keys, size_t n, hipStream_t s)
{
    using KD = typename TypeKey<K>::KD;
    using Plan = SortPlan<KD, PL_IDX>;
    if (n > MAX_N) return set_error(XSS_ERR_INVALID, "n = %zu exceeds the per-call limit %u", n, MAX_N);
    if (n <= 1) return XSS_OK;
    if (!objs || !keys || obj_bytes == 0) return set_error(XSS_ERR_INVALID, "null pointer or zero object size");
    const size_t sb = Plan::scratch_bytes(n);
    const size_t need = align256(sb) + align256(n * 8) + align256(n * obj_bytes);
    Workspace *ws = nullptr;
    std::unique_lock<std::mutex> lock;
    int rc = acquire_workspace(s, need, &ws, &lock);
    if (rc) return rc;
    char *scratch = ws->data;
    uint64_t *idx = reinterpret_cast<uint64_t *>(scratch + align256(sb));
    char *tmp = reinterpret_cast<char *>(idx) + align256(n * 8);
    // (key, index) sort: ascending, hasnan = false (lib/x86simdsort.h:90)
    rc = enqueue_sort<KD, PL_IDX>(ws, scratch, keys, nullptr, nullptr, nullptr, idx, n, false, false, false, s);
    if (rc) return rc;
    return enqueue_gather(objs, obj_bytes, idx, tmp, n, s);
}

}  // namespace

#define XSS_DEF_OBJ(tag, K)                                                                             \
    extern "C" int xss_gpu_object_sort_by_keys_##tag(void *objs, size_t obj_bytes, const K *keys, size_t n, \
                                                     xss_stream_t s)                                    \
    {                                                                                                   \
        return run_object_sort<K>(objs, obj_bytes, keys, n, s);                                         \
    }

XSS_DEF_OBJ(f64, double)
XSS_DEF_OBJ(f32, float)
XSS_DEF_OBJ(u64, uint64_t)
XSS_DEF_OBJ(i64, int64_t)
XSS_DEF_OBJ(u32, uint32_t)
XSS_DEF_OBJ(i32, int32_t)

extern "C" int xss_gpu_object_qsort_xyz_norm(double *xyz, size_t n, double *keys_out, xss_stream_t s)
{
    using KD = TypeKey<double>::KD;
    using Plan = SortPlan<KD, PL_IDX>;
    if (n > MAX_N) return set_error(XSS_ERR_INVALID, "n = %zu exceeds the per-call limit %u", n, MAX_N);
    if (n == 0) return XSS_OK;
    if (!xyz) return set_error(XSS_ERR_INVALID, "null pointer argument");
    const size_t obj_bytes = 3 * sizeof(double);
    const size_t sb = Plan::scratch_bytes(n);
    const size_t need = align256(sb) + align256(n * 8) + align256(n * obj_bytes) + align256(n * 8);
    Workspace *ws = nullptr;
    std::unique_lock<std::mutex> lock;
    int rc = acquire_workspace(s, need, &ws, &lock);
    if (rc) return rc;
    char *scratch = ws->data;
    uint64_t *idx = reinterpret_cast<uint64_t *>(scratch + align256(sb));
    char *tmp = reinterpret_cast<char *>(idx) + align256(n * 8);
    double *keys = reinterpret_cast<double *>(tmp + align256(n * obj_bytes));
    {
        KernelTimer t("xyz_norm", s);
        hipLaunchKernelGGL(xyz_norm_kernel, dim3(grid_for(n)), dim3(256), 0, s, xyz, keys, uint64_t(n));
    }
    XSS_HIP_CHECK(hipGetLastError());
    if (n == 1) {
        if (keys_out) XSS_HIP_CHECK(hipMemcpyAsync(keys_out, keys, 8, hipMemcpyDeviceToDevice, s));
        return XSS_OK;
    }
    rc = enqueue_sort<KD, PL_IDX>(ws, scratch, keys, nullptr, nullptr, nullptr, idx, n, false, false, false, s);
    if (rc) return rc;
    if (keys_out) {
        hipLaunchKernelGGL(gather_keys_kernel, dim3(grid_for(n)), dim3(256), 0, s, keys, idx, keys_out, uint64_t(n));
        XSS_HIP_CHECK(hipGetLastError());
    }
    return enqueue_gather(xyz, obj_bytes, idx, tmp, n, s);
}
