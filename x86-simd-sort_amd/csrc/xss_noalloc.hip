// xss_noalloc.hip -- qsort<float> that allocates nothing.
//
// xss_gpu_qsort_f32 takes its scratch (digit tables, the ping-pong key
// buffer) from a per-stream workspace that grows with hipMalloc.  Callers
// that cannot allocate -- code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc -- call xss_gpu_qsort_f32_noalloc instead.
// Its scratch is one file-scope device array of NOALLOC_SCRATCH_BYTES, so it
// serves arrays of at most NOALLOC_MAX_N keys.  Each GPU holds its own copy
// of the array for the life of the process; every call on a GPU shares it,
// so calls that overlap on one GPU (on different streams, or from different
// threads) race and must be serialised by the caller.
//
// The kernels are those of xss_gpu_qsort_f32's LSD path (xss_sort.hpp):
// one LDS block sort up to 4096 keys, four 8-bit passes above.  Nothing in
// the scratch needs to start at zero: the AND/OR words are reset by a kernel
// at the start of every call, every pass writes its own "moved" flag before
// a later pass reads it, and each pass's histograms are written in full by
// its upsweep before its scan reads them.
#include <mutex>

#include "xss_sort.hpp"

using namespace xss;

namespace {

constexpr size_t NOALLOC_MAX_N = size_t(1) << 22;  // 4,194,304 keys (16 MiB of float)
constexpr size_t NOALLOC_DATA_BYTES = (NOALLOC_MAX_N * sizeof(float) + 255) & ~size_t(255);
constexpr size_t NOALLOC_META_BYTES = (META_BYTES + 255) & ~size_t(255);
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_META_BYTES + NOALLOC_DATA_BYTES;  // ~24 MiB

// The scratch.  In host code `scratch` names a host-side stand-in, never the
// device memory: only the address hipGetSymbolAddress returns for the
// current device is passed to kernels.
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
char *g_ptr[MAX_DEVICES] = {};  // device address of `scratch`, per device

// The device address of this device's scratch, looked up on the first call
// made with the device current.
int scratch_of(int dev, char **out)
{
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return set_error(XSS_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_ptr[dev] = static_cast<char *>(p);
    }
    *out = g_ptr[dev];
    return XSS_OK;
}

}  // namespace

extern "C" int xss_gpu_qsort_f32_noalloc(float *arr, size_t n, int hasnan, int descending, xss_stream_t s)
{
    if (n > NOALLOC_MAX_N)
        return set_error(XSS_ERR_INVALID, "xss_gpu_qsort_f32_noalloc: n = %zu exceeds its scratch limit of %zu keys",
                         n, NOALLOC_MAX_N);
    if (n < 2) return XSS_OK;  // n <= 1 is a no-op (src/xss-common-qsort.h:660)
    if (!arr) return set_error(XSS_ERR_INVALID, "null pointer argument");
    int dev = 0;
    XSS_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(XSS_ERR_INVALID, "device %d out of range", dev);
    char *base = nullptr;
    int rc = scratch_of(dev, &base);
    if (rc) return rc;
    // a workspace view of the static scratch: meta tables, then the data buffer
    Workspace ws;
    ws.device = dev;
    ws.stream = s;
    ws.meta = base;
    ws.data = base + NOALLOC_META_BYTES;
    ws.data_cap = NOALLOC_DATA_BYTES;
    TimingSuppressed quiet;  // the profiling switch's events are an allocation too
    return enqueue_sort<TypeKey<float>::KD, PL_NONE>(&ws, ws.data, arr, arr, nullptr, nullptr, nullptr, n,
                                                     hasnan != 0, descending != 0, true, s);
}
