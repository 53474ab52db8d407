// xss_runtime.hpp -- device scratch workspaces and error reporting.
//
// The reference sorts in place on caller memory and allocates only the
// argsort index vector (src/x86simdsort-static-incl.h:98-121).  A GPU radix
// sort needs scratch: ping-pong key/payload buffers, digit histograms and
// look-back words.  Each (device, stream) pair owns one workspace that grows
// on demand and is reused by every later call on that stream, so steady-state
// calls allocate nothing and work on one stream never races on scratch.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <mutex>

#include "xss_kernels.hpp"

namespace xss {

constexpr size_t CTRL_BYTES = 256;
constexpr size_t RSUM_BYTES = size_t(MAX_RANGES) * 256 * sizeof(uint32_t);
constexpr size_t ANDOR_SLOTS_BYTES = size_t(MAX_RANGES) * 16;
constexpr size_t META_BYTES = CTRL_BYTES + 2 * RSUM_BYTES + 1024 + 256 + ANDOR_SLOTS_BYTES;

struct Workspace {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;             // held while a call enqueues work on this workspace
    char *meta = nullptr;      // ctrl (select state) | rsum | rpre | digit totals | pass moves
    char *data = nullptr;      // ping-pong scratch
    size_t data_cap = 0;
    std::mutex aux_mu;         // held while a call uses `aux` (taken before `mu`)
    char *aux = nullptr;       // staging of an entry point that then calls a sort or select
    size_t aux_cap = 0;

    uint32_t *rsum() const { return reinterpret_cast<uint32_t *>(meta + CTRL_BYTES); }
    uint32_t *rpre() const { return reinterpret_cast<uint32_t *>(meta + CTRL_BYTES + RSUM_BYTES); }
    uint32_t *total() const { return reinterpret_cast<uint32_t *>(meta + CTRL_BYTES + 2 * RSUM_BYTES); }
    // per pass of an LSD sort that skips constant digits: 1 if it moved the keys
    uint32_t *ran() const { return reinterpret_cast<uint32_t *>(meta + CTRL_BYTES + 2 * RSUM_BYTES + 1024); }
    // AND / OR of the order keys of such a sort (after the 8 pass flags)
    uint64_t *andor() const { return reinterpret_cast<uint64_t *>(meta + CTRL_BYTES + 2 * RSUM_BYTES + 1024 + 64); }
    // per-workgroup AND / OR of pass 0's upsweep, folded into andor() by the scan
    uint64_t *andor_slots() const
    {
        return reinterpret_cast<uint64_t *>(meta + CTRL_BYTES + 2 * RSUM_BYTES + 1024 + 256);
    }
};

// Returns the workspace of (current device, stream), locked, with at least
// `data_bytes` of scratch.  On error returns a status code and sets the
// thread's error text.
int acquire_workspace(hipStream_t stream, size_t data_bytes, Workspace **out, std::unique_lock<std::mutex> *lock);

// The aux buffer of (current device, stream), locked, with at least `bytes`.
// A caller holding it may then acquire the workspace (never the reverse).
int acquire_aux(hipStream_t stream, size_t bytes, char **out, std::unique_lock<std::mutex> *lock);

int set_error(int code, const char *fmt, ...);

// Largest workspace a stream may grow to (xss_gpu_set_tuning
// "workspace_limit", 0 = no limit): growth beyond it fails with
// XSS_ERR_NOMEM, as a refused allocation does.
std::atomic<uint64_t> &workspace_limit();

// Optional per-kernel timing (xss_gpu_timing_enable): HIP events recorded
// on the launch stream around each kernel, summed per kernel name.
bool timing_on();
// While an instance lives, the calling thread records no timing events
// (entry points that must not create runtime objects).
struct TimingSuppressed {
    TimingSuppressed();
    ~TimingSuppressed();
};
void timing_begin(const char *name, hipStream_t s);
void timing_end(hipStream_t s);
struct KernelTimer {
    hipStream_t s;
    bool on;
    KernelTimer(const char *name, hipStream_t st) : s(st), on(timing_on()) { if (on) timing_begin(name, s); }
    ~KernelTimer() { if (on) timing_end(s); }
};
int hip_fail(hipError_t e, const char *what);

}  // namespace xss

#define XSS_HIP_CHECK(call)                                        \
    do {                                                           \
        hipError_t xss_e_ = (call);                                \
        if (xss_e_ != hipSuccess) return ::xss::hip_fail(xss_e_, #call); \
    } while (0)
