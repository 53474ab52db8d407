/*
 * xss_gpu.h -- C ABI of the MI355X (gfx950) sort/select library.
 *
 * Device-pointer entry points.  Every function is stream-ordered and
 * asynchronous: it enqueues kernels on `stream` (NULL = the default stream)
 * and returns; results are valid once the stream has drained.  Scratch memory
 * comes from a per-(device, stream) workspace the library keeps and grows.
 * Return value: XSS_OK or an error code; xss_gpu_last_error() describes the
 * last error of the calling thread.  Semantics (NaN placement, canonical NaN
 * bits, ascending/descending, k handling) are those of the x86simdsort::
 * functions each entry point replaces:
 *
 *   xss_gpu_qsort_<T>              x86simdsort::qsort<T>           lib/x86simdsort.h:15-18, lib/x86simdsort.cpp:60-66
 *   xss_gpu_qselect_<T>            x86simdsort::qselect<T>         lib/x86simdsort.h:20-26, lib/x86simdsort.cpp:68-76
 *   xss_gpu_partial_qsort_<T>      x86simdsort::partial_qsort<T>   lib/x86simdsort.h:28-34, lib/x86simdsort.cpp:78-87
 *   xss_gpu_argsort_<T>            x86simdsort::argsort<T>         lib/x86simdsort.h:36-39, lib/x86simdsort.cpp:89-98
 *   xss_gpu_argselect_<T>          x86simdsort::argselect<T>       lib/x86simdsort.h:41-44, lib/x86simdsort.cpp:100-109
 *   xss_gpu_keyvalue_qsort_<K>_<V>        x86simdsort::keyvalue_qsort<K,V>        lib/x86simdsort.h:46-52, lib/x86simdsort.cpp:207-221
 *   xss_gpu_keyvalue_select_<K>_<V>       x86simdsort::keyvalue_select<K,V>       lib/x86simdsort.h:54-61, lib/x86simdsort.cpp:222-233
 *   xss_gpu_keyvalue_partial_sort_<K>_<V> x86simdsort::keyvalue_partial_sort<K,V> lib/x86simdsort.h:63-70, lib/x86simdsort.cpp:234-246
 *
 * <T> in {u16, i16, f16, u32, i32, f32, u64, i64, f64}; <K>,<V> in
 * {u32, i32, f32, u64, i64, f64} (the 36 pairs of lib/x86simdsort.cpp:280-293),
 * and <K> in {u16, i16, f16} with the same six <V>: 16-bit keys with
 * payloads, which the reference does not provide (it sorts 16-bit keys
 * only alone or for argsort, lib/x86simdsort.cpp:167-204).
 * argsort/argselect write n size_t (uint64_t) indices to `arg` and leave the
 * keys untouched, like the reference (src/x86simdsort-static-incl.h:92-121).
 *
 * Limits: n <= 4,278,190,080 (2^32 - 2^24) keys per array (positions are
 * 32-bit on the device); a larger n returns XSS_ERR_INVALID.  The reference
 * takes any size_t.  Host-pointer forms are the C++ templates of
 * x86simdsort.h and the 8 keyvalue_qsort_* symbols declared at the end.
 */
#ifndef XSS_GPU_H
#define XSS_GPU_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define XSS_API __attribute__((visibility("default")))
#else
#define XSS_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *xss_stream_t; /* == hipStream_t */
typedef uint16_t xss_f16_bits;             /* _Float16 passed as its bit pattern */

enum {
    XSS_OK = 0,
    XSS_ERR_INVALID = 1,     /* bad argument (null pointer, k >= n, n too large) */
    XSS_ERR_HIP = 2,         /* a HIP runtime call failed */
    XSS_ERR_NOMEM = 3,       /* device scratch could not be allocated */
    XSS_ERR_CAPTURE = 4,     /* workspace must grow while the stream is being captured */
    XSS_ERR_NO_DEVICE = 5    /* no HIP device */
};

XSS_API const char *xss_gpu_last_error(void);
XSS_API const char *xss_gpu_version(void);
/* Release the calling thread's workspaces on every stream (synchronises). */
XSS_API int xss_gpu_release_workspaces(void);
/* Bytes of device scratch currently held by all workspaces. */
XSS_API size_t xss_gpu_workspace_bytes(void);
/* Per-kernel timing for profiling: when enabled, HIP events are recorded on
 * the launch stream around every kernel the library enqueues.  collect()
 * waits for them, returns up to max_names (name, total ms, launches) rows
 * and clears the record.  No counterpart in the reference. */
XSS_API int xss_gpu_timing_enable(int on);
/* Runtime tuning switches (no counterpart in the reference):
 *   "hybrid_min_n"  qsort / keyvalue_qsort / argsort of 4- and 8-byte
 *                   keys sort arrays of at least this many keys MSD-first
 *                   (two passes over HBM, then every bucket of the two
 *                   most significant varying digits in LDS); default 2^25.
 *   "count_min_n"   qsort of 16-bit keys counts keys per value (a counting
 *                   sort: one read, one write) for arrays of at least this
 *                   many keys; default 2^23.
 *   "local1_max_n"  qsort / argsort of 4- and 8-byte keys sort an array above the
 *                   LDS bitonic size (4096 keys; 2048 for 8-byte keys) and
 *                   up to this many keys (at most 6144) in one workgroup, in
 *                   one launch; default 6144, 0 = the LSD plan.
 *   "big1_max_n"    qsort of 4- and 8-byte keys above that and up to this
 *                   many keys runs its LSD passes in one workgroup through
 *                   the cache (two launches); default 16384, 0 = off.
 *   "workspace_limit" bytes a stream's workspace may grow to (0, the
 *                   default: no limit); a call that would need more fails
 *                   with XSS_ERR_NOMEM, like a refused allocation. */
XSS_API int xss_gpu_set_tuning(const char *name, uint64_t value);
XSS_API int xss_gpu_timing_collect(int max_names, const char **names, double *total_ms, uint64_t *counts);

#define XSS_DECLARE_T(tag, T)                                                                      \
    XSS_API int xss_gpu_qsort_##tag(T *arr, size_t n, int hasnan, int descending, xss_stream_t stream);    \
    XSS_API int xss_gpu_qselect_##tag(T *arr, size_t k, size_t n, int hasnan, int descending,              \
                              xss_stream_t stream);                                                \
    XSS_API int xss_gpu_partial_qsort_##tag(T *arr, size_t k, size_t n, int hasnan, int descending,        \
                                    xss_stream_t stream);                                          \
    XSS_API int xss_gpu_argsort_##tag(const T *arr, uint64_t *arg, size_t n, int hasnan, int descending,   \
                              xss_stream_t stream);                                                \
    XSS_API int xss_gpu_argselect_##tag(const T *arr, uint64_t *arg, size_t k, size_t n, int hasnan,       \
                                xss_stream_t stream);

XSS_DECLARE_T(u16, uint16_t)
XSS_DECLARE_T(i16, int16_t)
XSS_DECLARE_T(f16, xss_f16_bits)
XSS_DECLARE_T(u32, uint32_t)
XSS_DECLARE_T(i32, int32_t)
XSS_DECLARE_T(f32, float)
XSS_DECLARE_T(u64, uint64_t)
XSS_DECLARE_T(i64, int64_t)
XSS_DECLARE_T(f64, double)

/*
 * qsort<float> that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * semantics as xss_gpu_qsort_f32, but its scratch is a fixed device array
 * inside the library (about 24 MiB per GPU, held until the process ends), so
 * n is limited to 4,194,304 keys (larger n: XSS_ERR_INVALID).  Every call on
 * one GPU uses that GPU's one array: calls that overlap on a GPU (several
 * streams or threads at once) race, and the caller must serialise them.
 */
XSS_API int xss_gpu_qsort_f32_noalloc(float *arr, size_t n, int hasnan, int descending, xss_stream_t stream);

/*
 * NumPy-style arg forms (x86simdsortStatic::argsort(T *, size_t *arg, size,
 * hasnan, descending) and argselect(T *, size_t *arg, k, size, hasnan),
 * src/x86simdsort-static-incl.h:31-36,43-46,92-97,108-113).  `arg` holds n
 * indices into `arr` on entry (NumPy passes 0..n-1) and is reordered in place
 * so that arr[arg[i]] is sorted, or has rank k at position k with smaller
 * keys before it and larger after (argselect), with the NaN and descending
 * rules of argsort / argselect.  `arr` is not modified.  Indices must be
 * < n; an index >= n (undefined input for the reference) is read as index 0.
 * The xss_gpu_ forms take device pointers and are stream-ordered; the
 * xss_host_ forms take host pointers, stage them through the GPU and return
 * when `arg` holds the result.
 */
#define XSS_DECLARE_ARGFORM(tag, T)                                                                     \
    XSS_API int xss_gpu_argsort_arg_##tag(const T *arr, uint64_t *arg, size_t n, int hasnan, int descending,   \
                                          xss_stream_t stream);                                         \
    XSS_API int xss_gpu_argselect_arg_##tag(const T *arr, uint64_t *arg, size_t k, size_t n, int hasnan,       \
                                            xss_stream_t stream);                                       \
    XSS_API int xss_host_argsort_arg_##tag(const T *arr, size_t *arg, size_t n, int hasnan, int descending);   \
    XSS_API int xss_host_argselect_arg_##tag(const T *arr, size_t *arg, size_t k, size_t n, int hasnan);

XSS_DECLARE_ARGFORM(u16, uint16_t)
XSS_DECLARE_ARGFORM(i16, int16_t)
XSS_DECLARE_ARGFORM(f16, xss_f16_bits)
XSS_DECLARE_ARGFORM(u32, uint32_t)
XSS_DECLARE_ARGFORM(i32, int32_t)
XSS_DECLARE_ARGFORM(f32, float)
XSS_DECLARE_ARGFORM(u64, uint64_t)
XSS_DECLARE_ARGFORM(i64, int64_t)
XSS_DECLARE_ARGFORM(f64, double)

#define XSS_DECLARE_KV(ktag, K, vtag, V)                                                           \
    XSS_API int xss_gpu_keyvalue_qsort_##ktag##_##vtag(K *key, V *val, size_t n, int hasnan,               \
                                               int descending, xss_stream_t stream);               \
    XSS_API int xss_gpu_keyvalue_select_##ktag##_##vtag(K *key, V *val, size_t k, size_t n, int hasnan,    \
                                                int descending, xss_stream_t stream);              \
    XSS_API int xss_gpu_keyvalue_partial_sort_##ktag##_##vtag(K *key, V *val, size_t k, size_t n,          \
                                                      int hasnan, int descending,                  \
                                                      xss_stream_t stream);

#define XSS_DECLARE_KV_ALLV(ktag, K)     \
    XSS_DECLARE_KV(ktag, K, u32, uint32_t) \
    XSS_DECLARE_KV(ktag, K, i32, int32_t)  \
    XSS_DECLARE_KV(ktag, K, f32, float)    \
    XSS_DECLARE_KV(ktag, K, u64, uint64_t) \
    XSS_DECLARE_KV(ktag, K, i64, int64_t)  \
    XSS_DECLARE_KV(ktag, K, f64, double)

XSS_DECLARE_KV_ALLV(u16, uint16_t)
XSS_DECLARE_KV_ALLV(i16, int16_t)
XSS_DECLARE_KV_ALLV(f16, xss_f16_bits)
XSS_DECLARE_KV_ALLV(u32, uint32_t)
XSS_DECLARE_KV_ALLV(i32, int32_t)
XSS_DECLARE_KV_ALLV(f32, float)
XSS_DECLARE_KV_ALLV(u64, uint64_t)
XSS_DECLARE_KV_ALLV(i64, int64_t)
XSS_DECLARE_KV_ALLV(f64, double)

/*
 * object_qsort support (lib/x86simdsort.h:72-106).  The reference computes
 * one 32/64-bit key per object on the host, argsorts the keys with
 * keyvalue_qsort(keys, iota) and permutes the objects in place along cycles.
 * Here the permutation is a device gather: xss_gpu_object_sort_by_keys_<K>
 * argsorts the keys ascending (hasnan = false, as the reference) and gathers
 * the `obj_bytes`-sized objects into that order, in place in `objs`.  `keys`
 * is left unchanged.
 */
XSS_API int xss_gpu_object_sort_by_keys_f64(void *objs, size_t obj_bytes, const double *keys, size_t n,
                                            xss_stream_t stream);
XSS_API int xss_gpu_object_sort_by_keys_f32(void *objs, size_t obj_bytes, const float *keys, size_t n,
                                            xss_stream_t stream);
XSS_API int xss_gpu_object_sort_by_keys_u64(void *objs, size_t obj_bytes, const uint64_t *keys, size_t n,
                                            xss_stream_t stream);
XSS_API int xss_gpu_object_sort_by_keys_i64(void *objs, size_t obj_bytes, const int64_t *keys, size_t n,
                                            xss_stream_t stream);
XSS_API int xss_gpu_object_sort_by_keys_u32(void *objs, size_t obj_bytes, const uint32_t *keys, size_t n,
                                            xss_stream_t stream);
XSS_API int xss_gpu_object_sort_by_keys_i32(void *objs, size_t obj_bytes, const int32_t *keys, size_t n,
                                            xss_stream_t stream);
/* object_qsort over {double x, y, z} keyed by sqrt(x*x + y*y + z*z), with
 * the key computed on the device in IEEE double arithmetic without
 * contraction (the same value a host computes with -ffp-contract=off);
 * keys_out (n doubles, may be NULL) receives the keys in sorted order.
 * The Euclidean-distance workload of the reference's object_qsort
 * benchmark (benchmarks/bench-objsort.hpp). */
XSS_API int xss_gpu_object_qsort_xyz_norm(double *xyz, size_t n, double *keys_out, xss_stream_t stream);

/*
 * Segmented form (no counterpart in the reference, which sorts one array per
 * call): many independent arrays, stored back to back in one device buffer,
 * sorted / argsorted / selected by one pipeline.  `offsets` (host memory,
 * nseg + 1 entries, offsets[0] = 0) delimit the segments.  A plan buckets the
 * segments by size once and owns all the device memory its runs need, so
 * xss_gpu_segplan_run allocates nothing and may be captured in a hipGraph.
 *   op        "qsort"   in-place sort of every segment (qsort semantics);
 *             "argsort" writes segment-local uint32 indices to idx_out;
 *             "qselect" every segment partitioned at k = len / 2, as
 *                       qselect(seg, len / 2) leaves it (NaN bits kept): the
 *                       k-th element in place, smaller ones before, larger
 *                       ones after.  4- and 8-byte keys: segments above
 *                       one LDS workgroup are split by a sample sort and
 *                       only the bucket holding rank len / 2 is sorted;
 *                       2-byte keys sort every segment (valid for every k)
 *   key_type  "u16" "i16" "f16" "u32" "i32" "f32" "u64" "i64" "f64"
 *   small_max segments up to this length use one LDS workgroup each
 *             (0 = the largest the type allows); range_tiles: downsweep
 *             tiles per workgroup in larger segments (0 = 8).
 * 4- and 8-byte keys: segments up to 512 x items keys (8-byte keys: 4096)
 * are sorted by one LDS workgroup each; longer ones by a sample sort (splitters from a sorted
 * sample of the segment, one histogram read and one scatter into buckets,
 * every bucket sorted in LDS).  Plan knobs, read at creation
 * (xss_gpu_set_tuning, 0 = default): "seg_bucket_keys" average bucket
 * (default 2/3 of the LDS capacity), "seg_oversample" samples per bucket
 * (32), "seg_local_items" keys per LDS thread, 12 or 15 for 4-byte keys (12),
 * 8 or 12 for 8-byte keys (12; argsort: 8 or 11, default 11), "seg_wide_tiles" 1024-thread scatter tiles (0),
 * "seg_sample_off" 1 = the LSD plan (as 2-byte keys take) for every width.
 */
typedef struct xss_segplan *xss_segplan_t;
XSS_API int xss_gpu_segplan_create(xss_segplan_t *plan, const char *op, const char *key_type,
                                   const uint64_t *offsets, size_t nseg, uint32_t small_max, uint32_t range_tiles);
XSS_API int xss_gpu_segplan_run(xss_segplan_t plan, void *keys, uint32_t *idx_out, int hasnan, int descending,
                                xss_stream_t stream);
XSS_API int xss_gpu_segplan_info(xss_segplan_t plan, uint32_t *nsmall, uint32_t *nranges, uint32_t *nlarge);
XSS_API int xss_gpu_segplan_destroy(xss_segplan_t plan);
/* Baselines: one plain call per segment, enqueued in a C++ loop. */
XSS_API int xss_gpu_loop_qsort_f32(float *keys, const uint64_t *offsets, size_t nseg, int hasnan, int descending,
                                   xss_stream_t stream);
XSS_API int xss_gpu_loop_argsort_i32(const int32_t *keys, uint64_t *idx, const uint64_t *offsets, size_t nseg,
                                     xss_stream_t stream);
XSS_API int xss_gpu_loop_qselect_f32(float *keys, const uint64_t *offsets, size_t nseg, int hasnan,
                                     xss_stream_t stream);

/*
 * Serving form (no counterpart in the reference): a continuous-batching
 * request loop.  A worker thread owns `nstreams` HIP streams, admits queued
 * requests in arrival order while a stream is free and the admitted HBM
 * footprint (keys + values + scratch) stays within budget_bytes, and
 * coalesces queued qsort / keyvalue_qsort requests of at most coalesce_max
 * keys (capped at 2^24) with the same op, key type, value width and flags
 * into one launch group: a pointer-list LDS sort for the short ones and
 * shared segmented passes (gathered into a per-stream arena and scattered
 * back) for the others, up to 2^24 keys and 1024 requests per group.  Every
 * other request runs through its entry point above on its own stream, so
 * requests overlap on the device.  coalesce_max = 0 turns coalescing off;
 * otherwise each stream reserves an arena of 512 MiB (outside budget_bytes)
 * at creation.  reserve_bytes grows each stream's pooled workspace up
 * front.  one_at_a_time = 1 serves requests strictly one by one (the
 * plain-call baseline).  Requests work in place on the caller's device
 * buffers, which must stay valid until the request is done.
 */
typedef struct xss_request {
    int op;            /* 0 qsort, 1 keyvalue_qsort, 2 qselect, 3 partial_qsort */
    int key_type;      /* 0 u16, 1 i16, 2 f16, 3 u32, 4 i32, 5 f32, 6 u64, 7 i64, 8 f64 */
    int val_type;      /* keyvalue_qsort: value type, same codes */
    int hasnan;
    int descending;
    int reserved;
    void *keys;        /* device pointers */
    void *vals;
    uint64_t n;
    uint64_t k;        /* qselect / partial_qsort */
    double arrival_us; /* xss_server_run_trace: submit time after the call starts */
    double latency_us; /* out: submission to observed completion, -1 on failure */
    void *internal;
} xss_request;
typedef struct xss_server *xss_server_t;
XSS_API int xss_server_create(xss_server_t *server, size_t budget_bytes, int nstreams, uint32_t coalesce_max,
                              size_t reserve_bytes, int one_at_a_time);
XSS_API int xss_server_submit(xss_server_t server, xss_request *req);
XSS_API int xss_server_wait(xss_server_t server, xss_request *req);
/* Submits reqs[i] at arrival_us after the call starts (open loop), then
 * waits for all of them. */
XSS_API int xss_server_run_trace(xss_server_t server, xss_request *reqs, size_t count);
XSS_API int xss_server_stats(xss_server_t server, uint64_t *completed, uint64_t *launch_groups, uint64_t *coalesced);
XSS_API int xss_server_destroy(xss_server_t server);

/*
 * Host-pointer C symbols of the reference (lib/x86simdsort.cpp:298-339):
 * ascending key-value sort with hasnan = true, synchronous.
 */
XSS_API void keyvalue_qsort_float_uint32(float *key, uint32_t *val, size_t size);
XSS_API void keyvalue_qsort_float_uint64(float *key, uint64_t *val, size_t size);
XSS_API void keyvalue_qsort_uint64_uint32(uint64_t *key, uint32_t *val, size_t size);
XSS_API void keyvalue_qsort_uint64_uint64(uint64_t *key, uint64_t *val, size_t size);
XSS_API void keyvalue_qsort_int32_uint32(int32_t *key, uint32_t *val, size_t size);
XSS_API void keyvalue_qsort_int32_uint64(int32_t *key, uint64_t *val, size_t size);
XSS_API void keyvalue_qsort_uint32_uint32(uint32_t *key, uint32_t *val, size_t size);
XSS_API void keyvalue_qsort_uint32_uint64(uint32_t *key, uint64_t *val, size_t size);

#ifdef __cplusplus
}
#endif
#endif /* XSS_GPU_H */
