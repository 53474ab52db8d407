// Test library for include/x86simdsort-gpu.h: x86simdsort::gpu::object_qsort
// instantiated by this translation unit (built with hipcc, as a caller would)
// with device key functors of each key type.  Python tests call these
// entry points on device buffers and check the results on the host.
#include "x86simdsort-gpu.h"

#include <math.h>

namespace {

struct Point3 {
    double x, y, z;
};

// 20-byte record: objects of a size that is not a multiple of 8
struct Rec {
    uint32_t id;
    float score;
    int32_t rank;
    uint32_t a, b;
};

struct ByScore {  // functor with a device call operator
    __device__ float operator()(const Rec &r) const { return r.score; }
};

}  // namespace

extern "C" {

// {x, y, z} by the Euclidean norm (double key), as the reference's benchmark
int objtest_xyz_norm(double *d_xyz, size_t n, hipStream_t s)
{
    return x86simdsort::gpu::object_qsort(reinterpret_cast<Point3 *>(d_xyz), n,
                                          [] __device__(const Point3 &p) {
                                              return sqrt((p.x * p.x + p.y * p.y) + p.z * p.z);
                                          },
                                          s);
}

// records by score (float key, functor), by rank (int32), by a combined
// unsigned 64-bit key, and by id read through a 32-bit count
int objtest_rec_score(void *d, size_t n, hipStream_t s)
{
    return x86simdsort::gpu::object_qsort(static_cast<Rec *>(d), n, ByScore{}, s);
}

int objtest_rec_rank(void *d, size_t n, hipStream_t s)
{
    return x86simdsort::gpu::object_qsort(static_cast<Rec *>(d), int64_t(n),
                                          [] __device__(Rec r) { return r.rank; }, s);
}

int objtest_rec_ab(void *d, size_t n, hipStream_t s)
{
    return x86simdsort::gpu::object_qsort(static_cast<Rec *>(d), n,
                                          [] __device__(const Rec &r) { return (uint64_t(r.a) << 32) | r.b; }, s);
}

int objtest_rec_id(void *d, uint32_t n, hipStream_t s)
{
    return x86simdsort::gpu::object_qsort(static_cast<Rec *>(d), n,
                                          [] __device__(const Rec &r) { return r.id; }, s);
}

}  // extern "C"
