// Drop-in check: a program written against the reference API
// (lib/x86simdsort.h), built against include/x86simdsort.h and linked with
// libxss_amd.so.  Results are compared with std::sort / std::nth_element
// under the reference's NaN-aware order.  Prints "dropin ok" on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "x86simdsort.h"

#define CHECK(c)                                                     \
    do {                                                             \
        if (!(c)) {                                                  \
            std::printf("FAILED %s at line %d\n", #c, __LINE__);     \
            return 1;                                                \
        }                                                            \
    } while (0)

static bool nan_last_less(float a, float b)
{
    if (std::isnan(a)) return false;
    if (std::isnan(b)) return true;
    return a < b;
}

int main()
{
    std::mt19937 g(7);
    // qsort<float> with NaNs: canonical NaN at the end
    std::vector<float> a(300000);
    for (auto &x : a) x = std::uniform_real_distribution<float>(-1e6f, 1e6f)(g);
    for (int i = 0; i < 100; ++i) a[g() % a.size()] = std::nanf("");
    std::vector<float> ref = a;
    std::sort(ref.begin(), ref.end(), nan_last_less);
    x86simdsort::qsort(a.data(), a.size(), true);
    for (size_t i = 0; i < a.size(); ++i) {
        if (std::isnan(ref[i])) {
            uint32_t bits;
            std::memcpy(&bits, &a[i], 4);
            CHECK(bits == 0x7fc00000u);
        } else {
            CHECK(a[i] == ref[i]);
        }
    }
    // descending uint64
    std::vector<uint64_t> u(100000);
    for (auto &x : u) x = (uint64_t(g()) << 32) | g();
    std::vector<uint64_t> ur = u;
    std::sort(ur.begin(), ur.end(), std::greater<uint64_t>());
    x86simdsort::qsort(u.data(), u.size(), false, true);
    CHECK(u == ur);
    // argsort<int32_t>
    std::vector<int32_t> s(50000);
    for (auto &x : s) x = int32_t(g());
    std::vector<size_t> arg = x86simdsort::argsort(s.data(), s.size());
    CHECK(arg.size() == s.size());
    for (size_t i = 1; i < arg.size(); ++i) CHECK(s[arg[i - 1]] <= s[arg[i]]);
    // qselect<double>
    std::vector<double> d(77777);
    for (auto &x : d) x = std::uniform_real_distribution<double>(-1, 1)(g);
    std::vector<double> dr = d;
    const size_t k = 31337;
    std::nth_element(dr.begin(), dr.begin() + k, dr.end());
    x86simdsort::qselect(d.data(), k, d.size());
    CHECK(d[k] == dr[k]);
    for (size_t i = 0; i < k; ++i) CHECK(d[i] <= d[k]);
    for (size_t i = k + 1; i < d.size(); ++i) CHECK(d[i] >= d[k]);
    // partial_qsort<uint16_t>
    std::vector<uint16_t> h(40000);
    for (auto &x : h) x = uint16_t(g());
    std::vector<uint16_t> hr = h;
    std::sort(hr.begin(), hr.end());
    x86simdsort::partial_qsort(h.data(), 1000, h.size());
    CHECK(std::equal(h.begin(), h.begin() + 1000, hr.begin()));
    // keyvalue_qsort<float, uint64_t> via the C symbol (hasnan = true)
    std::vector<float> kk(60000);
    std::vector<uint64_t> vv(kk.size());
    for (size_t i = 0; i < kk.size(); ++i) {
        kk[i] = float(g() % 1000);
        vv[i] = i;
    }
    std::vector<float> k0 = kk;
    keyvalue_qsort_float_uint64(kk.data(), vv.data(), kk.size());
    for (size_t i = 0; i < kk.size(); ++i) CHECK(k0[vv[i]] == kk[i]);
    for (size_t i = 1; i < kk.size(); ++i) CHECK(kk[i - 1] <= kk[i]);
    // object_qsort by Euclidean norm
    struct P { double x, y, z; };
    std::vector<P> pts(20000);
    for (auto &p : pts) {
        p.x = std::uniform_real_distribution<double>(-1, 1)(g);
        p.y = std::uniform_real_distribution<double>(-1, 1)(g);
        p.z = std::uniform_real_distribution<double>(-1, 1)(g);
    }
    auto norm = [](P p) { return std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z); };
    x86simdsort::object_qsort(pts.data(), pts.size(), norm);
    for (size_t i = 1; i < pts.size(); ++i) CHECK(norm(pts[i - 1]) <= norm(pts[i]));
    std::printf("dropin ok\n");
    return 0;
}
