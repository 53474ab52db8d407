// Host API from many threads at once, and its error channel.
//
// Eight threads call the drop-in x86simdsort:: functions concurrently on
// their own arrays (each thread's calls run on its own stream and pooled
// workspace) and check every result against std::sort; then a workspace
// limit makes a call fail, which the C++ API reports by throwing
// std::runtime_error (the reference's functions return void).  Prints
// "threads ok" on success.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "x86simdsort.h"
#include "xss_gpu.h"

static void worker(int t, std::atomic<int> *bad)
{
    std::mt19937_64 g(1000 + t);
    for (int it = 0; it < 12; ++it) {
        const size_t n = 1000 + g() % 400000;
        std::vector<float> a(n);
        for (auto &x : a) x = std::uniform_real_distribution<float>(-1e3f, 1e3f)(g);
        std::vector<float> ref = a;
        std::sort(ref.begin(), ref.end());
        x86simdsort::qsort(a.data(), n, false, it % 2 == 1);
        if (it % 2 == 1) std::reverse(ref.begin(), ref.end());
        if (a != ref) ++*bad;

        std::vector<uint64_t> k(n);
        for (auto &x : k) x = g() % 5000;
        std::vector<uint32_t> v(n);
        std::iota(v.begin(), v.end(), 0u);
        const std::vector<uint64_t> k0 = k;
        x86simdsort::keyvalue_qsort(k.data(), v.data(), n);
        for (size_t i = 0; i < n; ++i)
            if (k0[v[i]] != k[i] || (i && k[i - 1] > k[i])) {
                ++*bad;
                break;
            }

        std::vector<double> d(n);
        for (auto &x : d) x = std::normal_distribution<double>()(g);
        std::vector<size_t> idx = x86simdsort::argsort(d.data(), n);
        for (size_t i = 1; i < n; ++i)
            if (d[idx[i - 1]] > d[idx[i]]) {
                ++*bad;
                break;
            }
    }
}

int main()
{
    std::atomic<int> bad{0};
    std::vector<std::thread> ts;
    for (int t = 0; t < 8; ++t) ts.emplace_back(worker, t, &bad);
    for (auto &t : ts) t.join();
    if (bad) {
        std::printf("FAILED: %d wrong results from concurrent calls\n", bad.load());
        return 1;
    }
    // a refused workspace (the limit stands in for an allocator that says
    // no): the call throws instead of returning unsorted data
    std::vector<float> big(1 << 20, 1.0f);
    big[7] = 0.0f;
    if (xss_gpu_set_tuning("workspace_limit", 1 << 20) != XSS_OK) return 1;
    bool threw = false;
    try {
        x86simdsort::qsort(big.data(), big.size());
    } catch (const std::runtime_error &e) {
        threw = true;
        std::printf("refused as expected: %s\n", e.what());
    }
    xss_gpu_set_tuning("workspace_limit", 0);
    if (!threw) {
        std::printf("FAILED: no exception for a refused workspace\n");
        return 1;
    }
    x86simdsort::qsort(big.data(), big.size());  // without the limit it sorts
    if (big[0] != 0.0f || big[1] != 1.0f) return 1;
    std::printf("threads ok\n");
    return 0;
}
