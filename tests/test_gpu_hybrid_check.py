"""The sampled spread check in front of the MSD-first sort (on by default):
spread data takes the MSD-first path, skewed or duplicate-heavy data the LSD
path, and both sort correctly against the oracle."""
import ctypes

import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dist", ["random", "dups", "constant", "nan"])
def test_sampled_check_routes_and_sorts(xss, dist):
    """With the sampled spread check on (the default), spread data takes the
    MSD-first path and skewed data the LSD path; both sort correctly."""
    L = xss.lib()
    L.xss_gpu_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
    L.xss_gpu_timing_collect.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]
    n = 1 << 21
    a = T.gen("f32", dist, n, 321)
    exp = T.ORACLE().qsort(a, "f32", dist == "nan", False)
    assert L.xss_gpu_set_tuning(b"hybrid_min_n", 1 << 20) == 0
    assert L.xss_gpu_set_tuning(b"hybrid_check", 1) == 0
    try:
        d = xss.DArray.from_numpy(a)
        names = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_double * 32)()
        cnt = (ctypes.c_uint64 * 32)()
        L.xss_gpu_timing_collect(32, names, ms, cnt)  # clear
        L.xss_gpu_timing_enable(1)
        xss.qsort(d, dist == "nan", False)
        got = d.numpy()
        L.xss_gpu_timing_enable(0)
        k = L.xss_gpu_timing_collect(32, names, ms, cnt)
        kernels = {names[i].decode() for i in range(k)}
    finally:
        L.xss_gpu_set_tuning(b"hybrid_min_n", 1 << 25)
    T.assert_sorted_equal(got, exp, "f32", f"qsort<f32> {dist} n=2^21 with the sampled check")
    took_hybrid = "hyb_local" in kernels
    assert took_hybrid == (dist in ("random", "nan")), (dist, sorted(kernels))  # 1% NaN: one shared bucket
