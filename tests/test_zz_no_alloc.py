"""xss_gpu_qsort_f32_noalloc (static device scratch, no allocation) against
xss_gpu_qsort_f32 on the same inputs: bit-identical output on two calls in a
row, the second on a smaller array, and the size limit as an error."""
import ctypes

import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

NOALLOC_MAX_N = 1 << 22


def _run(xss, name, a, hasnan, desc):
    import torch
    d = xss.DArray.from_numpy(a)
    torch.cuda.synchronize()
    rc = getattr(xss.lib(), name)(d.ptr, ctypes.c_size_t(a.size), int(hasnan), int(desc),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, xss.lib().xss_gpu_last_error().decode()
    return d.numpy()


@pytest.mark.parametrize("case", [
    # (first n, first recipe, second n, second recipe, hasnan, descending)
    (NOALLOC_MAX_N, "nan", 100_003, "signed_zeros", True, False),
    (3_000_001, "dups", 4096, "nan_payloads", True, True),
    (1_000_000, "random", 5000, "sorted", False, False),
])
def test_noalloc_matches_qsort(xss, case):
    n1, d1, n2, d2, hasnan, desc = case
    O = T.ORACLE()
    for i, (n, dist) in enumerate(((n1, d1), (n2, d2))):  # two calls in a row, the second smaller
        a = T.gen("f32", dist, n, 4242 + i)
        print(f"noalloc case n={n} {dist} hasnan={hasnan} desc={desc}", flush=True)
        got = _run(xss, "xss_gpu_qsort_f32_noalloc", a, hasnan, desc)
        exp = _run(xss, "xss_gpu_qsort_f32", a, hasnan, desc)
        assert np.array_equal(T.bits(got), T.bits(exp)), f"noalloc != qsort_f32 (n={n} {dist})"
        if n <= 200_000:
            T.assert_sorted_equal(got, O.qsort(a, "f32", hasnan, desc), "f32", f"noalloc vs oracle n={n}")


def test_noalloc_limits(xss):
    import torch
    L = xss.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # sizes beyond the scratch are refused on the host, before any launch
    rc = L.xss_gpu_qsort_f32_noalloc(ctypes.c_void_p(0x1000), ctypes.c_size_t(NOALLOC_MAX_N + 1), 0, 0, s)
    assert rc == 1 and b"scratch limit" in L.xss_gpu_last_error()
    assert L.xss_gpu_qsort_f32_noalloc(None, ctypes.c_size_t(10), 0, 0, s) == 1
    assert L.xss_gpu_qsort_f32_noalloc(None, ctypes.c_size_t(1), 0, 0, s) == 0  # n <= 1: no-op
    # the library's pooled workspaces do not grow for it
    before = L.xss_gpu_workspace_bytes()
    a = T.gen("f32", "random", 123_457, 9)
    got = _run(xss, "xss_gpu_qsort_f32_noalloc", a, False, False)
    assert L.xss_gpu_workspace_bytes() == before
    T.assert_sorted_equal(got, T.ORACLE().qsort(a, "f32"), "f32", "noalloc 123457")
