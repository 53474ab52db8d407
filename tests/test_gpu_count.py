"""GPU parity of the counting sort behind qsort of 16-bit keys (uint16_t,
int16_t, _Float16): per-value counts, then runs of each value.  By default
it serves arrays of 2^23+ keys; here xss_gpu_set_tuning("count_min_n", 2)
sends every 16-bit qsort through it, so small seeded arrays of every recipe
(constants and duplicates drive the bins' overflow counters, NaN runs the
canonical NaN) are checked against the oracle with the reference's rules."""
import ctypes

import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

SIZES = [2, 3, 17, 100, 4096, 4099, 65536, 300001, (1 << 20) + 5]


@pytest.fixture(scope="module")
def count_everywhere(xss):
    L = xss.lib()
    L.xss_gpu_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
    assert L.xss_gpu_set_tuning(b"count_min_n", 2) == 0
    yield xss
    assert L.xss_gpu_set_tuning(b"count_min_n", 1 << 23) == 0


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", ["u16", "i16", "f16"])
def test_count_qsort(count_everywhere, tag, desc):
    xss = count_everywhere
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in T.DISTS for n in SIZES):
        a = T.gen(tag, dist, n, 12000 + seed)
        hasnan = dist in T.NAN_DISTS
        exp = O.qsort(a, tag, hasnan, desc)
        d = xss.DArray.from_numpy(a)
        xss.qsort(d, hasnan, desc)
        T.assert_sorted_equal(d.numpy(), exp, tag, f"counting qsort<{tag}> {dist} n={n} desc={desc}")


def test_count_unaligned_subarray(count_everywhere):
    """A 16-bit array starting 2 bytes into an allocation (16-byte stores
    inside runs must follow the memory alignment, not the index)."""
    import numpy as np
    import torch
    xss = count_everywhere
    n = 300003
    a = T.gen("u16", "smallrange", n, 77)
    buf = torch.zeros(n + 8, dtype=torch.int16, device="cuda")
    buf[1:n + 1] = torch.from_numpy(a.view(np.int16)).cuda()
    L = xss.lib()
    rc = L.xss_gpu_qsort_u16(ctypes.c_void_p(buf.data_ptr() + 2), ctypes.c_size_t(n), 0, 0,
                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    got = buf.cpu().numpy().view(np.uint16)
    assert got[0] == 0 and (got[n + 1:] == 0).all(), "wrote outside the array"
    assert np.array_equal(got[1:n + 1], np.sort(a))
