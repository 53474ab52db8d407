"""GPU parity of the MSD-first sort path (two passes over HBM, then every
bucket of the top 16 bits in LDS; buckets too large for LDS take segmented
passes) for qsort, keyvalue_qsort and argsort.  By default only arrays of
2^25+ keys take it; here the switch xss_gpu_set_tuning("hybrid_min_n", 2)
sends every 4- and 8-byte sort through it, so small seeded arrays exercise
empty, single-key, LDS-sized and large buckets (duplicates, constants, NaN
runs) against the oracle with the reference's rules (bit-exact keys; +-0 and
NaN runs by value; payloads per equal-key run; argsort as a permutation
gathering the sorted keys)."""
import ctypes

import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

TAGS = ["u32", "i32", "f32", "u64", "i64", "f64"]
SIZES = [2, 3, 100, 4096, 6144, 6145, 20001, 300001, 1 << 20]
DISTS = ["random", "dups", "constant", "sorted", "reverse", "smallrange", "signed_zeros", "nan",
         "rand_with_max_and_nan", "nan_payloads", "random_5d"]


@pytest.fixture(scope="module")
def hybrid_everywhere(xss):
    L = xss.lib()
    L.xss_gpu_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
    # every size, and no sampled fallback to the LSD path
    assert L.xss_gpu_set_tuning(b"hybrid_min_n", 2) == 0
    assert L.xss_gpu_set_tuning(b"hybrid_check", 0) == 0
    yield xss
    assert L.xss_gpu_set_tuning(b"hybrid_min_n", 1 << 25) == 0
    assert L.xss_gpu_set_tuning(b"hybrid_check", 1) == 0


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", TAGS)
def test_hybrid_qsort(hybrid_everywhere, tag, desc):
    xss = hybrid_everywhere
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in DISTS for n in SIZES):
        a = T.gen(tag, dist, n, 9000 + seed)
        hasnan = dist in T.NAN_DISTS
        exp = O.qsort(a, tag, hasnan, desc)
        d = xss.DArray.from_numpy(a)
        xss.qsort(d, hasnan, desc)
        T.assert_sorted_equal(d.numpy(), exp, tag, f"hybrid qsort<{tag}> {dist} n={n} desc={desc}")


def test_hybrid_unknown_tuning_is_rejected(xss):
    L = xss.lib()
    L.xss_gpu_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
    assert L.xss_gpu_set_tuning(b"no_such_switch", 1) != 0


KV_PAIRS = [("u32", "u32"), ("f32", "u64"), ("i32", "f64"), ("u64", "u64"), ("f64", "i32"), ("i64", "f32")]
KV_SIZES = [2, 33, 4096, 4097, 6145, 70001, 300001]
KV_DISTS = ["random", "dups", "nan_payloads", "signed_zeros", "constant", "rand_with_max_and_nan", "reverse"]


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("ktag,vtag", KV_PAIRS)
def test_hybrid_keyvalue_qsort(hybrid_everywhere, ktag, vtag, desc):
    xss = hybrid_everywhere
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in KV_DISTS for n in KV_SIZES):
        k = T.gen(ktag, dist, n, 9500 + seed)
        v = T.gen(vtag, "random", n, 9700 + seed)
        hasnan = dist in T.NAN_DISTS
        ek, ev = O.keyvalue_qsort(k, v, ktag, vtag, hasnan, desc)
        dk, dv = xss.DArray.from_numpy(k), xss.DArray.from_numpy(v)
        xss.keyvalue_qsort(dk, dv, hasnan, desc)
        T.assert_kv_equal(dk.numpy(), dv.numpy(), ek, ev, ktag, f"hybrid kv<{ktag},{vtag}> {dist} n={n} desc={desc}")


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", TAGS)
def test_hybrid_argsort(hybrid_everywhere, tag, desc):
    xss = hybrid_everywhere
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in DISTS for n in SIZES[:-1]):
        a = T.gen(tag, dist, n, 9900 + seed)
        hasnan = dist in T.NAN_DISTS
        exp = a[O.argsort(a, tag, hasnan, desc).astype(np.int64)]
        d = xss.DArray.from_numpy(a)
        idx = xss.argsort(d, hasnan, desc).numpy()
        T.assert_argsort(idx, a, exp, tag, f"hybrid argsort<{tag}> {dist} n={n} desc={desc}")
        assert np.array_equal(T.bits(d.numpy()), T.bits(a)), "argsort must leave the keys untouched"

