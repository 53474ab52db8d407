"""GPU parity of the MSD-first sort path (two passes over HBM, then every
bucket of the two most significant varying digits in LDS; buckets too large
for LDS take segmented passes, buckets of equal keys none) for qsort,
keyvalue_qsort and argsort.  By default only arrays of
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
    assert L.xss_gpu_set_tuning(b"hybrid_min_n", 2) == 0  # every size
    yield xss
    assert L.xss_gpu_set_tuning(b"hybrid_min_n", 1 << 25) == 0


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



def narrow(tag, n, bits_, seed):
    """Keys whose order keys differ only in their low `bits_` bits: the device
    plan then starts the MSD passes lower (64-bit keys below 2^32 sort like
    32-bit keys)."""
    rng = np.random.default_rng(seed)
    nb = np.dtype(T.TAGS[tag]).itemsize * 8
    u = T.UINT[nb // 8]
    # non-negative values below 2^bits_ (for signed types the order key is
    # bits ^ sign, whose top bits are then constant too)
    x = rng.integers(0, 2 ** min(bits_, nb - 1), n, dtype=np.uint64).astype(u)
    return x.view(T.TAGS[tag])


@pytest.mark.parametrize("tag,bits_", [("u64", 32), ("u64", 40), ("u64", 24), ("i64", 33), ("u64", 12),
                                       ("u64", 7), ("u32", 24), ("u32", 16), ("i32", 9), ("u32", 0)])
@pytest.mark.parametrize("op", ["qsort", "kv", "argsort"])
def test_hybrid_narrow_keys(hybrid_everywhere, tag, bits_, op):
    """Keys with constant top digits: the plan made on the device skips them
    (in pairs), including the case with no low digit left and with a single
    value (bits_ = 0)."""
    xss = hybrid_everywhere
    O = T.ORACLE()
    for seed, n in enumerate([3, 5000, 70001, 1 << 20]):
        a = narrow(tag, n, bits_, 100 + seed)
        what = f"narrow {op}<{tag}> bits={bits_} n={n}"
        for desc in (False, True):
            d = xss.DArray.from_numpy(a)
            if op == "qsort":
                xss.qsort(d, False, desc)
                T.assert_sorted_equal(d.numpy(), O.qsort(a, tag, False, desc), tag, what)
            elif op == "kv":
                v = np.arange(n, dtype=np.uint64)
                dv = xss.DArray.from_numpy(v)
                xss.keyvalue_qsort(d, dv, False, desc)
                ek, ev = O.keyvalue_qsort(a, v, tag, "u64", False, desc)
                T.assert_kv_equal(d.numpy(), dv.numpy(), ek, ev, tag, what)
            else:
                idx = xss.argsort(d, False, desc).numpy()
                T.assert_argsort(idx, a, a[O.argsort(a, tag, False, desc).astype(np.int64)], tag, what)
