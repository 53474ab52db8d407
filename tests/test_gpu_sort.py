"""GPU parity of qsort / argsort / keyvalue_qsort against the oracle.

Inputs are seeded; every case is checked with the reference's own rules
(bit-exact keys; +-0 and NaN runs as multisets; payloads as multisets per
equal-key run; argsort as a permutation with equal gathered keys).  Sizes
cover the small-array LDS path, the tile boundaries of the onesweep path and
multi-tile arrays.
"""
import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

SMALL = [0, 1, 2, 3, 5, 17, 64, 100, 255, 1000, 2047, 2048, 2049, 4095, 4096]
LARGE = [4097, 6144, 6145, 8191, 8193, 65537, 300001]
HUGE = [1 << 20, (1 << 22) + 3]
HUGE_DISTS = ["random", "dups", "nan", "sorted"]


def cases(sizes, dists):
    for d in dists:
        for n in sizes:
            yield d, n


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", list(T.TAGS))
@pytest.mark.parametrize("sizes", ["small", "large", "huge"])
def test_qsort(xss, tag, desc, sizes):
    O = T.ORACLE()
    sz = {"small": SMALL, "large": LARGE, "huge": HUGE}[sizes]
    dists = T.DISTS if sizes != "huge" else HUGE_DISTS
    for seed, (dist, n) in enumerate(cases(sz, dists)):
        a = T.gen(tag, dist, n, 1000 + seed)
        hasnan = dist in T.NAN_DISTS
        exp = O.qsort(a, tag, hasnan, desc)
        d = xss.DArray.from_numpy(a)
        xss.qsort(d, hasnan, desc)
        T.assert_sorted_equal(d.numpy(), exp, tag, f"qsort<{tag}> {dist} n={n} desc={desc}")


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", list(T.TAGS))
@pytest.mark.parametrize("sizes", ["small", "large"])
def test_argsort(xss, tag, desc, sizes):
    O = T.ORACLE()
    sz = {"small": SMALL, "large": LARGE}[sizes]
    for seed, (dist, n) in enumerate(cases(sz, T.DISTS)):
        a = T.gen(tag, dist, n, 2000 + seed)
        hasnan = dist in T.NAN_DISTS
        exp = a[O.argsort(a, tag, hasnan, desc).astype(np.int64)]
        d = xss.DArray.from_numpy(a)
        idx = xss.argsort(d, hasnan, desc).numpy()
        T.assert_argsort(idx, a, exp, tag, f"argsort<{tag}> {dist} n={n} desc={desc}")
        # the keys are left untouched
        assert np.array_equal(T.bits(d.numpy()), T.bits(a))


# the 36 pairs of lib/x86simdsort.cpp:280-293, then 16-bit keys with the six value types
KV_PAIRS = [(k, v) for k in T.KV_TAGS + T.KV16_TAGS for v in T.KV_TAGS]


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("ktag,vtag", KV_PAIRS)
def test_keyvalue_qsort(xss, ktag, vtag, desc):
    O = T.ORACLE()
    rng = np.random.default_rng(7)
    for seed, (dist, n) in enumerate(cases([0, 1, 2, 33, 1000, 2048, 4096, 4097, 70001],
                                          ["random", "dups", "nan_payloads", "signed_zeros", "constant",
                                           "rand_with_max_and_nan", "reverse"])):
        k = T.gen(ktag, dist, n, 3000 + seed)
        v = T.gen(vtag, "random", n, 4000 + seed)
        hasnan = dist in T.NAN_DISTS
        ek, ev = O.keyvalue_qsort(k, v, ktag, vtag, hasnan, desc)
        dk, dv = xss.DArray.from_numpy(k), xss.DArray.from_numpy(v)
        xss.keyvalue_qsort(dk, dv, hasnan, desc)
        T.assert_kv_equal(dk.numpy(), dv.numpy(), ek, ev, ktag, f"kv<{ktag},{vtag}> {dist} n={n} desc={desc}")
    del rng


@pytest.mark.parametrize("tag,n", [("f32", 5000), ("f32", 10000), ("f64", 3000), ("f64", 10000), ("f32", 100000)])
def test_qsort_all_nan_payloads_canonical(xss, tag, n):
    """An array of NaNs with assorted payloads, hasnan: every key comes back
    as the canonical NaN, also where every order key is equal and no radix
    pass runs (the one-workgroup paths)."""
    rng = np.random.default_rng(n)
    u = T.TAGS[tag]
    bits = np.uint32 if tag == "f32" else np.uint64
    exp = np.uint32(0x7F800000) if tag == "f32" else np.uint64(0x7FF0000000000000)
    mant = rng.integers(1, 1 << 20, n).astype(bits)
    sign = (rng.integers(0, 2, n).astype(bits) << bits(31 if tag == "f32" else 63))
    keys = (exp | mant | sign).view(u)
    assert np.isnan(keys).all()
    d = xss.DArray.from_numpy(keys)
    xss.qsort(d, hasnan=True)
    got = T.bits(d.numpy())
    assert (got == T.QNAN[tag]).all(), f"{(got != T.QNAN[tag]).sum()} keys are not the canonical NaN"
