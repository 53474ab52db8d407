"""GPU parity of qselect / partial_qsort / argselect / keyvalue_select /
keyvalue_partial_sort against the oracle (partition property for selects,
bit-exact sorted prefix for partial sorts, as tests/test-qsort-common.h:52-104
and tests/test-keyvalue.cpp:102-171 check the reference)."""
import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

SIZES = [2, 3, 100, 4096, 4097, 20001, 300001]
DISTS = ["random", "dups", "constant", "signed_zeros", "nan", "rand_with_max_and_nan", "nan_payloads", "sorted"]


def ks(n):
    return sorted({0, 1 % n, n // 2, n - 1, (7 * n) // 10})


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", list(T.TAGS))
def test_qselect(xss, tag, desc):
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in DISTS for n in SIZES):
        a = T.gen(tag, dist, n, 5000 + seed)
        hasnan = dist in T.NAN_DISTS
        srt = O.qselect(a, tag, 0, hasnan, desc)  # the oracle's select is a full sort
        for k in ks(n):
            d = xss.DArray.from_numpy(a)
            xss.qselect(d, k, hasnan, desc)
            T.assert_selected(d.numpy(), a, srt, k, tag, f"qselect<{tag}> {dist} n={n} k={k} desc={desc}")


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", list(T.TAGS))
def test_partial_qsort(xss, tag, desc):
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in DISTS for n in SIZES):
        a = T.gen(tag, dist, n, 6000 + seed)
        hasnan = dist in T.NAN_DISTS
        for k in sorted({1, 2, min(n, 1000), max(1, n // 100), n}):
            exp = O.partial_qsort(a, tag, k, hasnan, desc)
            d = xss.DArray.from_numpy(a)
            xss.partial_qsort(d, k, hasnan, desc)
            got = d.numpy()
            what = f"partial_qsort<{tag}> {dist} n={n} k={k} desc={desc}"
            T.assert_prefix_equal(got, exp, k, tag, what)
            assert np.array_equal(np.sort(T.vclass(got, tag)), np.sort(T.vclass(exp, tag))), what


@pytest.mark.parametrize("tag", list(T.TAGS))
def test_argselect(xss, tag):
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in DISTS for n in SIZES):
        a = T.gen(tag, dist, n, 7000 + seed)
        hasnan = dist in T.NAN_DISTS
        srt = a[O.argsort(a, tag, hasnan, False).astype(np.int64)]
        d = xss.DArray.from_numpy(a)
        for k in ks(n):
            idx = xss.argselect(d, k, hasnan).numpy()
            what = f"argselect<{tag}> {dist} n={n} k={k}"
            assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint64)), what + ": not a permutation"
            T.assert_selected(a[idx.astype(np.int64)], a, srt, k, tag, what)
        assert np.array_equal(T.bits(d.numpy()), T.bits(a)), "argselect modified its input"


# the 36 pairs (lib/x86simdsort.cpp:280-293), then 16-bit keys with the six value types
KV = [(k, v) for k in T.KV_TAGS + T.KV16_TAGS for v in T.KV_TAGS]
KV_DISTS = ["random", "dups", "constant", "signed_zeros", "nan", "rand_with_max_and_nan", "sorted"]


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("ktag,vtag", KV)
def test_keyvalue_select_partial(xss, ktag, vtag, desc):
    """keyvalue_select: the k-th key and the partition property, pairs kept;
    keyvalue_partial_sort: the sorted prefix with its payloads checked as
    multisets per run of equal keys (tests/test-keyvalue.cpp:102-171)."""
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in KV_DISTS for n in [3, 4097, 30011]):
        k = T.gen(ktag, dist, n, 8000 + seed)
        v = T.gen(vtag, "random", n, 9000 + seed) if seed % 2 else np.arange(n).astype(T.TAGS[vtag])
        hasnan = dist in T.NAN_DISTS
        ek, ev = O.keyvalue_qsort(k, v, ktag, vtag, hasnan, desc)
        for kk in ks(n):
            dk, dv = xss.DArray.from_numpy(k), xss.DArray.from_numpy(v)
            xss.keyvalue_select(dk, dv, kk, hasnan, desc)
            gk, gv = dk.numpy(), dv.numpy()
            what = f"kvselect<{ktag},{vtag}> {dist} n={n} k={kk} desc={desc}"
            T.assert_selected(gk, k, ek, kk, ktag, what)
            T.assert_pairs_preserved(gk, gv, k, v, what)
        for kk in sorted({1, 2, min(n, 1000), n // 2, n}):
            dk, dv = xss.DArray.from_numpy(k), xss.DArray.from_numpy(v)
            xss.keyvalue_partial_sort(dk, dv, kk, hasnan, desc)
            gk, gv = dk.numpy(), dv.numpy()
            what = f"kvpartial<{ktag},{vtag}> {dist} n={n} k={kk} desc={desc}"
            T.assert_kv_partial(gk, gv, kk, k, v, ek, ev, ktag, what)
