"""Parity at the full workload sizes: 2^28 32-bit keys, 2^27 64-bit keys,
2^29 16-bit keys (1 GiB of keys each), the five input recipes (uniform over
the full bit range, 1024 distinct values, sorted, reverse sorted, 1% NaN with
hasnan), ascending and descending.

At these sizes the oracle's single-threaded restatement takes minutes, so
results are checked through properties that fix a sort's output:
  * sortedness under the reference's order (tests/test-qsort-common.h:29-36),
    every NaN one class placed last (first when descending), -0.0 == +0.0;
  * equality of the output's value classes with the input's, sorted (a
    permutation), and the same count of each zero's bit pattern;
  * qsort(hasnan) writes only the canonical NaN (src/xss-custom-float.h:34-38,81-85);
  * argsort / keyvalue_qsort with values = original index: the indices are a
    permutation and gather the output keys from the input;
  * select / partial sort: the k-th value equals numpy's partition of the
    value classes, with the partition property around it.
Where the reference library was built (oracle/_ref), the 32-bit sorts are
also compared with the reference's own output on the same input.
"""
import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

N_OF = {2: 1 << 29, 4: 1 << 28, 8: 1 << 27}
RECIPES = ["random", "dups", "sorted", "reverse", "nan"]


def n_for(tag):
    return N_OF[np.dtype(T.TAGS[tag]).itemsize]


def ordv(a, tag, desc=False):
    """Order key of the reference's comparator as unsigned integers: value
    classes (every NaN one class above +inf, -0 == +0) mapped monotonically."""
    dt = np.dtype(T.TAGS[tag])
    u = T.UINT[dt.itemsize]
    nb = dt.itemsize * 8
    sign = u(1) << u(nb - 1)
    b = T.bits(np.ascontiguousarray(a)).copy()
    if tag in T.FLOAT_TAGS:
        b[b == sign] = 0                          # -0 -> +0
        nan = T.is_nan(a, tag)
        neg = (b & sign) != 0
        b = np.where(neg, ~b, b | sign)
        b[nan] = u(~u(0))                         # all NaNs above +inf
    elif tag.startswith("i"):
        b ^= sign
    if desc:
        b = ~b
    return b


def check_sort(got, inp, tag, hasnan, desc, canon_nan, what):
    og = ordv(got, tag, desc)
    assert (og[1:] >= og[:-1]).all(), f"{what}: not sorted"
    assert np.array_equal(og, np.sort(ordv(inp, tag, desc))), f"{what}: not a permutation of the input"
    if tag in T.FLOAT_TAGS:
        nb = got.dtype.itemsize * 8
        negz = T.UINT[got.dtype.itemsize](1) << T.UINT[got.dtype.itemsize](nb - 1)
        assert (T.bits(got) == negz).sum() == (T.bits(inp) == negz).sum(), f"{what}: -0.0 count differs"
        if hasnan and canon_nan:
            nanbits = T.bits(got)[T.is_nan(got, tag)]
            assert (nanbits == T.QNAN[tag]).all(), f"{what}: NaN not canonical"


def gather_ok(keys_out, inp, idx, what):
    n = inp.size
    idx = idx.astype(np.int64)
    assert idx.min() >= 0 and idx.max() < n, f"{what}: index out of range"
    seen = np.zeros(n, bool)
    seen[idx] = True
    assert seen.all(), f"{what}: indices are not a permutation"
    assert np.array_equal(T.bits(inp[idx]), T.bits(keys_out)), f"{what}: indices do not gather the keys"


CASES_32 = [("f32", r) for r in RECIPES] + [("u32", r) for r in RECIPES[:4]] + [("i32", "random")]
CASES_64_16 = [(t, "random") for t in ("f64", "u64", "i64", "f16", "u16", "i16")] + [("f64", "nan"), ("f16", "nan")]


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag,recipe", CASES_32)
def test_qsort_32bit_full(xss, tag, recipe, desc):
    n = n_for(tag)
    inp = T.gen(tag, recipe, n, 4242)
    hasnan = recipe == "nan"
    d = xss.DArray.from_numpy(inp)
    xss.qsort(d, hasnan, desc)
    got = d.numpy()
    del d
    what = f"qsort<{tag}> {recipe} n=2^28 desc={desc}"
    check_sort(got, inp, tag, hasnan, desc, True, what)
    R = T.REF()
    if R is not None:
        exp = R.qsort(inp, tag, hasnan, desc)
        if not np.array_equal(T.bits(got), T.bits(exp)):
            T.assert_sorted_equal(got, exp, tag, what + " vs reference", zeros_as_values=True)


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag,recipe", CASES_64_16)
def test_qsort_64_16bit_full(xss, tag, recipe, desc):
    n = n_for(tag)
    inp = T.gen(tag, recipe, n, 4343)
    hasnan = recipe == "nan"
    d = xss.DArray.from_numpy(inp)
    xss.qsort(d, hasnan, desc)
    got = d.numpy()
    del d
    check_sort(got, inp, tag, hasnan, desc, True, f"qsort<{tag}> {recipe} n={n} desc={desc}")


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", ["f32", "u64", "f16"])
def test_argsort_full(xss, tag, desc):
    n = n_for(tag)
    inp = T.gen(tag, "nan" if tag in T.FLOAT_TAGS else "dups", n, 4444)
    hasnan = tag in T.FLOAT_TAGS
    d = xss.DArray.from_numpy(inp)
    idx = xss.argsort(d, hasnan, desc).numpy()
    del d
    what = f"argsort<{tag}> n={n} desc={desc}"
    keys = inp[idx.astype(np.int64)]
    gather_ok(keys, inp, idx, what)
    check_sort(keys, inp, tag, hasnan, desc, False, what)


def zipf_keys(n, seed):
    rng = np.random.default_rng(seed)
    z = rng.zipf(1.1, n).astype(np.uint64)
    bad = z >= (1 << 32)
    while bad.any():
        z[bad] = rng.zipf(1.1, int(bad.sum())).astype(np.uint64)
        bad = z >= (1 << 32)
    return z


@pytest.mark.parametrize("desc", [False, True])
def test_keyvalue_u64_zipf_full(xss, desc):
    n = 1 << 27
    k = zipf_keys(n, 11)
    v = np.arange(n, dtype=np.uint64)
    dk, dv = xss.DArray.from_numpy(k), xss.DArray.from_numpy(v)
    xss.keyvalue_qsort(dk, dv, False, desc)
    gk, gv = dk.numpy(), dv.numpy()
    del dk, dv
    what = f"keyvalue_qsort<u64,u64> zipf n=2^27 desc={desc}"
    gather_ok(gk, k, gv, what)
    check_sort(gk, k, "u64", False, desc, False, what)


@pytest.mark.parametrize("desc", [False, True])
def test_keyvalue_f32_normal_full(xss, desc):
    n = 1 << 28
    rng = np.random.default_rng(12)
    k = rng.standard_normal(n, dtype=np.float32)
    v = np.arange(n, dtype=np.uint32)
    dk, dv = xss.DArray.from_numpy(k), xss.DArray.from_numpy(v)
    xss.keyvalue_qsort(dk, dv, False, desc)
    gk, gv = dk.numpy(), dv.numpy()
    del dk, dv
    what = f"keyvalue_qsort<f32,u32> normal n=2^28 desc={desc}"
    gather_ok(gk, k, gv, what)
    check_sort(gk, k, "f32", False, desc, False, what)


@pytest.mark.parametrize("tag", ["f32", "u64"])
def test_select_partial_full(xss, tag):
    n = n_for(tag)
    inp = T.gen(tag, "nan" if tag == "f32" else "random", n, 4545)
    hasnan = tag == "f32"
    o_in = ordv(inp, tag)
    for k in (n // 2, n - 1):
        d = xss.DArray.from_numpy(inp)
        xss.qselect(d, k, hasnan)
        got = d.numpy()
        del d
        what = f"qselect<{tag}> n={n} k={k}"
        og = ordv(got, tag)
        kth = np.partition(o_in, k)[k]
        assert og[k] == kth, f"{what}: k-th value differs"
        assert (og[:k] <= kth).all() and (og[k + 1:] >= kth).all(), f"{what}: partition property"
        assert np.array_equal(np.sort(T.bits(got)), np.sort(T.bits(inp))), f"{what}: not a permutation"
    for k in (1000, n // 100):
        d = xss.DArray.from_numpy(inp)
        xss.partial_qsort(d, k, hasnan)
        got = d.numpy()
        del d
        what = f"partial_qsort<{tag}> n={n} k={k}"
        og = ordv(got, tag)
        exp = np.sort(np.partition(o_in, k - 1)[:k])
        assert np.array_equal(og[:k], exp), f"{what}: sorted prefix differs"
        assert (og[k:] >= og[k - 1]).all(), f"{what}: partition property"
        assert np.array_equal(np.sort(T.bits(got)), np.sort(T.bits(inp))), f"{what}: not a permutation"
