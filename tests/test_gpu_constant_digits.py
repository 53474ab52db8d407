"""GPU parity of LSD sorts that skip passes over a digit every key shares
(in-place qsort and keyvalue_qsort): keys confined to narrow ranges so that
1, 2, 3 or 4 of the digits vary, which leaves the keys after an odd or even
number of executed passes in the scratch or the caller's buffer.  The
arrays stay below the MSD-first size, so the LSD path runs; checked against
the oracle with the reference's rules (bit-exact keys, payload multisets
per equal-key run)."""
import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

SIZES = [5000, 100003, (1 << 20) + 7]
BITS = [8, 16, 24, 32]  # varying low bits of the key


def narrow(tag, n, bits, seed):
    rng = np.random.default_rng(seed)
    dt = T.TAGS[tag]
    if tag in ("f32", "f64"):
        # positive floats sharing sign, exponent and the top mantissa bits
        u = np.uint32 if tag == "f32" else np.uint64
        base = u(0x3F800000) if tag == "f32" else u(0x3FF0000000000000)
        low = rng.integers(0, 1 << min(bits, 23 if tag == "f32" else 52), n, dtype=np.uint64).astype(u)
        return (base | low).view(dt)
    lo = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    if tag.startswith("i"):
        return (lo.astype(np.int64) - (1 << (bits - 1))).astype(dt)
    return lo.astype(dt)


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", ["u32", "i32", "f32", "u64", "i64", "f64"])
def test_qsort_narrow_keys(xss, tag, desc):
    O = T.ORACLE()
    for seed, (bits, n) in enumerate((b, n) for b in BITS for n in SIZES):
        a = narrow(tag, n, bits, 17000 + seed)
        exp = O.qsort(a, tag, False, desc)
        d = xss.DArray.from_numpy(a)
        xss.qsort(d, False, desc)
        T.assert_sorted_equal(d.numpy(), exp, tag, f"qsort<{tag}> {bits} varying bits n={n} desc={desc}")


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("ktag,vtag", [("u64", "u64"), ("u32", "u32"), ("i64", "f32"), ("f32", "u64")])
def test_keyvalue_narrow_keys(xss, ktag, vtag, desc):
    O = T.ORACLE()
    for seed, (bits, n) in enumerate((b, n) for b in BITS for n in SIZES):
        k = narrow(ktag, n, bits, 18000 + seed)
        v = T.gen(vtag, "random", n, 18500 + seed)
        ek, ev = O.keyvalue_qsort(k, v, ktag, vtag, False, desc)
        dk, dv = xss.DArray.from_numpy(k), xss.DArray.from_numpy(v)
        xss.keyvalue_qsort(dk, dv, False, desc)
        T.assert_kv_equal(dk.numpy(), dv.numpy(), ek, ev, ktag,
                          f"kv<{ktag},{vtag}> {bits} varying bits n={n} desc={desc}")


@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_qsort_constant_and_nan_keeps_canonical_nan(xss, tag):
    """Every digit constant except where NaNs sit: the last pass still runs
    and writes the canonical NaN."""
    O = T.ORACLE()
    for seed, n in enumerate(SIZES):
        a = np.full(n, 1.5, T.TAGS[tag])
        rng = np.random.default_rng(19000 + seed)
        a[rng.integers(0, n, n // 100)] = np.nan
        for desc in (False, True):
            exp = O.qsort(a, tag, True, desc)
            d = xss.DArray.from_numpy(a)
            xss.qsort(d, True, desc)
            T.assert_sorted_equal(d.numpy(), exp, tag, f"qsort<{tag}> constant+NaN n={n} desc={desc}")
        c = np.full(n, -7, T.TAGS[tag])
        d = xss.DArray.from_numpy(c)
        xss.qsort(d, True, False)
        assert np.array_equal(T.bits(d.numpy()), T.bits(c)), "constant array changed"
