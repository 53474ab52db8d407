"""CPU tests that pin the oracle (oracle/xss_oracle.c) before it is trusted:

  * the comparator known-answer table of the reference's own tests
    (tests/test-qsort.cpp:261-292);
  * golden fixtures written by the reference library itself
    (tests/golden/make_golden.py);
  * the reference library live, where oracle/_ref was built;
  * closed forms: every permutation of a small set with NaN, +-inf, +-0 and
    duplicates sorts to one hand-written answer; all 2^16 patterns of each
    16-bit type sort to the order of their values;
  * planted mistakes: the checkers must reject corrupted outputs.
"""
import itertools
import os

import numpy as np
import pytest

import xss_testlib as T

GOLD = os.path.join(T.REPO, "tests", "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ------------------------------------------------------------ known answers

def test_comparator_truth_table():
    # ASSERT_EQs of tests/test-qsort.cpp:271-290, op codes: 0 <, 1 >, 2 <=, 3 >=, 4 ==
    c = T.oracle().xss_oracle_compare_f64
    nan, inf = float("nan"), float("inf")
    table = [(0, nan, inf, 0), (0, nan, nan, 0), (0, inf, nan, 1), (0, inf, inf, 0),
             (2, nan, inf, 0), (2, nan, nan, 1), (2, inf, nan, 1), (2, inf, inf, 1),
             (3, nan, inf, 1), (3, nan, nan, 1), (3, inf, nan, 0), (3, inf, inf, 1),
             (1, nan, inf, 1), (1, nan, nan, 0), (1, inf, nan, 0), (1, inf, inf, 0),
             (4, nan, inf, 0), (4, nan, nan, 1), (4, inf, nan, 0), (4, inf, inf, 1)]
    for op, a, b, want in table:
        assert c(a, b, op) == want, (op, a, b)


def test_permutations_sort_to_one_answer():
    """Every order of {NaN, +inf, -inf, -2.5, -0.0, +0.0, 1, 1} sorts to the
    hand-written answer (NaN largest, canonical NaN bits with hasnan)."""
    O = T.ORACLE()
    s = np.array([np.nan, np.inf, -np.inf, -2.5, -0.0, 0.0, 1.0, 1.0], np.float32)
    asc = np.array([-np.inf, -2.5, 0.0, 0.0, 1.0, 1.0, np.inf, np.nan], np.float32)
    asc_nz = T.vclass(asc, "f32")
    for perm in itertools.permutations(range(8)):
        a = s[list(perm)]
        for desc in (False, True):
            out = O.qsort(a, "f32", True, desc)
            want = asc_nz[::-1] if desc else asc_nz
            assert np.array_equal(T.vclass(out, "f32"), want)
            nanpos = 0 if desc else 7
            assert T.bits(out)[nanpos] == T.QNAN["f32"]
            # the zeros keep their bits: one -0.0 and one +0.0
            assert sorted(T.bits(out)[[4, 5] if desc else [2, 3]]) == [0, 0x80000000]


@pytest.mark.parametrize("tag", ["u16", "i16", "f16"])
def test_all_16bit_patterns(tag):
    O = T.ORACLE()
    allp = np.arange(1 << 16, dtype=np.uint16).view(T.TAGS[tag])
    rng = np.random.default_rng(3)
    a = allp[rng.permutation(allp.size)]
    for desc in (False, True):
        out = O.qsort(a, tag, True, desc)
        if tag == "f16":
            isnan = T.is_nan(out, tag)
            assert isnan.sum() == 2046
            assert (isnan[:2046] if desc else isnan[-2046:]).all()
            assert (T.bits(out)[isnan] == T.QNAN["f16"]).all()
            vals = out[~isnan].astype(np.float64)
            d = np.diff(vals)
            assert ((d <= 0) if desc else (d >= 0)).all()
            assert np.array_equal(np.sort(T.bits(out)[~isnan]), np.sort(T.bits(allp)[~T.is_nan(allp, tag)]))
        else:
            want = np.sort(allp)
            assert np.array_equal(out, want[::-1] if desc else want)


# ------------------------------------------------------------ golden fixtures

def _parse(key):
    parts = key.split("|")
    return parts


def test_golden_sort():
    O = T.ORACLE()
    g = gold("golden_sort.npz")
    ins = [k for k in g.files if k.endswith("|in")]
    assert len(ins) > 200
    for k in ins:
        tag, dist, n, hasnan = _parse(k)[:4]
        base = k[: -len("|in")]
        a = g[k]
        for desc in (0, 1):
            T.assert_sorted_equal(O.qsort(a, tag, int(hasnan), desc), g[f"{base}|qsort{desc}"], tag,
                                  f"golden qsort {base} desc={desc}", zeros_as_values=True)
            ok = a[O.argsort(a, tag, int(hasnan), desc).astype(np.int64)]
            T.assert_sorted_equal(ok, g[f"{base}|argkeys{desc}"], tag, f"golden argsort {base}", zeros_as_values=True)


def test_golden_select():
    O = T.ORACLE()
    g = gold("golden_select.npz")
    for k in [k for k in g.files if k.endswith("|in")]:
        tag, dist, n, hasnan = _parse(k)[:4]
        base = k[: -len("|in")]
        a, n, hn = g[k], int(n), int(hasnan)
        for key in [x for x in g.files if x.startswith(base + "|") and x != k]:
            op = key[len(base) + 1:]
            if op.startswith("qselect"):
                kk, desc = op[len("qselect"):].split("|")
                kk, desc = int(kk), int(desc)
                srt = O.qselect(a, tag, kk, hn, desc)
                T.assert_selected(g[key], a, srt, kk, tag, f"golden qselect {key}", zeros_as_values=True)
            elif op.startswith("partial"):
                kk, desc = op[len("partial"):].split("|")
                kk, desc = int(kk), int(desc)
                exp = O.partial_qsort(a, tag, kk, hn, desc)
                assert np.array_equal(T.vclass(exp[:kk], tag), T.vclass(g[key], tag)), f"golden partial {key}"
            elif op.startswith("argselect"):
                kk = int(op[len("argselect"):])
                srt = a[O.argsort(a, tag, hn, False).astype(np.int64)]
                T.assert_selected(g[key], a, srt, kk, tag, f"golden argselect {key}", zeros_as_values=True)


def test_golden_keyvalue():
    O = T.ORACLE()
    g = gold("golden_kv.npz")
    for k in [k for k in g.files if k.endswith("|k") and k.count("|") == 5]:
        ktag, vtag, dist, n, hasnan = _parse(k)[:5]
        base = k[: -len("|k")]
        kin, vin, hn = g[k], g[base + "|v"], int(hasnan)
        for desc in (0, 1):
            ek, ev = O.keyvalue_qsort(kin, vin, ktag, vtag, hn, desc)
            T.assert_kv_equal(g[f"{base}|qsort{desc}|k"], g[f"{base}|qsort{desc}|v"], ek, ev, ktag,
                              f"golden kv {base} desc={desc}", zeros_as_values=True)
            kk = int(n) // 2
            T.assert_selected(g[f"{base}|select{kk}|{desc}|k"], kin, ek, kk, ktag, f"golden kvselect {base}",
                              zeros_as_values=True)
            assert np.array_equal(T.vclass(g[f"{base}|partial{kk + 1}|{desc}|k"], ktag),
                                  T.vclass(ek[: kk + 1], ktag)), f"golden kvpartial {base}"


# ------------------------------------------------------------ live reference

needs_ref = pytest.mark.skipif(T.ref() is None, reason="oracle/_ref not built (needs the reference tree)")


@needs_ref
@pytest.mark.parametrize("tag", list(T.TAGS))
def test_oracle_matches_reference(tag):
    O, R = T.ORACLE(), T.REF()
    for seed, (dist, n) in enumerate((d, n) for d in T.DISTS for n in [0, 1, 2, 5, 33, 513, 4000, 20000]):
        a = T.gen(tag, dist, n, 40000 + seed)
        hasnan = dist in T.NAN_DISTS
        for desc in (False, True):
            T.assert_sorted_equal(O.qsort(a, tag, hasnan, desc), R.qsort(a, tag, hasnan, desc), tag,
                                  f"qsort {tag} {dist} {n} {desc}", zeros_as_values=True)
            ra = a[R.argsort(a, tag, hasnan, desc).astype(np.int64)]
            oa = a[O.argsort(a, tag, hasnan, desc).astype(np.int64)]
            T.assert_sorted_equal(oa, ra, tag, f"argsort {tag} {dist} {n}", zeros_as_values=True)
            if n >= 2:
                k = (7 * n) // 10
                T.assert_selected(R.qselect(a, tag, k, hasnan, desc), a, O.qselect(a, tag, k, hasnan, desc), k, tag,
                                  f"qselect {tag} {dist} {n}", zeros_as_values=True)


@needs_ref
@pytest.mark.parametrize("ktag", T.KV_TAGS)
def test_oracle_kv_matches_reference(ktag):
    O, R = T.ORACLE(), T.REF()
    for vtag in T.KV_TAGS:
        for seed, (dist, n) in enumerate((d, n) for d in ["random", "dups", "nan_payloads", "signed_zeros"]
                                         for n in [2, 100, 3000]):
            k = T.gen(ktag, dist, n, 41000 + seed)
            v = T.gen(vtag, "random", n, 42000 + seed)
            hasnan = dist in T.NAN_DISTS
            for desc in (False, True):
                rk, rv = R.keyvalue_qsort(k, v, ktag, vtag, hasnan, desc)
                ok, ov = O.keyvalue_qsort(k, v, ktag, vtag, hasnan, desc)
                T.assert_kv_equal(rk, rv, ok, ov, ktag, f"kv {ktag} {vtag} {dist} {n}", zeros_as_values=True)


@needs_ref
@pytest.mark.parametrize("ktag", T.KV16_TAGS)
def test_oracle_kv16_matches_reference_argsort(ktag):
    """16-bit keys with payloads have no reference entry point; the oracle's
    kv sort of them must equal the reference's argsort applied to keys and
    values (payloads compared as multisets per run of equal keys)."""
    O, R = T.ORACLE(), T.REF()
    for vtag in T.KV_TAGS:
        for seed, (dist, n) in enumerate((d, n) for d in ["random", "dups", "nan_payloads", "signed_zeros"]
                                         for n in [2, 100, 3000]):
            k = T.gen(ktag, dist, n, 43000 + seed)
            v = T.gen(vtag, "random", n, 44000 + seed)
            hasnan = dist in T.NAN_DISTS
            for desc in (False, True):
                idx = R.argsort(k, ktag, hasnan, desc).astype(np.int64)
                ok, ov = O.keyvalue_qsort(k, v, ktag, vtag, hasnan, desc)
                T.assert_kv_equal(k[idx], v[idx], ok, ov, ktag, f"kv16 {ktag} {vtag} {dist} {n}", zeros_as_values=True)


@needs_ref
def test_oracle_object_sort_matches_reference():
    import ctypes
    O, R = T.oracle(), T.ref()
    rng = np.random.default_rng(5)
    for n in [1, 2, 17, 5000]:
        xyz = rng.uniform(-1, 1, (n, 3))
        a, b = xyz.copy(), xyz.copy()
        keys = np.zeros(n)
        O.xss_oracle_object_qsort_xyz(ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(n), ctypes.c_void_p(keys.ctypes.data))
        R.xssref_object_qsort_xyz(ctypes.c_void_p(b.ctypes.data), ctypes.c_size_t(n))
        kb = np.sqrt(b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1] + b[:, 2] * b[:, 2])
        assert np.array_equal(keys, kb)
        # objects with equal keys may come in either order
        assert np.array_equal(np.sort(a.view(np.uint64), axis=0), np.sort(b.view(np.uint64), axis=0))


# ------------------------------------------------------------ planted mistakes

def test_checkers_reject_planted_mistakes():
    O = T.ORACLE()
    a = T.gen("f32", "nan", 2000, 1)
    good = O.qsort(a, "f32", True, False)
    bad = good.copy()
    bad[[10, 11]] = bad[[11, 10]] if bad[10] != bad[11] else bad[[10, 12]]
    with pytest.raises(AssertionError):
        T.assert_sorted_equal(bad, good, "f32")
    bad = good.copy()
    T.bits(bad)[-1] = 0x7FC00001  # a NaN with other bits where the canonical one belongs
    with pytest.raises(AssertionError):
        T.assert_sorted_equal(bad, good, "f32")
    k = T.gen("u32", "dups", 3000, 2)
    v = np.arange(k.size, dtype=np.uint32)
    ek, ev = O.keyvalue_qsort(k, v, "u32", "u32")
    i = int(np.nonzero(ek[1:] != ek[:-1])[0][0])  # last index of the first run
    bv = ev.copy()
    bv[[i, i + 1]] = bv[[i + 1, i]]  # a value moved to another key's run
    with pytest.raises(AssertionError):
        T.assert_kv_equal(ek, bv, ek, ev, "u32")
    idx = O.argsort(k, "u32")
    idx[5] = idx[6]
    with pytest.raises(AssertionError):
        T.assert_argsort(idx, k, k[O.argsort(k, "u32").astype(np.int64)], "u32")
    srt = O.qselect(k, "u32", 1500)
    sel = srt.copy()
    sel[[0, -1]] = sel[[-1, 0]]  # the largest key moved left of k
    with pytest.raises(AssertionError):
        T.assert_selected(sel, k, srt, 1500, "u32")
