"""Shared test helpers: seeded input generators, oracle/reference wrappers and
the parity checkers.

The checkers restate the reference's own test checkers
(tests/test-qsort-common.h:29-104, tests/test-keyvalue.cpp:44-171):
  * sorted keys are compared position by position on value classes (every NaN
    one class, -0.0 == +0.0) and the whole output must be a bit-pattern
    permutation of the expected one -- so keys are bit-exact except inside
    the +-0 run and the NaN run, which are compared as multisets;
  * key-value payloads are compared as a multiset within each run of equal
    keys;
  * argsort must be a permutation whose gathered keys equal the sorted keys;
  * select must put the k-th value at k with the partition property, which
    is checked as multiset equality of the value classes left of k, at k and
    right of k against a full sort.
"""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libxss_ref.so")

TAGS = {
    "u16": np.uint16, "i16": np.int16, "f16": np.float16,
    "u32": np.uint32, "i32": np.int32, "f32": np.float32,
    "u64": np.uint64, "i64": np.int64, "f64": np.float64,
}
UINT = {2: np.uint16, 4: np.uint32, 8: np.uint64}
FLOAT_TAGS = ("f16", "f32", "f64")
KV_TAGS = ["u32", "i32", "f32", "u64", "i64", "f64"]
KV16_TAGS = ["u16", "i16", "f16"]  # 16-bit kv keys: beyond the reference's 36 pairs
# canonical NaN written by qsort(hasnan) (src/xss-custom-float.h:34-38,81-85)
QNAN = {"f16": 0x7C01, "f32": 0x7FC00000, "f64": 0x7FF8000000000000}


def bits(a):
    return a.view(UINT[a.dtype.itemsize])


def is_nan(a, tag):
    if tag not in FLOAT_TAGS:
        return np.zeros(a.shape, bool)
    return np.isnan(a.astype(np.float64)) if tag == "f16" else np.isnan(a)


# ---------------------------------------------------------------- generators

def _finite_random(tag, n, rng):
    dt = np.dtype(TAGS[tag])
    u = UINT[dt.itemsize]
    nbits = dt.itemsize * 8
    out = rng.integers(0, 2 ** nbits, size=n, dtype=np.uint64 if nbits == 64 else np.int64).astype(u)
    if tag in FLOAT_TAGS:
        exp_mask = {16: 0x7C00, 32: 0x7F800000, 64: 0x7FF0000000000000}[nbits]
        bad = (out & u(exp_mask)) == u(exp_mask)
        while bad.any():
            out[bad] = rng.integers(0, 2 ** nbits, size=int(bad.sum()),
                                    dtype=np.uint64 if nbits == 64 else np.int64).astype(u)
            bad = (out & u(exp_mask)) == u(exp_mask)
    return out.view(dt)


def _sorted_np(a, tag):
    if tag in FLOAT_TAGS:
        return np.sort(a.astype(np.float64) if tag == "f16" else a, kind="stable").astype(a.dtype)
    return np.sort(a, kind="stable")


def _maxval(tag):
    dt = TAGS[tag]
    return dt(np.inf) if tag in FLOAT_TAGS else np.iinfo(dt).max


def _nan(tag):
    return np.array([QNAN[tag]], dtype=UINT[np.dtype(TAGS[tag]).itemsize]).view(TAGS[tag])[0]


DISTS = ["random", "dups", "sorted", "reverse", "constant", "smallrange", "max_at_the_end", "random_5d",
         "rand_max", "nan", "rand_with_nan", "rand_with_max_and_nan", "signed_zeros", "nan_payloads"]
NAN_DISTS = ("nan", "rand_with_nan", "rand_with_max_and_nan", "nan_payloads")


def gen(tag, dist, n, seed):
    """Seeded input of `n` keys of type `tag` (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    dt = TAGS[tag]
    if n == 0:
        return np.zeros(0, dt)
    if dist in ("nan", "rand_with_nan", "rand_with_max_and_nan", "nan_payloads") and tag not in FLOAT_TAGS:
        dist = {"nan": "random", "rand_with_nan": "rand_max", "rand_with_max_and_nan": "rand_max",
                "nan_payloads": "random"}[dist]
    if dist == "random":
        return _finite_random(tag, n, rng)
    if dist == "dups":  # heavy duplicates: 1024 distinct values
        pool = _finite_random(tag, 1024, rng)
        return pool[rng.integers(0, 1024, n)]
    if dist == "sorted":
        return _sorted_np(_finite_random(tag, n, rng), tag)
    if dist == "reverse":
        return _sorted_np(_finite_random(tag, n, rng), tag)[::-1].copy()
    if dist == "constant":
        return np.full(n, _finite_random(tag, 1, rng)[0], dt)
    if dist == "smallrange":
        return rng.integers(1, 21, n).astype(dt)
    if dist == "max_at_the_end":
        a = _finite_random(tag, n, rng)
        a[-1] = _maxval(tag)
        return a
    if dist == "random_5d":
        a = _finite_random(tag, n, rng)
        a[rng.random(n) < 0.5] = 0
        return a
    if dist == "rand_max":
        a = _finite_random(tag, n, rng)
        a[rng.random(n) < 0.5] = _maxval(tag)
        return a
    if dist == "nan":  # 1% canonical NaN
        a = _finite_random(tag, n, rng)
        a[rng.random(n) < 0.01] = _nan(tag)
        return a
    if dist == "rand_with_nan":
        a = _finite_random(tag, n, rng)
        a[rng.integers(0, n, max(1, 10 % n))] = _nan(tag)
        return a
    if dist == "rand_with_max_and_nan":
        a = _finite_random(tag, n, rng)
        r = rng.integers(0, 4, n)
        a[r == 2] = _maxval(tag)
        a[r == 3] = _nan(tag)
        return a
    if dist == "signed_zeros":
        a = _finite_random(tag, n, rng)
        r = rng.integers(0, 3, n)
        if tag in FLOAT_TAGS:
            a[r == 0] = dt(0.0)
            a[r == 1] = dt(-0.0)
        else:
            a[r == 0] = 0
        return a
    if dist == "nan_payloads":  # NaNs of both signs with random payloads
        a = _finite_random(tag, n, rng)
        u = bits(a)
        nb = a.dtype.itemsize * 8
        exp_mask = {16: 0x7C00, 32: 0x7F800000, 64: 0x7FF0000000000000}[nb]
        pick = rng.random(n) < 0.05
        k = int(pick.sum())
        pay = rng.integers(1, 2 ** min(nb - 12, 50), k).astype(u.dtype)
        sign = (rng.integers(0, 2, k).astype(u.dtype) << u.dtype.type(nb - 1))
        u[pick] = sign | u.dtype.type(exp_mask) | pay
        return a
    raise ValueError(dist)


# ---------------------------------------------------------------- oracle / reference

_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "liboracle.so"], check=True,
                           capture_output=True)
        L = ctypes.CDLL(ORACLE_SO)
        L.xss_oracle_compare_f64.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _oracle = L
    return _oracle


def ref():
    """The reference library built from its own sources (None if absent)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        L = ctypes.CDLL(REF_SO)
        L.xssref_isa.restype = ctypes.c_char_p
        _ref = L
    return _ref


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _z(x):
    return ctypes.c_size_t(int(x))


class CpuImpl:
    """Calls into the oracle (prefix xss_oracle_) or the reference (xssref_)."""

    def __init__(self, L, prefix):
        self.L, self.pre = L, prefix

    def qsort(self, a, tag, hasnan=False, desc=False):
        b = a.copy()
        getattr(self.L, f"{self.pre}qsort_{tag}")(_p(b), _z(b.size), int(hasnan), int(desc))
        return b

    def qselect(self, a, tag, k, hasnan=False, desc=False):
        b = a.copy()
        getattr(self.L, f"{self.pre}qselect_{tag}")(_p(b), _z(k), _z(b.size), int(hasnan), int(desc))
        return b

    def partial_qsort(self, a, tag, k, hasnan=False, desc=False):
        b = a.copy()
        getattr(self.L, f"{self.pre}partial_qsort_{tag}")(_p(b), _z(k), _z(b.size), int(hasnan), int(desc))
        return b

    def argsort(self, a, tag, hasnan=False, desc=False):
        b = a.copy()
        out = np.zeros(max(1, a.size), np.uint64)
        getattr(self.L, f"{self.pre}argsort_{tag}")(_p(b), _z(b.size), int(hasnan), int(desc), _p(out))
        return out[: a.size]

    def argselect(self, a, tag, k, hasnan=False):
        b = a.copy()
        out = np.zeros(max(1, a.size), np.uint64)
        getattr(self.L, f"{self.pre}argselect_{tag}")(_p(b), _z(k), _z(b.size), int(hasnan), _p(out))
        return out[: a.size]

    def _kv(self, op, k, v, ktag, vtag, kk, hasnan, desc):
        kb, vb = k.copy(), v.copy()
        if self.pre == "xss_oracle_":
            f = getattr(self.L, f"{self.pre}{op}_{ktag}")
            args = [_p(kb), _p(vb), _z(vb.dtype.itemsize)]
        else:
            f = getattr(self.L, f"{self.pre}{op}_{ktag}_{vtag}")
            args = [_p(kb), _p(vb)]
        if kk is not None:
            args.append(_z(kk))
        f(*args, _z(kb.size), int(hasnan), int(desc))
        return kb, vb

    def keyvalue_qsort(self, k, v, ktag, vtag, hasnan=False, desc=False):
        return self._kv("keyvalue_qsort", k, v, ktag, vtag, None, hasnan, desc)

    def keyvalue_select(self, k, v, ktag, vtag, kk, hasnan=False, desc=False):
        return self._kv("keyvalue_select", k, v, ktag, vtag, kk, hasnan, desc)

    def keyvalue_partial_sort(self, k, v, ktag, vtag, kk, hasnan=False, desc=False):
        return self._kv("keyvalue_partial_sort", k, v, ktag, vtag, kk, hasnan, desc)


def ORACLE():
    return CpuImpl(oracle(), "xss_oracle_")


def REF():
    L = ref()
    return CpuImpl(L, "xssref_") if L is not None else None


# ---------------------------------------------------------------- checkers

def vclass(a, tag):
    """Value classes: every NaN one class, -0.0 and +0.0 one class."""
    u = bits(np.ascontiguousarray(a)).astype(np.uint64)
    if tag in FLOAT_TAGS:
        nb = a.dtype.itemsize * 8
        sign = np.uint64(1) << np.uint64(nb - 1)
        u = np.where(u == sign, np.uint64(0), u)
        u = np.where(is_nan(a, tag), np.uint64(QNAN[tag]), u)
    return u


def assert_sorted_equal(got, exp, tag, what="", zeros_as_values=False):
    """Position-wise equal value classes and equal bit-pattern multisets.

    zeros_as_values: compare the +-0 run by value only.  Needed when `exp` or
    `got` comes from the reference itself: its AVX-512 min/max networks may
    return +0.0 for -0.0 (and the reverse), so its zero run is not a
    permutation of the input's zeros.
    """
    assert got.shape == exp.shape, f"{what}: shape {got.shape} != {exp.shape}"
    cg, ce = vclass(got, tag), vclass(exp, tag)
    if not np.array_equal(cg, ce):
        i = int(np.nonzero(cg != ce)[0][0])
        raise AssertionError(f"{what}: first mismatch at {i}: got {got[max(0,i-2):i+3]} "
                             f"expected {exp[max(0,i-2):i+3]} (n={got.size})")
    bg, be = bits(got), bits(exp)
    if zeros_as_values and tag in FLOAT_TAGS:
        bg, be = bg[cg != 0], be[ce != 0]
    assert np.array_equal(np.sort(bg), np.sort(be)), f"{what}: bit patterns differ as multisets (+-0 / NaN runs)"


def assert_prefix_equal(got, exp, k, tag, what=""):
    """The first k of a partial sort.  The run of values equal to the k-th
    element (a +-0 run, a NaN run with several payloads) may hold any of its
    members before k and at k-1, so that run is compared by value; everything
    else bit-exactly."""
    cg, ce = vclass(got[:k], tag), vclass(exp[:k], tag)
    if not np.array_equal(cg, ce):
        i = int(np.nonzero(cg != ce)[0][0])
        raise AssertionError(f"{what}: first mismatch at {i}: got {got[max(0,i-2):i+3]} "
                             f"expected {exp[max(0,i-2):i+3]}")
    # the run holding the k-th element may be cut anywhere: compare it by value
    keep = ce != ce[-1] if k > 0 else np.ones(0, bool)
    assert np.array_equal(np.sort(bits(got[:k])[keep]), np.sort(bits(exp[:k])[keep])), \
        f"{what}: bit patterns of the sorted prefix differ"


def assert_kv_equal(gk, gv, ek, ev, tag, what="", zeros_as_values=False):
    """Keys as assert_sorted_equal; (key, value) pairs as a multiset within
    each run of equal keys (tests/test-keyvalue.cpp:44-100).  With
    zeros_as_values the key bits inside the +-0 run are compared by value
    (the reference's min/max networks may rewrite -0.0 as +0.0)."""
    assert_sorted_equal(gk, ek, tag, what + " keys", zeros_as_values)
    ce = vclass(ek, tag)
    run = np.concatenate([[0], np.cumsum(ce[1:] != ce[:-1])]) if ce.size else ce
    kg, ke = bits(gk).astype(np.uint64), bits(ek).astype(np.uint64)
    if zeros_as_values:
        kg = np.where(vclass(gk, tag) == 0, np.uint64(0), kg)
        ke = np.where(ce == 0, np.uint64(0), ke)
    # (run, key bits, value bits) multisets must agree
    og = np.lexsort((bits(gv), kg, run))
    oe = np.lexsort((bits(ev), ke, run))
    assert np.array_equal(kg[og], ke[oe]), f"{what}: key/value pairing differs"
    assert np.array_equal(bits(gv)[og], bits(ev)[oe]), f"{what}: values differ within equal-key runs"


def assert_argsort(idx, a, exp_sorted, tag, what=""):
    n = a.size
    assert idx.shape == (n,), f"{what}: index count {idx.shape} != {n}"
    assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint64)), f"{what}: not a permutation"
    assert_sorted_equal(a[idx.astype(np.int64)], exp_sorted, tag, what + " gathered keys")


def assert_selected(got, inp, exp_sorted, k, tag, what="", zeros_as_values=False):
    """got is a valid select(k) result for input `inp` whose full sort is exp_sorted."""
    if zeros_as_values:
        assert np.array_equal(np.sort(vclass(got, tag)), np.sort(vclass(inp, tag))), f"{what}: not a permutation"
    else:
        assert np.array_equal(np.sort(bits(got)), np.sort(bits(inp))), f"{what}: not a permutation of the input"
    cg, ce = vclass(got, tag), vclass(exp_sorted, tag)
    assert cg[k] == ce[k], f"{what}: k-th element {got[k]} != {exp_sorted[k]}"
    assert np.array_equal(np.sort(cg[:k]), np.sort(ce[:k])), f"{what}: left partition differs"
    assert np.array_equal(np.sort(cg[k + 1:]), np.sort(ce[k + 1:])), f"{what}: right partition differs"


def assert_pairs_preserved(gk, gv, k, v, what=""):
    """Every (key, value) pair of the output is an input pair, each used once
    (bit patterns)."""
    pin = np.lexsort((bits(v), bits(k)))
    pout = np.lexsort((bits(gv), bits(gk)))
    assert np.array_equal(bits(k)[pin], bits(gk)[pout]) and np.array_equal(bits(v)[pin], bits(gv)[pout]), \
        f"{what}: (key, value) pairs are not the input's"


def assert_kv_partial(gk, gv, k, ink, inv, ek, ev, tag, what=""):
    """keyvalue_partial_sort(k) result (gk, gv) of input (ink, inv), against
    the oracle's full key-value sort (ek, ev), with the reference's checker
    rules (tests/test-keyvalue.cpp:102-171): keys of the first k as the
    sorted prefix; the values of every run of equal keys that ends inside the
    prefix as a multiset; the run cut at k holds values of that run's keys;
    and pairs are never split."""
    assert_pairs_preserved(gk, gv, ink, inv, what)
    assert_prefix_equal(gk, ek, k, tag, what)
    if k == 0:
        return
    ce = vclass(ek, tag)
    kfull = k
    if k < ek.size and ce[k] == ce[k - 1]:  # the run at k-1 goes on past the prefix
        kfull = int(np.argmax(ce == ce[k - 1]))
    assert_kv_equal(gk[:kfull], gv[:kfull], ek[:kfull], ev[:kfull], tag, what + " runs inside the prefix")
    if kfull < k:  # values of the cut run: a sub-multiset of that run's values
        pv, pc = np.unique(bits(ev)[ce == ce[k - 1]], return_counts=True)
        gvv, gc = np.unique(bits(gv)[kfull:k], return_counts=True)
        j = np.minimum(np.searchsorted(pv, gvv), pv.size - 1)
        assert bool(((pv[j] == gvv) & (pc[j] >= gc)).all()), f"{what}: a value in the cut run is not from that run"
