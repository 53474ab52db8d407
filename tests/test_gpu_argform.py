"""GPU parity of the NumPy-style arg forms, argsort(T *arr, size_t *arg, ...)
and argselect(T *arr, size_t *arg, k, ...) (src/x86simdsort-static-incl.h:
31-46,92-113): `arg` holds indices into `arr` on entry and is reordered in
place.  With arg = 0..n-1 (what NumPy passes) the result must satisfy the
argsort / argselect rules; with any other index list the indices are a
permutation of the ones given and the keys they gather, a[arg], must equal
the oracle's argsort (or select) of the gathered input a[arg_in]."""
import ctypes

import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

SIZES = [2, 3, 100, 4097, 70001, (1 << 20) + 3]
DISTS = ["random", "dups", "constant", "sorted", "nan", "rand_with_max_and_nan", "signed_zeros"]


def arg_lists(n, seed):
    rng = np.random.default_rng(seed)
    yield "identity", np.arange(n, dtype=np.uint64)
    yield "permutation", rng.permutation(n).astype(np.uint64)
    yield "with repeats", rng.integers(0, n, n).astype(np.uint64)


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", list(T.TAGS))
def test_argsort_arg(xss, tag, desc):
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in DISTS for n in SIZES):
        if dist == "signed_zeros" and tag not in ("f16", "f32", "f64"):
            continue
        a = T.gen(tag, dist, n, 13000 + seed)
        hasnan = dist in T.NAN_DISTS
        d = xss.DArray.from_numpy(a)
        for name, arg_in in arg_lists(n, seed):
            kin = a[arg_in.astype(np.int64)]
            exp = kin[O.argsort(kin, tag, hasnan, desc).astype(np.int64)]
            g = xss.DArray.from_numpy(arg_in.copy())
            out = xss.argsort_arg(d, g, hasnan, desc).numpy()
            what = f"argsort_arg<{tag}> {dist} n={n} desc={desc} arg={name}"
            assert np.array_equal(np.sort(out), np.sort(arg_in)), what + ": indices are not those given"
            T.assert_sorted_equal(a[out.astype(np.int64)], exp, tag, what + " gathered keys")
        assert np.array_equal(T.bits(d.numpy()), T.bits(a)), "argsort_arg modified the keys"


@pytest.mark.parametrize("tag", list(T.TAGS))
def test_argselect_arg(xss, tag):
    O = T.ORACLE()
    for seed, (dist, n) in enumerate((d, n) for d in DISTS for n in SIZES):
        if dist == "signed_zeros" and tag not in ("f16", "f32", "f64"):
            continue
        a = T.gen(tag, dist, n, 14000 + seed)
        hasnan = dist in T.NAN_DISTS
        d = xss.DArray.from_numpy(a)
        for name, arg_in in arg_lists(n, seed):
            kin = a[arg_in.astype(np.int64)]
            srt = kin[O.argsort(kin, tag, hasnan, False).astype(np.int64)]
            for k in sorted({0, n // 2, n - 1}):
                g = xss.DArray.from_numpy(arg_in.copy())
                out = xss.argselect_arg(d, g, k, hasnan).numpy()
                what = f"argselect_arg<{tag}> {dist} n={n} k={k} arg={name}"
                assert np.array_equal(np.sort(out), np.sort(arg_in)), what + ": indices are not those given"
                T.assert_selected(a[out.astype(np.int64)], kin, srt, k, tag, what)


def test_argform_rejects_k_out_of_range(xss):
    a = xss.DArray.from_numpy(np.arange(10, dtype=np.float32))
    g = xss.DArray.from_numpy(np.arange(10, dtype=np.uint64))
    with pytest.raises(RuntimeError):
        xss.argselect_arg(a, g, 10)


def test_host_static_argsort_arg(xss):
    """The host-pointer form behind x86simdsortStatic::argsort(T *, size_t *arg)."""
    L = xss.lib()
    a = T.gen("f32", "rand_with_max_and_nan", 50001, 15000)
    arg = np.random.default_rng(1).permutation(a.size).astype(np.uint64)
    arg_in = arg.copy()
    f = L.xss_host_argsort_arg_f32
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    assert f(a.ctypes.data, arg.ctypes.data, a.size, 1, 1) == 0
    kin = a[arg_in.astype(np.int64)]
    exp = kin[T.ORACLE().argsort(kin, "f32", True, True).astype(np.int64)]
    assert np.array_equal(np.sort(arg), np.sort(arg_in))
    T.assert_sorted_equal(a[arg.astype(np.int64)], exp, "f32", "host argsort_arg")
