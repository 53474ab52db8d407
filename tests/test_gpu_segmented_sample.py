"""GPU parity of the segmented form's sample-sort plan (4- and 8-byte keys)
against the oracle, segment by segment.

Segment lengths sit on both sides of every class boundary of the plan (LDS
bitonic up to 4096 keys -- 2048 for 8-byte keys --, one-workgroup LDS sort
up to 512 x items -- 6144 keys, 8-byte argsort 5632 --, sample sort above),
each segment with its own distribution (constant, two values,
duplicates, sorted, reverse, NaN, signed zeros), and the plan knobs are
forced so that every kernel of the path runs: buckets far above the LDS
capacity (the large-bucket passes), 15-item LDS sorts (8-byte keys: the
8-item ones), wide scatter tiles, one sample per bucket."""
import ctypes

import numpy as np
import pytest

import xss_testlib as T
from test_gpu_segmented import Plan, offsets_of

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 3, 100, 4095, 4096, 4097, 6143, 6144, 6145, 8000, 12289, 50000, 100003, 250000, 4099, 9000, 2048,
         2049, 5632, 5633]
DISTS = ["random", "dups", "sorted", "reverse", "constant", "two", "smallrange", "nan", "signed_zeros", "rand_max"]
KNOBS = {
    "default": {},
    "big_buckets": {"seg_bucket_keys": 40000},
    "items15": {"seg_local_items": 15},
    "wide": {"seg_wide_tiles": 1},
    "oversample1": {"seg_oversample": 1},
}


def batch(tag, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for i, n in enumerate(SIZES):
        dist = DISTS[i % len(DISTS)]
        if dist == "two":
            pool = T.gen(tag, "random", 2, seed + i)
            parts.append(pool[rng.integers(0, 2, n)])
        else:
            parts.append(T.gen(tag, dist, n, seed + i))
    keys = np.concatenate(parts).astype(T.TAGS[tag])
    return keys, offsets_of(np.array(SIZES, dtype=np.int64))


@pytest.fixture
def knobs(xss, request):
    L = xss.lib()
    cfg = KNOBS[request.param]
    for k, v in cfg.items():
        assert L.xss_gpu_set_tuning(k.encode(), ctypes.c_uint64(v)) == 0
    yield request.param
    for k in cfg:
        L.xss_gpu_set_tuning(k.encode(), ctypes.c_uint64(0))


@pytest.mark.parametrize("knobs", list(KNOBS), indirect=True)
@pytest.mark.parametrize("tag", ["u32", "i32", "f32", "u64", "i64", "f64"])
@pytest.mark.parametrize("desc", [False, True])
def test_sample_qsort(xss, knobs, tag, desc):
    O = T.ORACLE()
    keys, off = batch(tag, 31)
    hasnan = tag[0] == "f"
    d = xss.DArray.from_numpy(keys)
    Plan(xss, "qsort", tag, off, range_tiles=2).run(d, hasnan=hasnan, desc=desc)
    got = d.numpy()
    for i in range(len(SIZES)):
        a, b = int(off[i]), int(off[i + 1])
        T.assert_sorted_equal(got[a:b], O.qsort(keys[a:b], tag, hasnan, desc), tag,
                              f"{knobs} segment {i} len {b - a} ({DISTS[i % len(DISTS)]})")


@pytest.mark.parametrize("knobs", list(KNOBS), indirect=True)
@pytest.mark.parametrize("tag", ["u32", "i32", "f32", "u64", "i64", "f64"])
def test_sample_argsort(xss, knobs, tag):
    """Indices are segment-local, a permutation whose gathered keys equal the
    oracle's sort; the path is stable, so equal keys keep the input order."""
    O = T.ORACLE()
    keys, off = batch(tag, 41)
    d = xss.DArray.from_numpy(keys)
    idx = xss.DArray.empty("u32", keys.size)
    Plan(xss, "argsort", tag, off, range_tiles=2).run(d, idx)
    got = idx.numpy()
    assert np.array_equal(T.bits(d.numpy()), T.bits(keys)), "argsort modified the keys"
    for i in range(len(SIZES)):
        a, b = int(off[i]), int(off[i + 1])
        seg = keys[a:b]
        what = f"{knobs} segment {i} len {b - a} ({DISTS[i % len(DISTS)]})"
        T.assert_argsort(got[a:b].astype(np.uint64), seg, seg[O.argsort(seg, tag).astype(np.int64)], tag, what)
        g = T.bits(seg[got[a:b].astype(np.int64)])
        tie = g[1:] == g[:-1]
        assert np.all(got[a:b][1:][tie] > got[a:b][:-1][tie]), f"{what}: equal keys out of input order"


@pytest.mark.parametrize("knobs", list(KNOBS), indirect=True)
@pytest.mark.parametrize("tag", ["u32", "f32", "i64", "f64"])
def test_sample_qselect(xss, knobs, tag):
    """Every segment partitioned at k = len // 2 (NaN bits kept)."""
    O = T.ORACLE()
    keys, off = batch(tag, 51)
    hasnan = tag[0] == "f"
    d = xss.DArray.from_numpy(keys)
    Plan(xss, "qselect", tag, off, range_tiles=2).run(d, hasnan=hasnan)
    got = d.numpy()
    for i in range(len(SIZES)):
        a, b = int(off[i]), int(off[i + 1])
        if b - a < 2:
            continue
        k = (b - a) // 2
        T.assert_selected(got[a:b], keys[a:b], O.qselect(keys[a:b], tag, k, hasnan), k, tag,
                          f"{knobs} segment {i} len {b - a} ({DISTS[i % len(DISTS)]})")


@pytest.mark.parametrize("tag,sizes", [("f32", [10, 5000, 6144, 6145, 100000]),
                                       ("f64", [10, 3000, 6144, 6145, 100000])])
def test_sample_plan_info(xss, tag, sizes):
    """The plan sorts segments above the bitonic limit and up to the LDS radix
    capacity with one workgroup each (no ranges), the rest by sample sort."""
    off = offsets_of(np.array(sizes, dtype=np.int64))
    p = Plan(xss, "qsort", tag, off)
    nsmall, nr, nlarge = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    assert xss.lib().xss_gpu_segplan_info(p.p, ctypes.byref(nsmall), ctypes.byref(nr), ctypes.byref(nlarge)) == 0
    assert (nsmall.value, nlarge.value) == (1, 2)
    assert nr.value >= 2


@pytest.fixture
def sample_off(xss):
    L = xss.lib()
    assert L.xss_gpu_set_tuning(b"seg_sample_off", ctypes.c_uint64(1)) == 0
    yield
    L.xss_gpu_set_tuning(b"seg_sample_off", ctypes.c_uint64(0))


@pytest.mark.parametrize("op,tag", [("qsort", "f32"), ("qsort", "u64"), ("argsort", "i64"), ("qselect", "f64")])
def test_lsd_plan_with_sample_off(xss, sample_off, op, tag):
    """With the sample sort switched off, 4- and 8-byte keys take the LSD
    plan of 2-byte keys; it stays a drop-in for the same calls."""
    O = T.ORACLE()
    keys, off = batch(tag, 61)
    hasnan = tag[0] == "f"
    d = xss.DArray.from_numpy(keys)
    p = Plan(xss, op, tag, off, range_tiles=2)
    nsmall, nr, nlarge = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    assert xss.lib().xss_gpu_segplan_info(p.p, ctypes.byref(nsmall), ctypes.byref(nr), ctypes.byref(nlarge)) == 0
    idx = xss.DArray.empty("u32", keys.size) if op == "argsort" else None
    p.run(d, idx, hasnan=hasnan)
    got = d.numpy()
    for i in range(len(SIZES)):
        a, b = int(off[i]), int(off[i + 1])
        what = f"segment {i} len {b - a} ({DISTS[i % len(DISTS)]})"
        seg = keys[a:b]
        if op == "qsort":
            T.assert_sorted_equal(got[a:b], O.qsort(seg, tag, hasnan, False), tag, what)
        elif op == "argsort":
            gi = idx.numpy()[a:b].astype(np.uint64)
            T.assert_argsort(gi, seg, seg[O.argsort(seg, tag).astype(np.int64)], tag, what)
        elif b - a >= 2:
            k = (b - a) // 2
            T.assert_selected(got[a:b], seg, O.qselect(seg, tag, k, hasnan), k, tag, what)


@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_sample_all_nan_segment_canonical(xss, tag):
    """A segment of 20000 NaNs with assorted payloads (one bucket above the
    LDS capacity, every order key equal) comes back as canonical NaNs."""
    rng = np.random.default_rng(7)
    bits = np.uint32 if tag == "f32" else np.uint64
    exp = bits(0x7F800000) if tag == "f32" else bits(0x7FF0000000000000)
    nan = (exp | rng.integers(1, 1 << 20, 20000).astype(bits)).view(T.TAGS[tag])
    other = T.gen(tag, "random", 9000, 3)
    keys = np.concatenate([nan, other]).astype(T.TAGS[tag])
    off = offsets_of(np.array([20000, 9000], dtype=np.int64))
    d = xss.DArray.from_numpy(keys)
    Plan(xss, "qsort", tag, off, range_tiles=2).run(d, hasnan=True)
    got = d.numpy()
    assert (T.bits(got[:20000]) == T.QNAN[tag]).all()
    O = T.ORACLE()
    T.assert_sorted_equal(got[20000:], O.qsort(keys[20000:], tag, True, False), tag, "second segment")
