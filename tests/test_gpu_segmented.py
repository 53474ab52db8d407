"""GPU parity of the segmented form (xss_gpu_segplan_*) against the oracle,
segment by segment, and of its hipGraph replay against the eager run."""
import ctypes

import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu


def seg_sizes(seed, nseg, lo=0, hi=300_000):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(max(lo, 1)), np.log(hi), nseg)).astype(np.int64)
    sizes[:6] = [0, 1, 2, 3, 4096, 4097][: min(6, nseg)]
    return sizes


def offsets_of(sizes):
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)


class Plan:
    def __init__(self, xss, op, tag, off, small_max=0, range_tiles=0):
        self.L = xss.lib()
        self.p = ctypes.c_void_p()
        rc = self.L.xss_gpu_segplan_create(ctypes.byref(self.p), op.encode(), tag.encode(),
                                           off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                           ctypes.c_size_t(len(off) - 1), ctypes.c_uint32(small_max),
                                           ctypes.c_uint32(range_tiles))
        assert rc == 0, self.L.xss_gpu_last_error()

    def run(self, keys, idx=None, hasnan=False, desc=False):
        import torch
        rc = self.L.xss_gpu_segplan_run(self.p, keys.ptr, idx.ptr if idx is not None else None, int(hasnan),
                                        int(desc), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == 0, self.L.xss_gpu_last_error()

    def __del__(self):
        self.L.xss_gpu_segplan_destroy(self.p)


@pytest.mark.parametrize("tag", list(T.TAGS))
@pytest.mark.parametrize("desc", [False, True])
def test_segmented_qsort(xss, tag, desc):
    O = T.ORACLE()
    sizes = seg_sizes(1 + len(tag), 60)
    off = offsets_of(sizes)
    dist = "nan" if tag in T.FLOAT_TAGS else "dups"
    keys = T.gen(tag, dist, int(off[-1]), 11)
    hasnan = tag in T.FLOAT_TAGS
    d = xss.DArray.from_numpy(keys)
    Plan(xss, "qsort", tag, off, range_tiles=2).run(d, hasnan=hasnan, desc=desc)
    got = d.numpy()
    for i in range(len(sizes)):
        a, b = int(off[i]), int(off[i + 1])
        T.assert_sorted_equal(got[a:b], O.qsort(keys[a:b], tag, hasnan, desc), tag, f"segment {i} len {b - a}")


@pytest.mark.parametrize("tag", ["i32", "f32", "u64", "i16"])
def test_segmented_argsort(xss, tag):
    O = T.ORACLE()
    sizes = seg_sizes(7, 50)
    off = offsets_of(sizes)
    keys = T.gen(tag, "random", int(off[-1]), 12)
    d = xss.DArray.from_numpy(keys)
    idx = xss.DArray.empty("u32", int(off[-1]))
    Plan(xss, "argsort", tag, off).run(d, idx)
    got = idx.numpy()
    assert np.array_equal(T.bits(d.numpy()), T.bits(keys)), "argsort modified the keys"
    for i in range(len(sizes)):
        a, b = int(off[i]), int(off[i + 1])
        seg = keys[a:b]
        T.assert_argsort(got[a:b].astype(np.uint64), seg, seg[O.argsort(seg, tag).astype(np.int64)], tag,
                         f"segment {i} len {b - a}")


def test_segmented_qselect(xss):
    O = T.ORACLE()
    sizes = seg_sizes(9, 40)
    off = offsets_of(sizes)
    keys = T.gen("f32", "nan_payloads", int(off[-1]), 13)
    d = xss.DArray.from_numpy(keys)
    Plan(xss, "qselect", "f32", off).run(d, hasnan=True)
    got = d.numpy()
    for i in range(len(sizes)):
        a, b = int(off[i]), int(off[i + 1])
        if b - a < 2:
            continue
        k = (b - a) // 2
        T.assert_selected(got[a:b], keys[a:b], O.qselect(keys[a:b], "f32", k, True), k, "f32", f"segment {i}")


def test_segmented_graph_replay(xss):
    """The plan's launches captured in a graph and replayed three times, with
    eager work between the replays; every replay equals the eager result."""
    import torch
    sizes = seg_sizes(21, 80)
    off = offsets_of(sizes)
    keys = T.gen("f32", "nan", int(off[-1]), 14)
    plan = Plan(xss, "qsort", "f32", off)
    eager = xss.DArray.from_numpy(keys)
    plan.run(eager, hasnan=True)
    ref = eager.numpy()
    pristine = torch.from_numpy(keys.view(np.uint8).copy()).cuda()
    work = xss.DArray("f32", keys.size, pristine.clone())
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.run(work, hasnan=True)  # warm-up outside capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.run(work, hasnan=True)
    for rep in range(3):
        work.t.copy_(pristine)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(T.bits(work.numpy()), T.bits(ref)), f"replay {rep} differs from the eager run"
        # eager library work between replays
        other = xss.DArray.from_numpy(T.gen("u32", "random", 5000, rep))
        xss.qsort(other)
        torch.cuda.synchronize()


# ---------------------------------------------------------------- the batch workload at full size

def c4_offsets():
    """8192 segments, lengths log-uniform in [16, 2^20] (~760M keys), the
    batch workload's seeded lengths."""
    rng = np.random.default_rng(2024)
    sizes = np.exp(rng.uniform(np.log(16), np.log(1 << 20), 8192)).astype(np.int64)
    return offsets_of(sizes)


@pytest.fixture(scope="module")
def c4(xss):
    import torch
    import xss_devcheck as D
    off = c4_offsets()
    n = int(off[-1])
    seg = D.seg_ids(off, "cuda")
    yield off, n, seg
    del seg
    torch.cuda.empty_cache()


def _c4_floats(n):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    v = torch.rand(n, device="cuda", generator=g) * 2 - 1
    v[torch.rand(n, device="cuda", generator=g) < 0.001] = float("nan")
    return v.view(torch.int32)


def test_segmented_c4_qsort_all_segments(xss, c4):
    """qsort of all 8192 segments (U(-1,1) floats, 0.1% NaN, hasnan): every
    segment ordered and a permutation of its input, NaNs canonical."""
    import torch
    import xss_devcheck as D
    off, n, seg = c4
    inp = _c4_floats(n)
    for desc in (False, True):
        got = inp.clone()
        Plan(xss, "qsort", "f32", off).run(xss.DArray("f32", n, got.view(torch.uint8)), hasnan=True, desc=desc)
        D.check_segments_sorted_perm(got, inp, seg, "f32", desc, f"c4 qsort desc={desc}")
        nan = D.is_nan(got, "f32")
        assert int(nan.sum()) == int(D.is_nan(inp, "f32").sum()) and bool((got[nan] == 0x7FC00000).all())


def test_segmented_c4_argsort_all_segments(xss, c4):
    """argsort of int32 keys (full range) in all 8192 segments: segment-local
    uint32 indices that gather each segment's keys in order."""
    import torch
    import xss_devcheck as D
    off, n, seg = c4
    g = torch.Generator(device="cuda")
    g.manual_seed(78)
    inp = torch.randint(-2 ** 31, 2 ** 31, (n,), dtype=torch.int32, device="cuda", generator=g)
    keep = inp.clone()
    idx = torch.empty(n, dtype=torch.int32, device="cuda")
    Plan(xss, "argsort", "i32", off).run(xss.DArray("i32", n, inp.view(torch.uint8)),
                                         xss.DArray("u32", n, idx.view(torch.uint8)))
    assert torch.equal(inp, keep), "argsort modified the keys"
    start = torch.as_tensor(off[:-1].astype(np.int64), device="cuda")
    lens = torch.as_tensor((off[1:] - off[:-1]).astype(np.int64), device="cuda")
    local = idx.to(torch.int64)
    assert bool((local >= 0).all()) and bool((local < lens[seg]).all()), "index outside its segment"
    glob = local + start[seg]
    D.check_permutation(glob, n, "c4 argsort")
    D.check_segments_sorted_perm(inp[glob], inp, seg, "i32", False, "c4 argsort gathered keys")


def test_segmented_c4_qselect_all_segments(xss, c4):
    """qselect at k = len/2 in all 8192 segments: the k-th value class at k,
    nothing larger before it and nothing smaller after, a permutation."""
    import torch
    import xss_devcheck as D
    off, n, seg = c4
    inp = _c4_floats(n)
    got = inp.clone()
    Plan(xss, "qselect", "f32", off).run(xss.DArray("f32", n, got.view(torch.uint8)), hasnan=True)
    start = torch.as_tensor(off[:-1].astype(np.int64), device="cuda")
    lens = torch.as_tensor((off[1:] - off[:-1]).astype(np.int64), device="cuda")
    seg64 = seg.to(torch.int64)
    comp = lambda x: (seg64 << 32) | (D.order_key(x, "f32").to(torch.int64) + (1 << 31))  # noqa: E731
    srt = torch.sort(comp(inp)).values
    cg = comp(got)
    assert torch.equal(torch.sort(cg).values, srt), "segments are not permutations of their inputs"
    kpos = start + lens // 2
    kth = srt[kpos]
    assert torch.equal(cg[kpos], kth), "k-th value differs"
    pos = torch.arange(n, device="cuda")
    left = pos < kpos[seg64]
    assert bool((cg[left] <= kth[seg64[left]]).all()), "a larger value left of k"
    assert bool((cg[~left] >= kth[seg64[~left]]).all()), "a smaller value right of k"
    assert torch.equal(torch.sort(got).values, torch.sort(inp).values), "NaN bits or keys changed"
