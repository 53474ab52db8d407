"""The MSD-first sort plans on the device: no host round trip decides its
route, so a sort of skewed data captured in a hipGraph does the same work as
the eager call.  Each replay (with eager library work between replays) must
equal the eager result, and the result must be a sorted permutation."""
import ctypes

import pytest
import torch

import xss_devcheck as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tag,recipe,kv", [("f32", "dups", False), ("f32", "nan", False), ("u64", "narrow", False),
                                           ("u64", "narrow", True), ("f32", "random", True)])
def test_captured_msd_sort_matches_eager(xss, tag, recipe, kv):
    n = 1 << 25
    if recipe == "narrow":  # 64-bit keys below 2^32 (constant top digits)
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        keys = torch.randint(0, 2 ** 32, (n,), dtype=torch.int64, device="cuda", generator=g)
    else:
        keys = D.gen(tag, recipe, n, 6)
    hasnan = recipe == "nan"
    vtag = "u64" if D.WIDTH[tag] == 64 else "u32"
    vals = torch.arange(n, device="cuda", dtype=torch.int64 if vtag == "u64" else torch.int32)
    s = torch.cuda.Stream()
    work, wv = keys.clone(), vals.clone()
    dk = xss.DArray(tag, n, work.view(torch.uint8))
    dv = xss.DArray(vtag, n, wv.view(torch.uint8))

    def run():
        if kv:
            xss.keyvalue_qsort(dk, dv, hasnan, False)
        else:
            xss.qsort(dk, hasnan, False)

    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # eager run; also grows this stream's workspace before the capture
        run()
    s.synchronize()
    eager_k, eager_v = work.clone(), wv.clone()
    D.check_sort(eager_k, keys, tag, hasnan, False, not kv, "eager")
    if kv:
        D.check_gather(eager_k, keys, eager_v.to(torch.int64), "eager")
    graph = torch.cuda.CUDAGraph()
    work.copy_(keys)
    wv.copy_(vals)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        run()
    for rep in range(3):
        work.copy_(keys)
        wv.copy_(vals)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(work, eager_k), f"replay {rep}: keys differ from the eager run"
        if kv:
            D.check_gather(work, keys, wv.to(torch.int64), f"replay {rep}")
        other = xss.DArray.from_numpy(__import__("numpy").arange(5000, dtype="uint32")[::-1].copy())
        xss.qsort(other)  # eager library work between replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("tag,n,op", [("f32", 5000, "qsort"), ("u64", 6144, "qsort"), ("i32", 6000, "argsort"),
                                      ("f64", 3000, "argsort")])
def test_captured_one_workgroup_sort_matches_eager(xss, tag, n, op):
    """Arrays above the LDS bitonic size and up to 6144 keys are sorted by
    one workgroup in one launch, with no scratch: a captured call replays
    to the eager result."""
    keys = D.gen(tag, "random", n, 8)
    s = torch.cuda.Stream()
    work = keys.clone()
    dk = xss.DArray(tag, n, work.view(torch.uint8))
    idx = xss.DArray.empty("u64", n)

    def run():
        if op == "qsort":
            xss.qsort(dk)
        else:
            xss.argsort(dk, out=idx)

    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run()
    s.synchronize()
    eager_k, eager_i = work.clone(), idx.numpy().copy()
    if op == "qsort":
        D.check_sort(eager_k, keys, tag, False, False, True, "eager")
    graph = torch.cuda.CUDAGraph()
    work.copy_(keys)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        run()
    for rep in range(3):
        work.copy_(keys)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(work, eager_k), f"replay {rep}: keys differ from the eager run"
        if op == "argsort":
            assert (idx.numpy() == eager_i).all(), f"replay {rep}: indices differ from the eager run"
        other = xss.DArray.from_numpy(__import__("numpy").arange(5000, dtype="uint32")[::-1].copy())
        xss.qsort(other)
        torch.cuda.synchronize()
