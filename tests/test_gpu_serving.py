"""Serving form: every request's result in a concurrent, coalesced mix equals
its result when run alone through the plain call.

A seeded mix of qsort (all key types, short and long, NaN / descending
flags), keyvalue_qsort, qselect and partial_qsort requests is served four
ways -- a burst into 4 streams with coalescing of requests up to 2^20 keys
(short ones by the pointer-list LDS sort, longer ones by shared segmented
passes), the same with coalescing of LDS-sized ones only, a 2-stream server
whose admission budget holds only a few requests at a time, and one request
at a time -- and compared with the same requests run alone:
  * qsort: bit-exact (a keys-only result is fixed by its keys, and
    qsort(hasnan) writes the canonical NaN);
  * keyvalue_qsort: keys bit-exact, payloads as a multiset within each run
    of equal keys (the reference's kv sort is not stable either);
  * qselect: the same value class at k, the same multisets left and right of
    k, and a permutation of the input (the order inside a partition is not
    fixed);
  * partial_qsort: the sorted prefix as the partial-sort checker compares it,
    and the same multiset overall.
Every served result is also checked against the oracle with the reference's
checker rules.  A trace shaped like the served workload (70% qsort<float>
of 1K-16M keys, 20% keyvalue_qsort<u64,u64> of Zipf keys of 1K-4M, 10%
qselect / partial_qsort<u32> of 1K-16M, Poisson arrivals at 10k req/s) is
checked through the full-size properties on the GPU, with the oracle for the
requests of at most 2^20 keys.
"""
import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

SMALL_TAGS = ["f32", "u64", "i16"]
KV_VAL = ["u32", "u64", "f64", "i32"]
DISTS = ["random", "dups", "signed_zeros", "nan", "nan_payloads", "sorted"]


def _mix(count, seed):
    """[(op, tag, vtag, keys, vals, k, hasnan, desc)]"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        u = rng.random()
        dist = DISTS[int(rng.integers(0, len(DISTS)))]
        desc = bool(rng.integers(0, 2))
        if u < 0.45:  # short qsorts (coalesced), a few type/flag combinations
            tag = SMALL_TAGS[int(rng.integers(0, len(SMALL_TAGS)))]
            n = int(rng.integers(0, 4097)) if tag != "u64" else int(rng.integers(0, 2049))
            op, vtag, k = "qsort", None, 0
        elif u < 0.65:
            tag = list(T.TAGS)[int(rng.integers(0, 9))]
            n = int(np.exp(rng.uniform(np.log(4097), np.log(1 << 18))))
            op, vtag, k = "qsort", None, 0
        elif u < 0.8:
            kt = T.KV_TAGS + T.KV16_TAGS
            tag = kt[int(rng.integers(0, len(kt)))]
            vtag = KV_VAL[int(rng.integers(0, len(KV_VAL)))]
            n = int(np.exp(rng.uniform(np.log(2), np.log(1 << 17))))
            op, k = "keyvalue_qsort", 0
        else:
            tag = list(T.TAGS)[int(rng.integers(0, 9))]
            n = int(np.exp(rng.uniform(np.log(2), np.log(1 << 18))))
            op, vtag = ("qselect", None) if rng.random() < 0.5 else ("partial_qsort", None)
            k = int(rng.integers(0, n)) if op == "qselect" else int(rng.integers(1, n + 1))
        keys = T.gen(tag, dist, n, seed * 1000 + i)
        hasnan = dist in T.NAN_DISTS and tag in T.FLOAT_TAGS
        vals = None
        if vtag is not None:
            vals = rng.integers(0, 2 ** 31, n).astype(T.TAGS[vtag])
        out.append((op, tag, vtag, keys, vals, k, hasnan, desc))
    return out


def _alone(xss, item):
    op, tag, vtag, keys, vals, k, hasnan, desc = item
    d = xss.DArray.from_numpy(keys)
    if op == "qsort":
        xss.qsort(d, hasnan, desc)
        return d.numpy(), None
    if op == "keyvalue_qsort":
        v = xss.DArray.from_numpy(vals)
        xss.keyvalue_qsort(d, v, hasnan, desc)
        return d.numpy(), v.numpy()
    if op == "qselect":
        xss.qselect(d, k, hasnan, desc)
    else:
        xss.partial_qsort(d, k, hasnan, desc)
    return d.numpy(), None


def _serve(xss, mix, **server_args):
    import torch
    arrays = []
    reqs = []
    for op, tag, vtag, keys, vals, k, hasnan, desc in mix:
        d = xss.DArray.from_numpy(keys)
        v = xss.DArray.from_numpy(vals) if vals is not None else None
        arrays.append((d, v))
        reqs.append(xss.make_request(op, d, v, k, hasnan, desc))
    torch.cuda.synchronize()  # inputs ready before the server's streams read them
    with xss.Server(**server_args) as sv:
        done = sv.run_trace(reqs)
        stats = sv.stats()
    assert all(r.latency_us >= 0 for r in done)
    results = [(d.numpy(), v.numpy() if v is not None else None) for d, v in arrays]
    return results, stats


def _oracle_check(item, sk, sv, w):
    """The served result against the oracle, with the reference's checker
    rules (tests/test-qsort-common.h:29-104, tests/test-keyvalue.cpp:44-100)."""
    op, tag, vtag, keys, vals, k, hasnan, desc = item
    O = T.ORACLE()
    if op == "qsort":
        T.assert_sorted_equal(sk, O.qsort(keys, tag, hasnan, desc), tag, w + " vs oracle")
    elif op == "keyvalue_qsort":
        ek, ev = O.keyvalue_qsort(keys, vals, tag, vtag, hasnan, desc)
        T.assert_kv_equal(sk, sv, ek, ev, tag, w + " vs oracle")
    elif op == "qselect":
        if keys.size:
            T.assert_selected(sk, keys, O.qselect(keys, tag, 0, hasnan, desc), k, tag, w + " vs oracle")
    else:
        exp = O.partial_qsort(keys, tag, k, hasnan, desc)
        T.assert_prefix_equal(sk, exp, k, tag, w + " vs oracle")
        assert np.array_equal(np.sort(T.vclass(sk, tag)), np.sort(T.vclass(exp, tag))), w + " vs oracle"


def _compare(mix, alone, served, what):
    for i, (item, (ak, av), (sk, sv)) in enumerate(zip(mix, alone, served)):
        op, tag, vtag, keys, vals, k, hasnan, desc = item
        w = f"{what} request {i}: {op}<{tag}{',' + vtag if vtag else ''}> n={keys.size} k={k} " \
            f"hasnan={hasnan} desc={desc}"
        _oracle_check(item, sk, sv, w)
        if op == "qsort":
            assert np.array_equal(T.bits(sk), T.bits(ak)), w + ": keys differ"
        elif op == "keyvalue_qsort":
            # the same keys bit for bit; payloads as a multiset per run of
            # equal keys (a coalesced launch may order a run differently)
            assert np.array_equal(T.bits(sk), T.bits(ak)), w + ": keys differ"
            T.assert_kv_equal(sk, sv, ak, av, tag, w + " vs alone")
        elif op == "qselect":
            assert np.array_equal(np.sort(T.bits(sk)), np.sort(T.bits(keys))), w + ": not a permutation"
            cs, ca = T.vclass(sk, tag), T.vclass(ak, tag)
            assert cs[k] == ca[k], w + ": k-th element differs"
            assert np.array_equal(np.sort(cs[:k]), np.sort(ca[:k])), w + ": left partition differs"
            assert np.array_equal(np.sort(cs[k + 1:]), np.sort(ca[k + 1:])), w + ": right partition differs"
        else:
            T.assert_prefix_equal(sk, ak, k, tag, w)
            assert np.array_equal(np.sort(T.vclass(sk, tag)), np.sort(T.vclass(ak, tag))), w


@pytest.fixture(scope="module")
def mix_and_alone(xss):
    mix = _mix(240, 31)
    return mix, [_alone(xss, it) for it in mix]


def test_serving_burst_coalesced(xss, mix_and_alone):
    mix, alone = mix_and_alone
    served, stats = _serve(xss, mix, budget_bytes=64 << 30, streams=4, coalesce_max=1 << 20)
    _compare(mix, alone, served, "burst")
    assert stats["completed"] == len(mix)
    assert stats["coalesced"] > 0, stats  # requests queued together shared launch groups
    assert stats["launch_groups"] < len(mix)


def test_serving_burst_short_only(xss, mix_and_alone):
    """Coalescing limited to what one LDS workgroup sorts (pointer-list form)."""
    mix, alone = mix_and_alone
    served, stats = _serve(xss, mix, budget_bytes=64 << 30, streams=4, coalesce_max=4096)
    _compare(mix, alone, served, "burst short")
    assert stats["completed"] == len(mix) and stats["coalesced"] > 0, stats


def _medium_mix(count, seed):
    """Medium requests only (longer than an LDS sort): qsort<f32> with NaNs
    and qsort<u16> descending of 4097..2^20 keys, keyvalue_qsort<u64,u64>
    and <f16,u32> of 2049..2^19 keys."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        c = i % 4
        if c < 2:
            tag, hasnan, desc = ("f32", True, False) if c == 0 else ("u16", False, True)
            n = int(np.exp(rng.uniform(np.log(4097), np.log(1 << 20))))
            keys = T.gen(tag, "nan" if c == 0 else "dups", n, 500 + i)
            out.append(("qsort", tag, None, keys, None, 0, hasnan, desc))
        else:
            tag, vtag = ("u64", "u64") if c == 2 else ("f16", "u32")
            n = int(np.exp(rng.uniform(np.log(2049), np.log(1 << 19))))
            keys = T.gen(tag, "dups" if c == 2 else "nan_payloads", n, 600 + i)
            vals = rng.integers(0, 2 ** 31, n).astype(T.TAGS[vtag])
            out.append(("keyvalue_qsort", tag, vtag, keys, vals, 0, c == 3, bool(i % 8 == 7)))
    return out


def test_serving_medium_groups(xss):
    """Medium requests queued together share gathered segmented passes: every
    result equals the plain call's and the oracle's."""
    mix = _medium_mix(48, 91)
    alone = [_alone(xss, it) for it in mix]
    served, stats = _serve(xss, mix, budget_bytes=64 << 30, streams=2, coalesce_max=1 << 20)
    _compare(mix, alone, served, "medium groups")
    assert stats["completed"] == len(mix)
    assert stats["coalesced"] >= len(mix) // 2, stats
    assert stats["launch_groups"] <= len(mix) // 2, stats


def test_serving_tight_budget(xss, mix_and_alone):
    mix, alone = mix_and_alone
    # 8 MiB admits a few requests at a time; larger ones run alone when nothing else is in flight
    served, stats = _serve(xss, mix, budget_bytes=8 << 20, streams=2, coalesce_max=1 << 20)
    _compare(mix, alone, served, "tight budget")
    assert stats["completed"] == len(mix)


def test_serving_one_at_a_time(xss, mix_and_alone):
    mix, alone = mix_and_alone
    served, stats = _serve(xss, mix, budget_bytes=64 << 30, streams=1, coalesce_max=0, one_at_a_time=True)
    _compare(mix, alone, served, "one at a time")
    assert stats["launch_groups"] == len(mix) and stats["coalesced"] == 0


def test_serving_poisson_open_loop(xss):
    """Requests submitted at Poisson arrival times while earlier ones run."""
    import torch
    mix = _mix(60, 77)
    alone = [_alone(xss, it) for it in mix]
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(200.0, len(mix)))  # 5k req/s
    arrays, reqs = [], []
    for (op, tag, vtag, keys, vals, k, hasnan, desc), t in zip(mix, arrivals):
        d = xss.DArray.from_numpy(keys)
        v = xss.DArray.from_numpy(vals) if vals is not None else None
        arrays.append((d, v))
        reqs.append(xss.make_request(op, d, v, k, hasnan, desc, arrival_us=t))
    torch.cuda.synchronize()
    with xss.Server(streams=4) as sv:
        done = sv.run_trace(reqs)
    lat = np.array([r.latency_us for r in done])
    assert (lat >= 0).all()
    served = [(d.numpy(), v.numpy() if v is not None else None) for d, v in arrays]
    _compare(mix, alone, served, "poisson")


def _c5_trace(count, seed):
    """Requests shaped like the served workload: (op, tag, vtag, keys, vals, k)
    with keys and values as torch tensors on the GPU."""
    import torch
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = []
    for _ in range(count):
        u = rng.random()
        if u < 0.7:
            n = int(np.exp(rng.uniform(np.log(1 << 10), np.log(16 << 20))))
            keys = (torch.rand(n, device="cuda", generator=g) * 2 - 1).view(torch.int32)
            out.append(("qsort", "f32", None, keys, None, 0))
        elif u < 0.9:
            n = int(np.exp(rng.uniform(np.log(1 << 10), np.log(4 << 20))))
            z = rng.zipf(1.1, n).astype(np.uint64)
            z[z >= (1 << 32)] = np.uint64((1 << 32) - 1)
            keys = torch.from_numpy(z.view(np.int64)).cuda()
            vals = torch.arange(n, dtype=torch.int64, device="cuda")
            out.append(("keyvalue_qsort", "u64", "u64", keys, vals, 0))
        else:
            n = int(np.exp(rng.uniform(np.log(1 << 10), np.log(16 << 20))))
            keys = torch.randint(-2 ** 31, 2 ** 31, (n,), dtype=torch.int32, device="cuda", generator=g)
            op = "qselect" if rng.random() < 0.5 else "partial_qsort"
            k = int(rng.integers(1, n + 1))
            out.append((op, "u32", None, keys, None, k - 1 if op == "qselect" else k))
    return out


def test_serving_c5_trace(xss):
    """An open-loop Poisson trace at 10k req/s of the served workload's mix;
    every result checked: against the oracle up to 2^20 keys, through the
    full-size properties above."""
    import torch
    import xss_devcheck as D
    trace = _c5_trace(120, 505)
    rng = np.random.default_rng(6)
    arrivals = np.cumsum(rng.exponential(100.0, len(trace)))
    work, reqs = [], []
    for (op, tag, vtag, keys, vals, k), t in zip(trace, arrivals):
        gk = keys.clone()
        gv = vals.clone() if vals is not None else None
        dk = xss.DArray(tag, gk.numel(), gk.view(torch.uint8))
        dv = xss.DArray(vtag, gv.numel(), gv.view(torch.uint8)) if gv is not None else None
        work.append((gk, gv))
        reqs.append(xss.make_request(op, dk, dv, k, False, False, arrival_us=t))
    torch.cuda.synchronize()
    with xss.Server(budget_bytes=64 << 30, streams=4) as sv:
        done = sv.run_trace(reqs)
        stats = sv.stats()
    assert stats["completed"] == len(trace)
    assert all(r.latency_us >= 0 for r in done)
    O = T.ORACLE()
    for i, ((op, tag, vtag, keys, vals, k), (gk, gv)) in enumerate(zip(trace, work)):
        n = keys.numel()
        w = f"c5 request {i}: {op}<{tag}> n={n} k={k}"
        if n <= (1 << 20):
            hk = keys.cpu().numpy().view(T.TAGS[tag])
            sk = gk.cpu().numpy().view(T.TAGS[tag])
            if op == "keyvalue_qsort":
                hv = vals.cpu().numpy().view(np.uint64)
                _oracle_check((op, tag, vtag, hk, hv, k, False, False), sk, gv.cpu().numpy().view(np.uint64), w)
            else:
                _oracle_check((op, tag, vtag, hk, None, k, False, False), sk, None, w)
            continue
        if op == "qsort":
            D.check_sort(gk, keys, tag, False, False, True, w)
        elif op == "keyvalue_qsort":
            D.check_gather(gk, keys, gv, w)
            D.check_sort(gk, keys, tag, False, False, False, w)
        else:
            srt = torch.sort(D.order_key(keys, tag)).values
            og = D.order_key(gk, tag)
            assert torch.equal(torch.sort(og).values, srt), w + ": not a permutation"
            if op == "qselect":
                assert og[k] == srt[k], w + ": k-th value"
                assert bool((og[:k] <= srt[k]).all()) and bool((og[k + 1:] >= srt[k]).all()), w + ": partition"
            else:
                assert torch.equal(og[:k], srt[:k]), w + ": sorted prefix"
                assert bool((og[k:] >= srt[k - 1]).all()), w + ": partition"
