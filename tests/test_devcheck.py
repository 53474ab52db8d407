"""The device-side checkers of the full-size tests (xss_devcheck), run on the
CPU: their order key against the oracle's comparator, the recipes against
their definitions, and planted mistakes that each checker must reject."""
import numpy as np
import pytest
import torch

import xss_devcheck as D
import xss_testlib as T


def t_of(a):
    nb = a.dtype.itemsize * 8
    return torch.from_numpy(T.bits(np.ascontiguousarray(a)).view({16: np.int16, 32: np.int32, 64: np.int64}[nb]).copy())


def np_of(x, tag):
    return x.numpy().view(T.TAGS[tag])


def edge_values(tag, rng, n=4000):
    a = T.gen(tag, "random", n, 3)
    extra = T.gen(tag, "nan_payloads", n, 4) if tag in T.FLOAT_TAGS else T.gen(tag, "rand_max", n, 4)
    z = T.gen(tag, "signed_zeros", n, 5)
    return np.concatenate([a, extra, z, T.gen(tag, "dups", n, 6)])


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", list(T.TAGS))
def test_order_key_matches_oracle(tag, desc):
    """Sorting by the order key gives the oracle's (reference comparator)
    order of value classes; all 2^16 patterns for 16-bit types."""
    O = T.ORACLE()
    if D.WIDTH[tag] == 16:
        a = np.arange(1 << 16, dtype=np.uint16).view(T.TAGS[tag])
    else:
        a = edge_values(tag, np.random.default_rng(1))
    hasnan = tag in T.FLOAT_TAGS
    exp = a[O.argsort(a, tag, hasnan, desc).astype(np.int64)]
    order = torch.argsort(D.order_key(t_of(a), tag, desc), stable=True).numpy()
    got = a[order]
    assert np.array_equal(T.vclass(got, tag), T.vclass(exp, tag))


@pytest.mark.parametrize("tag", list(T.TAGS))
def test_recipes(tag):
    n = 50_000
    r = D.gen(tag, "random", n, 1, device="cpu")
    assert not bool(D.is_nan(r, tag).any())
    if tag in T.FLOAT_TAGS:  # finite
        nb = D.WIDTH[tag]
        assert not bool(((r & D.EXP[nb]) == D.EXP[nb]).any())
    assert len(torch.unique(D.gen(tag, "dups", n, 2, device="cpu"))) <= 1024
    s = D.gen(tag, "sorted", n, 3, device="cpu")
    o = D.order_key(s, tag)
    assert bool((o[1:] >= o[:-1]).all())
    rv = D.order_key(D.gen(tag, "reverse", n, 3, device="cpu"), tag)
    assert bool((rv[1:] <= rv[:-1]).all())
    if tag in T.FLOAT_TAGS:
        x = D.gen(tag, "nan", n, 4, device="cpu")
        frac = float(D.is_nan(x, tag).float().mean())
        assert 0.005 < frac < 0.015
        assert bool((x[D.is_nan(x, tag)] == D.QNAN[D.WIDTH[tag]]).all())


@pytest.mark.parametrize("tag", ["f32", "u64", "i16", "f16", "f64"])
def test_check_sort_accepts_and_rejects(tag):
    """check_sort accepts the oracle's sort and rejects planted mistakes."""
    O = T.ORACLE()
    dist = "nan" if tag in T.FLOAT_TAGS else "dups"
    a = T.gen(tag, dist, 20_000, 9)
    z = T.gen(tag, "signed_zeros", 2000, 10)
    a = np.concatenate([a, z])
    hasnan = tag in T.FLOAT_TAGS
    for desc in (False, True):
        good = O.qsort(a, tag, hasnan, desc)
        inp, g = t_of(a), t_of(good)
        D.check_sort(g, inp, tag, hasnan, desc, True, "oracle")
        # two neighbours of different value swapped
        o = D.order_key(g, tag, desc)
        i = int(torch.nonzero(o[1:] != o[:-1])[0]) if bool((o[1:] != o[:-1]).any()) else 0
        bad = g.clone()
        bad[i], bad[i + 1] = g[i + 1], g[i]
        with pytest.raises(AssertionError):
            D.check_sort(bad, inp, tag, hasnan, desc, True, "swap")
        # one key replaced by its neighbour's value (no longer a permutation)
        bad = g.clone()
        bad[i] = g[i + 1]
        with pytest.raises(AssertionError):
            D.check_sort(bad, inp, tag, hasnan, desc, True, "dup")
        if tag in T.FLOAT_TAGS:
            nb = D.WIDTH[tag]
            bad = g.clone()
            j = int(torch.nonzero(g == -(1 << (nb - 1)))[0])  # a -0.0 written as +0.0
            bad[j] = 0
            with pytest.raises(AssertionError):
                D.check_sort(bad, inp, tag, hasnan, desc, True, "zero")
            bad = g.clone()
            j = int(torch.nonzero(D.is_nan(g, tag))[0])  # a NaN with a payload
            bad[j] = D.QNAN[nb] + 1
            with pytest.raises(AssertionError):
                D.check_sort(bad, inp, tag, hasnan, desc, True, "nan bits")


def test_check_gather_rejects():
    k = torch.randint(-100, 100, (5000,), dtype=torch.int32)
    idx = torch.argsort(k, stable=True)
    D.check_gather(k[idx], k, idx, "good")
    bad = idx.clone()
    bad[3] = bad[4]
    with pytest.raises(AssertionError):
        D.check_gather(k[bad], k, bad, "duplicate index")
    with pytest.raises(AssertionError):
        D.check_gather(k[idx].flip(0), k, idx, "keys not gathered")


def test_check_segments():
    rng = np.random.default_rng(3)
    sizes = rng.integers(0, 300, 40)
    off = np.concatenate([[0], np.cumsum(sizes)])
    a = T.gen("f32", "nan", int(off[-1]), 5)
    O = T.ORACLE()
    good = np.concatenate([O.qsort(a[off[i]:off[i + 1]], "f32", True) for i in range(40)])
    seg = D.seg_ids(off, "cpu")
    D.check_segments_sorted_perm(t_of(good), t_of(a), seg, "f32", False, "good")
    # the whole array sorted is ordered, but keys crossed segment boundaries
    whole = O.qsort(a, "f32", True)
    with pytest.raises(AssertionError):
        D.check_segments_sorted_perm(t_of(whole), t_of(a), seg, "f32", False, "crossed")


def test_kv_partial_checker():
    """assert_kv_partial accepts the oracle's partial sort and rejects values
    moved between runs of the prefix or split from their keys."""
    O = T.ORACLE()
    k = T.gen("u32", "dups", 5000, 21) % np.uint32(50)
    v = np.arange(5000, dtype=np.uint64)
    ek, ev = O.keyvalue_qsort(k, v, "u32", "u64")
    for kk in (1, 777, 2500, 5000):
        gk, gv = O.keyvalue_partial_sort(k, v, "u32", "u64", kk)
        T.assert_kv_partial(gk, gv, kk, k, v, ek, ev, "u32", "oracle")
    gk, gv = O.keyvalue_partial_sort(k, v, "u32", "u64", 2500)
    ce = T.vclass(gk, "u32")
    i = int(np.nonzero(ce[1:2400] != ce[:2399])[0][0])  # last of a run / first of the next
    bad = gv.copy()
    bad[i], bad[i + 1] = gv[i + 1], gv[i]  # two values traded across a run boundary
    with pytest.raises(AssertionError):
        T.assert_kv_partial(gk, bad, 2500, k, v, ek, ev, "u32", "traded")
    # a value in the cut run swapped with one beyond the prefix that has another key
    j = 2499
    far = int(np.nonzero(ce != ce[j])[0][-1])
    bk, bv = gk.copy(), gv.copy()
    bv[j], bv[far] = gv[far], gv[j]
    with pytest.raises(AssertionError):
        T.assert_kv_partial(bk, bv, 2500, k, v, ek, ev, "u32", "cut run")


def test_check_sorted_perm_in_chunks(monkeypatch):
    """The chunked 16-bit path (arrays beyond one torch.sort) accepts a sort
    and rejects a swap and a changed key; run here with small chunks."""
    monkeypatch.setattr(D, "CHUNK", 1000)
    monkeypatch.setattr(D, "SORT_MAX", 999)
    O = T.ORACLE()
    a = T.gen("i16", "dups", 5555, 3)
    good = t_of(O.qsort(a, "i16"))
    D.check_sorted_perm(good, t_of(a), "i16", False, "good")
    o = D.order_key(good, "i16")
    i = int(torch.nonzero(o[1:] != o[:-1])[2])
    bad = good.clone()
    bad[i], bad[i + 1] = good[i + 1], good[i]
    with pytest.raises(AssertionError):
        D.check_sorted_perm(bad, t_of(a), "i16", False, "swap")
    bad = good.clone()
    bad[-1] = bad[0]
    with pytest.raises(AssertionError):
        D.check_sorted_perm(bad, t_of(a), "i16", False, "changed")
    idx = torch.randperm(5555)
    assert torch.equal(D.gather(good, idx), good[idx])
