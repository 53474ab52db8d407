"""GPU checks of the PyTorch shim (torch.ops.xss.sort / argsort, in
x86-simd-sort_amd/xss_torch.py) against torch.sort's contract, computed on
the host with numpy: the values equal the sorted input numerically (NaN last
ascending, first descending; -0.0 and +0.0 equal), and the indices are a
permutation of each row that gathers exactly those values, bit for bit.
One row and many rows (segmented plan), every dtype, both directions."""
import numpy as np
import pytest
import torch

import xss_testlib as T

pytestmark = pytest.mark.gpu

DTYPES = {"u16": torch.uint16, "i16": torch.int16, "f16": torch.float16, "u32": torch.uint32, "i32": torch.int32,
          "f32": torch.float32, "u64": torch.uint64, "i64": torch.int64, "f64": torch.float64}
SHAPES = [(1,), (2,), (1000,), (100003,), (1, 5000), (7, 3001), (300, 17), (3, 1 << 18)]
DISTS = ["random", "dups", "signed_zeros", "rand_with_max_and_nan", "nan_payloads"]
UINT = {2: np.uint16, 4: np.uint32, 8: np.uint64}


@pytest.fixture(scope="module")
def xt(xss):
    import xss_torch
    return xss_torch


def _expect(a, desc):
    s = np.sort(a, axis=-1, kind="stable")
    return s[..., ::-1] if desc else s


def _check(a, vals, idx, desc, what):
    exp = _expect(a, desc)
    assert vals.shape == a.shape and idx.shape == a.shape, what
    if a.dtype.kind == "f":
        assert np.array_equal(vals, exp, equal_nan=True), what + ": values"
    else:
        assert np.array_equal(vals, exp), what + ": values"
    assert (np.sort(idx, axis=-1) == np.arange(a.shape[-1])).all(), what + ": not a permutation of each row"
    g = np.take_along_axis(a, idx, axis=-1)
    u = UINT[a.dtype.itemsize]
    assert np.array_equal(g.view(u), vals.view(u)), what + ": indices do not gather the values"


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("tag", list(DTYPES))
def test_torch_sort_and_argsort(xt, tag, desc):
    for seed, (dist, shape) in enumerate((d, s) for d in DISTS for s in SHAPES):
        if dist in ("signed_zeros", "nan_payloads") and tag not in ("f16", "f32", "f64"):
            continue
        n = int(np.prod(shape))
        a = T.gen(tag, dist, n, 16000 + seed).reshape(shape)
        x = torch.from_numpy(a.copy()).cuda()
        what = f"xss::sort<{tag}> {dist} shape={shape} desc={desc}"
        vals, idx = torch.ops.xss.sort(x, desc)
        _check(a, vals.cpu().numpy(), idx.cpu().numpy(), desc, what)
        idx2 = torch.ops.xss.argsort(x, desc).cpu().numpy()
        _check(a, np.take_along_axis(a, idx2, axis=-1), idx2, desc, what.replace("sort", "argsort"))
        assert np.array_equal(x.cpu().numpy().view(UINT[a.dtype.itemsize]), a.view(UINT[a.dtype.itemsize])), \
            what + ": input modified"


def test_torch_ops_reject_what_they_cannot_sort(xt):
    with pytest.raises(TypeError):
        torch.ops.xss.sort(torch.zeros(4, dtype=torch.bool, device="cuda"), False)
    with pytest.raises(ValueError):
        torch.ops.xss.argsort(torch.zeros((), device="cuda"), False)


def test_torch_sort_matches_torch_on_floats(xt):
    """On data without ties, -0.0 or NaN the result is torch.sort's exactly."""
    x = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
    v, i = torch.ops.xss.sort(x, True)
    rv, ri = torch.sort(x, descending=True)
    assert torch.equal(v, rv) and torch.equal(i, ri)
    x = torch.randn(64, 4096, device="cuda", dtype=torch.float32)
    v, i = torch.ops.xss.sort(x, False)
    rv, ri = torch.sort(x)
    assert torch.equal(v, rv) and torch.equal(i, ri)
