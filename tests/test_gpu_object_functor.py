"""x86simdsort::gpu::object_qsort (include/x86simdsort-gpu.h): the key
functor runs on the device, instantiated by a caller's hipcc translation
unit (tests/cpp/object_gpu.hip, built into tests/cpp/libxss_objtest.so).

The reference's object_qsort (lib/x86simdsort.h:72-106) orders the objects
by key_func ascending; its key-value sort leaves equal keys in any order.
This form argsorts stably, so the expected result is fixed: the objects
gathered by a stable argsort of the keys, computed on the host by the same
IEEE operations (the test library is built with -ffp-contract=off).  The
Euclidean-norm case is also compared with the oracle's object sort.
"""
import ctypes
import os

import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

LIB = os.path.join(T.REPO, "tests", "cpp", "libxss_objtest.so")
REC = np.dtype([("id", "<u4"), ("score", "<f4"), ("rank", "<i4"), ("a", "<u4"), ("b", "<u4")])


@pytest.fixture(scope="module")
def objlib(xss):
    assert os.path.exists(LIB), "tests/cpp/libxss_objtest.so is not built (run build())"
    L = ctypes.CDLL(LIB)
    for f in ("objtest_xyz_norm", "objtest_rec_score", "objtest_rec_rank", "objtest_rec_ab"):
        getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    L.objtest_rec_id.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    return L


def _run(objlib, fn, host, n_arg=None):
    import torch
    d = torch.from_numpy(host.view(np.uint8).copy()).cuda()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(objlib, fn)(ctypes.c_void_p(d.data_ptr()), host.shape[0] if n_arg is None else n_arg, s)
    torch.cuda.synchronize()
    assert rc == 0, f"{fn}: status {rc}"
    return d.cpu().numpy().view(host.dtype).reshape(host.shape)


def _records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, REC)
    r["id"] = rng.permutation(n).astype(np.uint32) if n else r["id"]
    r["score"] = rng.integers(-500, 500, n).astype(np.float32) / np.float32(8)  # ties, no -0.0
    r["score"][r["score"] == 0] = np.float32(0.5)
    r["rank"] = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32)
    r["a"] = rng.integers(0, 64, n).astype(np.uint32)
    r["b"] = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    return r


@pytest.mark.parametrize("n", [0, 1, 2, 17, 4096, 4097, 100_000, 1 << 22])
def test_xyz_norm_device_key(objlib, n):
    xyz = np.random.default_rng(n + 3).uniform(-1, 1, (n, 3))
    got = _run(objlib, "objtest_xyz_norm", xyz)
    keys = np.sqrt((xyz[:, 0] * xyz[:, 0] + xyz[:, 1] * xyz[:, 1]) + xyz[:, 2] * xyz[:, 2])
    exp = xyz[np.argsort(keys, kind="stable")]
    assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), f"xyz n={n}"
    if 0 < n <= 100_000:  # the oracle's object_qsort of the same workload
        o = xyz.copy()
        ok = np.zeros(n)
        T.oracle().xss_oracle_object_qsort_xyz(ctypes.c_void_p(o.ctypes.data), ctypes.c_size_t(n),
                                               ctypes.c_void_p(ok.ctypes.data))
        gk = np.sqrt((got[:, 0] * got[:, 0] + got[:, 1] * got[:, 1]) + got[:, 2] * got[:, 2])
        assert np.array_equal(gk, ok), "keys in sorted order differ from the oracle's"


@pytest.mark.parametrize("n", [2, 3000, 4097, 250_000])
@pytest.mark.parametrize("fn,key", [("objtest_rec_score", lambda r: r["score"]),
                                    ("objtest_rec_rank", lambda r: r["rank"]),
                                    ("objtest_rec_ab", lambda r: (r["a"].astype(np.uint64) << np.uint64(32))
                                     | r["b"].astype(np.uint64)),
                                    ("objtest_rec_id", lambda r: r["id"])])
def test_records_device_key(objlib, fn, key, n):
    rec = _records(n, n + len(fn))
    got = _run(objlib, fn, rec)
    exp = rec[np.argsort(key(rec), kind="stable")]
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), f"{fn} n={n}"
