"""GPU tests of the drop-in C++ API (host pointers, the reference's
signatures) and of object_qsort."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

PKG = os.path.join(T.REPO, "x86-simd-sort_amd")


def test_cpp_dropin_program(tmp_path):
    exe = tmp_path / "dropin"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(T.REPO, "include"),
                    os.path.join(T.REPO, "tests", "cpp", "dropin.cpp"), "-o", str(exe), "-L", PKG, "-lxss_amd",
                    f"-Wl,-rpath,{PKG}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "dropin ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("n", [1, 2, 100, 4097, 300001])
def test_object_qsort_xyz_norm(xss, n):
    """Device key functor: keys equal the host's IEEE sqrt(x*x+y*y+z*z) bit
    for bit, objects permuted like the oracle's object_qsort."""
    import torch
    rng = np.random.default_rng(n)
    xyz = rng.uniform(-1, 1, (n, 3))
    a = xyz.copy()
    keys = np.zeros(n)
    T.oracle().xss_oracle_object_qsort_xyz(ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(n),
                                          ctypes.c_void_p(keys.ctypes.data))
    d = torch.from_numpy(xyz.copy()).cuda()
    dk = torch.empty(n, dtype=torch.float64, device="cuda")
    rc = xss.lib().xss_gpu_object_qsort_xyz_norm(ctypes.c_void_p(d.data_ptr()), ctypes.c_size_t(n),
                                                 ctypes.c_void_p(dk.data_ptr()),
                                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    got, gk = d.cpu().numpy(), dk.cpu().numpy()
    assert np.array_equal(gk.view(np.uint64), keys.view(np.uint64)), "device keys differ from host IEEE keys"
    # objects with equal keys may come in either order: compare (key, object) multisets per position run
    assert np.array_equal(np.sort(got.view(np.uint64), axis=0), np.sort(a.view(np.uint64), axis=0))
    kg = np.sqrt(got[:, 0] * got[:, 0] + got[:, 1] * got[:, 1] + got[:, 2] * got[:, 2])
    assert np.array_equal(kg, keys)


@pytest.mark.parametrize("ktag", T.KV_TAGS)
def test_object_sort_by_keys(xss, ktag):
    import torch
    for n in [2, 1000, 70001]:
        keys = T.gen(ktag, "dups", n, 77 + n)
        objs = np.random.default_rng(n).integers(0, 255, (n, 20), dtype=np.uint8)  # 20-byte objects
        d = torch.from_numpy(objs.copy()).cuda()
        dk = xss.DArray.from_numpy(keys)
        xss.object_sort_by_keys(d, 20, dk)
        got = d.cpu().numpy()
        order = T.ORACLE().argsort(keys, ktag).astype(np.int64)
        exp = objs[order]
        sk = keys[order]
        # within runs of equal keys the order of objects is free: compare per run as multisets
        cls = T.vclass(sk, ktag)
        run = np.concatenate([[0], np.cumsum(cls[1:] != cls[:-1])])
        g2 = np.column_stack([run, got]).astype(np.int64)
        e2 = np.column_stack([run, exp]).astype(np.int64)
        assert np.array_equal(g2[np.lexsort(g2.T[::-1])], e2[np.lexsort(e2.T[::-1])])
        assert np.array_equal(T.bits(dk.numpy()), T.bits(keys)), "keys must be left unchanged"
