"""The error channel of the C ABI (the reference has none: its functions
return void and leave k >= n undefined, src/xss-common-qsort.h:740):
XSS_ERR_INVALID for sizes over the limit, null pointers and k out of range,
XSS_ERR_CAPTURE for a workspace that would have to grow inside a graph
capture, XSS_ERR_NOMEM for a refused workspace; and the host API from many
threads at once, whose errors surface as std::runtime_error."""
import ctypes
import os
import subprocess

import pytest

import xss_testlib as T

pytestmark = pytest.mark.gpu

MAX_N = 0xFF000000
OK, INVALID, HIP, NOMEM, CAPTURE = 0, 1, 2, 3, 4
PKG = os.path.join(T.REPO, "x86-simd-sort_amd")
DUMMY = ctypes.c_void_p(0x1000)  # never dereferenced: every case is refused before any launch


def _z(x):
    return ctypes.c_size_t(x)


@pytest.fixture
def L(xss):
    L = xss.lib()
    L.xss_gpu_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
    return L


def _s():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("tag", list(T.TAGS))
def test_sizes_over_the_limit(L, tag):
    n = _z(MAX_N + 1)
    for rc in (getattr(L, f"xss_gpu_qsort_{tag}")(DUMMY, n, 0, 0, _s()),
               getattr(L, f"xss_gpu_qselect_{tag}")(DUMMY, _z(5), n, 0, 0, _s()),
               getattr(L, f"xss_gpu_partial_qsort_{tag}")(DUMMY, _z(5), n, 0, 0, _s()),
               getattr(L, f"xss_gpu_argsort_{tag}")(DUMMY, DUMMY, n, 0, 0, _s()),
               getattr(L, f"xss_gpu_argselect_{tag}")(DUMMY, DUMMY, _z(5), n, 0, _s()),
               getattr(L, f"xss_gpu_argsort_arg_{tag}")(DUMMY, DUMMY, n, 0, 0, _s())):
        assert rc == INVALID
        assert b"limit" in L.xss_gpu_last_error()
    if tag in T.KV_TAGS:
        for vtag in T.KV_TAGS:
            assert getattr(L, f"xss_gpu_keyvalue_qsort_{tag}_{vtag}")(DUMMY, DUMMY, n, 0, 0, _s()) == INVALID
            assert getattr(L, f"xss_gpu_keyvalue_select_{tag}_{vtag}")(DUMMY, DUMMY, _z(1), n, 0, 0, _s()) == INVALID


def test_null_pointers_and_k(L):
    n = _z(1000)
    assert L.xss_gpu_qsort_f32(None, n, 0, 0, _s()) == INVALID
    assert L.xss_gpu_argsort_u64(DUMMY, None, n, 0, 0, _s()) == INVALID
    assert L.xss_gpu_keyvalue_qsort_f32_u32(DUMMY, None, n, 0, 0, _s()) == INVALID
    assert L.xss_gpu_qselect_i32(None, _z(3), n, 0, 0, _s()) == INVALID
    assert L.xss_gpu_qselect_i32(DUMMY, _z(1000), n, 0, 0, _s()) == INVALID   # k == n
    assert L.xss_gpu_argselect_f64(DUMMY, DUMMY, _z(5000), n, 0, _s()) == INVALID
    assert L.xss_gpu_partial_qsort_u16(DUMMY, _z(1001), n, 0, 0, _s()) == INVALID  # k > n
    assert L.xss_gpu_keyvalue_select_u64_u64(DUMMY, DUMMY, _z(1000), n, 0, 0, _s()) == INVALID
    assert b"out of range" in L.xss_gpu_last_error()
    # n <= 1 and partial k == 0 are no-ops (src/xss-common-qsort.h:660,757), even with null pointers
    assert L.xss_gpu_qsort_f32(None, _z(1), 0, 0, _s()) == OK
    assert L.xss_gpu_partial_qsort_f32(None, _z(0), n, 0, 0, _s()) == OK


def test_workspace_growth_refused_during_capture(xss, L):
    """A call whose stream's workspace would have to grow while the stream
    is captured returns XSS_ERR_CAPTURE instead of allocating; after one
    eager call on that stream the same call captures."""
    import torch
    a = xss.DArray.from_numpy(T.gen("f32", "random", 1 << 20, 3))
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    rc = None
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        rc = L.xss_gpu_qsort_f32(a.ptr, _z(a.n), 0, 0, ctypes.c_void_p(s.cuda_stream))
        # capture needs some work recorded
        a.t[:1].add_(0)
    assert rc == CAPTURE, L.xss_gpu_last_error()
    assert b"capture" in L.xss_gpu_last_error()
    with torch.cuda.stream(s):
        xss.qsort(a)
    s.synchronize()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=s):
        rc = L.xss_gpu_qsort_f32(a.ptr, _z(a.n), 0, 0, ctypes.c_void_p(s.cuda_stream))
    assert rc == OK


def test_workspace_limit_gives_nomem(xss, L):
    import torch
    a = xss.DArray.from_numpy(T.gen("u32", "random", 1 << 21, 4))
    s = torch.cuda.Stream()  # a fresh stream: its workspace has to grow
    assert L.xss_gpu_set_tuning(b"workspace_limit", 1 << 20) == OK
    try:
        rc = L.xss_gpu_qsort_u32(a.ptr, _z(a.n), 0, 0, ctypes.c_void_p(s.cuda_stream))
        msg = L.xss_gpu_last_error()
    finally:
        L.xss_gpu_set_tuning(b"workspace_limit", 0)
    assert rc == NOMEM and b"limit" in msg, msg
    rc = L.xss_gpu_qsort_u32(a.ptr, _z(a.n), 0, 0, ctypes.c_void_p(s.cuda_stream))
    s.synchronize()
    assert rc == OK
    got = a.numpy()
    assert (got[1:] >= got[:-1]).all()


def test_cpp_threads_and_exceptions(tmp_path):
    exe = tmp_path / "threads"
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-I", os.path.join(T.REPO, "include"),
                    os.path.join(T.REPO, "tests", "cpp", "threads.cpp"), "-o", str(exe), "-L", PKG, "-lxss_amd",
                    f"-Wl,-rpath,{PKG}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "threads ok" in r.stdout, r.stdout + r.stderr
