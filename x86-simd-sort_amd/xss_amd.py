"""ctypes binding of libxss_amd.so (the C ABI declared in include/xss_gpu.h).

Host-side mirror of the x86simdsort:: API for Python callers, tests and the
benchmark.  Arrays live on the GPU as torch tensors; every call enqueues on
torch's current stream and is asynchronous like the C entry points.  The
functions keep the reference's names and argument meaning
(lib/x86simdsort.h:15-70): qsort, qselect, partial_qsort, argsort, argselect,
keyvalue_qsort, keyvalue_select, keyvalue_partial_sort, object_qsort.

There is no CPU fallback: if the library is missing or a call fails, a
RuntimeError is raised.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxss_amd.so")

# type tag <-> numpy dtype (_Float16 is numpy float16)
TAGS = {
    "u16": np.uint16, "i16": np.int16, "f16": np.float16,
    "u32": np.uint32, "i32": np.int32, "f32": np.float32,
    "u64": np.uint64, "i64": np.int64, "f64": np.float64,
}
KV_TAGS = ["u32", "i32", "f32", "u64", "i64", "f64"]
# keyvalue_* key types: the reference's six, and 16-bit keys (no reference counterpart)
KV_KEY_TAGS = ["u16", "i16", "f16"] + KV_TAGS
TAG_OF = {np.dtype(v): k for k, v in TAGS.items()}

XSS_OK = 0

_lib = None


def lib():
    """Load libxss_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run x86-simd-sort_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        L.xss_gpu_last_error.restype = ctypes.c_char_p
        L.xss_gpu_version.restype = ctypes.c_char_p
        L.xss_gpu_workspace_bytes.restype = ctypes.c_size_t
        _lib = L
    return _lib


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(rc, what):
    if rc != XSS_OK:
        msg = lib().xss_gpu_last_error().decode()
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


class DArray:
    """A typed 1-D array in GPU memory (raw bytes in a torch tensor)."""

    def __init__(self, tag, n, tensor):
        self.tag, self.n, self.t = tag, n, tensor

    @classmethod
    def from_numpy(cls, a):
        import torch
        a = np.ascontiguousarray(a)
        tag = TAG_OF[a.dtype]
        t = torch.from_numpy(a.view(np.uint8).reshape(-1).copy()).cuda()
        return cls(tag, a.size, t)

    @classmethod
    def empty(cls, tag, n):
        import torch
        item = np.dtype(TAGS[tag]).itemsize
        return cls(tag, n, torch.empty(max(1, n * item), dtype=torch.uint8, device="cuda"))

    @property
    def ptr(self):
        return ctypes.c_void_p(self.t.data_ptr())

    def numpy(self):
        import torch
        torch.cuda.current_stream().synchronize()
        item = np.dtype(TAGS[self.tag]).itemsize
        return self.t[: self.n * item].cpu().numpy().view(TAGS[self.tag]).copy()


def _sz(x):
    return ctypes.c_size_t(int(x))


def qsort(a: DArray, hasnan=False, descending=False):
    _check(getattr(lib(), f"xss_gpu_qsort_{a.tag}")(a.ptr, _sz(a.n), int(hasnan), int(descending), _stream()),
           f"qsort<{a.tag}>")


def qselect(a: DArray, k, hasnan=False, descending=False):
    _check(getattr(lib(), f"xss_gpu_qselect_{a.tag}")(a.ptr, _sz(k), _sz(a.n), int(hasnan), int(descending),
                                                      _stream()), f"qselect<{a.tag}>")


def partial_qsort(a: DArray, k, hasnan=False, descending=False):
    _check(getattr(lib(), f"xss_gpu_partial_qsort_{a.tag}")(a.ptr, _sz(k), _sz(a.n), int(hasnan), int(descending),
                                                            _stream()), f"partial_qsort<{a.tag}>")


def argsort(a: DArray, hasnan=False, descending=False, out: DArray = None):
    out = out or DArray.empty("u64", a.n)
    _check(getattr(lib(), f"xss_gpu_argsort_{a.tag}")(a.ptr, out.ptr, _sz(a.n), int(hasnan), int(descending),
                                                      _stream()), f"argsort<{a.tag}>")
    return out


def argselect(a: DArray, k, hasnan=False, out: DArray = None):
    out = out or DArray.empty("u64", a.n)
    _check(getattr(lib(), f"xss_gpu_argselect_{a.tag}")(a.ptr, out.ptr, _sz(k), _sz(a.n), int(hasnan), _stream()),
           f"argselect<{a.tag}>")
    return out


def argsort_arg(a: DArray, arg: DArray, hasnan=False, descending=False):
    """NumPy-style argsort (x86simdsortStatic::argsort(T *, size_t *arg, ...)):
    reorders the u64 indices in `arg` in place so that a[arg] is sorted."""
    _check(getattr(lib(), f"xss_gpu_argsort_arg_{a.tag}")(a.ptr, arg.ptr, _sz(a.n), int(hasnan), int(descending),
                                                          _stream()), f"argsort_arg<{a.tag}>")
    return arg


def argselect_arg(a: DArray, arg: DArray, k, hasnan=False):
    """NumPy-style argselect (x86simdsortStatic::argselect(T *, size_t *arg, k, ...)):
    reorders `arg` in place so that a[arg[k]] has rank k, smaller keys before it."""
    _check(getattr(lib(), f"xss_gpu_argselect_arg_{a.tag}")(a.ptr, arg.ptr, _sz(k), _sz(a.n), int(hasnan),
                                                            _stream()), f"argselect_arg<{a.tag}>")
    return arg


def keyvalue_qsort(key: DArray, val: DArray, hasnan=False, descending=False):
    _check(getattr(lib(), f"xss_gpu_keyvalue_qsort_{key.tag}_{val.tag}")(
        key.ptr, val.ptr, _sz(key.n), int(hasnan), int(descending), _stream()), f"keyvalue_qsort<{key.tag},{val.tag}>")


def keyvalue_select(key: DArray, val: DArray, k, hasnan=False, descending=False):
    _check(getattr(lib(), f"xss_gpu_keyvalue_select_{key.tag}_{val.tag}")(
        key.ptr, val.ptr, _sz(k), _sz(key.n), int(hasnan), int(descending), _stream()),
        f"keyvalue_select<{key.tag},{val.tag}>")


def keyvalue_partial_sort(key: DArray, val: DArray, k, hasnan=False, descending=False):
    _check(getattr(lib(), f"xss_gpu_keyvalue_partial_sort_{key.tag}_{val.tag}")(
        key.ptr, val.ptr, _sz(k), _sz(key.n), int(hasnan), int(descending), _stream()),
        f"keyvalue_partial_sort<{key.tag},{val.tag}>")


def object_sort_by_keys(objs, obj_bytes, keys: DArray):
    """Sort `objs` (a CUDA uint8 tensor of n*obj_bytes) by `keys` ascending."""
    _check(getattr(lib(), f"xss_gpu_object_sort_by_keys_{keys.tag}")(
        ctypes.c_void_p(objs.data_ptr()), _sz(obj_bytes), keys.ptr, _sz(keys.n), _stream()),
        f"object_sort_by_keys<{keys.tag}>")


# ---------------------------------------------------------------- serving form

# plan knobs of the segmented form (xss_gpu_set_tuning names; read when a
# plan is created) besides the two create() arguments
SEG_KNOBS = ("seg_bucket_keys", "seg_oversample", "seg_local_items", "seg_wide_tiles", "seg_sample_off")


class SegPlan:
    """A segmented plan (xss_gpu_segplan_*): op "qsort" / "argsort" /
    "qselect" over the segments offsets[i] .. offsets[i + 1] (host uint64)
    of one device buffer of key type `tag`.  `tuning` (optional dict, e.g. a
    tuning.json entry) sets small_max, range_tiles and the SEG_KNOBS while
    the plan is created."""

    def __init__(self, op, tag, offsets, tuning=None):
        L = lib()
        t = dict(tuning or {})
        self._off = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.nseg = len(self._off) - 1
        for k in SEG_KNOBS:
            if t.get(k):
                _check(L.xss_gpu_set_tuning(k.encode(), ctypes.c_uint64(int(t[k]))), "xss_gpu_set_tuning")
        self.p = ctypes.c_void_p()
        try:
            _check(L.xss_gpu_segplan_create(ctypes.byref(self.p), op.encode(), tag.encode(),
                                            self._off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                            ctypes.c_size_t(self.nseg), ctypes.c_uint32(int(t.get("small_max", 0))),
                                            ctypes.c_uint32(int(t.get("range_tiles", 0)))), "xss_gpu_segplan_create")
        finally:
            for k in SEG_KNOBS:
                if t.get(k):
                    L.xss_gpu_set_tuning(k.encode(), ctypes.c_uint64(0))

    def run(self, keys_ptr, idx_ptr=None, hasnan=False, descending=False):
        """Enqueue the plan on torch's current stream (device pointers)."""
        _check(lib().xss_gpu_segplan_run(self.p, ctypes.c_void_p(keys_ptr), ctypes.c_void_p(idx_ptr or 0),
                                         int(hasnan), int(descending), _stream()), "xss_gpu_segplan_run")

    def close(self):
        if self.p:
            lib().xss_gpu_segplan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Request(ctypes.Structure):
    """struct xss_request of include/xss_gpu.h."""
    _fields_ = [("op", ctypes.c_int), ("key_type", ctypes.c_int), ("val_type", ctypes.c_int),
                ("hasnan", ctypes.c_int), ("descending", ctypes.c_int), ("reserved", ctypes.c_int),
                ("keys", ctypes.c_void_p), ("vals", ctypes.c_void_p), ("n", ctypes.c_uint64),
                ("k", ctypes.c_uint64), ("arrival_us", ctypes.c_double), ("latency_us", ctypes.c_double),
                ("internal", ctypes.c_void_p)]


SERVE_OPS = {"qsort": 0, "keyvalue_qsort": 1, "qselect": 2, "partial_qsort": 3}
TYPE_CODE = {tag: i for i, tag in enumerate(TAGS)}  # u16 .. f64 -> 0 .. 8


def make_request(op, keys: DArray, vals: DArray = None, k=0, hasnan=False, descending=False, arrival_us=0.0):
    """A request on device arrays.  k is qselect's 0-based rank or
    partial_qsort's count, as in the synchronous calls."""
    r = Request()
    r.op = SERVE_OPS[op]
    r.key_type = TYPE_CODE[keys.tag]
    r.val_type = TYPE_CODE[vals.tag] if vals is not None else 0
    r.hasnan, r.descending = int(hasnan), int(descending)
    r.keys = keys.t.data_ptr()
    r.vals = vals.t.data_ptr() if vals is not None else None
    r.n, r.k, r.arrival_us = keys.n, int(k), float(arrival_us)
    return r


class Server:
    """Continuous-batching request loop (xss_server_* of include/xss_gpu.h).

    Requests run on the server's own streams: the caller makes their inputs
    ready (e.g. torch.cuda.synchronize()) before submitting them."""

    def __init__(self, budget_bytes=64 << 30, streams=4, coalesce_max=1 << 20, reserve_bytes=0, one_at_a_time=False):
        L = lib()
        self.h = ctypes.c_void_p()
        _check(L.xss_server_create(ctypes.byref(self.h), _sz(budget_bytes), int(streams),
                                   ctypes.c_uint32(coalesce_max), _sz(reserve_bytes), int(one_at_a_time)),
               "xss_server_create")

    def submit(self, req: Request):
        _check(lib().xss_server_submit(self.h, ctypes.byref(req)), "xss_server_submit")

    def wait(self, req: Request):
        _check(lib().xss_server_wait(self.h, ctypes.byref(req)), "xss_server_wait")

    def run_trace(self, reqs):
        """Submit every request at its arrival_us (open loop) and wait for all;
        returns the request array with latency_us filled in."""
        arr = reqs if isinstance(reqs, ctypes.Array) else (Request * len(reqs))(*reqs)
        _check(lib().xss_server_run_trace(self.h, arr, _sz(len(reqs))), "xss_server_run_trace")
        return arr

    def stats(self):
        s = (ctypes.c_uint64 * 3)()
        _check(lib().xss_server_stats(self.h, ctypes.byref(s, 0), ctypes.byref(s, 8), ctypes.byref(s, 16)),
               "xss_server_stats")
        return {"completed": s[0], "launch_groups": s[1], "coalesced": s[2]}

    def close(self):
        if self.h:
            lib().xss_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
