"""PyTorch custom ops over libxss_amd for device tensors: the shim a
framework uses in place of the reference's CPU sort (PyTorch calls
x86-simd-sort for torch.sort on the CPU, README.md:201-205).

    import xss_torch
    values, indices = torch.ops.xss.sort(t, False)       # like torch.sort(t, dim=-1)
    indices = torch.ops.xss.argsort(t, True)             # like torch.argsort(t, dim=-1, descending=True)

Both sort along the last dimension of a CUDA tensor of uint16, int16,
float16, uint32, int32, float32, uint64, int64 or float64.  Semantics are
torch.sort's with stable=False: NaN sorts above +inf (last ascending, first
descending) and keeps its bits; the order inside a run of equal keys is
unspecified; -0.0 sorts before +0.0 (they compare equal, so the values equal
torch's).  One row is one library call (keyvalue_qsort with 32-bit index
payloads, or argsort and a gather for 16-bit keys); several rows are one
segmented argsort (xss_gpu_segplan_*, plan cached per shape) and a gather.
Everything runs on the tensor's current stream; nothing falls back to the
CPU.
"""
import ctypes
import os
import sys
from collections import OrderedDict

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import xss_amd  # noqa: E402

_TAG = {torch.uint16: "u16", torch.int16: "i16", torch.float16: "f16", torch.uint32: "u32", torch.int32: "i32",
        torch.float32: "f32", torch.uint64: "u64", torch.int64: "i64", torch.float64: "f64"}
# same-width signed views for gathers (torch gathers every signed width on the GPU)
_SIGNED = {1: None, 2: torch.int16, 4: torch.int32, 8: torch.int64}
_PLANS = OrderedDict()  # (tag, rows, cols) -> xss_segplan_t, least recently used first
_MAX_PLANS = 16


def _stream(x):
    return ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)


MAX_ROW = 0xFF000000  # the library's per-array limit (2^32 - 2^24 keys)


def _check_input(x):
    if not x.is_cuda:
        raise ValueError("xss ops take CUDA tensors")
    if x.dtype not in _TAG:
        raise TypeError(f"xss ops do not sort {x.dtype}")
    if x.dim() == 0:
        raise ValueError("xss ops sort along the last dimension of a tensor with at least one")
    if x.shape[-1] > MAX_ROW:
        raise ValueError(f"rows of more than {MAX_ROW} elements are beyond the library's per-array limit")


def _gather(x, idx):
    """x gathered along the last dimension by int64 idx (through a signed view of the same width)."""
    v = _SIGNED[x.element_size()]
    src = x.view(v) if x.dtype != v else x
    out = torch.gather(src, -1, idx)
    return out.view(x.dtype) if x.dtype != v else out


def _plan(tag, rows, cols):
    key = (tag, rows, cols)
    p = _PLANS.get(key)
    if p is not None:
        _PLANS.move_to_end(key)
        return p
    L = xss_amd.lib()
    offs = np.arange(rows + 1, dtype=np.uint64) * np.uint64(cols)
    p = ctypes.c_void_p()
    xss_amd._check(L.xss_gpu_segplan_create(ctypes.byref(p), b"argsort", tag.encode(),
                                            offs.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(rows), 0, 0),
                   f"segplan_create<{tag}>")
    _PLANS[key] = p
    if len(_PLANS) > _MAX_PLANS:
        _, old = _PLANS.popitem(last=False)
        torch.cuda.synchronize()  # a run of the evicted plan may still be in flight
        L.xss_gpu_segplan_destroy(old)
    return p


def _argsort(x, descending):
    """int64 indices that sort x along its last dimension."""
    x = x.contiguous()
    tag, cols = _TAG[x.dtype], x.shape[-1]
    rows = x.numel() // cols if cols else 0
    if x.numel() == 0:
        return torch.empty(x.shape, dtype=torch.int64, device=x.device)
    L = xss_amd.lib()
    if rows == 1:
        idx = torch.empty(x.shape, dtype=torch.int64, device=x.device)
        xss_amd._check(getattr(L, f"xss_gpu_argsort_{tag}")(ctypes.c_void_p(x.data_ptr()),
                                                             ctypes.c_void_p(idx.data_ptr()), ctypes.c_size_t(cols),
                                                             1, int(descending), _stream(x)), f"argsort<{tag}>")
        return idx
    idx32 = torch.empty(x.shape, dtype=torch.int32, device=x.device)
    xss_amd._check(L.xss_gpu_segplan_run(_plan(tag, rows, cols), ctypes.c_void_p(x.data_ptr()),
                                         ctypes.c_void_p(idx32.data_ptr()), 1, int(descending), _stream(x)),
                   f"segplan_run<{tag}>")
    return idx32.to(torch.int64)


@torch.library.custom_op("xss::argsort", mutates_args=(), device_types="cuda")
def argsort(x: torch.Tensor, descending: bool = False) -> torch.Tensor:
    """torch.argsort(x, dim=-1, descending=descending) on the GPU library."""
    _check_input(x)
    return _argsort(x, descending)


@argsort.register_fake
def _argsort_fake(x, descending=False):
    return torch.empty(x.shape, dtype=torch.int64, device=x.device)


@torch.library.custom_op("xss::sort", mutates_args=(), device_types="cuda")
def sort(x: torch.Tensor, descending: bool = False) -> tuple[torch.Tensor, torch.Tensor]:
    """torch.sort(x, dim=-1, descending=descending) -> (values, int64 indices)."""
    _check_input(x)
    x = x.contiguous()
    tag, cols = _TAG[x.dtype], x.shape[-1]
    if x.numel() == 0:
        return x.clone(), torch.empty(x.shape, dtype=torch.int64, device=x.device)
    if x.numel() == cols and x.element_size() >= 4 and cols < (1 << 31):
        # one row of 32/64-bit keys: one key-value sort carries 32-bit
        # indices (int32 holds them), widened afterwards -- measured 1.45x faster
        # than carrying int64 indices at 2^28 float32 keys
        vals = x.clone()
        idx32 = torch.arange(cols, dtype=torch.int32, device=x.device).view(x.shape)
        L = xss_amd.lib()
        xss_amd._check(getattr(L, f"xss_gpu_keyvalue_qsort_{tag}_u32")(
            ctypes.c_void_p(vals.data_ptr()), ctypes.c_void_p(idx32.data_ptr()), ctypes.c_size_t(cols), 1,
            int(descending), _stream(x)), f"keyvalue_qsort<{tag},u32>")
        return vals, idx32.to(torch.int64)
    idx = _argsort(x, descending)
    return _gather(x, idx), idx


@sort.register_fake
def _sort_fake(x, descending=False):
    return torch.empty_like(x), torch.empty(x.shape, dtype=torch.int64, device=x.device)
