"""Autotuner for the segmented form.

    python x86-simd-sort_amd/autotune.py [--segments 8192] [--out x86-simd-sort_amd/tuning.json]

Sweeps the plan parameters of xss_gpu_segplan_create -- the small-segment
threshold (segments up to it are sorted by one LDS bitonic workgroup each)
and the range size (downsweep tiles per workgroup in larger segments) -- on
the batch workload of the benchmark (float32 arrays, lengths log-uniform in
[16, 2^20], 0.1% NaN, hasnan), timing each setting as a graph replay, and
writes the fastest setting per (op, key type) to tuning.json.  The benchmark
and callers that want the tuned values read that file.
"""
import argparse
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--segments", type=int, default=8192)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    args = ap.parse_args()
    import numpy as np
    import torch
    import xss_amd
    L = xss_amd.lib()
    rng = np.random.default_rng(2024)
    sizes = np.exp(rng.uniform(np.log(16), np.log(1 << 20), args.segments)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    vals = torch.rand(n, device="cuda", generator=g) * 2 - 1
    vals[torch.rand(n, device="cuda", generator=g) < 0.001] = float("nan")
    pristine = vals.view(torch.int32).clone()
    del vals
    ints = torch.randint(-2 ** 31, 2 ** 31, (n,), dtype=torch.int32, device="cuda", generator=g)
    work = torch.empty_like(pristine)
    idx = torch.empty_like(pristine)
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    results = {}
    for key, op, tag, src, hasnan in [("segmented_f32", "qsort", "f32", pristine, 1),
                                      ("segmented_argsort_i32", "argsort", "i32", ints, 0)]:
        best = None
        rows = []
        for small_max in [512, 1024, 2048, 4096]:
            for range_tiles in [1, 2, 4, 8]:
                p = ctypes.c_void_p()
                rc = L.xss_gpu_segplan_create(ctypes.byref(p), op.encode(), tag.encode(), offp,
                                              ctypes.c_size_t(args.segments), ctypes.c_uint32(small_max),
                                              ctypes.c_uint32(range_tiles))
                assert rc == 0, L.xss_gpu_last_error()
                ts = []
                for _ in range(3):
                    work.copy_(src)
                    s = torch.cuda.current_stream()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(s)
                    rc = L.xss_gpu_segplan_run(p, ctypes.c_void_p(work.data_ptr()),
                                               ctypes.c_void_p(idx.data_ptr()), hasnan, 0,
                                               ctypes.c_void_p(s.cuda_stream))
                    e1.record(s)
                    torch.cuda.synchronize()
                    assert rc == 0, L.xss_gpu_last_error()
                    ts.append(e0.elapsed_time(e1))
                L.xss_gpu_segplan_destroy(p)
                t = min(ts)
                rows.append({"small_max": small_max, "range_tiles": range_tiles, "ms": round(t, 3)})
                print(key, rows[-1], flush=True)
                if best is None or t < best["ms"]:
                    best = rows[-1]
        results[key] = {"small_max": best["small_max"], "range_tiles": best["range_tiles"], "ms": best["ms"],
                        "sweep": rows}
    results["workload"] = (f"{args.segments} arrays, len log-uniform [16, 2^20], {n} keys; "
                           f"tuned {time.strftime('%Y-%m-%d')} on {torch.cuda.get_device_name(0)}")
    with open(args.out, "w") as f:
        json.dump(results, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
