"""Autotuner for the segmented form.

    python x86-simd-sort_amd/autotune.py [--segments 8192] [--out x86-simd-sort_amd/tuning.json]

Tunes the plan knobs of xss_gpu_segplan_create on the batch workload of the
benchmark (8192 arrays, lengths log-uniform in [16, 2^20]) for every
(op, key type) the benchmark and the batch tests run:

  qsort_f32     float32 U(-1, 1) with 0.1% NaN, hasnan (the sample sort)
  argsort_i32   int32 over the full range, uint32 indices (sample sort)
  qselect_f32   the float32 arrays, k = len / 2 (sample sort)
  qsort_u64     uint64 over the full range (sample sort)
  argsort_u64   the uint64 keys, uint32 indices (sample sort)
  qsort_u16     uint16 over the full range (the LSD plan)

Knobs (include/xss_gpu.h): small_max (segments up to it: one LDS bitonic
workgroup each), range_tiles (scatter tiles per workgroup range) and, for
4- and 8-byte keys, seg_bucket_keys (the average bucket of the sample sort:
its bucket threshold), seg_oversample (samples per bucket), seg_local_items
(keys per thread of the LDS sort: 12 or 15; 8-byte keys 8 or 12, argsort 8
or 11) and seg_wide_tiles (scatter tiles of 1024 instead of 512 threads).
The digit of every scatter is one byte (256 buckets or radix digits per
pass).

Coordinate descent from the defaults: each knob in turn is swept with the
others at the best values so far, for two rounds; every setting is timed as
the best of three replays of the plan captured in a hipGraph.  The result
(the best setting per (op, key type), with the sweep) is written to
tuning.json, which bench.py reads.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

SAMPLE_KNOBS = {
    "small_max": [1024, 2048, 4096],
    "range_tiles": [2, 4, 8],
    "seg_bucket_keys": [2048, 3072, 4096, 5120],
    "seg_oversample": [8, 16, 32],
    "seg_local_items": [12, 15],
    "seg_wide_tiles": [0, 1],
}
SAMPLE_KNOBS_W8 = {
    "small_max": [512, 1024, 2048],
    "range_tiles": [2, 4, 8],
    "seg_bucket_keys": [3072, 4096, 5120],
    "seg_oversample": [16, 32, 64],
    "seg_local_items": [8, 12],
    "seg_wide_tiles": [0, 1],
}
LSD_KNOBS = {
    "small_max": [512, 1024, 2048],
    "range_tiles": [2, 4, 8, 16],
}
DEFAULTS = {"small_max": 0, "range_tiles": 0, "seg_bucket_keys": 0, "seg_oversample": 0, "seg_local_items": 0,
            "seg_wide_tiles": 0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--segments", type=int, default=8192)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--only", default=None, help="comma-separated case names")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    args = ap.parse_args()
    import numpy as np
    import torch
    import xss_amd
    rng = np.random.default_rng(2024)
    nseg = args.segments
    sizes = np.exp(rng.uniform(np.log(16), np.log(1 << 20), nseg)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    stream = torch.cuda.current_stream()

    def data(case):
        if case in ("qsort_f32", "qselect_f32"):
            v = torch.rand(n, device="cuda", generator=g) * 2 - 1
            v[torch.rand(n, device="cuda", generator=g) < 0.001] = float("nan")
            return v.view(torch.uint8)
        if case == "argsort_i32":
            return torch.randint(-2 ** 31, 2 ** 31, (n,), dtype=torch.int32, device="cuda", generator=g).view(torch.uint8)
        if case in ("qsort_u64", "argsort_u64"):
            return torch.randint(-2 ** 63, 2 ** 63 - 1, (n,), dtype=torch.int64, device="cuda",
                                 generator=g).view(torch.uint8)
        if case == "qsort_u16":
            return torch.randint(-2 ** 15, 2 ** 15, (n,), dtype=torch.int16, device="cuda",
                                 generator=g).view(torch.uint8)
        raise ValueError(case)

    cases = [("qsort_f32", "qsort", "f32", True, SAMPLE_KNOBS), ("argsort_i32", "argsort", "i32", False, SAMPLE_KNOBS),
             ("qselect_f32", "qselect", "f32", True, SAMPLE_KNOBS),
             ("qsort_u64", "qsort", "u64", False, SAMPLE_KNOBS_W8),
             ("argsort_u64", "argsort", "u64", False, dict(SAMPLE_KNOBS_W8, seg_local_items=[8, 11])),
             ("qsort_u16", "qsort", "u16", False, LSD_KNOBS)]
    if args.only:
        keep = set(args.only.split(","))
        cases = [c for c in cases if c[0] in keep]
    results = {}
    if os.path.exists(args.out):
        try:
            results = json.load(open(args.out))
        except ValueError:
            results = {}
    for name, op, tag, hasnan, knobs in cases:
        src = data(name)
        work = torch.empty_like(src)
        idx = torch.empty(n, dtype=torch.int32, device="cuda") if op == "argsort" else None
        cache = {}

        def timed(setting):
            key = json.dumps(setting, sort_keys=True)
            if key in cache:
                return cache[key]
            plan = xss_amd.SegPlan(op, tag, off, setting)
            run = lambda: plan.run(work.data_ptr(), idx.data_ptr() if idx is not None else None, hasnan)  # noqa: E731
            work.copy_(src)
            run()
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                run()
            ts = []
            for _ in range(3):
                work.copy_(src)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                graph.replay()
                e1.record(stream)
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
            del graph
            plan.close()
            cache[key] = min(ts)
            print(name, setting, round(cache[key], 3), flush=True)
            return cache[key]

        best = {k: DEFAULTS[k] for k in knobs}
        best_ms = timed(best)
        sweep = [{"setting": dict(best), "ms": round(best_ms, 3)}]
        for _ in range(args.rounds):
            for knob, values in knobs.items():
                for v in values:
                    trial = dict(best)
                    trial[knob] = v
                    t = timed(trial)
                    sweep.append({"setting": trial, "ms": round(t, 3)})
                    if t < best_ms:
                        best, best_ms = trial, t
        results[name] = dict(best, ms=round(best_ms, 3), sweep=sweep)
        del src, work, idx
        torch.cuda.empty_cache()
    results["workload"] = (f"{nseg} arrays, len log-uniform [16, 2^20], {n} keys; coordinate descent, "
                           f"{args.rounds} rounds, best of 3 graph replays; tuned {time.strftime('%Y-%m-%d')} on "
                           f"{torch.cuda.get_device_name(0)}")
    with open(args.out, "w") as f:
        json.dump(results, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
