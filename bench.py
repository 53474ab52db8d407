#!/usr/bin/env python3
"""Sort throughput of the HIP radix sort on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--dtype f32|u32] [--n N]

A step is one x86simdsort::qsort (device-pointer form, xss_gpu_qsort_<T>) of
N = 2^28 keys already resident in HBM: uniform over the full bit range of the
type (finite for floats), no NaN, hasnan = false, ascending.  The sort is in
place, so the input is restored from a pristine copy before every step; like
the reference's benchmark (benchmarks/bench-qsort.hpp:12-16, PauseTiming)
the restore is outside the timed sort: each step is bracketed by HIP events
on the stream the library launches on, and the JSON reports their sum.

The sort of 2^28 4-byte keys runs MSD-first: an LSD pass on the top digit,
a segmented pass on the second digit inside every top bucket, then every
bucket of the top 16 bits sorted by its low 16 bits in LDS (buckets too large
for LDS take segmented passes).

The JSON line carries:
  value/ms_per_step   whole-sort throughput (Gkeys/s) and time per sort
  roofline            the dominant kernel: its algorithmic bytes per launch
                      (N * 2 * key bytes for kernels that read and write
                      every key once, N * key bytes for histogram kernels)
                      over its average launch time from the library's
                      per-kernel HIP events on the launch stream, against
                      MI355X's 8 TB/s; `traffic` is the measured HBM bytes per
                      launch from profiles/ when a rocprofv3 --pmc summary is
                      present there
  pipeline            the whole sort against the roofline at the algorithmic
                      (1 + 2P) * key bytes per key of an 8-bit LSD sort; the
                      MSD-first design moves about 7 * key bytes per key (one
                      two-digit histogram read, two passes of read + write,
                      then one read and one write of each bucket); per-kernel
                      ms per sort; and the same sort timed on the LSD path
                      (four passes) for reference
  segmented           the batch form: 8192 uneven float32 arrays (~775M keys)
                      sorted by one graph-captured plan against a loop of
                      plain per-array calls on the same GPU; int32 argsort,
                      select at len / 2, and uint64 sort and argsort of the
                      same batch shape
  serving             the served workload (serve_bench.py) at 10k and 50k
                      requests/s: p50 / p99 latency, keys/s, launch groups
                      and coalesced requests of the server and of one plain
                      call at a time
  cpu_baseline        the reference library (oracle/_ref, AVX-512 dispatch +
                      OpenMP, <= 16 threads) timed on this host: `value` on
                      the headline array itself (2^28 keys), `rows` on
                      bounded samples of the other workloads (1M-key qsort,
                      argsort, qselect, Zipf key-value sort, object_qsort,
                      the segmented arrays as a loop of calls), each beside
                      the GPU's rate on the full workload
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "x86-simd-sort_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "Gkeys/s + % HBM roofline at 256M keys; segmented-batch vs per-call speedup"


def device_input(torch, dtype, n, seed):
    """n keys uniform over the 32-bit patterns; for f32, uniform over the
    finite ones (an infinity or NaN pattern is drawn again)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randint(-2 ** 31, 2 ** 31, (n,), dtype=torch.int32, device="cuda", generator=g)
    if dtype == "f32":
        bad = ((x & 0x7F800000) == 0x7F800000).nonzero().flatten()
        while bad.numel():
            x[bad] = torch.randint(-2 ** 31, 2 ** 31, (bad.numel(),), dtype=torch.int32, device="cuda", generator=g)
            bad = bad[(x[bad] & 0x7F800000) == 0x7F800000]
    return x


def lib_timings(L):
    names = (ctypes.c_char_p * 32)()
    ms = (ctypes.c_double * 32)()
    cnt = (ctypes.c_uint64 * 32)()
    k = L.xss_gpu_timing_collect(32, names, ms, cnt)
    return {names[i].decode(): (ms[i], cnt[i]) for i in range(max(0, k))}


def _ref_or_port():
    """(library, prefix, kind, threads, isa): the reference itself when
    oracle/_ref is built, else the oracle's C restatement (one thread)."""
    import xss_testlib as T
    L = T.ref()
    if L is not None:
        threads = min(16, os.cpu_count() or 1)
        L.xssref_set_threads(threads)
        return L, "xssref_", "reference", threads, L.xssref_isa().decode()
    return T.oracle(), "xss_oracle_", "port", 1, "scalar C"


def _cpu_time(fn, make, reps, seconds):
    """Median seconds of fn(buf) over up to `reps` runs (at least 2) within
    `seconds`; make() builds each run's input outside the timed call."""
    times = []
    t_end = time.time() + seconds
    while len(times) < 2 or (len(times) < reps and time.time() < t_end):
        buf = make()
        t0 = time.perf_counter()
        fn(buf)
        times.append(time.perf_counter() - t0)
    return statistics.median(times), len(times)


def _gpu_ms(torch, restore, call, reps=5):
    """Median ms of call() between HIP events on the current stream."""
    ts = []
    for _ in range(reps + 1):
        restore()
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        rc = call()
        e1.record(s)
        torch.cuda.synchronize()
        assert rc in (0, None), rc
        ts.append(e0.elapsed_time(e1))
    return statistics.median(ts[1:])


def _zipf_u64(n, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    z = rng.zipf(1.1, n).astype(np.uint64)
    bad = z >= (1 << 32)
    while bad.any():
        z[bad] = rng.zipf(1.1, int(bad.sum())).astype(np.uint64)
        bad = z >= (1 << 32)
    return z


def cpu_baseline(torch, L, dtype, gpu_ms_headline, seg=None, quick=False):
    """The reference (kind "reference": oracle/_ref, its AVX-512 dispatch and
    OpenMP team, <= 16 threads) timed on this host, the input of every run
    built outside the timed call as the reference's benchmarks do
    (benchmarks/bench-qsort.hpp:12-16).  `value` is the headline workload,
    qsort of 2^28 uniform keys, on the full array; `rows` time the other
    workloads on bounded samples and give the GPU's rate on the full
    workload beside them."""
    import numpy as np
    import xss_testlib as T
    R, pre, kind, threads, isa = _ref_or_port()
    p = ctypes.c_void_p
    z = ctypes.c_size_t
    n = 1 << 24 if quick else 1 << 28
    a = T.gen(dtype, "random", n, 99)
    f = getattr(R, f"{pre}qsort_{dtype}")
    sec, runs = _cpu_time(lambda b: f(p(b.ctypes.data), z(n), 0, 0), a.copy, 3, 20.0)
    out = {"value": round(n / sec / 1e9, 4), "unit": "Gkeys/s", "cores": threads, "kind": kind,
           "sample": f"qsort<{dtype}> of 2^{n.bit_length() - 1} uniform full-range keys (the headline array), "
                     f"median of {runs} sorts; dispatch {isa}",
           "gpu_over_cpu": round((float(1 << 28) / (gpu_ms_headline * 1e-3)) / (n / sec), 1)}
    del a
    if kind != "reference":
        out["rows"] = "the oracle port times the headline only"
        return out
    rows = []
    cur = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)  # noqa: E731

    def row(workload, cpu_n, cpu_s, cpu_runs, gpu_n, gpu_ms, unit_note=""):
        c = cpu_n / cpu_s / 1e9
        g = gpu_n / (gpu_ms * 1e-3) / 1e9
        rows.append({"workload": workload, "cpu_sample_n": cpu_n, "cpu_runs": cpu_runs, "cpu_Gkeys_s": round(c, 4),
                     "gpu_n": gpu_n, "gpu_ms": round(gpu_ms, 4), "gpu_Gkeys_s": round(g, 3),
                     "gpu_over_cpu": round(g / c, 1), **({"note": unit_note} if unit_note else {})})

    # C1: qsort<float> of 1,000,000 keys U(-1e6, 1e6)
    m = 1_000_000
    c1 = (np.random.default_rng(1).uniform(-1e6, 1e6, m)).astype(np.float32)
    sec, runs = _cpu_time(lambda b: R.xssref_qsort_f32(p(b.ctypes.data), z(m), 0, 0), c1.copy, 30, 3.0)
    src = torch.from_numpy(c1).cuda()
    w = torch.empty_like(src)
    gms = _gpu_ms(torch, lambda: w.copy_(src),
                  lambda: L.xss_gpu_qsort_f32(p(w.data_ptr()), z(m), 0, 0, cur()), 20)
    row("qsort<float>, 1,000,000 keys U(-1e6, 1e6)", m, sec, runs, m, gms)
    # argsort<float>, uniform bits
    cn, gn = (1 << 22, 1 << 24) if quick else (1 << 25, 1 << 28)
    ca = T.gen("f32", "random", cn, 5)
    idx = np.zeros(cn, np.uint64)
    sec, runs = _cpu_time(lambda b: R.xssref_argsort_f32(p(b.ctypes.data), z(cn), 0, 0, p(idx.ctypes.data)),
                          lambda: ca, 3, 10.0)
    gsrc = device_input(torch, "f32", gn, 6)
    gidx = torch.empty(gn, dtype=torch.int64, device="cuda")
    gms = _gpu_ms(torch, lambda: None, lambda: L.xss_gpu_argsort_f32(p(gsrc.data_ptr()), p(gidx.data_ptr()), z(gn),
                                                                     0, 0, cur()))
    row(f"argsort<float>, uniform full-range keys (GPU 2^{gn.bit_length() - 1})", cn, sec, runs, gn, gms)
    # qselect<float> at k = n/2
    sec, runs = _cpu_time(lambda b: R.xssref_qselect_f32(p(b.ctypes.data), z(cn // 2), z(cn), 0, 0), ca.copy, 5, 5.0)
    w = torch.empty_like(gsrc)
    gms = _gpu_ms(torch, lambda: w.copy_(gsrc),
                  lambda: L.xss_gpu_qselect_f32(p(w.data_ptr()), z(gn // 2), z(gn), 0, 0, cur()))
    row(f"qselect<float> at k = n/2, uniform full-range keys (GPU 2^{gn.bit_length() - 1})", cn, sec, runs, gn, gms)
    del gsrc, gidx, w, ca, idx
    torch.cuda.empty_cache()
    # keyvalue_qsort<uint64_t, uint64_t>, Zipf(1.1) keys below 2^32, values = index
    cn, gn = (1 << 22, 1 << 23) if quick else (1 << 25, 1 << 27)
    zk = _zipf_u64(cn, 3)
    zv = np.arange(cn, dtype=np.uint64)
    vb = zv.copy()

    def kv_make():
        vb[:] = zv
        return zk.copy()
    sec, runs = _cpu_time(lambda b: R.xssref_keyvalue_qsort_u64_u64(p(b.ctypes.data), p(vb.ctypes.data), z(cn), 0, 0),
                          kv_make, 3, 10.0)
    gk0 = torch.from_numpy(_zipf_u64(gn, 4).view(np.int64)).cuda()
    gv0 = torch.arange(gn, dtype=torch.int64, device="cuda")
    gk, gv = torch.empty_like(gk0), torch.empty_like(gv0)
    gms = _gpu_ms(torch, lambda: (gk.copy_(gk0), gv.copy_(gv0)),
                  lambda: L.xss_gpu_keyvalue_qsort_u64_u64(p(gk.data_ptr()), p(gv.data_ptr()), z(gn), 0, 0, cur()))
    row(f"keyvalue_qsort<uint64_t, uint64_t>, Zipf(1.1) keys < 2^32 (GPU 2^{gn.bit_length() - 1})",
        cn, sec, runs, gn, gms)
    del gk0, gv0, gk, gv, zk, zv, vb
    torch.cuda.empty_cache()
    # object_qsort of {double x, y, z} by the Euclidean norm
    on = 1 << 20 if quick else 1 << 24
    xyz = np.random.default_rng(7).uniform(-1, 1, (on, 3))
    sec, runs = _cpu_time(lambda b: R.xssref_object_qsort_xyz(p(b.ctypes.data), z(on)), xyz.copy, 2, 8.0)
    gx0 = torch.from_numpy(xyz.reshape(-1)).cuda()
    gx = torch.empty_like(gx0)
    keys = torch.empty(on, dtype=torch.float64, device="cuda")
    gms = _gpu_ms(torch, lambda: gx.copy_(gx0),
                  lambda: L.xss_gpu_object_qsort_xyz_norm(p(gx.data_ptr()), z(on), p(keys.data_ptr()), cur()))
    row(f"object_qsort of 2^{on.bit_length() - 1} {{x, y, z}} doubles by the Euclidean norm", on, sec, runs, on, gms,
        "the reference publishes 2.208 s at 16M objects (misc/bench-simdobjsort.txt)")
    del gx0, gx, keys, xyz
    torch.cuda.empty_cache()
    # the segmented workload as a loop of x86simdsort::qsort calls over its first 1024 arrays
    if seg is not None:
        sizes, vals_host = seg["sizes"][:1024], seg["host_sample"]
        off = np.concatenate([[0], np.cumsum(sizes)])
        cn = int(off[-1])

        def seg_loop(b):
            base = b.ctypes.data
            for i in range(len(sizes)):
                R.xssref_qsort_f32(p(base + 4 * int(off[i])), z(int(sizes[i])), 1, 0)
        sec, runs = _cpu_time(seg_loop, vals_host.copy, 3, 8.0)
        row("8192 float arrays (len log-uniform [16, 2^20], 0.1% NaN): x86simdsort::qsort per array "
            "(CPU: the first 1024 arrays); GPU: one segmented plan, graph replay",
            cn, sec, runs, seg["n_keys"], seg["segmented_ms"])
    out["rows"] = rows
    return out


def segmented_batch(torch, xss_amd, L, tuning, reps=3, loops=True):
    """8192 float32 arrays, lengths log-uniform in [16, 2^20] (seeded), values
    U(-1, 1) with 0.1% NaN, hasnan = true: one plan captured in a graph and
    replayed, against a C++ loop of plain xss_gpu_qsort_f32 calls on the same
    GPU (inputs restored outside the timed region).  Also per-segment int32
    argsort (uint32 indices) and select at k = len/2."""
    import numpy as np
    rng = np.random.default_rng(2024)
    nseg = 8192
    sizes = np.exp(rng.uniform(np.log(16), np.log(1 << 20), nseg)).astype(np.int64)
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
    stream = torch.cuda.current_stream()
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))

    def cur():  # the stream torch launches on now (a capture stream inside torch.cuda.graph)
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def plan(op, tag):
        """the plan with the autotuner's setting for (op, key type) (tuning.json)"""
        t = {k: v for k, v in tuning.get(f"{op}_{tag}", {}).items() if k not in ("ms", "sweep")}
        return xss_amd.SegPlan(op, tag, off, t)

    def timed(fn, src=None):
        ts = []
        for _ in range(reps):
            work.copy_(pristine if src is None else src)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return min(ts)

    def check_sorted():
        v = work.view(torch.float32)
        bad = 0
        for i in range(0, nseg, 512):  # spot-check every 512th segment
            a, b = int(off[i]), int(off[i + 1])
            seg = v[a:b]
            fin = seg[~torch.isnan(seg)]
            bad += int((fin[1:] < fin[:-1]).sum().item()) + int(torch.isnan(seg[: fin.numel()]).sum().item())
        return bad == 0

    out = {"workload": "8192 f32 arrays, len log-uniform [16, 2^20], U(-1,1) with 0.1% NaN, hasnan=true",
           "n_keys": n, "n_segments": nseg,
           "tuning": {k: {x: y for x, y in v.items() if x != "sweep"} if isinstance(v, dict) else v
                      for k, v in tuning.items()}}
    p = plan("qsort", "f32")
    ptr = ctypes.c_void_p(work.data_ptr())
    run = lambda: p.run(work.data_ptr(), None, True)  # noqa: E731
    work.copy_(pristine)
    run()  # warm-up on the real input (outside capture)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        run()
    t_graph = timed(graph.replay)
    ok = check_sorted()
    t_loop = timed(lambda: L.xss_gpu_loop_qsort_f32(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_float)), offp,
                                                    ctypes.c_size_t(nseg), 1, 0, cur())) if loops else float("nan")
    out.update({"segmented_ms": round(t_graph, 3), "loop_ms": round(t_loop, 3),
                "speedup_vs_loop": round(t_loop / t_graph, 2), "segmented_Gkeys_s": round(n / t_graph / 1e6, 2),
                "sorted_ok": ok})
    p.close()
    # argsort of int32 keys with uint32 indices
    idx = torch.empty(n, dtype=torch.int32, device="cuda")
    p = plan("argsort", "i32")
    ia = lambda: p.run(work.data_ptr(), idx.data_ptr(), False)  # noqa: E731
    work.copy_(ints)
    ia()
    t_arg = timed(ia, ints)
    idx64 = torch.empty(n, dtype=torch.int64, device="cuda")
    t_arg_loop = timed(lambda: L.xss_gpu_loop_argsort_i32(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int32)),
                                                          ctypes.cast(ctypes.c_void_p(idx64.data_ptr()),
                                                                      ctypes.POINTER(ctypes.c_uint64)),
                                                          offp, ctypes.c_size_t(nseg), cur()), ints) \
        if loops else float("nan")
    p.close()
    p = plan("qselect", "f32")
    sel = lambda: p.run(work.data_ptr(), None, True)  # noqa: E731
    work.copy_(pristine)
    sel()
    t_sel = timed(sel)
    t_sel_loop = timed(lambda: L.xss_gpu_loop_qselect_f32(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_float)), offp,
                                                          ctypes.c_size_t(nseg), 1, cur())) if loops else float("nan")
    p.close()
    out.update({"argsort_i32_ms": round(t_arg, 3), "argsort_i32_loop_ms": round(t_arg_loop, 3),
                "qselect_ms": round(t_sel, 3), "qselect_loop_ms": round(t_sel_loop, 3)})
    # the same batch shape with uint64 keys over the full range: sort and
    # argsort (uint32 indices)
    del work
    torch.cuda.empty_cache()
    keys64 = torch.randint(-2 ** 63, 2 ** 63 - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    work = torch.empty_like(keys64)
    p = plan("qsort", "u64")
    q64 = lambda: p.run(work.data_ptr(), None, False)  # noqa: E731
    work.copy_(keys64)
    q64()
    t_q64 = timed(q64, keys64)
    ok64 = True
    for i in range(0, nseg, 512):
        a, b = int(off[i]), int(off[i + 1])
        seg = work[a:b] ^ (-2 ** 63)  # unsigned order as signed
        ok64 &= bool((seg[1:] >= seg[:-1]).all().item()) if b - a > 1 else True
    p.close()
    p = plan("argsort", "u64")
    a64 = lambda: p.run(work.data_ptr(), idx.data_ptr(), False)  # noqa: E731
    work.copy_(keys64)
    a64()
    t_a64 = timed(a64, keys64)
    p.close()
    out.update({"qsort_u64_ms": round(t_q64, 3), "qsort_u64_Gkeys_s": round(n / t_q64 / 1e6, 2),
                "qsort_u64_sorted_ok": ok64, "argsort_u64_ms": round(t_a64, 3)})
    del keys64
    # the CPU loop's sample: the first 1024 arrays as the plan sorted them
    m = int(off[1024])
    extra = {"sizes": sizes, "n_keys": n, "segmented_ms": t_graph,
             "host_sample": pristine[:m].cpu().numpy().view(np.float32)}
    del work, pristine, ints, idx, idx64
    torch.cuda.empty_cache()
    return out, extra


def traffic_from_profiles(kernel, dtype, n):
    """Measured HBM bytes per launch of `kernel` from a rocprofv3 --pmc
    summary in profiles/ (None when there is none for this size)."""
    p = os.path.join(REPO, "profiles", f"{kernel}_{dtype}_256M_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        if int(d.get("n", 0)) != n:
            return None
        return d.get("bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dtype", default="f32", choices=["f32", "u32"])
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-seg", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--quick-cpu", action="store_true", help="smaller CPU samples (a quick check of the legs)")
    ap.add_argument("--seg-only", action="store_true",
                    help="only the segmented plans (no loop baselines; for profiling their kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X, and the sort of one array "
              f"does not span GPUs (--gpus {args.gpus} not supported)", file=sys.stderr)
        sys.exit(2)

    import torch
    import xss_amd
    L = xss_amd.lib()
    torch.cuda.set_device(0)
    if args.seg_only:
        tp = os.path.join(REPO, "x86-simd-sort_amd", "tuning.json")
        tuning = json.load(open(tp)) if os.path.exists(tp) else {}
        seg, _ = segmented_batch(torch, xss_amd, L, tuning, loops=False)
        print(json.dumps({k: v for k, v in seg.items() if k != "tuning"}))
        return
    n, dt = args.n, args.dtype
    key_bytes = 4
    pristine = device_input(torch, dt, n, 1234)
    work = torch.empty_like(pristine)
    arr = xss_amd.DArray(dt, n, work.view(torch.uint8))

    def restore():
        work.copy_(pristine)

    for _ in range(args.warmup):
        restore()
        xss_amd.qsort(arr)
    torch.cuda.synchronize()
    lib_timings(L)

    stream = torch.cuda.current_stream()
    evs = []
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        restore()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        xss_amd.qsort(arr)
        e1.record(stream)
        evs.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    # per-kernel breakdown from the library's HIP events, on separate steps
    # (the events add a little time, so they stay out of the timed loop)
    L.xss_gpu_timing_enable(1)
    kt_steps = 3
    for _ in range(kt_steps):
        restore()
        xss_amd.qsort(arr)
    torch.cuda.synchronize()
    L.xss_gpu_timing_enable(0)
    kt = lib_timings(L)
    sort_ms = [a.elapsed_time(b) for a, b in evs]
    ms_per_step = sum(sort_ms) / len(sort_ms)

    # correctness of the last step (ascending order of the float/uint values)
    out = work if dt == "f32" else work
    if dt == "f32":
        v = out.view(torch.float32)
        ok = bool((v[1:] >= v[:-1]).all().item())
    else:
        v = out.to(torch.int64) & 0xFFFFFFFF
        ok = bool((v[1:] >= v[:-1]).all().item())

    # The dominant kernel of the sort, priced at its algorithmic bytes per
    # launch: scatter-type kernels read and write every key once, histogram
    # kernels read every key once (the hybrid's per-bucket LDS sort also
    # reads and writes each key once).
    # (hyb_msd_upsweep, pass B's counts, reads only the ~5% of keys in runs
    # that its ranges cut, so it is no candidate)
    alg_per_launch = {"downsweep": 2 * n * key_bytes, "hyb_msd_downsweep": 2 * n * key_bytes,
                      "hyb_local": 2 * n * key_bytes, "upsweep": n * key_bytes}
    dom = max((k for k in alg_per_launch if k in kt), key=lambda k: kt[k][0])
    dom_ms, dom_cnt = kt[dom]
    dom_avg_ms = dom_ms / max(1, dom_cnt)
    dom_bytes = alg_per_launch[dom]
    achieved = dom_bytes / (dom_avg_ms / 1e3) / 1e9 if dom_avg_ms > 0 else 0.0
    P = 4
    hybrid = "hyb_local" in kt
    # algorithmic minimum of an 8-bit LSD sort: one histogram read + P x (read + write)
    pipe_bytes = n * (1 + 2 * P) * key_bytes
    # what this design moves: MSD-first = one histogram read, two passes of
    # read + write and the LDS bucket sort's read + write (7 x key bytes;
    # pass B's counts read ~5% more, buckets too large for LDS add segmented
    # passes); LSD = P x (histogram read + read + write)
    moved_bytes = n * (7 if hybrid else 3 * P) * key_bytes
    pipe_gbs = pipe_bytes / (ms_per_step / 1e3) / 1e9
    gkeys = n / (ms_per_step / 1e3) / 1e9
    kernel_desc = {"downsweep": ("downsweep_kernel (pass A: scatter on the top varying digit)" if hybrid
                                 else "downsweep_kernel (LSD scatter pass)"),
                   "hyb_msd_downsweep": "seg_downsweep_kernel (pass B: second digit inside each top bucket)",
                   "hyb_local": "hyb_local_kernel (LDS sort of each 16-bit bucket by its low 16 bits)",
                   "upsweep": ("hyb_h2_kernel (two-digit histogram read)" if hybrid
                               else "upsweep_kernel (digit histograms)")}[dom]
    # the same sort on the LSD path (four 8-bit passes), for reference
    lsd_ms = None
    if hybrid:
        L.xss_gpu_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
        L.xss_gpu_set_tuning(b"hybrid_min_n", 1 << 62)
        ts = []
        for _ in range(4):
            restore()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            xss_amd.qsort(arr)
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        L.xss_gpu_set_tuning(b"hybrid_min_n", 1 << 25)
        lsd_ms = round(statistics.median(ts[1:]), 4)
    line = {
        "metric": METRIC,
        "value": round(gkeys, 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if dt == "f32" else "u32",
        "data": "synthetic (uniform full-range bits, finite, seeded on device)",
        "config": {"workload": f"qsort<{'float' if dt == 'f32' else 'uint32_t'}> of 2^28 keys in HBM, "
                               f"ascending, hasnan=false", "n": n, "key_bytes": key_bytes,
                   "digit_bits": 8, "design": "msd_hybrid" if hybrid else "lsd", "restore_excluded": True},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(dom, dt, n),
                     "kernel": kernel_desc, "avg_launch_ms": round(dom_avg_ms, 4),
                     "bytes_per_launch_alg": dom_bytes},
        "pipeline": {"bytes_per_sort_alg_lsd8": pipe_bytes, "bytes_per_sort_moved": moved_bytes,
                     "achieved_GBs": round(pipe_gbs, 1),
                     "frac": round(pipe_gbs / HBM_PEAK_GBS, 4),
                     "kernels_ms_per_sort": {k: round(v[0] / kt_steps, 4) for k, v in kt.items()},
                     "lsd_path_ms_per_sort": lsd_ms},
        "wall_s_timed_loop_incl_restore": round(wall, 4),
        "sorted_ok": ok,
    }
    seg_extra = None
    if not args.no_seg:
        tuning = {}
        tp = os.path.join(REPO, "x86-simd-sort_amd", "tuning.json")
        if os.path.exists(tp):
            tuning = json.load(open(tp))
        line["segmented"], seg_extra = segmented_batch(torch, xss_amd, L, tuning)
        line["segmented_speedup_vs_loop"] = line["segmented"]["speedup_vs_loop"]
    if not args.no_serve:
        from serve_bench import serving_leg
        legs = serving_leg(torch)
        line["serving"] = {
            "workload": "open-loop Poisson traces of 1000 requests: 70% qsort<float> 1K-16M keys, 20% "
                        "keyvalue_qsort<u64,u64> Zipf keys 1K-4M, 10% qselect/partial_qsort<u32> 1K-16M; "
                        "server: 4 streams, coalescing up to 2^20 keys; baseline: one plain call at a time",
            "rates": [{"rate_req_s": r["rate_req_s"], "total_keys": r["total_keys"],
                       **{f"{m}_{k}": r[m][k] for m in ("server", "plain_call_baseline")
                          for k in ("p50_ms", "p99_ms", "keys_per_s", "launch_groups", "coalesced_requests")}}
                      for r in legs]}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(torch, L, dt, ms_per_step, seg_extra, args.quick_cpu)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
