"""Throughput of every single-array workload on one GPU, device-resident.

    python x86-simd-sort_amd/bench_configs.py [--reps 3] [--only c1,c2,c3]
        [--types u16,f32,...] [--recipes random,sorted,...]

c1  qsort float32, N = 1,000,000, uniform in [-1e6, 1e6], ascending.
c2  qsort and argsort, each type at 1 GiB of keys (2^28 32-bit, 2^27 64-bit,
    2^29 16-bit), five recipes -- uniform over the full bit range (finite
    floats), 1024 distinct values, sorted, reverse sorted, 1% NaN with hasnan
    (floats) -- ascending and descending; and for the float types one more
    recipe, "unit": uniform in (-1, 1), whose exponents are skewed (a few
    top digits hold most keys).
c3  keyvalue_qsort u64/u64 (2^27, Zipf s=1.1 keys below 2^32, values = index)
    and f32/u32 (2^28, normal keys); qselect / keyvalue_select at k = N/2 and
    N-1 and partial_qsort / keyvalue_partial_sort at k = 1000 and N/100 on
    the same arrays; object_qsort of 2^24 {double x, y, z} by the Euclidean
    norm (device key evaluation).
c6  (with --only c6) the PyTorch shim torch.ops.xss.sort beside torch.sort:
    2^28 float32 in one row and 4096 rows of 65536, values + int64 indices
    (bytes per key counted as one read and write of keys and indices).
Each measurement restores the input from a pristine device copy outside the
timed region and times the call alone with HIP events on the stream it runs
on; the median of --reps is reported.  One JSON line per measurement with
Gkeys/s, the path the library takes (`path`), the bytes per key that path
moves (`moved_bytes_per_key`) and the fraction of the 8 TB/s HBM peak they
reach (`achieved_frac`), beside the fraction at the algorithmic bytes per
key of the 8-bit LSD plan (`lsd_equiv_frac`, an effective rate: above
achieved_frac where the path moves fewer bytes):
  sort      s_k (1 + 2P)
  sort      2 s_k for 16-bit keys (the counting sort: one read, one write)
  argsort   s_k + (2 s_k + s_i) + 2 (P - 2)(s_k + s_i) + (s_k + s_i + s_o)
  kv sort   s_k + 2P (s_k + s_v)
  select    4 s_k (+ 2 s_v for key-value)
  object    24 (read xyz) + 8 (write key) + argsort(f64) + 48 (gather + copy back)
(P passes, s_i = 4-byte internal index, s_o = 8-byte output index).
Paths and their moved bytes per key:
  lsd       the plan above
  count     qsort of 2^23+ 16-bit keys: one read, one write (2 s_k)
  msd       qsort / argsort / kv sort of 2^25+ 4- or 8-byte keys, MSD-first:
            a two-digit histogram read, two scatter passes and the LDS bucket
            sort -- sort 7 s_k, argsort 6 s_k + 24, kv s_k + 6 (s_k + s_v)
            (passes over constant digits and large-bucket passes not counted)
  select    radix select: as above
"""
import argparse
import ctypes
import json
import os
import statistics
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import xss_amd  # noqa: E402

PEAK = 8000.0
BYTES = {"u16": 2, "i16": 2, "f16": 2, "u32": 4, "i32": 4, "f32": 4, "u64": 8, "i64": 8, "f64": 8}
N_FULL = {2: 1 << 29, 4: 1 << 28, 8: 1 << 27}
RECIPES = ["random", "dups", "sorted", "reverse", "nan", "unit"]


def passes(kb):
    return kb  # 8-bit digits: one pass per key byte


def alg_bytes(op, kb, vb=0):
    P = passes(kb)
    if op == "qsort":
        return 2 * kb if kb == 2 else kb * (1 + 2 * P)
    if op == "argsort":
        si, so = 4, 8
        return kb + (2 * kb + si) + 2 * (P - 2) * (kb + si) + (kb + si + so)
    if op == "kv":
        return kb + 2 * P * (kb + vb)
    if op in ("select", "partial"):
        return 4 * kb + 2 * vb
    raise ValueError(op)


def torch_dtype(torch, tag):
    return {"u16": torch.int16, "i16": torch.int16, "f16": torch.int16, "u32": torch.int32, "i32": torch.int32,
            "f32": torch.int32, "u64": torch.int64, "i64": torch.int64, "f64": torch.int64}[tag]


def rand_bits(torch, tag, n, g):
    """Uniform over the type's bit patterns, as raw integers; for floats,
    uniform over the finite patterns (an infinity or NaN pattern is drawn
    again until it is finite)."""
    kb = BYTES[tag]

    def draw(m):
        if kb == 2:
            return torch.randint(-2 ** 15, 2 ** 15, (m,), dtype=torch.int16, device="cuda", generator=g)
        if kb == 4:
            return torch.randint(-2 ** 31, 2 ** 31, (m,), dtype=torch.int32, device="cuda", generator=g)
        hi = torch.randint(-2 ** 31, 2 ** 31, (m,), dtype=torch.int64, device="cuda", generator=g)
        lo = torch.randint(0, 2 ** 32, (m,), dtype=torch.int64, device="cuda", generator=g)
        return (hi << 32) | lo

    x = draw(n)
    if tag in ("f16", "f32", "f64"):
        exp = {2: 0x7C00, 4: 0x7F800000, 8: 0x7FF0000000000000}[kb]
        e = torch.tensor(exp, dtype=x.dtype, device="cuda")
        bad = ((x & e) == e).nonzero().flatten()
        while bad.numel():
            x[bad] = draw(bad.numel())
            bad = bad[(x[bad] & e) == e]
    return x


QNAN_INT = {"f16": 0x7C01, "f32": 0x7FC00000, "f64": 0x7FF8000000000000}


def make_input(torch, L, tag, recipe, n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = rand_bits(torch, tag, n, g)
    if recipe == "dups":
        pool = rand_bits(torch, tag, 1024, g)
        x = pool[torch.randint(0, 1024, (n,), device="cuda", generator=g)]
    elif recipe in ("sorted", "reverse"):
        rc = getattr(L, f"xss_gpu_qsort_{tag}")(ctypes.c_void_p(x.data_ptr()), ctypes.c_size_t(n), 0,
                                                 int(recipe == "reverse"), stream(torch))
        assert rc == 0, L.xss_gpu_last_error()
    elif recipe == "nan":
        m = torch.rand(n, device="cuda", generator=g) < 0.01
        x[m] = QNAN_INT[tag]
    elif recipe == "unit":
        dt = {"f16": torch.float16, "f32": torch.float32, "f64": torch.float64}[tag]
        u = torch.rand(n, device="cuda", generator=g, dtype=torch.float64 if tag == "f64" else torch.float32)
        x = (u * 2 - 1).to(dt).view(x.dtype)
        del u
    torch.cuda.synchronize()
    return x


def stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def timed(torch, restore, call, reps):
    """Median ms of `call` over reps; restore() runs outside the events."""
    ts = []
    for _ in range(reps + 1):
        restore()
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        rc = call()
        e1.record(s)
        torch.cuda.synchronize()
        assert rc == 0, xss_amd.lib().xss_gpu_last_error()
        ts.append(e0.elapsed_time(e1))
    return statistics.median(ts[1:])  # the first run warms the workspace


HYB_MIN_N, COUNT_MIN_N = 1 << 25, 1 << 23  # the library's default routing thresholds


def route(op, kb, n, vb=0):
    """(path, moved bytes per key) of the call as the library routes it"""
    if op == "qsort" and kb == 2 and n >= COUNT_MIN_N:
        return "count", 2 * kb
    if op in ("qsort", "argsort", "kv") and kb >= 4 and n >= HYB_MIN_N:
        if op == "qsort":
            return "msd", 7 * kb
        if op == "argsort":
            return "msd", 6 * kb + 24
        return "msd", kb + 6 * (kb + vb)
    if op in ("select", "partial"):
        return "select", alg_bytes(op, kb, vb)
    return "lsd", alg_bytes(op, kb, vb)


def emit(cfg, op, tag, n, ms, bpk, moved=None, path=None, **kw):
    gk = n / (ms * 1e-3) / 1e9
    eff = n * bpk / (ms * 1e-3) / 1e9
    moved = bpk if moved is None else moved
    ach = n * moved / (ms * 1e-3) / 1e9
    rec = {"config": cfg, "op": op, "type": tag, "n": n, "ms": round(ms, 4), "Gkeys_s": round(gk, 2),
           "path": path or "lsd", "moved_bytes_per_key": moved, "achieved_GBs": round(ach, 1),
           "achieved_frac": round(ach / PEAK, 4), "alg_bytes_per_key": bpk, "lsd_equiv_frac": round(eff / PEAK, 4)}
    rec.update(kw)
    print(json.dumps(rec), flush=True)


def emit_op(cfg, op, tag, n, ms, kb, vb=0, route_op=None, **kw):
    """emit with the LSD-plan bytes and the routed path's bytes of op"""
    path, moved = route(route_op or op, kb, n, vb)
    emit(cfg, op, tag, n, ms, alg_bytes(route_op or op, kb, vb), moved=moved, path=path, **kw)


def c1(torch, L, reps):
    n = 1_000_000
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    src = ((torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1) * 1e6).float()
    w = torch.empty_like(src)
    ms = timed(torch, lambda: w.copy_(src),
               lambda: L.xss_gpu_qsort_f32(ctypes.c_void_p(w.data_ptr()), ctypes.c_size_t(n), 0, 0, stream(torch)),
               max(reps, 20))
    emit_op("c1", "qsort", "f32", n, ms, 4)


def c2(torch, L, reps, types=None, recipes=None):
    for tag in xss_amd.TAGS:
        if types and tag not in types:
            continue
        kb = BYTES[tag]
        n = N_FULL[kb]
        idx = torch.empty(n, dtype=torch.int64, device="cuda")
        for recipe in RECIPES:
            if recipes and recipe not in recipes:
                continue
            if recipe in ("nan", "unit") and tag not in ("f16", "f32", "f64"):
                continue
            hasnan = int(recipe == "nan")
            src = make_input(torch, L, tag, recipe, n, 100 + RECIPES.index(recipe))
            w = torch.empty_like(src)
            for desc in (0, 1):
                f = getattr(L, f"xss_gpu_qsort_{tag}")
                ms = timed(torch, lambda: w.copy_(src),
                           lambda: f(ctypes.c_void_p(w.data_ptr()), ctypes.c_size_t(n), hasnan, desc, stream(torch)),
                           reps)
                emit_op("c2", "qsort", tag, n, ms, kb, recipe=recipe, descending=bool(desc))
                fa = getattr(L, f"xss_gpu_argsort_{tag}")
                ms = timed(torch, lambda: None,
                           lambda: fa(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(idx.data_ptr()),
                                      ctypes.c_size_t(n), hasnan, desc, stream(torch)), reps)
                emit_op("c2", "argsort", tag, n, ms, kb, recipe=recipe, descending=bool(desc))
            del src, w
        del idx
        torch.cuda.empty_cache()


def c3(torch, L, reps):
    import numpy as np
    # key-value u64/u64, Zipf keys, values = original index
    n = 1 << 27
    rng = np.random.default_rng(3)
    z = rng.zipf(1.1, n).astype(np.uint64)
    bad = z >= (1 << 32)
    while bad.any():
        z[bad] = rng.zipf(1.1, int(bad.sum())).astype(np.uint64)
        bad = z >= (1 << 32)
    arrays = [("u64", "u64", torch.from_numpy(z.view(np.int64)).cuda(), torch.arange(n, device="cuda"))]
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    m = 1 << 28
    arrays.append(("f32", "u32", torch.randn(m, device="cuda", generator=g).view(torch.int32),
                   torch.randint(-2 ** 31, 2 ** 31, (m,), dtype=torch.int32, device="cuda", generator=g)))
    for kt, vt, ks, vs in arrays:
        n = ks.numel()
        kb, vb = BYTES[kt], BYTES[vt]
        wk, wv = torch.empty_like(ks), torch.empty_like(vs)

        def restore():
            wk.copy_(ks)
            wv.copy_(vs)

        pk, pv = ctypes.c_void_p(wk.data_ptr()), ctypes.c_void_p(wv.data_ptr())
        f = getattr(L, f"xss_gpu_keyvalue_qsort_{kt}_{vt}")
        ms = timed(torch, restore, lambda: f(pk, pv, ctypes.c_size_t(n), 0, 0, stream(torch)), reps)
        emit_op("c3", "keyvalue_qsort", f"{kt}/{vt}", n, ms, kb, vb, route_op="kv")
        for k in (n // 2, n - 1):
            f = getattr(L, f"xss_gpu_keyvalue_select_{kt}_{vt}")
            ms = timed(torch, restore, lambda: f(pk, pv, ctypes.c_size_t(k), ctypes.c_size_t(n), 0, 0, stream(torch)),
                       reps)
            emit("c3", "keyvalue_select", f"{kt}/{vt}", n, ms, alg_bytes("select", kb, vb), path="select", k=k)
            f = getattr(L, f"xss_gpu_qselect_{kt}")
            ms = timed(torch, lambda: wk.copy_(ks),
                       lambda: f(pk, ctypes.c_size_t(k), ctypes.c_size_t(n), 0, 0, stream(torch)), reps)
            emit("c3", "qselect", kt, n, ms, alg_bytes("select", kb), path="select", k=k)
        for k in (1000, n // 100):
            f = getattr(L, f"xss_gpu_keyvalue_partial_sort_{kt}_{vt}")
            ms = timed(torch, restore, lambda: f(pk, pv, ctypes.c_size_t(k), ctypes.c_size_t(n), 0, 0, stream(torch)),
                       reps)
            emit("c3", "keyvalue_partial_sort", f"{kt}/{vt}", n, ms, alg_bytes("select", kb, vb), path="select", k=k,
                 note="bytes exclude the sort of the k prefix")
            f = getattr(L, f"xss_gpu_partial_qsort_{kt}")
            ms = timed(torch, lambda: wk.copy_(ks),
                       lambda: f(pk, ctypes.c_size_t(k), ctypes.c_size_t(n), 0, 0, stream(torch)), reps)
            emit("c3", "partial_qsort", kt, n, ms, alg_bytes("select", kb), path="select", k=k,
                 note="bytes exclude the sort of the k prefix")
        del wk, wv
    del arrays
    torch.cuda.empty_cache()
    # object_qsort: 2^24 {double x, y, z} in [-1, 1], keyed by the Euclidean norm
    n = 1 << 24
    xyz = torch.rand(n * 3, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
    w = torch.empty_like(xyz)
    keys = torch.empty(n, dtype=torch.float64, device="cuda")
    ms = timed(torch, lambda: w.copy_(xyz),
               lambda: L.xss_gpu_object_qsort_xyz_norm(ctypes.c_void_p(w.data_ptr()), ctypes.c_size_t(n),
                                                       ctypes.c_void_p(keys.data_ptr()), stream(torch)), reps)
    obj = 24 + 8 + alg_bytes("argsort", 8) + 48
    emit("c3", "object_qsort", "xyz_f64_norm", n, ms, obj, path="object")


def c6(torch, L, reps):
    """The PyTorch shim beside torch.sort on the same GPU: one row of 2^28
    float32 (random bits, finite) and 4096 rows of 65536 float32."""
    import xss_torch  # noqa: F401  (registers torch.ops.xss)
    g = torch.Generator(device="cuda")
    g.manual_seed(600)
    for shape in ((1 << 28,), (4096, 1 << 16)):
        n = 1
        for d in shape:
            n *= d
        x = rand_bits(torch, "f32", n, g).view(torch.float32).view(shape)
        # each call returns 0 for timed(); its outputs are dropped after the events
        for name, fn in (("xss::sort", lambda: 0 if torch.ops.xss.sort(x, False) else 1),
                         ("torch.sort", lambda: 0 if torch.sort(x, dim=-1) else 1)):
            ms = timed(torch, lambda: None, fn, reps)
            emit("c6", name, "f32", n, ms, 2 * 4 + 2 * 8, path="torch", shape="x".join(map(str, shape)))
        del x
        torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--only", default="c1,c2,c3")
    ap.add_argument("--types", default="", help="c2: comma-separated key types (default all)")
    ap.add_argument("--recipes", default="", help="c2: comma-separated recipes (default all)")
    args = ap.parse_args()
    import torch
    L = xss_amd.lib()
    types = [t for t in args.types.split(",") if t]
    recipes = [r for r in args.recipes.split(",") if r]
    for c in args.only.split(","):
        if c == "c2":
            c2(torch, L, args.reps, types, recipes)
        else:
            {"c1": c1, "c3": c3, "c6": c6}[c](torch, L, args.reps)


if __name__ == "__main__":
    main()
