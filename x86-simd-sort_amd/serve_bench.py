"""Open-loop serving benchmark of the continuous-batching request loop.

    python x86-simd-sort_amd/serve_bench.py [--rates 1000,10000,50000] [--requests 2000]

Request mix (per request, seeded):
  70% qsort of float32, length log-uniform in [1K, 16M], values U(-1, 1);
  20% keyvalue_qsort uint64 keys / uint64 values, length log-uniform in
      [1K, 4M], keys Zipf(s = 1.1) ranks over [1, 2^32), values = index;
  10% qselect (half) / partial_qsort (half) of uint32, length log-uniform in
      [1K, 16M], k uniform in [1, len].
Arrivals are Poisson at each rate (open loop: a C++ submitter issues every
request at its arrival time whatever the backlog).  A 200-request trace
through both modes warms the kernels up first.  Inputs are device
resident before the trace starts.  The server runs with 4 streams, a 64 GB
admission budget and coalescing of qsort / keyvalue_qsort requests of at
most --coalesce keys (default 2^20: LDS-sized ones by the pointer-list sort,
longer ones by shared segmented passes); the baseline serves the same trace
one request at a time with the plain call.  Prints one JSON line per rate:
p50/p99 latency and keys/s for both.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from xss_amd import SERVE_OPS, TYPE_CODE, Request, Server  # noqa: E402

OP_QSORT, OP_KV, OP_SELECT, OP_PARTIAL = (SERVE_OPS[o] for o in ("qsort", "keyvalue_qsort", "qselect", "partial_qsort"))


def log_uniform(rng, lo, hi, size):
    import numpy as np
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size)).astype(np.int64)


class Pools:
    """Device-resident source data; every request copies a slice of it."""

    def __init__(self, torch, seed=5, size=1 << 24):
        import numpy as np
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        self.f32 = (torch.rand(size, device="cuda", generator=g) * 2 - 1).contiguous()
        self.u32 = torch.randint(-2 ** 31, 2 ** 31, (size,), dtype=torch.int32, device="cuda", generator=g)
        rng = np.random.default_rng(seed)
        z = rng.zipf(1.1, 4 << 20).astype(np.uint64)
        bad = z >= (1 << 32)
        while bad.any():
            z[bad] = rng.zipf(1.1, int(bad.sum())).astype(np.uint64)
            bad = z >= (1 << 32)
        self.zipf = torch.from_numpy(z.view(np.int64)).cuda()
        self.size = size


def make_trace(torch, pools, n_req, rate, seed):
    """Requests with their own device buffers (filled from the pools)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    kinds = rng.choice(3, size=n_req, p=[0.7, 0.2, 0.1])
    arrivals = np.cumsum(rng.exponential(1e6 / rate, n_req))
    reqs = (Request * n_req)()
    bufs = []
    meta = []
    for i in range(n_req):
        r = reqs[i]
        r.arrival_us = float(arrivals[i])
        if kinds[i] == 0:
            n = int(log_uniform(rng, 1 << 10, 1 << 24, 1)[0])
            off = int(rng.integers(0, pools.size - n + 1))
            k = pools.f32[off:off + n].clone()
            r.op, r.key_type, r.keys, r.n = OP_QSORT, TYPE_CODE["f32"], k.data_ptr(), n
            bufs.append((k,))
            meta.append(("qsort", "f32", n, 0))
        elif kinds[i] == 1:
            n = int(log_uniform(rng, 1 << 10, 1 << 22, 1)[0])
            off = int(rng.integers(0, pools.zipf.numel() - n + 1))
            k = pools.zipf[off:off + n].clone()
            v = torch.arange(n, dtype=torch.int64, device="cuda")
            r.op, r.key_type, r.val_type = OP_KV, TYPE_CODE["u64"], TYPE_CODE["u64"]
            r.keys, r.vals, r.n = k.data_ptr(), v.data_ptr(), n
            bufs.append((k, v))
            meta.append(("kv", "u64", n, 0))
        else:
            n = int(log_uniform(rng, 1 << 10, 1 << 24, 1)[0])
            off = int(rng.integers(0, pools.size - n + 1))
            k = pools.u32[off:off + n].clone()
            kk = int(rng.integers(1, n + 1))
            sel = rng.random() < 0.5
            r.op = OP_SELECT if sel else OP_PARTIAL
            r.key_type, r.keys, r.n = TYPE_CODE["u32"], k.data_ptr(), n
            r.k = kk - 1 if sel else kk  # qselect takes a 0-based rank, partial_qsort a count
            bufs.append((k,))
            meta.append(("qselect" if sel else "partial", "u32", n, int(r.k)))
    torch.cuda.synchronize()
    return reqs, bufs, meta


def run_trace(reqs, streams, coalesce_max, one_at_a_time, budget=64 << 30, reserve=600 << 20):
    with Server(budget, streams, coalesce_max, reserve, one_at_a_time) as sv:
        t0 = time.perf_counter()
        sv.run_trace(reqs)
        wall = time.perf_counter() - t0
        return wall, sv.stats()


def summarize(reqs, wall):
    import numpy as np
    lat = np.array([r.latency_us for r in reqs]) / 1e3
    keys = sum(int(r.n) for r in reqs)
    return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "keys_per_s": round(keys / wall, 1), "wall_s": round(wall, 4)}


MODES = lambda coalesce: [("server", 4, coalesce, 0), ("plain_call_baseline", 1, 0, 1)]  # noqa: E731


def warm_up(torch, pools, coalesce):
    """One short trace through both modes first: HIP loads a kernel's code
    on its first launch, which would otherwise land in the first requests'
    latencies."""
    for name, streams, coal, one in MODES(coalesce):
        reqs, bufs, _ = make_trace(torch, pools, 200, 50000, seed=1)
        run_trace(reqs, streams, coal, one)
        del bufs
    torch.cuda.empty_cache()


def measure(torch, pools, rate, requests, coalesce):
    out = {"rate_req_s": rate, "requests": requests}
    for name, streams, coal, one in MODES(coalesce):
        reqs, bufs, meta = make_trace(torch, pools, requests, rate, seed=rate)
        wall, stats = run_trace(reqs, streams, coal, one)
        s = summarize(reqs, wall)
        s.update({"launch_groups": stats["launch_groups"], "coalesced_requests": stats["coalesced"]})
        out[name] = s
        del bufs
        torch.cuda.empty_cache()
    out["total_keys"] = sum(int(m[2]) for m in meta)
    return out


def serving_leg(torch, rates=(10000, 50000), requests=1000, coalesce=1 << 20):
    """The served workload at each rate, server vs one plain call at a time."""
    pools = Pools(torch)
    warm_up(torch, pools, coalesce)
    res = [measure(torch, pools, r, requests, coalesce) for r in rates]
    del pools
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="1000,10000,50000")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--coalesce", type=int, default=1 << 20)
    args = ap.parse_args()
    import torch
    pools = Pools(torch)
    warm_up(torch, pools, args.coalesce)
    for rate in [int(x) for x in args.rates.split(",")]:
        print(json.dumps(measure(torch, pools, rate, args.requests, args.coalesce)), flush=True)


if __name__ == "__main__":
    main()
