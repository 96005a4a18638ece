"""Serving benchmark: the serving configuration on one MI355X.

Index: 10,000,000 x 96 float32 L2 vectors from a 4,096-centre Gaussian mixture
(centres U[0,1)^96, sigma 0.05, seed 4); IVF nlist 16384 with 4-bit PQ M=48
fast-scan; a RefineIndex (k_factor 10) whose raw vectors stay resident in HBM.
Load: a Poisson stream of requests, each a batch of 1-256 queries (log-uniform),
k from {10, 100}, nprobe from {16, 64}; 30% of requests carry an allowed-id
set of 10% of the ids (one of eight seeded sets). The plain path is measured
first on the same request mix: each request one batch_query of its own, one
at a time, FP32 coarse GEMM, no graph capture. Then the serving loop's
saturated throughput (requests offered far faster than served), then p50 /
p99 latency at 25%, 50% and 90% of it. recall@10 is measured on 1,000 queries
against exact search, and a sample of served results is checked against the
same queries run alone.

Run from the repository root:  python3 -m qann_amd.serving_bench [--scale 0.1]
Prints one JSON object (last line of stdout); progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(n, nq, dim, centres, sigma, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((centres, dim), dtype=np.float32)
    x = np.empty((n, dim), np.float32)
    for s in range(0, n, 1_000_000):
        e = min(n, s + 1_000_000)
        x[s:e] = c[rng.integers(0, centres, e - s)] + sigma * rng.standard_normal((e - s, dim), dtype=np.float32)
    q = c[rng.integers(0, centres, nq)] + sigma * rng.standard_normal((nq, dim), dtype=np.float32)
    return x, q.astype(np.float32)


def log_uniform_mean(max_nq, seed=0):
    """The mean request size of the log-uniform [1, max_nq] mix the load draws."""
    u = np.random.default_rng(seed).random(200_000)
    return float(np.mean(np.clip(np.rint(np.exp(u * np.log(max_nq))), 1, max_nq)))


def plain_path(index, ivf, pool, ks, nprobes, seconds, filters=(), filtered_fraction=0.0, max_nq=256, seed=12):
    """The plain FP32 one-launch-per-request path on a request mix: each request
    (1..max_nq queries, log-uniform; k and nprobe drawn from ks / nprobes; a
    filter with probability filtered_fraction) one synchronous batch_query of
    its own, one at a time, with the FP32 coarse GEMM and no graph capture,
    back to back for `seconds` (its saturated rate)."""
    saved = {k: os.environ.get(k) for k in ("QANN_AMD_COARSE", "QANN_AMD_GRAPHS")}
    os.environ["QANN_AMD_COARSE"], os.environ["QANN_AMD_GRAPHS"] = "fp32", "0"
    prng = np.random.default_rng(seed)
    lat, queries = [], 0
    try:
        start = time.perf_counter()
        while time.perf_counter() - start < seconds:
            nq = int(np.clip(np.rint(np.exp(prng.random() * np.log(max_nq))), 1, max_nq))
            lo = int(prng.integers(0, pool.shape[0] - nq))
            k, npb = int(prng.choice(ks)), int(prng.choice(nprobes))
            allowed = filters[int(prng.integers(0, len(filters)))] if filters and prng.random() < filtered_fraction else None
            if ivf is not None:
                ivf.nprobe = npb
            a = time.perf_counter()
            index.batch_query(pool[lo:lo + nq], k, allowed=allowed)
            lat.append(time.perf_counter() - a)
            queries += nq
        elapsed = time.perf_counter() - start
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return {"qps": round(queries / elapsed, 1), "requests": len(lat),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3)}


def served_levels(qann_amd, server, pool, spec, seconds, fractions=(0.25, 0.5, 0.9), offered=20_000.0):
    """The serving loop's saturated rate (`offered` requests per second, far
    more than it serves, for seconds / 2), then open-loop Poisson loads at the
    given fractions of it: latency percentiles per level."""
    mean_nq = log_uniform_mean(spec.get("max_nq", 256))
    sat = qann_amd.serve_load(server, pool, rate=offered, seconds=seconds / 2, seed=11, **spec)
    levels = {}
    for frac in fractions:
        rate = frac * sat["qps"] / mean_nq
        st = qann_amd.serve_load(server, pool, rate=rate, seconds=seconds, seed=int(frac * 100), **spec)
        levels[f"{int(frac * 100)}%"] = {k: round(v, 3) if isinstance(v, float) else v for k, v in st.items()}
        log(f"{int(frac * 100)}% load ({rate:,.0f} requests/s): {st['qps']:,.0f} queries/s, p50 {st['p50_ms']:.2f} ms, "
            f"p99 {st['p99_ms']:.2f} ms, {st['launches']} launches for {st['requests']} requests")
    return sat, levels, mean_nq


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the 10M vectors (and of nlist)")
    ap.add_argument("--seconds", type=float, default=6.0, help="length of each load level")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=4096, help="queries one launch may take")
    ap.add_argument("--stage-seconds", type=float, default=2.0, help="length of the stage-profiled load (0: none)")
    ap.add_argument("--offered", type=float, default=60_000.0,
                    help="requests per second offered to find the saturated rate (far more than are served)")
    args = ap.parse_args()

    import torch  # noqa: F401  (first: torch and the library share the HIP runtime)
    import qann_amd

    n = int(10_000_000 * args.scale)
    nlist = max(256, int(16384 * args.scale))
    dim, m, kf = 96, 48, 10
    t0 = time.perf_counter()
    x, pool = make_data(n, 20_000, dim, 4096, 0.05, 4)
    log(f"data {x.shape} in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    ivf = qann_amd.IVFIndex(dim, nlist, 16)
    ivf.enable_pq(m, fast_scan=True)
    rng = np.random.default_rng(4)
    ivf.train(x[rng.choice(n, min(n, 64 * nlist), replace=False)], max_iters=10, seed=4)
    index = qann_amd.RefineIndex(ivf, kf)
    index.add(x)
    index.batch_query(pool[:8], 10)  # uploads the index
    log(f"index built in {time.perf_counter() - t0:.1f}s")

    # recall@10 of 1,000 queries at each nprobe against exact search (torch, FP32)
    import torch as th

    xs = th.from_numpy(x).cuda()
    xsq = (xs * xs).sum(1)[None, :]
    truth = np.empty((1000, 10), dtype=np.int64)
    for s in range(0, 1000, 100):
        qq = th.from_numpy(pool[s:s + 100]).cuda()
        d2 = (qq * qq).sum(1, keepdim=True) - 2 * qq @ xs.T + xsq
        truth[s:s + 100] = d2.topk(10, dim=1, largest=False).indices.cpu().numpy()
    del xs, xsq, qq, d2
    th.cuda.empty_cache()
    recall = {}
    for npb in (16, 64):
        ivf.nprobe = npb
        ids, _ = index.batch_query(pool[:1000], 10)
        recall[npb] = float(np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, truth)]))
    log(f"recall@10: {recall}")

    ids_all = np.arange(n, dtype=np.int64)
    frng = np.random.default_rng(5)
    filters = [np.sort(frng.choice(ids_all, n // 10, replace=False)) for _ in range(8)]
    spec = dict(ks=[10, 100], nprobes=[16, 64], filtered_fraction=0.3, filters=filters, max_nq=256)

    plain = plain_path(index, ivf, pool, [10, 100], [16, 64], args.seconds / 2, filters=filters,
                       filtered_fraction=0.3)
    log(f"plain FP32, one launch per request: {plain['qps']:,.0f} queries/s over {plain['requests']} requests "
        f"(p50 {plain['p50_ms']:.2f} ms, p99 {plain['p99_ms']:.2f} ms per request)")

    server = qann_amd.Server(index, workers=args.workers, max_batch=args.max_batch)
    # served results are each query's own: a sample checked against single calls
    checks = []
    for r in range(40):
        lo = int(frng.integers(0, 19_000))
        nq = int(frng.integers(1, 257))
        k = int(frng.choice([10, 100]))
        npb = int(frng.choice([16, 64]))
        allowed = filters[r % 8] if r % 3 == 0 else None
        checks.append((lo, nq, k, npb, allowed, server.submit(pool[lo:lo + nq], k, nprobe=npb, allowed=allowed)))
    same = 0
    for lo, nq, k, npb, allowed, f in checks:
        got_ids, got_d, _ = f.result()
        ivf.nprobe = npb
        want_ids, want_d = index.batch_query(pool[lo:lo + nq], k, allowed=allowed)
        same += int(np.array_equal(got_ids, want_ids) and np.array_equal(got_d.view(np.uint32), want_d.view(np.uint32)))
    log(f"served results identical to single calls: {same}/{len(checks)}")

    sat, levels, mean_nq = served_levels(qann_amd, server, pool, spec, args.seconds, offered=args.offered)
    sat_qps = sat["qps"]
    # saturated only if the loop served clearly fewer requests than were offered
    served_rate = sat_qps / mean_nq
    saturated = served_rate < 0.9 * args.offered
    log(f"saturated: {sat_qps:,.0f} queries/s over {sat['requests']} requests in {sat['launches']} launches "
        f"({sat_qps / plain['qps']:.2f}x the plain path; {served_rate:,.0f} of {args.offered:,.0f} requests/s offered"
        f"{'' if saturated else ' -- NOT saturated: raise --offered'})")
    # per-stage kernel time of the searches (the library's stage profiler: each
    # stage is synchronized and timed with HIP events, so this pass is slower)
    stages = {}
    if args.stage_seconds > 0:
        import ctypes
        import pathlib

        lib = ctypes.CDLL(str(pathlib.Path(qann_amd.__file__).parent / "libqann_hip.so"))
        lib.qann_hip_profile_enable.argtypes = [ctypes.c_int]
        lib.qann_hip_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.c_int]
        lib.qann_hip_stage_name.restype = ctypes.c_char_p
        ms = (ctypes.c_double * 32)()
        cnt = (ctypes.c_uint64 * 32)()
        lib.qann_hip_profile_read(ms, cnt, 1)
        lib.qann_hip_profile_enable(1)
        qann_amd.serve_load(server, pool, rate=0.25 * sat_qps / mean_nq, seconds=args.stage_seconds, seed=7, **spec)
        lib.qann_hip_profile_enable(0)
        lib.qann_hip_profile_read(ms, cnt, 1)
        total = sum(ms[i] for i in range(32))
        for i in range(32):
            name = lib.qann_hip_stage_name(i).decode()
            if name and cnt[i]:
                stages[name] = {"ms_per_launch": round(ms[i] / cnt[i], 4), "share": round(ms[i] / total, 4)}
        log("stages (ms per launch, share of stage time): " +
            ", ".join(f"{k} {v['ms_per_launch']:.3f} ({100 * v['share']:.0f}%)" for k, v in stages.items()))
    server.close()
    print(json.dumps({
        "workload": f"serving: {n:,} x {dim} L2, IVF{nlist} + PQ{m}x4 fast-scan + RefineIndex(k_factor={kf}) in HBM; "
                    "Poisson requests of 1-256 queries, k in {10, 100}, nprobe in {16, 64}, 30% filtered at 10%",
        "saturated_qps": round(sat_qps, 1),
        "offered_requests_per_s": args.offered,
        "saturated": saturated,
        "plain_fp32_one_launch_per_request": plain,
        "saturated_over_plain": round(sat_qps / plain["qps"], 2),
        "levels": levels,
        "recall@10": recall,
        "served_equal_to_single_calls": f"{same}/{len(checks)}",
        "workers": args.workers,
        "max_batch": args.max_batch,
        "stage_ms_per_launch_at_25pct": stages,
    }), flush=True)


if __name__ == "__main__":
    main()
