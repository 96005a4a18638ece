"""Scale benchmark: the 100M-row configuration on one MI355X.

Rows: 100,000,000 x 128 float32, L2, from a 16,384-centre Gaussian mixture
whose cluster sizes follow Zipf(1.1) (centres U[0,1)^128, sigma 0.05), drawn on
the GPU with torch's seeded generator (seed 5) in chunks of 10M rows, rows then
queries; 100,000 queries the same way. IVF nlist 65535 (the reference's
largest), 4-bit PQ M=64 fast-scan (32 bytes per vector; --storage sq8: 8-bit
SQ, 128 bytes per vector) trained on 4.19M of the rows, a RefineIndex (k_factor 10) whose 51.2 GB of rows stay in HBM, k 10,
nprobe 32 and 128. The 100,000 queries run as one batch with queries and
results in HBM; recall@10 is measured on the first 1,000 against exact FP32
search over all rows.

Run from the repository root:  python3 -m qann_amd.scale_bench [--n 100000000]
Prints one JSON object (last line of stdout); progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=100_000_000)
    ap.add_argument("--nq", type=int, default=100_000)
    ap.add_argument("--nlist", type=int, default=65535)
    ap.add_argument("--centres", type=int, default=16384)
    ap.add_argument("--m", type=int, default=64)
    ap.add_argument("--storage", choices=("pq4fs", "sq8"), default="pq4fs",
                    help="list storage: 4-bit PQ fast-scan blocks, or 8-bit SQ rows (dim bytes per vector)")
    ap.add_argument("--nprobes", default="32,128")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--chunk", type=int, default=10_000_000)
    args = ap.parse_args()

    import torch
    import qann_amd

    qann_amd.set_num_threads(16)
    dim, k, kf = 128, 10, 10
    t0 = time.perf_counter()
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    centres = torch.rand((args.centres, dim), generator=g, device="cuda")
    w = torch.arange(1, args.centres + 1, device="cuda", dtype=torch.float64) ** -1.1
    w = (w / w.sum()).float()

    def draw(n):
        out = np.empty((n, dim), np.float32)
        for s in range(0, n, args.chunk):
            e = min(n, s + args.chunk)
            lab = torch.multinomial(w, e - s, replacement=True, generator=g)
            rows = centres[lab] + 0.05 * torch.randn((e - s, dim), generator=g, device="cuda")
            out[s:e] = rows.cpu().numpy()
        return out

    x = draw(args.n)
    q = draw(args.nq)
    log(f"data {x.shape} + {q.shape} in {time.perf_counter() - t0:.1f}s")

    t0 = time.perf_counter()
    ivf = qann_amd.IVFIndex(dim, args.nlist, 32)
    if args.storage == "sq8":
        ivf.enable_sq(8)
    else:
        ivf.enable_pq(args.m, fast_scan=True)
    rng = np.random.default_rng(5)
    train_rows = min(args.n, 64 * args.nlist)
    ivf.train(x[np.sort(rng.choice(args.n, train_rows, replace=False))], max_iters=10, seed=5)
    log(f"trained on {train_rows} rows in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    index = qann_amd.RefineIndex(ivf, kf)
    for s in range(0, args.n, args.chunk):
        index.add(x[s:s + args.chunk])
        log(f"added {min(args.n, s + args.chunk):,} rows ({time.perf_counter() - t0:.1f}s)")
    sizes = ivf.list_sizes()
    log(f"lists: min {sizes.min()} median {int(np.median(sizes))} max {sizes.max()}")
    t0 = time.perf_counter()
    index.batch_query(q[:8], k)  # uploads the index
    log(f"index uploaded in {time.perf_counter() - t0:.1f}s")

    # exact top-10 of the first 1,000 queries (FP32 on the GPU, rows in chunks)
    t0 = time.perf_counter()
    nt = min(1000, args.nq)
    qt = torch.from_numpy(q[:nt]).cuda()
    best_d = torch.full((nt, k), float("inf"), device="cuda")
    best_i = torch.zeros((nt, k), dtype=torch.int64, device="cuda")
    torch.backends.cuda.matmul.allow_tf32 = False
    for s in range(0, args.n, args.chunk):
        xs = torch.from_numpy(x[s:s + args.chunk]).cuda()
        d2 = (xs * xs).sum(1)[None, :] - 2.0 * (qt @ xs.T)
        dd, ii = d2.topk(k, dim=1, largest=False)
        cd = torch.cat([best_d, dd], 1)
        ci = torch.cat([best_i, ii + s], 1)
        o = cd.topk(k, dim=1, largest=False).indices
        best_d, best_i = torch.gather(cd, 1, o), torch.gather(ci, 1, o)
        del xs, d2
    truth = best_i.cpu().numpy()
    torch.cuda.empty_cache()
    log(f"exact top-{k} of {nt} queries in {time.perf_counter() - t0:.1f}s")

    import ctypes
    import pathlib

    lib = ctypes.CDLL(str(pathlib.Path(qann_amd.__file__).parent / "libqann_hip.so"))
    lib.qann_hip_profile_enable.argtypes = [ctypes.c_int]
    lib.qann_hip_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.c_int]
    lib.qann_hip_stage_name.restype = ctypes.c_char_p
    ms = (ctypes.c_double * 32)()
    cnt = (ctypes.c_uint64 * 32)()
    cents = torch.from_numpy(ivf.centroids()).cuda()
    csq = (cents * cents).sum(1)
    # bytes the scan must read per probed list: its code blocks (fast-scan),
    # or one dim-byte row per vector (SQ8)
    if args.storage == "sq8":
        lbytes = torch.from_numpy(sizes.astype(np.int64) * dim).cuda()
    else:
        lbytes = torch.from_numpy(((sizes + 31) // 32).astype(np.int64) * args.m * 16).cuda()

    dq = torch.from_numpy(q).cuda()
    ids = torch.empty((args.nq, k), dtype=torch.int64, device="cuda")
    dists = torch.empty((args.nq, k), dtype=torch.float32, device="cuda")
    sptr = torch.cuda.current_stream().cuda_stream
    results = {}
    for npb in (int(v) for v in args.nprobes.split(",")):
        ivf.nprobe = npb
        index.search_device(dq.data_ptr(), args.nq, k, ids.data_ptr(), dists.data_ptr(), sptr)
        torch.cuda.synchronize()
        times = []
        for _ in range(args.steps):
            a = time.perf_counter()
            index.search_device(dq.data_ptr(), args.nq, k, ids.data_ptr(), dists.data_ptr(), sptr)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - a)
        got = ids[:nt].cpu().numpy()
        rec = float(np.mean([len(set(a) & set(b)) / k for a, b in zip(got, truth)]))
        best = min(times)
        # one more batch with the library's stage timers (HIP events on each
        # stage's stream), and the code-block bytes its probes cover
        lib.qann_hip_profile_read(ms, cnt, 1)
        lib.qann_hip_profile_enable(1)
        index.search_device(dq.data_ptr(), args.nq, k, ids.data_ptr(), dists.data_ptr(), sptr)
        torch.cuda.synchronize()
        lib.qann_hip_profile_enable(0)
        lib.qann_hip_profile_read(ms, cnt, 1)
        stages = {lib.qann_hip_stage_name(i).decode(): {"ms": round(ms[i], 3), "launches": int(cnt[i])}
                  for i in range(32) if cnt[i]}
        scan_bytes = 0
        for s in range(0, args.nq, 1000):
            d2 = csq[None, :] - 2.0 * (dq[s:s + 1000] @ cents.T)
            scan_bytes += int(lbytes[d2.topk(npb, dim=1, largest=False).indices].sum().item())
        scan = stages.get("scan", {"ms": 0.0, "launches": 0})
        gbs = scan_bytes / (scan["ms"] * 1e-3) / 1e9 if scan["ms"] else 0.0
        results[str(npb)] = {"qps": round(args.nq / best, 1), "ms_per_batch": round(1e3 * best, 2),
                             "recall@10": round(rec, 4), "stages": stages,
                             "scan_code_bytes": scan_bytes, "scan_gb_per_s": round(gbs, 1)}
        log(f"nprobe {npb}: {args.nq / best:,.0f} QPS, {1e3 * best:.1f} ms per {args.nq}-query batch, "
            f"recall@10 {rec:.4f}")
        log("  stages (ms per batch, launches): " +
            ", ".join(f"{n} {v['ms']:.2f} ({v['launches']})" for n, v in stages.items()))
        log(f"  scan: {scan_bytes / 1e9:.1f} GB of codes, {gbs:,.0f} GB/s over the scan launches' time")
    print(json.dumps({
        "workload": f"{args.n:,} x {dim} L2, Zipf(1.1) {args.centres}-centre mixture; IVF{args.nlist} + "
                    + ("SQ8" if args.storage == "sq8" else f"PQ{args.m}x4 fast-scan")
                    + f" + RefineIndex(k_factor={kf}) in HBM; {args.nq:,}-query batch, k={k}",
        "results": results,
        "lists": {"min": int(sizes.min()), "median": int(np.median(sizes)), "max": int(sizes.max())},
    }), flush=True)


if __name__ == "__main__":
    main()
