"""Precision benchmark: exact FlatIndex search of the precision configuration
on one MI355X.

Rows: 1,000,000 x 768 float32, x = z * lambda^(1/2) with lambda_i = i^-0.8 and
z ~ N(0, I), each row unit-normalised; 4,096 queries drawn the same way
(numpy PCG64, seed 3, rows then queries). Inner product, k = 100.

Three searches of the same batch, queries and results resident in HBM:

* ``bf16``: FlatIndex's BF16 MFMA GEMM with FP32 accumulation, a proven bound
  on its error, and the exact FP32 rescoring of the rows that bound leaves in
  play (the k-th smallest approximate score plus twice the bound): exact
  results, identical to the reference's FlatIndex;
* ``fp32``: the same index kept on the FP32 MFMA GEMM (QANN_AMD_FLAT_FP32=1);
* ``torch``: plain torch FP32 GEMM + topk over the rows (not tie-exact).

Against an FP64 oracle on a sample of queries: recall@100 of the returned ids
and the largest distance error; the BF16 GEMM's own score error beside its
bound. Prints one JSON object (last line of stdout); progress goes to stderr.

Run from the repository root:  python3 -m qann_amd.flat_bench [--n 1000000]
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


def make_rows(rng, n, dim, lam_sqrt):
    out = np.empty((n, dim), np.float32)
    for s in range(0, n, 100_000):
        e = min(n, s + 100_000)
        z = rng.standard_normal((e - s, dim), dtype=np.float32) * lam_sqrt
        out[s:e] = z / np.linalg.norm(z, axis=1, keepdims=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=4096)
    ap.add_argument("--dim", type=int, default=768)
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--oracle-queries", type=int, default=256)
    ap.add_argument("--library-gemm-rows", type=int, default=0,
                    help="only time the vendor BF16 GEMM (torch.mm, hipBLASLt) on this many rows x nq x dim")
    args = ap.parse_args()

    import torch
    import qann_amd

    if args.library_gemm_rows:
        # the same GEMM shape through the library: BF16 operands, BF16 and FP32 outputs
        a = torch.randn(args.library_gemm_rows, args.dim, device="cuda", dtype=torch.bfloat16)
        b = torch.randn(args.dim, args.nq, device="cuda", dtype=torch.bfloat16)
        out = {}
        for name, fn in (("bf16_out", lambda: torch.mm(a, b)), ("fp32_out", lambda: torch.mm(a, b, out_dtype=torch.float32))):
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(10):
                fn()
            torch.cuda.synchronize()
            s = (time.perf_counter() - t) / 10
            out[name] = {"ms": round(1e3 * s, 3), "tflops": round(2 * args.library_gemm_rows * args.nq * args.dim / s / 1e12, 1)}
            log(f"torch.mm {args.library_gemm_rows} x {args.nq} x {args.dim} {name}: {1e3 * s:.3f} ms, "
                f"{out[name]['tflops']} TFLOP/s")
        print(json.dumps({"library_gemm": out}))
        return

    rng = np.random.Generator(np.random.PCG64(3))
    lam_sqrt = np.sqrt(np.arange(1, args.dim + 1, dtype=np.float64) ** -0.8).astype(np.float32)
    t0 = time.perf_counter()
    x = make_rows(rng, args.n, args.dim, lam_sqrt)
    q = make_rows(rng, args.nq, args.dim, lam_sqrt)
    log(f"data {x.shape} + {q.shape} in {time.perf_counter() - t0:.1f}s")

    dq = torch.from_numpy(q).cuda()
    sptr = torch.cuda.current_stream().cuda_stream
    results = {}

    def run(name, index):
        ids = torch.empty((args.nq, args.k), dtype=torch.int64, device="cuda")
        d = torch.empty((args.nq, args.k), dtype=torch.float32, device="cuda")

        def step():
            index.search_device(dq.data_ptr(), args.nq, args.k, ids.data_ptr(), d.data_ptr(), sptr)

        step()  # uploads the index
        torch.cuda.synchronize()
        best = None
        for _ in range(args.steps):
            t = time.perf_counter()
            step()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        results[name] = {"ms_per_batch": round(1e3 * best, 3), "qps": round(args.nq / best, 1)}
        log(f"{name}: {1e3 * best:.2f} ms per {args.nq}-query batch, {args.nq / best:,.0f} QPS")
        return ids.cpu().numpy(), d.cpu().numpy()

    t0 = time.perf_counter()
    flat = qann_amd.FlatIndex(args.dim, metric=qann_amd.Metric.InnerProduct)
    flat.add(x)
    ids_b, d_b = run("bf16", flat)
    log(f"bf16 path (upload included in the first call) in {time.perf_counter() - t0:.1f}s")
    del flat
    os.environ["QANN_AMD_FLAT_FP32"] = "1"
    flat32 = qann_amd.FlatIndex(args.dim, metric=qann_amd.Metric.InnerProduct)
    flat32.add(x)
    ids_f, d_f = run("fp32", flat32)
    del flat32
    os.environ.pop("QANN_AMD_FLAT_FP32")
    identical = bool(np.array_equal(ids_b, ids_f) and np.array_equal(d_b.view(np.uint32), d_f.view(np.uint32)))
    log(f"bf16 and fp32 paths identical: {identical}")

    # plain torch FP32 GEMM + topk
    xs = torch.from_numpy(x).cuda()
    torch.backends.cuda.matmul.allow_tf32 = False

    def plain():
        for s in range(0, args.nq, 1024):
            (dq[s:s + 1024] @ xs.T).topk(args.k, dim=1)

    plain()
    torch.cuda.synchronize()
    best = None
    for _ in range(args.steps):
        t = time.perf_counter()
        plain()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    results["torch"] = {"ms_per_batch": round(1e3 * best, 3), "qps": round(args.nq / best, 1)}
    log(f"torch: {1e3 * best:.2f} ms per batch, {args.nq / best:,.0f} QPS")

    # FP64 oracle on a sample of queries
    no = min(args.oracle_queries, args.nq)
    x64 = xs.double()
    q64 = dq[:no].double()
    s64 = q64 @ x64.T                      # exact inner products (FP64)
    top64 = s64.topk(args.k, dim=1).indices.cpu().numpy()
    rec = float(np.mean([len(set(a) & set(b)) / args.k for a, b in zip(ids_b[:no], top64)]))
    got = torch.from_numpy(ids_b[:no]).cuda()
    exact_d = -torch.gather(s64, 1, got)   # the index's distance is -q.x
    derr = (torch.from_numpy(d_b[:no]).cuda().double() - exact_d).abs().max().item()
    # the score error of BF16 inputs with FP32 accumulation (what the MFMA GEMM
    # computes), and the bound the search uses
    xb = xs.bfloat16().float()
    qb = dq[:no].bfloat16().float()
    sb = (qb @ xb.T).double()
    gemm_err = (sb - s64).abs().max().item()
    rx = (xs - xb).norm(dim=1).max().item()
    rq = (dq[:no] - qb).norm(dim=1).max().item()
    u = 2.0 ** -24
    bound = (qb.norm(dim=1).max().item() * rx + rq * 1.0 + (args.dim + 2) * u * 1.01 + (args.dim + 3) * u) * 1.0625
    log(f"FP64 oracle ({no} queries): recall@{args.k} {rec:.4f}, max distance error {derr:.3g}; "
        f"BF16 GEMM score error max {gemm_err:.3g} (bound about {bound:.3g})")
    print(json.dumps({
        "workload": f"exact FlatIndex, {args.n:,} x {args.dim} unit rows (i^-0.8 spectrum), inner product, "
                    f"{args.nq} queries, k={args.k}",
        "results": results,
        "speedup_bf16_over_fp32": round(results["fp32"]["ms_per_batch"] / results["bf16"]["ms_per_batch"], 2),
        "speedup_bf16_over_torch": round(results["torch"]["ms_per_batch"] / results["bf16"]["ms_per_batch"], 2),
        "bf16_equals_fp32_bit_exact": identical,
        "fp64_oracle": {"queries": no, f"recall@{args.k}": rec, "max_distance_error": derr,
                        "bf16_gemm_score_error_max": gemm_err, "bf16_bound_approx": bound},
    }), flush=True)


if __name__ == "__main__":
    main()
