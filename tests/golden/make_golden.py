"""Writes the synthetic golden fixtures under tests/golden/synth/.

Each case is an index trained, filled and saved by the reference library
itself (oracle/_ref/libqann_ref.so, maxtmiller/qann's vecengine core compiled
from its own sources by oracle/Makefile), 40 queries from the same
distribution, and the reference's batch_query results for several k, nprobe
and allowed-id filters. The tests compare the oracle and the HIP path with
these outputs bit for bit.

Run in the build container (where the reference sources are): python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle.refpy import Reference  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent / "synth"

L2, COS, IP = 0, 1, 2

CASES = [
    # name, dict(build parameters)
    ("l2_fs", dict(d=32, n=3000, nlist=100, nprobe=8, metric=L2, m=16, kf=0, remove=60)),
    ("l2_fs_noprecomp", dict(d=32, n=3000, nlist=100, nprobe=8, metric=L2, m=16, kf=0, precomputed=False)),
    ("l2_fs_refine", dict(d=32, n=2500, nlist=100, nprobe=6, metric=L2, m=16, kf=4, remove=50)),
    ("ip_fs_custom", dict(d=16, n=3000, nlist=100, nprobe=10, metric=IP, m=8, kf=0, custom=True, remove=40)),
    ("cos_fs_refine", dict(d=40, n=2500, nlist=128, nprobe=12, metric=COS, m=20, kf=3)),
    ("cos_l2_refine", dict(d=40, n=2500, nlist=128, nprobe=12, metric=COS, m=20, kf=3, cosine_ip=False)),
    ("rot_refine", dict(d=32, n=2500, nlist=100, nprobe=8, metric=L2, m=16, kf=2, rotate=True)),
    ("flat_l2_custom", dict(d=24, n=2000, flat=True, metric=L2, custom=True, remove=30)),
    ("flat_ip", dict(d=16, n=2000, flat=True, metric=IP)),
    ("flat_cos", dict(d=20, n=2000, flat=True, metric=COS)),
    # IVF lists stored as 8-bit PQ codes (ADC), SQ codes or float rows; "half"
    # adds a filter above 3 nlist slots, so SQ / float lists go through the
    # probes rather than IVFIndex::direct_scan
    ("l2_pq8", dict(d=32, n=3000, nlist=100, nprobe=8, metric=L2, m=8, pq_k=256, fast=False, kf=0, remove=60,
                    half=True)),
    ("l2_pq8_noprecomp", dict(d=32, n=3000, nlist=100, nprobe=8, metric=L2, m=8, pq_k=256, fast=False, kf=0,
                              precomputed=False)),
    ("ip_pq6_custom", dict(d=24, n=3000, nlist=100, nprobe=10, metric=IP, m=6, pq_k=64, fast=False, kf=0,
                           custom=True, remove=40)),
    ("cos_pq8_refine", dict(d=40, n=2500, nlist=128, nprobe=12, metric=COS, m=10, pq_k=256, fast=False, kf=3)),
    ("cos_pq8_l2", dict(d=40, n=2500, nlist=128, nprobe=12, metric=COS, m=10, pq_k=256, fast=False, kf=0,
                        cosine_ip=False)),
    ("l2_sq8", dict(d=32, n=3000, nlist=100, nprobe=8, metric=L2, m=0, sq=8, kf=0, remove=40, half=True)),
    ("ip_sq4_custom", dict(d=30, n=3000, nlist=100, nprobe=8, metric=IP, m=0, sq=4, kf=0, custom=True, half=True)),
    ("cos_sq8_refine", dict(d=20, n=2500, nlist=100, nprobe=8, metric=COS, m=0, sq=8, kf=3, half=True)),
    ("l2_float", dict(d=19, n=3000, nlist=100, nprobe=8, metric=L2, m=0, kf=0, remove=30, half=True)),
    ("ip_float_refine", dict(d=24, n=2500, nlist=100, nprobe=6, metric=IP, m=0, kf=2, half=True)),
    # PQ codes scored by SDC (the query's residual encoded, then code-to-code
    # table sums); 256 and 16 centroids per subspace, L2 and Cosine
    ("l2_pq8_sdc", dict(d=32, n=3000, nlist=100, nprobe=8, metric=L2, m=8, pq_k=256, fast=False, kf=0, remove=60,
                        half=True, sdc=True)),
    ("l2_pq16x16_sdc", dict(d=32, n=3000, nlist=100, nprobe=8, metric=L2, m=16, pq_k=16, fast=False, kf=0,
                            sdc=True)),
    ("cos_pq8_sdc_refine", dict(d=40, n=2500, nlist=128, nprobe=12, metric=COS, m=10, pq_k=256, fast=False, kf=3,
                                sdc=True)),
]


def mixture(rng, n, d, centers, sigma, metric):
    c = rng.random((centers, d)).astype(np.float32)
    if metric == IP:
        # lengths over 0.2x..5x so inner product differs from cosine
        scale = np.exp(rng.uniform(np.log(0.2), np.log(5.0), size=(centers, 1))).astype(np.float32)
        c = c * scale
    lab = rng.integers(0, centers, n)
    return (c[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32)


def main():
    ref = Reference()
    ref.set_threads(1)
    OUT.mkdir(exist_ok=True)
    manifest = {}
    for ci, (name, p) in enumerate(CASES):
        rng = np.random.default_rng(1000 + ci)
        data = mixture(rng, p["n"] + 40, p["d"], 24, 0.06, p["metric"])
        x, queries = data[: p["n"]], data[p["n"]:]
        ids = None
        if p.get("custom"):
            # unique ids, a few beyond 2^32 so results sort by (distance, id) as floats
            ids = rng.choice(10**6, size=p["n"], replace=False).astype(np.int64) * 7919
            ids[:: 97] += np.int64(1) << 33
        if p.get("flat"):
            index = ref.build_flat(x, metric=p["metric"], ids=ids)
        else:
            index = ref.build_ivf(x, p["nlist"], p["nprobe"], metric=p["metric"], m=p["m"], pq_k=p.get("pq_k", 16),
                                  fast_scan=p.get("fast", True), rotate=p.get("rotate", False), sq_bits=p.get("sq", 0),
                                  seed=7, k_factor=p["kf"], ids=ids)
            if "cosine_ip" in p:
                ref._check(ref.set_cosine_ip(index._h, int(p["cosine_ip"])))
            if "precomputed" in p:
                ref._check(ref.set_precomputed(index._h, int(p["precomputed"])))
            if p.get("sdc"):
                index.set_pq_distance(True)
        if p.get("remove"):
            pool = ids if ids is not None else np.arange(p["n"])
            ref.remove(index, rng.choice(pool, size=p["remove"], replace=False))
        path = OUT / f"{name}.qann"
        ref.save(index, path)
        index = ref.load(path)
        np.save(OUT / f"{name}_queries.npy", queries)
        live = index.info()["size"]
        pool = ids if ids is not None else np.arange(p["n"])
        filters = {
            "none": None,
            "tenth": np.sort(rng.choice(pool, size=p["n"] // 10, replace=False)),
            "few": np.sort(rng.choice(pool, size=25, replace=False)),
        }
        if p.get("half"):
            filters["half"] = np.sort(rng.choice(pool, size=p["n"] // 2, replace=False))
        results = {}
        nprobes = [None] if p.get("flat") else [p["nprobe"], 1, p["nlist"] // 3]
        for npb in nprobes:
            if npb is not None:
                index.set_nprobe(npb)
            for k in (1, 10, 50):
                for fname, allowed in filters.items():
                    got_ids, got_d = index.batch_query(queries, k, allowed=allowed)
                    key = f"p{npb or 0}_k{k}_{fname}"
                    results[key + "_ids"] = got_ids
                    results[key + "_dists"] = got_d
        for fname, allowed in filters.items():
            if allowed is not None:
                results[f"filter_{fname}"] = allowed
        np.savez_compressed(OUT / f"{name}_expected.npz", **results)
        manifest[name] = {k: v for k, v in p.items()}
        manifest[name]["live"] = live
        print(name, "live", live, "keys", len(results))
    (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
