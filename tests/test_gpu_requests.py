"""Several requests in one launch (qann_hip_search_requests, the engine of the
serving loop): each query keeps its own k, nprobe and filter, and every
request's rows equal what the reference recorded for them, and what the same
request returns searched alone."""
import numpy as np
import pytest

import qann_amd
from cases import SYNTH_CASES, synth_case, synth_queries

pytestmark = pytest.mark.gpu


def assert_same(got, want, msg=""):
    np.testing.assert_array_equal(got[0], want[0], err_msg=msg)
    np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32), err_msg=msg)


@pytest.mark.parametrize("name", SYNTH_CASES)
def test_gpu_requests_in_one_launch_match_goldens(name):
    # every recorded (nprobe, k, filter) result of a golden index, split into
    # two requests each, all in one call in a shuffled order
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    reqs, rows, want = [], [], []
    half = len(q) // 2
    for npb, k, fname, allowed, stem in synth_queries(exp, params):
        for lo, hi in ((0, half), (half, len(q))):
            reqs.append((hi - lo, k, npb or 0, allowed))
            rows.append(q[lo:hi])
            want.append((exp[stem + "_ids"][lo:hi], exp[stem + "_dists"][lo:hi], stem))
    order = np.random.default_rng(len(name)).permutation(len(reqs))
    got = idx.search_requests(np.concatenate([rows[i] for i in order]), [reqs[i] for i in order])
    assert len(got) == len(reqs)
    for j, i in enumerate(order):
        assert_same(got[j], want[i][:2], f"{name} {want[i][2]}")


def mixture(rng, n, d, centers=60, sigma=0.08):
    c = rng.random((centers, d)).astype(np.float32)
    lab = rng.integers(0, centers, n)
    return (c[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32)


@pytest.mark.parametrize("storage,kf", [("fs", 4), ("fs", 0), ("sq8", 0), ("pq8", 3)])
def test_gpu_mixed_requests_equal_single_searches(storage, kf):
    # requests of k 1 / 10 / 100, nprobe 2 / 16 / 64 and filters of 0.1% (too
    # few allowed rows in the first guess of lists: searched again with every
    # list), 5%, 50% or none, mixed in one launch, against each alone
    rng = np.random.default_rng(17 + kf)
    n, d = 60000, 32
    x = mixture(rng, n, d)
    q = mixture(rng, 400, d)
    ivf = qann_amd.IVFIndex(d, 1024, 8)
    if storage == "fs":
        ivf.enable_pq(16, 16, fast_scan=True)
    elif storage == "pq8":
        ivf.enable_pq(8, 256)
    else:
        ivf.enable_sq(8)
    ivf.train(x[:30000], seed=4)
    idx = qann_amd.RefineIndex(ivf, kf) if kf else ivf
    idx.add(x)
    idx.remove(rng.choice(n, 300, replace=False))
    live = idx.ids()
    filters = [None, rng.choice(live, 60, replace=False), rng.choice(live, 3000, replace=False),
               rng.choice(live, 30000, replace=False), np.array([], np.int64)]
    reqs, rows = [], []
    lo = 0
    while lo < len(q):
        m = int(rng.integers(1, 40))
        hi = min(len(q), lo + m)
        reqs.append((hi - lo, int(rng.choice([1, 10, 100])), int(rng.choice([2, 16, 64])),
                     filters[int(rng.integers(0, len(filters)))]))
        rows.append(q[lo:hi])
        lo = hi
    got = idx.search_requests(q, reqs)
    for (nq, k, nprobe, allowed), qq, g in zip(reqs, rows, got):
        ivf.nprobe = nprobe
        want = idx.batch_query(qq, k, allowed=allowed)
        nf = "none" if allowed is None else len(allowed)
        assert_same(g, want, f"{storage} kf {kf} k {k} nprobe {nprobe} filter {nf}")
