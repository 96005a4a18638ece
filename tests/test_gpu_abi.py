"""Searches through the C ABI of include/qann_hip.h itself, with no host index
objects in between: captured searches replayed after the index's data has
changed give what a fresh handle gives."""
import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu


def float_index(rng, n=6000, dim=24, nlist=100):
    x = rng.standard_normal((n, dim)).astype(np.float32)
    cent = x[rng.choice(n, nlist, replace=False)].copy()
    d2 = (x ** 2).sum(1)[:, None] - 2 * x @ cent.T + (cent ** 2).sum(1)[None, :]
    assign = d2.argmin(1)
    lists = [np.flatnonzero(assign == l) for l in range(nlist)]
    return x, cent, lists


def test_gpu_graph_replays_follow_set_ids_and_set_lists():
    # a shape searched 6 times is captured (after 3 plain runs) and replayed;
    # set_ids (new deletions) and set_lists (new lists) must make the capture
    # stale: every search after them equals a fresh handle's first search
    rng = np.random.default_rng(21)
    lib = capi.lib()
    x, cent, lists = float_index(rng)
    q = x[rng.choice(len(x), 40, replace=False)] + 0.01 * rng.standard_normal((40, x.shape[1])).astype(np.float32)
    h = capi.float_ivf(lib, cent, lists, x)
    h.set_ids(len(x))
    for _ in range(6):
        ids, d = h.search(q, 10, nprobe=4)
    fresh = capi.float_ivf(lib, cent, lists, x)
    fresh.set_ids(len(x))
    want = fresh.search(q, 10, nprobe=4)
    np.testing.assert_array_equal(ids, want[0])
    # delete every row the queries found
    deleted = np.zeros(len(x), np.uint8)
    deleted[ids[ids >= 0]] = 1
    h.set_ids(len(x), deleted=deleted)
    fresh2 = capi.float_ivf(lib, cent, lists, x)
    fresh2.set_ids(len(x), deleted=deleted)
    want2 = fresh2.search(q, 10, nprobe=4)
    for rep in range(3):
        got = h.search(q, 10, nprobe=4)
        np.testing.assert_array_equal(got[0], want2[0], err_msg=f"after set_ids, rep {rep}")
        np.testing.assert_array_equal(got[1].view(np.uint32), want2[1].view(np.uint32))
        assert not np.isin(got[0][got[0] >= 0], np.flatnonzero(deleted)).any()
    # custom labels through set_ids
    labels = (np.int64(1) << 40) + np.arange(len(x), dtype=np.int64) * 7
    h.set_ids(len(x), labels=labels, deleted=deleted)
    got = h.search(q, 10, nprobe=4)
    np.testing.assert_array_equal(got[0], np.where(want2[0] >= 0, labels[np.maximum(want2[0], 0)], -1))
    # new lists: half of every list moved to the next list
    h.set_ids(len(x))
    for _ in range(5):
        h.search(q, 10, nprobe=4)
    moved = [np.concatenate([lists[l][: len(lists[l]) // 2], lists[l - 1][len(lists[l - 1]) // 2:]])
             for l in range(len(lists))]
    capi_sizes = np.array([len(v) for v in moved], np.uint32)
    slots = np.ascontiguousarray(np.concatenate(moved).astype(np.uint32))
    rows = np.ascontiguousarray(x[slots.astype(np.int64)])
    h.check(lib.qann_hip_ivf_set_lists(h.h, capi.ptr(capi_sizes), capi.ptr(slots), capi.ptr(rows)))
    fresh3 = capi.float_ivf(lib, cent, moved, x)
    fresh3.set_ids(len(x))
    want3 = fresh3.search(q, 10, nprobe=4)
    for rep in range(3):
        got = h.search(q, 10, nprobe=4)
        np.testing.assert_array_equal(got[0], want3[0], err_msg=f"after set_lists, rep {rep}")
        np.testing.assert_array_equal(got[1].view(np.uint32), want3[1].view(np.uint32))


def test_gpu_abi_search_matches_exact_over_probed_lists():
    # the float-list IVF through the C ABI against numpy: the k nearest by
    # exact squared L2 among the rows of each query's nprobe nearest lists
    rng = np.random.default_rng(22)
    lib = capi.lib()
    x, cent, lists = float_index(rng, n=4000, dim=16)
    h = capi.float_ivf(lib, cent, lists, x)
    h.set_ids(len(x))
    q = rng.standard_normal((30, 16)).astype(np.float32)
    ids, d = h.search(q, 8, nprobe=5)
    x64, q64, c64 = x.astype(np.float64), q.astype(np.float64), cent.astype(np.float64)
    for i in range(len(q)):
        probe = np.argsort(((q64[i] - c64) ** 2).sum(1), kind="stable")[:5]
        cand = np.concatenate([lists[l] for l in probe])
        dist = ((x64[cand] - q64[i]) ** 2).sum(1)
        top = cand[np.argsort(dist, kind="stable")[:8]]
        assert set(ids[i]) == set(top), i
        np.testing.assert_allclose(d[i], np.sort(dist)[:8], rtol=1e-5, atol=1e-6)
