"""Training and adding on the GPU against the reference library.

IVFIndex::train (src/ivf_index.cpp:725-806: e-means or Lloyd's k-means for the
coarse centroids, residual centres, PQ codebooks with seed + 1, SQ ranges) and
IVFIndex::add_batch_impl (:111-195: nearest list, PQ / SQ encoding, fast-scan
packing) run by qann_amd -- every assign step, the encoding and the packing on
the GPU -- and by the reference library built from its sources, on the same
sample and seed: the saved files must be the same bytes. The Cosine table mode
is this library's own choice, so it is copied from the reference's file before
the comparison (the trained arrays are not touched).
"""
import numpy as np
import pytest

import qann_amd

pytestmark = pytest.mark.gpu

METRIC_ID = {qann_amd.Metric.L2: 0, qann_amd.Metric.Cosine: 1, qann_amd.Metric.InnerProduct: 2}


def mixture(rng, n, d, centers=40, sigma=0.1):
    c = rng.random((centers, d)).astype(np.float32)
    lab = rng.integers(0, centers, n)
    return (c[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32)


TRAINS = [
    # name, metric, dim, nlist, storage, m, clustering, iters or passes, k_factor, custom ids, rotate
    ("l2_fs_emeans", qann_amd.Metric.L2, 64, 200, "fs", 16, "emeans", 2, 0, False, False),
    ("cos_fs_emeans_refine", qann_amd.Metric.Cosine, 100, 256, "fs", 50, "emeans", 2, 4, False, False),
    ("ip_fs_lloyd_custom", qann_amd.Metric.InnerProduct, 48, 150, "fs", 24, "lloyd", 6, 0, True, False),
    ("l2_fs_d16_emeans", qann_amd.Metric.L2, 16, 128, "fs", 8, "emeans", 3, 0, False, False),
    ("l2_pq8_lloyd", qann_amd.Metric.L2, 32, 128, "pq8", 8, "lloyd", 5, 0, False, False),
    ("cos_pq8_emeans_refine", qann_amd.Metric.Cosine, 40, 100, "pq8", 10, "emeans", 2, 3, False, False),
    ("l2_sq8_emeans", qann_amd.Metric.L2, 48, 150, "sq8", 0, "emeans", 2, 0, False, False),
    ("cos_sq4_lloyd", qann_amd.Metric.Cosine, 24, 100, "sq4", 0, "lloyd", 4, 0, False, False),
    ("ip_float_emeans_refine", qann_amd.Metric.InnerProduct, 20, 100, "float", 0, "emeans", 3, 2, False, False),
    ("l2_fs_rot_emeans_refine", qann_amd.Metric.L2, 64, 128, "fs", 16, "emeans", 2, 3, False, True),
    ("cos_pq8_rot_lloyd", qann_amd.Metric.Cosine, 48, 100, "pq8", 12, "lloyd", 4, 0, False, True),
]


def build_ours(x, train_n, metric, dim, nlist, storage, m, clustering, iters, kf, ids, rot=False):
    ivf = qann_amd.IVFIndex(dim, nlist, 8, metric=metric, custom_ids=ids is not None and not kf)
    if storage == "fs":
        ivf.enable_pq(m, 16, fast_scan=True, rotate=rot)
    elif storage == "pq8":
        ivf.enable_pq(m, 256, rotate=rot)
    elif storage.startswith("sq"):
        ivf.enable_sq(int(storage[2:]))
    kind = qann_amd.Clustering.EMeans if clustering == "emeans" else qann_amd.Clustering.Lloyd
    if clustering == "emeans":
        ivf.train(x[:train_n], seed=11, clustering=kind, passes=iters)
    else:
        ivf.train(x[:train_n], iters, seed=11, clustering=kind)
    return ivf


@pytest.mark.parametrize("name,metric,dim,nlist,storage,m,clustering,iters,kf,custom,rot", TRAINS)
def test_gpu_trained_and_added_index_is_the_reference_bytes(tmp_path, reference, name, metric, dim, nlist, storage, m,
                                                            clustering, iters, kf, custom, rot):
    rng = np.random.default_rng(len(name) * 7 + dim)
    n, train_n = 12000, 9000
    x = mixture(rng, n, dim)
    more = mixture(rng, 700, dim)
    ids = more_ids = None
    if custom:
        ids = (np.int64(1) << 36) + np.arange(n, dtype=np.int64) * 5
        more_ids = (np.int64(1) << 37) + np.arange(700, dtype=np.int64) * 3
    # the reference: train on the first train_n rows, add all n, then 700 more
    ref = reference.build_ivf(x, nlist, 8, metric=METRIC_ID[metric], m=m if storage in ("fs", "pq8") else 0,
                              pq_k=16 if storage == "fs" else 256, fast_scan=storage == "fs", rotate=rot,
                              sq_bits=int(storage[2:]) if storage.startswith("sq") else 0, seed=11, k_factor=kf,
                              ids=ids, clustering=0 if clustering == "emeans" else 1, iters_or_passes=iters,
                              train_n=train_n)
    ref_first = tmp_path / "ref1.qann"
    reference.save(ref, ref_first)
    reference.add(ref, more, more_ids)
    reference.save(ref, tmp_path / "ref2.qann")

    ivf = build_ours(x, train_n, metric, dim, nlist, storage, m, clustering, iters, kf, ids, rot)
    if metric == qann_amd.Metric.Cosine and storage in ("fs", "pq8"):
        loaded = qann_amd.load(str(ref_first))
        ivf.cosine_inner_product = (loaded.base if kf else loaded).cosine_inner_product
    idx = qann_amd.RefineIndex(ivf, kf, custom_ids=custom) if kf else ivf
    idx.add(x, ids=ids) if custom else idx.add(x)
    idx.save(str(tmp_path / "ours1.qann"))
    a, b = (tmp_path / "ours1.qann").read_bytes(), ref_first.read_bytes()
    assert len(a) == len(b) and a == b, f"{name}: trained + added files differ ({len(a)} vs {len(b)} bytes)"
    # a second add continues the lists' partly filled blocks
    idx.add(more, ids=more_ids) if custom else idx.add(more)
    idx.save(str(tmp_path / "ours2.qann"))
    a, b = (tmp_path / "ours2.qann").read_bytes(), (tmp_path / "ref2.qann").read_bytes()
    assert len(a) == len(b) and a == b, f"{name}: files after a second add differ"


def test_gpu_trained_index_searches_as_the_reference(tmp_path, reference):
    # the index trained on the GPU answers queries as the reference's own
    # trained index does: ids and distance bits
    rng = np.random.default_rng(5)
    x = mixture(rng, 20000, 64)
    q = mixture(rng, 200, 64)
    ref = reference.build_ivf(x, 256, 12, metric=0, m=32, seed=2, k_factor=5)
    ivf = qann_amd.IVFIndex(64, 256, 12)
    ivf.enable_pq(32, 16, fast_scan=True)
    ivf.train(x, seed=2)
    idx = qann_amd.RefineIndex(ivf, 5)
    idx.add(x)
    for k in (1, 10, 50):
        got = idx.batch_query(q, k)
        want = ref.batch_query(q, k)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32))


@pytest.mark.parametrize("metric", [qann_amd.Metric.L2, qann_amd.Metric.Cosine])
def test_gpu_sdc_index_searches_as_the_reference(reference, metric):
    # PQ codes scored by SDC (src/ivf_index.cpp:626-632), set before training
    # here and after it in the reference (SDC leaves L2 training alone, and
    # Cosine's table mode only matters under ADC): the same ids and distance
    # bits, filtered too; then both switched back to ADC
    rng = np.random.default_rng(17 + METRIC_ID[metric])
    x = mixture(rng, 6000, 32)
    q = mixture(rng, 120, 32)
    ref = reference.build_ivf(x, 128, 10, metric=METRIC_ID[metric], m=8, pq_k=256, fast_scan=False, seed=4)
    ref.set_pq_distance(True)
    ivf = qann_amd.IVFIndex(32, 128, 10, metric=metric)
    ivf.pq_distance = qann_amd.PQDistance.SDC
    ivf.enable_pq(8, 256)
    ivf.train(x, seed=4)
    ivf.add(x)
    allowed = np.sort(rng.choice(6000, 900, replace=False)).astype(np.int64)
    for mode in ("sdc", "adc"):
        if mode == "adc":
            ref.set_pq_distance(False)
            ivf.pq_distance = qann_amd.PQDistance.ADC
            if metric == qann_amd.Metric.Cosine:
                ivf.cosine_inner_product = False
                reference._check(reference.set_cosine_ip(ref._h, 0))
        for k, al in ((1, None), (10, None), (50, None), (10, allowed)):
            got = ivf.batch_query(q, k, allowed=al)
            want = ref.batch_query(q, k, allowed=al)
            np.testing.assert_array_equal(got[0], want[0], err_msg=f"{mode} k={k}")
            np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32), err_msg=f"{mode} k={k}")


def test_gpu_subspace_assign_and_pack_through_the_abi():
    # qann_hip_assign_subspaces against a numpy restatement of nearest_centroid
    # on sub-vectors of 2 floats (l2_distance's scalar tail: two fused steps,
    # exact here since every product and sum below is a small dyadic value),
    # and qann_hip_fs_pack against fs_set's layout
    import ctypes as C

    import capi

    lib = capi.lib()
    lib.qann_hip_assign_subspaces.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                                              C.c_void_p]
    lib.qann_hip_fs_pack.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p]
    rng = np.random.default_rng(3)
    n, dim, m, K = 300, 8, 4, 16
    rows = (rng.integers(-8, 8, (n, dim)) / 4.0).astype(np.float32)
    cents = (rng.integers(-8, 8, (m, K, dim // m)) / 4.0).astype(np.float32)
    out = np.zeros((n, m), np.uint8)
    rc = lib.qann_hip_assign_subspaces(capi.ptr(rows), n, dim, m, capi.ptr(cents), K, capi.ptr(out))
    assert rc == 0
    sub = rows.reshape(n, m, 1, dim // m)
    dist = ((sub - cents[None]) ** 2).sum(-1)  # n x m x K, exact in float32 here
    np.testing.assert_array_equal(out, dist.argmin(-1))  # argmin: first on ties

    codes = rng.integers(0, 16, (50, 6)).astype(np.uint8)
    row_at = np.full(3 * 32, 0xFFFFFFFF, np.uint32)
    row_at[5:55] = np.arange(50, dtype=np.uint32)
    packed = np.zeros(3 * 6 * 16, np.uint8)
    rc = lib.qann_hip_fs_pack(capi.ptr(codes), 50, 6, capi.ptr(row_at), 3, capi.ptr(packed))
    assert rc == 0
    want = np.zeros((3, 6, 16), np.uint8)
    for p in range(5, 55):
        b, v = divmod(p, 32)
        for s in range(6):
            want[b, s, v % 16] |= codes[p - 5, s] << (4 * (v >= 16))
    np.testing.assert_array_equal(packed.reshape(3, 6, 16), want)
    bad = row_at.copy()
    bad[0] = 50  # a row beyond n is refused before any launch
    assert lib.qann_hip_fs_pack(capi.ptr(codes), 50, 6, capi.ptr(bad), 3, capi.ptr(packed)) == capi.EINVAL
