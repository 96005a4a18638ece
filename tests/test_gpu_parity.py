"""Parity of the HIP search path with the reference.

Every search here runs on the GPU through libqann_hip.so (the C ABI of
include/qann_hip.h, called by the qann_amd host objects). Results are
compared with
  * the results the reference library itself produced for the synthetic
    golden indexes (tests/golden/synth): ids and distance bits exact;
  * the results the reference recorded for its own shipped indexes: ids exact,
    distances within 1e-5 (its tests' bar), and bit-exact with the oracle;
  * the CPU oracle on indexes built here, at sizes it finishes in seconds.
"""
import numpy as np
import pytest

import qann_amd
from cases import REF_FIXTURES, REF_FIXTURES_OTHER, SYNTH_CASES, ref_fixture, synth_case, synth_queries

pytestmark = pytest.mark.gpu


def assert_same(got, want, msg=""):
    np.testing.assert_array_equal(got[0], want[0], err_msg=msg)
    np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32), err_msg=msg)


@pytest.mark.parametrize("name", SYNTH_CASES)
def test_gpu_bit_exact_on_reference_goldens(name):
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    for npb, k, fname, allowed, stem in synth_queries(exp, params):
        if npb is not None:
            ivf.nprobe = npb
        got = idx.batch_query(q, k, allowed=allowed)
        assert_same(got, (exp[stem + "_ids"], exp[stem + "_dists"]), f"{name} {stem}")


@pytest.mark.parametrize("name,queries", REF_FIXTURES + REF_FIXTURES_OTHER)
def test_gpu_reproduces_reference_fixtures(oracle, name, queries):
    path, q, ids, dists = ref_fixture(name, queries)
    idx = qann_amd.load(str(path))
    got_ids, got_d = idx.batch_query(q, ids.shape[1])
    np.testing.assert_array_equal(got_ids, ids)
    np.testing.assert_allclose(got_d, dists, rtol=1e-5, atol=1e-5)
    assert_same((got_ids, got_d), oracle.load(path).batch_query(q, ids.shape[1]), name)


@pytest.mark.parametrize("name", [n for n in SYNTH_CASES if not n.startswith("flat")])
def test_gpu_large_batch_matches_goldens(name):
    # the goldens' queries repeated past 2,048 rows: batches that large replay
    # the TopK without staging its loads (smaller ones stage them)
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    reps = 2048 // q.shape[0] + 1
    qq = np.tile(q, (reps, 1))
    for npb, k, fname, allowed, stem in synth_queries(exp, params):
        if npb is not None:
            ivf.nprobe = npb
        got = idx.batch_query(qq, k, allowed=allowed)
        want = (np.tile(exp[stem + "_ids"], (reps, 1)), np.tile(exp[stem + "_dists"], (reps, 1)))
        assert_same(got, want, f"{name} {stem}")


@pytest.mark.parametrize("name", ["l2_fs_refine", "ip_fs_custom", "cos_fs_refine", "l2_fs_noprecomp"])
def test_gpu_pipelined_batch_matches_goldens(monkeypatch, name):
    # unfiltered fast-scan batches of at least 8,192 queries run as a pipeline
    # of parts, each on a stream of its own (the first such search runs its
    # parts in turn, later ones overlap them): parts of uneven size, 4 and 3
    # parts, three searches each, through the synchronous entry point and, for
    # an L2 index, the stream-ordered one on device buffers
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    reps = 8200 // q.shape[0] + 1
    qq = np.ascontiguousarray(np.tile(q, (reps, 1)))
    runs = [(npb, k, stem) for npb, k, fname, allowed, stem in synth_queries(exp, params) if allowed is None]
    assert runs
    for parts in ("4", "3"):
        monkeypatch.setenv("QANN_AMD_PIPE", parts)
        for rep in range(3):
            for npb, k, stem in runs:
                if npb is not None:
                    ivf.nprobe = npb
                want = (np.tile(exp[stem + "_ids"], (reps, 1)), np.tile(exp[stem + "_dists"], (reps, 1)))
                assert_same(idx.batch_query(qq, k), want, f"{name} {stem} parts {parts} rep {rep}")
    if name == "l2_fs_refine":
        import torch

        npb, k, stem = runs[0]
        if npb is not None:
            ivf.nprobe = npb
        dq = torch.from_numpy(qq).cuda()
        ids = torch.empty((qq.shape[0], k), dtype=torch.int64, device="cuda")
        d = torch.empty((qq.shape[0], k), dtype=torch.float32, device="cuda")
        sptr = torch.cuda.current_stream().cuda_stream
        want = (np.tile(exp[stem + "_ids"], (reps, 1)), np.tile(exp[stem + "_dists"], (reps, 1)))
        for rep in range(3):
            ids.fill_(-7)
            idx.search_device(dq.data_ptr(), qq.shape[0], k, ids.data_ptr(), d.data_ptr(), sptr)
            torch.cuda.synchronize()
            assert_same((ids.cpu().numpy(), d.cpu().numpy()), want, f"{name} {stem} device rep {rep}")


@pytest.mark.parametrize("form", ["plain", "stage", "win", "win64", "gather", "gather64"])
@pytest.mark.parametrize("name", [n for n in SYNTH_CASES if not n.startswith("flat")])
def test_gpu_replay_forms_match_goldens(monkeypatch, name, form):
    # the TopK replay with its loads one block pair at a time, staged through
    # LDS, and with windows of block minima in LDS, grouped candidate loads and
    # bulk pushes (entries in the TopK, slots looked up at the end) -- a probe's
    # blocks at a time (win) or several probes' gathered into one batch
    # (gather) -- each forced, on every golden index: filtered, deleted,
    # custom ids, PQ / SQ / float lists; small and large batches
    # (win64 / gather64: windows of 64 block minima, restaged many times a query)
    monkeypatch.setenv("QANN_AMD_REPLAY", form[:-2] if form.endswith("64") else form)
    if form.endswith("64"):
        monkeypatch.setenv("QANN_AMD_REPLAY_WINDOW", "64")
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    reps = 2048 // q.shape[0] + 1
    for qq, r in ((q, 1), (np.tile(q, (reps, 1)), reps)):
        for npb, k, fname, allowed, stem in synth_queries(exp, params):
            if npb is not None:
                ivf.nprobe = npb
            got = idx.batch_query(qq, k, allowed=allowed)
            want = (np.tile(exp[stem + "_ids"], (r, 1)), np.tile(exp[stem + "_dists"], (r, 1)))
            assert_same(got, want, f"{name} {form} {stem}")


@pytest.mark.parametrize("form", ["bf16", "fp32"])
@pytest.mark.parametrize("name", [n for n in SYNTH_CASES if not n.startswith("flat")])
def test_gpu_coarse_forms_match_goldens(monkeypatch, name, form):
    # coarse scores by the split BF16 product (one BF16 MFMA GEMM over three
    # times K) with its own shortlist bound, or by the FP32 GEMM, each forced
    # (unforced, the first search of a shape times both): the same lists, so
    # the goldens; small and large batches
    monkeypatch.setenv("QANN_AMD_COARSE", form)
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    reps = 2048 // q.shape[0] + 1
    for qq, r in ((q, 1), (np.tile(q, (reps, 1)), reps)):
        for npb, k, fname, allowed, stem in synth_queries(exp, params):
            if npb is not None:
                ivf.nprobe = npb
            got = idx.batch_query(qq, k, allowed=allowed)
            want = (np.tile(exp[stem + "_ids"], (r, 1)), np.tile(exp[stem + "_dists"], (r, 1)))
            assert_same(got, want, f"{name} {form} coarse {stem}")


@pytest.mark.parametrize("form", ["thread", "wave"])
@pytest.mark.parametrize("name", [n for n in SYNTH_CASES if not n.startswith("flat")])
def test_gpu_probe_plan_forms_match_goldens(monkeypatch, name, form):
    # the probe plan (probes taken, their block offsets, the filtered stop
    # rule) by one thread per query and by one wave per query
    monkeypatch.setenv("QANN_AMD_PROBE_PLAN", form)
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    for npb, k, fname, allowed, stem in synth_queries(exp, params):
        if npb is not None:
            ivf.nprobe = npb
        assert_same(idx.batch_query(q, k, allowed=allowed), (exp[stem + "_ids"], exp[stem + "_dists"]),
                    f"{name} {form} {stem}")


@pytest.mark.parametrize("form", ["grid", "grid4", "grid6", "queue", "mfma"])
@pytest.mark.parametrize("name", ["l2_fs_refine", "ip_fs_custom", "cos_fs_refine", "l2_fs_noprecomp", "cos_l2_refine",
                                  "l2_fs"])
def test_gpu_sums_launch_forms_match_goldens(monkeypatch, name, form):
    # phase 1 as one workgroup per (query, window) (registers for four or six
    # waves per SIMD, or by batch size), as the persistent work-queue kernel,
    # and on the matrix cores (pairs grouped by list, int8 MFMA; inner-product
    # and precomputed L2 tables -- per-list residual tables keep the per-query
    # form), each forced, small and large (XCD-ordered) batches
    monkeypatch.setenv("QANN_AMD_SUMS_FORM", form if form in ("queue", "mfma") else "grid")
    if form in ("grid4", "grid6"):
        monkeypatch.setenv("QANN_AMD_SUMS_OCC", form[-1])
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    reps = 2048 // q.shape[0] + 1
    for qq, r in ((q, 1), (np.tile(q, (reps, 1)), reps)):
        for npb, k, fname, allowed, stem in synth_queries(exp, params):
            if npb is not None:
                ivf.nprobe = npb
            got = idx.batch_query(qq, k, allowed=allowed)
            want = (np.tile(exp[stem + "_ids"], (r, 1)), np.tile(exp[stem + "_dists"], (r, 1)))
            assert_same(got, want, f"{name} {form} {stem}")


@pytest.mark.parametrize("form", ["mfma", "grid"])
def test_gpu_zero_row_minimums_match_oracle(tmp_path, monkeypatch, oracle, form):
    # table rows whose minimum is exactly zero, whose sign the offset keeps:
    # half of each cluster sits on its centre and the rest at +-o around it
    # (integer entries, so the centres are the lists' exact means), so every
    # PQ codebook holds an all-zero entry; queries near the centres then give
    # rows of non-negative entries with a zero among them
    monkeypatch.setenv("QANN_AMD_SUMS_FORM", form)
    rng = np.random.default_rng(77)
    d, per = 16, 80
    centers = rng.integers(-6, 7, (100, d)).astype(np.float32) * 4.0
    off = rng.integers(-1, 2, (per // 4, d)).astype(np.float32)
    x = np.concatenate([np.concatenate([np.repeat(c[None], per // 2, 0), c + off, c - off]) for c in centers])
    x = np.ascontiguousarray(x[rng.permutation(len(x))])
    q = (centers[rng.integers(0, len(centers), 200)] + 0.01 * rng.standard_normal((200, d))).astype(np.float32)
    ivf = qann_amd.IVFIndex(d, 100, 16)
    ivf.enable_pq(8, fast_scan=True)
    ivf.precomputed_tables = True
    ivf.train(x, max_iters=6, seed=3)
    ivf.add(x)
    path = tmp_path / "zero_rows.qann"
    ivf.save(str(path))
    orc = oracle.load(path)
    for npb in (1, 3):
        ivf.nprobe = npb
        orc.set_nprobe(npb)
        reps = 2048 // q.shape[0] + 1
        qq = np.tile(q, (reps, 1))
        got = ivf.batch_query(qq, 10)
        want = orc.batch_query(q, 10)
        assert_same(got, (np.tile(want[0], (reps, 1)), np.tile(want[1], (reps, 1))), f"{form} nprobe {npb}")


@pytest.mark.parametrize("name", ["l2_fs_refine", "cos_fs_refine", "ip_pq6_custom", "l2_sq8"])
def test_gpu_graph_replays_match_goldens(name):
    # unfiltered searches run plainly three times, are captured as a HIP graph
    # the fourth and replayed after that: every call, with other searches,
    # filtered ones and host synchronizations between them, gives the goldens
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    runs = [(npb, k, stem) for npb, k, fname, allowed, stem in synth_queries(exp, params) if allowed is None]
    filt = [(npb, k, allowed, stem) for npb, k, fname, allowed, stem in synth_queries(exp, params) if allowed is not None]
    for rep in range(8):
        for npb, k, stem in runs:
            if npb is not None:
                ivf.nprobe = npb
            # 40 queries and 13 of them: two shapes, each captured once
            assert_same(idx.batch_query(q, k), (exp[stem + "_ids"], exp[stem + "_dists"]), f"{name} {stem} rep {rep}")
            assert_same(idx.batch_query(q[:13], k), (exp[stem + "_ids"][:13], exp[stem + "_dists"][:13]),
                        f"{name} {stem} 13 rows rep {rep}")
        if filt:
            npb, k, allowed, stem = filt[rep % len(filt)]
            if npb is not None:
                ivf.nprobe = npb
            assert_same(idx.batch_query(q, k, allowed=allowed), (exp[stem + "_ids"], exp[stem + "_dists"]),
                        f"{name} {stem} filtered rep {rep}")


def test_gpu_single_query_matches_batch_row():
    path, q, _, _ = synth_case("l2_fs_refine")
    idx = qann_amd.load(str(path))
    ids, d = idx.batch_query(q, 10)
    for i in range(5):
        si, sd = idx.query(q[i], 10)
        np.testing.assert_array_equal(si, ids[i])
        np.testing.assert_array_equal(sd, d[i])
    with pytest.raises(TypeError):
        idx.query(q[:2], 10)
    with pytest.raises(ValueError):
        idx.batch_query(q[:, :5], 10)


def test_gpu_edge_cases():
    path, q, exp, _ = synth_case("l2_fs")
    idx = qann_amd.load(str(path))
    # k above the number of vectors: padded with -1 / inf
    ids, d = idx.batch_query(q[:3], idx.size() + 10)
    assert (ids[:, -10:] == -1).all() and np.isinf(d[:, -10:]).all()
    # an empty or all-unknown filter returns nothing
    ids, d = idx.batch_query(q[:3], 5, allowed=[])
    assert (ids == -1).all()
    ids, d = idx.batch_query(q[:3], 5, allowed=[10**9, -5])
    assert (ids == -1).all()
    # k = 0 and no queries
    assert idx.batch_query(q[:3], 0)[0].shape == (3, 0)
    assert idx.batch_query(q[:0], 5)[0].shape == (0, 5)
    # every list probed
    idx.nprobe = 100
    ids, _ = idx.batch_query(q, 10)
    assert (ids >= 0).all()


def mixture(rng, n, d, centers=50, sigma=0.08):
    c = rng.random((centers, d)).astype(np.float32)
    lab = rng.integers(0, centers, n)
    return (c[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32)


BUILDS = [
    # metric, dim, m, nlist, k_factor (0: plain IVF)
    (qann_amd.Metric.L2, 48, 24, 128, 0),
    (qann_amd.Metric.L2, 48, 24, 128, 5),
    (qann_amd.Metric.InnerProduct, 32, 16, 100, 0),
    (qann_amd.Metric.InnerProduct, 32, 16, 100, 3),
    (qann_amd.Metric.Cosine, 100, 50, 150, 10),
    (qann_amd.Metric.Cosine, 64, 32, 100, 0),
]


@pytest.mark.parametrize("metric,dim,m,nlist,kf", BUILDS)
def test_gpu_built_index_matches_oracle(tmp_path, oracle, metric, dim, m, nlist, kf):
    rng = np.random.default_rng(dim * 7 + m + kf)
    x = mixture(rng, 12000, dim)
    q = mixture(rng, 200, dim)
    ivf = qann_amd.IVFIndex(dim, nlist, 12, metric=metric)
    ivf.enable_pq(m, fast_scan=True)
    ivf.train(x, seed=3)
    idx = qann_amd.RefineIndex(ivf, kf) if kf else ivf
    idx.add(x)
    idx.remove(rng.choice(12000, 300, replace=False))
    path = tmp_path / "built.qann"
    idx.save(str(path))
    orc = oracle.load(path)
    for npb in (1, 12, 40):
        ivf.nprobe = npb
        orc.set_nprobe(npb)
        for k in (1, 10, 100):
            assert_same(idx.batch_query(q, k), orc.batch_query(q, k), f"nprobe {npb} k {k}")
    allowed = rng.choice(12000, 1200, replace=False)
    ivf.nprobe = 12
    orc.set_nprobe(12)
    assert_same(idx.batch_query(q, 10, allowed=allowed), orc.batch_query(q, 10, allowed=allowed), "filtered")


@pytest.mark.parametrize("sample", ["1", "0"])
@pytest.mark.parametrize("nlist", [3000, 5000, 30000])
def test_gpu_many_lists_matches_oracle(tmp_path, monkeypatch, oracle, nlist, sample):
    # the coarse select's shapes: a row of 3,000 scores in registers of 256
    # threads, 5,000 in registers of 1,024 threads (up to 16,384 lists), 30,000
    # read from memory; each with its threshold from a sample of the row first
    # (the default above 16,384 lists) and without
    monkeypatch.setenv("QANN_AMD_SELECT_SAMPLE", sample)
    rng = np.random.default_rng(nlist)
    x = mixture(rng, 100_000, 32, centers=400, sigma=0.05)
    q = mixture(rng, 300, 32, centers=400, sigma=0.05)
    ivf = qann_amd.IVFIndex(32, nlist, 16)
    ivf.enable_pq(16, fast_scan=True)
    ivf.train(x, max_iters=4, seed=2)
    idx = qann_amd.RefineIndex(ivf, 4)
    idx.add(x)
    path = tmp_path / "many_lists.qann"
    idx.save(str(path))
    orc = oracle.load(path)
    for npb in (1, 16, 300):
        ivf.nprobe = npb
        orc.set_nprobe(npb)
        assert_same(idx.batch_query(q, 10), orc.batch_query(q, 10), f"nprobe {npb}")
    allowed = rng.choice(100_000, 20_000, replace=False)
    ivf.nprobe = 16
    orc.set_nprobe(16)
    assert_same(idx.batch_query(q, 10, allowed=allowed), orc.batch_query(q, 10, allowed=allowed), "filtered")


def test_gpu_flat_matches_oracle(tmp_path, oracle):
    rng = np.random.default_rng(11)
    for metric in (qann_amd.Metric.L2, qann_amd.Metric.InnerProduct, qann_amd.Metric.Cosine):
        x = mixture(rng, 5000, 37)
        q = mixture(rng, 64, 37)
        flat = qann_amd.FlatIndex(37, metric=metric, custom_ids=True)
        flat.add(x, ids=rng.choice(10**7, 5000, replace=False))
        path = tmp_path / "flat.qann"
        flat.save(str(path))
        orc = oracle.load(path)
        for k in (1, 16, 300):
            assert_same(flat.batch_query(q, k), orc.batch_query(q, k), f"{metric} k {k}")


def test_gpu_flat_gemm_precision_shape(tmp_path, oracle):
    # the precision configuration's shape at a size the oracle scores in seconds:
    # 768-dim unit rows with an i^-0.8 spectrum, inner product, k = 100 -- scored
    # by the MFMA path (8 or more queries), with deletions and a large filter
    rng = np.random.default_rng(768)
    lam = np.arange(1, 769, dtype=np.float64) ** -0.8
    x = (rng.standard_normal((6000, 768)) * np.sqrt(lam)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    q = (rng.standard_normal((48, 768)) * np.sqrt(lam)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    for metric in (qann_amd.Metric.InnerProduct, qann_amd.Metric.L2):
        flat = qann_amd.FlatIndex(768, metric=metric)
        flat.add(x)
        flat.remove(rng.choice(6000, 200, replace=False))
        path = tmp_path / "flat768.qann"
        flat.save(str(path))
        orc = oracle.load(path)
        for k in (1, 100):
            assert_same(flat.batch_query(q, k), orc.batch_query(q, k), f"{metric} k {k}")
        allowed = rng.choice(6000, 3500, replace=False)
        assert_same(flat.batch_query(q, 100, allowed=allowed), orc.batch_query(q, 100, allowed=allowed),
                    f"{metric} filtered")


def test_gpu_flat_gemm_many_ties(tmp_path, oracle):
    # 1500 copies of one row nearest to every query: more rows within the
    # shortlist bound than it holds, so every row is scored exactly; ties go by slot
    rng = np.random.default_rng(5)
    base = rng.random(32).astype(np.float32)
    x = np.concatenate([np.repeat(base[None], 1500, axis=0), rng.random((2500, 32)).astype(np.float32) + 0.5])
    q = (base[None] + 0.01 * rng.standard_normal((16, 32))).astype(np.float32)
    flat = qann_amd.FlatIndex(32, metric=qann_amd.Metric.L2, custom_ids=True)
    flat.add(x, ids=rng.choice(10**6, 4000, replace=False))
    path = tmp_path / "ties.qann"
    flat.save(str(path))
    orc = oracle.load(path)
    for k in (10, 1200):
        assert_same(flat.batch_query(q, k), orc.batch_query(q, k), f"k {k}")


def _stage_launches():
    """Launches per search stage since the last read (the library's profiler,
    which times every stage while it is on)."""
    import ctypes
    import pathlib

    lib = ctypes.CDLL(str(pathlib.Path(qann_amd.__file__).parent / "libqann_hip.so"))
    lib.qann_hip_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.c_int]
    lib.qann_hip_stage_name.restype = ctypes.c_char_p
    ms = (ctypes.c_double * 32)()
    n = (ctypes.c_uint64 * 32)()
    lib.qann_hip_profile_read(ms, n, 1)
    names = [lib.qann_hip_stage_name(i).decode() for i in range(32)]
    return {name: int(n[i]) for i, name in enumerate(names) if name}


def _profiling(on):
    import ctypes
    import pathlib

    lib = ctypes.CDLL(str(pathlib.Path(qann_amd.__file__).parent / "libqann_hip.so"))
    lib.qann_hip_profile_enable.argtypes = [ctypes.c_int]
    lib.qann_hip_profile_enable(int(on))


def spectrum_rows(rng, n, dim=768):
    lam = np.arange(1, dim + 1, dtype=np.float64) ** -0.8
    x = (rng.standard_normal((n, dim)) * np.sqrt(lam)).astype(np.float32)
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def test_gpu_flat_bf16_matches_oracle(tmp_path, oracle):
    # 70,000 rows (the BF16 GEMM path starts at 65,536) of the precision
    # configuration's kind: every query finishes on the BF16 path, and the
    # results are the reference's, bit for bit
    rng = np.random.default_rng(769)
    x = spectrum_rows(rng, 70_000)
    q = spectrum_rows(rng, 24)
    _profiling(True)
    try:
        for metric in (qann_amd.Metric.InnerProduct, qann_amd.Metric.L2):
            flat = qann_amd.FlatIndex(768, metric=metric)
            flat.add(x)
            flat.remove(rng.choice(70_000, 500, replace=False))
            path = tmp_path / "flat_bf16.qann"
            flat.save(str(path))
            orc = oracle.load(path)
            _stage_launches()
            for k in (1, 10, 100):
                assert_same(flat.batch_query(q, k), orc.batch_query(q, k), f"{metric} k {k}")
            allowed = rng.choice(70_000, 40_000, replace=False)
            assert_same(flat.batch_query(q, 100, allowed=allowed), orc.batch_query(q, 100, allowed=allowed),
                        f"{metric} filtered")
            launches = _stage_launches()
            assert launches["flat"] == 4 and launches["flat_fp32"] == 0, launches
    finally:
        _profiling(False)


@pytest.mark.parametrize("shape", ["0", "1", "2", "3", "4", "5", "6"])
def test_gpu_bf16_gemm_shapes_match_oracle(tmp_path, oracle, monkeypatch, shape):
    # each tile shape of the BF16 GEMM forced (3 and 4: the 256 x 256 tiles fed
    # by direct loads through the LDS ring, stages of K 32 and 64; K = 96 of
    # the coarse split product takes stages of 32 either way) on the flat
    # search's filter GEMM, with
    # row and query counts that leave partial tiles, and on the split-BF16
    # coarse scores of a golden index
    monkeypatch.setenv("QANN_AMD_GEMM_SHAPE", shape)
    rng = np.random.default_rng(771)
    x = spectrum_rows(rng, 70_001)
    q = spectrum_rows(rng, 300)
    for metric in (qann_amd.Metric.InnerProduct, qann_amd.Metric.L2):
        flat = qann_amd.FlatIndex(768, metric=metric)
        flat.add(x)
        path = tmp_path / "flat_shape.qann"
        flat.save(str(path))
        orc = oracle.load(path)
        for k in (1, 100):
            assert_same(flat.batch_query(q, k), orc.batch_query(q, k), f"shape {shape} {metric} k {k}")
    monkeypatch.setenv("QANN_AMD_COARSE", "bf16")
    path, qg, exp, params = synth_case("cos_fs_refine")
    idx = qann_amd.load(str(path))
    ivf = idx.base
    qq = np.tile(qg, (2048 // qg.shape[0] + 1, 1))
    for npb, k, fname, allowed, stem in synth_queries(exp, params):
        if npb is not None:
            ivf.nprobe = npb
        r = qq.shape[0] // qg.shape[0]
        want = (np.tile(exp[stem + "_ids"], (r, 1)), np.tile(exp[stem + "_dists"], (r, 1)))
        assert_same(idx.batch_query(qq, k, allowed=allowed), want, f"shape {shape} coarse {stem}")


def test_gpu_flat_bf16_hands_queries_to_fp32(tmp_path, oracle):
    # 4,500 copies of one row: the queries next to it have more rows within the
    # BF16 bound than the candidate buffer holds and go to the FP32 path, the
    # others finish on the BF16 path; all match the reference
    rng = np.random.default_rng(770)
    x = spectrum_rows(rng, 70_000)
    x[rng.choice(70_000, 4500, replace=False)] = x[7]
    near = x[7][None] + 0.01 * spectrum_rows(rng, 4)
    q = np.concatenate([spectrum_rows(rng, 12), near / np.linalg.norm(near, axis=1, keepdims=True)]).astype(np.float32)
    flat = qann_amd.FlatIndex(768, metric=qann_amd.Metric.InnerProduct, custom_ids=True)
    flat.add(x, ids=rng.choice(10**9, 70_000, replace=False))
    path = tmp_path / "flat_dup.qann"
    flat.save(str(path))
    orc = oracle.load(path)
    _profiling(True)
    try:
        _stage_launches()
        for k in (10, 100):
            assert_same(flat.batch_query(q, k), orc.batch_query(q, k), f"k {k}")
        launches = _stage_launches()
        assert launches["flat"] == 2 and launches["flat_fp32"] == 2, launches
    finally:
        _profiling(False)


def test_gpu_flat_bf16_loose_bound_uses_fp32(tmp_path, oracle):
    # rows far from the origin and close together: the BF16 bound admits whole
    # clusters, so the queries go to the FP32 path (and, after a search where
    # most did, the index keeps to it); the results match the reference
    rng = np.random.default_rng(771)
    x = mixture(rng, 70_000, 64, centers=8, sigma=0.02)
    q = mixture(rng, 16, 64, centers=8, sigma=0.02)
    flat = qann_amd.FlatIndex(64, metric=qann_amd.Metric.L2)
    flat.add(x)
    path = tmp_path / "flat_far.qann"
    flat.save(str(path))
    orc = oracle.load(path)
    for k in (10, 100):
        assert_same(flat.batch_query(q, k), orc.batch_query(q, k), f"k {k}")


def test_gpu_config_100k_x_128_matches_oracle(tmp_path, oracle):
    # the CPU-runnable configuration at its full size: 100,000 x 128 L2 from a
    # 64-centre mixture (centres U[0,1)^128, sigma 0.05), seed 1, IVF nlist 256
    # nprobe 16, PQ32 fast-scan, RefineIndex k_factor 10, k = 10, 1,000 queries
    rng = np.random.Generator(np.random.PCG64(1))
    centres = rng.random((64, 128), dtype=np.float32)
    x = (centres[rng.integers(0, 64, 100_000)] + 0.05 * rng.standard_normal((100_000, 128), dtype=np.float32))
    q = (centres[rng.integers(0, 64, 1000)] + 0.05 * rng.standard_normal((1000, 128), dtype=np.float32))
    ivf = qann_amd.IVFIndex(128, 256, 16)
    ivf.enable_pq(32, fast_scan=True)
    ivf.train(x, seed=1)
    idx = qann_amd.RefineIndex(ivf, 10)
    idx.add(x)
    path = tmp_path / "config1.qann"
    idx.save(str(path))
    got = idx.batch_query(q, 10)
    assert_same(got, oracle.load(path).batch_query(q, 10), "100k x 128")
    # recall@10 against exact search (float64): 0.7635 for this index -- the
    # 128-dim clusters' distances concentrate, so PQ32x4 and a k_factor 10
    # re-rank keep about three quarters of the exact 10; a guard against
    # regressions in training, not a target
    d2 = ((q.astype(np.float64) ** 2).sum(1)[:, None] - 2 * q.astype(np.float64) @ x.T.astype(np.float64)
          + (x.astype(np.float64) ** 2).sum(1)[None, :])
    truth = np.argpartition(d2, 10, axis=1)[:, :10]
    recall = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(got[0], truth)])
    assert recall > 0.74, recall


@pytest.mark.parametrize("name", ["l2_fs", "ip_fs_custom", "cos_fs_refine", "cos_l2_refine", "rot_refine", "l2_pq8",
                                  "ip_pq6_custom", "cos_pq8_refine", "l2_sq8", "ip_sq4_custom", "cos_sq8_refine",
                                  "l2_float", "ip_float_refine"])
def test_gpu_add_writes_the_reference_bytes(tmp_path, reference, name):
    # IVFIndex::add_batch_impl (src/ivf_index.cpp:111-195): the same vectors added
    # to the same reference-built index by qann_amd (nearest list on the GPU, PQ
    # encoding and fast-scan packing on the host) and by the reference library
    # save to the same bytes
    path, q, _, params = synth_case(name)
    rng = np.random.default_rng(len(name))
    new = (q[rng.integers(0, len(q), 400)] + 0.03 * rng.standard_normal((400, q.shape[1]))).astype(np.float32)
    ids = None
    if params.get("custom"):
        ids = (np.int64(1) << 40) + np.arange(400, dtype=np.int64) * 3
    ours = qann_amd.load(str(path))
    ours.add(new, ids=ids) if ids is not None else ours.add(new)
    ours.save(str(tmp_path / "ours.qann"))
    ref = reference.load(path)
    reference.add(ref, new, ids)
    reference.save(ref, tmp_path / "ref.qann")
    a = (tmp_path / "ours.qann").read_bytes()
    b = (tmp_path / "ref.qann").read_bytes()
    assert len(a) == len(b) and a == b, f"{name}: files differ ({len(a)} vs {len(b)} bytes)"


STORAGES = [
    # metric, dim, how the lists store vectors, k_factor
    (qann_amd.Metric.L2, 32, "pq8", 0),
    (qann_amd.Metric.Cosine, 40, "pq8", 4),
    (qann_amd.Metric.InnerProduct, 24, "sq4", 0),
    (qann_amd.Metric.L2, 21, "sq8", 3),
    (qann_amd.Metric.Cosine, 19, "float", 0),
    (qann_amd.Metric.InnerProduct, 16, "float", 2),
    # 8-bit SQ rows of a multiple of 16 bytes take the multi-row scan kernel
    (qann_amd.Metric.L2, 64, "sq8", 0),
    (qann_amd.Metric.InnerProduct, 96, "sq8", 3),
    (qann_amd.Metric.Cosine, 128, "sq8", 0),
]


@pytest.mark.parametrize("metric,dim,storage,kf", STORAGES)
def test_gpu_built_storages_match_oracle(tmp_path, oracle, metric, dim, storage, kf):
    # indexes of PQ (one byte per subspace), SQ or float lists trained and
    # filled by qann_amd, searched on the GPU and by the oracle from the file
    rng = np.random.default_rng(dim * 13 + kf)
    x = mixture(rng, 8000, dim)
    q = mixture(rng, 120, dim)
    ivf = qann_amd.IVFIndex(dim, 100, 8, metric=metric)
    if storage == "pq8":
        ivf.enable_pq(dim // 4)
    elif storage.startswith("sq"):
        ivf.enable_sq(int(storage[2:]))
    ivf.train(x, seed=5)
    idx = qann_amd.RefineIndex(ivf, kf) if kf else ivf
    idx.add(x)
    idx.remove(rng.choice(8000, 150, replace=False))
    path = tmp_path / "built.qann"
    idx.save(str(path))
    orc = oracle.load(path)
    for npb in (1, 8, 30):
        ivf.nprobe = npb
        orc.set_nprobe(npb)
        for k in (1, 10, 64):
            assert_same(idx.batch_query(q, k), orc.batch_query(q, k), f"{storage} nprobe {npb} k {k}")
    for size in (200, 4000):  # a direct scan (no PQ, at most 3 nlist) and a probed filter
        allowed = rng.choice(8000, size, replace=False)
        assert_same(idx.batch_query(q, 10, allowed=allowed), orc.batch_query(q, 10, allowed=allowed),
                    f"{storage} filtered {size}")


@pytest.mark.parametrize("form", ["rows", "thread"])
@pytest.mark.parametrize("metric", [qann_amd.Metric.L2, qann_amd.Metric.InnerProduct])
def test_gpu_sq8_scan_forms_match_oracle(tmp_path, monkeypatch, oracle, metric, form):
    # 8-bit SQ lists of 128-byte rows, scored four rows per lane (the default)
    # or one byte load per dimension (QANN_AMD_SQ8=thread); lists of about 600
    # rows, so a window spans many block groups; removed rows and filters
    if form == "thread":
        monkeypatch.setenv("QANN_AMD_SQ8", "thread")
    rng = np.random.default_rng(128 + int(metric == qann_amd.Metric.InnerProduct))
    x = mixture(rng, 60000, 128)
    q = mixture(rng, 300, 128)
    ivf = qann_amd.IVFIndex(128, 100, 8, metric=metric)
    ivf.enable_sq(8)
    ivf.train(x, seed=3)
    ivf.add(x)
    ivf.remove(rng.choice(60000, 500, replace=False))
    path = tmp_path / "sq8.qann"
    ivf.save(str(path))
    orc = oracle.load(path)
    for npb in (3, 20):
        ivf.nprobe = npb
        orc.set_nprobe(npb)
        for k in (10, 100):
            assert_same(ivf.batch_query(q, k), orc.batch_query(q, k), f"{form} nprobe {npb} k {k}")
    allowed = rng.choice(60000, 24000, replace=False)
    assert_same(ivf.batch_query(q, 10, allowed=allowed), orc.batch_query(q, 10, allowed=allowed), f"{form} filtered")


def test_gpu_stage_profiler_times_without_changing_results():
    # the stage timers record events around each stage without synchronizing:
    # results under them are the goldens, every stage of the IVF + refine path
    # is counted once per search, times are positive, and a read resets
    import ctypes
    import pathlib

    lib = ctypes.CDLL(str(pathlib.Path(qann_amd.__file__).parent / "libqann_hip.so"))
    lib.qann_hip_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.c_int]
    lib.qann_hip_stage_name.restype = ctypes.c_char_p
    path, q, exp, params = synth_case("l2_fs_refine")
    idx = qann_amd.load(str(path))
    ivf = idx.base
    runs = [(npb, k, stem) for npb, k, fname, allowed, stem in synth_queries(exp, params) if allowed is None]
    _stage_launches()
    _profiling(True)
    try:
        for npb, k, stem in runs:
            if npb is not None:
                ivf.nprobe = npb
            assert_same(idx.batch_query(q, k), (exp[stem + "_ids"], exp[stem + "_dists"]), f"profiled {stem}")
    finally:
        _profiling(False)
    ms = (ctypes.c_double * 32)()
    n = (ctypes.c_uint64 * 32)()
    lib.qann_hip_profile_read(ms, n, 1)
    stages = {lib.qann_hip_stage_name(i).decode(): (ms[i], int(n[i])) for i in range(32)
              if lib.qann_hip_stage_name(i)}
    for name in ("coarse", "select", "scan", "replay", "rerank"):
        assert stages[name][1] == len(runs), (name, stages[name])
        assert stages[name][0] > 0.0, (name, stages[name])
    assert all(v == 0 for v in _stage_launches().values())


@pytest.mark.parametrize("dim", [8193, 10000])
def test_gpu_long_queries_match_oracle(tmp_path, monkeypatch, oracle, dim):
    # queries longer than the kernels' LDS copy (8,192 floats) are read from
    # global memory: flat search by the FP32 GEMM, the BF16 GEMM (its row
    # threshold lowered) and the per-slot select (fewer than 8 queries, and
    # small filters); IVF over float lists with refine (coarse select, list
    # distances, re-rank) -- against the oracle, ids and distance bits
    rng = np.random.default_rng(dim)
    x = mixture(rng, 3000, dim, centers=20, sigma=0.05)
    q = mixture(rng, 20, dim, centers=20, sigma=0.05)
    for metric in (qann_amd.Metric.L2, qann_amd.Metric.InnerProduct):
        for form in ("fp32", "bf16"):
            monkeypatch.setenv("QANN_AMD_FLAT_BF16_ROWS", "1000" if form == "bf16" else "1000000000")
            flat = qann_amd.FlatIndex(dim, metric=metric)
            flat.add(x)
            flat.remove(rng.choice(3000, 50, replace=False))
            path = tmp_path / "flat_long.qann"
            flat.save(str(path))
            orc = oracle.load(path)
            for qq in (q, q[:3]):
                for k in (1, 10, 100):
                    assert_same(flat.batch_query(qq, k), orc.batch_query(qq, k), f"{metric} {form} nq {len(qq)} k {k}")
            allowed = rng.choice(3000, 40, replace=False)
            assert_same(flat.batch_query(q, 10, allowed=allowed), orc.batch_query(q, 10, allowed=allowed),
                        f"{metric} {form} filtered")
    ivf = qann_amd.IVFIndex(dim, 100, 6, metric=qann_amd.Metric.L2)
    ivf.train(x, seed=1)
    idx = qann_amd.RefineIndex(ivf, 3)
    idx.add(x)
    path = tmp_path / "ivf_long.qann"
    idx.save(str(path))
    orc = oracle.load(path)
    for k in (1, 10):
        assert_same(idx.batch_query(q, k), orc.batch_query(q, k), f"ivf k {k}")
    allowed = rng.choice(3000, 900, replace=False)
    assert_same(idx.batch_query(q, 10, allowed=allowed), orc.batch_query(q, 10, allowed=allowed), "ivf filtered")


PQ8_WIDE = [
    # metric, dim, m (K = 256: a table of m x 256 floats, 64 KB at m = 64), precomputed, SDC
    (qann_amd.Metric.L2, 128, 64, True, False),
    (qann_amd.Metric.L2, 128, 64, False, False),
    (qann_amd.Metric.InnerProduct, 128, 64, False, False),
    (qann_amd.Metric.Cosine, 192, 96, True, False),
    (qann_amd.Metric.L2, 192, 96, False, True),
    (qann_amd.Metric.L2, 256, 256, True, False),
    (qann_amd.Metric.InnerProduct, 256, 256, False, False),
    (qann_amd.Metric.Cosine, 256, 256, False, True),
]


@pytest.mark.parametrize("limit", ["device", "24576"])
@pytest.mark.parametrize("metric,dim,m,pre,sdc", PQ8_WIDE)
def test_gpu_wide_pq8_tables_match_oracle(tmp_path, monkeypatch, oracle, metric, dim, m, pre, sdc, limit):
    # one-byte PQ lists whose per-(query, list) table alone exceeds 64 KB take
    # gfx950's larger LDS, one probe per workgroup; tables beyond that (256 KB
    # at PQ256x8), or beyond a lowered limit, are built and summed a run of
    # subspaces at a time
    if limit != "device":
        monkeypatch.setenv("QANN_AMD_LDS_LIMIT", limit)
    rng = np.random.default_rng(dim + m)
    x = mixture(rng, 6000, dim, centers=30)
    q = mixture(rng, 60, dim, centers=30)
    ivf = qann_amd.IVFIndex(dim, 100, 8, metric=metric)
    if sdc:
        ivf.pq_distance = qann_amd.PQDistance.SDC
    ivf.enable_pq(m, 256)
    ivf.train(x, seed=5)
    ivf.precomputed_tables = pre
    if metric == qann_amd.Metric.Cosine:
        ivf.cosine_inner_product = False
    ivf.add(x)
    path = tmp_path / "wide_pq8.qann"
    ivf.save(str(path))
    orc = oracle.load(path)
    for npb in (1, 8, 30):
        ivf.nprobe = npb
        orc.set_nprobe(npb)
        for k in (1, 10, 50):
            assert_same(ivf.batch_query(q, k), orc.batch_query(q, k), f"nprobe {npb} k {k}")
    allowed = rng.choice(6000, 1500, replace=False)
    ivf.nprobe = 8
    orc.set_nprobe(8)
    assert_same(ivf.batch_query(q, 10, allowed=allowed), orc.batch_query(q, 10, allowed=allowed), "filtered")


@pytest.mark.parametrize("limit", ["device", "65536"])
def test_gpu_large_topk_and_rerank_buffers_match_oracle(tmp_path, monkeypatch, oracle, limit):
    # k' = 4,000 and 10,000 candidates: TopK buffers of 96 and 240 KB in HBM;
    # re-rank keys of 32 and 128 KB in LDS under the device's limit, and with
    # the limit lowered to 64 KB the larger in HBM
    if limit != "device":
        monkeypatch.setenv("QANN_AMD_LDS_LIMIT", limit)
    rng = np.random.default_rng(44)
    x = mixture(rng, 20000, 32, centers=40)
    q = mixture(rng, 30, 32, centers=40)
    ivf = qann_amd.IVFIndex(32, 100, 20)
    ivf.enable_pq(16, fast_scan=True)
    ivf.train(x, seed=2)
    idx = qann_amd.RefineIndex(ivf, 10)
    idx.add(x)
    path = tmp_path / "big_k.qann"
    idx.save(str(path))
    orc = oracle.load(path)
    orc.set_nprobe(20)
    for k in (400, 1000):
        assert_same(idx.batch_query(q, k), orc.batch_query(q, k), f"{limit} k {k}")
