"""Probe counts above 4,096 and filtered searches whose stopping rule needs
more lists than the first guess (IVFIndex takes nprobe in [1, nlist],
src/ivf_index.cpp:51-54, and a filtered scan keeps probing until its rule
holds, :510-521), against the oracle: ids and distance bits equal."""
import numpy as np
import pytest

import qann_amd

pytestmark = pytest.mark.gpu


def assert_same(got, want, msg=""):
    np.testing.assert_array_equal(got[0], want[0], err_msg=msg)
    np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32), err_msg=msg)


def clustered(rng, n, dim, centres):
    c = rng.random((centres, dim), dtype=np.float32)
    return (c[rng.integers(0, centres, n)] + 0.04 * rng.standard_normal((n, dim), dtype=np.float32)).astype(np.float32)


def build(tmp_path, oracle, n, dim, nlist, m, seed, refine=0):
    rng = np.random.default_rng(seed)
    x = clustered(rng, n + 64, dim, 2 * nlist)
    x, q = np.ascontiguousarray(x[:n]), np.ascontiguousarray(x[n:])
    ivf = qann_amd.IVFIndex(dim, nlist, 16)
    ivf.enable_pq(m, fast_scan=True)
    ivf.train(x, max_iters=3, seed=seed)
    idx = qann_amd.RefineIndex(ivf, refine) if refine else ivf
    idx.add(x)
    path = tmp_path / f"probes_{nlist}.qann"
    idx.save(str(path))
    return rng, idx, ivf, q, oracle.load(path)


def test_gpu_probes_above_4096_and_narrow_filters_16384_lists(tmp_path, oracle):
    rng, idx, ivf, q, o = build(tmp_path, oracle, 200_000, 32, 16384, 16, 81)
    for npb in (8192, 16384):
        ivf.nprobe = npb
        o.set_nprobe(npb)
        assert_same(idx.batch_query(q, 10), o.batch_query(q, 10), f"nprobe {npb}")
    # 0.1% and 1% filters at nprobe 16: the first guess of probes (twice
    # min_probes) is too few for most (0.1%) or a few (1%) queries, which are
    # searched again with every list; and a 0.1% filter at nprobe 8192
    from test_gpu_parity import _profiling, _stage_launches

    for frac, npb in ((0.001, 16), (0.01, 16), (0.001, 8192)):
        allowed = rng.choice(200_000, int(200_000 * frac), replace=False)
        ivf.nprobe = npb
        o.set_nprobe(npb)
        for k in (10, 50):
            _profiling(True)
            try:
                _stage_launches()
                got = idx.batch_query(q, k, allowed=allowed)
                launches = _stage_launches()
            finally:
                _profiling(False)
            assert_same(got, o.batch_query(q, k, allowed=allowed), f"filter {frac} nprobe {npb} k {k}")
            if (frac, npb) == (0.001, 16):
                # the queries the first guess could not finish were searched again
                assert launches["scan"] == 2, launches
    # without probe scaling the rule still holds
    ivf.filter_probe_scaling = False
    o.set_filter_probe_scaling(False)
    ivf.nprobe = 16
    o.set_nprobe(16)
    allowed = rng.choice(200_000, 200, replace=False)
    assert_same(idx.batch_query(q, 10, allowed=allowed), o.batch_query(q, 10, allowed=allowed), "no scaling")


def test_gpu_probes_above_16384_at_65535_lists(tmp_path, oracle):
    rng, idx, ivf, q, o = build(tmp_path, oracle, 150_000, 16, 65535, 8, 82)
    for npb in (20000, 65535):
        ivf.nprobe = npb
        o.set_nprobe(npb)
        assert_same(idx.batch_query(q[:24], 10), o.batch_query(q[:24], 10), f"nprobe {npb}")
    allowed = rng.choice(150_000, 30_000, replace=False)
    ivf.nprobe = 40
    o.set_nprobe(40)
    assert_same(idx.batch_query(q, 10, allowed=allowed), o.batch_query(q, 10, allowed=allowed), "filtered")
