"""The serving loop (qann_amd.Server): requests coalesced into launches and
run concurrently return exactly what each returns alone."""
import numpy as np
import pytest

import qann_amd
from cases import synth_case
from conftest import has_gpu


def test_server_needs_a_gpu():
    if has_gpu():
        pytest.skip("a GPU is present")
    path, _, _, _ = synth_case("l2_fs_refine")
    idx = qann_amd.load(str(path))
    with pytest.raises(RuntimeError):
        qann_amd.Server(idx)


def test_server_rejects_bad_arguments():
    path, _, _, _ = synth_case("l2_fs_refine")
    idx = qann_amd.load(str(path))
    with pytest.raises((ValueError, RuntimeError)):
        qann_amd.Server(idx, workers=0)
    with pytest.raises((ValueError, RuntimeError)):
        qann_amd.Server(idx, memory_fraction=1.5)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["l2_fs_refine", "cos_l2_refine", "ip_fs_custom", "l2_sq8", "flat_l2_custom"])
def test_gpu_results_alone_and_in_the_mix(name):
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    is_ivf = isinstance(ivf, qann_amd.IVFIndex)
    rng = np.random.default_rng(3)
    ids_all = idx.ids()
    filters = [None, rng.choice(ids_all, size=len(ids_all) // 2, replace=False), rng.choice(ids_all, 30, replace=False)]
    server = qann_amd.Server(idx, workers=3, max_batch=64)
    reqs = []
    for r in range(60):
        lo = int(rng.integers(0, len(q) - 1))
        n = int(rng.integers(1, len(q) - lo + 1))
        k = int(rng.choice([1, 10, 50]))
        nprobe = int(rng.choice([1, 4, 9])) if is_ivf else 0
        allowed = filters[int(rng.integers(0, 3))]
        reqs.append((q[lo:lo + n], k, nprobe, allowed, server.submit(q[lo:lo + n], k, nprobe=nprobe, allowed=allowed)))
    for qq, k, nprobe, allowed, fut in reqs:
        ids, d, lat = fut.result()
        assert lat >= 0
        if is_ivf:
            ivf.nprobe = nprobe
        want = idx.batch_query(qq, k, allowed=allowed)
        np.testing.assert_array_equal(ids, want[0])
        np.testing.assert_array_equal(d.view(np.uint32), want[1].view(np.uint32))
    st = server.stats()
    assert st["requests"] == 60 and st["launches"] <= 60 and st["p99_ms"] >= st["p50_ms"] > 0
    server.close()


@pytest.mark.gpu
def test_gpu_poisson_load_reports_latency():
    path, q, _, _ = synth_case("l2_fs_refine")
    idx = qann_amd.load(str(path))
    server = qann_amd.Server(idx, workers=2)
    st = qann_amd.serve_load(server, q, rate=400.0, seconds=0.5, ks=[10, 50], nprobes=[2, 6], max_nq=16,
                             filtered_fraction=0.3, filters=[idx.ids()[::10]])
    assert st["requests"] > 50 and st["queries"] >= st["requests"]
    assert 0 < st["p50_ms"] <= st["p90_ms"] <= st["p99_ms"] <= st["max_ms"]
    server.close()
