"""The saved-index format: files the reference wrote load here, and files
written here load in the CPU checkers with the same search results. Host
only (no GPU work)."""
import numpy as np
import pytest

import qann_amd
from cases import REF_FIXTURES, REF_FIXTURES_OTHER, SYNTH_CASES, ref_fixture, synth_case


@pytest.mark.parametrize("name,queries", REF_FIXTURES + REF_FIXTURES_OTHER)
def test_reference_files_round_trip(tmp_path, oracle, name, queries):
    path, q, _, _ = ref_fixture(name, queries)
    idx = qann_amd.load(str(path))
    out = tmp_path / "x.qann"
    idx.save(str(out))
    a = oracle.load(path).batch_query(q, 10)
    b = oracle.load(out).batch_query(q, 10)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@pytest.mark.parametrize("name", SYNTH_CASES)
def test_synthetic_files_round_trip_in_reference(tmp_path, reference, name):
    path, q, _, _ = synth_case(name)
    idx = qann_amd.load(str(path))
    out = tmp_path / "x.qann"
    idx.save(str(out))
    a = reference.load(path).batch_query(q, 10)
    b = reference.load(out).batch_query(q, 10)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def test_loaded_properties():
    idx = qann_amd.load(str(ref_fixture("v12_ivf_cos_l2", "v12_cos_queries")[0]))
    assert isinstance(idx, qann_amd.IVFIndex)
    assert idx.metric == qann_amd.Metric.Cosine and not idx.cosine_inner_product and idx.precomputed_tables
    assert idx.fast_scan and idx.size() == 3000 and idx.dim() == 32
    idx = qann_amd.load(str(ref_fixture("v11_ivf_cos_fs", "v11_cos_queries")[0]))
    assert idx.cosine_inner_product and not idx.precomputed_tables
    r = qann_amd.load(str(ref_fixture("v9_ip_refine_ivf_fs", "v9_ip_queries")[0]))
    assert isinstance(r, qann_amd.RefineIndex) and isinstance(r.base, qann_amd.IVFIndex)
    assert r.k_factor == 6 and r.base.pq_rotate and r.metric == qann_amd.Metric.InnerProduct
    s = qann_amd.load(str(synth_case("cos_pq8_sdc_refine")[0]))
    assert s.base.pq_distance == qann_amd.PQDistance.SDC and not s.base.fast_scan


def test_corrupt_and_truncated_files_raise(tmp_path):
    path = ref_fixture("v4_ivf_fs", "v4_queries")[0]
    data = path.read_bytes()
    bad = tmp_path / "bad.qann"
    bad.write_bytes(b"XXXX" + data[4:])
    with pytest.raises(RuntimeError, match="not a qann index file"):
        qann_amd.load(str(bad))
    bad.write_bytes(data[: len(data) // 2])
    with pytest.raises(RuntimeError):
        qann_amd.load(str(bad))
    bad.write_bytes(data + b"\0")
    with pytest.raises(RuntimeError, match="unexpected data"):
        qann_amd.load(str(bad))


def test_api_validation_without_gpu():
    with pytest.raises(ValueError):
        qann_amd.IVFIndex(16, 50)
    ivf = qann_amd.IVFIndex(16, 100, 8)
    with pytest.raises(ValueError):
        ivf.nprobe = 0
    with pytest.raises(ValueError):
        ivf.enable_pq(6, fast_scan=True, centroids_per_subspace=8)
    ivf.enable_pq(8, fast_scan=True)
    assert ivf.fast_scan
    # SDC: not with fast-scan, not with inner product (ivf_index.hpp:94-100, ivf_index.cpp:89-90)
    assert ivf.pq_distance == qann_amd.PQDistance.ADC
    with pytest.raises(ValueError, match="fast-scan"):
        ivf.pq_distance = qann_amd.PQDistance.SDC
    ip = qann_amd.IVFIndex(16, 100, 8, metric=qann_amd.Metric.InnerProduct)
    with pytest.raises(ValueError, match="InnerProduct"):
        ip.pq_distance = qann_amd.PQDistance.SDC
    sdc = qann_amd.IVFIndex(16, 100, 8)
    sdc.pq_distance = qann_amd.PQDistance.SDC
    with pytest.raises(ValueError, match="fast-scan"):
        sdc.enable_pq(8, fast_scan=True)
    sdc.enable_pq(8, fast_scan=False, centroids_per_subspace=256)
    assert sdc.pq_distance == qann_amd.PQDistance.SDC and not sdc.precomputed_tables
    with pytest.raises(ValueError):
        qann_amd.RefineIndex(ivf, 0)
