"""where= filtered searches on the GPU: a metadata filter gives what the same
rows passed as allowed= ids give (the reference's own check,
tests/test_python.py:1246-1290), for flat, IVF fast-scan and refine indexes;
the reference's saved metadata index (tests/data/v9_refine_metadata.qann)
searched as its tests/test_python.py:1729-1749 demands; and the serving loop
under where= filters, against the oracle."""
import numpy as np
import pytest

import qann_amd
from cases import REF_DATA

pytestmark = pytest.mark.gpu

COLORS = np.array(["red", "green", "blue", "black"])


def make_index(kind, data, ids):
    if kind == "flat":
        return qann_amd.FlatIndex(32, custom_ids=True)
    ivf = qann_amd.IVFIndex(32, 100, 100, custom_ids=(kind == "ivf"))
    ivf.enable_pq(16, fast_scan=True)
    ivf.train(data, seed=1)
    return ivf if kind == "ivf" else qann_amd.RefineIndex(ivf, 5, custom_ids=True)


@pytest.mark.parametrize("kind", ["flat", "ivf", "refine"])
def test_gpu_where_equals_allowed_of_the_matching_rows(tmp_path, kind):
    rng = np.random.default_rng(7)
    data = rng.random((4000, 32), dtype=np.float32)
    ids = np.arange(len(data)) * 10 + 7
    index = make_index(kind, data, ids)
    i = np.arange(len(data))
    half = len(data) // 2
    index.add(data[:half], ids=ids[:half],
              metadata={"price": i[:half] % 100, "weight": i[:half] * 0.5, "color": COLORS[i[:half] % 4]})
    index.add(data[half:], ids=ids[half:], metadata={"price": i[half:] % 100})  # weight, color null
    assert index.metadata_schema == {"price": "int", "weight": "float", "color": "str"}
    f = qann_amd.field
    where = ((f("price") < 30) & ~(f("color") == "red")) | f("weight").isin([1.0, 2.5])
    match = ((i % 100 < 30) & (i < half) & (i % 4 != 0)) | ((i % 100 < 30) & (i >= half)) | np.isin(i * 0.5, [1.0, 2.5])
    queries = data[:40] + 0.01
    got, dists = index.batch_query(queries, 10, where=where)
    expected, expected_dists = index.batch_query(queries, 10, allowed=ids[match])
    np.testing.assert_array_equal(got, expected)
    np.testing.assert_array_equal(dists.view(np.uint32), expected_dists.view(np.uint32))
    one, _ = index.query(queries[0], 10, where=where)
    np.testing.assert_array_equal(one, got[0])
    # where and allowed together: only vectors passing both
    evens = ids[::2]
    got, _ = index.batch_query(queries, 10, where=where, allowed=evens)
    np.testing.assert_array_equal(got, index.batch_query(queries, 10, allowed=ids[match & (i % 2 == 0)])[0])
    # nulls: rows added without color
    got, _ = index.batch_query(queries, 5, where=f("color").is_null())
    assert np.isin(got[got >= 0], ids[half:]).all()
    # the same through SQL
    got_sql, _ = index.batch_query(queries, 10, where="(price < 30 AND color != 'red') OR weight IN (1.0, 2.5)")
    expected_sql = index.batch_query(queries, 10, allowed=ids[((i % 100 < 30) & (i < half) & (i % 4 != 0)) |
                                                              np.isin(i * 0.5, [1.0, 2.5])])[0]
    np.testing.assert_array_equal(got_sql, expected_sql)
    # deletions and a saved round trip
    index.remove(ids[:10])
    index.save(str(tmp_path / "md.qann"))
    loaded = qann_amd.load(str(tmp_path / "md.qann"))
    assert loaded.metadata_schema == index.metadata_schema
    np.testing.assert_array_equal(loaded.batch_query(queries, 10, where=where)[0],
                                  index.batch_query(queries, 10, where=where)[0])
    with pytest.raises(ValueError):
        index.batch_query(queries, 10, where="missing = 1")


def test_gpu_where_on_the_references_saved_metadata_index():
    # a RefineIndex over IVF-PQ saved by the reference (format 9) with int,
    # float and string columns, nulls from a second add, and deletions
    index = qann_amd.load(str(REF_DATA / "v9_refine_metadata.qann"))
    assert index.metadata_schema == {"price": "int", "weight": "float", "color": "str", "tag": "str"}
    f = qann_amd.field
    where = ((f("price") < 40) & (f("color") != "red")) | (f("tag") == "new")
    queries = np.load(REF_DATA / "v9_queries.npy")
    ids, dists = index.batch_query(queries, 10, where=where)
    saved_ids, saved_dists = np.load(REF_DATA / "v9_refine_metadata_ids.npy"), np.load(REF_DATA / "v9_refine_metadata_dists.npy")
    # the filter allows 340 vectors, which a RefineIndex over IVF scores
    # directly and exactly, so results can beat the probed ones recorded at
    # save time but never lose to them; ids in both rows score the same
    matching = index.query(queries[0], index.size(), where=where)[0]
    assert len(matching) == 340 and np.isin(ids, matching).all()
    assert (dists <= saved_dists + 1e-5).all()
    for row, saved_row, d, saved_d in zip(ids, saved_ids, dists, saved_dists):
        for a, b in zip(*np.nonzero(row[:, None] == saved_row[None, :])):
            np.testing.assert_allclose(d[a], saved_d[b], rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(np.sort(matching), index.ids(where=where))


def test_gpu_where_matches_the_oracle_and_serves(tmp_path, oracle):
    # an IVF fast-scan + refine index with metadata: where= searches equal the
    # oracle's searches over the matching ids, alone and through the server
    rng = np.random.default_rng(12)
    centres = rng.random((40, 48), dtype=np.float32)
    x = centres[rng.integers(0, 40, 20000)] + 0.06 * rng.standard_normal((20000, 48), dtype=np.float32)
    q = centres[rng.integers(0, 40, 64)] + 0.06 * rng.standard_normal((64, 48), dtype=np.float32)
    ivf = qann_amd.IVFIndex(48, 128, 12)
    ivf.enable_pq(24, fast_scan=True)
    ivf.train(x, seed=2)
    idx = qann_amd.RefineIndex(ivf, 6)
    i = np.arange(len(x))
    idx.add(x, metadata={"shard": i % 7, "score": (i % 1000) / 10.0, "lang": np.array(["en", "de", "fr"])[i % 3]})
    idx.remove(rng.choice(len(x), 300, replace=False))
    path = tmp_path / "meta.qann"
    idx.save(str(path))
    o = oracle.load(path)
    o.set_nprobe(12)
    filters = ["shard = 3", "lang = 'de' AND score < 20", "NOT shard IN (1, 2) AND score BETWEEN 10 AND 90",
               "shard = 1 OR lang IS NULL"]
    for sql in filters:
        want = o.batch_query(q, 10, allowed=idx.ids(where=sql))
        got = idx.batch_query(q, 10, where=sql)
        np.testing.assert_array_equal(got[0], want[0], err_msg=sql)
        np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32), err_msg=sql)
    server = qann_amd.Server(idx, workers=3)
    futs = [(sql, server.submit(q[s:s + 20], 10, nprobe=12, where=sql)) for s in (0, 20, 40) for sql in filters]
    for n, (sql, fut) in enumerate(futs):
        ids, d, _ = fut.result()
        s = 20 * (n // len(filters))
        want = o.batch_query(q[s:s + 20], 10, allowed=idx.ids(where=sql))
        np.testing.assert_array_equal(ids, want[0], err_msg=sql)
        np.testing.assert_array_equal(d.view(np.uint32), want[1].view(np.uint32), err_msg=sql)
    server.close()
