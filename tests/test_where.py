"""where= metadata filters on the host (no GPU): filters parse, print and fail
as the reference's do (the expectations of its tests/test_python.py:1246-1418),
and the rows a filter selects are the reference library's own, bit for bit,
over random SQL filters on an index saved here and loaded there."""
import numpy as np
import pytest

import qann_amd

COLORS = np.array(["red", "green", "blue", "black"])


def test_filters_print_as_the_reference_prints_them():
    f = qann_amd.field
    assert repr((f("n") >= 2) & ~f("s").isin(["a"])) == "qann.Expr(((n >= 2) & ~s in ['a']))"
    assert repr(f("price").between(1, 2.5)) == "qann.Expr(((price >= 1) & (price <= 2.5)))"
    assert repr(qann_amd.where("NOT (price < 5 OR color IS NULL)")) == "qann.Expr(((price >= 5) & ~is_null(color)))"
    assert repr(f("w") == 1.0) == "qann.Expr((w == 1.0))"
    assert repr(qann_amd.any_of([f("a") == 1, f("b") == "x"])) == "qann.Expr(((a == 1) | (b == 'x')))"
    assert repr(f("price")) == "qann.field('price')"


def test_filter_errors():
    f = qann_amd.field
    with pytest.raises(ValueError, match="unexpected .color. at column 11"):
        qann_amd.where("price < 5 color = 'red'")
    with pytest.raises(ValueError, match="no parameter named 'hi'"):
        qann_amd.where("price < :hi", {"lo": 1})
    with pytest.raises(ValueError, match="placeholders"):
        qann_amd.where("price < ?", [1, 2])
    with pytest.raises(ValueError, match="number out of range"):
        qann_amd.where("price < 1e400")
    with pytest.raises(ValueError, match="non-ASCII characters needs double quotes"):
        qann_amd.where("prïce < 5")
    with pytest.raises(ValueError, match="nested more than 256 levels"):
        qann_amd.where("(" * 100000 + "price = 1" + ")" * 100000)
    with pytest.raises(ValueError, match="nested more than 256 levels"):
        e = f("price") == 1
        for _ in range(300):
            e = ~e
    with pytest.raises(ValueError, match="more than 1024 conditions"):
        qann_amd.where(" OR ".join(f"price = {p}" for p in range(2000)))
    with pytest.raises(TypeError):
        qann_amd.where("price = ?", [None])
    for misuse in (lambda: bool(f("n") < 2), lambda: 0 < f("n") < 3, lambda: (f("n") < 2) and (f("n") > 0)):
        with pytest.raises(TypeError):
            misuse()
    assert qann_amd.get_max_filter_conditions() == 1024
    qann_amd.set_max_filter_conditions(1)
    try:
        qann_amd.where("price < 5")
        with pytest.raises(ValueError, match="more than 1 conditions"):
            qann_amd.where("price < 5 AND color = 'red'")
    finally:
        qann_amd.set_max_filter_conditions(1024)
    with pytest.raises(ValueError, match="at least 1"):
        qann_amd.set_max_filter_conditions(0)
    with pytest.raises(TypeError):
        qann_amd.set_max_filter_conditions(-1)


def metadata_flat(rng, n=3000, dim=4):
    """A FlatIndex (host only, never searched here) with int, float and string
    columns, nulls from a second add and from None values, and deletions."""
    idx = qann_amd.FlatIndex(dim, custom_ids=True)
    ids = rng.choice(10**9, n, replace=False).astype(np.int64)
    x = rng.random((n, dim), dtype=np.float32)
    i = np.arange(n)
    half = n // 2
    weight = (i[:half] * 0.5).astype(object)
    weight[::7] = None
    color = COLORS[i[:half] % 4].astype(object)
    color[::11] = None
    idx.add(x[:half], ids=ids[:half], metadata={"price": i[:half] % 100, "weight": list(weight), "color": list(color)})
    idx.add(x[half:], ids=ids[half:], metadata={"price": i[half:] % 100, "tag": np.array(["a", "b", "c"])[i[half:] % 3]})
    idx.remove(ids[rng.choice(n, n // 20, replace=False)])
    return idx, x


def random_condition(rng):
    kind = rng.integers(0, 12)
    p = int(rng.integers(-5, 105))
    w = float(rng.integers(0, 1600)) / 2
    c = COLORS[rng.integers(0, 4)]
    ops = ["=", "!=", "<>", "<", "<=", ">", ">="]
    op = ops[rng.integers(0, len(ops))]
    if kind == 0:
        return f"price {op} {p}"
    if kind == 1:
        return f"price {op} {p + 0.5}"
    if kind == 2:
        return f"weight {op} {w}"
    if kind == 3:
        return f"color {op} '{c}'"
    if kind == 4:
        vals = ", ".join(str(int(v)) for v in rng.integers(0, 100, rng.integers(1, 6)))
        return f"price {'NOT ' if rng.random() < 0.4 else ''}IN ({vals}, 2.0)"
    if kind == 5:
        vals = ", ".join(f"'{v}'" for v in COLORS[rng.integers(0, 4, rng.integers(1, 3))])
        return f"color {'NOT ' if rng.random() < 0.4 else ''}IN ({vals})"
    if kind == 6:
        return f"{['price', 'weight', 'color', 'tag'][rng.integers(0, 4)]} IS {'NOT ' if rng.random() < 0.5 else ''}NULL"
    if kind == 7:
        return f"price {'NOT ' if rng.random() < 0.4 else ''}BETWEEN {p} AND {p + int(rng.integers(-3, 30))}"
    if kind == 8:
        return f"weight {'NOT ' if rng.random() < 0.4 else ''}BETWEEN {w} AND {w + 50}"
    if kind == 9:
        return f"tag {op} '{'abc'[rng.integers(0, 3)]}'"
    if kind == 10:
        return f"weight IN ({w}, {w + 0.5}, 3)"
    return f"color {op} 'c'"


def random_sql(rng, depth=0):
    r = rng.random()
    if depth >= 3 or r < 0.35:
        return random_condition(rng)
    if r < 0.5:
        return f"NOT ({random_sql(rng, depth + 1)})"
    joiner = " AND " if rng.random() < 0.5 else " OR "
    return "(" + joiner.join(random_sql(rng, depth + 1) for _ in range(rng.integers(2, 4))) + ")"


def test_filters_select_the_reference_librarys_rows(tmp_path, reference):
    rng = np.random.default_rng(46)
    idx, x = metadata_flat(rng)
    path = tmp_path / "md.qann"
    idx.save(str(path))
    ref = reference.load(path)
    live = idx.size()
    q = x[:1]
    checked = 0
    for t in range(400):
        sql = random_sql(rng)
        ours = idx.ids(where=sql)
        got, _ = ref.batch_query(q, live, where=sql)
        theirs = np.sort(got[0][got[0] >= 0])
        np.testing.assert_array_equal(ours, theirs, err_msg=sql)
        checked += len(ours) > 0
    assert checked > 200
    # filters that fail when evaluated fail in both
    for sql in ("size = 1", "color = 3", "price = 'x'", "color IN ('a', 1)"):
        with pytest.raises(ValueError):
            idx.ids(where=sql)
        with pytest.raises((ValueError, RuntimeError)):
            ref.batch_query(q, 5, where=sql)


def test_filters_built_in_code_equal_their_sql(tmp_path):
    rng = np.random.default_rng(3)
    idx, _ = metadata_flat(rng, n=800)
    f = qann_amd.field
    pairs = [
        ("price < 30 AND color != 'red'", (f("price") < 30) & (f("color") != "red")),
        ("color IN ('red', 'blue') OR price BETWEEN 10 AND 20",
         f("color").isin(["red", "blue"]) | ((f("price") >= 10) & (f("price") <= 20))),
        (("price < ? AND color = ?", [30, "red"]), (f("price") < 30) & (f("color") == "red")),
        (("price >= :lo AND price < :hi", {"lo": 5, "hi": 9.5}), (f("price") >= 5) & (f("price") < 9.5)),
        (("price < ? AND color = ?", [np.int32(30), np.str_("red")]), "price < 30 AND color = 'red'"),
        (("price < ?", [np.float32(2.5)]), "price < 2.5"),
        (("price = ?", [True]), "price = 1"),
        (("price < ?", (30,)), "price < 30"),
        (qann_amd.where("price < ?", [30]) & (f("color") == "red"), (f("price") < 30) & (f("color") == "red")),
        (f("price").between(10, 20), "price BETWEEN 10 AND 20"),
        (f("price") < np.int64(30), "price < 30"),
    ]
    for a, b in pairs:
        np.testing.assert_array_equal(idx.ids(where=a), idx.ids(where=b), err_msg=str(a))
    # SQL's NOT excludes nulls, ~ includes them
    nulls = idx.ids(where="color IS NULL")
    assert not np.isin(nulls, idx.ids(where="NOT color = 'red'")).any()
    assert np.isin(nulls, idx.ids(where=~(f("color") == "red"))).all()
    # a parameter is one value, whatever it holds
    assert idx.ids(where=("color = ?", ["red' OR 'a' = 'a"])).size == 0
    for bad in (("price < 5",), ("price < ?", [5], 1), 5, b"price < 5"):
        with pytest.raises(TypeError):
            idx.ids(where=bad)


def test_add_metadata_checks_and_schema():
    rng = np.random.default_rng(4)
    data = rng.random((10, 32), dtype=np.float32)
    index = qann_amd.FlatIndex(32)
    index.add(data[:3], metadata={"n": [1, 2, 3], "x": np.array([0.5, 1.5, 2.5], dtype=np.float32),
                                  "s": np.array(["a", "b", "c"], dtype=object), "flag": np.array([True, False, True])})
    assert index.metadata_schema == {"n": "int", "x": "float", "s": "str", "flag": "int"}
    np.testing.assert_array_equal(index.ids(where=qann_amd.field("flag") == 1), [0, 2])
    assert index.ids(where=qann_amd.field("x") > np.float32(1.0)).size == 2
    with pytest.raises(ValueError):  # one value per row
        index.add(data[3:5], metadata={"n": [1]})
    with pytest.raises(ValueError):  # a column keeps its type
        index.add(data[3:4], metadata={"n": [1.5]})
    with pytest.raises(TypeError):  # mixed kinds
        index.add(data[3:5], metadata={"s": ["a", 1]})
    with pytest.raises(TypeError):  # not ints, floats or strings
        index.add(data[3:4], metadata={"d": np.array(["2026-01-01"], dtype="datetime64[D]")})
    with pytest.raises(ValueError):  # would wrap to a negative int64
        index.add(data[3:5], metadata={"u": np.array([1, 2**63], dtype=np.uint64)})
    assert index.size() == 3
    # ints go into a float column; an add of no rows creates no column
    index.add(data[3:4], metadata={"x": [7]})
    index.add(data[:0], metadata={"tag": []})
    assert "tag" not in index.metadata_schema
    index.add(data[4:5], metadata={"tag": ["x"]})
    assert index.metadata_schema["tag"] == "str"
    np.testing.assert_array_equal(index.ids(where="x = 7"), [3])
    np.testing.assert_array_equal(index.ids(where="tag IS NULL"), [0, 1, 2, 3])
    with pytest.raises(ValueError):
        index.ids(where="missing = 1")


def test_metadata_round_trips_through_a_saved_file(tmp_path):
    rng = np.random.default_rng(5)
    idx, _ = metadata_flat(rng, n=500)
    idx.save(str(tmp_path / "a.qann"))
    loaded = qann_amd.load(str(tmp_path / "a.qann"))
    assert loaded.metadata_schema == idx.metadata_schema
    for sql in ("price < 40 AND color != 'red'", "tag = 'b' OR weight IS NULL", "NOT color IN ('red', 'blue')"):
        np.testing.assert_array_equal(loaded.ids(where=sql), idx.ids(where=sql))
    loaded.save(str(tmp_path / "b.qann"))
    assert (tmp_path / "a.qann").read_bytes() == (tmp_path / "b.qann").read_bytes()
