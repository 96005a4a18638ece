// The qann_amd._qann extension: the reference's Python API for the dense
// indexes on this path (bindings/python_bindings.cpp:803-1199), rebuilt on
// pybind11 over the host index objects. Queries release the GIL and hold the
// index (and a RefineIndex's base) shared; writes hold them exclusively
// (read_locked / write_locked, python_bindings.cpp:97-131).
#include "indexes.hpp"
#include "serve.hpp"
#include "train.hpp"

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <optional>
#include <shared_mutex>
#include <unordered_map>

namespace py = pybind11;
using namespace qh;

namespace {

using FloatArray = py::array_t<float, py::array::c_style | py::array::forcecast>;

// vecengine::PQDistance, as IvfData::pq_distance stores it (0 ADC, 1 SDC)
enum class PQDistanceKind : uint8_t { ADC = 0, SDC = 1 };

// lock_chain (python_bindings.cpp:97-113): the index, then the base it wraps.
template <class Lock>
std::vector<Lock> lock_chain(const DenseIndex& idx) {
    std::vector<Lock> locks;
    locks.emplace_back(idx.mu);
    if (auto* r = dynamic_cast<const RefineIndex*>(&idx)) locks.emplace_back(r->base()->mu);
    return locks;
}

template <class F>
auto read_locked(const DenseIndex& idx, F&& f) {
    py::gil_scoped_release release;
    auto locks = lock_chain<std::shared_lock<std::shared_mutex>>(idx);
    return f();
}

template <class F>
auto write_locked(const DenseIndex& idx, F&& f) {
    py::gil_scoped_release release;
    auto locks = lock_chain<std::unique_lock<std::shared_mutex>>(idx);
    return f();
}

// as_ids (python_bindings.cpp:163-172): any integer sequence as contiguous int64.
py::array_t<int64_t> as_ids(py::handle obj) {
    py::module_ np = py::module_::import("numpy");
    py::object arr = np.attr("asarray")(obj);
    if (arr.attr("size").cast<size_t>() > 0) {
        const std::string kind = arr.attr("dtype").attr("kind").cast<std::string>();
        if (kind != "i" && kind != "u") throw py::type_error("ids must be integers");
    }
    if (arr.attr("ndim").cast<size_t>() != 1) throw py::type_error("ids must be one-dimensional");
    return np.attr("ascontiguousarray")(arr, py::arg("dtype") = "int64").cast<py::array_t<int64_t>>();
}

void refuse_if_wrapped(const DenseIndex& idx, const char* what) {
    if (idx.wrapped)
        throw std::invalid_argument(std::string("cannot ") + what +
                                    " on an index wrapped by a RefineIndex; use the RefineIndex instead");
}

// A Python value as a filter value (the reference's Expr::Value): a str (or
// NumPy str) as a string, a bool or an int (Python or NumPy) as int64, a float
// (Python or NumPy) as double; anything else, None included, is a TypeError.
FilterValue to_value(py::handle v) {
    if (py::isinstance<py::str>(v)) return v.cast<std::string>();
    py::module_ np = py::module_::import("numpy");
    if (py::isinstance<py::bool_>(v) || py::isinstance(v, np.attr("bool_")))
        return static_cast<int64_t>(v.cast<bool>() ? 1 : 0);
    if (py::isinstance<py::int_>(v) || py::isinstance(v, np.attr("integer")))
        return py::reinterpret_steal<py::object>(PyNumber_Index(v.ptr())).cast<int64_t>();
    if (py::isinstance<py::float_>(v) || py::isinstance(v, np.attr("floating")))
        return PyFloat_AsDouble(py::reinterpret_steal<py::object>(PyNumber_Float(v.ptr())).ptr());
    throw py::type_error("filter values must be ints, floats or strings");
}

std::vector<FilterValue> to_values(py::handle seq) {
    if (py::isinstance<py::str>(seq) || py::isinstance<py::bytes>(seq) || !py::isinstance<py::sequence>(seq))
        throw py::type_error("filter values must be a sequence of ints, floats or strings");
    std::vector<FilterValue> out;
    for (py::handle v : seq) out.push_back(to_value(v));
    return out;
}

// qann.where(sql, params): params a sequence for ? placeholders or a dict for
// :name ones (python_bindings.cpp:330-347).
Expr parse_sql(const std::string& sql, py::handle params) {
    if (params.is_none()) return Expr::parse(sql, nullptr, nullptr);
    if (py::isinstance<py::dict>(params)) {
        std::unordered_map<std::string, FilterValue> named;
        for (auto kv : py::reinterpret_borrow<py::dict>(params)) {
            if (!py::isinstance<py::str>(kv.first)) throw py::type_error("parameter names must be strings");
            named[kv.first.cast<std::string>()] = to_value(kv.second);
        }
        return Expr::parse(sql, nullptr, &named);
    }
    const std::vector<FilterValue> list = to_values(params);
    return Expr::parse(sql, &list, nullptr);
}

// where=: a qann_amd.Expr, a SQL string, or (SQL string, params)
std::optional<Expr> to_where(const py::object& where) {
    if (where.is_none()) return std::nullopt;
    if (py::isinstance<Expr>(where)) return where.cast<Expr>();
    if (py::isinstance<py::str>(where)) return parse_sql(where.cast<std::string>(), py::none());
    if (py::isinstance<py::tuple>(where)) {
        py::tuple t = py::reinterpret_borrow<py::tuple>(where);
        if (t.size() == 2 && py::isinstance<py::str>(t[0])) return parse_sql(t[0].cast<std::string>(), t[1]);
    }
    throw py::type_error("where must be a qann_amd.Expr, a SQL string or (sql, params)");
}

// One metadata column of n values (as_column, python_bindings.cpp:349-409):
// values go through numpy.asarray; integer and bool arrays make int columns,
// float arrays float columns, str arrays (or object arrays of str) string
// columns; None marks a null and the other values decide the type.
MetaInput to_column(const std::string& name, py::handle value, size_t n) {
    py::module_ np = py::module_::import("numpy");
    py::object arr = np.attr("asarray")(value);
    if (arr.attr("ndim").cast<size_t>() != 1)
        throw py::type_error("metadata column '" + name + "' must be one-dimensional");
    if (arr.attr("shape").cast<py::tuple>()[0].cast<size_t>() != n)
        throw std::invalid_argument("metadata column '" + name + "' must have one value per row of data");
    const std::string kind = arr.attr("dtype").attr("kind").cast<std::string>();
    // numpy.asarray(["a", 1]) makes "1" a string: a list's own items are checked
    py::object items = py::hasattr(value, "dtype") ? arr.attr("tolist")() : py::reinterpret_borrow<py::object>(value);
    MetaInput col;
    col.name = name;
    if (kind == "O") {
        std::vector<uint8_t> valid;
        py::list present;
        for (py::handle it : items) {
            valid.push_back(it.is_none() ? 0 : 1);
            if (!it.is_none()) present.append(it);
        }
        if (std::find(valid.begin(), valid.end(), 0) != valid.end()) {
            const size_t m = py::len(present);
            MetaInput some = m ? to_column(name, present, m) : MetaInput{name, 0, {}, {}, {}, {}};
            // the m values spread over the n rows; null rows keep a zero value
            MetaInput full;
            full.name = name;
            full.type = some.type;
            full.ints.resize(some.type == 0 ? n : 0);
            full.floats.resize(some.type == 1 ? n : 0);
            full.strings.resize(some.type == 2 ? n : 0);
            for (size_t i = 0, j = 0; i < n; ++i) {
                if (!valid[i]) continue;
                if (some.type == 0) full.ints[i] = some.ints[j];
                else if (some.type == 1) full.floats[i] = some.floats[j];
                else full.strings[i] = std::move(some.strings[j]);
                ++j;
            }
            full.valid = std::move(valid);
            return full;
        }
    }
    if (kind == "u" && arr.attr("size").cast<size_t>() > 0 &&
        arr.attr("max")().cast<uint64_t>() > static_cast<uint64_t>(INT64_MAX))
        throw std::invalid_argument("metadata column '" + name + "' has values above the int64 maximum");
    if (kind == "i" || kind == "u" || kind == "b") {
        auto a = np.attr("ascontiguousarray")(arr, py::arg("dtype") = "int64").cast<py::array_t<int64_t>>();
        col.type = 0;
        col.ints.assign(a.data(), a.data() + a.size());
        return col;
    }
    if (kind == "f") {
        auto a = np.attr("ascontiguousarray")(arr, py::arg("dtype") = "float64").cast<py::array_t<double>>();
        col.type = 1;
        col.floats.assign(a.data(), a.data() + a.size());
        return col;
    }
    if (kind == "U" || kind == "O") {
        col.type = 2;
        for (py::handle it : items) {
            if (!py::isinstance<py::str>(it))
                throw py::type_error("metadata column '" + name + "' mixes strings with other values");
            col.strings.push_back(it.cast<std::string>());
        }
        return col;
    }
    throw py::type_error("metadata column '" + name + "' must hold ints, floats or strings");
}

std::vector<MetaInput> to_metadata(const py::dict& md, size_t n) {
    std::vector<MetaInput> out;
    for (auto kv : md) {
        if (!py::isinstance<py::str>(kv.first)) throw py::type_error("metadata column names must be strings");
        out.push_back(to_column(kv.first.cast<std::string>(), kv.second, n));
    }
    return out;
}

py::tuple batch_query(DenseIndex& self, FloatArray q, uint32_t k, const py::object& allowed, const py::object& where_arg) {
    const std::optional<Expr> where = to_where(where_arg);
    if (q.ndim() != 2) throw py::type_error("query must be a 2-D array of shape (num_queries, dim)");
    if (static_cast<uint64_t>(q.shape(1)) != self.dim())
        throw std::invalid_argument("query dimension does not match index dimension");
    const uint64_t nq = static_cast<uint64_t>(q.shape(0));
    py::array_t<int64_t> ids({static_cast<py::ssize_t>(nq), static_cast<py::ssize_t>(k)});
    py::array_t<float> dists({static_cast<py::ssize_t>(nq), static_cast<py::ssize_t>(k)});
    py::array_t<int64_t> al;
    const bool filtered = !allowed.is_none();
    if (filtered) al = as_ids(allowed);
    const float* qp = q.data();
    int64_t* ip = ids.mutable_data();
    float* dp = dists.mutable_data();
    const int64_t* ap = filtered ? al.data() : nullptr;
    const size_t an = filtered ? static_cast<size_t>(al.shape(0)) : 0;
    const Expr* wp = where ? &*where : nullptr;
    read_locked(self, [&] {
        self.batch_query(qp, nq, k, ap, an, filtered, ip, dp, wp);
        return 0;
    });
    return py::make_tuple(ids, dists);
}

// Several batch_query calls answered in one launch: requests is a list of
// (nq, k, nprobe, allowed) over consecutive rows of queries (nprobe 0: the
// index's own; allowed: ids or None). Returns one (ids, distances) per request.
py::list search_requests(DenseIndex& self, FloatArray q, const py::list& requests) {
    if (q.ndim() != 2) throw py::type_error("queries must be a 2-D array of shape (num_queries, dim)");
    if (static_cast<uint64_t>(q.shape(1)) != self.dim())
        throw std::invalid_argument("query dimension does not match index dimension");
    const size_t nreq = requests.size();
    std::vector<DenseIndex::Request> reqs(nreq);
    std::vector<py::array_t<int64_t>> allowed(nreq);
    std::vector<bool> filtered(nreq, false);
    uint64_t rows = 0, entries = 0;
    for (size_t r = 0; r < nreq; ++r) {
        const py::tuple t = requests[r].cast<py::tuple>();
        if (t.size() != 4) throw py::type_error("each request is (nq, k, nprobe, allowed)");
        reqs[r].nq = t[0].cast<uint64_t>();
        reqs[r].k = t[1].cast<uint32_t>();
        reqs[r].nprobe = t[2].cast<uint32_t>();
        if (!t[3].is_none()) {
            allowed[r] = as_ids(t[3]);
            filtered[r] = true;
        }
        rows += reqs[r].nq;
        entries += reqs[r].nq * reqs[r].k;
    }
    if (rows != static_cast<uint64_t>(q.shape(0))) throw std::invalid_argument("the requests must cover every query row");
    std::vector<int64_t> ids(entries);
    std::vector<float> dists(entries);
    const float* qp = q.data();
    read_locked(self, [&] {
        std::vector<SlotFilter> filters(nreq);
        for (size_t r = 0; r < nreq; ++r)
            if (filtered[r]) {
                filters[r] = self.filter_of(allowed[r].data(), static_cast<size_t>(allowed[r].shape(0)), true, nullptr);
                reqs[r].filter = &filters[r];
            }
        const std::vector<float> prepared = self.prepare(qp, rows);
        self.search_requests(prepared.data(), reqs.data(), nreq, ids.data(), dists.data());
        return 0;
    });
    py::list out;
    uint64_t at = 0;
    for (size_t r = 0; r < nreq; ++r) {
        const auto n = static_cast<py::ssize_t>(reqs[r].nq), k = static_cast<py::ssize_t>(reqs[r].k);
        py::array_t<int64_t> ri({n, k});
        py::array_t<float> rd({n, k});
        std::copy_n(ids.data() + at, reqs[r].nq * reqs[r].k, ri.mutable_data());
        std::copy_n(dists.data() + at, reqs[r].nq * reqs[r].k, rd.mutable_data());
        at += reqs[r].nq * reqs[r].k;
        out.append(py::make_tuple(ri, rd));
    }
    return out;
}

py::tuple query(DenseIndex& self, FloatArray q, uint32_t k, const py::object& allowed, const py::object& where_arg) {
    const std::optional<Expr> where = to_where(where_arg);
    if (q.ndim() != 1) throw py::type_error("query must be a 1-D array of shape (dim,)");
    if (static_cast<uint64_t>(q.shape(0)) != self.dim())
        throw std::invalid_argument("query dimension does not match index dimension");
    std::vector<int64_t> ids(k);
    std::vector<float> dists(k);
    py::array_t<int64_t> al;
    const bool filtered = !allowed.is_none();
    if (filtered) al = as_ids(allowed);
    const float* qp = q.data();
    const int64_t* ap = filtered ? al.data() : nullptr;
    const size_t an = filtered ? static_cast<size_t>(al.shape(0)) : 0;
    const Expr* wp = where ? &*where : nullptr;
    read_locked(self, [&] {
        self.batch_query(qp, 1, k, ap, an, filtered, ids.data(), dists.data(), wp);
        return 0;
    });
    size_t n = 0;
    while (n < k && ids[n] >= 0) ++n;
    py::array_t<int64_t> oi(static_cast<py::ssize_t>(n));
    py::array_t<float> od(static_cast<py::ssize_t>(n));
    std::copy_n(ids.data(), n, oi.mutable_data());
    std::copy_n(dists.data(), n, od.mutable_data());
    return py::make_tuple(oi, od);
}

py::array_t<int64_t> add(DenseIndex& self, FloatArray data, const py::object& ids, const py::object& metadata) {
    if (data.ndim() != 2) throw py::type_error("data must be a 2-D array of shape (n, dim)");
    if (static_cast<uint64_t>(data.shape(1)) != self.dim())
        throw std::invalid_argument("data dimension does not match index dimension");
    refuse_if_wrapped(self, "add");
    py::array_t<int64_t> idv;
    const bool has_ids = !ids.is_none();
    if (has_ids) {
        idv = as_ids(ids);
        if (idv.shape(0) != data.shape(0)) throw std::invalid_argument("ids must hold one id per row");
    }
    const float* dp = data.data();
    const uint64_t n = static_cast<uint64_t>(data.shape(0));
    const int64_t* ip = has_ids ? idv.data() : nullptr;
    std::vector<MetaInput> md;
    const bool has_md = !metadata.is_none();
    if (has_md) {
        if (!py::isinstance<py::dict>(metadata)) throw py::type_error("metadata must be a dict of columns");
        md = to_metadata(py::reinterpret_borrow<py::dict>(metadata), n);
    }
    std::vector<int64_t> out = write_locked(self, [&] { return self.add(dp, n, ip, has_md ? &md : nullptr); });
    py::array_t<int64_t> res(static_cast<py::ssize_t>(out.size()));
    std::copy(out.begin(), out.end(), res.mutable_data());
    return res;
}

// A pending serving result: result() waits (without the GIL) and returns
// (ids, distances, latency_ms) with the request's shape.
struct ServeFuture {
    std::shared_ptr<std::future<ServeResult>> f;
    uint64_t nq;
    uint32_t k;
    py::tuple result() {
        ServeResult r;
        {
            py::gil_scoped_release release;
            r = f->get();
        }
        py::array_t<int64_t> ids({static_cast<py::ssize_t>(nq), static_cast<py::ssize_t>(k)});
        py::array_t<float> dists({static_cast<py::ssize_t>(nq), static_cast<py::ssize_t>(k)});
        std::copy(r.ids.begin(), r.ids.end(), ids.mutable_data());
        std::copy(r.dists.begin(), r.dists.end(), dists.mutable_data());
        return py::make_tuple(ids, dists, r.latency_ms);
    }
};

std::shared_ptr<const SlotFilter> slot_filter(const DenseIndex& idx, py::handle allowed,
                                              const py::object& where_arg = py::none()) {
    const std::optional<Expr> where = to_where(where_arg);
    const bool filtered = !allowed.is_none();
    py::array_t<int64_t> a;
    if (filtered) a = as_ids(allowed);
    auto f = std::make_shared<SlotFilter>();
    read_locked(idx, [&] {
        *f = idx.filter_of(filtered ? a.data() : nullptr, filtered ? static_cast<size_t>(a.size()) : 0, filtered,
                           where ? &*where : nullptr);
        return 0;
    });
    return f;
}

// qann.field(name): a column to compare; its operators build filters
struct Field {
    std::string name;
};

py::dict stats_dict(const ServeStats& s) {
    py::dict d;
    d["requests"] = s.requests;
    d["queries"] = s.queries;
    d["launches"] = s.launches;
    d["p50_ms"] = s.p50_ms;
    d["p90_ms"] = s.p90_ms;
    d["p99_ms"] = s.p99_ms;
    d["mean_ms"] = s.mean_ms;
    d["max_ms"] = s.max_ms;
    d["seconds"] = s.seconds;
    d["qps"] = s.seconds > 0 ? static_cast<double>(s.queries) / s.seconds : 0.0;
    return d;
}

} // namespace

PYBIND11_MODULE(_qann, m) {
    m.doc() = "qann_amd: IVF 4-bit PQ fast-scan search with exact re-ranking on AMD MI355X";

    py::enum_<Metric>(m, "Metric")
        .value("L2", Metric::L2)
        .value("Cosine", Metric::Cosine)
        .value("InnerProduct", Metric::InnerProduct);

    // vecengine::Clustering (include/vecengine/ivf_index.hpp:27-30)
    py::enum_<ClusteringKind>(m, "Clustering", "How IVFIndex.train learns the coarse centroids")
        .value("EMeans", ClusteringKind::EMeans, "e-means: mini-batch k-means with EMA centroids (default)")
        .value("Lloyd", ClusteringKind::Lloyd, "Lloyd's k-means");

    // detail::balanced_pca_rotation on a sample, as IVFIndex.train learns it
    m.def(
        "_balanced_pca_rotation",
        [](FloatArray data, uint64_t num_subspaces) {
            if (data.ndim() != 2) throw std::invalid_argument("data must be (n, dim)");
            const uint64_t n = static_cast<uint64_t>(data.shape(0)), d = static_cast<uint64_t>(data.shape(1));
            std::vector<float> r;
            {
                py::gil_scoped_release nogil;
                r = balanced_pca_rotation(data.data(), n, d, num_subspaces);
            }
            py::array_t<float> out({static_cast<py::ssize_t>(d), static_cast<py::ssize_t>(d)});
            std::copy(r.begin(), r.end(), out.mutable_data());
            return out;
        },
        py::arg("data"), py::arg("num_subspaces"));

    // vecengine::PQDistance (include/vecengine/pq.hpp:18)
    py::enum_<PQDistanceKind>(m, "PQDistance", "How IVF queries score PQ codes")
        .value("ADC", PQDistanceKind::ADC, "the query's own (residual) against each code's centroids (default)")
        .value("SDC", PQDistanceKind::SDC, "the query's residual encoded too, then code-to-code distances");

    m.def("set_num_threads", &set_num_threads, py::arg("n"), "Host threads for adds and training (0 = all cores)");
    m.def("get_num_threads", &num_threads);

    // Filters for where=: qann_amd.field("price") < 50, combined with & | ~
    py::class_<Expr>(m, "Expr", "A metadata filter for where=; build one with qann_amd.field() or qann_amd.where()")
        .def("__and__", [](const Expr& a, const Expr& b) { return Expr::join(Expr::Op::And, {a, b}); })
        .def("__or__", [](const Expr& a, const Expr& b) { return Expr::join(Expr::Op::Or, {a, b}); })
        .def("__invert__", [](const Expr& a) { return Expr::negate(a); })
        // and / or / not and chained comparisons would drop part of a filter: refused
        .def("__bool__",
             [](const Expr&) -> bool {
                 throw py::type_error("a filter has no truth value: combine filters with & | ~ (not and / or / "
                                      "not), and write ranges as f.between(0, 5) or (f > 0) & (f < 5)");
             })
        .def("__repr__", [](const Expr& e) { return "qann.Expr(" + e.to_string() + ")"; });

    auto cmp = [](Expr::Op op) {
        return [op](const Field& f, py::handle v) { return Expr::compare(f.name, op, to_value(v)); };
    };
    py::class_<Field>(m, "Field", "A metadata column to compare in a where= filter; see qann_amd.field()")
        .def_readonly("name", &Field::name)
        .def("__eq__", cmp(Expr::Op::Eq), py::arg("value"))
        .def("__ne__", cmp(Expr::Op::Ne), py::arg("value"))
        .def("__lt__", cmp(Expr::Op::Lt), py::arg("value"))
        .def("__le__", cmp(Expr::Op::Le), py::arg("value"))
        .def("__gt__", cmp(Expr::Op::Gt), py::arg("value"))
        .def("__ge__", cmp(Expr::Op::Ge), py::arg("value"))
        .def("isin", [](const Field& f, py::handle values) { return Expr::isin(f.name, to_values(values)); },
             py::arg("values"), "Matches rows whose value is one of values")
        .def("is_null", [](const Field& f) { return Expr::is_null(f.name); }, "Matches rows with no value in this column")
        .def("between",
             [](const Field& f, py::handle lo, py::handle hi) { return Expr::between(f.name, to_value(lo), to_value(hi)); },
             py::arg("lo"), py::arg("hi"), "Matches rows with lo <= value <= hi: (field >= lo) & (field <= hi)")
        .def("__repr__", [](const Field& f) { return "qann.field('" + f.name + "')"; });
    m.def("field", [](std::string name) { return Field{std::move(name)}; }, py::arg("name"),
          "A metadata column, for filters such as (qann_amd.field('price') < 50) & (qann_amd.field('color') == 'red')");
    m.def("where", [](const std::string& sql, const py::object& params) { return parse_sql(sql, params); },
          py::arg("sql"), py::arg("params") = py::none(),
          "A filter from a SQL WHERE clause: comparisons (= != <> < <= > >=), [NOT] IN (...), IS [NOT] NULL, "
          "[NOT] BETWEEN x AND y, AND, OR, NOT and parentheses; ? placeholders take a list of params, :name ones "
          "a dict. NOT follows SQL: NOT (price < 50) excludes rows with no price, unlike ~");
    m.def("any_of",
          [](std::vector<Expr> filters) {
              if (filters.empty()) throw std::invalid_argument("any_of needs at least one filter");
              return Expr::join(Expr::Op::Or, std::move(filters));
          },
          py::arg("filters"), "Matches rows matching any of filters, built in one step");
    m.def("all_of",
          [](std::vector<Expr> filters) {
              if (filters.empty()) throw std::invalid_argument("all_of needs at least one filter");
              return Expr::join(Expr::Op::And, std::move(filters));
          },
          py::arg("filters"), "Matches rows matching all of filters, built in one step");
    m.def("get_max_filter_conditions", &Expr::max_conditions,
          "The most conditions a SQL where= string may have (default 1024)");
    m.def("set_max_filter_conditions", &Expr::set_max_conditions, py::arg("n"),
          "Set the most conditions (comparisons, IN, IS NULL, BETWEEN; an IN list counts once) a SQL where= string "
          "may have, for the whole process");

    py::class_<DenseIndex, std::shared_ptr<DenseIndex>>(m, "DenseIndex", "An index of dense float vectors")
        .def("add", &add, py::arg("data"), py::arg("ids") = py::none(), py::kw_only(),
             py::arg("metadata") = py::none(), "Add an (n, dim) array of vectors; returns their ids")
        .def("query", &query, py::arg("query"), py::arg("k"), py::kw_only(), py::arg("allowed") = py::none(),
             py::arg("where") = py::none(), "Query k-NN for a (dim,) vector; returns 1D (indices, distances)")
        .def("batch_query", &batch_query, py::arg("query"), py::arg("k"), py::kw_only(),
             py::arg("allowed") = py::none(), py::arg("where") = py::none(),
             "Query k-NN for each row; returns (indices, distances) of shape (num_queries, k), padded with -1 / inf")
        .def("search_requests", &search_requests, py::arg("queries"), py::arg("requests"),
             "Several batch_query calls at once: requests is a list of (nq, k, nprobe, allowed) over consecutive "
             "rows of queries (nprobe 0: the index's own; allowed: ids or None). IVF and refine indexes answer "
             "them in one launch, each query with its own k, nprobe and filter. Returns one (indices, distances) "
             "per request, each as batch_query would return it")
        .def("remove",
             [](DenseIndex& self, const py::object& ids) {
                 refuse_if_wrapped(self, "remove");
                 py::array_t<int64_t> idv = as_ids(ids);
                 const int64_t* p = idv.data();
                 const size_t n = static_cast<size_t>(idv.shape(0));
                 return write_locked(self, [&] { return self.remove(p, n); });
             },
             py::arg("ids"), "Delete the vectors with these ids; returns how many were deleted")
        .def("save",
             [](const DenseIndex& self, const std::string& path) {
                 read_locked(self, [&] {
                     self.save(path);
                     return 0;
                 });
             },
             py::arg("path"), "Save the index in the qann file format")
        .def("size", [](const DenseIndex& self) { return read_locked(self, [&] { return self.size(); }); })
        .def("dim", &DenseIndex::dim)
        .def("ids",
             [](const DenseIndex& self, const py::object& where_arg) {
                 const std::optional<Expr> where = to_where(where_arg);
                 std::vector<int64_t> v = read_locked(self, [&] { return self.live_ids(where ? &*where : nullptr); });
                 py::array_t<int64_t> out(static_cast<py::ssize_t>(v.size()));
                 std::copy(v.begin(), v.end(), out.mutable_data());
                 return out;
             },
             py::kw_only(), py::arg("where") = py::none(),
             "The ids of the live vectors in ascending order; with where=, those whose metadata matches it")
        .def_property_readonly("metadata_schema",
                               [](const DenseIndex& self) {
                                   auto schema = read_locked(self, [&] { return meta_schema(self.meta()); });
                                   py::dict out;
                                   for (const auto& kv : schema) out[py::str(kv.first)] = kv.second;
                                   return out;
                               },
                               "Each metadata column's name and type ('int', 'float' or 'str'), in creation order")
        .def("get_vectors",
             [](const DenseIndex& self, const py::object& ids) {
                 py::array_t<int64_t> idv = as_ids(ids);
                 const int64_t* p = idv.data();
                 const size_t n = static_cast<size_t>(idv.shape(0));
                 std::vector<float> v = read_locked(self, [&] { return self.get_vectors(p, n); });
                 py::array_t<float> out({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(self.dim())});
                 std::copy(v.begin(), v.end(), out.mutable_data());
                 return out;
             },
             py::arg("ids"))
        .def("search_device",
             [](DenseIndex& self, uintptr_t q, uint64_t nq, uint32_t k, uintptr_t ids, uintptr_t dists,
                uintptr_t stream) {
                 read_locked(self, [&] {
                     self.search_device(reinterpret_cast<const float*>(q), nq, k, reinterpret_cast<int64_t*>(ids),
                                        reinterpret_cast<float*>(dists), reinterpret_cast<void*>(stream));
                     return 0;
                 });
             },
             py::arg("queries_ptr"), py::arg("nq"), py::arg("k"), py::arg("ids_ptr"), py::arg("dists_ptr"),
             py::arg("stream") = 0,
             "batch_query on prepared queries already in device memory (pointers as ints); results stay on the "
             "device, enqueued on the given HIP stream. Distances are the internal form (not halved for Cosine)")
        .def("device_handle",
             [](const DenseIndex& self) {
                 return read_locked(self, [&] { return reinterpret_cast<uintptr_t>(self.device_handle()); });
             },
             "The qann_hip_index* of the index's device copy as an int, for calling include/qann_hip.h "
             "directly; valid until the index is changed or freed")
        .def_property_readonly("metric", &DenseIndex::metric)
        .def_property_readonly("stores_vectors", &DenseIndex::stores_vectors);

    py::class_<FlatIndex, DenseIndex, std::shared_ptr<FlatIndex>>(m, "FlatIndex")
        .def(py::init([](uint64_t dim, bool custom_ids, Metric metric) {
                 return std::make_shared<FlatIndex>(dim, custom_ids, metric);
             }),
             py::arg("dim"), py::kw_only(), py::arg("custom_ids") = false, py::arg("metric") = Metric::L2);

    py::class_<IVFIndex, DenseIndex, std::shared_ptr<IVFIndex>>(m, "IVFIndex")
        .def(py::init([](uint64_t dim, uint64_t nlist, uint64_t nprobe, bool custom_ids, Metric metric) {
                 return std::make_shared<IVFIndex>(dim, nlist, nprobe, custom_ids, metric);
             }),
             py::arg("dim"), py::arg("nlist"), py::arg("nprobe") = 10, py::kw_only(), py::arg("custom_ids") = false,
             py::arg("metric") = Metric::L2)
        .def("enable_pq",
             [](IVFIndex& self, uint64_t m, const py::object& K, bool fast_scan, bool rotate) {
                 const uint64_t k = K.is_none() ? (fast_scan ? 16 : 256) : K.cast<uint64_t>();
                 write_locked(self, [&] {
                     self.enable_pq(m, k, fast_scan, rotate);
                     return 0;
                 });
             },
             py::arg("num_subspaces"), py::arg("centroids_per_subspace") = py::none(), py::kw_only(),
             py::arg("fast_scan") = false, py::arg("rotate") = false)
        .def("enable_sq",
             [](IVFIndex& self, int bits) {
                 write_locked(self, [&] {
                     self.enable_sq(bits);
                     return 0;
                 });
             },
             py::arg("bits") = 8)
        .def("train",
             [](IVFIndex& self, FloatArray data, uint64_t max_iters, const py::object& seed, ClusteringKind clustering,
                uint64_t passes) {
                 if (data.ndim() != 2) throw py::type_error("data must be a 2-D array of shape (n, dim)");
                 if (static_cast<uint64_t>(data.shape(1)) != self.dim())
                     throw std::invalid_argument("training vector dimension does not match index dimension");
                 std::optional<uint32_t> s;
                 if (!seed.is_none()) s = seed.cast<uint32_t>();
                 const float* p = data.data();
                 const uint64_t n = static_cast<uint64_t>(data.shape(0));
                 write_locked(self, [&] {
                     self.train(p, n, max_iters, s, clustering == ClusteringKind::EMeans, passes);
                     return 0;
                 });
             },
             py::arg("data"), py::arg("max_iters") = 10, py::arg("seed") = py::none(), py::kw_only(),
             py::arg("clustering") = ClusteringKind::EMeans, py::arg("passes") = 2,
             "Learn the coarse centroids (e-means in `passes` passes, or Lloyd's k-means for max_iters "
             "iterations), then the PQ codebook (25 Lloyd iterations per subspace) or SQ ranges, as qann "
             "trains them; every assign step runs on the GPU")
        .def_property("nprobe", [](const IVFIndex& self) { return self.nprobe(); },
                      [](IVFIndex& self, uint64_t v) { write_locked(self, [&] { self.set_nprobe(v); return 0; }); })
        .def("list_sizes",
             [](const IVFIndex& self) {
                 const auto& lists = self.data().lists;
                 py::array_t<int64_t> out(static_cast<py::ssize_t>(lists.size()));
                 for (size_t l = 0; l < lists.size(); ++l) out.mutable_data()[l] = static_cast<int64_t>(lists[l].ids.size());
                 return out;
             },
             "Entries stored in each inverted list")
        .def("centroids",
             [](const IVFIndex& self) {
                 const auto& c = self.data().centroids;
                 py::array_t<float> out({static_cast<py::ssize_t>(self.nlist()), static_cast<py::ssize_t>(self.dim())});
                 std::copy(c.begin(), c.end(), out.mutable_data());
                 return out;
             },
             "The coarse centroids, (nlist, dim)")
        .def("rotation",
             [](const IVFIndex& self) -> py::object {
                 const auto& r = self.data().rotation;
                 if (!self.pq_rotate() || r.empty()) return py::none();
                 py::array_t<float> out({static_cast<py::ssize_t>(self.dim()), static_cast<py::ssize_t>(self.dim())});
                 std::copy(r.begin(), r.end(), out.mutable_data());
                 return std::move(out);
             },
             "The learnt rotation R, (dim, dim), rows rotated as x R; None without one")
        .def_property_readonly("nlist", &IVFIndex::nlist)
        .def_property_readonly("fast_scan", &IVFIndex::fast_scan)
        .def_property_readonly("pq_rotate", &IVFIndex::pq_rotate)
        .def_property_readonly("sq_bits", &IVFIndex::sq_bits)
        .def_property_readonly("mapped", [](const IVFIndex&) { return false; })
        .def_property("precomputed_tables", &IVFIndex::precomputed_tables,
                      [](IVFIndex& self, bool v) { write_locked(self, [&] { self.set_precomputed_tables(v); return 0; }); })
        .def_property("cosine_inner_product", &IVFIndex::cosine_inner_product,
                      [](IVFIndex& self, bool v) { write_locked(self, [&] { self.set_cosine_inner_product(v); return 0; }); })
        .def_property(
            "pq_distance", [](const IVFIndex& self) { return static_cast<PQDistanceKind>(self.pq_distance()); },
            [](IVFIndex& self, PQDistanceKind v) {
                write_locked(self, [&] { self.set_pq_distance(static_cast<uint8_t>(v)); return 0; });
            },
            "How queries score PQ codes: PQDistance.ADC (default, more accurate) or PQDistance.SDC")
        .def_property("filter_probe_scaling", &IVFIndex::filter_probe_scaling,
                      [](IVFIndex& self, bool v) { write_locked(self, [&] { self.set_filter_probe_scaling(v); return 0; }); });

    py::class_<RefineIndex, DenseIndex, std::shared_ptr<RefineIndex>>(m, "RefineIndex")
        .def(py::init([](std::shared_ptr<DenseIndex> base, uint64_t k_factor, bool custom_ids) {
                 std::unique_lock<std::shared_mutex> g(base->mu);
                 return std::make_shared<RefineIndex>(base, k_factor, custom_ids);
             }),
             py::arg("base"), py::arg("k_factor") = 10, py::kw_only(), py::arg("custom_ids") = false)
        .def_property("k_factor", [](const RefineIndex& self) { return self.k_factor(); },
                      [](RefineIndex& self, uint64_t v) { write_locked(self, [&] { self.set_k_factor(v); return 0; }); })
        .def_property_readonly("base", &RefineIndex::base)
        .def_property_readonly("mapped", [](const RefineIndex&) { return false; });

    m.def("load",
          [](const std::string& path, bool mmap) {
              (void)mmap; // the device keeps its own copy in HBM either way
              py::gil_scoped_release release;
              return load_index(path);
          },
          py::arg("path"), py::arg("mmap") = false, "Load an index saved by qann or qann_amd");

    py::class_<ServeFuture>(m, "ServeFuture", "A pending serving result")
        .def("result", &ServeFuture::result, "Waits; returns (ids, distances, latency_ms)");

    py::class_<Server, std::shared_ptr<Server>>(m, "Server",
                                                "A serving loop: requests of any size, k, nprobe and filter, coalesced "
                                                "into launches and run concurrently within a device-memory budget")
        .def(py::init([](std::shared_ptr<DenseIndex> index, int workers, uint64_t max_batch, double memory_fraction) {
                 py::gil_scoped_release release;
                 return std::make_shared<Server>(std::move(index), workers, max_batch, memory_fraction);
             }),
             py::arg("index"), py::kw_only(), py::arg("workers") = 4, py::arg("max_batch") = 4096,
             py::arg("memory_fraction") = 0.5)
        .def("submit",
             [](Server& self, py::handle queries, uint32_t k, uint32_t nprobe, py::handle allowed,
                const py::object& where) {
                 FloatArray q = FloatArray::ensure(queries);
                 if (!q || q.ndim() != 2 || static_cast<uint64_t>(q.shape(1)) != self.index().dim())
                     throw py::value_error("queries must be a 2-D array with the index's dimension");
                 ServeRequest r;
                 r.nq = static_cast<uint64_t>(q.shape(0));
                 r.k = k;
                 r.nprobe = nprobe;
                 r.queries.assign(q.data(), q.data() + q.size());
                 if (!allowed.is_none() || !where.is_none()) r.filter = slot_filter(self.index(), allowed, where);
                 const uint64_t nq = r.nq;
                 auto f = std::make_shared<std::future<ServeResult>>(self.submit(std::move(r)));
                 return ServeFuture{f, nq, k};
             },
             py::arg("queries"), py::arg("k"), py::kw_only(), py::arg("nprobe") = 0, py::arg("allowed") = py::none(),
             py::arg("where") = py::none())
        .def("stats", [](Server& self, bool reset) { return stats_dict(self.stats(reset)); }, py::arg("reset") = false)
        .def_property_readonly("memory_budget", &Server::memory_budget)
        .def("close", [](Server& self) {
            py::gil_scoped_release release;
            self.close();
        });

    m.def("serve_load",
          [](Server& server, py::handle pool, double rate, double seconds, uint64_t seed, std::vector<uint32_t> ks,
             std::vector<uint32_t> nprobes, double filtered_fraction, py::list filters, uint32_t max_nq) {
              FloatArray p = FloatArray::ensure(pool);
              if (!p || p.ndim() != 2 || static_cast<uint64_t>(p.shape(1)) != server.index().dim())
                  throw py::value_error("pool must be a 2-D array with the index's dimension");
              LoadSpec spec;
              spec.rate = rate;
              spec.seconds = seconds;
              spec.seed = seed;
              spec.ks = std::move(ks);
              spec.nprobes = std::move(nprobes);
              spec.filtered_fraction = filtered_fraction;
              spec.max_nq = max_nq;
              for (py::handle f : filters) spec.filters.push_back(slot_filter(server.index(), f));
              ServeStats st;
              {
                  py::gil_scoped_release release;
                  st = run_load(server, p.data(), static_cast<uint64_t>(p.shape(0)), spec);
              }
              return stats_dict(st);
          },
          py::arg("server"), py::arg("pool"), py::arg("rate"), py::arg("seconds"), py::kw_only(), py::arg("seed") = 1,
          py::arg("ks") = std::vector<uint32_t>{10}, py::arg("nprobes") = std::vector<uint32_t>{0},
          py::arg("filtered_fraction") = 0.0, py::arg("filters") = py::list(), py::arg("max_nq") = 256,
          "An open-loop Poisson load of requests on a server; returns its latency percentiles and throughput");
}
