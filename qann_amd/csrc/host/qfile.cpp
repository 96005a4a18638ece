#include "qfile.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <unordered_set>

namespace qh {

namespace {

std::runtime_error corrupt(const std::string& what) { return std::runtime_error("corrupt file: " + what); }

class Reader {
public:
    explicit Reader(const std::string& path) : in_(path, std::ios::binary) {
        if (!in_) throw std::runtime_error("cannot open " + path + " for reading");
    }
    template <class T>
    T pod() {
        T v;
        in_.read(reinterpret_cast<char*>(&v), sizeof(T));
        if (!in_) throw std::runtime_error("truncated file");
        return v;
    }
    // read_vec (serialize.hpp:71-86): a u64 count, at most max_len, then the elements
    template <class T>
    vector<T> vec(uint64_t max_len) {
        const uint64_t n = pod<uint64_t>();
        if (n > max_len) throw corrupt("vector length exceeds expected size");
        vector<T> v(n);
        if (n) in_.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
        if (!in_) throw std::runtime_error("truncated file");
        return v;
    }
    void raw(void* dst, uint64_t bytes) {
        if (bytes) in_.read(static_cast<char*>(dst), static_cast<std::streamsize>(bytes));
        if (!in_) throw std::runtime_error("truncated file");
    }
    // read_pad (src/vector_store.cpp:83-88)
    void pad(uint64_t align) {
        const uint32_t p = pod<uint32_t>();
        if (p >= align) throw corrupt("invalid padding");
        in_.ignore(p);
        if (!in_ || in_.gcount() != static_cast<std::streamsize>(p)) throw std::runtime_error("truncated file");
    }
    bool at_end() { return in_.peek() == std::ifstream::traits_type::eof(); }

private:
    std::ifstream in_;
};

class Writer {
public:
    explicit Writer(const std::string& path) : out_(path, std::ios::binary | std::ios::trunc) {
        if (!out_) throw std::runtime_error("cannot open " + path + " for writing");
    }
    template <class T>
    void pod(const T& v) {
        out_.write(reinterpret_cast<const char*>(&v), sizeof(T));
    }
    template <class T>
    void vec(const vector<T>& v) {
        pod(static_cast<uint64_t>(v.size()));
        if (!v.empty()) out_.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
    }
    void raw(const void* p, uint64_t bytes) {
        if (bytes) out_.write(static_cast<const char*>(p), static_cast<std::streamsize>(bytes));
    }
    // write_pad (src/vector_store.cpp:67-81): a u32 pad length then zeros, so the
    // data after it starts at a multiple of align in the file
    void pad(uint64_t align) {
        const auto pos = static_cast<int64_t>(out_.tellp());
        const uint64_t after = pos < 0 ? 0 : static_cast<uint64_t>(pos) + sizeof(uint32_t);
        const auto p = static_cast<uint32_t>(pos < 0 ? 0 : (align - after % align) % align);
        pod(p);
        static const char zeros[4096] = {};
        for (uint32_t left = p; left > 0;) {
            const uint32_t c = std::min<uint32_t>(left, sizeof(zeros));
            out_.write(zeros, c);
            left -= c;
        }
    }
    void header(Kind kind) {
        pod(kMagic);
        pod(kFormatVersion);
        pod(static_cast<uint8_t>(kind));
    }
    void close() {
        out_.close();
        if (!out_) throw std::runtime_error("failed to write index file");
    }

private:
    std::ofstream out_;
};

void read_header(Reader& r, Kind& kind, uint32_t& version) {
    if (r.pod<uint32_t>() != kMagic) throw std::runtime_error("not a qann index file");
    version = r.pod<uint32_t>();
    if (version < 1 || version > kFormatVersion) throw std::runtime_error("unsupported format version");
    const uint8_t k = r.pod<uint8_t>();
    if (k == 4) throw std::runtime_error("HNSW indexes are not supported by qann_amd");
    if (k >= 5 && k <= 7) throw std::runtime_error("sparse and hybrid indexes are not supported by qann_amd");
    if (k == 8)
        throw std::runtime_error("this file holds a PQCodebook, not an index; load it with load_codebook "
                                 "(qann.PQCodebook.load)");
    if (k < 1 || k > 3) throw std::runtime_error("unknown index kind");
    kind = static_cast<Kind>(k);
}

Metric read_metric(Reader& r) {
    const uint8_t m = r.pod<uint8_t>();
    if (m > 2) throw corrupt("unknown metric");
    return static_cast<Metric>(m);
}

// IdMap::load (id_map.cpp:199-239)
IdMap read_ids(Reader& r, uint64_t slots, uint32_t version) {
    const uint8_t custom = r.pod<uint8_t>();
    const uint64_t slot_num = r.pod<uint64_t>();
    const int64_t next_id = version >= 3 ? r.pod<int64_t>() : static_cast<int64_t>(slot_num);
    vector<int64_t> labels = r.vec<int64_t>(slot_num);
    if (slots != slot_num || slots > 0xFFFFFFFFull) throw corrupt("id map slot count mismatch");
    if (custom > 1 || next_id < 0) throw corrupt("invalid id map shape");
    if (custom) {
        if (labels.size() != slots) throw corrupt("invalid id map shape");
    } else if (labels.empty()) {
        if (next_id != static_cast<int64_t>(slots) && slots != 0) throw corrupt("invalid id map shape");
    } else {
        if (version < 3 || labels.size() != slots || labels[0] < 0 || labels.back() >= next_id)
            throw corrupt("invalid id map shape");
        for (size_t i = 1; i < labels.size(); ++i)
            if (labels[i] <= labels[i - 1]) throw corrupt("auto ids out of order");
    }
    vector<uint8_t> deleted = r.vec<uint8_t>(slot_num);
    if (deleted.size() != slots) throw corrupt("invalid deleted flags length");
    IdMap m;
    m.custom = custom == 1;
    m.slots = slots;
    m.next_id = next_id;
    m.labels = std::move(labels);
    m.deleted = std::move(deleted);
    for (uint64_t s = 0; s < slots; ++s) {
        if (m.deleted[s] > 1) throw corrupt("invalid deleted flag value");
        if (m.deleted[s]) {
            ++m.deleted_count;
            continue;
        }
        if (m.custom && !m.slot_of.emplace(m.labels[s], static_cast<uint32_t>(s)).second)
            throw corrupt("duplicate live id");
    }
    return m;
}

IdMap plain_ids(uint64_t slots) {
    IdMap m;
    m.add(slots, nullptr);
    return m;
}

void write_ids(Writer& w, const IdMap& m) {
    w.pod(static_cast<uint8_t>(m.custom));
    w.pod(static_cast<uint64_t>(m.slots));
    w.pod(m.next_id);
    w.vec(m.labels);
    w.vec(m.deleted);
}

std::string read_string(Reader& r) {
    const vector<char> v = r.vec<char>(uint64_t{1} << 24);
    return {v.begin(), v.end()};
}

void write_string(Writer& w, const std::string& s) { w.vec(vector<char>(s.begin(), s.end())); }

// MetadataStore::load (src/metadata.cpp:1560-1610)
Metadata read_meta(Reader& r, uint64_t slots) {
    Metadata md;
    const uint64_t count = r.pod<uint64_t>();
    if (count > (uint64_t{1} << 16)) throw corrupt("too many metadata columns");
    std::unordered_set<std::string> names;
    for (uint64_t c = 0; c < count; ++c) {
        MetaColumn col;
        col.name = read_string(r);
        if (col.name.empty() || !names.insert(col.name).second) throw corrupt("invalid metadata column name");
        col.type = r.pod<uint8_t>();
        if (col.type > 2) throw corrupt("invalid metadata column type");
        col.valid = r.vec<uint8_t>(slots);
        if (col.valid.size() != slots) throw corrupt("metadata column length mismatch");
        for (uint8_t f : col.valid)
            if (f > 1) throw corrupt("invalid metadata null flag");
        if (col.type == 0) {
            col.ints = r.vec<int64_t>(slots);
            if (col.ints.size() != slots) throw corrupt("metadata column length mismatch");
        } else if (col.type == 1) {
            col.floats = r.vec<double>(slots);
            if (col.floats.size() != slots) throw corrupt("metadata column length mismatch");
        } else {
            col.codes = r.vec<uint32_t>(slots);
            if (col.codes.size() != slots) throw corrupt("metadata column length mismatch");
            const uint64_t dict = r.pod<uint64_t>();
            if (dict > 0xFFFFFFFFull) throw corrupt("metadata dictionary too large");
            std::unordered_set<std::string> seen;
            for (uint64_t d = 0; d < dict; ++d) {
                std::string s = read_string(r);
                if (!seen.insert(s).second) throw corrupt("duplicate metadata dictionary entry");
                col.dict.push_back(std::move(s));
            }
            for (uint64_t i = 0; i < slots; ++i)
                if (col.valid[i] && col.codes[i] >= col.dict.size()) throw corrupt("metadata string code out of range");
        }
        md.columns.push_back(std::move(col));
    }
    return md;
}

void write_meta(Writer& w, const Metadata& md) {
    w.pod(static_cast<uint64_t>(md.columns.size()));
    for (const MetaColumn& c : md.columns) {
        write_string(w, c.name);
        w.pod(c.type);
        w.vec(c.valid);
        if (c.type == 0) w.vec(c.ints);
        else if (c.type == 1) w.vec(c.floats);
        else {
            w.vec(c.codes);
            w.pod(static_cast<uint64_t>(c.dict.size()));
            for (const std::string& s : c.dict) write_string(w, s);
        }
    }
}

std::unique_ptr<FlatData> read_flat(Reader& r, uint32_t version) {
    auto x = std::make_unique<FlatData>();
    x->dim = r.pod<uint64_t>();
    const uint64_t count = r.pod<uint64_t>();
    if (x->dim == 0 || count > UINT64_MAX / x->dim) throw corrupt("invalid FlatIndex shape");
    x->rows = r.vec<float>(count * x->dim);
    if (x->rows.size() != count * x->dim) throw corrupt("FlatIndex data length mismatch");
    x->ids = version >= 2 ? read_ids(r, count, version) : plain_ids(count);
    if (version >= 5) x->metric = read_metric(r);
    if (version >= 9) x->meta = read_meta(r, count);
    return x;
}

// An orthogonal-matrix check as IVFIndex::load_body applies (detail::is_orthogonal:
// R^T R = I within 1e-3 per entry).
bool is_orthogonal(const vector<float>& R, size_t d) {
    for (size_t i = 0; i < d; ++i)
        for (size_t j = i; j < d; ++j) {
            double s = 0;
            for (size_t k = 0; k < d; ++k) s += double(R[k * d + i]) * double(R[k * d + j]);
            if (std::fabs(s - (i == j ? 1.0 : 0.0)) > 1e-3) return false;
        }
    return true;
}

std::unique_ptr<IvfData> read_ivf(Reader& r, uint32_t version) {
    auto x = std::make_unique<IvfData>();
    x->dim = r.pod<uint64_t>();
    x->nlist = r.pod<uint64_t>();
    x->nprobe = r.pod<uint64_t>();
    const uint64_t total = r.pod<uint64_t>();
    const uint8_t trained = r.pod<uint8_t>();
    x->pq_distance = r.pod<uint8_t>();
    if (x->dim == 0 || total > 0xFFFFFFFFull || trained > 1) throw corrupt("invalid IVFIndex shape");
    if (x->nlist < 100 || x->nlist > 65535) throw corrupt("nlist must be in [100, 65535]");
    if (x->dim > UINT64_MAX / x->nlist) throw corrupt("invalid IVFIndex shape");
    if (x->nprobe == 0 || x->nprobe > x->nlist) throw corrupt("nprobe must be in [1, nlist]");
    const uint64_t d = x->dim;
    x->centroids = r.vec<float>(x->nlist * d);
    if (x->pq_distance > 1) throw corrupt("invalid IVFIndex shape");
    if (!trained && total > 0) throw corrupt("invalid IVFIndex shape");
    if ((trained && x->centroids.size() != x->nlist * d) || (!trained && !x->centroids.empty()))
        throw corrupt("invalid IVFIndex shape");
    x->trained = trained != 0;
    if (r.pod<uint8_t>()) {
        x->has_pq = true;
        PQ& pq = x->pq;
        pq.dim = r.pod<uint64_t>();
        pq.m = r.pod<uint64_t>();
        pq.K = r.pod<uint64_t>();
        if (pq.dim == 0 || pq.m == 0) throw corrupt("dim and num_subspaces must be >= 1");
        if (pq.dim % pq.m != 0) throw corrupt("dim must be divisible by num_subspaces");
        if (pq.K == 0 || pq.K > 256) throw corrupt("centroids_per_subspace must be in [1, 256]");
        pq.sub = pq.dim / pq.m;
        const uint64_t expected = pq.m * pq.K * pq.sub;
        pq.centroids = r.vec<float>(expected);
        if (pq.centroids.size() != expected) throw corrupt("unexpected number of centroids");
        if (pq.dim != d) throw corrupt("invalid IVFIndex shape");
        if (version >= 4) {
            const uint8_t fs = r.pod<uint8_t>();
            if (fs > 1 || (fs && (pq.K != 16 || pq.m % 2 != 0 || pq.m > 1024 || x->pq_distance != 0)))
                throw corrupt("invalid IVFIndex fast-scan settings");
            x->fast_scan = fs != 0;
        }
        if (version >= 8) {
            const uint8_t rot = r.pod<uint8_t>();
            if (rot > 1 || d > 0xFFFFFFFFull) throw corrupt("invalid IVFIndex rotation");
            x->rotate = rot != 0;
            x->rotation = r.vec<float>(d * d);
            if (x->rotation.size() != (x->rotate && trained ? d * d : 0)) throw corrupt("IVFIndex rotation length mismatch");
            if (!x->rotation.empty() && !is_orthogonal(x->rotation, d)) throw corrupt("IVFIndex rotation is not orthogonal");
        }
    }
    if (version >= 7) {
        const uint8_t bits = r.pod<uint8_t>();
        if ((bits != 0 && bits != 4 && bits != 8) || (bits && x->has_pq)) throw corrupt("invalid IVFIndex SQ settings");
        x->sq_bits = bits;
        if (bits) {
            x->sq_vmin = r.vec<float>(d);
            x->sq_step = r.vec<float>(d);
            const size_t ranges = trained ? d : 0;
            if (x->sq_vmin.size() != ranges || x->sq_step.size() != ranges) throw corrupt("IVF SQ range length mismatch");
        }
    }
    const uint64_t sq_bytes = x->sq_bits == 8 ? d : (d + 1) / 2;
    uint64_t seen = 0;
    x->lists.resize(x->nlist);
    for (uint64_t l = 0; l < x->nlist; ++l) {
        IvfList& list = x->lists[l];
        list.ids = r.vec<uint32_t>(total);
        const uint64_t n = list.ids.size();
        if (!x->has_pq && version >= 11) {
            const uint64_t row_bytes = x->sq_bits ? sq_bytes : d * 4;
            const uint64_t expected = n * (x->sq_bits ? row_bytes : d);
            if (r.pod<uint64_t>() != expected) throw corrupt("IVF list length mismatch");
            r.pad(64);
            if (x->sq_bits) {
                list.codes.resize(n * row_bytes);
                r.raw(list.codes.data(), n * row_bytes);
            } else {
                list.vecs.resize(expected);
                r.raw(list.vecs.data(), n * row_bytes);
            }
        } else if (x->sq_bits) {
            list.codes = r.vec<uint8_t>(n * sq_bytes);
            if (list.codes.size() != n * sq_bytes) throw corrupt("IVF code length mismatch");
        } else if (x->has_pq) {
            const uint64_t blocks = (n + 31) / 32;
            const uint64_t expected = x->fast_scan ? blocks * x->pq.m * 16 : n * x->pq.m;
            list.codes = r.vec<uint8_t>(expected);
            if (list.codes.size() != expected) throw corrupt("IVF code length mismatch");
        } else {
            list.vecs = r.vec<float>(n * d);
            if (list.vecs.size() != n * d) throw corrupt("IVF vector length mismatch");
        }
        for (uint32_t id : list.ids)
            if (id >= total) throw corrupt("IVF id out of range");
        seen += n;
    }
    if (seen != total) throw corrupt("invalid IVFIndex shape");
    x->ids = version >= 2 ? read_ids(r, total, version) : plain_ids(total);
    if (version >= 3) {
        const uint8_t up = r.pod<uint8_t>();
        if (up > 1) throw corrupt("invalid IVFIndex shape");
        x->use_precomputed = up != 0;
    }
    if (version >= 5) x->metric = read_metric(r);
    if (version >= 9) x->meta = read_meta(r, total);
    if (version >= 12) {
        const uint8_t ci = r.pod<uint8_t>();
        if (ci > 1) throw corrupt("invalid IVFIndex shape");
        x->cosine_ip = ci != 0;
    }
    if (version >= 13) {
        x->residual_centers = r.vec<float>(x->nlist * d);
        const bool expected = x->trained && x->has_pq && x->metric == Cosine;
        if (x->residual_centers.size() != (expected ? x->nlist * d : 0))
            throw corrupt("IVFIndex residual centers length mismatch");
    }
    return x;
}

void write_ivf(Writer& w, const IvfData& x) {
    w.header(KIVF);
    w.pod(x.dim);
    w.pod(x.nlist);
    w.pod(x.nprobe);
    w.pod(static_cast<uint64_t>(x.ids.slots));
    w.pod(static_cast<uint8_t>(x.trained));
    w.pod(x.pq_distance);
    w.vec(x.centroids);
    w.pod(static_cast<uint8_t>(x.has_pq));
    if (x.has_pq) {
        w.pod(x.pq.dim);
        w.pod(x.pq.m);
        w.pod(x.pq.K);
        w.vec(x.pq.centroids);
        w.pod(static_cast<uint8_t>(x.fast_scan));
        w.pod(static_cast<uint8_t>(x.rotate));
        w.vec(x.rotation);
    }
    w.pod(static_cast<uint8_t>(x.sq_bits));
    if (x.sq_bits) {
        w.vec(x.sq_vmin);
        w.vec(x.sq_step);
    }
    const uint64_t sq_bytes = x.sq_bits == 8 ? x.dim : (x.dim + 1) / 2;
    for (uint64_t l = 0; l < x.nlist; ++l) {
        const IvfList& list = x.lists[l];
        if (x.has_pq) {
            w.vec(list.ids);
            w.vec(list.codes);
            continue;
        }
        w.vec(list.ids);
        const uint64_t per_row = x.sq_bits ? sq_bytes : x.dim;
        w.pod(static_cast<uint64_t>(list.ids.size() * per_row));
        w.pad(64);
        if (x.sq_bits) w.raw(list.codes.data(), list.ids.size() * sq_bytes);
        else w.raw(list.vecs.data(), list.ids.size() * x.dim * 4);
    }
    write_ids(w, x.ids);
    w.pod(static_cast<uint8_t>(x.use_precomputed));
    w.pod(static_cast<uint8_t>(x.metric));
    write_meta(w, x.meta);
    w.pod(static_cast<uint8_t>(x.cosine_ip));
    // Format 13 requires residual centres for a trained Cosine PQ index. A
    // file from before format 13 has none and takes residuals against the
    // coarse centroids (IVFIndex::residual_center, ivf_index.hpp:278-280), so
    // those centroids are written as its residual centres: every term the
    // search derives from them is the same.
    if (x.residual_centers.empty() && x.trained && x.has_pq && x.metric == Cosine) w.vec(x.centroids);
    else w.vec(x.residual_centers);
}

void write_flat(Writer& w, const FlatData& x) {
    w.header(KFlat);
    w.pod(x.dim);
    w.pod(static_cast<uint64_t>(x.ids.slots));
    w.vec(x.rows);
    write_ids(w, x.ids);
    w.pod(static_cast<uint8_t>(x.metric));
    write_meta(w, x.meta);
}

} // namespace

SlotFilter IdMap::filter(const int64_t* ids, size_t n) const {
    SlotFilter f;
    f.words.assign((slots + 63) / 64, 0);
    for (size_t i = 0; i < n; ++i) {
        const int64_t id = ids[i];
        uint32_t slot;
        if (custom) {
            auto it = slot_of.find(id);
            if (it == slot_of.end()) continue;
            slot = it->second;
        } else if (labels.empty()) {
            if (id < 0 || static_cast<uint64_t>(id) >= slots) continue;
            slot = static_cast<uint32_t>(id);
        } else {
            auto it = std::lower_bound(labels.begin(), labels.end(), id);
            if (it == labels.end() || *it != id) continue;
            slot = static_cast<uint32_t>(it - labels.begin());
        }
        if (deleted[slot]) continue;
        uint64_t& word = f.words[slot >> 6];
        const uint64_t bit = uint64_t{1} << (slot & 63);
        if (!(word & bit)) {
            word |= bit;
            ++f.count;
        }
    }
    return f;
}

void IdMap::check(size_t n, const int64_t* ids) const {
    if (n > 0xFFFFFFFFull - slots) throw std::invalid_argument("an index holds at most 2^32 - 1 vectors");
    if (custom && ids == nullptr) throw std::invalid_argument("ids are required: this index was created with custom ids");
    if (!custom && ids != nullptr)
        throw std::invalid_argument("ids are not accepted: this index was created without custom ids");
    if (!custom) return;
    std::unordered_set<int64_t> seen;
    for (size_t i = 0; i < n; ++i) {
        const int64_t id = ids[i];
        if (id < 0) throw std::invalid_argument("ids must be >= 0, got " + std::to_string(id));
        if (!seen.insert(id).second) throw std::invalid_argument("duplicate id " + std::to_string(id) + " in this batch");
        if (slot_of.count(id)) throw std::invalid_argument("id " + std::to_string(id) + " is already in the index");
    }
}

void IdMap::add(size_t n, const int64_t* ids) {
    check(n, ids);
    if (!custom) {
        if (labels.empty() && static_cast<uint64_t>(next_id) == slots) {
            deleted.resize(slots + n, 0);
            slots += n;
            next_id += static_cast<int64_t>(n);
        } else {
            for (size_t i = 0; i < n; ++i) {
                labels.push_back(next_id++);
                ++slots;
                deleted.push_back(0);
            }
        }
        return;
    }
    for (size_t i = 0; i < n; ++i) {
        labels.push_back(ids[i]);
        slot_of[ids[i]] = static_cast<uint32_t>(slots);
        deleted.push_back(0);
        ++slots;
    }
}

size_t IdMap::remove(const int64_t* ids, size_t n, vector<uint32_t>* removed) {
    size_t counter = 0;
    for (size_t i = 0; i < n; ++i) {
        const int64_t id = ids[i];
        uint32_t slot;
        if (custom) {
            auto it = slot_of.find(id);
            if (it == slot_of.end()) continue;
            slot = it->second;
            slot_of.erase(it);
        } else if (labels.empty()) {
            if (id < 0 || static_cast<uint64_t>(id) >= slots || deleted[id]) continue;
            slot = static_cast<uint32_t>(id);
        } else {
            auto it = std::lower_bound(labels.begin(), labels.end(), id);
            if (it == labels.end() || *it != id) continue;
            slot = static_cast<uint32_t>(it - labels.begin());
            if (deleted[slot]) continue;
        }
        deleted[slot] = 1;
        ++deleted_count;
        ++counter;
        if (removed) removed->push_back(slot);
    }
    return counter;
}

void Metadata::append_nulls(size_t n) {
    for (MetaColumn& c : columns) {
        c.valid.resize(c.valid.size() + n, 0);
        if (c.type == 0) c.ints.resize(c.ints.size() + n, 0);
        else if (c.type == 1) c.floats.resize(c.floats.size() + n, 0.0);
        else c.codes.resize(c.codes.size() + n, 0);
    }
}

LoadedIndex load_file(const std::string& path) {
    Reader r(path);
    LoadedIndex out;
    Kind kind;
    uint32_t version;
    read_header(r, kind, version);
    out.kind = kind;
    out.version = version;
    if (kind == KFlat) {
        out.flat = read_flat(r, version);
    } else if (kind == KIVF) {
        out.ivf = read_ivf(r, version);
    } else {
        // RefineIndex::load_body (src/refine_index.cpp:180-213)
        auto ref = std::make_unique<RefineData>();
        ref->k_factor = r.pod<uint64_t>();
        const uint64_t count = r.pod<uint64_t>();
        if (ref->k_factor == 0) throw corrupt("invalid RefineIndex shape");
        Kind bk;
        uint32_t bv;
        read_header(r, bk, bv);
        if (bv != version) throw corrupt("RefineIndex base has a different format version");
        if (bk == KFlat) out.flat = read_flat(r, bv);
        else if (bk == KIVF) out.ivf = read_ivf(r, bv);
        else throw corrupt("RefineIndex base must be Flat, IVF or HNSW");
        const IdMap& bids = out.flat ? out.flat->ids : out.ivf->ids;
        if (!bids.ids_are_slots()) throw corrupt("RefineIndex base ids must equal its slots");
        const uint64_t dim = out.flat ? out.flat->dim : out.ivf->dim;
        if (count != bids.slots || count > UINT64_MAX / dim) throw corrupt("invalid RefineIndex shape");
        // VectorStore::load (src/vector_store.cpp:164-189)
        const uint64_t len = r.pod<uint64_t>();
        if (len != count * dim) throw corrupt("unexpected number of vectors");
        if (version >= 10) r.pad(64 * 1024);
        ref->rows.resize(len);
        r.raw(ref->rows.data(), len * 4);
        ref->ids = version >= 2 ? read_ids(r, count, version) : plain_ids(count);
        if (version >= 9) ref->meta = read_meta(r, count);
        out.refine = std::move(ref);
    }
    if (!r.at_end()) throw corrupt("unexpected data after the index");
    return out;
}

void save_file(const std::string& path, const FlatData* flat, const IvfData* ivf, const RefineData* refine) {
    const std::string tmp = path + ".tmp";
    try {
        Writer w(tmp);
        if (refine) {
            w.header(KRefine);
            w.pod(refine->k_factor);
            w.pod(static_cast<uint64_t>(refine->ids.slots));
        }
        if (flat) write_flat(w, *flat);
        else write_ivf(w, *ivf);
        if (refine) {
            const uint64_t dim = flat ? flat->dim : ivf->dim;
            w.pod(static_cast<uint64_t>(refine->ids.slots * dim));
            w.pad(64 * 1024);
            w.raw(refine->rows.data(), refine->rows.size() * 4);
            write_ids(w, refine->ids);
            write_meta(w, refine->meta);
        }
        w.close();
        std::filesystem::rename(tmp, path);
    } catch (...) {
        std::error_code ignored;
        std::filesystem::remove(tmp, ignored);
        throw;
    }
}

vector<float> pq_transpose(const PQ& pq) {
    vector<float> t(pq.m * pq.K * pq.sub);
    for (uint64_t i = 0; i < pq.m; ++i)
        for (uint64_t j = 0; j < pq.sub; ++j)
            for (uint64_t k = 0; k < pq.K; ++k) t[(i * pq.sub + j) * pq.K + k] = pq.centroids[(i * pq.K + k) * pq.sub + j];
    return t;
}

bool wants_precomputed(const IvfData& x) {
    if (!x.use_precomputed || !x.trained || !x.has_pq || x.metric == InnerProduct || (x.metric == Cosine && x.cosine_ip))
        return false;
    return x.nlist * x.pq.m * x.pq.K * sizeof(float) <= (size_t{256} << 20);
}

} // namespace qh
