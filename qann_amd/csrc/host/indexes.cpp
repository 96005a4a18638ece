#include "indexes.hpp"
#include "parallel.hpp"
#include "train.hpp"

#include "../../../include/qann_hip.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <stdexcept>
#include <thread>
#include <unordered_set>

namespace qh {

namespace {

std::atomic<size_t> g_threads{0};

// IdMap::slots_of (src/id_map.cpp:93-112)
std::vector<uint32_t> slots_of(const IdMap& m, const int64_t* ids, size_t n) {
    std::vector<uint32_t> out;
    out.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const int64_t id = ids[i];
        uint32_t slot = 0;
        bool found;
        if (m.custom) {
            auto it = m.slot_of.find(id);
            found = it != m.slot_of.end();
            if (found) slot = it->second;
        } else if (m.labels.empty()) {
            found = id >= 0 && static_cast<uint64_t>(id) < m.slots;
            if (found) slot = static_cast<uint32_t>(id);
        } else {
            auto it = std::lower_bound(m.labels.begin(), m.labels.end(), id);
            found = it != m.labels.end() && *it == id;
            if (found) slot = static_cast<uint32_t>(it - m.labels.begin());
        }
        if (!found || m.deleted[slot]) throw std::invalid_argument("id " + std::to_string(id) + " is not in the index");
        out.push_back(slot);
    }
    std::vector<uint32_t> sorted = out;
    std::sort(sorted.begin(), sorted.end());
    auto dup = std::adjacent_find(sorted.begin(), sorted.end());
    if (dup != sorted.end()) throw std::invalid_argument("id " + std::to_string(m.label(*dup)) + " appears twice");
    return out;
}

// detail::rotate (src/rotation.cpp:165-185): an fma chain per output in k order.
std::vector<float> rotate_rows(const std::vector<float>& R, const float* in, uint64_t n, uint64_t d) {
    std::vector<float> out(n * d);
    parallel_for(n, [&](size_t i) {
        const float* x = in + i * d;
        for (uint64_t j = 0; j < d; ++j) {
            float acc = 0.0f;
            for (uint64_t k = 0; k < d; ++k) acc = std::fma(x[k], R[k * d + j], acc);
            out[i * d + j] = acc;
        }
    });
    return out;
}

// PQCodebook::encode with 16 centroids (src/pq.cpp:96-130): per subspace the
// squared distance to each centroid (the first term rounded, the rest fused),
// and the first index of the minimum.
void pq_encode(const std::vector<float>& cent_t, uint64_t m, uint64_t sub, uint64_t K, const float* r, uint8_t* out) {
    float dist[256];
    for (uint64_t s = 0; s < m; ++s) {
        const float* c = cent_t.data() + s * sub * K;
        for (uint64_t k = 0; k < K; ++k) {
            const float d0 = r[s * sub] - c[k];
            dist[k] = d0 * d0;
        }
        for (uint64_t j = 1; j < sub; ++j) {
            const float q = r[s * sub + j];
            for (uint64_t k = 0; k < K; ++k) {
                const float d = q - c[j * K + k];
                dist[k] = std::fma(d, d, dist[k]);
            }
        }
        uint64_t best = 0;
        for (uint64_t k = 1; k < K; ++k)
            if (dist[k] < dist[best]) best = k;
        out[s] = static_cast<uint8_t>(best);
    }
}

void assign_gpu(const float* rows, uint64_t n, uint64_t dim, const float* cents, uint64_t k, uint32_t* out) {
    check_status(qann_hip_assign_nearest(rows, n, static_cast<uint32_t>(dim), cents, static_cast<uint32_t>(k), out));
}

// The trainer's assign steps on the GPU (train.hpp).
void nearest_gpu(const float* rows, uint64_t n, uint64_t d, const float* cents, uint64_t k, uint32_t* out) {
    assign_gpu(rows, n, d, cents, k, out);
}

void nearest_subspaces_gpu(const float* rows, uint64_t n, uint64_t dim, uint64_t m, const float* cents, uint64_t K,
                           uint8_t* out) {
    check_status(qann_hip_assign_subspaces(rows, n, static_cast<uint32_t>(dim), static_cast<uint32_t>(m), cents,
                                           static_cast<uint32_t>(K), out));
}

} // namespace

void check_status(int rc) {
    if (rc == QANN_OK) return;
    const std::string msg = qann_hip_last_error();
    if (rc == QANN_EINVAL) throw std::invalid_argument(msg);
    if (rc == QANN_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

void set_num_threads(size_t n) { g_threads.store(n); }

size_t num_threads() {
    const size_t n = g_threads.load();
    return n ? n : std::max<size_t>(std::thread::hardware_concurrency(), 1);
}

// ---------------------------------------------------------------------------
// DenseIndex
// ---------------------------------------------------------------------------

DenseIndex::~DenseIndex() = default;

// DenseIndex::normalize_rows (src/index.cpp:193-218): the squared length summed
// in double, one float scale 1 / sqrt per row.
std::vector<float> DenseIndex::prepare(const float* q, uint64_t n) const {
    const uint64_t d = dim();
    std::vector<float> out(q, q + n * d);
    if (metric() != Cosine) return out;
    for (uint64_t i = 0; i < n; ++i) {
        double sq = 0;
        for (uint64_t j = 0; j < d; ++j) {
            const float c = q[i * d + j];
            sq += double(c) * c;
        }
        const double len = std::sqrt(sq);
        if (len == 0) throw std::invalid_argument("cosine metric needs non-zero vectors");
        const float scale = static_cast<float>(1.0 / len);
        for (uint64_t j = 0; j < d; ++j) out[i * d + j] = q[i * d + j] * scale;
    }
    return out;
}

SlotFilter DenseIndex::filter_of(const int64_t* allowed, size_t n_allowed, bool filtered, const Expr* where) const {
    if (!where) return ids().filter(allowed, n_allowed);
    const IdMap& m = ids();
    SlotFilter f;
    f.words = evaluate(meta(), *where, m.slots);
    if (filtered) {
        const SlotFilter also = m.filter(allowed, n_allowed);
        for (size_t w = 0; w < f.words.size(); ++w) f.words[w] &= also.words[w];
    }
    // IdMap::live: deleted slots never pass
    for (uint64_t s = 0; s < m.slots; ++s)
        if (m.deleted[s]) f.words[s >> 6] &= ~(uint64_t{1} << (s & 63));
    for (uint64_t w : f.words) f.count += static_cast<uint64_t>(__builtin_popcountll(w));
    return f;
}

void DenseIndex::batch_query(const float* queries, uint64_t nq, uint32_t k, const int64_t* allowed, size_t n_allowed,
                             bool filtered, int64_t* out_ids, float* out_dists, const Expr* where) {
    const size_t total = static_cast<size_t>(nq) * k;
    std::fill_n(out_ids, total, int64_t{-1});
    std::fill_n(out_dists, total, INFINITY);
    // an unknown column or a value of the wrong kind raises even without queries
    SlotFilter f;
    if (filtered || where) f = filter_of(allowed, n_allowed, filtered, where);
    if (nq == 0 || k == 0) return;
    if (filtered || where) {
        filtered = true;
        if (f.count == 0) return; // query_prepared: nothing may be returned
    }
    const std::vector<float> q = prepare(queries, nq);
    qann_hip_search_params p{};
    fill_params(&p);
    p.k = k;
    if (filtered) {
        p.allowed_words = f.words.data();
        p.allowed_count = f.count;
    }
    qann_hip_index* dev = device();
    check_status(qann_hip_search(dev, q.data(), nq, &p, out_ids, out_dists));
    // l2_to_cosine (src/index.cpp:23-25)
    if (metric() == Cosine)
        for (size_t i = 0; i < total; ++i) out_dists[i] *= 0.5f;
}

void DenseIndex::search_prepared(const float* prepared, uint64_t nq, uint32_t k, uint32_t nprobe,
                                 const SlotFilter* filter, int64_t* out_ids, float* out_dists) const {
    const size_t total = static_cast<size_t>(nq) * k;
    std::fill_n(out_ids, total, int64_t{-1});
    std::fill_n(out_dists, total, INFINITY);
    if (nq == 0 || k == 0 || (filter && filter->count == 0)) return;
    qann_hip_search_params p{};
    fill_params(&p);
    p.k = k;
    if (nprobe) p.nprobe = nprobe;
    if (filter) {
        p.allowed_words = filter->words.data();
        p.allowed_count = filter->count;
    }
    check_status(qann_hip_search(device(), prepared, nq, &p, out_ids, out_dists));
    if (metric() == Cosine)
        for (size_t i = 0; i < total; ++i) out_dists[i] *= 0.5f;
}

void DenseIndex::search_requests(const float* prepared, const Request* reqs, size_t nreq, int64_t* out_ids,
                                 float* out_dists) const {
    qann_hip_search_params base{};
    fill_params(&base);
    std::vector<qann_hip_request> r(nreq);
    size_t total = 0;
    for (size_t i = 0; i < nreq; ++i) {
        r[i].nq = reqs[i].nq;
        r[i].k = reqs[i].k;
        r[i].nprobe = reqs[i].nprobe ? reqs[i].nprobe : base.nprobe;
        if (reqs[i].filter) {
            r[i].allowed_words = reqs[i].filter->words.data();
            r[i].allowed_count = reqs[i].filter->count;
        }
        total += static_cast<size_t>(reqs[i].nq) * reqs[i].k;
    }
    std::fill_n(out_ids, total, int64_t{-1});
    std::fill_n(out_dists, total, INFINITY);
    check_status(qann_hip_search_requests(device(), prepared, r.data(), static_cast<uint32_t>(nreq), base.k_factor,
                                          base.filter_probe_scaling, out_ids, out_dists));
    if (metric() == Cosine)
        for (size_t i = 0; i < total; ++i) out_dists[i] *= 0.5f;
}

uint64_t DenseIndex::scratch_bytes(uint64_t nq, uint32_t k, uint32_t nprobe, bool filtered) const {
    qann_hip_search_params p{};
    fill_params(&p);
    p.k = k;
    if (nprobe) p.nprobe = nprobe;
    static const uint64_t word = 1;
    if (filtered) p.allowed_words = &word;
    uint64_t b = 0;
    check_status(qann_hip_search_scratch_bytes(device(), nq, &p, &b));
    return b;
}

void DenseIndex::search_device(const float* d_queries, uint64_t nq, uint32_t k, int64_t* d_ids, float* d_dists,
                               void* stream) {
    qann_hip_search_params p{};
    fill_params(&p);
    p.k = k;
    check_status(qann_hip_search_async(device(), d_queries, nq, &p, d_ids, d_dists, stream));
}

std::vector<int64_t> DenseIndex::add(const float* vecs, uint64_t n, const int64_t* idv,
                                     const std::vector<MetaInput>* md) {
    const uint64_t first = slots();
    if (md) check_meta(meta(), n, *md);
    const std::vector<float> prepared = prepare(vecs, n);
    add_prepared(prepared.data(), n, idv);
    if (md) append_meta(meta(), first, n, *md);
    invalidate();
    std::vector<int64_t> out(n);
    for (uint64_t i = 0; i < n; ++i) out[i] = ids().label(static_cast<uint32_t>(first + i));
    return out;
}

std::vector<int64_t> DenseIndex::live_ids(const Expr* where) const {
    const IdMap& m = ids();
    std::vector<uint64_t> match;
    if (where) match = evaluate(meta(), *where, m.slots);
    std::vector<int64_t> out;
    out.reserve(m.live());
    for (uint64_t s = 0; s < m.slots; ++s)
        if (!m.deleted[s] && (!where || (match[s >> 6] >> (s & 63) & 1))) out.push_back(m.label(static_cast<uint32_t>(s)));
    std::sort(out.begin(), out.end());
    return out;
}

std::vector<float> DenseIndex::get_vectors(const int64_t* idv, size_t n) const {
    if (!stores_vectors())
        throw std::invalid_argument("this index stores PQ or SQ codes, not the vectors; a RefineIndex over it keeps them");
    const std::vector<uint32_t> s = slots_of(ids(), idv, n);
    const uint64_t d = dim();
    std::vector<float> out(s.size() * d);
    for (size_t i = 0; i < s.size(); ++i) std::copy_n(stored_row(s[i]), d, out.data() + i * d);
    return out;
}

// ---------------------------------------------------------------------------
// FlatIndex (src/flat_index.cpp)
// ---------------------------------------------------------------------------

FlatIndex::FlatIndex(uint64_t dim, bool custom_ids, Metric metric) : d_(std::make_unique<FlatData>()) {
    if (dim == 0) throw std::invalid_argument("dim must be >= 1");
    d_->dim = dim;
    d_->metric = metric;
    d_->ids.custom = custom_ids;
}

FlatIndex::FlatIndex(std::unique_ptr<FlatData> data) : d_(std::move(data)) {}

FlatIndex::~FlatIndex() { invalidate(); }

void FlatIndex::add_prepared(const float* vecs, uint64_t n, const int64_t* idv) {
    d_->ids.check(n, idv);
    d_->rows.insert(d_->rows.end(), vecs, vecs + n * d_->dim);
    d_->ids.add(n, idv);
    d_->meta.append_nulls(n);
}

size_t FlatIndex::remove(const int64_t* idv, size_t n) {
    const size_t r = d_->ids.remove(idv, n, nullptr);
    invalidate();
    return r;
}

void FlatIndex::save(const std::string& path) const { save_file(path, d_.get(), nullptr, nullptr); }

qann_hip_index* FlatIndex::device() const {
    std::lock_guard<std::mutex> g(dev_mu);
    if (dev_) return dev_;
    qann_hip_index* h = nullptr;
    check_status(qann_hip_flat_create(static_cast<uint32_t>(d_->dim), d_->metric, d_->rows.data(), d_->ids.slots, &h));
    const int rc = qann_hip_set_ids(h, d_->ids.slots, d_->ids.labels.empty() ? nullptr : d_->ids.labels.data(),
                                    d_->ids.deleted.data());
    if (rc != QANN_OK) {
        qann_hip_index_destroy(h);
        check_status(rc);
    }
    dev_ = h;
    return dev_;
}

void FlatIndex::fill_params(void*) const {}

void FlatIndex::invalidate() const {
    std::lock_guard<std::mutex> g(dev_mu);
    if (dev_) qann_hip_index_destroy(dev_);
    dev_ = nullptr;
}

// ---------------------------------------------------------------------------
// IVFIndex (src/ivf_index.cpp)
// ---------------------------------------------------------------------------

IVFIndex::IVFIndex(uint64_t dim, uint64_t nlist, uint64_t nprobe, bool custom_ids, Metric metric)
    : d_(std::make_unique<IvfData>()) {
    if (nlist < 100 || nlist > 65535) throw std::invalid_argument("nlist must be in [100, 65535]");
    if (dim == 0) throw std::invalid_argument("dim must be >= 1");
    d_->dim = dim;
    d_->nlist = nlist;
    d_->metric = metric;
    d_->ids.custom = custom_ids;
    d_->lists.resize(nlist);
    set_nprobe(nprobe);
}

IVFIndex::IVFIndex(std::unique_ptr<IvfData> data) : d_(std::move(data)) {}

IVFIndex::~IVFIndex() { invalidate(); }

void IVFIndex::set_nprobe(uint64_t n) {
    if (n == 0 || n > d_->nlist) throw std::invalid_argument("nprobe must be in [1, nlist]");
    d_->nprobe = n;
}

void IVFIndex::set_precomputed_tables(bool on) {
    d_->use_precomputed = on;
    invalidate();
}

void IVFIndex::set_cosine_inner_product(bool on) {
    d_->cosine_ip = on;
    invalidate();
}

// IVFIndex::set_pq_distance (include/vecengine/ivf_index.hpp:94-100).
void IVFIndex::set_pq_distance(uint8_t mode) {
    if (mode > 1) throw std::invalid_argument("pq_distance must be ADC (0) or SDC (1)");
    if (d_->fast_scan && mode == 1) throw std::invalid_argument("cannot use SDC with fast-scan");
    if (d_->metric == InnerProduct && mode == 1) throw std::invalid_argument("cannot use SDC with Metric::InnerProduct");
    d_->pq_distance = mode;
    invalidate();
}

void IVFIndex::enable_pq(uint64_t m, uint64_t K, bool fast_scan, bool rotate) {
    if (d_->ids.slots > 0 || d_->trained) throw std::logic_error("enable_pq must be called before train()/add()");
    if (d_->sq_bits) throw std::logic_error("enable_pq cannot be combined with enable_sq");
    if (fast_scan && (K != 16 || m % 2 != 0 || m > 1024))
        throw std::invalid_argument("fast_scan needs 16 centroids per subspace and an even num_subspaces <= 1024");
    if (fast_scan && d_->pq_distance != 0) throw std::invalid_argument("cannot use SDC with fast-scan");
    if (d_->dim == 0 || m == 0) throw std::invalid_argument("dim and num_subspaces must be >= 1");
    if (d_->dim % m != 0) throw std::invalid_argument("dim must be divisible by num_subspaces");
    if (K == 0 || K > 256) throw std::invalid_argument("centroids_per_subspace must be in [1, 256]");
    d_->has_pq = true;
    d_->fast_scan = fast_scan;
    d_->rotate = rotate;
    d_->pq.dim = d_->dim;
    d_->pq.m = m;
    d_->pq.K = K;
    d_->pq.sub = d_->dim / m;
    d_->pq.centroids.assign(m * K * d_->pq.sub, 0.0f);
}

// IVFIndex::enable_sq (src/ivf_index.cpp:103-109).
void IVFIndex::enable_sq(int bits) {
    if (bits != 4 && bits != 8) throw std::invalid_argument("bits must be 4 or 8");
    if (d_->ids.slots > 0 || d_->trained) throw std::logic_error("enable_sq must be called before train()/add()");
    if (d_->has_pq) throw std::logic_error("enable_sq cannot be combined with enable_pq");
    d_->sq_bits = bits;
}

void IVFIndex::train(const float* vecs, uint64_t n, uint64_t iters, std::optional<uint32_t> seed, bool emeans_coarse,
                     uint64_t passes) {
    if (d_->ids.slots > 0) throw std::logic_error("train() cannot be called after add()");
    if (d_->trained) throw std::logic_error("train() has already been called");
    if (n < d_->nlist) throw std::invalid_argument("training set must have at least nlist vectors");
    const uint64_t d = d_->dim;
    // the clusters are learnt on the unit vectors that adds and queries use (Cosine),
    // rotated first by the learnt rotation (ivf_index.cpp:738-745)
    std::vector<float> x = prepare(vecs, n);
    if (d_->rotate) {
        d_->rotation = balanced_pca_rotation(x.data(), n, d, d_->pq.m);
        x = rotate_rows(d_->rotation, x.data(), n, d);
    }
    const bool spherical = d_->metric == Cosine;
    Clusters coarse = emeans_coarse ? emeans(x.data(), n, d, d_->nlist, passes, seed, spherical, nearest_gpu)
                                    : lloyd_kmeans(x.data(), n, d, d_->nlist, iters, seed, spherical, nearest_gpu);
    d_->centroids = std::move(coarse.centroids);
    const std::vector<uint32_t>& assign = coarse.assign;
    if (d_->has_pq && d_->metric == Cosine) {
        // each list's plain mean as its residual centre (ivf_index.cpp:769-785)
        d_->residual_centers.assign(d_->nlist * d, 0.0f);
        std::vector<uint64_t> cnt(d_->nlist, 0);
        for (uint64_t i = 0; i < n; ++i) {
            float* e = d_->residual_centers.data() + size_t{assign[i]} * d;
            for (uint64_t t = 0; t < d; ++t) e[t] += x[i * d + t];
            ++cnt[assign[i]];
        }
        for (uint64_t l = 0; l < d_->nlist; ++l) {
            float* e = d_->residual_centers.data() + l * d;
            if (cnt[l] == 0) std::copy_n(d_->centroids.data() + l * d, d, e);
            else
                for (uint64_t t = 0; t < d; ++t) e[t] /= static_cast<float>(cnt[l]);
        }
    }
    if (d_->has_pq) {
        // PQ codebooks on the residuals to each row's list centre (:787-802),
        // seeded with seed + 1
        PQ& pq = d_->pq;
        const float* rc = d_->residual_centers.empty() ? d_->centroids.data() : d_->residual_centers.data();
        std::vector<float> resid(n * d);
        parallel_for(n, [&](size_t i) {
            const float* c = rc + size_t{assign[i]} * d;
            for (uint64_t t = 0; t < d; ++t) resid[i * d + t] = x[i * d + t] - c[t];
        });
        pq.centroids = pq_codebooks(resid.data(), n, d, pq.m, pq.K, 25,
                                    seed ? std::optional<uint32_t>(*seed + 1) : std::nullopt, nearest_subspaces_gpu);
    }
    if (d_->has_pq && d_->metric == Cosine && d_->pq_distance == 0)
        d_->cosine_ip = cosine_tables_by_recall(x.data(), n, assign, seed ? *seed : 0u);
    if (d_->sq_bits) {
        // detail::SQ::train (src/sq.cpp): each dimension's range over the sample
        d_->sq_vmin.assign(x.begin(), x.begin() + d);
        std::vector<float> hi(d_->sq_vmin);
        for (uint64_t i = 1; i < n; ++i)
            for (uint64_t t = 0; t < d; ++t) {
                d_->sq_vmin[t] = std::min(d_->sq_vmin[t], x[i * d + t]);
                hi[t] = std::max(hi[t], x[i * d + t]);
            }
        const float top = static_cast<float>((1 << d_->sq_bits) - 1);
        d_->sq_step.resize(d);
        for (uint64_t t = 0; t < d; ++t) d_->sq_step[t] = (hi[t] - d_->sq_vmin[t]) / top;
    }
    d_->trained = true;
    invalidate();
}

// Fast-scan table rounding as the search uses it (fast_scan.cpp quantize_lut):
// one scale for the widest row, bytes rounded half to even from each row's
// minimum, offset the sum of the row minimums.
static void round_table(const float* table, uint64_t m, uint8_t* lut, float& scale, float& offset) {
    std::vector<float> lo(m);
    float widest = 0.0f, part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint64_t s = 0; s < m; ++s) {
        const float* row = table + s * 16;
        float a = row[0], b = row[0];
        for (int j = 1; j < 16; ++j) {
            a = std::min(a, row[j]);
            b = std::max(b, row[j]);
        }
        lo[s] = a;
        part[s % 4] += a;
        widest = std::max(widest, b - a);
    }
    offset = (part[0] + part[1]) + (part[2] + part[3]);
    scale = widest > 0.0f ? 63.0f / widest : 1.0f;
    for (uint64_t s = 0; s < m; ++s)
        for (int j = 0; j < 16; ++j) lut[s * 16 + j] = static_cast<uint8_t>(std::nearbyint((table[s * 16 + j] - lo[s]) * scale));
}

// The table mode of a Cosine PQ index, chosen by simulating its refined
// search on the training data: true keeps the query's one inner-product table,
// false per-list L2 tables of the residual q - c. 256 training rows (a seeded
// sample) query a pool of up to 32,768 other rows. Each query takes pool rows
// from its lists, nearest centroid first, until it holds 1/32 of the pool; each
// mode scores them as the search does (rounded byte tables for fast-scan,
// float tables otherwise) and shortlists its best 40, as a re-rank would; a
// mode earns a point for every row of the query's exact top 10 over the whole
// pool that its shortlist holds, and inner product wins ties. (The reference
// makes the same choice by its own test, IVFIndex::pick_cosine_inner_product,
// src/ivf_index.cpp:808-902; an index trained here is not the reference's
// either way, and the mode is saved with it.)
bool IVFIndex::cosine_tables_by_recall(const float* x, uint64_t n, const std::vector<uint32_t>& assign,
                                       uint32_t seed) const {
    constexpr uint64_t kQueries = 256, kPoolMax = 32768, kShort = 40, kTrue = 10;
    if (n < 4 * kQueries) return true;
    const uint64_t d = d_->dim, nlist = d_->nlist;
    const PQ& pq = d_->pq;
    const uint64_t m = pq.m, K = pq.K;
    const bool bytes = d_->fast_scan;
    const float* centres = d_->residual_centers.empty() ? d_->centroids.data() : d_->residual_centers.data();
    // queries first, then the pool: a partial shuffle of the rows
    std::vector<uint64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937_64 pick(0x9e3779b97f4a7c15ull ^ seed);
    const uint64_t pool_n = std::min<uint64_t>(kPoolMax, n - kQueries), taken = kQueries + pool_n;
    for (uint64_t i = 0; i < taken; ++i) std::swap(order[i], order[i + pick() % (n - i)]);
    const uint64_t* queries = order.data();
    const uint64_t* pool = order.data() + kQueries;

    auto ip = [d](const float* a, const float* b) {
        float acc = 0.0f;
        for (uint64_t t = 0; t < d; ++t) acc = std::fma(a[t], b[t], acc);
        return acc;
    };
    // the pool's PQ codes (residuals to their own list's centre) and the pool rows of every list
    const std::vector<float> cent_t = pq_transpose(pq);
    std::vector<uint8_t> code(pool_n * m);
    parallel_for(pool_n, [&](size_t j) {
        std::vector<float> r(d);
        const float* row = x + pool[j] * d;
        const float* c = centres + size_t{assign[pool[j]]} * d;
        for (uint64_t t = 0; t < d; ++t) r[t] = row[t] - c[t];
        pq_encode(cent_t, m, pq.sub, K, r.data(), code.data() + j * m);
    });
    std::vector<uint64_t> first(nlist + 1, 0);
    for (uint64_t j = 0; j < pool_n; ++j) ++first[assign[pool[j]] + 1];
    for (uint64_t l = 0; l < nlist; ++l) first[l + 1] += first[l];
    std::vector<uint32_t> members(pool_n);
    {
        std::vector<uint64_t> fill(first.begin(), first.end() - 1);
        for (uint64_t j = 0; j < pool_n; ++j) members[fill[assign[pool[j]]]++] = static_cast<uint32_t>(j);
    }
    // one subspace table (m x K) of a query vector: -2 q.y (inner product) or |r - y|^2
    auto subspace_table = [&](const float* v, bool inner, std::vector<float>& out) {
        for (uint64_t s = 0; s < m; ++s)
            for (uint64_t c = 0; c < K; ++c) {
                const float* y = pq.centroids.data() + (s * K + c) * pq.sub;
                const float* vs = v + s * pq.sub;
                float acc = 0.0f;
                for (uint64_t t = 0; t < pq.sub; ++t) acc += inner ? -2.0f * vs[t] * y[t] : (vs[t] - y[t]) * (vs[t] - y[t]);
                out[s * K + c] = acc;
            }
    };
    // a code's approximate score from a table: the rounded bytes' sum as the
    // fast-scan decodes it, or the float sum
    struct Scorer {
        std::vector<float> table;
        std::vector<uint8_t> lut;
        float scale = 1.0f, offset = 0.0f;
    };
    auto prepare_scorer = [&](Scorer& sc) {
        if (bytes) round_table(sc.table.data(), m, sc.lut.data(), sc.scale, sc.offset);
    };
    auto score = [&](const Scorer& sc, const uint8_t* c) {
        if (bytes) {
            uint32_t sum = 0;
            for (uint64_t s = 0; s < m; ++s) sum += sc.lut[s * 16 + c[s]];
            return static_cast<float>(sum) / sc.scale + sc.offset;
        }
        float sum = 0.0f;
        for (uint64_t s = 0; s < m; ++s) sum += sc.table[s * K + c[s]];
        return sum;
    };
    std::vector<uint32_t> points_inner(kQueries, 0), points_l2(kQueries, 0);
    parallel_for(kQueries, [&](size_t qi) {
        const float* q = x + queries[qi] * d;
        // the exact top 10 of the pool (rows are unit length: largest q.x first)
        std::vector<std::pair<float, uint32_t>> exact(pool_n);
        for (uint64_t j = 0; j < pool_n; ++j) exact[j] = {-ip(q, x + pool[j] * d), static_cast<uint32_t>(j)};
        const uint64_t want = std::min<uint64_t>(kTrue, pool_n);
        std::partial_sort(exact.begin(), exact.begin() + want, exact.end());
        std::vector<uint8_t> in_truth(pool_n, 0);
        for (uint64_t t = 0; t < want; ++t) in_truth[exact[t].second] = 1;
        // its lists, nearest centroid first, until 1/32 of the pool is in hand
        std::vector<std::pair<float, uint32_t>> near(nlist);
        for (uint64_t l = 0; l < nlist; ++l) near[l] = {-ip(q, d_->centroids.data() + l * d), static_cast<uint32_t>(l)};
        std::sort(near.begin(), near.end());
        Scorer inner{std::vector<float>(m * K), std::vector<uint8_t>(bytes ? m * 16 : 0)};
        Scorer l2{std::vector<float>(m * K), std::vector<uint8_t>(bytes ? m * 16 : 0)};
        subspace_table(q, true, inner.table);
        prepare_scorer(inner);
        std::vector<std::pair<float, uint32_t>> by_inner, by_l2;
        std::vector<float> r(d);
        for (uint64_t p = 0; p < nlist && by_inner.size() * 32 < pool_n; ++p) {
            const uint32_t l = near[p].second;
            if (first[l] == first[l + 1]) continue;
            const float* c = centres + size_t{l} * d;
            const float base = 2.0f - 2.0f * ip(q, c);
            for (uint64_t t = 0; t < d; ++t) r[t] = q[t] - c[t];
            subspace_table(r.data(), false, l2.table);
            prepare_scorer(l2);
            for (uint64_t e = first[l]; e < first[l + 1]; ++e) {
                const uint32_t j = members[e];
                const uint8_t* c8 = code.data() + size_t{j} * m;
                by_inner.emplace_back(base + score(inner, c8), j);
                by_l2.emplace_back(score(l2, c8), j);
            }
        }
        auto points = [&](std::vector<std::pair<float, uint32_t>>& cands) {
            const size_t keep = std::min<size_t>(kShort, cands.size());
            std::partial_sort(cands.begin(), cands.begin() + keep, cands.end());
            uint32_t got = 0;
            for (size_t t = 0; t < keep; ++t) got += in_truth[cands[t].second];
            return got;
        };
        points_inner[qi] = points(by_inner);
        points_l2[qi] = points(by_l2);
    });
    uint64_t inner_total = 0, l2_total = 0;
    for (uint64_t qi = 0; qi < kQueries; ++qi) {
        inner_total += points_inner[qi];
        l2_total += points_l2[qi];
    }
    return inner_total >= l2_total;
}

// detail::SQ::encode (src/sq.cpp): per dimension round((x - vmin) / step)
// (0 where step is 0), clamped to the code range; 4-bit codes two to a byte.
static void sq_encode(const IvfData& x, const float* v, uint8_t* out) {
    const uint64_t d = x.dim;
    const float top = static_cast<float>((1 << x.sq_bits) - 1);
    if (x.sq_bits == 4) std::fill_n(out, (d + 1) / 2, uint8_t{0});
    for (uint64_t t = 0; t < d; ++t) {
        const float c = x.sq_step[t] > 0 ? std::round((v[t] - x.sq_vmin[t]) / x.sq_step[t]) : 0.0f;
        const auto code = static_cast<uint8_t>(std::clamp(c, 0.0f, top));
        if (x.sq_bits == 8) out[t] = code;
        else out[t / 2] |= static_cast<uint8_t>(code << ((t & 1) * 4));
    }
}

// The device descriptor of an index's trained parts (no lists).
static qann_hip_ivf_desc ivf_desc(const IvfData& x) {
    qann_hip_ivf_desc desc{};
    desc.dim = static_cast<uint32_t>(x.dim);
    desc.nlist = static_cast<uint32_t>(x.nlist);
    desc.metric = x.metric;
    desc.cosine_ip = x.cosine_ip;
    desc.centroids = x.centroids.data();
    desc.residual_centers = x.residual_centers.empty() ? nullptr : x.residual_centers.data();
    desc.rotation = x.rotate && !x.rotation.empty() ? x.rotation.data() : nullptr;
    if (x.sq_bits) {
        desc.storage = QANN_IVF_SQ;
        desc.sq_bits = static_cast<uint32_t>(x.sq_bits);
        desc.sq_vmin = x.sq_vmin.data();
        desc.sq_step = x.sq_step.data();
    } else if (x.has_pq) {
        desc.storage = x.fast_scan ? QANN_IVF_FASTSCAN : QANN_IVF_PQ;
        desc.m = static_cast<uint32_t>(x.pq.m);
        desc.pq_k = static_cast<uint32_t>(x.pq.K);
        desc.pq_centroids = x.pq.centroids.data();
        desc.pq_sdc = x.pq_distance != 0 && x.metric != InnerProduct ? 1u : 0u;
    } else {
        desc.storage = QANN_IVF_FLOAT;
    }
    return desc;
}

// IVFIndex::add_batch_impl (src/ivf_index.cpp:111-195): rotation, the nearest
// list of each vector (on the GPU), its PQ code of the residual, SQ code or the
// raw row, and each list filled in input order.
void IVFIndex::add_prepared(const float* vecs, uint64_t n, const int64_t* idv) {
    if (!d_->trained) throw std::logic_error("train() has not be run yet");
    d_->ids.check(n, idv);
    const uint64_t d = d_->dim;
    std::vector<float> rot;
    if (d_->rotate) {
        rot = rotate_rows(d_->rotation, vecs, n, d);
        vecs = rot.data();
    }
    // nearest lists and codes on the GPU
    const PQ& pq = d_->pq;
    const uint64_t w = d_->sq_bits ? (d_->sq_bits == 8 ? d : (d + 1) / 2) : d_->has_pq ? pq.m : 0;
    std::vector<uint32_t> assign(n);
    std::vector<uint8_t> codes(n * w);
    const qann_hip_ivf_desc desc = ivf_desc(*d_);
    check_status(qann_hip_ivf_encode(&desc, vecs, n, assign.data(), codes.data()));

    // each list's new rows in input order
    const uint64_t nlist = d_->nlist;
    std::vector<uint64_t> start(nlist + 1, 0);
    for (uint64_t i = 0; i < n; ++i) ++start[assign[i] + 1];
    for (uint64_t l = 0; l < nlist; ++l) start[l + 1] += start[l];
    std::vector<uint32_t> order(n);
    {
        std::vector<uint64_t> cursor(start.begin(), start.end() - 1);
        for (uint64_t i = 0; i < n; ++i) order[cursor[assign[i]]++] = static_cast<uint32_t>(i);
    }
    const uint32_t first = static_cast<uint32_t>(d_->ids.slots);

    // fast-scan: the blocks each list's new rows touch, packed on the GPU, from
    // the block holding the list's next position on
    const bool fs = d_->has_pq && d_->fast_scan;
    const uint64_t bb = pq.m * 16;
    std::vector<uint64_t> block_at(nlist + 1, 0);
    std::vector<uint8_t> packed;
    if (fs) {
        for (uint64_t l = 0; l < nlist; ++l) {
            const uint64_t pos0 = d_->lists[l].ids.size(), cnt = start[l + 1] - start[l];
            block_at[l + 1] = block_at[l] + (cnt ? (pos0 + cnt - 1) / 32 - pos0 / 32 + 1 : 0);
        }
        const uint64_t nblocks = block_at[nlist];
        std::vector<uint32_t> row_at(nblocks * 32, 0xFFFFFFFFu);
        parallel_for(nlist, [&](size_t l) {
            const uint64_t pos0 = d_->lists[l].ids.size(), fb = pos0 / 32;
            for (uint64_t r = start[l]; r < start[l + 1]; ++r) {
                const uint64_t p = pos0 + (r - start[l]);
                row_at[(block_at[l] + p / 32 - fb) * 32 + p % 32] = order[r];
            }
        });
        packed.resize(nblocks * bb);
        check_status(qann_hip_fs_pack(codes.data(), n, static_cast<uint32_t>(pq.m), row_at.data(), nblocks,
                                      packed.data()));
    }

    parallel_for(nlist, [&](size_t l) {
        if (start[l + 1] == start[l]) return;
        IvfList& list = d_->lists[l];
        const uint64_t pos0 = list.ids.size(), cnt = start[l + 1] - start[l];
        for (uint64_t r = start[l]; r < start[l + 1]; ++r) list.ids.push_back(first + order[r]);
        if (fs) {
            // the first touched block may hold earlier rows: only the new
            // positions' nibbles are taken from the packed block (fs_set)
            const uint8_t* src = packed.data() + block_at[l] * bb;
            uint64_t nb = block_at[l + 1] - block_at[l];
            const uint64_t a = pos0 % 32;
            if (a != 0) {
                uint8_t* dst = list.codes.data() + (pos0 / 32) * bb;
                const uint64_t e = std::min<uint64_t>(32, a + cnt);
                for (uint64_t t = 0; t < bb; ++t) {
                    const uint64_t j = t % 16;
                    const uint8_t mask = static_cast<uint8_t>((j >= a && j < e ? 0x0F : 0) | (j + 16 >= a && j + 16 < e ? 0xF0 : 0));
                    dst[t] = static_cast<uint8_t>((dst[t] & ~mask) | (src[t] & mask));
                }
                src += bb;
                --nb;
            }
            list.codes.insert(list.codes.end(), src, src + nb * bb);
        } else if (w) {
            for (uint64_t r = start[l]; r < start[l + 1]; ++r)
                list.codes.insert(list.codes.end(), codes.begin() + size_t{order[r]} * w, codes.begin() + (size_t{order[r]} + 1) * w);
        } else {
            for (uint64_t r = start[l]; r < start[l + 1]; ++r)
                list.vecs.insert(list.vecs.end(), vecs + size_t{order[r]} * d, vecs + (size_t{order[r]} + 1) * d);
        }
    });
    d_->ids.add(n, idv);
    d_->meta.append_nulls(n);
}

size_t IVFIndex::remove(const int64_t* idv, size_t n) {
    const size_t r = d_->ids.remove(idv, n, nullptr);
    invalidate();
    return r;
}

void IVFIndex::save(const std::string& path) const { save_file(path, nullptr, d_.get(), nullptr); }

qann_hip_index* IVFIndex::device() const {
    std::lock_guard<std::mutex> g(dev_mu);
    if (dev_) return dev_;
    if (!d_->trained) throw std::logic_error("train() has not be run yet");
    qann_hip_ivf_desc desc = ivf_desc(*d_);
    desc.use_precomputed = d_->has_pq && !d_->sq_bits && !desc.pq_sdc ? wants_precomputed(*d_) : 0;
    qann_hip_index* h = nullptr;
    check_status(qann_hip_ivf_create(&desc, &h));
    std::vector<uint32_t> sizes(d_->nlist);
    uint64_t total = 0, cb = 0;
    const bool raw = desc.storage == QANN_IVF_FLOAT;
    for (uint64_t l = 0; l < d_->nlist; ++l) {
        sizes[l] = static_cast<uint32_t>(d_->lists[l].ids.size());
        total += sizes[l];
        cb += raw ? d_->lists[l].vecs.size() * sizeof(float) : d_->lists[l].codes.size();
    }
    std::vector<uint32_t> slots;
    std::vector<uint8_t> codes;
    slots.reserve(total);
    codes.reserve(cb);
    for (const IvfList& list : d_->lists) {
        slots.insert(slots.end(), list.ids.begin(), list.ids.end());
        if (raw) {
            const uint8_t* b = reinterpret_cast<const uint8_t*>(list.vecs.data());
            codes.insert(codes.end(), b, b + list.vecs.size() * sizeof(float));
        } else {
            codes.insert(codes.end(), list.codes.begin(), list.codes.end());
        }
    }
    int rc = qann_hip_ivf_set_lists(h, sizes.data(), slots.data(), codes.data());
    if (rc == QANN_OK)
        rc = qann_hip_set_ids(h, d_->ids.slots, d_->ids.labels.empty() ? nullptr : d_->ids.labels.data(),
                              d_->ids.deleted.data());
    if (rc != QANN_OK) {
        qann_hip_index_destroy(h);
        check_status(rc);
    }
    dev_ = h;
    return dev_;
}

void IVFIndex::fill_params(void* params) const {
    auto* p = static_cast<qann_hip_search_params*>(params);
    p->nprobe = static_cast<uint32_t>(d_->nprobe);
    p->filter_probe_scaling = filter_probe_scaling_;
}

void IVFIndex::invalidate() const {
    if (wrapper) wrapper->invalidate();
    std::lock_guard<std::mutex> g(dev_mu);
    if (dev_) qann_hip_index_destroy(dev_);
    dev_ = nullptr;
}

// ---------------------------------------------------------------------------
// RefineIndex (src/refine_index.cpp)
// ---------------------------------------------------------------------------

RefineIndex::RefineIndex(std::shared_ptr<DenseIndex> base, uint64_t k_factor, bool custom_ids)
    : base_(std::move(base)), d_(std::make_unique<RefineData>()) {
    if (base_->slots() != 0) throw std::invalid_argument("base index must be empty");
    if (k_factor == 0) throw std::invalid_argument("k_factor must be >= 1");
    if (!base_->ids().ids_are_slots())
        throw std::invalid_argument("base index must not use custom ids or have been compacted");
    if (base_->wrapped) throw std::invalid_argument("base index is already wrapped by another index");
    if (dynamic_cast<RefineIndex*>(base_.get()))
        throw std::invalid_argument("a RefineIndex's base must not be a RefineIndex; wrap the approximate index directly");
    d_->k_factor = k_factor;
    d_->ids.custom = custom_ids;
    base_->wrapped = true;
    base_->wrapper = this;
}

RefineIndex::RefineIndex(std::shared_ptr<DenseIndex> base, std::unique_ptr<RefineData> data)
    : base_(std::move(base)), d_(std::move(data)) {
    base_->wrapped = true;
    base_->wrapper = this;
}

RefineIndex::~RefineIndex() {
    invalidate();
    base_->wrapped = false;
    base_->wrapper = nullptr;
}

void RefineIndex::set_k_factor(uint64_t kf) {
    if (kf == 0) throw std::invalid_argument("k_factor must be >= 1");
    d_->k_factor = kf;
}

void RefineIndex::add_prepared(const float* vecs, uint64_t n, const int64_t* idv) {
    d_->ids.check(n, idv);
    base_->add_prepared(vecs, n, nullptr);
    base_->invalidate();
    d_->ids.add(n, idv);
    d_->rows.insert(d_->rows.end(), vecs, vecs + n * dim());
    d_->meta.append_nulls(n);
}

size_t RefineIndex::remove(const int64_t* idv, size_t n) {
    std::vector<uint32_t> slots;
    const size_t removed = d_->ids.remove(idv, n, &slots);
    const std::vector<int64_t> base_ids(slots.begin(), slots.end());
    base_->remove(base_ids.data(), base_ids.size());
    invalidate();
    return removed;
}

void RefineIndex::save(const std::string& path) const {
    const auto* flat = dynamic_cast<const FlatIndex*>(base_.get());
    const auto* ivf = dynamic_cast<const IVFIndex*>(base_.get());
    save_file(path, flat ? &flat->data() : nullptr, ivf ? &ivf->data() : nullptr, d_.get());
}

qann_hip_index* RefineIndex::device() const {
    std::lock_guard<std::mutex> g(dev_mu);
    if (dev_) return dev_;
    qann_hip_index* h = nullptr;
    if (auto* ivf = dynamic_cast<IVFIndex*>(base_.get())) {
        qann_hip_index* b = ivf->device();
        check_status(qann_hip_refine_create(b, d_->rows.data(), d_->ids.slots, &h));
    } else {
        // over a Flat base the candidates are already exact: the k nearest of
        // the raw rows by (distance, slot) are the re-ranked result
        check_status(qann_hip_flat_create(static_cast<uint32_t>(dim()), metric(), d_->rows.data(), d_->ids.slots, &h));
    }
    const int rc = qann_hip_set_ids(h, d_->ids.slots, d_->ids.labels.empty() ? nullptr : d_->ids.labels.data(),
                                    d_->ids.deleted.data());
    if (rc != QANN_OK) {
        qann_hip_index_destroy(h);
        check_status(rc);
    }
    dev_ = h;
    return dev_;
}

void RefineIndex::fill_params(void* params) const {
    base_->fill_params(params);
    static_cast<qann_hip_search_params*>(params)->k_factor = static_cast<uint32_t>(d_->k_factor);
}

void RefineIndex::invalidate() const {
    std::lock_guard<std::mutex> g(dev_mu);
    if (dev_) qann_hip_index_destroy(dev_);
    dev_ = nullptr;
}

// ---------------------------------------------------------------------------

std::shared_ptr<DenseIndex> load_index(const std::string& path) {
    LoadedIndex li = load_file(path);
    std::shared_ptr<DenseIndex> base;
    if (li.flat) base = std::make_shared<FlatIndex>(std::move(li.flat));
    else base = std::make_shared<IVFIndex>(std::move(li.ivf));
    if (!li.refine) return base;
    return std::make_shared<RefineIndex>(base, std::move(li.refine));
}

} // namespace qh
