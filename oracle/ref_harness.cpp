// Test-infrastructure harness over the reference library (maxtmiller/qann's
// vecengine core), compiled together with the reference's own sources by
// oracle/Makefile into oracle/_ref/libqann_ref.so.
//
// It is NOT part of the product. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it, to check this project's CPU oracle and
// HIP path against the reference itself and to time the reference on the host.
//
// Every entry point is a plain C function over pointers and sizes so Python can
// bind it with ctypes. Status: 0 ok, 1 invalid argument, 2 runtime error; the
// message of the last failure is returned by ref_last_error().

#include "vecengine/flat_index.hpp"
#include "vecengine/index.hpp"
#include "vecengine/index_factory.hpp"
#include "vecengine/ivf_index.hpp"
#include "vecengine/metadata.hpp"
#include "vecengine/refine_index.hpp"
#include "vecengine/threads.hpp"
#include "vecengine/distance.hpp"
#include "fast_scan.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace vecengine;

namespace {

thread_local std::string g_err;

// An index the harness owns: either loaded from a file (owned), or built here
// (a caller-owned IVF base kept alive next to the RefineIndex wrapping it).
struct Handle {
    std::unique_ptr<IndexBase> owned;
    std::unique_ptr<IVFIndex> base;      // built IVF under a RefineIndex
    std::unique_ptr<RefineIndex> refine; // built RefineIndex
    IndexBase* index() const {
        if (refine) return refine.get();
        if (owned) return owned.get();
        return base.get();
    }
    DenseIndex* dense() const { return dynamic_cast<DenseIndex*>(index()); }
};

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 2;
    }
}

IVFIndex* ivf_of(const Handle* h) {
    if (auto* r = dynamic_cast<RefineIndex*>(h->index())) return dynamic_cast<IVFIndex*>(&r->base());
    return dynamic_cast<IVFIndex*>(h->index());
}

void pad_results(const vector<vector<Neighbor>>& res, uint64_t k, int64_t* ids, float* dists) {
    for (size_t i = 0; i < res.size(); ++i) {
        for (uint64_t j = 0; j < k; ++j) {
            ids[i * k + j] = -1;
            dists[i * k + j] = INFINITY;
        }
        for (size_t j = 0; j < res[i].size() && j < k; ++j) {
            ids[i * k + j] = static_cast<int64_t>(res[i][j].index);
            dists[i * k + j] = res[i][j].distance;
        }
    }
}

} // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void ref_set_threads(uint64_t n) { set_num_threads(n); }

void* ref_load(const char* path) {
    Handle* h = nullptr;
    int rc = guarded([&] {
        auto idx = load_index(std::filesystem::path(path), false);
        h = new Handle();
        h->owned = std::move(idx);
    });
    return rc == 0 ? h : nullptr;
}

void ref_free(void* p) { delete static_cast<Handle*>(p); }

// kind: 1 flat, 2 ivf, 3 refine; info[0]=dim info[1]=size info[2]=slots
int ref_info(void* p, int64_t* info) {
    auto* h = static_cast<Handle*>(p);
    return guarded([&] {
        IndexBase* b = h->index();
        int kind = dynamic_cast<RefineIndex*>(b) ? 3 : dynamic_cast<IVFIndex*>(b) ? 2 : dynamic_cast<FlatIndex*>(b) ? 1 : 0;
        info[0] = kind;
        info[1] = static_cast<int64_t>(h->dense()->dim());
        info[2] = static_cast<int64_t>(b->size());
        info[3] = static_cast<int64_t>(b->slots());
        info[4] = static_cast<int64_t>(h->dense()->metric());
    });
}

int ref_set_nprobe(void* p, uint64_t nprobe) {
    return guarded([&] {
        IVFIndex* ivf = ivf_of(static_cast<Handle*>(p));
        if (!ivf) throw std::invalid_argument("not an IVF index");
        ivf->set_nprobe(nprobe);
    });
}

int ref_set_k_factor(void* p, uint64_t kf) {
    return guarded([&] {
        auto* r = dynamic_cast<RefineIndex*>(static_cast<Handle*>(p)->index());
        if (!r) throw std::invalid_argument("not a RefineIndex");
        r->set_k_factor(kf);
    });
}

int ref_set_precomputed(void* p, int on) {
    return guarded([&] {
        IVFIndex* ivf = ivf_of(static_cast<Handle*>(p));
        if (!ivf) throw std::invalid_argument("not an IVF index");
        ivf->set_precomputed_tables(on != 0);
    });
}

int ref_set_cosine_ip(void* p, int on) {
    return guarded([&] {
        IVFIndex* ivf = ivf_of(static_cast<Handle*>(p));
        if (!ivf) throw std::invalid_argument("not an IVF index");
        ivf->set_cosine_inner_product(on != 0);
    });
}

int ref_set_pq_distance(void* p, int sdc) {
    return guarded([&] {
        IVFIndex* ivf = ivf_of(static_cast<Handle*>(p));
        if (!ivf) throw std::invalid_argument("not an IVF index");
        ivf->set_pq_distance(sdc ? PQDistance::SDC : PQDistance::ADC);
    });
}

int ref_set_filter_probe_scaling(void* p, int on) {
    return guarded([&] {
        IVFIndex* ivf = ivf_of(static_cast<Handle*>(p));
        if (!ivf) throw std::invalid_argument("not an IVF index");
        ivf->set_filter_probe_scaling(on != 0);
    });
}

// batch_query as the Python binding does it: padded (nq, k) results. allowed
// may be null (no filter); where may be null or an SQL WHERE string.
int ref_batch_query(void* p, const float* q, uint64_t nq, uint64_t k, const int64_t* allowed, int64_t nallowed,
                    const char* where, int64_t* ids, float* dists) {
    auto* h = static_cast<Handle*>(p);
    return guarded([&] {
        DenseIndex* d = h->dense();
        std::span<const float> qs(q, nq * d->dim());
        vector<vector<Neighbor>> res;
        if (where) {
            Expr e = Expr::parse(where);
            if (allowed) res = d->query_batch(qs, nq, k, e, std::span<const int64_t>(allowed, static_cast<size_t>(nallowed)));
            else res = d->query_batch(qs, nq, k, e);
        } else if (allowed) {
            res = d->query_batch(qs, nq, k, std::span<const int64_t>(allowed, static_cast<size_t>(nallowed)));
        } else {
            res = d->query_batch(qs, nq, k);
        }
        pad_results(res, k, ids, dists);
    });
}

// query() for one vector: writes up to k results and their count.
int ref_query(void* p, const float* q, uint64_t k, const int64_t* allowed, int64_t nallowed, int64_t* ids, float* dists,
              uint64_t* count) {
    auto* h = static_cast<Handle*>(p);
    return guarded([&] {
        DenseIndex* d = h->dense();
        std::span<const float> qs(q, d->dim());
        vector<Neighbor> res = allowed ? d->query(qs, k, std::span<const int64_t>(allowed, static_cast<size_t>(nallowed)))
                                       : d->query(qs, k);
        *count = res.size();
        for (size_t j = 0; j < res.size(); ++j) {
            ids[j] = static_cast<int64_t>(res[j].index);
            dists[j] = res[j].distance;
        }
    });
}

// Builds (trains and fills) an IVF fast-scan or PQ index, optionally wrapped in
// a RefineIndex (k_factor > 0). metric: 0 L2, 1 Cosine, 2 InnerProduct.
// pq_k: centroids per subspace (16 for fast-scan). sq_bits: 0, 4 or 8 (then m = 0).
void* ref_build_ivf(const float* data, uint64_t n, uint64_t dim, uint64_t nlist, uint64_t nprobe, int metric,
                    uint64_t m, uint64_t pq_k, int fast_scan, int rotate, int sq_bits, uint32_t seed, uint64_t k_factor,
                    const int64_t* custom_ids, int clustering, uint64_t iters_or_passes, uint64_t train_n) {
    Handle* h = nullptr;
    int rc = guarded([&] {
        auto hh = std::make_unique<Handle>();
        const bool custom = custom_ids != nullptr;
        hh->base = std::make_unique<IVFIndex>(dim, nlist, nprobe, k_factor == 0 && custom, static_cast<Metric>(metric));
        if (m > 0) hh->base->enable_pq(m, pq_k, fast_scan != 0, rotate != 0);
        if (sq_bits) hh->base->enable_sq(sq_bits);
        const uint64_t tn = train_n ? train_n : n;
        if (clustering == 1)
            hh->base->train(std::span<const float>(data, tn * dim), tn, iters_or_passes, seed, Clustering::Lloyd, 2);
        else
            hh->base->train(std::span<const float>(data, tn * dim), tn, 10, seed, Clustering::EMeans, iters_or_passes);
        if (k_factor > 0) {
            hh->refine = std::make_unique<RefineIndex>(*hh->base, k_factor, custom);
            hh->refine->add_batch(std::span<const float>(data, n * dim), n, custom_ids);
        } else {
            hh->base->add_batch(std::span<const float>(data, n * dim), n, custom_ids);
        }
        h = hh.release();
    });
    return rc == 0 ? h : nullptr;
}

void* ref_build_flat(const float* data, uint64_t n, uint64_t dim, int metric, const int64_t* custom_ids) {
    Handle* h = nullptr;
    int rc = guarded([&] {
        auto hh = std::make_unique<Handle>();
        auto flat = std::make_unique<FlatIndex>(dim, n, custom_ids != nullptr, static_cast<Metric>(metric));
        flat->add_batch(std::span<const float>(data, n * dim), n, custom_ids);
        hh->owned = std::move(flat);
        h = hh.release();
    });
    return rc == 0 ? h : nullptr;
}

int ref_add(void* p, const float* data, uint64_t n, const int64_t* custom_ids) {
    auto* h = static_cast<Handle*>(p);
    return guarded([&] { h->dense()->add_batch(std::span<const float>(data, n * h->dense()->dim()), n, custom_ids); });
}

int ref_remove(void* p, const int64_t* ids, uint64_t n, uint64_t* removed) {
    auto* h = static_cast<Handle*>(p);
    return guarded([&] { *removed = h->index()->remove(std::span<const int64_t>(ids, n)); });
}

int ref_save(void* p, const char* path) {
    auto* h = static_cast<Handle*>(p);
    return guarded([&] { save_index(*h->index(), std::filesystem::path(path)); });
}

// --- the reference's kernels on their own, for known-answer checks ---------

void ref_quantize_lut(const float* table, uint64_t m, uint8_t* lut, float* scale, float* offset) {
    detail::quantize_lut(table, m, lut, *scale, *offset);
}

void ref_build_quantize_lut(const float* B, const float* C, float A, uint64_t m, float* table, uint8_t* lut,
                            float* scale, float* offset) {
    detail::build_quantize_lut(B, C, A, m, table, lut, *scale, *offset);
}

uint32_t ref_fast_scan_block_mask(const uint8_t* lut, const uint8_t* block, uint64_t m, uint16_t limit, uint16_t* out) {
    return detail::fast_scan_block_mask(lut, block, m, limit, out);
}

float ref_l2_distance(const float* a, const float* b, uint64_t n) { return l2_distance(a, b, n); }
float ref_inner_product(const float* a, const float* b, uint64_t n) { return inner_product(a, b, n); }

} // extern "C"
