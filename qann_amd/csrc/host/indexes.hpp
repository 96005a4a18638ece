// Host-side index objects with the reference's API (include/vecengine/
// index.hpp, flat_index.hpp, ivf_index.hpp, refine_index.hpp): they own the
// index contents as the saved format holds them, keep a device copy in HBM
// through the C ABI of include/qann_hip.h, and run every query there.
#pragma once

#include "metadata.hpp"
#include "qfile.hpp"

#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <random>
#include <shared_mutex>
#include <string>
#include <vector>

struct qann_hip_index;

namespace qh {

// Throws the exception the reference would for a C-ABI status: ValueError
// (std::invalid_argument), MemoryError (std::bad_alloc) or RuntimeError.
void check_status(int rc);

class DenseIndex {
public:
    virtual ~DenseIndex();
    virtual uint64_t dim() const = 0;
    virtual Metric metric() const = 0;
    virtual const IdMap& ids() const = 0;
    virtual IdMap& ids() = 0;
    uint64_t size() const { return ids().live(); }
    uint64_t slots() const { return ids().slots; }

    // DenseIndex::query_batch (src/index.cpp:117-159, query_prepared :170-184):
    // padded nq x k results. allowed / n_allowed: the id filter, or null;
    // where: a metadata filter, or null (both: slots passing both).
    void batch_query(const float* queries, uint64_t nq, uint32_t k, const int64_t* allowed, size_t n_allowed,
                     bool filtered, int64_t* out_ids, float* out_dists, const Expr* where = nullptr);
    // The slot filter of a query (IndexBase::where_filter, src/index.cpp:161-168,
    // or IdMap::filter): the rows where matches, AND the allowed ids when
    // given, without deleted slots.
    SlotFilter filter_of(const int64_t* allowed, size_t n_allowed, bool filtered, const Expr* where) const;

    // batch_query on prepared queries already in HBM, results left in HBM,
    // enqueued on stream (qann_hip_search_async). No filter.
    void search_device(const float* d_queries, uint64_t nq, uint32_t k, int64_t* d_ids, float* d_dists, void* stream);

    // batch_query's search on queries already prepared, with a slot filter (or
    // null) and an nprobe for IVF (0: the index's own); distances converted and
    // rows padded as batch_query leaves them. The caller holds the locks.
    void search_prepared(const float* prepared, uint64_t nq, uint32_t k, uint32_t nprobe, const SlotFilter* filter,
                         int64_t* out_ids, float* out_dists) const;
    // An upper estimate of the device scratch such a search takes.
    uint64_t scratch_bytes(uint64_t nq, uint32_t k, uint32_t nprobe, bool filtered) const;
    // Several search_prepared calls answered at once (qann_hip_search_requests):
    // request r takes the next reqs[r].nq prepared rows and writes its nq * k
    // results after those of the requests before it. The caller holds the locks.
    struct Request {
        uint64_t nq = 0;
        uint32_t k = 0, nprobe = 0; // nprobe 0: the index's own
        const SlotFilter* filter = nullptr;
    };
    void search_requests(const float* prepared, const Request* reqs, size_t nreq, int64_t* out_ids,
                         float* out_dists) const;
    // The C-ABI handle of the device copy (built if needed; the caller holds
    // the lock shared), valid until the index changes.
    qann_hip_index* device_handle() const { return device(); }
    // the queries prepared for the metric (unit length for Cosine)
    std::vector<float> prepare(const float* q, uint64_t n) const;

    // add_batch (src/index.cpp:33-49): returns the added ids; md: the rows'
    // metadata columns, or null.
    std::vector<int64_t> add(const float* vecs, uint64_t n, const int64_t* ids,
                             const std::vector<MetaInput>* md = nullptr);
    virtual size_t remove(const int64_t* ids, size_t n) = 0;
    // the live ids in ascending order; with where, those whose rows match it
    std::vector<int64_t> live_ids(const Expr* where = nullptr) const;
    // the metadata columns of this index (MetadataStore, one row per slot)
    virtual const Metadata& meta() const = 0;
    virtual Metadata& meta() = 0;
    virtual bool stores_vectors() const { return false; }
    std::vector<float> get_vectors(const int64_t* ids, size_t n) const;
    virtual void save(const std::string& path) const = 0;

    // readers take it shared, writers exclusive (python_bindings.cpp:97-131)
    mutable std::shared_mutex mu;
    bool wrapped = false;
    // the RefineIndex wrapping this index, whose device copy points into ours
    const DenseIndex* wrapper = nullptr;

    // frees the device copy after a change (called with mu held exclusively);
    // a wrapped index also frees its wrapper's copy, which points into its own
    virtual void invalidate() const = 0;

protected:
    friend class RefineIndex;
    // the vectors already prepared (unit length for Cosine)
    virtual void add_prepared(const float* vecs, uint64_t n, const int64_t* ids) = 0;
    virtual const float* stored_row(uint32_t) const { return nullptr; }
    // builds or refreshes the device copy; called with mu held shared
    virtual qann_hip_index* device() const = 0;
    virtual void fill_params(void* params) const = 0;
    mutable std::mutex dev_mu;
};

class FlatIndex : public DenseIndex {
public:
    FlatIndex(uint64_t dim, bool custom_ids, Metric metric);
    explicit FlatIndex(std::unique_ptr<FlatData> data);
    ~FlatIndex() override;
    uint64_t dim() const override { return d_->dim; }
    Metric metric() const override { return d_->metric; }
    const IdMap& ids() const override { return d_->ids; }
    IdMap& ids() override { return d_->ids; }
    size_t remove(const int64_t* ids, size_t n) override;
    bool stores_vectors() const override { return true; }
    void save(const std::string& path) const override;
    const FlatData& data() const { return *d_; }
    const Metadata& meta() const override { return d_->meta; }
    Metadata& meta() override { return d_->meta; }

protected:
    void add_prepared(const float* vecs, uint64_t n, const int64_t* ids) override;
    const float* stored_row(uint32_t s) const override { return d_->rows.data() + static_cast<size_t>(s) * d_->dim; }
    qann_hip_index* device() const override;
    void fill_params(void* params) const override;
    void invalidate() const override;

private:
    std::unique_ptr<FlatData> d_;
    mutable qann_hip_index* dev_ = nullptr;
};

class IVFIndex : public DenseIndex {
public:
    IVFIndex(uint64_t dim, uint64_t nlist, uint64_t nprobe, bool custom_ids, Metric metric);
    explicit IVFIndex(std::unique_ptr<IvfData> data);
    ~IVFIndex() override;
    uint64_t dim() const override { return d_->dim; }
    Metric metric() const override { return d_->metric; }
    const IdMap& ids() const override { return d_->ids; }
    IdMap& ids() override { return d_->ids; }
    size_t remove(const int64_t* ids, size_t n) override;
    void save(const std::string& path) const override;

    uint64_t nlist() const { return d_->nlist; }
    uint64_t nprobe() const { return d_->nprobe; }
    void set_nprobe(uint64_t n);
    bool fast_scan() const { return d_->fast_scan; }
    bool pq_rotate() const { return d_->rotate; }
    int sq_bits() const { return d_->sq_bits; }
    bool trained() const { return d_->trained; }
    bool precomputed_tables() const { return wants_precomputed(*d_); }
    void set_precomputed_tables(bool on);
    bool cosine_inner_product() const { return d_->cosine_ip; }
    void set_cosine_inner_product(bool on);
    bool filter_probe_scaling() const { return filter_probe_scaling_; }
    void set_filter_probe_scaling(bool on) { filter_probe_scaling_ = on; }
    uint8_t pq_distance() const { return d_->pq_distance; }
    void set_pq_distance(uint8_t mode); // 0 ADC, 1 SDC

    // IVFIndex::enable_pq (src/ivf_index.cpp:83-94)
    void enable_pq(uint64_t m, uint64_t K, bool fast_scan, bool rotate);
    // IVFIndex::enable_sq (src/ivf_index.cpp:103-109)
    void enable_sq(int bits);
    // IVFIndex::train (src/ivf_index.cpp:725-806): e-means in `passes` passes
    // (emeans_coarse) or Lloyd's k-means for `iters` iterations, then the PQ
    // codebooks and SQ ranges, as the reference trains them, with every
    // assign step on the GPU; the cosine table mode is this library's own
    // choice (cosine_tables_by_recall)
    void train(const float* vecs, uint64_t n, uint64_t iters, std::optional<uint32_t> seed, bool emeans_coarse = true,
               uint64_t passes = 2);
    const IvfData& data() const { return *d_; }
    const Metadata& meta() const override { return d_->meta; }
    Metadata& meta() override { return d_->meta; }

protected:
    void add_prepared(const float* vecs, uint64_t n, const int64_t* ids) override;
    qann_hip_index* device() const override;
    void fill_params(void* params) const override;
    void invalidate() const override;

private:
    friend class RefineIndex;
    bool cosine_tables_by_recall(const float* x, uint64_t n, const std::vector<uint32_t>& assign, uint32_t seed) const;
    std::unique_ptr<IvfData> d_;
    bool filter_probe_scaling_ = true;
    mutable qann_hip_index* dev_ = nullptr;
};

class RefineIndex : public DenseIndex {
public:
    RefineIndex(std::shared_ptr<DenseIndex> base, uint64_t k_factor, bool custom_ids);
    RefineIndex(std::shared_ptr<DenseIndex> base, std::unique_ptr<RefineData> data);
    ~RefineIndex() override;
    uint64_t dim() const override { return base_->dim(); }
    Metric metric() const override { return base_->metric(); }
    const IdMap& ids() const override { return d_->ids; }
    IdMap& ids() override { return d_->ids; }
    size_t remove(const int64_t* ids, size_t n) override;
    bool stores_vectors() const override { return true; }
    void save(const std::string& path) const override;
    uint64_t k_factor() const { return d_->k_factor; }
    void set_k_factor(uint64_t kf);
    std::shared_ptr<DenseIndex> base() const { return base_; }
    const Metadata& meta() const override { return d_->meta; }
    Metadata& meta() override { return d_->meta; }

protected:
    void add_prepared(const float* vecs, uint64_t n, const int64_t* ids) override;
    const float* stored_row(uint32_t s) const override { return d_->rows.data() + static_cast<size_t>(s) * dim(); }
    qann_hip_index* device() const override;
    void fill_params(void* params) const override;
    void invalidate() const override;

private:
    std::shared_ptr<DenseIndex> base_;
    std::unique_ptr<RefineData> d_;
    mutable qann_hip_index* dev_ = nullptr;
};

// load_index (src/index_factory.cpp:120-139)
std::shared_ptr<DenseIndex> load_index(const std::string& path);

// set_num_threads / num_threads (src/threads.cpp): host threads for adds and training.
void set_num_threads(size_t n);
size_t num_threads();

} // namespace qh
