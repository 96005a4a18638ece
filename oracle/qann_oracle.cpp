// qann_oracle.cpp -- CPU restatement of the reference's search path.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle for the HIP
// product path in qann_amd/. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it (through oracle/refpy.py); the product
// never links, imports or calls it.
//
// It restates, in plain scalar C++, what maxtmiller/qann (the reference) does
// for IVFIndex / RefineIndex / FlatIndex batch_query and query, with every
// floating-point step spelled out so the float bits are fixed:
//   * the reference is built with g++ -O3 -mavx2 -mfma, whose default
//     -ffp-contract=fast fuses a*b+c into fma; where it does, this file calls
//     fmaf explicitly, and it is itself built with -ffp-contract=off;
//   * exact distances follow the AVX2 kernels' order (4 x 8-lane
//     accumulators, fixed horizontal sum, scalar fma tail), i.e. the
//     reference with QANN_DISABLE_AVX512=1.
//
// Parity pinning (see tests/test_oracle.py): the reference's shipped fixtures
// (tests/data: v4_ivf_fs, v11_ivf_cos_fs, v12_ivf_cos_l2, v9_ip_refine_ivf_fs,
// v8_ivf_rot, flats, refines, ivf_pq, sq8, float), bit-equality with the
// reference library compiled from its own sources (oracle/_ref) on synthetic
// indexes (SDC-scored PQ among them), and the known-answer cases of
// tests/test_fast_scan.cpp.
//
// Exported as plain C functions (orc_*) for ctypes.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace orc {

using std::size_t;
using std::vector;

enum Metric : uint8_t { L2 = 0, Cosine = 1, InnerProduct = 2 };
enum Kind : uint8_t { KFlat = 1, KIVF = 2, KRefine = 3 };

// ---------------------------------------------------------------------------
// Exact distances in the reference's AVX2 order (src/distance.cpp:69-153).
// ---------------------------------------------------------------------------

// hsum256(add(add(acc0, acc1), add(acc2, acc3))) of distance.cpp:57-65, 97.
static float hsum_4x8(const float acc[4][8]) {
    float v[8];
    for (int l = 0; l < 8; ++l) v[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);
    float lo[4];
    for (int j = 0; j < 4; ++j) lo[j] = v[j] + v[j + 4];
    return (lo[0] + lo[1]) + (lo[2] + lo[3]);
}

// l2_distance_avx2 (distance.cpp:69-101): main loop of 32, then 8-wide steps
// into acc0, then the scalar tail as g++ 11 -O3 -mavx2 -mfma compiles it.
static float l2_exact(const float* a, const float* b, size_t n) {
    float acc[4][8] = {};
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int r = 0; r < 4; ++r)
            for (int l = 0; l < 8; ++l) {
                const float d = a[i + 8 * r + l] - b[i + 8 * r + l];
                acc[r][l] = fmaf(d, d, acc[r][l]);
            }
    for (; i + 8 <= n; i += 8)
        for (int l = 0; l < 8; ++l) {
            const float d = a[i + l] - b[i + l];
            acc[0][l] = fmaf(d, d, acc[0][l]);
        }
    float result = hsum_4x8(acc);
    // g++ -O3 vectorizes the scalar tail (fewer than 8 left): with 4 or more
    // left, the first 4 squares are rounded and added in order, unfused; the
    // rest are fused.
    size_t j = 0;
    const size_t r = n - i;
    if (r >= 4)
        for (; j < 4; ++j) {
            const float d = a[i + j] - b[i + j];
            const float sq = d * d;
            result = result + sq;
        }
    for (; j < r; ++j) {
        const float d = a[i + j] - b[i + j];
        result = fmaf(d, d, result);
    }
    return result;
}

// inner_product_avx2 (distance.cpp:135-148).
static float ip_exact(const float* a, const float* b, size_t n) {
    float acc[4][8] = {};
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int r = 0; r < 4; ++r)
            for (int l = 0; l < 8; ++l) acc[r][l] = fmaf(a[i + 8 * r + l], b[i + 8 * r + l], acc[r][l]);
    for (; i + 8 <= n; i += 8)
        for (int l = 0; l < 8; ++l) acc[0][l] = fmaf(a[i + l], b[i + l], acc[0][l]);
    float result = hsum_4x8(acc);
    // the tail as in l2_exact
    size_t j = 0;
    const size_t r = n - i;
    if (r >= 4)
        for (; j < 4; ++j) {
            const float p = a[i + j] * b[i + j];
            result = result + p;
        }
    for (; j < r; ++j) result = fmaf(a[i + j], b[i + j], result);
    return result;
}

// ---------------------------------------------------------------------------
// Fast-scan primitives (src/fast_scan.hpp:19-43, src/fast_scan.cpp:22-46,
// 243-262, 379-385).
// ---------------------------------------------------------------------------

constexpr size_t kBlock = 32;
constexpr uint8_t kLutMax = 63;
constexpr size_t kMaxSubspaces = 1024;

static inline uint8_t fs_code(const uint8_t* block, size_t v, size_t s) {
    const uint8_t b = block[s * 16 + (v & 15)];
    return v < 16 ? static_cast<uint8_t>(b & 0x0F) : static_cast<uint8_t>(b >> 4);
}

// quantize_lut_scalar: row minimum and maximum, one scale for the widest row,
// offset = (p0 + p1) + (p2 + p3) with p[i % 4] summing the row minimums in
// row order, bytes = nearbyint((t - min) * scale) (half to even).
static void quantize_lut(const float* table, size_t m, uint8_t* lut, float& scale, float& offset) {
    vector<float> row_min(m);
    float widest = 0.0f, part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (size_t i = 0; i < m; ++i) {
        const float* row = table + i * 16;
        float lo = row[0], hi = row[0];
        for (size_t j = 1; j < 16; ++j) {
            lo = std::min(lo, row[j]);
            hi = std::max(hi, row[j]);
        }
        row_min[i] = lo;
        part[i % 4] += lo;
        widest = std::max(widest, hi - lo);
    }
    offset = (part[0] + part[1]) + (part[2] + part[3]);
    scale = widest > 0.0f ? static_cast<float>(kLutMax) / widest : 1.0f;
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < 16; ++j)
            lut[i * 16 + j] = static_cast<uint8_t>(std::nearbyint((table[i * 16 + j] - row_min[i]) * scale));
}

// fast_scan_block_mask_scalar: 32 sums of picked lut bytes and the mask of
// sums below limit.
static uint32_t fast_scan_block(const uint8_t* lut, const uint8_t* block, size_t m, uint16_t limit, uint16_t* out) {
    uint32_t mask = 0;
    for (size_t v = 0; v < kBlock; ++v) {
        uint32_t sum = 0;
        for (size_t s = 0; s < m; ++s) sum += lut[s * 16 + fs_code(block, v, s)];
        out[v] = static_cast<uint16_t>(sum);
        if (out[v] < limit) mask |= 1u << v;
    }
    return mask;
}

// ---------------------------------------------------------------------------
// TopK (src/topk.hpp:26-138): the k smallest by (distance, push order), kept in
// a buffer of 2k trimmed to k whenever it fills; the threshold is the k-th
// smallest distance at the last trim; the final sort is by (key(dist), id).
// ---------------------------------------------------------------------------

static inline uint32_t fkey(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b ^ (static_cast<uint32_t>(static_cast<int32_t>(b) >> 31) | 0x80000000u);
}
static inline float from_fkey(uint32_t k) {
    const uint32_t b = (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

struct Neighbor {
    uint64_t id;
    float dist;
};

class TopK {
public:
    explicit TopK(size_t k) : k_(k) {}
    float threshold() const { return worst_; }
    size_t size() const { return std::min(dist_.size(), k_); }
    void push(float d, uint64_t id) {
        dist_.push_back(d);
        id_.push_back(id);
        if (dist_.size() >= 2 * k_) trim();
    }
    vector<Neighbor> take_sorted() {
        if (dist_.size() > k_) trim();
        const size_t n = dist_.size();
        vector<size_t> order(n);
        std::iota(order.begin(), order.end(), 0);
        const bool small_ids = std::all_of(id_.begin(), id_.end(), [](uint64_t i) { return i <= UINT32_MAX; });
        if (small_ids) {
            std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
                const uint32_t ka = fkey(dist_[a]), kb = fkey(dist_[b]);
                return ka != kb ? ka < kb : id_[a] < id_[b];
            });
        } else {
            std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
                return dist_[a] < dist_[b] || (!(dist_[b] < dist_[a]) && id_[a] < id_[b]);
            });
        }
        vector<Neighbor> out(n);
        for (size_t i = 0; i < n; ++i) {
            // the packed sort returns from_key(key(d)), which is d itself
            out[i] = {id_[order[i]], dist_[order[i]]};
        }
        dist_.clear();
        id_.clear();
        return out;
    }

private:
    void trim() {
        const size_t n = dist_.size();
        vector<uint32_t> keys(n);
        for (size_t i = 0; i < n; ++i) keys[i] = fkey(dist_[i]);
        vector<uint32_t> sorted = keys;
        std::nth_element(sorted.begin(), sorted.begin() + (k_ - 1), sorted.end());
        const uint32_t t = sorted[k_ - 1];
        size_t below = 0;
        for (uint32_t key : keys) below += key < t;
        size_t ties = k_ - below, w = 0;
        for (size_t i = 0; i < n; ++i) {
            const bool tie = keys[i] == t && ties > 0;
            if (keys[i] < t || tie) {
                dist_[w] = dist_[i];
                id_[w] = id_[i];
                ++w;
            }
            ties -= tie;
        }
        dist_.resize(k_);
        id_.resize(k_);
        worst_ = from_fkey(t);
    }

    size_t k_;
    float worst_ = INFINITY;
    vector<float> dist_;
    vector<uint64_t> id_;
};

// ---------------------------------------------------------------------------
// Ids and filters (include/vecengine/id_map.hpp:20-35, 81-89; src/id_map.cpp).
// ---------------------------------------------------------------------------

struct SlotFilter {
    vector<uint64_t> words;
    size_t count = 0;
    bool test(uint32_t s) const { return (words[s >> 6] >> (s & 63)) & 1; }
};

struct IdMap {
    bool custom = false;
    uint64_t slots = 0;
    int64_t next_id = 0;
    vector<int64_t> labels;
    vector<uint8_t> deleted;
    size_t deleted_count = 0;
    std::unordered_map<int64_t, uint32_t> slot_of;

    int64_t label(uint32_t s) const { return labels.empty() ? static_cast<int64_t>(s) : labels[s]; }
    bool is_deleted(uint32_t s) const { return deleted[s] != 0; }
    size_t live() const { return slots - deleted_count; }

    // IdMap::filter (id_map.cpp:114-142): unknown, deleted and repeated ids ignored.
    SlotFilter filter(const int64_t* ids, size_t n) const {
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
            uint64_t& w = f.words[slot >> 6];
            const uint64_t bit = uint64_t{1} << (slot & 63);
            if (!(w & bit)) {
                w |= bit;
                ++f.count;
            }
        }
        return f;
    }
};

// ---------------------------------------------------------------------------
// The .qann reader (src/serialize.hpp, index bodies' load_body).
// ---------------------------------------------------------------------------

constexpr uint32_t kMagic = 0x4E4E4151;
constexpr uint32_t kMaxVersion = 13;

struct Reader {
    std::ifstream in;
    explicit Reader(const char* path) : in(path, std::ios::binary) {
        if (!in) throw std::runtime_error(std::string("cannot open ") + path + " for reading");
    }
    template <class T>
    T pod() {
        T v;
        in.read(reinterpret_cast<char*>(&v), sizeof(T));
        if (!in) throw std::runtime_error("truncated file");
        return v;
    }
    template <class T>
    vector<T> vec(uint64_t max_len) {
        const uint64_t n = pod<uint64_t>();
        if (n > max_len) throw std::runtime_error("corrupt file: vector length exceeds expected size");
        vector<T> v(n);
        if (n) in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
        if (!in) throw std::runtime_error("truncated file");
        return v;
    }
    void pad(uint64_t align) {
        const uint32_t p = pod<uint32_t>();
        if (p >= align) throw std::runtime_error("corrupt file: invalid padding");
        in.ignore(p);
        if (!in || in.gcount() != static_cast<std::streamsize>(p)) throw std::runtime_error("truncated file");
    }
    void bytes(void* dst, uint64_t n) {
        if (n) in.read(static_cast<char*>(dst), static_cast<std::streamsize>(n));
        if (!in) throw std::runtime_error("truncated file");
    }
};

static IdMap read_id_map(Reader& r, uint64_t slots, uint32_t version) {
    IdMap m;
    const uint8_t custom = r.pod<uint8_t>();
    const uint64_t n = r.pod<uint64_t>();
    const int64_t next_id = version >= 3 ? r.pod<int64_t>() : static_cast<int64_t>(n);
    m.labels = r.vec<int64_t>(n);
    if (n != slots) throw std::runtime_error("corrupt file: id map slot count mismatch");
    m.deleted = r.vec<uint8_t>(n);
    if (m.deleted.size() != slots) throw std::runtime_error("corrupt file: invalid deleted flags length");
    m.custom = custom != 0;
    m.slots = slots;
    m.next_id = next_id;
    for (uint64_t s = 0; s < slots; ++s) {
        if (m.deleted[s]) {
            ++m.deleted_count;
            continue;
        }
        if (m.custom) m.slot_of.emplace(m.labels[s], static_cast<uint32_t>(s));
    }
    return m;
}

static IdMap plain_id_map(uint64_t slots) {
    IdMap m;
    m.slots = slots;
    m.next_id = static_cast<int64_t>(slots);
    m.deleted.assign(slots, 0);
    return m;
}

// MetadataStore::load (src/metadata.cpp:1560-1610): parsed and dropped; the
// oracle does not evaluate where filters.
static void skip_metadata(Reader& r, uint64_t slots) {
    const uint64_t cols = r.pod<uint64_t>();
    if (cols > (uint64_t{1} << 16)) throw std::runtime_error("corrupt file: too many metadata columns");
    for (uint64_t c = 0; c < cols; ++c) {
        r.vec<char>(uint64_t{1} << 24);
        const uint8_t type = r.pod<uint8_t>();
        r.vec<uint8_t>(slots);
        if (type == 0) r.vec<int64_t>(slots);
        else if (type == 1) r.vec<double>(slots);
        else if (type == 2) {
            r.vec<uint32_t>(slots);
            const uint64_t dict = r.pod<uint64_t>();
            for (uint64_t d = 0; d < dict; ++d) r.vec<char>(uint64_t{1} << 24);
        } else throw std::runtime_error("corrupt file: invalid metadata column type");
    }
}

struct PQ {
    size_t dim = 0, m = 0, K = 0, sub = 0;
    vector<float> centroids;   // [s][c][sub]
    vector<float> centroids_t; // [s][sub][c]
};

struct List {
    vector<uint32_t> ids;
    vector<float> vecs;
    vector<uint8_t> codes;
};

struct Index {
    Kind kind = KFlat;
    uint32_t version = 0;
    Metric metric = L2;
    size_t dim = 0;
    IdMap ids;
    // Flat rows, or Refine raw rows
    vector<float> rows;
    // Refine
    size_t k_factor = 0;
    std::unique_ptr<Index> base;
    // IVF
    size_t nlist = 0, nprobe = 0;
    bool trained = false;
    uint8_t pq_distance = 0;
    vector<float> centroids, centroid_sqnorms, residual_centers, residual_sqnorms;
    bool has_pq = false, fast_scan = false, rotate = false, use_precomputed = true, cosine_ip = true;
    bool filter_probe_scaling = true;
    PQ pq;
    vector<float> rotation;
    int sq_bits = 0;
    vector<float> sq_vmin, sq_step;
    vector<List> lists;
    vector<float> precomputed; // B, nlist * m * K
    vector<float> sdc;         // SDC table, m * K * K, when pq_distance is SDC
    vector<uint16_t> slot_list;
    vector<uint32_t> slot_pos;

    const float* residual_center(size_t l) const {
        return (residual_centers.empty() ? centroids.data() : residual_centers.data()) + l * dim;
    }
    size_t sq_code_bytes() const { return sq_bits == 8 ? dim : (dim + 1) / 2; }
    float exact_score(const float* a, const float* b) const {
        return metric == InnerProduct ? -ip_exact(a, b, dim) : l2_exact(a, b, dim);
    }
};

static void read_header(Reader& r, Kind& kind, uint32_t& version) {
    if (r.pod<uint32_t>() != kMagic) throw std::runtime_error("not a qann index file");
    version = r.pod<uint32_t>();
    if (version < 1 || version > kMaxVersion) throw std::runtime_error("unsupported format version");
    const uint8_t k = r.pod<uint8_t>();
    if (k < 1 || k > 3) throw std::runtime_error("index kind not supported by the oracle");
    kind = static_cast<Kind>(k);
}

static Metric read_metric(Reader& r) {
    const uint8_t m = r.pod<uint8_t>();
    if (m > 2) throw std::runtime_error("corrupt file: unknown metric");
    return static_cast<Metric>(m);
}

// PQCodebook::build_tables (src/pq.cpp:266-275): the [s][sub][c] transpose.
static void pq_build(PQ& pq) {
    pq.centroids_t.assign(pq.m * pq.K * pq.sub, 0.0f);
    for (size_t i = 0; i < pq.m; ++i)
        for (size_t j = 0; j < pq.sub; ++j)
            for (size_t k = 0; k < pq.K; ++k)
                pq.centroids_t[(i * pq.sub + j) * pq.K + k] = pq.centroids[(i * pq.K + k) * pq.sub + j];
}

// compute_inner_table (pq.cpp:186-204): C[s][k] accumulates fma(-2 q_d, y, C)
// over the subspace's dimensions in order (g++ fuses the two statements).
static void inner_table(const PQ& pq, const float* q, float* out) {
    for (size_t i = 0; i < pq.m; ++i) {
        float* o = out + i * pq.K;
        std::fill_n(o, pq.K, 0.0f);
        for (size_t j = 0; j < pq.sub; ++j) {
            const float qd = q[i * pq.sub + j];
            const float* c = pq.centroids_t.data() + (i * pq.sub + j) * pq.K;
            for (size_t k = 0; k < pq.K; ++k) o[k] = fmaf(-2.0f * qd, c[k], o[k]);
        }
    }
}

// compute_adc_table (pq.cpp:165-184): fma(diff, diff, acc) per dimension.
static void adc_table(const PQ& pq, const float* r, float* out) {
    for (size_t i = 0; i < pq.m; ++i) {
        float* o = out + i * pq.K;
        std::fill_n(o, pq.K, 0.0f);
        for (size_t j = 0; j < pq.sub; ++j) {
            const float qd = r[i * pq.sub + j];
            const float* c = pq.centroids_t.data() + (i * pq.sub + j) * pq.K;
            for (size_t k = 0; k < pq.K; ++k) {
                const float d = qd - c[k];
                o[k] = fmaf(d, d, o[k]);
            }
        }
    }
}

// The SDC table (pq.cpp:275-292): sdc[i][k][l] = sum over the subspace's
// dimensions, from the first, of fma(y_k - y_l, y_k - y_l, acc).
static vector<float> sdc_table(const PQ& pq) {
    vector<float> t(pq.m * pq.K * pq.K, 0.0f);
    for (size_t i = 0; i < pq.m; ++i)
        for (size_t j = 0; j < pq.sub; ++j) {
            const float* y = pq.centroids_t.data() + (i * pq.sub + j) * pq.K;
            for (size_t k = 0; k < pq.K; ++k) {
                float* row = t.data() + (i * pq.K + k) * pq.K;
                for (size_t l = 0; l < pq.K; ++l) {
                    const float d = y[k] - y[l];
                    row[l] = fmaf(d, d, row[l]);
                }
            }
        }
    return t;
}

// PQCodebook::encode (pq.cpp:88-156): per subspace, the first centroid of
// least (r_0 - y_0)^2 + fma-accumulated (r_j - y_j)^2 over the rest.
static void pq_encode(const PQ& pq, const float* r, uint8_t* out) {
    for (size_t i = 0; i < pq.m; ++i) {
        float best = 0.0f;
        size_t pick = 0;
        for (size_t c = 0; c < pq.K; ++c) {
            float d = r[i * pq.sub] - pq.centroids_t[(i * pq.sub) * pq.K + c];
            float acc = d * d;
            for (size_t j = 1; j < pq.sub; ++j) {
                d = r[i * pq.sub + j] - pq.centroids_t[(i * pq.sub + j) * pq.K + c];
                acc = fmaf(d, d, acc);
            }
            if (c == 0 || acc < best) {
                best = acc;
                pick = c;
            }
        }
        out[i] = static_cast<uint8_t>(pick);
    }
}

// compute_list_term (pq.cpp:206-223): B += y * (y + 2 c_d), fused.
static void list_term(const PQ& pq, const float* c, float* out) {
    for (size_t i = 0; i < pq.m; ++i) {
        float* o = out + i * pq.K;
        std::fill_n(o, pq.K, 0.0f);
        for (size_t j = 0; j < pq.sub; ++j) {
            const float cd = c[i * pq.sub + j];
            const float* y = pq.centroids_t.data() + (i * pq.sub + j) * pq.K;
            for (size_t k = 0; k < pq.K; ++k) o[k] = fmaf(y[k], y[k] + 2.0f * cd, o[k]);
        }
    }
}

// IVFIndex::build_precomputed (ivf_index.cpp:66-81).
static void build_precomputed(Index& x) {
    x.precomputed.clear();
    if (!x.use_precomputed || !x.trained || !x.has_pq || x.metric == InnerProduct ||
        (x.metric == Cosine && x.cosine_ip))
        return;
    const size_t per = x.pq.m * x.pq.K;
    if (x.nlist * per * sizeof(float) > (size_t{256} << 20)) return;
    x.precomputed.resize(x.nlist * per);
    for (size_t l = 0; l < x.nlist; ++l) list_term(x.pq, x.residual_center(l), x.precomputed.data() + l * per);
}

static std::unique_ptr<Index> load_flat(Reader& r, uint32_t version) {
    auto x = std::make_unique<Index>();
    x->kind = KFlat;
    x->version = version;
    x->dim = r.pod<uint64_t>();
    const uint64_t count = r.pod<uint64_t>();
    if (x->dim == 0 || count > UINT64_MAX / x->dim) throw std::runtime_error("corrupt file: invalid FlatIndex shape");
    x->rows = r.vec<float>(count * x->dim);
    if (x->rows.size() != count * x->dim) throw std::runtime_error("corrupt file: FlatIndex data length mismatch");
    x->ids = version >= 2 ? read_id_map(r, count, version) : plain_id_map(count);
    if (version >= 5) x->metric = read_metric(r);
    if (version >= 9) skip_metadata(r, count);
    return x;
}

static std::unique_ptr<Index> load_ivf(Reader& r, uint32_t version) {
    auto x = std::make_unique<Index>();
    x->kind = KIVF;
    x->version = version;
    x->dim = r.pod<uint64_t>();
    x->nlist = r.pod<uint64_t>();
    x->nprobe = r.pod<uint64_t>();
    const uint64_t total = r.pod<uint64_t>();
    x->trained = r.pod<uint8_t>() != 0;
    x->pq_distance = r.pod<uint8_t>();
    if (x->dim == 0 || total > UINT32_MAX) throw std::runtime_error("corrupt file: invalid IVFIndex shape");
    if (x->nlist < 100 || x->nlist > 65535) throw std::runtime_error("corrupt file: nlist must be in [100, 65535]");
    if (x->nprobe == 0 || x->nprobe > x->nlist) throw std::runtime_error("corrupt file: nprobe must be in [1, nlist]");
    const size_t d = x->dim;
    x->centroids = r.vec<float>(x->nlist * d);
    x->centroid_sqnorms.resize(x->centroids.size() / d);
    for (size_t c = 0; c < x->centroid_sqnorms.size(); ++c)
        x->centroid_sqnorms[c] = ip_exact(x->centroids.data() + c * d, x->centroids.data() + c * d, d);
    if (r.pod<uint8_t>()) {
        x->has_pq = true;
        PQ& pq = x->pq;
        pq.dim = r.pod<uint64_t>();
        pq.m = r.pod<uint64_t>();
        pq.K = r.pod<uint64_t>();
        if (pq.dim != d || pq.m == 0 || d % pq.m != 0 || pq.K == 0 || pq.K > 256)
            throw std::runtime_error("corrupt file: invalid IVFIndex shape");
        pq.sub = d / pq.m;
        pq.centroids = r.vec<float>(pq.m * pq.K * pq.sub);
        if (pq.centroids.size() != pq.m * pq.K * pq.sub) throw std::runtime_error("corrupt file: unexpected number of centroids");
        pq_build(pq);
        if (version >= 4) x->fast_scan = r.pod<uint8_t>() != 0;
        if (version >= 8) {
            x->rotate = r.pod<uint8_t>() != 0;
            x->rotation = r.vec<float>(d * d);
        }
    }
    if (version >= 7) {
        x->sq_bits = r.pod<uint8_t>();
        if (x->sq_bits) {
            x->sq_vmin = r.vec<float>(d);
            x->sq_step = r.vec<float>(d);
        }
    }
    x->lists.resize(x->nlist);
    for (size_t l = 0; l < x->nlist; ++l) {
        List& list = x->lists[l];
        list.ids = r.vec<uint32_t>(total);
        const size_t n = list.ids.size();
        if (!x->has_pq && version >= 11) {
            const size_t row_bytes = x->sq_bits ? x->sq_code_bytes() : d * sizeof(float);
            const uint64_t expected = n * (x->sq_bits ? row_bytes : d);
            if (r.pod<uint64_t>() != expected) throw std::runtime_error("corrupt file: IVF list length mismatch");
            r.pad(64);
            if (x->sq_bits) {
                list.codes.resize(n * row_bytes);
                r.bytes(list.codes.data(), n * row_bytes);
            } else {
                list.vecs.resize(n * d);
                r.bytes(list.vecs.data(), n * row_bytes);
            }
        } else if (x->sq_bits) {
            list.codes = r.vec<uint8_t>(n * x->sq_code_bytes());
        } else if (x->has_pq) {
            const uint64_t blocks = (n + kBlock - 1) / kBlock;
            list.codes = r.vec<uint8_t>(x->fast_scan ? blocks * x->pq.m * 16 : n * x->pq.m);
        } else {
            list.vecs = r.vec<float>(n * d);
        }
        for (uint32_t id : list.ids)
            if (id >= total) throw std::runtime_error("corrupt file: IVF id out of range");
    }
    x->ids = version >= 2 ? read_id_map(r, total, version) : plain_id_map(total);
    if (version >= 3) x->use_precomputed = r.pod<uint8_t>() != 0;
    if (version >= 5) x->metric = read_metric(r);
    if (version >= 9) skip_metadata(r, total);
    if (version >= 12) x->cosine_ip = r.pod<uint8_t>() != 0;
    if (version >= 13) {
        x->residual_centers = r.vec<float>(x->nlist * d);
        x->residual_sqnorms.resize(x->residual_centers.size() / d);
        for (size_t l = 0; l < x->residual_sqnorms.size(); ++l)
            x->residual_sqnorms[l] = ip_exact(x->residual_center(l), x->residual_center(l), d);
    }
    build_precomputed(*x);
    if (x->has_pq && x->pq_distance != 0) x->sdc = sdc_table(x->pq);
    if (!x->has_pq) {
        x->slot_list.assign(total, 0);
        x->slot_pos.assign(total, 0);
        for (size_t l = 0; l < x->nlist; ++l)
            for (size_t j = 0; j < x->lists[l].ids.size(); ++j) {
                const uint32_t s = x->lists[l].ids[j];
                if (s == UINT32_MAX) continue;
                x->slot_list[s] = static_cast<uint16_t>(l);
                x->slot_pos[s] = static_cast<uint32_t>(j);
            }
    }
    return x;
}

static std::unique_ptr<Index> load_body(Reader& r, Kind kind, uint32_t version);

static std::unique_ptr<Index> load_refine(Reader& r, uint32_t version) {
    auto x = std::make_unique<Index>();
    x->kind = KRefine;
    x->version = version;
    x->k_factor = r.pod<uint64_t>();
    const uint64_t count = r.pod<uint64_t>();
    if (x->k_factor == 0) throw std::runtime_error("corrupt file: invalid RefineIndex shape");
    Kind bk;
    uint32_t bv;
    read_header(r, bk, bv);
    if (bv != version) throw std::runtime_error("corrupt file: RefineIndex base has a different format version");
    if (bk == KRefine) throw std::runtime_error("corrupt file: RefineIndex base must be Flat, IVF or HNSW");
    x->base = load_body(r, bk, bv);
    x->dim = x->base->dim;
    x->metric = x->base->metric;
    const uint64_t len = r.pod<uint64_t>();
    if (len != count * x->dim) throw std::runtime_error("corrupt file: unexpected number of vectors");
    if (version >= 10) r.pad(64 * 1024);
    x->rows.resize(len);
    r.bytes(x->rows.data(), len * sizeof(float));
    x->ids = version >= 2 ? read_id_map(r, count, version) : plain_id_map(count);
    if (version >= 9) skip_metadata(r, count);
    return x;
}

static std::unique_ptr<Index> load_body(Reader& r, Kind kind, uint32_t version) {
    switch (kind) {
        case KFlat: return load_flat(r, version);
        case KIVF: return load_ivf(r, version);
        case KRefine: return load_refine(r, version);
    }
    throw std::runtime_error("unknown index kind");
}

static std::unique_ptr<Index> load(const char* path) {
    Reader r(path);
    Kind kind;
    uint32_t version;
    read_header(r, kind, version);
    auto x = load_body(r, kind, version);
    if (r.in.peek() != std::ifstream::traits_type::eof())
        throw std::runtime_error("corrupt file: unexpected data after the index");
    return x;
}

// ---------------------------------------------------------------------------
// Search.
// ---------------------------------------------------------------------------

// rotate (src/rotation.cpp:165-185): out[j] = fma(x[k], R[k][j], acc) for k in order.
static vector<float> rotated(const Index& x, const float* q) {
    const size_t d = x.dim;
    vector<float> out(d);
    for (size_t j = 0; j < d; ++j) {
        float acc = 0.0f;
        for (size_t k = 0; k < d; ++k) acc = fmaf(q[k], x.rotation[k * d + j], acc);
        out[j] = acc;
    }
    return out;
}

struct Probe {
    float dist;
    uint16_t list;
    bool operator<(const Probe& o) const { return dist < o.dist || (dist == o.dist && list < o.list); }
};

// coarse (ivf_index.cpp:305-352): the nprobe nearest lists by (A, list),
// nearest first; coarse_all (:362-383): every list in that order.
static vector<Probe> coarse(const Index& x, const float* q, bool all) {
    vector<Probe> p(x.nlist);
    for (size_t l = 0; l < x.nlist; ++l) p[l] = {x.exact_score(q, x.centroids.data() + l * x.dim), static_cast<uint16_t>(l)};
    if (!all) {
        std::partial_sort(p.begin(), p.begin() + x.nprobe, p.end());
        p.resize(x.nprobe);
    } else {
        std::sort(p.begin(), p.end());
    }
    return p;
}

// SQ::score, AVX2 order (sq.cpp:126-211): 8-lane fma accumulators, a fixed
// fold, then the scalar tail with g++'s fusions.
template <bool IP>
static float sq_score(const Index& x, const float* q, const uint8_t* code) {
    const size_t n = x.dim;
    const float* lo = x.sq_vmin.data();
    const float* st = x.sq_step.data();
    const int bits = x.sq_bits;
    auto code_at = [&](size_t d) -> float {
        return bits == 8 ? code[d] : static_cast<float>((code[d / 2] >> ((d & 1) * 4)) & 0x0F);
    };
    float acc[8] = {};
    size_t d = 0;
    auto step8 = [&](size_t at) {
        for (int l = 0; l < 8; ++l) {
            const float v = fmaf(code_at(at + l), st[at + l], lo[at + l]);
            if (IP) acc[l] = fmaf(q[at + l], v, acc[l]);
            else {
                const float diff = q[at + l] - v;
                acc[l] = fmaf(diff, diff, acc[l]);
            }
        }
    };
    if (bits == 8) {
        for (; d + 8 <= n; d += 8) step8(d);
    } else {
        for (; d + 16 <= n; d += 16) {
            step8(d);
            step8(d + 8);
        }
    }
    float half[4];
    for (int j = 0; j < 4; ++j) half[j] = acc[j] + acc[j + 4];
    float sum = (half[0] + half[2]) + (half[1] + half[3]);
    for (; d < n; ++d) {
        const float v = fmaf(code_at(d), st[d], lo[d]);
        if (IP) sum = fmaf(q[d], v, sum);
        else {
            const float diff = q[d] - v;
            sum = fmaf(diff, diff, sum);
        }
    }
    return sum;
}

// IVFIndex::scan (ivf_index.cpp:421-641).
static vector<Neighbor> ivf_scan(const Index& x, const float* q, vector<Probe>& probes, size_t k, const SlotFilter* filter) {
    const size_t m = x.has_pq ? x.pq.m : 0, K = x.has_pq ? x.pq.K : 0;
    vector<float> table(m * K);
    vector<uint8_t> lut(x.fast_scan ? m * K : 0);
    TopK top(k);
    const bool use_pre = x.has_pq && x.pq_distance == 0 && !x.precomputed.empty();
    vector<float> C(use_pre ? m * K : 0);
    if (use_pre) inner_table(x.pq, q, C.data());

    const bool cos_pq = x.has_pq && x.metric == Cosine && x.pq_distance == 0 && x.cosine_ip;
    const bool ip_pq = (x.has_pq && x.metric == InnerProduct) || cos_pq;
    float ip_scale = 1.0f, ip_offset = 0.0f;
    if (ip_pq) {
        inner_table(x.pq, q, table.data());
        if (!cos_pq)
            for (float& t : table) t *= 0.5f;
        if (x.fast_scan) quantize_lut(table.data(), m, lut.data(), ip_scale, ip_offset);
    }
    const bool block_filter = filter && filter->count * 16 < x.ids.slots;

    auto scan_fast = [&](const List& list, float scale, float offset, auto&& push) {
        const size_t bb = m * 16, n = list.ids.size();
        uint16_t sums[kBlock];
        for (size_t start = 0; start < n; start += kBlock) {
            const size_t count = std::min(kBlock, n - start);
            uint32_t allowed = count < kBlock ? (1u << count) - 1 : ~0u;
            if (block_filter) {
                allowed = 0;
                for (size_t v = 0; v < count; ++v) allowed |= uint32_t{filter->test(list.ids[start + v])} << v;
                if (allowed == 0) continue;
            }
            const float limit = (top.threshold() - offset) * scale;
            const uint16_t limit16 = limit >= 65535.0f ? 65535 : limit <= 0.0f ? 0 : static_cast<uint16_t>(std::ceil(limit));
            uint32_t mask = fast_scan_block(lut.data(), list.codes.data() + (start / kBlock) * bb, m, limit16, sums);
            mask &= allowed;
            while (mask) {
                const unsigned v = static_cast<unsigned>(__builtin_ctz(mask));
                mask &= mask - 1;
                push(static_cast<float>(sums[v]) / scale + offset, start + v);
            }
        }
    };

    const size_t wanted = filter ? std::min(k, filter->count) : 0;
    size_t min_probes = x.nprobe;
    if (filter && x.filter_probe_scaling && filter->count > 0) {
        const double want = std::ceil(0.75 * double(x.nprobe) * std::cbrt(double(x.ids.live()) / double(filter->count)));
        min_probes = std::clamp<size_t>(static_cast<size_t>(std::min(want, double(x.nlist))), x.nprobe, x.nlist);
    }

    for (size_t p = 0; p < probes.size(); ++p) {
        if (p >= x.nprobe) {
            if (p >= min_probes && top.size() >= wanted) break;
            if (p == x.nprobe) std::sort(probes.begin() + p, probes.end());
        }
        const Probe& probe = probes[p];
        const float A = !x.residual_centers.empty() ? x.exact_score(q, x.residual_center(probe.list)) : probe.dist;
        const List& list = x.lists[probe.list];
        const size_t n = list.ids.size();
        if (n == 0) continue;

        auto push = [&](float dist, size_t j) {
            const uint32_t slot = list.ids[j];
            if (dist < top.threshold() && (filter ? filter->test(slot) : !x.ids.is_deleted(slot)))
                top.push(dist, static_cast<uint64_t>(x.ids.label(slot)));
        };
        auto passes = [&](size_t j) {
            const uint32_t slot = list.ids[j];
            return slot != UINT32_MAX && (!filter || filter->test(slot));
        };

        if (x.sq_bits) {
            const size_t w = x.sq_code_bytes();
            for (size_t j = 0; j < n; ++j)
                if (passes(j)) {
                    const uint8_t* c = list.codes.data() + j * w;
                    push(x.metric == InnerProduct ? -sq_score<true>(x, q, c) : sq_score<false>(x, q, c), j);
                }
            continue;
        }
        if (!x.has_pq) {
            for (size_t j = 0; j < n; ++j)
                if (passes(j)) push(x.exact_score(q, list.vecs.data() + j * x.dim), j);
            continue;
        }
        if (x.pq_distance != 0 && !ip_pq) {
            // SDC (ivf_index.cpp:626-632): the residual's own code against each
            // row's, summed over subspaces from the first
            const float* c = x.residual_center(probe.list);
            vector<float> residual(x.dim);
            for (size_t l = 0; l < x.dim; ++l) residual[l] = q[l] - c[l];
            vector<uint8_t> qc(m);
            pq_encode(x.pq, residual.data(), qc.data());
            for (size_t j = 0; j < n; ++j) {
                const uint8_t* code = list.codes.data() + j * m;
                float d = 0.0f;
                for (size_t s = 0; s < m; ++s) d += x.sdc[(s * K + qc[s]) * K + code[s]];
                push(d, j);
            }
            continue;
        }
        auto adc_sum = [&](size_t j) {
            const uint8_t* c = list.codes.data() + j * m;
            float d = 0.0f;
            for (size_t s = 0; s < m; ++s) d += table[s * K + c[s]];
            return d;
        };
        if (ip_pq) {
            const float base = cos_pq ? (A + 1.0f) - (x.residual_centers.empty() ? x.centroid_sqnorms[probe.list]
                                                                                 : x.residual_sqnorms[probe.list])
                                      : A;
            if (x.fast_scan) scan_fast(list, ip_scale, ip_offset + base, push);
            else
                for (size_t j = 0; j < n; ++j) push(adc_sum(j) + base, j);
            continue;
        }
        const float* B = use_pre ? x.precomputed.data() + size_t{probe.list} * m * K : nullptr;
        if (use_pre) {
            for (size_t t = 0; t < m * K; ++t) table[t] = B[t] + C[t];
            for (size_t j = 0; j < K; ++j) table[j] += A;
        } else {
            const float* c = x.residual_center(probe.list);
            vector<float> residual(x.dim);
            for (size_t l = 0; l < x.dim; ++l) residual[l] = q[l] - c[l];
            adc_table(x.pq, residual.data(), table.data());
        }
        if (x.fast_scan) {
            float scale = 1.0f, offset = 0.0f;
            quantize_lut(table.data(), m, lut.data(), scale, offset);
            scan_fast(list, scale, offset, push);
        } else {
            for (size_t j = 0; j < n; ++j) push(adc_sum(j), j);
        }
    }
    return top.take_sorted();
}

static vector<Neighbor> ivf_query(const Index& x, const float* q0, size_t k, const SlotFilter* filter) {
    if (!x.trained) throw std::logic_error("train() has not be run yet");
    vector<float> rot;
    const float* q = q0;
    if (x.rotate) {
        rot = rotated(x, q0);
        q = rot.data();
    }
    if (filter && k > 0 && !x.has_pq && filter->count <= 3 * x.nlist) {
        // direct_scan (ivf_index.cpp:248-256)
        TopK top(k);
        for (size_t w = 0; w < filter->words.size(); ++w)
            for (uint64_t bits = filter->words[w]; bits; bits &= bits - 1) {
                const uint32_t slot = static_cast<uint32_t>(w * 64 + __builtin_ctzll(bits));
                const List& list = x.lists[x.slot_list[slot]];
                const size_t j = x.slot_pos[slot];
                float dist;
                if (x.sq_bits) {
                    const uint8_t* c = list.codes.data() + j * x.sq_code_bytes();
                    dist = x.metric == InnerProduct ? -sq_score<true>(x, q, c) : sq_score<false>(x, q, c);
                } else {
                    dist = x.exact_score(q, list.vecs.data() + j * x.dim);
                }
                if (dist < top.threshold()) top.push(dist, static_cast<uint64_t>(x.ids.label(slot)));
            }
        return top.take_sorted();
    }
    vector<Probe> probes = coarse(x, q, filter != nullptr);
    return ivf_scan(x, q, probes, k, filter);
}

// FlatIndex::query_impl (flat_index.cpp:33-64): k smallest by (dist, slot).
static vector<Neighbor> flat_query(const Index& x, const float* q, size_t k, const SlotFilter* filter) {
    vector<std::pair<float, size_t>> cand;
    auto score = [&](size_t s) { cand.emplace_back(x.exact_score(x.rows.data() + s * x.dim, q), s); };
    if (filter) {
        for (size_t w = 0; w < filter->words.size(); ++w)
            for (uint64_t bits = filter->words[w]; bits; bits &= bits - 1) score(w * 64 + __builtin_ctzll(bits));
    } else {
        for (size_t s = 0; s < x.ids.slots; ++s)
            if (!x.ids.is_deleted(static_cast<uint32_t>(s))) score(s);
    }
    const size_t keep = std::min(k, cand.size());
    std::partial_sort(cand.begin(), cand.begin() + keep, cand.end());
    vector<Neighbor> out(keep);
    for (size_t i = 0; i < keep; ++i)
        out[i] = {static_cast<uint64_t>(x.ids.label(static_cast<uint32_t>(cand[i].second))), cand[i].first};
    return out;
}

static vector<Neighbor> query_impl(const Index& x, const float* q, size_t k, const SlotFilter* filter);

// RefineIndex::query_impl (refine_index.cpp:84-90), use_direct (:59-69),
// direct_scan (:71-82), rerank (:112-140).
static vector<Neighbor> refine_query(const Index& x, const float* q, size_t k, const SlotFilter* filter) {
    if (x.ids.live() == 0 || k == 0) return {};
    if (filter && x.base->kind == KIVF && filter->count <= 6 * x.base->nlist) {
        TopK top(k);
        for (size_t w = 0; w < filter->words.size(); ++w)
            for (uint64_t bits = filter->words[w]; bits; bits &= bits - 1) {
                const uint32_t slot = static_cast<uint32_t>(w * 64 + __builtin_ctzll(bits));
                const float dist = x.exact_score(q, x.rows.data() + size_t{slot} * x.dim);
                if (dist < top.threshold()) top.push(dist, static_cast<uint64_t>(x.ids.label(slot)));
            }
        return top.take_sorted();
    }
    const size_t kc = std::min(k * x.k_factor, filter ? filter->count : x.ids.live());
    vector<Neighbor> fetched = query_impl(*x.base, q, kc, filter);
    vector<std::pair<float, size_t>> cand;
    cand.reserve(fetched.size());
    for (const Neighbor& nb : fetched)
        cand.emplace_back(x.exact_score(q, x.rows.data() + nb.id * x.dim), static_cast<size_t>(nb.id));
    const size_t keep = std::min(k, cand.size());
    std::partial_sort(cand.begin(), cand.begin() + keep, cand.end());
    vector<Neighbor> out(keep);
    for (size_t i = 0; i < keep; ++i)
        out[i] = {static_cast<uint64_t>(x.ids.label(static_cast<uint32_t>(cand[i].second))), cand[i].first};
    return out;
}

static vector<Neighbor> query_impl(const Index& x, const float* q, size_t k, const SlotFilter* filter) {
    if (k == 0) return {};
    switch (x.kind) {
        case KFlat: return flat_query(x, q, k, filter);
        case KIVF: return ivf_query(x, q, k, filter);
        case KRefine: return refine_query(x, q, k, filter);
    }
    return {};
}

// DenseIndex::normalize_rows (index.cpp:193-218): a double sum per row, one
// float scale 1 / sqrt(sum).
static vector<float> normalized(const float* q, size_t n, size_t d) {
    vector<float> out(n * d);
    for (size_t i = 0; i < n; ++i) {
        double sq = 0;
        for (size_t j = 0; j < d; ++j) {
            const float c = q[i * d + j];
            sq += double(c) * c;
        }
        const double len = std::sqrt(sq);
        if (len == 0) throw std::invalid_argument("cosine metric needs non-zero vectors");
        const float scale = static_cast<float>(1.0 / len);
        for (size_t j = 0; j < d; ++j) out[i * d + j] = q[i * d + j] * scale;
    }
    return out;
}

static size_t g_threads = 1;

// DenseIndex::query_prepared (index.cpp:170-184) over a batch, in parallel.
static void batch_query(const Index& x, const float* q, size_t nq, size_t k, const SlotFilter* filter,
                        vector<vector<Neighbor>>& out) {
    out.assign(nq, {});
    if (filter && filter->count == 0) return;
    vector<float> unit;
    if (x.metric == Cosine) {
        unit = normalized(q, nq, x.dim);
        q = unit.data();
    }
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    auto worker = [&] {
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= nq) return;
            try {
                out[i] = query_impl(x, q + i * x.dim, k, filter);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu);
                if (!err) err = std::current_exception();
                return;
            }
        }
    };
    const size_t nt = std::max<size_t>(1, std::min(g_threads, nq));
    vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    if (err) std::rethrow_exception(err);
    if (x.metric == Cosine)
        for (auto& row : out)
            for (Neighbor& nb : row) nb.dist *= 0.5f;
}

static const Index& ivf_of(const Index& x) {
    if (x.kind == KRefine && x.base->kind == KIVF) return *x.base;
    if (x.kind == KIVF) return x;
    throw std::invalid_argument("not an IVF index");
}

} // namespace orc

using namespace orc;

namespace {
thread_local std::string g_err;

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
} // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(uint64_t n) { g_threads = n ? n : std::max(1u, std::thread::hardware_concurrency()); }

void* orc_load(const char* path) {
    Index* out = nullptr;
    const int rc = guarded([&] { out = load(path).release(); });
    return rc == 0 ? out : nullptr;
}

void orc_free(void* p) { delete static_cast<Index*>(p); }

int orc_info(void* p, int64_t* info) {
    return guarded([&] {
        const Index& x = *static_cast<Index*>(p);
        info[0] = x.kind;
        info[1] = static_cast<int64_t>(x.dim);
        info[2] = static_cast<int64_t>(x.ids.live());
        info[3] = static_cast<int64_t>(x.ids.slots);
        info[4] = x.metric;
    });
}

int orc_set_nprobe(void* p, uint64_t nprobe) {
    return guarded([&] {
        Index& x = const_cast<Index&>(ivf_of(*static_cast<Index*>(p)));
        if (nprobe == 0 || nprobe > x.nlist) throw std::invalid_argument("nprobe must be in [1, nlist]");
        x.nprobe = nprobe;
    });
}

int orc_set_k_factor(void* p, uint64_t kf) {
    return guarded([&] {
        Index& x = *static_cast<Index*>(p);
        if (x.kind != KRefine) throw std::invalid_argument("not a RefineIndex");
        if (kf == 0) throw std::invalid_argument("k_factor must be >= 1");
        x.k_factor = kf;
    });
}

int orc_set_filter_probe_scaling(void* p, int on) {
    return guarded([&] { const_cast<Index&>(ivf_of(*static_cast<Index*>(p))).filter_probe_scaling = on != 0; });
}

int orc_batch_query(void* p, const float* q, uint64_t nq, uint64_t k, const int64_t* allowed, int64_t nallowed,
                    int64_t* ids, float* dists) {
    return guarded([&] {
        const Index& x = *static_cast<Index*>(p);
        SlotFilter f;
        if (allowed) f = x.ids.filter(allowed, static_cast<size_t>(nallowed));
        vector<vector<Neighbor>> res;
        batch_query(x, q, nq, k, allowed ? &f : nullptr, res);
        for (size_t i = 0; i < nq; ++i) {
            for (size_t j = 0; j < k; ++j) {
                ids[i * k + j] = -1;
                dists[i * k + j] = INFINITY;
            }
            for (size_t j = 0; j < res[i].size() && j < k; ++j) {
                ids[i * k + j] = static_cast<int64_t>(res[i][j].id);
                dists[i * k + j] = res[i][j].dist;
            }
        }
    });
}

// Known-answer hooks on the restated kernels.
void orc_quantize_lut(const float* table, uint64_t m, uint8_t* lut, float* scale, float* offset) {
    quantize_lut(table, m, lut, *scale, *offset);
}

uint32_t orc_fast_scan_block_mask(const uint8_t* lut, const uint8_t* block, uint64_t m, uint16_t limit, uint16_t* out) {
    return fast_scan_block(lut, block, m, limit, out);
}

float orc_l2_distance(const float* a, const float* b, uint64_t n) { return l2_exact(a, b, n); }
float orc_inner_product(const float* a, const float* b, uint64_t n) { return ip_exact(a, b, n); }

// The probes query i of a batch scans first (nprobe of them, in scan order):
// prepared queries (normalized / not rotated: rotation is applied here).
int orc_probes(void* p, const float* q, uint64_t nq, uint16_t* lists, float* dists) {
    return guarded([&] {
        const Index& x = ivf_of(*static_cast<Index*>(p));
        for (size_t i = 0; i < nq; ++i) {
            vector<float> rot;
            const float* qq = q + i * x.dim;
            if (x.rotate) {
                rot = rotated(x, qq);
                qq = rot.data();
            }
            vector<Probe> pr = coarse(x, qq, false);
            for (size_t j = 0; j < pr.size(); ++j) {
                lists[i * x.nprobe + j] = pr[j].list;
                dists[i * x.nprobe + j] = pr[j].dist;
            }
        }
    });
}

} // extern "C"
