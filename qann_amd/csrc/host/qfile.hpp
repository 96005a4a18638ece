// The reference's saved-index format (src/serialize.hpp and each index's
// save / load_body), read and written by this project's own code so a file
// saved by the reference loads here unchanged and a file saved here loads in
// the reference. Covers Flat, IVF (fast-scan PQ, plain PQ, SQ, float lists)
// and Refine bodies, format versions 1 to 13; this writer emits version 13.
#pragma once

#include <cstdint>
#include <iosfwd>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace qh {

using std::size_t;
using std::vector;

enum Metric : uint8_t { L2 = 0, Cosine = 1, InnerProduct = 2 };
enum Kind : uint8_t { KFlat = 1, KIVF = 2, KRefine = 3 };

constexpr uint32_t kMagic = 0x4E4E4151;   // "QANN" little-endian (serialize.hpp:23)
constexpr uint32_t kFormatVersion = 13;   // serialize.hpp:24
constexpr uint32_t kRemovedSlot = 0xFFFFFFFFu; // IdMap::kRemoved (id_map.hpp:93)

// detail::SlotFilter (id_map.hpp:20-35)
struct SlotFilter {
    vector<uint64_t> words;
    uint64_t count = 0;
};

// detail::IdMap (id_map.hpp, src/id_map.cpp): slot <-> id and deleted flags.
struct IdMap {
    bool custom = false;
    uint64_t slots = 0;
    int64_t next_id = 0;
    vector<int64_t> labels; // empty while ids equal slots
    vector<uint8_t> deleted;
    uint64_t deleted_count = 0;
    std::unordered_map<int64_t, uint32_t> slot_of; // live id -> slot, custom ids only

    int64_t label(uint32_t s) const { return labels.empty() ? static_cast<int64_t>(s) : labels[s]; }
    uint64_t live() const { return slots - deleted_count; }
    bool ids_are_slots() const { return !custom && labels.empty() && static_cast<uint64_t>(next_id) == slots; }
    // IdMap::filter (id_map.cpp:114-142)
    SlotFilter filter(const int64_t* ids, size_t n) const;
    // IdMap::check / add (id_map.cpp:11-53)
    void check(size_t n, const int64_t* ids) const;
    void add(size_t n, const int64_t* ids);
    // IdMap::remove (id_map.cpp:55-91); appends newly deleted slots to removed when given
    size_t remove(const int64_t* ids, size_t n, vector<uint32_t>* removed);
};

// MetadataStore's saved columns (src/metadata.cpp:1542-1610), kept so a file
// round-trips; where filters are not evaluated by this library.
struct MetaColumn {
    std::string name;
    uint8_t type = 0; // 0 int64, 1 float64, 2 string
    vector<uint8_t> valid;
    vector<int64_t> ints;
    vector<double> floats;
    vector<uint32_t> codes;
    vector<std::string> dict;
};
struct Metadata {
    vector<MetaColumn> columns;
    void append_nulls(size_t n); // rows added without metadata
};

struct PQ {
    uint64_t dim = 0, m = 0, K = 0, sub = 0;
    vector<float> centroids; // [s][c][sub], as saved
};

struct IvfList {
    vector<uint32_t> ids;
    vector<uint8_t> codes; // PQ (fast-scan blocks or m bytes per vector) or SQ codes
    vector<float> vecs;    // float lists
};

struct IvfData {
    uint64_t dim = 0, nlist = 0, nprobe = 1;
    bool trained = false;
    uint8_t pq_distance = 0; // 0 ADC, 1 SDC
    vector<float> centroids;
    bool has_pq = false, fast_scan = false, rotate = false;
    PQ pq;
    vector<float> rotation;
    int sq_bits = 0;
    vector<float> sq_vmin, sq_step;
    vector<IvfList> lists;
    IdMap ids;
    bool use_precomputed = true;
    Metric metric = L2;
    Metadata meta;
    bool cosine_ip = true;
    vector<float> residual_centers;
};

struct FlatData {
    uint64_t dim = 0;
    vector<float> rows;
    IdMap ids;
    Metric metric = L2;
    Metadata meta;
};

struct RefineData {
    uint64_t k_factor = 10;
    vector<float> rows;
    IdMap ids;
    Metadata meta;
};

// One loaded file: exactly one of flat / ivf is the base; refine is set for a RefineIndex.
struct LoadedIndex {
    Kind kind = KFlat;
    uint32_t version = 0;
    std::unique_ptr<FlatData> flat;
    std::unique_ptr<IvfData> ivf;
    std::unique_ptr<RefineData> refine; // with flat or ivf as its base
};

// load_index(path) (src/index_factory.cpp:120-139). Throws std::runtime_error
// ("corrupt file: ..." / "truncated file" / "not a qann index file") as the reference does.
LoadedIndex load_file(const std::string& path);

// save_index(index, path): writes path + ".tmp" then renames it into place
// (src/index_factory.cpp:73-96).
void save_file(const std::string& path, const FlatData* flat, const IvfData* ivf, const RefineData* refine);

// PQ codebook transpose [s][sub][c] (PQCodebook::build_tables, src/pq.cpp:266-275).
vector<float> pq_transpose(const PQ& pq);

// Whether the reference builds precomputed tables for this IVF (build_precomputed, ivf_index.cpp:66-81).
bool wants_precomputed(const IvfData& x);

} // namespace qh
