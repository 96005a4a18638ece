// Device-side views of an index and the launch wrappers of the search kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qd {

// Probes per k_fs_sums workgroup (a window of one query's probe order) and its waves.
constexpr int kSumsWin = 16;
constexpr int kSumsWaves = 4;
// Most probes one query can be given: every list of the largest index.
constexpr int kSelectMax = 65535;

// An IVF fast-scan index in HBM. Lists are stored back to back: list l's
// slots at list_ids[list_off[l] .. list_off[l + 1]), its code blocks at
// codes + code_off[l] (256-byte aligned). A block of 32 vectors takes m4 * 16
// bytes (m4 = m rounded up to a multiple of 4, the extra subspaces zero): for
// each group of 4 subspaces, 4 columns of 16 bytes, column g holding bytes
// 4g..4g+3 of the reference's 16-byte rows of those subspaces (src/fast_scan.hpp:19-28).
struct DevIvf {
    int dim, nlist, metric, m, m4, sub, cosine_ip;
    int storage;                   // QANN_IVF_FASTSCAN / PQ / SQ / FLOAT (see qann_hip.h)
    int K;                         // PQ centroids per subspace
    int sq_bits;                   // SQ: 4 or 8
    int vbytes;                    // bytes per entry in codes for PQ / SQ / FLOAT storage
    int sq_chunked;                // 1: 8-bit SQ rows stored chunk-major (sq_row, device_common.hpp)
    const float* sq_vmin;          // dim, SQ only
    const float* sq_step;          // dim, SQ only
    const uint32_t* entry_of_slot; // slot -> list entry (list_off[l] + j), PQ-less storages only
    const float* centroids;        // nlist * dim
    const float* centroid_sqnorms; // nlist
    float cnorm_max;               // an upper bound on every centroid's norm
    const float* residual_centers; // nlist * dim or null
    const float* residual_sqnorms; // nlist or null
    const float* pq_cent_t;        // m * sub * K
    const float* B;                // nlist * m * K, or null (no precomputed tables)
    const float* sdc;              // m * K * K code-to-code table when PQ lists are scored by SDC, else null
    const float* rotation;         // dim * dim or null
    const uint32_t* list_off;      // nlist + 1
    const uint64_t* code_off;      // nlist
    const uint32_t* list_ids;      // slots of every list entry
    const uint8_t* codes;          // fast-scan blocks (below), or vbytes per entry in entry order
    const int64_t* labels;         // slot -> id, or null for ids == slots
    const uint64_t* deleted;       // bitmap over slots, or null when none is deleted
};

// The (query, probe) pairs of a scan grouped by list, for the matrix-core form
// of phase 1: pair_off / unit_off (nlist + 1 each) are per-list prefix counts
// of pairs and of 32-pair units, unit_list (nq * M / 32 + nlist) each unit's
// list, ghist (pairs_ghist_words(nlist, nq)) per-group counts and cursors, tot
// (nlist + 1) the pairs of each list,
// items (nq * M) each pair's item index i * M + p, list by list.
struct PairPlan {
    uint32_t* pair_off;
    uint32_t* unit_off;
    uint32_t* unit_list;
    uint32_t* ghist;
    uint32_t* tot; // nlist + 1: pairs per list
    uint32_t* items;
};
size_t pairs_ghist_words(int nlist, int nq);

// Phase 1 of the fast-scan: the LUT and the u16 sums of every (query, probe).
struct SumsArgs {
    DevIvf ix;
    int M;                     // probe slots per query in the probe arrays
    const float* q;            // nq * dim, prepared (and rotated)
    const uint16_t* probe_idx; // nq * M, sorted nearest first
    const float* probe_A;      // nq * M
    const float* probe_base;   // nq * M
    const int* nprobe_q;       // probes each query scans
    const uint32_t* item_blk;  // nq * M: first block of each probe's sums in its query's region
    const float* C;            // nq * m * 16 or null
    const uint8_t* ip_lut;     // nq * m * 16 or null
    const float* ip_scale;
    const float* ip_offset;
    uint16_t* sums;            // query i's region starts at i * stride_slots
    uint64_t stride_slots;
    uint16_t* bmin;            // smallest sum of each block: query i's region at i * stride_slots / 32
    float* item_scale;         // nq * M
    float* item_offset;
    uint32_t* item_min;        // smallest sum of the probe's list
    int win, nwin;             // probes per workgroup and workgroups per query (set by the launcher)
    int resid_in_lds;
    // Optional order of the queries (query_order): each XCD then takes a
    // contiguous range of it, so queries that probe the same lists share
    // that XCD's L2. nq and per_xcd are set by the launcher.
    const uint32_t* order;
    int nq, per_xcd;
    // 8 counters (one per XCD) for the persistent form, or null for the plain launch only
    uint32_t* queue;
    // buffers of the matrix-core form, or null pointers when it may not run
    PairPlan pairs;
    hipEvent_t grouped; // or null: recorded after the matrix-core form's pair grouping
};

// Phase 1 for PQ (ADC), SQ and float lists: the float distance of every
// vector of every (query, probe), +inf for vectors a query may not return.
// Per-query search parameters of a launch that carries several requests (each
// request its own k, nprobe and filter). Every array has one entry per query of
// the launch (offset with at() for a chunk); a null array means the launch's
// single value. kp and k are at most the launch's kp and k, which size the
// buffers; filter[i] and list_allowed[i] are null for an unfiltered query.
struct PerQuery {
    const int* kp;                       // TopK capacity k' (candidates)
    const int* k;                        // results
    const uint64_t* out_off;             // first entry of the query's result row
    const int* nprobe;
    const int* min_probes;
    const int* wanted;
    const int* block_filter;
    const uint64_t* const* filter;       // allowed-slot bitmap
    const uint32_t* const* list_allowed; // allowed entries per list under that bitmap
    __host__ __device__ PerQuery at(size_t q0) const {
        PerQuery r = *this;
        if (kp) r.kp += q0;
        if (k) r.k += q0;
        if (out_off) r.out_off += q0;
        if (nprobe) r.nprobe += q0;
        if (min_probes) r.min_probes += q0;
        if (wanted) r.wanted += q0;
        if (block_filter) r.block_filter += q0;
        if (filter) r.filter += q0;
        if (list_allowed) r.list_allowed += q0;
        return r;
    }
};

struct DistArgs {
    DevIvf ix;
    int M, win, nwin;
    int mode;                  // kDist* below
    const float* q;
    const uint16_t* probe_idx;
    const float* probe_A;
    const float* probe_base;
    const int* nprobe_q;
    const uint32_t* item_blk;
    const float* C;            // nq * m * K: the query's inner table (kDistIp: the scoring table)
    const uint64_t* filter;    // allowed slots or null
    PerQuery pq;               // per-query filters of a launch of several requests (pq.filter)
    float* dists;              // query i's region starts at i * stride_slots
    uint64_t stride_slots;
    uint32_t* bmin;            // fkey of each block's smallest distance
    uint32_t* item_min;        // fkey of each probe's smallest distance
    int resid_in_lds;
    int mc;                    // ADC: subspaces per table in LDS (m unless a whole table does not fit)
};
enum { kDistIp = 0, kDistL2Pre = 1, kDistL2Resid = 2, kDistSq = 3, kDistFloat = 4, kDistSdc = 5 };

// Phase 2: the reference's TopK replayed over the sums.
struct ReplayArgs {
    DevIvf ix;
    int kp, M, block_filter;
    int unsorted;              // 1: the candidates go to a re-rank, which needs only their set
    const uint16_t* probe_idx;
    const int* nprobe_q;
    const uint32_t* item_blk;
    const uint16_t* sums;      // fast-scan sums, or null when fdists is set
    const float* fdists;       // float distances (PQ / SQ / float lists), or null
    uint64_t stride_slots;
    const uint16_t* bmin;      // fast-scan: smallest sum per block
    const uint32_t* fbmin;     // float: fkey of the smallest distance per block
    const float* item_scale;
    const float* item_offset;
    const uint32_t* item_min;  // fast-scan: smallest sum; float: fkey of the smallest distance
    const uint64_t* filter;    // allowed slots bitmap or null
    PerQuery pq;               // per-query kp, filter and block_filter of a launch of several requests
    float* out_dist;           // nq * kp
    int64_t* out_label;        // nq * kp
    int* out_count;            // nq
    float* g_dist;             // nq * 2kp TopK buffers in HBM when they do not fit LDS, else null
    int64_t* g_label;
    int window;                // block minima the window replay keeps in LDS (set by the launcher)
    int gather;                // 1: the window replay gathers several probes' blocks a batch
};

// LDS budget of a workgroup where more LDS buys fewer workgroups per CU:
// the windows of probe tables and of block minima are sized to it.
constexpr size_t kLdsMax = 64 * 1024;

// Dynamic LDS a launch of the kernel `fn` may ask for on the current device:
// the device's per-workgroup limit (160 KB on gfx950; QANN_AMD_LDS_LIMIT=<bytes>
// lowers it) less the kernel's static LDS. Buffers that do not fit go to HBM.
size_t lds_room(const void* fn);
// lds_room of the kernels that hold a query's TopK buffers (the replays), its
// re-rank keys, its exact-select buffers, its coarse shortlist and its PQ tables
size_t replay_room();
size_t rerank_room();
size_t block_select_room();
size_t coarse_select_room();
size_t ivf_dists_room();

// Replays of batches below this many queries stage their loads through LDS.
constexpr int kReplayStageBelow = 2048;

// LDS bytes of k_fs_replay's TopK buffer.
inline size_t replay_lds_bytes(int kp) { return static_cast<size_t>(2 * kp) * 12; }

void launch_sqnorms(const float* rows, int n, int dim, float* out, hipStream_t s);
void launch_pq_transpose(const float* cent, int m, int K, int sub, float* cent_t, hipStream_t s);
void launch_sdc_table(const float* cent_t, int m, int sub, int K, float* sdc, hipStream_t s);
void launch_list_term(const float* centers, const float* cent_t, int nlist, int dim, int m, int sub, int K, float* B,
                      hipStream_t s);
void launch_rotate(const float* q, const float* R, int nq, int dim, float* out, hipStream_t s);
// Coarse quantizer in two launches: approximate scores of every (query, list)
// into D (nq * nlist), then per query the M nearest lists by (exact A, list),
// sorted (D is overwritten when a query needs every list scored exactly).
// The select's LDS: dim floats + coarse_select_cap(M) keys.
size_t coarse_select_cap(int M);
void launch_coarse_gemm(const float* q, int nq, const float* cents, const float* csq, int nlist, int dim, int ip,
                        float* D, hipStream_t s, const uint64_t* excl = nullptr, const uint64_t* incl = nullptr);
// split_kp > 0: D holds the split BF16 product's scores over K = 3 split_kp
// (launch_coarse_split), and the shortlist bound is that product's. When
// coarse_select_global(M, dim), the shortlist does not fit LDS and gkeys must
// hold nq * coarse_select_cap(M) keys of global memory.
bool coarse_select_global(int M, int dim);
// Each probe's A and base as the scan takes them (src/ivf_index.cpp:526-529,
// 567-570), written by the select beside the probe (null A: not wanted): A is
// the list's own exact term when the index has residual centres, else the
// coarse distance; base = A, or for Cosine inner-product tables
// (A + 1) - |c|^2.
struct ProbeBaseOut {
    const float* residual_centers = nullptr; // nlist * dim or null
    const float* residual_sqnorms = nullptr;
    const float* centroid_sqnorms = nullptr;
    int ip = 0;          // residual terms by inner product (Metric::InnerProduct)
    int cosine_base = 0; // base = (A + 1) - |c|^2
    float* A = nullptr;  // nq * M
    float* base = nullptr;
};
ProbeBaseOut probe_base_out(const DevIvf& ix, float* A, float* base);
void launch_coarse_select(float* D, const float* q, int nq, const float* cents, float cnorm_max, int nlist, int dim,
                          int ip, int M, uint16_t* out_idx, float* out_val, hipStream_t s, int split_kp = 0,
                          uint64_t* gkeys = nullptr, const ProbeBaseOut& pb = ProbeBaseOut{});
// FlatIndex over D from launch_coarse_gemm (rows as the "centroids", masked by
// excl / incl): the k nearest rows per query, exact, in (distance, slot) order.
void launch_flat_select(float* D, const float* q, int nq, const float* rows, int n, int dim, int ip, float xnorm_max,
                        int k, const uint64_t* excl, const uint64_t* incl, const int64_t* labels, int64_t* out_ids,
                        float* out_dists, hipStream_t s);
// Per query: kTableInner the inner table C (m * K), kTableIp the inner-product
// scoring table (C, halved for Metric::InnerProduct), kTableIpLut that table
// rounded to fast-scan bytes (K = 16) with its scale and offset.
enum { kTableInner = 0, kTableIp = 1, kTableIpLut = 2 };
void launch_query_tables(const float* q, const DevIvf& ix, int nq, int mode, float* C, uint8_t* lut, float* scale,
                         float* offset, hipStream_t s);
// Dynamic LDS of k_fs_sums (0 when a table does not fit); sets win, nwin and the LDS flags.
size_t fs_sums_lds(SumsArgs& a);
// Phase 1, as one workgroup per (query, window), as the persistent
// work-queue kernel, or on the matrix cores (pairs grouped by list, int8
// MFMA); which one a problem class runs is chosen by timing each on its first
// launch (QANN_AMD_TUNE_LOG=1 prints the choice; QANN_AMD_SUMS_FORM=grid,
// queue or mfma forces one).
// Returns the form launched: 0 one workgroup per window, 1 persistent
// queue, 2 matrix cores (which records a.grouped, when set, after its pair
// grouping).
// Whether the sums form launch_fs_sums will run reads a.order (the query order).
bool fs_sums_reads_order(SumsArgs a, int nq);
// Whether it reads a.C (the per-query L2 tables).
bool fs_sums_reads_tables(SumsArgs a, int nq);
int launch_fs_sums(SumsArgs a, int nq, hipStream_t s);
// Whether the matrix-core form can run these sums: inner-product LUTs or
// precomputed L2 tables, m4 <= 64, the pair buffers given.
bool fs_mfma_ok(const SumsArgs& a);
// The queries ordered by their nearest probed list (a counting sort over the
// lists; hist takes nlist counters): consecutive queries in it probe many of
// the same lists.
void launch_query_order(const uint16_t* probe_idx, int M, int nq, int nlist, uint32_t* hist, uint32_t* order,
                        hipStream_t s);
void launch_probe_plan(const DevIvf& ix, int nq, int M, int nprobe, int min_probes, int wanted,
                       const uint16_t* probe_idx, const uint32_t* list_allowed, int* nprobe_q, uint32_t* item_blk,
                       int* status, hipStream_t s, const PerQuery& pq = PerQuery{});
void launch_list_allowed(const DevIvf& ix, const uint64_t* filter, uint32_t* out, hipStream_t s);
void launch_fs_replay(const ReplayArgs& a, int nq, hipStream_t s);
// The LUT build and the block lookups of k_fs_sums on their own, for checking
// them against quantize_lut and fast_scan_block: n tables of m x 16 floats at
// ix.B (negzero: m x 16 floats of -0) rounded to plain LUT bytes with their
// scales and offsets; nblocks blocks (device layout) summed against one LUT
// each (plain bytes), 32 u16 sums per block.
void launch_fs_lut_check(const DevIvf& ix, int n, const float* negzero, uint8_t* lut, float* scale, float* offset,
                         hipStream_t s);
void launch_fs_block_check(const uint8_t* luts, int m, int m4, const uint8_t* blocks, int nblocks, uint16_t* out,
                           hipStream_t s);
// Dynamic LDS of k_ivf_dists (0 when not even one subspace's table fits);
// sets win, nwin, mc and resid_in_lds.
size_t ivf_dists_lds(DistArgs& a);
void launch_ivf_dists(DistArgs a, int nq, hipStream_t s);
// Keys of one query's sort: in LDS, or gkeys (nq * rerank_keys(kc)) in HBM when too many.
size_t rerank_keys(int kc);
void launch_rerank(const float* q, int nq, int dim, int ip, const float* rows, const int64_t* cand,
                   const int* cand_count, int kc, int k, const int64_t* labels, int64_t* out_ids, float* out_dists,
                   uint64_t* gkeys, hipStream_t s, const PerQuery& pq = PerQuery{});
// How k_block_select scores a slot: exact distance to its float row, found
// at rows + slot * dim or through entry_of_slot (an IVF's float lists), or
// the SQ score of its code (an IVF's SQ lists).
struct SlotScorer {
    const float* rows;
    const uint32_t* entry_of_slot; // null: rows are indexed by slot
    const uint8_t* sq_codes;       // non-null: SQ scoring of sq_codes + entry * vbytes
    const float* sq_vmin;
    const float* sq_step;
    int sq_bits, vbytes, sq_chunked;
};
// Bytes of one query's selection buffers; in HBM (gbuf, nq of them) when above kLdsMax.
size_t block_select_bytes(int k, bool by_label);
void launch_block_select(const float* q, int nq, int dim, int ip, const SlotScorer& sc, const uint32_t* slot_list,
                         uint32_t n_slots, const uint64_t* deleted, int use_fkey, int final_by_label, int k,
                         const int64_t* labels, int64_t* out_ids, float* out_dists, unsigned char* gbuf,
                         hipStream_t s);

} // namespace qd
