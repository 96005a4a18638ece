/*
 * qann_hip.h -- the C ABI between the host index objects and the MI355X
 * (gfx950) search kernels of qann_amd.
 *
 * The reference (maxtmiller/qann) has no device boundary: its search runs
 * in-process behind the C++ virtual hooks DenseIndex::query_batch_impl /
 * query_impl (include/vecengine/index.hpp:277-297). This header is the seam a
 * host implementation of those hooks calls instead of its CPU loops; every
 * entry point below names the reference code whose work it takes over.
 *
 * Conventions
 *   - Plain pointers and sizes only. Pointers are host memory unless a
 *     parameter says "device".
 *   - Every int-returning function returns a status: QANN_OK, or an error
 *     whose message qann_hip_last_error() returns on the calling thread.
 *     QANN_EINVAL maps to the reference's std::invalid_argument (Python
 *     ValueError), QANN_ESTATE / QANN_EDEVICE to std::runtime_error /
 *     std::logic_error (RuntimeError), QANN_ENOMEM to MemoryError.
 *   - Queries arrive prepared as DenseIndex::query_prepared prepares them
 *     (src/index.cpp:170-184): unit length for Metric::Cosine. Rotation of a
 *     rotated IVF index (src/ivf_index.cpp:200-204) is applied on the device.
 *   - Distances come back in the reference's internal form: squared L2 between
 *     prepared vectors, or -(q . x) for inner product; the caller halves them
 *     for Metric::Cosine (l2_to_cosine, src/index.cpp:23-25).
 *   - Result rows are padded with id -1 and distance +inf, as results_padded
 *     does (bindings/python_bindings.cpp:582-594).
 *   - Host input buffers are borrowed for the duration of the call; the index
 *     keeps its own device copies.
 *   - qann_hip_search / qann_hip_search_async may run concurrently on one
 *     handle (the reference's shared-lock queries); the create / set calls need
 *     exclusive access, which the host holds as write_locked does.
 */
#ifndef QANN_HIP_H
#define QANN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    QANN_OK = 0,
    QANN_EINVAL = 1,
    QANN_ESTATE = 2,
    QANN_ENOMEM = 3,
    QANN_EDEVICE = 4,
};

/* Metric values are the reference's file-format values (index.hpp:57-62). */
enum { QANN_METRIC_L2 = 0, QANN_METRIC_COSINE = 1, QANN_METRIC_IP = 2 };

typedef struct qann_hip_index qann_hip_index;

/* Message of the last failed call on this thread ("" if none). */
const char* qann_hip_last_error(void);

/* Number of visible GPUs (0 without a usable device). */
int qann_hip_device_count(int* count);

/* How an IVF index stores its lists (include/vecengine/ivf_index.hpp:32-50):
 * 4-bit PQ fast-scan blocks (enable_pq(m, 16, fast_scan=true)), PQ codes of
 * one byte per subspace scored by ADC (enable_pq without fast-scan), SQ codes
 * (enable_sq, src/sq.cpp), or the raw float vectors. */
enum { QANN_IVF_FASTSCAN = 0, QANN_IVF_PQ = 1, QANN_IVF_SQ = 2, QANN_IVF_FLOAT = 3 };

/*
 * An IVF index: what IVFIndex::load_body builds (src/ivf_index.cpp:976-1122),
 * moved into HBM. The device derives what IVFIndex derives on load: centroid
 * and residual-centre squared norms (update_centroid_sqnorms, :297-303), the
 * PQ transpose (PQCodebook::build_tables, src/pq.cpp:266-275) and, when
 * use_precomputed is set, the per-list term B (build_precomputed, :66-81 with
 * PQCodebook::compute_list_term, src/pq.cpp:206-223).
 */
typedef struct {
    uint32_t dim;
    uint32_t nlist;           /* 100 .. 65535 */
    uint32_t metric;          /* QANN_METRIC_* */
    uint32_t m;               /* PQ subspaces (fast-scan: even, <= 1024); dim % m == 0; 0 without PQ */
    uint32_t cosine_ip;       /* Metric::Cosine scored by inner product (ivf_index.hpp:116) */
    uint32_t use_precomputed; /* build B (the caller applies the reference's rule and 256 MiB cap) */
    const float* centroids;        /* nlist * dim, coarse centroids */
    const float* residual_centers; /* nlist * dim, or NULL (format 13, Cosine PQ) */
    const float* pq_centroids;     /* m * pq_k * (dim / m), the saved [s][c][sub] layout; NULL without PQ */
    const float* rotation;         /* dim * dim row-major R (x * R), or NULL */
    uint32_t storage;              /* QANN_IVF_*; 0 (fast-scan) for zero-initialised descriptors */
    uint32_t pq_k;                 /* PQ centroids per subspace, 1 .. 256 (fast-scan: 16; 0 reads as 16) */
    uint32_t sq_bits;              /* QANN_IVF_SQ: 4 or 8 */
    const float* sq_vmin;          /* QANN_IVF_SQ: dim floats, detail::SQ::vmin */
    const float* sq_step;          /* QANN_IVF_SQ: dim floats, detail::SQ::step */
    uint32_t pq_sdc;               /* QANN_IVF_PQ, L2 or Cosine: score by SDC (PQDistance::SDC, ivf_index.hpp:94-99;
                                      scan at src/ivf_index.cpp:626-632); 0 = ADC. Ignored for inner product. */
} qann_hip_ivf_desc;

/* Creates the device index on the current GPU (replaces IVFIndex::load_body's
 * in-RAM structures for the query path). */
int qann_hip_ivf_create(const qann_hip_ivf_desc* desc, qann_hip_index** out);

/*
 * Uploads every inverted list at once (InvertedList::ids / codes / vecs,
 * ivf_index.hpp:32-50): sizes[l] entries for list l; slots holds the lists'
 * slot ids back to back (sum(sizes) of them); codes holds each list's payload
 * back to back: ceil(sizes[l] / 32) * m * 16 fast-scan bytes in the
 * reference's block layout (src/fast_scan.hpp:19-34), sizes[l] * m PQ code
 * bytes, sizes[l] * SQ code bytes (dim, or ceil(dim / 2) for 4 bits), or
 * sizes[l] * dim floats, by the index's storage.
 */
int qann_hip_ivf_set_lists(qann_hip_index* index, const uint32_t* sizes, const uint32_t* slots, const uint8_t* codes);

/* A flat index of raw rows (FlatIndex, src/flat_index.cpp): slots rows of dim floats. */
int qann_hip_flat_create(uint32_t dim, uint32_t metric, const float* rows, uint64_t slots, qann_hip_index** out);

/*
 * A RefineIndex over an IVF device index (src/refine_index.cpp):
 * holds the raw rows it re-ranks with (VectorStore, one row per slot). base
 * must outlive the refine handle.
 */
int qann_hip_refine_create(qann_hip_index* base, const float* rows, uint64_t slots, qann_hip_index** out);

/*
 * The slot -> id map and deleted flags (detail::IdMap, include/vecengine/
 * id_map.hpp:81-89): labels is NULL when ids equal slots; deleted holds one
 * byte (0/1) per slot, or NULL for none deleted.
 */
int qann_hip_set_ids(qann_hip_index* index, uint64_t slots, const int64_t* labels, const uint8_t* deleted);

typedef struct {
    uint32_t k;                    /* results per query */
    uint32_t nprobe;               /* IVF: lists probed (IVFIndex::nprobe) */
    uint32_t k_factor;             /* Refine: candidates per result (RefineIndex::k_factor) */
    uint32_t filter_probe_scaling; /* IVFIndex::filter_probe_scaling */
    /* Filter (detail::SlotFilter, id_map.hpp:20-35): NULL for none, else
     * ceil(slots / 64) words of allowed live slots (host memory for
     * qann_hip_search, device memory for qann_hip_search_async) and their count. */
    const uint64_t* allowed_words;
    uint64_t allowed_count;
} qann_hip_search_params;

/*
 * Top-k search of nq prepared queries (nq * dim floats): the work of
 * DenseIndex::query_batch -> {IVFIndex, RefineIndex, FlatIndex}::
 * query_batch_impl (src/ivf_index.cpp:258-295, src/refine_index.cpp:92-110,
 * src/index.cpp:186-191). Writes nq * k ids and distances. Synchronous.
 */
int qann_hip_search(qann_hip_index* index, const float* queries, uint64_t nq, const qann_hip_search_params* params,
                    int64_t* out_ids, float* out_dists);

/* One request of qann_hip_search_requests: nq query rows, its own k, nprobe and
 * filter (host memory, as in qann_hip_search_params). */
typedef struct {
    uint64_t nq;
    uint32_t k;
    uint32_t nprobe;
    const uint64_t* allowed_words;
    uint64_t allowed_count;
} qann_hip_request;

/*
 * Several requests at once, each answered as qann_hip_search would answer it
 * alone: the work of concurrent DenseIndex::query_batch calls on one index
 * (each under the shared lock, bindings/python_bindings.cpp:97-113). No single
 * reference function does this; it is for callers that batch requests, such
 * as a serving loop. queries holds the requests' rows back to back; request
 * r's nq * k results start after those of requests 0 .. r-1 in out_ids /
 * out_dists. On an IVF or refine index the requests share one launch, each
 * query with its own k, nprobe and filter (requests that take a direct scan
 * or have an empty filter run on their own); on a FlatIndex they run one after
 * another. k_factor and filter_probe_scaling are the index's. Synchronous.
 */
int qann_hip_search_requests(qann_hip_index* index, const float* queries, const qann_hip_request* requests,
                             uint32_t nreq, uint32_t k_factor, uint32_t filter_probe_scaling, int64_t* out_ids,
                             float* out_dists);

/*
 * qann_hip_search with queries, results and allowed words in device memory,
 * enqueued on stream (a hipStream_t; NULL = the default stream) and returning
 * before the work finishes. Calls on one stream run in order; calls on
 * different streams may overlap on the device. Each stream has its own
 * scratch, held by one call at a time (host threads sharing a stream take
 * turns). The host waits on the device only where it must read a result to
 * go on: a FlatIndex search on its BF16 path (it reads which queries
 * overflowed the candidate cap and re-scores them on the FP32 path), the first
 * search of a problem class (the autotuners time their forms with HIP
 * events), and a call that grows its stream's scratch. A filtered IVF search
 * here probes every list its stopping rule may need up front, so it does not
 * wait either.
 */
int qann_hip_search_async(qann_hip_index* index, const float* d_queries, uint64_t nq,
                          const qann_hip_search_params* params, int64_t* d_out_ids, float* d_out_dists, void* stream);

/*
 * qann_hip_search_async that allocates no device memory (no hipMalloc, no
 * hipFree, no allocating library call). No counterpart in the reference: it is
 * for callers that cannot allocate -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid it. The caller still passes the
 * queries, results and allowed words in device memory; the search's working
 * memory is one scratch array of qann_hip_noalloc_scratch_bytes() bytes (64 MiB)
 * that each GPU holds from library load until the process ends. Searches that
 * need more than it holds (too many queries at once, very large k * k_factor
 * or filters) are refused with QANN_EINVAL and a message giving the size
 * needed; the sums of large batches are taken in chunks that fit. IVF and
 * RefineIndex handles only (QANN_EINVAL for a FlatIndex); the index's GPU must
 * be current. Every call on one GPU shares that GPU's array, which keeps the
 * last call's contents: calls that overlap on one GPU race, so run them on one
 * stream or one after another. Results equal qann_hip_search_async's.
 */
int qann_hip_search_noalloc(qann_hip_index* index, const float* d_queries, uint64_t nq,
                            const qann_hip_search_params* params, int64_t* d_out_ids, float* d_out_dists, void* stream);
uint64_t qann_hip_noalloc_scratch_bytes(void);

/*
 * Nearest centroid of each of n rows (detail::assign_nearest without BLAS,
 * src/kmeans.cpp:34-100 / nearest_centroid): the smallest exact squared L2
 * distance, the lowest centroid index on ties. Used when vectors are added
 * to an IVF index (IVFIndex::add_batch_impl, src/ivf_index.cpp:124-125) and by
 * training. Host buffers; synchronous.
 */
int qann_hip_assign_nearest(const float* rows, uint64_t n, uint32_t dim, const float* centroids, uint32_t k,
                            uint32_t* out);

/*
 * The assign step of PQ codebook training (PQCodebook::train, src/pq.cpp:41-73,
 * runs detail::kmeans on each subspace; its assign_nearest without BLAS is
 * nearest_centroid, src/kmeans.cpp:102-114) for all m subspaces at once: for
 * each of n rows (dim floats) and each subspace s, the centroid of
 * centroids[s][0..K)[dim / m] at the smallest exact squared L2 distance of the
 * row's sub-vector s (l2_distance's order), the first on ties. out: n * m
 * codes, row-major. Host buffers; synchronous.
 */
int qann_hip_assign_subspaces(const float* rows, uint64_t n, uint32_t dim, uint32_t m, const float* centroids,
                              uint32_t K, uint8_t* out);

/*
 * The encoding half of IVFIndex::add_batch_impl (src/ivf_index.cpp:124-150)
 * for n prepared (and, for a rotated index, rotated) rows: each row's nearest
 * list (detail::assign_nearest, exact squared L2, lowest list on ties) into
 * out_lists, and for PQ / fast-scan storage the PQ code of its residual to the
 * list's residual centre (format 13) or centroid (PQCodebook::encode,
 * src/pq.cpp:87-163), for SQ storage its SQ code (SQ::encode, src/sq.cpp), into
 * out_codes (n * m, or n * SQ code bytes; untouched for float storage). desc
 * as for qann_hip_ivf_create (use_precomputed and cosine_ip are not read).
 * Host buffers; synchronous.
 */
int qann_hip_ivf_encode(const qann_hip_ivf_desc* desc, const float* rows, uint64_t n, uint32_t* out_lists,
                        uint8_t* out_codes);

/*
 * detail::fs_set (src/fast_scan.hpp:25-28) for nblocks fast-scan blocks of 32
 * positions: position p of block b holds the 4-bit codes (m bytes, row-major
 * in codes, n rows) of row row_at[b * 32 + p], or zero nibbles where that is
 * 0xFFFFFFFF. Writes nblocks * m * 16 bytes in the reference's block layout,
 * the bytes IVFIndex::add_batch_impl (:171-175) writes into a list's blocks.
 * Host buffers; synchronous.
 */
int qann_hip_fs_pack(const uint8_t* codes, uint64_t n, uint32_t m, const uint32_t* row_at, uint64_t nblocks,
                     uint8_t* out);

/*
 * The two integer stages of the fast-scan on their own, run by the same
 * device code the search kernels use, so that each can be checked against the
 * reference's functions directly. Host buffers; synchronous.
 *
 * qann_hip_fs_quantize_lut: detail::quantize_lut (src/fast_scan.cpp:22-46,
 * as build_quantize_lut :223-254 calls it) of n tables of m x 16 floats
 * (m <= 1024): m x 16 LUT bytes, the scale and the offset of each.
 *
 * qann_hip_fs_block_sums: detail::fast_scan_block (src/fast_scan.cpp:375-377,
 * whose arithmetic fast_scan_block_scalar :256-262 spells out; the sums
 * fast_scan_block_mask, :387-397, compares with its limit) of
 * nblocks code blocks in the reference's layout (m x 16 bytes each,
 * src/fast_scan.hpp:19-34), block b against LUT b (m x 16 bytes): 32 u16
 * sums per block.
 */
int qann_hip_fs_quantize_lut(const float* tables, uint64_t n, uint32_t m, uint8_t* lut, float* scale, float* offset);
int qann_hip_fs_block_sums(const uint8_t* luts, uint32_t m, const uint8_t* blocks, uint64_t nblocks, uint16_t* sums);

/*
 * Per-stage kernel timing, the device counterpart of the reference's
 * compile-time section timers (src/profile.hpp: coarse, table, scan, drain,
 * rerank). QANN_STAGE_SCAN is the fast-scan sums kernel, QANN_STAGE_REPLAY the
 * kernel replaying the TopK pushes over them. QANN_STAGE_FLAT is a FlatIndex
 * search by the BF16 GEMM (or, for fewer than 8 queries, by direct exact
 * scoring), QANN_STAGE_FLAT_FP32 one by the FP32 GEMM, queries the BF16 path
 * hands over included. QANN_STAGE_GROUP is the grouping of (query, probe)
 * pairs by list that the matrix-core form of the sums kernel runs first (its
 * QANN_STAGE_SCAN time then excludes it); no other form records it.
 * While enabled, every search records HIP events around each stage on the
 * stream the stage runs on (no synchronization: stages of concurrent streams
 * stay concurrent, and a stage's time includes what it waited for while others
 * ran); read waits for the recorded stages and returns the process-wide totals
 * in milliseconds and launch counts for QANN_STAGE_COUNT stages, in the order
 * of qann_hip_stage_name. Searches run while enabled are not captured as
 * graphs.
 */
enum {
    QANN_STAGE_ROTATE = 0,
    QANN_STAGE_COARSE = 1,
    QANN_STAGE_SELECT = 2,
    QANN_STAGE_TABLES = 3,
    QANN_STAGE_SCAN = 4,
    QANN_STAGE_RERANK = 5,
    QANN_STAGE_FLAT = 6,
    QANN_STAGE_REPLAY = 7,
    QANN_STAGE_FLAT_FP32 = 8,
    QANN_STAGE_GROUP = 9,
    QANN_STAGE_COUNT = 10,
};
int qann_hip_profile_enable(int on);
int qann_hip_profile_read(double* ms, uint64_t* launches, int reset);
const char* qann_hip_stage_name(int stage);

/* Device bytes the index holds. */
int qann_hip_index_bytes(qann_hip_index* index, uint64_t* bytes);

/*
 * Free and total bytes of the current GPU's memory (hipMemGetInfo). No
 * counterpart in the reference: for callers that admit work by free memory,
 * as a serving loop does.
 */
int qann_hip_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);

/*
 * An upper estimate of the device scratch one qann_hip_search of nq queries
 * with these parameters takes (it is held only during the call). No
 * counterpart in the reference: a serving loop sums it over the launches it
 * runs at once to stay within free memory.
 */
int qann_hip_search_scratch_bytes(qann_hip_index* index, uint64_t nq, const qann_hip_search_params* params,
                                  uint64_t* bytes);

void qann_hip_index_destroy(qann_hip_index* index);

#ifdef __cplusplus
}
#endif

#endif /* QANN_HIP_H */
