// The C ABI of include/qann_hip.h: device indexes in HBM and the search
// pipelines that run on them.
#include "../../../include/qann_hip.h"
#include "device_common.hpp"
#include "gemm_bf16.hpp"
#include "ivf_kernels.hpp"
#include "build_kernels.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    const int code = e == hipErrorOutOfMemory ? QANN_ENOMEM : QANN_EDEVICE;
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

void check_launch(const char* what) { check(hipGetLastError(), what); }

template <class F>
int guarded(F&& f) {
    try {
        f();
        return QANN_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return QANN_ENOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return QANN_ESTATE;
    }
}

void invalid(const std::string& m) { throw Error(QANN_EINVAL, m); }

// A bump allocator over one device buffer, with the stream its work runs on.
struct Scratch {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    char* base = nullptr;
    size_t cap = 0, used = 0;
    char* io = nullptr; // staging for the synchronous entry point's inputs and outputs
    size_t io_cap = 0;
    // searches captured as HIP graphs, by shape (queries padded to a power of
    // two, k, nprobe, k_factor), with the buffers they were captured over
    struct Graph {
        hipGraphExec_t exec = nullptr;
        char* base = nullptr;
        char* io = nullptr;
        // data generations of the index and of its base at capture: a setter
        // that replaces lists, ids or deleted flags makes the capture stale
        uint64_t gen = 0, base_gen = 0;
    };
    std::map<std::tuple<uint64_t, uint32_t, uint32_t, uint32_t>, Graph> graphs;
    std::map<std::tuple<uint64_t, uint32_t, uint32_t, uint32_t>, int> graph_seen;
    // a pipelined search's parts: one Scratch (own stream) per part, and the
    // events that order them (inputs ready; per part: scan done, part done)
    std::vector<Scratch*> parts;
    hipEvent_t ev_in = nullptr;
    std::vector<hipEvent_t> ev_scan, ev_done;
    bool pipe_warm = false; // a first pipelined search runs its parts one after another
    // held for a whole search on this Scratch: two host threads calling
    // qann_hip_search_async on one stream share it one after the other
    std::mutex mu;
    // base / cap are a buffer the Scratch does not own and may not grow (the
    // allocation-free search's): a plan that needs more is refused
    bool fixed = false;
    [[noreturn]] void too_small(size_t bytes) const {
        throw Error(QANN_EINVAL, "this search needs " + std::to_string(bytes) + " bytes of scratch; the allocation-free "
                                 "search holds " + std::to_string(cap) + " (search fewer queries at a time, or use "
                                 "qann_hip_search_async)");
    }

    void drop_graphs() {
        for (auto& kv : graphs)
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        graphs.clear();
    }

    // Frees the buffers, graphs, parts and events (and the stream when owned),
    // after the stream's work has finished.
    void release() {
        (void)hipStreamSynchronize(stream);
        for (Scratch* s : parts) {
            s->release();
            delete s;
        }
        parts.clear();
        if (ev_in) (void)hipEventDestroy(ev_in);
        for (hipEvent_t e : ev_scan) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_done) (void)hipEventDestroy(e);
        ev_in = nullptr;
        ev_scan.clear();
        ev_done.clear();
        drop_graphs();
        if (base) (void)hipFree(base);
        if (io) (void)hipFree(io);
        base = io = nullptr;
        cap = io_cap = 0;
        if (own_stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    char* reserve_io(size_t bytes) {
        if (bytes > io_cap) {
            if (io) check(hipFree(io), "hipFree");
            io = nullptr;
            io_cap = 0;
            check(hipMalloc(&io, bytes), "hipMalloc staging");
            io_cap = bytes;
        }
        return io;
    }

    void reserve(size_t bytes) {
        if (bytes <= cap) return;
        if (fixed) too_small(bytes);
        if (base) {
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            check(hipFree(base), "hipFree");
            base = nullptr;
            cap = 0;
        }
        const size_t want = std::max(bytes, cap * 3 / 2);
        check(hipMalloc(&base, want), "hipMalloc scratch");
        cap = want;
    }
    template <class T>
    T* take(size_t n) {
        const size_t bytes = (n * sizeof(T) + 255) & ~size_t{255};
        if (used + bytes > cap) throw Error(QANN_ESTATE, "scratch plan too small");
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        return p;
    }
};

size_t pad256(size_t b) { return (b + 255) & ~size_t{255}; }

// Stage timing (qann_hip_profile_*).
std::atomic<int> g_prof_on{0};
std::mutex g_prof_mu;
double g_prof_ms[QANN_STAGE_COUNT] = {};
// bytes of u16 sums one search keeps at once (queries are scanned in chunks within it)
uint64_t g_sums_budget = uint64_t{8} << 30;
uint64_t g_prof_n[QANN_STAGE_COUNT] = {};

// Stage timings recorded but not yet read: events around a stage's launches on
// its stream, resolved when the timings are read.
struct PendingStage {
    int stage;
    hipEvent_t a, b;
    bool own_b = true; // false: b starts a later pending stage, which destroys it
};
std::vector<PendingStage> g_prof_pending; // guarded by g_prof_mu

// Adds the pending timings to the totals (waits for their end events); the
// caller holds g_prof_mu.
void resolve_stages_locked() {
    for (const PendingStage& p : g_prof_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g_prof_ms[p.stage] += ms;
            g_prof_n[p.stage] += 1;
        }
        (void)hipEventDestroy(p.a);
        if (p.own_b) (void)hipEventDestroy(p.b);
    }
    g_prof_pending.clear();
}

// Times one stage when profiling is on: events around its launches on its
// stream, with no synchronization, so that stages of concurrent streams keep
// running concurrently while they are timed.
struct StageTimer {
    int stage;
    hipStream_t st;
    hipEvent_t a = nullptr;
    StageTimer(int s, hipStream_t stream) : stage(s), st(stream) {
        if (!g_prof_on.load()) return;
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventRecord(a, st), "hipEventRecord");
    }
    // An event recorded inside the stage (or null) for a first part timed as
    // stage `first`; the rest stays this stage's.
    hipEvent_t mark() {
        hipEvent_t m = nullptr;
        if (a) check(hipEventCreate(&m), "hipEventCreate");
        return m;
    }
    void split(int first, hipEvent_t m, bool recorded) {
        if (!m) return;
        if (!recorded) {
            (void)hipEventDestroy(m);
            return;
        }
        std::lock_guard<std::mutex> g(g_prof_mu);
        g_prof_pending.push_back(PendingStage{first, a, m, false});
        a = m;
    }
    void done() {
        if (!a) return;
        hipEvent_t b = nullptr;
        check(hipEventCreate(&b), "hipEventCreate");
        check(hipEventRecord(b, st), "hipEventRecord");
        std::lock_guard<std::mutex> g(g_prof_mu);
        g_prof_pending.push_back(PendingStage{stage, a, b});
        a = nullptr;
        if (g_prof_pending.size() >= 8192) resolve_stages_locked();
    }
    ~StageTimer() {
        if (a) (void)hipEventDestroy(a);
    }
};

} // namespace

struct qann_hip_index {
    int kind = 0; // 1 flat, 2 ivf, 3 refine
    int device = 0;
    uint32_t dim = 0, metric = 0;
    // IVF
    qd::DevIvf ivf{};
    std::vector<uint32_t> entry_of_slot_host; // host copy of ivf.entry_of_slot
    bool use_pre = false;
    std::vector<uint32_t> host_list_sizes;
    // blocks_desc_prefix[j]: blocks of the j largest lists together
    std::vector<uint64_t> blocks_desc_prefix{0};
    // ids
    uint64_t slots = 0, live = 0;
    int64_t* labels = nullptr;
    uint64_t* deleted = nullptr;
    // raw rows (flat, refine)
    float* rows = nullptr;
    float* row_sq = nullptr; // flat: squared norms (exact order), for the GEMM path
    float xnorm_max = 0.0f;  // flat: an upper bound on every row's norm
    // flat: BF16 copy of the rows (zero-padded to kp) and, on the device, the
    // largest rounding residual and BF16 norm; null below kFlatBf16MinRows
    uint16_t* rows_bf16 = nullptr;
    float* bf16_maxes = nullptr;
    int kp = 0;
    // cleared when most queries of a search needed the FP32 path
    std::atomic<bool> bf16_ok{true};
    // IVF: the centroids as split BF16 rows [hi | lo | hi] of split_kp each,
    // for the coarse scores by one BF16 GEMM
    uint16_t* cent_split = nullptr;
    int split_kp = 0;
    qann_hip_index* base = nullptr;
    // bumped by every setter that replaces device data a captured search reads
    std::atomic<uint64_t> gen{0};
    std::vector<void*> allocs;
    uint64_t bytes = 0;

    std::mutex mu;
    std::vector<Scratch*> free_scratch;         // for synchronous calls
    std::map<hipStream_t, Scratch*> stream_scratch; // for async calls, per caller stream

    template <class T>
    T* upload(const T* host, size_t n) {
        if (n == 0) return nullptr;
        T* d = alloc<T>(n);
        check(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        return d;
    }
    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        check(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 1)), "hipMalloc");
        allocs.push_back(p);
        bytes += n * sizeof(T);
        return static_cast<T*>(p);
    }
    void free_ptr(void* p) {
        if (!p) return;
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) {
            (void)hipFree(p);
            allocs.erase(it);
        }
    }
    ~qann_hip_index() {
        for (Scratch* s : free_scratch) {
            s->release();
            delete s;
        }
        for (auto& kv : stream_scratch) {
            kv.second->release();
            delete kv.second;
        }
        for (void* p : allocs) (void)hipFree(p);
    }
};

namespace {

// One fast-scan block from the reference's layout (m rows of 16 bytes,
// src/fast_scan.hpp:19-28) to the device's: word (subspace s, column g) at
// (s / 4) * 64 + g * 16 + (s % 4) * 4. dst holds m4 * 16 zeroed bytes.
void pack_fs_block(const uint8_t* src, uint8_t* dst, size_t m) {
    for (size_t sp = 0; sp < m; ++sp)
        for (size_t g = 0; g < 4; ++g) std::memcpy(dst + (sp / 4) * 64 + g * 16 + (sp % 4) * 4, src + sp * 16 + 4 * g, 4);
}

// Device buffers freed when a synchronous helper call ends.
struct TempBuffers {
    std::vector<void*> ptrs;
    template <class T>
    T* get(size_t n) {
        void* p = nullptr;
        check(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 4)), "hipMalloc");
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~TempBuffers() {
        (void)hipDeviceSynchronize();
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// Deleted flags (one byte per slot) as a bitmap.
std::vector<uint64_t> to_bitmap(const uint8_t* flags, uint64_t slots, uint64_t& set) {
    std::vector<uint64_t> w((slots + 63) / 64, 0);
    set = 0;
    if (!flags) return w;
    for (uint64_t s = 0; s < slots; ++s)
        if (flags[s]) {
            w[s >> 6] |= uint64_t{1} << (s & 63);
            ++set;
        }
    return w;
}

// The set slots of a bitmap, in any order (the selections that read them key
// every candidate by its slot). Writes at most cap.
__global__ void k_bitmap_slots(const uint64_t* words, uint64_t nwords, uint32_t* out, uint32_t* count, uint32_t cap) {
    const uint64_t w = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (w >= nwords) return;
    uint64_t bits = words[w];
    if (!bits) return;
    uint32_t pos = atomicAdd(count, static_cast<uint32_t>(__popcll(bits)));
    while (bits && pos < cap) {
        const int b = __builtin_ctzll(bits);
        bits &= bits - 1;
        out[pos++] = static_cast<uint32_t>(w * 64 + b);
    }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Compacts the filter bitmap into slot_list (cap entries).
void bitmap_slots(const uint64_t* d_words, uint64_t nwords, uint32_t* slot_list, uint32_t* cnt, uint32_t cap,
                  hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, cnt, 1u);
    hipLaunchKernelGGL(k_bitmap_slots, dim3(static_cast<unsigned>((nwords + 255) / 256)), dim3(256), 0, st, d_words,
                       nwords, slot_list, cnt, cap);
    check_launch("k_bitmap_slots");
}

__global__ void k_fill_results(int64_t* ids, float* dists, size_t n) {
    const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    ids[i] = -1;
    dists[i] = INFINITY;
}

__global__ void k_copy_results(const int64_t* src_ids, const float* src_d, int kp, int k, int nq, int64_t* ids,
                               float* dists) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * k) return;
    const int i = t / k, j = t % k;
    ids[t] = j < kp ? src_ids[static_cast<size_t>(i) * kp + j] : -1;
    dists[t] = j < kp ? src_d[static_cast<size_t>(i) * kp + j] : INFINITY;
}

void fill_empty(int64_t* d_ids, float* d_dists, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_results, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, d_ids, d_dists, n);
    check_launch("k_fill_results");
}

__global__ void k_gather_rows(const float* src, const uint32_t* idx, uint32_t cnt, uint32_t dim, float* dst) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= static_cast<uint64_t>(cnt) * dim) return;
    const uint64_t r = t / dim, e = t % dim;
    dst[t] = src[static_cast<uint64_t>(idx[r]) * dim + e];
}

__global__ void k_scatter_results(const int64_t* sids, const float* sd, const uint32_t* idx, uint32_t cnt, uint32_t k,
                                  int64_t* ids, float* dists) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= static_cast<uint64_t>(cnt) * k) return;
    const uint64_t r = t / k, j = t % k, o = static_cast<uint64_t>(idx[r]) * k + j;
    ids[o] = sids[t];
    dists[o] = sd[t];
}

// Fast-scan chunks of at least this many queries are scanned in the order of
// their nearest list (query_order), XCD by XCD.
constexpr uint64_t kOrderMinQueries = 1024;

// FlatIndex rows below this count keep to the FP32 GEMM path.
constexpr uint64_t kFlatBf16MinRows = 65536;

// FlatIndex::query_impl by the FP32 MFMA GEMM: scores of a chunk of queries
// against every row into D (chunk * n floats), then per query the exact k
// nearest from a provably sufficient shortlist.
void flat_fp32(qann_hip_index* h, const float* d_q, uint64_t nq, uint32_t k, const uint64_t* excl,
               const uint64_t* incl, int64_t* d_ids, float* d_dists, float* D, uint64_t chunk, hipStream_t st) {
    const uint64_t n = h->slots;
    const int ip = h->metric == QANN_METRIC_IP;
    for (uint64_t q0 = 0; q0 < nq; q0 += chunk) {
        const uint64_t nc = std::min(chunk, nq - q0);
        const float* qc = d_q + q0 * h->dim;
        qd::launch_coarse_gemm(qc, static_cast<int>(nc), h->rows, h->row_sq, static_cast<int>(n),
                               static_cast<int>(h->dim), ip, D, st, excl, incl);
        check_launch("k_coarse_gemm");
        qd::launch_flat_select(D, qc, static_cast<int>(nc), h->rows, static_cast<int>(n), static_cast<int>(h->dim), ip,
                               h->xnorm_max, static_cast<int>(k), excl, incl, h->labels, d_ids + q0 * k,
                               d_dists + q0 * k, st);
        check_launch("k_flat_select");
    }
}

uint64_t flat_fp32_chunk(uint64_t nq, uint64_t n) {
    return std::max<uint64_t>(1, std::min<uint64_t>(nq, (uint64_t{1} << 32) / (n * 4)));
}

// The BF16 plan of a flat search: candidates per query come from rows under
// a limit that a strided sample of the rows gives (about want * stride rows
// each). The stride keeps them within a quarter of the candidate buffer, and
// leaves the sample at least 4 * want tiles of 128 rows, whose minima bound
// the want-th score tightly.
struct FlatBf16Plan {
    uint32_t stride = 0, ns = 0;
    uint64_t chunk = 0, per_query = 0;
};

FlatBf16Plan flat_bf16_plan(const qann_hip_index* h, uint64_t nq, uint32_t want) {
    FlatBf16Plan p;
    const uint64_t w = std::max<uint32_t>(want, 1);
    const uint64_t by_cap = static_cast<uint64_t>(qd::kFlatCandCap) / (4 * w), by_tiles = h->slots / (512 * w);
    p.stride = static_cast<uint32_t>(std::clamp<uint64_t>(std::min(by_cap, by_tiles), 1, 64));
    p.ns = static_cast<uint32_t>((h->slots + p.stride - 1) / p.stride);
    p.per_query = pad256(static_cast<uint64_t>(h->kp) * 2) + 12 + static_cast<uint64_t>(qd::flat_sample_tiles(p.ns)) * 4 +
                  8 + static_cast<uint64_t>(qd::kFlatCandCap) * 8;
    p.chunk = std::max<uint64_t>(1, std::min<uint64_t>(nq, (uint64_t{1} << 30) / p.per_query));
    return p;
}

// FlatIndex::query_impl by the BF16 GEMM with exact rescoring (gemm_bf16.hpp).
// Returns the queries it could not finish, for flat_fp32_subset.
std::vector<uint32_t> flat_bf16(qann_hip_index* h, const float* d_q, uint64_t nq, uint32_t k, uint32_t want,
                                const uint64_t* excl, const uint64_t* incl, int64_t* d_ids, float* d_dists,
                                Scratch& sc) {
    hipStream_t st = sc.stream;
    const FlatBf16Plan pl = flat_bf16_plan(h, nq, want);
    const uint64_t ch = pl.chunk;
    sc.used = 0;
    const uint64_t ntile = static_cast<uint64_t>(qd::flat_sample_tiles(static_cast<int>(pl.ns)));
    sc.reserve(pad256(ch * h->kp * 2) + pad256(ch * 12) + pad256(ch * ntile * 4) + 2 * pad256(ch * 4) +
               2 * pad256(ch * qd::kFlatCandCap * 4) + pad256(nq * 4) + 256);
    uint16_t* qb = sc.take<uint16_t>(ch * h->kp);
    float* qinfo = sc.take<float>(ch * 3);
    float* Ds = sc.take<float>(ch * ntile);
    float* lim = sc.take<float>(ch);
    uint32_t* cnt = sc.take<uint32_t>(ch);
    float* cd = sc.take<float>(ch * qd::kFlatCandCap);
    uint32_t* ci = sc.take<uint32_t>(ch * qd::kFlatCandCap);
    uint32_t* ovf = sc.take<uint32_t>(nq);
    uint32_t* ovf_cnt = sc.take<uint32_t>(1);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ovf_cnt, 1u);
    qd::FlatBf16 f{};
    f.xb = h->rows_bf16;
    f.rows = h->rows;
    f.xsq = h->row_sq;
    f.maxes = h->bf16_maxes;
    f.xnorm_max = h->xnorm_max;
    f.n = static_cast<int>(h->slots);
    f.dim = static_cast<int>(h->dim);
    f.kp = h->kp;
    f.ip = h->metric == QANN_METRIC_IP;
    f.excl = excl;
    f.incl = incl;
    f.labels = h->labels;
    for (uint64_t q0 = 0; q0 < nq; q0 += ch) {
        const int nc = static_cast<int>(std::min(ch, nq - q0));
        const float* qc = d_q + q0 * h->dim;
        qd::launch_queries_bf16(qc, nc, f.dim, f.kp, qb, qinfo, st);
        check_launch("k_queries_bf16");
        qd::launch_flat_bf16(f, qc, qb, qinfo, nc, static_cast<int>(k), static_cast<int>(want),
                             static_cast<int>(pl.stride), static_cast<int>(pl.ns), Ds, lim, cnt, cd, ci,
                             d_ids + q0 * k, d_dists + q0 * k, ovf, ovf_cnt, static_cast<int>(q0), st);
        check_launch("k_flat_bf16");
    }
    uint32_t novf = 0;
    check(hipMemcpyAsync(&novf, ovf_cnt, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    std::vector<uint32_t> list(novf);
    if (novf == 0) return list;
    if (2ull * novf > nq) h->bf16_ok = false;
    check(hipMemcpyAsync(list.data(), ovf, novf * 4ull, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return list;
}

// flat_fp32 for the listed queries of a batch, gathered and scattered back.
void flat_fp32_subset(qann_hip_index* h, const float* d_q, const std::vector<uint32_t>& list, uint32_t k,
                      const uint64_t* excl, const uint64_t* incl, int64_t* d_ids, float* d_dists, Scratch& sc) {
    hipStream_t st = sc.stream;
    const uint32_t novf = static_cast<uint32_t>(list.size());
    const uint64_t fch = flat_fp32_chunk(novf, h->slots);
    sc.used = 0;
    sc.reserve(pad256(novf * 4ull) + pad256(novf * static_cast<uint64_t>(h->dim) * 4) + pad256(novf * k * 8ull) +
               pad256(novf * k * 4ull) +
               pad256(fch * h->slots * 4));
    uint32_t* d_list = sc.take<uint32_t>(novf);
    float* qg = sc.take<float>(novf * static_cast<uint64_t>(h->dim));
    int64_t* gids = sc.take<int64_t>(novf * static_cast<uint64_t>(k));
    float* gd = sc.take<float>(novf * static_cast<uint64_t>(k));
    float* D = sc.take<float>(fch * h->slots);
    check(hipMemcpyAsync(d_list, list.data(), novf * 4ull, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    const uint64_t ge = static_cast<uint64_t>(novf) * h->dim;
    hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>((ge + 255) / 256)), dim3(256), 0, st, d_q, d_list,
                       novf, h->dim, qg);
    check_launch("k_gather_rows");
    flat_fp32(h, qg, novf, k, excl, incl, gids, gd, D, fch, st);
    const uint64_t re = static_cast<uint64_t>(novf) * k;
    hipLaunchKernelGGL(k_scatter_results, dim3(static_cast<unsigned>((re + 255) / 256)), dim3(256), 0, st, gids, gd,
                       d_list, novf, k, d_ids, d_dists);
    check_launch("k_scatter_results");
    // list and the staging above must outlive the copies
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// The indexes i < n whose status is set, appended to list (in any order) and
// counted in *count.
__global__ void k_list_flagged(const int* status, uint32_t n, int* count, int* list) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && status[i]) list[atomicAdd(count, 1)] = static_cast<int>(i);
}

// The device search of an index for nq prepared queries already in HBM.
// A filtered IVF search selects the nearest M lists of each query and probes
// them in order until IVFIndex::scan's stopping rule holds. With d_more set,
// M is a first guess (twice the probes the rule asks for at least), and the
// queries that would need more are counted in d_more[0] and listed from
// d_more[1] on (nq ints); the caller searches those again with every list.
// Without it, M is every list at once.
// Coarse scores by the split BF16 product (hi.hi + hi.lo + lo.hi, one BF16
// MFMA GEMM over three times K) or by the FP32 MFMA GEMM. Both lead
// k_coarse_select to the same probes. The split product's proven error is
// about three times the FP32 GEMM's, so its exact shortlist is wider: where
// many lists crowd around a query (Zipf-sized clusters split over many lists)
// queries overflow into scoring every list exactly, and the FP32 GEMM is
// faster; elsewhere the BF16 GEMM is. QANN_AMD_COARSE=bf16 or =fp32 forces a
// form; otherwise the first search of a problem class times both (GEMM and
// select) and the class keeps the faster (QANN_AMD_TUNE_LOG=1 prints it).
// -1: not forced.
int coarse_forced(const qann_hip_index* ivfh) {
    if (!ivfh->cent_split) return 0;
    const char* e = std::getenv("QANN_AMD_COARSE");
    if (e && std::strcmp(e, "bf16") == 0) return 1;
    if (e && std::strcmp(e, "fp32") == 0) return 0;
    return -1;
}

struct CoarseKey {
    int nlist, dim, ip, M, nclass;
    bool operator<(const CoarseKey& o) const {
        return std::tie(nlist, dim, ip, M, nclass) < std::tie(o.nlist, o.dim, o.ip, o.M, o.nclass);
    }
};
std::mutex g_coarse_mu;
std::map<CoarseKey, int> g_coarse_form;

// The coarse scores, by the FP32 GEMM (split false) or the split BF16 GEMM
// (q3: the queries' split rows, scratch).
void coarse_scores(const qann_hip_index* ivfh, const float* qr, uint64_t nq, int ip, bool split, uint16_t* q3,
                   float* D, hipStream_t st) {
    const qd::DevIvf& ix = ivfh->ivf;
    if (split) {
        qd::launch_split_bf16(qr, nq, ix.dim, ivfh->split_kp, 0, q3, st);
        qd::launch_coarse_split(q3, static_cast<int>(nq), ivfh->cent_split, ix.nlist, ivfh->split_kp,
                                ix.centroid_sqnorms, ip, D, st);
        check_launch("k_gemm_bf16 (coarse)");
    } else {
        qd::launch_coarse_gemm(qr, static_cast<int>(nq), ix.centroids, ix.centroid_sqnorms, ix.nlist, ix.dim, ip, D,
                               st);
        check_launch("k_coarse_gemm");
    }
}

// The M nearest lists of each query from those scores (the bound of the form
// that wrote them).
void coarse_probes(const qann_hip_index* ivfh, const float* qr, uint64_t nq, int ip, uint64_t M, bool split,
                   float* D, uint16_t* pidx, float* pval, uint64_t* gkeys, hipStream_t st,
                   const qd::ProbeBaseOut& pb = qd::ProbeBaseOut{}) {
    const qd::DevIvf& ix = ivfh->ivf;
    qd::launch_coarse_select(D, qr, static_cast<int>(nq), ix.centroids, ix.cnorm_max, ix.nlist, ix.dim, ip,
                             static_cast<int>(M), pidx, pval, st, split ? ivfh->split_kp : 0, gkeys, pb);
    check_launch("k_coarse_select");
}

// The coarse form of this search: forced, known for its class, or timed now
// (both forms run on the real problem; their outputs are the same).
bool coarse_split(const qann_hip_index* ivfh, const float* qr, uint64_t nq, int ip, uint64_t M, uint16_t* q3,
                  float* D, uint16_t* pidx, float* pval, uint64_t* gkeys, hipStream_t st) {
    const int forced = coarse_forced(ivfh);
    if (forced >= 0) return forced == 1;
    const qd::DevIvf& ix = ivfh->ivf;
    int nclass = 0;
    while ((uint64_t{1} << nclass) < nq) ++nclass;
    const CoarseKey key{ix.nlist, ix.dim, ip, static_cast<int>(M), nclass};
    {
        std::lock_guard<std::mutex> g(g_coarse_mu);
        auto it = g_coarse_form.find(key);
        if (it != g_coarse_form.end()) return it->second == 1;
    }
    // the better of two runs of each form, alternated
    float t[2] = {1e30f, 1e30f};
    int form = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        for (int run = 0; run < 4; ++run) {
            const int f = run & 1;
            float ms = 0.0f;
            (void)hipEventRecord(e0, st);
            coarse_scores(ivfh, qr, nq, ip, f == 1, q3, D, st);
            coarse_probes(ivfh, qr, nq, ip, M, f == 1, D, pidx, pval, gkeys, st);
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                t[f] = std::min(t[f], ms);
        }
        form = t[1] < t[0] ? 1 : 0;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const char* log = std::getenv("QANN_AMD_TUNE_LOG");
    if (log && *log && *log != '0')
        std::fprintf(stderr, "qann_amd tune: coarse nq %llu nlist %d dim %d probes %llu -> %s (ms: %.3f fp32, %.3f bf16)\n",
                     static_cast<unsigned long long>(nq), ix.nlist, ix.dim, static_cast<unsigned long long>(M),
                     form ? "split bf16" : "fp32", t[0], t[1]);
    std::lock_guard<std::mutex> g(g_coarse_mu);
    g_coarse_form[key] = form;
    return form == 1;
}

// One search on sc's stream. scan_wait / scan_done (fast-scan lists only, may
// be null): the scan waits for scan_wait and records scan_done after its last
// launch, so that a pipelined search's parts scan one after another.
// A launch that carries several requests of one IVF / refine index, each with
// its own k, nprobe and filter: the per-query arrays (device memory) and the
// largest values over the launch, which size its buffers. The caller has
// ruled out what such a launch does not run (FlatIndex, empty filters, the
// direct scans) and filled pq.list_allowed.
struct MixPlan {
    qd::PerQuery pq;
    uint32_t k_max = 0, kp_max = 0;
    uint32_t M = 0; // probes selected per query
    bool any_filtered = false;
};

// The first k[i] entries of each query's candidate row (kl wide) into its result row.
__global__ void k_copy_results_pq(const int64_t* src_ids, const float* src_d, int kl, int kmax, int nq, const int* kq,
                                  const uint64_t* off, int64_t* ids, float* dists) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<uint64_t>(nq) * kmax) return;
    const int i = static_cast<int>(t / kmax), j = static_cast<int>(t % kmax);
    if (j >= kq[i]) return;
    ids[off[i] + j] = src_ids[static_cast<size_t>(i) * kl + j];
    dists[off[i] + j] = src_d[static_cast<size_t>(i) * kl + j];
}

void run_search(qann_hip_index* h, const float* d_q, uint64_t nq, const qann_hip_search_params& p,
                const uint64_t* d_filter, int64_t* d_ids, float* d_dists, Scratch& sc, int* d_more = nullptr,
                hipEvent_t scan_wait = nullptr, hipEvent_t scan_done = nullptr, const MixPlan* mix = nullptr) {
    hipStream_t st = sc.stream;
    const uint32_t k = mix ? mix->k_max : p.k;
    if (nq == 0 || k == 0) return;
    const bool filtered = mix ? mix->any_filtered : p.allowed_words != nullptr;
    if (mix && (h->kind == 1 || d_filter)) invalid("a launch of several requests runs on IVF and refine indexes only");
    // query_prepared: an empty filter returns empty rows (src/index.cpp:172)
    if (!mix && filtered && p.allowed_count == 0) {
        fill_empty(d_ids, d_dists, nq * k, st);
        return;
    }
    const int ip = h->metric == QANN_METRIC_IP;
    const uint64_t nwords = (h->slots + 63) / 64;

    const size_t qlds = qd::query_lds_bytes(static_cast<int>(h->dim)); // 0: kernels read the query from global memory
    if (h->kind == 1) { // FlatIndex::query_impl
        const uint64_t n = h->slots;
        const size_t fcap = qd::coarse_select_cap(static_cast<int>(k));
        const bool gemm = nq >= 8 && n > 0 && n <= 0x7FFFFFFFull && k <= 4096 && qlds + fcap * 8 <= qd::kLdsMax &&
                          (!filtered || p.allowed_count * 8 >= n);
        if (gemm) {
            const uint64_t* excl = filtered ? nullptr : (h->live < h->slots ? h->deleted : nullptr);
            const uint64_t* incl = filtered ? d_filter : nullptr;
            const uint64_t eligible = filtered ? p.allowed_count : h->live;
            const uint32_t want = static_cast<uint32_t>(std::min<uint64_t>(k, eligible));
            StageTimer t(QANN_STAGE_FLAT, st);
            if (want == 0) {
                fill_empty(d_ids, d_dists, nq * k, st);
                t.done();
            } else if (h->rows_bf16 && h->bf16_ok && want <= 256) {
                const std::vector<uint32_t> rest = flat_bf16(h, d_q, nq, k, want, excl, incl, d_ids, d_dists, sc);
                t.done();
                if (!rest.empty()) {
                    StageTimer t2(QANN_STAGE_FLAT_FP32, st);
                    flat_fp32_subset(h, d_q, rest, k, excl, incl, d_ids, d_dists, sc);
                    t2.done();
                }
            } else {
                t.done();
                StageTimer t2(QANN_STAGE_FLAT_FP32, st);
                const uint64_t chunk = flat_fp32_chunk(nq, n);
                sc.used = 0;
                sc.reserve(pad256(chunk * n * 4));
                float* D = sc.take<float>(chunk * n);
                flat_fp32(h, d_q, nq, k, excl, incl, d_ids, d_dists, D, chunk, st);
                t2.done();
            }
            return;
        }
        sc.used = 0;
        const bool f = filtered;
        const size_t selb = qd::block_select_bytes(static_cast<int>(k), false);
        const bool gsel = qlds + selb > qd::block_select_room();
        sc.reserve((f ? pad256(h->slots * 4) + 256 : 256) + (gsel ? pad256(nq * selb) : 0));
        unsigned char* gbuf = gsel ? sc.take<unsigned char>(nq * selb) : nullptr;
        uint32_t* slot_list = nullptr;
        uint32_t n_slots = static_cast<uint32_t>(h->slots);
        if (f) {
            slot_list = sc.take<uint32_t>(h->slots);
            uint32_t* cnt = sc.take<uint32_t>(1);
            n_slots = static_cast<uint32_t>(std::min<uint64_t>(p.allowed_count, h->slots));
            bitmap_slots(d_filter, nwords, slot_list, cnt, n_slots, st);
        }
        StageTimer t(QANN_STAGE_FLAT, st);
        qd::SlotScorer scorer{};
        scorer.rows = h->rows;
        qd::launch_block_select(d_q, static_cast<int>(nq), h->dim, ip, scorer, slot_list, n_slots,
                                f ? nullptr : h->deleted, 0, 0, static_cast<int>(k), h->labels, d_ids, d_dists, gbuf,
                                st);
        check_launch("k_block_select");
        t.done();
        return;
    }

    qann_hip_index* ivfh = h->kind == 3 ? h->base : h;
    const qd::DevIvf& ix = ivfh->ivf;
    const bool refine = h->kind == 3;
    uint32_t kp = k;
    if (refine) {
        if (h->live == 0) {
            fill_empty(d_ids, d_dists, nq * k, st);
            return;
        }
        const uint64_t cap = filtered ? p.allowed_count : h->live;
        kp = mix ? mix->kp_max : static_cast<uint32_t>(std::min<uint64_t>(static_cast<uint64_t>(k) * p.k_factor, cap));
        // RefineIndex::use_direct (src/refine_index.cpp:59-69): score the allowed rows directly
        if (!mix && filtered && p.allowed_count <= 6ull * ix.nlist) {
            sc.used = 0;
            const size_t selb = qd::block_select_bytes(static_cast<int>(k), h->labels != nullptr);
            const bool gsel = qlds + selb > qd::block_select_room();
            sc.reserve(pad256(p.allowed_count * 4) + 256 + (gsel ? pad256(nq * selb) : 0));
            unsigned char* gbuf = gsel ? sc.take<unsigned char>(nq * selb) : nullptr;
            uint32_t* slot_list = sc.take<uint32_t>(p.allowed_count);
            uint32_t* cnt = sc.take<uint32_t>(1);
            bitmap_slots(d_filter, nwords, slot_list, cnt, static_cast<uint32_t>(p.allowed_count), st);
            StageTimer t(QANN_STAGE_RERANK, st);
            qd::SlotScorer scorer{};
            scorer.rows = h->rows;
            qd::launch_block_select(d_q, static_cast<int>(nq), h->dim, ip, scorer, slot_list,
                                    static_cast<uint32_t>(p.allowed_count), nullptr, 1, 1, static_cast<int>(k),
                                    h->labels, d_ids, d_dists, gbuf, st);
            check_launch("k_block_select");
            t.done();
            return;
        }
    }
    if (mix && !refine) kp = mix->kp_max;
    if (kp == 0) {
        fill_empty(d_ids, d_dists, nq * k, st);
        return;
    }
    const bool flat_lists = ix.storage == QANN_IVF_SQ || ix.storage == QANN_IVF_FLOAT;
    // IVFIndex::direct_scan (src/ivf_index.cpp:248-256): without PQ, a filter of at
    // most 3 nlist slots is scored slot by slot, pushed in slot order
    if (!mix && !refine && flat_lists && filtered && p.allowed_count <= 3ull * ix.nlist) {
        sc.used = 0;
        const size_t selb = qd::block_select_bytes(static_cast<int>(k), h->labels != nullptr);
        const bool gsel = qlds + selb > qd::block_select_room();
        sc.reserve(pad256(p.allowed_count * 4) + 256 + (gsel ? pad256(nq * selb) : 0));
        unsigned char* gbuf = gsel ? sc.take<unsigned char>(nq * selb) : nullptr;
        uint32_t* slot_list = sc.take<uint32_t>(p.allowed_count);
        uint32_t* cnt = sc.take<uint32_t>(1);
        bitmap_slots(d_filter, nwords, slot_list, cnt, static_cast<uint32_t>(p.allowed_count), st);
        qd::SlotScorer scorer{};
        scorer.entry_of_slot = ix.entry_of_slot;
        if (ix.storage == QANN_IVF_SQ) {
            scorer.sq_codes = ix.codes;
            scorer.sq_vmin = ix.sq_vmin;
            scorer.sq_step = ix.sq_step;
            scorer.sq_bits = ix.sq_bits;
            scorer.vbytes = ix.vbytes;
            scorer.sq_chunked = ix.sq_chunked;
        } else {
            scorer.rows = reinterpret_cast<const float*>(ix.codes);
        }
        StageTimer t(QANN_STAGE_FLAT, st);
        qd::launch_block_select(d_q, static_cast<int>(nq), h->dim, ip, scorer, slot_list,
                                static_cast<uint32_t>(p.allowed_count), nullptr, 1, 1, static_cast<int>(k), h->labels,
                                d_ids, d_dists, gbuf, st);
        check_launch("k_block_select");
        t.done();
        return;
    }
    const uint32_t nlist = static_cast<uint32_t>(ix.nlist);
    const uint32_t nprobe = mix ? mix->M : p.nprobe;
    if (nprobe == 0 || nprobe > nlist) invalid("nprobe must be in [1, nlist]");
    const bool fs = ix.storage == QANN_IVF_FASTSCAN;
    const bool has_pq = fs || ix.storage == QANN_IVF_PQ;
    // SDC scores Cosine by L2 over residuals and never uses B (src/ivf_index.cpp:430, 448-449)
    const bool cos_pq = has_pq && ix.metric == QANN_METRIC_COSINE && ix.cosine_ip && !ix.sdc;
    const bool ip_pq = has_pq && (ix.metric == QANN_METRIC_IP || cos_pq);
    const bool use_pre = has_pq && !ip_pq && !ix.sdc && ix.B != nullptr;
    const uint32_t m = static_cast<uint32_t>(ix.m);
    const size_t tsz = static_cast<size_t>(m) * static_cast<size_t>(ix.K); // table floats
    // the TopK buffers go to HBM when they do not fit LDS
    // TopK buffers stay within the 64 KB budget: above it one replay wave per
    // CU in LDS ran slower than HBM buffers at more waves (k' = 4,000 on
    // 8,192 queries: 18.0 against 8.8 ms); the re-rank's and the selects' key
    // sorts gain from the device's whole LDS (k' = 10,000: 29.4 against 44.3 ms)
    const bool gtopk = qd::replay_lds_bytes(static_cast<int>(kp)) > std::min(qd::kLdsMax, qd::replay_room());
    const size_t rkeys = qd::rerank_keys(static_cast<int>(kp));
    const bool grerank = refine && qlds + rkeys * 8 > qd::rerank_room();
    // the replay gathers several probes' candidate blocks into one load batch
    // where a batch of 16 spans several probes: short lists and small TopK
    // buffers (1.2M x 100, lists of about 10 blocks, k' = 100: 9% faster;
    // 10M x 96, about 19 blocks a list, k' up to 1,000: 20% slower)
    const int gather_replay = kp <= 256 && ivfh->blocks_desc_prefix.back() <= 16ull * nlist ? 1 : 0;

    // IVFIndex::scan's filter parameters (src/ivf_index.cpp:462, 502, 510-514)
    int min_probes = static_cast<int>(nprobe), wanted = 0, block_filter = 0;
    if (filtered && !mix) {
        wanted = static_cast<int>(std::min<uint64_t>(kp, p.allowed_count));
        block_filter = p.allowed_count * 16 < ivfh->slots;
        if (p.filter_probe_scaling && p.allowed_count > 0) {
            const double want = std::ceil(0.75 * double(nprobe) * std::cbrt(double(ivfh->live) / double(p.allowed_count)));
            min_probes = static_cast<int>(std::clamp<size_t>(static_cast<size_t>(std::min(want, double(nlist))), nprobe, nlist));
        }
    }
    const uint32_t M_all = nlist;
    const uint32_t M = mix       ? mix->M
                       : !filtered ? nprobe
                       : d_more  ? std::min<uint32_t>(M_all, std::max<uint32_t>(64, 2 * static_cast<uint32_t>(min_probes)))
                                 : M_all;
    // sums of one query take at most the blocks of its M largest probed lists;
    // queries go through the two scan phases in chunks that bound that buffer
    const uint64_t wblocks = ivfh->blocks_desc_prefix[std::min<size_t>(M, ivfh->blocks_desc_prefix.size() - 1)];
    const uint64_t stride = std::max<uint64_t>(wblocks, 1) * qd::kFsBlock;
    const uint64_t vbytes = fs ? 2 : 4; // u16 sums or float distances per vector
    uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(nq, g_sums_budget / (stride * vbytes)));

    // scratch plan: what every query holds, then the sums of one chunk of queries
    const size_t d = ix.dim;
    size_t need = 0;
    need += ix.rotation ? pad256(nq * d * 4) : 0;
    need += pad256(nq * nlist * 4); // coarse distances
    const bool may_split = coarse_forced(ivfh) != 0;
    need += may_split ? pad256(nq * 3 * static_cast<size_t>(ivfh->split_kp) * 2) : 0;
    need += pad256(nq * M * 2) + 3 * pad256(nq * M * 4);
    const bool gsel_keys = qd::coarse_select_global(static_cast<int>(M), ix.dim);
    need += gsel_keys ? pad256(nq * qd::coarse_select_cap(static_cast<int>(M)) * 8) : 0;
    need += use_pre || (ip_pq && !fs) ? pad256(nq * tsz * 4) : 0;
    need += ip_pq && fs ? pad256(nq * m * 16) + 2 * pad256(nq * 4) : 0;
    need += filtered && !mix ? pad256(nlist * 4) : 0;
    need += 2 * pad256(nq * 4) + pad256(nq * M * 4);    // plan
    need += 3 * pad256(nq * M * 4);                     // per-probe scale, offset, min
    need += pad256(nq * kp * 4) + pad256(nq * kp * 8) + pad256(nq * 4);
    need += gtopk ? pad256(nq * 2 * kp * 4) + pad256(nq * 2 * kp * 8) : 0;
    need += grerank ? pad256(nq * rkeys * 8) : 0;
    need += fs ? pad256(nlist * 4) : 0; // the query order's histogram
    need += fs ? pad256(8 * 4) : 0;     // work-queue counters of the persistent sums kernel
    need += fs ? 3 * pad256((nlist + 1) * 4) : 0; // the matrix-core form's pair totals and offsets, unit offsets
    if (sc.fixed) {
        // a scratch of fixed size: chunks as large as what is left of it holds
        // (each query of a chunk: its sums or distances, block minimums, order entry)
        const uint64_t per_query = stride * vbytes + (stride / qd::kFsBlock) * vbytes + (fs ? 4 + 5ull * M + 256 : 0);
        const uint64_t left = sc.cap > need + 3 * 256 ? sc.cap - need - 3 * 256 : 0;
        chunk = std::min<uint64_t>(chunk, left / per_query);
        if (chunk == 0) sc.too_small(need + 3 * 256 + per_query);
    }
    need += pad256(chunk * stride * vbytes);                  // sums or distances
    need += pad256(chunk * (stride / qd::kFsBlock) * vbytes); // block minimums
    const bool order_queries = fs && chunk >= kOrderMinQueries;
    need += order_queries ? pad256(chunk * 4) : 0;
    // the matrix-core form's pairs grouped by list, their units and the grouping's counts
    need += fs ? pad256(chunk * M * 4) + pad256((chunk * M / 32 + nlist + 1) * 4) +
                     pad256(qd::pairs_ghist_words(static_cast<int>(nlist), static_cast<int>(chunk)) * 4)
               : 0;
    sc.used = 0;
    sc.reserve(need);
    const float* qr = d_q;
    if (ix.rotation) {
        float* r = sc.take<float>(nq * d);
        StageTimer t(QANN_STAGE_ROTATE, st);
        qd::launch_rotate(d_q, ix.rotation, static_cast<int>(nq), ix.dim, r, st);
        check_launch("k_rotate");
        t.done();
        qr = r;
    }
    float* D = sc.take<float>(nq * nlist);
    uint16_t* q3 = may_split ? sc.take<uint16_t>(nq * 3 * static_cast<size_t>(ivfh->split_kp)) : nullptr;
    uint16_t* pidx = sc.take<uint16_t>(nq * M);
    float* pval = sc.take<float>(nq * M);
    uint64_t* gkeys = gsel_keys ? sc.take<uint64_t>(nq * qd::coarse_select_cap(static_cast<int>(M))) : nullptr;
    const bool split = coarse_split(ivfh, qr, nq, ip, M, q3, D, pidx, pval, gkeys, st);
    StageTimer t_coarse(QANN_STAGE_COARSE, st);
    coarse_scores(ivfh, qr, nq, ip, split, q3, D, st);
    t_coarse.done();
    float* pA = sc.take<float>(nq * M);
    float* pbase = sc.take<float>(nq * M);
    StageTimer t_sel(QANN_STAGE_SELECT, st);
    // the probes, and each probe's A and base for the scan
    coarse_probes(ivfh, qr, nq, ip, M, split, D, pidx, pval, gkeys, st, qd::probe_base_out(ix, pA, pbase));
    uint32_t* list_allowed = nullptr;
    if (filtered && !mix) {
        list_allowed = sc.take<uint32_t>(nlist);
        qd::launch_list_allowed(ix, d_filter, list_allowed, st);
        check_launch("k_list_allowed");
    }
    int* nprobe_q = sc.take<int>(nq);
    int* status = sc.take<int>(nq);
    uint32_t* item_blk = sc.take<uint32_t>(nq * M);
    qd::launch_probe_plan(ix, static_cast<int>(nq), static_cast<int>(M), static_cast<int>(nprobe), min_probes, wanted,
                          pidx, list_allowed, nprobe_q, item_blk, status, st, mix ? mix->pq : qd::PerQuery{});
    check_launch("k_probe_plan");
    t_sel.done();
    if (filtered && M < nlist && d_more) {
        // the queries whose stopping rule needs more lists than the first guess
        hipLaunchKernelGGL(k_list_flagged, dim3(static_cast<unsigned>((nq + 255) / 256)), dim3(256), 0, st, status,
                           static_cast<uint32_t>(nq), d_more, d_more + 1);
        check_launch("k_list_flagged");
    }
    float* C = nullptr;
    uint8_t* lut = nullptr;
    float *lscale = nullptr, *loff = nullptr;
    // a fast-scan search's L2 tables are made chunk by chunk below, only for
    // the sums forms that read them (the matrix-core form builds each pair's
    // table itself from B and the PQ codebook)
    const bool defer_tables = fs && use_pre;
    StageTimer t_tab(QANN_STAGE_TABLES, st);
    if (use_pre || (ip_pq && !fs)) {
        C = sc.take<float>(nq * tsz);
        if (!defer_tables) {
            qd::launch_query_tables(qr, ix, static_cast<int>(nq), use_pre ? qd::kTableInner : qd::kTableIp, C,
                                    nullptr, nullptr, nullptr, st);
            check_launch("k_query_tables");
        }
    } else if (ip_pq) {
        lut = sc.take<uint8_t>(nq * m * 16);
        lscale = sc.take<float>(nq);
        loff = sc.take<float>(nq);
        qd::launch_query_tables(qr, ix, static_cast<int>(nq), qd::kTableIpLut, nullptr, lut, lscale, loff, st);
        check_launch("k_query_tables");
    }
    t_tab.done();
    float* item_scale = sc.take<float>(nq * M);
    float* item_offset = sc.take<float>(nq * M);
    uint32_t* item_min = sc.take<uint32_t>(nq * M);
    uint16_t* sums = nullptr;
    uint16_t* bmin = nullptr;
    float* fdists = nullptr;
    uint32_t* fbmin = nullptr;
    if (fs) {
        sums = sc.take<uint16_t>(chunk * stride);
        bmin = sc.take<uint16_t>(chunk * (stride / qd::kFsBlock));
    } else {
        fdists = sc.take<float>(chunk * stride);
        fbmin = sc.take<uint32_t>(chunk * (stride / qd::kFsBlock));
    }
    uint32_t* sums_queue = fs ? sc.take<uint32_t>(8) : nullptr;
    qd::PairPlan pairs{};
    if (fs) {
        pairs.pair_off = sc.take<uint32_t>(nlist + 1);
        pairs.unit_off = sc.take<uint32_t>(nlist + 1);
        pairs.tot = sc.take<uint32_t>(nlist + 1);
        pairs.items = sc.take<uint32_t>(chunk * M);
        pairs.unit_list = sc.take<uint32_t>(chunk * M / 32 + nlist + 1);
        pairs.ghist = sc.take<uint32_t>(qd::pairs_ghist_words(static_cast<int>(nlist), static_cast<int>(chunk)));
    }
    uint32_t* ord_hist = order_queries ? sc.take<uint32_t>(nlist) : nullptr;
    uint32_t* order = order_queries ? sc.take<uint32_t>(chunk) : nullptr;
    float* cd = sc.take<float>(nq * kp);
    int64_t* cl = sc.take<int64_t>(nq * kp);
    int* ccount = sc.take<int>(nq);
    float* gd = gtopk ? sc.take<float>(nq * 2 * kp) : nullptr;
    int64_t* gl = gtopk ? sc.take<int64_t>(nq * 2 * kp) : nullptr;
    qd::DevIvf ixs = ix;
    if (!use_pre) ixs.B = nullptr;
    const size_t bb = static_cast<size_t>(m) * 16;
    const int dist_mode = ip_pq ? qd::kDistIp
                          : ix.sdc ? qd::kDistSdc
                          : use_pre ? qd::kDistL2Pre
                          : has_pq  ? qd::kDistL2Resid
                          : ix.storage == QANN_IVF_SQ ? qd::kDistSq
                                                      : qd::kDistFloat;
    for (uint64_t q0 = 0; q0 < nq && !fs; q0 += chunk) {
        // PQ (ADC), SQ and float lists: distances, then the replay over them
        const uint64_t nc = std::min<uint64_t>(chunk, nq - q0);
        qd::DistArgs a{};
        a.ix = ixs;
        a.M = static_cast<int>(M);
        a.mode = dist_mode;
        a.q = qr + q0 * d;
        a.probe_idx = pidx + q0 * M;
        a.probe_A = pA + q0 * M;
        a.probe_base = pbase + q0 * M;
        a.nprobe_q = nprobe_q + q0;
        a.item_blk = item_blk + q0 * M;
        a.C = C ? C + q0 * tsz : nullptr;
        a.filter = filtered ? d_filter : nullptr;
        if (mix) a.pq = mix->pq.at(q0);
        a.dists = fdists;
        a.stride_slots = stride;
        a.bmin = fbmin;
        a.item_min = item_min + q0 * M;
        if (qd::ivf_dists_lds(a) == 0) throw Error(QANN_EINVAL, "the PQ tables of this index do not fit in LDS");
        StageTimer t_scan(QANN_STAGE_SCAN, st);
        qd::launch_ivf_dists(a, static_cast<int>(nc), st);
        check_launch("k_ivf_dists");
        t_scan.done();
        qd::ReplayArgs r{};
        r.ix = ixs;
        r.kp = static_cast<int>(kp);
        r.M = static_cast<int>(M);
        r.probe_idx = a.probe_idx;
        r.nprobe_q = a.nprobe_q;
        r.item_blk = a.item_blk;
        r.fdists = fdists;
        r.stride_slots = stride;
        r.fbmin = fbmin;
        r.item_min = a.item_min;
        r.filter = filtered ? d_filter : nullptr;
        if (mix) r.pq = mix->pq.at(q0);
        r.out_dist = cd + q0 * kp;
        r.out_label = cl + q0 * kp;
        r.out_count = ccount + q0;
        r.unsorted = refine ? 1 : 0;
        r.g_dist = gd ? gd + q0 * 2 * kp : nullptr;
        r.g_label = gl ? gl + q0 * 2 * kp : nullptr;
        r.gather = gather_replay;
        StageTimer t_rep(QANN_STAGE_REPLAY, st);
        qd::launch_fs_replay(r, static_cast<int>(nc), st);
        check_launch("k_fl_replay");
        t_rep.done();
    }
    for (uint64_t q0 = 0; q0 < nq && fs; q0 += chunk) {
        const uint64_t nc = std::min<uint64_t>(chunk, nq - q0);
        qd::SumsArgs a{};
        a.ix = ixs;
        a.M = static_cast<int>(M);
        a.q = qr + q0 * d;
        a.probe_idx = pidx + q0 * M;
        a.probe_A = pA + q0 * M;
        a.probe_base = pbase + q0 * M;
        a.nprobe_q = nprobe_q + q0;
        a.item_blk = item_blk + q0 * M;
        a.C = C ? C + q0 * bb : nullptr;
        a.ip_lut = lut ? lut + q0 * bb : nullptr;
        a.ip_scale = lscale ? lscale + q0 : nullptr;
        a.ip_offset = loff ? loff + q0 : nullptr;
        a.sums = sums;
        a.stride_slots = stride;
        a.bmin = bmin;
        a.item_scale = item_scale + q0 * M;
        a.item_offset = item_offset + q0 * M;
        a.item_min = item_min + q0 * M;
        if (qd::fs_sums_lds(a) == 0) throw Error(QANN_EINVAL, "the fast-scan tables of this index do not fit in LDS");
        a.queue = sums_queue;
        a.pairs = pairs;
        if (q0 == 0 && scan_wait) check(hipStreamWaitEvent(st, scan_wait, 0), "hipStreamWaitEvent");
        if (order && nc >= kOrderMinQueries) {
            a.order = order;
            if (qd::fs_sums_reads_order(a, static_cast<int>(nc))) {
                qd::launch_query_order(a.probe_idx, static_cast<int>(M), static_cast<int>(nc), static_cast<int>(nlist),
                                       ord_hist, order, st);
                check_launch("k_order");
            }
        }
        if (defer_tables && qd::fs_sums_reads_tables(a, static_cast<int>(nc))) {
            StageTimer t_ctab(QANN_STAGE_TABLES, st);
            qd::launch_query_tables(a.q, ix, static_cast<int>(nc), qd::kTableInner, C + q0 * bb, nullptr, nullptr,
                                    nullptr, st);
            check_launch("k_query_tables");
            t_ctab.done();
        }
        StageTimer t_scan(QANN_STAGE_SCAN, st);
        a.grouped = t_scan.mark();
        const int form = qd::launch_fs_sums(a, static_cast<int>(nc), st);
        check_launch("k_fs_sums");
        t_scan.split(QANN_STAGE_GROUP, a.grouped, form == 2);
        t_scan.done();
        if (q0 + nc >= nq && scan_done) check(hipEventRecord(scan_done, st), "hipEventRecord");
        qd::ReplayArgs r{};
        r.ix = ixs;
        r.kp = static_cast<int>(kp);
        r.M = static_cast<int>(M);
        r.block_filter = block_filter;
        r.probe_idx = a.probe_idx;
        r.nprobe_q = a.nprobe_q;
        r.item_blk = a.item_blk;
        r.sums = sums;
        r.stride_slots = stride;
        r.bmin = bmin;
        r.item_scale = a.item_scale;
        r.item_offset = a.item_offset;
        r.item_min = a.item_min;
        r.filter = filtered ? d_filter : nullptr;
        if (mix) r.pq = mix->pq.at(q0);
        r.out_dist = cd + q0 * kp;
        r.out_label = cl + q0 * kp;
        r.out_count = ccount + q0;
        r.unsorted = refine ? 1 : 0;
        r.g_dist = gd ? gd + q0 * 2 * kp : nullptr;
        r.g_label = gl ? gl + q0 * 2 * kp : nullptr;
        r.gather = gather_replay;
        StageTimer t_rep(QANN_STAGE_REPLAY, st);
        qd::launch_fs_replay(r, static_cast<int>(nc), st);
        check_launch("k_fs_replay");
        t_rep.done();
    }
    if (refine) {
        uint64_t* gkeys = grerank ? sc.take<uint64_t>(nq * rkeys) : nullptr;
        StageTimer t_re(QANN_STAGE_RERANK, st);
        qd::launch_rerank(d_q, static_cast<int>(nq), h->dim, ip, h->rows, cl, ccount, static_cast<int>(kp),
                          static_cast<int>(k), h->labels, d_ids, d_dists, gkeys, st, mix ? mix->pq : qd::PerQuery{});
        check_launch("k_rerank");
        t_re.done();
    } else if (mix) {
        hipLaunchKernelGGL(k_copy_results_pq, dim3(static_cast<unsigned>((nq * k + 255) / 256)), dim3(256), 0, st, cl,
                           cd, static_cast<int>(kp), static_cast<int>(k), static_cast<int>(nq), mix->pq.k,
                           mix->pq.out_off, d_ids, d_dists);
        check_launch("k_copy_results_pq");
    } else {
        hipLaunchKernelGGL(k_copy_results, dim3(static_cast<unsigned>((nq * k + 255) / 256)), dim3(256), 0, st, cl, cd,
                           static_cast<int>(kp), static_cast<int>(k), static_cast<int>(nq), d_ids, d_dists);
        check_launch("k_copy_results");
    }
}

// Unfiltered fast-scan searches of at least kPipeMinQueries queries run as a
// pipeline of parts (at most kPipeParts, each at least kPipeMinPart queries),
// each part a whole search on a stream of its own. The parts' scans run one
// after another; a part's coarse quantizer, probe selection and tables overlap
// the previous part's scan, and its replay and re-rank overlap the next one's.
// Above kGraphMaxQueries, so graph capture never meets a pipelined search.
constexpr uint64_t kPipeMinQueries = 8192;
constexpr uint64_t kPipeMinPart = 2048;
constexpr int kPipeParts = 4;

// Parts for a search (1: not pipelined). QANN_AMD_PIPE=n forces at most n
// parts for searches the pipeline applies to (1 turns it off).
int pipe_parts(const qann_hip_index* h, uint64_t nq, const qann_hip_search_params& p) {
    if (p.allowed_words || h->kind == 1) return 1;
    const qann_hip_index* ivfh = h->kind == 3 ? h->base : h;
    if (!ivfh || ivfh->ivf.storage != QANN_IVF_FASTSCAN) return 1;
    int parts = kPipeParts;
    if (const char* e = std::getenv("QANN_AMD_PIPE")) {
        if (*e) parts = std::clamp(std::atoi(e), 1, 64);
    } else {
        return 1;
    }
    if (nq < kPipeMinQueries) return 1;
    return static_cast<int>(std::clamp<uint64_t>(nq / kPipeMinPart, 1, static_cast<uint64_t>(parts)));
}

// A search as a pipeline of parts on the parts' streams; the results are in
// d_ids / d_dists once sc's stream reaches its end.
void run_pipelined(qann_hip_index* h, const float* d_q, uint64_t nq, const qann_hip_search_params& p, int parts,
                   int64_t* d_ids, float* d_dists, Scratch& sc) {
    while (static_cast<int>(sc.parts.size()) < parts) {
        auto* s = new Scratch();
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
            delete s;
            throw Error(QANN_EDEVICE, "hipStreamCreate failed");
        }
        s->own_stream = true;
        sc.parts.push_back(s);
        hipEvent_t a = nullptr, b = nullptr;
        check(hipEventCreateWithFlags(&a, hipEventDisableTiming), "hipEventCreate");
        sc.ev_scan.push_back(a);
        check(hipEventCreateWithFlags(&b, hipEventDisableTiming), "hipEventCreate");
        sc.ev_done.push_back(b);
    }
    if (!sc.ev_in) check(hipEventCreateWithFlags(&sc.ev_in, hipEventDisableTiming), "hipEventCreate");
    // the first pipelined search on this Scratch runs its parts in turn (the
    // launch autotuners time their candidates with the chip to themselves)
    const bool serial = !sc.pipe_warm;
    check(hipEventRecord(sc.ev_in, sc.stream), "hipEventRecord");
    const uint64_t per = (nq + parts - 1) / parts;
    const uint64_t d = h->dim, k = p.k;
    for (int i = 0; i < parts; ++i) {
        const uint64_t q0 = static_cast<uint64_t>(i) * per;
        if (q0 >= nq) break;
        const uint64_t nc = std::min(per, nq - q0);
        Scratch& ps = *sc.parts[i];
        check(hipStreamWaitEvent(ps.stream, i > 0 && serial ? sc.ev_done[i - 1] : sc.ev_in, 0), "hipStreamWaitEvent");
        run_search(h, d_q + q0 * d, nc, p, nullptr, d_ids + q0 * k, d_dists + q0 * k, ps, nullptr,
                   i > 0 ? sc.ev_scan[i - 1] : nullptr, sc.ev_scan[i]);
        check(hipEventRecord(sc.ev_done[i], ps.stream), "hipEventRecord");
        check(hipStreamWaitEvent(sc.stream, sc.ev_done[i], 0), "hipStreamWaitEvent");
    }
    sc.pipe_warm = true;
}

// run_search, or run_pipelined for the searches the pipeline applies to.
void search_any(qann_hip_index* h, const float* d_q, uint64_t nq, const qann_hip_search_params& p,
                const uint64_t* d_filter, int64_t* d_ids, float* d_dists, Scratch& sc, int* d_more = nullptr) {
    const int parts = pipe_parts(h, nq, p);
    if (parts > 1)
        run_pipelined(h, d_q, nq, p, parts, d_ids, d_dists, sc);
    else
        run_search(h, d_q, nq, p, d_filter, d_ids, d_dists, sc, d_more);
}

Scratch* checkout(qann_hip_index* h) {
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->free_scratch.empty()) {
        Scratch* s = h->free_scratch.back();
        h->free_scratch.pop_back();
        return s;
    }
    auto* s = new Scratch();
    check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    s->own_stream = true;
    return s;
}

void checkin(qann_hip_index* h, Scratch* s) {
    std::lock_guard<std::mutex> g(h->mu);
    h->free_scratch.push_back(s);
}

// Searches of at most this many queries are captured as HIP graphs.
constexpr uint64_t kGraphMaxQueries = 4096;

// QANN_AMD_GRAPHS=0 turns the capture off.
bool graphs_enabled() {
    const char* e = std::getenv("QANN_AMD_GRAPHS");
    return !(e && std::strcmp(e, "0") == 0);
}

constexpr size_t kGraphsPerScratch = 64;
constexpr int kGraphAfter = 3;

// run_search of an unfiltered IVF / refine batch through a HIP graph: a shape
// runs plainly a few times, is then captured, and later searches of it replay
// the capture -- as long as the scratch and staging buffers it was captured
// over are the ones the Scratch still holds.
void run_graphed(qann_hip_index* h, const float* dq, uint64_t nr, const qann_hip_search_params& p, int64_t* dids,
                 float* dd, Scratch& io) {
    const auto key = std::make_tuple(nr, p.k, p.nprobe, p.k_factor);
    const uint64_t gen = h->gen.load(), base_gen = h->base ? h->base->gen.load() : 0;
    auto it = io.graphs.find(key);
    if (it != io.graphs.end()) {
        if (it->second.base == io.base && it->second.io == io.io && it->second.gen == gen &&
            it->second.base_gen == base_gen) {
            check(hipGraphLaunch(it->second.exec, io.stream), "hipGraphLaunch");
            return;
        }
        // captured over buffers or index data since replaced: capture again
        (void)hipGraphExecDestroy(it->second.exec);
        io.graphs.erase(it);
        io.graph_seen[key] = kGraphAfter;
    }
    // a shape is captured once it has run plainly kGraphAfter times (the first
    // sizes the scratch and lets the launch autotuners settle; the wait keeps
    // one-off shapes from paying for captures); -1: never (cache full or failed)
    int& seen = io.graph_seen[key];
    if (seen == kGraphAfter && io.graphs.size() >= kGraphsPerScratch) seen = -1;
    if (seen != kGraphAfter) {
        if (seen >= 0) ++seen;
        run_search(h, dq, nr, p, nullptr, dids, dd, io);
        return;
    }
    seen = -1;
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    bool ok = hipStreamBeginCapture(io.stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        try {
            run_search(h, dq, nr, p, nullptr, dids, dd, io);
        } catch (...) {
            ok = false;
        }
        ok = hipStreamEndCapture(io.stream, &g) == hipSuccess && ok && g;
        ok = ok && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
    }
    (void)hipGetLastError();
    if (!ok) {
        // not capturable: this shape runs plainly from now on
        if (ex) (void)hipGraphExecDestroy(ex);
        run_search(h, dq, nr, p, nullptr, dids, dd, io);
        return;
    }
    io.graphs[key] = Scratch::Graph{ex, io.base, io.io, gen, base_gen};
    check(hipGraphLaunch(ex, io.stream), "hipGraphLaunch");
}

} // namespace

extern "C" {

const char* qann_hip_last_error(void) { return g_err.c_str(); }

int qann_hip_device_count(int* count) {
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

// An upper bound on the norms of n rows (the largest, from double sums, rounded up).
static float norm_bound(const float* rows, uint64_t n, uint32_t dim) {
    double mx = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (uint32_t t = 0; t < dim; ++t) acc += double(rows[i * dim + t]) * double(rows[i * dim + t]);
        mx = std::max(mx, acc);
    }
    return static_cast<float>(std::sqrt(mx) * (1.0 + 1e-6)) + 1e-30f;
}

int qann_hip_ivf_create(const qann_hip_ivf_desc* desc, qann_hip_index** out) {
    return guarded([&] {
        const qann_hip_ivf_desc& dsc = *desc;
        if (dsc.dim == 0) invalid("dim must be >= 1");
        if (dsc.nlist < 100 || dsc.nlist > 65535) invalid("nlist must be in [100, 65535]");
        if (dsc.metric > 2) invalid("unknown metric");
        if (dsc.storage > QANN_IVF_FLOAT) invalid("unknown list storage");
        const bool pq = dsc.storage == QANN_IVF_FASTSCAN || dsc.storage == QANN_IVF_PQ;
        const uint32_t K = dsc.pq_k ? dsc.pq_k : 16;
        if (dsc.storage == QANN_IVF_FASTSCAN && (dsc.m == 0 || dsc.m % 2 != 0 || dsc.m > 1024 || dsc.dim % dsc.m != 0 || K != 16))
            invalid("fast_scan needs 16 centroids per subspace and an even num_subspaces <= 1024 dividing dim");
        if (dsc.storage == QANN_IVF_PQ && (dsc.m == 0 || dsc.dim % dsc.m != 0 || K > 256))
            invalid("PQ needs num_subspaces dividing dim and at most 256 centroids per subspace");
        if (dsc.pq_sdc && dsc.storage == QANN_IVF_FASTSCAN) invalid("cannot use SDC with fast-scan");
        if (dsc.storage == QANN_IVF_SQ && ((dsc.sq_bits != 4 && dsc.sq_bits != 8) || !dsc.sq_vmin || !dsc.sq_step))
            invalid("SQ needs 4 or 8 bits and its ranges");
        if (!dsc.centroids || (pq && !dsc.pq_centroids)) invalid("centroids (and the PQ codebook with PQ) are required");
        auto h = std::make_unique<qann_hip_index>();
        check(hipGetDevice(&h->device), "hipGetDevice");
        h->kind = 2;
        h->dim = dsc.dim;
        h->metric = dsc.metric;
        qd::DevIvf& ix = h->ivf;
        ix.dim = static_cast<int>(dsc.dim);
        ix.nlist = static_cast<int>(dsc.nlist);
        ix.metric = static_cast<int>(dsc.metric);
        ix.storage = static_cast<int>(dsc.storage);
        ix.m = pq ? static_cast<int>(dsc.m) : 0;
        ix.m4 = (ix.m + 3) & ~3;
        ix.K = pq ? static_cast<int>(K) : 0;
        ix.sub = pq ? static_cast<int>(dsc.dim / dsc.m) : 0;
        ix.cosine_ip = dsc.cosine_ip ? 1 : 0;
        ix.sq_bits = dsc.storage == QANN_IVF_SQ ? static_cast<int>(dsc.sq_bits) : 0;
        ix.sq_chunked = ix.sq_bits == 8 && dsc.dim % 16 == 0 ? 1 : 0;
        ix.vbytes = dsc.storage == QANN_IVF_PQ     ? static_cast<int>(dsc.m)
                    : dsc.storage == QANN_IVF_SQ    ? static_cast<int>(dsc.sq_bits == 8 ? dsc.dim : (dsc.dim + 1) / 2)
                    : dsc.storage == QANN_IVF_FLOAT ? static_cast<int>(dsc.dim * 4)
                                                    : 0;
        const size_t d = dsc.dim, nl = dsc.nlist, m = static_cast<size_t>(ix.m), sub = static_cast<size_t>(ix.sub);
        hipStream_t st = nullptr;
        float* cent = h->upload(dsc.centroids, nl * d);
        ix.centroids = cent;
        float* csq = h->alloc<float>(nl);
        qd::launch_sqnorms(cent, static_cast<int>(nl), static_cast<int>(d), csq, st);
        ix.centroid_sqnorms = csq;
        ix.cnorm_max = norm_bound(dsc.centroids, nl, static_cast<uint32_t>(d));
        if (dsc.residual_centers) {
            float* rc = h->upload(dsc.residual_centers, nl * d);
            float* rsq = h->alloc<float>(nl);
            qd::launch_sqnorms(rc, static_cast<int>(nl), static_cast<int>(d), rsq, st);
            ix.residual_centers = rc;
            ix.residual_sqnorms = rsq;
        }
        if (pq) {
            float* pqc = h->upload(dsc.pq_centroids, m * K * sub);
            float* pqt = h->alloc<float>(m * K * sub);
            qd::launch_pq_transpose(pqc, static_cast<int>(m), static_cast<int>(K), static_cast<int>(sub), pqt, st);
            ix.pq_cent_t = pqt;
            if (dsc.pq_sdc && dsc.metric != QANN_METRIC_IP) {
                // SDC takes no precomputed tables (src/ivf_index.cpp:430)
                float* sdc = h->alloc<float>(m * K * K);
                qd::launch_sdc_table(pqt, static_cast<int>(m), static_cast<int>(sub), static_cast<int>(K), sdc, st);
                ix.sdc = sdc;
            } else if (dsc.use_precomputed) {
                float* B = h->alloc<float>(nl * m * K);
                const float* centers = ix.residual_centers ? ix.residual_centers : ix.centroids;
                qd::launch_list_term(centers, pqt, static_cast<int>(nl), static_cast<int>(d), static_cast<int>(m),
                                     static_cast<int>(sub), static_cast<int>(K), B, st);
                ix.B = B;
                h->use_pre = true;
            }
        }
        if (dsc.storage == QANN_IVF_SQ) {
            ix.sq_vmin = h->upload(dsc.sq_vmin, d);
            ix.sq_step = h->upload(dsc.sq_step, d);
        }
        if (dsc.rotation) ix.rotation = h->upload(dsc.rotation, d * d);
        h->split_kp = qd::bf16_kpad(static_cast<int>(d));
        h->cent_split = h->alloc<uint16_t>(nl * 3 * static_cast<size_t>(h->split_kp));
        qd::launch_split_bf16(cent, nl, static_cast<int>(d), h->split_kp, 1, h->cent_split, st);
        check_launch("index setup kernels");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        *out = h.release();
    });
}

int qann_hip_ivf_set_lists(qann_hip_index* h, const uint32_t* sizes, const uint32_t* slots, const uint8_t* codes) {
    return guarded([&] {
        if (!h || h->kind != 2) invalid("not an IVF index");
        qd::DevIvf& ix = h->ivf;
        const size_t nl = static_cast<size_t>(ix.nlist), m = static_cast<size_t>(ix.m);
        const size_t bb = m * 16, lb = static_cast<size_t>(ix.m4) * 16;
        h->gen.fetch_add(1);
        std::vector<uint32_t> off(nl + 1, 0);
        std::vector<uint64_t> coff(nl, 0), src_off(nl, 0);
        uint64_t total = 0, cbytes = 0, src_bytes = 0;
        for (size_t l = 0; l < nl; ++l) {
            off[l] = static_cast<uint32_t>(total);
            total += sizes[l];
            if (total > 0xFFFFFFFFull) invalid("an index holds at most 2^32 - 1 vectors");
            coff[l] = cbytes;
            src_off[l] = src_bytes;
            const uint64_t nblk = (sizes[l] + 31) / 32;
            cbytes += (nblk * lb + 255) & ~uint64_t{255};
            src_bytes += nblk * bb;
        }
        off[nl] = static_cast<uint32_t>(total);
        std::vector<uint8_t> packed;
        if (ix.storage == QANN_IVF_FASTSCAN) {
            // the device block layout: word (subspace s, column g) of a block at
            // (s / 4) * 64 + g * 16 + (s % 4) * 4; subspaces m .. m4 - 1 stay zero
            packed.assign(std::max<uint64_t>(cbytes, 1), 0);
            for (size_t l = 0; l < nl; ++l) {
                const uint64_t nblk = (sizes[l] + 31) / 32;
                for (uint64_t b = 0; b < nblk; ++b)
                    pack_fs_block(codes + src_off[l] + b * bb, packed.data() + coff[l] + b * lb, m);
            }
        } else if (ix.sq_chunked) {
            // 8-bit SQ rows of a multiple of 16 dimensions: each group of 32
            // entries chunk-major, chunk c of entry e at
            // (e / 32) * 32 * dim + c * 512 + (e % 32) * 16 (sq_row)
            const size_t vb = static_cast<size_t>(ix.vbytes);
            packed.assign(std::max<uint64_t>((total + 31) / 32 * 32 * vb, 4), 0);
            for (uint64_t e = 0; e < total; ++e)
                for (size_t c = 0; c < vb / 16; ++c)
                    std::memcpy(packed.data() + (e / 32) * 32 * vb + c * 512 + (e % 32) * 16, codes + e * vb + c * 16, 16);
            std::fill(coff.begin(), coff.end(), 0);
        } else {
            // PQ / SQ codes or float rows: vbytes per entry, in entry order
            const size_t vb = static_cast<size_t>(ix.vbytes);
            packed.assign(std::max<uint64_t>(total * vb, 4), 0);
            if (total) std::memcpy(packed.data(), codes, total * vb);
            std::fill(coff.begin(), coff.end(), 0);
        }        h->free_ptr(const_cast<uint32_t*>(ix.list_off));
        h->free_ptr(const_cast<uint64_t*>(ix.code_off));
        h->free_ptr(const_cast<uint32_t*>(ix.list_ids));
        h->free_ptr(const_cast<uint8_t*>(ix.codes));
        ix.list_off = h->upload(off.data(), nl + 1);
        ix.code_off = h->upload(coff.data(), nl);
        ix.list_ids = total ? h->upload(slots, total) : h->alloc<uint32_t>(1);
        ix.codes = h->upload(packed.data(), packed.size());
        h->free_ptr(const_cast<uint32_t*>(ix.entry_of_slot));
        ix.entry_of_slot = nullptr;
        if (ix.storage == QANN_IVF_SQ || ix.storage == QANN_IVF_FLOAT) {
            // slot -> entry, for scans over a filter's slots (IVFIndex::direct_scan, ivf_index.cpp:248-256)
            uint32_t top = 0;
            for (uint64_t e = 0; e < total; ++e)
                if (slots[e] != 0xFFFFFFFFu) top = std::max(top, slots[e] + 1);
            std::vector<uint32_t> eos(std::max<uint32_t>(top, 1), 0xFFFFFFFFu);
            for (uint64_t e = 0; e < total; ++e)
                if (slots[e] != 0xFFFFFFFFu) eos[slots[e]] = static_cast<uint32_t>(e);
            ix.entry_of_slot = h->upload(eos.data(), eos.size());
            h->entry_of_slot_host = std::move(eos);
        }
        h->host_list_sizes.assign(sizes, sizes + nl);
        std::vector<uint64_t> nb(nl);
        for (size_t l = 0; l < nl; ++l) nb[l] = (sizes[l] + 31) / 32;
        std::sort(nb.begin(), nb.end(), std::greater<uint64_t>());
        h->blocks_desc_prefix.assign(nl + 1, 0);
        for (size_t l = 0; l < nl; ++l) h->blocks_desc_prefix[l + 1] = h->blocks_desc_prefix[l] + nb[l];
    });
}

int qann_hip_flat_create(uint32_t dim, uint32_t metric, const float* rows, uint64_t slots, qann_hip_index** out) {
    return guarded([&] {
        if (dim == 0) invalid("dim must be >= 1");
        if (metric > 2) invalid("unknown metric");
        if (slots > 0xFFFFFFFFull) invalid("an index holds at most 2^32 - 1 vectors");
        auto h = std::make_unique<qann_hip_index>();
        check(hipGetDevice(&h->device), "hipGetDevice");
        h->kind = 1;
        h->dim = dim;
        h->metric = metric;
        h->rows = slots ? h->upload(rows, slots * dim) : h->alloc<float>(1);
        h->slots = slots;
        h->live = slots;
        if (slots) {
            h->row_sq = h->alloc<float>(slots);
            qd::launch_sqnorms(h->rows, static_cast<int>(std::min<uint64_t>(slots, 0x7FFFFFFF)), static_cast<int>(dim),
                               h->row_sq, nullptr);
            check_launch("k_sqnorms");
            h->xnorm_max = norm_bound(rows, slots, dim);
            // QANN_AMD_FLAT_FP32=1 keeps the index on the FP32 GEMM path;
            // QANN_AMD_FLAT_BF16_ROWS=<n> moves the BF16 path's row threshold
            const char* fp32_only = std::getenv("QANN_AMD_FLAT_FP32");
            const bool bf16 = !(fp32_only && *fp32_only && std::strcmp(fp32_only, "0") != 0);
            const char* min_rows = std::getenv("QANN_AMD_FLAT_BF16_ROWS");
            const uint64_t bf16_rows = min_rows && *min_rows ? std::strtoull(min_rows, nullptr, 10) : kFlatBf16MinRows;
            if (bf16 && slots >= bf16_rows &&
                qd::query_lds_bytes(static_cast<int>(dim)) + qd::kFlatCandCap * 8 <= qd::kLdsMax) {
                h->kp = qd::bf16_kpad(static_cast<int>(dim));
                h->rows_bf16 = h->alloc<uint16_t>(slots * static_cast<uint64_t>(h->kp));
                h->bf16_maxes = h->alloc<float>(2);
                check(hipMemset(h->bf16_maxes, 0, 8), "hipMemset");
                qd::launch_rows_bf16(h->rows, slots, static_cast<int>(dim), h->kp, h->rows_bf16, h->bf16_maxes, nullptr);
                check_launch("k_rows_bf16");
                check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            }
        }
        uint64_t set = 0;
        std::vector<uint64_t> bm = to_bitmap(nullptr, slots, set);
        h->deleted = h->upload(bm.data(), std::max<size_t>(bm.size(), 1));
        *out = h.release();
    });
}

int qann_hip_refine_create(qann_hip_index* base, const float* rows, uint64_t slots, qann_hip_index** out) {
    return guarded([&] {
        if (!base || base->kind != 2) invalid("the refine base must be an IVF index");
        auto h = std::make_unique<qann_hip_index>();
        check(hipGetDevice(&h->device), "hipGetDevice");
        h->kind = 3;
        h->dim = base->dim;
        h->metric = base->metric;
        h->base = base;
        h->rows = slots ? h->upload(rows, slots * static_cast<size_t>(base->dim)) : h->alloc<float>(1);
        h->slots = slots;
        h->live = slots;
        *out = h.release();
    });
}

int qann_hip_set_ids(qann_hip_index* h, uint64_t slots, const int64_t* labels, const uint8_t* deleted) {
    return guarded([&] {
        if (!h) invalid("null index");
        if (slots > 0xFFFFFFFFull) invalid("an index holds at most 2^32 - 1 vectors");
        h->gen.fetch_add(1);
        h->free_ptr(h->labels);
        h->free_ptr(h->deleted);
        h->labels = labels ? h->upload(labels, slots) : nullptr;
        uint64_t set = 0;
        std::vector<uint64_t> bm = to_bitmap(deleted, slots, set);
        h->deleted = h->upload(bm.data(), std::max<size_t>(bm.size(), 1));
        h->slots = slots;
        h->live = slots - set;
        if (h->kind == 2) {
            h->ivf.labels = h->labels;
            h->ivf.deleted = set ? h->deleted : nullptr;
            auto& eos = h->entry_of_slot_host;
            if (h->ivf.entry_of_slot && eos.size() < slots) {
                // every slot gets an entry (0xFFFFFFFF: in no list)
                eos.resize(slots, 0xFFFFFFFFu);
                h->free_ptr(const_cast<uint32_t*>(h->ivf.entry_of_slot));
                h->ivf.entry_of_slot = h->upload(eos.data(), eos.size());
            }
        }
    });
}

int qann_hip_search(qann_hip_index* h, const float* queries, uint64_t nq, const qann_hip_search_params* params,
                    int64_t* out_ids, float* out_dists) {
    return guarded([&] {
        if (!h || !params) invalid("null argument");
        const uint32_t k = params->k;
        if (nq == 0 || k == 0) return;
        if (params->allowed_words) {
            uint64_t set = 0;
            for (uint64_t w = 0; w < (h->slots + 63) / 64; ++w) set += static_cast<uint64_t>(__builtin_popcountll(params->allowed_words[w]));
            if (set != params->allowed_count) invalid("allowed_count does not match the allowed bitmap");
        }
        check(hipSetDevice(h->device), "hipSetDevice");
        Scratch* io = checkout(h);
        try {
            // unfiltered IVF / refine searches of up to kGraphMaxQueries queries whose shape
            // repeats run as HIP graphs (at most kGraphsPerScratch shapes per stream)
            const bool graph = graphs_enabled() && !params->allowed_words && h->kind != 1 && !g_prof_on.load() &&
                               nq <= kGraphMaxQueries;
            const uint64_t nr = nq;
            // inputs and outputs in a second buffer so run_search owns the scratch
            const size_t qb = pad256(nr * h->dim * 4), ob = pad256(nr * k * 8), db = pad256(nr * k * 4);
            const size_t fb = params->allowed_words ? pad256((h->slots + 63) / 64 * 8) : 0;
            // filtered: the count and list of queries to search again, their
            // gathered rows and their results
            const size_t mb = fb ? pad256((nr + 1) * 4) + qb + ob + db : 0;
            char* buf = io->reserve_io(qb + ob + db + fb + mb + 256);
            float* dq = reinterpret_cast<float*>(buf);
            int64_t* dids = reinterpret_cast<int64_t*>(buf + qb);
            float* dd = reinterpret_cast<float*>(buf + qb + ob);
            uint64_t* df = fb ? reinterpret_cast<uint64_t*>(buf + qb + ob + db) : nullptr;
            check(hipMemcpyAsync(dq, queries, nq * h->dim * 4, hipMemcpyHostToDevice, io->stream), "hipMemcpyAsync");
            if (graph) {
                run_graphed(h, dq, nr, *params, dids, dd, *io);
                check(hipMemcpyAsync(out_ids, dids, nq * k * 8, hipMemcpyDeviceToHost, io->stream), "hipMemcpyAsync");
                check(hipMemcpyAsync(out_dists, dd, nq * k * 4, hipMemcpyDeviceToHost, io->stream), "hipMemcpyAsync");
                check(hipStreamSynchronize(io->stream), "hipStreamSynchronize");
                checkin(h, io);
                return;
            }
            if (df)
                check(hipMemcpyAsync(df, params->allowed_words, (h->slots + 63) / 64 * 8, hipMemcpyHostToDevice,
                                     io->stream),
                      "hipMemcpyAsync");
            int* dmore = df ? reinterpret_cast<int*>(buf + qb + ob + db + fb) : nullptr;
            if (dmore) {
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, io->stream, reinterpret_cast<uint32_t*>(dmore), 1);
                check_launch("k_zero_u32");
            }
            search_any(h, dq, nq, *params, df, dids, dd, *io, dmore);
            int more = 0;
            if (dmore) check(hipMemcpyAsync(&more, dmore, 4, hipMemcpyDeviceToHost, io->stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(io->stream), "hipStreamSynchronize");
            if (more > 0) {
                // the queries whose stopping rule needs more lists than the first
                // guess: those alone again, with every list, results scattered back
                char* mbuf = buf + qb + ob + db + fb;
                const uint32_t* list = reinterpret_cast<const uint32_t*>(dmore + 1);
                float* gq = reinterpret_cast<float*>(mbuf + pad256((nr + 1) * 4));
                int64_t* gids = reinterpret_cast<int64_t*>(mbuf + pad256((nr + 1) * 4) + qb);
                float* gd = reinterpret_cast<float*>(mbuf + pad256((nr + 1) * 4) + qb + ob);
                const uint32_t cnt = static_cast<uint32_t>(more);
                const uint64_t ge = static_cast<uint64_t>(cnt) * h->dim;
                hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>((ge + 255) / 256)), dim3(256), 0,
                                   io->stream, dq, list, cnt, h->dim, gq);
                check_launch("k_gather_rows");
                run_search(h, gq, cnt, *params, df, gids, gd, *io);
                const uint64_t re = static_cast<uint64_t>(cnt) * k;
                hipLaunchKernelGGL(k_scatter_results, dim3(static_cast<unsigned>((re + 255) / 256)), dim3(256), 0,
                                   io->stream, gids, gd, list, cnt, k, dids, dd);
                check_launch("k_scatter_results");
            }
            check(hipMemcpyAsync(out_ids, dids, nq * k * 8, hipMemcpyDeviceToHost, io->stream), "hipMemcpyAsync");
            check(hipMemcpyAsync(out_dists, dd, nq * k * 4, hipMemcpyDeviceToHost, io->stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(io->stream), "hipStreamSynchronize");
        } catch (...) {
            (void)hipStreamSynchronize(io->stream);
            checkin(h, io);
            throw;
        }
        checkin(h, io);
    });
}

int qann_hip_search_requests(qann_hip_index* h, const float* queries, const qann_hip_request* reqs, uint32_t nreq,
                             uint32_t k_factor, uint32_t filter_probe_scaling, int64_t* out_ids, float* out_dists) {
    return guarded([&] {
        if (!h || (nreq && (!reqs || !queries || !out_ids || !out_dists))) invalid("null argument");
        const uint64_t dim = h->dim, nwords = (h->slots + 63) / 64;
        std::vector<uint64_t> qoff(nreq + 1, 0), ooff(nreq + 1, 0);
        for (uint32_t r = 0; r < nreq; ++r) {
            const qann_hip_request& q = reqs[r];
            if (q.allowed_words) {
                uint64_t set = 0;
                for (uint64_t w = 0; w < nwords; ++w) set += static_cast<uint64_t>(__builtin_popcountll(q.allowed_words[w]));
                if (set != q.allowed_count) invalid("allowed_count does not match the allowed bitmap");
            }
            qoff[r + 1] = qoff[r] + q.nq;
            ooff[r + 1] = ooff[r] + q.nq * q.k;
        }
        auto params_of = [&](const qann_hip_request& q) {
            qann_hip_search_params p{};
            p.k = q.k;
            p.nprobe = q.nprobe;
            p.k_factor = k_factor;
            p.filter_probe_scaling = filter_probe_scaling;
            p.allowed_words = q.allowed_words;
            p.allowed_count = q.allowed_count;
            return p;
        };
        auto alone = [&](uint32_t r) {
            const qann_hip_search_params p = params_of(reqs[r]);
            const int rc = qann_hip_search(h, queries + qoff[r] * dim, reqs[r].nq, &p, out_ids + ooff[r],
                                           out_dists + ooff[r]);
            if (rc != QANN_OK) throw Error(rc, g_err);
        };
        if (h->kind == 1) { // FlatIndex: no probes to share; each request on its own
            for (uint32_t r = 0; r < nreq; ++r) alone(r);
            return;
        }
        qann_hip_index* ivfh = h->kind == 3 ? h->base : h;
        const qd::DevIvf& ix = ivfh->ivf;
        const uint32_t nlist = static_cast<uint32_t>(ix.nlist);
        const bool refine = h->kind == 3;
        const bool flat_lists = ix.storage == QANN_IVF_SQ || ix.storage == QANN_IVF_FLOAT;

        // which requests share the launch, and each query's parameters
        std::vector<uint32_t> mixed;
        std::vector<int> kp_q, k_q, nprobe_q, minp_q, wanted_q, bf_q;
        std::vector<uint64_t> off_q;
        std::vector<int> filt_q; // index of the query's filter among the distinct ones, -1 for none
        std::vector<const uint64_t*> filters;
        std::map<const uint64_t*, int> filter_index;
        uint64_t nm = 0, mixed_entries = 0;
        uint32_t k_max = 0, kp_max = 0, M = 0, minp_max = 0;
        bool any_filtered = false;
        for (uint32_t r = 0; r < nreq; ++r) {
            const qann_hip_request& q = reqs[r];
            if (q.nq == 0 || q.k == 0) continue;
            if (q.nprobe == 0 || q.nprobe > nlist) invalid("nprobe must be in [1, nlist]");
            const bool f = q.allowed_words != nullptr;
            const uint64_t cap = f ? q.allowed_count : h->live;
            const uint32_t kp = refine ? static_cast<uint32_t>(std::min<uint64_t>(static_cast<uint64_t>(q.k) * k_factor, cap))
                                       : q.k;
            // what the launch does not run: empty results, the direct scans
            if ((f && q.allowed_count == 0) || kp == 0 || (refine && f && q.allowed_count <= 6ull * nlist) ||
                (!refine && flat_lists && f && q.allowed_count <= 3ull * nlist)) {
                alone(r);
                continue;
            }
            int min_probes = static_cast<int>(q.nprobe), wanted = 0, block_filter = 0, fi = -1;
            if (f) {
                wanted = static_cast<int>(std::min<uint64_t>(kp, q.allowed_count));
                block_filter = q.allowed_count * 16 < ivfh->slots;
                if (filter_probe_scaling) {
                    const double want = std::ceil(0.75 * double(q.nprobe) * std::cbrt(double(ivfh->live) / double(q.allowed_count)));
                    min_probes = static_cast<int>(std::clamp<size_t>(static_cast<size_t>(std::min(want, double(nlist))), q.nprobe, nlist));
                }
                auto it = filter_index.find(q.allowed_words);
                if (it == filter_index.end()) {
                    it = filter_index.emplace(q.allowed_words, static_cast<int>(filters.size())).first;
                    filters.push_back(q.allowed_words);
                }
                fi = it->second;
                any_filtered = true;
                minp_max = std::max<uint32_t>(minp_max, static_cast<uint32_t>(min_probes));
            }
            mixed.push_back(r);
            for (uint64_t i = 0; i < q.nq; ++i) {
                kp_q.push_back(static_cast<int>(kp));
                k_q.push_back(static_cast<int>(q.k));
                off_q.push_back(mixed_entries + i * q.k);
                nprobe_q.push_back(static_cast<int>(q.nprobe));
                minp_q.push_back(min_probes);
                wanted_q.push_back(wanted);
                bf_q.push_back(block_filter);
                filt_q.push_back(fi);
            }
            nm += q.nq;
            mixed_entries += q.nq * q.k;
            k_max = std::max(k_max, q.k);
            kp_max = std::max(kp_max, kp);
            M = std::max(M, q.nprobe);
        }
        if (nm == 0) return;
        // probes selected per query: every request's nprobe, and for filtered
        // ones the first guess of how many their stopping rule takes (queries
        // it falls short for are searched again with every list)
        if (any_filtered) M = std::max(M, std::min<uint32_t>(nlist, std::max<uint32_t>(64, 2 * minp_max)));
        M = std::min(M, nlist);

        check(hipSetDevice(h->device), "hipSetDevice");
        Scratch* io = checkout(h);
        try {
            hipStream_t st = io->stream;
            const uint64_t F = filters.size();
            // staging: queries, results, filters and their list counts, the
            // per-query arrays (twice: the launch and the queries searched again)
            const size_t qb = pad256(nm * dim * 4), ob = pad256(mixed_entries * 8), db = pad256(mixed_entries * 4);
            const size_t fb = pad256(F * nwords * 8), lb = pad256(F * nlist * 4);
            const size_t ab = 7 * pad256(nm * 8) + 2 * pad256(nm * 8); // 7 arrays (<= 8 bytes) + 2 pointer arrays
            const size_t mb = pad256((nm + 1) * 4) + qb;
            char* buf = io->reserve_io(qb + ob + db + fb + lb + 2 * ab + mb + 256);
            char* at = buf;
            auto carve = [&](size_t bytes) {
                char* p = at;
                at += pad256(bytes);
                return p;
            };
            float* dq = reinterpret_cast<float*>(carve(nm * dim * 4));
            int64_t* dids = reinterpret_cast<int64_t*>(carve(mixed_entries * 8));
            float* dd = reinterpret_cast<float*>(carve(mixed_entries * 4));
            uint64_t* dfilt = reinterpret_cast<uint64_t*>(carve(F * nwords * 8));
            uint32_t* dla = reinterpret_cast<uint32_t*>(carve(F * nlist * 4));
            int* dmore = reinterpret_cast<int*>(carve((nm + 1) * 4));
            float* gq = reinterpret_cast<float*>(carve(nm * dim * 4));
            // the queries of the shared requests, in order
            std::vector<float> hq(nm * dim);
            {
                uint64_t row = 0;
                for (uint32_t r : mixed) {
                    std::memcpy(hq.data() + row * dim, queries + qoff[r] * dim, reqs[r].nq * dim * 4);
                    row += reqs[r].nq;
                }
            }
            check(hipMemcpyAsync(dq, hq.data(), nm * dim * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
            for (uint64_t f = 0; f < F; ++f) {
                check(hipMemcpyAsync(dfilt + f * nwords, filters[f], nwords * 8, hipMemcpyHostToDevice, st),
                      "hipMemcpyAsync");
                qd::launch_list_allowed(ix, dfilt + f * nwords, dla + f * nlist, st);
                check_launch("k_list_allowed");
            }
            // per-query arrays of a set of the shared queries (all, or those searched again)
            auto upload_plan = [&](const std::vector<uint32_t>& rows, MixPlan& mp) {
                const size_t n = rows.size();
                std::vector<int> a_kp(n), a_k(n), a_np(n), a_mp(n), a_w(n), a_bf(n);
                std::vector<uint64_t> a_off(n);
                std::vector<const uint64_t*> a_f(n);
                std::vector<const uint32_t*> a_la(n);
                for (size_t j = 0; j < n; ++j) {
                    const uint32_t i = rows[j];
                    a_kp[j] = kp_q[i];
                    a_k[j] = k_q[i];
                    a_off[j] = off_q[i];
                    a_np[j] = nprobe_q[i];
                    a_mp[j] = minp_q[i];
                    a_w[j] = wanted_q[i];
                    a_bf[j] = bf_q[i];
                    a_f[j] = filt_q[i] >= 0 ? dfilt + static_cast<uint64_t>(filt_q[i]) * nwords : nullptr;
                    a_la[j] = filt_q[i] >= 0 ? dla + static_cast<uint64_t>(filt_q[i]) * nlist : nullptr;
                }
                auto put = [&](const void* src, size_t bytes) {
                    void* dst = carve(bytes);
                    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
                    return dst;
                };
                mp.pq.kp = static_cast<const int*>(put(a_kp.data(), n * 4));
                mp.pq.k = static_cast<const int*>(put(a_k.data(), n * 4));
                mp.pq.out_off = static_cast<const uint64_t*>(put(a_off.data(), n * 8));
                mp.pq.nprobe = static_cast<const int*>(put(a_np.data(), n * 4));
                mp.pq.min_probes = static_cast<const int*>(put(a_mp.data(), n * 4));
                mp.pq.wanted = static_cast<const int*>(put(a_w.data(), n * 4));
                mp.pq.block_filter = static_cast<const int*>(put(a_bf.data(), n * 4));
                mp.pq.filter = static_cast<const uint64_t* const*>(put(a_f.data(), n * 8));
                mp.pq.list_allowed = static_cast<const uint32_t* const*>(put(a_la.data(), n * 8));
                mp.k_max = k_max;
                mp.kp_max = kp_max;
                mp.any_filtered = any_filtered;
                // the host arrays are read by the copies: wait for them before they go
                check(hipStreamSynchronize(st), "hipStreamSynchronize");
            };
            std::vector<uint32_t> all(nm);
            std::iota(all.begin(), all.end(), 0u);
            MixPlan mp;
            upload_plan(all, mp);
            mp.M = M;
            qann_hip_search_params p{};
            p.k = k_max;
            p.nprobe = M;
            p.k_factor = k_factor;
            p.filter_probe_scaling = filter_probe_scaling;
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(dmore), 1);
            check_launch("k_zero_u32");
            run_search(h, dq, nm, p, nullptr, dids, dd, *io, any_filtered ? dmore : nullptr, nullptr, nullptr, &mp);
            int more = 0;
            if (any_filtered) check(hipMemcpyAsync(&more, dmore, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            if (more > 0) {
                // the queries whose stopping rule needs more lists than the first
                // guess: those again with every list, their rows written in place
                std::vector<uint32_t> again(static_cast<size_t>(more));
                check(hipMemcpyAsync(again.data(), dmore + 1, again.size() * 4, hipMemcpyDeviceToHost, st),
                      "hipMemcpyAsync");
                check(hipStreamSynchronize(st), "hipStreamSynchronize");
                const uint32_t cnt = static_cast<uint32_t>(again.size());
                for (uint32_t i : again)
                    if (i >= nm) throw Error(QANN_EDEVICE, "a flagged query index is out of range");
                // their rows and parameters, in the order the device listed them
                hipLaunchKernelGGL(k_gather_rows, dim3(static_cast<unsigned>((uint64_t{cnt} * dim + 255) / 256)),
                                   dim3(256), 0, st, dq, reinterpret_cast<const uint32_t*>(dmore + 1), cnt,
                                   static_cast<uint32_t>(dim), gq);
                check_launch("k_gather_rows");
                MixPlan mp2;
                upload_plan(again, mp2);
                mp2.M = nlist;
                qann_hip_search_params p2 = p;
                p2.nprobe = nlist;
                run_search(h, gq, cnt, p2, nullptr, dids, dd, *io, nullptr, nullptr, nullptr, &mp2);
            }
            uint64_t moff = 0;
            for (uint32_t r : mixed) {
                const uint64_t e = reqs[r].nq * reqs[r].k;
                check(hipMemcpyAsync(out_ids + ooff[r], dids + moff, e * 8, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
                check(hipMemcpyAsync(out_dists + ooff[r], dd + moff, e * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
                moff += e;
            }
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
        } catch (...) {
            (void)hipStreamSynchronize(io->stream);
            checkin(h, io);
            throw;
        }
        checkin(h, io);
    });
}

int qann_hip_search_async(qann_hip_index* h, const float* d_queries, uint64_t nq, const qann_hip_search_params* params,
                          int64_t* d_out_ids, float* d_out_dists, void* stream) {
    return guarded([&] {
        if (!h || !params) invalid("null argument");
        hipStream_t st = static_cast<hipStream_t>(stream);
        Scratch* sc;
        {
            std::lock_guard<std::mutex> g(h->mu);
            auto it = h->stream_scratch.find(st);
            if (it == h->stream_scratch.end()) {
                sc = new Scratch();
                sc->stream = st;
                h->stream_scratch[st] = sc;
            } else {
                sc = it->second;
            }
        }
        std::lock_guard<std::mutex> g(sc->mu);
        search_any(h, d_queries, nq, *params, params->allowed_words, d_out_ids, d_out_dists, *sc);
    });
}

}  // extern "C"

// The allocation-free search's scratch. Each GPU the library runs on holds its
// own copy of it until the process ends; host code reaches a device's copy only
// through the address hipGetSymbolAddress gives for that device (the array's
// name is a host-side stand-in). Calls that overlap on one GPU share it, so
// they race: a caller runs them one at a time per GPU (one stream, or waits).
static constexpr size_t kNoAllocScratchBytes = size_t{64} << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_scratch[kMaxDevices] = {};

// The scratch's address on the current device, looked up once per device.
unsigned char* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(QANN_EDEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (!g_noalloc_scratch[device]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (!p) throw Error(QANN_EDEVICE, "hipGetSymbolAddress gave no address for the scratch");
        g_noalloc_scratch[device] = static_cast<unsigned char*>(p);
    }
    return g_noalloc_scratch[device];
}

} // namespace

extern "C" {

uint64_t qann_hip_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int qann_hip_search_noalloc(qann_hip_index* h, const float* d_queries, uint64_t nq,
                            const qann_hip_search_params* params, int64_t* d_out_ids, float* d_out_dists,
                            void* stream) {
    return guarded([&] {
        if (!h || !params) invalid("null argument");
        if (h->kind == 1) invalid("the allocation-free search takes IVF and RefineIndex handles");
        int cur = -1;
        check(hipGetDevice(&cur), "hipGetDevice");
        if (cur != h->device) invalid("the allocation-free search runs with the index's GPU current");
        Scratch sc;
        sc.stream = static_cast<hipStream_t>(stream);
        sc.base = reinterpret_cast<char*>(noalloc_scratch(cur));
        sc.cap = kNoAllocScratchBytes;
        sc.fixed = true;
        run_search(h, d_queries, nq, *params, params->allowed_words, d_out_ids, d_out_dists, sc);
    });
}

namespace {

// A non-blocking stream of one call, synchronized and destroyed at its end.
struct OwnStream {
    hipStream_t st = nullptr;
    OwnStream() { check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate"); }
    ~OwnStream() {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    OwnStream(const OwnStream&) = delete;
    OwnStream& operator=(const OwnStream&) = delete;
};

// The nearest of k centroids for blocks of rows on one stream: the FP32 MFMA
// GEMM of approximate scores, then the exact shortlist of k_coarse_select
// (exact squared L2 in l2_distance's order, the lowest index on ties). Owns
// its device buffers; rows in blocks so the block x k score matrix stays near
// 1 GiB.
struct NearestRunner {
    TempBuffers t;
    hipStream_t st = nullptr;
    uint32_t dim = 0, k = 0;
    uint64_t block = 0;
    float cmax = 0.0f;
    float *dc = nullptr, *csq = nullptr, *dq = nullptr, *D = nullptr, *val = nullptr;
    uint16_t* idx = nullptr;

    NearestRunner(const float* centroids, uint32_t k_, uint32_t dim_, uint64_t n) : dim(dim_), k(k_) {
        if (dim == 0 || k == 0) invalid("dim and k must be >= 1");
        if (k > 65535) invalid("at most 65535 centroids");
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        block = std::max<uint64_t>(1, std::min<uint64_t>(n, (uint64_t{1} << 28) / k));
        dc = t.get<float>(static_cast<size_t>(k) * dim);
        csq = t.get<float>(k);
        dq = t.get<float>(block * dim);
        D = t.get<float>(block * k);
        idx = t.get<uint16_t>(block);
        val = t.get<float>(block);
        check(hipMemcpyAsync(dc, centroids, static_cast<size_t>(k) * dim * 4, hipMemcpyHostToDevice, st),
              "hipMemcpyAsync");
        qd::launch_sqnorms(dc, static_cast<int>(k), static_cast<int>(dim), csq, st);
        check_launch("k_sqnorms");
        cmax = norm_bound(centroids, k, dim);
    }
    ~NearestRunner() {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    // b <= block rows of host memory: uploads them to dq and leaves each
    // row's nearest centroid in idx (device).
    void run(const float* rows, uint64_t b) {
        check(hipMemcpyAsync(dq, rows, b * dim * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        qd::launch_coarse_gemm(dq, static_cast<int>(b), dc, csq, static_cast<int>(k), static_cast<int>(dim), 0, D, st);
        check_launch("k_coarse_gemm");
        qd::launch_coarse_select(D, dq, static_cast<int>(b), dc, cmax, static_cast<int>(k), static_cast<int>(dim), 0, 1,
                                 idx, val, st);
        check_launch("k_coarse_select");
    }
};

} // namespace

int qann_hip_assign_nearest(const float* rows, uint64_t n, uint32_t dim, const float* centroids, uint32_t k,
                            uint32_t* out) {
    return guarded([&] {
        if (dim == 0 || k == 0) invalid("dim and k must be >= 1");
        if (k > 65535) invalid("at most 65535 centroids");
        if (n == 0) return;
        NearestRunner nr(centroids, k, dim, n);
        std::vector<uint16_t> hidx(nr.block);
        for (uint64_t b0 = 0; b0 < n; b0 += nr.block) {
            const uint64_t b = std::min(nr.block, n - b0);
            nr.run(rows + b0 * dim, b);
            check(hipMemcpyAsync(hidx.data(), nr.idx, b * 2, hipMemcpyDeviceToHost, nr.st), "hipMemcpyAsync");
            check(hipStreamSynchronize(nr.st), "hipStreamSynchronize");
            for (uint64_t i = 0; i < b; ++i) out[b0 + i] = hidx[i];
        }
    });
}

int qann_hip_assign_subspaces(const float* rows, uint64_t n, uint32_t dim, uint32_t m, const float* centroids,
                              uint32_t K, uint8_t* out) {
    return guarded([&] {
        if (dim == 0 || m == 0 || dim % m != 0) invalid("dim must be a positive multiple of m");
        if (K == 0 || K > 256) invalid("K must be in [1, 256]");
        if (n == 0) return;
        const uint32_t sub = dim / m;
        TempBuffers t;
        OwnStream os;
        hipStream_t st = os.st;
        const uint64_t block = std::max<uint64_t>(1, std::min<uint64_t>(n, (uint64_t{1} << 27) / dim));
        float* dc = t.get<float>(static_cast<size_t>(m) * K * sub);
        float* dx = t.get<float>(block * dim);
        uint8_t* dout = t.get<uint8_t>(block * m);
        check(hipMemcpyAsync(dc, centroids, static_cast<size_t>(m) * K * sub * 4, hipMemcpyHostToDevice, st),
              "hipMemcpyAsync");
        for (uint64_t b0 = 0; b0 < n; b0 += block) {
            const uint64_t b = std::min(block, n - b0);
            check(hipMemcpyAsync(dx, rows + b0 * dim, b * dim * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
            qd::launch_assign_subspaces(dx, b, static_cast<int>(dim), static_cast<int>(m), static_cast<int>(sub), dc,
                                        static_cast<int>(K), dout, st);
            check_launch("k_assign_subspaces");
            check(hipMemcpyAsync(out + b0 * m, dout, b * m, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        }
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
    });
}

int qann_hip_ivf_encode(const qann_hip_ivf_desc* desc, const float* rows, uint64_t n, uint32_t* out_lists,
                        uint8_t* out_codes) {
    return guarded([&] {
        if (!desc) invalid("null descriptor");
        const qann_hip_ivf_desc& d = *desc;
        if (d.dim == 0) invalid("dim must be >= 1");
        if (d.nlist == 0 || d.nlist > 65535) invalid("nlist must be in [1, 65535]");
        if (!d.centroids) invalid("no centroids");
        const bool pq = d.storage == QANN_IVF_FASTSCAN || d.storage == QANN_IVF_PQ;
        const bool sq = d.storage == QANN_IVF_SQ;
        if (d.storage > QANN_IVF_FLOAT) invalid("unknown storage");
        const uint32_t K = d.pq_k ? d.pq_k : 16;
        if (pq) {
            if (d.m == 0 || d.dim % d.m != 0) invalid("dim must be a positive multiple of m");
            if (K > 256) invalid("pq_k must be in [1, 256]");
            if (d.storage == QANN_IVF_FASTSCAN && K != 16) invalid("fast-scan needs 16 centroids per subspace");
            if (!d.pq_centroids) invalid("no PQ centroids");
        }
        if (sq && ((d.sq_bits != 4 && d.sq_bits != 8) || !d.sq_vmin || !d.sq_step)) invalid("SQ needs 4 or 8 bits, vmin and step");
        if (n == 0) return;
        NearestRunner nr(d.centroids, d.nlist, d.dim, n);
        const uint32_t sub = pq ? d.dim / d.m : 0;
        const uint64_t vbytes = pq ? d.m : sq ? (d.sq_bits == 8 ? d.dim : (d.dim + 1) / 2) : 0;
        float *centers = nullptr, *cent = nullptr, *cent_t = nullptr, *vmin = nullptr, *step = nullptr;
        uint8_t* dcodes = nullptr;
        if (pq) {
            // residuals to each list's residual centre where the index keeps them
            // (Cosine PQ), else to its centroid (IVFIndex::residual_center)
            centers = nr.dc;
            if (d.residual_centers) {
                centers = nr.t.get<float>(static_cast<size_t>(d.nlist) * d.dim);
                check(hipMemcpyAsync(centers, d.residual_centers, static_cast<size_t>(d.nlist) * d.dim * 4,
                                     hipMemcpyHostToDevice, nr.st),
                      "hipMemcpyAsync");
            }
            const size_t ncent = static_cast<size_t>(d.m) * K * sub;
            cent = nr.t.get<float>(ncent);
            cent_t = nr.t.get<float>(ncent);
            check(hipMemcpyAsync(cent, d.pq_centroids, ncent * 4, hipMemcpyHostToDevice, nr.st), "hipMemcpyAsync");
            qd::launch_pq_transpose(cent, static_cast<int>(d.m), static_cast<int>(K), static_cast<int>(sub), cent_t, nr.st);
            check_launch("k_pq_transpose");
        }
        if (sq) {
            vmin = nr.t.get<float>(d.dim);
            step = nr.t.get<float>(d.dim);
            check(hipMemcpyAsync(vmin, d.sq_vmin, d.dim * 4, hipMemcpyHostToDevice, nr.st), "hipMemcpyAsync");
            check(hipMemcpyAsync(step, d.sq_step, d.dim * 4, hipMemcpyHostToDevice, nr.st), "hipMemcpyAsync");
        }
        if (vbytes) dcodes = nr.t.get<uint8_t>(nr.block * vbytes);
        std::vector<uint16_t> hidx(nr.block);
        for (uint64_t b0 = 0; b0 < n; b0 += nr.block) {
            const uint64_t b = std::min(nr.block, n - b0);
            nr.run(rows + b0 * d.dim, b);
            if (pq) {
                qd::launch_pq_encode(nr.dq, b, static_cast<int>(d.dim), nr.idx, centers, cent_t, static_cast<int>(d.m),
                                     static_cast<int>(sub), static_cast<int>(K), dcodes, nr.st);
                check_launch("k_pq_encode");
            } else if (sq) {
                qd::launch_sq_encode(nr.dq, b, static_cast<int>(d.dim), vmin, step, static_cast<int>(d.sq_bits), dcodes,
                                     nr.st);
                check_launch("k_sq_encode");
            }
            check(hipMemcpyAsync(hidx.data(), nr.idx, b * 2, hipMemcpyDeviceToHost, nr.st), "hipMemcpyAsync");
            if (vbytes)
                check(hipMemcpyAsync(out_codes + b0 * vbytes, dcodes, b * vbytes, hipMemcpyDeviceToHost, nr.st),
                      "hipMemcpyAsync");
            check(hipStreamSynchronize(nr.st), "hipStreamSynchronize");
            for (uint64_t i = 0; i < b; ++i) out_lists[b0 + i] = hidx[i];
        }
    });
}

int qann_hip_fs_pack(const uint8_t* codes, uint64_t n, uint32_t m, const uint32_t* row_at, uint64_t nblocks,
                     uint8_t* out) {
    return guarded([&] {
        if (m == 0 || m > 1024) invalid("m must be in [1, 1024]");
        if (nblocks == 0) return;
        // every position names a row of codes or none: the kernel reads codes[row * m + s]
        for (uint64_t p = 0; p < nblocks * 32; ++p)
            if (row_at[p] != 0xFFFFFFFFu && row_at[p] >= n) invalid("row_at names a row beyond n");
        TempBuffers t;
        OwnStream os;
        hipStream_t st = os.st;
        const size_t bb = static_cast<size_t>(m) * 16;
        uint8_t* dc = t.get<uint8_t>(std::max<uint64_t>(n, 1) * m);
        uint32_t* dr = t.get<uint32_t>(nblocks * 32);
        uint8_t* dout = t.get<uint8_t>(nblocks * bb);
        if (n) check(hipMemcpyAsync(dc, codes, n * m, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        check(hipMemcpyAsync(dr, row_at, nblocks * 32 * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        qd::launch_fs_pack(dc, static_cast<int>(m), dr, nblocks, dout, st);
        check_launch("k_fs_pack");
        check(hipMemcpyAsync(out, dout, nblocks * bb, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
    });
}

int qann_hip_fs_quantize_lut(const float* tables, uint64_t n, uint32_t m, uint8_t* lut, float* scale, float* offset) {
    return guarded([&] {
        if (m == 0 || m > 1024) invalid("m must be in [1, 1024]");
        if (n == 0) return;
        if (n > 0x7FFFFFFFull) invalid("too many tables");
        TempBuffers t;
        const size_t tf = static_cast<size_t>(m) * 16;
        float* dt = t.get<float>(n * tf);
        float* nz = t.get<float>(tf);
        uint8_t* dl = t.get<uint8_t>(n * tf);
        float* ds = t.get<float>(n);
        float* dof = t.get<float>(n);
        check(hipMemcpy(dt, tables, n * tf * 4, hipMemcpyHostToDevice), "hipMemcpy");
        const std::vector<float> negzero(tf, -0.0f);
        check(hipMemcpy(nz, negzero.data(), tf * 4, hipMemcpyHostToDevice), "hipMemcpy");
        qd::DevIvf ix{};
        ix.m = static_cast<int>(m);
        ix.m4 = (ix.m + 3) & ~3;
        ix.B = dt;
        ix.centroids = dt;
        qd::launch_fs_lut_check(ix, static_cast<int>(n), nz, dl, ds, dof, nullptr);
        check_launch("k_fs_lut_check");
        check(hipMemcpy(lut, dl, n * tf, hipMemcpyDeviceToHost), "hipMemcpy");
        check(hipMemcpy(scale, ds, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        check(hipMemcpy(offset, dof, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int qann_hip_fs_block_sums(const uint8_t* luts, uint32_t m, const uint8_t* blocks, uint64_t nblocks, uint16_t* sums) {
    return guarded([&] {
        if (m == 0 || m > 1024) invalid("m must be in [1, 1024]");
        if (nblocks == 0) return;
        if (nblocks > 0x7FFFFFFFull) invalid("too many blocks");
        const size_t bb = static_cast<size_t>(m) * 16, m4 = (m + 3) & ~3u, lb = m4 * 16;
        std::vector<uint8_t> packed(nblocks * lb, 0);
        for (uint64_t b = 0; b < nblocks; ++b) pack_fs_block(blocks + b * bb, packed.data() + b * lb, m);
        TempBuffers t;
        uint8_t* dl = t.get<uint8_t>(nblocks * bb);
        uint8_t* db = t.get<uint8_t>(nblocks * lb);
        uint16_t* ds = t.get<uint16_t>(nblocks * qd::kFsBlock);
        check(hipMemcpy(dl, luts, nblocks * bb, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(db, packed.data(), nblocks * lb, hipMemcpyHostToDevice), "hipMemcpy");
        qd::launch_fs_block_check(dl, static_cast<int>(m), static_cast<int>(m4), db, static_cast<int>(nblocks), ds,
                                  nullptr);
        check_launch("k_fs_block_check");
        check(hipMemcpy(sums, ds, nblocks * qd::kFsBlock * 2, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int qann_hip_profile_enable(int on) {
    g_prof_on.store(on ? 1 : 0);
    return QANN_OK;
}

int qann_hip_profile_read(double* ms, uint64_t* launches, int reset) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    resolve_stages_locked();
    for (int i = 0; i < QANN_STAGE_COUNT; ++i) {
        if (ms) ms[i] = g_prof_ms[i];
        if (launches) launches[i] = g_prof_n[i];
        if (reset) {
            g_prof_ms[i] = 0.0;
            g_prof_n[i] = 0;
        }
    }
    return QANN_OK;
}

const char* qann_hip_stage_name(int stage) {
    static const char* names[QANN_STAGE_COUNT] = {"rotate", "coarse", "select", "tables", "scan", "rerank", "flat",
                                                  "replay", "flat_fp32", "group"};
    return stage >= 0 && stage < QANN_STAGE_COUNT ? names[stage] : "";
}

int qann_hip_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        size_t f = 0, t = 0;
        check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

int qann_hip_search_scratch_bytes(qann_hip_index* h, uint64_t nq, const qann_hip_search_params* p, uint64_t* bytes) {
    return guarded([&] {
        if (!h || !p || !bytes) invalid("null argument");
        const uint64_t k = p->k;
        uint64_t b = 0;
        if (h->kind == 1) {
            b = std::min<uint64_t>(nq * h->slots * 4, uint64_t{1} << 32) + nq * k * 16;
            if (h->rows_bf16 && k <= 256) {
                const FlatBf16Plan pl = flat_bf16_plan(h, nq, static_cast<uint32_t>(k));
                b = pl.chunk * pl.per_query + nq * (k * 16 + 4) + (uint64_t{1} << 20);
            }
        } else {
            const qann_hip_index* ivfh = h->kind == 3 ? h->base : h;
            const qd::DevIvf& ix = ivfh->ivf;
            const uint64_t nlist = static_cast<uint64_t>(ix.nlist);
            // filtered: the first guess of probes (a few times nprobe; the queries
            // needing more are searched again on their own)
            const uint64_t M = p->allowed_words ? std::min<uint64_t>(nlist, std::max<uint64_t>(256, 16ull * p->nprobe))
                                                : std::max<uint32_t>(p->nprobe, 1);
            const uint64_t kp = h->kind == 3 ? k * std::max<uint32_t>(p->k_factor, 1) : k;
            const uint64_t vb = ix.storage == QANN_IVF_FASTSCAN ? 2 : 4;
            const uint64_t wblocks = ivfh->blocks_desc_prefix[std::min<size_t>(M, ivfh->blocks_desc_prefix.size() - 1)];
            const uint64_t stride = std::max<uint64_t>(wblocks, 1) * qd::kFsBlock;
            // a pipelined search's parts hold a sums buffer each
            const uint64_t parts = static_cast<uint64_t>(pipe_parts(h, nq, *p));
            const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(nq, parts * (g_sums_budget / (stride * vb))));
            b = nq * nlist * 4 + nq * 3 * static_cast<uint64_t>(ivfh->split_kp) * 2 + nq * M * 34 +
                nq * static_cast<uint64_t>(ix.m) * std::max(ix.K, 16) * 4 +
                chunk * stride * vb * 33 / 32 + nq * kp * 24 + nq * ix.dim * 4 + (uint64_t{1} << 20);
        }
        *bytes = b;
    });
}

int qann_hip_index_bytes(qann_hip_index* h, uint64_t* bytes) {
    return guarded([&] { *bytes = h ? h->bytes : 0; });
}

void qann_hip_index_destroy(qann_hip_index* h) { delete h; }

} // extern "C"
