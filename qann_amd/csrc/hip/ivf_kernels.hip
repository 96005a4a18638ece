// IVF fast-scan search kernels for MI355X (gfx950).
//
// One batch of prepared queries runs as:
//   k_rotate          q * R for rotated indexes           (src/rotation.cpp:165-185)
//   k_coarse_gemm     approximate coarse scores for every (query, list), MFMA
//   k_coarse_select   the M nearest lists by exact A, sorted by (A, list)
//                     (IVFIndex::coarse / coarse_batch, src/ivf_index.cpp:305-419)
//   k_query_tables    C = -2<q_s, y_sj> (PQCodebook::compute_inner_table, src/pq.cpp:186-204),
//                     or the inner-product LUT quantized once per query (scan, :448-456)
//   k_probe_plan      probes each query scans (filters: min_probes / wanted) and sum offsets
//   k_order_*         the queries ordered by nearest list, for L2 sharing in k_fs_sums
//   k_fs_sums         per (query, window of 16 probes): the u8 LUTs in LDS
//                     (build_quantize_lut, src/fast_scan.cpp:243-254) and the u16 sums of
//                     every 4-bit code of those lists (fast_scan_block, :256-262) with
//                     v_perm lookups, plus block minima (k_fs_sums_queue: the same work
//                     pulled from per-XCD queues)
//   k_fs_replay_gather one wave per query: the reference's scan replayed push by push over
//                     those sums -- limit16, survivors in bit order, TopK trims
//                     (src/topk.hpp:26-138, scan_fast / push, src/ivf_index.cpp:466-543),
//                     candidate blocks of several probes loaded as one batch;
//                     k_fs_replay_win a probe's blocks at a time; k_fs_replay the same
//                     without LDS windows (large TopK buffers)
//   k_ivf_dists,      PQ (ADC) / SQ / float lists: float distances, then the same replay
//   k_fl_replay_win     over them (:552-611)
//   k_rerank          exact FP32 re-rank of the candidates (RefineIndex::rerank, src/refine_index.cpp:112-140)
//   k_block_select    exact top-k over a slot set (FlatIndex::query_impl, RefineIndex::direct_scan)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>
#include <type_traits>

#include "device_common.hpp"
#include "ivf_kernels.hpp"

typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef int v16i_t __attribute__((ext_vector_type(16)));
#include "select.hpp"

namespace qd {

// ---------------------------------------------------------------------------
// Setup kernels (run once at index creation).
// ---------------------------------------------------------------------------

// ||c||^2 per row with the reference's inner_product order.
__global__ void k_sqnorms(const float* rows, int n, int dim, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rows + static_cast<size_t>(i) * dim;
    out[i] = exact_avx2<true>(r, r, dim);
}

// centroids_t [s][sub][c] from the saved [s][c][sub] (PQCodebook::build_tables).
__global__ void k_pq_transpose(const float* cent, int m, int K, int sub, float* cent_t) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * K * sub) return;
    const int s = idx / (K * sub), rem = idx % (K * sub), c = rem / sub, j = rem % sub;
    cent_t[(static_cast<size_t>(s) * sub + j) * K + c] = cent[idx];
}

// sdc[s][k][l] = sum over the subspace's dimensions j, from the first, of
// fma(y_skj - y_slj, y_skj - y_slj, acc) (PQCodebook::build_tables' SDC table,
// src/pq.cpp:275-292). One thread per entry.
__global__ void k_sdc_table(const float* cent_t, int m, int sub, int K, float* sdc) {
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t kk = static_cast<size_t>(K) * K;
    if (idx >= static_cast<size_t>(m) * kk) return;
    const int s = static_cast<int>(idx / kk), k = static_cast<int>(idx % kk) / K, l = static_cast<int>(idx % K);
    const float* y = cent_t + static_cast<size_t>(s) * sub * K;
    float acc = 0.0f;
    for (int j = 0; j < sub; ++j) {
        const float d = y[static_cast<size_t>(j) * K + k] - y[static_cast<size_t>(j) * K + l];
        acc = fmaf(d, d, acc);
    }
    sdc[idx] = acc;
}

// B[l][s][c] = sum_j y_scj * (y_scj + 2 c_l,sj), fused (compute_list_term, src/pq.cpp:206-223).
__global__ void k_list_term(const float* centers, const float* cent_t, int nlist, int dim, int m, int sub, int K,
                            float* B) {
    const int tiles = (m * K + blockDim.x - 1) / blockDim.x;
    const int e = (blockIdx.x % tiles) * blockDim.x + threadIdx.x; // (s, c) entry
    const int l = blockIdx.x / tiles;
    if (e >= m * K || l >= nlist) return;
    const int s = e / K, c = e % K;
    const float* cl = centers + static_cast<size_t>(l) * dim + s * sub;
    float acc = 0.0f;
    for (int j = 0; j < sub; ++j) {
        const float y = cent_t[(static_cast<size_t>(s) * sub + j) * K + c];
        acc = fmaf(y, y + 2.0f * cl[j], acc);
    }
    B[static_cast<size_t>(l) * m * K + e] = acc;
}

// ---------------------------------------------------------------------------
// Query-side kernels.
// ---------------------------------------------------------------------------

// out[i][j] = fma chain over k of q[i][k] * R[k][j] (detail::rotate).
__global__ void k_rotate(const float* q, const float* R, int nq, int dim, float* out) {
    const size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (t >= static_cast<size_t>(nq) * dim) return;
    const int i = static_cast<int>(t / dim), j = static_cast<int>(t % dim);
    const float* x = q + static_cast<size_t>(i) * dim;
    float acc = 0.0f;
    for (int k = 0; k < dim; ++k) acc = fmaf(x[k], R[static_cast<size_t>(k) * dim + j], acc);
    out[static_cast<size_t>(i) * dim + j] = acc;
}

// ---------------------------------------------------------------------------
// Coarse quantizer (IVFIndex::coarse / coarse_all, src/ivf_index.cpp:305-383):
// the M nearest lists of each query by (exact A, list), A = exact_score(q, c)
// in the reference's AVX2 order. An FP32 MFMA GEMM scores every (query, list)
// approximately; k_coarse_select then recomputes A exactly only for the lists
// whose approximate score is within a proven rounding bound of the M-th
// smallest, a set that holds every one of the M nearest.
// ---------------------------------------------------------------------------

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kGemmTile = 64; // queries x lists per workgroup
constexpr int kGemmK = 16;    // K staged through LDS per step

// D[i][l] = |c_l|^2 - 2 q_i . c_l, or -q_i . c_l for inner product (+inf for
// a row excluded by the excl bitmap or missing from the incl bitmap), by
// v_mfma_f32_16x16x4f32: four waves of 32 x 32 (2 x 2 tiles of 16 x 16). Lane
// t gives A[t % 16][t / 16] and B[t / 16][t % 16] of a 16 x 4 x 16 step and
// holds D[4 (t / 16) + v][t % 16] in accumulator v.
__global__ void __launch_bounds__(256) k_coarse_gemm(const float* q, int nq, const float* cents, const float* csq,
                                                     int nlist, int dim, int ip, float* D, const uint64_t* excl,
                                                     const uint64_t* incl) {
    __shared__ float qs[kGemmTile][kGemmK + 1];
    __shared__ float cs[kGemmTile][kGemmK + 1];
    const int l0 = blockIdx.x * kGemmTile, q0 = blockIdx.y * kGemmTile;
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const int wq = (wave >> 1) * 32, wl = (wave & 1) * 32;
    v4f acc00 = {0.0f, 0.0f, 0.0f, 0.0f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    for (int k0 = 0; k0 < dim; k0 += kGemmK) {
        for (int e = tid; e < kGemmTile * kGemmK; e += blockDim.x) {
            const int r = e / kGemmK, k = e % kGemmK, gk = k0 + k;
            const int gq = q0 + r, gl = l0 + r;
            qs[r][k] = (gq < nq && gk < dim) ? q[static_cast<size_t>(gq) * dim + gk] : 0.0f;
            cs[r][k] = (gl < nlist && gk < dim) ? cents[static_cast<size_t>(gl) * dim + gk] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGemmK; kk += 4) {
            const int kr = kk + lane / 16, rr = lane % 16;
            const float a0 = qs[wq + rr][kr], a1 = qs[wq + 16 + rr][kr];
            const float b0 = cs[wl + rr][kr], b1 = cs[wl + 16 + rr][kr];
            acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc11, 0, 0, 0);
        }
        __syncthreads();
    }
    auto put = [&](const v4f& acc, int ti, int tj) {
        const int l = l0 + wl + 16 * tj + lane % 16;
        if (l >= nlist) return;
        const float c2 = ip ? 0.0f : csq[l];
        // a row the query may not return scores +inf
        const bool out = (excl && bit_test(excl, static_cast<uint32_t>(l))) || (incl && !bit_test(incl, static_cast<uint32_t>(l)));
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = q0 + wq + 16 * ti + 4 * (lane / 16) + v;
            if (i < nq) D[static_cast<size_t>(i) * nlist + l] = out ? INFINITY : ip ? -acc[v] : c2 - 2.0f * acc[v];
        }
    };
    put(acc00, 0, 0);
    put(acc01, 0, 1);
    put(acc10, 1, 0);
    put(acc11, 1, 1);
}

// The same scores by v_mfma_f32_32x32x2f32 over 128 x 128 tiles: four waves
// of 64 x 64 (2 x 2 tiles of 32 x 32), K staged through LDS 16 at a time in
// K-major order (a wave's 32 lanes read 32 consecutive rows of one k), the
// next step's tiles loaded into registers while the current one's MFMAs run.
// Lane t gives A[t % 32][t / 32] and B[t / 32][t % 32] of each 32 x 2 x 32
// step and holds D[8 (v / 4) + 4 (t / 32) + v % 4][t % 32] in accumulator v.
// The scores differ from k_coarse_gemm's only by summation order, which the
// select's bound on |A - approx - const| covers.
constexpr int kG2Tile = 128, kG2K = 16, kG2Pad = 4;
__global__ void __launch_bounds__(256) k_coarse_gemm32(const float* q, int nq, const float* cents, const float* csq,
                                                       int nlist, int dim, int ip, float* D, const uint64_t* excl,
                                                       const uint64_t* incl) {
    __shared__ float qs[2][kG2K][kG2Tile + kG2Pad];
    __shared__ float cs[2][kG2K][kG2Tile + kG2Pad];
    const int l0 = blockIdx.x * kG2Tile, q0 = blockIdx.y * kG2Tile;
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const int wq = (wave >> 1) * 64, wl = (wave & 1) * 64;
    // each thread stages 8 values of each operand per K step: element e =
    // tid + 256 j is row e / 16, k e % 16 (16 threads read 64 contiguous bytes)
    constexpr int kPer = kG2Tile * kG2K / 256;
    float pa[kPer], pb[kPer];
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int e = tid + 256 * j, r = e / kG2K, k = k0 + e % kG2K;
            const int gq = q0 + r, gl = l0 + r;
            pa[j] = (gq < nq && k < dim) ? q[static_cast<size_t>(gq) * dim + k] : 0.0f;
            pb[j] = (gl < nlist && k < dim) ? cents[static_cast<size_t>(gl) * dim + k] : 0.0f;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int e = tid + 256 * j, r = e / kG2K, k = e % kG2K;
            qs[buf][k][r] = pa[j];
            cs[buf][k][r] = pb[j];
        }
    };
    typedef float v16f __attribute__((ext_vector_type(16)));
    v16f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    const int nk = (dim + kG2K - 1) / kG2K;
    gload(0);
    lstore(0);
    __syncthreads();
    const int rr = lane % 32, kh = lane / 32;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * kG2K);
#pragma unroll
        for (int kk = 0; kk < kG2K; kk += 2) {
            const float a0 = qs[buf][kk + kh][wq + rr], a1 = qs[buf][kk + kh][wq + 32 + rr];
            const float b0 = cs[buf][kk + kh][wl + rr], b1 = cs[buf][kk + kh][wl + 32 + rr];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (kt + 1 < nk) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
        const int l = l0 + wl + 32 * tj + rr;
        if (l >= nlist) continue;
        const float c2 = ip ? 0.0f : csq[l];
        // a row the query may not return scores +inf
        const bool out = (excl && bit_test(excl, static_cast<uint32_t>(l))) ||
                         (incl && !bit_test(incl, static_cast<uint32_t>(l)));
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int i = q0 + wq + 32 * ti + 8 * (v / 4) + 4 * kh + (v % 4);
                if (i < nq)
                    D[static_cast<size_t>(i) * nlist + l] = out ? INFINITY : ip ? -acc[ti][tj][v] : c2 - 2.0f * acc[ti][tj][v];
            }
    }
}

// Per query (one workgroup): the M smallest (A, list), A exact, sorted. The
// candidates are the lists whose approximate score is at most tau + 2 delta,
// tau an approximate score with at least M at or below it and delta twice the
// bound on |A - approx - const| (exact side: (n + 3) u (|q| + |c|)^2; GEMM
// side with |c|^2 rounded: (n + 2) u (|q| + |c|)^2; inner product:
// 2 (n + 1) u |q| |c|; plus n 2^-125 for flushed denormals). tau + delta holds
// every list of the M nearest (the M lists at or below tau have A at most
// tau + const + delta / 2, so a list of the M nearest has an approximate score
// at most that + delta / 2); the second delta is margin. Before the exact
// scores the shortlist is narrowed the same way from the M-th smallest
// approximate score among the candidates. If more than `cap` lists qualify (or
// fewer than M, with NaN scores), every list is scored exactly and selected by
// the full 64-bit key.
// CD: values of the row each thread keeps in registers (the row, nlist <=
// CD * NT, is then read once; CD 0: read from memory at every pass). SAMPLE:
// tau comes first from a sample of at most kSelectSample scores of the row.
constexpr int kSelectSample = 4096;
template <int NT, int CD, bool SAMPLE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(SAMPLE ? 4 : 5))) k_coarse_select(float* D, const float* q, const float* cents, int nlist,
                                                       int dim, int ip, float cnorm_max, int M, int cap,
                                                       uint16_t* out_idx, float* out_val, int split_kp,
                                                       uint64_t* gkeys, const ProbeBaseOut pb) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_wsum[4], s_digit, s_below, s_count, s_count2;
    __shared__ float s_red[NT / kWave], s_red8[2 * NT / kWave];
    float* qsh = reinterpret_cast<float*>(lds);
    // the shortlist's keys: in LDS after the query, or (gkeys, when more are
    // wanted than LDS holds) cap keys per query in global memory, which the
    // block's barriers order as they order LDS
    uint64_t* keys = gkeys ? gkeys + static_cast<size_t>(blockIdx.x) * cap
                           : reinterpret_cast<uint64_t*>(lds + query_lds_bytes(dim));
    const int i = blockIdx.x, tid = threadIdx.x;
    float* row = D + static_cast<size_t>(i) * nlist;
    const bool qcopy = dim <= kQueryLdsDims;
    const float* qv = qcopy ? qsh : q + static_cast<size_t>(i) * dim;
    // the M probes of this query, by quads of lanes: each probe's list and
    // coarse distance, and with pb.A its A and base (the residual-centre term
    // by the quad's four lanes, the query already in LDS)
    auto put_probes = [&]() {
        const int g = tid & 3;
        for (int t = tid >> 2; t < M; t += NT / 4) {
            const uint64_t key = keys[t];
            const uint32_t l = key_index(key);
            const float v = key_dist(key);
            const size_t o = static_cast<size_t>(i) * M + t;
            float A = v;
            if (pb.A && pb.residual_centers)
                A = exact_score_quad(pb.ip != 0, qv, pb.residual_centers + static_cast<size_t>(l) * dim, dim, g);
            if (g == 0) {
                out_idx[o] = static_cast<uint16_t>(l);
                out_val[o] = v;
                if (pb.A) {
                    pb.A[o] = A;
                    pb.base[o] = pb.cosine_base ? (A + 1.0f) - (pb.residual_centers ? pb.residual_sqnorms[l]
                                                                                     : pb.centroid_sqnorms[l])
                                                : A;
                }
            }
        }
    };
    float ss = 0.0f;
    for (int t = tid; t < dim; t += blockDim.x) {
        const float v = q[static_cast<size_t>(i) * dim + t];
        if (qcopy) qsh[t] = v;
        ss += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, kWave);
    if (lane_id() == 0) s_red[tid / kWave] = ss;
    if (tid == 0) s_count = 0;
    __syncthreads();
    ss = 0.0f;
    for (int w = 0; w < NT / kWave; ++w) ss += s_red[w];

    const int target = std::min(cap, M + std::max(64, M / 4));
    const float u = 5.9604645e-8f, n = static_cast<float>(dim);
    const float qn = sqrtf(ss * (1.0f + 2.0f * (n + 2.0f) * u)) * 1.0001f;
    float delta = 2.0f * (ip ? 2.0f * (n + 1.0f) * u * qn * cnorm_max
                             : (2.0f * n + 5.0f) * u * (qn + cnorm_max) * (qn + cnorm_max)) +
                  n * 2.3509887e-38f;
    if (split_kp > 0) {
        // scores from the split BF16 product over K = 3 kp: the terms it leaves
        // out (hi.r + lo.lo + lo.r + r.c) are at most 3.02 2^-16 |q||c|, its
        // FP32 accumulation at most (3 kp + 2) u times 1.02 |q||c|; for L2 the
        // rounding of |c|^2, of the subtraction and of the exact side add
        // (2n + 6) u (|q| + |c|)^2, for inner product the exact side (n + 1) u |q||c|
        const float k3 = 3.0f * static_cast<float>(split_kp);
        const float es = (3.05f * 1.52587890625e-5f + (k3 + 2.0f) * u * 1.02f) * qn * cnorm_max;
        delta = 2.0f * (ip ? es + (n + 1.0f) * u * qn * cnorm_max
                           : 2.0f * es + (2.0f * n + 6.0f) * u * (qn + cnorm_max) * (qn + cnorm_max)) +
                k3 * 2.3509887e-38f;
    }
    // every list scoring at most `limit` into keys (up to cap of them)
    auto collect = [&](float v, int l, float limit) {
        if (v <= limit) {
            const uint32_t pos = atomicAdd(&s_count, 1u);
            if (pos < static_cast<uint32_t>(cap)) keys[pos] = static_cast<uint64_t>(l);
        }
    };
    // the row: kept in registers (CD values a thread) or read from memory
    constexpr int kC = CD > 0 ? CD : 1;
    float cached[kC];
    if (CD > 0) {
#pragma unroll
        for (int j = 0; j < kC; ++j) {
            const int l = tid + NT * j;
            cached[j] = l < nlist ? row[l] : INFINITY;
        }
    }
    auto each_vl = [&](auto&& fn) {
        if (CD > 0) {
#pragma unroll
            for (int j = 0; j < kC; ++j)
                if (tid + NT * j < nlist) fn(cached[j], tid + NT * j);
        } else {
            for (int l0 = 0; l0 < nlist; l0 += 16 * NT) {
                float v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int l = l0 + tid + NT * j;
                    v[j] = l < nlist ? row[l] : INFINITY;
                }
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (l0 + tid + NT * j < nlist) fn(v[j], l0 + tid + NT * j);
            }
        }
    };
    bool found = false;
    if (SAMPLE) {
        // tau from a sample of ns scores, runs of 16 consecutive lists one every
        // 16 nlist / ns lists (a run is one 64-byte read; single scores that far
        // apart would each cost a memory sector, as much as the whole row): the
        // rank r in the sample that puts about tc scores of the row at or below it. One
        // pass over the row counts them and collects the candidates; when fewer
        // than M are at or below (or more than cap are candidates) the row's own
        // threshold is found below.
        constexpr int kPer = kSelectSample / NT;
        const int ns = std::min(kSelectSample, nlist / 16);
        const int tc = 2 * M + 96;
        if (ns >= 64 && tc <= cap * 3 / 4) {
            const int stride = nlist / ns;
            float smp[kPer];
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int si = tid + NT * j;
                smp[j] = si < ns ? row[static_cast<size_t>(si >> 4) * 16 * stride + (si & 15)] : INFINITY;
            }
            const int r = std::min(ns, (tc * ns + nlist - 1) / nlist);
            auto each_smp = [&](auto&& fn) {
#pragma unroll
                for (int j = 0; j < kPer; ++j)
                    if (tid + NT * j < ns) fn(smp[j]);
            };
            const float ts = select_threshold_each<false, NT>(each_smp, r, r + std::max(16, r / 4), hist, s_wsum,
                                                              &s_digit, &s_below, s_red8);
            const float limit = ts + 2.0f * delta;
            uint32_t le = 0;
            each_vl([&](float v, int l) {
                le += v <= ts ? 1u : 0u;
                collect(v, l, limit);
            });
            for (int off = 32; off > 0; off >>= 1) le += static_cast<uint32_t>(__shfl_xor(static_cast<int>(le), off, kWave));
            if (lane_id() == 0) s_red[tid / kWave] = __uint_as_float(le);
            __syncthreads();
            uint32_t total = 0;
            for (int w = 0; w < NT / kWave; ++w) total += __float_as_uint(s_red[w]);
            found = total >= static_cast<uint32_t>(M) && s_count <= static_cast<uint32_t>(cap);
            __syncthreads();
            if (!found && tid == 0) s_count = 0;
            __syncthreads();
        }
    }
    if (!found) {
        auto each_all = [&](auto&& fn) { each_vl([&](float v, int) { fn(v); }); };
        const float tau = select_threshold_each<false, NT>(each_all, M, target, hist, s_wsum, &s_digit, &s_below, s_red8);
        const float limit = tau + 2.0f * delta;
        each_vl([&](float v, int l) { collect(v, l, limit); });
    }
    __syncthreads();
    int cnt = static_cast<int>(s_count);
    if (cnt >= M && cnt <= cap) { // fewer than M only when scores are NaN
        if (cnt > M + 32 && cnt <= 16 * NT) {
            // narrower: the M-th smallest approximate score among the candidates
            // has M scores at or below it, so it is a tau of the same bound; only
            // the candidates within it + 2 delta are scored exactly
            auto akey = [&](int t) { return fkey(row[static_cast<uint32_t>(keys[t])]); };
            const float lim2 = from_fkey(block_radix_select<uint32_t, true, decltype(akey), NT>(
                                   cnt, static_cast<uint32_t>(M), akey, hist, s_wsum, &s_digit, &s_below)) +
                               2.0f * delta;
            uint32_t lj[16];
            bool kj[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int t = tid + NT * j;
                lj[j] = t < cnt ? static_cast<uint32_t>(keys[t]) : 0u;
                kj[j] = t < cnt && row[lj[j]] <= lim2;
            }
            if (tid == 0) s_count2 = 0;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (kj[j]) keys[atomicAdd(&s_count2, 1u)] = lj[j];
            __syncthreads();
            cnt = static_cast<int>(s_count2);
        }
        int np = 1;
        while (np < cnt) np <<= 1;
        for (int t = tid; t < np; t += blockDim.x) {
            if (t < cnt) {
                const uint32_t l = static_cast<uint32_t>(keys[t]);
                keys[t] = dist_key(exact_score(ip != 0, qv, cents + static_cast<size_t>(l) * dim, dim), l);
            } else {
                keys[t] = ~0ull;
            }
        }
        block_bitonic_sort(keys, np);
        put_probes();
        return;
    }
    // every list exactly, then the M smallest (ckey(A), list) keys
    for (int l = tid; l < nlist; l += blockDim.x)
        row[l] = exact_score(ip != 0, qv, cents + static_cast<size_t>(l) * dim, dim);
    __syncthreads();
    auto key64 = [&](int l) { return dist_key(row[l], static_cast<uint32_t>(l)); };
    const uint64_t kth =
        block_radix_select<uint64_t, false, decltype(key64), NT>(nlist, static_cast<uint32_t>(M), key64, hist, s_wsum, &s_digit, &s_below);
    int Mp = 1;
    while (Mp < M) Mp <<= 1;
    if (tid == 0) s_count = 0;
    for (int t = tid; t < Mp; t += blockDim.x) keys[t] = ~0ull;
    __syncthreads();
    for (int l = tid; l < nlist; l += blockDim.x) {
        const uint64_t k = key64(l);
        if (k <= kth) keys[atomicAdd(&s_count, 1u)] = k;
    }
    block_bitonic_sort(keys, Mp);
    put_probes();
}

// FlatIndex::query_impl (src/flat_index.cpp:33-64) from approximate scores:
// per query (one workgroup), the k smallest (exact distance, slot) over the
// rows it may return -- float order, ties by slot -- with ids and exact
// distances. Rows scored +inf by k_coarse_gemm's masks are not eligible. The
// shortlist is k_coarse_select's: every eligible row within tau + 3 delta of
// the k-th approximate score, so it holds the k nearest; when it overflows
// (or holds fewer than k) every eligible row is scored exactly.
__global__ void __launch_bounds__(256) k_flat_select(float* D, const float* q, const float* rows, int n, int dim,
                                                     int ip, float xnorm_max, int k, int cap, const uint64_t* excl,
                                                     const uint64_t* incl, const int64_t* labels, int64_t* out_ids,
                                                     float* out_dists) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_wsum[4], s_digit, s_below, s_count;
    __shared__ float s_red[4], s_red8[8];
    float* qsh = reinterpret_cast<float*>(lds);
    uint64_t* keys = reinterpret_cast<uint64_t*>(lds + query_lds_bytes(dim));
    const int i = blockIdx.x, tid = threadIdx.x;
    float* row = D + static_cast<size_t>(i) * n;
    auto eligible = [&](int l) {
        return !((excl && bit_test(excl, static_cast<uint32_t>(l))) || (incl && !bit_test(incl, static_cast<uint32_t>(l))));
    };
    const bool qcopy = dim <= kQueryLdsDims;
    const float* qv = qcopy ? qsh : q + static_cast<size_t>(i) * dim;
    float ss = 0.0f;
    for (int t = tid; t < dim; t += blockDim.x) {
        const float v = q[static_cast<size_t>(i) * dim + t];
        if (qcopy) qsh[t] = v;
        ss += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, kWave);
    if (lane_id() == 0) s_red[tid / kWave] = ss;
    if (tid == 0) s_count = 0;
    __syncthreads();
    ss = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
    const int M = min(k, n);
    const int target = min(cap, M + max(64, M / 4));
    // rows the query may not return score +inf and take no part in tau
    const float tau = n <= 16 * 256
                          ? select_threshold<true, true>(row, n, M, target, hist, s_wsum, &s_digit, &s_below, s_red8)
                          : select_threshold<false, true>(row, n, M, target, hist, s_wsum, &s_digit, &s_below, s_red8);
    const float u = 5.9604645e-8f, nd = static_cast<float>(dim);
    const float qn = sqrtf(ss * (1.0f + 2.0f * (nd + 2.0f) * u)) * 1.0001f;
    const float delta = 2.0f * (ip ? 2.0f * (nd + 1.0f) * u * qn * xnorm_max
                                   : (2.0f * nd + 5.0f) * u * (qn + xnorm_max) * (qn + xnorm_max)) +
                        nd * 2.3509887e-38f;
    const float limit = tau + 3.0f * delta;
    for (int l0 = 0; l0 < n; l0 += 16 * 256) {
        float v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int l = l0 + tid + 256 * j;
            v[j] = l < n ? row[l] : INFINITY;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int l = l0 + tid + 256 * j;
            if (v[j] <= limit && l < n && eligible(l)) {
                const uint32_t pos = atomicAdd(&s_count, 1u);
                if (pos < static_cast<uint32_t>(cap)) keys[pos] = static_cast<uint64_t>(l);
            }
        }
    }
    __syncthreads();
    const int cnt = static_cast<int>(s_count);
    int kept = M;
    if (cnt >= M && cnt <= cap) {
        int np = 1;
        while (np < cnt) np <<= 1;
        for (int t = tid; t < np; t += blockDim.x) {
            if (t < cnt) {
                const uint32_t l = static_cast<uint32_t>(keys[t]);
                keys[t] = dist_key(exact_score(ip != 0, qv, rows + static_cast<size_t>(l) * dim, dim), l);
            } else {
                keys[t] = ~0ull;
            }
        }
        block_bitonic_sort(keys, np);
    } else {
        // every eligible row exactly, then the M smallest (ckey, slot) keys
        for (int l = tid; l < n; l += blockDim.x)
            if (eligible(l)) row[l] = exact_score(ip != 0, qv, rows + static_cast<size_t>(l) * dim, dim);
        __syncthreads();
        auto key64 = [&](int l) { return eligible(l) ? dist_key(row[l], static_cast<uint32_t>(l)) : ~0ull; };
        const uint64_t kth =
            block_radix_select<uint64_t, false>(n, static_cast<uint32_t>(M), key64, hist, s_wsum, &s_digit, &s_below);
        int Mp = 1;
        while (Mp < M) Mp <<= 1;
        if (tid == 0) s_count = 0;
        for (int t = tid; t < Mp; t += blockDim.x) keys[t] = ~0ull;
        __syncthreads();
        for (int l = tid; l < n; l += blockDim.x) {
            const uint64_t kk = key64(l);
            if (kk != ~0ull && kk <= kth) keys[atomicAdd(&s_count, 1u)] = kk;
        }
        block_bitonic_sort(keys, Mp);
        kept = min(M, static_cast<int>(s_count));
    }
    for (int t = tid; t < k; t += blockDim.x) {
        int64_t id = -1;
        float d = INFINITY;
        if (t < kept && keys[t] != ~0ull) {
            const uint32_t l = key_index(keys[t]);
            id = labels ? labels[l] : static_cast<int64_t>(l);
            d = key_dist(keys[t]);
        }
        out_ids[static_cast<size_t>(i) * k + t] = id;
        out_dists[static_cast<size_t>(i) * k + t] = d;
    }
}

// Row minimum / maximum / widest / offset and the bytes of one m x 16 table
// held in LDS (quantize_lut_scalar, src/fast_scan.cpp:22-46), by one workgroup.
__device__ void block_quantize(const float* table, int m, float* row_min, uint8_t* lut, float* scale_out,
                               float* offset_out, float* red) {
    float wide = 0.0f;
    for (int r = threadIdx.x; r < m; r += blockDim.x) {
        const float* row = table + r * 16;
        float lo = row[0], hi = row[0];
        for (int j = 1; j < 16; ++j) {
            lo = fminf(lo, row[j]);
            hi = fmaxf(hi, row[j]);
        }
        row_min[r] = lo;
        wide = fmaxf(wide, hi - lo);
    }
    wide = wave_max(wide);
    if (lane_id() == 0) red[threadIdx.x / kWave] = wide;
    __syncthreads();
    if (threadIdx.x < 4) {
        float p = 0.0f;
        for (int r = threadIdx.x; r < m; r += 4) p += row_min[r];
        red[8 + threadIdx.x] = p;
    }
    __syncthreads();
    float widest = 0.0f;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) widest = fmaxf(widest, red[w]);
    const float scale = widest > 0.0f ? __fdiv_rn(kLutMax, widest) : 1.0f;
    for (int e = threadIdx.x; e < m * 16; e += blockDim.x)
        lut[e] = static_cast<uint8_t>(__builtin_rintf((table[e] - row_min[e >> 4]) * scale));
    if (threadIdx.x == 0) {
        *scale_out = scale;
        *offset_out = (red[8] + red[9]) + (red[10] + red[11]);
    }
    __syncthreads();
}

// Per query: C = compute_inner_table(q) (src/pq.cpp:186-204) for precomputed
// L2 tables, the inner-product scoring table (C, halved for
// Metric::InnerProduct), or that table rounded to fast-scan bytes once per
// query (scan, src/ivf_index.cpp:448-456).
__global__ void k_query_tables(const float* q, const DevIvf ix, int mode, float* C_out, uint8_t* lut_out,
                               float* scale_out, float* offset_out) {
    extern __shared__ float smem[];
    const int m = ix.m, sub = ix.sub, K = ix.K;
    const int i = blockIdx.x;
    const float* qi = q + static_cast<size_t>(i) * ix.dim;
    const bool half = mode != kTableInner && ix.metric == 2;
    float* table = smem; // m * 16 (LUT mode)
    for (int e = threadIdx.x; e < m * K; e += blockDim.x) {
        const int s = e / K, c = e % K;
        float acc = 0.0f;
        for (int j = 0; j < sub; ++j) {
            const float qd = qi[s * sub + j];
            acc = fmaf(-2.0f * qd, ix.pq_cent_t[(static_cast<size_t>(s) * sub + j) * K + c], acc);
        }
        if (half) acc *= 0.5f;
        if (mode == kTableIpLut) table[e] = acc;
        else C_out[static_cast<size_t>(i) * m * K + e] = acc;
    }
    if (mode != kTableIpLut) return;
    __syncthreads();
    float* row_min = table + m * 16; // m
    float* red = row_min + m;        // 16
    block_quantize(table, m, row_min, lut_out + static_cast<size_t>(i) * m * 16, scale_out + i, offset_out + i, red);
}

// ---------------------------------------------------------------------------
// The fast-scan scan with the reference's TopK, one wave per query.
// ---------------------------------------------------------------------------

struct WaveTopK {
    float* dist;    // LDS, 2k
    int64_t* label; // LDS, 2k
    int size;       // wave-uniform
    int k;
    float thr;      // threshold(): +inf until the first trim
    uint32_t* hist; // LDS, 256 counters for the trim's radix select
};

// TopK::trim (src/topk.hpp:94-131): keep the k smallest keys, ties at the
// k-th in push order; the threshold becomes the k-th smallest distance.
// The k-th smallest key t is found by radix select: the bits every key shares
// (those above the highest bit where the smallest and largest keys differ)
// are t's, then 8 bits a pass from a histogram of the keys still matching
// the prefix found so far.
__device__ void wave_trim(WaveTopK& tk) {
    const int n = tk.size, lane = lane_id();
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (int i = lane; i < n; i += kWave) {
        const uint32_t key = fkey(tk.dist[i]);
        lo = min(lo, key);
        hi = max(hi, key);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, static_cast<uint32_t>(__shfl_xor(static_cast<int>(lo), off, kWave)));
        hi = max(hi, static_cast<uint32_t>(__shfl_xor(static_cast<int>(hi), off, kWave)));
    }
    uint32_t t = lo, below = 0;
    if (lo != hi) {
        int top = 31 - __clz(static_cast<int>(lo ^ hi));
        uint32_t prefix = top == 31 ? 0u : lo & ~((2u << top) - 1u);
        uint32_t remain = static_cast<uint32_t>(tk.k);
        while (top >= 0) {
            const int low = max(top - 7, 0);
            const uint32_t above = top == 31 ? 0u : ~((2u << top) - 1u);
            const uint32_t dmask = (2u << (top - low)) - 1u;
            for (int b = lane; b < 256; b += kWave) tk.hist[b] = 0u;
            wave_sync();
            for (int i = lane; i < n; i += kWave) {
                const uint32_t key = fkey(tk.dist[i]);
                if ((key & above) == prefix) atomicAdd(&tk.hist[(key >> low) & dmask], 1u);
            }
            wave_sync();
            const uint32_t c0 = tk.hist[4 * lane], c1 = tk.hist[4 * lane + 1], c2 = tk.hist[4 * lane + 2],
                           c3 = tk.hist[4 * lane + 3];
            const uint32_t mine = c0 + c1 + c2 + c3;
            uint32_t incl = mine;
            for (int off = 1; off < kWave; off <<= 1) {
                const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), off, kWave));
                if (lane >= off) incl += y;
            }
            const uint32_t excl = incl - mine;
            // the lane whose four bins hold the remain-th key, then the bin
            const uint64_t at = ballot(excl < remain && incl >= remain);
            const int src = __builtin_ctzll(at);
            uint32_t acc = excl, d = 0;
            if (acc + c0 < remain) {
                acc += c0;
                d = 1;
                if (acc + c1 < remain) {
                    acc += c1;
                    d = 2;
                    if (acc + c2 < remain) {
                        acc += c2;
                        d = 3;
                    }
                }
            }
            const uint32_t digit = static_cast<uint32_t>(__shfl(static_cast<int>(4 * lane + d), src, kWave));
            const uint32_t nb = static_cast<uint32_t>(__shfl(static_cast<int>(acc), src, kWave));
            prefix |= digit << low;
            remain -= nb;
            below += nb;
            top = low - 1;
            wave_sync();
        }
        t = prefix;
    }
    int ties = tk.k - static_cast<int>(below);
    int w = 0;
    for (int base = 0; base < n; base += kWave) {
        const int i = base + lane_id();
        const bool valid = i < n;
        const float d = valid ? tk.dist[i] : 0.0f;
        const int64_t lab = valid ? tk.label[i] : 0;
        const uint32_t key = valid ? fkey(d) : 0xFFFFFFFFu;
        const bool lt = valid && key < t;
        const bool eq = valid && key == t;
        const uint64_t eqm = ballot(eq);
        const bool tie_keep = eq && static_cast<int>(lanes_below(eqm)) < ties;
        const bool keep = lt || tie_keep;
        const uint64_t km = ballot(keep);
        const int pos = w + static_cast<int>(lanes_below(km));
        wave_sync();
        if (keep) {
            tk.dist[pos] = d;
            tk.label[pos] = lab;
        }
        wave_sync();
        w += __popcll(km);
        ties -= __popcll(ballot(tie_keep));
    }
    tk.size = tk.k;
    tk.thr = from_fkey(t);
}

// Cross-lane moves within 16-lane DPP rows (CDNA data-parallel primitives).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}

// Minimum / maximum over each row of 16 lanes, left in every lane of the row:
// quad swaps (quad_perm [1,0,3,2] then [2,3,0,1]), half-row mirror, row mirror.
__device__ __forceinline__ float row16_min(float v) {
    v = fminf(v, dpp_f<0xB1>(v));
    v = fminf(v, dpp_f<0x4E>(v));
    v = fminf(v, dpp_f<0x141>(v));
    return fminf(v, dpp_f<0x140>(v));
}

__device__ __forceinline__ float row16_max(float v) {
    v = fmaxf(v, dpp_f<0xB1>(v));
    v = fmaxf(v, dpp_f<0x4E>(v));
    v = fmaxf(v, dpp_f<0x141>(v));
    return fmaxf(v, dpp_f<0x140>(v));
}

// Orders LDS writes and reads between the lanes of one wave (LDS operations
// of a wave complete in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The 16 entries of row r of one (query, list) L2 table: B + C (+ A on row 0)
// from the precomputed list terms (build_quantize_lut, src/fast_scan.cpp:243-254),
// or the table of the residual q - c (compute_adc_table, src/pq.cpp:165-184:
// out[j] += (r_d - y_jd)^2 over the subspace's dimensions, fused).
__device__ __forceinline__ void l2_row(const DevIvf& ix, const float* B, const float* Cq, float A, const float* q,
                                       const float* c, const float* resid, int r, float t[16]) {
    if (B) {
        const float4* b4 = reinterpret_cast<const float4*>(B + r * 16);
        const float4* c4 = reinterpret_cast<const float4*>(Cq + r * 16);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float4 x = b4[i], y = c4[i];
            t[4 * i + 0] = x.x + y.x;
            t[4 * i + 1] = x.y + y.y;
            t[4 * i + 2] = x.z + y.z;
            t[4 * i + 3] = x.w + y.w;
        }
        if (r == 0) {
#pragma unroll
            for (int j = 0; j < 16; ++j) t[j] = t[j] + A;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) t[j] = 0.0f;
    for (int s = 0; s < ix.sub; ++s) {
        const int x = r * ix.sub + s;
        const float rs = resid ? resid[x] : q[x] - c[x];
        const float4* y4 = reinterpret_cast<const float4*>(ix.pq_cent_t + static_cast<size_t>(x) * 16);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float4 y = y4[i];
            float d;
            d = rs - y.x; t[4 * i + 0] = fmaf(d, d, t[4 * i + 0]);
            d = rs - y.y; t[4 * i + 1] = fmaf(d, d, t[4 * i + 1]);
            d = rs - y.z; t[4 * i + 2] = fmaf(d, d, t[4 * i + 2]);
            d = rs - y.w; t[4 * i + 3] = fmaf(d, d, t[4 * i + 3]);
        }
    }
}

// The LUT as k_fs_sums keeps it in LDS: every byte complemented except those
// of entries 4 and 12 (a table row's four words XORed as lut_code does).
// A nibble x then reads as perm(entries 0..7, x) ^ perm(entries 8..15, x ^ 8):
// the half the nibble does not index answers 0x00 (v_perm selector 12, or
// selectors 8..11, the sign of an entry that was not complemented) or 0xFF
// (selectors 13..15, or the sign of a complemented entry), and the half it
// indexes holds the entry complemented exactly where the other answers 0xFF.
__device__ __forceinline__ uint4 lut_code(uint4 w) {
    return make_uint4(w.x ^ 0xFFFFFFFFu, w.y ^ 0xFFFFFF00u, w.z ^ 0xFFFFFFFFu, w.w ^ 0xFFFFFF00u);
}
__device__ __forceinline__ uint32_t lut_code_word(uint32_t w, int i) { return w ^ ((i & 1) ? 0xFFFFFF00u : 0xFFFFFFFFu); }

// The L2 table of one (query, probed list), rounded to bytes (quantize_lut,
// src/fast_scan.cpp:22-46), by one wave, one table row per lane: the row's
// minimum and maximum taken in entry order as std::min / std::max do, one
// scale 63 / widest over all rows, offset (p0 + p1) + (p2 + p3) with p[i % 4]
// summing row minimums in row order, bytes rounded half to even. Rows m ..
// m4 - 1 of the LUT are zero. The rows are written coded (lut_code).
__device__ void wave_lut_l2(const DevIvf& ix, const float* q, int list, float A, const float* Cq, float* rowmin,
                            float* resid, uint8_t* lut, float& scale_out, float& offset_out) {
    const int m = ix.m, lane = lane_id();
    const float* B = ix.B ? ix.B + static_cast<size_t>(list) * m * 16 : nullptr;
    const float* c = (ix.residual_centers ? ix.residual_centers : ix.centroids) + static_cast<size_t>(list) * ix.dim;
    if (!B && resid) {
        for (int e = lane; e < ix.dim; e += kWave) resid[e] = q[e] - c[e];
        wave_lds_sync();
    }
    float t[16];
    float wide = 0.0f, lo_keep = 0.0f;
    for (int r0 = 0; r0 < m; r0 += kWave) {
        const int r = r0 + lane;
        if (r < m) {
            l2_row(ix, B, Cq, A, q, c, resid, r, t);
            float lo = t[0], hi = t[0];
#pragma unroll
            for (int j = 1; j < 16; ++j) {
                lo = fminf(lo, t[j]);
                hi = fmaxf(hi, t[j]);
            }
            // std::min / std::max in entry order give the same values except
            // for a NaN first entry (kept) and the sign of a zero minimum (the
            // first zero's); the maximum's zero sign never reaches a result
            if (t[0] != t[0]) {
                lo = t[0];
                hi = t[0];
            }
            if (ballot(lo == 0.0f)) { // rare: taken by the wave only when some row needs it
                if (lo == 0.0f) {
                    for (int j = 15; j >= 0; --j)
                        if (t[j] == 0.0f) lo = t[j];
                }
            }
            rowmin[r] = lo;
            lo_keep = lo;
            wide = fmaxf(wide, hi - lo);
        }
    }
    const float widest = wave_max(wide);
    const float scale = widest > 0.0f ? __fdiv_rn(kLutMax, widest) : 1.0f;
    wave_lds_sync();
    float part = 0.0f;
    if (lane < 4)
        for (int r = lane; r < m; r += 4) part += rowmin[r];
    const float p0 = __shfl(part, 0, kWave), p1 = __shfl(part, 1, kWave);
    const float p2 = __shfl(part, 2, kWave), p3 = __shfl(part, 3, kWave);
    for (int r0 = 0; r0 < ix.m4; r0 += kWave) {
        const int r = r0 + lane;
        if (r < m) {
            float lo = lo_keep;
            if (m > kWave) {
                // a row of an earlier group: its entries again
                l2_row(ix, B, Cq, A, q, c, resid, r, t);
                lo = rowmin[r];
            }
            uint32_t w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                uint32_t v = 0;
                // the byte conversion rounds half to even, as nearbyint does; the
                // values lie in [0, 63 + rounding]
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    v = __builtin_amdgcn_cvt_pk_u8_f32((t[4 * i + j] - lo) * scale, j, v);
                w[i] = v;
            }
            *reinterpret_cast<uint4*>(lut + r * 16) = lut_code(make_uint4(w[0], w[1], w[2], w[3]));
        } else if (r < ix.m4) {
            *reinterpret_cast<uint4*>(lut + r * 16) = lut_code(make_uint4(0u, 0u, 0u, 0u));
        }
    }
    scale_out = scale;
    offset_out = (p0 + p1) + (p2 + p3);
}

// Eight LUT lookups of one code word (4 bytes = vectors 4g..4g+3 in the low
// nibbles, 16+4g.. in the high) against one coded 16-byte LUT row held in
// four registers; the results are added bytewise.
__device__ __forceinline__ void lookup8(uint32_t w, const uint4 r, uint32_t& lo8, uint32_t& hi8) {
    const uint32_t xl = w & 0x0F0F0F0Fu;
    const uint32_t xh = (w >> 4) & 0x0F0F0F0Fu;
    lo8 += __builtin_amdgcn_perm(r.y, r.x, xl) ^ __builtin_amdgcn_perm(r.w, r.z, xl ^ 0x08080808u);
    hi8 += __builtin_amdgcn_perm(r.y, r.x, xh) ^ __builtin_amdgcn_perm(r.w, r.z, xh ^ 0x08080808u);
}

__device__ __forceinline__ uint4 load_code16(const uint8_t* p) { return *reinterpret_cast<const uint4*>(p); }

// The u16 sums of one lane's eight vectors of a code block (fast_scan_block,
// src/fast_scan.cpp:256-262): lane g of the block's four reads, at cp, column
// g's 16 bytes of each group of four subspaces (n4 groups, 64 bytes apart),
// and looks them up in the coded LUT rows at lut. Sums of vectors 4g..4g+3
// come back as u16 pairs (4g, 4g+1) in l01 and (4g+2, 4g+3) in l23; those of
// vectors 16+4g.. in h01 / h23.
__device__ __forceinline__ void fs_block_sums(const uint8_t* cp, const uint8_t* lut, int n4, uint32_t& l01,
                                              uint32_t& l23, uint32_t& h01, uint32_t& h23) {
    uint32_t lo_a = 0, lo_b = 0, hi_a = 0, hi_b = 0; // u16 pairs: vectors (4g, 4g+2), (4g+1, 4g+3), +16
    // 16 subspaces per step of the main loop, their four 16-byte loads in flight together
    auto step4 = [&](const uint4 w, int s4) {
        const uint8_t* rp = lut + s4 * 64;
        uint32_t lo8 = 0, hi8 = 0; // four subspaces' bytes add to at most 4 * 63
        lookup8(w.x, *reinterpret_cast<const uint4*>(rp), lo8, hi8);
        lookup8(w.y, *reinterpret_cast<const uint4*>(rp + 16), lo8, hi8);
        lookup8(w.z, *reinterpret_cast<const uint4*>(rp + 32), lo8, hi8);
        lookup8(w.w, *reinterpret_cast<const uint4*>(rp + 48), lo8, hi8);
        lo_a += lo8 & 0x00FF00FFu;
        lo_b += __builtin_amdgcn_perm(0u, lo8, 0x0C030C01u);
        hi_a += hi8 & 0x00FF00FFu;
        hi_b += __builtin_amdgcn_perm(0u, hi8, 0x0C030C01u);
    };
    int s4 = 0;
    for (; s4 + 4 <= n4; s4 += 4) {
        const uint4 w0 = load_code16(cp + 64 * s4), w1 = load_code16(cp + 64 * (s4 + 1));
        const uint4 w2 = load_code16(cp + 64 * (s4 + 2)), w3 = load_code16(cp + 64 * (s4 + 3));
        step4(w0, s4);
        step4(w1, s4 + 1);
        step4(w2, s4 + 2);
        step4(w3, s4 + 3);
    }
    for (; s4 < n4; ++s4) step4(load_code16(cp + 64 * s4), s4);
    l01 = (lo_a & 0xFFFFu) | (lo_b << 16);
    l23 = (lo_a >> 16) | (lo_b & 0xFFFF0000u);
    h01 = (hi_a & 0xFFFFu) | (hi_b << 16);
    h23 = (hi_a >> 16) | (hi_b & 0xFFFF0000u);
}

// Phase 1: one workgroup of kSumsWaves waves per (query, window of up to
// kSumsWin consecutive probes). The waves build the window's LUTs in LDS (one
// per probe for L2 tables; the query's one inner-product LUT otherwise), then
// the window's code blocks are taken as one sequence, four lanes per block:
// lane g of a block reads 16 bytes per 4 subspaces (the device layout keeps
// column g's words of subspaces 4i..4i+3 together) and writes the u16 sums of
// vectors 4g..4g+3 and 16+4g..16+4g+3 (fast_scan_block, src/fast_scan.cpp:256-262).
// Each probe's smallest sum goes to item_min for the replay's list skip.
// One (query, window) of phase 1; the workgroup's shared state is passed in.
__device__ __forceinline__ void sums_item(const SumsArgs& a, int qi, int p0, unsigned char* lds, uint32_t* w_start,
                                          int* w_list, uint32_t* w_min) {
    const DevIvf& ix = a.ix;
    const int m = ix.m, lb = ix.m4 * 16, win = a.win;
    const int P = a.nprobe_q[qi];
    if (p0 >= P) return;
    const int np = min(win, P - p0);
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const bool ip = a.ip_lut != nullptr;
    const size_t item0 = static_cast<size_t>(qi) * a.M + p0;
    const float* q = a.q + static_cast<size_t>(qi) * ix.dim;
    uint8_t* luts = lds;
    float* rowmin = reinterpret_cast<float*>(lds + static_cast<size_t>(ip ? 1 : win) * lb);
    float* resid = rowmin + kSumsWaves * m;
    const uint32_t blk0 = a.item_blk[item0];
    if (tid < np) {
        const int list = a.probe_idx[item0 + tid];
        w_list[tid] = list;
        w_min[tid] = 0xFFFFFFFFu;
        const uint32_t rel = a.item_blk[item0 + tid] - blk0;
        w_start[tid] = rel;
        if (tid == np - 1) w_start[np] = rel + (ix.list_off[list + 1] - ix.list_off[list] + kFsBlock - 1) / kFsBlock;
    }
    if (ip) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.ip_lut + static_cast<size_t>(qi) * m * 16);
        for (int e = tid; e < lb / 4; e += blockDim.x)
            reinterpret_cast<uint32_t*>(luts)[e] = lut_code_word(e < m * 4 ? src[e] : 0u, e);
        if (tid < np) {
            a.item_scale[item0 + tid] = a.ip_scale[qi];
            a.item_offset[item0 + tid] = a.ip_offset[qi] + a.probe_base[item0 + tid];
        }
    }
    __syncthreads();
    if (!ip) {
        const float* Cq = a.C ? a.C + static_cast<size_t>(qi) * m * 16 : nullptr;
        for (int p = wave; p < np; p += kSumsWaves) {
            if (w_start[p + 1] == w_start[p]) continue; // an empty list: the replay skips it
            const size_t item = item0 + p;
            float scale, offset;
            wave_lut_l2(ix, q, w_list[p], a.probe_A[item], Cq, rowmin + wave * m,
                        a.resid_in_lds ? resid + wave * ix.dim : nullptr,
                        luts + static_cast<size_t>(p) * lb, scale, offset);
            if (lane == 0) {
                a.item_scale[item] = scale;
                a.item_offset[item] = offset;
            }
        }
    }
    __syncthreads();

    const uint32_t T = w_start[np];
    const int g = tid & 3, n4 = ix.m4 >> 2;
    uint16_t* qsums = a.sums + static_cast<size_t>(qi) * a.stride_slots + static_cast<size_t>(blk0) * kFsBlock;
    uint16_t* qbmin = a.bmin + static_cast<size_t>(qi) * (a.stride_slots / kFsBlock) + blk0;
    for (uint32_t base = 0; base < T; base += blockDim.x / 4) {
        const uint32_t t = base + (tid >> 2);
        uint32_t vmin = 0xFFFFFFFFu;
        int p = 0;
        if (t < T) {
            // the probe of block t: the last p with w_start[p] <= t (empty lists share a start)
#pragma unroll
            for (int step = kSumsWin / 2; step > 0; step >>= 1) {
                const int c = min(p + step, np - 1);
                p = w_start[c] <= t ? c : p;
            }
            const int list = w_list[p];
            const uint32_t b = t - w_start[p];
            const uint32_t n = ix.list_off[list + 1] - ix.list_off[list];
            const uint8_t* cp = ix.codes + ix.code_off[list] + static_cast<size_t>(b) * lb + g * 16;
            const uint8_t* lut = luts + (ip ? 0 : static_cast<size_t>(p) * lb);
            uint32_t l01, l23, h01, h23;
            fs_block_sums(cp, lut, n4, l01, l23, h01, h23);
            uint16_t* o = qsums + static_cast<size_t>(t) * kFsBlock;
            *reinterpret_cast<uint2*>(o + 4 * g) = make_uint2(l01, l23);
            *reinterpret_cast<uint2*>(o + 16 + 4 * g) = make_uint2(h01, h23);
            const uint32_t v0 = b * kFsBlock;
            const uint32_t vals[8] = {l01 & 0xFFFFu, l01 >> 16, l23 & 0xFFFFu, l23 >> 16,
                                      h01 & 0xFFFFu, h01 >> 16, h23 & 0xFFFFu, h23 >> 16};
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t v = v0 + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
                if (v < n) vmin = min(vmin, vals[i]);
            }
        }
        // the block's four lanes hold one probe: their minimum, then one LDS atomic
        vmin = min(vmin, dpp_u<0xB1>(vmin));
        vmin = min(vmin, dpp_u<0x4E>(vmin));
        if (t < T && g == 0) {
            atomicMin(&w_min[p], vmin);
            qbmin[t] = static_cast<uint16_t>(min(vmin, 0xFFFFu));
        }
    }
    __syncthreads();
    if (tid < np) a.item_min[item0 + tid] = w_min[tid];
}

// The (query, window) of work item `item` of the XCD-ordered form: XCD x owns
// ranks x * per_xcd .. of the order, window by window.
__device__ __forceinline__ void ordered_item(const SumsArgs& a, int x, int i, int& qi, int& p0, bool& ok) {
    const int rank = x * a.per_xcd + i / a.nwin;
    ok = rank < a.nq;
    qi = ok ? static_cast<int>(a.order ? a.order[rank] : static_cast<uint32_t>(rank)) : 0;
    p0 = (i % a.nwin) * a.win;
}

// OCC: waves per SIMD the register allocation leaves room for.
constexpr int kSumsHighOccQueries = 1024;
template <int OCC>
__global__ void __launch_bounds__(kSumsWaves * kWave, OCC) k_fs_sums(const SumsArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t w_start[kSumsWin + 1];
    __shared__ int w_list[kSumsWin];
    __shared__ uint32_t w_min[kSumsWin];
    int qi, p0;
    if (a.order) {
        // workgroup b runs on XCD b % 8: XCD x takes ranks x * per_xcd .. of the order
        bool ok;
        ordered_item(a, static_cast<int>(blockIdx.x % 8), static_cast<int>(blockIdx.x / 8), qi, p0, ok);
        if (!ok) return;
    } else {
        qi = static_cast<int>(blockIdx.x) / a.nwin;
        p0 = (static_cast<int>(blockIdx.x) % a.nwin) * a.win;
    }
    sums_item(a, qi, p0, lds, w_start, w_list, w_min);
}

// Phase 1 as a persistent kernel: a grid that fills the chip once, each
// workgroup pulling (query, window) items from a device queue per XCD (the
// XCD-ordered items of that XCD; a drained queue sends it to the others') until
// every queue is drained.
// One pull from the queues, starting at the home XCD's: (queue, index) or x < 0.
__device__ __forceinline__ void pull_item(const SumsArgs& a, int home, int per_queue, int& x_out, int& i_out) {
    x_out = -1;
    i_out = 0;
    for (int y = 0; y < 8; ++y) {
        const int x = (home + y) % 8;
        if (__atomic_load_n(&a.queue[x], __ATOMIC_RELAXED) >= static_cast<uint32_t>(per_queue)) continue;
        const uint32_t i = atomicAdd(&a.queue[x], 1u);
        if (i < static_cast<uint32_t>(per_queue)) {
            x_out = x;
            i_out = static_cast<int>(i);
            return;
        }
    }
}

__global__ void __launch_bounds__(kSumsWaves * kWave) k_fs_sums_queue(const SumsArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t w_start[kSumsWin + 1];
    __shared__ int w_list[kSumsWin];
    __shared__ uint32_t w_min[kSumsWin];
    __shared__ int s_x[2], s_i[2];
    // the XCD this workgroup runs on (hardware register XCC_ID), whose queue it drains first
    const int home = static_cast<int>(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 7);
    const int per_queue = a.per_xcd * a.nwin;
    if (threadIdx.x == 0) pull_item(a, home, per_queue, s_x[0], s_i[0]);
    __syncthreads();
    for (int cur = 0;; cur ^= 1) {
        const int x = s_x[cur], i = s_i[cur];
        if (x < 0) return;
        // the next item's pull is in flight while this one runs
        if (threadIdx.x == 0) pull_item(a, home, per_queue, s_x[cur ^ 1], s_i[cur ^ 1]);
        int qi, p0;
        bool ok;
        ordered_item(a, x, i, qi, p0, ok);
        if (ok) sums_item(a, qi, p0, lds, w_start, w_list, w_min);
        __syncthreads();
    }
}

// quantize_lut (src/fast_scan.cpp:22-46) of table t (m x 16 floats at
// ix.B + t * m * 16) by the search kernels' own LUT build, one wave per
// table: C is m x 16 entries of -0 and A is -0, so the entries reach the
// rounding unchanged (x + -0 = x, zeros keep their sign). The LUT goes out
// decoded (plain bytes), with its scale and offset.
__global__ void __launch_bounds__(kWave) k_fs_lut_check(const DevIvf ix, const float* negzero, uint8_t* lut_out,
                                                      float* scale_out, float* offset_out) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint8_t* lut = lds;
    float* rowmin = reinterpret_cast<float*>(lds + static_cast<size_t>(ix.m4) * 16);
    const int t = blockIdx.x, lane = lane_id();
    float scale, offset;
    wave_lut_l2(ix, nullptr, t, -0.0f, negzero, rowmin, nullptr, lut, scale, offset);
    wave_lds_sync();
    const uint32_t* words = reinterpret_cast<const uint32_t*>(lut);
    uint32_t* out = reinterpret_cast<uint32_t*>(lut_out + static_cast<size_t>(t) * ix.m * 16);
    for (int e = lane; e < ix.m * 4; e += kWave) out[e] = lut_code_word(words[e], e);
    if (lane == 0) {
        scale_out[t] = scale;
        offset_out[t] = offset;
    }
}

// fast_scan_block (src/fast_scan.cpp:256-262) of block b (the device layout,
// m4 * 16 bytes at blocks + b * m4 * 16) against LUT b (m x 16 plain bytes at
// luts + b * m * 16) by the search kernels' own lookups: the LUT is coded into
// LDS as k_fs_sums keeps it, and the block's four lanes each sum their eight
// vectors. out: 32 u16 sums per block.
__global__ void __launch_bounds__(kWave) k_fs_block_check(const uint8_t* luts, int m, int m4, const uint8_t* blocks,
                                                        uint16_t* out) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int b = blockIdx.x, lane = lane_id();
    uint32_t* words = reinterpret_cast<uint32_t*>(lds);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(luts + static_cast<size_t>(b) * m * 16);
    for (int e = lane; e < m4 * 4; e += kWave) words[e] = lut_code_word(e < m * 4 ? src[e] : 0u, e);
    wave_lds_sync();
    if (lane < 4) {
        uint32_t l01, l23, h01, h23;
        fs_block_sums(blocks + static_cast<size_t>(b) * m4 * 16 + lane * 16, lds, m4 >> 2, l01, l23, h01, h23);
        uint16_t* o = out + static_cast<size_t>(b) * kFsBlock;
        *reinterpret_cast<uint2*>(o + 4 * lane) = make_uint2(l01, l23);
        *reinterpret_cast<uint2*>(o + 16 + 4 * lane) = make_uint2(h01, h23);
    }
}

// ---------------------------------------------------------------------------
// Phase 1 on the matrix cores: the u16 sums of a list's code blocks for 32
// (query, probe) pairs at once, as an int8 MFMA product. For one block the
// sums are S[v][j] = sum over subspaces s of LUT_j[s][code(v, s)]: a product
// of the block's codes written one-hot (32 vectors x 16m, 0 or 1) with the 32
// pairs' LUTs (16m x 32, bytes <= 63), accumulated exactly in i32. The pairs
// are grouped by list, so one wave reads a list's blocks once for 32 queries
// (the per-query form reads them once per query).
// ---------------------------------------------------------------------------

// The pairs (query i, probe p < nprobe_q[i]) of non-empty lists, counted per
// list; pairs of empty lists get item_min = 0xFFFFFFFF (the replay skips them).
// Two ways: with per-list counters in LDS (nlist <= kPairsLdsLists), each of
// kPairGroups workgroups counting a contiguous range of the queries, the
// counts combined without global atomics; or with global atomics.
constexpr int kPairsLdsLists = 16384;
constexpr int kPairGroups = 64;

__device__ __forceinline__ bool pair_of(const SumsArgs& a, uint64_t t, int& list) {
    const int qi = static_cast<int>(t / a.M), p = static_cast<int>(t % a.M);
    if (p >= a.nprobe_q[qi]) return false;
    list = a.probe_idx[t];
    if (a.ix.list_off[list + 1] == a.ix.list_off[list]) {
        a.item_min[t] = 0xFFFFFFFFu;
        return false;
    }
    return true;
}

// queries [g nq / G, (g + 1) nq / G) of workgroup g: counts into ghist[g][list]
__global__ void __launch_bounds__(1024) k_pairs_hist(const SumsArgs a, int nq, uint32_t* ghist) {
    extern __shared__ uint32_t h[];
    const int nlist = a.ix.nlist, g = blockIdx.x;
    for (int l = threadIdx.x; l < nlist; l += blockDim.x) h[l] = 0u;
    __syncthreads();
    const uint64_t q0 = static_cast<uint64_t>(nq) * g / gridDim.x, q1 = static_cast<uint64_t>(nq) * (g + 1) / gridDim.x;
    for (uint64_t t = q0 * a.M + threadIdx.x; t < q1 * a.M; t += blockDim.x) {
        int list;
        if (pair_of(a, t, list)) atomicAdd(&h[list], 1u);
    }
    __syncthreads();
    for (int l = threadIdx.x; l < nlist; l += blockDim.x) ghist[static_cast<size_t>(g) * nlist + l] = h[l];
}

// From the per-workgroup counts ghist[g][list]: per list, each group's first
// position within the list's range (ghist, in place) and the list's total (tot).
// One wave per list, lane j holding groups kPerLane j .. kPerLane j + kPerLane - 1.
constexpr int kPerLane = kPairGroups / kWave;
static_assert(kPairGroups % kWave == 0, "whole groups per lane");
__global__ void __launch_bounds__(256) k_pairs_lists(uint32_t* ghist, int groups, int n, uint32_t* tot) {
    const int l = blockIdx.x * (256 / kWave) + static_cast<int>(threadIdx.x) / kWave, lane = lane_id();
    if (l >= n) return;
    uint32_t c[kPerLane], own = 0;
#pragma unroll
    for (int i = 0; i < kPerLane; ++i) {
        const int g = lane * kPerLane + i;
        c[i] = g < groups ? ghist[static_cast<size_t>(g) * n + l] : 0u;
        own += c[i];
    }
    uint32_t run = own; // inclusive prefix of the lanes' sums
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(run), off, kWave));
        if (lane >= off) run += o;
    }
    uint32_t at = run - own;
#pragma unroll
    for (int i = 0; i < kPerLane; ++i) {
        const int g = lane * kPerLane + i;
        if (g < groups) ghist[static_cast<size_t>(g) * n + l] = at;
        at += c[i];
    }
    if (lane == kWave - 1) tot[l] = run;
}

// pair_off / unit_off: exclusive prefix sums over the lists of the pair totals
// and of their 32-pair units, with the totals at index n. One 1024-thread
// block over chunks of 4,096 lists: a chunk's totals are read into LDS in one
// coalesced pass, thread t takes lists 4t .. 4t + 3 of it, an inclusive scan
// within each wave and then over the 16 waves' sums gives each thread its
// place, and the running totals carry to the next chunk.
constexpr int kOffChunk = 4096;
__global__ void __launch_bounds__(1024) k_pairs_offsets(const uint32_t* tot, int n, uint32_t* pair_off,
                                                        uint32_t* unit_off) {
    __shared__ uint32_t s_tot[kOffChunk];
    __shared__ uint32_t w_p[16], w_u[16];
    const int t = threadIdx.x, lane = lane_id(), w = t / kWave;
    uint32_t carry_p = 0, carry_u = 0;
    for (int base = 0; base < n; base += kOffChunk) {
        const int cnt = min(kOffChunk, n - base);
        for (int j = t; j < cnt; j += 1024) s_tot[j] = tot[base + j];
        __syncthreads();
        uint32_t c[4], sp = 0, su = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int l = 4 * t + k;
            c[k] = l < cnt ? s_tot[l] : 0u;
            sp += c[k];
            su += (c[k] + 31) / 32;
        }
        uint32_t ip = sp, iu = su; // inclusive, within the wave
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t yp = static_cast<uint32_t>(__shfl_up(static_cast<int>(ip), off, kWave));
            const uint32_t yu = static_cast<uint32_t>(__shfl_up(static_cast<int>(iu), off, kWave));
            if (lane >= off) {
                ip += yp;
                iu += yu;
            }
        }
        if (lane == kWave - 1) {
            w_p[w] = ip;
            w_u[w] = iu;
        }
        __syncthreads();
        uint32_t bp = carry_p, bu = carry_u, tp = 0, tu = 0;
        for (int v = 0; v < 16; ++v) {
            if (v < w) {
                bp += w_p[v];
                bu += w_u[v];
            }
            tp += w_p[v];
            tu += w_u[v];
        }
        uint32_t rp = bp + ip - sp, ru = bu + iu - su;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int l = 4 * t + k;
            if (l < cnt) {
                pair_off[base + l] = rp;
                unit_off[base + l] = ru;
                rp += c[k];
                ru += (c[k] + 31) / 32;
            }
        }
        carry_p += tp;
        carry_u += tu;
        __syncthreads(); // s_tot and the wave sums are rewritten by the next chunk
    }
    if (t == 0) {
        pair_off[n] = carry_p;
        unit_off[n] = carry_u;
    }
}

// Each unit's list: one wave per list, its lanes over the list's units.
__global__ void __launch_bounds__(256) k_pairs_units(const uint32_t* unit_off, int n, uint32_t* unit_list) {
    const int l = blockIdx.x * (256 / kWave) + static_cast<int>(threadIdx.x) / kWave;
    if (l >= n) return;
    const uint32_t end = unit_off[l + 1];
    for (uint32_t k = unit_off[l] + static_cast<uint32_t>(lane_id()); k < end; k += kWave)
        unit_list[k] = static_cast<uint32_t>(l);
}

// Each pair's item index (i * M + p) into its list's range, the workgroup's
// cursors in LDS.
__global__ void __launch_bounds__(1024) k_pairs_place(const SumsArgs a, int nq, const uint32_t* ghist,
                                                      const uint32_t* pair_off, uint32_t* items) {
    extern __shared__ uint32_t cur[];
    const int nlist = a.ix.nlist, g = blockIdx.x;
    for (int l = threadIdx.x; l < nlist; l += blockDim.x)
        cur[l] = pair_off[l] + ghist[static_cast<size_t>(g) * nlist + l];
    __syncthreads();
    const uint64_t q0 = static_cast<uint64_t>(nq) * g / gridDim.x, q1 = static_cast<uint64_t>(nq) * (g + 1) / gridDim.x;
    for (uint64_t t = q0 * a.M + threadIdx.x; t < q1 * a.M; t += blockDim.x) {
        int list;
        if (pair_of(a, t, list)) items[atomicAdd(&cur[list], 1u)] = static_cast<uint32_t>(t);
    }
}

__global__ void k_pairs_cursors(const uint32_t* pair_off, int n, uint32_t* cur) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) cur[l] = pair_off[l];
}

// The same grouping with global counters, for more lists than LDS holds.
__global__ void k_pairs_count(const SumsArgs a, int nq, uint32_t* cnt) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= static_cast<uint64_t>(nq) * a.M) return;
    int list;
    if (pair_of(a, t, list)) atomicAdd(&cnt[list], 1u);
}

__global__ void k_pairs_scatter(const SumsArgs a, int nq, uint32_t* fill, uint32_t* items) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= static_cast<uint64_t>(nq) * a.M) return;
    const int qi = static_cast<int>(t / a.M), p = static_cast<int>(t % a.M);
    if (p >= a.nprobe_q[qi]) return;
    const int list = a.probe_idx[t];
    if (a.ix.list_off[list + 1] == a.ix.list_off[list]) return;
    items[atomicAdd(&fill[list], 1u)] = static_cast<uint32_t>(t);
}

// A table row's minimum and maximum as std::min / std::max take them in entry
// order (quantize_lut, src/fast_scan.cpp:22-46): fminf / fmaxf agree except for
// a NaN first entry (kept) and the sign of a zero minimum (the first zero's).
__device__ __forceinline__ void row_lo_hi(const float t[16], float& lo, float& hi) {
    lo = t[0];
    hi = t[0];
#pragma unroll
    for (int c = 1; c < 16; ++c) {
        lo = fminf(lo, t[c]);
        hi = fmaxf(hi, t[c]);
    }
    if (t[0] != t[0]) {
        lo = t[0];
        hi = t[0];
    }
    if (lo == 0.0f) {
        for (int c = 15; c >= 0; --c)
            if (t[c] == 0.0f) lo = t[c];
    }
}

// row_lo_hi for every lane of a wave, with no branch but the zero-sign pass,
// which the wave takes only when one of its rows has a minimum of zero;
// lanes whose row is past the table get lo = hi = 0.
__device__ __forceinline__ void row_lo_hi_wave(const float t[16], bool row, float& lo, float& hi) {
    lo = t[0];
    hi = t[0];
#pragma unroll
    for (int c = 1; c < 16; ++c) {
        lo = fminf(lo, t[c]);
        hi = fmaxf(hi, t[c]);
    }
    const bool nan0 = t[0] != t[0];
    lo = nan0 ? t[0] : lo;
    hi = nan0 ? t[0] : hi;
    if (__ballot(row && lo == 0.0f)) {
        if (lo == 0.0f) {
            for (int c = 15; c >= 0; --c)
                if (t[c] == 0.0f) lo = t[c];
        }
    }
    lo = row ? lo : 0.0f;
    hi = row ? hi : 0.0f;
}

// Units of 32 pairs: unit u is a chunk of list unit_list[u]'s pairs.
constexpr int kMfmaWaves = 2;
constexpr int kMfmaSubMax = 4; // PQ sub-vector lengths whose codebook rows stay in registers

// LDS of one wave of k_fs_sums_mfma: the rounded LUT rows of 16 pairs on
// their way into MFMA operands, those pairs' row minimums, and the 32 pairs'
// scales and offsets; later two staged code blocks; then the 16 one-hot rows
// (256 bytes) the blocks' MFMA A operands are read from
__host__ __device__ inline size_t mfma_wave_lds(int m4) {
    const size_t build = static_cast<size_t>(16) * m4 * 16 + 16 * 64 * 4 + 32 * 8;
    const size_t blocks = static_cast<size_t>(2) * m4 * 16;
    return (((build > blocks ? build : blocks) + 15) & ~size_t{15}) + 256;
}

// One unit on one wave. NK = m4 / 2 MFMA steps of 2 subspaces; IP: the
// query's one LUT (ip_lut), else the precomputed L2 table built here per
// pair, for PQ sub-vectors of SUB floats (1, 2; 4 for 3 and 4). The 32 pairs'
// LUT columns stay in registers (NK x 16 bytes a lane) for all the list's
// blocks.
template <int NK, bool IP, int SUB>
__device__ __forceinline__ void mfma_unit(const SumsArgs& a, const PairPlan& pp, unsigned char* mine,
                                          const unsigned char* onehot, uint32_t u, int lane) {
    const DevIvf& ix = a.ix;
    const int m = ix.m, lb = ix.m4 * 16;
    const int list = static_cast<int>(pp.unit_list[u]);
    const uint32_t first = pp.pair_off[list] + (u - pp.unit_off[list]) * 32;
    const int npairs = static_cast<int>(min(32u, pp.pair_off[list + 1] - first));
    const int j = lane & 31, h = lane >> 5;
    const bool valid = j < npairs;
    // lanes past the unit's pairs work on its first pair and store nothing
    const uint32_t item = pp.items[first + (valid ? j : 0)];
    const int qi = static_cast<int>(item / a.M);

    // the pair's LUT rows 2t + h, as MFMA B operands
    v4i_t lut[NK];
    float scale = 1.0f, offset = 0.0f;
    if (IP) {
        const uint8_t* src = a.ip_lut + static_cast<size_t>(qi) * m * 16;
#pragma unroll
        for (int t = 0; t < NK; ++t) {
            const int r = 2 * t + h;
            v4i_t v = {0, 0, 0, 0};
            if (r < m) v = *reinterpret_cast<const v4i_t*>(src + r * 16);
            lut[t] = v;
        }
        scale = a.ip_scale[qi];
        offset = a.ip_offset[qi] + a.probe_base[item];
    } else {
        // one table row per lane (row r = lane): the list's B row and the PQ
        // transpose's rows of subspace r stay in registers for all 32 pairs
        const int sub = ix.sub, r = lane;
        const bool row = r < m;
        // loads are unconditional, from row 0 and entry sub - 1 where the
        // lane or k is past the table, so that they issue together (a guarded
        // load waits for itself before the next one)
        const int rs = row ? r : 0;
        float brow[16], crow[SUB][16];
        const float* Bl = ix.B + static_cast<size_t>(list) * m * 16;
#pragma unroll
        for (int c = 0; c < 16; ++c) brow[c] = Bl[rs * 16 + c];
#pragma unroll
        for (int k = 0; k < SUB; ++k)
#pragma unroll
            for (int c = 0; c < 16; ++c) crow[k][c] = ix.pq_cent_t[(rs * sub + min(k, sub - 1)) * 16 + c];
#pragma unroll
        for (int c = 0; c < 16; ++c) brow[c] = row ? brow[c] : 0.0f;
#pragma unroll
        for (int k = 0; k < SUB; ++k)
#pragma unroll
            for (int c = 0; c < 16; ++c) crow[k][c] = row && k < sub ? crow[k][c] : 0.0f;
        uint8_t* rows16 = mine;                                                // 16 pairs x m4 rows x 16 bytes
        float* rmin = reinterpret_cast<float*>(mine + static_cast<size_t>(16) * lb); // 16 pairs x 64 rows
        float* so = rmin + 16 * 64;                                            // 32 x (scale, offset)
        const float pairA = a.probe_A[item];
        // each lane's entries of the query, for PF pairs at a time, loaded together
        constexpr int PF = SUB <= 2 ? 16 : 4;
#pragma unroll
        for (int t = 0; t < NK; ++t) lut[t] = v4i_t{0, 0, 0, 0};
        for (int half = 0; half < 2; ++half) {
            for (int p0 = 0; p0 < 16; p0 += PF) {
                float qv[PF][SUB];
#pragma unroll
                for (int pl = 0; pl < PF; ++pl) {
                    const uint32_t it = static_cast<uint32_t>(
                        __builtin_amdgcn_readlane(static_cast<int>(item), half * 16 + p0 + pl));
                    const float* q = a.q + static_cast<size_t>(it / a.M) * ix.dim + rs * sub;
#pragma unroll
                    for (int k = 0; k < SUB; ++k) qv[pl][k] = q[min(k, sub - 1)];
                }
#pragma unroll
                for (int pl = 0; pl < PF; ++pl)
#pragma unroll
                    for (int k = 0; k < SUB; ++k) qv[pl][k] = row && k < sub ? qv[pl][k] : 0.0f;
#pragma unroll
                for (int pl = 0; pl < PF; ++pl) {
                    const int pk = p0 + pl, pj = half * 16 + pk;
                    const float A = lane_f32(pairA, pj);
                    // row r of the table: B + C (+ A on row 0), C = compute_inner_table
                    // (src/pq.cpp:186-204) as k_query_tables makes it
                    float e[16];
#pragma unroll
                    for (int c = 0; c < 16; ++c) e[c] = 0.0f;
#pragma unroll
                    for (int k = 0; k < SUB; ++k) {
                        if (SUB <= 2 || k < sub) {
                            const float m2q = -2.0f * qv[pl][k];
#pragma unroll
                            for (int c = 0; c < 16; ++c) e[c] = fmaf(m2q, crow[k][c], e[c]);
                        }
                    }
#pragma unroll
                    for (int c = 0; c < 16; ++c) e[c] = brow[c] + e[c];
                    if (r == 0) {
#pragma unroll
                        for (int c = 0; c < 16; ++c) e[c] = e[c] + A;
                    }
                    float rlo, rhi;
                    row_lo_hi_wave(e, row, rlo, rhi);
                    // the widest row by the bits of the positive widths (ordered as
                    // the floats are): a width of 0, below or NaN leaves the scale
                    // what fmaxf over them gives, so it counts as 0
                    const float wd = rhi - rlo;
                    const uint32_t widest = wave_max_u32_dpp(row && wd > 0.0f ? __float_as_uint(wd) : 0u);
                    const float sc = widest ? __fdiv_rn(kLutMax, __uint_as_float(widest)) : 1.0f;
                    rmin[pk * 64 + r] = rlo;
                    if (lane == 0) so[2 * pj] = sc;
                    // rows m .. m4 - 1 hold zeros (e and rlo are 0 there), so bytes 0
                    if (lane < ix.m4) {
                        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int c = 0; c < 4; ++c)
                                w[i] = __builtin_amdgcn_cvt_pk_u8_f32((e[4 * i + c] - rlo) * sc, c, w[i]);
                        *reinterpret_cast<uint4*>(rows16 + static_cast<size_t>(pk) * lb + r * 16) =
                            make_uint4(w[0], w[1], w[2], w[3]);
                    }
                }
            }
            wave_lds_sync();
            // the half's offsets: lane 4 pk + i sums pair pk's row minimums
            // of rows i, i + 4, ... in row order, then (p0 + p1) + (p2 + p3)
            {
                const int pk = lane >> 2;
                float part = 0.0f;
                for (int i = lane & 3; i < m; i += 4) part += rmin[pk * 64 + i];
                const int b4 = lane & ~3;
                const float s0 = __shfl(part, b4, kWave), s1 = __shfl(part, b4 + 1, kWave);
                const float s2 = __shfl(part, b4 + 2, kWave), s3 = __shfl(part, b4 + 3, kWave);
                if ((lane & 3) == 0) so[2 * (half * 16 + pk) + 1] = (s0 + s1) + (s2 + s3);
            }
            // this half's pairs take their rows 2t + h as operands
            if ((j >> 4) == half) {
                const unsigned char* src = rows16 + static_cast<size_t>(j & 15) * lb;
#pragma unroll
                for (int t = 0; t < NK; ++t) lut[t] = *reinterpret_cast<const v4i_t*>(src + (2 * t + h) * 16);
            }
            wave_lds_sync();
        }
        scale = so[2 * j];
        offset = so[2 * j + 1];
        wave_lds_sync(); // the LDS is the blocks' from here on
    }

    // the list's blocks, each staged in LDS a block ahead
    const uint32_t n = ix.list_off[list + 1] - ix.list_off[list];
    const uint32_t nblk = (n + kFsBlock - 1) / kFsBlock;
    const uint8_t* codes = ix.codes + ix.code_off[list];
    const bool loader = lane * 16 < lb;
    unsigned char* stage = mine;
    // Block loads are unconditional, of block min(b, nblk - 1) and column
    // min(lane * 16, lb - 16), so that none waits at a branch join; only the
    // loader lanes stage what they load. Blocks 1, 2 and 3 wait in registers:
    // each block's load is made three blocks before the loop stages it.
    const uint32_t last = nblk > 0 ? nblk - 1 : 0;
    const int col = min(lane * 16, lb - 16);
    auto load_block = [&](uint32_t bb) {
        return *reinterpret_cast<const uint4*>(codes + static_cast<size_t>(min(bb, last)) * lb + col);
    };
    uint4 n1 = make_uint4(0u, 0u, 0u, 0u), n2 = n1, n3 = n1;
    if (nblk > 0) {
        const uint4 c = load_block(0);
        if (loader) *reinterpret_cast<uint4*>(stage + lane * 16) = c;
        n1 = load_block(1);
        n2 = load_block(2);
        n3 = load_block(3);
    }
    const uint32_t blk0 = a.item_blk[item];
    uint16_t* qsums = a.sums + static_cast<size_t>(qi) * a.stride_slots + static_cast<size_t>(blk0) * kFsBlock;
    uint16_t* qbmin = a.bmin + static_cast<size_t>(qi) * (a.stride_slots / kFsBlock) + blk0;
    // vector v = lane % 32 of a block: byte v % 16 of each subspace row, the
    // low nibble for v < 16; in the device layout, word (s, g = (v % 16) / 4)
    // sits at (s / 4) * 64 + g * 16 + (s % 4) * 4
    const int v = lane & 31, g = (v & 15) >> 2, shift = 8 * (v & 3) + 4 * (v >> 4);
    uint32_t run_min = 0xFFFFFFFFu;
    // block b: its MFMAs, then nb (the register holding block b + 1) staged
    // and loaded again with block b + 4, then the sums stored
    auto block = [&](uint32_t b, uint4& nb) {
        const unsigned char* buf = stage + (b & 1) * lb;
        wave_lds_sync();
        // two accumulators (even and odd subspace groups) halve the chain of dependent MFMAs
        v16i_t acc0 = {}, acc1 = {};
#pragma unroll
        for (int s4 = 0; s4 < NK / 2; ++s4) {
            // subspaces 4 s4 + h and 4 s4 + 2 + h
            const uint32_t* wp = reinterpret_cast<const uint32_t*>(buf + s4 * 64 + g * 16) + h;
            const uint32_t w0 = wp[0], w1 = wp[2];
            const v4i_t a0 = *reinterpret_cast<const v4i_t*>(onehot + (__builtin_amdgcn_ubfe(w0, shift, 4) << 4));
            const v4i_t a1 = *reinterpret_cast<const v4i_t*>(onehot + (__builtin_amdgcn_ubfe(w1, shift, 4) << 4));
            if (s4 & 1) {
                acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, lut[2 * s4], acc1, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, lut[2 * s4 + 1], acc1, 0, 0, 0);
            } else {
                acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, lut[2 * s4], acc0, 0, 0, 0);
                acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, lut[2 * s4 + 1], acc0, 0, 0, 0);
            }
        }
        if (b + 1 < nblk && loader) *reinterpret_cast<uint4*>(stage + ((b + 1) & 1) * lb + lane * 16) = nb;
        nb = load_block(b + 4);
        // lane (pair j, half h) holds vectors 8 (i / 4) + 4 h + i % 4 of the
        // block. Stores are unconditional (no branch, so that the waits for
        // the block loads stay exact): lanes past the unit's pairs hold its
        // first pair and store that pair's values again.
        uint32_t sv[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) sv[i] = static_cast<uint32_t>(acc0[i] + acc1[i]);
#pragma unroll
        for (int v4 = 0; v4 < 4; ++v4)
            *reinterpret_cast<uint2*>(qsums + static_cast<size_t>(b) * kFsBlock + 8 * v4 + 4 * h) =
                make_uint2(sv[4 * v4] | (sv[4 * v4 + 1] << 16), sv[4 * v4 + 2] | (sv[4 * v4 + 3] << 16));
        // the block's smallest sum over its vectors (all 32 but in a list's last block)
        uint32_t vmin = 0xFFFFFFFFu;
        const uint32_t left = n - b * kFsBlock;
        if (left >= kFsBlock) {
#pragma unroll
            for (int i = 0; i < 16; ++i) vmin = min(vmin, sv[i]);
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if (static_cast<uint32_t>(8 * (i / 4) + 4 * h + i % 4) < left) vmin = min(vmin, sv[i]);
        }
        vmin = min(vmin, static_cast<uint32_t>(__shfl_xor(static_cast<int>(vmin), 32, kWave)));
        qbmin[b] = static_cast<uint16_t>(min(vmin, 0xFFFFu)); // the same value from both halves
        run_min = min(run_min, vmin);
    };
    // unrolled by three so that each of n1, n2, n3 keeps one role: block b
    // stages the register that holds block b + 1 (b + 1 = 1, 2, 0 mod 3).
    // The first three are peeled, so that the loop is entered with the loads
    // and stores in flight that its back edge has, and its waits for each
    // block's load leave the two younger loads in flight.
    uint32_t b = 0;
    if (nblk >= 3) {
        block(0, n1);
        block(1, n2);
        block(2, n3);
        for (b = 3; b + 3 <= nblk; b += 3) {
            block(b, n1);
            block(b + 1, n2);
            block(b + 2, n3);
        }
    }
    if (b < nblk) block(b, n1);
    if (b + 1 < nblk) block(b + 1, n2);
    if (valid && h == 0) {
        a.item_min[item] = run_min;
        a.item_scale[item] = scale;
        a.item_offset[item] = offset;
    }
}

// One wave per unit. Workgroups are laid out XCD by XCD (workgroup b runs
// on XCD b % 8), so consecutive units -- mostly one list's -- share an XCD's
// L2.
template <int NK, bool IP, int SUB>
__global__ void __launch_bounds__(kMfmaWaves * kWave, 2) k_fs_sums_mfma(const SumsArgs a, const PairPlan pp) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    unsigned char* mine = lds + static_cast<size_t>(wave) * mfma_wave_lds(a.ix.m4);
    const uint32_t groups = gridDim.x / 8;
    const uint32_t wg = (blockIdx.x % 8) * groups + blockIdx.x / 8;
    const uint32_t u = wg * kMfmaWaves + wave;
    if (u >= pp.unit_off[a.ix.nlist]) return;
    // one-hot row c (byte c is 1): dword lane % 4 of row lane / 4; the block
    // loop syncs before it reads them
    unsigned char* onehot = mine + mfma_wave_lds(a.ix.m4) - 256;
    reinterpret_cast<uint32_t*>(onehot)[lane] = (lane & 3) == (lane >> 4) ? 1u << (8 * ((lane >> 2) & 3)) : 0u;
    mfma_unit<NK, IP, SUB>(a, pp, mine, onehot, u, lane);
}

// Phase 1 for lists stored as PQ codes (ADC), SQ codes or raw vectors: one
// workgroup per (query, window of consecutive probes), one thread per vector.
// The distance of each vector is what scan pushes for it
// (src/ivf_index.cpp:552-611): the ADC sum over subspaces of the (query, list)
// table, in subspace order (PQCodebook::distances_adc, include/vecengine/
// pq.hpp:75-98), plus the list's base for inner product; SQ::distance /
// -SQ::inner_product; or exact_score on the raw row. A vector the query may
// not return (removed, filtered out, deleted) gets +inf, which never passes
// TopK's `dist < threshold`. Each block's and each probe's smallest distance
// (as fkey) let the replay skip what cannot be pushed.
__global__ void __launch_bounds__(kSumsWaves * kWave) k_ivf_dists(const DistArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t w_start[kSumsWin + 1];
    __shared__ int w_list[kSumsWin];
    __shared__ uint32_t w_min[kSumsWin];
    __shared__ float w_base[kSumsWin];
    const DevIvf& ix = a.ix;
    const int m = ix.m, K = ix.K, dim = ix.dim, win = a.win, mode = a.mode;
    const int qi = blockIdx.x / a.nwin, p0 = (blockIdx.x % a.nwin) * win;
    const int P = a.nprobe_q[qi];
    if (p0 >= P) return;
    const uint64_t* filter = a.pq.filter ? a.pq.filter[qi] : a.filter;
    const int np = min(win, P - p0);
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const size_t item0 = static_cast<size_t>(qi) * a.M + p0;
    const bool adc = mode <= kDistL2Resid || mode == kDistSdc; // a table summed over the codes
    const size_t tsz = static_cast<size_t>(m) * K; // table floats
    // the query (and SQ's vmin and step) in LDS up to kQueryLdsDims floats,
    // else read from global memory
    const int qfl = static_cast<int>(query_lds_bytes(dim) / 4);
    float* qsl = reinterpret_cast<float*>(lds);                     // qfl
    float* sqv = qsl + qfl;                                         // SQ: vmin, step (2 qfl)
    float* tables = sqv + (mode == kDistSq ? 2 * qfl : 0);          // ADC: 1 (inner product) or win tables of mc K
    float* resid = tables + (adc ? static_cast<size_t>(mode == kDistIp ? 1 : win) * a.mc * K : 0); // L2Resid / Sdc: a residual per wave
    const float* q = a.q + static_cast<size_t>(qi) * dim;
    const float* qs = qfl ? qsl : q;
    const float* vmin = qfl ? sqv : ix.sq_vmin;
    const float* vstep = qfl ? sqv + dim : ix.sq_step;
    if (qfl) {
        for (int t = tid; t < dim; t += blockDim.x) qsl[t] = q[t];
        if (mode == kDistSq)
            for (int t = tid; t < dim; t += blockDim.x) {
                sqv[t] = ix.sq_vmin[t];
                sqv[dim + t] = ix.sq_step[t];
            }
    }
    const uint32_t blk0 = a.item_blk[item0];
    if (tid < np) {
        const int list = a.probe_idx[item0 + tid];
        w_list[tid] = list;
        w_min[tid] = 0xFFFFFFFFu;
        w_base[tid] = a.probe_base[item0 + tid];
        const uint32_t rel = a.item_blk[item0 + tid] - blk0;
        w_start[tid] = rel;
        if (tid == np - 1) w_start[np] = rel + (ix.list_off[list + 1] - ix.list_off[list] + kFsBlock - 1) / kFsBlock;
    }
    float* qd = a.dists + static_cast<size_t>(qi) * a.stride_slots + static_cast<size_t>(blk0) * kFsBlock;
    uint32_t* qbmin = a.bmin + static_cast<size_t>(qi) * (a.stride_slots / kFsBlock) + blk0;
    const int v = tid & 31;
    // ADC tables hold mc subspaces at a time (all m unless a table does not
    // fit LDS): each vector's sum runs over the subspaces in order, carried
    // from one run of subspaces to the next through its slot in dists
    const int mc = adc ? a.mc : 1;
    int s0 = 0;
    do {
        const int s1 = adc ? min(m, s0 + mc) : 0;
        const size_t csz = static_cast<size_t>(s1 - s0) * K; // this run's table floats
        const bool last = !adc || s1 == m;
        __syncthreads(); // the previous run's tables are read
        if (mode == kDistIp)
            for (size_t e = tid; e < csz; e += blockDim.x)
                tables[e] = a.C[static_cast<size_t>(qi) * tsz + static_cast<size_t>(s0) * K + e];
        if (mode == kDistL2Pre || mode == kDistL2Resid || mode == kDistSdc) {
            // the (query, list) L2 table: B + C (+ A on row 0), compute_adc_table of
            // q - c, or for SDC the rows of the SDC table picked by the code of q - c
            for (int p = wave; p < np; p += kSumsWaves) {
                if (w_start[p + 1] == w_start[p]) continue;
                const int list = w_list[p];
                float* tb = tables + static_cast<size_t>(p) * csz;
                if (mode == kDistL2Pre) {
                    const float* B = ix.B + static_cast<size_t>(list) * tsz + static_cast<size_t>(s0) * K;
                    const float* Cq = a.C + static_cast<size_t>(qi) * tsz + static_cast<size_t>(s0) * K;
                    const float A = a.probe_A[item0 + p];
                    for (size_t e = lane; e < csz; e += kWave) {
                        float t = B[e] + Cq[e];
                        if (s0 == 0 && e < static_cast<size_t>(K)) t = t + A;
                        tb[e] = t;
                    }
                } else {
                    const float* c = (ix.residual_centers ? ix.residual_centers : ix.centroids) + static_cast<size_t>(list) * dim;
                    float* r = a.resid_in_lds ? resid + wave * dim : nullptr;
                    if (r) {
                        for (int t = lane; t < dim; t += kWave) r[t] = qs[t] - c[t];
                        wave_lds_sync();
                    }
                    if (mode == kDistSdc) {
                        // PQCodebook::encode of the residual (src/pq.cpp:88-156): per
                        // subspace the first centroid of least (r_0 - y_0)^2 +
                        // fma-accumulated (r_j - y_j)^2, compared as int32 bits as
                        // the reference does; lanes over centroids, then the wave's
                        // minimum by (bits, index), the same on every lane
                        for (int sidx = s0; sidx < s1; ++sidx) {
                            const int x0 = sidx * ix.sub;
                            int bd = 0x7FFFFFFF;
                            int bc = K;
                            for (int cc = lane; cc < K; cc += kWave) {
                                float dd = (r ? r[x0] : qs[x0] - c[x0]) - ix.pq_cent_t[static_cast<size_t>(x0) * K + cc];
                                float acc = dd * dd;
                                for (int j = 1; j < ix.sub; ++j) {
                                    const int x = x0 + j;
                                    dd = (r ? r[x] : qs[x] - c[x]) - ix.pq_cent_t[static_cast<size_t>(x) * K + cc];
                                    acc = fmaf(dd, dd, acc);
                                }
                                if (cc == lane || __float_as_int(acc) < bd) {
                                    bd = __float_as_int(acc);
                                    bc = cc;
                                }
                            }
                            for (int off = kWave / 2; off > 0; off >>= 1) {
                                const int od = __shfl_xor(bd, off);
                                const int oc = __shfl_xor(bc, off);
                                if (od < bd || (od == bd && oc < bc)) {
                                    bd = od;
                                    bc = oc;
                                }
                            }
                            const float* row = ix.sdc + (static_cast<size_t>(sidx) * K + bc) * K;
                            for (int cc = lane; cc < K; cc += kWave) tb[static_cast<size_t>(sidx - s0) * K + cc] = row[cc];
                        }
                    } else {
                        for (size_t e = lane; e < csz; e += kWave) {
                            const int sidx = s0 + static_cast<int>(e / K), cc = static_cast<int>(e % K);
                            float acc = 0.0f;
                            for (int j = 0; j < ix.sub; ++j) {
                                const int x = sidx * ix.sub + j;
                                const float rs = r ? r[x] : qs[x] - c[x];
                                const float d = rs - ix.pq_cent_t[static_cast<size_t>(x) * K + cc];
                                acc = fmaf(d, d, acc);
                            }
                            tb[e] = acc;
                        }
                    }
                    wave_lds_sync();
                }
            }
        }
        __syncthreads();

        const uint32_t nT = w_start[np];
        for (uint32_t base = 0; base < nT; base += blockDim.x / kFsBlock) {
            const uint32_t t = base + (tid >> 5);
            uint32_t key = 0xFFFFFFFFu;
            int p = 0;
            if (t < nT) {
#pragma unroll
                for (int step = kSumsWin / 2; step > 0; step >>= 1) {
                    const int c = min(p + step, np - 1);
                    p = w_start[c] <= t ? c : p;
                }
                const int list = w_list[p];
                const uint32_t n = ix.list_off[list + 1] - ix.list_off[list];
                const uint32_t j = (t - w_start[p]) * kFsBlock + v;
                float dist = INFINITY;
                bool scored = false;
                float* out = qd + static_cast<size_t>(t) * kFsBlock + v;
                if (j < n) {
                    const uint32_t entry = ix.list_off[list] + j;
                    const uint32_t slot = ix.list_ids[entry];
                    const bool ok = slot != 0xFFFFFFFFu &&
                                    (filter ? bit_test(filter, slot) : (ix.deleted == nullptr || !bit_test(ix.deleted, slot)));
                    if (ok) {
                        if (adc) {
                            const uint8_t* code = ix.codes + static_cast<size_t>(entry) * m;
                            const float* tb = tables + (mode == kDistIp ? 0 : static_cast<size_t>(p) * csz);
                            float d = s0 == 0 ? 0.0f : *out;
                            for (int sidx = s0; sidx < s1; ++sidx) d += tb[(sidx - s0) * K + code[sidx]];
                            dist = mode == kDistIp && last ? d + w_base[p] : d;
                        } else if (mode == kDistSq) {
                            int cs;
                            const uint8_t* code = sq_row(ix.codes, entry, ix.vbytes, ix.sq_chunked, cs);
                            dist = ix.metric == 2 ? -sq_score<true>(qs, code, vmin, vstep, dim, ix.sq_bits, cs)
                                                  : sq_score<false>(qs, code, vmin, vstep, dim, ix.sq_bits, cs);
                        } else {
                            const float* row = reinterpret_cast<const float*>(ix.codes) + static_cast<size_t>(entry) * dim;
                            dist = exact_score(ix.metric == 2, qs, row, dim);
                        }
                        key = fkey(dist);
                        scored = true;
                    }
                }
                // a run before the last leaves the partial sums; the last the distances
                if (last || scored) *out = dist;
            }
            if (last) {
                // the block's 32 threads: their smallest key
                for (int off = 16; off > 0; off >>= 1)
                    key = min(key, static_cast<uint32_t>(__shfl_xor(static_cast<int>(key), off, 32)));
                if (t < nT && v == 0) {
                    qbmin[t] = key;
                    atomicMin(&w_min[p], key);
                }
            }
        }
        s0 = s1;
    } while (s0 < m);
    __syncthreads();
    if (tid < np) a.item_min[item0 + tid] = w_min[tid];
}

// Phase 1 for 8-bit SQ lists stored chunk-major (sq_row): the same work item
// and outputs as k_ivf_dists. Each lane scores kSq8Rows vectors, vector
// v = lane % 32 of blocks g + 2r + lane / 32, one 16-byte chunk at a time: a
// wave's load of chunk c covers 64 consecutive entries' chunks, contiguous
// 512-byte runs, and every LDS read of q / vmin / step serves kSq8Rows
// vectors. Each distance keeps SQ::distance's order (src/sq.cpp:126-211), as
// sq_score does: 8 accumulators over dimensions d = l (mod 8),
// v = fma(code, step, vmin), then (q - v)^2 or q * v fused, folded as
// ((a0+a4)+(a2+a6)) + ((a1+a5)+(a3+a7)).
constexpr int kSq8Rows = 2;
template <bool IP>
__global__ void __launch_bounds__(kSumsWaves * kWave) k_sq8_dists(const DistArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t w_start[kSumsWin + 1];
    __shared__ int w_list[kSumsWin];
    __shared__ uint32_t w_min[kSumsWin];
    __shared__ uint32_t w_off[kSumsWin], w_len[kSumsWin];
    const DevIvf& ix = a.ix;
    const int dim = ix.dim, win = a.win;
    const int qi = blockIdx.x / a.nwin, p0 = (blockIdx.x % a.nwin) * win;
    const int P = a.nprobe_q[qi];
    if (p0 >= P) return;
    const uint64_t* filter = a.pq.filter ? a.pq.filter[qi] : a.filter;
    const int np = min(win, P - p0);
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const size_t item0 = static_cast<size_t>(qi) * a.M + p0;
    // q, vmin, step: dim each in LDS up to kQueryLdsDims floats, else read
    // from global memory (16-byte aligned: dim is a multiple of 16)
    const float* q = a.q + static_cast<size_t>(qi) * dim;
    const bool qcopy = dim <= kQueryLdsDims;
    float* qsl = reinterpret_cast<float*>(lds);
    const float* qs = qcopy ? qsl : q;
    const float* vmin = qcopy ? qsl + dim : ix.sq_vmin;
    const float* step = qcopy ? qsl + 2 * dim : ix.sq_step;
    if (qcopy)
        for (int t = tid; t < dim; t += blockDim.x) {
            qsl[t] = q[t];
            qsl[dim + t] = ix.sq_vmin[t];
            qsl[2 * dim + t] = ix.sq_step[t];
        }
    const uint32_t blk0 = a.item_blk[item0];
    if (tid < np) {
        const int list = a.probe_idx[item0 + tid];
        const uint32_t off = ix.list_off[list], n = ix.list_off[list + 1] - off;
        w_list[tid] = list;
        w_off[tid] = off;
        w_len[tid] = n;
        w_min[tid] = 0xFFFFFFFFu;
        const uint32_t rel = a.item_blk[item0 + tid] - blk0;
        w_start[tid] = rel;
        if (tid == np - 1) w_start[np] = rel + (n + kFsBlock - 1) / kFsBlock;
    }
    __syncthreads();

    const uint32_t T = w_start[np];
    float* qd = a.dists + static_cast<size_t>(qi) * a.stride_slots + static_cast<size_t>(blk0) * kFsBlock;
    uint32_t* qbmin = a.bmin + static_cast<size_t>(qi) * (a.stride_slots / kFsBlock) + blk0;
    const int v = lane & 31;
    constexpr uint32_t kGroup = 2 * kSq8Rows; // blocks of one wave's pass
    for (uint32_t g = wave * kGroup; g < T; g += kSumsWaves * kGroup) {
        const uint4* row[kSq8Rows];
        uint32_t t[kSq8Rows], entry[kSq8Rows];
        int pr[kSq8Rows];
        bool ok[kSq8Rows];
#pragma unroll
        for (int r = 0; r < kSq8Rows; ++r) {
            t[r] = g + 2 * r + (lane >> 5);
            // a block past the window's end scores a valid row of the last block, unused
            const uint32_t tt = min(t[r], T - 1);
            int p = 0;
#pragma unroll
            for (int s2 = kSumsWin / 2; s2 > 0; s2 >>= 1) {
                const int c = min(p + s2, np - 1);
                p = w_start[c] <= tt ? c : p;
            }
            pr[r] = p;
            const uint32_t n = w_len[p]; // >= 1: block tt is probe p's
            const uint32_t j = (tt - w_start[p]) * kFsBlock + v;
            entry[r] = w_off[p] + min(j, n - 1);
            ok[r] = t[r] < T && j < n;
            int cs;
            row[r] = reinterpret_cast<const uint4*>(sq_row(ix.codes, entry[r], dim, 1, cs));
        }
        // every row's slot, then its filter or deleted bit, loaded together
        uint32_t slot[kSq8Rows];
#pragma unroll
        for (int r = 0; r < kSq8Rows; ++r) slot[r] = ix.list_ids[entry[r]];
        const uint64_t* bits = filter ? filter : ix.deleted;
        if (bits) {
            uint64_t word[kSq8Rows];
#pragma unroll
            for (int r = 0; r < kSq8Rows; ++r) word[r] = bits[(slot[r] == 0xFFFFFFFFu ? 0u : slot[r]) >> 6];
#pragma unroll
            for (int r = 0; r < kSq8Rows; ++r) {
                const bool bit = (word[r] >> (slot[r] & 63)) & 1ull;
                ok[r] = ok[r] && (filter ? bit : !bit);
            }
        }
#pragma unroll
        for (int r = 0; r < kSq8Rows; ++r) ok[r] = ok[r] && slot[r] != 0xFFFFFFFFu;
        float acc[kSq8Rows][8];
#pragma unroll
        for (int r = 0; r < kSq8Rows; ++r)
#pragma unroll
            for (int l = 0; l < 8; ++l) acc[r][l] = 0.0f;
        // 16 bytes of each row: chunk c, dimensions 16c .. 16c + 15
        auto score = [&](const uint4 (&cw)[kSq8Rows], int c) {
#pragma unroll
            for (int w4 = 0; w4 < 4; ++w4) {
                const int d0 = c * 16 + w4 * 4;
                const float4 q4 = *reinterpret_cast<const float4*>(qs + d0);
                const float4 m4 = *reinterpret_cast<const float4*>(vmin + d0);
                const float4 s4 = *reinterpret_cast<const float4*>(step + d0);
                const float qv[4] = {q4.x, q4.y, q4.z, q4.w};
                const float mv[4] = {m4.x, m4.y, m4.z, m4.w};
                const float sv[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
                for (int r = 0; r < kSq8Rows; ++r) {
                    const uint32_t word = w4 == 0 ? cw[r].x : w4 == 1 ? cw[r].y : w4 == 2 ? cw[r].z : cw[r].w;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int l = (w4 * 4 + b) & 7;
                        const float val = fmaf(static_cast<float>((word >> (8 * b)) & 0xFFu), sv[b], mv[b]);
                        if (IP) {
                            acc[r][l] = fmaf(qv[b], val, acc[r][l]);
                        } else {
                            const float diff = qv[b] - val;
                            acc[r][l] = fmaf(diff, diff, acc[r][l]);
                        }
                    }
                }
            }
        };
        // two register buffers: chunk c + 1 of every row is in flight while
        // chunk c is scored (512 bytes between chunks; the empty asm keeps each
        // prefetch where it is, which would otherwise be sunk to its use; a
        // last pass of an even count reloads the final chunk, unused)
        const int nc = dim / 16;
        uint4 ca[kSq8Rows], cb[kSq8Rows];
#pragma unroll
        for (int r = 0; r < kSq8Rows; ++r) ca[r] = row[r][0];
        int c = 0;
        for (; c + 2 <= nc; c += 2) {
#pragma unroll
            for (int r = 0; r < kSq8Rows; ++r) cb[r] = row[r][(c + 1) * 32];
            asm volatile("" ::: "memory");
            score(ca, c);
            const int cn = min(c + 2, nc - 1);
#pragma unroll
            for (int r = 0; r < kSq8Rows; ++r) ca[r] = row[r][cn * 32];
            asm volatile("" ::: "memory");
            score(cb, c + 1);
        }
        if (c < nc) score(ca, c); // an odd count's last chunk
#pragma unroll
        for (int r = 0; r < kSq8Rows; ++r) {
            float h[4];
#pragma unroll
            for (int l = 0; l < 4; ++l) h[l] = acc[r][l] + acc[r][l + 4];
            const float sum = (h[0] + h[2]) + (h[1] + h[3]);
            const float dist = ok[r] ? (IP ? -sum : sum) : INFINITY;
            uint32_t key = ok[r] ? fkey(dist) : 0xFFFFFFFFu;
            if (t[r] < T) qd[static_cast<size_t>(t[r]) * kFsBlock + v] = dist;
            // the block's 32 lanes: their smallest key
            for (int off = 16; off > 0; off >>= 1)
                key = min(key, static_cast<uint32_t>(__shfl_xor(static_cast<int>(key), off, 32)));
            if (t[r] < T && v == 0) {
                qbmin[t[r]] = key;
                atomicMin(&w_min[pr[r]], key);
            }
        }
    }
    __syncthreads();
    if (tid < np) a.item_min[item0 + tid] = w_min[tid];
}

// query_order: a counting sort of the queries by their nearest probed list.
__global__ void k_order_zero(uint32_t* hist, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hist[i] = 0u;
}

__global__ void k_order_count(const uint16_t* probe_idx, int M, int nq, uint32_t* hist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) atomicAdd(&hist[probe_idx[static_cast<size_t>(i) * M]], 1u);
}

// Exclusive prefix sum of the nlist counters, in place, by one 1024-thread block.
__global__ void __launch_bounds__(1024) k_order_scan(uint32_t* hist, int n) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x, per = (n + 1023) / 1024;
    const int b = min(n, t * per), e = min(n, b + per);
    uint32_t sum = 0;
    for (int i = b; i < e; ++i) sum += hist[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (int i = b; i < e; ++i) {
        const uint32_t c = hist[i];
        hist[i] = run;
        run += c;
    }
}

__global__ void k_order_scatter(const uint16_t* probe_idx, int M, int nq, uint32_t* offs, uint32_t* order) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) order[atomicAdd(&offs[probe_idx[static_cast<size_t>(i) * M]], 1u)] = static_cast<uint32_t>(i);
}

// Probes each query scans and where each probe's sums start: nprobe lists
// unfiltered; with a filter, at least min_probes and then until the lists
// scanned hold `wanted` allowed vectors (scan, ivf_index.cpp:510-521: before
// the TopK first trims, every allowed vector is pushed, so top.size() >=
// wanted exactly when that many allowed vectors were scanned).
__global__ void k_probe_plan(const DevIvf ix, int nq, int M, int nprobe, int min_probes, int wanted,
                             const uint16_t* probe_idx, const uint32_t* list_allowed, int* nprobe_q, uint32_t* item_blk,
                             int* status, const PerQuery pq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    if (pq.nprobe) {
        nprobe = pq.nprobe[i];
        min_probes = pq.min_probes[i];
        wanted = pq.wanted[i];
        list_allowed = pq.list_allowed[i];
    }
    const uint16_t* pl = probe_idx + static_cast<size_t>(i) * M;
    uint64_t seen = 0;
    uint32_t blk = 0;
    int p = 0, st = 0;
    for (;; ++p) {
        if (p >= nprobe && p >= min_probes && seen >= static_cast<uint64_t>(wanted)) break;
        if (p >= M) {
            if (p < ix.nlist) st = 1; // needs probes beyond those selected
            break;
        }
        const int list = pl[p];
        const uint32_t n = ix.list_off[list + 1] - ix.list_off[list];
        item_blk[static_cast<size_t>(i) * M + p] = blk;
        blk += (n + kFsBlock - 1) / kFsBlock;
        if (list_allowed) seen += list_allowed[list];
    }
    nprobe_q[i] = p;
    status[i] = st;
}

// k_probe_plan with one wave per query: 64 probes at a time, the block
// offsets and allowed counts by wave prefix sums, and the loop's stop rule
// (checked before probe p is taken: p >= nprobe, p >= min_probes and the
// allowed entries of probes before p at least `wanted`) found by a ballot.
__global__ void __launch_bounds__(256) k_probe_plan_wave(const DevIvf ix, int nq, int M, int nprobe, int min_probes,
                                                          int wanted, const uint16_t* probe_idx,
                                                          const uint32_t* list_allowed, int* nprobe_q,
                                                          uint32_t* item_blk, int* status, const PerQuery pq) {
    const int i = blockIdx.x * 4 + static_cast<int>(threadIdx.x) / kWave;
    const int lane = lane_id();
    if (i >= nq) return;
    if (pq.nprobe) {
        nprobe = pq.nprobe[i];
        min_probes = pq.min_probes[i];
        wanted = pq.wanted[i];
        list_allowed = pq.list_allowed[i];
    }
    const uint16_t* pl = probe_idx + static_cast<size_t>(i) * M;
    uint32_t blk = 0;  // blocks of the probes taken so far
    uint64_t seen = 0; // their allowed entries
    auto incl_scan = [&](uint32_t v) {
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(v), off, kWave));
            if (lane >= off) v += y;
        }
        return v;
    };
    for (int c0 = 0; c0 < M; c0 += kWave) {
        const int p = c0 + lane;
        uint32_t nb = 0, al = 0;
        if (p < M) {
            const int list = pl[p];
            nb = (ix.list_off[list + 1] - ix.list_off[list] + kFsBlock - 1) / kFsBlock;
            if (list_allowed) al = list_allowed[list];
        }
        const uint32_t inb = incl_scan(nb), ial = incl_scan(al);
        const uint64_t seen_before = seen + (ial - al);
        const bool stop = p < M && p >= nprobe && p >= min_probes && seen_before >= static_cast<uint64_t>(wanted);
        const uint64_t sm = ballot(stop);
        const int first = sm ? __builtin_ctzll(sm) : kWave;
        if (p < M && lane < first) item_blk[static_cast<size_t>(i) * M + p] = blk + (inb - nb);
        if (sm) {
            if (lane == 0) {
                nprobe_q[i] = c0 + first;
                status[i] = 0;
            }
            return;
        }
        blk += static_cast<uint32_t>(__shfl(static_cast<int>(inb), kWave - 1, kWave));
        seen += static_cast<uint32_t>(__shfl(static_cast<int>(ial), kWave - 1, kWave));
    }
    // every selected probe taken: the loop stops at M, needing more lists
    // when its rule does not hold there and lists remain
    if (lane == 0) {
        const bool stop = M >= nprobe && M >= min_probes && seen >= static_cast<uint64_t>(wanted);
        nprobe_q[i] = M;
        status[i] = (!stop && M < ix.nlist) ? 1 : 0;
    }
}

// Allowed entries of each list under a filter.
__global__ void k_list_allowed(const DevIvf ix, const uint64_t* filter, uint32_t* out) {
    const int l = blockIdx.x;
    const uint32_t b = ix.list_off[l], e = ix.list_off[l + 1];
    uint32_t c = 0;
    for (uint32_t j = b + threadIdx.x; j < e; j += blockDim.x) c += bit_test(filter, ix.list_ids[j]) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) c += static_cast<uint32_t>(__shfl_xor(static_cast<int>(c), off, kWave));
    __shared__ uint32_t part[16];
    if (lane_id() == 0) part[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) t += part[w];
        out[l] = t;
    }
}

// The end of a replay: take_sorted (src/topk.hpp:53-76) of the TopK (which
// holds slots) into the query's output row, or, for a re-rank
// (RefineIndex::rerank, src/refine_index.cpp:112-140, orders the candidates
// itself by (exact distance, slot)), just the set.
__device__ void replay_finish(WaveTopK& tk, const ReplayArgs& a, int qi) {
    const DevIvf& ix = a.ix;
    const int kl = a.kp;                            // row stride of the outputs
    const int kp = a.pq.kp ? a.pq.kp[qi] : a.kp;    // this query's capacity
    const int lane = lane_id();

    const int n = tk.size;
    if (a.unsorted) {
        float* od = a.out_dist + static_cast<size_t>(qi) * kl;
        int64_t* ol = a.out_label + static_cast<size_t>(qi) * kl;
        for (int i = lane; i < kp; i += kWave) {
            od[i] = i < n ? tk.dist[i] : INFINITY;
            ol[i] = i < n ? (ix.labels ? ix.labels[static_cast<uint32_t>(tk.label[i])] : tk.label[i]) : -1;
        }
        if (lane == 0) a.out_count[qi] = n;
        return;
    }
    // slots become ids; (key, id) order when every id fits 32 bits, else (float, id) order
    bool big = false;
    for (int base = 0; base < n; base += kWave) {
        const int i = base + lane;
        if (i < n && ix.labels) tk.label[i] = ix.labels[static_cast<uint32_t>(tk.label[i])];
        big |= __popcll(ballot(i < n && static_cast<uint64_t>(tk.label[i]) > 0xFFFFFFFFull)) != 0;
    }
    int np = 1;
    while (np < n) np <<= 1;
    for (int i = n + lane; i < np; i += kWave) {
        tk.dist[i] = __uint_as_float(0x7FC00000u);
        tk.label[i] = -1;
    }
    wave_sync();
    // padding entries (label -1) sort last wherever the network moves them
    auto sort_key = [&](int i, uint32_t& k32, uint64_t& id) {
        if (tk.label[i] < 0) {
            k32 = 0xFFFFFFFFu;
            id = ~0ull;
            return;
        }
        const float d = tk.dist[i];
        k32 = big ? ckey(d) : fkey(d);
        id = static_cast<uint64_t>(tk.label[i]);
    };
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < (np >> 1); t += kWave) {
                const int x = 2 * t - (t & (stride - 1));
                const int y = x + stride;
                const bool up = (x & size) == 0;
                uint32_t kx, ky;
                uint64_t ix_, iy;
                sort_key(x, kx, ix_);
                sort_key(y, ky, iy);
                const bool gt = kx != ky ? kx > ky : ix_ > iy;
                if (gt == up) {
                    const float dx = tk.dist[x], dy = tk.dist[y];
                    const int64_t lx = tk.label[x], ly = tk.label[y];
                    tk.dist[x] = dy;
                    tk.dist[y] = dx;
                    tk.label[x] = ly;
                    tk.label[y] = lx;
                }
            }
            wave_sync();
        }
    }
    float* od = a.out_dist + static_cast<size_t>(qi) * kl;
    int64_t* ol = a.out_label + static_cast<size_t>(qi) * kl;
    for (int i = lane; i < kp; i += kWave) {
        od[i] = i < n ? tk.dist[i] : INFINITY;
        ol[i] = i < n ? tk.label[i] : -1;
    }
    if (lane == 0) a.out_count[qi] = n;
}

// Phase 2: one wave per query replays the reference's scan in probe order.
// Fast-scan (FL false): over the u16 sums -- limit16 from the threshold at
// each block (scan_fast, ivf_index.cpp:482-485), survivors in bit order
// through push (:539-543). PQ / SQ / float lists (FL true): over the float
// distances, each pushed in list order when below the threshold (:552-611).
// Either way into the TopK (src/topk.hpp), trimmed when its buffer fills.
// The probes' metadata is read 64 at a time, one probe per lane. A list whose
// smallest sum (distance) cannot pass the current limit (threshold) is
// skipped whole, and within a list only blocks whose smallest can are read
// (the threshold only falls). The TopK holds slots; they become ids (labels)
// for the final sort. With STAGE, a list's candidate blocks have their sums
// and slots loaded 16 blocks at a time into LDS before they are replayed, so
// one query waits for one round of loads per 16 blocks rather than per block:
// the form for small batches, whose few waves cannot hide the latency.
template <bool FL, bool STAGE>
__device__ void fs_replay(const ReplayArgs& a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const DevIvf& ix = a.ix;
    const int qi = blockIdx.x;
    const int kl = a.kp;                         // buffer layout: the launch's largest capacity
    const int kp = a.pq.kp ? a.pq.kp[qi] : kl;   // this query's TopK capacity
    const uint64_t* filter = a.pq.filter ? a.pq.filter[qi] : a.filter;
    const int block_filter = a.pq.block_filter ? a.pq.block_filter[qi] : a.block_filter;
    const bool tk_lds = a.g_dist == nullptr;
    int64_t* tk_label = tk_lds ? reinterpret_cast<int64_t*>(lds) : a.g_label + static_cast<size_t>(qi) * 2 * kl;
    float* tk_dist = tk_lds ? reinterpret_cast<float*>(tk_label + 2 * kl) : a.g_dist + static_cast<size_t>(qi) * 2 * kl;
    const int lane = lane_id();
    const bool filtered = filter != nullptr;
    __shared__ uint32_t trim_hist[256];
    WaveTopK tk{tk_dist, tk_label, 0, kp, INFINITY, trim_hist};
    const int P = a.nprobe_q[qi];
    const size_t qoff = static_cast<size_t>(qi) * a.stride_slots;
    const size_t qboff = static_cast<size_t>(qi) * (a.stride_slots / kFsBlock);
    for (int c0 = 0; c0 < P; c0 += kWave) {
        int m_list = 0;
        uint32_t m_n = 0, m_blk = 0, m_min = 0xFFFFFFFFu;
        float m_scale = 1.0f, m_off = 0.0f;
        if (c0 + lane < P) {
            const size_t item = static_cast<size_t>(qi) * a.M + c0 + lane;
            m_list = a.probe_idx[item];
            m_n = ix.list_off[m_list + 1] - ix.list_off[m_list];
            m_blk = a.item_blk[item];
            m_min = a.item_min[item];
            if (!FL) {
                m_scale = a.item_scale[item];
                m_off = a.item_offset[item];
            }
        }
        const int cn = min(kWave, P - c0);
        for (int j = 0; j < cn; ++j) {
            const uint32_t n = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_n), j));
            if (n == 0) continue;
            const float scale = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m_scale), j));
            const float offset = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m_off), j));
            // the bound a block's or list's smallest must be below to hold a push
            auto limit_now = [&]() -> uint32_t {
                if (FL) return fkey(tk.thr);
                const float limit = (tk.thr - offset) * scale;
                return limit >= 65535.0f ? 65535u : limit <= 0.0f ? 0u : static_cast<uint32_t>(ceilf(limit));
            };
            if (static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_min), j)) >= limit_now()) continue;
            const int list = __builtin_amdgcn_readlane(m_list, j);
            const uint32_t blk0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_blk), j));
            const uint32_t* ids = ix.list_ids + ix.list_off[list];
            const size_t voff = qoff + static_cast<size_t>(blk0) * kFsBlock;
            const uint32_t nb = (n + kFsBlock - 1) / kFsBlock;
            if constexpr (STAGE) {
                constexpr int kStage = 16; // candidate blocks staged at a time
                __shared__ int st_blk[kWave];
                __shared__ uint32_t st_val[kStage * kFsBlock];
                __shared__ uint32_t st_id[kStage * kFsBlock];
                for (uint32_t bc = 0; bc < nb; bc += kWave) {
                    const uint32_t bl = bc + lane;
                    uint32_t bm = 0xFFFFFFFFu;
                    if (bl < nb) bm = FL ? a.fbmin[qboff + blk0 + bl] : a.bmin[qboff + blk0 + bl];
                    const uint64_t cand = ballot(bm < limit_now());
                    if (cand == 0) continue;
                    // the candidate blocks in order; their sums (distances) and slots
                    // are staged kStage blocks at a time, all loads of a stage in flight
                    const int nc = __popcll(cand);
                    if ((cand >> lane) & 1ull) st_blk[lanes_below(cand)] = lane;
                    wave_sync();
                    const int v = lane & 31;
                    for (int g0 = 0; g0 < nc; g0 += kStage) {
                        const int gn = min(kStage, nc - g0);
                        uint32_t sv[kStage / 2], iv[kStage / 2];
    #pragma unroll
                        for (int u = 0; u < kStage / 2; ++u) {
                            const int r = 2 * u + (lane >> 5);
                            sv[u] = FL ? __float_as_uint(INFINITY) : 0xFFFFu;
                            iv[u] = 0u;
                            if (r < gn) {
                                const uint32_t e = (bc + static_cast<uint32_t>(st_blk[g0 + r])) * kFsBlock + v;
                                if (e < n) {
                                    sv[u] = FL ? __float_as_uint(a.fdists[voff + e]) : static_cast<uint32_t>(a.sums[voff + e]);
                                    iv[u] = ids[e];
                                }
                            }
                        }
    #pragma unroll
                        for (int u = 0; u < kStage / 2; ++u) {
                            const int r = 2 * u + (lane >> 5);
                            st_val[r * kFsBlock + v] = sv[u];
                            st_id[r * kFsBlock + v] = iv[u];
                        }
                        wave_sync();
                        // two candidate blocks at a time: lanes 0-31 and 32-63
                        for (int r = 0; r < gn; r += 2) {
                            const int rr = r + (lane >> 5);
                            const bool has2 = r + 1 < gn;
                            const int mb = rr < gn ? st_blk[g0 + rr] : -1;
                            const uint32_t e = mb >= 0 ? (bc + static_cast<uint32_t>(mb)) * kFsBlock + v : 0u;
                            const bool in = mb >= 0 && e < n;
                            uint32_t sum = 0xFFFFu;
                            float fd = INFINITY;
                            uint32_t sid = 0u;
                            if (in) {
                                const uint32_t x = st_val[rr * kFsBlock + v];
                                if (FL) fd = __uint_as_float(x);
                                else sum = x;
                                sid = st_id[rr * kFsBlock + v];
                            }
                            auto passes_now = [&]() -> bool { return FL ? fd < tk.thr : sum < limit_now(); };
                            if (ballot(in && passes_now()) == 0) continue;
                            for (int h = 0; h < 2; ++h) {
                                if (h == 1 && !has2) break;
                                const bool mine = (lane >> 5) == h;
                                const bool valid = mine && in;
                                const uint32_t slot = sid;
                                bool allowed = valid;
                                if (!FL && block_filter) {
                                    allowed = valid && bit_test(filter, slot);
                                    if (ballot(allowed) == 0) continue;
                                }
                                const bool surv = allowed && passes_now();
                                uint64_t rem = ballot(surv);
                                if (rem == 0) continue;
                                float dist;
                                bool ok;
                                if (FL) {
                                    dist = fd; // +inf already marks what may not be returned
                                    ok = surv;
                                } else {
                                    dist = __fadd_rn(__fdiv_rn(static_cast<float>(sum), scale), offset);
                                    ok = surv && (filtered ? bit_test(filter, slot)
                                                           : (ix.deleted == nullptr || !bit_test(ix.deleted, slot)));
                                }
                                while (rem) {
                                    const bool cond = ((rem >> lane) & 1ull) && ok && dist < tk.thr;
                                    const uint64_t cm = ballot(cond);
                                    if (cm == 0) break;
                                    const int space = 2 * kp - tk.size;
                                    const int cnt = __popcll(cm);
                                    if (cnt < space) {
                                        if (cond) {
                                            const int pos = tk.size + static_cast<int>(lanes_below(cm));
                                            tk.dist[pos] = dist;
                                            tk.label[pos] = slot;
                                        }
                                        tk.size += cnt;
                                        wave_sync();
                                        break;
                                    }
                                    // the space-th candidate fills the buffer and triggers a trim
                                    const uint32_t rank = lanes_below(cm);
                                    const uint64_t at = ballot(cond && static_cast<int>(rank) == space - 1);
                                    const int bpos = __builtin_ctzll(at);
                                    if (cond && lane <= bpos) {
                                        const int pos = tk.size + static_cast<int>(rank);
                                        tk.dist[pos] = dist;
                                        tk.label[pos] = slot;
                                    }
                                    tk.size = 2 * kp;
                                    wave_sync();
                                    wave_trim(tk);
                                    rem &= (bpos >= 63) ? 0ull : (~0ull << (bpos + 1));
                                }
                            }
                        }
                        wave_sync(); // the next stage overwrites st_val / st_id
                    }
                }
            } else {
                for (uint32_t bc = 0; bc < nb; bc += kWave) {
                    const uint32_t bl = bc + lane;
                    uint32_t bm = 0xFFFFFFFFu;
                    if (bl < nb) bm = FL ? a.fbmin[qboff + blk0 + bl] : a.bmin[qboff + blk0 + bl];
                    uint64_t cand = ballot(bm < limit_now());
                    while (cand) {
                        // the next two candidate blocks: lanes 0-31 and 32-63
                        const int b1 = __builtin_ctzll(cand);
                        cand &= cand - 1;
                        int b2 = -1;
                        if (cand) {
                            b2 = __builtin_ctzll(cand);
                            cand &= cand - 1;
                        }
                        const int mb = (lane >> 5) == 0 ? b1 : b2;
                        const int v = lane & 31;
                        const uint32_t e = (bc + static_cast<uint32_t>(mb)) * kFsBlock + v;
                        const bool in = mb >= 0 && e < n;
                        uint32_t sum = 0xFFFFu;
                        float fd = INFINITY;
                        if (in) {
                            if (FL) fd = a.fdists[voff + e];
                            else sum = a.sums[voff + e];
                        }
                        auto passes_now = [&]() -> bool { return FL ? fd < tk.thr : sum < limit_now(); };
                        if (ballot(in && passes_now()) == 0) continue;
                        for (int h = 0; h < 2; ++h) {
                            if (h == 1 && b2 < 0) break;
                            const bool mine = (lane >> 5) == h;
                            const bool valid = mine && in;
                            uint32_t slot = 0;
                            bool allowed = valid;
                            if (!FL && block_filter) {
                                slot = valid ? ids[e] : 0u;
                                allowed = valid && bit_test(filter, slot);
                                if (ballot(allowed) == 0) continue;
                            }
                            const bool surv = allowed && passes_now();
                            uint64_t rem = ballot(surv);
                            if (rem == 0) continue;
                            if ((FL || !block_filter) && surv) slot = ids[e];
                            float dist;
                            bool ok;
                            if (FL) {
                                dist = fd; // +inf already marks what may not be returned
                                ok = surv;
                            } else {
                                dist = __fadd_rn(__fdiv_rn(static_cast<float>(sum), scale), offset);
                                ok = surv && (filtered ? bit_test(filter, slot)
                                                       : (ix.deleted == nullptr || !bit_test(ix.deleted, slot)));
                            }
                            while (rem) {
                                const bool cond = ((rem >> lane) & 1ull) && ok && dist < tk.thr;
                                const uint64_t cm = ballot(cond);
                                if (cm == 0) break;
                                const int space = 2 * kp - tk.size;
                                const int cnt = __popcll(cm);
                                if (cnt < space) {
                                    if (cond) {
                                        const int pos = tk.size + static_cast<int>(lanes_below(cm));
                                        tk.dist[pos] = dist;
                                        tk.label[pos] = slot;
                                    }
                                    tk.size += cnt;
                                    wave_sync();
                                    break;
                                }
                                // the space-th candidate fills the buffer and triggers a trim
                                const uint32_t rank = lanes_below(cm);
                                const uint64_t at = ballot(cond && static_cast<int>(rank) == space - 1);
                                const int bpos = __builtin_ctzll(at);
                                if (cond && lane <= bpos) {
                                    const int pos = tk.size + static_cast<int>(rank);
                                    tk.dist[pos] = dist;
                                    tk.label[pos] = slot;
                                }
                                tk.size = 2 * kp;
                                wave_sync();
                                wave_trim(tk);
                                rem &= (bpos >= 63) ? 0ull : (~0ull << (bpos + 1));
                            }
                        }
                    }
                }
            }
        }
    }

    if (tk.size > kp) wave_trim(tk);
    replay_finish(tk, a, qi);
}

// TopK::push of one block's survivors in bit order (src/topk.hpp:44-50): the
// lanes of rem whose entry may be pushed (ok) and is below the threshold go
// into the buffer; the one that fills it triggers a trim, and the lanes after
// it are tested again against the new threshold.
__device__ __forceinline__ void push_survivors(WaveTopK& tk, uint64_t rem, bool ok, float dist, int64_t label) {
    const int lane = lane_id();
    while (rem) {
        const bool cond = ((rem >> lane) & 1ull) && ok && dist < tk.thr;
        const uint64_t cm = ballot(cond);
        if (cm == 0) break;
        const int space = 2 * tk.k - tk.size;
        const int cnt = __popcll(cm);
        if (cnt < space) {
            if (cond) {
                const int pos = tk.size + static_cast<int>(lanes_below(cm));
                tk.dist[pos] = dist;
                tk.label[pos] = label;
            }
            tk.size += cnt;
            wave_sync();
            return;
        }
        const uint32_t rank = lanes_below(cm);
        const uint64_t at = ballot(cond && static_cast<int>(rank) == space - 1);
        const int bpos = __builtin_ctzll(at);
        if (cond && lane <= bpos) {
            const int pos = tk.size + static_cast<int>(rank);
            tk.dist[pos] = dist;
            tk.label[pos] = label;
        }
        tk.size = 2 * tk.k;
        wave_sync();
        wave_trim(tk);
        rem &= (bpos >= 63) ? 0ull : (~0ull << (bpos + 1));
    }
}

// Candidate blocks a windowed replay loads at once: kPfRows rows of two.
constexpr int kPfRows = 8;

// Block minima a windowed replay keeps in LDS at once.
constexpr uint32_t kReplayWindow = 512; // 2,048 measured 10% slower at 10k queries (fewer waves per CU)

// The replay of fs_replay with the TopK always in LDS and its per-probe costs
// cut (the same pushes in the same order, so the same TopK):
//  * the query's block minima are copied into LDS a window of kReplayWindow
//    blocks at a time (a query's probes own consecutive blocks), so that a
//    probe's candidate blocks are found without a trip to HBM;
//  * of each 64 probes, the ones whose list minimum passes its limit are
//    found at once, one probe per lane, and only those are visited (a probe
//    that fails now fails at its turn: the limit only falls, and nothing is
//    pushed between);
//  * a probe's candidate blocks go 16 at a time: their sums (distances) are
//    loaded together and staged in LDS; the TopK holds list entries, whose
//    slots are read when the query ends, so without a filter or deleted rows
//    no slot is read on the way; with them, the slots and filter / deleted
//    bits of every entry passing the limit at the group's start (a superset
//    of what a later, lower limit passes) are read for the group at once.
template <bool FL>
__device__ void fs_replay_win(const ReplayArgs& a) {
    using BT = typename std::conditional<FL, uint32_t, uint16_t>::type;
    extern __shared__ __align__(16) unsigned char lds[];
    const DevIvf& ix = a.ix;
    const int qi = blockIdx.x;
    const int kl = a.kp;                         // LDS layout: the launch's largest capacity
    const int kp = a.pq.kp ? a.pq.kp[qi] : kl;   // this query's TopK capacity
    const uint64_t* filter = a.pq.filter ? a.pq.filter[qi] : a.filter;
    const int block_filter = a.pq.block_filter ? a.pq.block_filter[qi] : a.block_filter;
    int64_t* tk_label = reinterpret_cast<int64_t*>(lds);
    float* tk_dist = reinterpret_cast<float*>(tk_label + 2 * kl);
    BT* sb = reinterpret_cast<BT*>(tk_dist + 2 * kl);
    const int lane = lane_id();
    const int v = lane & 31;
    const bool filtered = filter != nullptr;
    const bool need_slot = !FL && (filtered || ix.deleted != nullptr);
    __shared__ uint32_t trim_hist[256];
    __shared__ int st_blk[kWave];
    __shared__ uint32_t st_val[2 * kPfRows * kFsBlock];
    WaveTopK tk{tk_dist, tk_label, 0, kp, INFINITY, trim_hist};
    const int P = a.nprobe_q[qi];
    const size_t qoff = static_cast<size_t>(qi) * a.stride_slots;
    const size_t qboff = static_cast<size_t>(qi) * (a.stride_slots / kFsBlock);
    const BT* gbmin;
    if constexpr (FL) gbmin = a.fbmin + qboff;
    else gbmin = a.bmin + qboff;
    uint32_t win0 = 0, win1 = 0; // staged blocks [win0, win1) of the query
    for (int c0 = 0; c0 < P; c0 += kWave) {
        int m_list = 0;
        uint32_t m_n = 0, m_blk = 0, m_min = 0xFFFFFFFFu, m_first = 0;
        float m_scale = 1.0f, m_off = 0.0f;
        if (c0 + lane < P) {
            const size_t item = static_cast<size_t>(qi) * a.M + c0 + lane;
            m_list = a.probe_idx[item];
            m_first = ix.list_off[m_list];
            m_n = ix.list_off[m_list + 1] - m_first;
            m_blk = a.item_blk[item];
            m_min = a.item_min[item];
            if (!FL) {
                m_scale = a.item_scale[item];
                m_off = a.item_offset[item];
            }
        }
        const int cn = min(kWave, P - c0);
        // the end of this chunk's blocks (a window never reaches past it)
        const uint32_t cend = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_blk), cn - 1)) +
                              (static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_n), cn - 1)) +
                               kFsBlock - 1) / kFsBlock;
        // the probes whose list minimum passes their limit now, one per lane
        auto live_now = [&]() -> uint64_t {
            bool live;
            if (FL) {
                live = m_min < fkey(tk.thr);
            } else {
                const float limit = (tk.thr - m_off) * m_scale;
                const uint32_t l16 = limit >= 65535.0f ? 65535u : limit <= 0.0f ? 0u : static_cast<uint32_t>(ceilf(limit));
                live = m_min < l16;
            }
            return ballot(c0 + lane < P && m_n != 0 && live);
        };
        uint64_t todo = live_now();
        while (todo) {
            const int j = __builtin_ctzll(todo);
            todo &= todo - 1;
            const uint32_t n = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_n), j));
            const uint32_t nb = (n + kFsBlock - 1) / kFsBlock;
            const float scale = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m_scale), j));
            const float offset = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m_off), j));
            const uint32_t first = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_first), j));
            const uint32_t blk0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_blk), j));
            auto limit_now = [&]() -> uint32_t {
                if (FL) return fkey(tk.thr);
                const float limit = (tk.thr - offset) * scale;
                return limit >= 65535.0f ? 65535u : limit <= 0.0f ? 0u : static_cast<uint32_t>(ceilf(limit));
            };
            // the window: restaged from this probe on when it does not hold the probe's blocks
            if ((blk0 < win0 || blk0 + nb > win1) && nb <= static_cast<uint32_t>(a.window)) {
                win0 = blk0;
                win1 = min(blk0 + static_cast<uint32_t>(a.window), cend);
                wave_sync();
                for (uint32_t i = static_cast<uint32_t>(lane); i < win1 - win0; i += kWave) sb[i] = gbmin[win0 + i];
                wave_sync();
            }
            const bool staged = blk0 >= win0 && blk0 + nb <= win1;
            const float thr_in = tk.thr;
            const size_t voff = qoff + static_cast<size_t>(blk0) * kFsBlock;
            for (uint32_t bc = 0; bc < nb; bc += kWave) {
                const uint32_t bl = bc + static_cast<uint32_t>(lane);
                uint32_t bm = 0xFFFFFFFFu;
                if (bl < nb) bm = staged ? static_cast<uint32_t>(sb[blk0 - win0 + bl]) : static_cast<uint32_t>(gbmin[blk0 + bl]);
                uint64_t cand = ballot(bm < limit_now());
                while (cand) {
                    // the next (up to) 2 * kPfRows candidate blocks, in order; row r
                    // holds candidates 2r (lanes 0-31) and 2r + 1 (lanes 32-63)
                    const int ncand = __popcll(cand);
                    const int nc = min(ncand, 2 * kPfRows);
                    if ((cand >> lane) & 1ull) st_blk[lanes_below(cand)] = lane;
                    wave_sync();
                    const int rows = (nc + 1) / 2;
                    const uint32_t lim0 = limit_now();
                    uint32_t sv[kPfRows];
#pragma unroll
                    for (int r = 0; r < kPfRows; ++r) {
                        sv[r] = FL ? __float_as_uint(INFINITY) : 0xFFFFu;
                        if (r < rows) {
                            const int rr = 2 * r + (lane >> 5);
                            const int mb = rr < nc ? st_blk[rr] : -1;
                            const uint32_t e = (bc + static_cast<uint32_t>(mb)) * kFsBlock + v;
                            if (mb >= 0 && e < n)
                                sv[r] = FL ? __float_as_uint(a.fdists[voff + e]) : static_cast<uint32_t>(a.sums[voff + e]);
                        }
                    }
                    if (need_slot) {
                        // bit 16 set: the entry may not be returned (filtered out or deleted)
                        uint32_t sl[kPfRows];
#pragma unroll
                        for (int r = 0; r < kPfRows; ++r) {
                            sl[r] = 0u;
                            if (r < rows && sv[r] < lim0) {
                                const int rr = 2 * r + (lane >> 5);
                                const uint32_t e = (bc + static_cast<uint32_t>(st_blk[rr])) * kFsBlock + v;
                                sl[r] = ix.list_ids[first + e];
                            }
                        }
#pragma unroll
                        for (int r = 0; r < kPfRows; ++r) {
                            if (r < rows && sv[r] < lim0) {
                                const bool okb = filtered ? bit_test(filter, sl[r]) : !bit_test(ix.deleted, sl[r]);
                                if (!okb) sv[r] |= 0x10000u;
                            }
                        }
                    }
#pragma unroll
                    for (int r = 0; r < kPfRows; ++r)
                        if (r < rows) st_val[r * kWave + lane] = sv[r];
                    wave_sync();
                    // row r's entry in this lane: its value, whether it is in the list and
                    // may be returned, and its distance
                    auto row_entry = [&](int r, uint32_t& e, bool& in, bool& okb, uint32_t& sum, float& fd) {
                        const int rr = 2 * r + (lane >> 5);
                        const int mb = rr < nc ? st_blk[rr] : -1;
                        e = (bc + static_cast<uint32_t>(mb)) * kFsBlock + v;
                        in = mb >= 0 && e < n;
                        const uint32_t x = st_val[r * kWave + lane];
                        okb = FL || !(x & 0x10000u);
                        sum = 0xFFFFu;
                        fd = INFINITY;
                        if (in) {
                            if (FL) fd = __uint_as_float(x);
                            else sum = x & 0xFFFFu;
                        }
                    };
                    // Rows whose pushes cannot fill the buffer leave the threshold
                    // (so every block's limit) as it is: their survivors are counted
                    // and written at once, in push order (row by row; in a row, the
                    // first block's entries, lanes 0-31, then the second's). The row
                    // that would fill it replays push by push, trims included.
                    int r = 0;
                    while (r < rows) {
                        const uint32_t lim = limit_now();
                        const float thr = tk.thr;
                        for (; r < rows; ++r) {
                            uint32_t e, sum;
                            bool in, okb;
                            float fd;
                            row_entry(r, e, in, okb, sum, fd);
                            const float dist = FL ? fd : __fadd_rn(__fdiv_rn(static_cast<float>(sum), scale), offset);
                            const bool surv = in && okb && (FL ? fd < thr : sum < lim) && dist < thr;
                            const uint64_t m = ballot(surv);
                            const int c = __popcll(m);
                            if (tk.size + c >= 2 * kp) break;
                            if (surv) {
                                const int pos = tk.size + static_cast<int>(lanes_below(m));
                                tk.dist[pos] = dist;
                                tk.label[pos] = static_cast<int64_t>(first + e);
                            }
                            tk.size += c;
                        }
                        wave_sync();
                        if (r >= rows) break;
                        // row r fills the buffer: its two blocks push by push
                        {
                            uint32_t e, sum;
                            bool in, okb;
                            float fd;
                            row_entry(r, e, in, okb, sum, fd);
                            auto passes_now = [&]() -> bool { return FL ? fd < tk.thr : sum < limit_now(); };
                            for (int h = 0; h < 2; ++h) {
                                if (h == 1 && 2 * r + 1 >= nc) break;
                                const bool valid = (lane >> 5) == h && in;
                                bool allowed = valid;
                                if (!FL && block_filter) {
                                    allowed = valid && okb;
                                    if (ballot(allowed) == 0) continue;
                                }
                                const bool surv = allowed && passes_now();
                                const uint64_t rem = ballot(surv);
                                if (rem == 0) continue;
                                float dist;
                                bool ok;
                                if (FL) {
                                    dist = fd; // +inf already marks what may not be returned
                                    ok = surv;
                                } else {
                                    dist = __fadd_rn(__fdiv_rn(static_cast<float>(sum), scale), offset);
                                    ok = surv && okb;
                                }
                                push_survivors(tk, rem, ok, dist, static_cast<int64_t>(first + e));
                            }
                        }
                        ++r;
                    }
                    // drop the candidates just replayed
                    if (nc == ncand) {
                        cand = 0;
                    } else {
                        const int last = __builtin_amdgcn_readfirstlane(st_blk[nc - 1]);
                        cand &= ~((2ull << last) - 1ull);
                    }
                    wave_sync(); // the next group overwrites st_blk / st_val
                }
            }
            // a push lowered the threshold: drop the probes that now fail
            if (tk.thr != thr_in) todo &= live_now();
        }
    }
    // entries become slots
    if (tk.size > kp) wave_trim(tk);
    for (int i = lane; i < tk.size; i += kWave) tk.label[i] = ix.list_ids[static_cast<uint32_t>(tk.label[i])];
    wave_sync();
    replay_finish(tk, a, qi);
}

template <bool STAGE>
__global__ void __launch_bounds__(64) k_fs_replay(const ReplayArgs a) { fs_replay<false, STAGE>(a); }
// fs_replay_win with the candidate blocks of consecutive probes gathered into
// one batch of up to 2 kPfRows blocks before their sums are loaded: a probe
// no longer costs a round trip to memory of its own. Each gathered block keeps
// its probe's scale, offset and entry range, and the batch is pushed block by
// block in probe order, so the pushes and the TopK are those of the probe-by-
// probe replay. A block gathered under a limit that a later push lowers is
// still visited; its entries fail the new limit, as they would have at its turn.
// Rows of two candidate blocks a gathered batch holds.
constexpr int kGatherRows = 8;

template <bool FL>
__device__ void fs_replay_gather(const ReplayArgs& a) {
    using BT = typename std::conditional<FL, uint32_t, uint16_t>::type;
    extern __shared__ __align__(16) unsigned char lds[];
    const DevIvf& ix = a.ix;
    const int qi = blockIdx.x;
    const int kl = a.kp;                         // LDS layout: the launch's largest capacity
    const int kp = a.pq.kp ? a.pq.kp[qi] : kl;   // this query's TopK capacity
    const uint64_t* filter = a.pq.filter ? a.pq.filter[qi] : a.filter;
    const int block_filter = a.pq.block_filter ? a.pq.block_filter[qi] : a.block_filter;
    int64_t* tk_label = reinterpret_cast<int64_t*>(lds);
    float* tk_dist = reinterpret_cast<float*>(tk_label + 2 * kl);
    BT* sb = reinterpret_cast<BT*>(tk_dist + 2 * kl);
    const int lane = lane_id();
    const int v = lane & 31, h = lane >> 5;
    const bool filtered = filter != nullptr;
    const bool need_slot = !FL && (filtered || ix.deleted != nullptr);
    constexpr int kB = 2 * kGatherRows; // blocks a batch holds
    __shared__ uint32_t trim_hist[256];
    __shared__ uint32_t g_vbase[kB];  // the block's first sum / distance, within the query's region
    __shared__ uint32_t g_ebase[kB];  // its first list entry
    __shared__ int g_cnt[kB];         // its entries (32, fewer in a list's last block)
    __shared__ float g_scale[kB], g_off[kB];
    __shared__ uint32_t st_val[kGatherRows * kWave];
    WaveTopK tk{tk_dist, tk_label, 0, kp, INFINITY, trim_hist};
    const int P = a.nprobe_q[qi];
    const size_t qoff = static_cast<size_t>(qi) * a.stride_slots;
    const size_t qboff = static_cast<size_t>(qi) * (a.stride_slots / kFsBlock);
    const BT* gbmin;
    if constexpr (FL) gbmin = a.fbmin + qboff;
    else gbmin = a.bmin + qboff;
    auto limit_of = [&](float scale, float offset) -> uint32_t {
        if (FL) return fkey(tk.thr);
        const float limit = (tk.thr - offset) * scale;
        return limit >= 65535.0f ? 65535u : limit <= 0.0f ? 0u : static_cast<uint32_t>(ceilf(limit));
    };
    // the batch's blocks pushed in order: rows of two blocks (lanes 0-31, 32-63)
    int nbatch = 0;
    auto flush = [&]() {
        const int rows = (nbatch + 1) / 2;
        const int bi0 = h; // this lane's block in row r is 2 r + h
        auto bl = [&](int r) { return 2 * r + bi0; };
        uint32_t sv[kGatherRows];
#pragma unroll
        for (int r = 0; r < kGatherRows; ++r) {
            sv[r] = FL ? __float_as_uint(INFINITY) : 0xFFFFu;
            if (r < rows && bl(r) < nbatch && v < g_cnt[bl(r)]) {
                const size_t e = qoff + g_vbase[bl(r)] + static_cast<uint32_t>(v);
                sv[r] = FL ? __float_as_uint(a.fdists[e]) : static_cast<uint32_t>(a.sums[e]);
            }
        }
        if (need_slot) {
            // bit 16 set: the entry may not be returned (filtered out or deleted);
            // read for the entries passing their limit now (later limits only fall)
            uint32_t sl[kGatherRows];
#pragma unroll
            for (int r = 0; r < kGatherRows; ++r) {
                sl[r] = 0u;
                if (r < rows && bl(r) < nbatch && v < g_cnt[bl(r)] && sv[r] < limit_of(g_scale[bl(r)], g_off[bl(r)]))
                    sl[r] = ix.list_ids[g_ebase[bl(r)] + static_cast<uint32_t>(v)];
            }
#pragma unroll
            for (int r = 0; r < kGatherRows; ++r) {
                if (r < rows && bl(r) < nbatch && v < g_cnt[bl(r)] && sv[r] < limit_of(g_scale[bl(r)], g_off[bl(r)])) {
                    const bool okb = filtered ? bit_test(filter, sl[r]) : !bit_test(ix.deleted, sl[r]);
                    if (!okb) sv[r] |= 0x10000u;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kGatherRows; ++r)
            if (r < rows) st_val[r * kWave + lane] = sv[r];
        wave_sync();
        // row r's entry in this lane: in its block and list, may be returned,
        // its sum or distance, its block's scale and offset
        auto row_entry = [&](int r, bool& in, bool& okb, uint32_t& sum, float& fd, float& scale, float& offset,
                             int64_t& label) {
            const int b = bl(r);
            in = b < nbatch && v < g_cnt[b];
            const uint32_t x = st_val[r * kWave + lane];
            okb = FL || !(x & 0x10000u);
            sum = 0xFFFFu;
            fd = INFINITY;
            scale = 1.0f;
            offset = 0.0f;
            label = 0;
            if (in) {
                if (FL) fd = __uint_as_float(x);
                else sum = x & 0xFFFFu;
                scale = g_scale[b];
                offset = g_off[b];
                label = static_cast<int64_t>(g_ebase[b] + static_cast<uint32_t>(v));
            }
        };
        // rows whose pushes cannot fill the buffer leave the threshold (so every
        // limit) as it is: their survivors are written at once, in push order;
        // the row that would fill it replays push by push, trims included
        int r = 0;
        while (r < rows) {
            const float thr = tk.thr;
            for (; r < rows; ++r) {
                bool in, okb;
                uint32_t sum;
                float fd, scale, offset;
                int64_t label;
                row_entry(r, in, okb, sum, fd, scale, offset, label);
                const float dist = FL ? fd : __fadd_rn(__fdiv_rn(static_cast<float>(sum), scale), offset);
                const bool surv = in && okb && (FL ? fd < thr : sum < limit_of(scale, offset)) && dist < thr;
                const uint64_t m = ballot(surv);
                const int c = __popcll(m);
                if (tk.size + c >= 2 * kp) break;
                if (surv) {
                    const int pos = tk.size + static_cast<int>(lanes_below(m));
                    tk.dist[pos] = dist;
                    tk.label[pos] = label;
                }
                tk.size += c;
            }
            wave_sync();
            if (r >= rows) break;
            {
                bool in, okb;
                uint32_t sum;
                float fd, scale, offset;
                int64_t label;
                row_entry(r, in, okb, sum, fd, scale, offset, label);
                for (int hh = 0; hh < 2; ++hh) {
                    if (2 * r + hh >= nbatch) break;
                    const bool valid = h == hh && in;
                    bool allowed = valid;
                    if (!FL && block_filter) {
                        allowed = valid && okb;
                        if (ballot(allowed) == 0) continue;
                    }
                    const bool surv = allowed && (FL ? fd < tk.thr : sum < limit_of(scale, offset));
                    const uint64_t rem = ballot(surv);
                    if (rem == 0) continue;
                    float dist;
                    bool ok;
                    if (FL) {
                        dist = fd; // +inf already marks what may not be returned
                        ok = surv;
                    } else {
                        dist = __fadd_rn(__fdiv_rn(static_cast<float>(sum), scale), offset);
                        ok = surv && okb;
                    }
                    push_survivors(tk, rem, ok, dist, label);
                }
            }
            ++r;
        }
        wave_sync(); // the next batch overwrites the g_* arrays and st_val
        nbatch = 0;
    };

    uint32_t win0 = 0, win1 = 0; // staged blocks [win0, win1) of the query
    for (int c0 = 0; c0 < P; c0 += kWave) {
        int m_list = 0;
        uint32_t m_n = 0, m_blk = 0, m_min = 0xFFFFFFFFu, m_first = 0;
        float m_scale = 1.0f, m_off = 0.0f;
        if (c0 + lane < P) {
            const size_t item = static_cast<size_t>(qi) * a.M + c0 + lane;
            m_list = a.probe_idx[item];
            m_first = ix.list_off[m_list];
            m_n = ix.list_off[m_list + 1] - m_first;
            m_blk = a.item_blk[item];
            m_min = a.item_min[item];
            if (!FL) {
                m_scale = a.item_scale[item];
                m_off = a.item_offset[item];
            }
        }
        const int cn = min(kWave, P - c0);
        const uint32_t cend = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_blk), cn - 1)) +
                              (static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_n), cn - 1)) +
                               kFsBlock - 1) / kFsBlock;
        // the probes whose list minimum passes their limit now, one per lane
        auto live_now = [&]() -> uint64_t {
            return ballot(c0 + lane < P && m_n != 0 && m_min < limit_of(m_scale, m_off));
        };
        uint64_t todo = live_now();
        while (todo) {
            const int j = __builtin_ctzll(todo);
            todo &= todo - 1;
            const uint32_t n = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_n), j));
            const uint32_t nb = (n + kFsBlock - 1) / kFsBlock;
            const float scale = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m_scale), j));
            const float offset = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m_off), j));
            const uint32_t first = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_first), j));
            const uint32_t blk0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m_blk), j));
            // the window: restaged from this probe on when it does not hold the
            // probe's blocks (the gathered blocks keep their sums' places, not
            // their minima, so restaging leaves the batch as it is)
            if ((blk0 < win0 || blk0 + nb > win1) && nb <= static_cast<uint32_t>(a.window)) {
                win0 = blk0;
                win1 = min(blk0 + static_cast<uint32_t>(a.window), cend);
                wave_sync();
                for (uint32_t i = static_cast<uint32_t>(lane); i < win1 - win0; i += kWave) sb[i] = gbmin[win0 + i];
                wave_sync();
            }
            const bool staged = blk0 >= win0 && blk0 + nb <= win1;
            for (uint32_t bc = 0; bc < nb; bc += kWave) {
                const uint32_t b = bc + static_cast<uint32_t>(lane);
                uint32_t bm = 0xFFFFFFFFu;
                if (b < nb) bm = staged ? static_cast<uint32_t>(sb[blk0 - win0 + b]) : static_cast<uint32_t>(gbmin[blk0 + b]);
                uint64_t cand = ballot(bm < limit_of(scale, offset));
                while (cand) {
                    const int take = min(__popcll(cand), kB - nbatch);
                    const int rank = static_cast<int>(lanes_below(cand));
                    if (((cand >> lane) & 1ull) && rank < take) {
                        const int gi = nbatch + rank;
                        g_vbase[gi] = (blk0 + b) * kFsBlock;
                        g_ebase[gi] = first + b * kFsBlock;
                        g_cnt[gi] = static_cast<int>(min(n - b * kFsBlock, static_cast<uint32_t>(kFsBlock)));
                        g_scale[gi] = scale;
                        g_off[gi] = offset;
                    }
                    nbatch += take;
                    // drop the blocks just gathered
                    if (take == __popcll(cand)) {
                        cand = 0;
                    } else {
                        const uint64_t taken = ballot(((cand >> lane) & 1ull) && rank < take);
                        cand &= ~taken;
                    }
                    if (nbatch == kB) {
                        wave_sync();
                        flush();
                        // a lower threshold drops blocks, and probes, that no longer pass
                        cand &= ballot(bm < limit_of(scale, offset));
                        todo &= live_now();
                    }
                }
            }
        }
    }
    // the last batch (a batch may span 64-probe chunks: its blocks carry their own places)
    if (nbatch) {
        wave_sync();
        flush();
    }
    // entries become slots
    if (tk.size > kp) wave_trim(tk);
    for (int i = lane; i < tk.size; i += kWave) tk.label[i] = ix.list_ids[static_cast<uint32_t>(tk.label[i])];
    wave_sync();
    replay_finish(tk, a, qi);
}

__global__ void __launch_bounds__(64) k_fs_replay_win(const ReplayArgs a) { fs_replay_win<false>(a); }
__global__ void __launch_bounds__(64) k_fs_replay_gather(const ReplayArgs a) { fs_replay_gather<false>(a); }
__global__ void __launch_bounds__(64) k_fl_replay_gather(const ReplayArgs a) { fs_replay_gather<true>(a); }
__global__ void __launch_bounds__(64) k_fl_replay_win(const ReplayArgs a) { fs_replay_win<true>(a); }
template <bool STAGE>
__global__ void __launch_bounds__(64) k_fl_replay(const ReplayArgs a) { fs_replay<true, STAGE>(a); }

// ---------------------------------------------------------------------------
// Exact re-rank (RefineIndex::rerank): the candidates' exact distances, the k
// smallest by (distance, slot) in pair order, mapped to the refine labels.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_rerank(const float* q, int dim, int ip, const float* rows, const int64_t* cand, const int* cand_count,
                         int kc, int k, const int64_t* labels, int64_t* out_ids, float* out_dists, uint64_t* gkeys,
                         int gstride, const PerQuery pq) {
    extern __shared__ __align__(16) unsigned char lds[];
    float* sq = reinterpret_cast<float*>(lds);
    int np = 1;
    const int i = blockIdx.x;
    const int n = cand_count[i];
    // this query's result count and row (a launch of several requests), or the launch's
    if (pq.k) k = pq.k[i];
    const size_t row = pq.out_off ? pq.out_off[i] : static_cast<size_t>(i) * k;
    while (np < n) np <<= 1;
    uint64_t* keys = gkeys ? gkeys + static_cast<size_t>(i) * gstride
                           : reinterpret_cast<uint64_t*>(lds + query_lds_bytes(dim));
    const bool qcopy = dim <= kQueryLdsDims;
    if (qcopy)
        for (int t = threadIdx.x; t < dim; t += blockDim.x) sq[t] = q[static_cast<size_t>(i) * dim + t];
    else
        sq = const_cast<float*>(q + static_cast<size_t>(i) * dim); // read only
    __syncthreads();
    for (int c = threadIdx.x; c < np; c += blockDim.x) {
        if (c < n) {
            const uint32_t slot = static_cast<uint32_t>(cand[static_cast<size_t>(i) * kc + c]);
            const float d = exact_score(ip != 0, sq, rows + static_cast<size_t>(slot) * dim, dim);
            keys[c] = (static_cast<uint64_t>(ckey(d)) << 32) | slot;
        } else {
            keys[c] = ~0ull;
        }
    }
    block_bitonic_sort(keys, np);
    const int keep = min(k, n);
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
        int64_t id = -1;
        float d = INFINITY;
        if (t < keep) {
            const uint32_t slot = static_cast<uint32_t>(keys[t] & 0xFFFFFFFFu);
            id = labels ? labels[slot] : static_cast<int64_t>(slot);
            d = exact_score(ip != 0, sq, rows + static_cast<size_t>(slot) * dim, dim);
        }
        out_ids[row + t] = id;
        out_dists[row + t] = d;
    }
}

// ---------------------------------------------------------------------------
// Exact top-k over a set of slots, by one workgroup per query: the k smallest
// 64-bit keys (key32(dist) << 32 | slot), which is FlatIndex::query_impl's
// max-heap result (ckey: float order) and TopK's result for pushes in slot
// order (fkey). Candidates come from a slot list, or all slots minus the
// deleted ones. Output sorted by (key32, label) -- the final order.
// ---------------------------------------------------------------------------
// The score of one slot under a SlotScorer (exact_score on its row, or SQ).
__device__ __forceinline__ float slot_score(const SlotScorer& sc, bool ip, const float* q, uint32_t slot, int dim) {
    const uint32_t e = sc.entry_of_slot ? sc.entry_of_slot[slot] : slot;
    if (e == 0xFFFFFFFFu) return INFINITY; // a slot in no list
    if (sc.sq_codes) {
        int cs;
        const uint8_t* code = sq_row(sc.sq_codes, e, sc.vbytes, sc.sq_chunked, cs);
        return ip ? -sq_score<true>(q, code, sc.sq_vmin, sc.sq_step, dim, sc.sq_bits, cs)
                  : sq_score<false>(q, code, sc.sq_vmin, sc.sq_step, dim, sc.sq_bits, cs);
    }
    return exact_score(ip, q, sc.rows + static_cast<size_t>(e) * dim, dim);
}

__global__ void __launch_bounds__(256) k_block_select(const float* q, int dim, int ip, const SlotScorer sc, const uint32_t* slot_list,
                               uint32_t n_slots, const uint64_t* deleted, int use_fkey, int final_by_label, int k,
                               int cap, const int64_t* labels, int64_t* out_ids, float* out_dists,
                               unsigned char* gbuf, size_t gstride) {
    extern __shared__ __align__(16) unsigned char lds[];
    float* sq = reinterpret_cast<float*>(lds);
    // cap (pow2) keys, then the final sort's arrays: in LDS after the query, or in HBM
    uint64_t* keys = gbuf ? reinterpret_cast<uint64_t*>(gbuf + blockIdx.x * gstride)
                          : reinterpret_cast<uint64_t*>(lds + query_lds_bytes(dim));
    __shared__ int count_s;
    __shared__ uint64_t thr_s;
    __shared__ int big_s;
    const int i = blockIdx.x;
    if (dim <= kQueryLdsDims)
        for (int t = threadIdx.x; t < dim; t += blockDim.x) sq[t] = q[static_cast<size_t>(i) * dim + t];
    else
        sq = const_cast<float*>(q + static_cast<size_t>(i) * dim); // read only
    if (threadIdx.x == 0) {
        count_s = 0;
        thr_s = ~0ull;
        big_s = 0;
    }
    __syncthreads();
    for (uint32_t base = 0; base < n_slots; base += blockDim.x) {
        // room for a full round: keep the k smallest so far
        if (count_s + static_cast<int>(blockDim.x) > cap) {
            for (int t = count_s + threadIdx.x; t < cap; t += blockDim.x) keys[t] = ~0ull;
            block_bitonic_sort(keys, cap);
            if (threadIdx.x == 0) {
                count_s = k;
                thr_s = keys[k - 1];
            }
            __syncthreads();
        }
        const uint32_t c = base + threadIdx.x;
        if (c < n_slots) {
            const uint32_t slot = slot_list ? slot_list[c] : c;
            if (!deleted || !bit_test(deleted, slot)) {
                const float d = slot_score(sc, ip != 0, sq, slot, dim);
                const uint64_t key = (static_cast<uint64_t>(use_fkey ? fkey(d) : ckey(d)) << 32) | slot;
                if (key < thr_s) keys[atomicAdd(&count_s, 1)] = key;
            }
        }
        __syncthreads();
    }
    const int n = count_s;
    for (int t = n + threadIdx.x; t < cap; t += blockDim.x) keys[t] = ~0ull;
    block_bitonic_sort(keys, cap);
    const int keep = min(k, n);
    // TopK::take_sorted order: (fkey, id) when every id fits 32 bits, else (float, id)
    if (final_by_label && labels) {
        int kp2 = 1;
        while (kp2 < keep) kp2 <<= 1;
        uint32_t* k32 = reinterpret_cast<uint32_t*>(keys + cap);         // kp2
        int64_t* lab = reinterpret_cast<int64_t*>(k32 + ((kp2 + 1) & ~1)); // kp2
        uint32_t* slt = reinterpret_cast<uint32_t*>(lab + kp2);           // kp2
        for (int t = threadIdx.x; t < keep; t += blockDim.x)
            if (static_cast<uint64_t>(labels[static_cast<uint32_t>(keys[t] & 0xFFFFFFFFu)]) > 0xFFFFFFFFull) big_s = 1;
        __syncthreads();
        const bool big = big_s != 0;
        for (int t = threadIdx.x; t < kp2; t += blockDim.x) {
            if (t < keep) {
                const uint32_t slot = static_cast<uint32_t>(keys[t] & 0xFFFFFFFFu);
                const float d = from_fkey(static_cast<uint32_t>(keys[t] >> 32)); // fkey mode: exact inverse
                k32[t] = big ? ckey(d) : static_cast<uint32_t>(keys[t] >> 32);
                lab[t] = labels[slot];
                slt[t] = slot;
            } else {
                k32[t] = 0xFFFFFFFFu;
                lab[t] = -1;
                slt[t] = 0xFFFFFFFFu;
            }
        }
        for (int size = 2; size <= kp2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (int t = threadIdx.x; t < (kp2 >> 1); t += blockDim.x) {
                    const int x = 2 * t - (t & (stride - 1));
                    const int y = x + stride;
                    const bool up = (x & size) == 0;
                    const bool gt = k32[x] != k32[y] ? k32[x] > k32[y]
                                                     : static_cast<uint64_t>(lab[x]) > static_cast<uint64_t>(lab[y]);
                    if (gt == up) {
                        const uint32_t tk = k32[x];
                        k32[x] = k32[y];
                        k32[y] = tk;
                        const int64_t tl = lab[x];
                        lab[x] = lab[y];
                        lab[y] = tl;
                        const uint32_t ts = slt[x];
                        slt[x] = slt[y];
                        slt[y] = ts;
                    }
                }
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < k; t += blockDim.x) {
            int64_t id = -1;
            float d = INFINITY;
            if (t < keep) {
                id = lab[t];
                d = slot_score(sc, ip != 0, sq, slt[t], dim);
            }
            out_ids[static_cast<size_t>(i) * k + t] = id;
            out_dists[static_cast<size_t>(i) * k + t] = d;
        }
        return;
    }
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
        int64_t id = -1;
        float d = INFINITY;
        if (t < keep) {
            const uint32_t slot = static_cast<uint32_t>(keys[t] & 0xFFFFFFFFu);
            id = labels ? labels[slot] : static_cast<int64_t>(slot);
            d = slot_score(sc, ip != 0, sq, slot, dim);
        }
        out_ids[static_cast<size_t>(i) * k + t] = id;
        out_dists[static_cast<size_t>(i) * k + t] = d;
    }
}

} // namespace qd

// ---------------------------------------------------------------------------
// Launch wrappers.
// ---------------------------------------------------------------------------
namespace qd {

static inline unsigned cdiv(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

void launch_sqnorms(const float* rows, int n, int dim, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sqnorms, dim3(cdiv(n, 256)), dim3(256), 0, s, rows, n, dim, out);
}

void launch_pq_transpose(const float* cent, int m, int K, int sub, float* cent_t, hipStream_t s) {
    hipLaunchKernelGGL(k_pq_transpose, dim3(cdiv(static_cast<size_t>(m) * K * sub, 256)), dim3(256), 0, s, cent, m,
                       K, sub, cent_t);
}

void launch_sdc_table(const float* cent_t, int m, int sub, int K, float* sdc, hipStream_t s) {
    hipLaunchKernelGGL(k_sdc_table, dim3(cdiv(static_cast<size_t>(m) * K * K, 256)), dim3(256), 0, s, cent_t, m, sub, K,
                       sdc);
}

void launch_list_term(const float* centers, const float* cent_t, int nlist, int dim, int m, int sub, int K, float* B,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_list_term, dim3(cdiv(static_cast<size_t>(m) * K, 256) * static_cast<unsigned>(nlist)),
                       dim3(256), 0, s, centers, cent_t, nlist, dim, m, sub, K, B);
}

void launch_rotate(const float* q, const float* R, int nq, int dim, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_rotate, dim3(cdiv(static_cast<size_t>(nq) * dim, 256)), dim3(256), 0, s, q, R, nq, dim, out);
}

size_t coarse_select_cap(int M) {
    size_t cap = 1024;
    while (cap < static_cast<size_t>(M)) cap <<= 1;
    return cap;
}

void launch_coarse_gemm(const float* q, int nq, const float* cents, const float* csq, int nlist, int dim, int ip,
                        float* D, hipStream_t s, const uint64_t* excl, const uint64_t* incl) {
    // QANN_AMD_COARSE_GEMM=16 keeps the 64 x 64 tiles of 16 x 16 x 4 MFMAs
    const char* e = std::getenv("QANN_AMD_COARSE_GEMM");
    const bool small = e && std::strcmp(e, "16") == 0;
    const int tile = small ? kGemmTile : kG2Tile;
    const int rows = 65535 * tile; // grid y limit
    for (int q0 = 0; q0 < nq; q0 += rows) {
        const int n = std::min(rows, nq - q0);
        if (small)
            hipLaunchKernelGGL(k_coarse_gemm, dim3(cdiv(nlist, tile), cdiv(n, tile)), dim3(256), 0, s,
                               q + static_cast<size_t>(q0) * dim, n, cents, csq, nlist, dim, ip,
                               D + static_cast<size_t>(q0) * nlist, excl, incl);
        else
            hipLaunchKernelGGL(k_coarse_gemm32, dim3(cdiv(nlist, tile), cdiv(n, tile)), dim3(256), 0, s,
                               q + static_cast<size_t>(q0) * dim, n, cents, csq, nlist, dim, ip,
                               D + static_cast<size_t>(q0) * nlist, excl, incl);
    }
}

size_t lds_room(const void* fn) {
    // the device's limit and each kernel's static LDS, looked up once
    static std::mutex mu;
    static std::map<int, size_t> dev_limit;
    static std::map<const void*, size_t> static_lds;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    size_t limit = 0, fixed = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        auto d = dev_limit.find(dev);
        if (d == dev_limit.end()) {
            int v = 0;
            if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
                v = static_cast<int>(kLdsMax);
            d = dev_limit.emplace(dev, static_cast<size_t>(v)).first;
        }
        limit = d->second;
        auto f = static_lds.find(fn);
        if (f == static_lds.end()) {
            hipFuncAttributes fa{};
            const size_t v = hipFuncGetAttributes(&fa, fn) == hipSuccess ? fa.sharedSizeBytes : limit;
            f = static_lds.emplace(fn, v).first;
        }
        fixed = f->second;
    }
    if (const char* e = std::getenv("QANN_AMD_LDS_LIMIT")) {
        const long v = std::atol(e);
        if (v >= 16 * 1024 && static_cast<size_t>(v) < limit) limit = static_cast<size_t>(v);
    }
    return fixed < limit ? limit - fixed : 0;
}

size_t replay_room() {
    return std::min({lds_room(reinterpret_cast<const void*>(&k_fs_replay<false>)),
                     lds_room(reinterpret_cast<const void*>(&k_fs_replay<true>)),
                     lds_room(reinterpret_cast<const void*>(&k_fl_replay<false>)),
                     lds_room(reinterpret_cast<const void*>(&k_fl_replay<true>))});
}

size_t rerank_room() { return lds_room(reinterpret_cast<const void*>(&k_rerank)); }

size_t block_select_room() { return lds_room(reinterpret_cast<const void*>(&k_block_select)); }

size_t coarse_select_room() {
    return std::min({lds_room(reinterpret_cast<const void*>(&k_coarse_select<256, 16, true>)),
                     lds_room(reinterpret_cast<const void*>(&k_coarse_select<256, 16, false>)),
                     lds_room(reinterpret_cast<const void*>(&k_coarse_select<1024, 16, true>)),
                     lds_room(reinterpret_cast<const void*>(&k_coarse_select<1024, 16, false>)),
                     lds_room(reinterpret_cast<const void*>(&k_coarse_select<256, 0, true>)),
                     lds_room(reinterpret_cast<const void*>(&k_coarse_select<256, 0, false>))});
}

size_t ivf_dists_room() { return lds_room(reinterpret_cast<const void*>(&k_ivf_dists)); }

bool coarse_select_global(int M, int dim) {
    return query_lds_bytes(dim) + coarse_select_cap(M) * 8 > coarse_select_room();
}

ProbeBaseOut probe_base_out(const DevIvf& ix, float* A, float* base) {
    ProbeBaseOut pb;
    pb.residual_centers = ix.residual_centers;
    pb.residual_sqnorms = ix.residual_sqnorms;
    pb.centroid_sqnorms = ix.centroid_sqnorms;
    pb.ip = ix.metric == 2 ? 1 : 0;
    pb.cosine_base = ix.metric == 1 && ix.cosine_ip ? 1 : 0;
    pb.A = A;
    pb.base = base;
    return pb;
}

void launch_coarse_select(float* D, const float* q, int nq, const float* cents, float cnorm_max, int nlist, int dim,
                          int ip, int M, uint16_t* out_idx, float* out_val, hipStream_t s, int split_kp,
                          uint64_t* gkeys, const ProbeBaseOut& pb) {
    if (nq <= 0) return;
    size_t cap = coarse_select_cap(M);
    const bool global = coarse_select_global(M, dim);
    if (global && !gkeys) throw std::invalid_argument("the coarse select needs global keys for this many probes");
    if (!global) gkeys = nullptr;
    // rows of more than 4,096 lists hold up to 4,096 candidates in LDS: where
    // many lists lie about as far from a query (a large cluster split into many
    // lists), the rounding window around the threshold holds hundreds of them,
    // and past the cap a query scores every list exactly
    if (!global && nlist > 4096) {
        const size_t wide = std::max<size_t>(cap, 4096);
        if (query_lds_bytes(dim) + wide * 8 <= coarse_select_room()) cap = wide;
    }
    const size_t lds = query_lds_bytes(dim) + (global ? 0 : cap * 8);
    // each query's row of scores stays in registers, 16 values a thread: with
    // 256 threads up to 4,096 lists, with 1,024 up to 16,384; larger rows are
    // read from memory, and take a threshold from a sample first (a row in
    // registers costs less to search than the sample's larger candidate set).
    // QANN_AMD_SELECT_SAMPLE=1 samples every row of more than 2,048 lists, 0 none.
    const char* se = std::getenv("QANN_AMD_SELECT_SAMPLE");
    const bool smp = se && std::strcmp(se, "1") == 0   ? nlist > 2048
                     : se && std::strcmp(se, "0") == 0 ? false
                                                        : nlist > 16 * 1024;
    if (nlist <= 16 * 256) {
        if (smp)
            hipLaunchKernelGGL((k_coarse_select<256, 16, true>), dim3(nq), dim3(256), lds, s, D, q, cents, nlist, dim,
                               ip, cnorm_max, M, static_cast<int>(cap), out_idx, out_val, split_kp, gkeys, pb);
        else
            hipLaunchKernelGGL((k_coarse_select<256, 16, false>), dim3(nq), dim3(256), lds, s, D, q, cents, nlist, dim,
                               ip, cnorm_max, M, static_cast<int>(cap), out_idx, out_val, split_kp, gkeys, pb);
    } else if (nlist <= 16 * 1024) {
        if (smp)
            hipLaunchKernelGGL((k_coarse_select<1024, 16, true>), dim3(nq), dim3(1024), lds, s, D, q, cents, nlist, dim,
                               ip, cnorm_max, M, static_cast<int>(cap), out_idx, out_val, split_kp, gkeys, pb);
        else
            hipLaunchKernelGGL((k_coarse_select<1024, 16, false>), dim3(nq), dim3(1024), lds, s, D, q, cents, nlist,
                               dim, ip, cnorm_max, M, static_cast<int>(cap), out_idx, out_val, split_kp, gkeys, pb);
    } else {
        if (smp)
            hipLaunchKernelGGL((k_coarse_select<256, 0, true>), dim3(nq), dim3(256), lds, s, D, q, cents, nlist, dim, ip,
                               cnorm_max, M, static_cast<int>(cap), out_idx, out_val, split_kp, gkeys, pb);
        else
            hipLaunchKernelGGL((k_coarse_select<256, 0, false>), dim3(nq), dim3(256), lds, s, D, q, cents, nlist, dim,
                               ip, cnorm_max, M, static_cast<int>(cap), out_idx, out_val, split_kp, gkeys, pb);
    }
}

void launch_flat_select(float* D, const float* q, int nq, const float* rows, int n, int dim, int ip, float xnorm_max,
                        int k, const uint64_t* excl, const uint64_t* incl, const int64_t* labels, int64_t* out_ids,
                        float* out_dists, hipStream_t s) {
    if (nq <= 0) return;
    const size_t cap = coarse_select_cap(k);
    const size_t lds = query_lds_bytes(dim) + cap * 8;
    hipLaunchKernelGGL(k_flat_select, dim3(nq), dim3(256), lds, s, D, q, rows, n, dim, ip, xnorm_max, k,
                       static_cast<int>(cap), excl, incl, labels, out_ids, out_dists);
}


void launch_query_tables(const float* q, const DevIvf& ix, int nq, int mode, float* C, uint8_t* lut, float* scale,
                         float* offset, hipStream_t s) {
    const size_t lds = mode == kTableIpLut ? (static_cast<size_t>(ix.m) * 16 + ix.m + 16) * 4 : 0;
    hipLaunchKernelGGL(k_query_tables, dim3(nq), dim3(256), lds, s, q, ix, mode, C, lut, scale, offset);
}

size_t fs_sums_lds(SumsArgs& a) {
    const DevIvf& ix = a.ix;
    const size_t lb = static_cast<size_t>(ix.m4) * 16;
    const bool ip = a.ip_lut != nullptr;
    size_t fixed = 0;
    a.resid_in_lds = 0;
    if (!ip) {
        fixed += static_cast<size_t>(kSumsWaves) * ix.m * 4;
        if (!a.C && static_cast<size_t>(kSumsWaves) * ix.dim * 4 + fixed + lb <= kLdsMax) {
            a.resid_in_lds = 1;
            fixed += static_cast<size_t>(kSumsWaves) * ix.dim * 4;
        }
    }
    if (fixed + lb > kLdsMax) return 0;
    a.win = ip ? kSumsWin : static_cast<int>(std::min<size_t>(kSumsWin, (kLdsMax - fixed) / lb));
    a.nwin = (a.M + a.win - 1) / a.win;
    return (ip ? lb : static_cast<size_t>(a.win) * lb) + fixed;
}

namespace {

struct SumsKey {
    int m4, win, nwin, nclass, ordered, mfma;
    bool operator<(const SumsKey& o) const {
        return std::tie(m4, win, nwin, nclass, ordered, mfma) < std::tie(o.m4, o.win, o.nwin, o.nclass, o.ordered, o.mfma);
    }
};
std::mutex g_sums_mu;
std::map<SumsKey, int> g_sums_form;

// workgroups of k_fs_sums_queue that fill the chip once (a multiple of 8, one per XCD slot)
unsigned persistent_grid(size_t lds) {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_fs_sums_queue, kSumsWaves * kWave, lds) != hipSuccess ||
        per <= 0)
        per = 1;
    return static_cast<unsigned>((cus * per + 7) / 8 * 8);
}

size_t mfma_lds(const SumsArgs& a) { return static_cast<size_t>(kMfmaWaves) * mfma_wave_lds(a.ix.m4); }

template <int NK, bool IP, int SUB>
void launch_mfma_grid(const SumsArgs& a, unsigned grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_fs_sums_mfma<NK, IP, SUB>), dim3(grid), dim3(kMfmaWaves * kWave), lds, s, a, a.pairs);
}

template <bool IP>
void launch_mfma_nk(const SumsArgs& a, unsigned grid, size_t lds, hipStream_t s) {
    switch (a.ix.m4 / 2) {
#define QANN_NK(N)                                                                                              \
    case N:                                                                                                   \
        if (IP) launch_mfma_grid<N, true, 1>(a, grid, lds, s);                                                \
        else if (a.ix.sub == 1) launch_mfma_grid<N, false, 1>(a, grid, lds, s);                               \
        else if (a.ix.sub == 2) launch_mfma_grid<N, false, 2>(a, grid, lds, s);                               \
        else launch_mfma_grid<N, false, 4>(a, grid, lds, s);                                                  \
        break;
        QANN_NK(2) QANN_NK(4) QANN_NK(6) QANN_NK(8) QANN_NK(10) QANN_NK(12) QANN_NK(14) QANN_NK(16)
        QANN_NK(18) QANN_NK(20) QANN_NK(22) QANN_NK(24) QANN_NK(26) QANN_NK(28) QANN_NK(30) QANN_NK(32)
#undef QANN_NK
    default: break;
    }
}

// The matrix-core form: the pairs grouped by list (counted, offsets, placed),
// then one wave per unit of 32 pairs of one list.
void launch_mfma(const SumsArgs& a, int nq, hipStream_t s) {
    const int nlist = a.ix.nlist;
    const uint64_t items = static_cast<uint64_t>(nq) * a.M;
    const PairPlan& pp = a.pairs;
    if (nlist <= kPairsLdsLists) {
        const int groups = static_cast<int>(std::min<uint64_t>(kPairGroups, static_cast<uint64_t>(nq)));
        const size_t hl = static_cast<size_t>(nlist) * 4;
        hipLaunchKernelGGL(k_pairs_hist, dim3(groups), dim3(1024), hl, s, a, nq, pp.ghist);
        hipLaunchKernelGGL(k_pairs_lists, dim3(cdiv(nlist, 256 / kWave)), dim3(256), 0, s, pp.ghist, groups, nlist,
                           pp.tot);
        hipLaunchKernelGGL(k_pairs_offsets, dim3(1), dim3(1024), 0, s, pp.tot, nlist, pp.pair_off, pp.unit_off);
        hipLaunchKernelGGL(k_pairs_units, dim3(cdiv(nlist, 256 / kWave)), dim3(256), 0, s, pp.unit_off, nlist,
                           pp.unit_list);
        hipLaunchKernelGGL(k_pairs_place, dim3(groups), dim3(1024), hl, s, a, nq, pp.ghist, pp.pair_off, pp.items);
    } else {
        hipLaunchKernelGGL(k_order_zero, dim3(cdiv(nlist + 1, 256)), dim3(256), 0, s, pp.tot, nlist + 1);
        hipLaunchKernelGGL(k_pairs_count, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s, a, nq,
                           pp.tot);
        hipLaunchKernelGGL(k_pairs_offsets, dim3(1), dim3(1024), 0, s, pp.tot, nlist, pp.pair_off, pp.unit_off);
        hipLaunchKernelGGL(k_pairs_units, dim3(cdiv(nlist, 256 / kWave)), dim3(256), 0, s, pp.unit_off, nlist,
                           pp.unit_list);
        // the scatter's cursors start at each list's offset
        hipLaunchKernelGGL(k_pairs_cursors, dim3(cdiv(nlist, 256)), dim3(256), 0, s, pp.pair_off, nlist, pp.ghist);
        hipLaunchKernelGGL(k_pairs_scatter, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s, a, nq,
                           pp.ghist, pp.items);
    }
    if (a.grouped) (void)hipEventRecord(a.grouped, s);
    // units are at most one per 32 pairs plus one partial unit per list
    const uint64_t units = (items + 31) / 32 + static_cast<uint64_t>(nlist);
    const uint64_t groups = ((units + kMfmaWaves - 1) / kMfmaWaves + 7) / 8;
    const unsigned grid = static_cast<unsigned>(groups * 8);
    const size_t lds = mfma_lds(a);
    if (a.ip_lut) launch_mfma_nk<true>(a, grid, lds, s);
    else launch_mfma_nk<false>(a, grid, lds, s);
}

void sums_form(int form, const SumsArgs& a, int nq, size_t lds, hipStream_t s) {
    if (form == 2) {
        launch_mfma(a, nq, s);
        return;
    }
    if (form == 1) {
        hipLaunchKernelGGL(k_order_zero, dim3(1), dim3(64), 0, s, a.queue, 8);
        hipLaunchKernelGGL(k_fs_sums_queue, dim3(persistent_grid(lds)), dim3(kSumsWaves * kWave), lds, s, a);
        return;
    }
    const size_t grid = a.order ? static_cast<size_t>(8) * a.per_xcd * a.nwin : static_cast<size_t>(nq) * a.nwin;
    // large batches: registers for six waves per SIMD, whose extra waves hide
    // the table build's and the code loads' latency (12-15% faster at 10k
    // queries); small ones keep four, whose longer code load runs suit them
    // (QANN_AMD_SUMS_OCC=4|6 forces either)
    const char* e = std::getenv("QANN_AMD_SUMS_OCC");
    const int occ = e ? std::atoi(e) : (nq >= kSumsHighOccQueries ? 6 : 4);
    if (occ == 6)
        hipLaunchKernelGGL(k_fs_sums<6>, dim3(static_cast<unsigned>(grid)), dim3(kSumsWaves * kWave), lds, s, a);
    else
        hipLaunchKernelGGL(k_fs_sums<4>, dim3(static_cast<unsigned>(grid)), dim3(kSumsWaves * kWave), lds, s, a);
}

} // namespace

size_t pairs_ghist_words(int nlist, int nq) {
    return nlist <= kPairsLdsLists ? static_cast<size_t>(std::min(kPairGroups, std::max(nq, 1))) * nlist
                                   : static_cast<size_t>(nlist) + 1;
}

bool fs_mfma_ok(const SumsArgs& a) {
    if (!a.pairs.items || a.ix.m4 > 64 || a.ix.m4 < 4) return false;
    // per-list residual tables, and codebooks of long sub-vectors, stay on the per-query form
    if (!a.ip_lut && (!a.ix.B || a.ix.sub > kMfmaSubMax)) return false;
    return mfma_lds(a) <= kLdsMax;
}

bool fs_sums_reads_order(SumsArgs a, int nq) {
    // the matrix-core form groups the pairs by list itself: when it runs
    // (forced, or kept by the tuner for this problem class) the order is unused
    a.nq = nq;
    const bool mfma = fs_mfma_ok(a);
    const char* force = std::getenv("QANN_AMD_SUMS_FORM");
    if (force && std::strcmp(force, "mfma") == 0 && mfma) return false;
    if (force && *force) return true;
    if (!mfma || !a.queue) return true;
    int nclass = 0;
    while ((1 << nclass) < nq) ++nclass;
    const SumsKey key{a.ix.m4, a.win, a.nwin, nclass, a.order != nullptr, a.ip_lut ? 1 : 2};
    std::lock_guard<std::mutex> g(g_sums_mu);
    const auto it = g_sums_form.find(key);
    return it == g_sums_form.end() || it->second != 2;
}

bool fs_sums_reads_tables(SumsArgs a, int nq) {
    // the matrix-core form builds each pair's L2 table from B and the PQ
    // codebook: when it runs (forced, or kept by the tuner for this problem
    // class) the per-query tables are unused
    a.nq = nq;
    if (a.ip_lut) return true;
    const bool mfma = fs_mfma_ok(a);
    const char* force = std::getenv("QANN_AMD_SUMS_FORM");
    if (force && std::strcmp(force, "mfma") == 0 && mfma) return false;
    if (force && *force) return true;
    if (!mfma || !a.queue) return true;
    int nclass = 0;
    while ((1 << nclass) < nq) ++nclass;
    const SumsKey key{a.ix.m4, a.win, a.nwin, nclass, a.order != nullptr, 2};
    std::lock_guard<std::mutex> g(g_sums_mu);
    const auto it = g_sums_form.find(key);
    return it == g_sums_form.end() || it->second != 2;
}

int launch_fs_sums(SumsArgs a, int nq, hipStream_t s) {
    const size_t lds = fs_sums_lds(a);
    a.nq = nq;
    a.per_xcd = (nq + 7) / 8;
    const bool mfma = fs_mfma_ok(a);
    // QANN_AMD_SUMS_FORM=grid, =queue or =mfma fixes the form (tests run each)
    const char* force = std::getenv("QANN_AMD_SUMS_FORM");
    if (force && std::strcmp(force, "mfma") == 0 && mfma) {
        sums_form(2, a, nq, lds, s);
        return 2;
    }
    if (!a.queue) {
        sums_form(0, a, nq, lds, s);
        return 0;
    }
    if (force && (std::strcmp(force, "grid") == 0 || std::strcmp(force, "queue") == 0)) {
        const int f = std::strcmp(force, "queue") == 0 ? 1 : 0;
        sums_form(f, a, nq, lds, s);
        return f;
    }
    int nclass = 0;
    while ((1 << nclass) < nq) ++nclass;
    const SumsKey key{a.ix.m4, a.win, a.nwin, nclass, a.order != nullptr, mfma ? (a.ip_lut ? 1 : 2) : 0};
    int form = -1;
    {
        std::lock_guard<std::mutex> g(g_sums_mu);
        auto it = g_sums_form.find(key);
        if (it != g_sums_form.end()) form = it->second;
    }
    if (form < 0) {
        // the forms write the same outputs, so the trials can run on the real problem
        float t[3] = {0.0f, 0.0f, 1e30f};
        hipEvent_t e0, e1;
        form = 0;
        if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
            SumsArgs trial = a;
            trial.grouped = nullptr;
            for (int f = 0; f < (mfma ? 3 : 2); ++f) {
                (void)hipEventRecord(e0, s);
                sums_form(f, trial, nq, lds, s);
                (void)hipEventRecord(e1, s);
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t[f], e0, e1) != hipSuccess)
                    t[f] = 1e30f;
            }
            form = t[1] < t[0] ? 1 : 0;
            if (t[2] < t[form]) form = 2;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        const char* log = std::getenv("QANN_AMD_TUNE_LOG");
        if (log && *log && *log != '0')
            std::fprintf(stderr,
                         "qann_amd tune: fast-scan sums nq %d m4 %d windows %d -> %s (ms: %.3f grid, %.3f queue of %u "
                         "workgroups, %.3f matrix cores)\n",
                         nq, a.ix.m4, a.nwin,
                         form == 2 ? "matrix cores" : form ? "persistent queue" : "one workgroup per window", t[0],
                         t[1], persistent_grid(lds), t[2]);
        std::lock_guard<std::mutex> g(g_sums_mu);
        g_sums_form[key] = form;
    }
    sums_form(form, a, nq, lds, s);
    return form;
}

void launch_query_order(const uint16_t* probe_idx, int M, int nq, int nlist, uint32_t* hist, uint32_t* order,
                        hipStream_t s) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_order_zero, dim3(cdiv(nlist, 256)), dim3(256), 0, s, hist, nlist);
    hipLaunchKernelGGL(k_order_count, dim3(cdiv(nq, 256)), dim3(256), 0, s, probe_idx, M, nq, hist);
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, s, hist, nlist);
    hipLaunchKernelGGL(k_order_scatter, dim3(cdiv(nq, 256)), dim3(256), 0, s, probe_idx, M, nq, hist, order);
}

void launch_fs_lut_check(const DevIvf& ix, int n, const float* negzero, uint8_t* lut, float* scale, float* offset,
                         hipStream_t s) {
    const size_t lds = static_cast<size_t>(ix.m4) * 16 + static_cast<size_t>(ix.m) * 4;
    hipLaunchKernelGGL(k_fs_lut_check, dim3(n), dim3(kWave), lds, s, ix, negzero, lut, scale, offset);
}

void launch_fs_block_check(const uint8_t* luts, int m, int m4, const uint8_t* blocks, int nblocks, uint16_t* out,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_fs_block_check, dim3(nblocks), dim3(kWave), static_cast<size_t>(m4) * 16, s, luts, m, m4,
                       blocks, out);
}

size_t ivf_dists_lds(DistArgs& a) {
    const DevIvf& ix = a.ix;
    const size_t dim4 = static_cast<size_t>(ix.dim) * 4, qb = query_lds_bytes(ix.dim);
    const size_t tbytes = static_cast<size_t>(ix.m) * ix.K * 4, sbytes = static_cast<size_t>(ix.K) * 4;
    const bool adc = a.mode <= kDistL2Resid || a.mode == kDistSdc;
    size_t fixed = qb + (a.mode == kDistSq ? 2 * qb : 0);
    a.resid_in_lds = 0;
    a.win = kSumsWin;
    a.mc = ix.m;
    // the 64 KB budget, or less under a lowered device limit
    const size_t room = ivf_dists_room(), budget = std::min(kLdsMax, room);
    const bool resid = a.mode == kDistL2Resid || a.mode == kDistSdc;
    if (resid && fixed + kSumsWaves * dim4 + tbytes <= budget) {
        a.resid_in_lds = 1;
        fixed += kSumsWaves * dim4;
    }
    if (adc) {
        // windows of whole tables within the budget; a table that alone
        // exceeds it takes what the device allows, one probe per workgroup,
        // and beyond that the table is built mc subspaces at a time
        if (fixed + tbytes <= budget) {
            if (a.mode != kDistIp) a.win = static_cast<int>(std::min<size_t>(kSumsWin, (budget - fixed) / tbytes));
        } else {
            if (a.mode != kDistIp) a.win = 1;
            if (fixed + sbytes > room) return 0;
            a.mc = static_cast<int>(std::min<size_t>(static_cast<size_t>(ix.m), (room - fixed) / sbytes));
        }
        fixed += static_cast<size_t>(a.mode == kDistIp ? 1 : a.win) * a.mc * sbytes;
    }
    if (fixed > room) return 0;
    a.nwin = (a.M + a.win - 1) / a.win;
    // float and SQ lists of queries read from global memory need none; 0 is
    // the failure value, so they ask for 16 bytes
    return std::max<size_t>(fixed, 16);
}

// Chunk-major 8-bit SQ lists go to k_sq8_dists unless QANN_AMD_SQ8=thread
// keeps them on k_ivf_dists (one thread per vector, one byte load per dimension).
static bool sq8_rows_form(const DistArgs& a) {
    if (a.mode != kDistSq || !a.ix.sq_chunked) return false;
    const char* e = std::getenv("QANN_AMD_SQ8");
    return !(e && std::strcmp(e, "thread") == 0);
}

void launch_ivf_dists(DistArgs a, int nq, hipStream_t s) {
    const size_t lds = ivf_dists_lds(a);
    const dim3 grid(static_cast<unsigned>(static_cast<size_t>(nq) * a.nwin)), block(kSumsWaves * kWave);
    if (sq8_rows_form(a)) {
        if (a.ix.metric == 2)
            hipLaunchKernelGGL(k_sq8_dists<true>, grid, block, lds, s, a);
        else
            hipLaunchKernelGGL(k_sq8_dists<false>, grid, block, lds, s, a);
        return;
    }
    hipLaunchKernelGGL(k_ivf_dists, grid, block, lds, s, a);
}

void launch_probe_plan(const DevIvf& ix, int nq, int M, int nprobe, int min_probes, int wanted,
                       const uint16_t* probe_idx, const uint32_t* list_allowed, int* nprobe_q, uint32_t* item_blk,
                       int* status, hipStream_t s, const PerQuery& pq) {
    // QANN_AMD_PROBE_PLAN=thread keeps one thread per query
    const char* e = std::getenv("QANN_AMD_PROBE_PLAN");
    if (e && std::strcmp(e, "thread") == 0)
        hipLaunchKernelGGL(k_probe_plan, dim3(cdiv(nq, 128)), dim3(128), 0, s, ix, nq, M, nprobe, min_probes, wanted,
                           probe_idx, list_allowed, nprobe_q, item_blk, status, pq);
    else
        hipLaunchKernelGGL(k_probe_plan_wave, dim3(cdiv(nq, 4)), dim3(256), 0, s, ix, nq, M, nprobe, min_probes,
                           wanted, probe_idx, list_allowed, nprobe_q, item_blk, status, pq);
}

void launch_list_allowed(const DevIvf& ix, const uint64_t* filter, uint32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_list_allowed, dim3(ix.nlist), dim3(256), 0, s, ix, filter, out);
}

void launch_fs_replay(const ReplayArgs& a, int nq, hipStream_t s) {
    const size_t lds = a.g_dist ? 0 : static_cast<size_t>(2 * a.kp) * 12;
    // QANN_AMD_REPLAY=plain / stage / win forces a form (where it applies)
    int forced = -1;
    if (const char* e = std::getenv("QANN_AMD_REPLAY")) {
        if (std::strcmp(e, "plain") == 0) forced = 0;
        else if (std::strcmp(e, "stage") == 0) forced = 1;
        else if (std::strcmp(e, "win") == 0) forced = 2;
        else if (std::strcmp(e, "gather") == 0) forced = 3;
    }
    // the window form, whenever its TopK and window fit LDS (its fixed arrays
    // take about 3.3 KB more)
    // the window's blocks (QANN_AMD_REPLAY_WINDOW overrides kReplayWindow)
    ReplayArgs aw = a;
    aw.window = static_cast<int>(kReplayWindow);
    if (const char* w = std::getenv("QANN_AMD_REPLAY_WINDOW")) {
        const int v = std::atoi(w);
        if (v >= 64 && v <= 8192) aw.window = v;
    }
    const size_t wl = lds + static_cast<size_t>(aw.window) * (a.fdists ? 4 : 2);
    if ((forced < 0 || forced >= 2) && !a.g_dist && wl + 4096 <= kLdsMax) {
        // the gathered form where the search asks for it (QANN_AMD_REPLAY=gather
        // forces it), else a probe's blocks at a time
        if (forced == 2 || (forced < 0 && !a.gather)) {
            if (a.fdists) hipLaunchKernelGGL(k_fl_replay_win, dim3(nq), dim3(kWave), wl, s, aw);
            else hipLaunchKernelGGL(k_fs_replay_win, dim3(nq), dim3(kWave), wl, s, aw);
        } else {
            if (a.fdists) hipLaunchKernelGGL(k_fl_replay_gather, dim3(nq), dim3(kWave), wl, s, aw);
            else hipLaunchKernelGGL(k_fs_replay_gather, dim3(nq), dim3(kWave), wl, s, aw);
        }
        return;
    }
    // staged loads pay off while the batch leaves most of the chip's wave slots idle
    const bool stage = forced == 1 || (forced < 0 && nq < kReplayStageBelow);
    if (a.fdists) {
        if (stage) hipLaunchKernelGGL(k_fl_replay<true>, dim3(nq), dim3(kWave), lds, s, a);
        else hipLaunchKernelGGL(k_fl_replay<false>, dim3(nq), dim3(kWave), lds, s, a);
    } else {
        if (stage) hipLaunchKernelGGL(k_fs_replay<true>, dim3(nq), dim3(kWave), lds, s, a);
        else hipLaunchKernelGGL(k_fs_replay<false>, dim3(nq), dim3(kWave), lds, s, a);
    }
}

size_t rerank_keys(int kc) {
    size_t np = 1;
    while (np < static_cast<size_t>(kc)) np <<= 1;
    return np;
}

void launch_rerank(const float* q, int nq, int dim, int ip, const float* rows, const int64_t* cand,
                   const int* cand_count, int kc, int k, const int64_t* labels, int64_t* out_ids, float* out_dists,
                   uint64_t* gkeys, hipStream_t s, const PerQuery& pq) {
    const size_t np = rerank_keys(kc);
    const size_t lds = query_lds_bytes(dim) + (gkeys ? 0 : np * 8);
    // one thread per candidate up to 256 (k' = 100: two waves, not four with
    // more than half of them idle)
    const unsigned threads = np <= 64 ? 64u : np <= 128 ? 128u : 256u;
    hipLaunchKernelGGL(k_rerank, dim3(nq), dim3(threads), lds, s, q, dim, ip, rows, cand, cand_count, kc, k, labels,
                       out_ids, out_dists, gkeys, static_cast<int>(np), pq);
}

size_t block_select_bytes(int k, bool by_label) {
    const int threads = 256;
    size_t cap = 1;
    while (cap < static_cast<size_t>(k) + threads) cap <<= 1;
    size_t kp2 = 1;
    while (kp2 < static_cast<size_t>(k)) kp2 <<= 1;
    size_t b = cap * 8;
    if (by_label) b += ((kp2 + 1) & ~size_t{1}) * 4 + kp2 * 12;
    return (b + 255) & ~size_t{255};
}

void launch_block_select(const float* q, int nq, int dim, int ip, const SlotScorer& sc, const uint32_t* slot_list,
                         uint32_t n_slots, const uint64_t* deleted, int use_fkey, int final_by_label, int k,
                         const int64_t* labels, int64_t* out_ids, float* out_dists, unsigned char* gbuf,
                         hipStream_t s) {
    const int threads = 256;
    int cap = 1;
    while (cap < k + threads) cap <<= 1;
    const size_t buf = block_select_bytes(k, final_by_label && labels);
    const size_t lds = query_lds_bytes(dim) + (gbuf ? 0 : buf);
    hipLaunchKernelGGL(k_block_select, dim3(nq), dim3(threads), lds, s, q, dim, ip, sc, slot_list, n_slots, deleted,
                       use_fkey, final_by_label, k, cap, labels, out_ids, out_dists, gbuf, buf);
}

} // namespace qd
