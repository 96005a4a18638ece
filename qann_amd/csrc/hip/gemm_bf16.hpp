// BF16 MFMA scoring for exact FlatIndex search (and other dense score
// matrices), gfx950.
//
// FlatIndex::query_impl (src/flat_index.cpp:33-64) returns the k nearest rows
// by exact FP32 distance. Here a BF16 GEMM with FP32 accumulation scores every
// (query, row) approximately; a proven bound e on |approximate - exact| (rows
// and queries rounded to BF16, FP32 accumulation, the exact score's own
// rounding) then keeps exactly the rows that can still be among the k nearest:
// those whose approximate score is within 2e of the k-th smallest. They are
// rescored exactly, in the reference's summation order, and selected by
// (distance, slot): the result is the reference's, bit for bit.
//
// The k-th smallest approximate score is not known while the GEMM streams the
// rows, so a first, strided GEMM over a sample of the rows gives an upper
// bound for it; the full GEMM keeps only rows under that bound plus 2e (a few
// thousand per query at most), and the exact selection runs on those. A query
// whose candidates overflow their buffer is reported, and the caller runs it
// through the FP32 path instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace qd {

constexpr int kBf16K = 32;          // K granularity: rows are zero-padded to a multiple
constexpr int kFlatCandCap = 4096;  // candidates per query kept by the filtered GEMM

inline int bf16_kpad(int dim) { return (dim + kBf16K - 1) / kBf16K * kBf16K; }

// Rows to BF16 (round to nearest even), zero-padded to kp, with maxes[0] =
// max ||x - bf16(x)|| and maxes[1] = max ||bf16(x)|| over the rows (upper
// bounds; maxes must be zeroed first).
void launch_rows_bf16(const float* rows, uint64_t n, int dim, int kp, uint16_t* out, float* maxes, hipStream_t s);

// Per query: BF16 copy (zero-padded) and qinfo[3i..3i+2] = upper bounds on
// ||q||, ||bf16(q)|| and ||q - bf16(q)||.
void launch_queries_bf16(const float* q, int nq, int dim, int kp, uint16_t* out, float* qinfo, hipStream_t s);

// Rows as three BF16 segments of kp for the split coarse product (order 0:
// [hi | hi | lo] for queries, order 1: [hi | lo | hi] for centroids), out
// n * 3 kp.
void launch_split_bf16(const float* x, uint64_t n, int dim, int kp, int order, uint16_t* out, hipStream_t s);

// Coarse scores by the split product: D[i][l] = csq[l] - 2 s (or -s for inner
// product) with s = hi.hi + hi.lo + lo.hi of query i and centroid l, one BF16
// GEMM over K = 3 kp with FP32 accumulation.
void launch_coarse_split(const uint16_t* q3, int nq, const uint16_t* c3, int nlist, int kp, const float* csq, int ip,
                         float* D, hipStream_t s);

// The arguments shared by the flat-search launches.
struct FlatBf16 {
    const uint16_t* xb;    // rows, n * kp BF16
    const float* rows;     // rows, n * dim FP32 (the exact rescoring)
    const float* xsq;      // squared norms in the exact order (L2)
    const float* maxes;    // device: max ||x - bf16(x)||, max ||bf16(x)||
    float xnorm_max;       // upper bound on every row's norm
    int n, dim, kp, ip;
    const uint64_t* excl;  // rows that may not be returned (deleted), or null
    const uint64_t* incl;  // rows that may be returned (a filter), or null
    const int64_t* labels; // slot -> id, or null
};

// Tile minima per query of a sample of ns rows (the sample GEMM's row tiles).
inline int flat_sample_tiles(int ns) { return (ns + 127) / 128; }

// One chunk of queries (qb, qinfo from launch_queries_bf16; q the FP32 rows):
// sample GEMM (every stride-th row, stride * ns >= n) into Ds, its tile minima
// (nq * flat_sample_tiles(ns) floats), limits,
// filtered GEMM into the candidate buffers (cnt nq, cd / ci nq * kFlatCandCap),
// then the exact selection into out_ids / out_dists (nq * k). Queries it
// cannot finish go to ovf (ovf_cnt counts them; the caller zeroes it).
void launch_flat_bf16(const FlatBf16& f, const float* q, const uint16_t* qb, const float* qinfo, int nq, int k,
                      int want, int stride, int ns, float* Ds, float* lim, uint32_t* cnt, float* cd, uint32_t* ci,
                      int64_t* out_ids, float* out_dists, uint32_t* ovf, uint32_t* ovf_cnt, int ovf_base,
                      hipStream_t s);

// Autotuned GEMM classes, and the tile shape chosen for one (index 0: 64x64
// wave tiles, 1: 64x128, 2: 128x64; -1: not tuned yet).
enum { kGemmScores = 0, kGemmFilter = 1 };
int gemm_tuned_variant(int kind, int kp, int M, int N);

} // namespace qd
