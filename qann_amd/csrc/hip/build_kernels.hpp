// Launch wrappers of the index-building kernels: the assign step of PQ
// codebook training, PQ and SQ encoding of added vectors, and fast-scan block
// packing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qd {

// For each of n rows (dim floats, contiguous) and each of m subspaces, the
// centroid (of K, cents [s][c][sub]) at the smallest exact squared L2 distance
// in l2_distance's order, the first on ties (nearest_centroid,
// src/kmeans.cpp:102-114). out: n * m codes.
void launch_assign_subspaces(const float* rows, uint64_t n, int dim, int m, int sub, const float* cents, int K,
                             uint8_t* out, hipStream_t s);

// PQCodebook::encode (src/pq.cpp:87-163) of each row's residual to the centre
// of its list: out n * m codes. centers: nlist * dim; cent_t [s][sub][K].
void launch_pq_encode(const float* rows, uint64_t n, int dim, const uint16_t* list, const float* centers,
                      const float* cent_t, int m, int sub, int K, uint8_t* out, hipStream_t s);

// SQ::encode (src/sq.cpp): out n * vbytes (dim, or ceil(dim / 2) for 4 bits).
void launch_sq_encode(const float* rows, uint64_t n, int dim, const float* vmin, const float* step, int bits,
                      uint8_t* out, hipStream_t s);

// fs_set (src/fast_scan.hpp:25-28) over nblocks blocks of 32 positions:
// position p of block b holds the row at row_at[b * 32 + p] (0xFFFFFFFF:
// none, its nibble zero). Writes nblocks * m * 16 bytes.
void launch_fs_pack(const uint8_t* codes, int m, const uint32_t* row_at, uint64_t nblocks, uint8_t* out,
                    hipStream_t s);

} // namespace qd
