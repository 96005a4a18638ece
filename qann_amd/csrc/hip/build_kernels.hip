// Index-building kernels for gfx950: the assign step of PQ codebook training,
// PQ / SQ encoding of added vectors and fast-scan block packing. Each keeps
// the reference's arithmetic order, so the codes and codebooks it produces are
// the reference's bytes.
#include "build_kernels.hpp"
#include "device_common.hpp"

namespace qd {

namespace {

constexpr float kFloatMax = 3.402823466e+38f;

// Grids are capped and the kernels stride over their items, so that item
// counts above 2^32 (100M rows x 64 subspaces) launch as well.
inline unsigned grid_for(uint64_t items, int block) {
    const uint64_t g = (items + static_cast<uint64_t>(block) - 1) / static_cast<uint64_t>(block);
    return static_cast<unsigned>(g < (uint64_t{1} << 20) ? g : (uint64_t{1} << 20));
}

__device__ __forceinline__ uint64_t first_item() {
    return static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ uint64_t item_stride() { return static_cast<uint64_t>(gridDim.x) * blockDim.x; }

} // namespace

// One thread per (row, subspace): the K distances in l2_distance's order
// (exact_avx2), the first strict minimum below FLT_MAX (nearest_centroid,
// src/kmeans.cpp:102-114, starts from {FLT_MAX, 0}). The subspace index runs
// fastest, so a wave reads a row's consecutive sub-vectors.
__global__ void k_assign_subspaces(const float* __restrict__ rows, uint64_t n, int dim, int m, int sub,
                                   const float* __restrict__ cents, int K, uint8_t* __restrict__ out) {
    const uint64_t items = n * static_cast<uint64_t>(m);
    for (uint64_t t = first_item(); t < items; t += item_stride()) {
        const uint64_t row = t / static_cast<uint64_t>(m);
        const int s = static_cast<int>(t % static_cast<uint64_t>(m));
        const float* x = rows + row * static_cast<uint64_t>(dim) + static_cast<uint64_t>(s) * sub;
        const float* c = cents + static_cast<uint64_t>(s) * K * sub;
        float best = kFloatMax;
        int pick = 0;
        for (int j = 0; j < K; ++j) {
            const float dist = exact_avx2<false>(x, c + static_cast<uint64_t>(j) * sub, sub);
            if (dist < best) {
                best = dist;
                pick = j;
            }
        }
        out[t] = static_cast<uint8_t>(pick);
    }
}

// One thread per (row, subspace). The residual r = x - centre(list) component
// by component, then per centroid c: (r_0 - y_0c)^2 unfused and the rest fused,
// fma(r_j - y_jc, r_j - y_jc, acc), as encode_subspaces computes it; the first
// index of the smallest sum.
__global__ void k_pq_encode(const float* __restrict__ rows, uint64_t n, int dim, const uint16_t* __restrict__ list,
                            const float* __restrict__ centers, const float* __restrict__ cent_t, int m, int sub, int K,
                            uint8_t* __restrict__ out) {
    const uint64_t items = n * static_cast<uint64_t>(m);
    for (uint64_t t = first_item(); t < items; t += item_stride()) {
        const uint64_t row = t / static_cast<uint64_t>(m);
        const int s = static_cast<int>(t % static_cast<uint64_t>(m));
        const float* x = rows + row * static_cast<uint64_t>(dim) + static_cast<uint64_t>(s) * sub;
        const float* e = centers + static_cast<uint64_t>(list[row]) * dim + static_cast<uint64_t>(s) * sub;
        const float* y = cent_t + static_cast<uint64_t>(s) * sub * K;
        float best = 0.0f;
        int pick = 0;
        for (int c = 0; c < K; ++c) {
            float d = (x[0] - e[0]) - y[c];
            float acc = d * d;
            for (int j = 1; j < sub; ++j) {
                d = (x[j] - e[j]) - y[static_cast<uint64_t>(j) * K + c];
                acc = fmaf(d, d, acc);
            }
            if (c == 0 || acc < best) {
                best = acc;
                pick = c;
            }
        }
        out[t] = static_cast<uint8_t>(pick);
    }
}

// One thread per code byte: round((x - vmin) / step) (0 where step is 0),
// clamped to [0, 2^bits - 1]; 4-bit codes two to a byte, even dimension low.
__global__ void k_sq_encode(const float* __restrict__ rows, uint64_t n, int dim, const float* __restrict__ vmin,
                            const float* __restrict__ step, int bits, uint8_t* __restrict__ out) {
    const int vbytes = bits == 8 ? dim : (dim + 1) / 2;
    const float top = static_cast<float>((1 << bits) - 1);
    const uint64_t items = n * static_cast<uint64_t>(vbytes);
    for (uint64_t t = first_item(); t < items; t += item_stride()) {
        const uint64_t row = t / static_cast<uint64_t>(vbytes);
        const int b = static_cast<int>(t % static_cast<uint64_t>(vbytes));
        const float* x = rows + row * static_cast<uint64_t>(dim);
        auto code = [&](int d) -> uint32_t {
            const float c = step[d] > 0.0f ? roundf(__fdiv_rn(x[d] - vmin[d], step[d])) : 0.0f;
            return static_cast<uint32_t>(fminf(fmaxf(c, 0.0f), top));
        };
        uint32_t v;
        if (bits == 8) {
            v = code(b);
        } else {
            v = code(2 * b);
            if (2 * b + 1 < dim) v |= code(2 * b + 1) << 4;
        }
        out[t] = static_cast<uint8_t>(v);
    }
}

// One thread per output byte (block, subspace s, byte j): the code of position
// j in the low nibble and of position j + 16 in the high nibble.
__global__ void k_fs_pack(const uint8_t* __restrict__ codes, int m, const uint32_t* __restrict__ row_at,
                          uint64_t nblocks, uint8_t* __restrict__ out) {
    const uint64_t bb = static_cast<uint64_t>(m) * 16;
    const uint64_t items = nblocks * bb;
    for (uint64_t t = first_item(); t < items; t += item_stride()) {
        const uint64_t b = t / bb;
        const int within = static_cast<int>(t % bb), s = within / 16, j = within % 16;
        const uint32_t lo = row_at[b * 32 + j], hi = row_at[b * 32 + j + 16];
        uint32_t v = 0;
        if (lo != 0xFFFFFFFFu) v = codes[static_cast<uint64_t>(lo) * m + s] & 0xFu;
        if (hi != 0xFFFFFFFFu) v |= (codes[static_cast<uint64_t>(hi) * m + s] & 0xFu) << 4;
        out[t] = static_cast<uint8_t>(v);
    }
}

void launch_assign_subspaces(const float* rows, uint64_t n, int dim, int m, int sub, const float* cents, int K,
                             uint8_t* out, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_assign_subspaces, dim3(grid_for(n * static_cast<uint64_t>(m), 256)), dim3(256), 0, s, rows, n,
                       dim, m, sub, cents, K, out);
}

void launch_pq_encode(const float* rows, uint64_t n, int dim, const uint16_t* list, const float* centers,
                      const float* cent_t, int m, int sub, int K, uint8_t* out, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_pq_encode, dim3(grid_for(n * static_cast<uint64_t>(m), 256)), dim3(256), 0, s, rows, n, dim,
                       list, centers, cent_t, m, sub, K, out);
}

void launch_sq_encode(const float* rows, uint64_t n, int dim, const float* vmin, const float* step, int bits,
                      uint8_t* out, hipStream_t s) {
    if (n == 0) return;
    const uint64_t vbytes = bits == 8 ? static_cast<uint64_t>(dim) : static_cast<uint64_t>(dim + 1) / 2;
    hipLaunchKernelGGL(k_sq_encode, dim3(grid_for(n * vbytes, 256)), dim3(256), 0, s, rows, n, dim, vmin, step, bits,
                       out);
}

void launch_fs_pack(const uint8_t* codes, int m, const uint32_t* row_at, uint64_t nblocks, uint8_t* out,
                    hipStream_t s) {
    if (nblocks == 0) return;
    hipLaunchKernelGGL(k_fs_pack, dim3(grid_for(nblocks * static_cast<uint64_t>(m) * 16, 256)), dim3(256), 0, s,
                       codes, m, row_at, nblocks, out);
}

} // namespace qd
