// Device helpers shared by the qann_amd search kernels (gfx950, wave64).
//
// All float arithmetic here is written out operation by operation: the
// library is compiled with -ffp-contract=off, and every fused multiply-add
// is an explicit fmaf, so each result has the bits of the reference's CPU
// build (g++ -O3 -mavx2 -mfma, AVX2 kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qd {

constexpr int kWave = 64;
constexpr int kFsBlock = 32;   // vectors per fast-scan block (src/fast_scan.hpp:19)
constexpr float kLutMax = 63.0f; // largest LUT byte (src/fast_scan.hpp:39)

// Kernels that score rows against one query copy it into LDS when it has at
// most kQueryLdsDims floats (padded to 16 bytes); a longer query is read from
// global memory, where every thread of the workgroup reads the same cached row.
constexpr int kQueryLdsDims = 8192;
__host__ __device__ constexpr size_t query_lds_bytes(int dim) {
    return dim <= kQueryLdsDims ? (static_cast<size_t>(dim) * 4 + 15) & ~static_cast<size_t>(15) : 0;
}

// Orderable key of a float (src/topk.hpp:80-83): unsigned order = float order,
// with -0.0 below +0.0.
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t b = __float_as_uint(f);
    return b ^ (static_cast<uint32_t>(static_cast<int32_t>(b) >> 31) | 0x80000000u);
}

// Inverse of fkey (src/topk.hpp:86-88).
__device__ __forceinline__ float from_fkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k ^ 0x80000000u) : ~k);
}

// Key for orders that compare floats with operator< (std::pair / heap
// orders of the reference): -0.0 and +0.0 tie.
__device__ __forceinline__ uint32_t ckey(float f) { return fkey(f == 0.0f ? 0.0f : f); }

// Lanes of this wave below the current one with their bit set in mask.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ int lane_id() { return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u))); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Orders the LDS traffic of a kernel whose workgroup is exactly one wave.
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}

template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), Ctrl, 0xF, 0xF, false));
}

__device__ __forceinline__ uint32_t lane_u32(uint32_t x, int l) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), l));
}

__device__ __forceinline__ float lane_f32(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}

// The wave's largest value, for a wave whose 64 lanes are all active, in a
// scalar register: DPP maxima within each row of 16 lanes (lanes i ^ 1,
// i ^ 2, then the half-row and row mirrors), then the four rows' maxima.
__device__ __forceinline__ uint32_t wave_max_u32_dpp(uint32_t v) {
    v = max(v, dpp_u32<0xB1>(v));  // quad_perm [1, 0, 3, 2]
    v = max(v, dpp_u32<0x4E>(v));  // quad_perm [2, 3, 0, 1]
    v = max(v, dpp_u32<0x141>(v)); // row_half_mirror
    v = max(v, dpp_u32<0x140>(v)); // row_mirror
    return max(max(lane_u32(v, 0), lane_u32(v, 16)), max(lane_u32(v, 32), lane_u32(v, 48)));
}

// ---------------------------------------------------------------------------
// Exact distances in the reference's AVX2 order (src/distance.cpp:69-153 as
// g++ compiles them): 4 x 8 lane accumulators over chunks of 32, 8-wide steps
// into the first, the fixed horizontal sum, then the tail of n % 8 elements
// (first 4 unfused when at least 4 remain, the rest fused). One thread
// computes one distance. When both rows are 16-byte aligned the chunks are
// read 16 bytes at a time (a quarter of the load instructions, which bound
// the scattered row reads of the re-rank and the selects); every accumulator
// still takes its elements in the same order.
// ---------------------------------------------------------------------------

template <bool IP>
__device__ __forceinline__ void exact_step(float& acc, float x, float y) {
    if (IP) {
        acc = fmaf(x, y, acc);
    } else {
        const float d = x - y;
        acc = fmaf(d, d, acc);
    }
}

template <bool IP>
__device__ __forceinline__ void exact_step4(float* acc, float4 x, float4 y) {
    exact_step<IP>(acc[0], x.x, y.x);
    exact_step<IP>(acc[1], x.y, y.y);
    exact_step<IP>(acc[2], x.z, y.z);
    exact_step<IP>(acc[3], x.w, y.w);
}

template <bool IP>
__device__ __forceinline__ float exact_avx2(const float* __restrict__ a, const float* __restrict__ b, int n) {
    float acc[32];
#pragma unroll
    for (int t = 0; t < 32; ++t) acc[t] = 0.0f;
    int i = 0;
    if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15u) == 0) {
        const float4* a4 = reinterpret_cast<const float4*>(a);
        const float4* b4 = reinterpret_cast<const float4*>(b);
        for (; i + 32 <= n; i += 32) {
#pragma unroll
            for (int t = 0; t < 8; ++t) exact_step4<IP>(acc + 4 * t, a4[i / 4 + t], b4[i / 4 + t]);
        }
        for (; i + 8 <= n; i += 8) {
            exact_step4<IP>(acc, a4[i / 4], b4[i / 4]);
            exact_step4<IP>(acc + 4, a4[i / 4 + 1], b4[i / 4 + 1]);
        }
    } else {
        for (; i + 32 <= n; i += 32) {
#pragma unroll
            for (int t = 0; t < 32; ++t) exact_step<IP>(acc[t], a[i + t], b[i + t]);
        }
        for (; i + 8 <= n; i += 8) {
#pragma unroll
            for (int t = 0; t < 8; ++t) exact_step<IP>(acc[t], a[i + t], b[i + t]);
        }
    }
    float v[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) v[l] = (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]);
    float res = ((v[0] + v[4]) + (v[1] + v[5])) + ((v[2] + v[6]) + (v[3] + v[7]));
    const int r = n - i;
    int j = 0;
    if (r >= 4) {
        for (; j < 4; ++j) {
            const float x = a[i + j], y = b[i + j];
            float p;
            if (IP) {
                p = x * y;
            } else {
                const float d = x - y;
                p = d * d;
            }
            res = res + p;
        }
    }
    for (; j < r; ++j) {
        const float x = a[i + j], y = b[i + j];
        if (IP) {
            res = fmaf(x, y, res);
        } else {
            const float d = x - y;
            res = fmaf(d, d, res);
        }
    }
    return res;
}

// DenseIndex::exact_score (include/vecengine/index.hpp:269-271): squared L2,
// or -(a . b) for inner product.
__device__ __forceinline__ float exact_score(bool ip, const float* a, const float* b, int n) {
    return ip ? -exact_avx2<true>(a, b, n) : exact_avx2<false>(a, b, n);
}

// A float from lane (this lane ^ 1) or (this lane ^ 2) of its quad.
__device__ __forceinline__ float quad_xor1(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float quad_xor2(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x4E, 0xF, 0xF, true));
}

// exact_avx2 computed by the four lanes of a quad (g = lane % 4, all four
// active): lane g holds accumulators 8 g .. 8 g + 7 and reads elements
// 8 g .. 8 g + 7 of each chunk of 32, lane 0 the 8-wide steps after the
// chunks; the fold (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]) pairs
// lanes g ^ 1, then g ^ 2 (each add commutes, so every lane gets the same
// bits); the horizontal sum and the tail are computed by every lane. Every
// lane returns exact_avx2's result, with a quarter of its loads and
// accumulators.
template <bool IP>
__device__ __forceinline__ float exact_avx2_quad(const float* __restrict__ a, const float* __restrict__ b, int n,
                                                 int g) {
    float acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = 0.0f;
    int i = 0;
    if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15u) == 0) {
        const float4* a4 = reinterpret_cast<const float4*>(a);
        const float4* b4 = reinterpret_cast<const float4*>(b);
        for (; i + 32 <= n; i += 32) {
            const int o = (i + 8 * g) / 4;
            exact_step4<IP>(acc, a4[o], b4[o]);
            exact_step4<IP>(acc + 4, a4[o + 1], b4[o + 1]);
        }
        for (; i + 8 <= n; i += 8) {
            if (g == 0) {
                exact_step4<IP>(acc, a4[i / 4], b4[i / 4]);
                exact_step4<IP>(acc + 4, a4[i / 4 + 1], b4[i / 4 + 1]);
            }
        }
    } else {
        for (; i + 32 <= n; i += 32) {
#pragma unroll
            for (int t = 0; t < 8; ++t) exact_step<IP>(acc[t], a[i + 8 * g + t], b[i + 8 * g + t]);
        }
        for (; i + 8 <= n; i += 8) {
            if (g == 0) {
#pragma unroll
                for (int t = 0; t < 8; ++t) exact_step<IP>(acc[t], a[i + t], b[i + t]);
            }
        }
    }
    float v[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const float s = acc[l] + quad_xor1(acc[l]);
        v[l] = s + quad_xor2(s);
    }
    float res = ((v[0] + v[4]) + (v[1] + v[5])) + ((v[2] + v[6]) + (v[3] + v[7]));
    const int r = n - i;
    int j = 0;
    if (r >= 4) {
        for (; j < 4; ++j) {
            const float x = a[i + j], y = b[i + j];
            float p;
            if (IP) {
                p = x * y;
            } else {
                const float d = x - y;
                p = d * d;
            }
            res = res + p;
        }
    }
    for (; j < r; ++j) {
        const float x = a[i + j], y = b[i + j];
        if (IP) {
            res = fmaf(x, y, res);
        } else {
            const float d = x - y;
            res = fmaf(d, d, res);
        }
    }
    return res;
}

__device__ __forceinline__ float exact_score_quad(bool ip, const float* a, const float* b, int n, int g) {
    return ip ? -exact_avx2_quad<true>(a, b, n, g) : exact_avx2_quad<false>(a, b, n, g);
}

// SQ::score (src/sq.cpp:126-211) as the AVX2 build computes it: each value
// decoded as fma(code, step, vmin), 8 lane accumulators fed by fma (the
// difference squared, or q times the value), folded as ((l0+l4)+(l2+l6)) +
// ((l1+l5)+(l3+l7)), then the scalar tail, fused. 4-bit codes take 16
// dimensions per step of the vector loop.
template <bool IP>
__device__ __forceinline__ float sq_score(const float* q, const uint8_t* code, const float* vmin, const float* step,
                                          int n, int bits, int cstride = 16) {
    // 8-bit code of dimension d at code[(d / 16) * cstride + d % 16] (sq_row)
    auto code_at = [&](int d) -> float {
        return bits == 8 ? static_cast<float>(code[(d >> 4) * cstride + (d & 15)])
                         : static_cast<float>((code[d >> 1] >> ((d & 1) * 4)) & 0x0F);
    };
    float acc[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) acc[l] = 0.0f;
    int d = 0;
    auto step8 = [&](int at) {
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const float v = fmaf(code_at(at + l), step[at + l], vmin[at + l]);
            if (IP) {
                acc[l] = fmaf(q[at + l], v, acc[l]);
            } else {
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
    float h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = acc[j] + acc[j + 4];
    float sum = (h[0] + h[2]) + (h[1] + h[3]);
    for (; d < n; ++d) {
        const float v = fmaf(code_at(d), step[d], vmin[d]);
        if (IP) {
            sum = fmaf(q[d], v, sum);
        } else {
            const float diff = q[d] - v;
            sum = fmaf(diff, diff, sum);
        }
    }
    return sum;
}

// The SQ codes of list entry e and the stride between its 16-byte chunks.
// Rows of vbytes in entry order (stride 16: the row is contiguous), or, for
// 8-bit codes of a multiple of 16 dimensions (chunked), each group of 32
// entries stored chunk-major: chunk c of entry e at
// (e / 32) * 32 * dim + c * 512 + (e % 32) * 16, so one 16-byte load per
// lane of chunk c over consecutive entries reads contiguous bytes.
__device__ __forceinline__ const uint8_t* sq_row(const uint8_t* codes, uint32_t e, int vbytes, int chunked,
                                                 int& cstride) {
    cstride = chunked ? 512 : 16;
    return chunked ? codes + static_cast<size_t>(e >> 5) * 32 * vbytes + (e & 31) * 16
                   : codes + static_cast<size_t>(e) * vbytes;
}

// Filter bit of a slot (detail::SlotFilter::test, id_map.hpp:27).
__device__ __forceinline__ bool bit_test(const uint64_t* words, uint32_t s) {
    return (words[s >> 6] >> (s & 63)) & 1ull;
}

// ---------------------------------------------------------------------------
// Workgroup bitonic sort of n (power of two) 64-bit keys in LDS, ascending.
// ---------------------------------------------------------------------------
__device__ inline void block_bitonic_sort(uint64_t* keys, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
                const int i = 2 * t - (t & (stride - 1));
                const int j = i + stride;
                const bool up = (i & size) == 0;
                const uint64_t a = keys[i], b = keys[j];
                if ((a > b) == up) {
                    keys[i] = b;
                    keys[j] = a;
                }
            }
        }
    }
    __syncthreads();
}

} // namespace qd
