// Selection helpers shared by the search kernels: block-wide radix select,
// threshold search over a row of scores, and (distance, index) keys.
#pragma once

#include "device_common.hpp"

namespace qd {

// The digit d of a 256-bin histogram (one bin per thread of the first 256
// threads of an NT-thread block) whose running count first reaches `remain`,
// and the count below it: an inclusive scan of the bins within each of the
// first four waves, then across them.
template <int NT = 256>
__device__ __forceinline__ void block_find_digit(const uint32_t* hist, uint32_t remain, uint32_t* s_wsum, uint32_t* s_digit,
                                 uint32_t* s_below) {
    const int t = threadIdx.x, lane = lane_id(), w = t / kWave;
    const bool bins = NT == 256 || t < 256; // wave-uniform
    uint32_t h = 0, x = 0;
    if (bins) {
        h = hist[t];
        x = h;
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(x), off, kWave));
            if (lane >= off) x += y;
        }
        if (lane == kWave - 1) s_wsum[w] = x;
    }
    __syncthreads();
    if (bins) {
        uint32_t before = 0;
        for (int v = 0; v < w; ++v) before += s_wsum[v];
        const uint32_t incl = before + x, excl = incl - h;
        if (excl < remain && incl >= remain) {
            *s_digit = static_cast<uint32_t>(t);
            *s_below = excl;
        }
    }
    __syncthreads();
}

// Radix select over a row of n values by a 32- or 64-bit key (8 bits a pass):
// the key of rank `want` (1-based) among those key(l) gives. NT threads;
// with CACHE, the row's keys (n <= 16 * NT) stay in registers across passes.
template <typename Key, bool CACHE, typename KeyOf, int NT = 256>
__device__ __forceinline__ Key block_radix_select(int n, uint32_t want, KeyOf key, uint32_t* hist, uint32_t* s_wsum, uint32_t* s_digit,
                                  uint32_t* s_below) {
    Key cached[CACHE ? 16 : 1];
    if (CACHE) {
#pragma unroll
        for (int j = 0; j < (CACHE ? 16 : 1); ++j) {
            const int l = threadIdx.x + j * NT;
            cached[j] = l < n ? key(l) : ~Key(0);
        }
    }
    Key prefix = 0;
    uint32_t remain = want;
    constexpr int kBits = 8 * static_cast<int>(sizeof(Key));
    for (int shift = kBits - 8; shift >= 0; shift -= 8) {
        if (threadIdx.x < 256) hist[threadIdx.x] = 0;
        __syncthreads();
        const Key hi_mask = shift == kBits - 8 ? Key(0) : static_cast<Key>(~Key(0) << (shift + 8));
        if (CACHE) {
#pragma unroll
            for (int j = 0; j < (CACHE ? 16 : 1); ++j) {
                const int l = threadIdx.x + j * NT;
                if (l < n && (cached[j] & hi_mask) == prefix) atomicAdd(&hist[(cached[j] >> shift) & 255], 1u);
            }
        } else {
            for (int l = threadIdx.x; l < n; l += NT) {
                const Key k = key(l);
                if ((k & hi_mask) == prefix) atomicAdd(&hist[(k >> shift) & 255], 1u);
            }
        }
        __syncthreads();
        block_find_digit<NT>(hist, remain, s_wsum, s_digit, s_below);
        prefix |= static_cast<Key>(*s_digit) << shift;
        remain -= *s_below;
        __syncthreads();
    }
    return prefix;
}

// Selection keys of (distance, index): ckey(d) (float order, -0 tied with +0)
// above the index, with the index shifted up one bit so the lowest bit can
// record a -0 distance -- indexes are unique, so that bit never decides the
// order, and the key gives back the distance's exact bits.
__device__ __forceinline__ uint64_t dist_key(float d, uint32_t idx) {
    const uint32_t negzero = (d == 0.0f && (__float_as_uint(d) >> 31)) ? 1u : 0u;
    return (static_cast<uint64_t>(ckey(d)) << 32) | (static_cast<uint64_t>(idx) << 1) | negzero;
}
__device__ __forceinline__ uint32_t key_index(uint64_t key) { return static_cast<uint32_t>(key & 0xFFFFFFFFu) >> 1; }
__device__ __forceinline__ float key_dist(uint64_t key) {
    return (key & 1u) ? -0.0f : from_fkey(static_cast<uint32_t>(key >> 32));
}

// Minimum / maximum over an NT-thread block (every thread gets the result);
// s_red holds NT / 64 floats.
template <int NT = 256>
__device__ __forceinline__ float block_min256(float v, float* s_red) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, kWave));
    __syncthreads();
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = v;
    __syncthreads();
    float r = s_red[0];
    for (int w = 1; w < NT / kWave; ++w) r = fminf(r, s_red[w]);
    return r;
}

template <int NT = 256>
__device__ __forceinline__ float block_max256(float v, float* s_red) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    __syncthreads();
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = v;
    __syncthreads();
    float r = s_red[0];
    for (int w = 1; w < NT / kWave; ++w) r = fmaxf(r, s_red[w]);
    return r;
}

// A value tau of a row of scores with at least M values <= tau and, as a
// rule, few more than M: values are bucketed linearly over [min, max] (256
// buckets, the bucket a monotone function of the value), the bucket where the
// running count reaches M is split the same way once more when it holds many,
// and tau is the largest value in the buckets up to the one found. The row is
// given by each_all(fn), which calls fn on every value this thread holds (the
// block's threads together holding the row once). NT threads; s_red holds
// 2 * NT / 64 floats. With FINITE, +inf and NaN values (rows that may not be
// returned) take no part, and the result is +inf when fewer than M values
// remain.
template <bool FINITE, int NT, typename EachAll>
__device__ __forceinline__ float select_threshold_each(EachAll&& each_all, int M, int target, uint32_t* hist,
                                                       uint32_t* s_wsum, uint32_t* s_digit, uint32_t* s_below,
                                                       float* s_red) {
    const int tid = threadIdx.x;
    auto each = [&](auto&& fn) {
        each_all([&](float v) {
            if (!FINITE || v <= 3.40282347e38f) fn(v);
        });
    };
    float lo = INFINITY, hi = -INFINITY;
    uint32_t cnt = 0;
    each([&](float v) {
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
        ++cnt;
    });
    lo = block_min256<NT>(lo, s_red);
    hi = block_max256<NT>(hi, s_red + NT / kWave);
    if (FINITE) {
        if (tid < 256) hist[tid] = 0;
        __syncthreads();
        if (cnt) atomicAdd(&hist[0], cnt);
        __syncthreads();
        const uint32_t total = hist[0];
        __syncthreads();
        if (total < static_cast<uint32_t>(M)) return INFINITY;
    }
    if (!(hi > lo) || !isfinite(hi - lo)) return hi;
    const float s1 = 256.0f / (hi - lo);
    auto b1 = [&](float v) { return min(255, static_cast<int>((v - lo) * s1)); };
    if (tid < 256) hist[tid] = 0;
    __syncthreads();
    each([&](float v) { atomicAdd(&hist[b1(v)], 1u); });
    __syncthreads();
    block_find_digit<NT>(hist, static_cast<uint32_t>(M), s_wsum, s_digit, s_below);
    const int d1 = static_cast<int>(*s_digit);
    const uint32_t below1 = *s_below, c1 = hist[d1];
    __syncthreads();
    if (below1 + c1 <= static_cast<uint32_t>(target)) {
        float m = -INFINITY;
        each([&](float v) {
            if (b1(v) <= d1) m = fmaxf(m, v);
        });
        return block_max256<NT>(m, s_red);
    }
    float lo2 = INFINITY, hi2 = -INFINITY;
    each([&](float v) {
        if (b1(v) == d1) {
            lo2 = fminf(lo2, v);
            hi2 = fmaxf(hi2, v);
        }
    });
    lo2 = block_min256<NT>(lo2, s_red);
    hi2 = block_max256<NT>(hi2, s_red + NT / kWave);
    if (!(hi2 > lo2)) {
        float m = -INFINITY;
        each([&](float v) {
            if (b1(v) <= d1) m = fmaxf(m, v);
        });
        return block_max256<NT>(m, s_red);
    }
    const float s2 = 256.0f / (hi2 - lo2);
    auto b2 = [&](float v) { return min(255, max(0, static_cast<int>((v - lo2) * s2))); };
    if (tid < 256) hist[tid] = 0;
    __syncthreads();
    each([&](float v) {
        if (b1(v) == d1) atomicAdd(&hist[b2(v)], 1u);
    });
    __syncthreads();
    block_find_digit<NT>(hist, static_cast<uint32_t>(M) - below1, s_wsum, s_digit, s_below);
    const int d2 = static_cast<int>(*s_digit);
    __syncthreads();
    float m = -INFINITY;
    each([&](float v) {
        const int b = b1(v);
        if (b < d1 || (b == d1 && b2(v) <= d2)) m = fmaxf(m, v);
    });
    return block_max256<NT>(m, s_red);
}

// select_threshold_each over row[0 .. n): with CACHE the row (n <= 16 * NT)
// is read once into registers, else it is read from memory at every pass.
template <bool CACHE, bool FINITE = false, int NT = 256>
__device__ __forceinline__ float select_threshold(const float* row, int n, int M, int target, uint32_t* hist, uint32_t* s_wsum,
                                  uint32_t* s_digit, uint32_t* s_below, float* s_red) {
    const int tid = threadIdx.x;
    float cached[CACHE ? 16 : 1];
    if (CACHE) {
#pragma unroll
        for (int j = 0; j < (CACHE ? 16 : 1); ++j) {
            const int l = tid + NT * j;
            cached[j] = l < n ? row[l] : 0.0f;
        }
    }
    auto each_all = [&](auto&& fn) {
        if (CACHE) {
#pragma unroll
            for (int j = 0; j < (CACHE ? 16 : 1); ++j)
                if (tid + NT * j < n) fn(cached[j]);
        } else {
            for (int l = tid; l < n; l += NT) fn(row[l]);
        }
    };
    return select_threshold_each<FINITE, NT>(each_all, M, target, hist, s_wsum, s_digit, s_below, s_red);
}

} // namespace qd
