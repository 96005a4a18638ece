// BF16 MFMA scoring with a proven error bound, for exact FlatIndex search
// (FlatIndex::query_impl, src/flat_index.cpp:33-64) on gfx950.
//
//   k_rows_bf16 / k_queries_bf16   FP32 -> BF16 (round to nearest even) and the
//                                  norms the error bound needs
//   k_gemm_bf16<..., Epi>          C = A * B^T by v_mfma_f32_32x32x16_bf16,
//                                  LDS double-buffered, XCD-aware tile order;
//                                  the epilogue writes scores or keeps candidates
//   k_gemm_bf16_lds<..., Epi>      the same product, operands loaded directly
//                                  into a ring of LDS stages
//   k_flat_limit                   per query, an upper bound on its k-th smallest
//                                  approximate score from a strided row sample
//   k_flat_finish                  per query, the exact k nearest among the
//                                  candidates, or a hand-over to the FP32 path
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>

#include "device_common.hpp"
#include "gemm_bf16.hpp"
#include "select.hpp"

namespace qd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

namespace {

// Round to nearest even; infinities and NaN keep their class.
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// sqrt of a float sum of n squares, rounded up past the sum's own rounding
// (relative error at most n u, u = 2^-24; n <= 2^13 keeps that under 2^-11).
__device__ __forceinline__ float norm_up(float sumsq) { return sqrtf(sumsq) * (1.0f + 0x1p-10f); }

// One wave per row: BF16 copy, zero-padded to kp, and the row's norms.
__global__ void __launch_bounds__(256) k_rows_bf16(const float* rows, uint64_t n, int dim, int kp, uint16_t* out,
                                                   float* maxes) {
    const uint64_t r = blockIdx.x * 4ull + threadIdx.x / kWave;
    const int lane = lane_id();
    if (r >= n) return;
    const float* x = rows + r * dim;
    uint16_t* o = out + r * kp;
    float res = 0.0f, nt = 0.0f;
    for (int e = lane; e < kp; e += kWave) {
        uint16_t b = 0;
        if (e < dim) {
            b = to_bf16(x[e]);
            const float t = from_bf16(b), d = x[e] - t;
            res += d * d;
            nt += t * t;
        }
        o[e] = b;
    }
    res = wave_sum(res);
    nt = wave_sum(nt);
    if (lane == 0) {
        // non-negative floats order as their bit patterns
        atomicMax(reinterpret_cast<unsigned int*>(maxes), __float_as_uint(norm_up(res)));
        atomicMax(reinterpret_cast<unsigned int*>(maxes + 1), __float_as_uint(norm_up(nt)));
    }
}

__global__ void __launch_bounds__(256) k_queries_bf16(const float* q, int nq, int dim, int kp, uint16_t* out,
                                                      float* qinfo) {
    const int i = blockIdx.x * 4 + threadIdx.x / kWave;
    const int lane = lane_id();
    if (i >= nq) return;
    const float* x = q + static_cast<size_t>(i) * dim;
    uint16_t* o = out + static_cast<size_t>(i) * kp;
    float s = 0.0f, res = 0.0f, nt = 0.0f;
    for (int e = lane; e < kp; e += kWave) {
        uint16_t b = 0;
        if (e < dim) {
            b = to_bf16(x[e]);
            const float t = from_bf16(b), d = x[e] - t;
            s += x[e] * x[e];
            res += d * d;
            nt += t * t;
        }
        o[e] = b;
    }
    s = wave_sum(s);
    res = wave_sum(res);
    nt = wave_sum(nt);
    if (lane == 0) {
        qinfo[3 * i] = norm_up(s);
        qinfo[3 * i + 1] = norm_up(nt);
        qinfo[3 * i + 2] = norm_up(res);
    }
}

// Each row as three BF16 segments of kp (zero-padded) for a split product:
// x = hi + lo + r with hi = bf16(x), lo = bf16(x - hi), |r| <= 2^-16 |x|.
// order 0 writes [hi | hi | lo], order 1 [hi | lo | hi], so that one GEMM over
// K = 3 kp of an order-0 row against an order-1 row sums hi.hi + hi.lo + lo.hi.
__global__ void __launch_bounds__(256) k_split_bf16(const float* x, uint64_t n, int dim, int kp, int order,
                                                    uint16_t* out) {
    const uint64_t r = blockIdx.x * 4ull + threadIdx.x / kWave;
    if (r >= n) return;
    const float* xr = x + r * dim;
    uint16_t* o = out + r * 3 * static_cast<uint64_t>(kp);
    for (int e = lane_id(); e < kp; e += kWave) {
        uint16_t hi = 0, lo = 0;
        if (e < dim) {
            hi = to_bf16(xr[e]);
            lo = to_bf16(xr[e] - from_bf16(hi));
        }
        o[e] = hi;
        o[kp + e] = order == 0 ? hi : lo;
        o[2 * kp + e] = order == 0 ? lo : hi;
    }
}

// A bound on |approximate - exact + const| for one query, const = ||q||^2 for
// L2 (the approximate score leaves it out) and 0 for inner product. With
// q~, x~ the BF16 values and s~ the FP32-accumulated q~ . x~:
//   |s~ - q.x| <= |q~||x - x~| + |q - q~||x| + (n + 2) u |q~||x~|
// inner product: approximate = -s~; the exact score's own rounding adds
//   (n + 3) u |q||x|;
// L2: approximate = ||x||^2 (rounded, (n + 3) u |x|^2) - 2 s~, the subtraction
// rounds by u times its size, and the exact sum of squared differences is
// within (n + 4) u (|q| + |x|)^2 of ||q - x||^2. Each norm is the largest
// bound over the rows; the total is raised by 1/16, and by n 2^-100 for
// products flushed to zero.
__device__ __forceinline__ float flat_err(int ip, int dim, const float* qinfo, const float* maxes, float xmax) {
    const float u = 5.9604645e-8f, n = static_cast<float>(dim);
    const float qn = qinfo[0], qtn = qinfo[1], rq = qinfo[2];
    const float rx = maxes[0], xt = maxes[1];
    const float dot = qtn * rx + rq * xmax + (n + 2.0f) * u * qtn * xt;
    float e;
    if (ip) {
        e = dot + (n + 3.0f) * u * qn * xmax;
    } else {
        const float s = qn + xmax;
        e = (n + 3.0f) * u * xmax * xmax + 2.0f * dot + u * (xmax * xmax + 2.0f * qtn * xt) + (n + 4.0f) * u * s * s;
    }
    return e * 1.0625f + n * 7.9e-31f;
}

// Every approximate score within 2 e of tau: a limit that float rounding
// cannot pull below tau + 2 e.
__device__ __forceinline__ float within(float tau, float e) {
    return tau + (2.0f * e + fabsf(tau) * 0x1p-22f);
}

constexpr int kLdsRow = kBf16K + 8; // BF16 per LDS row: 64 bytes of data, 16 of padding

// C = A * B^T over K = kp (both row-major, BF16, rows zero-padded), 4 waves
// of (32 TM) x (32 TN), a workgroup tile of (64 TM) x (64 TN). Lane t gives
// A[t % 32][8 (t / 32) .. +7] and B[t % 32][...] of each 32 x 16 step and
// holds C[8 (v / 4) + 4 (t / 32) + v % 4][t % 32] in accumulator v. Tiles are
// ordered along the big operand, and a workgroup's tile is picked so that the
// workgroups of one XCD (every 8th) take consecutive tiles: the big operand's
// tile is read into that XCD's L2 once and reused across the small operand.
template <int TM, int TN, bool A_BIG, class Epi>
__global__ void __launch_bounds__(256, 2) k_gemm_bf16(const uint16_t* __restrict__ A, int M, size_t lda,
                                                   const uint16_t* __restrict__ B, int N, size_t ldb, int kp,
                                                   int per_xcd, Epi epi) {
    constexpr int BM = 64 * TM, BN = 64 * TN;
    __shared__ __align__(16) uint16_t As[2][BM * kLdsRow];
    __shared__ __align__(16) uint16_t Bs[2][BN * kLdsRow];
    const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    const int slot = blockIdx.x / 8, xcd = blockIdx.x % 8;
    if (slot >= per_xcd) return;
    const int t = xcd * per_xcd + slot;
    if (t >= tiles_m * tiles_n) return;
    const int tm = A_BIG ? t / tiles_n : t % tiles_m;
    const int tn = A_BIG ? t % tiles_n : t / tiles_m;
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const int wm = wave >> 1, wn = wave & 1;

    constexpr int CA = BM * 4 / 256, CB = BN * 4 / 256; // 16-byte chunks per thread per stage
    u32x4 ra[CA], rb[CB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int c = 0; c < CA; ++c) {
            const int e = tid + 256 * c, r = e >> 2, ch = e & 3;
            const int g = m0 + r;
            ra[c] = g < M ? *reinterpret_cast<const u32x4*>(A + static_cast<size_t>(g) * lda + k0 + 8 * ch)
                          : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const int e = tid + 256 * c, r = e >> 2, ch = e & 3;
            const int g = n0 + r;
            rb[c] = g < N ? *reinterpret_cast<const u32x4*>(B + static_cast<size_t>(g) * ldb + k0 + 8 * ch)
                          : u32x4{0u, 0u, 0u, 0u};
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int c = 0; c < CA; ++c) {
            const int e = tid + 256 * c, r = e >> 2, ch = e & 3;
            *reinterpret_cast<u32x4*>(&As[buf][r * kLdsRow + 8 * ch]) = ra[c];
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const int e = tid + 256 * c, r = e >> 2, ch = e & 3;
            *reinterpret_cast<u32x4*>(&Bs[buf][r * kLdsRow + 8 * ch]) = rb[c];
        }
    };

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    const int nk = kp / kBf16K;
    gload(0);
    lstore(0);
    __syncthreads();
    const int arow = wm * 32 * TM + (lane & 31), brow = wn * 32 * TN + (lane & 31), kofs = 8 * (lane >> 5);
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * kBf16K);
#pragma unroll
        for (int ks = 0; ks < kBf16K / 16; ++ks) {
            bf16x8 a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i)
                a[i] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const u32x4*>(&As[buf][(arow + 32 * i) * kLdsRow + 16 * ks + kofs]));
#pragma unroll
            for (int j = 0; j < TN; ++j)
                b[j] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const u32x4*>(&Bs[buf][(brow + 32 * j) * kLdsRow + 16 * ks + kofs]));
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) lstore(buf ^ 1);
        __syncthreads();
    }
    if constexpr (Epi::kTile) {
        // a whole-tile epilogue, with the (now idle) LDS for its reductions
        epi.template tile<TM, TN>(acc, m0, n0, wm, wn, lane, M, N, reinterpret_cast<float*>(&As[0][0]));
    } else {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int col = n0 + wn * 32 * TN + 32 * j + (lane & 31);
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int row = m0 + wm * 32 * TM + 32 * i + 8 * (v / 4) + 4 * (lane >> 5) + (v & 3);
                    if (row < M && col < N) epi(row, col, acc[i][j][v]);
                }
            }
    }
}

// The same product with operands loaded straight into LDS: workgroup tiles of
// 256 x 64 TN, four waves of 128 x 32 TN, fed through a ring of NST stages of
// KS in LDS that direct global-to-LDS loads (global_load_lds_dwordx4) fill
// NST - 1 stages ahead of the MFMAs, so no operand passes through registers on
// its way in. TN = 4 takes a whole SIMD per wave (its 256 accumulators) and
// the ring 128 KB of gfx950's 160 KB; TN = 2 runs two workgroups per CU, a
// ring of 3 x 24 KB each (the shape the autotuner keeps for the 1M x 4,096 x
// 768 filter GEMM). One barrier per stage. A stage holds 256 rows of A and 64
// TN of B, 2 KS bytes each; a wave's load instruction fills 1 KB of
// consecutive rows (lane l the 16-byte slot l % (KS / 8) of row l / (KS / 8)),
// and the 16-byte chunk of K a slot holds is its slot number XOR a function of
// the row, so that the 32 lanes reading one chunk of 32 consecutive rows
// (ds_read_b128) hit distinct banks. Rows past M (N) load the last row again;
// the epilogue drops them.

typedef __attribute__((address_space(3))) unsigned char lds_u8;

// One 16-byte global-to-LDS load per lane (global_load_lds_dwordx4): lane l's
// 16 bytes land at lds + 16 l. A plain (non-template) function: the builtin
// written straight into the templated kernel keeps the host pass from
// emitting the kernel's launch stub.
__device__ __forceinline__ void lds_load16(const uint16_t* g, lds_u8* lds) {
    __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
}

// Waits until at most N of this wave's vector-memory loads are outstanding.
// The LDS buffer's address goes into the asm, which clobbers memory: the
// compiler must assume the buffer written (it does not model the direct
// loads as LDS writes).
template <int N>
__device__ __forceinline__ void wait_loads(const lds_u8* lds) {
    asm volatile("s_waitcnt vmcnt(%0)" : : "n"(N), "v"(static_cast<uint32_t>(reinterpret_cast<uintptr_t>(lds)))
                 : "memory");
}

// Tile order across XCDs (each has its own L2): the small operand's tiles
// fall into `groups` groups (1, 2, 4 or 8) small enough to stay in one L2,
// and each XCD takes one group and a 1 / (8 / groups) share of the big
// operand's tiles, all of its group's tiles for each of them. Every small
// tile is then read into an XCD's L2 once, and every big tile by `groups`
// XCDs. slot -> (big tile, small tile) within the XCD's share.
struct XcdTiles {
    int groups, big_per, small_per;
};

// NW waves per workgroup (4 or 8) in a 2 x NW/2 grid of (32 TM) x (32 TN) wave tiles.
template <int KS, int TN, int NST, bool A_BIG, class Epi, int NW>
__global__ void __launch_bounds__(64 * NW, (TN == 4 || NW == 8) ? 1 : 2)
    k_gemm_bf16_lds(const uint16_t* __restrict__ A, int M, size_t lda, const uint16_t* __restrict__ B, int N,
                    size_t ldb, int kp, XcdTiles xt, Epi epi) {
    constexpr int TM = 4, WN = NW / 2, BM = 256, BN = WN * 32 * TN;
    constexpr int RB = 2 * KS;                     // bytes of one row in a stage
    constexpr int CPR = RB / 16;                   // 16-byte chunks per row
    constexpr int RPL = 64 / CPR;                  // rows one wave load fills
    constexpr int LPWA = BM / RPL / NW;            // loads of A per wave per stage
    constexpr int LPWB = BN / RPL / NW;            // loads of B per wave per stage
    constexpr int STAGE = (BM + BN) * RB;
    constexpr int LPS = LPWA + LPWB;               // loads per thread per stage
    constexpr int SWZ = 16 / CPR;                  // rows sharing one swizzle value
    static_assert(NST >= 2 && NST <= 4 && NST * STAGE <= 160 * 1024 && LPWB >= 1, "ring within gfx950's LDS");
    __shared__ __align__(16) unsigned char smem[NST * STAGE];
    lds_u8* const sbase = (lds_u8*)smem; // the ring as an LDS address (address space 3)
    const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    const int slot = blockIdx.x / 8, xcd = blockIdx.x % 8;
    if (slot >= xt.big_per * xt.small_per) return;
    const int group = xcd % xt.groups, part = xcd / xt.groups;
    const int big = part * xt.big_per + slot / xt.small_per, small = group * xt.small_per + slot % xt.small_per;
    const int tm = A_BIG ? big : small, tn = A_BIG ? small : big;
    if (tm >= tiles_m || tn >= tiles_n) return;
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const int wm = wave / WN, wn = wave % WN;

    // this lane's source rows: load c of the wave fills rows (NW c + wave) RPL .. + RPL - 1
    const uint16_t* ga[LPWA];
    const uint16_t* gb[LPWB];
#pragma unroll
    for (int c = 0; c < LPWA; ++c) {
        const int r = (NW * c + wave) * RPL + lane / CPR;
        ga[c] = A + static_cast<size_t>(min(m0 + r, M - 1)) * lda + 8 * ((lane % CPR) ^ ((r / SWZ) % CPR));
    }
#pragma unroll
    for (int c = 0; c < LPWB; ++c) {
        const int r = (NW * c + wave) * RPL + lane / CPR;
        gb[c] = B + static_cast<size_t>(min(n0 + r, N - 1)) * ldb + 8 * ((lane % CPR) ^ ((r / SWZ) % CPR));
    }
    auto issue = [&](int buf, int k0) {
        lds_u8* st = sbase + buf * STAGE;
#pragma unroll
        for (int c = 0; c < LPWA; ++c)
            lds_load16(ga[c] + k0, st + (NW * c + wave) * RPL * RB);
#pragma unroll
        for (int c = 0; c < LPWB; ++c)
            lds_load16(gb[c] + k0, st + BM * RB + (NW * c + wave) * RPL * RB);
    };

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    const int nk = kp / KS;
#pragma unroll
    for (int s = 0; s < NST - 1; ++s)
        if (s < nk) issue(s, s * KS);
    const int arow = wm * 32 * TM + (lane & 31), brow = wn * 32 * TN + (lane & 31), half = lane >> 5;
    // the swizzle of every fragment row a lane reads (rows 32 apart share it)
    const int aswz = (arow / SWZ) % CPR, bswz = (brow / SWZ) % CPR;
    for (int kt = 0; kt < nk; ++kt) {
        // stage kt landed; the later stages already issued may stay in flight
        const int ahead = min(NST - 2, nk - 1 - kt);
        if (ahead >= 2)
            wait_loads<2 * LPS>(sbase);
        else if (ahead == 1)
            wait_loads<LPS>(sbase);
        else
            wait_loads<0>(sbase);
        // every wave's loads of stage kt are in, and stage kt - 1 is read by all
        // (its LDS reads fed the MFMAs before the wait): a bare barrier, since
        // __syncthreads() would also wait for the later stages' loads; the
        // asm fences keep LDS accesses on their side of it
        __builtin_amdgcn_s_barrier();
        asm volatile("" : : : "memory");
        if (kt + NST - 1 < nk) issue((kt + NST - 1) % NST, (kt + NST - 1) * KS);
        const unsigned char* st = smem + (kt % NST) * STAGE;
        // the fragments of one 16-wide K slice; the next slice's are read
        // while this one's 16 MFMAs run
        bf16x8 fa[2][TM], fb[2][TN];
        auto read = [&](int ks, bf16x8 (&a)[TM], bf16x8 (&b)[TN]) {
            const int ch = 2 * ks + half;
#pragma unroll
            for (int i = 0; i < TM; ++i)
                a[i] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const u32x4*>(st + (arow + 32 * i) * RB + 16 * (ch ^ aswz)));
#pragma unroll
            for (int j = 0; j < TN; ++j)
                b[j] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const u32x4*>(st + BM * RB + (brow + 32 * j) * RB + 16 * (ch ^ bswz)));
        };
        read(0, fa[0], fb[0]);
#pragma unroll
        for (int ks = 0; ks < KS / 16; ++ks) {
            if (ks + 1 < KS / 16) read(ks + 1, fa[(ks + 1) & 1], fb[(ks + 1) & 1]);
            // keep the reads ahead of the MFMAs (the scheduler would otherwise
            // reuse one fragment's registers and wait on each read in turn)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[ks & 1][i], fb[ks & 1][j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads(); // the ring is free for the epilogue's reductions
    if constexpr (Epi::kTile) {
        epi.template tile<TM, TN>(acc, m0, n0, wm, wn, lane, M, N, reinterpret_cast<float*>(smem));
    } else {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int col = n0 + wn * 32 * TN + 32 * j + (lane & 31);
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int row = m0 + wm * 32 * TM + 32 * i + 8 * (v / 4) + 4 * (lane >> 5) + (v & 3);
                    if (row < M && col < N) epi(row, col, acc[i][j][v]);
                }
            }
    }
}

// Scores of queries (A rows) against every stride-th row (B rows, B's row j
// is row j * stride): ||x||^2 - 2 q.x, or -q.x, and +inf for rows the query
// may not return.
struct EpiScores {
    static constexpr bool kTile = false;
    float* D;
    size_t ldd;
    int ip, stride;
    const float* xsq;
    const uint64_t* excl;
    const uint64_t* incl;
    __device__ __forceinline__ void operator()(int i, int j, float s) const {
        const uint32_t r = static_cast<uint32_t>(j) * static_cast<uint32_t>(stride);
        const bool out = (excl && bit_test(excl, r)) || (incl && !bit_test(incl, r));
        D[static_cast<size_t>(i) * ldd + j] = out ? INFINITY : ip ? -s : xsq[r] - 2.0f * s;
    }
};

// Rows (A rows) against queries (B rows): a row whose score is at most the
// query's limit, and which the query may return, joins its candidates. Each
// lane holds TN query columns: their limits are read once per tile.
struct EpiFilter {
    static constexpr bool kTile = true;
    const float* lim;
    int ip, cap;
    const float* xsq;
    const uint64_t* excl;
    const uint64_t* incl;
    uint32_t* cnt;
    float* cd;
    uint32_t* ci;
    template <int TM, int TN>
    __device__ __forceinline__ void tile(const f32x16 (&acc)[TM][TN], int m0, int n0, int wm, int wn, int lane, int M,
                                         int N, float*) const {
        float lj[TN];
        int cj[TN];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            cj[j] = n0 + wn * 32 * TN + 32 * j + (lane & 31);
            lj[j] = cj[j] < N ? lim[cj[j]] : -INFINITY;
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = m0 + wm * 32 * TM + 32 * i + 8 * (v / 4) + 4 * (lane >> 5) + (v & 3);
                if (row >= M) continue;
                const float x2 = ip ? 0.0f : xsq[row];
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const float s = acc[i][j][v];
                    const float a = ip ? -s : x2 - 2.0f * s;
                    if (!(a <= lj[j])) continue;
                    const uint32_t r = static_cast<uint32_t>(row);
                    if ((excl && bit_test(excl, r)) || (incl && !bit_test(incl, r))) continue;
                    const uint32_t pos = atomicAdd(&cnt[cj[j]], 1u);
                    if (pos < static_cast<uint32_t>(cap)) {
                        cd[static_cast<size_t>(cj[j]) * cap + pos] = a;
                        ci[static_cast<size_t>(cj[j]) * cap + pos] = r;
                    }
                }
            }
    }
};

// Every stride-th row (A rows; A's row i is row i * stride) against queries
// (B rows): per query, the smallest approximate score over each tile of 128
// sampled rows that it may return (+inf if none) into out[query * ntile + tile].
// The want-th smallest of those minima is an upper bound on the want-th
// smallest score of the sample, and so of every row: the minima belong to
// distinct rows.
struct EpiTileMin {
    static constexpr bool kTile = true;
    float* out;
    int ntile, ip, stride;
    const float* xsq;
    const uint64_t* excl;
    const uint64_t* incl;
    template <int TM, int TN>
    __device__ __forceinline__ void tile(const f32x16 (&acc)[TM][TN], int m0, int n0, int wm, int wn, int lane, int M,
                                         int N, float* red) const {
        float mn[TN];
#pragma unroll
        for (int j = 0; j < TN; ++j) mn[j] = INFINITY;
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = m0 + wm * 32 * TM + 32 * i + 8 * (v / 4) + 4 * (lane >> 5) + (v & 3);
                if (row >= M) continue;
                const uint32_t r = static_cast<uint32_t>(row) * static_cast<uint32_t>(stride);
                if ((excl && bit_test(excl, r)) || (incl && !bit_test(incl, r))) continue;
                const float x2 = ip ? 0.0f : xsq[r];
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const float s = acc[i][j][v];
                    mn[j] = fminf(mn[j], ip ? -s : x2 - 2.0f * s);
                }
            }
        // lanes t and t + 32 hold the two row halves of one query column
#pragma unroll
        for (int j = 0; j < TN; ++j) mn[j] = fminf(mn[j], __shfl_xor(mn[j], 32, kWave));
        if (wm == 1 && lane < 32)
#pragma unroll
            for (int j = 0; j < TN; ++j) red[(wn * TN + j) * 32 + lane] = mn[j];
        __syncthreads();
        if (wm == 0 && lane < 32) {
            const int t = m0 / (64 * TM);
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int col = n0 + wn * 32 * TN + 32 * j + lane;
                if (col < N) out[static_cast<size_t>(col) * ntile + t] = fminf(mn[j], red[(wn * TN + j) * 32 + lane]);
            }
        }
    }
};

// Per query: tau, an upper bound on the want-th smallest approximate score of
// the rows it may return (the want-th smallest of its tile minima), and the
// candidate limit tau + 2 e.
__global__ void __launch_bounds__(256) k_flat_limit(const float* tmin, int ntile, int want, const float* qinfo,
                                                    const float* maxes, float xmax, int ip, int dim, float* lim,
                                                    uint32_t* cnt) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_wsum[4], s_digit, s_below;
    __shared__ float s_red8[8];
    const int i = blockIdx.x;
    const float* row = tmin + static_cast<size_t>(i) * ntile;
    const int target = want + max(16, want / 8);
    const float tau = ntile <= 16 * 256
                          ? select_threshold<true, true>(row, ntile, want, target, hist, s_wsum, &s_digit, &s_below, s_red8)
                          : select_threshold<false, true>(row, ntile, want, target, hist, s_wsum, &s_digit, &s_below, s_red8);
    if (threadIdx.x == 0) {
        const float e = flat_err(ip, dim, qinfo + 3 * i, maxes, xmax);
        lim[i] = tau <= 3.40282347e38f ? within(tau, e) : INFINITY;
        cnt[i] = 0;
    }
}

// Per query: the want-th smallest approximate score tau among the candidates
// (they hold every row under the limit, so it is the query's own), the rows
// within tau + 2 e rescored exactly in the reference's order, sorted by
// (distance, slot), the first k written. A query whose candidates overflowed,
// or fell short of want, goes to ovf instead.
__global__ void __launch_bounds__(256) k_flat_finish(const float* q, int dim, int ip, const float* rows,
                                                     const float* qinfo, const float* maxes, float xmax, int k,
                                                     int want, int cap, const uint32_t* cnt, const float* cd,
                                                     const uint32_t* ci, const int64_t* labels, int64_t* out_ids,
                                                     float* out_dists, uint32_t* ovf, uint32_t* ovf_cnt, int ovf_base) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_wsum[4], s_digit, s_below, s_count;
    __shared__ float s_red8[8];
    const int i = blockIdx.x, tid = threadIdx.x;
    const uint32_t c = cnt[i];
    if (c > static_cast<uint32_t>(cap) || c < static_cast<uint32_t>(want)) {
        if (tid == 0) ovf[atomicAdd(ovf_cnt, 1u)] = static_cast<uint32_t>(ovf_base + i);
        return;
    }
    float* qsl = reinterpret_cast<float*>(lds);
    uint64_t* keys = reinterpret_cast<uint64_t*>(lds + query_lds_bytes(dim));
    const bool qcopy = dim <= kQueryLdsDims;
    if (qcopy)
        for (int t = tid; t < dim; t += blockDim.x) qsl[t] = q[static_cast<size_t>(i) * dim + t];
    const float* qsh = qcopy ? qsl : q + static_cast<size_t>(i) * dim;
    if (tid == 0) s_count = 0;
    __syncthreads();
    const float* crow = cd + static_cast<size_t>(i) * cap;
    const uint32_t* irow = ci + static_cast<size_t>(i) * cap;
    const int n = static_cast<int>(c);
    float limit = INFINITY;
    if (want > 0) {
        const int target = want + max(64, want / 4);
        const float tau =
            select_threshold<true, true>(crow, n, want, target, hist, s_wsum, &s_digit, &s_below, s_red8);
        limit = within(tau, flat_err(ip, dim, qinfo + 3 * i, maxes, xmax));
    }
    for (int t = tid; t < n; t += blockDim.x)
        if (crow[t] <= limit) keys[atomicAdd(&s_count, 1u)] = irow[t];
    __syncthreads();
    const int m = static_cast<int>(s_count);
    int np = 1;
    while (np < m) np <<= 1;
    for (int t = tid; t < np; t += blockDim.x) {
        if (t < m) {
            const uint32_t l = static_cast<uint32_t>(keys[t]);
            keys[t] = dist_key(exact_score(ip != 0, qsh, rows + static_cast<size_t>(l) * dim, dim), l);
        } else {
            keys[t] = ~0ull;
        }
    }
    block_bitonic_sort(keys, np);
    for (int t = tid; t < k; t += blockDim.x) {
        int64_t id = -1;
        float d = INFINITY;
        if (t < m) {
            const uint32_t l = key_index(keys[t]);
            id = labels ? labels[l] : static_cast<int64_t>(l);
            d = key_dist(keys[t]);
        }
        out_ids[static_cast<size_t>(i) * k + t] = id;
        out_dists[static_cast<size_t>(i) * k + t] = d;
    }
}

inline unsigned cdiv(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

template <int TM, int TN, bool A_BIG, class Epi>
void gemm(const uint16_t* A, int M, size_t lda, const uint16_t* B, int N, size_t ldb, int kp, const Epi& epi,
          hipStream_t s) {
    if (M <= 0 || N <= 0) return;
    const size_t tiles = static_cast<size_t>(cdiv(M, 64 * TM)) * cdiv(N, 64 * TN);
    const int per_xcd = static_cast<int>((tiles + 7) / 8);
    hipLaunchKernelGGL((k_gemm_bf16<TM, TN, A_BIG, Epi>), dim3(static_cast<unsigned>(8 * per_xcd)), dim3(256), 0, s,
                       A, M, lda, B, N, ldb, kp, per_xcd, epi);
}

template <int KS, int TN, int NST, bool A_BIG, int NW = 4, class Epi>
void gemm_lds(const uint16_t* A, int M, size_t lda, const uint16_t* B, int N, size_t ldb, int kp, const Epi& epi,
              hipStream_t s) {
    if (M <= 0 || N <= 0) return;
    const int BN = NW / 2 * 32 * TN;
    const int big_tiles = static_cast<int>(A_BIG ? cdiv(M, 256) : cdiv(N, BN));
    const int small_tiles = static_cast<int>(A_BIG ? cdiv(N, BN) : cdiv(M, 256));
    // groups of small tiles of at most about 1.5 MB (an XCD's L2 holds 4 MB)
    const size_t tile_bytes = static_cast<size_t>(A_BIG ? BN : 256) * kp * 2;
    const int fit = static_cast<int>(std::max<size_t>(1, (size_t{3} << 19) / tile_bytes));
    XcdTiles xt{1, 0, 0};
    while (xt.groups < 8 && (small_tiles + xt.groups - 1) / xt.groups > fit) xt.groups *= 2;
    xt.small_per = (small_tiles + xt.groups - 1) / xt.groups;
    const int parts = 8 / xt.groups;
    xt.big_per = (big_tiles + parts - 1) / parts;
    const size_t per_xcd = static_cast<size_t>(xt.big_per) * xt.small_per;
    hipLaunchKernelGGL((k_gemm_bf16_lds<KS, TN, NST, A_BIG, Epi, NW>), dim3(static_cast<unsigned>(8 * per_xcd)),
                       dim3(64 * NW), 0, s, A, M, lda, B, N, ldb, kp, xt, epi);
}

__global__ void k_zero_cnt(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// The tile shapes the launches choose from: wave tiles of (32 TM) x (32 TN)
// with (TM, TN) = (2, 2), (2, 4), (4, 2), and (4, 4) fed by direct loads
// through the LDS ring (k_gemm_bf16_lds) in stages of K 32 or 64, with four
// waves per workgroup or (256 x 256 tiles of 128 x 64 wave tiles) eight.
constexpr int kVariants = 7;

template <bool A_BIG, class Epi>
void gemm_variant(int v, const uint16_t* A, int M, size_t lda, const uint16_t* B, int N, size_t ldb, int kp,
                  const Epi& epi, hipStream_t s) {
    switch (v) {
    case 1: gemm<2, 4, A_BIG>(A, M, lda, B, N, ldb, kp, epi, s); break;
    case 2: gemm<4, 2, A_BIG>(A, M, lda, B, N, ldb, kp, epi, s); break;
    case 3: gemm_lds<32, 4, 4, A_BIG>(A, M, lda, B, N, ldb, kp, epi, s); break;
    case 4:
        // stages of 64 need K a multiple of 64
        if (kp % 64 == 0)
            gemm_lds<64, 4, 2, A_BIG>(A, M, lda, B, N, ldb, kp, epi, s);
        else
            gemm_lds<32, 4, 4, A_BIG>(A, M, lda, B, N, ldb, kp, epi, s);
        break;
    case 5: gemm_lds<32, 2, 3, A_BIG>(A, M, lda, B, N, ldb, kp, epi, s); break;
    case 6: gemm_lds<32, 2, 4, A_BIG, 8>(A, M, lda, B, N, ldb, kp, epi, s); break;
    default: gemm<2, 2, A_BIG>(A, M, lda, B, N, ldb, kp, epi, s); break;
    }
}

// The autotuner: per problem class (which epilogue, K, the size classes of M
// and N), the first launch times every tile shape on the problem itself (HIP
// events on the caller's stream; reset() undoes an epilogue's effect between
// trials) and keeps the fastest for the life of the process.
struct TuneKey {
    int kind, kp, mclass, nclass;
    bool operator<(const TuneKey& o) const {
        return kind != o.kind ? kind < o.kind : kp != o.kp ? kp < o.kp : mclass != o.mclass ? mclass < o.mclass : nclass < o.nclass;
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;

int size_class(int x) {
    int c = 0;
    while ((1 << c) < x) ++c;
    return c;
}

template <bool A_BIG, class Epi, class Reset>
void gemm_tuned(int kind, const uint16_t* A, int M, size_t lda, const uint16_t* B, int N, size_t ldb, int kp,
                const Epi& epi, Reset reset, hipStream_t s) {
    const TuneKey key{kind, kp, size_class(M), size_class(N)};
    int best = -1;
    // QANN_AMD_GEMM_SHAPE=0..6 forces one tile shape (tests check each)
    const char* force = std::getenv("QANN_AMD_GEMM_SHAPE");
    if (force && *force >= '0' && *force < '0' + kVariants && force[1] == '\0') {
        gemm_variant<A_BIG>(*force - '0', A, M, lda, B, N, ldb, kp, epi, s);
        return;
    }
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) best = it->second;
    }
    if (best < 0) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            best = 0;
        } else {
            float best_ms = 0.0f, times[kVariants] = {};
            for (int v = 0; v < kVariants; ++v) {
                reset();
                (void)hipEventRecord(a, s);
                gemm_variant<A_BIG>(v, A, M, lda, B, N, ldb, kp, epi, s);
                (void)hipEventRecord(b, s);
                float ms = 0.0f;
                if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) continue;
                times[v] = ms;
                if (best < 0 || ms < best_ms) {
                    best = v;
                    best_ms = ms;
                }
            }
            // QANN_AMD_TUNE_LOG=1 prints each choice and its trial times
            const char* log = std::getenv("QANN_AMD_TUNE_LOG");
            if (log && *log && *log != '0')
                std::fprintf(stderr,
                             "qann_amd tune: gemm kind %d M %d N %d K %d -> shape %d (ms: %.3f %.3f %.3f %.3f %.3f %.3f)\n",
                             kind, M, N, kp, best, times[0], times[1], times[2], times[3], times[4], times[5]);
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            if (best < 0) best = 0;
            std::lock_guard<std::mutex> g(g_tune_mu);
            g_tuned[key] = best;
        }
        reset();
    }
    gemm_variant<A_BIG>(best, A, M, lda, B, N, ldb, kp, epi, s);
}

} // namespace

int gemm_tuned_variant(int kind, int kp, int M, int N) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find(TuneKey{kind, kp, size_class(M), size_class(N)});
    return it == g_tuned.end() ? -1 : it->second;
}

void launch_split_bf16(const float* x, uint64_t n, int dim, int kp, int order, uint16_t* out, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_split_bf16, dim3(cdiv(n, 4)), dim3(256), 0, s, x, n, dim, kp, order, out);
}

void launch_coarse_split(const uint16_t* q3, int nq, const uint16_t* c3, int nlist, int kp, const float* csq, int ip,
                         float* D, hipStream_t s) {
    if (nq <= 0 || nlist <= 0) return;
    EpiScores e{D, static_cast<size_t>(nlist), ip, 1, csq, nullptr, nullptr};
    const size_t k3 = 3 * static_cast<size_t>(kp);
    gemm_tuned<true>(kGemmScores, q3, nq, k3, c3, nlist, k3, static_cast<int>(k3), e, [] {}, s);
}

void launch_rows_bf16(const float* rows, uint64_t n, int dim, int kp, uint16_t* out, float* maxes, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_rows_bf16, dim3(cdiv(n, 4)), dim3(256), 0, s, rows, n, dim, kp, out, maxes);
}

void launch_queries_bf16(const float* q, int nq, int dim, int kp, uint16_t* out, float* qinfo, hipStream_t s) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_queries_bf16, dim3(cdiv(nq, 4)), dim3(256), 0, s, q, nq, dim, kp, out, qinfo);
}

void launch_flat_bf16(const FlatBf16& f, const float* q, const uint16_t* qb, const float* qinfo, int nq, int k,
                      int want, int stride, int ns, float* Ds, float* lim, uint32_t* cnt, float* cd, uint32_t* ci,
                      int64_t* out_ids, float* out_dists, uint32_t* ovf, uint32_t* ovf_cnt, int ovf_base,
                      hipStream_t s) {
    if (nq <= 0) return;
    // the sample: every stride-th row, one minimum per query per 128 of them
    const int ntile = flat_sample_tiles(ns);
    EpiTileMin et{Ds, ntile, f.ip, stride, f.xsq, f.excl, f.incl};
    gemm<2, 2, true>(f.xb, ns, static_cast<size_t>(stride) * f.kp, qb, nq, static_cast<size_t>(f.kp), f.kp, et, s);
    hipLaunchKernelGGL(k_flat_limit, dim3(static_cast<unsigned>(nq)), dim3(256), 0, s, Ds, ntile, want, qinfo,
                       f.maxes, f.xnorm_max, f.ip, f.dim, lim, cnt);
    EpiFilter ef{lim, f.ip, kFlatCandCap, f.xsq, f.excl, f.incl, cnt, cd, ci};
    // k_flat_limit zeroed the counters; a tuning trial zeroes them again
    auto reset = [&] { hipLaunchKernelGGL(k_zero_cnt, dim3(cdiv(nq, 256)), dim3(256), 0, s, cnt, nq); };
    gemm_tuned<true>(kGemmFilter, f.xb, f.n, static_cast<size_t>(f.kp), qb, nq, static_cast<size_t>(f.kp), f.kp, ef,
                     reset, s);
    const size_t lds = query_lds_bytes(f.dim) + static_cast<size_t>(kFlatCandCap) * 8;
    hipLaunchKernelGGL(k_flat_finish, dim3(static_cast<unsigned>(nq)), dim3(256), lds, s, q, f.dim, f.ip, f.rows,
                       qinfo, f.maxes, f.xnorm_max, k, want, kFlatCandCap, cnt, cd, ci, f.labels, out_ids, out_dists,
                       ovf, ovf_cnt, ovf_base);
}

} // namespace qd
