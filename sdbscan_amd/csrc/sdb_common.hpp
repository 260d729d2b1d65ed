// Shared device helpers for the sDbscan neighbourhood-discovery kernels
// (gfx950, 64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdb {

constexpr int kWave = 64;

enum Metric { kCosine = 0, kL1 = 1, kL2 = 2, kChi2 = 3, kJS = 4 };
enum Precision { kFP32 = 0, kBF16 = 1 };

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Order-preserving map of a float to uint32; -0 and +0 map to the same key so
// that they compare equal, as they do for the reference's float comparisons.
__device__ __forceinline__ uint32_t ord_key(float f) {
    uint32_t u = __float_as_uint(f);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_to_float(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, m, kWave);
    hi = __shfl_xor(hi, m, kWave);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t o = shfl_xor_u64(v, m);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ double wave_sum_f64_tree(double v) {
    // xor butterfly, offsets 32..1; every lane ends with the same value
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, kWave);
    return v;
}

// A counter that every wave of a large grid adds to, kept as kSpread copies
// 128 bytes apart: the workgroup adds to copy (blockIdx.x & spread), so with
// spread = kSpread - 1 no single address serialises every wave's atomic (one
// address took about 1.5 ms for the 125k waves of an 8M-point pass); spread =
// 0 uses copy 0 only, i.e. the counter itself.  k_spread_fold adds the copies
// into the final counter.
constexpr int kSpread = 64, kSpreadStride = 16;  // stride in 8-byte words
__device__ __forceinline__ void spread_add(unsigned long long* ctr, int spread, unsigned long long v) {
    atomicAdd(ctr + (size_t)(blockIdx.x & (unsigned)spread) * kSpreadStride, v);
}

// Group entries (the grouped (point, slot) entries the distance tiles walk):
// bits 0-30 hold the entry a * S + s (n * S < 2^31), bit 31 is set when the
// entry's first-occurrence mask is the full one (every list position, none
// repeated and not the point itself), which the tiles then do not load.
constexpr uint32_t kEntFull = 0x80000000u;
__device__ __forceinline__ uint32_t ent_id(uint32_t g) { return g & ~kEntFull; }
__device__ __forceinline__ bool ent_full(uint32_t g) { return (g & kEntFull) != 0u; }
// Word x of the full first-occurrence mask over M positions.
__device__ __forceinline__ uint32_t fmask_full_word(int x, int M) {
    const int nb = M - 32 * x;
    return nb >= 32 ? 0xFFFFFFFFu : (nb > 0 ? (1u << nb) - 1u : 0u);
}

// BF16 bits of a non-negative finite float rounded towards +infinity (an
// upper bound of the value in BF16).
__device__ __forceinline__ uint32_t bf16_up(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u >> 16) + ((u & 0xFFFFu) ? 1u : 0u);
}

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Atomic load served from this XCD's L2 (workgroup scope: no cross-XCD
// revalidation), used for union-find parents, the adjacency hints and the
// border minima.  A value read here may be stale, never newer than memory;
// those words only move one way (a parent only ever points to a smaller
// member of the same set, a minimum only falls), so a stale value costs an
// extra hop, a failed compare-and-swap that returns the fresh value, or a
// redundant atomicMin -- never a wrong merge.
__device__ __forceinline__ int load_relaxed(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Point index a = ent / S of an entry ent = a * S + s (ent < 2^31, S < 512)
// with magic = ceil(2^40 / S): bits 40 and up of the full product (a 64-bit
// product would overflow once ent >= S * 2^24, i.e. from 2^24 points on).
__device__ __forceinline__ unsigned long long ent_point(uint32_t ent, unsigned long long magic) {
    const unsigned long long lo = (unsigned long long)ent * magic;
    const unsigned long long hi = __umul64hi((unsigned long long)ent, magic);
    return (hi << 24) | (lo >> 40);
}

__device__ __forceinline__ uint32_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return r >> 16;
}

// Bound on the BF16 rounding of one row, for the per-pair error band of the
// BF16 distance tiles.  sq is the FP32 sum, in any order, of (b_i - s_i)^2
// over the row, b_i the BF16 rounding of the FP32 value s_i (the differences
// are exact).  Returns rho >= (1 + 2^-7) ||b - r||, r the exact row the tiles
// stand for: the transformed row itself (Cosine), or the exact square roots
// of the transformed Chi2 row, which the correctly rounded FP32 square roots s
// approach within 2^-24 relative (||s - r|| <= 2^-23 for a unit-sum row).
// The factor 1.001 covers the FP32 sum of at most 4096 squares, 1.0079 the
// factor 1 + 2^-7 and the rounding of these few steps.
__device__ __forceinline__ float bf16_row_err(float sq, bool sqrt_rows) {
    const float e = sqrtf(sq * 1.001f) + (sqrt_rows ? 0x1p-23f : 0.0f);
    return e * 1.0079f;
}

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, kWave);
    return v;
}

// FP64 distance of two FP32 rows in a fixed order: per lane a sequential sum
// over k = lane, lane + 64, ..., then the xor butterfly (offsets 32..1);
// every lane returns the same value.  The decision value of the recheck.
// s <- op(s, x[i], y[i]) for i = lane, lane + 64, ... < d, in that order, with
// the loads of U iterations issued before their terms: the arithmetic of the
// plain loop without one memory round trip per iteration.
template <int U, class T, class Op>
__device__ __forceinline__ T wave_row_fold(const float* x, const float* y, int d, int lane, T s, Op op) {
    int i = lane;
    for (; i + (U - 1) * 64 < d; i += U * 64) {
        float xv[U], yv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv[u] = x[i + 64 * u];
            yv[u] = y[i + 64 * u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) s = op(s, xv[u], yv[u]);
    }
    float xv[U], yv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = i + 64 * u;
        xv[u] = k < d ? x[k] : 0.0f;
        yv[u] = k < d ? y[k] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (i + 64 * u < d) s = op(s, xv[u], yv[u]);
    return s;
}

// The same over one row: s <- op(s, x[i]).
template <int U, class T, class Op>
__device__ __forceinline__ T wave_row_fold1(const float* x, int d, int lane, T s, Op op) {
    int i = lane;
    for (; i + (U - 1) * 64 < d; i += U * 64) {
        float xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[i + 64 * u];
#pragma unroll
        for (int u = 0; u < U; ++u) s = op(s, xv[u]);
    }
    float xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = i + 64 * u < d ? x[i + 64 * u] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (i + 64 * u < d) s = op(s, xv[u]);
    return s;
}

template <int METRIC>
__device__ __forceinline__ double sum64_wave(const float* x, const float* y, int d, int lane) {
    return wave_row_fold<8>(x, y, d, lane, 0.0, [](double s, float a, float b) {
        const double xv = a, yv = b;
        if constexpr (METRIC == kCosine) return s + xv * yv;
        else if constexpr (METRIC == kL1) return s + fabs(xv - yv);
        else if constexpr (METRIC == kL2) {
            const double t = xv - yv;
            return s + t * t;
        } else if constexpr (METRIC == kChi2) {
            return (a > 0.0f && b > 0.0f) ? s + (xv * yv) / (xv + yv) : s;
        } else {
            const double xe = xv + 1e-6, ye = yv + 1e-6, t = xe + ye, ln2 = 0.69314718055994530942;
            return s + (xe * (log(t / xe) / ln2) + ye * (log(t / ye) / ln2));
        }
    });
}

__device__ __forceinline__ double dist64_wave(const float* x, const float* y, int d, int metric, int lane) {
    double s;
    switch (metric) {
        case kCosine: s = sum64_wave<kCosine>(x, y, d, lane); break;
        case kL1: s = sum64_wave<kL1>(x, y, d, lane); break;
        case kL2: s = sum64_wave<kL2>(x, y, d, lane); break;
        case kChi2: s = sum64_wave<kChi2>(x, y, d, lane); break;
        default: s = sum64_wave<kJS>(x, y, d, lane); break;
    }
    s = wave_sum_f64_tree(s);
    if (metric == kCosine) return 1.0 - s;
    if (metric == kL1) return s;
    if (metric == kL2) return sqrt(s);
    if (metric == kChi2) return 1.0 - 2.0 * s;
    return 1.0 - s / 2.0;
}

// Cross-lane exchange with lane ^ S without the LDS crossbar: DPP quad
// permutes for S = 1, 2; two DPP row permutes (half-mirror / mirror, then a
// quad or half-mirror permute) for S = 4, 8; the gfx950 permlane swaps for
// S = 16, 32.  Every lane has a valid DPP source, so bound_ctrl leaves no
// lane to an old value and no old value is materialised.
template <int S>
__device__ __forceinline__ uint32_t xchg(uint32_t v, int lane) {
    const int iv = (int)v;
    if constexpr (S == 1) return (uint32_t)__builtin_amdgcn_mov_dpp(iv, 0xB1, 0xF, 0xF, true);
    if constexpr (S == 2) return (uint32_t)__builtin_amdgcn_mov_dpp(iv, 0x4E, 0xF, 0xF, true);
    if constexpr (S == 4) {
        const int h = __builtin_amdgcn_mov_dpp(iv, 0x141, 0xF, 0xF, true);
        return (uint32_t)__builtin_amdgcn_mov_dpp(h, 0x1B, 0xF, 0xF, true);
    }
    if constexpr (S == 8) {
        const int m = __builtin_amdgcn_mov_dpp(iv, 0x140, 0xF, 0xF, true);
        return (uint32_t)__builtin_amdgcn_mov_dpp(m, 0x141, 0xF, 0xF, true);
    }
    if constexpr (S == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16) ? p[0] : p[1];
    }
    if constexpr (S == 32) {
        const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (lane & 32) ? p[0] : p[1];
    }
    return v;
}

// Mirror lane layout of the index kernels.  Physical lane p holds logical
// element lane_logical(p): bits 4-5 are kept and the low nibble is mapped by
// the GF(2)-linear bijection with logical bit 0 -> p ^ 1, bit 1 -> p ^ 2,
// bit 2 -> p ^ 7 (DPP half-mirror), bit 3 -> p ^ 15 (DPP mirror), so the
// exchange with logical partner lane ^ S is a single DPP move for every S < 16.
__host__ __device__ constexpr int lane_logical(int p) {
    return (p & 48) | (((p ^ (p >> 2)) & 1)) | ((((p >> 1) ^ (p >> 2)) & 1) << 1) |
           ((((p >> 2) ^ (p >> 3)) & 1) << 2) | (((p >> 3) & 1) << 3);
}

// Exchange with the lane holding logical element lg ^ S (mirror layout).
template <int S>
__device__ __forceinline__ uint32_t xchg_l(uint32_t v, int lg) {
    const int iv = (int)v;
    if constexpr (S == 1) return (uint32_t)__builtin_amdgcn_mov_dpp(iv, 0xB1, 0xF, 0xF, true);
    if constexpr (S == 2) return (uint32_t)__builtin_amdgcn_mov_dpp(iv, 0x4E, 0xF, 0xF, true);
    if constexpr (S == 4) return (uint32_t)__builtin_amdgcn_mov_dpp(iv, 0x141, 0xF, 0xF, true);
    if constexpr (S == 8) return (uint32_t)__builtin_amdgcn_mov_dpp(iv, 0x140, 0xF, 0xF, true);
    return xchg<S>(v, lg);  // bits 4 and 5 of lg are those of the physical lane
}

// Exclusive scan of one int per thread over a workgroup of 1024 threads;
// *total gets the workgroup's sum (wsum: 16 ints of LDS).
__device__ __forceinline__ int block_scan_excl_1024(int v, int* wsum, int* total) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    if (w == 0) {
        int s = lane < 16 ? wsum[lane] : 0;
        int si = s;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(si, o, kWave);
            if (lane >= o) si += y;
        }
        if (lane < 16) wsum[lane] = si - s;  // exclusive wave offsets
        if (lane == 15) *total = si;
    }
    __syncthreads();
    const int r = wsum[w] + incl - v;
    __syncthreads();
    return r;
}

}  // namespace sdb
