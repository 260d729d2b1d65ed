// sDbscan-1NN labelling of noise points on gfx950.
//
// Reference: labelNoise with clusterNoise = 1 (src/sDbscan.cpp:685-807).  A
// sample of the core points is projected again (k_index, kPoints mode, into
// cproj: P x nS, column i = sampled core i).  Every noise point a scores each
// sampled core i with the CEOs estimate
//     est(i) = 0; for k < K: est += cproj[close_k][i]; est -= cproj[far_k][i]
// (that order, FP32), and takes the label of the first core with the largest
// estimate.  Scores are ranked by the 64-bit key (ord(est) << 32) | ~i, so the
// maximum key is the largest estimate at the smallest sample position.
#include "sdb_kernels.hpp"

namespace sdb {

__global__ void k_noise_list(const int* __restrict__ labels, long long n, int* __restrict__ noise,
                             int* __restrict__ count) {
    const long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool is_noise = a < n && labels[a] == -1;
    const unsigned long long b = __ballot(is_noise);
    if (!b) return;
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(count, __popcll(b));
    base = __shfl(base, 0, kWave);
    if (is_noise) {
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
        noise[base + below] = (int)a;
    }
}

// Workgroup (x, y): noise points [x * kNoisePerGroup, ...) against sampled
// cores [y * kCoresPerGroup, ...).  Each thread owns kCoresPerThread cores
// (strided by the workgroup size, so the reads of a projection row are
// coalesced) and walks the workgroup's noise points; the best key of a noise
// point is reduced over the workgroup and merged with an atomic max.
constexpr int kCoreThreads = 256;
constexpr int kCoresPerThread = 8;
constexpr int kCoresPerGroup = kCoreThreads * kCoresPerThread;
constexpr int kNoisePerGroup = 16;

__global__ __launch_bounds__(kCoreThreads) void k_noise_score(const float* __restrict__ cproj, long long nS,
                                                               const int* __restrict__ topK, int K,
                                                               const int* __restrict__ noise, int n_noise,
                                                               unsigned long long* __restrict__ best) {
    __shared__ unsigned long long red[kCoreThreads / kWave];
    const int tid = threadIdx.x;
    const long long c0 = (long long)blockIdx.y * kCoresPerGroup;
    const int y0 = blockIdx.x * kNoisePerGroup;
    for (int yi = 0; yi < kNoisePerGroup && y0 + yi < n_noise; ++yi) {
        const long long a = noise[y0 + yi];
        const int* tk = topK + a * 2 * K;
        float est[kCoresPerThread];
#pragma unroll
        for (int u = 0; u < kCoresPerThread; ++u) est[u] = 0.0f;
        for (int k = 0; k < K; ++k) {
            const float* rc = cproj + (size_t)tk[k] * nS;
            const float* rf = cproj + (size_t)tk[K + k] * nS;
#pragma unroll
            for (int u = 0; u < kCoresPerThread; ++u) {
                const long long i = c0 + tid + (long long)u * kCoreThreads;
                if (i < nS) {
                    est[u] = est[u] + rc[i];
                    est[u] = est[u] - rf[i];
                }
            }
        }
        unsigned long long key = 0ull;
#pragma unroll
        for (int u = 0; u < kCoresPerThread; ++u) {
            const long long i = c0 + tid + (long long)u * kCoreThreads;
            if (i < nS) {
                const unsigned long long kk = ((unsigned long long)ord_key(est[u]) << 32) |
                                              (unsigned long long)(0xffffffffu - (uint32_t)i);
                key = kk > key ? kk : key;
            }
        }
        key = wave_max_u64(key);
        if (lane_id() == 0) red[tid / kWave] = key;
        __syncthreads();
        if (tid == 0) {
            unsigned long long m = red[0];
            for (int w = 1; w < kCoreThreads / kWave; ++w) m = red[w] > m ? red[w] : m;
            if (m > best[y0 + yi]) atomicMax(&best[y0 + yi], m);
        }
        __syncthreads();
    }
}

__global__ void k_noise_assign(const int* __restrict__ noise, int n_noise, const unsigned long long* __restrict__ best,
                               const int* __restrict__ sampled, int* __restrict__ labels) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_noise) return;
    const unsigned long long b = best[t];
    if (b == 0ull) return;
    const uint32_t i = 0xffffffffu - (uint32_t)(b & 0xffffffffull);
    labels[noise[t]] = labels[sampled[i]];
}

int launch_noise_list(const int* labels, long long n, int* noise, int* count, hipStream_t st) {
    hipLaunchKernelGGL(k_noise_list, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, labels, n, noise, count);
    return (int)hipGetLastError();
}

int launch_noise_score(const float* cproj, long long nS, const int* topK, int K, const int* noise, int n_noise,
                       unsigned long long* best, hipStream_t st) {
    if (n_noise <= 0 || nS <= 0) return 0;
    const unsigned gx = (unsigned)((n_noise + kNoisePerGroup - 1) / kNoisePerGroup);
    const unsigned gy = (unsigned)((nS + kCoresPerGroup - 1) / kCoresPerGroup);
    if (gy > 65535u) return -2;
    hipLaunchKernelGGL(k_noise_score, dim3(gx, gy), dim3(kCoreThreads), 0, st, cproj, nS, topK, K, noise, n_noise,
                       best);
    return (int)hipGetLastError();
}

int launch_noise_assign(const int* noise, int n_noise, const unsigned long long* best, const int* sampled,
                        int* labels, hipStream_t st) {
    if (n_noise <= 0) return 0;
    hipLaunchKernelGGL(k_noise_assign, dim3((n_noise + 255) / 256), dim3(256), 0, st, noise, n_noise, best, sampled,
                       labels);
    return (int)hipGetLastError();
}

}  // namespace sdb
