// sOPTICS support and synthetic workloads on gfx950.
//
// Reference: rp_findCoreDist (src/sDbscan.cpp:1314-1474) keeps every
// deduplicated (neighbour, distance) pair and takes the (minPts-1)-th smallest
// distance as core distance; formOptics_scikit (src/sDbscan.cpp:1588-1703)
// sets reach(j) = min over core points i processed before j with j in N(i)
// of max(core_dist(i), d(i,j)), and -2^31 for points no core point reaches.
#include "sdb_kernels.hpp"

namespace sdb {

__device__ __forceinline__ int slot_list_o(const int* topK, long long a, int s, int K, int P) {
    int k = s >> 1;
    return (s & 1) ? P + topK[a * 2 * K + K + k] : topK[a * 2 * K + k];
}

// Neighbourhood CSR: every hit (a -> c) lands in a's row; it lands in c's
// row too unless a is itself a candidate of c (then c's own hit covers it).
__global__ void k_csr_fill(ClusterArgs c, const long long* ptr, int* cursor, int* idx) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.n * c.S) return;
    const long long a = t / c.S;
    const int s = (int)(t % c.S);
    const int l = slot_list_o(c.topK, a, s, c.K, c.P);
    const int h = c.hub_id[a];
    for (int w = 0; w < c.W; ++w) {
        uint32_t e = c.E[(size_t)t * c.W + w];
        if (!e) continue;
        uint32_t mm = 0;
        if (h >= 0)
            for (int q = c.hub_ptr[h]; q < c.hub_ptr[h + 1]; ++q)
                mm |= c.inv[((size_t)l * c.L2 + c.hub_mem[q] / c.M) * c.W + w];
        while (e) {
            const int b = __ffs(e) - 1;
            e &= e - 1;
            const int cp = c.lists[(size_t)l * c.M + w * 32 + b];
            int pos = atomicAdd(&cursor[a], 1);
            if (ptr[a] + pos < ptr[a + 1]) idx[ptr[a] + pos] = cp;
            if (!((mm >> b) & 1u)) {
                int pos2 = atomicAdd(&cursor[cp], 1);
                if (ptr[cp] + pos2 < ptr[cp + 1]) idx[ptr[cp] + pos2] = (int)a;
            }
        }
    }
}

// FP32 distance of every CSR entry, one wavefront per point; sums run per
// lane with stride 64 and combine through an xor butterfly, so d(a,c) and
// d(c,a) are bitwise equal for every metric but JS (whose two terms are added
// in operand order).
__global__ __launch_bounds__(256) void k_csr_dist(const float* X, int ldx, int d, int metric,
                                                  long long n, const long long* ptr, const int* idx,
                                                  float* dist) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long a = wv; a < n; a += nwv) {
        const float* x = X + (size_t)a * ldx;
        for (long long j = ptr[a]; j < ptr[a + 1]; ++j) {
            const float* y = X + (size_t)idx[j] * ldx;
            float s = wave_row_fold<8>(x, y, d, lane, 0.0f, [metric](float t, float xv, float yv) {
                if (metric == kCosine) return fmaf(xv, yv, t);
                if (metric == kL1) return t + fabsf(xv - yv);
                if (metric == kL2) {
                    const float df = xv - yv;
                    return fmaf(df, df, t);
                }
                if (metric == kChi2) return (xv > 0.0f && yv > 0.0f) ? t + xv * yv / (xv + yv) : t;
                const float xe = xv + 1e-6f, ye = yv + 1e-6f, u = xe + ye;
                return t + xe * (logf(u / xe) * 1.4426950408889634f) + ye * (logf(u / ye) * 1.4426950408889634f);
            });
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) s = s + __shfl_xor(s, m, kWave);
            float r;
            if (metric == kCosine) r = 1.0f - s;
            else if (metric == kL1) r = s;
            else if (metric == kL2) r = sqrtf(s);
            else if (metric == kChi2) r = 1.0f - 2.0f * s;
            else r = 1.0f - s * 0.5f;
            if (lane == 0) dist[j] = r;
        }
    }
}

// The per-lane term of the neighbourhood distances (as k_csr_dist's fold).
__device__ __forceinline__ float csr_term(int metric, float t, float xv, float yv) {
    if (metric == kCosine) return fmaf(xv, yv, t);
    if (metric == kL1) return t + fabsf(xv - yv);
    if (metric == kL2) {
        const float df = xv - yv;
        return fmaf(df, df, t);
    }
    if (metric == kChi2) return (xv > 0.0f && yv > 0.0f) ? t + xv * yv / (xv + yv) : t;
    const float xe = xv + 1e-6f, ye = yv + 1e-6f, u = xe + ye;
    return t + xe * (logf(u / xe) * 1.4426950408889634f) + ye * (logf(u / ye) * 1.4426950408889634f);
}

// The same distances as k_csr_dist, term for term and in the same order
// (lane l sums i = l, l + 64, ... < d in turn, then the xor butterfly), for
// rows of at most 64 NR elements: the point's row stays in registers and J
// neighbour rows are loaded together before their sums, so that a wave has
// J row loads in flight instead of one.
template <int NR, int J>
__global__ __launch_bounds__(256) void k_csr_dist_j(const float* X, int ldx, int d, int metric, long long n,
                                                    const long long* ptr, const int* idx, float* dist,
                                                    int upper) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long a = wv; a < n; a += nwv) {
        float xr[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int i = lane + 64 * k;
            xr[k] = i < d ? X[(size_t)a * ldx + i] : 0.0f;
        }
        const long long e = ptr[a + 1];
        long long jb = ptr[a];
        if (upper) {
            // skip the entries with id < a (the row is sorted by id): 64 ids a step
            for (;;) {
                const long long q = jb + lane;
                const int below = __popcll(__ballot(q < e && (long long)idx[q] < a));
                jb += below;
                if (below < 64) break;
            }
        }
        for (long long j = jb; j < e; j += J) {
            float yv[J][NR];
#pragma unroll
            for (int u = 0; u < J; ++u) {
                const long long c = j + u < e ? (long long)idx[j + u] : a;
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    const int i = lane + 64 * k;
                    yv[u][k] = i < d ? X[(size_t)c * ldx + i] : 0.0f;
                }
            }
            float s[J];
#pragma unroll
            for (int u = 0; u < J; ++u) {
                s[u] = 0.0f;
#pragma unroll
                for (int k = 0; k < NR; ++k)
                    if (lane + 64 * k < d) s[u] = csr_term(metric, s[u], xr[k], yv[u][k]);
            }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
                for (int u = 0; u < J; ++u) s[u] = s[u] + __shfl_xor(s[u], m, kWave);
#pragma unroll
            for (int u = 0; u < J; ++u) {
                float r;
                if (metric == kCosine) r = 1.0f - s[u];
                else if (metric == kL1) r = s[u];
                else if (metric == kL2) r = sqrtf(s[u]);
                else if (metric == kChi2) r = 1.0f - 2.0f * s[u];
                else r = 1.0f - s[u] * 0.5f;
                if (lane == u && j + u < e) dist[j + u] = r;
            }
        }
    }
}

// Core distance of each core point: the k-th smallest distance of its
// neighbourhood (k = minPts - 2; 0 for other points and for rows of at most k
// entries), the element k of the row sorted by distance, found without
// sorting.  A radix selection over the row's 32-bit order keys (the float's
// bits with the sign bit flipped when clear and every bit flipped when set)
// fixes four 8-bit digits from the most significant one; each digit comes
// from a 256-bin LDS histogram of the row's keys that match the digits so
// far.  One wavefront per workgroup and per point.
__device__ __forceinline__ uint32_t raw_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float raw_unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ __launch_bounds__(64) void k_kth_select(long long n, const long long* ptr, const float* dist,
                                                   const uint8_t* core, int k, float* out) {
    __shared__ unsigned int h[256];
    const int lane = (int)threadIdx.x;
    for (long long a = blockIdx.x; a < n; a += gridDim.x) {
        const long long b = ptr[a], e = ptr[a + 1];
        float v = 0.0f;
        if (core[a] && e - b > k) {
            uint32_t prefix = 0u;
            unsigned int kk = (unsigned int)k;  // rank inside the keys matching prefix
            for (int shift = 24; shift >= 0; shift -= 8) {
                for (int i = lane; i < 256; i += 64) h[i] = 0u;
                __syncthreads();
                const uint32_t hm = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
                for (long long j = b + lane; j < e; j += 64) {
                    const uint32_t u = raw_key(dist[j]);
                    if ((u & hm) == prefix) atomicAdd(&h[(u >> shift) & 255u], 1u);
                }
                __syncthreads();
                const unsigned int c0 = h[4 * lane], c1 = h[4 * lane + 1], c2 = h[4 * lane + 2], c3 = h[4 * lane + 3];
                const unsigned int sum = c0 + c1 + c2 + c3;
                unsigned int incl = sum;
#pragma unroll
                for (int m = 1; m < 64; m <<= 1) {
                    const unsigned int t = __shfl_up(incl, m, kWave);
                    if (lane >= m) incl += t;
                }
                const unsigned int excl = incl - sum;
                // exactly one lane's four bins hold rank kk (kk < the matching count)
                const bool mine = excl <= kk && kk < incl;
                const int src = __ffsll((long long)__ballot(mine)) - 1;
                unsigned int bin = 4u * (unsigned int)lane, rest = kk - excl;
                if (rest >= c0) {
                    rest -= c0; ++bin;
                    if (rest >= c1) {
                        rest -= c1; ++bin;
                        if (rest >= c2) { rest -= c2; ++bin; }
                    }
                }
                prefix |= (uint32_t)__shfl((int)bin, src, kWave) << shift;
                kk = (unsigned int)__shfl((int)rest, src, kWave);
                __syncthreads();  // every lane has read h before the next digit clears it
            }
            v = raw_unkey(prefix);
        }
        if (lane == 0) out[a] = v;
    }
}

// One wavefront per core point i, its lanes striding over i's neighbourhood
// (the minima do not depend on the order of the updates).
__global__ __launch_bounds__(256) void k_reach(long long n, const long long* ptr, const int* idx, const float* dist,
                                               const uint8_t* core, const float* core_dist, const int* pos,
                                               uint32_t* reach_key) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long i = wv; i < n; i += nwv) {
        if (!core[i]) continue;
        const int pi = pos[i];
        const float cd = core_dist[i];
        for (long long j = ptr[i] + lane; j < ptr[i + 1]; j += kWave) {
            const int q = idx[j];
            if (pos[q] <= pi) continue;
            const uint32_t k = ord_key(fmaxf(cd, dist[j]));
            if (k < (uint32_t)load_relaxed((const int*)&reach_key[q])) atomicMin(&reach_key[q], k);
        }
    }
}

__global__ void k_reach_final(long long n, uint32_t* reach_key) {
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const uint32_t k = reach_key[a];
    const float v = k == 0xffffffffu ? -2147483648.0f : key_to_float(k);
    reach_key[a] = __float_as_uint(v);
}

// ---------------------------------------------------------------------------
// Synthetic mixture (10 components): centres uniform in [0,1]^d with 80% of
// the coordinates zeroed, point = centre + N(0, sigma^2) clipped to [0,1],
// L2-normalised.  Counter-based (splitmix64) so any row can be regenerated.
__device__ __forceinline__ uint64_t smix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float u24(uint64_t h) { return (float)(h >> 40) * (1.0f / 16777216.0f); }

__global__ __launch_bounds__(256) void k_gen_mixture(float* X, long long n, int d, uint64_t seed,
                                                     float sigma) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long p = wv; p < n; p += nwv) {
        const int comp = (int)(smix(seed * 0x5851F42D4C957F2Dull + (uint64_t)p) >> 33) % 10;
        float* x = X + (size_t)p * d;
        float ss = 0.0f;
        for (int i = lane; i < d; i += 64) {
            const uint64_t hc = smix(seed ^ smix(0xC0FFEEull + (uint64_t)comp * 1000003ull + (uint64_t)i));
            const float cen = u24(smix(hc)) < 0.8f ? 0.0f : u24(hc);
            const uint64_t h1 = smix(seed ^ smix(((uint64_t)p * (uint64_t)d + (uint64_t)i) * 2ull + 1ull));
            const uint64_t h2 = smix(h1);
            const float u1 = ((float)(h1 >> 40) + 1.0f) * (1.0f / 16777216.0f);
            const float u2 = u24(h2);
            const float z = sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
            float v = cen + sigma * z;
            v = fminf(fmaxf(v, 0.0f), 1.0f);
            x[i] = v;
            ss = fmaf(v, v, ss);
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) ss = ss + __shfl_xor(ss, m, kWave);
        const float nr = sqrtf(ss);
        if (nr > 0.0f)
            for (int i = lane; i < d; i += 64) x[i] = x[i] / nr;
    }
}

// Dirichlet histograms: 10 components, each with kGenBoost bins (drawn on the
// host) at alpha = 2 and every other bin at alpha = 0.05; x_i = G_i / sum G
// with G_i ~ Gamma(alpha_i).  Gamma by Marsaglia and Tsang ("A simple method
// for generating gamma variables", ACM TOMS 26(3), 2000) with a bounded
// rejection loop, and Gamma(a) = Gamma(a + 1) U^(1/a) for a < 1.
struct GenBins {
    uint16_t bin[10][kGenBoost];
};

__device__ __forceinline__ float u01_open(uint64_t h) { return ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f); }

__device__ float gamma_draw(float alpha, uint64_t h) {
    const bool small = alpha < 1.0f;
    const float a = small ? alpha + 1.0f : alpha;
    const float dd = a - 1.0f / 3.0f, cc = 1.0f / sqrtf(9.0f * dd);
    float v = 1.0f;
    for (int t = 0; t < 16; ++t) {
        const uint64_t h1 = smix(h + 0x632BE59BD9B4E019ull * (uint64_t)(t + 1)), h2 = smix(h1), h3 = smix(h2);
        const float z = sqrtf(-2.0f * logf(u01_open(h1))) * cosf(6.283185307179586f * u24(h2));
        const float x1 = 1.0f + cc * z;
        if (x1 <= 0.0f) continue;
        v = x1 * x1 * x1;
        if (logf(u01_open(h3)) < 0.5f * z * z + dd - dd * v + dd * logf(v)) break;
    }
    float g = dd * v;
    if (small) g *= powf(u01_open(smix(h ^ 0xD1B54A32D192ED03ull)), 1.0f / alpha);
    return g;
}

__global__ __launch_bounds__(256) void k_gen_dirichlet(float* X, long long n, int d, uint64_t seed, GenBins tb) {
    extern __shared__ unsigned char boosted[];  // 10 x d
    for (int i = threadIdx.x; i < 10 * d; i += blockDim.x) boosted[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < 10 * kGenBoost; i += blockDim.x)
        boosted[(i / kGenBoost) * d + tb.bin[i / kGenBoost][i % kGenBoost]] = 1;
    __syncthreads();
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long p = wv; p < n; p += nwv) {
        const int comp = (int)(smix(seed * 0x5851F42D4C957F2Dull + (uint64_t)p) >> 33) % 10;
        float* x = X + (size_t)p * d;
        float s = 0.0f;
        for (int i = lane; i < d; i += 64) {
            const float alpha = boosted[comp * d + i] ? 2.0f : 0.05f;
            const float g = gamma_draw(alpha, smix(seed ^ smix(((uint64_t)p * (uint64_t)d + (uint64_t)i) * 2ull + 1ull)));
            x[i] = g;
            s = s + g;
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) s = s + __shfl_xor(s, m, kWave);
        if (s > 0.0f)
            for (int i = lane; i < d; i += 64) x[i] = x[i] / s;
    }
}

static inline unsigned nb(long long n, int b) { return (unsigned)((n + b - 1) / b); }

int launch_gen_dirichlet(float* X, long long n, int d, unsigned long long seed, const uint16_t* bins,
                         hipStream_t st) {
    GenBins tb;
    for (int c = 0; c < 10; ++c)
        for (int j = 0; j < kGenBoost; ++j) {
            if (bins[c * kGenBoost + j] >= d) return -1;
            tb.bin[c][j] = bins[c * kGenBoost + j];
        }
    long long blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_gen_dirichlet, dim3((unsigned)blocks), dim3(256), (size_t)10 * d, st, X, n, d,
                       (uint64_t)seed, tb);
    return (int)hipGetLastError();
}

int launch_gen_mixture(float* X, long long n, int d, unsigned long long seed, float sigma,
                       hipStream_t st) {
    long long blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_gen_mixture, dim3((unsigned)blocks), dim3(256), 0, st, X, n, d, (uint64_t)seed, sigma);
    return (int)hipGetLastError();
}

int launch_csr_fill(const ClusterArgs& c, const long long* ptr, int* cursor, int* idx, hipStream_t st) {
    hipLaunchKernelGGL(k_csr_fill, dim3(nb(c.n * c.S, 256)), dim3(256), 0, st, c, ptr, cursor, idx);
    return (int)hipGetLastError();
}

int launch_csr_dist(const float* X, int ldx, int d, int metric, long long n, const long long* ptr,
                    const int* idx, float* dist, int upper, hipStream_t st) {
    if (upper && (d > 1024 || metric == kJS)) return (int)hipErrorInvalidValue;
    long long blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    const dim3 g((unsigned)blocks), b(256);
#define SDB_CSR_ARGS X, ldx, d, metric, n, ptr, idx, dist, upper
    if (d <= 64) hipLaunchKernelGGL((k_csr_dist_j<1, 8>), g, b, 0, st, SDB_CSR_ARGS);
    else if (d <= 128) hipLaunchKernelGGL((k_csr_dist_j<2, 8>), g, b, 0, st, SDB_CSR_ARGS);
    else if (d <= 256) hipLaunchKernelGGL((k_csr_dist_j<4, 4>), g, b, 0, st, SDB_CSR_ARGS);
    else if (d <= 512) hipLaunchKernelGGL((k_csr_dist_j<8, 4>), g, b, 0, st, SDB_CSR_ARGS);
    else if (d <= 1024) hipLaunchKernelGGL((k_csr_dist_j<16, 2>), g, b, 0, st, SDB_CSR_ARGS);
    else hipLaunchKernelGGL(k_csr_dist, g, b, 0, st, X, ldx, d, metric, n, ptr, idx, dist);
#undef SDB_CSR_ARGS
    return (int)hipGetLastError();
}

// The lower entries (a, c), c < a, of the neighbourhoods after the upper pass
// of k_csr_dist_j: each takes the distance of the entry (c, a) of row c,
// found by binary search (rows are sorted by id).  The two are the same
// number: the sums run over the same products in the same lane order with
// the operands swapped, and every term but JS's is symmetric in them
// (fmaf(x, y, t) == fmaf(y, x, t), (x - y)^2 == (y - x)^2, |x - y|, xy / (x + y)).
// Entries without a mirror are counted in *missing; the caller then runs the
// full pass.  One wavefront per point, lanes over the row's lower entries.
__global__ __launch_bounds__(256) void k_csr_mirror(long long n, const long long* ptr, const int* idx, float* dist,
                                                    unsigned int* missing) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long a = wv; a < n; a += nwv) {
        const long long e = ptr[a + 1];
        for (long long j = ptr[a] + lane; j < e; j += 64) {
            const int c = idx[j];
            if ((long long)c >= a) break;
            const long long ce = ptr[c + 1];
            long long lo = ptr[c], hi = ce;
            while (lo < hi) {
                const long long mid = (lo + hi) >> 1;
                if ((long long)idx[mid] < a) lo = mid + 1;
                else hi = mid;
            }
            if (lo < ce && (long long)idx[lo] == a) dist[j] = dist[lo];
            else atomicAdd(missing, 1u);
        }
    }
}

int launch_csr_mirror(long long n, const long long* ptr, const int* idx, float* dist, unsigned int* missing,
                      hipStream_t st) {
    long long blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_csr_mirror, dim3((unsigned)blocks), dim3(256), 0, st, n, ptr, idx, dist, missing);
    return (int)hipGetLastError();
}

int launch_kth(long long n, const long long* ptr, const float* dist, const uint8_t* core, int k,
               float* out, hipStream_t st) {
    if (k < 0) return (int)hipErrorInvalidValue;
    const long long blocks = std::min<long long>(std::max<long long>(n, 1), 262144);
    hipLaunchKernelGGL(k_kth_select, dim3((unsigned)blocks), dim3(64), 0, st, n, ptr, dist, core, k, out);
    return (int)hipGetLastError();
}

int launch_reach(long long n, const long long* ptr, const int* idx, const float* dist,
                 const uint8_t* core, const float* core_dist, const int* pos, uint32_t* reach_key,
                 hipStream_t st) {
    hipError_t e = hipMemsetAsync(reach_key, 0xff, (size_t)n * 4, st);
    if (e != hipSuccess) return (int)e;
    const long long rblocks = std::min<long long>(65536, (n + 3) / 4);
    hipLaunchKernelGGL(k_reach, dim3((unsigned)std::max<long long>(rblocks, 1)), dim3(256), 0, st, n, ptr, idx, dist,
                       core, core_dist, pos, reach_key);
    hipLaunchKernelGGL(k_reach_final, dim3(nb(n, 256)), dim3(256), 0, st, n, reach_key);
    return (int)hipGetLastError();
}

}  // namespace sdb
