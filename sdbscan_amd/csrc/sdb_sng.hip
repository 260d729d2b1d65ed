// sngDbscan / sngOptics neighbourhoods on gfx950.
//
// Reference: sng_findCorePoints (src/sDbscan.cpp:416-504) and sng_findCoreDist
// (src/sDbscan.cpp:1484-1583).  Point a is compared with the samples
// s_0, s_1, ... of a draw sequence up to its window end (draws equal to a are
// skipped); hits (distance <= eps) enter N(a) and, for the sampled variants,
// a enters N(s_j) as well.  sngDbscan counts |N(a)| with duplicates, sngOptics
// deduplicates.
//
// The sample sequence is cut into chunks of Ms samples, and the chunks play
// the role of the top-M lists of the index: every (point, chunk) pair is an
// entry a * S2 + l (S2 = chunk count rounded up to a power of two), whose F
// mask holds the chunk positions inside a's window that are not a itself.
// The distance tiles of sdb_dist.hip then produce the hit masks E and the
// per-position hit counts (colsum) exactly as for the index.
#include "sdb_kernels.hpp"

namespace sdb {

// F masks: window bits of chunk l for point a, minus the positions where the
// sample is a (occ CSR: positions of each point in the sequence).
__global__ void k_sng_fmask(long long n, int L, int S2, int Ms, int W, int n_samples,
                            const int* __restrict__ win_end, const int* __restrict__ occ_ptr,
                            const int* __restrict__ occ_pos, uint32_t* __restrict__ F) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * L) return;
    const long long a = t / L;
    const int l = (int)(t - a * L);
    const int we = win_end ? win_end[a] : n_samples;
    const int j0 = l * Ms;
    uint32_t m[kMaxW];
#pragma unroll
    for (int w = 0; w < kMaxW; ++w) {
        int nb = min(we, n_samples) - j0 - w * 32;  // valid positions in this word
        nb = max(0, min(nb, min(32, Ms - w * 32)));
        m[w] = nb >= 32 ? 0xffffffffu : ((1u << nb) - 1u);
    }
    for (int q = occ_ptr[a]; q < occ_ptr[a + 1]; ++q) {
        const int i = occ_pos[q] - j0;
        if (i >= 0 && i < Ms) {
#pragma unroll
            for (int w = 0; w < kMaxW; ++w)
                if (w == (i >> 5)) m[w] &= ~(1u << (i & 31));
        }
    }
    uint32_t* Fo = F + ((size_t)a * S2 + l) * W;
#pragma unroll
    for (int w = 0; w < kMaxW; ++w)
        if (w < W) Fo[w] = m[w];
}

// group l = all points in order: grp[l * n + a] = a * S2 + l
__global__ void k_sng_grp(long long n, int L, int S2, uint32_t* __restrict__ grp) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * L) return;
    const int l = (int)(t / n);
    const long long a = t - (long long)l * n;
    grp[t] = (uint32_t)(a * S2 + l);
}

// reverse pushes: every hit at sequence position j adds a to N(s_j)
__global__ void k_sng_rev(int n_samples, const int* __restrict__ samples, const int* __restrict__ colsum,
                          int* __restrict__ rev) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_samples) return;
    const int v = colsum[j];
    if (v) atomicAdd(&rev[samples[j]], v);
}

// |N(a)| with duplicates (sngDbscan) and the core test
__global__ void k_sng_degree(long long n, int L, int S2, int W, const uint32_t* __restrict__ E,
                             const int* __restrict__ rev, int minPts, int* __restrict__ degree,
                             uint8_t* __restrict__ core, unsigned long long* __restrict__ n_core) {
    // n_core: kSpread copies (spread_add)
    const long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = a < n;
    bool is_core = false;
    if (live) {
        int deg = rev ? rev[a] : 0;
        const uint32_t* Ea = E + (size_t)a * S2 * W;
        for (int q = 0; q < L * W; ++q) deg += __popc(Ea[q]);
        degree[a] = deg;
        is_core = deg >= minPts - 1;
        core[a] = is_core ? 1 : 0;
    }
    const unsigned long long b = __ballot(is_core);
    if (lane_id() == 0 && b) spread_add(n_core, kSpread - 1, (unsigned long long)__popcll(b));
}

__device__ __forceinline__ int sng_find(int* par, int x) {
    int cur = load_relaxed(&par[x]);
    if (cur == x) return x;
    int prev = x;
    while (true) {
        int nxt = load_relaxed(&par[cur]);
        if (nxt == cur) break;
        __hip_atomic_store(&par[prev], nxt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        prev = cur;
        cur = nxt;
    }
    return cur;
}

__device__ __forceinline__ void sng_union(int* par, int a, int b) {
    int ra = sng_find(par, a), rb = sng_find(par, b);
    while (ra != rb) {
        if (ra < rb) { int t = ra; ra = rb; rb = t; }
        int old = atomicCAS(&par[ra], ra, rb);
        if (old == ra) return;
        ra = sng_find(par, old);
        rb = sng_find(par, rb);
    }
}

// Every hit is an edge a - s_j of the symmetric neighbourhood graph
// (formCluster, src/sDbscan.cpp:509-637): core-core edges are unions; the
// border pass gives a non-core end the smallest root among its core ends.
__global__ void k_sng_uf(long long n, int L, int S2, int Ms, int W, const uint32_t* __restrict__ E,
                         const int* __restrict__ samples, const uint8_t* __restrict__ core, int* __restrict__ par) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * L) return;
    const long long a = t / L;
    if (!core[a]) return;
    const int l = (int)(t - a * L);
    const uint32_t* Et = E + ((size_t)a * S2 + l) * W;
    for (int w = 0; w < W; ++w) {
        uint32_t e = Et[w];
        while (e) {
            const int b = __ffs(e) - 1;
            e &= e - 1;
            const int r = samples[l * Ms + w * 32 + b];
            if (core[r]) sng_union(par, (int)a, r);
        }
    }
}

__global__ void k_sng_border(long long n, int L, int S2, int Ms, int W, const uint32_t* __restrict__ E,
                             const int* __restrict__ samples, const uint8_t* __restrict__ core,
                             const int* __restrict__ par, int* __restrict__ border_min) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * L) return;
    const long long a = t / L;
    const int l = (int)(t - a * L);
    const bool ca = core[a] != 0;
    const uint32_t* Et = E + ((size_t)a * S2 + l) * W;
    int best = 0x7fffffff;
    for (int w = 0; w < W; ++w) {
        uint32_t e = Et[w];
        while (e) {
            const int b = __ffs(e) - 1;
            e &= e - 1;
            const int r = samples[l * Ms + w * 32 + b];
            const bool cr = core[r] != 0;
            if (ca && !cr) {
                const int root = par[a];
                if (load_relaxed(&border_min[r]) > root) atomicMin(&border_min[r], root);
            } else if (!ca && cr) {
                const int root = par[r];
                best = root < best ? root : best;
            }
        }
    }
    if (best != 0x7fffffff && load_relaxed(&border_min[a]) > best) atomicMin(&border_min[a], best);
}

// Neighbourhood CSR (sngOptics): counts, then the entries of every hit in
// both rows (a's own hit, and a in s_j's row when both_ways).
__global__ void k_sng_csr_count(long long n, int L, int S2, int Ms, int W, int both_ways,
                                const uint32_t* __restrict__ E, const int* __restrict__ samples,
                                long long* __restrict__ cnt) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * L) return;
    const long long a = t / L;
    const int l = (int)(t - a * L);
    const uint32_t* Et = E + ((size_t)a * S2 + l) * W;
    int own = 0;
    for (int w = 0; w < W; ++w) {
        uint32_t e = Et[w];
        own += __popc(e);
        if (both_ways)
            while (e) {
                const int b = __ffs(e) - 1;
                e &= e - 1;
                atomicAdd((unsigned long long*)&cnt[samples[l * Ms + w * 32 + b]], 1ull);
            }
    }
    if (own) atomicAdd((unsigned long long*)&cnt[a], (unsigned long long)own);
}

__global__ void k_sng_csr_fill(long long n, int L, int S2, int Ms, int W, int both_ways,
                               const uint32_t* __restrict__ E, const int* __restrict__ samples,
                               const long long* __restrict__ ptr, unsigned long long* __restrict__ cur,
                               int* __restrict__ idx) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * L) return;
    const long long a = t / L;
    const int l = (int)(t - a * L);
    const uint32_t* Et = E + ((size_t)a * S2 + l) * W;
    for (int w = 0; w < W; ++w) {
        uint32_t e = Et[w];
        while (e) {
            const int b = __ffs(e) - 1;
            e &= e - 1;
            const int r = samples[l * Ms + w * 32 + b];
            const unsigned long long p = atomicAdd(&cur[a], 1ull);
            if (ptr[a] + (long long)p < ptr[a + 1]) idx[ptr[a] + p] = r;
            if (both_ways) {
                const unsigned long long p2 = atomicAdd(&cur[r], 1ull);
                if (ptr[r] + (long long)p2 < ptr[r + 1]) idx[ptr[r] + p2] = (int)a;
            }
        }
    }
}

// distinct entries per row of a row-sorted CSR (dedup of sng_findCoreDist's
// unordered_map, src/sDbscan.cpp:1533-1537)
__global__ void k_uniq_count(long long n, const long long* __restrict__ ptr, const int* __restrict__ idx,
                             long long* __restrict__ ucnt) {
    const long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    long long u = 0;
    for (long long q = ptr[a]; q < ptr[a + 1]; ++q) u += (q == ptr[a] || idx[q] != idx[q - 1]) ? 1 : 0;
    ucnt[a] = u;
}

__global__ void k_uniq_fill(long long n, const long long* __restrict__ ptr, const int* __restrict__ idx,
                            const long long* __restrict__ uptr, int* __restrict__ uidx) {
    const long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    long long o = uptr[a];
    for (long long q = ptr[a]; q < ptr[a + 1]; ++q)
        if (q == ptr[a] || idx[q] != idx[q - 1]) uidx[o++] = idx[q];
}

__global__ void k_rowsize_core(long long n, const long long* __restrict__ ptr, int minPts, int* __restrict__ degree,
                               uint8_t* __restrict__ core, unsigned long long* __restrict__ n_core) {
    // n_core: kSpread copies (spread_add)
    const long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = a < n;
    bool is_core = false;
    if (live) {
        const int deg = (int)(ptr[a + 1] - ptr[a]);
        degree[a] = deg;
        is_core = deg >= minPts - 1;
        core[a] = is_core ? 1 : 0;
    }
    const unsigned long long b = __ballot(is_core);
    if (lane_id() == 0 && b) spread_add(n_core, kSpread - 1, (unsigned long long)__popcll(b));
}

static inline unsigned nb256(long long n) { return (unsigned)((n + 255) / 256); }

int launch_sng_fmask(long long n, int L, int S2, int Ms, int W, int n_samples, const int* win_end,
                     const int* occ_ptr, const int* occ_pos, uint32_t* F, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_fmask, dim3(nb256(n * L)), dim3(256), 0, st, n, L, S2, Ms, W, n_samples, win_end, occ_ptr,
                       occ_pos, F);
    return (int)hipGetLastError();
}
int launch_sng_grp(long long n, int L, int S2, uint32_t* grp, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_grp, dim3(nb256(n * L)), dim3(256), 0, st, n, L, S2, grp);
    return (int)hipGetLastError();
}
int launch_sng_rev(int n_samples, const int* samples, const int* colsum, int* rev, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_rev, dim3(nb256(n_samples)), dim3(256), 0, st, n_samples, samples, colsum, rev);
    return (int)hipGetLastError();
}
int launch_sng_degree(long long n, int L, int S2, int W, const uint32_t* E, const int* rev, int minPts, int* degree,
                      uint8_t* core, unsigned long long* n_core, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_degree, dim3(nb256(n)), dim3(256), 0, st, n, L, S2, W, E, rev, minPts, degree, core,
                       n_core);
    return (int)hipGetLastError();
}
int launch_sng_uf(long long n, int L, int S2, int Ms, int W, const uint32_t* E, const int* samples,
                  const uint8_t* core, int* par, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_uf, dim3(nb256(n * L)), dim3(256), 0, st, n, L, S2, Ms, W, E, samples, core, par);
    return (int)hipGetLastError();
}
int launch_sng_border(long long n, int L, int S2, int Ms, int W, const uint32_t* E, const int* samples,
                      const uint8_t* core, const int* par, int* border_min, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_border, dim3(nb256(n * L)), dim3(256), 0, st, n, L, S2, Ms, W, E, samples, core, par,
                       border_min);
    return (int)hipGetLastError();
}
int launch_sng_csr_count(long long n, int L, int S2, int Ms, int W, int both_ways, const uint32_t* E,
                         const int* samples, long long* cnt, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_csr_count, dim3(nb256(n * L)), dim3(256), 0, st, n, L, S2, Ms, W, both_ways, E, samples,
                       cnt);
    return (int)hipGetLastError();
}
int launch_sng_csr_fill(long long n, int L, int S2, int Ms, int W, int both_ways, const uint32_t* E,
                        const int* samples, const long long* ptr, unsigned long long* cur, int* idx, hipStream_t st) {
    hipLaunchKernelGGL(k_sng_csr_fill, dim3(nb256(n * L)), dim3(256), 0, st, n, L, S2, Ms, W, both_ways, E, samples,
                       ptr, cur, idx);
    return (int)hipGetLastError();
}
int launch_uniq(long long n, const long long* ptr, const int* idx, long long* ucnt, hipStream_t st) {
    hipLaunchKernelGGL(k_uniq_count, dim3(nb256(n)), dim3(256), 0, st, n, ptr, idx, ucnt);
    return (int)hipGetLastError();
}
int launch_uniq_fill(long long n, const long long* ptr, const int* idx, const long long* uptr, int* uidx,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_uniq_fill, dim3(nb256(n)), dim3(256), 0, st, n, ptr, idx, uptr, uidx);
    return (int)hipGetLastError();
}
int launch_rowsize_core(long long n, const long long* ptr, int minPts, int* degree, uint8_t* core,
                        unsigned long long* n_core, hipStream_t st) {
    hipLaunchKernelGGL(k_rowsize_core, dim3(nb256(n)), dim3(256), 0, st, n, ptr, minPts, degree, core, n_core);
    return (int)hipGetLastError();
}

}  // namespace sdb
