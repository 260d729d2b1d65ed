// Degrees, core test and cluster labels of sDbscan on gfx950.
//
// Reference: the dedup + core test of rp_findCorePoints
// (src/sDbscan.cpp:374-400) and formCluster (src/sDbscan.cpp:509-637).
//
// Neighbourhood of a point a in the reference: N(a) = O(a) u R(a) with
// O(a) = {c in Cand(a) : hit(a,c)} and R(a) = {b : a in Cand(b), hit(a,b)}
// (edges are pushed both ways, then deduplicated).  Here
//   |O(a)| = popcount of a's first-occurrence hit masks E,
//   |R(a)| = sum of the list-position hit counts (colsum) over a's list
//            memberships (non-zero only for points that sit in a top-M list),
//   |O(a) n R(a)| = hits c of a for which a is itself a candidate of c,
//            found with the inv masks.
// Labels follow formCluster exactly: clusters are the components of the
// core graph numbered by their smallest core index, a border point takes the
// smallest cluster id among its core neighbours, everything else is -1.
#include <hipcub/hipcub.hpp>
#include "sdb_kernels.hpp"

namespace sdb {

__device__ __forceinline__ int slot_list_c(const int* topK, long long a, int s, int K, int P) {
    int k = s >> 1;
    return (s & 1) ? P + topK[a * 2 * K + K + k] : topK[a * 2 * K + k];
}

// The W mask words of entry t (one 16-byte load when W = 4).
__device__ __forceinline__ void entry_words(const uint32_t* E, size_t t, int W, uint32_t (&m)[kMaxW]) {
    if (W == 4) {
        const uint4 v = ((const uint4*)E)[t];
        m[0] = v.x; m[1] = v.y; m[2] = v.z; m[3] = v.w;
#pragma unroll
        for (int w = 4; w < kMaxW; ++w) m[w] = 0u;
    } else {
#pragma unroll
        for (int w = 0; w < kMaxW; ++w) m[w] = w < W ? E[t * W + w] : 0u;
    }
}

// One workgroup per 256 consecutive points.  |O(a)|: the E words of the
// block's points are contiguous, so the workgroup reads them as coalesced
// 16-byte vectors and adds each vector's popcount to its point's LDS counter.
// |R(a)| and the mutual pairs only exist for points that sit in a list.
// Each kernel body below takes its ClusterArgs and block index, so that the
// single-dataset kernels and the batched set kernels (k_sets) share it.
__device__ __forceinline__ void degree_block(const ClusterArgs& c, long long blk) {
    __shared__ int own[256];
    const long long a0 = blk * 256;
    const int npt = (int)min((long long)256, c.n - a0);
    const int SW = c.S * c.W;
    own[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t* Eb = c.E + (size_t)a0 * SW;
    if ((SW & 15) == 0) {
        // four lanes per point, each summing a quarter of the point's E row
        // (consecutive lanes read consecutive 16-byte vectors), then a
        // reduction over the four lanes: no conflicting LDS atomics
        const int per = SW >> 4;
        const int part = threadIdx.x & 3;
        for (int pt = threadIdx.x >> 2; pt < 256; pt += 64) {
            int pc = 0;
            if (pt < npt) {
                const uint4* v = (const uint4*)(Eb + (size_t)pt * SW) + part * per;
                for (int k = 0; k < per; ++k) {
                    const uint4 e = v[k];
                    pc += __popc(e.x) + __popc(e.y) + __popc(e.z) + __popc(e.w);
                }
            }
            pc += __shfl_xor(pc, 1, kWave);
            pc += __shfl_xor(pc, 2, kWave);
            if (part == 0 && pt < npt) own[pt] = pc;
        }
    } else if ((SW & 3) == 0) {
        const int per = SW >> 2;
        const int nv = npt * per;
        for (int v = threadIdx.x; v < nv; v += 256) {
            const uint4 e = ((const uint4*)Eb)[v];
            const int pc = __popc(e.x) + __popc(e.y) + __popc(e.z) + __popc(e.w);
            if (pc) atomicAdd(&own[v / per], pc);
        }
    } else {
        const int nw = npt * SW;
        for (int v = threadIdx.x; v < nw; v += 256) {
            const int pc = __popc(Eb[v]);
            if (pc) atomicAdd(&own[v / SW], pc);
        }
    }
    const long long a = a0 + threadIdx.x;
    const bool live = threadIdx.x < npt;
    const int h = live ? c.hub_id[a] : -1;
    __syncthreads();
    int deg = live ? own[threadIdx.x] : 0;
    // Points in a list (a few percent): the whole wave takes one such point at
    // a time, one lane per (slot, word) of its E row and per membership, so
    // that its loads are issued together rather than one after another.
    unsigned long long hubs = __ballot(h >= 0);
    const int lane = lane_id();
    while (hubs) {
        const int src = __builtin_ctzll(hubs);
        hubs &= hubs - 1;
        const long long ah = a - lane + src;
        const int hh = __builtin_amdgcn_readlane(h, src);
        const int hb = c.hub_ptr[hh], he = c.hub_ptr[hh + 1];
        int part = 0;
        for (int q = hb + lane; q < he; q += kWave) part += c.colsum[c.hub_mem[q]];  // |R(a)|
        const uint32_t* Ea = c.E + (size_t)ah * SW;
        for (int i = lane; i < SW; i += kWave) {
            const uint32_t e = Ea[i];
            if (!e) continue;
            const int s = i / c.W, w = i - s * c.W;
            const int l = slot_list_c(c.topK, ah, s, c.K, c.P);
            uint32_t m = 0;
            for (int q = hb; q < he; ++q) {
                const int L = c.hub_mem[q] / c.M;
                m |= c.inv[((size_t)l * c.L2 + L) * c.W + w];
            }
            part -= __popc(e & m);  // |O(a) n R(a)|
        }
        part = wave_sum_i32(part);
        if (lane == src) deg += part;
    }
    const bool is_core = live && deg >= c.minPts - 1;
    if (live) {
        c.degree[a] = deg;
        c.core[a] = is_core ? 1 : 0;
    }
    const unsigned long long b = __ballot(is_core);
    if (lane_id() == 0 && b) spread_add(c.n_core, c.spread, (unsigned long long)__popcll(b));
}
__global__ __launch_bounds__(256) void k_degree(ClusterArgs c) { degree_block(c, blockIdx.x); }

// corehub[l][w]: bit i set when the point at position i of list l is core.
__device__ __forceinline__ void corehub_block(const ClusterArgs& c, long long blk) {
    const long long t = blk * 256 + threadIdx.x;
    if (t >= (long long)c.L2 * c.W) return;
    int l = t / c.W, w = t % c.W;
    uint32_t m = 0;
    for (int b = 0; b < 32; ++b) {
        int i = w * 32 + b;
        if (i < c.M && c.core[c.lists[(size_t)l * c.M + i]]) m |= 1u << b;
    }
    c.corehub[t] = m;
}
__global__ __launch_bounds__(256) void k_corehub(ClusterArgs c) { corehub_block(c, blockIdx.x); }

// Union-find with parent[x] <= x: roots are hooked under the smaller root, so
// every component ends with its smallest member as root.
__device__ __forceinline__ int uf_find(int* par, int x) {
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

__device__ __forceinline__ void uf_union(int* par, int a, int b) {
    int ra = uf_find(par, a), rb = uf_find(par, b);
    while (ra != rb) {
        if (ra < rb) { int t = ra; ra = rb; rb = t; }
        int old = atomicCAS(&par[ra], ra, rb);
        if (old == ra) return;
        ra = uf_find(par, old);
        rb = uf_find(par, rb);
    }
}

__device__ __forceinline__ void uf_init_block(const ClusterArgs& c, long long blk) {
    const long long a = blk * 256 + threadIdx.x;
    if (a >= c.n) return;
    c.parent[a] = (int)a;
    c.border_min[a] = 0x7fffffff;
}
__global__ __launch_bounds__(256) void k_uf_init(ClusterArgs c) { uf_init_block(c, blockIdx.x); }

// Per core point, slot by slot: all core hits of the slot's list are
// connected through the point.  The point is joined to the first core hit;
// the later ones are recorded as star edges from the first (adj[l][first])
// for the list-local unions of k_uf_lists.
// The star edges of core point a: per slot, the first core hit (after
// masking by corehub) is joined to a, and the other core hits of the slot are
// recorded as adjacent to that first one (adj, joined by k_uf_lists).
// words(s, m) yields the E words of slot s.
template <typename Words>
__device__ __forceinline__ void uf_point(const ClusterArgs& c, long long a, Words words) {
    int ra = uf_find(c.parent, (int)a);
    for (int s = 0; s < c.S; ++s) {
        const int l = slot_list_c(c.topK, a, s, c.K, c.P);
        uint32_t m[kMaxW];
        words(s, m);
        int first = -1;
#pragma unroll
        for (int w = 0; w < kMaxW; ++w) {
            if (w < c.W && m[w]) m[w] &= c.corehub[(size_t)l * c.W + w];
            if (first < 0 && m[w]) first = w * 32 + __ffs(m[w]) - 1;
        }
        if (first < 0) continue;
#pragma unroll
        for (int w = 0; w < kMaxW; ++w)
            if (w == (first >> 5)) m[w] &= ~(1u << (first & 31));
        uint32_t* ad = c.adj + ((size_t)l * c.M + first) * c.W;
#pragma unroll
        for (int w = 0; w < kMaxW; ++w)
            if (w < c.W && m[w] && (load_relaxed((const int*)&ad[w]) & m[w]) != m[w]) atomicOr(&ad[w], m[w]);
        const int x = c.lists[(size_t)l * c.M + first];
        if (uf_find(c.parent, x) != ra) {
            uf_union(c.parent, (int)a, x);
            ra = uf_find(c.parent, (int)a);
        }
    }
}

__device__ __forceinline__ void uf_points_block(const ClusterArgs& c, long long blk) {
    const long long a = blk * 256 + threadIdx.x;
    if (a >= c.n || !c.core[a]) return;
    uf_point(c, a, [&](int s, uint32_t (&m)[kMaxW]) { entry_words(c.E, (size_t)a * c.S + s, c.W, m); });
}
__global__ __launch_bounds__(256) void k_uf_points(ClusterArgs c) { uf_points_block(c, blockIdx.x); }

// The E rows of the workgroup's 128 points (S x W words each) are staged
// through LDS with coalesced loads (16-byte vectors when S W is a multiple of
// 4), instead of each thread striding through its own row in global memory
// across the slot loop.  LDS: 128 x S x W x 4 bytes.
__device__ __forceinline__ void uf_points_staged_block(const ClusterArgs& c, long long blk) {
    extern __shared__ __align__(16) uint32_t erow[];
    const int SW = c.S * c.W;
    const long long a0 = blk * 128;
    const int npt = (int)min((long long)128, c.n - a0);
    const uint32_t* Eb = c.E + (size_t)a0 * SW;
    if ((SW & 3) == 0) {
        const int nv = npt * (SW >> 2);
        for (int v = threadIdx.x; v < nv; v += 128) ((uint4*)erow)[v] = ((const uint4*)Eb)[v];
    } else {
        for (int v = threadIdx.x; v < npt * SW; v += 128) erow[v] = Eb[v];
    }
    __syncthreads();
    const long long a = a0 + threadIdx.x;
    if ((int)threadIdx.x >= npt || !c.core[a]) return;
    const uint32_t* mine = erow + (size_t)threadIdx.x * SW;
    if (c.W == 4) {
        uf_point(c, a, [&](int s, uint32_t (&m)[kMaxW]) {
            const uint4 v = ((const uint4*)mine)[s];
            m[0] = v.x; m[1] = v.y; m[2] = v.z; m[3] = v.w;
#pragma unroll
            for (int w = 4; w < kMaxW; ++w) m[w] = 0u;
        });
    } else {
        const int W = c.W;
        uf_point(c, a, [&](int s, uint32_t (&m)[kMaxW]) {
#pragma unroll
            for (int w = 0; w < kMaxW; ++w) m[w] = w < W ? mine[s * W + w] : 0u;
        });
    }
}
__global__ __launch_bounds__(128) void k_uf_points_staged(ClusterArgs c) { uf_points_staged_block(c, blockIdx.x); }

__device__ __forceinline__ void uf_lists_block(const ClusterArgs& c, long long blk) {
    const long long tl = blk * 256 + threadIdx.x;
    if (tl >= (long long)c.L2 * c.M) return;
    const int t = (int)tl;
    const int l = t / c.M;
    const int f = c.lists[t];
    for (int w = 0; w < c.W; ++w) {
        uint32_t m = c.adj[(size_t)t * c.W + w];
        while (m) {
            int b = __ffs(m) - 1;
            m &= m - 1;
            uf_union(c.parent, f, c.lists[(size_t)l * c.M + w * 32 + b]);
        }
    }
}
__global__ __launch_bounds__(256) void k_uf_lists(ClusterArgs c) { uf_lists_block(c, blockIdx.x); }

__device__ __forceinline__ void uf_compress_block(const ClusterArgs& c, long long blk) {
    const long long a = blk * 256 + threadIdx.x;
    if (a >= c.n) return;
    if (c.core[a]) c.parent[a] = uf_find(c.parent, (int)a);
}
__global__ __launch_bounds__(256) void k_uf_compress(ClusterArgs c) { uf_compress_block(c, blockIdx.x); }

__device__ __forceinline__ void root_flags_block(const ClusterArgs& c, int* flags, long long blk) {
    const long long a = blk * 256 + threadIdx.x;
    if (a >= c.n) return;
    flags[a] = (c.core[a] && c.parent[a] == (int)a) ? 1 : 0;
}
__global__ __launch_bounds__(256) void k_root_flags(ClusterArgs c, int* flags) { root_flags_block(c, flags, blockIdx.x); }

// Border points: the smallest root among adjacent core points, from both
// directions of every hit pair.
__device__ __forceinline__ void border_block(const ClusterArgs& c, long long blk) {
    const long long t = blk * 256 + threadIdx.x;
    if (t >= c.n * c.S) return;
    const long long a = t / c.S;
    const int s = (int)(t % c.S);
    const int l = slot_list_c(c.topK, a, s, c.K, c.P);
    const bool ca = c.core[a] != 0;
    int best = 0x7fffffff;
    uint32_t ew[kMaxW];
    entry_words(c.E, (size_t)t, c.W, ew);
    for (int w = 0; w < c.W; ++w) {
        const uint32_t e = ew[w];
        if (!e) continue;
        const uint32_t ch = c.corehub[(size_t)l * c.W + w];
        uint32_t m = ca ? (e & ~ch) : (e & ch);
        while (m) {
            const int b = __ffs(m) - 1;
            m &= m - 1;
            const int cp = c.lists[(size_t)l * c.M + w * 32 + b];
            if (ca) {
                const int r = c.parent[a];
                if (load_relaxed(&c.border_min[cp]) > r) atomicMin(&c.border_min[cp], r);
            } else {
                const int r = c.parent[cp];
                best = r < best ? r : best;
            }
        }
    }
    if (!ca && best != 0x7fffffff && load_relaxed(&c.border_min[a]) > best) atomicMin(&c.border_min[a], best);
}
__global__ __launch_bounds__(256) void k_border(ClusterArgs c) { border_block(c, blockIdx.x); }

__device__ __forceinline__ void labels_block(const ClusterArgs& c, const int* rank, long long blk) {
    const long long a = blk * 256 + threadIdx.x;
    if (a >= c.n) return;
    int lab;
    if (c.core[a]) lab = rank[c.parent[a]];
    else lab = c.border_min[a] == 0x7fffffff ? -1 : rank[c.border_min[a]];
    c.labels[a] = lab;
}
__global__ __launch_bounds__(256) void k_labels(ClusterArgs c, const int* rank) { labels_block(c, rank, blockIdx.x); }

// Core test of new minPts on degrees already counted for the same eps.
__device__ __forceinline__ void core_minpts_block(const int* __restrict__ degree, long long n, int minPts,
                                                  uint8_t* __restrict__ core, unsigned long long* __restrict__ n_core,
                                                  int spread, long long blk) {
    const long long a = blk * 256 + threadIdx.x;
    const bool is_core = a < n && degree[a] >= minPts - 1;
    if (a < n) core[a] = is_core ? 1 : 0;
    const unsigned long long b = __ballot(is_core);
    if (lane_id() == 0 && b) spread_add(n_core, spread, (unsigned long long)__popcll(b));
}
__global__ __launch_bounds__(256) void k_core_minpts(const int* __restrict__ degree, long long n, int minPts,
                                                     uint8_t* __restrict__ core,
                                                     unsigned long long* __restrict__ n_core, int spread) {
    core_minpts_block(degree, n, minPts, core, n_core, spread, blockIdx.x);
}

__global__ __launch_bounds__(64) void k_spread_fold(const unsigned long long* __restrict__ ctr,
                                                    unsigned long long* __restrict__ out) {
    unsigned long long v = ctr[(size_t)threadIdx.x * kSpreadStride];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    if (threadIdx.x == 0) *out += v;
}

static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

int launch_core_minpts(const int* degree, long long n, int minPts, uint8_t* core, unsigned long long* n_core,
                       int spread, hipStream_t st) {
    hipLaunchKernelGGL(k_core_minpts, dim3(nblk(n, 256)), dim3(256), 0, st, degree, n, minPts, core, n_core, spread);
    return (int)hipGetLastError();
}
int launch_spread_fold(const unsigned long long* ctr, unsigned long long* out, hipStream_t st) {
    static_assert(kSpread == kWave, "one lane per copy");
    hipLaunchKernelGGL(k_spread_fold, dim3(1), dim3(kSpread), 0, st, ctr, out);
    return (int)hipGetLastError();
}

int launch_degree(const ClusterArgs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_degree, dim3(nblk(c.n, 256)), dim3(256), 0, st, c);
    return (int)hipGetLastError();
}
int launch_corehub(const ClusterArgs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_corehub, dim3(nblk((long long)c.L2 * c.W, 256)), dim3(256), 0, st, c);
    return (int)hipGetLastError();
}
int launch_uf_init(const ClusterArgs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_uf_init, dim3(nblk(c.n, 256)), dim3(256), 0, st, c);
    return (int)hipGetLastError();
}
int launch_uf_points(const ClusterArgs& c, hipStream_t st) {
    const size_t sm = (size_t)128 * c.S * c.W * 4;
    if (sm <= 64 * 1024) {
        hipLaunchKernelGGL(k_uf_points_staged, dim3(nblk(c.n, 128)), dim3(128), sm, st, c);
    } else {
        hipLaunchKernelGGL(k_uf_points, dim3(nblk(c.n, 256)), dim3(256), 0, st, c);
    }
    return (int)hipGetLastError();
}
int launch_uf_lists(const ClusterArgs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_uf_lists, dim3(nblk((long long)c.L2 * c.M, 256)), dim3(256), 0, st, c);
    return (int)hipGetLastError();
}
int launch_uf_compress(const ClusterArgs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_uf_compress, dim3(nblk(c.n, 256)), dim3(256), 0, st, c);
    return (int)hipGetLastError();
}
int launch_root_flags(const ClusterArgs& c, int* flags, hipStream_t st) {
    hipLaunchKernelGGL(k_root_flags, dim3(nblk(c.n, 256)), dim3(256), 0, st, c, flags);
    return (int)hipGetLastError();
}
int launch_border(const ClusterArgs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_border, dim3(nblk(c.n * c.S, 256)), dim3(256), 0, st, c);
    return (int)hipGetLastError();
}
int launch_labels(const ClusterArgs& c, const int* rank, hipStream_t st) {
    hipLaunchKernelGGL(k_labels, dim3(nblk(c.n, 256)), dim3(256), 0, st, c, rank);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched chain (ChainSet): one launch per step for all sets.

// Degrees of all eps of a dataset (x is its first set): per eps the own hits
// are counted from the coalesced E rows as in degree_block; for the points in
// a list, the mutual-pair mask of every (slot, word) is built once and applied
// to every eps's E words.
__device__ __forceinline__ void degree_eps_block(const ChainSet& x, long long blk) {
    __shared__ int own[kSweepMaxEps][256];
    const ClusterArgs& c = x.a;
    const int ne = x.n_eps, nm = x.n_minpts;
    const long long a0 = blk * 256;
    const int npt = (int)min((long long)256, c.n - a0);
    const int SW = c.S * c.W;
    const size_t es = (size_t)c.n * SW, cs = (size_t)c.L2 * c.M;
    for (int e = 0; e < ne; ++e) own[e][threadIdx.x] = 0;
    __syncthreads();
    for (int e = 0; e < ne; ++e) {
        const uint32_t* Eb = c.E + e * es + (size_t)a0 * SW;
        if ((SW & 3) == 0) {
            const int per = SW >> 2;
            for (int v = threadIdx.x; v < npt * per; v += 256) {
                const uint4 w4 = ((const uint4*)Eb)[v];
                const int pc = __popc(w4.x) + __popc(w4.y) + __popc(w4.z) + __popc(w4.w);
                if (pc) atomicAdd(&own[e][v / per], pc);
            }
        } else {
            for (int v = threadIdx.x; v < npt * SW; v += 256) {
                const int pc = __popc(Eb[v]);
                if (pc) atomicAdd(&own[e][v / SW], pc);
            }
        }
    }
    __syncthreads();
    const long long a = a0 + threadIdx.x;
    const bool live = threadIdx.x < npt;
    int dx[kSweepMaxEps];
#pragma unroll
    for (int e = 0; e < kSweepMaxEps; ++e) dx[e] = 0;  // rev - mutual per eps
    if (live) {
        const int h = c.hub_id[a];
        if (h >= 0) {
            const int hb = c.hub_ptr[h], he = c.hub_ptr[h + 1];
            for (int q = hb; q < he; ++q) {
                const int pos = c.hub_mem[q];
#pragma unroll
                for (int e = 0; e < kSweepMaxEps; ++e)
                    if (e < ne) dx[e] += c.colsum[e * cs + pos];
            }
            for (int s = 0; s < c.S; ++s) {
                const int l = slot_list_c(c.topK, a, s, c.K, c.P);
                for (int w = 0; w < c.W; ++w) {
                    uint32_t m = 0;
                    for (int q = hb; q < he; ++q) {
                        const int L = c.hub_mem[q] / c.M;
                        m |= c.inv[((size_t)l * c.L2 + L) * c.W + w];
                    }
                    if (!m) continue;
                    const uint32_t* Ea = c.E + (size_t)a * SW + s * c.W + w;
#pragma unroll
                    for (int e = 0; e < kSweepMaxEps; ++e)
                        if (e < ne) dx[e] -= __popc(Ea[e * es] & m);
                }
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kSweepMaxEps; ++e) {
        if (e >= ne) break;
        const int deg = live ? own[e][threadIdx.x] + dx[e] : 0;
        const bool is_core = live && deg >= c.minPts - 1;
        if (live) {
            c.degree[(size_t)e * c.n + a] = deg;
            c.core[(size_t)e * nm * c.n + a] = is_core ? 1 : 0;
        }
        const unsigned long long b = __ballot(is_core);
        if (lane_id() == 0 && b) atomicAdd(c.n_core + (size_t)e * nm, (unsigned long long)__popcll(b));
    }
}

// The set of workgroup b: the largest s with pre[s] <= b (a set without
// workgroups in this step has pre[s] == pre[s + 1] and is never chosen).
__device__ __forceinline__ int set_of_block(const unsigned* __restrict__ pre, int nset, unsigned b) {
    int lo = 0, hi = nset;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int OP, int BS>
__global__ __launch_bounds__(BS) void k_sets(const ChainSet* __restrict__ sets, const unsigned* __restrict__ pre,
                                             int nset) {
    const int s = set_of_block(pre, nset, blockIdx.x);
    const long long blk = (long long)(blockIdx.x - pre[s]);
    const ChainSet& x = sets[s];
    const ClusterArgs& c = x.a;
    if constexpr (OP == kChDegree) {
        degree_eps_block(x, blk);
    } else if constexpr (OP == kChCore) {
        core_minpts_block(c.degree, c.n, c.minPts, c.core, c.n_core, c.spread, blk);
    } else if constexpr (OP == kChAdjClear) {
        const long long i = blk * BS + threadIdx.x;
        if (i < (long long)c.L2 * c.M * c.W) c.adj[i] = 0u;
    } else if constexpr (OP == kChCorehub) {
        corehub_block(c, blk);
    } else if constexpr (OP == kChUfInit) {
        uf_init_block(c, blk);
    } else if constexpr (OP == kChUfPoints) {
        uf_points_block(c, blk);
    } else if constexpr (OP == kChUfPointsStaged) {
        uf_points_staged_block(c, blk);
    } else if constexpr (OP == kChUfLists) {
        uf_lists_block(c, blk);
    } else if constexpr (OP == kChUfCompress) {
        uf_compress_block(c, blk);
    } else if constexpr (OP == kChRootFlags) {
        root_flags_block(c, x.rootf, blk);
    } else if constexpr (OP == kChScanBlocks) {
        __shared__ int wsum[16];
        __shared__ int tot;
        const long long i = blk * 1024 + threadIdx.x;
        (void)block_scan_excl_1024(i < c.n ? x.rootf[i] : 0, wsum, &tot);
        if (threadIdx.x == 0) x.bsum[blk] = tot;
    } else if constexpr (OP == kChScanTop) {
        __shared__ int wsum[16];
        __shared__ int tot;
        const int nb = (int)((c.n + 1023) / 1024);
        int carry = 0;
        for (int b0 = 0; b0 < nb; b0 += 1024) {
            const int i = b0 + (int)threadIdx.x;
            const int v = i < nb ? x.bsum[i] : 0;
            const int ex = block_scan_excl_1024(v, wsum, &tot);
            if (i < nb) x.bsum[i] = carry + ex;
            carry += tot;
            __syncthreads();
        }
        if (threadIdx.x == 0) *x.total = carry;
    } else if constexpr (OP == kChScanApply) {
        __shared__ int wsum[16];
        __shared__ int tot;
        const long long i = blk * 1024 + threadIdx.x;
        const int ex = block_scan_excl_1024(i < c.n ? x.rootf[i] : 0, wsum, &tot);
        if (i < c.n) x.rank[i] = x.bsum[blk] + ex;
    } else if constexpr (OP == kChBorder) {
        border_block(c, blk);
    } else if constexpr (OP == kChLabels) {
        labels_block(c, x.rank, blk);
    }
}

static constexpr int chain_bs(int op) {
    return op == kChUfPointsStaged ? 128 : (op == kChScanBlocks || op == kChScanTop || op == kChScanApply) ? 1024 : 256;
}
int chain_block_threads(int op) { return chain_bs(op); }

long long chain_set_blocks(int op, const ChainSet& x) {
    const ClusterArgs& a = x.a;
    const long long n = a.n;
    const long long bs = chain_block_threads(op);
    auto cdiv = [](long long u, long long v) { return (u + v - 1) / v; };
    switch (op) {
        case kChDegree: return x.eps_i == 0 && x.minpts_i == 0 ? cdiv(n, bs) : 0;
        case kChCore: return x.minpts_i > 0 ? cdiv(n, bs) : 0;
        case kChAdjClear: return cdiv((long long)a.L2 * a.M * a.W, bs);
        case kChCorehub: return cdiv((long long)a.L2 * a.W, bs);
        case kChUfLists: return cdiv((long long)a.L2 * a.M, bs);
        case kChScanTop: return 1;
        case kChBorder: return cdiv(n * a.S, bs);
        default: return cdiv(n, bs);
    }
}

template <int OP>
static int launch_op(const ChainSet* sets, const unsigned* pre, int nset, unsigned grid, size_t lds, hipStream_t st) {
    constexpr int BS = chain_bs(OP);
    hipLaunchKernelGGL((k_sets<OP, BS>), dim3(grid), dim3(BS), OP == kChUfPointsStaged ? lds : 0, st, sets, pre, nset);
    return (int)hipGetLastError();
}

int launch_chain_op(int op, const ChainSet* sets, const unsigned* pre, int nset, unsigned grid, size_t lds_staged,
                    hipStream_t st) {
    if (grid == 0 || nset <= 0) return 0;
    if (op == kChUfPointsStaged && lds_staged > 64 * 1024) return -2;
    switch (op) {
        case kChDegree: return launch_op<kChDegree>(sets, pre, nset, grid, lds_staged, st);
        case kChCore: return launch_op<kChCore>(sets, pre, nset, grid, lds_staged, st);
        case kChAdjClear: return launch_op<kChAdjClear>(sets, pre, nset, grid, lds_staged, st);
        case kChCorehub: return launch_op<kChCorehub>(sets, pre, nset, grid, lds_staged, st);
        case kChUfInit: return launch_op<kChUfInit>(sets, pre, nset, grid, lds_staged, st);
        case kChUfPoints: return launch_op<kChUfPoints>(sets, pre, nset, grid, lds_staged, st);
        case kChUfPointsStaged: return launch_op<kChUfPointsStaged>(sets, pre, nset, grid, lds_staged, st);
        case kChUfLists: return launch_op<kChUfLists>(sets, pre, nset, grid, lds_staged, st);
        case kChUfCompress: return launch_op<kChUfCompress>(sets, pre, nset, grid, lds_staged, st);
        case kChRootFlags: return launch_op<kChRootFlags>(sets, pre, nset, grid, lds_staged, st);
        case kChScanBlocks: return launch_op<kChScanBlocks>(sets, pre, nset, grid, lds_staged, st);
        case kChScanTop: return launch_op<kChScanTop>(sets, pre, nset, grid, lds_staged, st);
        case kChScanApply: return launch_op<kChScanApply>(sets, pre, nset, grid, lds_staged, st);
        case kChBorder: return launch_op<kChBorder>(sets, pre, nset, grid, lds_staged, st);
        case kChLabels: return launch_op<kChLabels>(sets, pre, nset, grid, lds_staged, st);
        default: return -3;
    }
}

// ---------------------------------------------------------------------------
// Exclusive prefix sums (hipcub) and hub ids.
size_t scan_tmp_bytes(long long n) {
    size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int*)nullptr, (int*)nullptr, (int)n);
    return bytes;
}

__global__ void k_total(const int* in, const int* excl, long long n, int* total) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *total = n > 0 ? excl[n - 1] + in[n - 1] : 0;
}

int scan_exclusive_i32(const int* in, int* out, long long n, int* total_out_dev, void* tmp,
                       size_t tmp_bytes, hipStream_t st) {
    if (n <= 0) return 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, (int)n, st);
    if (e != hipSuccess) return (int)e;
    if (total_out_dev) hipLaunchKernelGGL(k_total, dim3(1), dim3(1), 0, st, in, out, n, total_out_dev);
    return (int)hipGetLastError();
}

__global__ void k_hub_flags(const uint8_t* is_hub, int* flags, long long n) {
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) flags[a] = is_hub[a] ? 1 : 0;
}
__global__ void k_hub_id(const uint8_t* is_hub, const int* excl, int* hub_id, long long n) {
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) hub_id[a] = is_hub[a] ? excl[a] : -1;
}
int launch_is_hub_scan_prep(const uint8_t* is_hub, int* flags, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_hub_flags, dim3(nblk(n, 256)), dim3(256), 0, st, is_hub, flags, n);
    return (int)hipGetLastError();
}
int launch_hub_id(const uint8_t* is_hub, const int* excl, int* hub_id, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_hub_id, dim3(nblk(n, 256)), dim3(256), 0, st, is_hub, excl, hub_id, n);
    return (int)hipGetLastError();
}

}  // namespace sdb
