// sdb_ctx.hpp -- the C ABI's context (device buffers, stream, state of the
// pipeline) and the host helpers shared by the entry points (sdb_capi.hip)
// and the batched sweep (sdb_sweep.hip).  Private to the library.
#pragma once
#include <hip/hip_runtime.h>
#include <sys/mman.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sdbscan_amd.h"
#include "sdb_kernels.hpp"

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool in_arena = false;  // carved from a DevArena: never freed on its own
};

// Bump allocator over large device chunks, shared by the contexts of one
// batched sweep, so that their many buffers cost a few hipMalloc / hipFree
// calls (each hipFree waits for the device) instead of one each.  Buffers
// are not returned one by one (a grown buffer leaves its old space behind);
// the chunks go with the arena.  Thread-safe.
struct DevArena {
    static constexpr size_t kChunk = size_t(1) << 30;
    std::mutex mu;
    std::vector<void*> chunks;
    char* cur = nullptr;
    size_t left = 0;
    hipError_t alloc(size_t bytes, void** out) {
        bytes = (bytes + 255) & ~size_t(255);
        std::lock_guard<std::mutex> g(mu);
        if (bytes > kChunk / 4) {  // large buffers get a chunk of their own
            void* p = nullptr;
            const hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess) return e;
            chunks.push_back(p);
            *out = p;
            return hipSuccess;
        }
        if (bytes > left) {
            void* p = nullptr;
            const hipError_t e = hipMalloc(&p, kChunk);
            if (e != hipSuccess) return e;
            chunks.push_back(p);
            cur = (char*)p;
            left = kChunk;
        }
        *out = cur;
        cur += bytes;
        left -= bytes;
        return hipSuccess;
    }
    ~DevArena() {
        for (void* p : chunks) (void)hipFree(p);
    }
};

struct sdb_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[8];
    std::string err;
    sdb_params prm{};
    bool have_params = false;
    // input
    const float* X = nullptr;  // device, n x d
    DevBuf Xown;
    long long n = 0;
    int d = 0;
    int ldx = 0;
    // transformed
    DevBuf Xn, Xb;
    DevBuf emb;  // L1/L2 embeddings (n x fhtDim) of the index passes
    // index
    DevBuf bits, R, topK, tie_idx, tie_vals, tie_rows, tie_count, cmax, cmin, thi, tlo, cand_cnt, cand_key,
        cand_idx, topM, flags, lists, is_hub, hub_flags, hub_excl, hub_id, hub_cnt, hub_ptr,
        hub_cur, hub_mem, inv, pos_hub, pcnt, poff, pairs, pseg, F, wgcnt, wgoff, grp_ptr, grp, scan_tmp, total, err_flag,
        row_slot, row_buf;
    bool indexed = false;
    int H = 0, L2 = 0, W = 0, S = 0;
    int Mt = 0;        // candidates per list in the distance tiles (topM, or the sample chunk)
    bool sng = false;  // neighbourhoods from a sample sequence (sngDbscan / sngOptics)
    int sng_flags = 0, sngL = 0;
    long long qblk = 0;
    int nblk = 1;
    std::vector<int> h_grp_ptr;
    // find_core
    DevBuf E, colsum, queue, qcount, tiles, degree, core, stats64;
    // the queued pairs sorted by candidate point for the recheck: two key
    // buffers, the alternate value buffer and the sort's scratch
    DevBuf qkeys, qvals, qsort_tmp;
    // dense list intersections (inter[l][L], W words) when lists share many points
    DevBuf inter;
    DevBuf core_spread;  // kSpread copies of the core count (spread_add)
    DevBuf ffull;        // n x S: 1 where the first-occurrence mask is the full one
    // squared row norms for the L2 MFMA tiles and their maximum; whether the
    // last distance pass used those tiles (the accuracy sample follows it)
    DevBuf rnorm2, rnorm2_max;
    bool l2x_used = false;
    // the rows the norms were computed from (cleared whenever the rows are
    // transformed again) and their largest norm
    const float* rnorm2_of = nullptr;
    float rnorm2_maxv = 0.0f;
    long long qcap = 0;
    int tile_rows = 0;   // distance tile rows (0: the default for the row length)
    bool cored = false;
    float eps = 0.f;
    int minPts = 0;
    // clusters
    DevBuf corehub, parent, border, labels, adj, rootf, rank;
    // neighbourhoods / optics
    DevBuf nb_deg, nb_ptr, nb_cur, nb_idx, nb_dist, nb_sorted, core_dist, pos, reach;
    bool have_nb = false;
    bool clustered = false;
    // noise labelling
    DevBuf ns_ids, cproj, noise, nbest;
    // projections read back for listed points
    DevBuf pj_ids, pj_buf;
    // sample-sequence neighbourhoods
    DevBuf occ_ptr, occ_pos, win, rev, csr_cnt, csr_ptr, csr_cur, csr_idx;
    bool transformed = false;
    int transformed_prec = -1;
    // BF16 rounding bound of every row (bf16_row_err) and its maximum
    DevBuf rowerr, rowerr_max;
    float rho_max = 0.f;
    int sng_Ms = 64;
    sdb_stats stats{};
    void* scratch = nullptr;  // this device's address of the static scratch array
    DevArena* arena = nullptr;  // when set, buffers are carved from it (batched sweep)
    // batched sweep: hit masks and column counts of every eps, labels and
    // component counts of every (eps, minPts)
    DevBuf sw_E, sw_colsum, sw_labels, sw_comps;
};

static inline int fail(sdb_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
    return code;
}

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(c, SDB_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                                \
    } while (0)

#define LCHK(expr)                                                                       \
    do {                                                                                 \
        int _r = (expr);                                                                 \
        if (_r != 0)                                                                     \
            return fail(c, _r == -2 || _r == -3 ? SDB_EUNSUPPORTED : SDB_EHIP,           \
                        "launch %s failed (%d)%s", #expr, _r,                            \
                        _r > 0 ? hipGetErrorString((hipError_t)_r) : "");                \
    } while (0)

// Frees a buffer the context allocated itself (arena buffers stay with the
// arena).
static inline void release(DevBuf& b) {
    if (b.p && !b.in_arena) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    b.in_arena = false;
}

static inline int ensure(sdb_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return 0;
    release(b);
    const hipError_t e = c->arena ? c->arena->alloc(bytes, &b.p) : hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return fail(c, SDB_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    b.bytes = bytes;
    b.in_arena = c->arena != nullptr;
    return 0;
}

#define ENSURE(buf, bytes)                    \
    do {                                      \
        int _r = ensure(c, (buf), (bytes));   \
        if (_r) return _r;                    \
    } while (0)

template <typename T>
static inline T* P_(DevBuf& b) { return (T*)b.p; }

static inline float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}


// Error band of the fast distance, derived from the standard bound
// |fl(sum) - sum| <= gamma_k * sum |terms| (any summation order, one rounding of
// at most 2^-23 relative per operation), with a safety factor of 2.
// Total order on float bit patterns (host side of the device ord_key).
static inline uint32_t h_key(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static inline float h_unkey(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Cuts on the raw cosine dot s equivalent to the band tests on the distance
// fl(1 - s): fl(1 - s) is non-increasing in s, so "fl(1 - s) < lo" holds
// exactly for s > s_hit and "not fl(1 - s) > hi" exactly for s >= s_near.
// Found by bisection over the ordered finite floats.
static inline void cosine_cuts(double lo, double hi, float* s_hit, float* s_near) {
    auto is_hit = [&](float s) { return (double)(1.0f - s) < lo; };
    auto is_near = [&](float s) { return !((double)(1.0f - s) > hi); };
    auto bisect = [](auto pred, float* last_false, float* first_true) {
        if (pred(-FLT_MAX)) { *last_false = -INFINITY; *first_true = -INFINITY; return; }
        if (!pred(FLT_MAX)) { *last_false = FLT_MAX; *first_true = INFINITY; return; }
        uint32_t a = h_key(-FLT_MAX), b = h_key(FLT_MAX);  // pred(a) false, pred(b) true
        while (b - a > 1) {
            const uint32_t m = a + (b - a) / 2;
            if (pred(h_unkey(m))) b = m;
            else a = m;
        }
        *last_false = h_unkey(a);
        *first_true = h_unkey(b);
    };
    float lf, ft;
    bisect(is_hit, &lf, &ft);
    *s_hit = lf == FLT_MAX ? INFINITY : lf;
    bisect(is_near, &lf, &ft);
    *s_near = ft;
}

// Chi2 on BF16 rows (prec BF16): the tiles compute BC~, the dot of the BF16
// square roots of the sum-normalised rows (their Bhattacharyya coefficient).
// Term by term 2xy/(x+y) <= sqrt(xy), so chi2(x, y) = 1 - 2 sum xy/(x+y) >=
// 1 - BC: a pair with 1 - BC~ > eps + band has chi2 > eps.  And term by term
// 2 sqrt(xy) - 2xy/(x+y) <= (x+y)/2 (the difference is (x+y-2 sqrt(xy))^2 /
// (2(x+y))), so with sum x = sum y = 1 (within 4u after the transform)
// chi2 <= 2(1 - BC) + 4u: a pair with 2(1 - BC~ + band) + 1e-6 < eps has
// chi2 < eps and is a hit on the tiles.  Every other pair goes to the recheck,
// which decides with the FP32 chi2 outside band32 and with the FP64 chi2
// inside it.
static inline void compute_bands(const sdb_params& q, int prec, float eps, double* band, double* band32) {
    const double u = std::ldexp(1.0, -23);
    const double d = q.n_features;
    const double Smax = 1.0 + 1e-5;  // sum |x_i y_i| <= ||x|| ||y|| for unit rows
    const double b32 = 2.0 * ((d + 2.0) * u * Smax + 2.0 * u);
    *band32 = b32;
    switch (q.metric) {
        case SDB_COSINE:
            if (prec == SDB_BF16) {
                const double ub = std::ldexp(1.0, -8);
                *band = 2.0 * ((2.0 * ub + ub * ub) * Smax + (d + 2.0) * u * Smax * (1.0 + 2.0 * ub) + 2.0 * u);
            } else {
                *band = b32;
            }
            break;
        case SDB_L2: *band = 2.0 * (double)eps * (d + 8.0) * u + 1e-30; break;
        case SDB_L1: *band = 2.0 * (double)eps * (d + 4.0) * u + 1e-30; break;
        case SDB_CHI2:
            if (prec == SDB_BF16) {
                // per factor: sqrtf (<= 1 ulp) then BF16 rounding, relative e
                const double e = std::ldexp(1.0, -8) + 3.0 * u;
                *band = 2.0 * ((2.0 * e + e * e) * Smax + (d + 2.0) * u * Smax * (1.0 + 2.0 * e) + 2.0 * u) + 1e-12;
                *band32 = 2.0 * (d + 8.0) * u;
            } else {
                *band = 2.0 * (d + 8.0) * u;
            }
            break;
        default: *band = 2.0 * (2.0 * d + 200.0) * u; break;
    }
}

// Whether the context's BF16 rows carry their rounding bound in their last
// padding element: whenever a row has padding (d < ldx).  Every writer of the
// BF16 rows and every tile that reads them follow this one rule.
static inline int bf16_rho_pad(const sdb_ctx* c) { return c->d < c->ldx ? 1 : 0; }

static inline sdb::ClusterArgs cluster_args(sdb_ctx* c) {
    const sdb_params& q = c->prm;
    sdb::ClusterArgs ca{};
    ca.n = c->n; ca.S = c->S; ca.W = c->W; ca.M = c->Mt; ca.P = q.n_proj; ca.K = q.topK; ca.L2 = c->L2;
    ca.topK = P_<int>(c->topK); ca.lists = P_<int>(c->lists); ca.E = P_<uint32_t>(c->E);
    ca.colsum = P_<int>(c->colsum); ca.hub_id = P_<int>(c->hub_id); ca.hub_ptr = P_<int>(c->hub_ptr);
    ca.hub_mem = P_<int>(c->hub_mem); ca.inv = P_<uint32_t>(c->inv); ca.minPts = c->minPts;
    ca.degree = P_<int>(c->degree); ca.core = P_<uint8_t>(c->core);
    ca.corehub = P_<uint32_t>(c->corehub); ca.parent = P_<int>(c->parent);
    ca.border_min = P_<int>(c->border); ca.labels = P_<int>(c->labels); ca.adj = P_<uint32_t>(c->adj);
    return ca;
}


// Host memory for the neighbourhood lists the sOPTICS walks read (ids, then
// distances): an anonymous mapping aligned to 2 MiB with transparent huge
// pages asked for and every page faulted in when it is mapped
// (MADV_POPULATE_WRITE; one store per 4 KiB page where the kernel lacks it).
// The device-to-host copies then write into resident pages (copies into
// pageable memory from several host threads wait on each other, so a page
// fault inside one delays the rest), and no pass zero-fills the lists first.
// Grows, never shrinks.
struct HostLists {
    void* base = nullptr;  // the mapping
    size_t mapped = 0;
    void* p = nullptr;  // 2 MiB-aligned start inside it
    size_t cap = 0;
    HostLists() = default;
    HostLists(const HostLists&) = delete;
    HostLists& operator=(const HostLists&) = delete;
    ~HostLists() { drop(); }
    void drop() {
        if (base) munmap(base, mapped);
        base = p = nullptr;
        mapped = cap = 0;
    }
    // true when at least `bytes` are mapped and resident
    bool reserve(size_t bytes) {
        if (bytes <= cap) return true;
        drop();
        const size_t huge = (size_t)2 << 20;
        const size_t want = (bytes + huge - 1) / huge * huge;
        void* m = mmap(nullptr, want + huge, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (m == MAP_FAILED) return false;
        base = m;
        mapped = want + huge;
        p = (void*)(((uintptr_t)m + huge - 1) & ~(uintptr_t)(huge - 1));
        cap = want;
        madvise(p, cap, MADV_HUGEPAGE);  // advisory: ignored where THP is off
#ifdef MADV_POPULATE_WRITE
        if (madvise(p, cap, MADV_POPULATE_WRITE) == 0) return true;
#endif
        for (size_t o = 0; o < cap; o += 4096) ((volatile char*)p)[o] = 0;
        return true;
    }
};

// The lists of a pool of host threads that run the walks of many contexts in
// turn (the batched runs): each walk takes the free lists that fit best and
// gives them back, so pages faulted in for one dataset serve the next.
struct HostListPool {
    std::mutex mu;
    std::vector<std::unique_ptr<HostLists>> all;
    std::vector<HostLists*> idle;
    // the idle lists of the smallest capacity >= bytes, else the largest idle
    // ones (grown by the caller's reserve), else new ones
    HostLists* take(size_t bytes) {
        std::lock_guard<std::mutex> g(mu);
        int fit = -1, big = -1;
        for (int i = 0; i < (int)idle.size(); ++i) {
            const size_t cp = idle[i]->cap;
            if (cp >= bytes && (fit < 0 || cp < idle[fit]->cap)) fit = i;
            if (big < 0 || cp > idle[big]->cap) big = i;
        }
        const int k = fit >= 0 ? fit : big;
        if (k >= 0) {
            HostLists* h = idle[k];
            idle.erase(idle.begin() + k);
            return h;
        }
        all.emplace_back(new HostLists());
        return all.back().get();
    }
    void give(HostLists* h) {
        std::lock_guard<std::mutex> g(mu);
        idle.push_back(h);
    }
    // unmaps every list (none may be taken)
    void clear() {
        std::lock_guard<std::mutex> g(mu);
        idle.clear();
        all.clear();
    }
};

// sdb_optics_orders with the host lists taken from `pool` (a private set when
// null).
__attribute__((visibility("hidden"))) int optics_orders(sdb_ctx* c, int n_minpts, const int* minPts,
                                                      int* orderings, float* reaches, HostListPool* pool);
