// Batched eps / minPts sweeps in one distance launch (sdb_batch_sweep).
//
// Reference: test_sDbscan (src/sDbscan.cpp:980-1039) builds one index and
// runs rp_findCorePoints + formCluster for every eps, recomputing the
// candidate distances each time; fit_sOptics (:1711-1769) runs sOPTICS.  This
// entry point takes a batch of such datasets (jobs) and:
//   1. builds every job's index (host worker threads, one context and HIP
//      stream per job);
//   2. computes the candidate distances of all jobs once, in one persistent
//      launch (k_sweep_tiles) fed by a device work queue of (job, list, rows)
//      tiles, deciding every eps of each job from the same dot products;
//      tile rows per row length are picked by timing that kernel (autotune);
//   3. re-decides the pairs inside the FP32 band of an eps in FP64 (one
//      launch over all jobs);
//   4. runs the degree -> core test -> union-find -> labels chain of every
//      (job, eps, minPts) set together, one launch per step over all sets
//      (each workgroup finds its set from per-step workgroup offsets), the
//      steps captured as one hipGraph (clears by kernels, no memset nodes);
//   5. sOPTICS at the largest eps on the host walk, as sdb_optics_orders.
// The decisions are the single-eps path's (same cuts, same FP64 recheck), so
// the outputs equal sdb_batch_run's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>
#include <vector>

#include "sdb_ctx.hpp"

using namespace sdb;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct JobState {
    sdb_ctx* c = nullptr;
    std::vector<int> order;  // eps indices, ascending eps
    SweepJob desc{};
    size_t lds = 0;
    int qb = 0;              // tile rows
    double ms = 0.0;
};

void set_job_error(sdb_batch_job& j, int code, const char* what, const sdb_ctx* c) {
    j.status = code;
    std::snprintf(j.error, sizeof(j.error), "%s: %s", what, c ? sdb_last_error(c) : "");
}

// Runs fn(k) for k in order on n_workers host threads.
template <typename F>
void for_jobs(const std::vector<int>& order, int n_workers, F fn) {
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= (int)order.size()) break;
            fn(order[k]);
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < n_workers; ++w) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
}

// The job's descriptor, eps cuts and distance-tile geometry (FP32 k-steps of
// 16 elements; the LDS holds the list's rows of one k-chunk, the point ids and
// a column count per eps and candidate).
int prepare_job(sdb_batch_job& j, JobState& s) {
    sdb_ctx* c = s.c;
    const sdb_params& q = c->prm;
    const long long n = c->n;
    const int M = c->Mt, W = c->W, S = c->S, L2 = c->L2, ne = j.n_eps;
    s.order.resize(ne);
    std::iota(s.order.begin(), s.order.end(), 0);
    std::stable_sort(s.order.begin(), s.order.end(), [&](int a, int b) { return j.eps[a] < j.eps[b]; });
    const long long n_ent = n * S;
    ENSURE(c->sw_E, (size_t)ne * n_ent * W * 4);
    ENSURE(c->sw_colsum, (size_t)ne * L2 * M * 4);
    SweepJob& d = s.desc;
    d = SweepJob{};
    d.X = P_<float>(c->Xn);
    d.ldx = c->ldx;
    d.d = c->d;
    d.dk = (c->d + 15) / 16 * 16;
    const int steps = (int)(((kLdsBytes - (4 * ne + 4) * M - 64) / M - 16) / 64);
    if (steps <= 0) return fail(c, SDB_EUNSUPPORTED, "topM too large for the LDS tile");
    d.kc = std::min(d.dk, steps * 16);
    d.nchunk = (d.dk + d.kc - 1) / d.kc;
    d.M = M; d.W = W; d.S = S; d.L2 = L2;
    d.grp = P_<uint32_t>(c->grp);
    d.lists = P_<int>(c->lists);
    d.F = P_<uint32_t>(c->F);
    d.E = P_<uint32_t>(c->sw_E);
    d.colsum = P_<int>(c->sw_colsum);
    d.n_ent = n_ent;
    d.magicS = ((1ull << 40) + (unsigned long long)S - 1) / (unsigned long long)S;
    d.n_eps = ne;
    for (int oe = 0; oe < ne; ++oe) {
        const float eps = j.eps[s.order[oe]];
        double band = 0, band32 = 0;
        compute_bands(q, SDB_FP32, eps, &band, &band32);
        d.eps[oe] = eps;
        cosine_cuts((double)eps - band, (double)eps + band, &d.s_hit[oe], &d.s_near[oe]);
    }
    s.lds = (size_t)M * ((size_t)d.kc * 4 + 16) + (size_t)(ne + 1) * M * 4;
    s.qb = d.nchunk == 1 ? 2048 : 256;  // the single-eps FP32 tiles' default
    return SDB_OK;
}

// The job's tiles of at most qb rows; with max_rows > 0 only the groups
// (query block, list) that start within the first max_rows entries.
void append_tiles(int job, const JobState& s, int qb, std::vector<int4>& tiles, long long max_rows = 0) {
    const sdb_ctx* c = s.c;
    const int NG = c->nblk * c->L2;
    for (int g = 0; g < NG; ++g) {
        if (max_rows > 0 && c->h_grp_ptr[g] >= max_rows) break;
        for (int b = c->h_grp_ptr[g]; b < c->h_grp_ptr[g + 1]; b += qb)
            tiles.push_back(int4{job, g % c->L2, b, std::min(b + qb, c->h_grp_ptr[g + 1])});
    }
}

// Device side of the batch: descriptors, tiles, work-queue head, recheck
// queue (16-byte entries) and counters.
struct Shared {
    DevBuf desc, tiles, head, queue, qcount, stats, chain_sets, chain_pre;
    long long qcap = 0;
    ~Shared() {
        for (DevBuf* b : {&desc, &tiles, &head, &queue, &qcount, &stats, &chain_sets, &chain_pre}) release(*b);
    }
};

// One launch of the tiles of `tiles` (already on the device) and the queue
// count it leaves; clears the jobs' column counts and the head first.
int run_tiles(sdb_ctx* c, Shared& sh, const std::vector<JobState*>& jobs_used, int maxM, size_t lds, int ntiles,
              int grid, unsigned long long* nq_out, float* ms_out) {
    hipStream_t st = c->st;
    for (JobState* s : jobs_used) {
        const SweepJob& d = s->desc;
        HIPCHK(hipMemsetAsync(d.colsum, 0, (size_t)d.n_eps * d.L2 * d.M * 4, st));
    }
    HIPCHK(hipMemsetAsync(sh.head.p, 0, 4, st));
    HIPCHK(hipMemsetAsync(sh.qcount.p, 0, 8, st));
    HIPCHK(hipEventRecord(c->ev[4], st));
    LCHK(launch_sweep_tiles(P_<SweepJob>(sh.desc), maxM, P_<int4>(sh.tiles), ntiles, P_<int>(sh.head),
                            P_<uint4>(sh.queue), P_<unsigned long long>(sh.qcount), sh.qcap, lds, grid, st));
    HIPCHK(hipEventRecord(c->ev[5], st));
    HIPCHK(hipMemcpyAsync(nq_out, sh.qcount.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *ms_out = elapsed(c->ev[4], c->ev[5]);
    return SDB_OK;
}

// Buffers of the job's cluster chain: the degrees of every eps; per (eps,
// minPts) set the core flags, union-find parents, border minima, root flags,
// ranks, scan block sums, list adjacency, core-hub masks, core count, labels
// and cluster count.
int chain_buffers(sdb_batch_job& j, JobState& s) {
    sdb_ctx* c = s.c;
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)c->n;
    const size_t ns = (size_t)j.n_eps * j.n_minpts;
    const size_t nb = (n + 1023) / 1024;
    const SweepJob& d = s.desc;
    ENSURE(c->degree, (size_t)j.n_eps * n * 4);
    ENSURE(c->core, ns * n);
    ENSURE(c->parent, ns * n * 4);
    ENSURE(c->border, ns * n * 4);
    ENSURE(c->rootf, ns * n * 4);
    ENSURE(c->rank, ns * n * 4);
    ENSURE(c->scan_tmp, ns * nb * 4);
    ENSURE(c->adj, ns * d.L2 * d.M * d.W * 4);
    ENSURE(c->corehub, ns * d.L2 * d.W * 4);
    ENSURE(c->stats64, ns * 8);
    ENSURE(c->sw_labels, ns * n * 4);
    ENSURE(c->sw_comps, ns * 4);
    return SDB_OK;
}

// The job's sets in ascending eps, then minPts, order: set q = oe * n_minpts
// + m (the oe-th smallest eps) uses the q-th slice of the chain buffers and
// writes output slot ie * n_minpts + m (ie: that eps's position in the job).
void chain_sets(const sdb_batch_job& j, JobState& s, std::vector<ChainSet>& out) {
    sdb_ctx* c = s.c;
    const SweepJob& d = s.desc;
    const size_t n = (size_t)c->n;
    const size_t nb = (n + 1023) / 1024;
    const int ne = j.n_eps, nm = j.n_minpts;
    const size_t adj_words = (size_t)d.L2 * d.M * d.W, hub_words = (size_t)d.L2 * d.W;
    for (int oe = 0; oe < ne; ++oe) {
        const int ie = s.order[oe];
        for (int m = 0; m < nm; ++m) {
            const size_t o = (size_t)ie * nm + m, q = (size_t)oe * nm + m;
            ChainSet x{};
            x.a = cluster_args(c);
            x.a.E = d.E + (size_t)oe * d.n_ent * d.W;
            x.a.colsum = d.colsum + (size_t)oe * d.L2 * d.M;
            x.a.minPts = j.minPts[m];
            x.a.degree = P_<int>(c->degree) + (size_t)oe * n;
            x.a.core = P_<uint8_t>(c->core) + q * n;
            x.a.parent = P_<int>(c->parent) + q * n;
            x.a.border_min = P_<int>(c->border) + q * n;
            x.a.labels = P_<int>(c->sw_labels) + o * n;
            x.a.adj = P_<uint32_t>(c->adj) + q * adj_words;
            x.a.corehub = P_<uint32_t>(c->corehub) + q * hub_words;
            x.a.n_core = P_<unsigned long long>(c->stats64) + q;  // core counts: not read by the sweep
            x.rootf = P_<int>(c->rootf) + q * n;
            x.rank = P_<int>(c->rank) + q * n;
            x.bsum = P_<int>(c->scan_tmp) + q * nb;
            x.total = P_<int>(c->sw_comps) + o;
            x.n_eps = ne;
            x.n_minpts = nm;
            x.eps_i = oe;
            x.minpts_i = m;
            out.push_back(x);
        }
    }
}

// A run of consecutive sets launched together, and its workgroup offsets per
// step (kChOps x (count + 1), set k of step op at pre[op * (count + 1) + k]).
struct ChainGroup {
    int first = 0, count = 0;
    bool staged = false;  // union-find points with the E rows staged through LDS
    size_t lds = 0;
    std::vector<unsigned> pre;
};

// Splits the sets into groups whose every step launches fewer than 2^31
// threads, and computes their workgroup offsets.
std::vector<ChainGroup> chain_groups(const std::vector<ChainSet>& sets) {
    std::vector<ChainGroup> groups;
    const long long cap = 1ll << 31;
    size_t k = 0;
    while (k < sets.size()) {
        ChainGroup g;
        g.first = (int)k;
        std::vector<long long> tot(kChOps, 0);
        while (k < sets.size()) {
            bool fits = true;
            for (int op = 0; op < kChOps && fits; ++op) {
                const long long b = chain_set_blocks(op, sets[k]);
                fits = (tot[op] + b) * chain_block_threads(op) < cap;
            }
            if (!fits && g.count > 0) break;
            for (int op = 0; op < kChOps; ++op) tot[op] += chain_set_blocks(op, sets[k]);
            ++g.count;
            ++k;
        }
        int maxSW = 0;
        for (int i = g.first; i < g.first + g.count; ++i) maxSW = std::max(maxSW, sets[i].a.S * sets[i].a.W);
        g.lds = (size_t)128 * maxSW * 4;
        g.staged = g.lds <= 64 * 1024;
        g.pre.assign((size_t)kChOps * (g.count + 1), 0u);
        for (int op = 0; op < kChOps; ++op) {
            unsigned acc = 0;
            for (int i = 0; i < g.count; ++i) {
                g.pre[(size_t)op * (g.count + 1) + i] = acc;
                acc += (unsigned)chain_set_blocks(op, sets[g.first + i]);
            }
            g.pre[(size_t)op * (g.count + 1) + g.count] = acc;
        }
        groups.push_back(std::move(g));
    }
    return groups;
}

// The chain of every group, one launch per step, on stream st (captured by
// the caller: kernels only, the clears included).
int enqueue_chain(sdb_ctx* c, const ChainSet* d_sets, const unsigned* d_pre, const std::vector<ChainGroup>& groups,
                  hipStream_t st) {
    static const int steps[] = {kChAdjClear, kChDegree,     kChCore,       kChCorehub,  kChUfInit,
                                kChUfPoints, kChUfLists,    kChUfCompress, kChRootFlags, kChScanBlocks,
                                kChScanTop,  kChScanApply,  kChBorder,     kChLabels};
    size_t pre_off = 0;
    for (const ChainGroup& g : groups) {
        const ChainSet* gs = d_sets + g.first;
        for (int op : steps) {
            const int lop = (op == kChUfPoints && g.staged) ? (int)kChUfPointsStaged : op;
            const size_t row = (size_t)lop * (g.count + 1);
            const unsigned grid = g.pre[row + g.count];
            LCHK(launch_chain_op(lop, gs, d_pre + pre_off + row, g.count, grid, g.lds, st));
        }
        pre_off += g.pre.size();
    }
    return SDB_OK;
}

// Copies the job's labels and cluster counts to its host arrays.
int chain_copy_out(sdb_batch_job& j, JobState& s) {
    sdb_ctx* c = s.c;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->st;
    const size_t n = (size_t)c->n;
    const size_t ns = (size_t)j.n_eps * j.n_minpts;
    std::vector<int> comps(ns);
    if (j.labels) HIPCHK(hipMemcpyAsync(j.labels, c->sw_labels.p, ns * n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(comps.data(), c->sw_comps.p, ns * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (j.n_clusters)
        for (size_t o = 0; o < ns; ++o) j.n_clusters[o] = comps[o] == 0 ? 2 : comps[o] + 1;  // src/sDbscan.cpp:633
    return SDB_OK;
}

// sOPTICS at the largest eps: the context takes that eps's hit masks, column
// counts and degrees (the last eps of the chain's degree slices), then the
// host walks of sdb_optics_orders.
int optics_job(sdb_batch_job& j, JobState& s, HostListPool& lists) {
    sdb_ctx* c = s.c;
    HIPCHK(hipSetDevice(c->device));
    const SweepJob& d = s.desc;
    const int ne = j.n_eps, nm = j.n_minpts;
    const size_t ew = (size_t)d.n_ent * d.W, cw = (size_t)d.L2 * d.M;
    ENSURE(c->E, ew * 4);
    ENSURE(c->colsum, cw * 4);
    HIPCHK(hipMemcpyAsync(c->E.p, d.E + (size_t)(ne - 1) * ew, ew * 4, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->colsum.p, d.colsum + (size_t)(ne - 1) * cw, cw * 4, hipMemcpyDeviceToDevice, c->st));
    if (ne > 1)
        HIPCHK(hipMemcpyAsync(c->degree.p, P_<int>(c->degree) + (size_t)(ne - 1) * c->n, (size_t)c->n * 4,
                              hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->cored = true;
    c->eps = j.eps[s.order[ne - 1]];
    c->minPts = j.minPts[nm - 1];
    c->have_nb = false;
    c->clustered = false;
    std::vector<int> ord;
    int* op = j.ordering;
    if (!op) {
        ord.resize((size_t)nm * c->n);
        op = ord.data();
    }
    return optics_orders(c, nm, j.minPts, op, j.reach, &lists);
}

}  // namespace

extern "C" int sdb_batch_sweep(int device, sdb_batch_job* jobs, int n_jobs, int n_streams, int autotune,
                               double* stage_ms, double* wall_ms) {
    if (n_jobs < 0 || (n_jobs > 0 && !jobs) || n_streams < 1) return SDB_EINVAL;
    const double t0 = now_ms();
    // setup, index, tune, tiles, recheck, cluster buffers, capture, chain, copy-out, optics, teardown
    double st_ms[SDB_SWEEP_STAGES] = {};
    int bad = SDB_OK;
    for (int i = 0; i < n_jobs; ++i) {
        sdb_batch_job& j = jobs[i];
        j.status = SDB_OK;
        j.error[0] = '\0';
        j.tile_rows = 0;
        j.ms = 0.0;
        const sdb_params& p = j.params;
        const char* why = nullptr;
        int code = SDB_OK;
        if (!j.X || !j.hd3_bits || p.n_points <= 0 || p.n_features <= 0 || j.n_eps < 1 || !j.eps ||
            j.n_minpts < 1 || !j.minPts) {
            code = SDB_EINVAL;
            why = "job needs X, hd3_bits, n_points, n_features, eps and minPts";
        } else if (j.n_eps > kSweepMaxEps) {
            code = SDB_EUNSUPPORTED;
            why = "at most 16 eps values per job";
        } else if (p.metric != SDB_COSINE || p.precision != SDB_FP32 || p.topM > 128) {
            code = SDB_EUNSUPPORTED;
            why = "the one-launch sweep takes Cosine jobs with FP32 distances and topM <= 128 (sdb_batch_run takes the rest)";
        }
        if (code) {
            j.status = code;
            std::snprintf(j.error, sizeof(j.error), "%s", why);
            if (!bad) bad = code;
        }
    }
    if (bad) return bad;
    if (n_jobs == 0) {
        if (wall_ms) *wall_ms = 0.0;
        return SDB_OK;
    }
    if (hipSetDevice(device) != hipSuccess) return SDB_EHIP;

    // every job's buffers come from one arena (declared first: it outlives
    // the contexts and the shared buffers, and frees its chunks last)
    DevArena arena;
    std::vector<JobState> S(n_jobs);
    int rc = SDB_OK;
    double t = now_ms();
    for (int i = 0; i < n_jobs && rc == SDB_OK; ++i) {
        rc = sdb_create(device, &S[i].c);
        if (rc == SDB_OK) S[i].c->arena = &arena;
        if (rc) {
            jobs[i].status = rc;
            std::snprintf(jobs[i].error, sizeof(jobs[i].error), "sdb_create failed on device %d", device);
        }
    }
    auto cleanup = [&]() {
        for (JobState& s : S)
            if (s.c) sdb_destroy(s.c);
        S.clear();
    };
    if (rc) {
        cleanup();
        return rc;
    }
    // largest jobs first
    std::vector<int> order(n_jobs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (double)jobs[x].params.n_points * jobs[x].params.n_features >
               (double)jobs[y].params.n_points * jobs[y].params.n_features;
    });
    const int nw = std::min(n_streams, n_jobs);
    st_ms[0] = now_ms() - t;

    // 1. indexes
    t = now_ms();
    for_jobs(order, nw, [&](int i) {
        sdb_batch_job& j = jobs[i];
        sdb_ctx* c = S[i].c;
        const double tj = now_ms();
        int r = sdb_set_params(c, &j.params);
        if (r) { set_job_error(j, r, "set_params", c); return; }
        r = sdb_load_host(c, j.X, j.params.n_points, j.params.n_features);
        if (r) { set_job_error(j, r, "load", c); return; }
        r = sdb_build_index(c, j.hd3_bits, j.R);
        if (r) { set_job_error(j, r, "build_index", c); return; }
        r = prepare_job(j, S[i]);
        if (r) { set_job_error(j, r, "prepare", c); return; }
        S[i].ms += now_ms() - tj;
    });
    st_ms[1] = now_ms() - t;
    for (int i = 0; i < n_jobs; ++i)
        if (jobs[i].status) rc = rc ? rc : jobs[i].status;
    if (rc) {
        cleanup();
        return rc;
    }

    // 2. one launch over every job's tiles
    sdb_ctx* c0 = S[order[0]].c;  // its stream runs the tiles; errors report through it
    Shared sh;
    int maxM = 0;
    size_t lds = 0;
    std::vector<JobState*> used;
    for (JobState& s : S) {
        maxM = std::max(maxM, s.desc.M);
        lds = std::max(lds, s.lds);
        used.push_back(&s);
    }
    int n_cu = 256;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            n_cu = prop.multiProcessorCount;
    }
    const int per_cu = std::max(1, std::min(2, (int)(kLdsBytes / std::max<size_t>(lds, 1))));
    const int grid = n_cu * per_cu;
    auto fail_all = [&](int r, const char* what) {
        for (int i = 0; i < n_jobs; ++i) set_job_error(jobs[i], r, what, c0);
        cleanup();
        return r;
    };
    if (ensure(c0, sh.desc, sizeof(SweepJob) * n_jobs) || ensure(c0, sh.head, 16) || ensure(c0, sh.qcount, 16) ||
        ensure(c0, sh.stats, 64))
        return fail_all(SDB_ENOMEM, "sweep buffers");
    sh.qcap = 1 << 22;
    if (ensure(c0, sh.queue, (size_t)sh.qcap * 16)) return fail_all(SDB_ENOMEM, "sweep queue");
    {
        std::vector<SweepJob> hd(n_jobs);
        for (int i = 0; i < n_jobs; ++i) hd[i] = S[i].desc;
        if (hipMemcpy(sh.desc.p, hd.data(), sizeof(SweepJob) * n_jobs, hipMemcpyHostToDevice) != hipSuccess)
            return fail_all(SDB_EHIP, "descriptor upload");
    }
    // autotune: the first (largest) job of every row length times the sweep
    // kernel for each candidate tile height on the same entries, its groups
    // within the first 4M entries (enough tiles of every height for all
    // workgroups of the launch)
    t = now_ms();
    if (autotune) {
        std::map<int, int> best_by_d;
        static const int cand[] = {256, 1024, 2048, 4096};
        for (int i : order) {
            const int dk = S[i].desc.dk;
            auto it = best_by_d.find(dk);
            if (it != best_by_d.end()) {
                S[i].qb = it->second;
                continue;
            }
            float best_ms = 0.f;
            int best = S[i].qb;
            for (int qb : cand) {
                std::vector<int4> tl;
                append_tiles(i, S[i], qb, tl, 4ll << 20);
                if (ensure(c0, sh.tiles, tl.size() * sizeof(int4)) ||
                    hipMemcpy(sh.tiles.p, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
                    return fail_all(SDB_EHIP, "tile upload");
                unsigned long long nq = 0;
                float ms = 0.f;
                const int r = run_tiles(c0, sh, {&S[i]}, S[i].desc.M, S[i].lds, (int)tl.size(),
                                        std::min(grid, (int)tl.size()), &nq, &ms);
                if (r) return fail_all(r, "autotune");
                if (best_ms == 0.f || ms < best_ms) { best_ms = ms; best = qb; }
            }
            best_by_d[dk] = best;
            S[i].qb = best;
        }
    }
    st_ms[2] = now_ms() - t;
    std::vector<int4> tiles;
    for (int i : order) {
        append_tiles(i, S[i], S[i].qb, tiles);
        jobs[i].tile_rows = S[i].qb;
    }
    // the heaviest tiles first, so that the queue drains evenly
    std::stable_sort(tiles.begin(), tiles.end(), [&](const int4& a, const int4& b) {
        return (long long)(a.w - a.z) * S[a.x].desc.dk > (long long)(b.w - b.z) * S[b.x].desc.dk;
    });
    if (ensure(c0, sh.tiles, std::max<size_t>(tiles.size(), 1) * sizeof(int4)) ||
        hipMemcpy(sh.tiles.p, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
        return fail_all(SDB_EHIP, "tile upload");
    unsigned long long nq = 0;
    float tiles_ms = 0.f;
    for (int attempt = 0; attempt < 3; ++attempt) {
        const int r = run_tiles(c0, sh, used, maxM, lds, (int)tiles.size(), grid, &nq, &tiles_ms);
        if (r) return fail_all(r, "sweep tiles");
        if (nq <= (unsigned long long)sh.qcap) break;
        // more band pairs than the queue holds: a larger queue and the tiles again
        sh.qcap = (long long)nq * 2;
        if (ensure(c0, sh.queue, (size_t)sh.qcap * 16)) return fail_all(SDB_ENOMEM, "sweep queue");
    }
    if (nq > (unsigned long long)sh.qcap) return fail_all(SDB_ENOMEM, "sweep queue overflow");
    st_ms[3] = tiles_ms;

    // 3. FP64 recheck of the band pairs, all jobs at once
    t = now_ms();
    if (hipMemsetAsync(sh.stats.p, 0, 64, c0->st) != hipSuccess ||
        launch_sweep_recheck(P_<SweepJob>(sh.desc), P_<uint4>(sh.queue), (long long)nq,
                             P_<unsigned long long>(sh.stats), c0->st) != 0 ||
        hipStreamSynchronize(c0->st) != hipSuccess)
        return fail_all(SDB_EHIP, "sweep recheck");
    st_ms[4] = now_ms() - t;

    // 4. cluster chains: one launch per step over every (job, eps, minPts)
    // set, all steps captured as one graph (kernels only; no memset nodes)
    int replays = 1;
    if (const char* e = std::getenv("SDB_GRAPH_REPLAYS")) replays = std::max(1, std::min(16, std::atoi(e)));
    t = now_ms();
    for_jobs(order, nw, [&](int i) {
        const int r = chain_buffers(jobs[i], S[i]);
        if (r) set_job_error(jobs[i], r, "cluster buffers", S[i].c);
    });
    std::vector<ChainSet> sets;
    for (int i : order)
        if (!jobs[i].status) chain_sets(jobs[i], S[i], sets);
    const std::vector<ChainGroup> groups = chain_groups(sets);
    size_t pre_words = 0;
    for (const ChainGroup& g : groups) pre_words += g.pre.size();
    if (!sets.empty()) {
        if (ensure(c0, sh.chain_sets, sets.size() * sizeof(ChainSet)) ||
            ensure(c0, sh.chain_pre, pre_words * sizeof(unsigned)))
            return fail_all(SDB_ENOMEM, "cluster chain descriptors");
        std::vector<unsigned> pre_all;
        pre_all.reserve(pre_words);
        for (const ChainGroup& g : groups) pre_all.insert(pre_all.end(), g.pre.begin(), g.pre.end());
        if (hipMemcpy(sh.chain_sets.p, sets.data(), sets.size() * sizeof(ChainSet), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(sh.chain_pre.p, pre_all.data(), pre_words * sizeof(unsigned), hipMemcpyHostToDevice) !=
                hipSuccess)
            return fail_all(SDB_EHIP, "cluster chain descriptor upload");
    }
    st_ms[5] = now_ms() - t;
    t = now_ms();
    hipGraphExec_t exec = nullptr;
    if (!sets.empty()) {
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(c0->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail_all(SDB_EHIP, "cluster chain capture");
        const int re = enqueue_chain(c0, P_<ChainSet>(sh.chain_sets), P_<unsigned>(sh.chain_pre), groups, c0->st);
        const hipError_t ee = hipStreamEndCapture(c0->st, &graph);
        hipError_t ei = hipErrorUnknown;
        if (re == SDB_OK && ee == hipSuccess) ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (re != SDB_OK || ee != hipSuccess || ei != hipSuccess) {
            (void)hipGetLastError();
            return fail_all(SDB_EHIP, "cluster chain capture");
        }
    }
    st_ms[6] = now_ms() - t;
    t = now_ms();
    if (exec) {
        // every replay must give the first replay's labels
        std::vector<std::vector<int>> lab0;
        for (int k = 0; k < replays; ++k) {
            if (hipGraphLaunch(exec, c0->st) != hipSuccess || hipStreamSynchronize(c0->st) != hipSuccess) {
                (void)hipGraphExecDestroy(exec);
                return fail_all(SDB_EHIP, "cluster chain");
            }
            if (replays == 1) break;
            size_t q = 0;
            for (int i : order) {
                if (jobs[i].status) continue;
                sdb_ctx* c = S[i].c;
                std::vector<int> lab((size_t)jobs[i].n_eps * jobs[i].n_minpts * c->n);
                if (hipMemcpy(lab.data(), c->sw_labels.p, lab.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
                    (void)hipGraphExecDestroy(exec);
                    return fail_all(SDB_EHIP, "cluster chain replay read-back");
                }
                if (k == 0) {
                    lab0.push_back(std::move(lab));
                } else if (lab != lab0[q]) {
                    (void)hipGraphExecDestroy(exec);
                    char msg[96];
                    std::snprintf(msg, sizeof(msg), "graph replay %d gave other labels than replay 0", k);
                    return fail_all(SDB_EHIP, msg);
                }
                ++q;
            }
        }
        (void)hipGraphExecDestroy(exec);
    }
    st_ms[7] = now_ms() - t;
    t = now_ms();
    for_jobs(order, nw, [&](int i) {
        if (jobs[i].status) return;
        const double tj = now_ms();
        const int r = chain_copy_out(jobs[i], S[i]);
        if (r) set_job_error(jobs[i], r, "clusters", S[i].c);
        S[i].ms += now_ms() - tj;
    });
    st_ms[8] = now_ms() - t;

    // 5. sOPTICS at the largest eps; the walks and their neighbourhood lists
    // scale with the point count, so the most points go first
    t = now_ms();
    std::vector<int> order5(order);
    std::stable_sort(order5.begin(), order5.end(),
                     [&](int x, int y) { return jobs[x].params.n_points > jobs[y].params.n_points; });
    HostListPool lists;
    for_jobs(order5, nw, [&](int i) {
        sdb_batch_job& j = jobs[i];
        if (j.status || !j.optics || (!j.ordering && !j.reach)) return;
        const double tj = now_ms();
        const int r = optics_job(j, S[i], lists);
        if (r) set_job_error(j, r, "optics", S[i].c);
        S[i].ms += now_ms() - tj;
    });
    lists.clear();
    st_ms[9] = now_ms() - t;
    for (int i = 0; i < n_jobs; ++i) jobs[i].ms = S[i].ms;
    t = now_ms();
    cleanup();
    st_ms[10] = now_ms() - t;
    if (stage_ms) std::memcpy(stage_ms, st_ms, sizeof(st_ms));
    if (wall_ms) *wall_ms = now_ms() - t0;
    for (int i = 0; i < n_jobs; ++i)
        if (jobs[i].status != SDB_OK) return jobs[i].status;
    return SDB_OK;
}
