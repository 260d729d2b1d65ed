// Batched eps / minPts sweeps over many datasets (sdb_batch_run).
//
// Reference: test_sDbscan (src/sDbscan.cpp:980-1039) sweeps the eps values of
// one dataset on one index; fit_sOptics (src/sDbscan.cpp:1711-1769) runs
// sOPTICS.  A batch holds many such datasets of uneven size.  Host worker
// threads, each owning one context (one HIP stream, buffers grown to its
// largest job), take jobs from a shared counter, so the small launches and
// host round trips of one dataset overlap with the kernels of the others.
// Per job: one index, one distance pass per eps (ascending, so the largest eps
// is last and sOPTICS reuses it), one core test per minPts.  The tile rows of
// the distance pass are tuned once per (n_features, metric, precision) class
// by the first job of the class.  The walks' host lists come from one pool
// shared by the workers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <map>
#include <mutex>
#include <numeric>
#include <random>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/sdbscan_amd.h"
#include "sdb_ctx.hpp"

namespace {

typedef std::tuple<int, int, int> TuneKey;  // n_features, metric, precision

struct TuneTable {
    std::mutex mu;
    std::condition_variable cv;
    std::map<TuneKey, int> rows;  // > 0 tuned, 0 being tuned, -1 tuning failed (default tiles)
};

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int job_fail(sdb_batch_job& j, int code, const char* what, sdb_ctx* c) {
    j.status = code;
    std::snprintf(j.error, sizeof(j.error), "%s: %s", what, c ? sdb_last_error(c) : "no context");
    return code;
}

int check_job(sdb_batch_job& j) {
    j.status = SDB_OK;
    j.error[0] = '\0';
    j.tile_rows = 0;
    j.ms = 0.0;
    if (!j.X || !j.hd3_bits || j.params.n_points <= 0 || j.params.n_features <= 0 || j.n_eps < 1 || !j.eps ||
        j.n_minpts < 1 || !j.minPts) {
        j.status = SDB_EINVAL;
        std::snprintf(j.error, sizeof(j.error), "job needs X, hd3_bits, n_points, n_features, eps and minPts");
        return SDB_EINVAL;
    }
    return SDB_OK;
}

void run_job(sdb_ctx* c, sdb_batch_job& j, int autotune, TuneTable& tune, HostListPool& lists) {
    const double t0 = now_ms();
    const sdb_params& p = j.params;
    const long long n = p.n_points;
    int r = sdb_set_params(c, &p);
    if (r) { job_fail(j, r, "set_params", c); return; }
    r = sdb_load_host(c, j.X, n, p.n_features);
    if (r) { job_fail(j, r, "load", c); return; }
    r = sdb_build_index(c, j.hd3_bits, j.R);
    if (r) { job_fail(j, r, "build_index", c); return; }

    // eps values ascending; outputs keep the job's order
    std::vector<int> order(j.n_eps);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return j.eps[x] < j.eps[y]; });
    const float eps_max = j.eps[order.back()];

    // tile rows: tuned by the first job of the class, reused afterwards
    int rows = 0;
    const TuneKey key(p.n_features, p.metric, p.precision);
    bool tune_here = false;
    {
        // a job of a class being tuned waits for the tuner's choice
        std::unique_lock<std::mutex> lk(tune.mu);
        auto it = tune.rows.find(key);
        if (it == tune.rows.end()) {
            if (autotune) { tune.rows[key] = 0; tune_here = true; }
        } else {
            tune.cv.wait(lk, [&] { return tune.rows[key] != 0; });
            rows = std::max(0, tune.rows[key]);
        }
    }
    if (tune_here) {
        static const int cand[] = {256, 1024, 2048, 4096, 16384};
        int best = 0;
        double ms = 0.0;
        r = sdb_tune_tile_rows(c, eps_max, cand, (int)(sizeof(cand) / sizeof(cand[0])), &best, &ms);
        {
            std::lock_guard<std::mutex> lk(tune.mu);
            tune.rows[key] = r ? -1 : best;
        }
        tune.cv.notify_all();
        if (r) { job_fail(j, r, "tune_tile_rows", c); return; }
        rows = best;
    } else {
        r = sdb_set_tile_rows(c, rows);
        if (r) { job_fail(j, r, "set_tile_rows", c); return; }
    }
    j.tile_rows = rows;

    for (int oe = 0; oe < j.n_eps; ++oe) {
        const int ie = order[oe];
        for (int im = 0; im < j.n_minpts; ++im) {
            r = sdb_find_core(c, j.eps[ie], j.minPts[im], nullptr, nullptr);
            if (r) { job_fail(j, r, "find_core", c); return; }
            const size_t o = (size_t)ie * j.n_minpts + im;
            int nc = 0;
            r = sdb_form_clusters(c, j.labels ? j.labels + o * (size_t)n : nullptr, &nc);
            if (r) { job_fail(j, r, "form_clusters", c); return; }
            if (j.n_clusters) j.n_clusters[o] = nc;
        }
    }
    if (j.optics && (j.ordering || j.reach)) {
        // the last eps of the sweep is the largest: its distances stand
        r = sdb_find_core(c, eps_max, j.minPts[0], nullptr, nullptr);
        if (r) { job_fail(j, r, "find_core", c); return; }
        std::vector<int> ord;
        int* op = j.ordering;
        if (!op) { ord.resize((size_t)j.n_minpts * n); op = ord.data(); }
        r = optics_orders(c, j.n_minpts, j.minPts, op, j.reach, &lists);
        if (r) { job_fail(j, r, "optics_orders", c); return; }
    }
    j.ms = now_ms() - t0;
}

}  // namespace

extern "C" int sdb_batch_run(int device, sdb_batch_job* jobs, int n_jobs, int n_streams, int autotune,
                             double* wall_ms) {
    if (n_jobs < 0 || (n_jobs > 0 && !jobs) || n_streams < 1) return SDB_EINVAL;
    const double t0 = now_ms();
    for (int i = 0; i < n_jobs; ++i) check_job(jobs[i]);
    // largest jobs first, so that no big job starts last and holds the tail
    std::vector<int> queue(n_jobs);
    std::iota(queue.begin(), queue.end(), 0);
    std::stable_sort(queue.begin(), queue.end(), [&](int x, int y) {
        return (double)jobs[x].params.n_points * jobs[x].params.n_features >
               (double)jobs[y].params.n_points * jobs[y].params.n_features;
    });
    std::atomic<int> next{0};
    TuneTable tune;
    HostListPool lists;  // the walks' host lists, reused from job to job
    const int nw = std::min(n_streams, std::max(n_jobs, 1));
    std::vector<int> create_rc(nw, SDB_OK);
    auto worker = [&](int w) {
        sdb_ctx* c = nullptr;
        create_rc[w] = sdb_create(device, &c);
        while (true) {
            const int k = next.fetch_add(1);
            if (k >= n_jobs) break;
            sdb_batch_job& j = jobs[queue[k]];
            if (j.status != SDB_OK) continue;
            if (create_rc[w] != SDB_OK) {
                j.status = create_rc[w];
                std::snprintf(j.error, sizeof(j.error), "sdb_create failed on device %d", device);
                continue;
            }
            run_job(c, j, autotune, tune, lists);
        }
        if (c) sdb_destroy(c);
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nw; ++w) th.emplace_back(worker, w);
    worker(0);
    for (auto& t : th) t.join();
    lists.clear();
    if (wall_ms) *wall_ms = now_ms() - t0;
    for (int i = 0; i < n_jobs; ++i)
        if (jobs[i].status != SDB_OK) return jobs[i].status;
    return SDB_OK;
}

extern "C" int sdb_hd3_bits(int n_bits, unsigned int seed, uint8_t* out) {
    if (n_bits < 0 || (n_bits > 0 && !out)) return SDB_EINVAL;
    std::default_random_engine gen(seed);
    std::uniform_int_distribution<uint32_t> u(0, 1);
    for (int i = 0; i < n_bits; ++i) out[i] = (uint8_t)(u(gen) & 1u);
    return SDB_OK;
}

extern "C" int sdb_rff_matrix(int metric, int half, int d, float sigma, unsigned int seed, float* out) {
    if (half < 0 || d < 0 || (half > 0 && d > 0 && !out) || !(sigma > 0.0f)) return SDB_EINVAL;
    if (metric != SDB_L1 && metric != SDB_L2) return SDB_EINVAL;
    std::default_random_engine gen(seed);
    if (metric == SDB_L1) {
        std::cauchy_distribution<float> cd(0, 1.0 / sigma);
        for (int c = 0; c < d; ++c)
            for (int r = 0; r < half; ++r) out[(size_t)c * half + r] = cd(gen);
    } else {
        std::normal_distribution<float> nd(0, 1.0 / sigma);
        for (int c = 0; c < d; ++c)
            for (int r = 0; r < half; ++r) out[(size_t)c * half + r] = nd(gen);
    }
    return SDB_OK;
}
