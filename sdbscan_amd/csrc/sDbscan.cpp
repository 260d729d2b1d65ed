// sDbscan.cpp -- host side of the drop-in class: parameters, the reference's host
// RNG calls, phase timers and the sOPTICS ordering pass.  Every device stage goes
// through the C ABI (include/sdbscan_amd.h).
#include "sDbscan.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <queue>
#include <random>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace {

int metric_of(const std::string& dist) {
    if (dist == "Cosine") return SDB_COSINE;
    if (dist == "L1") return SDB_L1;
    if (dist == "L2") return SDB_L2;
    if (dist == "Chi2") return SDB_CHI2;
    if (dist == "JS") return SDB_JS;
    return -1;
}

std::string int2str(long long x) {
    std::stringstream ss;
    ss << x;
    return ss.str();
}

unsigned seed_of(int random_seed) {
    unsigned s = (unsigned)std::chrono::system_clock::now().time_since_epoch().count();
    if (random_seed >= 0) s = (unsigned)random_seed;
    return s;
}


}  // namespace

sDbscan::sDbscan(int n, int d) : n_points(n), n_features(d) {}

sDbscan::~sDbscan() {
    if (ctx) sdb_destroy(ctx);
}

void sDbscan::set_threads(int t) { n_threads = t <= 0 ? omp_get_max_threads() : t; }

// src/sDbscan.h:79-108
void sDbscan::set_params(int numProj, int k, int m, std::string dist, int kerDim, float kerSigma,
                         float kerSam, int noiseClustering, float prob, bool ver,
                         std::string filename, int numThreads, int randomSeed) {
    n_proj = numProj;
    topK = k;
    topM = m;
    distance = dist;
    ker_n_features = kerDim;
    ker_sigma = kerSigma;
    ker_intervalSampling = kerSam;
    samplingProb = prob;
    clusterNoise = noiseClustering;
    verbose = ver;
    set_threads(numThreads);
    seed = randomSeed;
    output = filename;
    if (distance == "Cosine") ker_n_features = n_features;
    fhtDim = 1 << int(std::ceil(std::log2(std::max(n_proj, ker_n_features))));
}

// src/sDbscan.h:110-124
void sDbscan::set_kernel_params(std::string dist, int kerDim, float kerSigma, float kerSam) {
    distance = dist;
    ker_n_features = kerDim;
    ker_sigma = kerSigma;
    ker_intervalSampling = kerSam;
    if (distance == "Cosine") ker_n_features = n_features;
    fhtDim = 1 << int(std::ceil(std::log2(std::max(n_proj, ker_n_features))));
}

// src/sDbscan.h:127-140
void sDbscan::set_sngParams(std::string dist, float prob, int noiseClustering, bool ver,
                            std::string filename, int numThreads, int randomSeed) {
    distance = dist;
    samplingProb = prob;
    if (noiseClustering > 0) clusterNoise = 2;
    verbose = ver;
    set_threads(numThreads);
    seed = randomSeed;
    output = filename;
}

// src/sDbscan.h:142-163
void sDbscan::clear() {
    labels_.clear();
    n_clusters_ = 0;
    ordering_.clear();
    reachability_.clear();
    matrix_R.clear();
    bitHD3.clear();
}

void sDbscan::set_precision(const std::string& p) {
    if (p == "FP32" || p == "fp32") precision = SDB_FP32;
    else if (p == "BF16" || p == "bf16") precision = SDB_BF16;
    else throw std::invalid_argument("precision must be FP32 or BF16");
}

void sDbscan::transform_data(float* X) {
    prepare(X);
    check(sdb_transform(ctx, X), "transform");
}

void sDbscan::set_device(int dev) {
    if (ctx) { sdb_destroy(ctx); ctx = nullptr; }
    device = dev;
}

void sDbscan::ensure_ctx() {
    if (ctx) return;
    int rc = sdb_create(device, &ctx);
    if (rc != SDB_OK) {
        ctx = nullptr;
        throw std::runtime_error("sdb_create failed (no usable HIP device " + int2str(device) + ")");
    }
}

void sDbscan::check(int rc, const char* what) {
    if (rc == SDB_OK) return;
    std::string msg = std::string(what) + ": " + (ctx ? sdb_last_error(ctx) : "no context");
    if (rc == SDB_ENEGATIVE || rc == SDB_EZERO) msg = ctx ? sdb_last_error(ctx) : what;
    if (rc == SDB_EINVAL || rc == SDB_EUNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Host RNG: bitHD3Generator, gaussGenerator, cauchyGenerator
// (src/Utilities.h:47-122), called in the order of rp_parIndex (src/sDbscan.cpp:25-34).
void sDbscan::make_random() {
    const int half = ker_n_features / 2;
    matrix_R.clear();
    if (distance == "L1" || distance == "L2") {
        std::default_random_engine gen(seed_of(seed));
        matrix_R.assign((size_t)half * n_features, 0.0f);
        if (distance == "L1") {
            std::cauchy_distribution<float> cd(0, 1.0 / ker_sigma);
            for (int c = 0; c < n_features; ++c)
                for (int r = 0; r < half; ++r) matrix_R[(size_t)c * half + r] = cd(gen);
        } else {
            std::normal_distribution<float> nd(0, 1.0 / ker_sigma);
            for (int c = 0; c < n_features; ++c)
                for (int r = 0; r < half; ++r) matrix_R[(size_t)c * half + r] = nd(gen);
        }
    }
    const int nbits = fhtDim * n_rotate;
    std::default_random_engine gen(seed_of(seed));
    std::uniform_int_distribution<uint32_t> u(0, 1);
    bitHD3.assign(nbits, 0);
    for (int i = 0; i < nbits; ++i) bitHD3[i] = (uint8_t)(u(gen) & 1u);
}

void sDbscan::prepare(const float* X) {
    const int metric = metric_of(distance);
    if (metric < 0) throw std::invalid_argument("Error: The distance is not support");
    if (fhtDim <= 0) set_params(n_proj, topK, topM, distance, ker_n_features, ker_sigma,
                                ker_intervalSampling, clusterNoise, samplingProb, verbose, output,
                                n_threads, seed);
    ensure_ctx();
    sdb_params p{};
    p.n_points = n_points;
    p.n_features = n_features;
    p.n_proj = n_proj;
    p.topK = topK;
    p.topM = topM;
    p.metric = metric;
    p.ker_n_features = metric == SDB_COSINE ? n_features : ker_n_features;
    p.ker_sigma = ker_sigma;
    p.ker_intervalSampling = ker_intervalSampling;
    p.fhtDim = fhtDim;
    p.precision = precision;
    check(sdb_set_params(ctx, &p), "set_params");
    check(sdb_load_host(ctx, X, n_points, n_features), "load");
    make_random();
}

void sDbscan::print_params(float eps, int minPts) const {
    std::cout << "eps: " << eps << std::endl;
    std::cout << "minPts: " << minPts << std::endl;
    std::cout << "n_points: " << n_points << std::endl;
    std::cout << "n_features: " << n_features << std::endl;
    std::cout << "n_proj: " << n_proj << std::endl;
    std::cout << "topK: " << topK << std::endl;
    std::cout << "topM: " << topM << std::endl;
    std::cout << "distance: " << distance << std::endl;
    std::cout << "cluster noise: " << clusterNoise << std::endl;
    std::cout << "kernel features: " << ker_n_features << std::endl;
    std::cout << "sigma: " << ker_sigma << std::endl;
    std::cout << "interval sampling: " << ker_intervalSampling << std::endl;
    std::cout << "sDbscan-1NN prob: " << samplingProb << std::endl;
    std::cout << "n_threads: " << n_threads << std::endl;
}

// File names of src/sDbscan.cpp:880-886 and :1759-1765.
std::string sDbscan::out_name(float eps, int minPts, bool optics) const {
    return output + "_" + distance + "_Eps_" + int2str(std::lround(1000 * eps)) + "_MinPts_" +
           int2str(minPts) + (optics ? "_NumEmbed_" : "_KerFeatures_") + int2str(ker_n_features) +
           "_NumProj_" + int2str(n_proj) + "_TopM_" + int2str(topM) + "_TopK_" + int2str(topK);
}

// Noise labelling of formCluster (src/sDbscan.cpp:625-627 -> labelNoise,
// :647-807).  clusterNoise = 2 (heuristic, :661-682) walks the kept
// neighbourhood of each noise point, but a noise point is never core and the
// neighbourhoods of non-core points are not kept (:374-400), so no label
// changes.  Any other nonzero value is sDbscan-1NN: the core points are
// sampled here exactly as the reference samples them (same expressions, the
// process's rand() in increasing index order, :688-709) and the device does
// the projections and the 1NN scoring.
void sDbscan::label_noise(const std::vector<uint8_t>& core) {
    if (clusterNoise == 0) return;
    if (verbose) {
        long long nn = 0;
        for (int v : labels_) nn += v == -1;
        std::cout << "Number of noisy points: " << nn << std::endl;
    }
    if (clusterNoise == 2) return;
    int iNumCore = 0;
    for (uint8_t c : core) iNumCore += c ? 1 : 0;
    float fProb = 1.0;
    if (iNumCore >= samplingProb * n_points) fProb = samplingProb * n_points / iNumCore;
    std::vector<int> sampled;
    for (int Xi = 0; Xi < n_points; ++Xi) {
        if (!core[Xi]) continue;
        if (fProb == 1.0)
            sampled.push_back(Xi);
        else if (rand() / (RAND_MAX + 1.) <= fProb)
            sampled.push_back(Xi);
    }
    if (verbose) std::cout << "Number of sampled core points: " << sampled.size() << std::endl;
    check(sdb_label_noise(ctx, sampled.data(), (int)sampled.size(), labels_.data()), "label_noise");
}

// loadtxtData (src/Utilities.cpp:60-95): n lines of d whitespace-separated
// floats read with fscanf("%f") into n rows of d floats (the reference's d x n
// column-major matrix); values the file lacks stay 0, as in the reference.
// The transform of the distance runs on the device with the rest of the fit.
std::vector<float> sDbscan::load_txt(const std::string& dataset) const {
    FILE* f = std::fopen(dataset.c_str(), "r");
    if (!f) throw std::runtime_error("Error: Data file does not exist !");
    std::vector<float> X((size_t)n_points * n_features, 0.0f);
    for (size_t i = 0; i < X.size(); ++i)
        if (std::fscanf(f, "%f", &X[i]) != 1) break;
    std::fclose(f);
    std::cout << "Finish reading data" << std::endl;
    std::cout << "Now checking the condition of data given the distance." << std::endl;
    return X;
}

// src/sDbscan.cpp:903-976
void sDbscan::load_fit_sDbscan(const std::string& dataset, float eps, int minPts) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<float> X = load_txt(dataset);
    if (verbose)
        std::cout << "Loading data time = " << std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count()
                  << "[ms]" << std::endl;
    fit_sDbscan(X.data(), eps, minPts);
}

// src/sDbscan.cpp:1053-1112
void sDbscan::load_test_sDbscan(const std::string& dataset, float eps, float rangeEps, int n_tests, int minPts) {
    std::vector<float> X = load_txt(dataset);
    test_sDbscan(X.data(), eps, rangeEps, n_tests, minPts);
}

// src/sDbscan.cpp:1777-1838
void sDbscan::load_fit_sOptics(const std::string& dataset, float eps, int minPts) {
    std::vector<float> X = load_txt(dataset);
    fit_sOptics(X.data(), eps, minPts);
}

// src/sDbscan.cpp:830-890
void sDbscan::fit_sDbscan(const float* X, float eps, int minPts) {
    if (verbose) print_params(eps, minPts);
    prepare(X);
    check(sdb_build_index(ctx, bitHD3.data(), matrix_R.empty() ? nullptr : matrix_R.data()), "build_index");
    std::vector<uint8_t> core(clusterNoise ? n_points : 0);
    check(sdb_find_core(ctx, eps, minPts, clusterNoise ? core.data() : nullptr, nullptr), "find_core");
    labels_.assign(n_points, -1);
    check(sdb_form_clusters(ctx, labels_.data(), &n_clusters_), "form_clusters");
    label_noise(core);
    if (verbose) {
        sdb_stats s{};
        sdb_get_stats(ctx, &s);
        std::cout << "Check X supporting distance time = " << (long long)s.t_transform_ms << "[ms]" << std::endl;
        std::cout << "Build index time = " << (long long)s.t_index_ms << "[ms]" << std::endl;
        std::cout << "Number of core points: " << s.n_core << std::endl;
        std::cout << "Find core points time = " << (long long)s.t_core_ms << "[ms]" << std::endl;
        std::cout << "Number of clusters: " << n_clusters_ << std::endl;
        std::cout << "Form clusters time  (including clustering noise) = " << (long long)s.t_cluster_ms << "[ms]" << std::endl;
        std::ofstream f(out_name(eps, minPts, false));
        for (int v : labels_) f << v << '\n';
    }
}

// src/sDbscan.cpp:980-1039: one index, n_tests x (core points + clusters)
void sDbscan::test_sDbscan(const float* X, float eps, float rangeEps, int n_tests, int minPts) {
    std::cout << "base eps: " << eps << std::endl;
    std::cout << "range eps: " << rangeEps << std::endl;
    std::cout << "n_tests: " << n_tests << std::endl;
    prepare(X);
    verbose = true;
    check(sdb_build_index(ctx, bitHD3.data(), matrix_R.empty() ? nullptr : matrix_R.data()), "build_index");
    sdb_stats s{};
    sdb_get_stats(ctx, &s);
    std::cout << "Build index time = " << (long long)(s.t_transform_ms + s.t_index_ms) << "[ms]" << std::endl;
    for (int i = 0; i < n_tests; ++i) {
        const float e = sweep_eps(eps, rangeEps, i);
        std::cout << "Eps: " << e << std::endl;
        std::vector<uint8_t> core(clusterNoise ? n_points : 0);
        check(sdb_find_core(ctx, e, minPts, clusterNoise ? core.data() : nullptr, nullptr), "find_core");
        labels_.assign(n_points, -1);
        check(sdb_form_clusters(ctx, labels_.data(), &n_clusters_), "form_clusters");
        label_noise(core);
        sdb_get_stats(ctx, &s);
        std::cout << "Find core points time = " << (long long)s.t_core_ms << "[ms]" << std::endl;
        std::cout << "Form clusters time (including clustering noise) = " << (long long)s.t_cluster_ms << "[ms]" << std::endl;
        std::ofstream f(out_name(e, minPts, false));
        for (int v : labels_) f << v << '\n';
    }
}

// src/sDbscan.cpp:1711-1769.  Neighbourhoods and core distances come from
// the device; the ordering pass is the sequential heap walk of
// formOptics_scikit (src/sDbscan.cpp:1588-1703) over neighbourhoods sorted by
// index; the reachability values are then computed on the device for that
// ordering.
void sDbscan::fit_sOptics(const float* X, float eps, int minPts) {
    if (verbose) print_params(eps, minPts);
    prepare(X);
    check(sdb_build_index(ctx, bitHD3.data(), matrix_R.empty() ? nullptr : matrix_R.data()), "build_index");
    std::vector<uint8_t> core(n_points);
    check(sdb_find_core(ctx, eps, minPts, core.data(), nullptr), "find_core");
    optics_order(core, minPts);
    if (!output.empty()) {
        std::ofstream f(out_name(eps, minPts, true));
        for (int k = 0; k < n_points; ++k) f << ordering_[k] << " " << reachability_[k] << '\n';
    }
}

// formOptics_scikit (src/sDbscan.cpp:1588-1703) over the context's
// neighbourhoods: the ordering is the sequential heap walk on the host, the
// reachability values are computed on the device for that ordering.
void sDbscan::optics_order(const std::vector<uint8_t>& core, int minPts) {
    (void)core;  // the device core flags of the last find_core are the same
    ordering_.assign(n_points, 0);
    reachability_.assign(n_points, 0.0f);
    check(sdb_optics_order(ctx, minPts, ordering_.data(), reachability_.data()), "optics_order");
}

// ---------------------------------------------------------------------------
// sngDbscan / sngOptics.  The sample draws are made here with the
// reference's std:: calls; the device compares every point with its samples.
//
// sng_findCorePoints (src/sDbscan.cpp:425-465): every point seeds its own
// default_random_engine with `seed` (random_seed >= 0) or with the clock, and
// draws uniform_int_distribution<>(0, n - 1) until it has iNumSamples draws
// other than itself.  With random_seed >= 0 all points therefore walk the
// same draw sequence; this build uses one shared sequence in both cases (with
// a clock seed the reference's per-thread clock reads are not reproducible).
// Point a's window is the shortest prefix holding iNumSamples draws != a.
void sDbscan::sng_samples(bool optics, std::vector<int>& samples, std::vector<int>& win_end) const {
    const int iNumSamples = (int)std::ceil(1.0 * n_points * samplingProb);
    samples.clear();
    win_end.clear();
    if (!optics && samplingProb >= 1.0f) {  // exact branch (src/sDbscan.cpp:466-478)
        samples.resize(n_points);
        for (int i = 0; i < n_points; ++i) samples[i] = i;
        return;
    }
    if (optics) {
        // sng_findCoreDist (src/sDbscan.cpp:1497-1505): mt19937, exactly
        // iNumSamples draws per point, draws equal to the point are skipped
        std::random_device rd;
        std::mt19937 gen(seed >= 0 ? (unsigned)seed : rd());
        std::uniform_int_distribution<> distrib(0, n_points - 1);
        samples.resize(iNumSamples);
        for (int j = 0; j < iNumSamples; ++j) samples[j] = distrib(gen);
        return;
    }
    unsigned sd = (unsigned)std::chrono::system_clock::now().time_since_epoch().count();
    if (seed >= 0) sd = (unsigned)seed;
    std::default_random_engine generator(sd);
    std::uniform_int_distribution<> distrib(0, n_points - 1);
    long long T = (long long)iNumSamples + 64;
    for (;;) {
        while ((long long)samples.size() < T) samples.push_back(distrib(generator));
        // window ends of the points that occur in the sequence
        std::vector<std::vector<int>> occ;
        std::vector<int> who;
        std::unordered_map<int, int> slot;
        for (int j = 0; j < (int)samples.size(); ++j) {
            auto it = slot.find(samples[j]);
            if (it == slot.end()) {
                slot.emplace(samples[j], (int)occ.size());
                occ.emplace_back(1, j);
                who.push_back(samples[j]);
            } else {
                occ[it->second].push_back(j);
            }
        }
        bool ok = true;
        win_end.assign(n_points, iNumSamples);
        for (size_t k = 0; k < occ.size(); ++k) {
            int cnt = 0;
            while (cnt < (int)occ[k].size() && occ[k][cnt] < iNumSamples + cnt) ++cnt;
            win_end[who[k]] = iNumSamples + cnt;
            if (iNumSamples + cnt > T) ok = false;
        }
        if (ok) break;
        T *= 2;
    }
}

// src/sDbscan.cpp:1131-1190
void sDbscan::fit_sngDbscan(const float* X, float eps, int minPts) {
    if (verbose) print_params(eps, minPts);
    clusterNoise = 0;  // the reference does not cluster noise for sngDbscan (:1159)
    prepare(X);
    std::vector<int> samples, win_end;
    sng_samples(false, samples, win_end);
    const bool exact = win_end.empty() && samplingProb >= 1.0f;
    check(sdb_sng_find_core(ctx, samples.data(), (int)samples.size(), win_end.empty() ? nullptr : win_end.data(),
                            exact ? 0 : 1, eps, minPts, nullptr, nullptr),
          "sng_find_core");
    labels_.assign(n_points, -1);
    check(sdb_form_clusters(ctx, labels_.data(), &n_clusters_), "form_clusters");
    if (verbose) {
        sdb_stats st{};
        sdb_get_stats(ctx, &st);
        std::cout << "Number of core points: " << st.n_core << std::endl;
        std::cout << "Find core points time = " << (long long)st.t_core_ms << "[ms]" << std::endl;
        std::cout << "Form clusters (including clustering noise) time = " << (long long)st.t_cluster_ms << "[ms]" << std::endl;
        std::ofstream f(output + "_" + distance + "_Eps_" + int2str(std::lround(1000 * eps)) + "_MinPts_" +
                        int2str(minPts) + "_Prob_" + int2str(std::lround(1000 * samplingProb)));
        for (int v : labels_) f << v << '\n';
    }
}

// src/sDbscan.cpp:1192-1252: one sampled neighbourhood per eps
void sDbscan::test_sngDbscan(const float* X, float eps, float rangeEps, int n_tests, int minPts) {
    std::cout << "base eps: " << eps << std::endl;
    std::cout << "range eps: " << rangeEps << std::endl;
    std::cout << "n_tests: " << n_tests << std::endl;
    const bool v = verbose;
    for (int i = 0; i < n_tests; ++i) {
        const float e = sweep_eps(eps, rangeEps, i);
        std::cout << "Eps: " << e << std::endl;
        verbose = true;
        fit_sngDbscan(X, e, minPts);
    }
    verbose = v;
}

// src/sDbscan.cpp:1846-1886
void sDbscan::fit_sngOptics(const float* X, float eps, int minPts) {
    if (verbose) print_params(eps, minPts);
    prepare(X);
    std::vector<int> samples, win_end;
    sng_samples(true, samples, win_end);
    std::vector<uint8_t> core(n_points);
    check(sdb_sng_find_core(ctx, samples.data(), (int)samples.size(), nullptr, 3, eps, minPts, core.data(), nullptr),
          "sng_find_core");
    optics_order(core, minPts);
    if (!output.empty()) {
        std::ofstream f(output + "_" + distance + "_Eps_" + int2str(std::lround(1000 * eps)) + "_MinPts_" +
                        int2str(minPts) + "_Prob_" + int2str(std::lround(1000 * samplingProb)));
        for (int k = 0; k < n_points; ++k) f << ordering_[k] << " " << reachability_[k] << '\n';
    }
}

void sDbscan::load_test_sngDbscan(const std::string& dataset, float eps, float rangeEps, int n_tests, int minPts) {
    std::vector<float> X = load_txt(dataset);
    test_sngDbscan(X.data(), eps, rangeEps, n_tests, minPts);
}

void sDbscan::load_fit_sngOptics(const std::string& dataset, float eps, int minPts) {
    std::vector<float> X = load_txt(dataset);
    fit_sngOptics(X.data(), eps, minPts);
}

sdb_stats sDbscan::get_stats() const {
    sdb_stats s{};
    if (ctx) sdb_get_stats(ctx, &s);
    return s;
}

void sDbscan::get_index(std::vector<int>& topK_out, std::vector<int>& topM_out) {
    ensure_ctx();
    topK_out.assign((size_t)n_points * 2 * topK, 0);
    topM_out.assign((size_t)n_proj * 2 * topM, 0);
    check(sdb_get_index(ctx, topK_out.data(), topM_out.data()), "get_index");
}
