// sdb_main.cpp -- command-line front end `sdbscan` of the drop-in class: the
// reference's src/main.cpp:15-85 with its argument reader readParam_sDbscan
// (src/Utilities.cpp:109-468) and its text loader loadtxtData
// (src/Utilities.cpp:60-95).  Same flags, defaults, printed lines and exit
// codes; the fits run on the device through the class.
//
//   sdbscan --n_points N --n_features D --eps E --minPts P --dataset FILE
//           [--distance Cosine|L1|L2|Chi2|JS] [--topK K] [--topM M] [--n_proj P]
//           [--ker_n_features F] [--ker_sigma S] [--ker_intervalSampling L]
//           [--clusterNoise C] [--samplingProb p] [--output PREFIX] [--verbose]
//           [--n_threads T] [--random_seed S]
//   extensions: [--precision FP32|BF16] [--device G]
//
// main runs ten fits of sDbscan and ten of sngDbscan at eps, eps + 0.01, ...,
// eps + 0.09, printing the wall clock of each (src/main.cpp:53-74).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sDbscan.h"

namespace {

// Fields of the reference's sDbscanParam (src/Header.h:77-98).  The values the
// reader leaves unset for some distances (kernel sigma for Cosine / Chi2 / JS,
// interval sampling for Cosine / L1 / L2) take the class defaults here.
struct CliParams {
    int n_points = 0;
    int n_features = 0;
    float eps = 0.0f;
    int minPts = 0;
    int n_proj = 0;
    int n_threads = -1;
    int topPoints = 0;   // topM
    int topVectors = 0;  // topK
    float ker_sigma = 1.0f;
    int ker_n_features = 0;
    float ker_intervalSampling = 0.4f;
    float samplingProb = 0.01f;
    bool verbose = false;
    std::string output;
    int clusterNoise = 0;
    std::string distance = "Cosine";
    int seed = -1;
    std::string precision = "FP32";
    int device = 0;
};

[[noreturn]] void die(const std::string& msg) {
    std::cerr << msg << std::endl;
    std::exit(1);
}

// Value following the first occurrence of `flag`, or nullptr.  The reference
// scans args[1..] for the first match and reads args[i + 1]; a flag given as
// the last argument has no value and counts as missing here.
const char* value_of(int nargs, char** args, const char* flag) {
    for (int i = 1; i < nargs; ++i)
        if (std::strcmp(args[i], flag) == 0) return i + 1 < nargs ? args[i + 1] : nullptr;
    return nullptr;
}

bool present(int nargs, char** args, const char* flag) {
    for (int i = 1; i < nargs; ++i)
        if (std::strcmp(args[i], flag) == 0) return true;
    return false;
}

int next_pow2_at_least_256(int x) {
    const int e = (int)std::ceil(std::log2(1.0 * x));
    return std::max(256, 1 << e);
}

// readParam_sDbscan (src/Utilities.cpp:109-468): required values first
// (points, features, minPts, eps), then the optional ones with the defaults
// they derive from the distance.
CliParams read_params(int nargs, char** args) {
    if (nargs < 6) die("Error: Not enough parameters !");
    CliParams p;
    const char* v;

    if (!(v = value_of(nargs, args, "--n_points"))) die("Error: Number of rows/points is missing !");
    p.n_points = std::atoi(v);
    std::cout << "Number of rows/points of X: " << p.n_points << std::endl;

    if (!(v = value_of(nargs, args, "--n_features"))) die("Error: Number of columns/dimensions is missing !");
    p.n_features = std::atoi(v);
    std::cout << "Number of columns/dimensions: " << p.n_features << std::endl;

    if (!(v = value_of(nargs, args, "--minPts"))) die("Error: minPts is missing !");
    p.minPts = std::atoi(v);
    std::cout << "minPts: " << p.minPts << std::endl;

    if (!(v = value_of(nargs, args, "--eps"))) die("Error: Eps is missing !");
    p.eps = (float)std::atof(v);
    std::cout << "Radius eps: " << p.eps << std::endl;

    if ((v = value_of(nargs, args, "--clusterNoise"))) {
        p.clusterNoise = std::atoi(v);
        std::cout << "We cluster noisy points." << std::endl;
    } else {
        std::cout << "Default: We do not cluster the noisy points." << std::endl;
    }

    if (present(nargs, args, "--verbose")) {
        p.verbose = true;
        std::cout << "verbose = true." << std::endl;
    }

    if ((v = value_of(nargs, args, "--distance"))) {
        static const char* const known[][2] = {{"Cosine", "Cosine distance - no kernel embedding"},
                                               {"L1", "L1 distance"},
                                               {"L2", "L2 distance"},
                                               {"Chi2", "Chi2 distance"},
                                               {"JS", "Jensen-Shannon distance"}};
        bool found = false;
        for (const auto& k : known)
            if (std::strcmp(v, k[0]) == 0) {
                p.distance = k[0];
                std::cout << k[1] << std::endl;
                found = true;
            }
        if (!found) {
            std::cout << "Use default cosine distance" << std::endl;
            p.distance = "Cosine";
        }
    } else {
        std::cout << "Distance is missing so we use default cosine distance" << std::endl;
    }
    const bool l12 = p.distance == "L1" || p.distance == "L2";
    const bool divergence = p.distance == "Chi2" || p.distance == "JS";

    if ((v = value_of(nargs, args, "--topK"))) {
        p.topVectors = std::atoi(v);
        std::cout << "TopK closest/furthest vectors: " << p.topVectors << std::endl;
    } else {
        p.topVectors = 5;
        std::cout << "TopK is missing. Use default topK: " << p.topVectors << std::endl;
    }

    if ((v = value_of(nargs, args, "--topM"))) {
        p.topPoints = std::atoi(v);
        std::cout << "TopM: " << p.topPoints << std::endl;
    } else {
        p.topPoints = p.minPts;
        std::cout << "TopM is missing. Use default TopM = minPts = " << p.topPoints << std::endl;
    }

    if ((v = value_of(nargs, args, "--ker_n_features"))) {
        p.ker_n_features = std::atoi(v);
        std::cout << "Number of kernel embedded dimensions: " << p.ker_n_features << std::endl;
        std::cout << "If using L1 and L2, it must be an even number. " << std::endl;
    } else {
        p.ker_n_features = l12 ? 2 * p.n_features : p.n_features;
        std::cout << "Kernel features is missing. Use default number of kernel features: " << p.ker_n_features
                  << std::endl;
    }

    if ((v = value_of(nargs, args, "--n_proj"))) {
        p.n_proj = std::atoi(v);
        std::cout << "Number of projections: " << p.n_proj << std::endl;
    } else {
        p.n_proj = next_pow2_at_least_256(p.distance == "Cosine" ? p.n_features : p.ker_n_features);
        std::cout << "Number of projections is missing. Use number of projections: " << p.n_proj << std::endl;
    }

    if ((v = value_of(nargs, args, "--ker_sigma"))) {
        p.ker_sigma = (float)std::atof(v);
        std::cout << "Sigma: " << p.ker_sigma << std::endl;
    } else if (p.distance == "L1") {
        p.ker_sigma = p.eps;
        std::cout << "Sigma is missing. Use default sigma = eps for L1: " << p.ker_sigma << std::endl;
    } else if (p.distance == "L2") {
        p.ker_sigma = 2 * p.eps;
        std::cout << "Sigma is missing. Use default sigma = 2 * eps for L2: " << p.ker_sigma << std::endl;
    }

    if ((v = value_of(nargs, args, "--ker_intervalSampling"))) {
        p.ker_intervalSampling = (float)std::atof(v);
        std::cout << "Sampling ratio for divergence distance: " << p.ker_intervalSampling << std::endl;
    } else if (divergence) {
        p.ker_intervalSampling = 0.4f;
        std::cout << "Interval sampling ratio is missing. Use default sampling ratio for Chi2 and JS distances: "
                  << p.ker_intervalSampling << std::endl;
    }

    if ((v = value_of(nargs, args, "--samplingProb"))) {
        p.samplingProb = (float)std::atof(v);
        std::cout << "Sampling ratio for sngDbscan or sDbscan-1NN: " << p.samplingProb << std::endl;
    } else {
        p.samplingProb = 0.01f;
        std::cout << "Sampling probability is missing. Use default sampling ratio: " << p.samplingProb << std::endl;
    }

    if ((v = value_of(nargs, args, "--output"))) p.output = v;
    if (p.output.empty()) std::cout << "No output file" << std::endl;

    if ((v = value_of(nargs, args, "--n_threads"))) {
        p.n_threads = std::atoi(v);
        std::cout << "Number of threads: " << p.n_threads << std::endl;
    } else {
        p.n_threads = -1;
        std::cout << "Number of threads is missing. Use all threads: " << p.n_threads << std::endl;
    }

    if ((v = value_of(nargs, args, "--random_seed"))) {
        p.seed = std::atoi(v);
        std::cout << "Seed: " << p.seed << std::endl;
    } else {
        p.seed = -1;
        std::cout << "Use a random seed: " << p.seed << std::endl;
    }

    // extensions of this build
    if ((v = value_of(nargs, args, "--precision"))) p.precision = v;
    if ((v = value_of(nargs, args, "--device"))) p.device = std::atoi(v);
    return p;
}

// loadtxtData (src/Utilities.cpp:60-95): n lines of d floats read with
// fscanf("%f") (values the file lacks stay 0), then transformData of the
// distance, here on the device.
std::vector<float> load_text(sDbscan& db, const std::string& dataset, int n, int d) {
    FILE* f = std::fopen(dataset.c_str(), "r");
    if (!f) die("Error: Data file does not exist !");
    std::vector<float> X((size_t)n * d, 0.0f);
    for (size_t i = 0; i < X.size(); ++i)
        if (std::fscanf(f, "%f", &X[i]) != 1) break;
    std::fclose(f);
    std::cout << "Finish reading data" << std::endl;
    std::cout << "Now checking the condition of data given the distance." << std::endl;
    db.transform_data(X.data());
    return X;
}

long long ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int nargs, char** args) {
    std::srand((unsigned)std::time(nullptr));  // rand() of sDbscan-1NN (src/main.cpp:17)
    const CliParams prm = read_params(nargs, args);
    try {
        sDbscan dbscan(prm.n_points, prm.n_features);
        dbscan.set_params(prm.n_proj, prm.topVectors, prm.topPoints, prm.distance, prm.ker_n_features,
                          prm.ker_sigma, prm.ker_intervalSampling, prm.clusterNoise, prm.samplingProb,
                          prm.verbose, prm.output, prm.n_threads, prm.seed);
        dbscan.set_precision(prm.precision);
        dbscan.set_device(prm.device);

        const char* dataset = value_of(nargs, args, "--dataset");
        if (!dataset || !*dataset) die("Error: Data file does not exist !");
        std::vector<float> X = load_text(dbscan, dataset, prm.n_points, prm.n_features);

        const float fRangeEps = 0.01f;
        for (int i = 0; i < 10; ++i) {
            const float new_eps = prm.eps + i * fRangeEps;  // float arithmetic, as src/main.cpp:55

            dbscan.clear();
            std::cout << "------------------" << std::endl;
            auto t0 = std::chrono::steady_clock::now();
            dbscan.fit_sDbscan(X.data(), new_eps, prm.minPts);
            std::cout << "sDBSCAN Wall Clock = " << ms_since(t0) << "[ms]" << std::endl;

            dbscan.clear();
            std::cout << "------------------" << std::endl;
            t0 = std::chrono::steady_clock::now();
            dbscan.fit_sngDbscan(X.data(), new_eps, prm.minPts);
            std::cout << "sngDBSCAN Wall Clock = " << ms_since(t0) << "[ms]" << std::endl;
        }
    } catch (const std::exception& e) {
        // the reference reports these with cerr and exit(1)
        die(e.what());
    }
    return 0;
}
