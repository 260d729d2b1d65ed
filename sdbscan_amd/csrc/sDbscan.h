// sDbscan.h -- drop-in C++ class for the reference's sDbscan (src/sDbscan.h:10-223)
// whose neighbourhood-discovery path runs on an MI355X through the C ABI of
// include/sdbscan_amd.h.
//
// Same class name, method names, parameter order and defaults.  The input
// matrix is the reference's d x n column-major float matrix passed as a raw
// pointer (n rows of d floats) instead of an Eigen::Ref, since Eigen is not a
// dependency here.  Errors that the reference reports with cerr + exit(1)
// (src/Utilities.h:308-324) throw std::runtime_error with the same message.
#ifndef SDBSCAN_AMD_CLASS_H
#define SDBSCAN_AMD_CLASS_H

#include <cstdint>
#include <string>
#include <vector>

#include "sdbscan_amd.h"

typedef std::vector<float> FVector;
typedef std::vector<int> IVector;

class sDbscan {
private:
    int n_rotate = 3;
    int fhtDim = 0;

protected:
    int n_points;
    int n_features;

    int n_proj = 1024;
    int topK = 10;
    int topM = 50;

    int n_threads = 8;

    float ker_sigma = 1.0;
    int ker_n_features = 1024;
    float ker_intervalSampling = 0.4;
    float samplingProb = 0.01;

    bool verbose = false;
    std::string output;
    int clusterNoise = 0;
    std::string distance = "Cosine";
    int seed = -1;

    // random structures made on the host with the reference's std:: calls
    std::vector<uint8_t> bitHD3;  // 3 x fhtDim sign bits (src/Utilities.h:47-64)
    std::vector<float> matrix_R;  // (ker_n_features/2) x n_features, column-major

    int device = 0;
    int precision = SDB_FP32;
    sdb_ctx* ctx = nullptr;

    void ensure_ctx();
    void check(int rc, const char* what);
    void prepare(const float* X);
    void make_random();
    void print_params(float eps, int minPts) const;
    std::string out_name(float eps, int minPts, bool optics) const;
    void label_noise(const std::vector<uint8_t>& core);
    std::vector<float> load_txt(const std::string& dataset) const;
    void optics_order(const std::vector<uint8_t>& core, int minPts);
    void sng_samples(bool optics, std::vector<int>& samples, std::vector<int>& win_end) const;

public:
    // Dbscan's output
    IVector labels_;
    int n_clusters_ = 0;

    // Optics's output
    IVector ordering_;
    FVector reachability_;

    sDbscan(int n, int d);
    ~sDbscan();
    sDbscan(const sDbscan&) = delete;
    sDbscan& operator=(const sDbscan&) = delete;

    void set_params(int numProj = 1024, int k = 5, int m = 50, std::string dist = "Cosine",
                    int kerDim = 1024, float kerSigma = 1.0, float kerSam = 0.4,
                    int noiseClustering = 0, float prob = 0.01, bool ver = false,
                    std::string filename = "", int numThreads = 8, int randomSeed = -1);
    void set_kernel_params(std::string dist = "L2", int kerDim = 1024, float kerSigma = 1.0,
                           float kerSam = 0.4);
    void set_sngParams(std::string dist = "Cosine", float prob = 0.01, int noiseClustering = 0,
                       bool ver = false, std::string filename = "", int numThreads = -1,
                       int randomSeed = -1);
    void clear();

    void set_topM(int m) { topM = m; }
    void set_topK(int k) { topK = k; }
    void set_proj(int p) { n_proj = p; }
    void set_clusterNoise(int c) { clusterNoise = c; }
    void set_threads(int t);

    // X: n_features x n_points column-major (n_points rows of n_features floats)
    void fit_sDbscan(const float* X, float eps, int minPts);
    void test_sDbscan(const float* X, float eps, float rangeEps, int n_tests, int minPts);
    void fit_sOptics(const float* X, float eps, int minPts);

    // Text datasets: n_points lines of n_features floats (src/Utilities.cpp:60-95)
    void load_fit_sDbscan(const std::string& dataset, float eps, int minPts);
    void load_test_sDbscan(const std::string& dataset, float eps, float rangeEps, int n_tests, int minPts);
    void load_fit_sOptics(const std::string& dataset, float eps, int minPts);

    // sngDbscan / sngOptics (sampled neighbourhoods, no index)
    void fit_sngDbscan(const float* X, float eps, int minPts);
    void test_sngDbscan(const float* X, float eps, float rangeEps, int n_tests, int minPts);
    void fit_sngOptics(const float* X, float eps, int minPts);
    void load_test_sngDbscan(const std::string& dataset, float eps, float rangeEps, int n_tests, int minPts);
    void load_fit_sngOptics(const std::string& dataset, float eps, int minPts);

    // eps of test i of an eps sweep: eps + 1.0 * i * rangeEps in double,
    // rounded once to float (src/sDbscan.cpp:1016, :1221)
    static float sweep_eps(float eps, float rangeEps, int i) { return (float)(eps + 1.0 * i * rangeEps); }

    // Extensions of this build (no counterpart in the reference)
    void set_precision(const std::string& p);  // "FP32" (default) or "BF16" distance tiles
    // transformData (src/Utilities.h:289-335) of an n_points x n_features host
    // matrix, in place, on the device: what loadtxtData applies to a text
    // dataset (src/Utilities.cpp:94) before main's fits transform it again
    void transform_data(float* X);
    void set_device(int dev);
    sdb_stats get_stats() const;
    void get_index(std::vector<int>& topK_out, std::vector<int>& topM_out);
    int get_topK() const { return topK; }
    int get_topM() const { return topM; }
    int get_n_proj() const { return n_proj; }
    int get_fhtDim() const { return fhtDim; }
    const std::vector<uint8_t>& get_bits() const { return bitHD3; }
    const std::vector<float>& get_R() const { return matrix_R; }
};

#endif
