/* sdbscan_amd.h -- C ABI of the MI355X (gfx950) sDbscan neighbourhood-discovery path.
 *
 * The reference (NinhPham/sDbscan) runs this path inside the C++ class sDbscan
 * (src/sDbscan.h:10-223).  Its device-free stages are replaced one to one:
 *
 *   sdb_set_params     <- sDbscan::set_params / set_kernel_params (src/sDbscan.h:79-124)
 *   sdb_load_host      <- the copy matrix_X = X of fit_sDbscan (src/sDbscan.cpp:857)
 *   sdb_build_index    <- transformData (src/Utilities.h:289-335) + rp_parIndex
 *                         (src/sDbscan.cpp:19-209); sign bits and RFF matrix come from
 *                         the host RNG calls of src/Utilities.h:47-122
 *   sdb_find_core      <- rp_findCorePoints (src/sDbscan.cpp:217-408)
 *   sdb_form_clusters  <- formCluster (src/sDbscan.cpp:509-637)
 *   sdb_label_noise    <- labelNoise, sDbscan-1NN (src/sDbscan.cpp:685-807)
 *   sdb_sng_find_core  <- sng_findCorePoints / sng_findCoreDist (src/sDbscan.cpp:416-504, 1484-1583)
 *   sdb_core_dist      <- rp_findCoreDist (src/sDbscan.cpp:1314-1474)
 *   sdb_optics_reach   <- the reachability values of formOptics_scikit
 *                         (src/sDbscan.cpp:1588-1703) for a given ordering
 *   sdb_optics_order   <- formOptics_scikit (src/sDbscan.cpp:1588-1703): ordering
 *                         and reachability
 *   sdb_batch_run      <- test_sDbscan's eps sweep (src/sDbscan.cpp:980-1039) over
 *                         many datasets at once (no batch entry in the reference)
 *   sdb_get_index      <- the protected matrix_topK / matrix_topM (src/sDbscan.h:54-55)
 *   sdb_get_projections <- columns of the local MATRIX_FHT of rp_parIndex
 *                         (src/sDbscan.cpp:31, 74-91)
 *   sdb_transform      <- transformData (src/Utilities.h:289-335) as loadtxtData
 *                         calls it (src/Utilities.cpp:94)
 *
 * All entry points return 0 on success and a nonzero SDB_E* code otherwise;
 * sdb_last_error() describes the last failure.  No exception or exit() crosses
 * this boundary.  Matrices use the reference's layouts: X is d x n column-major
 * (n rows of d floats), topK is 2K x n column-major, topM is 2M x n_proj
 * column-major.  A context is not thread safe (like the reference object). */
#ifndef SDBSCAN_AMD_H
#define SDBSCAN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sdb_metric { SDB_COSINE = 0, SDB_L1 = 1, SDB_L2 = 2, SDB_CHI2 = 3, SDB_JS = 4 };
enum sdb_precision { SDB_FP32 = 0, SDB_BF16 = 1 };
enum sdb_status {
    SDB_OK = 0,
    SDB_EINVAL = 1,     /* bad argument or parameter combination */
    SDB_EHIP = 2,       /* HIP runtime error */
    SDB_ENOMEM = 3,     /* device allocation failed */
    SDB_ENEGATIVE = 4,  /* Chi2/JS input with a negative entry (src/Utilities.h:308-312) */
    SDB_EZERO = 5,      /* zero point: zero sum (src/Utilities.h:320-324) or zero norm */
    SDB_ESTATE = 6,     /* call order violated (e.g. find_core before build_index) */
    SDB_EUNSUPPORTED = 7
};

typedef struct sdb_ctx sdb_ctx;

typedef struct sdb_params {
    long long n_points;
    int n_features;
    int n_proj;            /* numProj */
    int topK;              /* k */
    int topM;              /* m */
    int metric;            /* enum sdb_metric */
    int ker_n_features;    /* kerDim (Cosine: n_features) */
    float ker_sigma;
    float ker_intervalSampling;
    int fhtDim;            /* 2^ceil(log2(max(n_proj, ker_n_features))) */
    int precision;         /* enum sdb_precision: distance tiles.  SDB_BF16 for Cosine and
                              Chi2 (Chi2: bounds on the BF16 square-root rows, the rest
                              rechecked); the decisions equal SDB_FP32's */
} sdb_params;

typedef struct sdb_stats {
    double t_transform_ms, t_index_ms, t_core_ms, t_cluster_ms;  /* phase times (HIP events) */
    double t_dist_ms;             /* distance kernel, HIP events on the context stream */
    long long n_pairs;            /* (point, candidate) slots evaluated by the tiles */
    long long n_queue;            /* pairs inside the error band of eps */
    long long n_recheck32, n_recheck64, n_recheck_hits;
    long long n_hubs;             /* distinct points over all top-M lists */
    long long n_tie_points;       /* points whose top-K was replayed on the host */
    long long n_tie_rows;         /* top-M rows replayed on the host */
    long long n_core;
    int n_components;
    double band;                  /* error band of the fast distance */
    double band32;                /* error band of the FP32 recheck */
    long long n_near_eps;         /* FP64-rechecked pairs within 1e-6 of eps: all such pairs
                                     whenever the band exceeds 1e-6 (Cosine, Chi2, JS; L1/L2
                                     when 2 eps (d + 8) 2^-23 > 1e-6) */
} sdb_stats;

/* Host random structures, the reference's std:: calls (no device work):
 * sdb_hd3_bits  <- bitHD3Generator (src/Utilities.h:47-64): n_bits bytes of
 *                  uniform_int_distribution<uint32_t>(0, 1) over
 *                  default_random_engine(seed);
 * sdb_rff_matrix <- gaussGenerator / cauchyGenerator (src/Utilities.h:76-122):
 *                  the (ker_n_features / 2) x d matrix R (column c holds
 *                  half consecutive values) of N(0, 1/sigma) (SDB_L2) or
 *                  Cauchy(0, 1/sigma) (SDB_L1) draws. */
int sdb_hd3_bits(int n_bits, unsigned int seed, uint8_t* out);
int sdb_rff_matrix(int metric, int half, int d, float sigma, unsigned int seed, float* out);

/* Context lifecycle (one HIP stream per context). */
int sdb_create(int device, sdb_ctx** out);
void sdb_destroy(sdb_ctx* ctx);
const char* sdb_last_error(const sdb_ctx* ctx);

/* Validates and stores the parameters; invalidates the index.  Parameters of
 * other n_points / n_features than the loaded input also drop that input
 * (SDB_ESTATE until the next load); the same dimensions keep it. */
int sdb_set_params(sdb_ctx* ctx, const sdb_params* p);

/* Input points: host buffer (copied) or device buffer (borrowed; must stay
 * valid until the next load or destroy).  X is n rows of d floats. */
int sdb_load_host(sdb_ctx* ctx, const float* X, long long n, int d);
int sdb_load_device(sdb_ctx* ctx, const float* X_dev, long long n, int d);

/* Transform + embedding + HD3 projections + top-K + top-M.  hd3_bits holds
 * 3 * fhtDim bytes (0/1) in the order of bitHD3Generator; R is the
 * (ker_n_features/2) x n_features column-major RFF matrix for L1/L2 (NULL
 * otherwise). */
int sdb_build_index(sdb_ctx* ctx, const uint8_t* hd3_bits, const float* R);

/* Neighbourhood counts and core flags for (eps, minPts); core_out (n bytes)
 * and degree_out (n ints) are optional.  At an unchanged eps only the core
 * test runs again.  Pairs inside the error band of eps are re-decided in
 * FP32/FP64; when more than 2^30 are queued (SDB_QUEUE_MAX lowers the limit)
 * the pass is redone on the FP32 tiles. */
int sdb_find_core(sdb_ctx* ctx, float eps, int minPts, uint8_t* core_out, int* degree_out);

/* Cluster labels after sdb_find_core; n_clusters as the reference reports
 * it (components + 1, or 2 when there is none). */
int sdb_form_clusters(sdb_ctx* ctx, int* labels_out, int* n_clusters_out);

/* sDbscan-1NN noise labelling after sdb_form_clusters (labelNoise with
 * clusterNoise = 1, src/sDbscan.cpp:685-807): the sampled core points are
 * projected again, and every point labelled -1 takes the label of the
 * sampled core maximising sum_k proj[close_k] - proj[far_k] (the first such
 * core in the given order).  sampled_core lists the sampled core points in
 * increasing index order, drawn by the caller as the reference draws them
 * (src/sDbscan.cpp:695-709).  labels_out (n ints) may be NULL. */
int sdb_label_noise(sdb_ctx* ctx, const int* sampled_core, int n_sampled, int* labels_out);

/* sngDbscan / sngOptics neighbourhoods without the index
 * (sng_findCorePoints, src/sDbscan.cpp:416-504; sng_findCoreDist,
 * src/sDbscan.cpp:1484-1583).  Point a is compared with samples[j] for
 * j < win_end[a] (n_samples for every point when win_end is NULL), skipping
 * the positions where samples[j] == a; a pair is a hit when its distance is
 * <= eps (decided exactly as in sdb_find_core).
 *   flags bit 0: a hit also puts a into the neighbourhood of samples[j]
 *                (the sampled variants; the exact variant sets it to 0);
 *   flags bit 1: deduplicate the neighbourhoods (sngOptics); otherwise
 *                |N(a)| counts repeated entries, as sngDbscan does.
 * Core iff |N(a)| >= minPts - 1.  Afterwards sdb_form_clusters,
 * sdb_get_neighbors, sdb_core_dist and sdb_optics_reach work on these
 * neighbourhoods; sdb_label_noise needs the index and is refused. */
int sdb_sng_find_core(sdb_ctx* ctx, const int* samples, int n_samples, const int* win_end, int flags,
                      float eps, int minPts, uint8_t* core_out, int* degree_out);

/* Index matrices in the reference layouts (either pointer may be NULL). */
int sdb_get_index(sdb_ctx* ctx, int* topK_out, int* topM_out);

/* HD3 projections of listed points after sdb_build_index (the columns of
 * MATRIX_FHT, src/sDbscan.cpp:74-91, recomputed on the device from the
 * transformed rows and the index's sign bits): out is n_proj x n_ids, column
 * q (n_proj consecutive floats) for point ids[q].  The values equal the
 * reference's fht_float chain as IEEE values (the sign of an exact zero may
 * differ). */
int sdb_get_projections(sdb_ctx* ctx, const int* ids, int n_ids, float* out);

/* Transformed points (n x n_features) as used by the distance tiles. */
int sdb_get_transformed(sdb_ctx* ctx, float* X_out);

/* transformData (src/Utilities.h:289-335) of the loaded input under the
 * parameters' metric, as loadtxtData applies it to a text dataset
 * (src/Utilities.cpp:94): Cosine rows divided by their norm, Chi2 / JS rows
 * by their sum (SDB_ENEGATIVE / SDB_EZERO with the reference's messages),
 * L1 / L2 rows unchanged.  Writes the n x n_features result to X_out; the
 * context keeps it as its transformed rows. */
int sdb_transform(sdb_ctx* ctx, float* X_out);

/* Their BF16 copy (n x n_features, round to nearest even) read by the BF16
 * distance tiles; for Chi2 the BF16 of their square roots (the tiles bound
 * chi2 from below by 1 - sum sqrt(x_i y_i) and queue the rest for the exact
 * recheck).  SDB_ESTATE unless the precision is SDB_BF16.  No counterpart in
 * the reference (it has no BF16 path). */
int sdb_get_transformed_bf16(sdb_ctx* ctx, uint16_t* X_out);

/* Deduplicated neighbourhoods after sdb_find_core as CSR sorted by index:
 * ptr has n+1 entries; idx/dist hold ptr[n] entries (dist = FP32 distance,
 * may be NULL).  Call with idx == NULL to get ptr only. */
int sdb_get_neighbors(sdb_ctx* ctx, long long* ptr, int* idx, float* dist);

/* sOPTICS: core distances (the (minPts-1)-th smallest neighbour distance of
 * each core point, 0 otherwise) after sdb_find_core. */
int sdb_core_dist(sdb_ctx* ctx, int minPts, float* core_dist_out);

/* Reachability of formOptics_scikit for a given ordering (n ints), computed
 * on the device from the neighbourhoods and core distances. */
int sdb_optics_reach(sdb_ctx* ctx, const int* ordering, float* reach_out);

/* Accuracy sample (no counterpart in the reference): n_samples candidate
 * pairs (a, b) drawn from the index after sdb_build_index (a uniform, one of
 * its 2 topK lists uniform, a position in it uniform; a itself skipped), with
 * the quantity the distance tiles test computed in the tiles' precision
 * (fast_out) and in FP64 from the FP32 rows (exact_out).  The quantity is the
 * distance, except for Chi2 with SDB_BF16, where it is 1 - sum sqrt(x_i y_i).
 * |fast - exact| is what the error band of sdb_stats.band bounds. */
int sdb_sample_pair_dist(sdb_ctx* ctx, long long n_samples, unsigned long long seed, int* a_out, int* b_out,
                         float* fast_out, double* exact_out);

/* The same sample without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The sampled pairs are
 * staged in a static device array of SDB_NOALLOC_MAX_SAMPLES x 20 bytes
 * (5 MiB) that every GPU the library runs on holds until the process ends;
 * larger n_samples are refused with SDB_EINVAL.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several contexts or
 * threads) race.  Results equal sdb_sample_pair_dist's for the same
 * arguments. */
#define SDB_NOALLOC_MAX_SAMPLES (1 << 18)
int sdb_sample_pair_dist_noalloc(sdb_ctx* ctx, long long n_samples, unsigned long long seed, int* a_out,
                                 int* b_out, float* fast_out, double* exact_out);

int sdb_get_stats(const sdb_ctx* ctx, sdb_stats* out);

/* Synthetic workloads generated on the device into the context's input
 * buffer (no host copy): kind 0 = mixture of 10 sparse nonnegative centres
 * plus N(0, sigma^2) noise, clipped to [0,1] and L2-normalised; kind 1 =
 * Dirichlet histograms over n_features bins from 10 components, each with 20
 * bins at alpha 2 and the rest at alpha 0.05 (sigma unused). */
int sdb_generate(sdb_ctx* ctx, int kind, long long n, int d, unsigned long long seed, float sigma);

/* Copy rows [row0, row0 + rows) of the current input to the host. */
int sdb_copy_input(sdb_ctx* ctx, long long row0, long long rows, float* out);

/* sOPTICS at the eps of the last sdb_find_core: formOptics_scikit
 * (src/sDbscan.cpp:1588-1703) walks the device neighbourhoods on the host
 * (ascending start points, lazy min-heap on reach) and writes the ordering
 * (n ints); reach_out (n floats, may be NULL) gets the reachability values
 * of sdb_optics_reach.  minPts as in sdb_core_dist (at least 2 is used). */
int sdb_optics_order(sdb_ctx* ctx, int minPts, int* ordering_out, float* reach_out);

/* The same for several minPts at the eps of the last sdb_find_core: the
 * neighbourhoods are fetched once and the host walks of the settings run on
 * parallel threads.  orderings (and reaches, may be NULL) hold n_minpts
 * consecutive runs of n values.  After sdb_sng_find_core only n_minpts = 1
 * is accepted (its core flags stand). */
int sdb_optics_orders(sdb_ctx* ctx, int n_minpts, const int* minPts, int* orderings, float* reaches);

/* Rows of the (point, list) entries per distance tile; 0 restores the
 * default for the row length.  No counterpart in the reference. */
int sdb_set_tile_rows(sdb_ctx* ctx, int rows);

/* Tile-shape autotuner: runs the distance pass of the built index at eps
 * once per candidate row count, keeps the fastest on the context and returns
 * it (and its kernel time).  The next sdb_find_core recomputes distances.
 * No counterpart in the reference. */
int sdb_tune_tile_rows(sdb_ctx* ctx, float eps, const int* candidates, int n_candidates, int* best_rows,
                       double* best_ms);

/* Batched sweeps (no counterpart in the reference, whose test_sDbscan sweeps
 * the eps values of one dataset, src/sDbscan.cpp:980-1039).  Each job is one
 * dataset with its own parameters and random structures, swept over n_eps
 * eps values x n_minpts minPts values (labels of every pair, as fit_sDbscan)
 * and, with optics set, sOPTICS at the largest eps for every minPts (as
 * fit_sOptics).  The index is built once per job, the distances once per eps
 * (the minPts settings reuse them).  n_streams contexts, one HIP stream each,
 * take jobs from a shared queue; with autotune set the first job of every
 * (n_features, metric, precision) class tunes the tile rows for the rest.
 * Output arrays are caller-owned and may be NULL. */
typedef struct sdb_batch_job {
    const float* X;            /* n x d host rows (the reference's d x n column-major) */
    sdb_params params;         /* n_points / n_features are n and d */
    const uint8_t* hd3_bits;   /* 3 * fhtDim sign bytes (bitHD3Generator) */
    const float* R;            /* (ker_n_features / 2) x d for L1 / L2, else NULL */
    int n_eps;
    const float* eps;          /* n_eps values */
    int n_minpts;
    const int* minPts;         /* n_minpts values */
    int optics;                /* 1: sOPTICS at max(eps) for every minPts */
    int* labels;               /* [n_eps][n_minpts][n] */
    int* n_clusters;           /* [n_eps][n_minpts] */
    int* ordering;             /* [n_minpts][n] (optics) */
    float* reach;              /* [n_minpts][n] (optics) */
    int status;                /* out: SDB_OK or the job's error code */
    int tile_rows;             /* out: distance tile rows used */
    double ms;                 /* out: wall time of the job */
    char error[192];           /* out: message when status != SDB_OK */
} sdb_batch_job;

/* Runs the jobs; returns SDB_OK when every job succeeded, else the first
 * failing job's code (each job keeps its own status).  wall_ms (may be NULL)
 * gets the wall time of the whole batch. */
int sdb_batch_run(int device, sdb_batch_job* jobs, int n_jobs, int n_streams, int autotune, double* wall_ms);

/* The same sweeps with the candidate distances of all jobs computed once, in
 * one persistent launch fed by a device work queue of (job, list, rows)
 * tiles that decides every eps of a job from the same FP32 MFMA dot products
 * (no counterpart in the reference, whose test_sDbscan recomputes them per
 * eps, src/sDbscan.cpp:980-1039).  Band pairs of every eps are re-decided in
 * FP64 in one launch; the degree -> core -> union-find -> labels chain of
 * every (job, eps, minPts) set runs one launch per step over all sets, the
 * steps captured as one hipGraph; sOPTICS as in sdb_batch_run.  Outputs equal sdb_batch_run's.  Takes Cosine jobs with
 * FP32 distances, topM <= 128 and at most 16 eps values (SDB_EUNSUPPORTED
 * otherwise, before any work).  n_streams host threads build the indexes and
 * run the per-job stages; with autotune set the tile rows of every row length
 * are picked by timing the sweep kernel.  stage_ms (SDB_SWEEP_STAGES doubles,
 * may be NULL), host milliseconds unless noted: context set-up, index build,
 * tile autotune, distance launch (device time), recheck, cluster buffers,
 * chain capture, chain run, label copy-out, sOPTICS, teardown. */
#define SDB_SWEEP_STAGES 11
int sdb_batch_sweep(int device, sdb_batch_job* jobs, int n_jobs, int n_streams, int autotune, double* stage_ms,
                    double* wall_ms);

#ifdef __cplusplus
}
#endif

#endif /* SDBSCAN_AMD_H */
