// Kernel argument structs and host launchers of the sDbscan HIP path.
#pragma once
#include "sdb_common.hpp"

namespace sdb {

constexpr int kMaxW = 8;          // top-M lists up to 256 points (8 x 32-bit words)
constexpr int kLdsBytes = 163840;  // LDS per workgroup on gfx950
constexpr int kGenBoost = 20;      // boosted bins per component of the Dirichlet generator

struct IndexArgs {
    const float* X;   // n x d input rows (cosine passes read and transform them)
    const float* Xn;  // n x ldx transformed rows (other metrics; row replay)
    float* Xn_out;    // cosine full pass writes the transformed rows here
    uint16_t* Xb_out; // and their BF16 copy (optional)
    int* err;         // bit 2: zero point
    const uint32_t* signmask;  // [3][NW][64] per-lane sign bits (lanes in the mirror layout)
    long long sample_stride;
    long long n;
    int ldx, d, metric, ker;
    float kerL;        // ker_intervalSampling
    const float* R;    // (ker/2) x d column-major (L1/L2), else null
    const uint8_t* bits;  // 3 x D sign bits
    int P, K;
    int* topK;         // n x 2K (mode 0 output)
    int* tie_count;
    int tie_cap;       // points whose values are kept
    int* tie_idx;      // n: every point with a tied top-K
    float* tie_vals;   // tie_cap x P
    uint32_t* chunk_max;  // G x P keys
    uint32_t* chunk_min;
    const uint32_t* thr_hi;  // P keys
    const uint32_t* thr_lo;
    int* cand_cnt;     // 2P (row * 2 + side)
    int cand_cap;
    uint32_t* cand_key;
    int* cand_idx;
    const int* row_slot;  // P (mode 2)
    float* row_buf;       // slots x n (mode 2)
    const int* pt_ids;    // points whose projections are written (mode 3)
    long long n_ids;
    float* pt_buf;        // P x n_ids (mode 3)
    const float* emb;     // n x ldE precomputed L1/L2 embeddings (sample / full passes), or null
    int ldE;
    float* rowerr;        // n: BF16 rounding bound of each row (full pass with Xb_out), or null
    uint32_t* rowerr_max; // its maximum (float bits), or null
    int rho_pad;          // 1: each BF16 row also carries its bound, rounded up, in element ldx - 1
};

struct DistArgs {
    const void* X;       // rows (fp32 or bf16), stride ldx elements
    const float* X32;    // fp32 rows (recheck / VALU path)
    int ldx, d, metric;
    const uint32_t* grp;  // entries a * S + s grouped by list
    const int* tile_list; // per workgroup: list id
    const int* tile_beg;  // per workgroup: first group position
    const int* tile_end;  // per workgroup: one past last group position
    const int* lists;     // L2 x M point ids
    int M, W, S;
    const uint32_t* F;    // first-occurrence masks (n*S x W)
    uint32_t* E;          // hit masks (n*S x W)
    int* colsum;          // L2 x M
    unsigned long long* queue;  // (entry << 32) | column
    unsigned long long* qcount;  // 64-bit count of queued pairs (stored while < qcap)
    int qcap;
    float eps;
    // fast distance d_hat: hit iff d_hat < lo, queued for the recheck iff
    // lo <= d_hat <= hi (lo = -inf: the tiles only prefilter, never decide)
    double lo, hi;
    int kc;               // dims per LDS chunk
    int nchunk;
    int dk;               // dims covered by the tiles (multiple of the k-step)
    // single-chunk cosine tiles decide on the raw dot s: hit iff s > s_hit,
    // hit or inside the band iff s >= s_near (equivalent to the band test on
    // fl(1 - s), see cosine_cuts)
    float s_hit, s_near;
    // BF16 single-chunk tiles: per-pair band from the rows' BF16 rounding
    // bounds rho (bf16_row_err): hit iff s > th_hit + rho_c + rho_q, hit or
    // inside the band iff s >= th_near - rho_c - rho_q (rho_c: the largest
    // bound over the tile's list)
    const float* rowerr;
    float th_hit, th_near;
    // the BF16 rows carry their bound (rounded up to BF16) in their last
    // padding element, ldx - 1 (d < ldx): the tiles read it with the row and
    // zero it in the staged candidate rows, so the dot products stay exact
    int rho_pad;
    // L2 on the FP32 MFMA tiles: squared row norms (FP32) and the per-pair
    // decision on D = |x|^2 + |y|^2 - 2 x.y: hit iff D < eps2 - band, queued
    // iff not a hit and D <= eps2 + band, band = kappa (|x|^2 + |y|^2) + eps0
    const float* rnorm2;
    double eps2, kappa, eps0;
    unsigned long long magicS;  // ceil(2^40 / S): a = ent_point(ent, magicS)
    long long n_ent;            // entries (rows of E and F)
};

// Batched sweep (sdb_batch_sweep): one job = one dataset's index and its eps
// values; its FP32 Cosine tiles decide every eps in one pass.
constexpr int kSweepMaxEps = 16;
struct SweepJob {
    const float* X;         // transformed FP32 rows, stride ldx
    int ldx, d, dk, kc, nchunk;
    int M, W, S, L2;
    const uint32_t* grp;    // entries grouped by list
    const int* lists;       // L2 x M
    const uint32_t* F;      // first-occurrence masks
    uint32_t* E;            // n_eps slices of n_ent x W hit words
    int* colsum;            // n_eps slices of L2 x M
    long long n_ent;        // n * S
    unsigned long long magicS;
    int n_eps;
    float eps[kSweepMaxEps];                      // ascending
    float s_hit[kSweepMaxEps], s_near[kSweepMaxEps];  // cuts on the raw dot (cosine_cuts)
};
// tiles: int4 (job, list, first group position, one past the last)
int launch_sweep_tiles(const SweepJob* jobs, int maxM, const int4* tiles, int ntiles, int* head, uint4* queue,
                       unsigned long long* qcount, long long qcap, size_t lds_bytes, int grid, hipStream_t st);
int launch_sweep_recheck(const SweepJob* jobs, const uint4* queue, long long nq, unsigned long long* stats,
                         hipStream_t st);
int launch_fill_u32(uint32_t* p, long long count, uint32_t value, hipStream_t st);
// exclusive scan without library calls (graph-capturable); bsum: ceil(n / 1024) ints
int scan_exclusive_i32_graph(const int* in, int* out, long long n, int* total, int* bsum, hipStream_t st);
int launch_copy_i32(const int* src, int* dst, long long count, hipStream_t st);

// index
// embedding kind of the index kernels for (metric, n_features, ker_n_features):
// 0 and 1 fuse the transform into the index passes, 2 needs launch_transform
int index_embedding(int metric, int d, int ker);
// [cos(R x), sin(R x)] of every row (L1/L2 random features) into E (n x ldE)
int launch_rff_embed(const float* Xn, long long n, int d, int ldx, const float* R, int half, float* E, int ldE,
                     hipStream_t st);
int launch_transform(const float* X, long long n, int d, int ldx, int metric, float* Xn,
                     uint16_t* Xb, float* rowerr, uint32_t* rowerr_max, int rho_pad, int* err, hipStream_t st);
int launch_index(const IndexArgs& a, int logD, int mode, int grid, hipStream_t st);
int launch_scatter_rows(const int* idx, const int* rows, int count, int width, int* dst, hipStream_t st);
int launch_topm_threshold(const uint32_t* cmax, const uint32_t* cmin, int G, int P, int M,
                          uint32_t* thi, uint32_t* tlo, hipStream_t st);
int launch_topm_select(const uint32_t* key, const int* idx, const int* cnt, int cap, int P, int M,
                       int* topM, int* flags, hipStream_t st);
int launch_build_lists(const int* topM, int P, int M, int* lists, uint8_t* is_hub, hipStream_t st);
int launch_hub_count(const int* lists, int LM, const int* hub_id, int* hub_cnt, hipStream_t st);
int launch_hub_fill(const int* lists, int LM, const int* hub_id, const int* hub_ptr, int* hub_cur,
                    int* hub_mem, hipStream_t st);
int launch_invmask(const int* lists, const int* topK, int P, int K, int M, int L2, int W,
                   uint32_t* inv, hipStream_t st);
int launch_pair_count(const int* lists, const int* hub_id, const int* hub_ptr, int L2, int M, int* pos_hub,
                      int* pcnt, int* poff, int* total, void* tmp, size_t tmp_bytes, hipStream_t st);
int launch_pair_fill(const int* pos_hub, const int* hub_ptr, const int* hub_mem, int L2, int M, const int* poff,
                     uint32_t* pairs, hipStream_t st);
int launch_inter_fill(const int* pos_hub, const int* hub_ptr, const int* hub_mem, int L2, int M, int W,
                      uint32_t* inter, hipStream_t st);
int launch_first_mask(const int* topK, long long n, int P, int K, int M, int W, const int* poff,
                      uint32_t* pairs, int2* seg, const int* hub_id, const int* hub_ptr, const int* hub_mem,
                      const uint32_t* inter, uint32_t* F, uint8_t* full, hipStream_t st);
int launch_group_count(const int* topK, long long n, int P, int K, int L2, long long qblk, int nblk,
                       int Gx, int* wgcnt, hipStream_t st);
int launch_row_norm2(const float* X, long long n, int d, int ldx, float* out, uint32_t* maxbits, hipStream_t st);
int launch_queue_keys(const unsigned long long* queue, int nq, const int* lists, int M, uint32_t* keys,
                      hipStream_t st);
int launch_group_scatter(const int* topK, long long n, int P, int K, int L2, long long qblk, int nblk,
                         int Gx, const int* wgoff, const uint8_t* full, uint32_t* grp, int* grp_ptr, hipStream_t st);

// distances
int launch_dist_mfma(const DistArgs& a, int precision, int ntiles, hipStream_t st);
int launch_dist_valu(const DistArgs& a, int ntiles, hipStream_t st);

// scan helpers
int scan_exclusive_i32(const int* in, int* out, long long n, int* total_out_dev, void* tmp,
                       size_t tmp_bytes, hipStream_t st);
size_t scan_tmp_bytes(long long n);
int launch_is_hub_scan_prep(const uint8_t* is_hub, int* flags, long long n, hipStream_t st);
int launch_hub_id(const uint8_t* is_hub, const int* excl, int* hub_id, long long n, hipStream_t st);

// degrees, cores, clusters
struct ClusterArgs {
    long long n;
    int S, W, M, P, K, L2;
    const int* topK;
    const int* lists;
    const uint32_t* E;
    const int* colsum;
    const int* hub_id;
    const int* hub_ptr;
    const int* hub_mem;
    const uint32_t* inv;
    int minPts;
    int* degree;
    uint8_t* core;
    uint32_t* corehub;   // L2 x W: list positions holding core points
    int* parent;
    int* root_rank;
    int* border_min;
    int* labels;
    uint32_t* adj;       // L2 x M x W list-local co-hit adjacency
    unsigned long long* n_core;
    int spread;          // n_core copies (spread_add): kSpread - 1, or 0 for one counter
};
int launch_degree(const ClusterArgs& c, hipStream_t st);
int launch_corehub(const ClusterArgs& c, hipStream_t st);
int launch_uf_init(const ClusterArgs& c, hipStream_t st);
int launch_uf_lists(const ClusterArgs& c, hipStream_t st);
int launch_core_minpts(const int* degree, long long n, int minPts, uint8_t* core, unsigned long long* n_core,
                       int spread, hipStream_t st);
// *out += the kSpread copies of a spread counter
int launch_spread_fold(const unsigned long long* ctr, unsigned long long* out, hipStream_t st);
int launch_uf_points(const ClusterArgs& c, hipStream_t st);
int launch_uf_compress(const ClusterArgs& c, hipStream_t st);
int launch_root_flags(const ClusterArgs& c, int* flags, hipStream_t st);
int launch_border(const ClusterArgs& c, hipStream_t st);
int launch_labels(const ClusterArgs& c, const int* rank, hipStream_t st);

// Batched cluster chain: many independent settings (a dataset at one eps and
// one minPts) go through degree -> core -> union-find -> labels together, one
// launch per step for all of them.  A set carries its own ClusterArgs (its
// degree / core / parent / border / labels / adjacency buffers), root flags,
// ranks, scan block sums and the component count it writes.  The sets of a
// dataset are numbered eps-major (set eps_i * n_minpts + minpts_i), and its
// per-eps arrays are strided: E and colsum by one eps slice, degrees by n,
// core flags by n_minpts x n, core counts by n_minpts.  The dataset's first
// set counts the degrees of all its eps in one pass (kChDegree: the eps-free
// mutual-pair masks once per point; core test of the first minPts included);
// the sets of the other minPts test theirs on those degrees (kChCore).
struct ChainSet {
    ClusterArgs a;
    int* rootf;
    int* rank;
    int* bsum;   // ceil(n / 1024) block sums of the root-flag scan
    int* total;  // number of clusters (roots)
    int n_eps, n_minpts, eps_i, minpts_i;
};
enum ChainOp {
    kChDegree, kChCore, kChAdjClear, kChCorehub, kChUfInit, kChUfPoints, kChUfPointsStaged, kChUfLists,
    kChUfCompress, kChRootFlags, kChScanBlocks, kChScanTop, kChScanApply, kChBorder, kChLabels, kChOps
};
// Workgroups of one set in op's launch, and the workgroup size of op.
long long chain_set_blocks(int op, const ChainSet& x);
int chain_block_threads(int op);
// One step for all sets: pre[s] is the first workgroup of set s, pre[nset]
// the grid; lds_staged is the dynamic LDS of kChUfPointsStaged (128 points
// x the largest S x W x 4 bytes).
int launch_chain_op(int op, const ChainSet* sets, const unsigned* pre, int nset, unsigned grid, size_t lds_staged,
                    hipStream_t st);

}  // namespace sdb
