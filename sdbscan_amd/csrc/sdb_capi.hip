// C ABI of the sDbscan HIP path: context, device buffers, pipeline order and
// the host replay of top-K / top-M selections whose boundary is tied.
#include <hipcub/hipcub.hpp>
#include <immintrin.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sdbscan_amd.h"
#include "sdb_ctx.hpp"
#include "sdb_kernels.hpp"

namespace sdb {
int launch_recheck_k(const DistArgs& a, const int* topK, int K, int P, int precision,
                     double band32, int n_entries, unsigned long long* stats, hipStream_t st);
int launch_gen_mixture(float* X, long long n, int d, unsigned long long seed, float sigma,
                       hipStream_t st);
int launch_pair_sample(const DistArgs& a, const int* topK, int K, int P, long long n, int precision,
                       unsigned long long seed, long long n_samples, int* pa, int* pb, float* fast,
                       double* exact, hipStream_t st);
int launch_gen_dirichlet(float* X, long long n, int d, unsigned long long seed, const uint16_t* bins,
                         hipStream_t st);
int launch_csr_fill(const ClusterArgs& c, const long long* ptr, int* cursor, int* idx,
                    hipStream_t st);
int launch_csr_dist(const float* X, int ldx, int d, int metric, long long n, const long long* ptr,
                    const int* idx, float* dist, int upper, hipStream_t st);
int launch_csr_mirror(long long n, const long long* ptr, const int* idx, float* dist, unsigned int* missing,
                      hipStream_t st);
int launch_reach(long long n, const long long* ptr, const int* idx, const float* dist,
                 const uint8_t* core, const float* core_dist, const int* pos, uint32_t* reach_key,
                 hipStream_t st);
int launch_kth(long long n, const long long* ptr, const float* dist, const uint8_t* core,
               int k, float* out, hipStream_t st);
int launch_noise_list(const int* labels, long long n, int* noise, int* count, hipStream_t st);
int launch_sng_fmask(long long n, int L, int S2, int Ms, int W, int n_samples, const int* win_end,
                     const int* occ_ptr, const int* occ_pos, uint32_t* F, hipStream_t st);
int launch_sng_grp(long long n, int L, int S2, uint32_t* grp, hipStream_t st);
int launch_sng_rev(int n_samples, const int* samples, const int* colsum, int* rev, hipStream_t st);
int launch_sng_degree(long long n, int L, int S2, int W, const uint32_t* E, const int* rev, int minPts, int* degree,
                      uint8_t* core, unsigned long long* n_core, hipStream_t st);
int launch_sng_uf(long long n, int L, int S2, int Ms, int W, const uint32_t* E, const int* samples,
                  const uint8_t* core, int* par, hipStream_t st);
int launch_sng_border(long long n, int L, int S2, int Ms, int W, const uint32_t* E, const int* samples,
                      const uint8_t* core, const int* par, int* border_min, hipStream_t st);
int launch_sng_csr_count(long long n, int L, int S2, int Ms, int W, int both_ways, const uint32_t* E,
                         const int* samples, long long* cnt, hipStream_t st);
int launch_sng_csr_fill(long long n, int L, int S2, int Ms, int W, int both_ways, const uint32_t* E,
                        const int* samples, const long long* ptr, unsigned long long* cur, int* idx, hipStream_t st);
int launch_uniq(long long n, const long long* ptr, const int* idx, long long* ucnt, hipStream_t st);
int launch_uniq_fill(long long n, const long long* ptr, const int* idx, const long long* uptr, int* uidx,
                     hipStream_t st);
int launch_rowsize_core(long long n, const long long* ptr, int minPts, int* degree, uint8_t* core,
                        unsigned long long* n_core, hipStream_t st);
int launch_noise_score(const float* cproj, long long nS, const int* topK, int K, const int* noise, int n_noise,
                       unsigned long long* best, hipStream_t st);
int launch_noise_assign(const int* noise, int n_noise, const unsigned long long* best, const int* sampled,
                        int* labels, hipStream_t st);
}  // namespace sdb

using namespace sdb;

// Scratch of sdb_sample_pair_dist_noalloc: 20 bytes per sampled pair (see
// sample_pairs_into), held by every GPU the library runs on until the process
// ends.  Host code never uses the name: each device's address comes from
// hipGetSymbolAddress in sdb_create.  File scope with external linkage (in
// namespace sdb): a static one is a local symbol of the device code object,
// which the runtime cannot resolve by name.
namespace sdb {
__device__ __align__(256) unsigned char scratch[(size_t)SDB_NOALLOC_MAX_SAMPLES * 20];
}  // namespace sdb

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};
bool g_scratch_done[64] = {};

// Device address of this device's scratch array (the device must be
// current); resolved once per device.
hipError_t scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    if (!g_scratch_done[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(sdb::scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = ptr;
        g_scratch_done[device] = true;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
}  // namespace

namespace {

// Pointer kind as the HIP runtime sees it, for error messages.
static std::string ptr_desc(const void* p) {
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, p);
    (void)hipGetLastError();
    char buf[160];
    snprintf(buf, sizeof(buf), "%p: %s, type %d, device %d", p, hipGetErrorString(e), (int)at.type, at.device);
    return buf;
}

// Heap selection of src/sDbscan.cpp:97-143 / 151-208: min-heaps of size k,
// elements inserted in index order, replaced only by a strictly larger value.
struct HPair {
    int idx;
    float val;
    HPair(int i, double v) : idx(i), val((float)v) {}
    bool operator>(const HPair& o) const { return val > o.val; }
};
typedef std::priority_queue<HPair, std::vector<HPair>, std::greater<HPair>> HMin;

void heap_select(const float* v, long long len, int k, int* out_close, int* out_far) {
    HMin close, far;
    for (long long j = 0; j < len; ++j) {
        const float x = v[j];
        if ((int)close.size() < k) close.emplace((int)j, x);
        else if (x > close.top().val) { close.pop(); close.emplace((int)j, x); }
        if ((int)far.size() < k) far.emplace((int)j, -x);
        else if (-x > far.top().val) { far.pop(); far.emplace((int)j, -x); }
    }
    for (int m = k - 1; m >= 0; --m) {
        out_close[m] = close.top().idx; close.pop();
        out_far[m] = far.top().idx; far.pop();
    }
}

int ilog2(int x) { int l = 0; while ((1 << l) < x) ++l; return l; }

}  // namespace

extern "C" {

int sdb_create(int device, sdb_ctx** out) {
    if (!out) return SDB_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { (void)hipGetLastError(); return SDB_EHIP; }
    if (device < 0 || device >= ndev) return SDB_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return SDB_EHIP;
    void* scr = nullptr;
    if (scratch_address(device, &scr) != hipSuccess) { (void)hipGetLastError(); return SDB_EHIP; }
    sdb_ctx* c = new sdb_ctx();
    c->device = device;
    c->scratch = scr;
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; return SDB_EHIP; }
    for (int i = 0; i < 8; ++i) hipEventCreate(&c->ev[i]);
    *out = c;
    return SDB_OK;
}

void sdb_destroy(sdb_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->st);
    DevBuf* all[] = {&c->Xown, &c->Xn, &c->Xb, &c->emb, &c->bits, &c->R, &c->topK, &c->tie_idx, &c->tie_vals, &c->tie_rows,
                     &c->tie_count, &c->cmax, &c->cmin, &c->thi, &c->tlo, &c->cand_cnt, &c->cand_key,
                     &c->cand_idx, &c->topM, &c->flags, &c->lists, &c->is_hub, &c->hub_flags,
                     &c->hub_excl, &c->hub_id, &c->hub_cnt, &c->hub_ptr, &c->hub_cur, &c->hub_mem,
                     &c->inv, &c->pos_hub, &c->pcnt, &c->poff, &c->pairs, &c->pseg, &c->F, &c->wgcnt, &c->wgoff, &c->grp_ptr, &c->grp,
                     &c->scan_tmp, &c->total, &c->err_flag, &c->row_slot, &c->row_buf, &c->E,
                     &c->colsum, &c->queue, &c->qcount, &c->qkeys, &c->qvals, &c->qsort_tmp, &c->inter, &c->core_spread, &c->ffull, &c->rnorm2, &c->rnorm2_max, &c->tiles, &c->degree, &c->core,
                     &c->stats64, &c->corehub, &c->parent, &c->border, &c->labels, &c->adj,
                     &c->rootf, &c->rank, &c->nb_deg, &c->nb_ptr, &c->nb_cur, &c->nb_idx,
                     &c->nb_dist, &c->nb_sorted, &c->core_dist, &c->pos, &c->reach, &c->ns_ids,
                     &c->cproj, &c->noise, &c->nbest, &c->pj_ids, &c->pj_buf, &c->rowerr, &c->rowerr_max, &c->sw_E, &c->sw_colsum, &c->sw_labels, &c->sw_comps, &c->occ_ptr, &c->occ_pos, &c->win, &c->rev,
                     &c->csr_cnt, &c->csr_ptr, &c->csr_cur, &c->csr_idx};
    for (DevBuf* b : all) release(*b);
    for (int i = 0; i < 8; ++i) hipEventDestroy(c->ev[i]);
    hipStreamDestroy(c->st);
    delete c;
}

const char* sdb_last_error(const sdb_ctx* c) { return c ? c->err.c_str() : "null context"; }

int sdb_set_params(sdb_ctx* c, const sdb_params* p) {
    if (!c || !p) return SDB_EINVAL;
    const sdb_params& q = *p;
    if (q.n_points <= 0 || q.n_features <= 0) return fail(c, SDB_EINVAL, "n_points and n_features must be positive");
    if (q.metric < 0 || q.metric > 4) return fail(c, SDB_EINVAL, "unknown metric %d", q.metric);
    if (q.topK <= 0 || q.topM <= 0 || q.n_proj <= 0) return fail(c, SDB_EINVAL, "topK, topM and n_proj must be positive");
    if (q.topK > q.n_proj) return fail(c, SDB_EINVAL, "topK (%d) exceeds n_proj (%d)", q.topK, q.n_proj);
    if (q.topM > q.n_points) return fail(c, SDB_EINVAL, "topM (%d) exceeds n_points", q.topM);
    if (q.topM > 32 * kMaxW) return fail(c, SDB_EUNSUPPORTED, "topM above %d is not supported", 32 * kMaxW);
    if (q.topK > 64) return fail(c, SDB_EUNSUPPORTED, "topK above 64 is not supported");
    if (q.fhtDim < 64 || q.fhtDim > 4096 || (q.fhtDim & (q.fhtDim - 1)))
        return fail(c, SDB_EUNSUPPORTED, "fhtDim %d outside [64, 4096]", q.fhtDim);
    if (q.n_proj > q.fhtDim || q.ker_n_features > q.fhtDim)
        return fail(c, SDB_EINVAL, "n_proj and ker_n_features must not exceed fhtDim");
    if (q.metric == SDB_COSINE && q.ker_n_features != q.n_features)
        return fail(c, SDB_EINVAL, "Cosine requires ker_n_features == n_features");
    if (q.precision != SDB_FP32 && q.precision != SDB_BF16) return fail(c, SDB_EINVAL, "unknown precision");
    if (q.precision == SDB_BF16 && q.metric != SDB_COSINE && q.metric != SDB_CHI2)
        return fail(c, SDB_EUNSUPPORTED, "BF16 distance tiles are implemented for Cosine and Chi2 only");
    // the (point, slot) entries are counted, scanned and grouped in signed
    // 32-bit offsets
    if ((long long)q.n_points * 2 * q.topK >= (1ll << 31)) return fail(c, SDB_EUNSUPPORTED, "n_points * 2 * topK must stay below 2^31");
    if (2 * q.n_proj >= (1 << 24)) return fail(c, SDB_EUNSUPPORTED, "n_proj too large");
    // an input loaded under other dimensions does not fit these parameters:
    // it has to be loaded again (its buffer is kept for reuse)
    if (c->X && (q.n_points != c->n || q.n_features != c->d)) {
        c->X = nullptr;
        c->sng = c->clustered = false;
    }
    // the transformed rows depend on the metric, the precision and the kernel
    // parameters: made again by the next pass that needs them
    c->transformed = false;
    c->rnorm2_of = nullptr;
    c->prm = q;
    c->have_params = true;
    c->indexed = false;
    c->cored = false;
    c->have_nb = false;
    return SDB_OK;
}

static int set_input_dims(sdb_ctx* c, long long n, int d) {
    if (!c->have_params) return fail(c, SDB_ESTATE, "set_params first");
    if (n != c->prm.n_points || d != c->prm.n_features)
        return fail(c, SDB_EINVAL, "input is %lld x %d but params say %lld x %d", n, d, c->prm.n_points, c->prm.n_features);
    c->n = n;
    c->d = d;
    c->ldx = (d + 31) / 32 * 32;
    c->indexed = c->cored = c->have_nb = false;
    c->transformed = c->sng = c->clustered = false;
    c->rnorm2_of = nullptr;
    return 0;
}

int sdb_load_host(sdb_ctx* c, const float* X, long long n, int d) {
    if (!c || !X) return SDB_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    int r = set_input_dims(c, n, d);
    if (r) return r;
    ENSURE(c->Xown, (size_t)n * d * 4);
    HIPCHK(hipMemcpyAsync(c->Xown.p, X, (size_t)n * d * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->X = P_<float>(c->Xown);
    return SDB_OK;
}

int sdb_load_device(sdb_ctx* c, const float* X, long long n, int d) {
    if (!c || !X) return SDB_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    int r = set_input_dims(c, n, d);
    if (r) return r;
    c->X = X;
    return SDB_OK;
}

int sdb_generate(sdb_ctx* c, int kind, long long n, int d, unsigned long long seed, float sigma) {
    if (!c) return SDB_EINVAL;
    if (kind != 0 && kind != 1) return fail(c, SDB_EINVAL, "unknown generator kind %d", kind);
    if (kind == 1 && (d < kGenBoost || d > 4096))
        return fail(c, SDB_EINVAL, "Dirichlet generator needs %d <= n_features <= 4096", kGenBoost);
    HIPCHK(hipSetDevice(c->device));
    int r = set_input_dims(c, n, d);
    if (r) return r;
    ENSURE(c->Xown, (size_t)n * d * 4);
    if (kind == 0) {
        LCHK(launch_gen_mixture(P_<float>(c->Xown), n, d, seed, sigma, c->st));
    } else {
        // boosted bins of each component: a partial Fisher-Yates shuffle
        std::mt19937_64 g(seed);
        std::vector<int> perm(d);
        std::vector<uint16_t> bins(10 * kGenBoost);
        for (int k = 0; k < 10; ++k) {
            for (int i = 0; i < d; ++i) perm[i] = i;
            for (int j = 0; j < kGenBoost; ++j) {
                const int t = j + (int)(g() % (uint64_t)(d - j));
                std::swap(perm[j], perm[t]);
                bins[k * kGenBoost + j] = (uint16_t)perm[j];
            }
        }
        LCHK(launch_gen_dirichlet(P_<float>(c->Xown), n, d, seed, bins.data(), c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->X = P_<float>(c->Xown);
    return SDB_OK;
}

int sdb_copy_input(sdb_ctx* c, long long row0, long long rows, float* out) {
    if (!c || !out || !c->X) return SDB_EINVAL;
    if (row0 < 0 || rows < 0 || row0 + rows > c->n) return fail(c, SDB_EINVAL, "row range outside input");
    HIPCHK(hipMemcpy(out, c->X + (size_t)row0 * c->d, (size_t)rows * c->d * 4, hipMemcpyDeviceToHost));
    return SDB_OK;
}

int sdb_build_index(sdb_ctx* c, const uint8_t* hd3_bits, const float* Rh) {
    if (!c || !hd3_bits) return SDB_EINVAL;
    if (!c->X) return fail(c, SDB_ESTATE, "no input loaded");
    HIPCHK(hipSetDevice(c->device));
    const sdb_params& q = c->prm;
    const long long n = c->n;
    const int d = c->d, D = q.fhtDim, P = q.n_proj, K = q.topK, M = q.topM;
    const int ldx = c->ldx;
    const bool l12 = q.metric == SDB_L1 || q.metric == SDB_L2;
    if (l12 && !Rh) return fail(c, SDB_EINVAL, "L1/L2 need the random feature matrix R");
    c->L2 = 2 * P;
    c->W = (M + 31) / 32;
    c->S = 2 * K;
    c->Mt = M;
    c->sng = false;
    c->stats = sdb_stats{};
    hipStream_t st = c->st;

    HIPCHK(hipEventRecord(c->ev[0], st));
    // ---- transform (src/Utilities.h:289-335); fused into the index passes for
    // Cosine and for the Chi2 / JS features without sampled components
    const bool fused = index_embedding(q.metric, d, q.ker_n_features) != 2;
    ENSURE(c->Xn, (size_t)n * ldx * 4);
    if (q.precision == SDB_BF16) ENSURE(c->Xb, (size_t)n * ldx * 2);
    ENSURE(c->err_flag, 4);
    HIPCHK(hipMemsetAsync(c->err_flag.p, 0, 4, st));
    const bool bf = q.precision == SDB_BF16;
    if (bf) {
        ENSURE(c->rowerr, (size_t)n * 4);
        ENSURE(c->rowerr_max, 4);
        HIPCHK(hipMemsetAsync(c->rowerr_max.p, 0, 4, st));
    }
    const int rho_pad = bf ? bf16_rho_pad(c) : 0;
    int errf = 0;
    if (!fused) {
        LCHK(launch_transform(c->X, n, d, ldx, q.metric, P_<float>(c->Xn), bf ? P_<uint16_t>(c->Xb) : nullptr,
                              bf ? P_<float>(c->rowerr) : nullptr, bf ? P_<uint32_t>(c->rowerr_max) : nullptr,
                              rho_pad, P_<int>(c->err_flag), st));
        HIPCHK(hipMemcpyAsync(&errf, c->err_flag.p, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (errf & 1) return fail(c, SDB_ENEGATIVE, "Error: X is not non-negative !");
        if (errf & 2) return fail(c, SDB_EZERO, "Error: There is an zero point !");
    }
    HIPCHK(hipEventRecord(c->ev[1], st));

    // ---- random structures: per-lane sign masks of the 3 x D sign bits (lane
    // p holds elements lane_logical(p) + 64 j in the index kernels).  Rounds 1
    // and 2 carry the pending sign of the previous FWHT of the index kernels:
    // lanes with odd popcount(lane_logical(p) & 15) hold negated outputs.
    const int VPL = D / 64, NWs = (VPL + 31) / 32;
    std::vector<uint32_t> smask((size_t)3 * NWs * 64, 0u);
    for (int r = 0; r < 3; ++r)
        for (int w = 0; w < NWs; ++w)
            for (int lane = 0; lane < 64; ++lane) {
                const uint32_t pend = r > 0 ? (uint32_t)(__builtin_popcount(lane_logical(lane) & 15) & 1) : 0u;
                uint32_t m = 0;
                for (int b = 0; b < 32 && w * 32 + b < VPL; ++b)
                    m |= ((uint32_t)(hd3_bits[(size_t)r * D + lane_logical(lane) + 64 * (w * 32 + b)] & 1u) ^ pend) << b;
                smask[((size_t)r * NWs + w) * 64 + lane] = m;
            }
    ENSURE(c->bits, smask.size() * 4);
    HIPCHK(hipMemcpyAsync(c->bits.p, smask.data(), smask.size() * 4, hipMemcpyHostToDevice, st));
    if (l12) {
        ENSURE(c->R, (size_t)(q.ker_n_features / 2) * d * 4);
        HIPCHK(hipMemcpyAsync(c->R.p, Rh, (size_t)(q.ker_n_features / 2) * d * 4, hipMemcpyHostToDevice, st));
        // the embeddings of all rows at once (a tiled FP32 product with the
        // same sequential sums), read as rows by the index passes
        const int half = q.ker_n_features / 2;
        ENSURE(c->emb, (size_t)n * D * 4);
        if (2 * half < D) HIPCHK(hipMemsetAsync(c->emb.p, 0, (size_t)n * D * 4, st));
        LCHK(launch_rff_embed(P_<float>(c->Xn), n, d, ldx, P_<float>(c->R), half, P_<float>(c->emb), D, st));
    }

    // ---- sample pass: top-M thresholds from every stride-th point
    const long long stride = n >= 65536 ? 8 : 1;
    const long long npts = (n + stride - 1) / stride;
    const int Gs = (int)std::max<long long>(1, std::min<long long>(1024, (npts + 3) / 4));
    const int Gf = (int)std::max<long long>(1, std::min<long long>(8192, (n + 3) / 4));
    // values kept for up to 65,536 tied points per pass (SDB_TIE_CAP lowers
    // it); further tied points are projected again in batches of that size
    long long tie_keep = 65536;
    if (const char* e = std::getenv("SDB_TIE_CAP")) tie_keep = std::max(1ll, std::min(tie_keep, std::atoll(e)));
    const int tie_cap = (int)std::min<long long>(n, tie_keep);
    ENSURE(c->topK, (size_t)n * 2 * K * 4);
    ENSURE(c->tie_idx, (size_t)n * 4);
    ENSURE(c->tie_vals, (size_t)tie_cap * P * 4);
    ENSURE(c->tie_count, 4);
    ENSURE(c->cmax, (size_t)Gs * P * 4);
    ENSURE(c->cmin, (size_t)Gs * P * 4);
    ENSURE(c->thi, (size_t)P * 4);
    ENSURE(c->tlo, (size_t)P * 4);
    const int cap = 4096;
    ENSURE(c->cand_cnt, (size_t)2 * P * 4);
    ENSURE(c->cand_key, (size_t)2 * P * cap * 4);
    ENSURE(c->cand_idx, (size_t)2 * P * cap * 4);
    ENSURE(c->topM, (size_t)P * 2 * M * 4);
    ENSURE(c->flags, (size_t)2 * P * 4);
    HIPCHK(hipMemsetAsync(c->tie_count.p, 0, 4, st));
    HIPCHK(hipMemsetAsync(c->cand_cnt.p, 0, (size_t)2 * P * 4, st));

    IndexArgs ia{};
    ia.X = c->X;
    ia.Xn = P_<float>(c->Xn);
    ia.Xn_out = P_<float>(c->Xn);
    ia.Xb_out = bf ? P_<uint16_t>(c->Xb) : nullptr;
    ia.rowerr = bf ? P_<float>(c->rowerr) : nullptr;
    ia.rowerr_max = bf ? P_<uint32_t>(c->rowerr_max) : nullptr;
    ia.rho_pad = rho_pad;
    ia.err = P_<int>(c->err_flag);
    ia.signmask = P_<uint32_t>(c->bits);
    ia.sample_stride = stride;
    ia.n = n; ia.ldx = ldx; ia.d = d; ia.metric = q.metric; ia.ker = q.ker_n_features;
    ia.kerL = q.ker_intervalSampling;
    ia.R = l12 ? P_<float>(c->R) : nullptr;
    ia.emb = l12 ? P_<float>(c->emb) : nullptr;
    ia.ldE = D;
    ia.P = P; ia.K = K;
    ia.topK = P_<int>(c->topK);
    ia.tie_count = P_<int>(c->tie_count); ia.tie_cap = tie_cap;
    ia.tie_idx = P_<int>(c->tie_idx); ia.tie_vals = P_<float>(c->tie_vals);
    ia.chunk_max = P_<uint32_t>(c->cmax); ia.chunk_min = P_<uint32_t>(c->cmin);
    ia.thr_hi = P_<uint32_t>(c->thi); ia.thr_lo = P_<uint32_t>(c->tlo);
    ia.cand_cnt = P_<int>(c->cand_cnt); ia.cand_cap = cap;
    ia.cand_key = P_<uint32_t>(c->cand_key); ia.cand_idx = P_<int>(c->cand_idx);
    const int logD = ilog2(D);
    LCHK(launch_index(ia, logD, 0, Gs, st));
    LCHK(launch_topm_threshold(ia.chunk_max, ia.chunk_min, Gs, P, M, P_<uint32_t>(c->thi), P_<uint32_t>(c->tlo), st));
    // ---- full pass: (transform,) projections, top-K, top-M candidates
    LCHK(launch_index(ia, logD, 1, Gf, st));
    int ntie = 0;
    HIPCHK(hipMemcpyAsync(&ntie, c->tie_count.p, 4, hipMemcpyDeviceToHost, st));
    if (bf) HIPCHK(hipMemcpyAsync(&c->rho_max, c->rowerr_max.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&errf, c->err_flag.p, 4, hipMemcpyDeviceToHost, st));
    LCHK(launch_topm_select(ia.cand_key, ia.cand_idx, ia.cand_cnt, cap, P, M, P_<int>(c->topM), P_<int>(c->flags), st));
    std::vector<int> hflags(2 * P);
    HIPCHK(hipMemcpyAsync(hflags.data(), c->flags.p, (size_t)2 * P * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (errf & 1) return fail(c, SDB_ENEGATIVE, "Error: X is not non-negative !");
    if (errf & 2) return fail(c, SDB_EZERO, "Error: There is an zero point !");

    // ---- host replay of tied top-K points (reference heap semantics).  The
    // values of the first tie_cap points come with the full pass; the rest
    // are projected again (point mode) tie_cap at a time.  The replayed rows
    // go back through one scatter per batch.
    if (ntie > 0) {
        std::vector<int> tidx(ntie);
        HIPCHK(hipMemcpy(tidx.data(), c->tie_idx.p, (size_t)ntie * 4, hipMemcpyDeviceToHost));
        const int nbmax = std::min(ntie, tie_cap);
        ENSURE(c->tie_rows, (size_t)nbmax * 2 * K * 4);
        std::vector<float> tv((size_t)nbmax * P);
        std::vector<float> col(P);
        std::vector<int> rows((size_t)nbmax * 2 * K);
        for (int t0 = 0; t0 < ntie; t0 += tie_cap) {
            const int nb = std::min(tie_cap, ntie - t0);
            if (t0 > 0) {
                IndexArgs ib = ia;
                ib.pt_ids = P_<int>(c->tie_idx) + t0;
                ib.n_ids = nb;
                ib.pt_buf = P_<float>(c->tie_vals);  // P x nb, row e at e * nb
                const int G = (int)std::max<long long>(1, std::min<long long>(2048, ((long long)nb + 3) / 4));
                LCHK(launch_index(ib, logD, 3, G, st));
            }
            HIPCHK(hipMemcpyAsync(tv.data(), c->tie_vals.p, (size_t)nb * P * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int t = 0; t < nb; ++t) {
                const float* v = tv.data() + (size_t)t * P;
                if (t0 > 0) {
                    for (int e = 0; e < P; ++e) col[e] = tv[(size_t)e * nb + t];
                    v = col.data();
                }
                int* row = rows.data() + (size_t)t * 2 * K;
                heap_select(v, P, K, row, row + K);
            }
            HIPCHK(hipMemcpyAsync(c->tie_rows.p, rows.data(), (size_t)nb * 2 * K * 4, hipMemcpyHostToDevice, st));
            LCHK(launch_scatter_rows(P_<int>(c->tie_idx) + t0, P_<int>(c->tie_rows), nb, 2 * K, P_<int>(c->topK), st));
            HIPCHK(hipStreamSynchronize(st));
        }
    }
    c->stats.n_tie_points = ntie;

    // ---- host replay of flagged top-M rows
    std::vector<int> rows;
    for (int r = 0; r < P; ++r)
        if (hflags[2 * r] || hflags[2 * r + 1]) rows.push_back(r);
    c->stats.n_tie_rows = (long long)rows.size();
    if (!rows.empty()) {
        const int batch = (int)std::max<long long>(1, std::min<long long>(16, (1ll << 31) / (n * 4)));
        ENSURE(c->row_slot, (size_t)P * 4);
        ENSURE(c->row_buf, (size_t)batch * n * 4);
        std::vector<int> slot(P);
        std::vector<float> buf((size_t)batch * n);
        std::vector<int> out(2 * M);
        for (size_t b0 = 0; b0 < rows.size(); b0 += batch) {
            const size_t b1 = std::min(rows.size(), b0 + batch);
            std::fill(slot.begin(), slot.end(), -1);
            for (size_t i = b0; i < b1; ++i) slot[rows[i]] = (int)(i - b0);
            HIPCHK(hipMemcpyAsync(c->row_slot.p, slot.data(), (size_t)P * 4, hipMemcpyHostToDevice, st));
            ia.row_slot = P_<int>(c->row_slot);
            ia.row_buf = P_<float>(c->row_buf);
            LCHK(launch_index(ia, logD, 2, Gf, st));
            HIPCHK(hipMemcpyAsync(buf.data(), c->row_buf.p, (b1 - b0) * n * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (size_t i = b0; i < b1; ++i) {
                heap_select(buf.data() + (i - b0) * n, n, M, out.data(), out.data() + M);
                HIPCHK(hipMemcpy(P_<int>(c->topM) + (size_t)rows[i] * 2 * M, out.data(), (size_t)2 * M * 4, hipMemcpyHostToDevice));
            }
        }
    }

    // ---- list structures, hubs, first-occurrence masks, groups
    const int L2 = c->L2, W = c->W, S = c->S;
    const int LM = L2 * M;
    ENSURE(c->lists, (size_t)LM * 4);
    ENSURE(c->is_hub, (size_t)n);
    ENSURE(c->hub_flags, (size_t)n * 4);
    ENSURE(c->hub_excl, (size_t)n * 4);
    ENSURE(c->hub_id, (size_t)n * 4);
    ENSURE(c->total, 16);
    size_t stb = std::max(scan_tmp_bytes(n), scan_tmp_bytes((long long)std::max(L2 * 1024, LM) + 1));
    ENSURE(c->scan_tmp, stb);
    HIPCHK(hipMemsetAsync(c->is_hub.p, 0, (size_t)n, st));
    LCHK(launch_build_lists(P_<int>(c->topM), P, M, P_<int>(c->lists), P_<uint8_t>(c->is_hub), st));
    LCHK(launch_is_hub_scan_prep(P_<uint8_t>(c->is_hub), P_<int>(c->hub_flags), n, st));
    LCHK(scan_exclusive_i32(P_<int>(c->hub_flags), P_<int>(c->hub_excl), n, P_<int>(c->total), c->scan_tmp.p, c->scan_tmp.bytes, st));
    LCHK(launch_hub_id(P_<uint8_t>(c->is_hub), P_<int>(c->hub_excl), P_<int>(c->hub_id), n, st));
    int H = 0;
    HIPCHK(hipMemcpyAsync(&H, c->total.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c->H = H;
    c->stats.n_hubs = H;
    ENSURE(c->hub_cnt, (size_t)(H + 1) * 4);
    ENSURE(c->hub_ptr, (size_t)(H + 1) * 4);
    ENSURE(c->hub_cur, (size_t)(H + 1) * 4);
    ENSURE(c->hub_mem, (size_t)LM * 4);
    HIPCHK(hipMemsetAsync(c->hub_cnt.p, 0, (size_t)(H + 1) * 4, st));
    HIPCHK(hipMemsetAsync(c->hub_cur.p, 0, (size_t)(H + 1) * 4, st));
    LCHK(launch_hub_count(P_<int>(c->lists), LM, P_<int>(c->hub_id), P_<int>(c->hub_cnt), st));
    LCHK(scan_exclusive_i32(P_<int>(c->hub_cnt), P_<int>(c->hub_ptr), H + 1, nullptr, c->scan_tmp.p, c->scan_tmp.bytes, st));
    LCHK(launch_hub_fill(P_<int>(c->lists), LM, P_<int>(c->hub_id), P_<int>(c->hub_ptr), P_<int>(c->hub_cur), P_<int>(c->hub_mem), st));
    const size_t tbl = (size_t)L2 * L2 * W * 4;
    ENSURE(c->inv, tbl);
    HIPCHK(hipMemsetAsync(c->inv.p, 0, tbl, st));
    LCHK(launch_invmask(P_<int>(c->lists), P_<int>(c->topK), P, K, M, L2, W, P_<uint32_t>(c->inv), st));
    // (position, other list) pairs of every list, for the first-occurrence masks
    ENSURE(c->pos_hub, (size_t)LM * 4);
    ENSURE(c->pcnt, (size_t)(LM + 1) * 4);
    ENSURE(c->poff, (size_t)(LM + 1) * 4);
    LCHK(launch_pair_count(P_<int>(c->lists), P_<int>(c->hub_id), P_<int>(c->hub_ptr), L2, M, P_<int>(c->pos_hub),
                           P_<int>(c->pcnt), P_<int>(c->poff), P_<int>(c->total), c->scan_tmp.p, c->scan_tmp.bytes, st));
    int npairs = 0;
    HIPCHK(hipMemcpyAsync(&npairs, c->total.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ENSURE(c->pairs, (size_t)std::max(npairs, 1) * 4);
    LCHK(launch_pair_fill(P_<int>(c->pos_hub), P_<int>(c->hub_ptr), P_<int>(c->hub_mem), L2, M, P_<int>(c->poff),
                          P_<uint32_t>(c->pairs), st));
    ENSURE(c->F, (size_t)n * S * W * 4);
    ENSURE(c->ffull, (size_t)n * S);
    ENSURE(c->pseg, (size_t)L2 * 8);
    // lists that share many points (long pair segments, e.g. the extreme
    // points of blobs in many lists): the dense intersections replace one
    // binary search per earlier slot with one load (at most 256 MB)
    const size_t inter_bytes = (size_t)L2 * L2 * W * 4;
    const bool use_inter = (long long)npairs >= 16ll * L2 && inter_bytes <= ((size_t)256 << 20);
    if (use_inter) {
        ENSURE(c->inter, inter_bytes);
        HIPCHK(hipMemsetAsync(c->inter.p, 0, inter_bytes, st));
        LCHK(launch_inter_fill(P_<int>(c->pos_hub), P_<int>(c->hub_ptr), P_<int>(c->hub_mem), L2, M, W,
                               P_<uint32_t>(c->inter), st));
    }
    LCHK(launch_first_mask(P_<int>(c->topK), n, P, K, M, W, P_<int>(c->poff), P_<uint32_t>(c->pairs), P_<int2>(c->pseg),
                           P_<int>(c->hub_id), P_<int>(c->hub_ptr), P_<int>(c->hub_mem),
                           use_inter ? P_<uint32_t>(c->inter) : nullptr, P_<uint32_t>(c->F), P_<uint8_t>(c->ffull),
                           st));
    // group the (point, slot) entries by (query block, list): the distance
    // tiles then sweep the points block by block, and a 192 MB block of query
    // rows stays in the 256 MB last-level cache, so most of the 2K visits of a
    // row after the first hit there (larger blocks lose hits, smaller ones
    // restage the lists' candidate rows more often)
    // (FP32 tiles are MFMA-bound and only lose from smaller tiles: one block)
    long long qblk_mb = q.precision == SDB_BF16 ? 192 : (1ll << 30);
    if (const char* e = std::getenv("SDB_QBLK_MB")) qblk_mb = std::max(1ll, std::atoll(e));
    const long long row_bytes = (long long)ldx * (q.precision == SDB_BF16 ? 2 : 4);
    const long long qblk = std::max<long long>(1, (qblk_mb << 20) / row_bytes);
    const int nblk = (int)((n + qblk - 1) / qblk);
    const int Gx = (int)std::max<long long>(1, std::min<long long>(2048, 2048 / nblk));
    const int NG = nblk * L2;
    c->qblk = qblk;
    c->nblk = nblk;
    ENSURE(c->wgcnt, (size_t)NG * Gx * 4);
    ENSURE(c->wgoff, (size_t)NG * Gx * 4);
    ENSURE(c->grp_ptr, (size_t)(NG + 1) * 4);
    ENSURE(c->grp, (size_t)n * S * 4);
    ENSURE(c->scan_tmp, scan_tmp_bytes((long long)NG * Gx));
    LCHK(launch_group_count(P_<int>(c->topK), n, P, K, L2, qblk, nblk, Gx, P_<int>(c->wgcnt), st));
    LCHK(scan_exclusive_i32(P_<int>(c->wgcnt), P_<int>(c->wgoff), (long long)NG * Gx, nullptr, c->scan_tmp.p, c->scan_tmp.bytes, st));
    LCHK(launch_group_scatter(P_<int>(c->topK), n, P, K, L2, qblk, nblk, Gx, P_<int>(c->wgoff), P_<uint8_t>(c->ffull),
                              P_<uint32_t>(c->grp), P_<int>(c->grp_ptr), st));
    c->h_grp_ptr.resize(NG + 1);
    HIPCHK(hipMemcpyAsync(c->h_grp_ptr.data(), c->grp_ptr.p, (size_t)(NG + 1) * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev[2], st));
    HIPCHK(hipStreamSynchronize(st));
    c->stats.t_transform_ms = elapsed(c->ev[0], c->ev[1]);
    c->stats.t_index_ms = elapsed(c->ev[1], c->ev[2]);
    c->indexed = true;
    c->transformed = true;
    c->rnorm2_of = nullptr;  // the rows were written again
    c->transformed_prec = q.precision;
    c->cored = false;
    c->have_nb = false;
    return SDB_OK;
}

int sdb_get_index(sdb_ctx* c, int* topK_out, int* topM_out) {
    if (!c) return SDB_EINVAL;
    if (!c->indexed) return fail(c, SDB_ESTATE, "build_index first");
    HIPCHK(hipSetDevice(c->device));
    if (topK_out) HIPCHK(hipMemcpy(topK_out, c->topK.p, (size_t)c->n * 2 * c->prm.topK * 4, hipMemcpyDeviceToHost));
    if (topM_out) HIPCHK(hipMemcpy(topM_out, c->topM.p, (size_t)c->prm.n_proj * 2 * c->prm.topM * 4, hipMemcpyDeviceToHost));
    return SDB_OK;
}

// HD3 projections (src/sDbscan.cpp:74-91) of listed points, recomputed from
// the transformed rows and the index's sign bits by the index kernel's
// point mode: column q of out (n_proj floats) is the MATRIX_FHT column of
// point ids[q].
int sdb_get_projections(sdb_ctx* c, const int* ids, int n_ids, float* out) {
    if (!c || n_ids < 0 || (n_ids > 0 && (!ids || !out))) return SDB_EINVAL;
    if (!c->indexed) return fail(c, SDB_ESTATE, "build_index first");
    if (n_ids == 0) return SDB_OK;
    for (int i = 0; i < n_ids; ++i)
        if (ids[i] < 0 || ids[i] >= c->n) return fail(c, SDB_EINVAL, "point %d out of range", ids[i]);
    HIPCHK(hipSetDevice(c->device));
    const sdb_params& q = c->prm;
    const int P = q.n_proj;
    hipStream_t st = c->st;
    ENSURE(c->pj_ids, (size_t)n_ids * 4);
    ENSURE(c->pj_buf, (size_t)P * n_ids * 4);
    HIPCHK(hipMemcpyAsync(c->pj_ids.p, ids, (size_t)n_ids * 4, hipMemcpyHostToDevice, st));
    IndexArgs ia{};
    ia.X = c->X;
    ia.Xn = P_<float>(c->Xn);
    ia.err = P_<int>(c->err_flag);
    ia.signmask = P_<uint32_t>(c->bits);
    ia.n = c->n; ia.ldx = c->ldx; ia.d = c->d; ia.metric = q.metric; ia.ker = q.ker_n_features;
    ia.kerL = q.ker_intervalSampling;
    ia.R = (q.metric == SDB_L1 || q.metric == SDB_L2) ? P_<float>(c->R) : nullptr;
    ia.P = P; ia.K = q.topK;
    ia.pt_ids = P_<int>(c->pj_ids);
    ia.n_ids = n_ids;
    ia.pt_buf = P_<float>(c->pj_buf);
    const int G = (int)std::max<long long>(1, std::min<long long>(2048, ((long long)n_ids + 3) / 4));
    LCHK(launch_index(ia, ilog2(q.fhtDim), 3, G, st));
    // the point mode writes row e of the P x n_ids block at e * n_ids; the
    // caller gets one column of P values per point (MATRIX_FHT's layout)
    std::vector<float> rows((size_t)P * n_ids);
    HIPCHK(hipMemcpyAsync(rows.data(), c->pj_buf.p, rows.size() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int e = 0; e < P; ++e)
        for (int qi = 0; qi < n_ids; ++qi) out[(size_t)qi * P + e] = rows[(size_t)e * n_ids + qi];
    return SDB_OK;
}

int sdb_get_transformed(sdb_ctx* c, float* X_out) {
    if (!c || !X_out) return SDB_EINVAL;
    if (!c->transformed) return fail(c, SDB_ESTATE, "build_index or sng_find_core first");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpy2D(X_out, (size_t)c->d * 4, c->Xn.p, (size_t)c->ldx * 4, (size_t)c->d * 4, (size_t)c->n, hipMemcpyDeviceToHost));
    return SDB_OK;
}

int sdb_get_transformed_bf16(sdb_ctx* c, uint16_t* X_out) {
    if (!c || !X_out) return SDB_EINVAL;
    if (!c->transformed || c->transformed_prec != SDB_BF16 || !c->Xb.p)
        return fail(c, SDB_ESTATE, "no BF16 rows: build the index (or sng_find_core) with precision BF16 first");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpy2D(X_out, (size_t)c->d * 2, c->Xb.p, (size_t)c->ldx * 2, (size_t)c->d * 2, (size_t)c->n, hipMemcpyDeviceToHost));
    return SDB_OK;
}

// Distance tiles over the ctx's list groups (c->h_grp_ptr over nblk x L2
// groups of entries, lists of c->Mt points, masks F -> hits E, colsum) and the
// FP32/FP64 recheck of the pairs inside the error band.
static constexpr int kQueueFull = -100;  // internal: more queued pairs than the queue may hold

static int run_distances_p(sdb_ctx* c, float eps, int precision, int* nq_out, double* band_out,
                           double* band32_out) {
    const sdb_params& q = c->prm;
    const long long n = c->n;
    const int M = c->Mt, W = c->W, S = c->S, L2 = c->L2;
    hipStream_t st = c->st;
    // L2 on FP32 MFMA tiles through |x - y|^2 = |x|^2 + |y|^2 - 2 x.y, with a
    // band of kappa (|x|^2 + |y|^2) per pair (k_dist_mfma); rows far from the
    // origin relative to eps would widen that band, and the VALU tiles, whose
    // error is relative to the distance itself, keep those inputs
    bool l2x = false;
    double l2_kappa = 0.0, l2_maxn = 0.0;
    if (q.metric == SDB_L2 && M <= 128 && !std::getenv("SDB_L2_VALU")) {
        // the norms of the rows are kept until the rows change
        if (c->rnorm2_of != P_<float>(c->Xn) || !c->transformed) {
            ENSURE(c->rnorm2, (size_t)n * 4);
            ENSURE(c->rnorm2_max, 4);
            HIPCHK(hipMemsetAsync(c->rnorm2_max.p, 0, 4, st));
            LCHK(launch_row_norm2(P_<float>(c->Xn), n, c->d, c->ldx, P_<float>(c->rnorm2),
                                  P_<uint32_t>(c->rnorm2_max), st));
            uint32_t mb = 0;
            HIPCHK(hipMemcpyAsync(&mb, c->rnorm2_max.p, 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            std::memcpy(&c->rnorm2_maxv, &mb, 4);
            c->rnorm2_of = P_<float>(c->Xn);
        }
        l2_maxn = c->rnorm2_maxv;
        l2_kappa = 1.02 * (2.0 * c->d + 16.0) * std::ldexp(1.0, -23);
        const double eps2 = (double)eps * (double)eps;
        l2x = std::isfinite(l2_maxn) && 2.0 * l2_kappa * l2_maxn <= 0.05 * eps2;
    }
    // Cosine on MFMA tiles at either precision; Chi2 with BF16 on the BF16
    // tiles through bounds on chi2 (compute_bands); L2 as above; everything
    // else on the VALU tiles in FP32
    const bool mfma = (q.metric == SDB_COSINE || (q.metric == SDB_CHI2 && precision == SDB_BF16) || l2x) && M <= 128;
    c->l2x_used = l2x;
    const int prec = mfma ? precision : SDB_FP32;
    const bool prefilter = q.metric == SDB_CHI2 && prec == SDB_BF16;

    // tiles: contiguous ranges of one list's group
    int QB;
    DistArgs da{};
    da.ldx = c->ldx; da.d = c->d; da.metric = q.metric;
    da.M = M; da.W = W; da.S = S;
    if (mfma) {
        const int SE = prec == SDB_BF16 ? 32 : 16;
        da.dk = (c->d + SE - 1) / SE * SE;
        // LDS: M rows of steps x 64 B + 16 B, 8 B of ids / counts per
        // candidate and one word of the list's largest rounding bound
        const int steps = (int)(((kLdsBytes - 8 * M - 16) / M - 16) / 64);
        if (steps <= 0) return fail(c, SDB_EUNSUPPORTED, "topM too large for the LDS tile");
        da.kc = std::min(da.dk, steps * SE);
        da.nchunk = (da.dk + da.kc - 1) / da.kc;
        QB = da.nchunk == 1 ? 2048 : 256;  // single-chunk tiles: 1024-3072 rows measured best (8M x 784 BF16)
        if (const char* e = std::getenv("SDB_TILE_ROWS")) QB = std::max(256, std::atoi(e));
        if (c->tile_rows > 0) QB = c->tile_rows;
        QB = std::min(QB, 16 * 8 * 255);  // per-lane byte counters of the single-chunk tiles
    } else {
        int kc = (kLdsBytes - 8 * M) / (4 * M + 128);
        kc = std::min(kc, c->d);
        if (kc <= 0) return fail(c, SDB_EUNSUPPORTED, "topM too large for the LDS tile");
        da.kc = kc;
        da.nchunk = (c->d + kc - 1) / kc;
        da.dk = c->d;
        QB = 256;
    }
    std::vector<int> tl, tb, te;
    const int NG = c->nblk * L2;
    for (int g = 0; g < NG; ++g)
        for (int b = c->h_grp_ptr[g]; b < c->h_grp_ptr[g + 1]; b += QB) {
            tl.push_back(g % L2);
            tb.push_back(b);
            te.push_back(std::min(b + QB, c->h_grp_ptr[g + 1]));
        }
    const int ntiles = (int)tl.size();
    ENSURE(c->tiles, (size_t)3 * std::max(ntiles, 1) * 4);
    int* dt = P_<int>(c->tiles);
    HIPCHK(hipMemcpyAsync(dt, tl.data(), (size_t)ntiles * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dt + ntiles, tb.data(), (size_t)ntiles * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dt + 2 * ntiles, te.data(), (size_t)ntiles * 4, hipMemcpyHostToDevice, st));

    ENSURE(c->E, (size_t)n * S * W * 4);
    ENSURE(c->colsum, (size_t)L2 * M * 4);
    ENSURE(c->qcount, 8);
    ENSURE(c->stats64, 64);
    if (c->qcap == 0) c->qcap = 1 << 22;
    ENSURE(c->queue, (size_t)c->qcap * 8);

    double band = 0, band32 = 0;
    compute_bands(q, prec, eps, &band, &band32);
    da.X = prec == SDB_BF16 ? (const void*)c->Xb.p : (const void*)c->Xn.p;
    da.rho_pad = prec == SDB_BF16 ? bf16_rho_pad(c) : 0;
    da.X32 = P_<float>(c->Xn);
    da.grp = P_<uint32_t>(c->grp);
    da.tile_list = dt; da.tile_beg = dt + ntiles; da.tile_end = dt + 2 * ntiles;
    da.lists = P_<int>(c->lists);
    da.F = P_<uint32_t>(c->F);
    da.E = P_<uint32_t>(c->E);
    da.colsum = P_<int>(c->colsum);
    da.eps = eps;
    da.lo = prefilter ? ((double)eps - 1e-6) / 2.0 - band : (double)eps - band;
    da.hi = (double)eps + band;
    cosine_cuts(da.lo, da.hi, &da.s_hit, &da.s_near);
    if (l2x) {
        da.rnorm2 = P_<float>(c->rnorm2);
        da.eps2 = (double)eps * (double)eps;
        da.kappa = l2_kappa;
        // FP64 rounding of the recheck's distance stays far inside this margin
        da.eps0 = 1e-9 * da.eps2 + 1e-300;
        // the widest pair band, as a distance near eps
        band = (2.0 * l2_kappa * l2_maxn + da.eps0) / (2.0 * (double)eps);
    }
    da.magicS = ((1ull << 40) + (unsigned long long)S - 1) / (unsigned long long)S;
    da.n_ent = n * S;
    // BF16 rows in one LDS tile (the pipelined tiles): a band per pair from
    // the two rows' rounding bounds.  |s_bf16 - s| <= rho_q + rho_c + gam,
    // with s the exact dot of the rows the decision is made on and gam the
    // FP32 accumulation (2 (d + 2) u over unit rows, factor 2 for the unknown
    // order inside the MFMA unit) plus the rounding of the threshold
    // arithmetic; the exact FP64 decision is then certain outside the band.
    if (prec == SDB_BF16 && da.nchunk == 1 && da.kc / 32 <= 25) {
        const double u = std::ldexp(1.0, -23);
        const double gam = 2.0 * (c->d + 2.0) * u * 1.016 + 16.0 * u + 1e-9;
        // hit: Cosine 1 - s < eps; Chi2 (s = BC) 2 (1 - BC) + 4u < eps - 1e-6
        const double t_hit = prefilter ? 1.0 - ((double)eps - 1e-6) / 2.0 + 2.0 * u : 1.0 - (double)eps;
        const double t_near = 1.0 - (double)eps;  // not a certain miss
        float fh = (float)(t_hit + gam), fn = (float)(t_near - gam);
        if ((double)fh < t_hit + gam) fh = std::nextafter(fh, INFINITY);
        if ((double)fn > t_near - gam) fn = std::nextafter(fn, -INFINITY);
        da.rowerr = P_<float>(c->rowerr);
        da.th_hit = fh;
        da.th_near = fn;
        band = 2.0 * (double)c->rho_max + gam;  // the widest pair band
    }

    HIPCHK(hipEventRecord(c->ev[3], st));
    // queued pairs are counted in 64 bits (a pass may meet more than 2^31)
    // and stored while they fit; a pass that overflows is run again with a
    // queue of twice the count, up to kQueueMax entries (SDB_QUEUE_MAX lowers
    // the limit)
    long long kQueueMax = 1ll << 30;
    if (const char* e = std::getenv("SDB_QUEUE_MAX")) kQueueMax = std::max(1ll, std::min(kQueueMax, std::atoll(e)));
    unsigned long long nq64 = 0;
    for (int attempt = 0; attempt < 4; ++attempt) {
        da.queue = (unsigned long long*)c->queue.p;
        da.qcount = (unsigned long long*)c->qcount.p;
        da.qcap = (int)c->qcap;
        HIPCHK(hipMemsetAsync(c->colsum.p, 0, (size_t)L2 * M * 4, st));
        HIPCHK(hipMemsetAsync(c->qcount.p, 0, 8, st));
        HIPCHK(hipEventRecord(c->ev[4], st));
        if (mfma) LCHK(launch_dist_mfma(da, prec, ntiles, st));
        else LCHK(launch_dist_valu(da, ntiles, st));
        HIPCHK(hipEventRecord(c->ev[5], st));
        HIPCHK(hipMemcpyAsync(&nq64, c->qcount.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (nq64 > (unsigned long long)kQueueMax) {
            c->stats.n_queue = (long long)nq64;
            return kQueueFull;
        }
        if (nq64 <= (unsigned long long)c->qcap) break;
        c->qcap = std::min<long long>(kQueueMax, (long long)nq64 * 2);
        ENSURE(c->queue, (size_t)c->qcap * 8);
    }
    if (nq64 > (unsigned long long)c->qcap) return fail(c, SDB_ENOMEM, "recheck queue overflow");
    const int nq = (int)nq64;
    c->stats.t_dist_ms = elapsed(c->ev[4], c->ev[5]);
    // a long queue is sorted by candidate point first: the recheck then reads
    // a candidate's FP32 row from memory once and from L2 for the rest of its
    // pairs (the pairs' order does not change any result: hits are ORed into
    // E and added to colsum)
    // (an optimisation only: without the memory for it the pairs are
    // rechecked in queue order)
    if (nq >= (1 << 14)) {
        const int nbits = std::max(1, 32 - __builtin_clz((unsigned)std::max<long long>(1, n - 1)));
        size_t tb = 0;
        hipcub::DoubleBuffer<uint32_t> kb0(nullptr, nullptr);
        hipcub::DoubleBuffer<unsigned long long> vb0(nullptr, nullptr);
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb0, vb0, nq, 0, nbits, st));
        if (!ensure(c, c->qkeys, (size_t)nq * 8) && !ensure(c, c->qvals, (size_t)nq * 8) &&
            !ensure(c, c->qsort_tmp, std::max<size_t>(tb, 16))) {
            uint32_t* k0 = P_<uint32_t>(c->qkeys);
            LCHK(launch_queue_keys(da.queue, nq, da.lists, da.M, k0, st));
            hipcub::DoubleBuffer<uint32_t> kb(k0, k0 + nq);
            hipcub::DoubleBuffer<unsigned long long> vb(da.queue, (unsigned long long*)c->qvals.p);
            HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->qsort_tmp.p, tb, kb, vb, nq, 0, nbits, st));
            da.queue = vb.Current();
        } else {
            c->err.clear();  // the failed allocation is not an error of this call
        }
    }
    HIPCHK(hipMemsetAsync(c->stats64.p, 0, 64, st));
    LCHK(launch_recheck_k(da, P_<int>(c->topK), q.topK, q.n_proj, prec, band32, nq, (unsigned long long*)c->stats64.p, st));

    *nq_out = nq;
    *band_out = band;
    *band32_out = band32;
    return SDB_OK;
}

// The tiles of the context's precision; when they leave more pairs near eps
// than the recheck queue holds (BF16 Chi2 lower bounds at a large eps), the
// FP32 tiles, which decide all pairs outside a band of 2(d+8)u themselves.
static int run_distances(sdb_ctx* c, float eps, int* nq_out, double* band_out, double* band32_out) {
    int r = run_distances_p(c, eps, c->prm.precision, nq_out, band_out, band32_out);
    if (r == kQueueFull && c->prm.precision != SDB_FP32)
        r = run_distances_p(c, eps, SDB_FP32, nq_out, band_out, band32_out);
    if (r == kQueueFull)
        return fail(c, SDB_ENOMEM, "recheck queue of %lld pairs is too large", c->stats.n_queue);
    return r;
}

// Candidate-pair sample of sdb_sample_pair_dist into the device buffer base
// (20 bytes per sample: the FP64 values first, 8-byte aligned, then the FP32
// values and the two index arrays), copied to the caller's host arrays.
static int sample_pairs_into(sdb_ctx* c, long long n_samples, unsigned long long seed, void* base, int* pa_out,
                             int* pb_out, float* fast_out, double* exact_out) {
    const sdb_params& q = c->prm;
    hipStream_t st = c->st;
    // the precision the distance tiles run at (as in run_distances_p)
    const bool mfma = (q.metric == SDB_COSINE || (q.metric == SDB_CHI2 && q.precision == SDB_BF16)) && c->Mt <= 128;
    const int prec = mfma ? q.precision : SDB_FP32;
    DistArgs da{};
    da.X = prec == SDB_BF16 ? (const void*)c->Xb.p : (const void*)c->Xn.p;
    da.X32 = P_<float>(c->Xn);
    da.ldx = c->ldx; da.d = c->d; da.metric = q.metric;
    da.M = c->Mt;
    da.lists = P_<int>(c->lists);
    if (q.metric == SDB_L2 && c->l2x_used) da.rnorm2 = P_<float>(c->rnorm2);
    const size_t ns = (size_t)std::max<long long>(n_samples, 1);
    double* ex = (double*)base;
    float* fs = (float*)(ex + ns);
    int* pa = (int*)(fs + ns);
    int* pb = pa + ns;
    if (launch_pair_sample(da, P_<int>(c->topK), q.topK, q.n_proj, c->n, prec, seed, n_samples, pa, pb, fs, ex, st))
        return SDB_EHIP;
    if (n_samples > 0) {
        if (hipMemcpyAsync(pa_out, pa, (size_t)n_samples * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(pb_out, pb, (size_t)n_samples * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(fast_out, fs, (size_t)n_samples * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(exact_out, ex, (size_t)n_samples * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
            return SDB_EHIP;
    }
    return hipStreamSynchronize(st) == hipSuccess ? SDB_OK : SDB_EHIP;
}

int sdb_sample_pair_dist(sdb_ctx* c, long long n_samples, unsigned long long seed, int* pa_out, int* pb_out,
                         float* fast_out, double* exact_out) {
    if (!c || n_samples < 0 || (n_samples > 0 && (!pa_out || !pb_out || !fast_out || !exact_out))) return SDB_EINVAL;
    if (!c->indexed || c->sng) return fail(c, SDB_ESTATE, "build_index first");
    if (n_samples > (1ll << 26)) return fail(c, SDB_EINVAL, "at most 2^26 sampled pairs per call");
    HIPCHK(hipSetDevice(c->device));
    void* base = nullptr;
    if (hipMalloc(&base, (size_t)std::max<long long>(n_samples, 1) * 20) != hipSuccess)
        return fail(c, SDB_ENOMEM, "sample buffer");
    const int r = sample_pairs_into(c, n_samples, seed, base, pa_out, pb_out, fast_out, exact_out);
    (void)hipStreamSynchronize(c->st);
    (void)hipFree(base);
    if (r) return fail(c, r, "sampled pair distances failed");
    return SDB_OK;
}

int sdb_sample_pair_dist_noalloc(sdb_ctx* c, long long n_samples, unsigned long long seed, int* pa_out,
                                 int* pb_out, float* fast_out, double* exact_out) {
    if (!c || n_samples < 0 || (n_samples > 0 && (!pa_out || !pb_out || !fast_out || !exact_out))) return SDB_EINVAL;
    if (!c->indexed || c->sng) return fail(c, SDB_ESTATE, "build_index first");
    if (n_samples > SDB_NOALLOC_MAX_SAMPLES)
        return fail(c, SDB_EINVAL, "%lld sampled pairs exceed the %d of the static scratch buffer", n_samples,
                    SDB_NOALLOC_MAX_SAMPLES);
    if (!c->scratch) return fail(c, SDB_ESTATE, "no scratch buffer on device %d", c->device);
    HIPCHK(hipSetDevice(c->device));
    const int r = sample_pairs_into(c, n_samples, seed, c->scratch, pa_out, pb_out, fast_out, exact_out);
    if (r) return fail(c, r, "sampled pair distances failed");
    return SDB_OK;
}

int sdb_find_core(sdb_ctx* c, float eps, int minPts, uint8_t* core_out, int* degree_out) {
    if (!c) return SDB_EINVAL;
    if (!c->indexed || c->sng) return fail(c, SDB_ESTATE, "build_index first");
    HIPCHK(hipSetDevice(c->device));
    const sdb_params& q = c->prm;
    const long long n = c->n;
    const int M = c->Mt, W = c->W, S = c->S, L2 = c->L2;
    hipStream_t st = c->st;
    if (minPts < 1) return fail(c, SDB_EINVAL, "minPts must be positive");
    if (c->cored && c->eps == eps) {
        // same eps: the hit masks and degrees stand; only the core test moves
        HIPCHK(hipEventRecord(c->ev[3], st));
        HIPCHK(hipMemsetAsync((unsigned long long*)c->stats64.p + 3, 0, 8, st));
        ENSURE(c->core_spread, (size_t)kSpread * kSpreadStride * 8);
        HIPCHK(hipMemsetAsync(c->core_spread.p, 0, c->core_spread.bytes, st));
        LCHK(launch_core_minpts(P_<int>(c->degree), n, minPts, P_<uint8_t>(c->core),
                                P_<unsigned long long>(c->core_spread), kSpread - 1, st));
        LCHK(launch_spread_fold(P_<unsigned long long>(c->core_spread), (unsigned long long*)c->stats64.p + 3, st));
        HIPCHK(hipEventRecord(c->ev[6], st));
        unsigned long long nc = 0;
        HIPCHK(hipMemcpyAsync(&nc, (unsigned long long*)c->stats64.p + 3, 8, hipMemcpyDeviceToHost, st));
        if (core_out) HIPCHK(hipMemcpyAsync(core_out, c->core.p, (size_t)n, hipMemcpyDeviceToHost, st));
        if (degree_out) HIPCHK(hipMemcpyAsync(degree_out, c->degree.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        c->stats.t_core_ms = elapsed(c->ev[3], c->ev[6]);
        c->stats.t_dist_ms = 0.f;
        c->stats.n_core = (long long)nc;
        c->minPts = minPts;
        c->clustered = false;
        return SDB_OK;
    }
    int nq = 0;
    double band = 0, band32 = 0;
    int r = run_distances(c, eps, &nq, &band, &band32);
    if (r) return r;

    // degrees and core flags
    ENSURE(c->degree, (size_t)n * 4);
    ENSURE(c->core, (size_t)n);
    ClusterArgs ca{};
    ca.n = n; ca.S = S; ca.W = W; ca.M = M; ca.P = q.n_proj; ca.K = q.topK; ca.L2 = L2;
    ca.topK = P_<int>(c->topK); ca.lists = P_<int>(c->lists); ca.E = P_<uint32_t>(c->E);
    ca.colsum = P_<int>(c->colsum); ca.hub_id = P_<int>(c->hub_id); ca.hub_ptr = P_<int>(c->hub_ptr);
    ca.hub_mem = P_<int>(c->hub_mem); ca.inv = P_<uint32_t>(c->inv); ca.minPts = minPts;
    ca.degree = P_<int>(c->degree); ca.core = P_<uint8_t>(c->core);
    ENSURE(c->core_spread, (size_t)kSpread * kSpreadStride * 8);
    HIPCHK(hipMemsetAsync(c->core_spread.p, 0, c->core_spread.bytes, st));
    ca.n_core = P_<unsigned long long>(c->core_spread);
    ca.spread = kSpread - 1;
    LCHK(launch_degree(ca, st));
    LCHK(launch_spread_fold(P_<unsigned long long>(c->core_spread), (unsigned long long*)c->stats64.p + 3, st));
    HIPCHK(hipEventRecord(c->ev[6], st));
    unsigned long long hs[8] = {0};
    HIPCHK(hipMemcpyAsync(hs, c->stats64.p, 64, hipMemcpyDeviceToHost, st));
    if (core_out) HIPCHK(hipMemcpyAsync(core_out, c->core.p, (size_t)n, hipMemcpyDeviceToHost, st));
    if (degree_out) HIPCHK(hipMemcpyAsync(degree_out, c->degree.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c->stats.t_core_ms = elapsed(c->ev[3], c->ev[6]);
    c->stats.n_pairs = (long long)n * S * M;
    c->stats.n_queue = nq;
    c->stats.n_recheck32 = (long long)hs[0];
    c->stats.n_recheck64 = (long long)hs[1];
    c->stats.n_recheck_hits = (long long)hs[2];
    c->stats.n_near_eps = (long long)hs[4];
    c->stats.n_core = (long long)hs[3];
    c->stats.band = band;
    c->stats.band32 = band32;
    c->eps = eps;
    c->minPts = minPts;
    c->cored = true;
    c->have_nb = false;
    c->clustered = false;
    return SDB_OK;
}

int sdb_form_clusters(sdb_ctx* c, int* labels_out, int* n_clusters_out) {
    if (!c) return SDB_EINVAL;
    if (!c->cored) return fail(c, SDB_ESTATE, "find_core first");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->st;
    const long long n = c->n;
    const int M = c->Mt, W = c->W, L2 = c->L2;
    ENSURE(c->total, 16);
    ENSURE(c->corehub, (size_t)L2 * W * 4);
    ENSURE(c->parent, (size_t)n * 4);
    ENSURE(c->border, (size_t)n * 4);
    ENSURE(c->labels, (size_t)n * 4);
    ENSURE(c->adj, (size_t)L2 * M * W * 4);
    ENSURE(c->rootf, (size_t)n * 4);
    ENSURE(c->rank, (size_t)n * 4);
    ClusterArgs ca = cluster_args(c);
    HIPCHK(hipEventRecord(c->ev[0], st));
    if (c->sng) {
        // the hits are the edges of the neighbourhood graph
        const int L = c->sngL, S2 = c->S, Ms = c->Mt;
        LCHK(launch_uf_init(ca, st));
        LCHK(launch_sng_uf(n, L, S2, Ms, W, P_<uint32_t>(c->E), P_<int>(c->lists), P_<uint8_t>(c->core),
                           P_<int>(c->parent), st));
        LCHK(launch_uf_compress(ca, st));
        LCHK(launch_root_flags(ca, P_<int>(c->rootf), st));
        ENSURE(c->scan_tmp, scan_tmp_bytes(n));
        LCHK(scan_exclusive_i32(P_<int>(c->rootf), P_<int>(c->rank), n, P_<int>(c->total), c->scan_tmp.p, c->scan_tmp.bytes, st));
        LCHK(launch_sng_border(n, L, S2, Ms, W, P_<uint32_t>(c->E), P_<int>(c->lists), P_<uint8_t>(c->core),
                               P_<int>(c->parent), P_<int>(c->border), st));
        LCHK(launch_labels(ca, P_<int>(c->rank), st));
    } else {
        HIPCHK(hipMemsetAsync(c->adj.p, 0, (size_t)L2 * M * W * 4, st));
        LCHK(launch_corehub(ca, st));
        LCHK(launch_uf_init(ca, st));
        if (c->stats.n_core > 0) {  // without core points there is nothing to join
            LCHK(launch_uf_points(ca, st));
            LCHK(launch_uf_lists(ca, st));
        }
        LCHK(launch_uf_compress(ca, st));
        LCHK(launch_root_flags(ca, P_<int>(c->rootf), st));
        LCHK(scan_exclusive_i32(P_<int>(c->rootf), P_<int>(c->rank), n, P_<int>(c->total), c->scan_tmp.p, c->scan_tmp.bytes, st));
        LCHK(launch_border(ca, st));
        LCHK(launch_labels(ca, P_<int>(c->rank), st));
    }
    HIPCHK(hipEventRecord(c->ev[1], st));
    int comps = 0;
    {
        const hipError_t e = hipMemcpyAsync(&comps, c->total.p, 4, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess)
            return fail(c, SDB_EHIP, "component count read failed: %s; device %s (%zu bytes); host %s",
                        hipGetErrorString(e), ptr_desc(c->total.p).c_str(), c->total.bytes, ptr_desc(&comps).c_str());
    }
    if (labels_out) HIPCHK(hipMemcpyAsync(labels_out, c->labels.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c->stats.t_cluster_ms = elapsed(c->ev[0], c->ev[1]);
    c->stats.n_components = comps;
    c->clustered = true;
    if (n_clusters_out) *n_clusters_out = comps == 0 ? 2 : comps + 1;  // src/sDbscan.cpp:633
    return SDB_OK;
}

int sdb_label_noise(sdb_ctx* c, const int* sampled_core, int n_sampled, int* labels_out) {
    if (!c) return SDB_EINVAL;
    if (!c->clustered) return fail(c, SDB_ESTATE, "form_clusters first");
    if (c->sng) return fail(c, SDB_ESTATE, "noise labelling needs the random-projection index");
    if (n_sampled < 0 || (n_sampled > 0 && !sampled_core)) return fail(c, SDB_EINVAL, "bad sampled core list");
    HIPCHK(hipSetDevice(c->device));
    const sdb_params& q = c->prm;
    const long long n = c->n;
    hipStream_t st = c->st;
    for (int i = 0; i < n_sampled; ++i)
        if (sampled_core[i] < 0 || sampled_core[i] >= n) return fail(c, SDB_EINVAL, "sampled point %d out of range", sampled_core[i]);
    if (n_sampled > 0) {
        const int P = q.n_proj;
        ENSURE(c->ns_ids, (size_t)n_sampled * 4);
        ENSURE(c->cproj, (size_t)P * n_sampled * 4);
        HIPCHK(hipMemcpyAsync(c->ns_ids.p, sampled_core, (size_t)n_sampled * 4, hipMemcpyHostToDevice, st));
        // projections of the sampled cores (src/sDbscan.cpp:720-770)
        IndexArgs ia{};
        ia.X = c->X;
        ia.Xn = P_<float>(c->Xn);
        ia.err = P_<int>(c->err_flag);
        ia.signmask = P_<uint32_t>(c->bits);
        ia.n = n; ia.ldx = c->ldx; ia.d = c->d; ia.metric = q.metric; ia.ker = q.ker_n_features;
        ia.kerL = q.ker_intervalSampling;
        ia.R = (q.metric == SDB_L1 || q.metric == SDB_L2) ? P_<float>(c->R) : nullptr;
        ia.P = P; ia.K = q.topK;
        ia.pt_ids = P_<int>(c->ns_ids);
        ia.n_ids = n_sampled;
        ia.pt_buf = P_<float>(c->cproj);
        const int G = (int)std::max<long long>(1, std::min<long long>(2048, ((long long)n_sampled + 3) / 4));
        LCHK(launch_index(ia, ilog2(q.fhtDim), 3, G, st));
        // noise points and their best sampled core (src/sDbscan.cpp:775-806)
        ENSURE(c->noise, (size_t)n * 4);
        ENSURE(c->qcount, 4);
        HIPCHK(hipMemsetAsync(c->qcount.p, 0, 4, st));
        LCHK(launch_noise_list(P_<int>(c->labels), n, P_<int>(c->noise), P_<int>(c->qcount), st));
        int nn = 0;
        HIPCHK(hipMemcpyAsync(&nn, c->qcount.p, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (nn > 0) {
            ENSURE(c->nbest, (size_t)nn * 8);
            HIPCHK(hipMemsetAsync(c->nbest.p, 0, (size_t)nn * 8, st));
            LCHK(launch_noise_score(P_<float>(c->cproj), n_sampled, P_<int>(c->topK), q.topK, P_<int>(c->noise), nn,
                                    (unsigned long long*)c->nbest.p, st));
            LCHK(launch_noise_assign(P_<int>(c->noise), nn, (const unsigned long long*)c->nbest.p, P_<int>(c->ns_ids),
                                     P_<int>(c->labels), st));
        }
    }
    if (labels_out) HIPCHK(hipMemcpyAsync(labels_out, c->labels.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return SDB_OK;
}

static int build_neighbors_sng(sdb_ctx* c);

static int build_neighbors(sdb_ctx* c) {
    if (c->have_nb) return 0;
    if (c->sng) return build_neighbors_sng(c);
    hipStream_t st = c->st;
    const long long n = c->n;
    ENSURE(c->nb_ptr, (size_t)(n + 1) * 8);
    ENSURE(c->nb_cur, (size_t)n * 4);
    // ptr = exclusive scan of the degrees (64-bit)
    std::vector<int> deg(n);
    HIPCHK(hipMemcpyAsync(deg.data(), c->degree.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<long long> ptr(n + 1, 0);
    for (long long a = 0; a < n; ++a) ptr[a + 1] = ptr[a] + deg[a];
    const long long tot = ptr[n];
    // the segmented sorts count entries in 32 bits
    if (tot >= (1ll << 31)) return fail(c, SDB_EUNSUPPORTED, "%lld neighbourhood entries", tot);
    HIPCHK(hipMemcpyAsync(c->nb_ptr.p, ptr.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(c->nb_cur.p, 0, (size_t)n * 4, st));
    ENSURE(c->nb_idx, (size_t)std::max<long long>(tot, 1) * 4);
    ENSURE(c->nb_dist, (size_t)std::max<long long>(tot, 1) * 4);
    ENSURE(c->nb_sorted, (size_t)std::max<long long>(tot, 1) * 4);
    ClusterArgs ca = cluster_args(c);
    LCHK(launch_csr_fill(ca, P_<long long>(c->nb_ptr), P_<int>(c->nb_cur), P_<int>(c->nb_idx), st));
    {
        std::vector<int> cur(n);
        HIPCHK(hipMemcpyAsync(cur.data(), c->nb_cur.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (long long a = 0; a < n; ++a)
            if (cur[a] != deg[a])
                return fail(c, SDB_EHIP, "neighbourhood of point %lld has %d entries, degree says %d", a, cur[a], deg[a]);
    }
    // sort every neighbourhood by index (segmented radix sort of the ids)
    if (tot > 0) {
        int* keys_out = (int*)c->nb_sorted.p;
        size_t tb = 0;
        hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, P_<int>(c->nb_idx), keys_out, (int)tot, (int)n,
                                                   P_<long long>(c->nb_ptr), P_<long long>(c->nb_ptr) + 1, 0, 32, st);
        DevBuf tmp;
        if (ensure(c, tmp, tb)) return SDB_ENOMEM;
        hipError_t e = hipcub::DeviceSegmentedRadixSort::SortKeys(tmp.p, tb, P_<int>(c->nb_idx), keys_out, (int)tot, (int)n,
                                                                  P_<long long>(c->nb_ptr), P_<long long>(c->nb_ptr) + 1, 0, 32, st);
        HIPCHK(hipMemcpyAsync(c->nb_idx.p, keys_out, (size_t)tot * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        release(tmp);
        if (e != hipSuccess) return fail(c, SDB_EHIP, "segmented sort failed: %s", hipGetErrorString(e));
        // each symmetric pair's distance once: the entries with id >= the
        // row's point, then the others copied from their mirror entries; if
        // any entry has no mirror, every entry is computed.  Rows of up to
        // 128 features are computed whole: there the mirror's binary searches
        // cost more than the half pass saves (400k points: d = 32 68 -> 30 +
        // 49 ms, d = 128 128 -> 65 + 74 ms, d = 784 766 -> 369 + 74 ms).
        bool full = c->d <= 128 || c->d > 1024 || c->prm.metric == SDB_JS;
        if (!full) {
            DevBuf miss;
            if (ensure(c, miss, 16)) return SDB_ENOMEM;
            HIPCHK(hipMemsetAsync(miss.p, 0, 4, st));
            LCHK(launch_csr_dist(P_<float>(c->Xn), c->ldx, c->d, c->prm.metric, n, P_<long long>(c->nb_ptr),
                                 P_<int>(c->nb_idx), P_<float>(c->nb_dist), 1, st));
            LCHK(launch_csr_mirror(n, P_<long long>(c->nb_ptr), P_<int>(c->nb_idx), P_<float>(c->nb_dist),
                                   P_<unsigned int>(miss), st));
            unsigned int missing = 0;
            HIPCHK(hipMemcpyAsync(&missing, miss.p, 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            release(miss);
            full = missing != 0;
        }
        if (full)
            LCHK(launch_csr_dist(P_<float>(c->Xn), c->ldx, c->d, c->prm.metric, n, P_<long long>(c->nb_ptr),
                                 P_<int>(c->nb_idx), P_<float>(c->nb_dist), 0, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    c->have_nb = true;
    return 0;
}

int sdb_get_neighbors(sdb_ctx* c, long long* ptr, int* idx, float* dist) {
    if (!c || !ptr) return SDB_EINVAL;
    if (!c->cored) return fail(c, SDB_ESTATE, "find_core first");
    HIPCHK(hipSetDevice(c->device));
    int r = build_neighbors(c);
    if (r) return r;
    const long long n = c->n;
    hipStream_t st = c->st;
    HIPCHK(hipMemcpyAsync(ptr, c->nb_ptr.p, (size_t)(n + 1) * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const long long tot = ptr[n];
    if (idx && tot) HIPCHK(hipMemcpyAsync(idx, c->nb_idx.p, (size_t)tot * 4, hipMemcpyDeviceToHost, st));
    if (dist && tot) HIPCHK(hipMemcpyAsync(dist, c->nb_dist.p, (size_t)tot * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return SDB_OK;
}

int sdb_core_dist(sdb_ctx* c, int minPts, float* core_dist_out) {
    if (!c || !core_dist_out) return SDB_EINVAL;
    if (!c->cored) return fail(c, SDB_ESTATE, "find_core first");
    if (minPts < 2) return fail(c, SDB_EINVAL, "minPts must be at least 2 for core distances");
    HIPCHK(hipSetDevice(c->device));
    int r = build_neighbors(c);
    if (r) return r;
    hipStream_t st = c->st;
    const long long n = c->n;
    ENSURE(c->core_dist, (size_t)n * 4);
    // the (minPts-1)-th smallest distance of each core row, selected in place
    LCHK(launch_kth(n, P_<long long>(c->nb_ptr), P_<float>(c->nb_dist), P_<uint8_t>(c->core), minPts - 2,
                    P_<float>(c->core_dist), st));
    HIPCHK(hipMemcpyAsync(core_dist_out, c->core_dist.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return SDB_OK;
}

int sdb_optics_reach(sdb_ctx* c, const int* ordering, float* reach_out) {
    if (!c || !ordering || !reach_out) return SDB_EINVAL;
    if (!c->have_nb || !c->core_dist.p) return fail(c, SDB_ESTATE, "core_dist first");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->st;
    const long long n = c->n;
    std::vector<int> pos(n, -1);
    for (long long k = 0; k < n; ++k) {
        int p = ordering[k];
        if (p < 0 || p >= n || pos[p] >= 0) return fail(c, SDB_EINVAL, "ordering is not a permutation");
        pos[p] = (int)k;
    }
    ENSURE(c->pos, (size_t)n * 4);
    ENSURE(c->reach, (size_t)n * 4);
    HIPCHK(hipMemcpyAsync(c->pos.p, pos.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    LCHK(launch_reach(n, P_<long long>(c->nb_ptr), P_<int>(c->nb_idx), P_<float>(c->nb_dist), P_<uint8_t>(c->core),
                      P_<float>(c->core_dist), P_<int>(c->pos), P_<uint32_t>(c->reach), st));
    HIPCHK(hipMemcpyAsync(reach_out, c->reach.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return SDB_OK;
}

// formOptics_scikit (src/sDbscan.cpp:1588-1703) on the host over the device
// neighbourhoods: points in ascending order start a walk; a core point pushes
// its unprocessed neighbours with reach max(core_dist, d) into a lazy
// min-heap keyed on the FP32 reach only (ties in the heap's own order, as the
// reference's priority_queue of (index, value) pairs ordered by value).
namespace {
struct ReachPair {
    int idx;
    float val;
    ReachPair(int i, double v) : idx(i), val((float)v) {}
    bool operator>(const ReachPair& o) const { return val > o.val; }
};
typedef std::priority_queue<ReachPair, std::vector<ReachPair>, std::greater<ReachPair>> ReachHeap;
}  // namespace

// The neighbour scan of a core point: a neighbour xj with reach NEG (unseen)
// or above rr = max(core_dist, d) takes rr and is pushed; processed points
// hold -inf, which neither test accepts.
static inline void expand_scalar(long long j0, long long j1, float cd, const int* idx, const float* dd, float* reach,
                                 ReachHeap& seeds) {
    const float NEG = -2147483648.0f;  // src/Header.h:20
    for (long long j = j0; j < j1; ++j) {
        const int xj = idx[j];
        const float rj = reach[xj];
        const float rr = std::max(cd, dd[j]);
        if (rj == NEG || rj > rr) {
            reach[xj] = rr;
            seeds.emplace(xj, rr);
        }
    }
}

// The same scan 16 neighbours at a time: gathered reach values and the two
// tests in AVX-512, then the flagged neighbours in order through the scalar
// test against the current reach (so repeated ids behave as in the scalar
// loop) -- the pushes, and so the heap, are those of expand_scalar.
__attribute__((target("avx512f"))) static void expand_avx512(long long j0, long long j1, float cd, const int* idx,
                                                             const float* dd, float* reach, ReachHeap& seeds) {
    const float NEG = -2147483648.0f;
    const __m512 vcd = _mm512_set1_ps(cd), vneg = _mm512_set1_ps(NEG);
    long long j = j0;
    for (; j + 16 <= j1; j += 16) {
        const __m512i vi = _mm512_loadu_si512((const void*)(idx + j));
        const __m512 rj = _mm512_i32gather_ps(vi, reach, 4);
        const __m512 vd = _mm512_loadu_ps(dd + j);
        const __m512 rr = _mm512_mask_blend_ps(_mm512_cmp_ps_mask(vcd, vd, _CMP_LT_OQ), vcd, vd);  // std::max(cd, d)
        unsigned m = (unsigned)(_mm512_cmp_ps_mask(rj, vneg, _CMP_EQ_OQ) | _mm512_cmp_ps_mask(rj, rr, _CMP_GT_OQ));
        while (m) {
            const int b = __builtin_ctz(m);
            m &= m - 1;
            const int xj = idx[j + b];
            const float r = std::max(cd, dd[j + b]);
            if (reach[xj] == NEG || reach[xj] > r) {
                reach[xj] = r;
                seeds.emplace(xj, r);
            }
        }
    }
    expand_scalar(j, j1, cd, idx, dd, reach, seeds);
}

static void optics_walk(long long n, const long long* ptr, const int* idx, const float* dd, const uint8_t* core,
                        const float* cdist, int* ordering_out) {
    const float NEG = -2147483648.0f;  // src/Header.h:20
    // one array for the reach and the processed flag: a processed point holds
    // -inf, which no update test (reach == NEG or reach > rr) accepts
    const float DONE = -INFINITY;
    static const bool wide = __builtin_cpu_supports("avx512f");
    std::vector<float> reach(n, NEG);
    long long k = 0;
    auto expand = [&](int xi, ReachHeap& seeds) {
        if (wide) expand_avx512(ptr[xi], ptr[xi + 1], cdist[xi], idx, dd, reach.data(), seeds);
        else expand_scalar(ptr[xi], ptr[xi + 1], cdist[xi], idx, dd, reach.data(), seeds);
    };
    for (long long xi = 0; xi < n; ++xi) {
        if (reach[xi] == DONE) continue;
        reach[xi] = DONE;
        ordering_out[k++] = (int)xi;
        if (!core[xi]) continue;
        ReachHeap seeds;
        expand((int)xi, seeds);
        while (!seeds.empty()) {
            const int xj = seeds.top().idx;
            seeds.pop();
            if (reach[xj] == DONE) continue;
            reach[xj] = DONE;
            ordering_out[k++] = xj;
            if (core[xj]) expand(xj, seeds);
        }
    }
}

}  // extern "C"

// C++ linkage: shared with the batched drivers (declared in sdb_ctx.hpp)
int optics_orders(sdb_ctx* c, int n_minpts, const int* minPts, int* orderings, float* reaches, HostListPool* pool) {
    if (!c || n_minpts < 1 || !minPts || !orderings) return SDB_EINVAL;
    if (!c->cored) return fail(c, SDB_ESTATE, "find_core first");
    // sample-sequence neighbourhoods carry the core flags of their own call
    if (c->sng && n_minpts != 1) return fail(c, SDB_EUNSUPPORTED, "one minPts per call after sng_find_core");
    const long long n = c->n;
    const float eps = c->eps;
    const int keep = c->minPts;
    // the neighbourhoods of eps do not depend on minPts: fetched once
    std::vector<long long> ptr(n + 1);
    int r = sdb_get_neighbors(c, ptr.data(), nullptr, nullptr);
    if (r) return r;
    HostListPool own;
    HostListPool& lp = pool ? *pool : own;
    const size_t nent = (size_t)std::max<long long>(ptr[n], 1);
    HostLists* lists = lp.take(nent * 8);
    struct Give {
        HostListPool& lp;
        HostLists* h;
        ~Give() { lp.give(h); }
    } give{lp, lists};
    if (!lists->reserve(nent * 8))
        return fail(c, SDB_ENOMEM, "%zu bytes of host memory for the neighbourhood lists", nent * 8);
    int* idx = (int*)lists->p;
    float* dd = (float*)((char*)lists->p + nent * 4);
    r = sdb_get_neighbors(c, ptr.data(), idx, dd);
    if (r) return r;
    std::vector<std::vector<uint8_t>> core(n_minpts, std::vector<uint8_t>(n));
    std::vector<std::vector<float>> cdist(n_minpts, std::vector<float>(n));
    for (int m = 0; m < n_minpts; ++m) {
        if (c->sng) {
            HIPCHK(hipMemcpy(core[m].data(), c->core.p, (size_t)n, hipMemcpyDeviceToHost));
        } else {
            r = sdb_find_core(c, eps, minPts[m], core[m].data(), nullptr);
            if (r) return r;
        }
        r = sdb_core_dist(c, std::max(minPts[m], 2), cdist[m].data());
        if (r) return r;
    }
    // the walks are serial each; the settings run side by side
    std::vector<std::thread> th;
    for (int m = 1; m < n_minpts; ++m)
        th.emplace_back(optics_walk, n, ptr.data(), idx, dd, core[m].data(), cdist[m].data(),
                        orderings + (size_t)m * n);
    optics_walk(n, ptr.data(), idx, dd, core[0].data(), cdist[0].data(), orderings);
    for (auto& t : th) t.join();
    if (reaches) {
        for (int m = 0; m < n_minpts; ++m) {
            // core flags and core distances of this setting back on the device
            HIPCHK(hipMemcpyAsync(c->core.p, core[m].data(), (size_t)n, hipMemcpyHostToDevice, c->st));
            HIPCHK(hipMemcpyAsync(c->core_dist.p, cdist[m].data(), (size_t)n * 4, hipMemcpyHostToDevice, c->st));
            r = sdb_optics_reach(c, orderings + (size_t)m * n, reaches + (size_t)m * n);
            if (r) return r;
        }
    }
    if (c->sng) return SDB_OK;
    return sdb_find_core(c, eps, keep, nullptr, nullptr);  // the core flags of the last find_core again
}

extern "C" {

int sdb_optics_orders(sdb_ctx* c, int n_minpts, const int* minPts, int* orderings, float* reaches) {
    return optics_orders(c, n_minpts, minPts, orderings, reaches, nullptr);
}

int sdb_optics_order(sdb_ctx* c, int minPts, int* ordering_out, float* reach_out) {
    return sdb_optics_orders(c, 1, &minPts, ordering_out, reach_out);
}

int sdb_set_tile_rows(sdb_ctx* c, int rows) {
    if (!c) return SDB_EINVAL;
    if (rows < 0) return fail(c, SDB_EINVAL, "tile rows must be >= 0");
    c->tile_rows = rows == 0 ? 0 : std::min(std::max(rows, 64), 16 * 8 * 255);
    return SDB_OK;
}

int sdb_tune_tile_rows(sdb_ctx* c, float eps, const int* cand, int ncand, int* best_rows, double* best_ms) {
    if (!c || !cand || ncand < 1) return SDB_EINVAL;
    if (!c->indexed || c->sng) return fail(c, SDB_ESTATE, "build_index first");
    HIPCHK(hipSetDevice(c->device));
    int best = 0;
    double bms = 0.0;
    for (int i = 0; i < ncand; ++i) {
        if (cand[i] <= 0) return fail(c, SDB_EINVAL, "tile row candidates must be positive");
        int r = sdb_set_tile_rows(c, cand[i]);
        if (r) return r;
        int nq = 0;
        double band = 0, band32 = 0;
        r = run_distances(c, eps, &nq, &band, &band32);
        if (r) return r;
        HIPCHK(hipStreamSynchronize(c->st));
        const double ms = c->stats.t_dist_ms;
        if (i == 0 || ms < bms) { bms = ms; best = c->tile_rows; }
    }
    c->tile_rows = best;
    c->cored = false;  // the hit masks belong to the last candidate; degrees were not counted
    c->have_nb = false;
    c->clustered = false;
    if (best_rows) *best_rows = best;
    if (best_ms) *best_ms = bms;
    return SDB_OK;
}

// transformData (src/Utilities.h:289-335) on its own, for the paths that do
// not build the index (the index passes fuse it for Cosine).
static int ensure_transformed(sdb_ctx* c) {
    const sdb_params& q = c->prm;
    if (c->transformed && (q.precision != SDB_BF16 || c->transformed_prec == SDB_BF16)) return 0;
    hipStream_t st = c->st;
    const long long n = c->n;
    const bool bf = q.precision == SDB_BF16;
    ENSURE(c->Xn, (size_t)n * c->ldx * 4);
    if (bf) {
        ENSURE(c->Xb, (size_t)n * c->ldx * 2);
        ENSURE(c->rowerr, (size_t)n * 4);
        ENSURE(c->rowerr_max, 4);
        HIPCHK(hipMemsetAsync(c->rowerr_max.p, 0, 4, st));
    }
    ENSURE(c->err_flag, 4);
    HIPCHK(hipMemsetAsync(c->err_flag.p, 0, 4, st));
    LCHK(launch_transform(c->X, n, c->d, c->ldx, q.metric, P_<float>(c->Xn), bf ? P_<uint16_t>(c->Xb) : nullptr,
                          bf ? P_<float>(c->rowerr) : nullptr, bf ? P_<uint32_t>(c->rowerr_max) : nullptr,
                          bf ? bf16_rho_pad(c) : 0, P_<int>(c->err_flag), st));
    int errf = 0;
    HIPCHK(hipMemcpyAsync(&errf, c->err_flag.p, 4, hipMemcpyDeviceToHost, st));
    if (bf) HIPCHK(hipMemcpyAsync(&c->rho_max, c->rowerr_max.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (errf & 1) return fail(c, SDB_ENEGATIVE, "Error: X is not non-negative !");
    if (errf & 2) return fail(c, SDB_EZERO, "Error: There is an zero point !");
    c->transformed = true;
    c->rnorm2_of = nullptr;  // the rows were written again
    c->transformed_prec = q.precision;
    return 0;
}

int sdb_transform(sdb_ctx* c, float* X_out) {
    if (!c || !X_out) return SDB_EINVAL;
    if (!c->X) return fail(c, SDB_ESTATE, "no input loaded");
    HIPCHK(hipSetDevice(c->device));
    int r = ensure_transformed(c);
    if (r) return r;
    return sdb_get_transformed(c, X_out);
}

// Deduplicated neighbourhood CSR of the sample-sequence hits (sorted by index)
// with FP32 distances.
static int build_neighbors_sng(sdb_ctx* c) {
    if (c->have_nb) return 0;
    hipStream_t st = c->st;
    const long long n = c->n;
    const int L = c->sngL, S2 = c->S, Ms = c->Mt, W = c->W;
    const int both = c->sng_flags & 1;
    ENSURE(c->csr_cnt, (size_t)(n + 1) * 8);
    ENSURE(c->csr_ptr, (size_t)(n + 1) * 8);
    HIPCHK(hipMemsetAsync(c->csr_cnt.p, 0, (size_t)(n + 1) * 8, st));
    LCHK(launch_sng_csr_count(n, L, S2, Ms, W, both, P_<uint32_t>(c->E), P_<int>(c->lists), P_<long long>(c->csr_cnt),
                              st));
    std::vector<long long> cnt(n), ptr(n + 1, 0);
    HIPCHK(hipMemcpyAsync(cnt.data(), c->csr_cnt.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (long long a = 0; a < n; ++a) ptr[a + 1] = ptr[a] + cnt[a];
    const long long tot = ptr[n];
    if (tot >= (1ll << 31)) return fail(c, SDB_EUNSUPPORTED, "%lld neighbourhood entries", tot);
    HIPCHK(hipMemcpyAsync(c->csr_ptr.p, ptr.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st));
    ENSURE(c->csr_cur, (size_t)n * 8);
    HIPCHK(hipMemsetAsync(c->csr_cur.p, 0, (size_t)n * 8, st));
    ENSURE(c->csr_idx, (size_t)std::max<long long>(tot, 1) * 4);
    ENSURE(c->nb_sorted, (size_t)std::max<long long>(tot, 1) * 4);
    LCHK(launch_sng_csr_fill(n, L, S2, Ms, W, both, P_<uint32_t>(c->E), P_<int>(c->lists), P_<long long>(c->csr_ptr),
                             (unsigned long long*)c->csr_cur.p, P_<int>(c->csr_idx), st));
    int* sorted = (int*)c->nb_sorted.p;
    if (tot > 0) {
        size_t tb = 0;
        hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, P_<int>(c->csr_idx), sorted, (int)tot, (int)n,
                                                   P_<long long>(c->csr_ptr), P_<long long>(c->csr_ptr) + 1, 0, 32, st);
        DevBuf tmp;
        if (ensure(c, tmp, tb)) return SDB_ENOMEM;
        hipError_t e = hipcub::DeviceSegmentedRadixSort::SortKeys(tmp.p, tb, P_<int>(c->csr_idx), sorted, (int)tot, (int)n,
                                                                  P_<long long>(c->csr_ptr), P_<long long>(c->csr_ptr) + 1, 0, 32, st);
        HIPCHK(hipStreamSynchronize(st));
        release(tmp);
        if (e != hipSuccess) return fail(c, SDB_EHIP, "segmented sort failed: %s", hipGetErrorString(e));
    }
    // distinct ids per row
    LCHK(launch_uniq(n, P_<long long>(c->csr_ptr), sorted, P_<long long>(c->csr_cnt), st));
    HIPCHK(hipMemcpyAsync(cnt.data(), c->csr_cnt.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::fill(ptr.begin(), ptr.end(), 0);
    for (long long a = 0; a < n; ++a) ptr[a + 1] = ptr[a] + cnt[a];
    const long long utot = ptr[n];
    ENSURE(c->nb_ptr, (size_t)(n + 1) * 8);
    HIPCHK(hipMemcpyAsync(c->nb_ptr.p, ptr.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st));
    ENSURE(c->nb_idx, (size_t)std::max<long long>(utot, 1) * 4);
    ENSURE(c->nb_dist, (size_t)std::max<long long>(utot, 1) * 4);
    LCHK(launch_uniq_fill(n, P_<long long>(c->csr_ptr), sorted, P_<long long>(c->nb_ptr), P_<int>(c->nb_idx), st));
    if (utot > 0)
        LCHK(launch_csr_dist(P_<float>(c->Xn), c->ldx, c->d, c->prm.metric, n, P_<long long>(c->nb_ptr),
                             P_<int>(c->nb_idx), P_<float>(c->nb_dist), 0, st));
    HIPCHK(hipStreamSynchronize(st));
    c->have_nb = true;
    return 0;
}

int sdb_sng_find_core(sdb_ctx* c, const int* samples, int n_samples, const int* win_end, int flags, float eps,
                      int minPts, uint8_t* core_out, int* degree_out) {
    if (!c) return SDB_EINVAL;
    if (!c->X) return fail(c, SDB_ESTATE, "no input loaded");
    if (n_samples < 0 || (n_samples > 0 && !samples)) return fail(c, SDB_EINVAL, "bad sample sequence");
    HIPCHK(hipSetDevice(c->device));
    const long long n = c->n;
    hipStream_t st = c->st;
    for (int j = 0; j < n_samples; ++j)
        if (samples[j] < 0 || samples[j] >= n) return fail(c, SDB_EINVAL, "sample %d out of range", samples[j]);
    const bool both = (flags & 1) != 0, dedup = (flags & 2) != 0;
    int r = ensure_transformed(c);
    if (r) return r;

    // chunks of the sequence act as lists of Ms points
    const int Ms = c->sng_Ms, W = (Ms + 31) / 32;
    const int L = std::max(1, (n_samples + Ms - 1) / Ms);
    int S2 = 1;
    while (S2 < L) S2 <<= 1;
    if ((long long)n * S2 >= (1ll << 31)) return fail(c, SDB_EUNSUPPORTED, "n_points x sample chunks too large");
    std::vector<int> seq((size_t)L * Ms, n_samples > 0 ? samples[0] : 0);
    std::copy(samples, samples + n_samples, seq.begin());
    ENSURE(c->lists, seq.size() * 4);
    HIPCHK(hipMemcpyAsync(c->lists.p, seq.data(), seq.size() * 4, hipMemcpyHostToDevice, st));
    // positions of every point in the sequence (the point itself is skipped)
    std::vector<int> optr(n + 1, 0), opos(std::max(n_samples, 1));
    for (int j = 0; j < n_samples; ++j) ++optr[samples[j] + 1];
    for (long long a = 0; a < n; ++a) optr[a + 1] += optr[a];
    {
        std::vector<int> cur(optr.begin(), optr.end() - 1);
        for (int j = 0; j < n_samples; ++j) opos[cur[samples[j]]++] = j;
    }
    ENSURE(c->occ_ptr, (size_t)(n + 1) * 4);
    ENSURE(c->occ_pos, opos.size() * 4);
    HIPCHK(hipMemcpyAsync(c->occ_ptr.p, optr.data(), (size_t)(n + 1) * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->occ_pos.p, opos.data(), opos.size() * 4, hipMemcpyHostToDevice, st));
    if (win_end) {
        ENSURE(c->win, (size_t)n * 4);
        HIPCHK(hipMemcpyAsync(c->win.p, win_end, (size_t)n * 4, hipMemcpyHostToDevice, st));
    }
    ENSURE(c->F, (size_t)n * S2 * W * 4);
    LCHK(launch_sng_fmask(n, L, S2, Ms, W, n_samples, win_end ? P_<int>(c->win) : nullptr, P_<int>(c->occ_ptr),
                          P_<int>(c->occ_pos), P_<uint32_t>(c->F), st));
    ENSURE(c->grp, (size_t)n * L * 4);
    LCHK(launch_sng_grp(n, L, S2, P_<uint32_t>(c->grp), st));
    c->h_grp_ptr.resize(L + 1);
    for (int l = 0; l <= L; ++l) c->h_grp_ptr[l] = (int)((long long)l * n);
    c->nblk = 1;
    c->L2 = L;
    c->S = S2;
    c->W = W;
    c->Mt = Ms;
    c->sngL = L;
    c->sng = true;
    c->sng_flags = flags;
    c->indexed = false;
    c->have_nb = false;
    c->clustered = false;
    c->stats = sdb_stats{};

    int nq = 0;
    double band = 0, band32 = 0;
    r = run_distances(c, eps, &nq, &band, &band32);
    if (r) return r;
    ENSURE(c->degree, (size_t)n * 4);
    ENSURE(c->core, (size_t)n);
    ENSURE(c->core_spread, (size_t)kSpread * kSpreadStride * 8);
    HIPCHK(hipMemsetAsync(c->core_spread.p, 0, c->core_spread.bytes, st));
    unsigned long long* ncore = P_<unsigned long long>(c->core_spread);
    if (!dedup) {
        // sngDbscan: |N(a)| with duplicates (src/sDbscan.cpp:487-501)
        ENSURE(c->rev, (size_t)n * 4);
        HIPCHK(hipMemsetAsync(c->rev.p, 0, (size_t)n * 4, st));
        if (both) LCHK(launch_sng_rev(n_samples, P_<int>(c->lists), P_<int>(c->colsum), P_<int>(c->rev), st));
        LCHK(launch_sng_degree(n, L, S2, W, P_<uint32_t>(c->E), P_<int>(c->rev), minPts, P_<int>(c->degree),
                               P_<uint8_t>(c->core), ncore, st));
    } else {
        // sngOptics: deduplicated neighbourhoods (src/sDbscan.cpp:1528-1540)
        r = build_neighbors_sng(c);
        if (r) return r;
        LCHK(launch_rowsize_core(n, P_<long long>(c->nb_ptr), minPts, P_<int>(c->degree), P_<uint8_t>(c->core), ncore,
                                 st));
    }
    LCHK(launch_spread_fold(ncore, (unsigned long long*)c->stats64.p + 3, st));
    HIPCHK(hipEventRecord(c->ev[6], st));
    unsigned long long hs[8] = {0};
    HIPCHK(hipMemcpyAsync(hs, c->stats64.p, 64, hipMemcpyDeviceToHost, st));
    if (core_out) HIPCHK(hipMemcpyAsync(core_out, c->core.p, (size_t)n, hipMemcpyDeviceToHost, st));
    if (degree_out) HIPCHK(hipMemcpyAsync(degree_out, c->degree.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c->stats.t_core_ms = elapsed(c->ev[3], c->ev[6]);
    c->stats.n_pairs = (long long)n * n_samples;
    c->stats.n_queue = nq;
    c->stats.n_recheck32 = (long long)hs[0];
    c->stats.n_recheck64 = (long long)hs[1];
    c->stats.n_recheck_hits = (long long)hs[2];
    c->stats.n_near_eps = (long long)hs[4];
    c->stats.n_core = (long long)hs[3];
    c->stats.band = band;
    c->stats.band32 = band32;
    c->eps = eps;
    c->minPts = minPts;
    c->cored = true;
    return SDB_OK;
}

int sdb_get_stats(const sdb_ctx* c, sdb_stats* out) {
    if (!c || !out) return SDB_EINVAL;
    *out = c->stats;
    return SDB_OK;
}

}  // extern "C"
