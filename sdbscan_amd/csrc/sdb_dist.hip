// Candidate distances and the eps decision of sDbscan on gfx950.
//
// Reference: the candidate loop of rp_findCorePoints (src/sDbscan.cpp:243-359)
// and computeDist (src/Utilities.h:212-287).  Instead of one point at a time,
// work is grouped by top-M list: the (point, slot) entries that visit list l
// form the query rows of a tile whose candidate rows are the M points of l,
// kept in LDS.  Each pair is decided as "distance <= eps" exactly: a pair
// whose fast distance lies within a derived error band of eps is queued and
// re-decided by k_recheck (FP32, then FP64 in a fixed order), so every
// precision makes the same decision.  Hits are stored only for the first
// occurrence of a candidate among the point's slots (F masks), which is the
// reference's per-point dedup (src/sDbscan.cpp:260-282).
#include <cstdlib>

#include "sdb_kernels.hpp"

namespace sdb {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// The Chi2 term x y / (x + y) of the FP32 tiles through the hardware
// reciprocal (1 ulp): about 5u relative per term against 3u for a correctly
// rounded division, so a tile distance stays within (d + 6) u of the exact
// one, inside the 2 (d + 8) u band.  Sums x + y below 2^-100 (terms below
// 2^-102) count as 0, which no band of 2^-23 scale notices.
__device__ __forceinline__ float chi2_term(float x, float y) {
    if (!(x > 0.0f && y > 0.0f)) return 0.0f;
    const float s = x + y;
    return s > 0x1p-100f ? x * y * __builtin_amdgcn_rcpf(s) : 0.0f;
}

// Reserves cnt slots per lane of a wave-wide append with one atomic: returns
// the lane's first slot (an inclusive prefix sum over the lanes, the total
// added to *counter by the last lane).
__device__ __forceinline__ unsigned long long wave_reserve(unsigned long long* counter, int cnt, int lane) {
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += y;
    }
    unsigned long long base = 0;
    if (lane == kWave - 1 && incl > 0) base = atomicAdd(counter, (unsigned long long)incl);
    const uint32_t lo = __shfl((uint32_t)base, kWave - 1, kWave);
    const uint32_t hi = __shfl((uint32_t)(base >> 32), kWave - 1, kWave);
    return (((unsigned long long)hi << 32) | lo) + (unsigned long long)(incl - cnt);
}

__device__ __forceinline__ unsigned long long queue_entry(uint32_t ent, int l, int col) {
    return ((unsigned long long)ent << 32) | ((unsigned long long)(uint32_t)l << 8) | (unsigned)col;
}

// ---------------------------------------------------------------------------
// MFMA tiles (cosine).  A = 16 query rows x k, B = k x 16 candidates, via
// v_mfma_f32_16x16x32_bf16 (BF) or 4 x v_mfma_f32_16x16x4_f32.  Lane l loads
// 16 bytes of query row (l & 15) and of candidate column (l & 15) at k-group
// (l >> 4) per step; A and B use the same k assignment, so the sum over k is
// the dot product whatever order the unit uses internally.
// L2X (FP32 rows, metric L2): the tiles compute x.y and decide on
// D = fl(fl(|x|^2 + |y|^2) - 2 x.y) against a per-pair band.  With the
// squared norms rounded once to FP32 (u each) and the MFMA dot within
// 2(d + 2)u sum|x_i y_i| <= (d + 2)u (|x|^2 + |y|^2) of x.y, the two
// roundings of D add at most 3u (|x|^2 + |y|^2), so |D - ||x - y||^2| <=
// (2d + 8)u (|x|^2 + |y|^2); the host's kappa = 1.02 (2d + 16)u covers it.
template <bool BF, int QT, int NT, int WAVES, int PF, int BATCH, bool L2X>
__global__ __launch_bounds__(WAVES * 64) void k_dist_mfma(DistArgs a) {
    constexpr int SZ = BF ? 2 : 4;
    constexpr int SE = 64 / SZ;  // elements per 64-byte k-step
    constexpr int EW = (NT + 1) / 2;
    extern __shared__ __align__(16) unsigned char smem[];
    const int M = a.M;
    const size_t RSB = (size_t)a.kc * SZ + 16;
    int* col_lds = (int*)(smem + (size_t)M * RSB);
    const int tid = threadIdx.x;
    const int lane = lane_id(), w = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int l = a.tile_list[blockIdx.x];
    const int beg = a.tile_beg[blockIdx.x], end = a.tile_end[blockIdx.x];
    const int* Lm = a.lists + (size_t)l * M;
    const char* Xc = (const char*)a.X;
    const double lo = a.lo, hi = a.hi;

    int* pts_lds = col_lds + M;
    float* cn_lds = (float*)(pts_lds + M);  // L2X: squared norms of the candidates
    if (tid < M) {  // M <= 256 < WAVES * 64
        col_lds[tid] = 0;
        pts_lds[tid] = Lm[tid];
        if constexpr (L2X) cn_lds[tid] = a.rnorm2[Lm[tid]];
    }
    int colacc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) colacc[t] = 0;

    const int rows_per_set = WAVES * QT * 16;
    const int nsets = (end - beg + rows_per_set - 1) / rows_per_set;
    for (int set = 0; set < nsets; ++set) {
        uint32_t ent[QT];
        long long pa[QT];
        bool valid[QT], ffl[QT];
        float nq[QT];  // L2X: squared norm of the lane's query row
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            int rr = beg + set * rows_per_set + (w * QT + q) * 16 + r16;
            valid[q] = rr < end;
            const uint32_t gq = a.grp[valid[q] ? rr : beg];
            ent[q] = ent_id(gq);
            ffl[q] = ent_full(gq);
            pa[q] = (long long)(ent[q] / (uint32_t)a.S);
            nq[q] = L2X ? a.rnorm2[pa[q]] : 0.0f;
        }
        f32x4_t acc[QT][NT];
#pragma unroll
        for (int q = 0; q < QT; ++q)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[q][t] = f32x4_t{0.f, 0.f, 0.f, 0.f};

        for (int c = 0; c < a.nchunk; ++c) {
            const int k0 = c * a.kc;
            const int kn = min(a.kc, a.dk - k0);
            if (a.nchunk > 1 || set == 0) {
                // candidate rows of this k-chunk -> LDS, BATCH loads in flight
                __syncthreads();
                const int vpr = kn * SZ / 16, total = M * vpr;
                for (int v0 = tid; v0 < total; v0 += WAVES * 64 * BATCH) {
                    uint4 xs[BATCH];
#pragma unroll
                    for (int u = 0; u < BATCH; ++u) {
                        const int v = min(v0 + u * WAVES * 64, total - 1);  // clamped: loads stay unconditional
                        const int row = v / vpr, cv = v - row * vpr;
                        xs[u] = ((const uint4*)(Xc + ((size_t)pts_lds[row] * a.ldx + k0) * SZ))[cv];
                    }
#pragma unroll
                    for (int u = 0; u < BATCH; ++u) {
                        const int v = v0 + u * WAVES * 64;
                        const int row = v / vpr, cv = v - row * vpr;
                        uint4 x = xs[u];
                        // BF16 rows with their bound in element ldx - 1 (the
                        // high half of the row's last 16-byte vector): zero
                        // there, so that it adds nothing to the dot products
                        if constexpr (BF) {
                            if (a.rho_pad && k0 + kn == a.ldx && cv == vpr - 1) x.w &= 0xFFFFu;
                        }
                        if (v < total) *(uint4*)(smem + row * RSB + (size_t)cv * 16) = x;
                    }
                }
                __syncthreads();
            }
            const int nsteps = kn / SE;
            const char* arow[QT];
#pragma unroll
            for (int q = 0; q < QT; ++q) arow[q] = Xc + ((size_t)pa[q] * a.ldx + k0) * SZ + g * 16;
            // query fragments are kept PF k-steps ahead in a register ring so
            // that enough HBM reads are in flight per CU.  Ring loads are
            // unconditional (step index clamped to the last step) so the
            // compiler can wait on each load individually.
            uint4 af[PF][QT];
#pragma unroll
            for (int i = 0; i < PF; ++i)
#pragma unroll
                for (int q = 0; q < QT; ++q)
                    af[i][q] = *(const uint4*)(arow[q] + (size_t)min(i, nsteps - 1) * 64);
            auto mfma_step = [&](int st, const uint4 (&ar)[QT]) {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    int row = t * 16 + r16;
                    row = row < M ? row : M - 1;
                    const uint4 bfr = *(const uint4*)(smem + row * RSB + (size_t)st * 64 + g * 16);
#pragma unroll
                    for (int q = 0; q < QT; ++q) {
                        if constexpr (BF) {
                            const bf16x8_t av = __builtin_bit_cast(bf16x8_t, ar[q]);
                            const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, bfr);
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[q][t], 0, 0, 0);
                        } else {
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ar[q].x), __uint_as_float(bfr.x), acc[q][t], 0, 0, 0);
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ar[q].y), __uint_as_float(bfr.y), acc[q][t], 0, 0, 0);
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ar[q].z), __uint_as_float(bfr.z), acc[q][t], 0, 0, 0);
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ar[q].w), __uint_as_float(bfr.w), acc[q][t], 0, 0, 0);
                        }
                    }
                }
            };
            const int ngroups = nsteps / PF;
            const int rem = nsteps - ngroups * PF;
            for (int gi = 0; gi < ngroups; ++gi) {
#pragma unroll
                for (int i = 0; i < PF; ++i) {
                    const int st = gi * PF + i;
                    mfma_step(st, af[i]);
                    const int ns = min(st + PF, nsteps - 1);
#pragma unroll
                    for (int q = 0; q < QT; ++q) af[i][q] = *(const uint4*)(arow[q] + (size_t)ns * 64);
                }
            }
#pragma unroll
            for (int i = 0; i < PF - 1; ++i)
                if (i < rem) mfma_step(ngroups * PF + i, af[i]);
        }

        // epilogue: lane (r16, g) holds rows 4g..4g+3 of column t*16 + r16
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            uint32_t Fw[EW], Ew[EW];
#pragma unroll
            for (int x = 0; x < EW; ++x) {
                Fw[x] = x < a.W ? (ffl[q] ? fmask_full_word(x, M) : a.F[(size_t)ent[q] * a.W + x]) : 0u;
                Ew[x] = 0u;
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int col = t * 16 + r16;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 4 * g + i;
                    const uint32_t fword = __shfl(Fw[t >> 1], row, kWave);
                    const int vr = __shfl((int)valid[q], row, kWave);
                    const bool act = vr && col < M && ((fword >> (col & 31)) & 1u);
                    bool hit, unc;
                    if constexpr (L2X) {
                        const float nqr = __shfl(nq[q], row, kWave);
                        const float ncc = cn_lds[col < M ? col : M - 1];
                        const float D2 = (nqr + ncc) - 2.0f * acc[q][t][i];
                        const double band = a.kappa * ((double)nqr + (double)ncc) + a.eps0;
                        const double dv = (double)D2;
                        hit = act && dv < a.eps2 - band;
                        unc = act && !hit && !(dv > a.eps2 + band);
                    } else {
                        const float dd = 1.0f - acc[q][t][i];
                        const double dv = (double)dd;
                        hit = act && dv < lo;
                        unc = act && !(dv < lo) && !(dv > hi);
                    }
                    const unsigned long long bh = __ballot(hit);
                    if ((r16 & 3) == i)
                        Ew[t >> 1] |= (uint32_t)((bh >> (16 * (r16 >> 2))) & 0xffffull) << (16 * (t & 1));
                    colacc[t] += hit ? 1 : 0;
                    const uint32_t er = __shfl(ent[q], row, kWave);
                    if (unc) {
                        const unsigned long long pos = atomicAdd(a.qcount, 1ull);
                        if (pos < (unsigned long long)a.qcap) a.queue[pos] = queue_entry(er, l, col);
                    }
                }
            }
            if (g == 0 && valid[q]) {
#pragma unroll
                for (int x = 0; x < EW; ++x)
                    if (x < a.W) a.E[(size_t)ent[q] * a.W + x] = Ew[x];
            }
        }
    }
    // column hit counts -> colsum (reverse degrees of the list's points)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        int v = colacc[t];
        v += __shfl_xor(v, 16, kWave);
        v += __shfl_xor(v, 32, kWave);
        const int col = t * 16 + r16;
        if (g == 0 && col < M && v) atomicAdd(&col_lds[col], v);
    }
    __syncthreads();
    for (int i = tid; i < M; i += WAVES * 64) {
        int v = col_lds[i];
        if (v) atomicAdd(&a.colsum[(size_t)l * M + i], v);
    }
}

// ---------------------------------------------------------------------------
// Pipelined single-chunk BF16 tiles (cosine, every dim of a row in one LDS
// tile).  The list's candidate rows are the MFMA A operand (LDS), the query
// rows the B operand, so lane (r16, g) accumulates the dots of query r16 of
// its 16-row block with candidates t*16 + 4g + i (i < 4) and needs no
// cross-lane traffic for its query's masks.  Query fragments stream through a
// ring of PF k-steps that runs across blocks: the fragment of step st of the
// next block is loaded as soon as step st of the current block has been
// consumed, so the next block's rows are in flight during the MFMAs and the
// epilogue of the current one.
//   per-pair band from the rows' BF16 rounding bounds rho (bf16_row_err),
//   rho_c the largest over the list's candidates: hit iff s > th_hit +
//   rho_c + rho_q; queued for the recheck iff not a hit and s >= th_near -
//   rho_c - rho_q;
//   E words: per-lane nibbles OR-reduced across the 4 lanes of a query;
//   colsum: per-lane byte counters (<= 255 blocks per lane and tile).
__device__ __forceinline__ uint32_t spread4(uint32_t n) {
    // bit i of n -> byte i
    return (n * 0x00204081u) & 0x01010101u;
}

// F32L2: the same tiles on FP32 rows for L2, with 4 x v_mfma_f32_16x16x4f32
// per 64-byte k-step (lane (r16, g) feeds float c of its 16 bytes to the c-th
// MFMA, on both operands) and the decision of k_dist_mfma's L2X epilogue on
// D = fl(fl(|x|^2 + |y|^2) - 2 x.y) with the per-pair band kappa (|x|^2 +
// |y|^2) + eps0; the query's squared norm is lane-local, the candidates'
// come from LDS.
template <int NT, int WAVES, int NS, bool EXACT, int PADB, bool F32L2 = false>
__global__ __launch_bounds__(WAVES * 64) void k_dist_pipe(DistArgs a) {
    // NS k-steps per row; EXACT: the rows have exactly NS steps, otherwise
    // fewer, with the LDS rows zero-padded to NS steps and the query loads of
    // the missing steps clamped to the last step (their products are zero)
    constexpr int SZ = F32L2 ? 4 : 2;
    constexpr int EW = (NT + 1) / 2;
    constexpr int RSB = NS * 64 + PADB;  // row stride in LDS, padded against bank conflicts
    static_assert(NT <= 8, "candidate bits per lane exceed 32");
    extern __shared__ __align__(16) unsigned char smem[];
    const int M = a.M;
    int* col_lds = (int*)(smem + (size_t)M * RSB);
    const int tid = threadIdx.x;
    const int lane = lane_id(), w = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int l = a.tile_list[blockIdx.x];
    const int beg = a.tile_beg[blockIdx.x], end = a.tile_end[blockIdx.x];
    const int* Lm = a.lists + (size_t)l * M;
    const char* Xc = (const char*)a.X;
    const int nsteps = EXACT ? NS : a.dk / (64 / SZ);
    const size_t rowb = (size_t)a.ldx * SZ;

    const int nblk = (end - beg + 15) >> 4;
    auto entry_of = [&](int blk, uint32_t& e, bool& v, bool& f) {
        const int rr = beg + blk * 16 + r16;
        v = rr < end;
        const uint32_t gv = a.grp[v ? rr : beg];
        e = ent_id(gv);
        f = ent_full(gv);
    };
    auto row_of = [&](uint32_t e) {
        const size_t p = (size_t)ent_point(e, a.magicS);
        return Xc + p * rowb + g * 16;
    };
    auto step_off = [&](int st) { return (size_t)(EXACT ? st : min(st, nsteps - 1)) * 64; };
    // the list's point ids are requested first, then the first block's query
    // rows, so that both latencies overlap the staging of the candidate rows
    int* pts_lds = col_lds + M;
    // the largest rounding bound of the list's candidates (float bits)
    uint32_t* rc_lds = (uint32_t*)(pts_lds + M);
    // F32L2: the candidates' squared norms, NT * 16 of them (0 past M), at a
    // 16-byte boundary so that a lane reads its four as one vector
    float* cn_lds = (float*)(smem + (((size_t)M * RSB + (size_t)M * 8 + 16 + 15) & ~(size_t)15));
    int pid = 0;
    if (tid < M) pid = Lm[tid];
    // (waves without a block load a valid row they never use, which keeps the
    // loads unconditional and the compiler's wait counts exact)
    const bool active = w < nblk;
    int b = w;
    uint32_t ent, entn;
    bool valid, validn, full, fulln;
    entry_of(min(b, nblk - 1), ent, valid, full);
    entry_of(min(b + WAVES, nblk - 1), entn, validn, fulln);
    uint4 af[NS];
    {
        const char* rp = row_of(ent);
#pragma unroll
        for (int i = 0; i < NS; ++i) af[i] = *(const uint4*)(rp + step_off(i));
    }
    if (tid < M) {  // M <= 256 < WAVES * 64
        col_lds[tid] = 0;
        pts_lds[tid] = pid;
    }
    if constexpr (F32L2) {
        if (tid < NT * 16) cn_lds[tid] = tid < M ? a.rnorm2[pid] : 0.0f;
    }
    if (tid == 0) *rc_lds = 0u;

    // candidate rows -> LDS with BATCH 16-byte loads in flight per thread (the
    // list's point ids went through LDS first, so no row load waits on the
    // load of its id).  The candidate rows, the F masks and the E words are
    // touched once per tile: their accesses are non-temporal, so that they
    // do not push the query block's rows out of the last-level cache
    // (measured 0.7-1.1 ms faster at 8M x 784)
    __syncthreads();
    {
        constexpr int BATCH = 10;
        const int vpr = nsteps * 4, total = M * NS * 4;
        for (int v0 = tid; v0 < total; v0 += WAVES * 64 * BATCH) {
            uint4 xs[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                // clamped indices keep the loads unconditional; the zero fill
                // of the padding steps is a select
                const int v = min(v0 + u * WAVES * 64, total - 1);
                const int row = v / (NS * 4), cv = v - row * (NS * 4);
                const u32x4_t* src = ((const u32x4_t*)(Xc + (size_t)pts_lds[row] * rowb)) + min(cv, vpr - 1);
                const uint4 x = __builtin_bit_cast(uint4, __builtin_nontemporal_load(src));
                xs[u] = cv < vpr ? x : uint4{0u, 0u, 0u, 0u};
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int v = v0 + u * WAVES * 64;
                const int row = v / (NS * 4), cv = v - row * (NS * 4);
                uint4 x = xs[u];
                if constexpr (!F32L2) {
                    // rows with their bound in element ldx - 1 (the high half
                    // of the row's last 16-byte vector): the list's largest
                    // bound from there, and a zero in its place
                    if (a.rho_pad && cv == vpr - 1 && v < total) {
                        atomicMax(rc_lds, x.w & 0xFFFF0000u);  // bounds are >= 0
                        x.w &= 0xFFFFu;
                    }
                }
                if (v < total) *(uint4*)(smem + row * RSB + cv * 16) = x;
            }
        }
    }
    if constexpr (!F32L2) {
        if (!a.rho_pad && tid < M) atomicMax(rc_lds, __float_as_uint(a.rowerr[pts_lds[tid]]));  // bounds are >= 0
    }
    __syncthreads();
    // thresholds with the list's largest candidate bound; the query's own
    // bound is added per block
    const float rcm = __uint_as_float(*rc_lds);
    const float th_hit0 = a.th_hit + rcm, th_near0 = a.th_near - rcm;


    uint32_t colacc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) colacc[t] = 0u;
    int crow[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) crow[t] = min(t * 16 + r16, M - 1) * RSB + g * 16;

    if (active) {
        for (;;) {
            const bool has_next = b + WAVES < nblk;
            uint32_t Fw[EW];
#pragma unroll
            for (int x = 0; x < EW; ++x)
                Fw[x] = (valid && x < a.W)
                            ? (full ? fmask_full_word(x, M) : __builtin_nontemporal_load(&a.F[(size_t)ent * a.W + x]))
                            : 0u;
            // BF16: the query row's rounding bound (from the row's last
            // padding element, read below, when the rows carry it); F32L2:
            // its squared norm
            float eq = 0.0f;
            if (F32L2) eq = a.rnorm2[ent_point(ent, a.magicS)];
            else if (!a.rho_pad) eq = a.rowerr[ent_point(ent, a.magicS)];
            uint32_t eqb = 0u;  // the last k-step's high word of this lane's row piece
            uint32_t ent2;
            bool valid2, full2;
            entry_of(min(b + 2 * WAVES, nblk - 1), ent2, valid2, full2);
            const char* np = row_of(entn);

            // the candidate fragments are the same for every block; an empty
            // asm on their offsets keeps the compiler from hoisting all NS x NT
            // LDS reads out of the loop (they would not fit in registers)
#pragma unroll
            for (int t = 0; t < NT; ++t) asm volatile("" : "+v"(crow[t]));
            f32x4_t acc[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
            // candidate fragments are read one step ahead; a scheduling
            // barrier after every step keeps each ring reload next to the step
            // that consumed its slot (left alone, the scheduler sinks all NS
            // reloads below the MFMAs, which serialises HBM latency and math)
            uint4 cf[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) cf[t] = *(const uint4*)(smem + crow[t]);
#pragma unroll
            for (int st = 0; st < NS; ++st) {
                uint4 cn[NT];
                if (st + 1 < NS) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) cn[t] = *(const uint4*)(smem + crow[t] + (st + 1) * 64);
                }
                if constexpr (F32L2) {
                    const uint4 qv = af[st];
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(cf[t].x), __uint_as_float(qv.x), acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(cf[t].y), __uint_as_float(qv.y), acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(cf[t].z), __uint_as_float(qv.z), acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(cf[t].w), __uint_as_float(qv.w), acc[t], 0, 0, 0);
                    }
                } else {
                    const bf16x8_t qv = __builtin_bit_cast(bf16x8_t, af[st]);
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, cf[t]), qv,
                                                                         acc[t], 0, 0, 0);
                }
                if (st == NS - 1) eqb = af[st].w;  // (the last real step: step_off clamps)
                af[st] = *(const uint4*)(np + step_off(st));
                if (st + 1 < NS) __builtin_amdgcn_sched_group_barrier(0x100, NT, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, F32L2 ? 4 * NT : NT, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (st + 1 < NS) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) cf[t] = cn[t];
                }
            }

            // epilogue: lane (r16, g) holds s(query r16, candidate t*16 + 4g + i) in acc[t][i]
            if (!F32L2 && a.rho_pad) {
                // element ldx - 1 sits in the high half of lane (r16, 3)'s word
                eq = __uint_as_float((uint32_t)__shfl((int)eqb, (lane & 15) + 48, kWave) & 0xFFFF0000u);
            }
            uint32_t hw[EW];
#pragma unroll
            for (int x = 0; x < EW; ++x) hw[x] = 0u;
            // hit: s > th_hit + rho_c + rho_q; in the band: not a hit and
            // s >= th_near - rho_c - rho_q (rho_c: the list's largest bound)
            const float th_h = th_hit0 + eq, th_n = th_near0 - eq;
            uint32_t unc = 0u;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                uint32_t hr = 0u, nr = 0u;
                // F32L2: the squared norms of candidates t*16 + 4g + i
                float4 cnt4 = float4{0.f, 0.f, 0.f, 0.f};
                if constexpr (F32L2) cnt4 = *(const float4*)(cn_lds + t * 16 + 4 * g);
                const float cnv[4] = {cnt4.x, cnt4.y, cnt4.z, cnt4.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if constexpr (F32L2) {
                        // D = fl(fl(|q|^2 + |c|^2) - 2 s) against eps^2 -+ band
                        const float D2 = (eq + cnv[i]) - 2.0f * acc[t][i];
                        const double band = a.kappa * ((double)eq + (double)cnv[i]) + a.eps0;
                        const double dv = (double)D2;
                        hr |= (dv < a.eps2 - band ? 1u : 0u) << i;
                        nr |= (!(dv > a.eps2 + band) ? 1u : 0u) << i;
                    } else {
                        hr |= (acc[t][i] > th_h ? 1u : 0u) << i;
                        nr |= (acc[t][i] >= th_n ? 1u : 0u) << i;
                    }
                }
                const int base = (t & 1) * 16 + 4 * g;
                const uint32_t f = (Fw[t >> 1] >> base) & 0xFu;
                const uint32_t hn = hr & f;
                unc |= ((nr ^ hr) & f) << (4 * t);
                hw[t >> 1] |= hn << base;
                colacc[t] += spread4(hn);
            }
#pragma unroll
            for (int x = 0; x < EW; ++x) {
                hw[x] |= xchg<16>(hw[x], lane);
                hw[x] |= xchg<32>(hw[x], lane);
            }
            uint32_t mine = hw[0];
#pragma unroll
            for (int x = 1; x < EW; ++x) mine = g == x ? hw[x] : mine;
            if (valid && g < a.W) __builtin_nontemporal_store(mine, &a.E[(size_t)ent * a.W + g]);
            if (__any(unc != 0u)) {
                // one queue reservation per wave: lane offsets by a prefix sum
                unsigned long long pos = wave_reserve(a.qcount, __popc(unc), lane);
                while (unc) {
                    const int bp = __ffs(unc) - 1;
                    unc &= unc - 1;
                    const int col = (bp >> 2) * 16 + 4 * g + (bp & 3);
                    if (pos < (unsigned long long)a.qcap) a.queue[pos] = queue_entry(ent, l, col);
                    ++pos;
                }
            }
            if (!has_next) break;
            ent = entn; valid = validn; full = fulln;
            entn = ent2; validn = valid2; fulln = full2;
            b += WAVES;
        }
    }
    // column hit counts: bytes i of colacc[t] count candidate t*16 + 4g + i;
    // sum the 16 query lanes in 16-bit fields
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        uint32_t lo = colacc[t] & 0x00FF00FFu, hi = (colacc[t] >> 8) & 0x00FF00FFu;
        lo += xchg<1>(lo, lane); hi += xchg<1>(hi, lane);
        lo += xchg<2>(lo, lane); hi += xchg<2>(hi, lane);
        lo += xchg<4>(lo, lane); hi += xchg<4>(hi, lane);
        lo += xchg<8>(lo, lane); hi += xchg<8>(hi, lane);
        if (r16 == 0) {
            const int c0 = t * 16 + 4 * g;
            const int v[4] = {(int)(lo & 0xFFFFu), (int)(hi & 0xFFFFu), (int)(lo >> 16), (int)(hi >> 16)};
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (c0 + i < M && v[i]) atomicAdd(&col_lds[c0 + i], v[i]);
        }
    }
    __syncthreads();
    for (int i = tid; i < M; i += WAVES * 64) {
        const int v = col_lds[i];
        if (v) atomicAdd(&a.colsum[(size_t)l * M + i], v);
    }
}

template <int NT, int NS, bool EXACT, int PADB, bool F32L2 = false>
static int launch_pipe_p(const DistArgs& a, int ntiles, hipStream_t st) {
    constexpr int WAVES = 8;
    const size_t sm = (((size_t)a.M * (NS * 64 + PADB) + (size_t)a.M * 8 + 16 + 15) & ~(size_t)15) +
                      (F32L2 ? (size_t)NT * 16 * 4 : 0);
    if (sm > (size_t)kLdsBytes) return -2;
    if (F32L2 ? !a.rnorm2 : !a.rowerr) return -3;
    hipError_t e = hipFuncSetAttribute((const void*)k_dist_pipe<NT, WAVES, NS, EXACT, PADB, F32L2>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL((k_dist_pipe<NT, WAVES, NS, EXACT, PADB, F32L2>), dim3(ntiles), dim3(WAVES * 64), sm, st, a);
    return (int)hipGetLastError();
}

// L2 on FP32 rows in one LDS tile (rows of up to 256 dims at M <= 128):
// NS 64-byte k-steps of 16 floats
template <int NT>
static int launch_pipe_l2(const DistArgs& a, int ntiles, hipStream_t st) {
    const int nsteps = a.kc / 16;
    if (a.nchunk != 1) return -4;
    if (nsteps == 4) return launch_pipe_p<NT, 4, true, 16, true>(a, ntiles, st);
    if (nsteps <= 8) return launch_pipe_p<NT, 8, false, 16, true>(a, ntiles, st);
    if (nsteps <= 16) return launch_pipe_p<NT, 16, false, 16, true>(a, ntiles, st);
    return -4;  // (25 steps spill registers: the k-chunked tiles take those rows)
}

template <int NT, int NS, bool EXACT>
static int launch_pipe(const DistArgs& a, int ntiles, hipStream_t st) {
    return launch_pipe_p<NT, NS, EXACT, 16>(a, ntiles, st);
}

// (staging one 16-byte row piece per thread and step measured faster for these
// tiles than deeper batches: 271 vs 310 ms for the FP32 8M x 784 tiles)
template <bool BF, int NT, int QT, int PF, int BATCH = 1, bool L2X = false>
static int launch_mfma_v(const DistArgs& a, int ntiles, hipStream_t st) {
    constexpr int WAVES = 8;
    const size_t sm = (size_t)a.M * ((size_t)a.kc * (BF ? 2 : 4) + 16) + (size_t)a.M * (L2X ? 12 : 8);
    if (sm > (size_t)kLdsBytes) return -2;
    if (L2X && !a.rnorm2) return -3;
    hipError_t e = hipFuncSetAttribute((const void*)k_dist_mfma<BF, QT, NT, WAVES, PF, BATCH, L2X>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL((k_dist_mfma<BF, QT, NT, WAVES, PF, BATCH, L2X>), dim3(ntiles), dim3(WAVES * 64), sm, st, a);
    return (int)hipGetLastError();
}

template <bool BF, int NT>
static int launch_mfma_nt(const DistArgs& a, int ntiles, hipStream_t st) {
    if constexpr (BF) {
        // rows that fit one LDS tile: the pipelined kernel; longer rows:
        // k-chunked tiles
        const int nsteps = a.kc / 32;
        if (a.nchunk == 1 && nsteps == 25) return launch_pipe<NT, 25, true>(a, ntiles, st);
        if (a.nchunk == 1 && nsteps <= 8) return launch_pipe<NT, 8, false>(a, ntiles, st);
        if (a.nchunk == 1 && nsteps <= 25) return launch_pipe<NT, 25, false>(a, ntiles, st);
        return launch_mfma_v<BF, NT, 2, 5>(a, ntiles, st);
    } else {
        if (a.rnorm2) {
            // short rows: the pipelined tiles (the query lane owns its norm,
            // masks and entry); longer rows: k-chunked tiles
            const int r = std::getenv("SDB_L2_CHUNKED") ? -4 : launch_pipe_l2<NT>(a, ntiles, st);
            if (r != -4) return r;
            return launch_mfma_v<BF, NT, 2, 1, 1, true>(a, ntiles, st);
        }
        return launch_mfma_v<BF, NT, 2, 1>(a, ntiles, st);
    }
}

template <bool BF>
static int launch_mfma_bf(const DistArgs& a, int ntiles, hipStream_t st) {
    const int NT = (a.M + 15) / 16;
    switch (NT) {
        case 1: return launch_mfma_nt<BF, 1>(a, ntiles, st);
        case 2: return launch_mfma_nt<BF, 2>(a, ntiles, st);
        case 3: return launch_mfma_nt<BF, 3>(a, ntiles, st);
        case 4: return launch_mfma_nt<BF, 4>(a, ntiles, st);
        case 5: return launch_mfma_nt<BF, 5>(a, ntiles, st);
        case 6: return launch_mfma_nt<BF, 6>(a, ntiles, st);
        case 7: return launch_mfma_nt<BF, 7>(a, ntiles, st);
        case 8: return launch_mfma_nt<BF, 8>(a, ntiles, st);
        default: return -3;
    }
}

int launch_dist_mfma(const DistArgs& a, int precision, int ntiles, hipStream_t st) {
    if (ntiles <= 0) return 0;
    return precision == kBF16 ? launch_mfma_bf<true>(a, ntiles, st)
                              : launch_mfma_bf<false>(a, ntiles, st);
}

// ---------------------------------------------------------------------------
// VALU tiles for every metric (src/Utilities.h:212-287).  Candidate rows (fp32)
// in LDS with an odd row stride; each wave takes QV queries at a time, lane j
// owns candidates j, j + 64, ...; per-lane sums run sequentially over k.
template <int WAVES, int QV, int CM>
__global__ __launch_bounds__(WAVES * 64) void k_dist_valu(DistArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int M = a.M;
    const int RS = a.kc + 1;
    float* cand = (float*)smem;
    float* qbuf = cand + (size_t)M * RS;  // WAVES x QV x kc
    int* col_lds = (int*)(qbuf + (size_t)WAVES * QV * a.kc);
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const int l = a.tile_list[blockIdx.x];
    const int beg = a.tile_beg[blockIdx.x], end = a.tile_end[blockIdx.x];
    const int* Lm = a.lists + (size_t)l * M;
    const int metric = a.metric;
    const double lo = a.lo, hi = a.hi;
    float* qb = qbuf + (size_t)w * QV * a.kc;

    for (int i = tid; i < M; i += WAVES * 64) col_lds[i] = 0;
    const int rows_per_set = WAVES * QV;
    const int nsets = (end - beg + rows_per_set - 1) / rows_per_set;
    for (int set = 0; set < nsets; ++set) {
        uint32_t ent[QV];
        long long pa[QV];
        bool valid[QV], ffl[QV];
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            int rr = beg + set * rows_per_set + w * QV + q;
            valid[q] = rr < end;
            const uint32_t gq = a.grp[valid[q] ? rr : beg];
            ent[q] = ent_id(gq);
            ffl[q] = ent_full(gq);
            pa[q] = (long long)(ent[q] / (uint32_t)a.S);
        }
        float acc[QV][CM];
#pragma unroll
        for (int q = 0; q < QV; ++q)
#pragma unroll
            for (int c = 0; c < CM; ++c) acc[q][c] = 0.0f;
        for (int ch = 0; ch < a.nchunk; ++ch) {
            const int k0 = ch * a.kc;
            const int kn = min(a.kc, a.d - k0);
            if (a.nchunk > 1 || set == 0) {
                __syncthreads();
                for (int v = tid; v < M * kn; v += WAVES * 64) {
                    int row = v / kn, k = v - row * kn;
                    cand[(size_t)row * RS + k] = a.X32[(size_t)Lm[row] * a.ldx + k0 + k];
                }
            }
#pragma unroll
            for (int q = 0; q < QV; ++q)
                for (int k = lane; k < kn; k += 64) qb[q * a.kc + k] = a.X32[(size_t)pa[q] * a.ldx + k0 + k];
            __syncthreads();
#pragma unroll
            for (int c = 0; c < CM; ++c) {
                int ci = lane + 64 * c;
                int row = ci < M ? ci : M - 1;
                const float* cr = cand + (size_t)row * RS;
                for (int k = 0; k < kn; ++k) {
                    const float y = cr[k];
#pragma unroll
                    for (int q = 0; q < QV; ++q) {
                        const float x = qb[q * a.kc + k];
                        if (metric == kL2) { float t = x - y; acc[q][c] = acc[q][c] + t * t; }
                        else if (metric == kL1) { acc[q][c] = acc[q][c] + fabsf(x - y); }
                        else if (metric == kChi2) { acc[q][c] = acc[q][c] + chi2_term(x, y); }
                        else if (metric == kJS) {
                            const float xe = x + 1e-6f, ye = y + 1e-6f, s = xe + ye;
                            acc[q][c] = acc[q][c] + xe * (logf(s / xe) * 1.4426950408889634f) + ye * (logf(s / ye) * 1.4426950408889634f);
                        } else { acc[q][c] = acc[q][c] + x * y; }
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            uint32_t Fw[2 * CM];
#pragma unroll
            for (int x = 0; x < 2 * CM; ++x)
                Fw[x] = (valid[q] && x < a.W) ? (ffl[q] ? fmask_full_word(x, M) : a.F[(size_t)ent[q] * a.W + x]) : 0u;
#pragma unroll
            for (int c = 0; c < CM; ++c) {
                const int ci = lane + 64 * c;
                float dist;
                const float s = acc[q][c];
                if (metric == kL2) dist = sqrtf(s);
                else if (metric == kL1) dist = s;
                else if (metric == kChi2) dist = 1.0f - 2.0f * s;
                else if (metric == kJS) dist = 1.0f - s * 0.5f;
                else dist = 1.0f - s;
                const uint32_t fw = Fw[2 * c + (lane >> 5)];
                const bool act = valid[q] && ci < M && ((fw >> (lane & 31)) & 1u);
                const double dv = (double)dist;
                const bool hit = act && dv < lo;
                const bool unc = act && !(dv < lo) && !(dv > hi);
                const unsigned long long bh = __ballot(hit);
                if (valid[q] && lane == 0) {
                    if (2 * c < a.W) a.E[(size_t)ent[q] * a.W + 2 * c] = (uint32_t)bh;
                    if (2 * c + 1 < a.W) a.E[(size_t)ent[q] * a.W + 2 * c + 1] = (uint32_t)(bh >> 32);
                }
                if (hit) atomicAdd(&col_lds[ci], 1);
                if (unc) {
                    const unsigned long long pos = atomicAdd(a.qcount, 1ull);
                    if (pos < (unsigned long long)a.qcap) a.queue[pos] = queue_entry(ent[q], l, ci);
                }
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < M; i += WAVES * 64) {
        int v = col_lds[i];
        if (v) atomicAdd(&a.colsum[(size_t)l * M + i], v);
    }
}

// Rows that fit one LDS chunk (d <= 256): the candidate rows are staged once
// per tile; every wave then streams its own query sets through a private LDS
// slice, the next set's rows already loading into registers while the current
// set is summed, with no workgroup barrier between sets.  The metric is a
// template parameter; rows are zero padded to d4 = d rounded up to 4 (zeros
// add nothing for Cosine, L1, L2 and Chi2; JS sums exactly d terms).
template <int METRIC>
__device__ __forceinline__ float valu_term(float x, float y) {
    if constexpr (METRIC == kL2) { const float t = x - y; return t * t; }
    else if constexpr (METRIC == kL1) return fabsf(x - y);
    else if constexpr (METRIC == kChi2) return chi2_term(x, y);
    else if constexpr (METRIC == kJS) {
        const float xe = x + 1e-6f, ye = y + 1e-6f, s = xe + ye;
        return xe * (logf(s / xe) * 1.4426950408889634f) + ye * (logf(s / ye) * 1.4426950408889634f);
    } else return x * y;
}

template <int WAVES, int QV, int CM, int METRIC>
__global__ __launch_bounds__(WAVES * 64) void k_dist_valu1(DistArgs a) {
    constexpr int KPL = 4;  // query values per lane: d <= 256
    extern __shared__ __align__(16) unsigned char smem[];
    const int M = a.M, d = a.d;
    const int d4 = (d + 3) & ~3;
    const int kend = METRIC == kJS ? d : d4;
    const int RS = d4 + 1;
    float* cand = (float*)smem;
    float* qbuf = cand + (size_t)M * RS;
    int* col_lds = (int*)(qbuf + (size_t)WAVES * QV * d4);
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const int l = a.tile_list[blockIdx.x];
    const int beg = a.tile_beg[blockIdx.x], end = a.tile_end[blockIdx.x];
    const int* Lm = a.lists + (size_t)l * M;
    const double lo = a.lo, hi = a.hi;
    float* qb = qbuf + (size_t)w * QV * d4;

    for (int i = tid; i < M; i += WAVES * 64) col_lds[i] = 0;
    for (int v = tid; v < M * d4; v += WAVES * 64) {
        const int row = v / d4, k = v - row * d4;
        cand[(size_t)row * RS + k] = k < d ? a.X32[(size_t)Lm[row] * a.ldx + k] : 0.0f;
    }
    __syncthreads();

    const int per_set = WAVES * QV;
    const int nsets = (end - beg + per_set - 1) / per_set;
    float pre[QV][KPL];
    uint32_t entn[QV];
    bool validn[QV], ffn[QV];
    auto fetch = [&](int set) {
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            const int rr = beg + set * per_set + w * QV + q;
            validn[q] = rr < end;
            const uint32_t gq = a.grp[validn[q] ? rr : beg];
            entn[q] = ent_id(gq);
            ffn[q] = ent_full(gq);
            const long long pa = (long long)(entn[q] / (uint32_t)a.S);
            const float* xr = a.X32 + (size_t)pa * a.ldx;
#pragma unroll
            for (int t = 0; t < KPL; ++t) {
                const int k = lane + 64 * t;
                pre[q][t] = k < d ? xr[k] : 0.0f;
            }
        }
    };
    if (nsets > 0) fetch(0);
    for (int set = 0; set < nsets; ++set) {
        uint32_t ent[QV];
        bool valid[QV], ffl[QV];
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            ent[q] = entn[q];
            valid[q] = validn[q];
            ffl[q] = ffn[q];
#pragma unroll
            for (int t = 0; t < KPL; ++t) {
                const int k = lane + 64 * t;
                if (k < d4) qb[q * d4 + k] = pre[q][t];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (set + 1 < nsets) fetch(set + 1);
        float acc[QV][CM];
#pragma unroll
        for (int q = 0; q < QV; ++q)
#pragma unroll
            for (int c = 0; c < CM; ++c) acc[q][c] = 0.0f;
#pragma unroll
        for (int c = 0; c < CM; ++c) {
            const int ci = lane + 64 * c;
            const float* cr = cand + (size_t)(ci < M ? ci : M - 1) * RS;
            for (int k = 0; k < kend; ++k) {
                const float y = cr[k];
#pragma unroll
                for (int q = 0; q < QV; ++q) acc[q][c] = acc[q][c] + valu_term<METRIC>(qb[q * d4 + k], y);
            }
        }
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            uint32_t Fw[2 * CM];
#pragma unroll
            for (int x = 0; x < 2 * CM; ++x)
                Fw[x] = (valid[q] && x < a.W) ? (ffl[q] ? fmask_full_word(x, M) : a.F[(size_t)ent[q] * a.W + x]) : 0u;
#pragma unroll
            for (int c = 0; c < CM; ++c) {
                const int ci = lane + 64 * c;
                const float s = acc[q][c];
                float dist;
                if constexpr (METRIC == kL2) dist = sqrtf(s);
                else if constexpr (METRIC == kL1) dist = s;
                else if constexpr (METRIC == kChi2) dist = 1.0f - 2.0f * s;
                else if constexpr (METRIC == kJS) dist = 1.0f - s * 0.5f;
                else dist = 1.0f - s;
                const uint32_t fw = Fw[2 * c + (lane >> 5)];
                const bool act = valid[q] && ci < M && ((fw >> (lane & 31)) & 1u);
                const double dv = (double)dist;
                const bool hit = act && dv < lo;
                const bool unc = act && !(dv < lo) && !(dv > hi);
                const unsigned long long bh = __ballot(hit);
                if (valid[q] && lane == 0) {
                    if (2 * c < a.W) a.E[(size_t)ent[q] * a.W + 2 * c] = (uint32_t)bh;
                    if (2 * c + 1 < a.W) a.E[(size_t)ent[q] * a.W + 2 * c + 1] = (uint32_t)(bh >> 32);
                }
                if (hit) atomicAdd(&col_lds[ci], 1);
                if (unc) {
                    const unsigned long long pos = atomicAdd(a.qcount, 1ull);
                    if (pos < (unsigned long long)a.qcap) a.queue[pos] = queue_entry(ent[q], l, ci);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    for (int i = tid; i < M; i += WAVES * 64) {
        const int v = col_lds[i];
        if (v) atomicAdd(&a.colsum[(size_t)l * M + i], v);
    }
}

// Rows of 257 to 1024 dims: the candidates of a tile are split into parts of
// at most 64 (grid y), one candidate per lane, whose rows fill the LDS; each
// wave keeps QV query rows in registers (element k in lane k % 64, register
// k / 64, the next set already loading) and broadcasts element k with a
// readlane.  Parts share the hit words of an entry, so E is cleared first and
// the hits are OR-ed in.
template <int WAVES, int QV, int KPL, int METRIC>
__global__ __launch_bounds__(WAVES * 64) void k_dist_valu2(DistArgs a, int Mc) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int M = a.M, d = a.d;
    const int RS = d | 1;
    const int c0 = blockIdx.y * Mc;
    const int mc = min(Mc, M - c0);
    float* cand = (float*)smem;
    int* col_lds = (int*)(cand + (size_t)Mc * RS);
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const int l = a.tile_list[blockIdx.x];
    const int beg = a.tile_beg[blockIdx.x], end = a.tile_end[blockIdx.x];
    const int* Lm = a.lists + (size_t)l * M + c0;
    const double lo = a.lo, hi = a.hi;

    for (int i = tid; i < mc; i += WAVES * 64) col_lds[i] = 0;
    for (int v = tid; v < mc * d; v += WAVES * 64) {
        const int row = v / d, k = v - row * d;
        cand[(size_t)row * RS + k] = a.X32[(size_t)Lm[row] * a.ldx + k];
    }
    __syncthreads();
    const bool colok = lane < mc;
    const int ci = c0 + lane;
    const float* cr = cand + (size_t)(colok ? lane : 0) * RS;

    const int per_set = WAVES * QV;
    const int nsets = (end - beg + per_set - 1) / per_set;
    float pre[QV][KPL];
    uint32_t entn[QV];
    bool validn[QV], ffn[QV];
    auto fetch = [&](int set) {
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            const int rr = beg + set * per_set + w * QV + q;
            validn[q] = rr < end;
            const uint32_t gq = a.grp[validn[q] ? rr : beg];
            entn[q] = ent_id(gq);
            ffn[q] = ent_full(gq);
            const long long pa = (long long)(entn[q] / (uint32_t)a.S);
            const float* xr = a.X32 + (size_t)pa * a.ldx;
#pragma unroll
            for (int t = 0; t < KPL; ++t) {
                const int k = lane + 64 * t;
                pre[q][t] = k < d ? xr[k] : 0.0f;
            }
        }
    };
    if (nsets > 0) fetch(0);
    for (int set = 0; set < nsets; ++set) {
        float cur[QV][KPL];
        uint32_t ent[QV];
        bool valid[QV], ffl[QV];
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            ent[q] = entn[q];
            valid[q] = validn[q];
            ffl[q] = ffn[q];
#pragma unroll
            for (int t = 0; t < KPL; ++t) cur[q][t] = pre[q][t];
        }
        if (set + 1 < nsets) fetch(set + 1);
        float acc[QV];
#pragma unroll
        for (int q = 0; q < QV; ++q) acc[q] = 0.0f;
#pragma unroll
        for (int t = 0; t < KPL; ++t) {
            const int kn = min(64, d - 64 * t);
            for (int kk = 0; kk < kn; ++kk) {
                const float y = cr[64 * t + kk];
#pragma unroll
                for (int q = 0; q < QV; ++q) {
                    const float x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur[q][t]), kk));
                    acc[q] = acc[q] + valu_term<METRIC>(x, y);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < QV; ++q) {
            const float s = acc[q];
            float dist;
            if constexpr (METRIC == kL2) dist = sqrtf(s);
            else if constexpr (METRIC == kL1) dist = s;
            else if constexpr (METRIC == kChi2) dist = 1.0f - 2.0f * s;
            else if constexpr (METRIC == kJS) dist = 1.0f - s * 0.5f;
            else dist = 1.0f - s;
            const uint32_t fw = (valid[q] && colok)
                                    ? (ffl[q] ? fmask_full_word(ci >> 5, M) : a.F[(size_t)ent[q] * a.W + (ci >> 5)])
                                    : 0u;
            const bool act = valid[q] && colok && ((fw >> (ci & 31)) & 1u);
            const double dv = (double)dist;
            const bool hit = act && dv < lo;
            const bool unc = act && !(dv < lo) && !(dv > hi);
            const unsigned long long bh = __ballot(hit);
            if (bh && lane == 0) {
                // bit (c - c0) of bh is candidate c; OR it into word c >> 5
                for (int wi = c0 >> 5; wi <= (c0 + mc - 1) >> 5; ++wi) {
                    const int sh = 32 * wi - c0;
                    const uint32_t bits = sh >= 0 ? (uint32_t)(bh >> sh) : (uint32_t)(bh << -sh);
                    if (bits) atomicOr(&a.E[(size_t)ent[q] * a.W + wi], bits);
                }
            }
            if (hit) atomicAdd(&col_lds[lane], 1);
            if (unc) {
                const unsigned long long pos = atomicAdd(a.qcount, 1ull);
                if (pos < (unsigned long long)a.qcap) a.queue[pos] = queue_entry(ent[q], l, ci);
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < mc; i += WAVES * 64) {
        const int v = col_lds[i];
        if (v) atomicAdd(&a.colsum[(size_t)l * M + c0 + i], v);
    }
}

template <int KPL, int METRIC>
static int launch_valu2(const DistArgs& a, int ntiles, hipStream_t st) {
    constexpr int WAVES = 8, QV = 4;
    const int RS = a.d | 1;
    const int cap = std::min(64, (int)((kLdsBytes - 64 * 4) / (4 * RS)));
    if (cap < 1) return -2;
    const int parts = (a.M + cap - 1) / cap;
    const int Mc = (a.M + parts - 1) / parts;
    const size_t sm = (size_t)Mc * RS * 4 + (size_t)Mc * 4;
    if (sm > (size_t)kLdsBytes) return -2;
    hipError_t e = hipFuncSetAttribute((const void*)k_dist_valu2<WAVES, QV, KPL, METRIC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return (int)e;
    // the parts OR their hits into cleared words
    e = hipMemsetAsync(a.E, 0, (size_t)a.n_ent * a.W * 4, st);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL((k_dist_valu2<WAVES, QV, KPL, METRIC>), dim3(ntiles, parts), dim3(WAVES * 64), sm, st, a, Mc);
    return (int)hipGetLastError();
}

template <int KPL>
static int launch_valu2_m(const DistArgs& a, int ntiles, hipStream_t st) {
    switch (a.metric) {
        case kL2: return launch_valu2<KPL, kL2>(a, ntiles, st);
        case kL1: return launch_valu2<KPL, kL1>(a, ntiles, st);
        case kChi2: return launch_valu2<KPL, kChi2>(a, ntiles, st);
        case kJS: return launch_valu2<KPL, kJS>(a, ntiles, st);
        default: return launch_valu2<KPL, kCosine>(a, ntiles, st);
    }
}

template <int CM, int METRIC>
static int launch_valu1(const DistArgs& a, int ntiles, hipStream_t st) {
    constexpr int WAVES = 8, QV = 8;
    const size_t d4 = (size_t)((a.d + 3) & ~3);
    const size_t sm = ((size_t)a.M * (d4 + 1) + (size_t)WAVES * QV * d4) * 4 + (size_t)a.M * 4;
    if (sm > (size_t)kLdsBytes) return -2;
    hipError_t e = hipFuncSetAttribute((const void*)k_dist_valu1<WAVES, QV, CM, METRIC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL((k_dist_valu1<WAVES, QV, CM, METRIC>), dim3(ntiles), dim3(WAVES * 64), sm, st, a);
    return (int)hipGetLastError();
}

template <int CM>
static int launch_valu1_m(const DistArgs& a, int ntiles, hipStream_t st) {
    switch (a.metric) {
        case kL2: return launch_valu1<CM, kL2>(a, ntiles, st);
        case kL1: return launch_valu1<CM, kL1>(a, ntiles, st);
        case kChi2: return launch_valu1<CM, kChi2>(a, ntiles, st);
        case kJS: return launch_valu1<CM, kJS>(a, ntiles, st);
        default: return launch_valu1<CM, kCosine>(a, ntiles, st);
    }
}

template <int CM>
static int launch_valu_cm(const DistArgs& a, int ntiles, hipStream_t st) {
    constexpr int WAVES = 8, QV = 4;
    const size_t sm = ((size_t)a.M * (a.kc + 1) + (size_t)WAVES * QV * a.kc) * 4 + (size_t)a.M * 4;
    if (sm > (size_t)kLdsBytes) return -2;
    hipError_t e = hipFuncSetAttribute((const void*)k_dist_valu<WAVES, QV, CM>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL((k_dist_valu<WAVES, QV, CM>), dim3(ntiles), dim3(WAVES * 64), sm, st, a);
    return (int)hipGetLastError();
}

int launch_dist_valu(const DistArgs& a, int ntiles, hipStream_t st) {
    if (ntiles <= 0) return 0;
    const int CM = (a.M + 63) / 64;
    if (a.d <= 256 && CM <= 4) {
        int r = -2;
        switch (CM) {
            case 1: r = launch_valu1_m<1>(a, ntiles, st); break;
            case 2: r = launch_valu1_m<2>(a, ntiles, st); break;
            case 3: r = launch_valu1_m<3>(a, ntiles, st); break;
            case 4: r = launch_valu1_m<4>(a, ntiles, st); break;
        }
        if (r != -2) return r;  // -2: the tile does not fit LDS, take the chunked tiles
    } else if (a.d <= 1024 && a.n_ent > 0) {
        int r = a.d <= 512 ? launch_valu2_m<8>(a, ntiles, st)
              : a.d <= 768 ? launch_valu2_m<12>(a, ntiles, st)
                           : launch_valu2_m<16>(a, ntiles, st);
        if (r != -2) return r;
    }
    switch (CM) {
        case 1: return launch_valu_cm<1>(a, ntiles, st);
        case 2: return launch_valu_cm<2>(a, ntiles, st);
        case 3: return launch_valu_cm<3>(a, ntiles, st);
        case 4: return launch_valu_cm<4>(a, ntiles, st);
        default: return -3;
    }
}

// ---------------------------------------------------------------------------
// Boundary recheck, one wavefront per queued pair.  Sums run per lane over
// k = lane, lane + 64, ... and combine through an xor butterfly (offsets
// 32..1).  The BF16 paths (Cosine, and the Chi2 prefilter) first re-decide
// with the FP32 distance when it lies outside the FP32 band; everything else is
// decided with the FP64 value.
// stats[0]: FP32 rechecks, stats[1]: FP64 rechecks, stats[2]: hits added,
// stats[4]: FP64 distances within 1e-6 of eps.
__device__ __forceinline__ float dot32_wave(const float* x, const float* y, int d, int lane) {
    float s = wave_row_fold<8>(x, y, d, lane, 0.0f, [](float t, float a, float b) { return fmaf(a, b, t); });
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s = s + __shfl_xor(s, m, kWave);
    return s;
}

// sum of x y / (x + y) over the entries positive in both rows, FP32
__device__ __forceinline__ float chi2sum32_wave(const float* x, const float* y, int d, int lane) {
    float s = wave_row_fold<8>(x, y, d, lane, 0.0f, [](float t, float a, float b) {
        return (a > 0.0f && b > 0.0f) ? t + a * b / (a + b) : t;
    });
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s = s + __shfl_xor(s, m, kWave);
    return s;
}

__global__ __launch_bounds__(256) void k_recheck(DistArgs a, const int* topK, int K, int P,
                                                 int precision, double band32, int n_entries,
                                                 unsigned long long* stats) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    // the wave's counts, added to stats once at the end (one atomic per
    // counter and wave rather than per pair)
    unsigned long long c32 = 0, c64 = 0, chit = 0, cnear = 0;
    for (long long e = wv; e < n_entries; e += nwv) {
        const unsigned long long qe = a.queue[e];
        const uint32_t ent = (uint32_t)(qe >> 32);
        const int l = (int)((qe >> 8) & 0xffffffu);
        const int col = (int)(qe & 0xffu);
        const long long pa = ent / (uint32_t)a.S;
        const int c = a.lists[(size_t)l * a.M + col];
        const float* x = a.X32 + (size_t)pa * a.ldx;
        const float* y = a.X32 + (size_t)c * a.ldx;
        bool hit;
        bool decided = false;
        if (precision == kBF16 && (a.metric == kCosine || a.metric == kChi2)) {
            const float d32 = a.metric == kCosine ? 1.0f - dot32_wave(x, y, a.d, lane)
                                                  : 1.0f - 2.0f * chi2sum32_wave(x, y, a.d, lane);
            const double dv = (double)d32;
            ++c32;
            if (dv < (double)a.eps - band32) { hit = true; decided = true; }
            else if (dv > (double)a.eps + band32) { hit = false; decided = true; }
        }
        if (!decided) {
            const double d64 = dist64_wave(x, y, a.d, a.metric, lane);
            ++c64;
            // pairs decided in FP32 (tiles or recheck) lie at least a band from eps,
            // which exceeds 1e-6 except for L1/L2 at small eps * d
            if (fabs(d64 - (double)a.eps) < 1e-6) ++cnear;
            hit = d64 <= (double)a.eps;
        }
        if (hit) {
            ++chit;
            if (lane == 0) {
                atomicOr(&a.E[(size_t)ent * a.W + (col >> 5)], 1u << (col & 31));
                atomicAdd(&a.colsum[(size_t)l * a.M + col], 1);
            }
        }
    }
    if (lane == 0) {
        if (c32) atomicAdd(&stats[0], c32);
        if (c64) atomicAdd(&stats[1], c64);
        if (chit) atomicAdd(&stats[2], chit);
        if (cnear) atomicAdd(&stats[4], cnear);
    }
}

// ---------------------------------------------------------------------------
// Accuracy sample: candidate pairs drawn from the index (point a uniform,
// slot s uniform, list position uniform; the point itself skipped), one
// wavefront per pair.  fast: the quantity the tiles test, in the tiles'
// precision (BF16 operands or FP32, FP32 sums over lanes k = lane, lane + 64,
// ...); exact: the same quantity in FP64 from the FP32 rows.  The quantity is
// the distance, except for Chi2 on BF16 rows, where it is 1 - sum sqrt(x y).
__device__ __forceinline__ uint64_t acc_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_pair_sample(DistArgs a, const int* topK, int K, int P, long long n,
                                                     int precision, uint64_t seed, long long n_samples,
                                                     int* pa, int* pb, float* fast, double* exact) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    const bool sq = a.metric == kChi2 && precision == kBF16;
    for (long long e = wv; e < n_samples; e += nwv) {
        const uint64_t h = acc_mix(seed ^ acc_mix((uint64_t)e));
        const long long p = (long long)(h % (uint64_t)n);
        const int s = (int)((h >> 40) % (uint64_t)(2 * K));
        int i = (int)(acc_mix(h) % (uint64_t)a.M);
        const int l = (s & 1) ? P + topK[p * 2 * K + K + (s >> 1)] : topK[p * 2 * K + (s >> 1)];
        int c = a.lists[(size_t)l * a.M + i];
        if (c == p) c = a.lists[(size_t)l * a.M + (i + 1) % a.M];
        const float* x = a.X32 + (size_t)p * a.ldx;
        const float* y = a.X32 + (size_t)c * a.ldx;
        double ex;
        float fs;
        if (sq) {
            double t = 0.0;
            for (int k = lane; k < a.d; k += 64) t = t + sqrt((double)x[k]) * sqrt((double)y[k]);
            ex = 1.0 - wave_sum_f64_tree(t);
        } else {
            ex = dist64_wave(x, y, a.d, a.metric, lane);
        }
        if (precision == kBF16) {
            const uint16_t* xb = (const uint16_t*)a.X + (size_t)p * a.ldx;
            const uint16_t* yb = (const uint16_t*)a.X + (size_t)c * a.ldx;
            float t = 0.0f;
            for (int k = lane; k < a.d; k += 64)
                t = fmaf(__uint_as_float((uint32_t)xb[k] << 16), __uint_as_float((uint32_t)yb[k] << 16), t);
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) t = t + __shfl_xor(t, m, kWave);
            fs = 1.0f - t;
        } else if (a.metric == kCosine) {
            fs = 1.0f - dot32_wave(x, y, a.d, lane);
        } else if (a.metric == kChi2) {
            fs = 1.0f - 2.0f * chi2sum32_wave(x, y, a.d, lane);
        } else if (a.metric == kL2 && a.rnorm2) {
            // the L2 MFMA tiles' arithmetic, D = fl(fl(|x|^2 + |y|^2) - 2 x.y)
            // with the tiles' squared norms (the dot's order emulated)
            const float D2 = (a.rnorm2[p] + a.rnorm2[c]) - 2.0f * dot32_wave(x, y, a.d, lane);
            fs = sqrtf(fmaxf(D2, 0.0f));
        } else if (a.metric == kL1 || a.metric == kL2) {
            float t = 0.0f;
            for (int k = lane; k < a.d; k += 64) {
                const float df = x[k] - y[k];
                t = a.metric == kL2 ? fmaf(df, df, t) : t + fabsf(df);
            }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) t = t + __shfl_xor(t, m, kWave);
            fs = a.metric == kL2 ? sqrtf(t) : t;
        } else {
            fs = (float)ex;  // JS: its FP32 tiles are not sampled here
        }
        if (lane == 0) {
            pa[e] = (int)p;
            pb[e] = c;
            fast[e] = fs;
            exact[e] = ex;
        }
    }
}

int launch_pair_sample(const DistArgs& a, const int* topK, int K, int P, long long n, int precision,
                       unsigned long long seed, long long n_samples, int* pa, int* pb, float* fast,
                       double* exact, hipStream_t st) {
    if (n_samples <= 0) return 0;
    long long blocks = (n_samples + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pair_sample, dim3((unsigned)blocks), dim3(256), 0, st, a, topK, K, P, n, precision,
                       (uint64_t)seed, n_samples, pa, pb, fast, exact);
    return (int)hipGetLastError();
}

// Squared norms of the rows for the L2 MFMA tiles: 16 lanes per row, each
// summing float4 pieces of the zero-padded row in FP64, a 16-lane xor tree,
// one rounding to FP32 (the order of an FP64 sum of d squares moves it by
// far less than that rounding); and their maximum (float bits, norms >= 0).
// One workgroup per 16 rows: every row's load is in flight at once.
__global__ __launch_bounds__(256) void k_row_norm2(const float* __restrict__ X, long long n, int ldx,
                                                   float* __restrict__ out, uint32_t* __restrict__ maxbits) {
    const int lane = lane_id(), sub = lane & 15;
    const long long p = (long long)blockIdx.x * 16 + (threadIdx.x >> 4);
    double s = 0.0;
    if (p < n) {
        const float4* r = (const float4*)(X + (size_t)p * ldx);
        for (int k = sub; k < ldx / 4; k += 16) {
            const float4 v = r[k];
            s = fma((double)v.x, (double)v.x, s);
            s = fma((double)v.y, (double)v.y, s);
            s = fma((double)v.z, (double)v.z, s);
            s = fma((double)v.w, (double)v.w, s);
        }
    }
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) s = s + __shfl_xor(s, m, kWave);
    const float f = p < n ? (float)s : 0.0f;
    if (sub == 0 && p < n) out[p] = f;
    float mx = f;
#pragma unroll
    for (int m = 32; m >= 16; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m, kWave));
    // one address for every wave: the atomic only when it raises the maximum
    if (lane == 0 && mx > 0.0f && __hip_atomic_load(maxbits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < __float_as_uint(mx))
        atomicMax(maxbits, __float_as_uint(mx));
}

int launch_row_norm2(const float* X, long long n, int d, int ldx, float* out, uint32_t* maxbits, hipStream_t st) {
    if (n <= 0) return 0;
    if (ldx % 4 != 0 || d > ldx) return -1;
    const long long blocks = (n + 15) / 16;
    if (blocks > 0x7fffffffll) return -3;
    hipLaunchKernelGGL(k_row_norm2, dim3((unsigned)blocks), dim3(256), 0, st, X, n, ldx, out, maxbits);
    return (int)hipGetLastError();
}

// Sort key of every queued pair: its candidate point (lists[l][col]).
__global__ void k_queue_keys(const unsigned long long* __restrict__ queue, int nq, const int* __restrict__ lists,
                             int M, uint32_t* __restrict__ keys) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nq) return;
    const unsigned long long qe = queue[e];
    const int l = (int)((qe >> 8) & 0xffffffu), col = (int)(qe & 0xffu);
    keys[e] = (uint32_t)lists[(size_t)l * M + col];
}

int launch_queue_keys(const unsigned long long* queue, int nq, const int* lists, int M, uint32_t* keys,
                      hipStream_t st) {
    if (nq <= 0) return 0;
    hipLaunchKernelGGL(k_queue_keys, dim3((nq + 255) / 256), dim3(256), 0, st, queue, nq, lists, M, keys);
    return (int)hipGetLastError();
}

int launch_recheck_k(const DistArgs& a, const int* topK, int K, int P, int precision,
                     double band32, int n_entries, unsigned long long* stats, hipStream_t st) {
    if (n_entries <= 0) return 0;
    // 4 waves per block, grid-stride; 16,384 blocks (64 k waves, 64 per CU)
    // keep every CU busy while the per-wave counter atomics, all on the same
    // four words, stay few
    long long blocks = ((long long)n_entries + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_recheck, dim3((unsigned)blocks), dim3(256), 0, st, a, topK, K, P, precision, band32, n_entries, stats);
    return (int)hipGetLastError();
}

}  // namespace sdb
