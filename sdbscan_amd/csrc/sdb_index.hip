// Index construction of sDbscan on gfx950: transform, kernel embedding,
// HD3 Walsh-Hadamard projection, per-point top-K and per-random-vector top-M.
//
// Reference: rp_parIndex (src/sDbscan.cpp:19-209), transformData
// (src/Utilities.h:289-335), embeddings (src/sDbscan.cpp:53-68,
// src/Utilities.h:132-208), FFHT (src/fht_avx.c:1013-1256).
#include <algorithm>

#include "sdb_kernels.hpp"

namespace sdb {

// ---------------------------------------------------------------------------
// transformData (src/Utilities.h:289-335).  One wavefront per point.  Cosine:
// x / ||x|| with ||x||^2 accumulated in FP64 per lane (stride 64) and combined
// by an xor butterfly; Chi2/JS: nonnegativity check and x / sum(x).  Writes
// the FP32 rows (row stride ldx, zero padded) and, when requested, the BF16
// rows the BF16 distance tiles read (Chi2: BF16 of the square roots, whose
// dot product bounds chi2 from below).
// With the BF16 rows it also writes each row's rounding bound rho
// (bf16_row_err) and their maximum.
// err bit 1: negative entry (Chi2/JS); err bit 2: zero point.
__global__ __launch_bounds__(256) void k_transform(const float* __restrict__ X, long long n, int d,
                                                   int ldx, int metric, float* __restrict__ Xn,
                                                   uint16_t* __restrict__ Xb, float* __restrict__ rowerr,
                                                   uint32_t* __restrict__ rowerr_max, int rho_pad,
                                                   int* __restrict__ err) {
    const int lane = lane_id();
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const long long nwaves = (long long)gridDim.x * blockDim.x / kWave;
    float rmax = 0.0f;
    for (long long p = wave; p < n; p += nwaves) {
        const float* x = X + p * d;
        float scale = 1.0f;
        bool scaled = false;
        if (metric == kCosine) {
            const double acc = wave_row_fold1<8>(x, d, lane, 0.0, [](double s, float f) {
                const double v = f;
                return s + v * v;
            });
            double t = wave_sum_f64_tree(acc);
            scale = (float)sqrt(t);
            scaled = true;
            if (!(scale > 0.0f) && lane == 0) atomicOr(err, 2);
        } else if (metric == kChi2 || metric == kJS) {
            int neg = 0;
            const double acc = wave_row_fold1<8>(x, d, lane, 0.0, [&neg](double s, float v) {
                neg |= v < 0.0f;
                return s + (double)v;
            });
            if (__any(neg) && lane == 0) atomicOr(err, 1);
            double t = wave_sum_f64_tree(acc);
            scale = (float)t;
            scaled = true;
            if (!(scale > 0.0f) && lane == 0) atomicOr(err, 2);
        }
        float* o = Xn + p * ldx;
        uint16_t* ob = Xb ? Xb + p * ldx : nullptr;
        float sq = 0.0f;
#pragma unroll 4
        for (int i = lane; i < ldx; i += kWave) {
            float v = 0.0f;
            if (i < d) v = scaled ? x[i] / scale : x[i];
            o[i] = v;
            if (ob) {
                const float sv = metric == kChi2 ? sqrtf(v) : v;
                const uint32_t h = f32_to_bf16_rne(sv);
                if (!rho_pad || i != ldx - 1) ob[i] = (uint16_t)h;
                const float df = __uint_as_float(h << 16) - sv;
                sq = fmaf(df, df, sq);
            }
        }
        if (ob) {
            const float rho = bf16_row_err(wave_sum_f32(sq), metric == kChi2);
            if (lane == 0) {
                rowerr[p] = rho;
                if (rho_pad) ob[ldx - 1] = (uint16_t)bf16_up(rho);
            }
            rmax = fmaxf(rmax, rho);
        }
    }
    if (Xb && lane == 0 && (int)__float_as_uint(rmax) > load_relaxed((const int*)rowerr_max))
        atomicMax(rowerr_max, __float_as_uint(rmax));
}

// One Walsh-Hadamard level across lanes (logical stride S < 64, mirror lane
// layout, lg = logical index of the lane): the lower element becomes a + b,
// the upper a - b, computed as fma(+-1, mine, partner) with an exact product,
// i.e. one rounding of partner +- mine.
template <int S, int VPL>
__device__ __forceinline__ void fwht_cross(float (&v)[VPL], int lg) {
    const float sg = (lg & S) ? -1.0f : 1.0f;
#pragma unroll
    for (int j = 0; j < VPL; ++j) {
        const float o = __uint_as_float(xchg_l<S>(__float_as_uint(v[j]), lg));
        v[j] = fmaf(sg, v[j], o);
    }
}

// Strides 16 and 32 on register pairs: a permlane swap of (x, y) leaves each
// lane holding one butterfly pair (of x in the lower row / half, of y in the
// upper one); the sums and differences are swapped back into x and y.  Two
// swaps, one add and one subtract per two registers.
template <int S, int VPL>
__device__ __forceinline__ void fwht_cross_pairs(float (&v)[VPL]) {
#pragma unroll
    for (int j = 0; j + 1 < VPL; j += 2) {
        float x, y;
        if constexpr (S == 16) {
            const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[j]), __float_as_uint(v[j + 1]), false, false);
            x = __uint_as_float(p[0]);
            y = __uint_as_float(p[1]);
        } else {
            const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[j]), __float_as_uint(v[j + 1]), false, false);
            x = __uint_as_float(p[0]);
            y = __uint_as_float(p[1]);
        }
        const float sm = x + y, df = x - y;
        if constexpr (S == 16) {
            const auto q = __builtin_amdgcn_permlane16_swap(__float_as_uint(sm), __float_as_uint(df), false, false);
            v[j] = __uint_as_float(q[0]);
            v[j + 1] = __uint_as_float(q[1]);
        } else {
            const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(sm), __float_as_uint(df), false, false);
            v[j] = __uint_as_float(q[0]);
            v[j + 1] = __uint_as_float(q[1]);
        }
    }
}

template <int VPL>
__device__ __forceinline__ void fwht_inreg(float (&v)[VPL], int js) {
#pragma unroll
    for (int j = 0; j < VPL; ++j) {
        if ((j & js) == 0) {
            float a = v[j], b = v[j + js];
            v[j] = a + b;
            v[j + js] = a - b;
        }
    }
}

// Walsh-Hadamard levels at logical strides 1, 2, 4, 8 (DPP partners in the
// mirror layout) with one v_fmac_f32_dpp per element: v += sg * v[partner],
// sg = -1 on the upper lanes of the level.  An upper lane then holds
// mine - partner = -(partner - mine), the negated level output with the same
// rounding (round to nearest even is symmetric), so after the four levels a
// lane holds its outputs times (-1)^popcount(lg & 15); hd3_rotate folds that
// pending sign into the next sign flip.  The partners of a level have equal
// pending signs (they differ only in the level's bit), so sg is the plain
// per-level sign.
// Hazards: a DPP read needs 2 wait states after a VALU write of its source
// and 5 after a VALU write of EXEC.  Only the stride-1 blocks can follow
// compiler code that wrote v (the sign flip), so they start with s_nop 4; in
// blocks of 8 the other levels read registers written at least 7 DPP
// instructions earlier.  The stride-8 blocks end with s_nop 4 for the
// permlane swaps that follow.  Single-register blocks (fewer than 8 registers)
// wait on both sides.
#define SDB_FMAC_DPP(i, ctrl) "v_fmac_f32_dpp %" #i ", %" #i ", %8 " ctrl " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
#define SDB_FMAC_DPP8(lead, ctrl, trail)                                                               \
    lead SDB_FMAC_DPP(0, ctrl) SDB_FMAC_DPP(1, ctrl) SDB_FMAC_DPP(2, ctrl) SDB_FMAC_DPP(3, ctrl)     \
        SDB_FMAC_DPP(4, ctrl) SDB_FMAC_DPP(5, ctrl) SDB_FMAC_DPP(6, ctrl) SDB_FMAC_DPP(7, ctrl) trail
#define SDB_FMAC_DPP1(ctrl) "s_nop 4\n\tv_fmac_f32_dpp %0, %0, %1 " ctrl " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\ts_nop 4"

template <int S>
__device__ __forceinline__ void fmac_dpp8(float* v, float sg) {
#define SDB_OPS "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]), "+v"(v[7]) : "v"(sg)
    if constexpr (S == 1) asm volatile(SDB_FMAC_DPP8("s_nop 4\n\t", "quad_perm:[1,0,3,2]", "") : SDB_OPS);
    if constexpr (S == 2) asm volatile(SDB_FMAC_DPP8("", "quad_perm:[2,3,0,1]", "") : SDB_OPS);
    if constexpr (S == 4) asm volatile(SDB_FMAC_DPP8("", "row_half_mirror", "") : SDB_OPS);
    if constexpr (S == 8) asm volatile(SDB_FMAC_DPP8("", "row_mirror", "s_nop 4") : SDB_OPS);
#undef SDB_OPS
}
template <int S>
__device__ __forceinline__ void fmac_dpp1(float& v, float sg) {
    if constexpr (S == 1) asm volatile(SDB_FMAC_DPP1("quad_perm:[1,0,3,2]") : "+v"(v) : "v"(sg));
    if constexpr (S == 2) asm volatile(SDB_FMAC_DPP1("quad_perm:[2,3,0,1]") : "+v"(v) : "v"(sg));
    if constexpr (S == 4) asm volatile(SDB_FMAC_DPP1("row_half_mirror") : "+v"(v) : "v"(sg));
    if constexpr (S == 8) asm volatile(SDB_FMAC_DPP1("row_mirror") : "+v"(v) : "v"(sg));
}
template <int S, int VPL>
__device__ __forceinline__ void fwht_dpp(float (&v)[VPL], int lg) {
    const float sg = (lg & S) ? -1.0f : 1.0f;
    if constexpr (VPL % 8 == 0) {
#pragma unroll
        for (int j = 0; j < VPL; j += 8) fmac_dpp8<S>(&v[j], sg);
    } else {
#pragma unroll
        for (int j = 0; j < VPL; ++j) fmac_dpp1<S>(v[j], sg);
    }
}

// Unnormalised FWHT with increasing strides (the level order of FFHT,
// src/fht_avx.c:1013-1251).  Element e lives in register e >> 6 of the lane
// whose logical index (mirror layout, lane_logical) is e & 63.
// Outputs equal FFHT's as IEEE values times the pending sign of fwht_dpp;
// FFHT's (0 - v) butterflies at strides 1, 2, 4 and the negated levels can
// only change the sign of an exact zero, which no comparison of the index
// distinguishes.
template <int LOGD>
__device__ __forceinline__ void ffht_wave(float (&v)[(1 << LOGD) / 64], int lg) {
    constexpr int VPL = (1 << LOGD) / 64;
    fwht_dpp<1, VPL>(v, lg);
    fwht_dpp<2, VPL>(v, lg);
    fwht_dpp<4, VPL>(v, lg);
    fwht_dpp<8, VPL>(v, lg);
    if constexpr (VPL % 2 == 0) {
        fwht_cross_pairs<16, VPL>(v);
        fwht_cross_pairs<32, VPL>(v);
    } else {
        fwht_cross<16, VPL>(v, lg);
        fwht_cross<32, VPL>(v, lg);
    }
#pragma unroll
    for (int js = 1; js < VPL; js <<= 1) fwht_inreg<VPL>(v, js);
}

// Embedding of a transformed point for the metrics other than cosine, staged
// through the wave's LDS slice (element e at emb[e]); loops are not unrolled so
// the transcendental code is emitted once.  L1/L2: [cos(Rx), sin(Rx)] with Rx
// a sequential fmaf chain over the input dims (src/sDbscan.cpp:59-64);
// Chi2/JS: the deterministic sampled features (src/Utilities.h:132-208).
template <int D>
__device__ __noinline__ void embed_lds(const IndexArgs& a, long long p, int lane, float* emb) {
    const float* x = a.Xn + p * a.ldx;
    const int d = a.d;
    const float L = a.kerL;
    for (int e = lane; e < D; e += 64) {
        float v = 0.0f;
        if (a.metric == kL1 || a.metric == kL2) {
            const int half = a.ker / 2;
            if (e < 2 * half) {
                const int r = e < half ? e : e - half;
                float acc = 0.0f;
                for (int c = 0; c < d; ++c) acc = fmaf(a.R[(size_t)c * half + r], x[c], acc);
                v = e < half ? cosf(acc) : sinf(acc);
            }
        } else if (e < a.ker) {
            const int comps = ((a.ker / d) - 1) / 2;
            if (e < d) {
                const float xv = x[e];
                if (xv > 0.0f) {
                    if (a.metric == kChi2) v = sqrtf(xv * L);
                    else v = (float)sqrt((double)(xv * L) * 2.0 / log(4.0));
                }
            } else {
                const int t = e - d;
                const int i = t / (2 * d) + 1;
                const int rem = t % (2 * d);
                const float xv = x[rem < d ? rem : rem - d];
                if (i <= comps && xv > 0.0f) {
                    float f, ph;
                    if (a.metric == kChi2) {
                        f = (float)sqrt(2.0 * xv * L / cosh(M_PI * i * L));
                        ph = (float)i * L * logf(xv);
                    } else {
                        const float iL = (float)i * L;
                        const float kap = (float)(2.0 / (log(4.0) * (double)(1.0f + 4.0f * iL * iL) * cosh(M_PI * i * L)));
                        f = (float)sqrt(2.0 * xv * L * kap);
                        ph = iL * logf(xv);
                    }
                    v = rem < d ? f * cosf(ph) : f * sinf(ph);
                }
            }
        }
        emb[e] = v;
    }
}

// HD3 rotation in registers (src/sDbscan.cpp:77-88): 3 x {sign flip, FWHT};
// the (2 bit - 1) multiplier negates where the bit is 0.  sg[r][w] bit b is
// the sign bit of register 32 w + b in round r for this lane's elements; the
// masks of rounds 1 and 2 come with the pending sign of the previous FWHT
// folded in (bit flipped on lanes with odd popcount(lg & 15)), and the last
// FWHT's pending sign is removed at the end.
template <int LOGD>
__device__ __forceinline__ void hd3_rotate(float (&v)[(1 << LOGD) / 64], int lg,
                                           const uint32_t (&sg)[3][((1 << LOGD) / 64 + 31) / 32]) {
    constexpr int VPL = (1 << LOGD) / 64;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const uint32_t bit = (sg[r][j >> 5] >> (j & 31)) & 1u;
            v[j] = __uint_as_float(__float_as_uint(v[j]) ^ ((bit ^ 1u) << 31));
        }
        ffht_wave<LOGD>(v, lg);
    }
    const uint32_t pend = (uint32_t)(__popc(lg & 15) & 1) << 31;
#pragma unroll
    for (int j = 0; j < VPL; ++j) v[j] = __uint_as_float(__float_as_uint(v[j]) ^ pend);
}

// The same rotation with the sign words of each round read from an LDS table
// (one xor per element instead of a shift and a masked xor): tbl holds, per
// round r and register quad k, one 16-byte word per lane at
// tbl[(r * VPL / 4 + k) * 64 + lane] (hd3_sign_table), the sign bit of
// register 4 k + c in component c.
template <int LOGD>
__device__ __forceinline__ void hd3_rotate_tbl(float (&v)[(1 << LOGD) / 64], int lg, const uint4* tbl) {
    constexpr int VPL = (1 << LOGD) / 64;
    static_assert(VPL % 4 == 0, "sign table rows are register quads");
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int k = 0; k < VPL / 4; ++k) {
            const uint4 m = tbl[(r * (VPL / 4) + k) * 64];
            v[4 * k + 0] = __uint_as_float(__float_as_uint(v[4 * k + 0]) ^ m.x);
            v[4 * k + 1] = __uint_as_float(__float_as_uint(v[4 * k + 1]) ^ m.y);
            v[4 * k + 2] = __uint_as_float(__float_as_uint(v[4 * k + 2]) ^ m.z);
            v[4 * k + 3] = __uint_as_float(__float_as_uint(v[4 * k + 3]) ^ m.w);
        }
        ffht_wave<LOGD>(v, lg);
    }
    const uint32_t pend = (uint32_t)(__popc(lg & 15) & 1) << 31;
#pragma unroll
    for (int j = 0; j < VPL; ++j) v[j] = __uint_as_float(__float_as_uint(v[j]) ^ pend);
}

// Fills the sign table of hd3_rotate_tbl from the lane's sign masks (called
// by one wave of the workgroup; every lane writes its own words).
template <int LOGD>
__device__ __forceinline__ void hd3_sign_table(const uint32_t (&sg)[3][((1 << LOGD) / 64 + 31) / 32], int lane,
                                               uint4* tbl) {
    constexpr int VPL = (1 << LOGD) / 64;
    auto word = [&](int r, int j) { return ((((sg[r][j >> 5] >> (j & 31)) & 1u) ^ 1u) << 31); };
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < VPL / 4; ++k)
            tbl[(r * (VPL / 4) + k) * 64 + lane] =
                uint4{word(r, 4 * k), word(r, 4 * k + 1), word(r, 4 * k + 2), word(r, 4 * k + 3)};
}

// Bitonic sorting network across the 64 lanes of a wavefront, descending:
// stage k merges blocks of size k; lanes with (lane & k) == 0 sort their
// block descending, so after the last stage lane 0 holds the largest key.
template <int J>
__device__ __forceinline__ void bitonic_u32(uint32_t& x, int lane, int k) {
    const uint32_t p = xchg<J>(x, lane);
    const bool keep_max = ((lane & J) == 0) == ((lane & k) == 0);
    x = keep_max ? (p > x ? p : x) : (p < x ? p : x);
}
template <int J>
__device__ __forceinline__ void bitonic_u64(uint32_t& h, uint32_t& lo, int lane, int k) {
    const uint32_t ph = xchg<J>(h, lane), pl = xchg<J>(lo, lane);
    const bool pg = ph > h || (ph == h && pl > lo);
    const bool keep_max = ((lane & J) == 0) == ((lane & k) == 0);
    const bool take = keep_max ? pg : !pg && (ph != h || pl != lo);
    h = take ? ph : h;
    lo = take ? pl : lo;
}
template <int K2, int J>
__device__ __forceinline__ void bitonic_pass_u32(uint32_t& x, int lane) {
    bitonic_u32<J>(x, lane, K2);
    if constexpr (J > 1) bitonic_pass_u32<K2, J / 2>(x, lane);
}
template <int K2>
__device__ __forceinline__ void bitonic_sort_u32(uint32_t& x, int lane) {
    if constexpr (K2 > 2) bitonic_sort_u32<K2 / 2>(x, lane);
    bitonic_pass_u32<K2, K2 / 2>(x, lane);
}
template <int K2, int J>
__device__ __forceinline__ void bitonic_pass_u64(uint32_t& h, uint32_t& lo, int lane) {
    bitonic_u64<J>(h, lo, lane, K2);
    if constexpr (J > 1) bitonic_pass_u64<K2, J / 2>(h, lo, lane);
}
template <int K2>
__device__ __forceinline__ void bitonic_sort_u64(uint32_t& h, uint32_t& lo, int lane) {
    if constexpr (K2 > 2) bitonic_sort_u64<K2 / 2>(h, lo, lane);
    bitonic_pass_u64<K2, K2 / 2>(h, lo, lane);
}

__device__ __forceinline__ int lanes_below(unsigned long long b) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

// Fallback selection of the K largest keys (key desc, index asc) among the
// first P elements, used when more than 64 elements pass the candidate bound.
// Winners are written to out[0..K); returns true when the (K+1)-th largest
// key equals the K-th (a tie across the selection boundary).
template <int VPL>
__device__ __forceinline__ bool wave_topk_general(const uint32_t (&key)[VPL], int P, int K, int lane, int lg,
                                                  int* out) {
    // K + 1 rounds of a wave-wide argmax over the register keys (value desc,
    // index asc); the winner of a round is masked out in its lane.
    uint64_t removed = 0;
    uint64_t prev = 0;
    bool tie = false;
    for (int k = 0; k <= K; ++k) {
        uint64_t best = 0;
        int jb = -1;
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const int e = lg + 64 * j;
            if (e < P && !((removed >> j) & 1ull)) {
                const uint64_t c = ((uint64_t)key[j] << 32) | (uint32_t)(0xffffffffu - (uint32_t)e);
                if (c > best) { best = c; jb = j; }
            }
        }
        const uint64_t w = wave_max_u64(best);
        if (w == 0) break;
        if (k < K) {
            if (best == w && jb >= 0) removed |= 1ull << jb;
            if (lane == 0) out[k] = (int)(0xffffffffu - (uint32_t)w);
            prev = w;
        } else {
            tie = (uint32_t)(w >> 32) == (uint32_t)(prev >> 32);
        }
    }
    return tie;
}

// Both ends of the per-point selection (src/sDbscan.cpp:97-143) in one pass:
// the K largest (close) and the K smallest (far) projection values, each as
// (value desc / asc, index asc).  Far keys are the bitwise complements of the
// close keys (order reversed; +0 and -0 stay equal); elements at or beyond P
// get key 0 on both sides and are never selected.
//   1. T <= the K-th largest lane maximum (kth_bounds): at least K elements
//      are >= T, so every winner is a candidate.
//   2. Candidates (key >= T) are compacted into LDS with ballots.
//   3. Each candidate lane counts the candidates ranked before it (a readlane
//      loop over at most 64 entries); rank r < K goes to out[r].
// More than 64 candidates on a side take the K + 1 round wave argmax.
// Returns the tie flags: the (K+1)-th key equals the K-th.
// Lower bounds for the K-th largest lane maximum of the close and the far
// keys: a ballot bisection over the top NB bits of the key, keeping the
// largest prefix T with at least K lane maxima >= T.  Lanes whose maximum is
// 0 (no valid element) never count.
template <int NB>
__device__ __forceinline__ void kth_bounds(uint32_t mc, uint32_t mf, int K, uint32_t& tc, uint32_t& tf) {
    uint32_t pc = 0u, pf = 0u;
#pragma unroll
    for (int b = 31; b >= 32 - NB; --b) {
        const uint32_t xc = pc | (1u << b), xf = pf | (1u << b);
        if (__popcll(__ballot(mc >= xc)) >= K) pc = xc;
        if (__popcll(__ballot(mf >= xf)) >= K) pf = xf;
    }
    tc = pc;
    tf = pf;
}

// Winners of one side from its ranks: rank r < K goes to out[r]; the tie flag
// compares the K-th and (K+1)-th keys.
__device__ __forceinline__ void rank_emit(uint32_t h, uint32_t lo, int rank, int total, int K, int lane, int* out,
                                          bool& tie) {
    if (lane < total && rank < K) out[rank] = (int)~lo;
    tie = false;
    if (total > K) {
        const unsigned long long bk = __ballot(lane < total && rank == K - 1);
        const unsigned long long bn = __ballot(lane < total && rank == K);
        const uint32_t vk = (uint32_t)__builtin_amdgcn_readlane((int)h, __builtin_ctzll(bk));
        const uint32_t vn = (uint32_t)__builtin_amdgcn_readlane((int)h, __builtin_ctzll(bn));
        tie = vk == vn;
    }
}

// Ranks of the compacted candidates of both sides (at most 64 each; slot i
// in lane i, as (key, ~index) with the far keys stored as close keys): the
// rank of a candidate counts the candidates of its side ordered before it
// (key desc, index asc).  One readlane loop serves both sides; slots past a
// side's count hold (0, 0), which orders after every candidate.
__device__ __forceinline__ void rank_select2(const uint2* shc, int nc, const uint2* shf, int nf, int K, int lane,
                                             int* out_c, int* out_f, bool& tie_c, bool& tie_f) {
    uint32_t hc = 0u, lc = 0u, hf = 0u, lf = 0u;
    if (lane < nc) { const uint2 e = shc[lane]; hc = e.x; lc = e.y; }
    if (lane < nf) { const uint2 e = shf[lane]; hf = ~e.x; lf = e.y; }
    const uint64_t mc = ((uint64_t)hc << 32) | lc, mf = ((uint64_t)hf << 32) | lf;
    int rc = 0, rf = 0;
    const int nmax = nc > nf ? nc : nf;
    for (int i = 0; i < nmax; ++i) {
        const uint64_t oc = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hc, i) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)lc, i);
        const uint64_t of = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hf, i) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)lf, i);
        rc += oc > mc ? 1 : 0;
        rf += of > mf ? 1 : 0;
    }
    rank_emit(hc, lc, rc, nc, K, lane, out_c, tie_c);
    rank_emit(hf, lf, rf, nf, K, lane, out_f, tie_f);
    __builtin_amdgcn_wave_barrier();
}

// Candidate compaction of both sides into LDS slots [0, 64): only the
// candidate lanes store (a register with no candidate on a side skips its
// stores on an empty exec mask).  A side's slots wrap inside its own 128-slot
// area; past 64 candidates the side goes to the general selection, which does
// not read the slots.  FULL: every register holds P elements.
template <int VPL, bool FULL>
__device__ __forceinline__ void compact_both(const uint32_t (&kc)[VPL], int P, int lane, int lg, uint32_t tc,
                                             uint32_t tfn, uint32_t* shc, uint32_t* shf, int& nc_out, int& nf_out) {
    int nc = 0, nf = 0;
#pragma unroll
    for (int j = 0; j < VPL; ++j) {
        const uint32_t ne = ~(uint32_t)(lg + 64 * j);
        bool cc, cf;
        if constexpr (FULL) {
            cc = kc[j] >= tc;
            cf = kc[j] <= tfn;
        } else {
            const bool live = 64 * j + 63 < P || kc[j] != 0u;
            cc = live && kc[j] >= tc;
            cf = live && kc[j] <= tfn;
        }
        const unsigned long long bc = __ballot(cc), bf = __ballot(cf);
        if (cc) {
            const uint32_t pc = __builtin_amdgcn_mbcnt_hi((uint32_t)(bc >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bc, (uint32_t)nc)) & 127u;
            shc[2 * pc] = kc[j];
            shc[2 * pc + 1] = ne;
        }
        if (cf) {
            const uint32_t pf = __builtin_amdgcn_mbcnt_hi((uint32_t)(bf >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bf, (uint32_t)nf)) & 127u;
            shf[2 * pf] = kc[j];
            shf[2 * pf + 1] = ne;
        }
        nc += (int)__popcll(bc);
        nf += (int)__popcll(bf);
    }
    nc_out = nc;
    nf_out = nf;
}

template <int VPL>
__device__ __forceinline__ void wave_topk_both(const uint32_t (&kc)[VPL], int P, int K, int lane, int lg, uint32_t* lds,
                                               int* out_close, int* out_far, bool& tie_c, bool& tie_f) {
    // kc: close keys, 0 for elements at or beyond P; the far key of a valid
    // element is ~kc, so far comparisons run on kc with complemented bounds.
    const bool full = P >= VPL * 64;
    uint32_t mc = 0u, mn = 0xffffffffu;
#pragma unroll
    for (int j = 0; j < VPL; ++j) {
        mc = kc[j] > mc ? kc[j] : mc;
        const bool live = 64 * j + 63 < P || kc[j] != 0u;
        mn = live && kc[j] < mn ? kc[j] : mn;
    }
    uint32_t tc, tf;
    kth_bounds<16>(mc, mn == 0xffffffffu ? 0u : ~mn, K, tc, tf);
    const uint32_t tfn = ~tf;  // far candidate: kc <= ~tf
    uint32_t* shc = lds;
    uint32_t* shf = lds + 256;
    int nc, nf;
    if (full) compact_both<VPL, true>(kc, P, lane, lg, tc, tfn, shc, shf, nc, nf);
    else compact_both<VPL, false>(kc, P, lane, lg, tc, tfn, shc, shf, nc, nf);
    __builtin_amdgcn_wave_barrier();
    bool tc_, tf_;
    rank_select2((const uint2*)shc, nc <= 64 ? nc : 0, (const uint2*)shf, nf <= 64 ? nf : 0, K, lane, out_close,
                 out_far, tc_, tf_);
    tie_c = tc_;
    tie_f = tf_;
    if (nc > 64) tie_c = wave_topk_general<VPL>(kc, P, K, lane, lg, out_close);
    if (nf > 64) {
        uint32_t kf[VPL];
#pragma unroll
        for (int j = 0; j < VPL; ++j) kf[j] = kc[j] ? ~kc[j] : 0u;
        tie_f = wave_topk_general<VPL>(kf, P, K, lane, lg, out_far);
    }
}

// ---------------------------------------------------------------------------
// Index passes over the points, one wavefront per point:
//   kSample: every sample_stride-th point; HD3 projections and per-chunk row
//            maxima / minima for the top-M thresholds.
//   kFull:   every point; (cosine: transformData fused in) HD3 projections,
//            per-point top-K (src/sDbscan.cpp:97-143) with capture of points
//            whose top-K is tied at the boundary, and the values beyond the
//            per-row thresholds appended to the top-M candidate lists.
//   kRows:   recompute and store whole projection rows listed in row_slot
//            (host replay of tied top-M rows).
//   kPoints: projections of the points listed in pt_ids (P x n_ids, column
//            per listed point) for the noise labelling.
// Grid: G workgroups x 4 waves; wave w of group g takes the points
// q = (4 g + w) + t * 4 G, so chunk g is an interleaved subset.
enum { kSample = 0, kFull = 1, kRows = 2, kPoints = 3 };

// Sum over the wave in FP64: xor butterfly over the logical lane index,
// offsets 32, 16, ..., 1 (the order of k_transform's tree).
__device__ __forceinline__ double wave_sum_f64(double v, int lg) {
    uint32_t lo = (uint32_t)__double_as_longlong(v);
    uint32_t hi = (uint32_t)((unsigned long long)__double_as_longlong(v) >> 32);
#define SDB_SUM_STEP(S)                                                                   \
    {                                                                                     \
        const uint32_t plo = xchg_l<S>(lo, lg), phi = xchg_l<S>(hi, lg);                  \
        const double o = __longlong_as_double((long long)(((unsigned long long)phi << 32) | plo)); \
        v = v + o;                                                                        \
        lo = (uint32_t)__double_as_longlong(v);                                           \
        hi = (uint32_t)((unsigned long long)__double_as_longlong(v) >> 32);               \
    }
    SDB_SUM_STEP(32) SDB_SUM_STEP(16) SDB_SUM_STEP(8) SDB_SUM_STEP(4) SDB_SUM_STEP(2) SDB_SUM_STEP(1)
#undef SDB_SUM_STEP
    return v;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Row loads and stores through buffer resources (base = the row, num_records
// = its valid bytes): element e sits at byte 4 e, and accesses at or past the
// end read 0 / are dropped by the hardware range check, so the ragged last
// register of a row needs no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* row, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(row), (short)0, bytes, 0x00020000);
}
template <int VPL>
__device__ __forceinline__ void load_row(const float* row, int d, int lg, float (&out)[VPL]) {
    const __amdgpu_buffer_rsrc_t r = row_rsrc(row, d * 4);
#pragma unroll
    for (int j = 0; j < VPL; ++j)
        out[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, 4 * lg + 256 * j, 0, 0));
}

// Embedding kinds of k_index.  kEmbCos: the transformed row itself;
// kEmbSqrt: Chi2 / JS with no sampled components (ker_n_features < 3 d, so
// (ker / d - 1) / 2 = 0 in src/Utilities.h:134-135 and :176-177): element e
// is sqrt(x_e L) (JS: sqrt(2 x_e L / ln 4)) for e < min(d, ker) and x_e > 0,
// else 0 -- elementwise, in registers; kEmbLds: the general embeddings through
// the wave's LDS slice (embed_lds); kEmbRows: L1/L2 embeddings precomputed by
// k_rff_embed (the same values), read as rows.  kEmbCos and kEmbSqrt fuse
// transformData into the passes that read untransformed rows.
enum EmbKind { kEmbCos = 0, kEmbSqrt = 1, kEmbLds = 2, kEmbRows = 3 };

// L1/L2 random features of every row (src/sDbscan.cpp:59-64): E[p][r] =
// cos(s), E[p][half + r] = sin(s) with s = R[0][r] x_0 + ... + R[d-1][r] x_{d-1}
// as a sequential fmaf chain over the input dims, as in embed_lds, so the
// values are the same.  A 64-point x 64-feature tile per workgroup (4 x 4 per
// thread), its X and R slices staged through LDS 32 dims at a time.
__global__ __launch_bounds__(256) void k_rff_embed(const float* __restrict__ Xn, long long n, int d, int ldx,
                                                   const float* __restrict__ R, int half, float* __restrict__ E,
                                                   int ldE) {
    __shared__ float xs[64][33];
    __shared__ float rs[32][64];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const long long p0 = (long long)blockIdx.x * 64;
    const int r0 = blockIdx.y * 64;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int c0 = 0; c0 < d; c0 += 32) {
        const int cn = min(32, d - c0);
        for (int t = threadIdx.x; t < 64 * 32; t += 256) {
            const int pp = t >> 5, cc = t & 31;
            const long long p = p0 + pp;
            xs[pp][cc] = (p < n && cc < cn) ? Xn[p * ldx + c0 + cc] : 0.0f;
        }
        for (int t = threadIdx.x; t < 32 * 64; t += 256) {
            const int cc = t >> 6, rr = t & 63;
            rs[cc][rr] = (cc < cn && r0 + rr < half) ? R[(size_t)(c0 + cc) * half + r0 + rr] : 0.0f;
        }
        __syncthreads();
        for (int cc = 0; cc < cn; ++cc) {
            float xv[4], rv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) xv[i] = xs[4 * ty + i][cc];
#pragma unroll
            for (int j = 0; j < 4; ++j) rv[j] = rs[cc][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(rv[j], xv[i], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long long p = p0 + 4 * ty + i;
        if (p >= n) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + tx + 16 * j;
            if (r < half) {
                E[p * ldE + r] = cosf(acc[i][j]);
                E[p * ldE + half + r] = sinf(acc[i][j]);
            }
        }
    }
}

// FULLP: n_proj == fhtDim (every element of a lane's registers is a
// projection), which folds the ragged-end tests away at compile time.
template <int LOGD, int MODE, int EMB, bool FULLP>
// Occupancy: 3 workgroups per CU up to D = 1024 (at most 168 registers),
// fewer for the longer rows so that their registers do not spill.
__global__ __launch_bounds__(256, LOGD <= 10 ? 3 : (LOGD == 11 ? 2 : 1)) void k_index(IndexArgs a) {
    constexpr int D = 1 << LOGD;
    constexpr int VPL = D / 64;
    const int lane = lane_id();
    const int lg = lane_logical(lane);  // mirror layout: this lane holds elements lg + 64 j
    const int wib = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    const long long gw = (long long)blockIdx.x * 4 + wib;
    const long long nw = (long long)gridDim.x * 4;
    const int P = FULLP ? D : a.P;
    constexpr int NW = (VPL + 31) / 32;

    extern __shared__ __align__(16) unsigned char smem[];
    // LDS: [top-K scratch 4 x 2 KB (kFull)] [2P keys] [embedding 4 x D x 4 B (kEmbLds)]
    const size_t off_red = MODE == kFull ? (size_t)4 * 2048 : 0;
    // threshold area: kSample row max / min accumulators [P] + [P]; kFull
    // per-row (hi, lo) threshold pairs for all D elements (past P: never hit)
    const size_t off_emb = off_red + (size_t)8 * (MODE == kFull ? D : P);
    uint64_t* tks = (uint64_t*)smem + (size_t)wib * 256;
    uint32_t* red = (uint32_t*)(smem + off_red);
    float* emb = (float*)(smem + off_emb) + (size_t)wib * D;
    // sign table of the HD3 rounds (register quads: D >= 256), after the
    // embedding slices
    constexpr bool SGN_LDS = VPL % 4 == 0 && LOGD <= 11;
    const size_t off_sgn = (off_emb + (EMB == kEmbLds ? (size_t)4 * D * 4 : 0) + 15) & ~(size_t)15;
    uint4* sgn_tbl = (uint4*)(smem + off_sgn);

    uint32_t sg[3][NW];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int w = 0; w < NW; ++w) sg[r][w] = a.signmask[(r * NW + w) * 64 + lane];
    if constexpr (SGN_LDS) {
        if (wib == 0) hd3_sign_table<LOGD>(sg, lane, sgn_tbl);
    }


    if (MODE == kSample) {
        for (int i = threadIdx.x; i < P; i += blockDim.x) { red[i] = 0u; red[P + i] = 0xffffffffu; }
    } else if (MODE == kFull) {
        uint2* thr = (uint2*)red;
        for (int i = threadIdx.x; i < D; i += blockDim.x)
            thr[i] = i < P ? uint2{a.thr_hi[i], a.thr_lo[i]} : uint2{0xffffffffu, 0u};
    }
    __syncthreads();

    uint32_t cmax[MODE == kSample ? VPL : 1], cmin[MODE == kSample ? VPL : 1];
    if constexpr (MODE == kSample) {
#pragma unroll
        for (int j = 0; j < VPL; ++j) { cmax[j] = 0u; cmin[j] = 0xffffffffu; }
    }
    const long long stride = MODE == kSample ? a.sample_stride : 1;
    const long long npts = MODE == kPoints ? a.n_ids : (a.n + stride - 1) / stride;
    float rho_max = 0.0f;  // largest BF16 rounding bound of the wave's rows
    constexpr bool RAW = (EMB == kEmbCos || EMB == kEmbSqrt) && (MODE == kSample || MODE == kFull);  // reads untransformed rows
    for (long long q = gw; q < npts; q += nw) {
        const long long p = MODE == kPoints ? (long long)a.pt_ids[q] : q * stride;
        float v[VPL];
        // the sign masks are re-derived per point rather than kept expanded
        // across the loop (3 registers instead of 3 x VPL)
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int w = 0; w < NW; ++w) asm volatile("" : "+v"(sg[r][w]));
        if constexpr (EMB == kEmbRows) {
            load_row<VPL>(a.emb + p * a.ldE, a.ldE, lg, v);
        } else if constexpr (EMB != kEmbLds) {
            if constexpr (!RAW) {
                load_row<VPL>(a.Xn + p * a.ldx, a.d, lg, v);
            } else {
                // transformData (src/Utilities.h:295-310): Cosine x / ||x||, with
                // ||x||^2 summed in FP64 per lane (elements e = lg + 64 j) then by
                // the xor butterfly over lg -- the order of k_transform; Chi2 / JS
                // x / sum(x) with the sum in the same order, and the
                // nonnegativity check.
                load_row<VPL>(a.X + p * a.d, a.d, lg, v);
                double acc = 0.0;
                bool neg = false;
#pragma unroll
                for (int j = 0; j < VPL; ++j) {
                    const double t = v[j];
                    // t * t is exact in FP64, so the fma rounds acc + t * t once, as the sum does
                    if constexpr (EMB == kEmbCos) acc = fma(t, t, acc);
                    else { acc = acc + t; neg |= v[j] < 0.0f; }
                }
                const double tot = wave_sum_f64(acc, lg);
                const float nrm = EMB == kEmbCos ? (float)sqrt(tot) : (float)tot;
                if constexpr (EMB != kEmbCos) {
                    if (__any(neg) && lane == 0) atomicOr(a.err, 1);
                }
                if (!(nrm > 0.0f)) {
                    if (lane == 0) atomicOr(a.err, 2);
                }
                // x / nrm, correctly rounded, without a division per element:
                // the FP64 product with the FP64 reciprocal is within 2^-52
                // (relative) of the exact quotient, while an FP32 quotient of two
                // FP32 numbers is never closer than 2^-49 (relative) to an FP32
                // rounding boundary, so rounding the product to FP32 gives the
                // rounded quotient.  Elements past d are 0 and stay 0.
                const double rinv = 1.0 / (double)nrm;
#pragma unroll
                for (int j = 0; j < VPL; ++j) v[j] = (float)((double)v[j] * rinv);
                if (MODE == kFull) {
                    // stores past ldx are dropped by the range check
                    const __amdgpu_buffer_rsrc_t ro = row_rsrc(a.Xn_out + p * a.ldx, a.ldx * 4);
#pragma unroll
                    for (int j = 0; j < VPL; ++j)
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v[j]), ro, 4 * lg + 256 * j, 0, 0);
                    if (a.Xb_out) {
                        // BF16 rows, round to nearest even (hardware pair
                        // conversion); Chi2: of the square roots (the rows whose
                        // dot product bounds chi2 from below)
                        // (with the bound in element ldx - 1, the range ends
                        // before it: the zero of that padding element is
                        // dropped and the bound is stored below)
                        const __amdgpu_buffer_rsrc_t rb =
                            row_rsrc(a.Xb_out + p * a.ldx, (a.ldx - (a.rho_pad ? 1 : 0)) * 2);
                        auto bsrc = [&](int j) { return EMB == kEmbCos ? v[j] : sqrtf(v[j]); };
                        // with the squared rounding errors of the row (the
                        // per-pair band of the BF16 tiles); elements past d
                        // are 0 in both
                        float sq = 0.0f;
#pragma unroll
                        for (int j = 0; j + 1 < VPL; j += 2) {
                            const float s0 = bsrc(j), s1 = bsrc(j + 1);
                            const uint32_t u = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){s0, s1}, bf16x2));
                            __builtin_amdgcn_raw_buffer_store_b16((uint16_t)u, rb, 2 * lg + 128 * j, 0, 0);
                            __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(u >> 16), rb, 2 * lg + 128 * (j + 1), 0, 0);
                            const float d0 = __uint_as_float(u << 16) - s0, d1 = __uint_as_float(u & 0xffff0000u) - s1;
                            sq = fmaf(d0, d0, sq);
                            sq = fmaf(d1, d1, sq);
                        }
                        if constexpr (VPL % 2 == 1) {
                            const float s0 = bsrc(VPL - 1);
                            const uint32_t h = f32_to_bf16_rne(s0);
                            __builtin_amdgcn_raw_buffer_store_b16((uint16_t)h, rb, 2 * lg + 128 * (VPL - 1), 0, 0);
                            const float d0 = __uint_as_float(h << 16) - s0;
                            sq = fmaf(d0, d0, sq);
                        }
                        const float rho = bf16_row_err(wave_sum_f32(sq), EMB != kEmbCos);
                        if (lane == 0) {
                            a.rowerr[p] = rho;
                            if (a.rho_pad) a.Xb_out[p * a.ldx + a.ldx - 1] = (uint16_t)bf16_up(rho);
                        }
                        rho_max = fmaxf(rho_max, rho);
                    }
                }
            }
            if constexpr (EMB == kEmbSqrt) {
                // the features of src/Utilities.h:140-145 (Chi2) / :180-185 (JS)
                const float L = a.kerL;
                const int lim = a.d < a.ker ? a.d : a.ker;
#pragma unroll
                for (int j = 0; j < VPL; ++j) {
                    const float xv = v[j];
                    float f = 0.0f;
                    if (xv > 0.0f && lg + 64 * j < lim)
                        f = a.metric == kChi2 ? sqrtf(xv * L) : (float)sqrt((double)(xv * L) * 2.0 / log(4.0));
                    v[j] = f;
                }
            }
        } else {
            embed_lds<D>(a, p, lane, emb);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < VPL; ++j) v[j] = emb[lg + 64 * j];
            __builtin_amdgcn_wave_barrier();
        }
        if constexpr (SGN_LDS) {
            // the lane's table offset, opaque to the compiler so that the
            // table reads stay in the loop (48 registers if hoisted)
            int so = lane;
            asm volatile("" : "+v"(so));
            hd3_rotate_tbl<LOGD>(v, lg, sgn_tbl + so);
        } else {
            hd3_rotate<LOGD>(v, lg, sg);
        }

        if constexpr (MODE == kSample) {
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                if (lg + 64 * j < P) {
                    const uint32_t k = ord_key(v[j]);
                    cmax[j] = k > cmax[j] ? k : cmax[j];
                    cmin[j] = k < cmin[j] ? k : cmin[j];
                }
            }
        } else if constexpr (MODE == kFull) {
            uint32_t kc[VPL];
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                kc[j] = ord_key(v[j]);
                if (64 * j + 63 >= P && lg + 64 * j >= P) kc[j] = 0u;
            }
            int* tk = a.topK + p * 2 * a.K;
            bool t1, t2;
            wave_topk_both<VPL>(kc, P, a.K, lane, lg, (uint32_t*)tks, tk, tk + a.K, t1, t2);
            if (t1 || t2) {
                int slot = 0;
                if (lane == 0) slot = atomicAdd(a.tie_count, 1);
                slot = __shfl(slot, 0, kWave);
                // every tied point is listed (tie_idx holds n entries); the
                // values of the first tie_cap are kept, the others are
                // projected again for the replay
                if (lane == 0) a.tie_idx[slot] = (int)p;
                if (slot < a.tie_cap) {
                    float* tv = a.tie_vals + (size_t)slot * P;
#pragma unroll
                    for (int j = 0; j < VPL; ++j) {
                        const int e = lg + 64 * j;
                        if (e < P) tv[e] = key_to_float(kc[j]);
                    }
                }
            }
            // values beyond their row's threshold are rare: one unrolled test
            // against the per-row (hi, lo) pairs, then the appends.  Elements
            // past P have key 0 and the pair (~0, 0).
            const uint2* thr = (const uint2*)red;
            bool anyc = false;
            if (P >= VPL * 64) {
#pragma unroll
                for (int j = 0; j < VPL; ++j) {
                    const uint2 th = thr[lg + 64 * j];
                    anyc |= kc[j] >= th.x || kc[j] <= th.y;
                }
            } else {
#pragma unroll
                for (int j = 0; j < VPL; ++j) {
                    const uint2 th = thr[lg + 64 * j];
                    anyc |= kc[j] >= th.x || (kc[j] <= th.y && kc[j] != 0u);
                }
            }
            if (__any(anyc))
#pragma unroll 1
            for (int j = 0; j < VPL; ++j) {
                const int e = lg + 64 * j;
                const uint32_t k = kc[j];
                if (k == 0u) continue;
                const uint2 th = thr[e];
                if (k >= th.x) {
                    const int pos = atomicAdd(&a.cand_cnt[2 * e], 1);
                    if (pos < a.cand_cap) {
                        const size_t o = (size_t)(2 * e) * a.cand_cap + pos;
                        a.cand_key[o] = k;
                        a.cand_idx[o] = (int)p;
                    }
                }
                if (k <= th.y) {
                    const int pos = atomicAdd(&a.cand_cnt[2 * e + 1], 1);
                    if (pos < a.cand_cap) {
                        const size_t o = (size_t)(2 * e + 1) * a.cand_cap + pos;
                        a.cand_key[o] = k;
                        a.cand_idx[o] = (int)p;
                    }
                }
            }
        } else if constexpr (MODE == kRows) {
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const int e = lg + 64 * j;
                if (e < P) {
                    const int s = a.row_slot[e];
                    if (s >= 0) a.row_buf[(size_t)s * a.n + p] = v[j];
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const int e = lg + 64 * j;
                if (e < P) a.pt_buf[(size_t)e * a.n_ids + q] = v[j];
            }
        }
    }
    if constexpr (RAW && MODE == kFull) {
        // every wave ends here: the atomic only when it would raise the
        // maximum, so the waves do not queue on one address
        if (a.Xb_out && lane == 0 && (int)__float_as_uint(rho_max) > load_relaxed((const int*)a.rowerr_max))
            atomicMax(a.rowerr_max, __float_as_uint(rho_max));
    }
    if constexpr (MODE == kSample) {
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const int e = lg + 64 * j;
            if (e < P) { atomicMax(&red[e], cmax[j]); atomicMin(&red[P + e], cmin[j]); }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
            a.chunk_max[(size_t)blockIdx.x * P + i] = red[i];
            a.chunk_min[(size_t)blockIdx.x * P + i] = red[P + i];
        }
    }
}

// ---------------------------------------------------------------------------
// Per-row top-M thresholds: the (M+1)-th largest chunk maximum bounds the
// (M+1)-th largest projection value of the row from below (at least M+1
// chunks hold a value that large), likewise for minima.  One workgroup per
// row; rank by counting over the G chunk keys.
__global__ __launch_bounds__(256) void k_topm_threshold(const uint32_t* __restrict__ cmax,
                                                        const uint32_t* __restrict__ cmin, int G,
                                                        int P, int M, uint32_t* __restrict__ thi,
                                                        uint32_t* __restrict__ tlo) {
    const int r = blockIdx.x;
    __shared__ uint32_t smax[4096], smin[4096];
    __shared__ uint32_t res_hi, res_lo;
    if (threadIdx.x == 0) { res_hi = 0u; res_lo = 0xffffffffu; }
    for (int i = threadIdx.x; i < G; i += blockDim.x) {
        smax[i] = cmax[(size_t)i * P + r];
        smin[i] = cmin[(size_t)i * P + r];
    }
    __syncthreads();
    const int want = M + 1;
    if (G < want) {  // too few chunks: every value is a candidate
        if (threadIdx.x == 0) { thi[r] = 0u; tlo[r] = 0xffffffffu; }
        return;
    }
    for (int i = threadIdx.x; i < G; i += blockDim.x) {
        uint32_t x = smax[i], y = smin[i];
        int gx = 0, gy = 0;
        for (int j = 0; j < G; ++j) {
            uint32_t u = smax[j], w = smin[j];
            gx += (u > x || (u == x && j < i)) ? 1 : 0;
            gy += (w < y || (w == y && j < i)) ? 1 : 0;
        }
        if (gx == want - 1) res_hi = x;
        if (gy == want - 1) res_lo = y;
    }
    __syncthreads();
    if (threadIdx.x == 0) { thi[r] = res_hi; tlo[r] = res_lo; }
}

// Exact top-M per row and side from the candidate lists: bitonic sort of the
// (value, index) keys in LDS, descending value then ascending index (the far
// side sorts the complemented key, so the smallest values come first).
// flags[row*2+side]: 1 = tie across the boundary, 2 = candidate overflow;
// flagged rows are replayed on the host.
__global__ __launch_bounds__(1024) void k_topm_select(const uint32_t* __restrict__ cand_key,
                                                      const int* __restrict__ cand_idx,
                                                      const int* __restrict__ cand_cnt, int cap,
                                                      int P, int M, int* __restrict__ topM,
                                                      int* __restrict__ flags) {
    const int rs = blockIdx.x;  // row * 2 + side
    const int r = rs >> 1, side = rs & 1;
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t* s = (uint64_t*)smem;
    const int cnt = cand_cnt[rs];
    if (cnt > cap) {
        if (threadIdx.x == 0) flags[rs] = 2;
        return;
    }
    int sz = 1;
    while (sz < cnt) sz <<= 1;
    for (int i = threadIdx.x; i < sz; i += blockDim.x) {
        uint64_t k = 0;
        if (i < cnt) {
            uint32_t kv = cand_key[(size_t)rs * cap + i];
            if (side) kv = ~kv;
            k = ((uint64_t)kv << 32) | (uint32_t)(0xffffffffu - (uint32_t)cand_idx[(size_t)rs * cap + i]);
        }
        s[i] = k;
    }
    __syncthreads();
    for (int k = 2; k <= sz; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < sz; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    uint64_t x = s[i], y = s[ixj];
                    bool desc = (i & k) == 0;
                    if (desc ? (x < y) : (x > y)) { s[i] = y; s[ixj] = x; }
                }
            }
            __syncthreads();
        }
    }
    for (int m = threadIdx.x; m < M; m += blockDim.x)
        topM[(size_t)r * 2 * M + side * M + m] = m < cnt ? (int)(0xffffffffu - (uint32_t)s[m]) : -1;
    if (threadIdx.x == 0) {
        int f = 0;
        if (cnt < M) f = 2;
        else if (cnt > M && (uint32_t)(s[M - 1] >> 32) == (uint32_t)(s[M] >> 32)) f = 1;
        flags[rs] = f;
    }
}

// ---------------------------------------------------------------------------
// Candidate-list structures.  A "list" is a top-M list: list l < P is the
// close list of random vector l, list P + r the far list of r.  Point a's
// slot s (0 <= s < 2K) refers to list slot_list(a, s): slot 2k is the close
// list of its k-th closest vector, slot 2k+1 the far list of its k-th
// furthest vector (the reference's loop order, src/sDbscan.cpp:262-354).
__device__ __forceinline__ int slot_list(const int* topK, long long a, int s, int K, int P) {
    int k = s >> 1;
    return (s & 1) ? P + topK[a * 2 * K + K + k] : topK[a * 2 * K + k];
}

// lists[l * M + i] = point id of position i of list l (from matrix_topM).
__global__ void k_build_lists(const int* __restrict__ topM, int P, int M, int* __restrict__ lists,
                              uint8_t* __restrict__ is_hub) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * P * M) return;
    int l = t / M, i = t % M;
    int r = l < P ? l : l - P;
    int side = l < P ? 0 : 1;
    int pid = topM[(size_t)r * 2 * M + side * M + i];
    lists[t] = pid;
    is_hub[pid] = 1;
}

// Per hub membership count (hub id from the exclusive scan of is_hub).
__global__ void k_hub_count(const int* __restrict__ lists, int LM, const int* __restrict__ hub_id,
                            int* __restrict__ hub_cnt) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= LM) return;
    atomicAdd(&hub_cnt[hub_id[lists[t]]], 1);
}

__global__ void k_hub_fill(const int* __restrict__ lists, int LM, const int* __restrict__ hub_id,
                           const int* __restrict__ hub_ptr, int* __restrict__ hub_cur,
                           int* __restrict__ hub_mem) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= LM) return;
    int h = hub_id[lists[t]];
    int pos = atomicAdd(&hub_cur[h], 1);
    hub_mem[hub_ptr[h] + pos] = t;  // t = l * M + i
}

// inv[l][L] bit i set when list L is one of the 2K lists of the point at
// position i of list l.  Used for the mutual-pair correction of hub degrees.
__global__ void k_invmask(const int* __restrict__ lists, const int* __restrict__ topK, int P, int K,
                          int M, int L2, int W, uint32_t* __restrict__ inv) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long tot = (long long)L2 * M * 2 * K;
    if (t >= tot) return;
    int s = (int)(t % (2 * K));
    long long li = t / (2 * K);
    int l = (int)(li / M), i = (int)(li % M);
    int c = lists[li];
    int L = slot_list(topK, c, s, K, P);
    atomicOr(&inv[((size_t)l * L2 + L) * W + (i >> 5)], 1u << (i & 31));
}

// Per list position (l, i): pos_hub = hub id of the point there and pcnt =
// the number of its other list memberships (scanned into pair offsets).
__global__ void k_pos_hub(const int* __restrict__ lists, const int* __restrict__ hub_id,
                          const int* __restrict__ hub_ptr, int LM, int* __restrict__ pos_hub,
                          int* __restrict__ pcnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > LM) return;
    if (t == LM) { pcnt[LM] = 0; return; }
    const int h = hub_id[lists[t]];
    pos_hub[t] = h;
    pcnt[t] = hub_ptr[h + 1] - hub_ptr[h] - 1;
}

// pairs[poff[l M + i] ...]: (L << 8) | i for every other list L holding the
// point at position i of list l; the pairs of list l are contiguous.
__global__ void k_pair_fill(const int* __restrict__ pos_hub, const int* __restrict__ hub_ptr,
                            const int* __restrict__ hub_mem, int LM, int M, const int* __restrict__ poff,
                            uint32_t* __restrict__ pairs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= LM) return;
    const int l = t / M, i = t - l * M;
    const int h = pos_hub[t];
    int o = poff[t];
    for (int q = hub_ptr[h]; q < hub_ptr[h + 1]; ++q) {
        const int L = hub_mem[q] / M;
        if (L != l) pairs[o++] = ((uint32_t)L << 8) | (uint32_t)i;
    }
}

// Dense list intersections, for inputs whose lists share many points: word w
// of inter[l][L] has bit i set when the point at position 32w + i of list l
// is also a member of list L != l (the same facts as the pair table, indexed
// directly).  One thread per list position.
__global__ void k_inter_fill(const int* __restrict__ pos_hub, const int* __restrict__ hub_ptr,
                             const int* __restrict__ hub_mem, int LM, int M, int L2, int W,
                             uint32_t* __restrict__ inter) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= LM) return;
    const int l = t / M, i = t - l * M;
    const int h = pos_hub[t];
    for (int q = hub_ptr[h]; q < hub_ptr[h + 1]; ++q) {
        const int L = hub_mem[q] / M;
        if (L != l) atomicOr(&inter[((size_t)l * L2 + L) * W + (i >> 5)], 1u << (i & 31));
    }
}

// First-occurrence mask per (point, slot): bit i of F[a][s] is set when the
// point at position i of the slot's list is not a itself and does not occur
// in an earlier slot of a (the reference's per-point dedup bitset,
// src/sDbscan.cpp:260-282).  Only points held by several lists can repeat;
// the (position, other list) pairs of the slot's list are checked against the
// lists of the earlier slots.  A workgroup takes ppb consecutive points with one thread per (point, slot)
// entry: the slot lists and a bitmap of each point's lists go to LDS for the
// duplicate checks (a pair whose other list is not one of the point's costs
// one bit test, whatever the slot), and the W mask
// words of consecutive entries are written by consecutive threads (one
// 16-byte store each when W = 4).
// Each list's pair segment sorted by (other list, position) in LDS, one
// workgroup per list, so that k_first_mask can look up one other list by
// binary search; segments longer than kPairSortCap stay unsorted and are
// scanned.  seg[l] = (first pair, one past the last | 2^31 when sorted).
constexpr int kPairSortCap = 4096;
__global__ __launch_bounds__(256) void k_pair_sort(const int* __restrict__ poff, int M, uint32_t* __restrict__ pairs,
                                                   int2* __restrict__ seg) {
    __shared__ uint32_t keys[kPairSortCap];
    const int l = blockIdx.x;
    const int q0 = poff[(size_t)l * M], q1 = poff[(size_t)(l + 1) * M];
    const int len = q1 - q0;
    if (len > kPairSortCap) {
        if (threadIdx.x == 0) seg[l] = make_int2(q0, q1);
        return;
    }
    int np = 1;
    while (np < len) np <<= 1;
    for (int i = threadIdx.x; i < np; i += blockDim.x) keys[i] = i < len ? pairs[q0 + i] : 0xffffffffu;
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t x = keys[i], y = keys[ixj];
                    if ((x > y) == ((i & k) == 0)) { keys[i] = y; keys[ixj] = x; }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < len; i += blockDim.x) pairs[q0 + i] = keys[i];
    if (threadIdx.x == 0) seg[l] = make_int2(q0, (int)((uint32_t)q1 | 0x80000000u));
}

// The mask words of entry (a, s) with list l: mine = the lists of a's
// slots, bm = the bitmap of those lists, hs = hub_id[a]; full[entry] = 1 when
// the mask is the full one (the grouping marks those entries).
__device__ __forceinline__ void first_mask_entry(int l, const int* mine, const uint32_t* bm, int s, int S, int P,
                                                 int M, int W, const int2* __restrict__ seg,
                                                 const uint32_t* __restrict__ pairs,
                                                 const uint32_t* __restrict__ inter, long long a, int hs,
                                                 const int* __restrict__ hub_ptr, const int* __restrict__ hub_mem,
                                                 uint32_t* __restrict__ F, uint8_t* __restrict__ full) {
    uint32_t m[kMaxW];
#pragma unroll
    for (int w = 0; w < kMaxW; ++w) {
        const int nb = M - w * 32;
        m[w] = w >= W ? 0u : (nb >= 32 ? 0xffffffffu : ((1u << nb) - 1u));
    }
    if (s > 0) {
        // positions of l whose point is also in another list L: a duplicate
        // when L is one of the earlier slots' lists
        const int2 sg = seg[l];
        const int q0 = sg.x, q1 = sg.y & 0x7fffffff;
        if (q1 - q0 > 32 && inter) {
            // long segment with the dense intersections: one masked load per
            // earlier slot's list, independent of the others
            const int L2 = 2 * P;
            for (int s2 = 0; s2 < s; ++s2) {
                const uint32_t* iw = inter + ((size_t)l * L2 + mine[s2]) * W;
                if (W == 4) {
                    const uint4 v = *(const uint4*)iw;
                    m[0] &= ~v.x; m[1] &= ~v.y; m[2] &= ~v.z; m[3] &= ~v.w;
                } else {
#pragma unroll
                    for (int x = 0; x < kMaxW; ++x)
                        if (x < W) m[x] &= ~iw[x];
                }
            }
        } else if (q1 - q0 > 32 && sg.y < 0) {
            // long segment, sorted by (L, position): one lower bound per
            // earlier slot's list
            for (int s2 = 0; s2 < s; ++s2) {
                const uint32_t key = (uint32_t)mine[s2] << 8;
                int lo = q0, hi = q1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (pairs[mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                for (int q = lo; q < q1; ++q) {
                    const uint32_t pr = pairs[q];
                    if ((pr >> 8) != (uint32_t)mine[s2]) break;
                    const int i = (int)(pr & 255u);
#pragma unroll
                    for (int x = 0; x < kMaxW; ++x)
                        if (x == (i >> 5)) m[x] &= ~(1u << (i & 31));
                }
            }
        } else {
            // short segment: a pair's other list is looked up in the point's
            // bitmap first; only lists of the point are searched among the
            // earlier slots
            for (int q = q0; q < q1; ++q) {
                const uint32_t pr = pairs[q];
                const int L = (int)(pr >> 8), i = (int)(pr & 255u);
                if (!((bm[L >> 5] >> (L & 31)) & 1u)) continue;
                bool dup = false;
                for (int s2 = 0; s2 < s; ++s2) dup |= mine[s2] == L;
                if (dup) {
#pragma unroll
                    for (int x = 0; x < kMaxW; ++x)
                        if (x == (i >> 5)) m[x] &= ~(1u << (i & 31));
                }
            }
        }
    }
    if (hs >= 0) {
        for (int q = hub_ptr[hs]; q < hub_ptr[hs + 1]; ++q) {
            const int li = hub_mem[q];
            if (li / M == l) {
                const int i = li % M;
#pragma unroll
                for (int w = 0; w < kMaxW; ++w)
                    if (w == (i >> 5)) m[w] &= ~(1u << (i & 31));
            }
        }
    }
    // full masks are only flagged: the tiles never read F for a flagged
    // entry, so its words are not stored
    bool all = true;
#pragma unroll
    for (int w = 0; w < kMaxW; ++w)
        if (w < W) all &= m[w] == fmask_full_word(w, M);
    full[(size_t)a * S + s] = all ? 1 : 0;
    if (all) return;
    uint32_t* Fo = F + ((size_t)a * S + s) * W;
    if (W == 4) {
        *(uint4*)Fo = uint4{m[0], m[1], m[2], m[3]};
    } else {
#pragma unroll
        for (int w = 0; w < kMaxW; ++w)
            if (w < W) Fo[w] = m[w];
    }
}

__global__ __launch_bounds__(512) void k_first_mask(const int* __restrict__ topK, long long n, int P, int K,
                                                    int M, int W, int ppb, const uint32_t* __restrict__ pairs,
                                                    const int2* __restrict__ seg,
                                                    const int* __restrict__ hub_id,
                                                    const int* __restrict__ hub_ptr,
                                                    const int* __restrict__ hub_mem,
                                                    const uint32_t* __restrict__ inter,
                                                    uint32_t* __restrict__ F, uint8_t* __restrict__ full) {
    // LDS: the slot lists of the workgroup's points, then per point a bitmap
    // over the 2P list ids with the bits of its own slot lists set
    extern __shared__ int lids[];
    const int S = 2 * K, BW = (2 * P + 31) >> 5;
    uint32_t* bm = (uint32_t*)(lids + ppb * S);
    for (int i = threadIdx.x; i < ppb * BW; i += blockDim.x) bm[i] = 0u;
    const int tid = threadIdx.x;
    const int pl = tid / S, s = tid - pl * S;
    const long long a = (long long)blockIdx.x * ppb + pl;
    const bool live = pl < ppb && a < n;
    const int l = live ? slot_list(topK, a, s, K, P) : 0;
    const int hs = live ? hub_id[a] : -1;
    if (live) lids[tid] = l;
    __syncthreads();
    if (live) atomicOr(&bm[pl * BW + (l >> 5)], 1u << (l & 31));
    __syncthreads();
    if (!live) return;
    first_mask_entry(l, lids + pl * S, bm + pl * BW, s, S, P, M, W, seg, pairs, inter, a, hs, hub_ptr, hub_mem, F,
                     full);
}

// Grouping of the (point, slot) entries by (query block, list) without
// contended global atomics.  Query block y holds the points
// [y * qblk, (y + 1) * qblk); group y * L2 + l lists its entries that visit
// list l.  Workgroup (x, y) counts a contiguous part of block y's entries per
// list in LDS (k_group_count); the counts are scanned in (y, l, x) order and
// the workgroup scatters its entries from LDS cursors (k_group_scatter).
__device__ __forceinline__ void group_range(long long n, int K, long long qblk, int Gx, long long& t0,
                                            long long& t1) {
    const long long S = 2 * K;
    const long long b0 = (long long)blockIdx.y * qblk * S;
    const long long b1 = min((long long)(blockIdx.y + 1) * qblk, n) * S;
    const long long per = (b1 - b0 + Gx - 1) / Gx;
    t0 = b0 + (long long)blockIdx.x * per;
    t1 = min(t0 + per, b1);
}

__global__ __launch_bounds__(256) void k_group_count(const int* __restrict__ topK, long long n, int P,
                                                     int K, int L2, long long qblk, int Gx,
                                                     int* __restrict__ wgcnt) {
    extern __shared__ int hist[];
    long long t0, t1;
    group_range(n, K, qblk, Gx, t0, t1);
    for (int i = threadIdx.x; i < L2; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (long long t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
        const uint32_t a = (uint32_t)t / (uint32_t)(2 * K);  // n * 2K < 2^31 (sdb_set_params)
        const int s = (int)((uint32_t)t - a * (uint32_t)(2 * K));
        atomicAdd(&hist[slot_list(topK, a, s, K, P)], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L2; i += blockDim.x)
        wgcnt[((size_t)blockIdx.y * L2 + i) * Gx + blockIdx.x] = hist[i];
}

__global__ __launch_bounds__(256) void k_group_scatter(const int* __restrict__ topK, long long n, int P,
                                                       int K, int L2, long long qblk, int Gx,
                                                       const int* __restrict__ wgoff,
                                                       const uint8_t* __restrict__ full,
                                                       uint32_t* __restrict__ grp) {
    extern __shared__ int cur[];
    long long t0, t1;
    group_range(n, K, qblk, Gx, t0, t1);
    for (int i = threadIdx.x; i < L2; i += blockDim.x)
        cur[i] = wgoff[((size_t)blockIdx.y * L2 + i) * Gx + blockIdx.x];
    __syncthreads();
    // the entries in the order of k_group_count's pass (the order inside a
    // group changes how the distance tiles hit the caches: measured), with the
    // next entry's loads issued before the current entry's atomic
    auto load = [&](long long t, int& li, uint32_t& gv) {
        const uint32_t a = (uint32_t)t / (uint32_t)(2 * K);  // n * 2K < 2^31 (sdb_set_params)
        const int s = (int)((uint32_t)t - a * (uint32_t)(2 * K));
        li = slot_list(topK, a, s, K, P);
        gv = (uint32_t)t | (full[t] ? kEntFull : 0u);
    };
    long long t = t0 + threadIdx.x;
    int li = 0;
    uint32_t gv = 0u;
    if (t < t1) load(t, li, gv);
    while (t < t1) {
        const long long tn = t + blockDim.x;
        int lin = 0;
        uint32_t gvn = 0u;
        if (tn < t1) load(tn, lin, gvn);
        const int pos = atomicAdd(&cur[li], 1);
        grp[pos] = gv;
        t = tn;
        li = lin;
        gv = gvn;
    }
}

__global__ void k_group_ptr(const int* __restrict__ wgoff, int NG, int Gx, long long tot,
                            int* __restrict__ grp_ptr) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < NG) grp_ptr[g] = wgoff[(size_t)g * Gx];
    if (g == NG) grp_ptr[NG] = (int)tot;
}

// ---------------------------------------------------------------------------
// Host launchers
int launch_rff_embed(const float* Xn, long long n, int d, int ldx, const float* R, int half, float* E, int ldE,
                     hipStream_t st) {
    if (n <= 0 || half <= 0) return 0;
    if (2 * half > ldE) return -1;
    const long long gx = (n + 63) / 64;
    if (gx > 0x7fffffffll) return -3;
    hipLaunchKernelGGL(k_rff_embed, dim3((unsigned)gx, (unsigned)((half + 63) / 64)), dim3(256), 0, st, Xn, n, d,
                       ldx, R, half, E, ldE);
    return (int)hipGetLastError();
}

int launch_transform(const float* X, long long n, int d, int ldx, int metric, float* Xn,
                     uint16_t* Xb, float* rowerr, uint32_t* rowerr_max, int rho_pad, int* err, hipStream_t st) {
    if (Xb && (!rowerr || !rowerr_max)) return -3;
    if (rho_pad && d >= ldx) return -3;
    long long blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_transform, dim3((unsigned)blocks), dim3(256), 0, st, X, n, d, ldx, metric, Xn, Xb, rowerr,
                       rowerr_max, rho_pad, err);
    return (int)hipGetLastError();
}

template <int LOGD, int MODE, int EMB, bool FULLP>
static int launch_index_f(const IndexArgs& a, int grid, hipStream_t st) {
    constexpr int D = 1 << LOGD;
    constexpr bool SGN_LDS = (D / 64) % 4 == 0 && LOGD <= 11;
    size_t sm = (MODE == kFull ? (size_t)4 * 2048 + (size_t)8 * D : (size_t)8 * a.P) + (EMB == kEmbLds ? (size_t)4 * D * 4 : 0) +
                (SGN_LDS ? (size_t)3 * D * 4 + 16 : 0);
    if (sm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_index<LOGD, MODE, EMB, FULLP>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL((k_index<LOGD, MODE, EMB, FULLP>), dim3(grid), dim3(256), sm, st, a);
    return (int)hipGetLastError();
}

template <int LOGD, int MODE, int EMB>
static int launch_index_t(const IndexArgs& a, int grid, hipStream_t st) {
    // the bulk passes with n_proj == fhtDim take the folded variant
    if constexpr (MODE == kSample || MODE == kFull)
        if (a.P == (1 << LOGD)) return launch_index_f<LOGD, MODE, EMB, true>(a, grid, st);
    return launch_index_f<LOGD, MODE, EMB, false>(a, grid, st);
}

template <int LOGD, int MODE>
static int launch_index_m(const IndexArgs& a, int grid, hipStream_t st) {
    switch (index_embedding(a.metric, a.d, a.ker)) {
        case kEmbCos: return launch_index_t<LOGD, MODE, kEmbCos>(a, grid, st);
        case kEmbSqrt: return launch_index_t<LOGD, MODE, kEmbSqrt>(a, grid, st);
        default:
            // the bulk passes read the precomputed L1/L2 embeddings; the
            // replays of a few rows or points compute them in place
            if constexpr (MODE == kSample || MODE == kFull)
                if (a.emb && (a.metric == kL1 || a.metric == kL2)) return launch_index_t<LOGD, MODE, kEmbRows>(a, grid, st);
            return launch_index_t<LOGD, MODE, kEmbLds>(a, grid, st);
    }
}

template <int LOGD>
static int launch_index_d(const IndexArgs& a, int mode, int grid, hipStream_t st) {
    if (mode == kSample) return launch_index_m<LOGD, kSample>(a, grid, st);
    if (mode == kFull) return launch_index_m<LOGD, kFull>(a, grid, st);
    if (mode == kRows) return launch_index_m<LOGD, kRows>(a, grid, st);
    return launch_index_m<LOGD, kPoints>(a, grid, st);
}

int index_embedding(int metric, int d, int ker) {
    if (metric == kCosine) return kEmbCos;
    if ((metric == kChi2 || metric == kJS) && ker < 3 * d) return kEmbSqrt;
    return kEmbLds;
}

int launch_index(const IndexArgs& a, int logD, int mode, int grid, hipStream_t st) {
    if (a.rho_pad && (a.d >= a.ldx || !a.Xb_out)) return -3;
    switch (logD) {
        case 6: return launch_index_d<6>(a, mode, grid, st);
        case 7: return launch_index_d<7>(a, mode, grid, st);
        case 8: return launch_index_d<8>(a, mode, grid, st);
        case 9: return launch_index_d<9>(a, mode, grid, st);
        case 10: return launch_index_d<10>(a, mode, grid, st);
        case 11: return launch_index_d<11>(a, mode, grid, st);
        case 12: return launch_index_d<12>(a, mode, grid, st);
        default: return -3;
    }
}


// rows[i] (width ints) -> dst[idx[i] * width]: the host-replayed top-K rows
__global__ __launch_bounds__(256) void k_scatter_rows(const int* __restrict__ idx, const int* __restrict__ rows,
                                                      int count, int width, int* __restrict__ dst) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)count * width) return;
    const int i = (int)(t / width), w = (int)(t - (long long)i * width);
    dst[(size_t)idx[i] * width + w] = rows[t];
}

int launch_scatter_rows(const int* idx, const int* rows, int count, int width, int* dst, hipStream_t st) {
    if (count <= 0) return 0;
    const long long tot = (long long)count * width;
    hipLaunchKernelGGL(k_scatter_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, idx, rows, count, width,
                       dst);
    return (int)hipGetLastError();
}

int launch_topm_threshold(const uint32_t* cmax, const uint32_t* cmin, int G, int P, int M,
                          uint32_t* thi, uint32_t* tlo, hipStream_t st) {
    hipLaunchKernelGGL(k_topm_threshold, dim3(P), dim3(256), 0, st, cmax, cmin, G, P, M, thi, tlo);
    return (int)hipGetLastError();
}

int launch_topm_select(const uint32_t* key, const int* idx, const int* cnt, int cap, int P, int M,
                       int* topM, int* flags, hipStream_t st) {
    size_t sm = 0;
    int sz = 1;
    while (sz < cap) sz <<= 1;
    sm = (size_t)sz * 8;
    if (sm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_topm_select,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(k_topm_select, dim3(2 * P), dim3(1024), sm, st, key, idx, cnt, cap, P, M, topM, flags);
    return (int)hipGetLastError();
}

int launch_build_lists(const int* topM, int P, int M, int* lists, uint8_t* is_hub, hipStream_t st) {
    int tot = 2 * P * M;
    hipLaunchKernelGGL(k_build_lists, dim3((tot + 255) / 256), dim3(256), 0, st, topM, P, M, lists, is_hub);
    return (int)hipGetLastError();
}

int launch_hub_count(const int* lists, int LM, const int* hub_id, int* hub_cnt, hipStream_t st) {
    hipLaunchKernelGGL(k_hub_count, dim3((LM + 255) / 256), dim3(256), 0, st, lists, LM, hub_id, hub_cnt);
    return (int)hipGetLastError();
}

int launch_hub_fill(const int* lists, int LM, const int* hub_id, const int* hub_ptr, int* hub_cur,
                    int* hub_mem, hipStream_t st) {
    hipLaunchKernelGGL(k_hub_fill, dim3((LM + 255) / 256), dim3(256), 0, st, lists, LM, hub_id, hub_ptr, hub_cur, hub_mem);
    return (int)hipGetLastError();
}

int launch_invmask(const int* lists, const int* topK, int P, int K, int M, int L2, int W,
                   uint32_t* inv, hipStream_t st) {
    long long tot = (long long)L2 * M * 2 * K;
    hipLaunchKernelGGL(k_invmask, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, lists, topK, P, K, M, L2, W, inv);
    return (int)hipGetLastError();
}

int launch_pair_count(const int* lists, const int* hub_id, const int* hub_ptr, int L2, int M, int* pos_hub,
                      int* pcnt, int* poff, int* total, void* tmp, size_t tmp_bytes, hipStream_t st) {
    const int LM = L2 * M;
    hipLaunchKernelGGL(k_pos_hub, dim3((LM + 1 + 255) / 256), dim3(256), 0, st, lists, hub_id, hub_ptr, LM, pos_hub,
                       pcnt);
    return scan_exclusive_i32(pcnt, poff, (long long)LM + 1, total, tmp, tmp_bytes, st);
}

int launch_pair_fill(const int* pos_hub, const int* hub_ptr, const int* hub_mem, int L2, int M, const int* poff,
                     uint32_t* pairs, hipStream_t st) {
    const int LM = L2 * M;
    hipLaunchKernelGGL(k_pair_fill, dim3((LM + 255) / 256), dim3(256), 0, st, pos_hub, hub_ptr, hub_mem, LM, M, poff,
                       pairs);
    return (int)hipGetLastError();
}

int launch_inter_fill(const int* pos_hub, const int* hub_ptr, const int* hub_mem, int L2, int M, int W,
                      uint32_t* inter, hipStream_t st) {
    const int LM = L2 * M;
    hipLaunchKernelGGL(k_inter_fill, dim3((LM + 255) / 256), dim3(256), 0, st, pos_hub, hub_ptr, hub_mem, LM, M, L2, W,
                       inter);
    return (int)hipGetLastError();
}

int launch_first_mask(const int* topK, long long n, int P, int K, int M, int W, const int* poff,
                      uint32_t* pairs, int2* seg, const int* hub_id, const int* hub_ptr, const int* hub_mem,
                      const uint32_t* inter, uint32_t* F, uint8_t* full, hipStream_t st) {
    const int S = 2 * K;
    hipLaunchKernelGGL(k_pair_sort, dim3(2 * P), dim3(256), 0, st, poff, M, pairs, seg);
    const int ppb = std::max(1, 512 / S);
    const long long blocks = (n + ppb - 1) / ppb;
    // tables in LDS when they take at most 48 KB (with the slot lists): at
    // least three workgroups of 512 threads per CU
    const size_t lds = (size_t)ppb * S * 4 + (size_t)ppb * ((2 * P + 31) / 32) * 4;
    hipLaunchKernelGGL(k_first_mask, dim3((unsigned)blocks), dim3(ppb * S), lds, st, topK, n, P, K, M, W, ppb, pairs,
                       seg, hub_id, hub_ptr, hub_mem, inter, F, full);
    return (int)hipGetLastError();
}

int launch_group_count(const int* topK, long long n, int P, int K, int L2, long long qblk, int nblk,
                       int Gx, int* wgcnt, hipStream_t st) {
    hipLaunchKernelGGL(k_group_count, dim3(Gx, nblk), dim3(256), (size_t)L2 * 4, st, topK, n, P, K, L2, qblk, Gx, wgcnt);
    return (int)hipGetLastError();
}

int launch_group_scatter(const int* topK, long long n, int P, int K, int L2, long long qblk, int nblk,
                         int Gx, const int* wgoff, const uint8_t* full, uint32_t* grp, int* grp_ptr, hipStream_t st) {
    hipLaunchKernelGGL(k_group_scatter, dim3(Gx, nblk), dim3(256), (size_t)L2 * 4, st, topK, n, P, K, L2, qblk, Gx, wgoff,
                       full, grp);
    const int NG = nblk * L2;
    hipLaunchKernelGGL(k_group_ptr, dim3((NG + 1 + 255) / 256), dim3(256), 0, st, wgoff, NG, Gx, n * 2 * K, grp_ptr);
    return (int)hipGetLastError();
}

}  // namespace sdb
