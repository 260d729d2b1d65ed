// Device side of the batched eps sweep (sdb_batch_sweep).
//
// Reference: test_sDbscan (src/sDbscan.cpp:980-1039) sweeps eps over one
// index and recomputes the candidate distances for every eps
// (rp_findCorePoints, :217-408).  Here the candidate distances of every
// dataset of a batch are computed once: k_sweep_tiles is one persistent
// launch whose workgroups take (dataset, list, row range) tiles from a device
// work queue (an atomic head over the tile array) until it is drained, and
// decides every eps of the tile's dataset from the same FP32 MFMA dot
// products.  Pairs inside the FP32 error band of an eps are queued with that
// eps and decided in FP64 by k_sweep_recheck, so every eps gets the hit masks
// and column counts a single-eps pass would give.
#include "sdb_kernels.hpp"

namespace sdb {

typedef float f32x4_s __attribute__((ext_vector_type(4)));

// FP32 MFMA tiles (4 x v_mfma_f32_16x16x4_f32 per 16-element k-step) over a
// list's candidate rows staged in LDS, as the single-eps FP32 tiles: lane
// (r16, g) holds the dots of query rows 4g..4g+3 of its 16-row block with
// candidate column t*16 + r16.  QT blocks of 16 query rows per wave and set.
// Epilogue per eps e (ascending): hit iff s > s_hit[e]; queued iff not a hit
// and s >= s_near[e] (the cuts of the band test on fl(1 - s)); hits go to the
// eps's E slice (the first-occurrence mask F applied) and its column counts.
template <int NT>
__global__ __launch_bounds__(512) void k_sweep_tiles(const SweepJob* __restrict__ jobs, const int4* __restrict__ tiles,
                                                     int ntiles, int* __restrict__ head, uint4* __restrict__ queue,
                                                     unsigned long long* __restrict__ qcount, long long qcap) {
    constexpr int WAVES = 8, QT = 2, SE = 16, EW = (NT + 1) / 2;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_tile;
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    for (;;) {
        // every thread is done with the previous tile (and its s_tile read)
        __syncthreads();
        if (tid == 0) s_tile = atomicAdd(head, 1);
        __syncthreads();
        const int ti = __builtin_amdgcn_readfirstlane(s_tile);
        if (ti >= ntiles) break;  // the same value in every wave: all leave together
        const int4 T = tiles[ti];
        const int job = __builtin_amdgcn_readfirstlane(T.x), l = __builtin_amdgcn_readfirstlane(T.y);
        const int beg = __builtin_amdgcn_readfirstlane(T.z), end = __builtin_amdgcn_readfirstlane(T.w);
        const SweepJob* J = jobs + job;
        const int M = J->M, W = J->W, S = J->S, ne = J->n_eps, ldx = J->ldx;
        const int kc = J->kc, nchunk = J->nchunk, dk = J->dk;
        const size_t RSB = (size_t)kc * 4 + 16;
        int* col_lds = (int*)(smem + (size_t)M * RSB);  // ne x M column counts
        int* pts_lds = col_lds + ne * M;
        const int* Lm = J->lists + (size_t)l * M;
        const char* Xc = (const char*)J->X;
        if (tid < M) pts_lds[tid] = Lm[tid];  // M <= 128
        for (int i = tid; i < ne * M; i += WAVES * 64) col_lds[i] = 0;

        const int rows_per_set = WAVES * QT * 16;
        const int nsets = (end - beg + rows_per_set - 1) / rows_per_set;
        for (int set = 0; set < nsets; ++set) {
            uint32_t ent[QT];
            long long pa[QT];
            bool valid[QT], ffl[QT];
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                const int rr = beg + set * rows_per_set + (w * QT + q) * 16 + r16;
                valid[q] = rr < end;
                const uint32_t gq = J->grp[valid[q] ? rr : beg];
                ent[q] = ent_id(gq);
                ffl[q] = ent_full(gq);
                pa[q] = (long long)ent_point(ent[q], J->magicS);
            }
            f32x4_s acc[QT][NT];
#pragma unroll
            for (int q = 0; q < QT; ++q)
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[q][t] = f32x4_s{0.f, 0.f, 0.f, 0.f};

            for (int c = 0; c < nchunk; ++c) {
                const int k0 = c * kc;
                const int kn = min(kc, dk - k0);
                if (nchunk > 1 || set == 0) {
                    // candidate rows of this k-chunk -> LDS (the first pass also
                    // publishes pts_lds and the cleared counts)
                    __syncthreads();
                    const int vpr = kn / 4, total = M * vpr;
                    for (int v = tid; v < total; v += WAVES * 64) {
                        const int row = v / vpr, cv = v - row * vpr;
                        const uint4 x = ((const uint4*)(Xc + ((size_t)pts_lds[row] * ldx + k0) * 4))[cv];
                        *(uint4*)(smem + row * RSB + (size_t)cv * 16) = x;
                    }
                    __syncthreads();
                }
                const int nsteps = kn / SE;
                const char* arow[QT];
#pragma unroll
                for (int q = 0; q < QT; ++q) arow[q] = Xc + ((size_t)pa[q] * ldx + k0) * 4 + g * 16;
                uint4 af[QT];
#pragma unroll
                for (int q = 0; q < QT; ++q) af[q] = *(const uint4*)(arow[q]);
                for (int st = 0; st < nsteps; ++st) {
                    uint4 an[QT];
                    const int ns = min(st + 1, nsteps - 1);
#pragma unroll
                    for (int q = 0; q < QT; ++q) an[q] = *(const uint4*)(arow[q] + (size_t)ns * 64);
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const int row = min(t * 16 + r16, M - 1);
                        const uint4 bfr = *(const uint4*)(smem + row * RSB + (size_t)st * 64 + g * 16);
#pragma unroll
                        for (int q = 0; q < QT; ++q) {
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(af[q].x), __uint_as_float(bfr.x), acc[q][t], 0, 0, 0);
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(af[q].y), __uint_as_float(bfr.y), acc[q][t], 0, 0, 0);
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(af[q].z), __uint_as_float(bfr.z), acc[q][t], 0, 0, 0);
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(af[q].w), __uint_as_float(bfr.w), acc[q][t], 0, 0, 0);
                        }
                    }
#pragma unroll
                    for (int q = 0; q < QT; ++q) af[q] = an[q];
                }
            }

            // epilogue: the F bits and the entries of the lane's 4 rows do
            // not depend on eps
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                uint32_t Fw[EW];
#pragma unroll
                for (int x = 0; x < EW; ++x)
                    Fw[x] = (valid[q] && x < W) ? (ffl[q] ? fmask_full_word(x, M) : J->F[(size_t)ent[q] * W + x]) : 0u;
                uint32_t act = 0u;  // bit 4t + i: pair (row 4g + i, column t*16 + r16) is evaluated
                uint32_t er[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 4 * g + i;
                    er[i] = __shfl(ent[q], row, kWave);
                    const int vr = __shfl((int)valid[q], row, kWave);
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const uint32_t fword = __shfl(Fw[t >> 1], row, kWave);
                        const int col = t * 16 + r16;
                        if (vr && col < M && ((fword >> (col & 31)) & 1u)) act |= 1u << (4 * t + i);
                    }
                }
#pragma unroll 1
                for (int e = 0; e < ne; ++e) {
                    const float sh = J->s_hit[e], sn = J->s_near[e];
                    uint32_t Ew[EW];
#pragma unroll
                    for (int x = 0; x < EW; ++x) Ew[x] = 0u;
                    uint32_t unc = 0u;
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        int nh = 0;
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const bool a = (act >> (4 * t + i)) & 1u;
                            const float s = acc[q][t][i];
                            const bool hit = a && s > sh;
                            if (a && !hit && s >= sn) unc |= 1u << (4 * t + i);
                            const unsigned long long bh = __ballot(hit);
                            if ((r16 & 3) == i)
                                Ew[t >> 1] |= (uint32_t)((bh >> (16 * (r16 >> 2))) & 0xffffull) << (16 * (t & 1));
                            nh += hit ? 1 : 0;
                        }
                        nh += __shfl_xor(nh, 16, kWave);
                        nh += __shfl_xor(nh, 32, kWave);
                        const int col = t * 16 + r16;
                        if (g == 0 && col < M && nh) atomicAdd(&col_lds[e * M + col], nh);
                    }
                    if (g == 0 && valid[q]) {
                        uint32_t* Ee = J->E + (size_t)e * J->n_ent * W + (size_t)ent[q] * W;
#pragma unroll
                        for (int x = 0; x < EW; ++x)
                            if (x < W) Ee[x] = Ew[x];
                    }
                    if (__any(unc != 0u)) {
                        // one queue reservation per wave
                        int incl = __popc(unc);
                        const int cnt = incl;
#pragma unroll
                        for (int o = 1; o < kWave; o <<= 1) {
                            const int y = __shfl_up(incl, o, kWave);
                            if (lane >= o) incl += y;
                        }
                        unsigned long long base = 0;
                        if (lane == kWave - 1 && incl > 0) base = atomicAdd(qcount, (unsigned long long)incl);
                        const uint32_t blo = __shfl((uint32_t)base, kWave - 1, kWave);
                        const uint32_t bhi = __shfl((uint32_t)(base >> 32), kWave - 1, kWave);
                        unsigned long long pos = (((unsigned long long)bhi << 32) | blo) + (unsigned long long)(incl - cnt);
                        while (unc) {
                            const int bp = __ffs(unc) - 1;
                            unc &= unc - 1;
                            const int col = (bp >> 2) * 16 + r16, i = bp & 3;
                            if (pos < (unsigned long long)qcap)
                                queue[pos] = uint4{er[i], (uint32_t)job, (uint32_t)l, (uint32_t)col | ((uint32_t)e << 16)};
                            ++pos;
                        }
                    }
                }
            }
        }
        // the tile's column counts of every eps
        __syncthreads();
        for (int i = tid; i < ne * M; i += WAVES * 64) {
            const int v = col_lds[i];
            if (v) {
                const int e = i / M, col = i - e * M;
                atomicAdd(&J->colsum[((size_t)e * J->L2 + l) * M + col], v);
            }
        }
    }
}

// FP64 decision of the queued (pair, eps) entries, one wavefront per entry.
// stats[0]: entries, stats[1]: hits added.
__global__ __launch_bounds__(256) void k_sweep_recheck(const SweepJob* __restrict__ jobs,
                                                       const uint4* __restrict__ queue, long long nq,
                                                       unsigned long long* __restrict__ stats) {
    const int lane = lane_id();
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwv = ((long long)gridDim.x * blockDim.x) >> 6;
    unsigned long long cnt = 0, hits = 0;
    for (long long k = wv; k < nq; k += nwv) {
        const uint4 qe = queue[k];
        const SweepJob* J = jobs + qe.y;
        const uint32_t ent = qe.x;
        const int l = (int)qe.z, col = (int)(qe.w & 0xffffu), e = (int)(qe.w >> 16);
        const long long pa = (long long)ent_point(ent, J->magicS);
        const int c = J->lists[(size_t)l * J->M + col];
        const float* x = J->X + (size_t)pa * J->ldx;
        const float* y = J->X + (size_t)c * J->ldx;
        const double d64 = dist64_wave(x, y, J->d, kCosine, lane);
        ++cnt;
        if (d64 <= (double)J->eps[e]) {
            ++hits;
            if (lane == 0) {
                atomicOr(&J->E[(size_t)e * J->n_ent * J->W + (size_t)ent * J->W + (col >> 5)], 1u << (col & 31));
                atomicAdd(&J->colsum[((size_t)e * J->L2 + l) * J->M + col], 1);
            }
        }
    }
    if (lane == 0) {
        if (cnt) atomicAdd(&stats[0], cnt);
        if (hits) atomicAdd(&stats[1], hits);
    }
}

// Exclusive prefix sum of int32 in three launches without library calls or
// host round trips (graph-capturable): per-block sums of 1024 elements, one
// workgroup scanning those sums (and writing the total), then every block
// scanning its elements from its offset.
__global__ __launch_bounds__(1024) void k_scan_blocks(const int* __restrict__ in, long long n, int* __restrict__ bsum) {
    __shared__ int wsum[16];
    __shared__ int tot;
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    (void)block_scan_excl_1024(i < n ? in[i] : 0, wsum, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_top(int* __restrict__ bsum, int nb, int* __restrict__ total) {
    __shared__ int wsum[16];
    __shared__ int tot;
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int i = b0 + (int)threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        const int ex = block_scan_excl_1024(v, wsum, &tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void k_scan_apply(const int* __restrict__ in, int* __restrict__ out, long long n,
                                                     const int* __restrict__ bsum) {
    __shared__ int wsum[16];
    __shared__ int tot;
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    const int ex = block_scan_excl_1024(i < n ? in[i] : 0, wsum, &tot);
    if (i < n) out[i] = bsum[blockIdx.x] + ex;
}

int scan_exclusive_i32_graph(const int* in, int* out, long long n, int* total, int* bsum, hipStream_t st) {
    if (n <= 0) return 0;
    const long long nb = (n + 1023) / 1024;
    if (nb > 0x7fffffffll) return -3;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(1024), 0, st, in, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, (int)nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(1024), 0, st, in, out, n, bsum);
    return (int)hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, long long count, uint32_t value) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

__global__ void k_copy_i32(const int* src, int* dst, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i];
}

template <int NT>
static int launch_sweep_nt(const SweepJob* jobs, const int4* tiles, int ntiles, int* head, uint4* queue,
                           unsigned long long* qcount, long long qcap, size_t lds, int grid, hipStream_t st) {
    if (lds > (size_t)kLdsBytes) return -2;
    hipError_t e = hipFuncSetAttribute((const void*)k_sweep_tiles<NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_sweep_tiles<NT>, dim3(grid), dim3(512), lds, st, jobs, tiles, ntiles, head, queue, qcount,
                       qcap);
    return (int)hipGetLastError();
}

int launch_sweep_tiles(const SweepJob* jobs, int maxM, const int4* tiles, int ntiles, int* head, uint4* queue,
                       unsigned long long* qcount, long long qcap, size_t lds_bytes, int grid, hipStream_t st) {
    if (ntiles <= 0) return 0;
    switch ((maxM + 15) / 16) {
        case 1: return launch_sweep_nt<1>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        case 2: return launch_sweep_nt<2>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        case 3: return launch_sweep_nt<3>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        case 4: return launch_sweep_nt<4>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        case 5: return launch_sweep_nt<5>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        case 6: return launch_sweep_nt<6>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        case 7: return launch_sweep_nt<7>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        case 8: return launch_sweep_nt<8>(jobs, tiles, ntiles, head, queue, qcount, qcap, lds_bytes, grid, st);
        default: return -3;
    }
}

int launch_sweep_recheck(const SweepJob* jobs, const uint4* queue, long long nq, unsigned long long* stats,
                         hipStream_t st) {
    if (nq <= 0) return 0;
    long long blocks = (nq + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_sweep_recheck, dim3((unsigned)blocks), dim3(256), 0, st, jobs, queue, nq, stats);
    return (int)hipGetLastError();
}

int launch_fill_u32(uint32_t* p, long long count, uint32_t value, hipStream_t st) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count, value);
    return (int)hipGetLastError();
}

int launch_copy_i32(const int* src, int* dst, long long count, hipStream_t st) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_copy_i32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, src, dst, count);
    return (int)hipGetLastError();
}

}  // namespace sdb
