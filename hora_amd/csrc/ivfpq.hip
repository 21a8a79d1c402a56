// ivfpq.hip -- IVFPQIndex kernels: k-means steps, residual PQ training,
// LUT + ADC inverted-list scan, per-query merge and the tie replay.
//
// Reference: src/core/kmeans.rs (Kmeans), src/index/pq_idx.rs (PQIndex,
// IVFPQIndex).  Every distance that decides an assignment, a code or a
// search result is computed with the reference's f32 left fold (no FMA), so
// assignments, codes and ADC distances equal the reference's given the same
// centroids and codebooks.
#include <hipcub/hipcub.hpp>

#include "ivfpq.h"
#include "pq_common.h"
#include "profiler.h"
#include "wave_heap.h"

#pragma clang fp contract(off)

namespace hm {

namespace {

__device__ __forceinline__ uint32_t fkey(float f)
{
    f = f + 0.0f;
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float unkey(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(b);
}

// kmeans.rs:50-86: mean over the rows (left fold in row order, f32 or f64),
// then centre i = mean + init_offset(i, j).  One lane per dimension.
template <typename ACC_T>
__global__ void coarse_init_kernel(const float *__restrict__ X, uint32_t ldx, uint64_t n,
                                   uint32_t d, uint32_t nc, float *__restrict__ C, uint32_t ldc)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ldc) return;
    if (j >= d) {
        for (uint32_t c = 0; c < nc; ++c) C[(size_t)c * ldc + j] = 0.0f;
        return;
    }
    ACC_T acc = 0;
    uint64_t i = 0;
    for (; i + 8 <= n; i += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = X[(i + u) * ldx + j];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + (ACC_T)v[u];
    }
    for (; i < n; ++i) acc = acc + (ACC_T)X[i * ldx + j];
    float mean;
    if constexpr (sizeof(ACC_T) == 4) mean = acc / (float)n;
    else mean = (float)(acc / (double)n);
    for (uint32_t c = 0; c < nc; ++c) {
        float v = mean;
        v = init_offset(c, j) > 0 ? v + 1.0f : v - 1.0f;
        C[(size_t)c * ldc + j] = v;
    }
}

__global__ void unpack_assign_kernel(const unsigned long long *__restrict__ best, uint64_t n,
                                     uint32_t *__restrict__ assign)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        assign[i] = (uint32_t)(best[i] & 0xFFFFFFFFull);
}

__global__ void iota_kernel(uint32_t *__restrict__ v, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

// off[c] = first position of key c in the sorted keys (off[nc] = n).
__global__ void offsets_kernel(const uint32_t *__restrict__ sorted, uint64_t n, uint32_t nc,
                               uint32_t *__restrict__ off)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nc) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (sorted[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    off[c] = (uint32_t)lo;
}

// kmeans.rs:88-125: per centre and dimension, the members' values summed in
// point order (f32 or f64), divided by the count; an empty centre is zero.
template <typename ACC_T>
__global__ void update_kernel(const float *__restrict__ X, uint32_t ldx,
                              const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
                              uint32_t d, float *__restrict__ C, uint32_t ldc)
{
    const uint32_t c = blockIdx.x;
    const uint32_t j = blockIdx.y * blockDim.x + threadIdx.x;
    if (j >= ldc) return;
    const uint32_t a = off[c], b = off[c + 1];
    if (j >= d || a == b) {
        C[(size_t)c * ldc + j] = 0.0f;
        return;
    }
    ACC_T acc = 0;
    uint32_t r = a;
    for (; r + 4 <= b; r += 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = X[(size_t)ids[r + u] * ldx + j];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = acc + (ACC_T)v[u];
    }
    for (; r < b; ++r) acc = acc + (ACC_T)X[(size_t)ids[r] * ldx + j];
    const uint32_t cnt = b - a;
    float v;
    if constexpr (sizeof(ACC_T) == 4) v = acc / (float)cnt;
    else v = (float)(acc / (double)cnt);
    C[(size_t)c * ldc + j] = v;
}

// FP64 column sums with a fixed split (the fp64 accumulation mode of the
// centre steps, kmeans.rs:50-86 and 88-125 with f64 sums).  The positions of
// segment c (off[c]..off[c+1] of ids; rows 0..n when off is null) are cut
// into chunks of SEG_ROWS; seg_partial_kernel folds one chunk per workgroup
// (four row groups, each a contiguous quarter of the chunk, combined in
// group order) and seg_mean_kernel folds a segment's chunk sums in chunk
// order.  The split depends only on the segment sizes, so the sums are the
// same on every run; they differ from a point-order f64 fold only in the
// rounding of the f64 additions.
constexpr uint32_t SEG_ROWS = 512;

// nch[c] = chunks of segment c; nch[nseg] = 0 (the exclusive scan of nch
// then gives each segment's first chunk and, at nseg, the total).
__global__ void seg_chunk_count_kernel(const uint32_t *__restrict__ off, uint32_t nseg,
                                       uint32_t *__restrict__ nch)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nseg) return;
    nch[c] = c < nseg ? (off[c + 1] - off[c] + SEG_ROWS - 1) / SEG_ROWS : 0u;
}

__global__ __launch_bounds__(256) void seg_partial_kernel(
    const float *__restrict__ X, uint32_t ldx, const uint32_t *__restrict__ ids,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ cstart, uint32_t nseg,
    uint64_t n, uint32_t d, double *__restrict__ part)
{
    __shared__ double red[4][64];
    const uint32_t q = blockIdx.x;
    uint64_t a, b;
    if (cstart) {
        if (q >= cstart[nseg]) return;
        uint32_t lo = 0, hi = nseg - 1;  // the segment c with cstart[c] <= q < cstart[c+1]
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (cstart[mid + 1] <= q) lo = mid + 1;
            else hi = mid;
        }
        a = off[lo] + (uint64_t)(q - cstart[lo]) * SEG_ROWS;
        b = min((uint64_t)off[lo + 1], a + SEG_ROWS);
    } else {
        a = (uint64_t)q * SEG_ROWS;
        if (a >= n) return;
        b = min(n, a + SEG_ROWS);
    }
    const uint32_t lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const uint32_t j = blockIdx.y * 64 + lane;
    const uint64_t quarter = (b - a + 3) / 4;
    const uint64_t ga = min(b, a + g * quarter), gb = min(b, ga + quarter);
    double acc = 0.0;
    if (j < d) {
        uint64_t r = ga;
        for (; r + 8 <= gb; r += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint64_t row = ids ? ids[r + u] : r + u;
                v[u] = X[row * ldx + j];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + (double)v[u];
        }
        for (; r < gb; ++r) acc = acc + (double)X[(ids ? ids[r] : r) * ldx + j];
    }
    red[g][lane] = acc;
    __syncthreads();
    if (g == 0 && j < d)
        part[(size_t)q * d + j] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Segment c's sum, its chunk sums folded in chunk order (four groups over
// contiguous quarters, combined in group order), divided by its size.
// nc_init == 0: C[c] = the mean (zero for an empty segment).  nc_init > 0
// (one segment, the whole of X): C[i] = mean +- 1 by init_offset(i, j) for
// i < nc_init.  Columns d..ldc are zero.
__global__ __launch_bounds__(256) void seg_mean_kernel(
    const double *__restrict__ part, const uint32_t *__restrict__ cstart,
    const uint32_t *__restrict__ off, uint64_t n, uint32_t d, uint32_t nc_init,
    float *__restrict__ C, uint32_t ldc)
{
    __shared__ double red[4][64];
    const uint32_t c = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const uint32_t j = blockIdx.y * 64 + lane;
    const uint32_t q0 = cstart ? cstart[c] : 0u;
    const uint32_t q1 = cstart ? cstart[c + 1] : (uint32_t)((n + SEG_ROWS - 1) / SEG_ROWS);
    const uint64_t cnt = off ? (uint64_t)(off[c + 1] - off[c]) : n;
    const uint32_t quarter = (q1 - q0 + 3) / 4;
    const uint32_t ga = min(q1, q0 + g * quarter), gb = min(q1, ga + quarter);
    double acc = 0.0;
    if (j < d) {
        uint32_t q = ga;
        for (; q + 4 <= gb; q += 4) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = part[(size_t)(q + u) * d + j];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = acc + v[u];
        }
        for (; q < gb; ++q) acc = acc + part[(size_t)q * d + j];
    }
    red[g][lane] = acc;
    __syncthreads();
    if (g != 0 || j >= ldc) return;
    const double sum = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    if (nc_init == 0) {
        C[(size_t)c * ldc + j] = (j < d && cnt != 0) ? (float)(sum / (double)cnt) : 0.0f;
        return;
    }
    const float mean = (float)(sum / (double)cnt);
    for (uint32_t i = 0; i < nc_init; ++i)
        C[(size_t)i * ldc + j] = j < d ? (init_offset(i, j) > 0 ? mean + 1.0f : mean - 1.0f) : 0.0f;
}

// R[r] = X[ids[r]] - C[list(r)] (f32), rows in list order.
__global__ void residual_kernel(const float *__restrict__ X, uint32_t ldx,
                                const uint32_t *__restrict__ ids,
                                const uint32_t *__restrict__ lists, const float *__restrict__ C,
                                uint64_t n, uint32_t dpad, float *__restrict__ R)
{
    uint64_t total = n * dpad;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = e / dpad, j = e % dpad;
        float x = X[(size_t)ids[r] * ldx + j];
        float c = C[(size_t)lists[r] * dpad + j];
        R[e] = x - c;
    }
}

// One workgroup trains the k-means of sub-space s of list l on the list's
// residuals (pq_idx.rs:122-149 -> kmeans.rs:195-207) and writes its
// codebook and the final assignment (the codes, pq_idx.rs:143-146).
//
// Centres live in LDS at a stride of SDM floats with zero padding past sd:
// a padded term adds (0 - 0)^2 = +0 to a fold that is already >= +0, so the
// distances equal the sd-term fold.  update_center (kmeans.rs:88-125) sums
// each centre's members in point order: a stable counting sort of the items
// by assignment (LDS histogram, block scan, then a per-wave rank from
// ballots of the assignment bits) lists every centre's members in item
// order, and one lane per (centre, dimension) folds its run.  This is the
// same sequence of additions as scanning the items for each centre, in
// O(L) work instead of O(L * KS).  A list of at most stage_cap items keeps
// its residuals, assignments and the sorted order in LDS; a longer one
// reads R and keeps both arrays in global scratch laid out [M][n_rows].
// Workgroups take lists in the order lorder gives (longest first), so the
// long lists do not start last.
constexpr int PQ_THREADS = 256;
constexpr int PQ_MAXSD = 32;

// Exclusive prefix sum of v over the workgroup (all threads call it).
__device__ __forceinline__ uint32_t pq_block_excl_scan(uint32_t v, uint32_t *wtot)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wtot[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int u = 0; u < w; ++u) pre += wtot[u];
    __syncthreads();
    return pre + x - v;
}

template <typename ACC_T, int SDM>
__global__ __launch_bounds__(PQ_THREADS) void pq_train_kernel(
    const float *__restrict__ R, uint32_t ldr, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ lorder, uint32_t M, const uint32_t *__restrict__ sub_b,
    const uint32_t *__restrict__ sub_e, uint32_t KS, uint32_t sdm, uint32_t epochs,
    uint64_t seed, uint32_t stage_cap, uint64_t n_rows, uint16_t *__restrict__ asg_g,
    uint32_t *__restrict__ ord_g, float *__restrict__ codebooks, uint8_t *__restrict__ codes)
{
    extern __shared__ __align__(16) unsigned char pq_smem[];
    __shared__ uint32_t wtot[PQ_THREADS / 64];
    constexpr int NW = PQ_THREADS / 64;
    const uint32_t l = lorder[blockIdx.x / M], s = blockIdx.x % M;
    const uint32_t base = off[l], L = off[l + 1] - base;
    const uint32_t b = sub_b[s], sd = sub_e[s] - sub_b[s];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    float *cb = codebooks + ((size_t)l * M + s) * KS * sdm;
    if (L == 0) {
        for (uint32_t e = tid; e < KS * sdm; e += PQ_THREADS) cb[e] = 0.0f;
        return;
    }
    double *pk = reinterpret_cast<double *>(pq_smem);         // KS (split)
    float *cent = reinterpret_cast<float *>(pk + KS);         // KS * SDM
    uint32_t *cnt = reinterpret_cast<uint32_t *>(cent + KS * SDM);  // KS
    uint32_t *start = cnt + KS;                              // KS
    uint32_t *run = start + KS;                              // KS
    uint16_t *wcnt = reinterpret_cast<uint16_t *>(run + KS);  // NW * KS
    float *zs = reinterpret_cast<float *>(wcnt + NW * KS);    // L * sd when staged
    const bool staged = L <= stage_cap;
    uint32_t *ord = staged ? reinterpret_cast<uint32_t *>(zs + (size_t)L * sd)
                           : ord_g + (size_t)s * n_rows + base;
    uint16_t *asg = staged ? reinterpret_cast<uint16_t *>(ord + L)
                           : asg_g + (size_t)s * n_rows + base;

    auto Z = [&](uint32_t i, uint32_t j) -> float {
        return staged ? zs[i * sd + j] : R[(size_t)(base + i) * ldr + b + j];
    };
    if (staged)
        for (uint32_t e = tid; e < L * sd; e += PQ_THREADS) {
            uint32_t i = e / sd, j = e % sd;
            zs[e] = R[(size_t)(base + i) * ldr + b + j];
        }
    for (uint32_t e = tid; e < KS * SDM; e += PQ_THREADS) cent[e] = 0.0f;
    __syncthreads();

    // init_center (kmeans.rs:50-86)
    if (tid < (int)sd) {
        ACC_T acc = 0;
        for (uint32_t i = 0; i < L; ++i) acc = acc + (ACC_T)Z(i, tid);
        float mean;
        if constexpr (sizeof(ACC_T) == 4) mean = acc / (float)L;
        else mean = (float)(acc / (double)L);
        for (uint32_t c = 0; c < KS; ++c)
            cent[c * SDM + tid] = init_offset(c, tid) > 0 ? mean + 1.0f : mean - 1.0f;
    }
    __syncthreads();

    uint64_t rng = stream_seed(seed, 1ull + (uint64_t)l * M + s);
    const float EPS = 1.0f / 1024.0f;
    for (uint32_t ep = 0; ep <= epochs; ++ep) {
        // search_data (kmeans.rs:127-148): argmin, strict '<'
        for (uint32_t i = tid; i < L; i += PQ_THREADS) {
            float z[SDM];
#pragma unroll
            for (int j = 0; j < SDM; ++j) z[j] = (uint32_t)j < sd ? Z(i, j) : 0.0f;
            float bd = 0.0f;
            uint32_t best = 0;
            for (uint32_t c = 0; c < KS; ++c) {
                const float *cc = cent + c * SDM;
                float dc = -0.0f;
#pragma unroll
                for (int j = 0; j < SDM; ++j) {
                    const float df = z[j] - cc[j];
                    dc = dc + df * df;
                }
                if (c == 0 || dc < bd) {
                    bd = dc;
                    best = c;
                }
            }
            if (ep == epochs) codes[(size_t)(base + i) * M + s] = (uint8_t)best;
            else asg[i] = (uint16_t)best;
        }
        if (ep == epochs) break;  // the search after train (pq_idx.rs:143-144)
        // update_center (kmeans.rs:88-125), KS <= PQ_THREADS
        if ((uint32_t)tid < KS) {
            cnt[tid] = 0;
#pragma unroll
            for (int u = 0; u < NW; ++u) wcnt[u * KS + tid] = 0;
        }
        __syncthreads();
        for (uint32_t i = tid; i < L; i += PQ_THREADS) atomicAdd(&cnt[asg[i]], 1u);
        __syncthreads();
        {
            const uint32_t st = pq_block_excl_scan((uint32_t)tid < KS ? cnt[tid] : 0u, wtot);
            if ((uint32_t)tid < KS) {
                start[tid] = st;
                run[tid] = st;
            }
        }
        __syncthreads();
        // stable scatter: ord[start[c] ..] = members of c in item order
        for (uint32_t i0 = 0; i0 < L; i0 += PQ_THREADS) {
            const uint32_t i = i0 + tid;
            const bool valid = i < L;
            const uint32_t v = valid ? asg[i] : 0u;
            uint64_t eq = __ballot(valid);
            for (uint32_t bit = 1; bit < KS; bit <<= 1) {
                const uint64_t bb = __ballot((v & bit) != 0);
                eq &= (v & bit) ? bb : ~bb;
            }
            const uint32_t rk = __popcll(eq & ((1ull << lane) - 1ull));
            if (valid && rk == 0) wcnt[w * KS + v] = (uint16_t)__popcll(eq);
            __syncthreads();
            if (valid) {
                uint32_t p = run[v] + rk;
                for (int u = 0; u < w; ++u) p += wcnt[u * KS + v];
                ord[p] = i;
            }
            __syncthreads();
            if ((uint32_t)tid < KS) {
                uint32_t add = 0;
#pragma unroll
                for (int u = 0; u < NW; ++u) {
                    add += wcnt[u * KS + tid];
                    wcnt[u * KS + tid] = 0;
                }
                run[tid] += add;
            }
            __syncthreads();
        }
        // new centre = its members' point-order sum / count; empty -> 0
        for (uint32_t e = tid; e < KS * sd; e += PQ_THREADS) {
            const uint32_t c = e / sd, j = e % sd;
            const uint32_t p0 = start[c], n_c = cnt[c];
            ACC_T acc = 0;
            for (uint32_t p = p0; p < p0 + n_c; ++p) acc = acc + (ACC_T)Z(ord[p], j);
            float nv;
            if (n_c == 0) nv = 0.0f;
            else if constexpr (sizeof(ACC_T) == 4) nv = acc / (float)n_c;
            else nv = (float)(acc / (double)n_c);
            cent[c * SDM + j] = nv;
        }
        __syncthreads();
        // split_center (kmeans.rs:150-193), all epochs but the last, by wave
        // 0.  The walk for an empty centre i draws once per step: step t
        // looks at centre (i + 1 + t) % KS with the generator's (t + 1)-th
        // next draw and stops at the first u < count / L.  Each lane takes
        // one step of a block of 64, a ballot finds the first hit, and the
        // generator advances by the steps taken -- the same draws and the
        // same picks as the step-by-step walk.
        if (ep + 1 < epochs && w == 0) {
            // pk[c] = count / L, the walk's pick probability, updated as
            // the counts change
            for (uint32_t c = lane; c < KS; c += 64) pk[c] = (double)cnt[c] / (double)L;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t i = 0; i < KS; ++i) {
                if (cnt[i] != 0) continue;
                uint32_t steps = 0;
                for (uint32_t t0 = 0;; t0 += 64) {
                    const uint32_t t = t0 + lane;
                    const uint32_t c = (i + 1 + t) % KS;
                    const double pc = pk[c];
                    const bool hit = pc != 0.0 && splitmix_unit_at(rng, t) < pc;
                    const uint64_t m = __ballot(hit);
                    if (m) {
                        steps = t0 + (uint32_t)__builtin_ctzll(m);
                        break;
                    }
                }
                const uint32_t sc = (i + 1 + steps) % KS;
                rng += ((uint64_t)steps + 1) * 0x9E3779B97F4A7C15ull;
                if ((uint32_t)lane < sd) {
                    const uint32_t j = lane;
                    const float src = cent[sc * SDM + j];
                    if (j % 2 == 0) {
                        cent[i * SDM + j] = src * (1.0f - EPS);
                        cent[sc * SDM + j] = src * (1.0f + EPS);
                    } else {
                        cent[i * SDM + j] = src * (1.0f + EPS);
                        cent[sc * SDM + j] = src * (1.0f - EPS);
                    }
                }
                if (lane == 0) {
                    const uint32_t half = cnt[sc] / 2;
                    cnt[i] = half;
                    cnt[sc] -= half;
                    pk[i] = (double)half / (double)L;
                    pk[sc] = (double)cnt[sc] / (double)L;
                }
                // the next walk reads these counts from other lanes
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (uint32_t e = tid; e < KS * sdm; e += PQ_THREADS) {
        uint32_t c = e / sdm, j = e % sdm;
        cb[e] = j < sd ? cent[c * SDM + j] : 0.0f;
    }
}

// One workgroup per (query, probe): residual, LUT (pq_idx.rs:170-179), ADC
// over the list (181-188), top-k1 by (distance, global id) kept in LDS
// across chunks of the list.
__global__ __launch_bounds__(256) void ivf_scan_kernel(
    const float *__restrict__ Qp, uint32_t dpad, const float *__restrict__ C,
    const uint32_t *__restrict__ probes, uint32_t ldp, uint32_t nprobe,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
    const uint8_t *__restrict__ codes, const float *__restrict__ codebooks,
    const uint32_t *__restrict__ sub_b, const uint32_t *__restrict__ sub_e, uint32_t M,
    uint32_t KS, uint32_t sdm, uint32_t k1, uint32_t P, const uint32_t *__restrict__ nprobe_q,
    unsigned long long *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char sc_smem[];
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(sc_smem);  // P
    float *lut = reinterpret_cast<float *>(keys + P);                            // M*KS
    float *z = lut + (size_t)M * KS;                                             // dpad
    const int tid = threadIdx.x;
    const uint32_t q = blockIdx.y, p = blockIdx.x;
    if (nprobe_q && p >= nprobe_q[q]) {  // this query probes fewer lists
        unsigned long long *o = out + ((size_t)q * nprobe + p) * k1;
        for (uint32_t j = tid; j < k1; j += 256) o[j] = ~0ull;
        return;
    }
    const uint32_t l = probes[(size_t)q * ldp + p];
    for (uint32_t j = tid; j < dpad; j += 256)
        z[j] = Qp[(size_t)q * dpad + j] - C[(size_t)l * dpad + j];
    __syncthreads();
    const float *cb = codebooks + (size_t)l * M * KS * sdm;
    for (uint32_t e = tid; e < M * KS; e += 256) {
        uint32_t s = e / KS;
        uint32_t b = sub_b[s], sd = sub_e[s] - b;
        const float *w = cb + (size_t)e * sdm;
        float acc = -0.0f;
        for (uint32_t j = 0; j < sd; ++j) {
            float df = z[b + j] - w[j];
            acc = acc + df * df;
        }
        lut[e] = acc;
    }
    __syncthreads();
    const uint32_t base = off[l], L = off[l + 1] - base;
    const uint32_t CH = P - k1;
    uint32_t kept = 0;
    for (uint32_t c0 = 0; c0 < L || c0 == 0; c0 += CH) {
        const uint32_t m = L > c0 ? (L - c0 < CH ? L - c0 : CH) : 0;
        for (uint32_t i = tid; i < m; i += 256) {
            const uint32_t row = base + c0 + i;
            const uint8_t *cd = codes + (size_t)row * M;
            float dist = 0.0f;
            for (uint32_t j = 0; j < M; ++j) dist = dist + lut[j * KS + cd[j]];
            keys[kept + i] = ((unsigned long long)fkey(dist) << 32) | ids[row];
        }
        const uint32_t total = kept + m;
        uint32_t PP = 1;
        while (PP < total) PP <<= 1;
        for (uint32_t j = total + tid; j < PP; j += 256) keys[j] = ~0ull;
        __syncthreads();
        for (uint32_t size = 2; size <= PP; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t t = tid; t < PP / 2; t += 256) {
                    uint32_t lo = bitonic_lo(t, stride), hi = lo + stride;
                    bool asc = (lo & size) == 0;
                    unsigned long long a = keys[lo], bb = keys[hi];
                    if ((a > bb) == asc) {
                        keys[lo] = bb;
                        keys[hi] = a;
                    }
                }
                __syncthreads();
            }
        kept = total < k1 ? total : k1;
        if (c0 + CH >= L) break;
    }
    unsigned long long *o = out + ((size_t)q * nprobe + p) * k1;
    for (uint32_t j = tid; j < k1; j += 256) o[j] = j < kept ? keys[j] : ~0ull;
}

// ---- List-major scan.  The (query, probe) pairs are bucketed by list, and
// one workgroup per list stages that list's codebook in LDS once, then
// serves every query that probes the list: residual, LUT from the LDS
// codebook, ADC over the list, top-k1.  The per-pair arithmetic is the same
// as ivf_scan_kernel's (pq_idx.rs:151-194), so the candidates are the same.

__global__ void fill_u32_small_kernel(uint32_t *p, uint32_t n, uint32_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// Pairs per list: counts, then positions (cnt zeroed by the caller).
__global__ void pair_count_kernel(const uint32_t *__restrict__ probes, uint32_t ldp,
                                  uint32_t nprobe, uint32_t rows,
                                  const uint32_t *__restrict__ nprobe_q, uint32_t *cnt)
{
    const uint64_t total = (uint64_t)rows * nprobe;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(e / nprobe), p = (uint32_t)(e % nprobe);
        if (nprobe_q && p >= nprobe_q[q]) continue;
        atomicAdd(cnt + probes[(size_t)q * ldp + p], 1u);
    }
}

// Exclusive scans of cnt[0..nl) into off[0..nl] and of ceil(cnt / G)
// into uoff[0..nl]; one workgroup of 1024.
__global__ __launch_bounds__(1024) void pair_scan_kernel(const uint32_t *__restrict__ cnt,
                                                          uint32_t nl, uint32_t G,
                                                          uint32_t *__restrict__ off,
                                                          uint32_t *__restrict__ uoff)
{
    __shared__ uint32_t part[1024], upart[1024];
    const uint32_t tid = threadIdx.x, per = (nl + 1023) / 1024, b = tid * per;
    uint32_t sum = 0, usum = 0;
    for (uint32_t i = 0; i < per && b + i < nl; ++i) {
        sum += cnt[b + i];
        usum += (cnt[b + i] + G - 1) / G;
    }
    part[tid] = sum;
    upart[tid] = usum;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t v = tid >= o ? part[tid - o] : 0u;
        const uint32_t w = tid >= o ? upart[tid - o] : 0u;
        __syncthreads();
        part[tid] += v;
        upart[tid] += w;
        __syncthreads();
    }
    uint32_t run = part[tid] - sum, urun = upart[tid] - usum;
    for (uint32_t i = 0; i < per && b + i < nl; ++i) {
        off[b + i] = run;
        uoff[b + i] = urun;
        run += cnt[b + i];
        urun += (cnt[b + i] + G - 1) / G;
    }
    if (tid == 1023) {
        off[nl] = part[1023];
        uoff[nl] = upart[1023];
    }
}

__global__ void pair_fill_kernel(const uint32_t *__restrict__ probes, uint32_t ldp,
                                 uint32_t nprobe, uint32_t rows,
                                 const uint32_t *__restrict__ nprobe_q,
                                 const uint32_t *__restrict__ off, uint32_t *fill,
                                 uint32_t *__restrict__ pairs, unsigned long long *__restrict__ out,
                                 uint32_t k1)
{
    const uint64_t total = (uint64_t)rows * nprobe;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(e / nprobe), p = (uint32_t)(e % nprobe);
        if (nprobe_q && p >= nprobe_q[q]) {  // this query probes fewer lists
            if (out) {
                unsigned long long *o = out + e * k1;
                for (uint32_t j = 0; j < k1; ++j) o[j] = ~0ull;
            }
            continue;
        }
        const uint32_t l = probes[(size_t)q * ldp + p];
        pairs[off[l] + atomicAdd(fill + l, 1u)] = (uint32_t)e;  // e = q * nprobe + p
    }
}

constexpr int LM_THREADS = 512;
constexpr uint32_t LM_BINS = 2048;

// Block-wide: the bin of hist[0..LM_BINS) holding the need-th smallest
// (1-based) counted element, and how many are still needed inside it.
__device__ void lm_find_bin(const uint32_t *hist, uint32_t need, uint32_t *s_part,
                            uint32_t *bin, uint32_t *rem)
{
    const uint32_t tid = threadIdx.x, per = LM_BINS / LM_THREADS, b0 = tid * per;
    uint32_t sum = 0;
    for (uint32_t i = 0; i < per; ++i) sum += hist[b0 + i];
    s_part[tid] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < LM_THREADS; o <<= 1) {
        const uint32_t v = tid >= o ? s_part[tid - o] : 0u;
        __syncthreads();
        s_part[tid] += v;
        __syncthreads();
    }
    const uint32_t incl = s_part[tid], excl = incl - sum;
    if (excl < need && need <= incl) {
        uint32_t c = excl;
        for (uint32_t i = 0; i < per; ++i) {
            const uint32_t h = hist[b0 + i];
            if (c + h >= need) {
                s_part[LM_THREADS] = b0 + i;
                s_part[LM_THREADS + 1] = need - c;
                break;
            }
            c += h;
        }
    }
    __syncthreads();
    *bin = s_part[LM_THREADS];
    *rem = s_part[LM_THREADS + 1];
    __syncthreads();
}

// A work unit is up to G pairs of one list (lists probed by many queries
// are split over several workgroups).  The unit's workgroup stages the
// list's codebook in LDS once; per pair it builds the LUT, and selects the
// k1 smallest (fkey(adc) << 32 | id) of the list: lists up to P items are
// sorted whole; longer ones keep their ADC keys in registers (up to 16 per
// thread), take two 11-bit histogram passes over them to find a threshold,
// and sort the keys under it; longer lists still are sorted in chunks.
__global__ __launch_bounds__(LM_THREADS) void ivf_scan_lists_kernel(
    const float *__restrict__ Qp, uint32_t dpad, const float *__restrict__ C,
    const uint32_t *__restrict__ pair_off, const uint32_t *__restrict__ unit_off, uint32_t nlist,
    uint32_t G, const uint32_t *__restrict__ pairs, uint32_t nprobe,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
    const uint8_t *__restrict__ codes, const float *__restrict__ codebooks,
    const uint32_t *__restrict__ sub_b, const uint32_t *__restrict__ sub_e, uint32_t M,
    uint32_t KS, uint32_t sdm, uint32_t k1, uint32_t P, unsigned long long *__restrict__ out)
{
    const uint32_t u = blockIdx.x;
    if (u >= unit_off[nlist]) return;
    uint32_t lo = 0, hi = nlist;  // the list l with unit_off[l] <= u < unit_off[l + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (unit_off[mid] <= u) lo = mid;
        else hi = mid;
    }
    const uint32_t l = lo;
    const uint32_t p0 = pair_off[l] + (u - unit_off[l]) * G;
    const uint32_t p1 = min(p0 + G, pair_off[l + 1]);
    extern __shared__ __align__(16) unsigned char lm_smem[];
    __shared__ uint32_t s_part[LM_THREADS + 2];
    __shared__ uint32_t s_n;
    const uint32_t ncb = M * KS * sdm;
    float *cbs = reinterpret_cast<float *>(lm_smem);  // [M][KS][sdm]
    float *lut = cbs + ncb;                            // [M][KS]
    unsigned char *region = reinterpret_cast<unsigned char *>(lut + (size_t)M * KS);
    float *z = reinterpret_cast<float *>(region);                            // [dpad], LUT build
    uint32_t *hist = reinterpret_cast<uint32_t *>(region);                   // [LM_BINS], select
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(region);  // [P], sort
    const int tid = threadIdx.x;
    const float *cb = codebooks + (size_t)l * ncb;
    if ((ncb & 3) == 0) {  // whole float4s per list, 16-byte aligned from the allocation base
        for (uint32_t e = tid; e < ncb / 4; e += LM_THREADS)
            reinterpret_cast<float4 *>(cbs)[e] = reinterpret_cast<const float4 *>(cb)[e];
    } else {
        for (uint32_t e = tid; e < ncb; e += LM_THREADS) cbs[e] = cb[e];
    }
    const uint32_t base = off[l], L = off[l + 1] - base;
    // ADC of item i in the reference's order (a fold from +0.0 over the
    // sub-spaces); codes are read a word at a time when M is a multiple of 4
    auto adc_key = [&](uint32_t i) -> unsigned long long {
        const uint32_t row = base + i;
        const uint8_t *cd = codes + (size_t)row * M;
        float dist = 0.0f;
        if ((M & 3) == 0) {
            const uint32_t *cw = reinterpret_cast<const uint32_t *>(cd);
            for (uint32_t w = 0; w < M / 4; ++w) {
                const uint32_t v = cw[w];
                dist = dist + lut[(4 * w) * KS + (v & 0xFFu)];
                dist = dist + lut[(4 * w + 1) * KS + ((v >> 8) & 0xFFu)];
                dist = dist + lut[(4 * w + 2) * KS + ((v >> 16) & 0xFFu)];
                dist = dist + lut[(4 * w + 3) * KS + (v >> 24)];
            }
        } else {
            for (uint32_t j = 0; j < M; ++j) dist = dist + lut[j * KS + cd[j]];
        }
        return ((unsigned long long)fkey(dist) << 32) | ids[row];
    };
    auto sort_keys = [&](uint32_t total) {
        uint32_t PP = 1;
        while (PP < total) PP <<= 1;
        for (uint32_t j = total + tid; j < PP; j += LM_THREADS) keys[j] = ~0ull;
        __syncthreads();
        for (uint32_t size = 2; size <= PP; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t t = tid; t < PP / 2; t += LM_THREADS) {
                    const uint32_t a0 = bitonic_lo(t, stride), a1 = a0 + stride;
                    const bool asc = (a0 & size) == 0;
                    const unsigned long long a = keys[a0], bb = keys[a1];
                    if ((a > bb) == asc) {
                        keys[a0] = bb;
                        keys[a1] = a;
                    }
                }
                __syncthreads();
            }
    };
    constexpr uint32_t RK = 16;  // keys held in registers per thread (lists up to 8192)
    for (uint32_t pp = p0; pp < p1; ++pp) {
        const uint32_t e = pairs[pp], q = e / nprobe;
        __syncthreads();  // previous pair done with the region and the LUT
        for (uint32_t j = tid; j < dpad; j += LM_THREADS)
            z[j] = Qp[(size_t)q * dpad + j] - C[(size_t)l * dpad + j];
        __syncthreads();
        for (uint32_t t = tid; t < M * KS; t += LM_THREADS) {
            const uint32_t sub = t / KS, b = sub_b[sub], sd = sub_e[sub] - b;
            const float *w = cbs + (size_t)t * sdm;
            float acc = -0.0f;
            for (uint32_t j = 0; j < sd; ++j) {
                float df = z[b + j] - w[j];
                acc = acc + df * df;
            }
            lut[t] = acc;
        }
        __syncthreads();
        uint32_t kept = 0;
        if (L <= P) {
            for (uint32_t i = tid; i < L; i += LM_THREADS) keys[i] = adc_key(i);
            sort_keys(L);
            kept = L < k1 ? L : k1;
        } else if (L <= RK * LM_THREADS) {
            // every key of the list in registers: two 11-bit histogram
            // passes find a threshold, the keys under it are sorted
            unsigned long long kr[RK];
#pragma unroll
            for (uint32_t r = 0; r < RK; ++r) {
                const uint32_t i = tid + r * LM_THREADS;
                kr[r] = i < L ? adc_key(i) : ~0ull;
            }
            for (uint32_t j = tid; j < LM_BINS; j += LM_THREADS) hist[j] = 0;
            __syncthreads();
#pragma unroll
            for (uint32_t r = 0; r < RK; ++r)
                if (tid + r * LM_THREADS < L) atomicAdd(&hist[(uint32_t)(kr[r] >> 53)], 1u);
            __syncthreads();
            uint32_t b1, r1;
            lm_find_bin(hist, k1, s_part, &b1, &r1);
            for (uint32_t j = tid; j < LM_BINS; j += LM_THREADS) hist[j] = 0;
            __syncthreads();
#pragma unroll
            for (uint32_t r = 0; r < RK; ++r)
                if (tid + r * LM_THREADS < L && (uint32_t)(kr[r] >> 53) == b1)
                    atomicAdd(&hist[(uint32_t)(kr[r] >> 42) & 2047u], 1u);
            __syncthreads();
            uint32_t b2, r2;
            lm_find_bin(hist, r1, s_part, &b2, &r2);
            const uint32_t thr = (b1 << 11) | b2;  // keep keys with key >> 42 <= thr
            if (tid == 0) s_n = 0;
            __syncthreads();
#pragma unroll
            for (uint32_t r = 0; r < RK; ++r)
                if (tid + r * LM_THREADS < L && (uint32_t)(kr[r] >> 42) <= thr) {
                    const uint32_t pos = atomicAdd(&s_n, 1u);
                    if (pos < P) keys[pos] = kr[r];
                }
            __syncthreads();
            const uint32_t total = s_n;
            if (total <= P) {
                sort_keys(total);
                kept = total < k1 ? total : k1;
            } else {
                // the threshold bin holds more than P keys (massive ties):
                // chunked sort of the whole list keeping the k1 smallest
                __syncthreads();
                const uint32_t CH = P - k1;
                for (uint32_t c0 = 0; c0 < L; c0 += CH) {
                    const uint32_t m = L - c0 < CH ? L - c0 : CH;
                    for (uint32_t i = tid; i < m; i += LM_THREADS) keys[kept + i] = adc_key(c0 + i);
                    sort_keys(kept + m);
                    kept = kept + m < k1 ? kept + m : k1;
                }
            }
        } else {
            // very long list: chunked sort keeping the k1 smallest
            const uint32_t CH = P - k1;
            for (uint32_t c0 = 0; c0 < L; c0 += CH) {
                const uint32_t m = L - c0 < CH ? L - c0 : CH;
                for (uint32_t i = tid; i < m; i += LM_THREADS) keys[kept + i] = adc_key(c0 + i);
                sort_keys(kept + m);
                kept = kept + m < k1 ? kept + m : k1;
            }
        }
        unsigned long long *o = out + (size_t)e * k1;
        for (uint32_t j = tid; j < k1; j += LM_THREADS) o[j] = j < kept ? keys[j] : ~0ull;
    }
}

// ---- Row form of the list scan: the ADC distance of every item of every
// probed list is written to the query's row of D (the lists in probe order,
// each list's items in list order), and the k+1 smallest of each row come
// from the row select.  seg[q * nprobe + p] = the column of the first item
// of the query's p-th probed list, row_n[q] = the row's length.
__global__ void ivf_seg_kernel(const uint32_t *__restrict__ probes, uint32_t ldp, uint32_t nprobe,
                               uint32_t rows, const uint32_t *__restrict__ nprobe_q,
                               const uint32_t *__restrict__ off, uint32_t *__restrict__ seg,
                               uint32_t *__restrict__ row_n)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rows) return;
    const uint32_t np_q = nprobe_q ? nprobe_q[q] : nprobe;
    uint32_t acc = 0;
    for (uint32_t p = 0; p < np_q; ++p) {
        const uint32_t l = probes[(size_t)q * ldp + p];
        seg[(size_t)q * nprobe + p] = acc;
        acc += off[l + 1] - off[l];
    }
    row_n[q] = acc;
}

// Work units as in ivf_scan_lists_kernel (up to G pairs of one list).  The
// list's codebook is held in registers: thread t owns the LUT entries
// t + r * LM_THREADS (r < R), each with its sub-vector of up to SD floats.
// Per pair: the residual z = q - c_l in LDS, the LUT entries (the same fold
// as ivf_scan_kernel: -0.0 plus (z - w)^2 per dimension in order), then the
// ADC of every item (0.0 plus the LUT values in sub-space order) into D.
template <int R, int SD, bool FULL>
__global__ __launch_bounds__(LM_THREADS, 2) void ivf_adc_rows_kernel(
    const float *__restrict__ Qp, uint32_t dpad, const float *__restrict__ C,
    const uint32_t *__restrict__ pair_off, const uint32_t *__restrict__ unit_off, uint32_t nlist,
    uint32_t G, const uint32_t *__restrict__ pairs, uint32_t nprobe,
    const uint32_t *__restrict__ off, const uint8_t *__restrict__ codes,
    const float *__restrict__ codebooks, const uint32_t *__restrict__ sub_b,
    const uint32_t *__restrict__ sub_e, uint32_t M, uint32_t KS, uint32_t sdm,
    const uint32_t *__restrict__ seg, float *__restrict__ D, uint64_t ldd)
{
    const uint32_t u = blockIdx.x;
    if (u >= unit_off[nlist]) return;
    uint32_t lo = 0, hi = nlist;  // the list l with unit_off[l] <= u < unit_off[l + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (unit_off[mid] <= u) lo = mid;
        else hi = mid;
    }
    const uint32_t l = lo;
    const uint32_t p0 = pair_off[l] + (u - unit_off[l]) * G;
    const uint32_t p1 = min(p0 + G, pair_off[l + 1]);
    extern __shared__ __align__(16) unsigned char ar_smem[];
    const uint32_t nent = M * KS;
    float *lut = reinterpret_cast<float *>(ar_smem);  // [nent]
    float *z = lut + ((nent + 3) & ~3u);               // [dpad]
    const uint32_t tid = threadIdx.x;
    const float *cb = codebooks + (size_t)l * nent * sdm;
    float w[R][SD];
    uint32_t eb[R], esd[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t t = tid + (uint32_t)r * LM_THREADS;
        eb[r] = 0;
        esd[r] = 0;
        if (t < nent) {
            const uint32_t sub = t / KS;
            eb[r] = sub_b[sub];
            esd[r] = sub_e[sub] - eb[r];
        }
#pragma unroll
        for (int j = 0; j < SD; ++j) w[r][j] = (t < nent && (uint32_t)j < sdm) ? cb[(size_t)t * sdm + j] : 0.0f;
    }
    const uint32_t base = off[l], L = off[l + 1] - base;
    for (uint32_t pp = p0; pp < p1; ++pp) {
        const uint32_t e = pairs[pp], q = e / nprobe;
        for (uint32_t j = tid; j < dpad; j += LM_THREADS)
            z[j] = Qp[(size_t)q * dpad + j] - C[(size_t)l * dpad + j];
        __syncthreads();  // z written; every thread done reading the last pair's LUT
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t t = tid + (uint32_t)r * LM_THREADS;
            float acc = -0.0f;
            if constexpr (FULL) {
                // every sub-space SD wide (4-aligned): its residual slice
                // as float4 reads, no per-dimension guard
                const float4 *zp = reinterpret_cast<const float4 *>(z + eb[r]);
#pragma unroll
                for (int j4 = 0; j4 < SD / 4; ++j4) {
                    const float4 zv = zp[j4];
                    const float zz[4] = {zv.x, zv.y, zv.z, zv.w};
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const float df = zz[h] - w[r][4 * j4 + h];
                        acc = acc + df * df;
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < SD; ++j) {
                    if ((uint32_t)j < esd[r]) {
                        const float df = z[eb[r] + j] - w[r][j];
                        acc = acc + df * df;
                    }
                }
            }
            if (t < nent) lut[t] = acc;
        }
        __syncthreads();  // LUT complete; z free for the next pair
        float *drow = D + (size_t)q * ldd + seg[e];
        for (uint32_t i = tid; i < L; i += LM_THREADS) {
            const size_t row = (size_t)base + i;
            float dist = 0.0f;
            if ((M & 15) == 0) {
                const uint4 *cv = reinterpret_cast<const uint4 *>(codes + row * M);
                for (uint32_t v4 = 0; v4 < M / 16; ++v4) {
                    const uint4 c = cv[v4];
                    const uint32_t wd[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const uint32_t s0 = 16 * v4 + 4 * h;
                        dist = dist + lut[s0 * KS + (wd[h] & 0xFFu)];
                        dist = dist + lut[(s0 + 1) * KS + ((wd[h] >> 8) & 0xFFu)];
                        dist = dist + lut[(s0 + 2) * KS + ((wd[h] >> 16) & 0xFFu)];
                        dist = dist + lut[(s0 + 3) * KS + (wd[h] >> 24)];
                    }
                }
            } else if ((M & 3) == 0) {
                const uint32_t *cw = reinterpret_cast<const uint32_t *>(codes + row * M);
                for (uint32_t v = 0; v < M / 4; ++v) {
                    const uint32_t c = cw[v];
                    dist = dist + lut[(4 * v) * KS + (c & 0xFFu)];
                    dist = dist + lut[(4 * v + 1) * KS + ((c >> 8) & 0xFFu)];
                    dist = dist + lut[(4 * v + 2) * KS + ((c >> 16) & 0xFFu)];
                    dist = dist + lut[(4 * v + 3) * KS + (c >> 24)];
                }
            } else {
                const uint8_t *cd = codes + row * M;
                for (uint32_t j = 0; j < M; ++j) dist = dist + lut[j * KS + cd[j]];
            }
            drow[i] = dist;
        }
    }
}

// Per query, from the row select's first K = min(k+1, row_n) candidates
// (columns, ascending by (distance, column)): the first min(k, row_n) as
// item ids, ROW_TIE when the K are not pairwise distinct (the select's flag)
// or the coarse probe set was tied.  One wave per query.
// Canonical candidate set of the row form (the exact-rerank form's rule:
// the k smallest ADC distances, ties broken by row index).  The radix
// select broke ties by column (probe rank, then list position); when the
// k-th and (k+1)-th selected distances are equal, the entries of that
// distance in the set are replaced by the ones with the smallest indices:
// a bisection on the index threshold T with need = k - (entries below the
// distance) finds the smallest T such that `need` entries of the row with
// that distance have index <= T, and those entries fill the set's tail.
// One workgroup per row; rows without such a tie return at once.
__global__ __launch_bounds__(256) void ivf_edge_fixup_kernel(
    const float *__restrict__ Dr, uint64_t ldr, const uint32_t *__restrict__ row_n,
    const float *__restrict__ cand_d, uint32_t *__restrict__ cand_i, uint32_t ldc, uint32_t k,
    const uint32_t *__restrict__ probes, uint32_t ldp, uint32_t nprobe,
    const uint32_t *__restrict__ nprobe_q, const uint32_t *__restrict__ seg,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids)
{
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint32_t n = row_n[q];
    const float *cd = cand_d + (size_t)q * ldc;
    if (n <= k || k == 0 || !(cd[k - 1] == cd[k])) return;
    const float v = cd[k - 1];
    __shared__ uint32_t s_cnt, s_below;
    if (tid == 0) {
        s_cnt = 0;
        s_below = 0;
    }
    __syncthreads();
    uint32_t below = 0;
    for (uint32_t j = tid; j < k; j += 256) below += cd[j] < v ? 1u : 0u;
    atomicAdd(&s_below, below);
    const uint32_t np_q = nprobe_q ? nprobe_q[q] : nprobe;
    const uint32_t *sg = seg + (size_t)q * nprobe;
    const float *row = Dr + (size_t)q * ldr;
    auto id_of = [&](uint32_t col) -> uint32_t {
        uint32_t a = 0, b = np_q;  // the last probe p with sg[p] <= col
        while (b - a > 1) {
            const uint32_t mid = (a + b) / 2;
            if (sg[mid] <= col) a = mid;
            else b = mid;
        }
        const uint32_t l = probes[(size_t)q * ldp + a];
        return ids[off[l] + (col - sg[a])];
    };
    // the entries of that distance, with their indices, while they fit in
    // LDS: each takes its rank by index among them
    constexpr uint32_t CAP = 1024;
    __shared__ uint32_t s_col[CAP], s_id[CAP], s_m;
    if (tid == 0) s_m = 0;
    __syncthreads();
    for (uint32_t col = tid; col < n; col += 256)
        if (row[col] == v) {
            const uint32_t j = atomicAdd(&s_m, 1u);
            if (j < CAP) {
                s_col[j] = col;
                s_id[j] = id_of(col);
            }
        }
    __syncthreads();
    const uint32_t need = k - s_below, mcount = s_m;
    uint32_t *ci = cand_i + (size_t)q * ldc;
    if (mcount <= CAP) {
        for (uint32_t j = tid; j < mcount; j += 256) {
            uint32_t rank = 0;
            for (uint32_t t = 0; t < mcount; ++t) rank += s_id[t] < s_id[j] ? 1u : 0u;
            if (rank < need) ci[s_below + rank] = s_col[j];
        }
        return;
    }
    // more than CAP entries share the distance: bisection on the index
    uint32_t lo = 0, hi = 0xFFFFFFFFu;
    while (lo < hi) {  // uniform: every thread sees the same counts
        const uint32_t mid = lo + (hi - lo) / 2;
        uint32_t c = 0;
        for (uint32_t col = tid; col < n; col += 256)
            if (row[col] == v && id_of(col) <= mid) ++c;
        if (tid == 0) s_cnt = 0;
        __syncthreads();
        atomicAdd(&s_cnt, c);
        __syncthreads();
        const uint32_t total = s_cnt;
        __syncthreads();
        if (total >= need) hi = mid;
        else lo = mid + 1;
    }
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    for (uint32_t col = tid; col < n; col += 256)
        if (row[col] == v && id_of(col) <= lo) {
            const uint32_t slot = atomicAdd(&s_cnt, 1u);
            if (slot < need) ci[s_below + slot] = col;
        }
}

__global__ __launch_bounds__(64) void ivf_rows_finalize_kernel(
    const float *__restrict__ cand_d, const uint32_t *__restrict__ cand_i, uint32_t ldc,
    const uint32_t *__restrict__ sel_flags, const uint32_t *__restrict__ coarse_flags,
    const uint32_t *__restrict__ probes, uint32_t ldp, uint32_t nprobe,
    const uint32_t *__restrict__ nprobe_q, const uint32_t *__restrict__ seg,
    const uint32_t *__restrict__ row_n, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ ids, uint32_t k, uint64_t *__restrict__ out_idx,
    float *__restrict__ out_dist, uint32_t *__restrict__ out_count, uint32_t *__restrict__ flags)
{
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint32_t n = row_n[q], m = k < n ? k : n;
    const uint32_t np_q = nprobe_q ? nprobe_q[q] : nprobe;
    const uint32_t *sg = seg + (size_t)q * nprobe;
    const float inf = __int_as_float(0x7f800000);
    for (uint32_t j = tid; j < k; j += 64) {
        if (j < m) {
            const uint32_t col = cand_i[(size_t)q * ldc + j];
            uint32_t a = 0, b = np_q;  // the last probe p with sg[p] <= col
            while (b - a > 1) {
                const uint32_t mid = (a + b) / 2;
                if (sg[mid] <= col) a = mid;
                else b = mid;
            }
            const uint32_t l = probes[(size_t)q * ldp + a];
            out_idx[(size_t)q * k + j] = ids[off[l] + (col - sg[a])];
            out_dist[(size_t)q * k + j] = cand_d[(size_t)q * ldc + j];
        } else {
            out_idx[(size_t)q * k + j] = ~0ull;
            out_dist[(size_t)q * k + j] = inf;
        }
    }
    if (tid == 0) {
        out_count[q] = m;
        const uint32_t f = sel_flags[q];
        const uint32_t cf = coarse_flags[q];
        // the set of the first k is tie-order dependent only at its edge
        const bool edge = n > k && cand_d[(size_t)q * ldc + k - 1] == cand_d[(size_t)q * ldc + k];
        flags[q] = (((f & (ROW_TIE | ROW_BAD)) || (cf & ROW_TIE)) ? ROW_TIE : 0u) |
                   ((edge || (cf & ROW_TIE_EDGE) || (f & ROW_BAD)) ? ROW_TIE_EDGE : 0u);
    }
}

// Which ties among a query's coarse distances can change its search
// result (pq_idx.rs:413-426 pops the coarse heap nprobe times; a list's
// place in the probe order matters only through the items it holds, so an
// empty list is interchangeable with any other empty list).  From the
// Kc = nprobe + 1 smallest (distance, list) pairs and the full distance row:
//  * a run of equal distances wholly inside the first nprobe: the order of
//    its non-empty lists is heap-dependent -> ROW_TIE when it holds >= 2;
//  * the run at the edge (the nprobe-th and (nprobe+1)-th are equal): which
//    of its lists are probed is heap-dependent -> ROW_TIE | ROW_TIE_EDGE
//    when any list of that distance in the whole row is non-empty.
// A NaN keeps its flag.  One wave per query.
__global__ __launch_bounds__(64) void coarse_ties_kernel(
    const float *__restrict__ D, uint64_t ldd, uint32_t nlist, const float *__restrict__ pd,
    const uint32_t *__restrict__ pi, uint32_t Kc, uint32_t nprobe,
    const uint32_t *__restrict__ off, uint32_t *__restrict__ flags)
{
    const uint32_t q = blockIdx.x, lane = threadIdx.x;
    const uint32_t f = flags[q];
    if (!(f & ROW_TIE) || (f & ROW_NAN)) return;  // no tie among the Kc, or a NaN row
    const float *d = pd + (size_t)q * Kc;
    const uint32_t *li = pi + (size_t)q * Kc;
    const uint32_t np = nprobe < Kc ? nprobe : Kc;
    auto nonempty = [&](uint32_t l) { return off[l + 1] > off[l]; };
    uint32_t out = 0;
    // runs inside the first np (each lane walks the runs that start at its
    // positions; Kc is small)
    for (uint32_t a = lane; a < np; a += 64) {
        if (a > 0 && d[a - 1] == d[a]) continue;  // not a run start
        uint32_t b = a + 1, ne = nonempty(li[a]) ? 1u : 0u;
        while (b < Kc && d[b] == d[a]) {
            ne += nonempty(li[b]) ? 1u : 0u;
            ++b;
        }
        if (b <= np && ne >= 2) out |= ROW_TIE;
    }
    // the edge run: every list of the row at that distance
    if (np < Kc && np > 0 && d[np - 1] == d[np]) {
        const float v = d[np];
        const float *row = D + (size_t)q * ldd;
        bool any = false;
        for (uint32_t l = lane; l < nlist; l += 64) any |= row[l] == v && nonempty(l);
        if (__ballot(any)) out |= ROW_TIE | ROW_TIE_EDGE;
    }
    // combine the lanes' findings
    const uint64_t t = __ballot((out & ROW_TIE) != 0), e = __ballot((out & ROW_TIE_EDGE) != 0);
    if (lane == 0) flags[q] = (f & ~ROW_TIE) | (t ? ROW_TIE : 0u) | (e ? ROW_TIE_EDGE : 0u);
}

// Per query: the nprobe*k1 candidates -> first k ascending; ROW_TIE when
// the first min(k+1, valid) distances are not pairwise distinct or the
// coarse probe set was tied.
__global__ __launch_bounds__(256) void ivf_merge_kernel(
    const unsigned long long *__restrict__ cand, uint32_t per_query, uint32_t k,
    const uint32_t *__restrict__ coarse_flags, uint64_t *__restrict__ out_idx,
    float *__restrict__ out_dist, uint32_t *__restrict__ out_count, uint32_t *__restrict__ flags)
{
    extern __shared__ __align__(16) unsigned char mg_smem[];
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(mg_smem);
    __shared__ uint32_t s_valid, s_tie;
    const int tid = threadIdx.x;
    const uint32_t q = blockIdx.x;
    if (tid == 0) {
        s_valid = 0;
        s_tie = 0;
    }
    __syncthreads();
    uint32_t P = 1;
    while (P < per_query) P <<= 1;
    const unsigned long long *c = cand + (size_t)q * per_query;
    uint32_t valid = 0;
    for (uint32_t j = tid; j < P; j += 256) {
        unsigned long long v = j < per_query ? c[j] : ~0ull;
        keys[j] = v;
        valid += v != ~0ull;
    }
    atomicAdd(&s_valid, valid);
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < P / 2; t += 256) {
                uint32_t lo = bitonic_lo(t, stride), hi = lo + stride;
                bool asc = (lo & size) == 0;
                unsigned long long a = keys[lo], bb = keys[hi];
                if ((a > bb) == asc) {
                    keys[lo] = bb;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    const uint32_t nv = s_valid;
    const uint32_t m = k < nv ? k : nv;
    const uint32_t K1 = (k + 1 < nv) ? k + 1 : nv;
    const float inf = __int_as_float(0x7f800000);
    for (uint32_t j = tid; j < k; j += 256) {
        if (j < m) {
            out_idx[(size_t)q * k + j] = (uint32_t)keys[j];
            out_dist[(size_t)q * k + j] = unkey((uint32_t)(keys[j] >> 32));
        } else {
            out_idx[(size_t)q * k + j] = ~0ull;
            out_dist[(size_t)q * k + j] = inf;
        }
    }
    for (uint32_t j = tid; j + 1 < K1; j += 256)
        if ((keys[j] >> 32) == (keys[j + 1] >> 32)) s_tie = 1;
    __syncthreads();
    if (tid == 0) {
        out_count[q] = m;
        const uint32_t cf = coarse_flags[q];
        const bool edge = nv > k && (keys[k - 1] >> 32) == (keys[k] >> 32);
        flags[q] = ((s_tie || (cf & ROW_TIE)) ? ROW_TIE : 0u) |
                   ((edge || (cf & ROW_TIE_EDGE)) ? ROW_TIE_EDGE : 0u);
    }
}

// ---- std BinaryHeap (max-heap on distance) over a per-thread array
struct GHeap {
    float *d;
    uint32_t *ix;
    uint32_t len;
    __device__ void sift_up(uint32_t start, uint32_t pos)
    {
        float ed = d[pos];
        uint32_t ei = ix[pos];
        while (pos > start) {
            uint32_t parent = (pos - 1) / 2;
            if (ed <= d[parent]) break;
            d[pos] = d[parent];
            ix[pos] = ix[parent];
            pos = parent;
        }
        d[pos] = ed;
        ix[pos] = ei;
    }
    __device__ void sift_down_to_bottom(uint32_t pos)
    {
        uint32_t end = len, start = pos;
        float ed = d[pos];
        uint32_t ei = ix[pos];
        uint32_t child = 2 * pos + 1;
        while (end >= 2 && child < end - 1) {
            child += (d[child] <= d[child + 1]) ? 1u : 0u;
            d[pos] = d[child];
            ix[pos] = ix[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (end >= 1 && child == end - 1) {
            d[pos] = d[child];
            ix[pos] = ix[child];
            pos = child;
        }
        d[pos] = ed;
        ix[pos] = ei;
        sift_up(start, pos);
    }
    __device__ void push(float x, uint32_t i)
    {
        d[len] = x;
        ix[len] = i;
        len += 1;
        sift_up(0, len - 1);
    }
    __device__ void pop(float *od, uint32_t *oi)
    {
        float ld = d[len - 1];
        uint32_t li = ix[len - 1];
        len -= 1;
        if (len > 0) {
            *od = d[0];
            *oi = ix[0];
            d[0] = ld;
            ix[0] = li;
            sift_down_to_bottom(0);
        } else {
            *od = ld;
            *oi = li;
        }
    }
};

constexpr int RP_THREADS = 256;

// Exact replay of IVFPQIndex::search_knn_adc + node_search_k
// (pq_idx.rs:408-477) for tied queries, with the heaps in LDS (each run by
// one wave as a WaveHeap), split three ways so that a tied query's
// probed lists are simulated in parallel:
//   ivf_order_kernel: the coarse heap of pq_idx.rs:416-425 (or the given
//     probes) -> order[t][p], np[t]; one wave per tied query;
//   ivf_list_replay_kernel: one workgroup per (tied query, probed list):
//     LUT, ADC, and the list's push-all / pop-to-k heap (pq_idx.rs:181-191)
//     drained in pop order into seq[t][p] (a tie-free list by a sort);
//   ivf_merge_replay_kernel: one wave per tied query pushes every list's
//     sequence, in probe order, into the top-k heap (pq_idx.rs:426-437) and
//     drains it (node_search_k, pq_idx.rs:458-477).
__global__ __launch_bounds__(64) void ivf_order_kernel(
    const uint32_t *__restrict__ rows, const float *__restrict__ Dc, uint32_t ldd, uint32_t nlist,
    uint32_t nprobe, const uint32_t *__restrict__ probes, uint32_t ldp,
    const uint32_t *__restrict__ nprobe_q, uint32_t ldo, uint32_t *__restrict__ order,
    uint32_t *__restrict__ np_out, const uint32_t *__restrict__ coarse_flags)
{
    extern __shared__ __align__(16) unsigned char io_smem[];
    const uint32_t t = blockIdx.x, q = rows[t];
    const int tid = threadIdx.x;
    uint32_t *ord = order + (size_t)t * ldo;
    // The given probes (learned policy), or the coarse select's sorted
    // candidates when the coarse check found no tie that can change the
    // result (an empty list may swap places with another: it adds nothing)
    if (probes && (!coarse_flags || !(coarse_flags[q] & ROW_TIE))) {
        const uint32_t np_q = nprobe_q ? nprobe_q[q] : (nprobe < nlist ? nprobe : nlist);
        for (uint32_t p = tid; p < np_q; p += 64) ord[p] = probes[(size_t)q * ldp + p];
        if (tid == 0) np_out[t] = np_q;
        return;
    }
    WaveHeap<WhPackedLds> coarse({reinterpret_cast<uint2 *>(io_smem)}, (uint32_t)tid);  // [nlist + 2]
    const float *drow = Dc + (size_t)q * ldd;
    for (uint32_t i0 = 0; i0 < nlist; i0 += 64) {  // 64 distances per load round
        const float v = i0 + tid < nlist ? -drow[i0 + tid] : 0.0f;
        const uint32_t cnt = min(64u, nlist - i0);
        for (uint32_t u = 0; u < cnt; ++u)
            coarse.push(make_uint2(i0 + u, (uint32_t)__builtin_amdgcn_readlane(__float_as_int(v), (int)u)));
    }
    uint32_t p = 0;
    for (; p < nprobe && coarse.len > 0; ++p) {
        const uint2 x = coarse.pop();
        if (tid == 0) ord[p] = x.x;
    }
    if (tid == 0) np_out[t] = p;
}

__global__ __launch_bounds__(RP_THREADS) void ivf_list_replay_kernel(
    const uint32_t *__restrict__ rows, const float *__restrict__ Qp, uint32_t dpad,
    const float *__restrict__ C, const uint32_t *__restrict__ off, const uint32_t *__restrict__ ids,
    const uint8_t *__restrict__ codes, const float *__restrict__ codebooks,
    const uint32_t *__restrict__ sub_b, const uint32_t *__restrict__ sub_e, uint32_t M,
    uint32_t KS, uint32_t sdm, uint32_t k, uint32_t Lcap, uint32_t Pcap,
    const uint32_t *__restrict__ order, uint32_t ldo, const uint32_t *__restrict__ np,
    uint2 *__restrict__ seq, uint32_t *__restrict__ seq_cnt)
{
    extern __shared__ __align__(16) unsigned char lr_smem[];
    __shared__ uint32_t s_tied;
    const uint32_t p = blockIdx.x, t = blockIdx.y;
    const int tid = threadIdx.x;
    uint2 *out = seq + ((size_t)t * ldo + p) * k;
    if (p >= np[t]) {
        if (tid == 0) seq_cnt[(size_t)t * ldo + p] = 0;
        return;
    }
    const uint32_t q = rows[t], l = order[(size_t)t * ldo + p];
    unsigned long long *sk = reinterpret_cast<unsigned long long *>(lr_smem);  // [Pcap]
    uint2 *he = reinterpret_cast<uint2 *>(sk + ((Pcap + 1) & ~1u));            // [Lcap + 1], 16-byte aligned
    float *lut = reinterpret_cast<float *>(he + Lcap + 2);                     // [M*KS]
    const float *qv = Qp + (size_t)q * dpad;
    const float *cb = codebooks + (size_t)l * M * KS * sdm;
    for (uint32_t e = tid; e < M * KS; e += RP_THREADS) {
        const uint32_t sub = e / KS, b = sub_b[sub], sd = sub_e[sub] - b;
        float acc = -0.0f;
        for (uint32_t j = 0; j < sd; ++j) {
            float zj = qv[b + j] - C[(size_t)l * dpad + b + j];
            float df = zj - cb[(size_t)e * sdm + j];
            acc = acc + df * df;
        }
        lut[e] = acc;
    }
    if (tid == 0) s_tied = 0;
    __syncthreads();
    const uint32_t base = off[l], L = off[l + 1] - base;
    uint32_t P = 1;
    while (P < L) P <<= 1;
    for (uint32_t i = tid; i < P; i += RP_THREADS) {
        if (i < L) {
            const uint8_t *cdp = codes + (size_t)(base + i) * M;
            float dist = 0.0f;
            for (uint32_t j = 0; j < M; ++j) dist = dist + lut[j * KS + cdp[j]];
            he[i + 1] = make_uint2(i, __float_as_uint(dist));
            sk[i] = ((unsigned long long)fkey(dist) << 32) | i;
        } else {
            sk[i] = ~0ull;
        }
    }
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t u = tid; u < P / 2; u += RP_THREADS) {
                const uint32_t a0 = bitonic_lo(u, stride), a1 = a0 + stride;
                const bool asc = (a0 & size) == 0;
                const unsigned long long a = sk[a0], bb = sk[a1];
                if ((a > bb) == asc) {
                    sk[a0] = bb;
                    sk[a1] = a;
                }
            }
            __syncthreads();
        }
    const uint32_t K1 = k + 1 < L ? k + 1 : L;
    for (uint32_t j = tid; j + 1 < K1; j += RP_THREADS)
        if ((uint32_t)(sk[j] >> 32) == (uint32_t)(sk[j + 1] >> 32)) s_tied = 1;
    __syncthreads();
    if (tid >= 64) return;
    const uint32_t m = k < L ? k : L;
    if (!s_tied) {
        // survivors: the m smallest, popped in descending order
        for (uint32_t j = tid; j < m; j += 64) {
            const unsigned long long v = sk[m - 1 - j];
            out[j] = make_uint2(__float_as_uint(unkey((uint32_t)(v >> 32))),
                                ids[base + (uint32_t)v]);
        }
    } else {
        // item i sits at position i until pushed, and a push writes only
        // positions up to its own: 64 items are read per load round
        WaveHeap<WhPackedLds> lh({he}, (uint32_t)tid);
        for (uint32_t i0 = 0; i0 < L; i0 += 64) {
            const uint2 v = i0 + tid < L ? he[i0 + tid + 1] : make_uint2(0u, 0u);
            const uint32_t cnt = min(64u, L - i0);
            for (uint32_t u = 0; u < cnt; ++u) lh.push(wh_lane(v, u));
        }
        while (lh.len > k) (void)lh.pop();
        for (uint32_t j = 0; lh.len > 0; ++j) {
            const uint2 x = lh.pop();
            if (tid == 0) out[j] = make_uint2(x.y, ids[base + x.x]);
        }
    }
    if (tid == 0) seq_cnt[(size_t)t * ldo + p] = m;
}

__global__ __launch_bounds__(64) void ivf_merge_replay_kernel(
    const uint32_t *__restrict__ rows, uint32_t k, uint32_t ldo, const uint32_t *__restrict__ np,
    const uint2 *__restrict__ seq, const uint32_t *__restrict__ seq_cnt,
    uint64_t *__restrict__ out_idx, float *__restrict__ out_dist, uint32_t *__restrict__ out_count)
{
    extern __shared__ __align__(16) unsigned char mr_smem[];
    const uint32_t t = blockIdx.x, q = rows[t];
    const int tid = threadIdx.x;
    WaveHeap<WhPackedLds> top({reinterpret_cast<uint2 *>(mr_smem)}, (uint32_t)tid);  // [k + 3]
    const uint32_t lane = (uint32_t)tid;
    // Once the heap holds k entries, an element above the root climbs to the
    // root and is popped again: a map F of the layout that does not depend
    // on the element's value.  A list's sequence is in pop order (largest
    // first), so it starts with a run of j such elements and every later
    // element is at most the root.  F^j: nothing while the layout is stable
    // (the round trip retraces the push's path), else F, F^2, ... until the
    // layout is stable or repeats one of the last RT_HIST, then the rest by
    // the period.
    constexpr uint32_t RT_HIST = 8;
    auto round_trip = [&]() {
        top.push(make_uint2(~0u, 0x7f800000u));  // +inf beats every finite entry
        (void)top.pop();
    };
    // heap_noop_roundtrip for the full heap, lane j checking step j of the
    // walk from position L = len to the root
    auto noop = [&]() -> bool {
        const uint32_t L = top.len;
        if (L == 0) return true;
        const uint32_t steps = 32u - (uint32_t)__clz(L + 1);
        const bool act = lane < steps;
        const uint32_t a = act ? ((L + 1) >> lane) - 1 : 0u;
        const uint32_t pp = a > 0 ? (a - 1) / 2 : 0u;
        const float vp = __uint_as_float(top.s.ld(pp).y);
        const float vq = __uint_as_float(top.s.ld(lane == 0 ? L - 1 : min(a + 1, L - 1)).y);
        bool bad = false;
        if (act && a > 0) {
            if (lane == 0) bad = a == 2 * pp + 2 && !(vp > vq);
            else bad = a == 2 * pp + 1 && a + 1 < L && !(vp > vq);
        }
        return __ballot(bad) == 0;
    };
    auto layout = [&](uint32_t pos) { return pos < top.len ? top.s.ld(pos).x : 0u; };
    auto apply_runs = [&](uint32_t j) {
        if (j == 0 || noop()) return;
        // heaps of more than 128 (the layouts are compared up to 128), and
        // short runs, whose round trips cost less than the layouts
        if (top.len > 128 || j <= 2 * RT_HIST) {
            for (uint32_t t2 = 0; t2 < j && !noop(); ++t2) round_trip();
            return;
        }
        uint32_t h0[RT_HIST], h1[RT_HIST];  // items at positions lane, lane + 64
        h0[0] = layout(lane);
        h1[0] = layout(lane + 64);
        for (uint32_t t2 = 1; t2 <= j; ++t2) {
            round_trip();
            if (noop()) return;
            const uint32_t c0 = layout(lane), c1 = layout(lane + 64);
            const uint32_t seen = t2 < RT_HIST ? t2 : RT_HIST;
            uint32_t prev = RT_HIST;
            for (uint32_t q2 = 0; q2 < seen; ++q2)
                if (prev == RT_HIST && __ballot(h0[q2] != c0 || h1[q2] != c1) == 0) prev = q2;
            if (prev != RT_HIST) {
                const uint32_t left = (j - t2) % (t2 - prev);
                for (uint32_t u = 0; u < left; ++u) round_trip();
                return;
            }
            if (t2 < RT_HIST) {
                h0[t2] = c0;
                h1[t2] = c1;
            }
        }
    };
    const uint32_t np_q = np[t];
    for (uint32_t p = 0; p < np_q; ++p) {
        const uint2 *sq = seq + ((size_t)t * ldo + p) * k;
        const uint32_t m = seq_cnt[(size_t)t * ldo + p];
        uint32_t first = 0;
        if (k > 0 && top.len == k && m > 0) {  // the run above the root
            const float root = __uint_as_float(top.s.ld(0).y);
            uint32_t j = m;
            for (uint32_t c0 = 0; c0 < m; c0 += 64) {
                const bool in = c0 + lane < m;
                const uint64_t live =
                    __ballot(in && !(__uint_as_float(sq[min(c0 + lane, m - 1)].x) > root));
                if (live) {
                    j = c0 + (uint32_t)__builtin_ctzll(live);
                    break;
                }
            }
            apply_runs(j);
            first = j;
        }
        // 64 entries per load round, handed out through register reads
        for (uint32_t j0 = first; j0 < m; j0 += 64) {
            const uint2 v = j0 + tid < m ? sq[j0 + tid] : make_uint2(0u, 0u);
            const uint32_t cnt = min(64u, m - j0);
            for (uint32_t u = 0; u < cnt; ++u) {
                const uint32_t xd = (uint32_t)__builtin_amdgcn_readlane((int)v.x, (int)u);
                const uint32_t xi = (uint32_t)__builtin_amdgcn_readlane((int)v.y, (int)u);
                top.push(make_uint2(xi, xd));
                if (top.len > k) (void)top.pop();
            }
        }
    }
    const uint32_t cnt = top.len;
    for (uint32_t p = cnt; p > 0; --p) {
        const uint2 x = top.pop();
        if (tid == 0) {
            out_idx[(size_t)q * k + p - 1] = x.x;
            out_dist[(size_t)q * k + p - 1] = __uint_as_float(x.y);
        }
    }
    const float inf = __int_as_float(0x7f800000);
    for (uint32_t j = cnt + tid; j < k; j += 64) {
        out_idx[(size_t)q * k + j] = ~0ull;
        out_dist[(size_t)q * k + j] = inf;
    }
    if (tid == 0) out_count[q] = cnt;
}

// The same replay in global memory for lists too long for LDS, one lane per
// query: coarse heap of Neighbor(i, -dist) over all lists, nprobe pops, per
// list an unbounded heap trimmed to k, popped into the global heap trimmed
// to k.
__global__ void ivf_replay_kernel(const uint32_t *__restrict__ rows, uint32_t nrows,
                                  const float *__restrict__ Dc, uint32_t ldd, uint32_t nlist,
                                  uint32_t nprobe, const float *__restrict__ Qp, uint32_t dpad,
                                  const float *__restrict__ C, const uint32_t *__restrict__ off,
                                  const uint32_t *__restrict__ ids,
                                  const uint8_t *__restrict__ codes,
                                  const float *__restrict__ codebooks,
                                  const uint32_t *__restrict__ sub_b,
                                  const uint32_t *__restrict__ sub_e, uint32_t M, uint32_t KS,
                                  uint32_t sdm, uint32_t k, uint32_t lmax, float *scratch_f,
                                  uint32_t *scratch_u, size_t per_thread,
                                  const uint32_t *__restrict__ probes, uint32_t ldp,
                                  const uint32_t *__restrict__ nprobe_q, uint64_t *out_idx,
                                  float *out_dist, uint32_t *out_count)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const uint32_t q = rows[t];
    float *fbase = scratch_f + (size_t)t * per_thread;
    uint32_t *ubase = scratch_u + (size_t)t * per_thread;
    // layout: coarse heap [nlist] | list heap [lmax] | global heap [k+1] | lut [M*KS]
    GHeap coarse{fbase, ubase, 0};
    GHeap lh{fbase + nlist, ubase + nlist, 0};
    GHeap top{fbase + nlist + lmax, ubase + nlist + lmax, 0};
    float *lut = fbase + nlist + lmax + (k + 1);
    const float *qv = Qp + (size_t)q * dpad;
    uint32_t np_q = nprobe;
    if (probes) {
        np_q = nprobe_q[q];
    } else {
        const float *drow = Dc + (size_t)q * ldd;
        for (uint32_t i = 0; i < nlist; ++i) coarse.push(-drow[i], i);
    }
    for (uint32_t p = 0; p < np_q && (probes || coarse.len > 0); ++p) {
        float xd;
        uint32_t l;
        if (probes) l = probes[(size_t)q * ldp + p];
        else coarse.pop(&xd, &l);
        const float *cb = codebooks + (size_t)l * M * KS * sdm;
        for (uint32_t e = 0; e < M * KS; ++e) {
            uint32_t s = e / KS, b = sub_b[s], sd = sub_e[s] - b;
            float acc = -0.0f;
            for (uint32_t j = 0; j < sd; ++j) {
                float zj = qv[b + j] - C[(size_t)l * dpad + b + j];
                float df = zj - cb[(size_t)e * sdm + j];
                acc = acc + df * df;
            }
            lut[e] = acc;
        }
        const uint32_t base = off[l], L = off[l + 1] - base;
        lh.len = 0;
        for (uint32_t i = 0; i < L; ++i) {
            const uint8_t *cd = codes + (size_t)(base + i) * M;
            float dist = 0.0f;
            for (uint32_t j = 0; j < M; ++j) dist = dist + lut[j * KS + cd[j]];
            lh.push(dist, i);
        }
        while (lh.len > k) {
            float a;
            uint32_t b;
            lh.pop(&a, &b);
        }
        while (lh.len > 0) {
            float a;
            uint32_t b;
            lh.pop(&a, &b);
            top.push(a, ids[base + b]);
            if (top.len > k) {
                float a2;
                uint32_t b2;
                top.pop(&a2, &b2);
            }
        }
    }
    const uint32_t cnt = top.len;
    for (uint32_t p = cnt; p > 0; --p) {
        float a;
        uint32_t b;
        top.pop(&a, &b);
        out_idx[(size_t)q * k + p - 1] = b;
        out_dist[(size_t)q * k + p - 1] = a;
    }
    const float inf = __int_as_float(0x7f800000);
    for (uint32_t j = cnt; j < k; ++j) {
        out_idx[(size_t)q * k + j] = ~0ull;
        out_dist[(size_t)q * k + j] = inf;
    }
    out_count[q] = cnt;
}

__global__ void negate_kernel(const float *__restrict__ x, float *__restrict__ y, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        y[i] = -x[i];
}

// nprobe_q[q] = smallest m with sum of the m largest probabilities >= mass,
// clamped to [pmin, pmax]; the candidates hold -p sorted ascending.
__global__ void probe_count_kernel(const float *__restrict__ negp, uint32_t ld, uint32_t rows,
                                   uint32_t pmax, float mass, uint32_t pmin,
                                   uint32_t *__restrict__ nprobe_q, const float *__restrict__ lse)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rows) return;
    float acc = 0.f;
    uint32_t m = pmax;
    for (uint32_t j = 0; j < pmax; ++j) {
        // -p, or the negated logit z with p = exp(z - lse)
        const float v = negp[(size_t)q * ld + j];
        acc += lse ? expf(-v - lse[q]) : -v;
        if (acc >= mass) {
            m = j + 1;
            break;
        }
    }
    nprobe_q[q] = m < pmin ? pmin : m;
}

// ids[i] -> list of the item; T[q][list] += 1/k (dense targets of a batch).
__global__ void dense_targets_kernel(const uint32_t *__restrict__ lab, uint32_t k,
                                     uint32_t rows, uint32_t nlist, float *__restrict__ T)
{
    const uint32_t q = blockIdx.x;
    for (uint32_t j = threadIdx.x; j < nlist; j += blockDim.x) T[(size_t)q * nlist + j] = 0.f;
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t j = 0; j < k; ++j) {
            uint32_t l = lab[(size_t)q * k + j];
            if (l < nlist) T[(size_t)q * nlist + l] += 1.0f / (float)k;
        }
}

__global__ void ids_to_lists_kernel(const uint64_t *__restrict__ ids, uint64_t n,
                                    const uint32_t *__restrict__ assign, uint64_t n_items,
                                    uint32_t *__restrict__ out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t id = ids[i];
        out[i] = id < n_items ? assign[id] : 0xFFFFFFFFu;
    }
}

// split_center (kmeans.rs:150-193) of the coarse k-means, on the device,
// one 1,024-thread workgroup.  counts[c] = list sizes from off, pk[c] =
// counts[c] / n (the walk's pick probability), both in global scratch since
// nc may exceed LDS.  The empty centres are visited in order; the walk for
// centre i draws once per step, step t looking at centre (i + 1 + t) % nc
// with the generator's (t + 1)-th next draw, and stops at the first
// u < pk.  Each round every thread takes one step, the waves' first hits
// meet in LDS, and the generator then advances by the steps taken: the same
// draws and picks as the step-by-step walk.  *rng carries the generator
// from epoch to epoch.
constexpr int CS_THREADS = 1024;

__global__ __launch_bounds__(CS_THREADS) void coarse_split_kernel(
    float *__restrict__ C, uint32_t ldc, uint32_t nc, uint32_t dim,
    const uint32_t *__restrict__ off, uint64_t n, uint64_t *__restrict__ rng_state,
    uint32_t *__restrict__ cnt, double *__restrict__ pk)
{
    __shared__ uint32_t s_first[2][CS_THREADS / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr uint32_t NW = CS_THREADS / 64;
    for (uint32_t c = tid; c < nc; c += CS_THREADS) {
        const uint32_t v = off[c + 1] - off[c];
        cnt[c] = v;
        pk[c] = (double)v / (double)n;
    }
    __syncthreads();
    uint64_t rng = *rng_state;
    const float EPS = 1.0f / 1024.0f;
    uint32_t par = 0;
    for (uint32_t i = 0; i < nc; ++i) {
        if (cnt[i] != 0) continue;
        uint32_t steps = 0;
        for (uint32_t t0 = 0;; t0 += CS_THREADS) {
            const uint32_t t = t0 + tid;
            const uint32_t c = (uint32_t)(((uint64_t)i + 1 + t) % nc);
            const double pc = pk[c];
            const bool hit = pc != 0.0 && splitmix_unit_at(rng, t) < pc;
            const uint64_t m = __ballot(hit);
            if (lane == 0) s_first[par][wave] = m ? t0 + wave * 64 + (uint32_t)__builtin_ctzll(m) : ~0u;
            __syncthreads();
            uint32_t first = ~0u;
#pragma unroll
            for (uint32_t u = 0; u < NW; ++u) first = min(first, s_first[par][u]);
            par ^= 1;  // the next round writes the other buffer
            if (first != ~0u) {
                steps = first;
                break;
            }
        }
        const uint32_t sc = (uint32_t)(((uint64_t)i + 1 + steps) % nc);
        rng += ((uint64_t)steps + 1) * 0x9E3779B97F4A7C15ull;
        for (uint32_t j = tid; j < dim; j += CS_THREADS) {
            const float src = C[(size_t)sc * ldc + j];
            if (j % 2 == 0) {
                C[(size_t)i * ldc + j] = src * (1.0f - EPS);
                C[(size_t)sc * ldc + j] = src * (1.0f + EPS);
            } else {
                C[(size_t)i * ldc + j] = src * (1.0f + EPS);
                C[(size_t)sc * ldc + j] = src * (1.0f - EPS);
            }
        }
        if (tid == 0) {
            const uint32_t half = cnt[sc] / 2;
            cnt[i] = half;
            cnt[sc] -= half;
            pk[i] = (double)half / (double)n;
            pk[sc] = (double)cnt[sc] / (double)n;
        }
        __syncthreads();
    }
    if (tid == 0) *rng_state = rng;
}

inline unsigned grid_for(uint64_t work, unsigned cap = 8192)
{
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((work + 255) / 256, cap));
}

}  // namespace

void launch_coarse_init(bool fp64, const float *X, uint32_t ldx, uint64_t n, uint32_t d,
                        uint32_t nc, float *C, uint32_t ldc, SegScratch &ws, hipStream_t s)
{
    ProfScope prof("kmeans_init", s);
    if (!fp64) {
        coarse_init_kernel<float><<<ceil_div(ldc, 64), 64, 0, s>>>(X, ldx, n, d, nc, C, ldc);
        HM_CHECK_LAUNCH();
        return;
    }
    if (n == 0) fail(HM_ERR_EMPTY, "no rows to average");
    const uint64_t nch = (n + SEG_ROWS - 1) / SEG_ROWS;
    if (nch > 0x7FFFFFFFull) fail(HM_ERR_CAPACITY, "too many rows");
    ws.part.reserve((size_t)nch * std::max(d, 1u) * 8);
    if (d > 0) {
        seg_partial_kernel<<<dim3((unsigned)nch, ceil_div(d, 64)), 256, 0, s>>>(
            X, ldx, nullptr, nullptr, nullptr, 1, n, d, ws.part.as<double>());
        HM_CHECK_LAUNCH();
    }
    seg_mean_kernel<<<dim3(1, ceil_div(ldc, 64)), 256, 0, s>>>(ws.part.as<double>(), nullptr,
                                                               nullptr, n, d, nc, C, ldc);
    HM_CHECK_LAUNCH();
}

void launch_unpack_assign(const unsigned long long *best, uint64_t n, uint32_t *assign,
                          hipStream_t s)
{
    unpack_assign_kernel<<<grid_for(n), 256, 0, s>>>(best, n, assign);
    HM_CHECK_LAUNCH();
}

void build_lists(const uint32_t *assign, uint64_t n, uint32_t nlist, uint32_t *sorted_keys,
                 uint32_t *ids, uint32_t *off, DeviceBuffer &iota, DeviceBuffer &tmp,
                 hipStream_t s)
{
    ProfScope prof("build_lists", s);
    iota.reserve(n * sizeof(uint32_t));
    iota_kernel<<<grid_for(n), 256, 0, s>>>(iota.as<uint32_t>(), n);
    HM_CHECK_LAUNCH();
    size_t tb = 0;
    int bits = 1;
    while ((1u << bits) < nlist + 1 && bits < 32) ++bits;
    HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, assign, sorted_keys,
                                              iota.as<uint32_t>(), ids, (int)n, 0, bits, s));
    tmp.reserve(tb);
    HM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tb, assign, sorted_keys,
                                              iota.as<uint32_t>(), ids, (int)n, 0, bits, s));
    offsets_kernel<<<ceil_div(nlist + 1, 256), 256, 0, s>>>(sorted_keys, n, nlist, off);
    HM_CHECK_LAUNCH();
}

void launch_update(bool fp64, const float *X, uint32_t ldx, const uint32_t *off,
                   const uint32_t *ids, uint32_t nlist, uint32_t d, float *C, uint32_t ldc,
                   uint64_t n, SegScratch &ws, hipStream_t s)
{
    ProfScope prof("kmeans_update", s);
    if (!fp64) {
        update_kernel<float><<<dim3(nlist, ceil_div(ldc, 64)), 64, 0, s>>>(X, ldx, off, ids, d,
                                                                           C, ldc);
        HM_CHECK_LAUNCH();
        return;
    }
    // chunk table: cstart = exclusive scan of the per-list chunk counts
    ws.nch.reserve((size_t)(nlist + 1) * 4);
    ws.cstart.reserve((size_t)(nlist + 1) * 4);
    seg_chunk_count_kernel<<<ceil_div(nlist + 1, 256), 256, 0, s>>>(off, nlist,
                                                                      ws.nch.as<uint32_t>());
    HM_CHECK_LAUNCH();
    size_t tb = 0;
    HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ws.nch.as<uint32_t>(),
                                            ws.cstart.as<uint32_t>(), (int)(nlist + 1), s));
    ws.tmp.reserve(tb);
    HM_HIP(hipcub::DeviceScan::ExclusiveSum(ws.tmp.get(), tb, ws.nch.as<uint32_t>(),
                                            ws.cstart.as<uint32_t>(), (int)(nlist + 1), s));
    // every list holds at most one partial chunk, so the total is bounded by
    const uint64_t max_chunks = n / SEG_ROWS + nlist;
    if (max_chunks > 0x7FFFFFFFull) fail(HM_ERR_CAPACITY, "too many rows");
    if (d > 0 && max_chunks > 0) {
        ws.part.reserve((size_t)max_chunks * d * 8);
        seg_partial_kernel<<<dim3((unsigned)max_chunks, ceil_div(d, 64)), 256, 0, s>>>(
            X, ldx, ids, off, ws.cstart.as<uint32_t>(), nlist, n, d, ws.part.as<double>());
        HM_CHECK_LAUNCH();
    }
    seg_mean_kernel<<<dim3(nlist, ceil_div(ldc, 64)), 256, 0, s>>>(
        ws.part.as<double>(), ws.cstart.as<uint32_t>(), off, n, d, 0, C, ldc);
    HM_CHECK_LAUNCH();
}

void launch_residuals(const float *X, uint32_t ldx, const uint32_t *ids, const uint32_t *lists,
                      const float *C, uint64_t n, uint32_t dpad, float *R, hipStream_t s)
{
    residual_kernel<<<grid_for(n * dpad), 256, 0, s>>>(X, ldx, ids, lists, C, n, dpad, R);
    HM_CHECK_LAUNCH();
}

namespace {

template <typename ACC_T, int SDM>
void pq_train_launch(const float *R, uint32_t ldr, const uint32_t *off, const uint32_t *lorder,
                     uint32_t nlist, uint32_t M, const uint32_t *sub_b, const uint32_t *sub_e,
                     uint32_t KS, uint32_t sd_max, uint32_t sdm, uint32_t epochs, uint64_t seed,
                     uint32_t max_list, uint64_t n, uint16_t *asg, uint32_t *ord,
                     float *codebooks, uint8_t *codes, hipStream_t s)
{
    constexpr size_t LDS_BUDGET = 52 * 1024;  // three workgroups per CU
    const size_t fixed = (size_t)KS * 8 + (size_t)KS * SDM * 4 + (size_t)KS * 3 * 4 +
                         (size_t)(PQ_THREADS / 64) * KS * 2;
    const size_t per_item = (size_t)sd_max * 4 + 4 + 2;
    uint32_t cap = fixed + 16 < LDS_BUDGET ? (uint32_t)((LDS_BUDGET - fixed - 16) / per_item) : 0;
    cap = std::min(cap, max_list);
    const size_t lds = fixed + (size_t)cap * per_item + 16;
    pq_train_kernel<ACC_T, SDM><<<dim3(nlist * M), PQ_THREADS, lds, s>>>(
        R, ldr, off, lorder, M, sub_b, sub_e, KS, sdm, epochs, seed, cap, n, asg, ord, codebooks,
        codes);
    HM_CHECK_LAUNCH();
}

template <typename ACC_T>
void pq_train_dispatch(uint32_t sd_max, const float *R, uint32_t ldr, const uint32_t *off,
                       const uint32_t *lorder, uint32_t nlist, uint32_t M, const uint32_t *sub_b,
                       const uint32_t *sub_e, uint32_t KS, uint32_t sdm, uint32_t epochs,
                       uint64_t seed, uint32_t max_list, uint64_t n, uint16_t *asg, uint32_t *ord,
                       float *codebooks, uint8_t *codes, hipStream_t s)
{
#define HM_PQ_TRAIN(SDM)                                                                      \
    pq_train_launch<ACC_T, SDM>(R, ldr, off, lorder, nlist, M, sub_b, sub_e, KS, sd_max, sdm, \
                                epochs, seed, max_list, n, asg, ord, codebooks, codes, s)
    if (sd_max <= 4) HM_PQ_TRAIN(4);
    else if (sd_max <= 8) HM_PQ_TRAIN(8);
    else if (sd_max <= 16) HM_PQ_TRAIN(16);
    else HM_PQ_TRAIN(32);
#undef HM_PQ_TRAIN
}

}  // namespace

void launch_pq_train(bool fp64, const float *R, uint32_t ldr, const uint32_t *off,
                     const uint32_t *lorder, uint32_t nlist, uint32_t M, const uint32_t *sub_b,
                     const uint32_t *sub_e, uint32_t KS, uint32_t sd_max, uint32_t sdm,
                     uint32_t epochs, uint64_t seed, uint32_t max_list, uint64_t n,
                     uint16_t *asg_scratch, uint32_t *ord_scratch, float *codebooks,
                     uint8_t *codes, hipStream_t s)
{
    if (sd_max > (uint32_t)PQ_MAXSD) fail(HM_ERR_NOT_IMPLEMENTED, "sub-space dimension above 32");
    if (KS > 256) fail(HM_ERR_NOT_IMPLEMENTED, "sub_bits above 8");
    if (n > 0xFFFFFFFFull) fail(HM_ERR_CAPACITY, "too many rows");
    ProfScope prof("pq_train", s);
    if (fp64)
        pq_train_dispatch<double>(sd_max, R, ldr, off, lorder, nlist, M, sub_b, sub_e, KS, sdm,
                                  epochs, seed, max_list, n, asg_scratch, ord_scratch, codebooks,
                                  codes, s);
    else
        pq_train_dispatch<float>(sd_max, R, ldr, off, lorder, nlist, M, sub_b, sub_e, KS, sdm,
                                 epochs, seed, max_list, n, asg_scratch, ord_scratch, codebooks,
                                 codes, s);
}

uint32_t ivf_scan_pool(uint32_t k1) { return k1 <= 1024 ? 4096u : 8192u; }

void launch_ivf_scan(const float *Qp, uint32_t dpad, const float *C, const uint32_t *probes,
                     uint32_t ldp, uint32_t nprobe, uint32_t rows, const uint32_t *off,
                     const uint32_t *ids, const uint8_t *codes, const float *codebooks,
                     const uint32_t *sub_b, const uint32_t *sub_e, uint32_t M, uint32_t KS,
                     uint32_t sdm, uint32_t k1, const uint32_t *nprobe_q,
                     unsigned long long *out, hipStream_t s)
{
    if (rows == 0 || nprobe == 0) return;
    const uint32_t P = ivf_scan_pool(k1);
    if (k1 * 2 > P) fail(HM_ERR_INVALID_ARG, "k too large for the list scan (max 4095)");
    const size_t lds = (size_t)P * 8 + (size_t)M * KS * 4 + (size_t)dpad * 4;
    if (lds > 160 * 1024) fail(HM_ERR_INVALID_ARG, "list scan does not fit in LDS");
    ProfScope prof("ivf_scan", s);
    ivf_scan_kernel<<<dim3(nprobe, rows), 256, lds, s>>>(Qp, dpad, C, probes, ldp, nprobe, off,
                                                          ids, codes, codebooks, sub_b, sub_e, M,
                                                          KS, sdm, k1, P, nprobe_q, out);
    HM_CHECK_LAUNCH();
}

// LDS of the list-major scan: codebook, LUT, and one region shared by the
// residual, the select histogram and the pool of P candidates.
static size_t lists_lds(uint32_t M, uint32_t KS, uint32_t sdm, uint32_t dpad, uint32_t P)
{
    const size_t region = std::max<size_t>({(size_t)dpad * 4, (size_t)LM_BINS * 4, (size_t)P * 8});
    return ((size_t)M * KS * sdm + (size_t)M * KS) * 4 + region;
}

// Dynamic LDS available to the list-major scan: the per-workgroup limit
// (160 KiB on gfx950) less the kernel's own static LDS.
static size_t lists_lds_budget()
{
    static size_t budget = 0;
    if (!budget) {
        hipFuncAttributes attr{};
        HM_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(ivf_scan_lists_kernel)));
        int dev = 0, max_lds = 0;
        HM_HIP(hipGetDevice(&dev));
        HM_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t lim = std::min<size_t>((size_t)max_lds, 160 * 1024);
        budget = lim > attr.sharedSizeBytes ? lim - attr.sharedSizeBytes : 1;
    }
    return budget;
}

uint32_t ivf_scan_lists_pool(uint32_t M, uint32_t KS, uint32_t sdm, uint32_t dpad, uint32_t k1)
{
    const size_t budget = lists_lds_budget();
    uint32_t P = 8192;
    while (P > 512 && lists_lds(M, KS, sdm, dpad, P) > budget) P >>= 1;
    return (P >= 2 * k1 && lists_lds(M, KS, sdm, dpad, P) <= budget) ? P : 0;
}

// (query, probe) pairs bucketed by list: pairs[pair_off[l] ..] = q * nprobe
// + p, and work units of up to G pairs of one list (unit_off).  Returns G.
static uint32_t build_pairs(const uint32_t *probes, uint32_t ldp, uint32_t nprobe, uint32_t rows,
                            uint32_t nlist, const uint32_t *nprobe_q, uint32_t *pair_cnt,
                            uint32_t *pair_off, uint32_t *pair_fill, uint32_t *unit_off,
                            uint32_t *pairs, unsigned long long *out, uint32_t k1, hipStream_t s)
{
    const uint64_t total = (uint64_t)rows * nprobe;
    // about 2048 work units for the batch, whatever the list skew
    const uint32_t G = (uint32_t)std::max<uint64_t>(8, (total + 2047) / 2048);
    const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 4096);
    const unsigned fb = std::min<unsigned>(ceil_div(nlist, 256), 1024);
    fill_u32_small_kernel<<<fb, 256, 0, s>>>(pair_cnt, nlist, 0u);
    fill_u32_small_kernel<<<fb, 256, 0, s>>>(pair_fill, nlist, 0u);
    pair_count_kernel<<<blocks, 256, 0, s>>>(probes, ldp, nprobe, rows, nprobe_q, pair_cnt);
    pair_scan_kernel<<<1, 1024, 0, s>>>(pair_cnt, nlist, G, pair_off, unit_off);
    pair_fill_kernel<<<blocks, 256, 0, s>>>(probes, ldp, nprobe, rows, nprobe_q, pair_off,
                                            pair_fill, pairs, out, k1);
    HM_CHECK_LAUNCH();
    return G;
}

int ivf_adc_rows_variant(uint32_t M, uint32_t KS, uint32_t sdm)
{
    const uint64_t nent = (uint64_t)M * KS;
    if (sdm <= 4 && nent <= 16u * LM_THREADS) return 16;
    if (sdm <= 8 && nent <= 8u * LM_THREADS) return 8;
    if (sdm <= 16 && nent <= 4u * LM_THREADS) return 4;
    if (sdm <= 32 && nent <= 2u * LM_THREADS) return 2;
    return 0;
}

void launch_ivf_adc_rows(const float *Qp, uint32_t dpad, uint32_t dim, const float *C,
                         const uint32_t *probes, uint32_t ldp, uint32_t nprobe, uint32_t rows,
                         uint32_t nlist,
                         const uint32_t *off, const uint8_t *codes, const float *codebooks,
                         const uint32_t *sub_b, const uint32_t *sub_e, uint32_t M, uint32_t KS,
                         uint32_t sdm, const uint32_t *nprobe_q, uint32_t *pair_cnt,
                         uint32_t *pair_off, uint32_t *pair_fill, uint32_t *unit_off,
                         uint32_t *pairs, uint32_t *seg, uint32_t *row_n, float *D, uint64_t ldd,
                         hipStream_t s)
{
    if (rows == 0 || nprobe == 0) return;
    const int R = ivf_adc_rows_variant(M, KS, sdm);
    if (!R) fail(HM_ERR_INVALID_ARG, "row-form scan: codebook does not fit in registers");
    // every sub-space exactly R-variant wide (d a multiple of M, width a
    // multiple of 4): the guard-free LUT fold
    const bool full = dim % M == 0 && dim / M == (uint32_t)(64 / R) && (dim / M) % 4 == 0;
    const uint64_t total = (uint64_t)rows * nprobe;
    if (total > 0x7FFFFFFFull) fail(HM_ERR_INVALID_ARG, "too many probe pairs");
    const size_t lds = (((size_t)M * KS + 3) & ~(size_t)3) * 4 + (size_t)dpad * 4;
    if (lds > 64 * 1024) fail(HM_ERR_INVALID_ARG, "row-form scan: LUT does not fit in LDS");
    ProfScope prof("ivf_scan", s);
    ivf_seg_kernel<<<ceil_div(rows, 256), 256, 0, s>>>(probes, ldp, nprobe, rows, nprobe_q, off,
                                                       seg, row_n);
    const uint32_t G = build_pairs(probes, ldp, nprobe, rows, nlist, nprobe_q, pair_cnt, pair_off,
                                   pair_fill, unit_off, pairs, nullptr, 0, s);
    const unsigned units = (unsigned)((total + G - 1) / G + nlist);  // >= sum of ceil(cnt_l / G)
#define HM_ADC_ROWS(RR, SDD)                                                                     \
    if (full)                                                                                    \
        ivf_adc_rows_kernel<RR, SDD, true><<<units, LM_THREADS, lds, s>>>(                       \
            Qp, dpad, C, pair_off, unit_off, nlist, G, pairs, nprobe, off, codes, codebooks,     \
            sub_b, sub_e, M, KS, sdm, seg, D, ldd);                                              \
    else                                                                                         \
        ivf_adc_rows_kernel<RR, SDD, false><<<units, LM_THREADS, lds, s>>>(                      \
            Qp, dpad, C, pair_off, unit_off, nlist, G, pairs, nprobe, off, codes, codebooks,     \
            sub_b, sub_e, M, KS, sdm, seg, D, ldd)
    switch (R) {
    case 16: HM_ADC_ROWS(16, 4); break;
    case 8: HM_ADC_ROWS(8, 8); break;
    case 4: HM_ADC_ROWS(4, 16); break;
    default: HM_ADC_ROWS(2, 32); break;
    }
#undef HM_ADC_ROWS
    HM_CHECK_LAUNCH();
}

void launch_ivf_edge_fixup(const float *Dr, uint64_t ldr, const uint32_t *row_n,
                           const float *cand_d, uint32_t *cand_i, uint32_t ldc, uint32_t k,
                           const uint32_t *probes, uint32_t ldp, uint32_t nprobe,
                           const uint32_t *nprobe_q, const uint32_t *seg, const uint32_t *off,
                           const uint32_t *ids, uint32_t rows, hipStream_t s)
{
    if (rows == 0) return;
    ProfScope prof("ivf_merge", s);
    ivf_edge_fixup_kernel<<<rows, 256, 0, s>>>(Dr, ldr, row_n, cand_d, cand_i, ldc, k, probes, ldp,
                                               nprobe, nprobe_q, seg, off, ids);
    HM_CHECK_LAUNCH();
}

void launch_coarse_ties(const float *D, uint64_t ldd, uint32_t nlist, const float *pd,
                        const uint32_t *pi, uint32_t Kc, uint32_t nprobe, const uint32_t *off,
                        uint32_t *flags, uint32_t rows, hipStream_t s)
{
    if (rows == 0) return;
    coarse_ties_kernel<<<rows, 64, 0, s>>>(D, ldd, nlist, pd, pi, Kc, nprobe, off, flags);
    HM_CHECK_LAUNCH();
}

void launch_ivf_rows_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                              const uint32_t *sel_flags, const uint32_t *coarse_flags,
                              const uint32_t *probes, uint32_t ldp, uint32_t nprobe,
                              const uint32_t *nprobe_q, const uint32_t *seg,
                              const uint32_t *row_n, const uint32_t *off, const uint32_t *ids,
                              uint32_t rows, uint32_t k, uint64_t *out_idx, float *out_dist,
                              uint32_t *out_count, uint32_t *flags, hipStream_t s)
{
    if (rows == 0) return;
    ProfScope prof("ivf_merge", s);
    ivf_rows_finalize_kernel<<<rows, 64, 0, s>>>(cand_d, cand_i, ldc, sel_flags, coarse_flags,
                                                 probes, ldp, nprobe, nprobe_q, seg, row_n, off,
                                                 ids, k, out_idx, out_dist, out_count, flags);
    HM_CHECK_LAUNCH();
}

void launch_ivf_scan_lists(const float *Qp, uint32_t dpad, const float *C,
                           const uint32_t *probes, uint32_t ldp, uint32_t nprobe, uint32_t rows,
                           uint32_t nlist, const uint32_t *off, const uint32_t *ids,
                           const uint8_t *codes, const float *codebooks, const uint32_t *sub_b,
                           const uint32_t *sub_e, uint32_t M, uint32_t KS, uint32_t sdm,
                           uint32_t k1, const uint32_t *nprobe_q, uint32_t *pair_cnt,
                           uint32_t *pair_off, uint32_t *pair_fill, uint32_t *unit_off,
                           uint32_t *pairs, unsigned long long *out, hipStream_t s)
{
    if (rows == 0 || nprobe == 0) return;
    const uint32_t P = ivf_scan_lists_pool(M, KS, sdm, dpad, k1);
    if (!P || lists_lds(M, KS, sdm, dpad, P) > lists_lds_budget())
        fail(HM_ERR_INVALID_ARG, "list-major scan does not fit in LDS");
    const uint64_t total = (uint64_t)rows * nprobe;
    if (total > 0x7FFFFFFFull) fail(HM_ERR_INVALID_ARG, "too many probe pairs");
    ProfScope prof("ivf_scan", s);
    const uint32_t G = build_pairs(probes, ldp, nprobe, rows, nlist, nprobe_q, pair_cnt, pair_off,
                                   pair_fill, unit_off, pairs, out, k1, s);
    const uint64_t units_max = (total + G - 1) / G + nlist;  // sum of ceil(cnt_l / G)
    ivf_scan_lists_kernel<<<(unsigned)units_max, LM_THREADS, lists_lds(M, KS, sdm, dpad, P), s>>>(
        Qp, dpad, C, pair_off, unit_off, nlist, G, pairs, nprobe, off, ids, codes, codebooks,
        sub_b, sub_e, M, KS, sdm, k1, P, out);
    HM_CHECK_LAUNCH();
}

void launch_ivf_merge(const unsigned long long *cand, uint32_t per_query, uint32_t rows,
                      uint32_t k, const uint32_t *coarse_flags, uint64_t *out_idx,
                      float *out_dist, uint32_t *out_count, uint32_t *flags, hipStream_t s)
{
    if (rows == 0) return;
    uint32_t P = 1;
    while (P < per_query) P <<= 1;
    if (P > 16384) fail(HM_ERR_INVALID_ARG, "nprobe*(k+1) above 16384");
    ProfScope prof("ivf_merge", s);
    ivf_merge_kernel<<<rows, 256, (size_t)P * 8, s>>>(cand, per_query, k, coarse_flags, out_idx,
                                                      out_dist, out_count, flags);
    HM_CHECK_LAUNCH();
}

void launch_ivf_replay(const uint32_t *rows, uint32_t nrows, const float *Dc, uint32_t ldd,
                       uint32_t nlist, uint32_t nprobe, const float *Qp, uint32_t dpad,
                       const float *C, const uint32_t *off, const uint32_t *ids,
                       const uint8_t *codes, const float *codebooks, const uint32_t *sub_b,
                       const uint32_t *sub_e, uint32_t M, uint32_t KS, uint32_t sdm, uint32_t k,
                       uint32_t lmax, float *scratch_f, uint32_t *scratch_u, size_t per_thread,
                       const uint32_t *probes, uint32_t ldp, const uint32_t *nprobe_q,
                       uint64_t *out_idx, float *out_dist, uint32_t *out_count, hipStream_t s)
{
    if (nrows == 0) return;
    ProfScope prof("ivf_replay", s);
    ivf_replay_kernel<<<ceil_div(nrows, 64), 64, 0, s>>>(
        rows, nrows, Dc, ldd, nlist, nprobe, Qp, dpad, C, off, ids, codes, codebooks, sub_b,
        sub_e, M, KS, sdm, k, lmax, scratch_f, scratch_u, per_thread, probes, ldp, nprobe_q,
        out_idx, out_dist, out_count);
    HM_CHECK_LAUNCH();
}


bool launch_ivf_replay_lds(const uint32_t *rows, uint32_t nrows, const float *Dc, uint32_t ldd,
                           uint32_t nlist, uint32_t nprobe, const float *Qp, uint32_t dpad,
                           const float *C, const uint32_t *off, const uint32_t *ids,
                           const uint8_t *codes, const float *codebooks, const uint32_t *sub_b,
                           const uint32_t *sub_e, uint32_t M, uint32_t KS, uint32_t sdm,
                           uint32_t k, uint32_t lmax, const uint32_t *probes, uint32_t ldp,
                           const uint32_t *nprobe_q, uint32_t *order, uint32_t *np,
                           uint2 *seq, uint32_t *seq_cnt, uint64_t *out_idx, float *out_dist,
                           uint32_t *out_count, hipStream_t s, const uint32_t *coarse_flags)
{
    if (nrows == 0) return true;
    if (nrows > 65535) return false;
    const uint32_t Lcap = std::max<uint32_t>(lmax, 1);
    uint32_t Pcap = 1;
    while (Pcap < Lcap) Pcap <<= 1;
    hipFuncAttributes attr{};
    HM_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(ivf_list_replay_kernel)));
    const size_t lds = (size_t)((Pcap + 1) & ~1u) * 8 + ((size_t)Lcap + 2) * 8 + (size_t)M * KS * 4;
    const size_t lds_order = probes && !coarse_flags ? 0 : ((size_t)nlist + 2) * 8;
    if (lds + attr.sharedSizeBytes > 160 * 1024 || lds_order > 160 * 1024 ||
        (size_t)(k + 3) * 8 > 160 * 1024)
        return false;  // the caller uses the global-memory form
    ProfScope prof("ivf_replay", s);
    const uint32_t ldo = nprobe;
    ivf_order_kernel<<<nrows, 64, lds_order, s>>>(rows, Dc, ldd, nlist, nprobe, probes, ldp,
                                                  nprobe_q, ldo, order, np, coarse_flags);
    HM_CHECK_LAUNCH();
    ivf_list_replay_kernel<<<dim3(nprobe, nrows), RP_THREADS, lds, s>>>(
        rows, Qp, dpad, C, off, ids, codes, codebooks, sub_b, sub_e, M, KS, sdm, k, Lcap, Pcap,
        order, ldo, np, seq, seq_cnt);
    HM_CHECK_LAUNCH();
    ivf_merge_replay_kernel<<<nrows, 64, (size_t)(k + 3) * 8, s>>>(rows, k, ldo, np, seq, seq_cnt,
                                                                  out_idx, out_dist, out_count);
    HM_CHECK_LAUNCH();
    return true;
}

void launch_negate(const float *x, float *y, uint64_t n, hipStream_t s)
{
    negate_kernel<<<grid_for(n), 256, 0, s>>>(x, y, n);
    HM_CHECK_LAUNCH();
}

void launch_probe_count(const float *negp, uint32_t ld, uint32_t rows, uint32_t pmax, float mass,
                        uint32_t pmin, uint32_t *nprobe_q, hipStream_t s, const float *lse)
{
    if (rows == 0) return;
    probe_count_kernel<<<ceil_div(rows, 256), 256, 0, s>>>(negp, ld, rows, pmax, mass, pmin,
                                                            nprobe_q, lse);
    HM_CHECK_LAUNCH();
}

void launch_dense_targets(const uint32_t *lab, uint32_t k, uint32_t rows, uint32_t nlist,
                          float *T, hipStream_t s)
{
    if (rows == 0) return;
    dense_targets_kernel<<<rows, 256, 0, s>>>(lab, k, rows, nlist, T);
    HM_CHECK_LAUNCH();
}

void launch_ids_to_lists(const uint64_t *ids, uint64_t n, const uint32_t *assign, uint64_t n_items,
                         uint32_t *out, hipStream_t s)
{
    if (n == 0) return;
    ids_to_lists_kernel<<<grid_for(n), 256, 0, s>>>(ids, n, assign, n_items, out);
    HM_CHECK_LAUNCH();
}

void launch_coarse_split(float *C, uint32_t ldc, uint32_t nc, uint32_t dim, const uint32_t *off,
                         uint64_t n, uint64_t *rng, uint32_t *cnt_ws, double *pk_ws, hipStream_t s)
{
    if (n == 0) fail(HM_ERR_EMPTY, "None to assigned impossible split center");
    if (nc == 0) return;
    ProfScope prof("kmeans_split", s);
    coarse_split_kernel<<<1, CS_THREADS, 0, s>>>(C, ldc, nc, dim, off, n, rng, cnt_ws, pk_ws);
    HM_CHECK_LAUNCH();
}

uint64_t host_stream_seed(uint64_t seed, uint64_t instance) { return stream_seed(seed, instance); }

}  // namespace hm
