// pq_index.hip -- standalone PQIndex on the device (pq_idx.rs:21-265).
//
// HBM layout after build:
//   base       [cap][dpad] f32, rows in insertion order (item i = row i)
//   codebook   [M][KS][sdm] f32 (a sub-space with fewer dims is zero padded)
//   codes      [n][M] u8, item i's code in sub-space s at codes[i*M + s]
// Build (PQIndex::train_center, pq_idx.rs:122-149): per sub-space the
// reference's k-means with the build metric on the raw sub-vectors (no
// residual): column means, then per epoch an assignment kernel (grid over
// items x sub-spaces, the sub-space's codebook in LDS), a point-order update
// kernel (one workgroup per sub-space, item chunks staged in LDS, one lane
// per centre folding its members in item order) and the split step; a last
// assignment writes the codes.
// Search (search_knn_adc, pq_idx.rs:165-194): a workgroup builds the LUTs of
// a group of queries in LDS with the build metric and scans the items' codes
// (ADC: f32 sum over the sub-spaces in order), writing a distance block;
// the radix select keeps the k+1 smallest by (distance, item); rows whose
// k+1 smallest are pairwise distinct output them directly, the others
// replay the reference's push-all / pop-to-k heap over the row.
#include <algorithm>
#include <cstring>
#include <vector>

#include "index.h"
#include "ivfpq.h"
#include "pq_common.h"
#include "profiler.h"
#include "wave_heap.h"

#pragma clang fp contract(off)

namespace hm {

namespace {

enum { PM_EUCLID = 3, PM_MANHATTAN = 1, PM_DOT = 2, PM_COSINE = 4, PM_ANGULAR = 5 };

// metrics.rs:73-83 as this build defines Cosine: -s_zy / (sqrt(s_zz)*sqrt(s_yy)).
__device__ __forceinline__ float pm_cosine(float szy, float szz, float syy)
{
    float den = sqrtf(szz) * sqrtf(syy);
    return (-szy) / den;
}

// metrics.rs:88-103 with float_two() == 1.0 (node.rs:54-56).
__device__ __forceinline__ float pm_angular(float szy, float szz, float syy)
{
    float rhd = -szz, lhd = -syy, rldot = -szy;
    float rlmul = rhd * lhd;
    if (rlmul > 0.0f) {
        float t = 1.0f * rldot;
        t = t / sqrtf(rlmul);
        return 1.0f - t;
    }
    return 1.0f;
}

// One fold step of the metric's accumulated sum (simd_metrics.rs scalar
// branches): Euclid (z-y)^2, Manhattan |z-y|, the others z*y.
template <int MET>
__device__ __forceinline__ float pm_step(float acc, float z, float y)
{
    if constexpr (MET == PM_EUCLID) {
        float d = z - y;
        float sq = d * d;
        return acc + sq;
    } else if constexpr (MET == PM_MANHATTAN) {
        float d = z - y;
        return acc + fabsf(d);
    } else {
        float p = z * y;
        return acc + p;
    }
}

template <int MET>
__device__ __forceinline__ float pm_finish(float acc, float szz, float syy)
{
    if constexpr (MET == PM_COSINE) return pm_cosine(acc, szz, syy);
    else if constexpr (MET == PM_ANGULAR) return pm_angular(acc, szz, syy);
    else return acc;
}

constexpr bool pm_norms(int met) { return met == PM_COSINE || met == PM_ANGULAR; }

// kmeans.rs:50-86: mean of column t over the items (left fold in item
// order from +0.0, f32 or f64), then centre c of the sub-space holding t
// gets mean + 1 where init_offset(c, j) says so, else mean - 1 (j is the
// dimension's index inside its sub-space).  One lane per used dimension.
template <typename ACC_T>
__global__ void pq_init_kernel(const float *__restrict__ X, uint32_t ldx, uint64_t n,
                               const uint32_t *__restrict__ sub_b, const uint32_t *__restrict__ sub_e,
                               uint32_t M, uint32_t KS, uint32_t sdm, float *__restrict__ cb)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sub_e[M - 1]) return;
    uint32_t s = 0;
    while (t >= sub_e[s]) ++s;
    const uint32_t j = t - sub_b[s];
    ACC_T acc = 0;
    for (uint64_t i = 0; i < n; ++i) acc = acc + (ACC_T)X[i * ldx + t];
    float mean;
    if constexpr (sizeof(ACC_T) == 4) mean = acc / (float)n;
    else mean = (float)(acc / (double)n);
    for (uint32_t c = 0; c < KS; ++c)
        cb[((size_t)s * KS + c) * sdm + j] = init_offset(c, j) > 0 ? mean + 1.0f : mean - 1.0f;
}

// kmeans.rs:127-148 (search_data) under the build metric: item i of
// sub-space s goes to the lowest-index centre with the smallest distance,
// scanning with a strict '<' from centre 0 (a NaN distance never wins, and
// a NaN at the current best keeps it).  Writes out[i*si + s*ss].
constexpr int PA_THREADS = 256;
constexpr int PA_REG = 32;  // sub-vector dims held in registers

template <int MET, bool STAGED>
__global__ __launch_bounds__(PA_THREADS) void pq_assign_kernel(
    const float *__restrict__ X, uint32_t ldx, uint32_t n, const uint32_t *__restrict__ sub_b,
    const uint32_t *__restrict__ sub_e, uint32_t KS, uint32_t sdm, const float *__restrict__ cb,
    uint8_t *__restrict__ out, uint64_t si, uint64_t ss)
{
    extern __shared__ __align__(16) float pa_smem[];
    const uint32_t s = blockIdx.y;
    const uint32_t b = sub_b[s], sd = sub_e[s] - b;
    const float *cbs = cb + (size_t)s * KS * sdm;
    float *cn = pa_smem;           // [KS] codeword self-dots (Cosine / Angular)
    float *cl = pa_smem + KS;      // [KS][sd] staged codebook (STAGED)
    const float *C;
    uint32_t ldc;
    if constexpr (STAGED) {
        for (uint32_t e = threadIdx.x; e < KS * sd; e += PA_THREADS)
            cl[e] = cbs[(size_t)(e / sd) * sdm + e % sd];
        C = cl;
        ldc = sd;
    } else {
        C = cbs;
        ldc = sdm;
    }
    __syncthreads();
    if constexpr (pm_norms(MET)) {
        for (uint32_t c = threadIdx.x; c < KS; c += PA_THREADS) {
            float a = -0.0f;
            for (uint32_t j = 0; j < sd; ++j) a = pm_step<PM_DOT>(a, C[c * ldc + j], C[c * ldc + j]);
            cn[c] = a;
        }
        __syncthreads();
    }
    const uint32_t i = blockIdx.x * PA_THREADS + threadIdx.x;
    if (i >= n) return;
    const float *x = X + (size_t)i * ldx + b;
    uint32_t best = 0;
    float bd = 0.0f;
    if (sd <= PA_REG) {
        float z[PA_REG];
#pragma unroll
        for (int j = 0; j < PA_REG; ++j) {
            // clamped address, masked value: every load stays inside the row
            const float v = x[min((uint32_t)j, sd - 1)];
            z[j] = (uint32_t)j < sd ? v : 0.0f;
        }
        float zz = -0.0f;
        if constexpr (pm_norms(MET)) {
#pragma unroll
            for (int j = 0; j < PA_REG; ++j)
                if ((uint32_t)j < sd) zz = pm_step<PM_DOT>(zz, z[j], z[j]);
        }
        for (uint32_t c = 0; c < KS; ++c) {
            const float *y = C + (size_t)c * ldc;
            float acc = -0.0f;
#pragma unroll
            for (int j = 0; j < PA_REG; ++j)
                if ((uint32_t)j < sd) acc = pm_step<MET>(acc, z[j], y[j]);
            float dc = pm_finish<MET>(acc, zz, pm_norms(MET) ? cn[c] : 0.0f);
            if (c == 0) bd = dc;
            else if (dc < bd) {
                bd = dc;
                best = c;
            }
        }
    } else {
        float zz = -0.0f;
        if constexpr (pm_norms(MET))
            for (uint32_t j = 0; j < sd; ++j) zz = pm_step<PM_DOT>(zz, x[j], x[j]);
        for (uint32_t c = 0; c < KS; ++c) {
            const float *y = C + (size_t)c * ldc;
            float acc = -0.0f;
            for (uint32_t j = 0; j < sd; ++j) acc = pm_step<MET>(acc, x[j], y[j]);
            float dc = pm_finish<MET>(acc, zz, pm_norms(MET) ? cn[c] : 0.0f);
            if (c == 0) bd = dc;
            else if (dc < bd) {
                bd = dc;
                best = c;
            }
        }
    }
    out[(size_t)i * si + (size_t)s * ss] = (uint8_t)best;
}

// kmeans.rs:88-125 (update_center) in four parallel steps over all
// sub-spaces.  The sums must run over each centre's members in item order,
// so the items are first grouped by centre, stably:
//   pq_chunk_hist: centre counts per chunk of PU_CH items;
//   pq_chunk_scan: per centre, the exclusive prefix of those counts over
//     the chunks, the centre's size, and the centres' base offsets;
//   pq_scatter: item i goes to base[c] + prefix[chunk][c] + (items before
//     it in its chunk with the same centre);
//   pq_fold: one lane per (sub-space, centre, dimension) adds its members'
//     values in item order from +0.0 (f32 or f64) and divides by the count;
//     an empty centre becomes zero.
constexpr int PU_CH = 256;

__global__ __launch_bounds__(PU_CH) void pq_chunk_hist_kernel(const uint8_t *__restrict__ asg,
                                                              uint32_t n, uint32_t KS,
                                                              uint32_t nch,
                                                              uint32_t *__restrict__ ccnt)
{
    __shared__ uint32_t hist[256];
    const uint32_t ch = blockIdx.x, s = blockIdx.y, tid = threadIdx.x;
    for (uint32_t c = tid; c < KS; c += PU_CH) hist[c] = 0;
    __syncthreads();
    const uint32_t i = ch * PU_CH + tid;
    if (i < n) atomicAdd(&hist[asg[(size_t)s * n + i]], 1u);
    __syncthreads();
    for (uint32_t c = tid; c < KS; c += PU_CH) ccnt[((size_t)s * nch + ch) * KS + c] = hist[c];
}

__global__ __launch_bounds__(256) void pq_chunk_scan_kernel(const uint32_t *__restrict__ ccnt,
                                                            uint32_t nch, uint32_t KS,
                                                            uint32_t *__restrict__ coff,
                                                            uint32_t *__restrict__ cbase,
                                                            uint32_t *__restrict__ counts)
{
    __shared__ uint32_t tot[256];
    const uint32_t s = blockIdx.x, c = threadIdx.x;
    uint32_t run = 0;
    if (c < KS) {
        const uint32_t *src = ccnt + (size_t)s * nch * KS + c;
        uint32_t *dst = coff + (size_t)s * nch * KS + c;
        uint32_t ch = 0;
        for (; ch + 8 <= nch; ch += 8) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = src[(size_t)(ch + u) * KS];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                dst[(size_t)(ch + u) * KS] = run;
                run += v[u];
            }
        }
        for (; ch < nch; ++ch) {
            const uint32_t v = src[(size_t)ch * KS];
            dst[(size_t)ch * KS] = run;
            run += v;
        }
        counts[(size_t)s * KS + c] = run;
    }
    tot[c] = c < KS ? run : 0u;
    __syncthreads();
    if (c == 0) {
        uint32_t acc = 0;
        for (uint32_t j = 0; j < KS; ++j) {
            cbase[(size_t)s * (KS + 1) + j] = acc;
            acc += tot[j];
        }
        cbase[(size_t)s * (KS + 1) + KS] = acc;
    }
}

__global__ __launch_bounds__(PU_CH) void pq_scatter_kernel(const uint8_t *__restrict__ asg,
                                                           uint32_t n, uint32_t KS, uint32_t nch,
                                                           const uint32_t *__restrict__ coff,
                                                           const uint32_t *__restrict__ cbase,
                                                           uint32_t *__restrict__ members)
{
    __shared__ uint32_t a[PU_CH];
    const uint32_t ch = blockIdx.x, s = blockIdx.y, tid = threadIdx.x;
    const uint32_t i = ch * PU_CH + tid;
    a[tid] = i < n ? asg[(size_t)s * n + i] : 0xFFFFFFFFu;
    __syncthreads();
    if (i >= n) return;
    const uint32_t c = a[tid];
    uint32_t rank = 0;
    for (uint32_t t = 0; t < tid; ++t) rank += a[t] == c ? 1u : 0u;
    const uint32_t pos = cbase[(size_t)s * (KS + 1) + c] + coff[((size_t)s * nch + ch) * KS + c] + rank;
    members[(size_t)s * n + pos] = i;
}

template <typename ACC_T>
__global__ __launch_bounds__(256) void pq_fold_kernel(
    const float *__restrict__ X, uint32_t ldx, uint32_t n, const uint32_t *__restrict__ sub_b,
    const uint32_t *__restrict__ sub_e, uint32_t M, uint32_t KS, uint32_t sdm,
    const uint32_t *__restrict__ cbase, const uint32_t *__restrict__ members,
    float *__restrict__ cb)
{
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;  // (s, c, j) with j < sdm
    if (e >= M * KS * sdm) return;
    const uint32_t j = e % sdm, sc = e / sdm, c = sc % KS, s = sc / KS;
    const uint32_t b = sub_b[s], sd = sub_e[s] - b;
    if (j >= sd) return;
    const uint32_t lo = cbase[(size_t)s * (KS + 1) + c], hi = cbase[(size_t)s * (KS + 1) + c + 1];
    const uint32_t *mem = members + (size_t)s * n;
    const float *col = X + b + j;
    ACC_T acc = 0;
    uint32_t m = lo;
    for (; m + 8 <= hi; m += 8) {
        uint32_t id[8];
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) id[u] = mem[m + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = col[(size_t)id[u] * ldx];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + (ACC_T)v[u];
    }
    for (; m < hi; ++m) acc = acc + (ACC_T)col[(size_t)mem[m] * ldx];
    const uint32_t cnt = hi - lo;
    float r;
    if (cnt == 0) r = 0.0f;
    else if constexpr (sizeof(ACC_T) == 4) r = acc / (float)cnt;
    else r = (float)(acc / (double)cnt);
    cb[((size_t)s * KS + c) * sdm + j] = r;
}

// kmeans.rs:150-193 (split_center) with the seeded generator: one lane per
// sub-space walks the empty centres in order.
__global__ void pq_split_kernel(float *__restrict__ cb, uint32_t M, uint32_t KS, uint32_t sdm,
                                const uint32_t *__restrict__ sub_b,
                                const uint32_t *__restrict__ sub_e, uint32_t *__restrict__ counts,
                                uint64_t n, uint64_t *__restrict__ rng)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const uint32_t sd = sub_e[s] - sub_b[s];
    float *cent = cb + (size_t)s * KS * sdm;
    uint32_t *cnt = counts + (size_t)s * KS;
    uint64_t st = rng[s];
    const float EPS = 1.0f / 1024.0f;
    uint64_t words[4] = {0, 0, 0, 0};  // KS <= 256
    NonzeroSet nz{words, KS};
    for (uint32_t c = 0; c < KS; ++c)
        if (cnt[c] != 0) nz.put(c, true);
    for (uint32_t i = 0; i < KS; ++i) {
        if (cnt[i] != 0) continue;
        const uint32_t sc = split_pick(i, KS, n, [&](uint32_t c) { return cnt[c]; }, nz, &st);
        for (uint32_t j = 0; j < sd; ++j) {
            float src = cent[(size_t)sc * sdm + j];
            if (j % 2 == 0) {
                cent[(size_t)i * sdm + j] = src * (1.0f - EPS);
                cent[(size_t)sc * sdm + j] = src * (1.0f + EPS);
            } else {
                cent[(size_t)i * sdm + j] = src * (1.0f + EPS);
                cent[(size_t)sc * sdm + j] = src * (1.0f - EPS);
            }
        }
        cnt[i] = cnt[sc] / 2;
        cnt[sc] -= cnt[i];
        nz.put(i, cnt[i] != 0);
        nz.put(sc, cnt[sc] != 0);
    }
    rng[s] = st;
}

// out[s*KS + c] = sum_j cb[s][c][j]^2 over the sub-space's own dims
// (left fold from -0.0).
__global__ void pq_norms_kernel(const float *__restrict__ cb, uint32_t M, uint32_t KS,
                                uint32_t sdm, const uint32_t *__restrict__ sub_b,
                                const uint32_t *__restrict__ sub_e, float *__restrict__ out)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M * KS) return;
    const uint32_t s = e / KS, sd = sub_e[s] - sub_b[s];
    const float *y = cb + (size_t)e * sdm;
    float a = -0.0f;
    for (uint32_t j = 0; j < sd; ++j) a = pm_step<PM_DOT>(a, y[j], y[j]);
    out[e] = a;
}

// LUT entry (s, c) of query z (pq_idx.rs:170-179 with get_distance_from_vec_range
// 151-163: the build metric on the raw sub-vector, no residual).
template <int MET>
__device__ __forceinline__ float pq_lut_entry(const float *__restrict__ z,
                                              const float *__restrict__ y, uint32_t sd,
                                              float syy)
{
    float acc = -0.0f, zz = -0.0f;
    for (uint32_t j = 0; j < sd; ++j) {
        acc = pm_step<MET>(acc, z[j], y[j]);
        if constexpr (pm_norms(MET)) zz = pm_step<PM_DOT>(zz, z[j], z[j]);
    }
    return pm_finish<MET>(acc, zz, syy);
}

// Search: the LUTs of 4 queries in LDS as float4 lut4[s*KS + c] (one lane
// of the float4 per query), then ADC over this workgroup's item range
// (pq_idx.rs:181-188: dist = 0; dist += LUT[s][code_s] in sub-space order),
// one item per lane: the item's codes are read once for the 4 queries and
// each lookup is one 16-byte LDS read.  Rows of the distance block are
// written by consecutive lanes.
constexpr int PS_THREADS = 256;
constexpr int PS_Q = 4;
// Filter form of the search: sample size for the threshold, largest k + 1
// (the threshold is a k-th smallest of at most 32), fewest items.
constexpr uint32_t PQ_SAMPLE = 4096;
constexpr uint32_t PQ_FILTER_MAX_K = 32;
constexpr uint32_t PQ_FILTER_MIN_N = 16384;

// Order-preserving float -> uint32 map (-0.0 as +0.0) and its inverse.
__device__ __forceinline__ uint32_t pq_fkey(float f)
{
    f = f + 0.0f;
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float pq_unkey(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Filter form of the scan's output: instead of the distance block, each
// query keeps the items whose distance is <= tau[q] as (fkey(dist) << 32 |
// item) in list[q][0..cap) (cnt[q] counts them, past cap too; zeroed by the
// caller), and nanq[q] = 1 when one of its distances is NaN.
struct PqFilter {
    const float *tau;
    unsigned long long *list;
    uint32_t cap;
    uint32_t *cnt;
    uint32_t *nanq;
};

template <int MET, bool FILT>
__global__ __launch_bounds__(PS_THREADS) void pq_scan_kernel(
    const float *__restrict__ Qp, uint32_t dpad, uint32_t rows,
    const uint32_t *__restrict__ sub_b, const uint32_t *__restrict__ sub_e, uint32_t M,
    uint32_t KS, uint32_t sdm, const float *__restrict__ cb, const float *__restrict__ cbn,
    const uint8_t *__restrict__ codes, uint32_t n, uint32_t per_split, float *__restrict__ D,
    uint64_t ldd, PqFilter flt)
{
    extern __shared__ __align__(16) float4 lut4[];
    float *lut = reinterpret_cast<float *>(lut4);
    const uint32_t q0 = blockIdx.y * PS_Q;
    const uint32_t i_begin = blockIdx.x * per_split;
    const uint32_t i_end = min(n, i_begin + per_split);
    const uint32_t tid = threadIdx.x;
    for (uint32_t e = tid; e < M * KS * PS_Q; e += PS_THREADS) {
        const uint32_t qq = e % PS_Q, sc = e / PS_Q, s = sc / KS, b = sub_b[s], sd = sub_e[s] - b;
        const uint32_t q = q0 + qq;
        float v = 0.0f;
        if (q < rows)
            v = pq_lut_entry<MET>(Qp + (size_t)q * dpad + b, cb + (size_t)sc * sdm, sd,
                                  pm_norms(MET) ? cbn[sc] : 0.0f);
        lut[e] = v;
    }
    __syncthreads();
    const uint32_t nq = min((uint32_t)PS_Q, rows - q0);
    float tq[PS_Q];
    bool nanf[PS_Q];
#pragma unroll
    for (int qq = 0; qq < PS_Q; ++qq) {
        tq[qq] = (FILT && (uint32_t)qq < nq) ? flt.tau[q0 + qq] : 0.0f;
        nanf[qq] = false;
    }
    const uint32_t lane = tid & 63;
    const uint64_t below = (1ull << lane) - 1;
    auto store = [&](uint32_t i, float a0, float a1, float a2, float a3) {
        if constexpr (!FILT) {
            float *d0 = D + (size_t)q0 * ldd;
            d0[i] = a0;
            if (nq > 1) d0[ldd + i] = a1;
            if (nq > 2) d0[2 * ldd + i] = a2;
            if (nq > 3) d0[3 * ldd + i] = a3;
        } else {
            // one counter update per wave and query: the wave's hits take
            // consecutive slots in lane order
            const float av[PS_Q] = {a0, a1, a2, a3};
#pragma unroll
            for (int qq = 0; qq < PS_Q; ++qq) {
                const bool live = (uint32_t)qq < nq;
                if (live && av[qq] != av[qq]) nanf[qq] = true;
                const bool hit = live && av[qq] <= tq[qq];
                const uint64_t m = __ballot(hit);
                if (m) {
                    const uint32_t lead = (uint32_t)__builtin_ctzll(m);
                    uint32_t base = 0;
                    if (lane == lead) base = atomicAdd(&flt.cnt[q0 + qq], (uint32_t)__popcll(m));
                    base = (uint32_t)__builtin_amdgcn_readlane((int)base, (int)lead);
                    const uint32_t pos = base + (uint32_t)__popcll(m & below);
                    if (hit && pos < flt.cap)
                        flt.list[(size_t)(q0 + qq) * flt.cap + pos] =
                            ((unsigned long long)pq_fkey(av[qq]) << 32) | i;
                }
            }
        }
    };
    if ((M & 15u) == 0) {
        // 16 codes per 16-byte load, two items in flight per lane
        const uint32_t W = M / 16;
        for (uint32_t i = i_begin + tid; i < i_end; i += 2 * PS_THREADS) {
            const uint32_t i2 = i + PS_THREADS;
            const bool two = i2 < i_end;
            const uint4 *c1 = reinterpret_cast<const uint4 *>(codes + (size_t)i * M);
            const uint4 *c2 = reinterpret_cast<const uint4 *>(codes + (size_t)(two ? i2 : i) * M);
            float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
            float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
            for (uint32_t w = 0; w < W; ++w) {
                const uint4 v1 = c1[w], v2 = c2[w];
                const uint32_t x1[4] = {v1.x, v1.y, v1.z, v1.w};
                const uint32_t x2[4] = {v2.x, v2.y, v2.z, v2.w};
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const uint32_t s = 16 * w + u;
                    const float4 l1 = lut4[s * KS + ((x1[u >> 2] >> (8 * (u & 3))) & 0xFFu)];
                    const float4 l2 = lut4[s * KS + ((x2[u >> 2] >> (8 * (u & 3))) & 0xFFu)];
                    a0 = a0 + l1.x;
                    a1 = a1 + l1.y;
                    a2 = a2 + l1.z;
                    a3 = a3 + l1.w;
                    b0 = b0 + l2.x;
                    b1 = b1 + l2.y;
                    b2 = b2 + l2.z;
                    b3 = b3 + l2.w;
                }
            }
            store(i, a0, a1, a2, a3);
            if (two) store(i2, b0, b1, b2, b3);
        }
    } else {
        for (uint32_t i = i_begin + tid; i < i_end; i += PS_THREADS) {
            const uint8_t *cd = codes + (size_t)i * M;
            float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
            for (uint32_t s = 0; s < M; ++s) {
                const float4 l = lut4[s * KS + cd[s]];
                a0 = a0 + l.x;
                a1 = a1 + l.y;
                a2 = a2 + l.z;
                a3 = a3 + l.w;
            }
            store(i, a0, a1, a2, a3);
        }
    }
    if constexpr (FILT) {
#pragma unroll
        for (int qq = 0; qq < PS_Q; ++qq)
            if (nanf[qq]) flt.nanq[q0 + qq] = 1u;
    }
}

// Strided sample of the codes for the filter threshold: sample j is item
// floor(j * n / S).
__global__ void pq_sample_codes_kernel(const uint8_t *__restrict__ codes, uint32_t M, uint32_t n,
                                       uint32_t S, uint8_t *__restrict__ out)
{
    const uint64_t total = (uint64_t)S * M;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = e / M, c = e % M;
        out[e] = codes[(j * n / S) * M + c];
    }
}

// Filter lists -> the row select's candidate format: the first K entries
// by (distance, item) as cand_d / cand_i, ROW_TIE when they are not
// pairwise distinct, ROW_NAN from the scan, ROW_UNCERT when the list
// overflowed (or holds fewer than K, which only NaN thresholds allow).
__global__ __launch_bounds__(256) void pq_list_select_kernel(
    const unsigned long long *__restrict__ list, uint32_t cap, const uint32_t *__restrict__ cnt,
    const uint32_t *__restrict__ nanq, uint32_t K, float *__restrict__ cand_d,
    uint32_t *__restrict__ cand_i, uint32_t *__restrict__ flags)
{
    extern __shared__ unsigned long long lk[];  // [P], P a power of two >= min(cnt, cap)
    __shared__ uint32_t s_tie;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint32_t c = cnt[q], m = min(c, cap);
    const bool bad = c > cap || c < K;
    uint32_t P = 1;
    while (P < m) P <<= 1;
    if (tid == 0) s_tie = 0;
    for (uint32_t j = tid; j < P; j += 256) lk[j] = j < m ? list[(size_t)q * cap + j] : ~0ull;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < P / 2; t += 256) {
                const uint32_t lo = bitonic_lo(t, stride), hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const unsigned long long a = lk[lo], b = lk[hi];
                if ((a > b) == asc) {
                    lk[lo] = b;
                    lk[hi] = a;
                }
            }
            __syncthreads();
        }
    for (uint32_t j = tid; j < K; j += 256) {
        const unsigned long long v = j < m ? lk[j] : ~0ull;
        cand_i[(size_t)q * K + j] = (uint32_t)v;
        cand_d[(size_t)q * K + j] = pq_unkey((uint32_t)(v >> 32));
        if (j + 1 < K && j + 1 < m && (uint32_t)(v >> 32) == (uint32_t)(lk[j + 1] >> 32)) s_tie = 1;
    }
    __syncthreads();
    if (tid == 0)
        flags[q] = (s_tie ? ROW_TIE : 0u) | (nanq[q] ? ROW_NAN : 0u) | (bad ? ROW_UNCERT : 0u);
}

// The LUTs of codebooks too large for LDS, in global memory:
// lut_g[(q*M + s)*KS + c], one query per workgroup.
template <int MET>
__global__ __launch_bounds__(PS_THREADS) void pq_lut_global_kernel(
    const float *__restrict__ Qp, uint32_t dpad, const uint32_t *__restrict__ sub_b,
    const uint32_t *__restrict__ sub_e, uint32_t M, uint32_t KS, uint32_t sdm,
    const float *__restrict__ cb, const float *__restrict__ cbn, float *__restrict__ lut_g)
{
    const uint32_t q = blockIdx.x;
    float *lq = lut_g + (size_t)q * M * KS;
    for (uint32_t e = threadIdx.x; e < M * KS; e += PS_THREADS) {
        const uint32_t s = e / KS, b = sub_b[s], sd = sub_e[s] - b;
        lq[e] = pq_lut_entry<MET>(Qp + (size_t)q * dpad + b, cb + (size_t)e * sdm, sd,
                                  pm_norms(MET) ? cbn[e] : 0.0f);
    }
}

// ADC over global LUTs (one query per workgroup row).
__global__ __launch_bounds__(PS_THREADS) void pq_scan_glut_kernel(
    uint32_t rows, uint32_t M, uint32_t KS, const uint8_t *__restrict__ codes, uint32_t n,
    uint32_t per_split, const float *__restrict__ lut_g, float *__restrict__ D, uint64_t ldd)
{
    const uint32_t q = blockIdx.y;
    if (q >= rows) return;
    const uint32_t i_begin = blockIdx.x * per_split;
    const uint32_t i_end = min(n, i_begin + per_split);
    const float *lq = lut_g + (size_t)q * M * KS;
    for (uint32_t i = i_begin + threadIdx.x; i < i_end; i += PS_THREADS) {
        const uint8_t *cd = codes + (size_t)i * M;
        float dist = 0.0f;
        for (uint32_t s = 0; s < M; ++s) dist = dist + lq[s * KS + cd[s]];
        D[(size_t)q * ldd + i] = dist;
    }
}

// Rows whose k+1 smallest are pairwise distinct: the survivors of the
// reference's pop-to-k are the k smallest, drained in ascending order
// (pq_idx.rs:189-191, 212-233).  Copies them; tied rows are overwritten by
// the replay.  A NaN among n >= 2 items sets err bit 0 (neighbor.rs:32).
__global__ void pq_finalize_kernel(const float *__restrict__ cand_d,
                                   const uint32_t *__restrict__ cand_i, uint32_t ldc,
                                   const uint32_t *__restrict__ flags, uint32_t rows, uint32_t n,
                                   uint32_t k, uint64_t *__restrict__ out_idx,
                                   float *__restrict__ out_dist, uint32_t *__restrict__ out_count,
                                   uint32_t *__restrict__ err)
{
    const uint32_t r = blockIdx.x;
    if (r >= rows) return;
    const uint32_t m = min(k, n);
    const float inf = __int_as_float(0x7f800000);
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
        if (j < m) {
            out_idx[(size_t)r * k + j] = cand_i[(size_t)r * ldc + j];
            out_dist[(size_t)r * k + j] = cand_d[(size_t)r * ldc + j];
        } else {
            out_idx[(size_t)r * k + j] = ~0ull;
            out_dist[(size_t)r * k + j] = inf;
        }
    }
    if (threadIdx.x == 0) {
        out_count[r] = m;
        if ((flags[r] & ROW_NAN) && n >= 2) atomicOr(err, 1u);
    }
}

// Replay of pq_idx.rs:181-191 + 212-233 for tied rows: push every item of
// the distance row in item order, pop until k remain, then drain (the pops
// come out descending and are written from the back).  One wave per row
// runs the std heap (wave_heap.h); rows are taken in a grid-stride loop so
// each workgroup reuses its slice of the global heap buffer.  Heap positions
// below H live in LDS, the rest (HYB) in HBM.
constexpr int PR_THREADS = 64;

template <bool HYB> struct WhHybrid {
    uint2 *lds;
    uint2 *g;
    uint32_t H;
    __device__ uint2 ld(uint32_t p) const
    {
        if constexpr (HYB) {
            if (p >= H) return g[p - H];
        }
        return lds[p];
    }
    __device__ void st(uint32_t p, uint2 v) const
    {
        if constexpr (HYB) {
            if (p >= H) {
                g[p - H] = v;
                return;
            }
        }
        lds[p] = v;
    }
    __device__ uint4 pair(uint32_t first, uint32_t end) const
    {
        const uint2 a = ld(min(first, end)), b = ld(min(first + 1, end));
        return make_uint4(a.x, a.y, b.x, b.y);
    }
};

template <bool HYB>
__global__ __launch_bounds__(PR_THREADS) void pq_replay_kernel(
    const uint32_t *__restrict__ rows, uint32_t nrows, const float *__restrict__ D, uint64_t ldd,
    uint32_t n, uint32_t k, uint32_t H, uint2 *__restrict__ gbuf, uint64_t *__restrict__ out_idx,
    float *__restrict__ out_dist)
{
    extern __shared__ __align__(16) uint2 pr_heap[];
    const uint32_t lane = threadIdx.x;
    const size_t gslice = n > H ? (size_t)(n - H) : 0;
    uint2 *hg = HYB ? gbuf + (size_t)blockIdx.x * gslice : nullptr;
    for (uint32_t t = blockIdx.x; t < nrows; t += gridDim.x) {
        const uint32_t r = rows[t];
        const float *drow = D + (size_t)r * ldd;
        WaveHeap<WhHybrid<HYB>> h({pr_heap, hg, H}, lane);
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const float v = i0 + lane < n ? drow[i0 + lane] : 0.0f;
            const uint32_t cnt = min(64u, n - i0);
            for (uint32_t u = 0; u < cnt; ++u) {
                const uint32_t bits = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(v), (int)u);
                h.push(make_uint2(i0 + u, bits));
            }
        }
        while (h.len > k) (void)h.pop();
        const uint32_t m = h.len;
        for (uint32_t p = m; p > 0; --p) {
            const uint2 x = h.pop();
            if (lane == 0) {
                out_idx[(size_t)r * k + p - 1] = x.x;
                out_dist[(size_t)r * k + p - 1] = __uint_as_float(x.y);
            }
        }
        __syncthreads();
    }
}

}  // namespace

// ------------------------------------------------------------ host side

namespace {

size_t pq_lds_limit()
{
    static size_t lim = 0;
    if (!lim) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
                hipSuccess &&
            v > 0)
            lim = (size_t)v;
        else
            lim = 65536;
    }
    return lim;
}

class PQ final : public Index {
  public:
    PQ(uint64_t dim, const hm_params &p) : Index(HM_INDEX_PQ, dim), p_(p)
    {
        if (dim == 0 || dim > 4096) fail(HM_ERR_INVALID_ARG, "dimension out of range (1..4096)");
        if (p.n_sub == 0 || p.n_sub > 256 || dim / p.n_sub == 0)
            fail(HM_ERR_INVALID_ARG, "n_sub must be in 1..min(256, dimension)");
        // pq_idx.rs:54: assert!(sub_bits <= 32)
        if (p.sub_bits == 0 || p.sub_bits > 32) fail(HM_ERR_INVALID_ARG, "sub_bits out of range");
        if (p.sub_bits > 8)
            fail(HM_ERR_NOT_IMPLEMENTED, "sub_bits above 8 (codes are stored as u8)");
        dpad_ = (uint32_t)round_up(dim, 16);
        M_ = p.n_sub;
        KS_ = 1u << p.sub_bits;
        // pq_idx.rs:75-88
        const uint32_t sd = (uint32_t)(dim / M_), r = (uint32_t)(dim % sd);
        std::vector<uint32_t> b(M_), e(M_);
        for (uint32_t i = 0; i < M_; ++i) {
            if (i < r) {
                b[i] = i * (sd + 1);
                e[i] = (i + 1) * (sd + 1);
            } else {
                b[i] = r * (sd + 1) + (i - r) * sd;
                e[i] = r * (sd + 1) + (i + 1 - r) * sd;
            }
        }
        sdm_ = sd + (r ? 1 : 0);
        HM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        sub_b_.reserve(M_ * 4);
        sub_e_.reserve(M_ * 4);
        HM_HIP(hipMemcpyAsync(sub_b_.get(), b.data(), M_ * 4, hipMemcpyHostToDevice, stream_));
        HM_HIP(hipMemcpyAsync(sub_e_.get(), e.data(), M_ * 4, hipMemcpyHostToDevice, stream_));
        HM_HIP(hipStreamSynchronize(stream_));
    }
    ~PQ() override
    {
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    // pq_idx.rs:101-115: add_item (the 100,000-item limit is not reproduced)
    void add(const float *rows, uint64_t n_rows) override
    {
        if (n_rows == 0) return;
        if (!rows) fail(HM_ERR_INVALID_ARG, "rows is null");
        check_finite_host(rows, n_rows * dim_);
        grow(n_rows);
        DeviceBuffer stage;
        upload_rows(rows, n_rows, (uint32_t)dim_, base_.as<float>() + n_ * dpad_, dpad_, stage,
                    stream_);
        HM_HIP(hipStreamSynchronize(stream_));
        n_ += n_rows;
    }

    void add_device(const float *d_rows, uint64_t n_rows, uint64_t ld) override
    {
        if (n_rows == 0) return;
        grow(n_rows);
        launch_pad_rows(d_rows, ld, base_.as<float>() + n_ * dpad_, dpad_, n_rows, (uint32_t)dim_,
                        dpad_, stream_);
        HM_HIP(hipStreamSynchronize(stream_));
        n_ += n_rows;
    }

    // pq_idx.rs:198-202: build(mt) sets the metric and trains
    void build(int32_t metric) override
    {
        if (metric < HM_METRIC_UNKNOWN || metric > HM_METRIC_ANGULAR)
            fail(HM_ERR_INVALID_ARG, "metric out of range");
        if (metric == HM_METRIC_UNKNOWN) fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        metric_ = metric;
        train();
    }

    bool built() const override { return true; }  // pq_idx.rs:209-211
    int32_t metric() const override { return metric_; }
    const hm_params &params() const { return p_; }

    void copy_rows(uint64_t first, uint64_t count, float *out) const override
    {
        if (first + count > n_) fail(HM_ERR_INVALID_ARG, "row range out of bounds");
        if (count == 0) return;
        DeviceBuffer stage;
        download_rows(base_.as<float>() + first * dpad_, dpad_, count, (uint32_t)dim_, out, stage,
                      stream_);
        HM_HIP(hipStreamSynchronize(stream_));
    }

    void import_trained(int32_t metric, const float *codebook, const uint8_t *codes,
                        uint64_t n)
    {
        if (metric <= HM_METRIC_UNKNOWN || metric > HM_METRIC_ANGULAR)
            fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        if (n > n_) fail(HM_ERR_INVALID_ARG, "more coded items than rows");
        for (uint64_t i = 0; i < n * M_; ++i)
            if (codes[i] >= KS_) fail(HM_ERR_INVALID_ARG, "code out of range");
        hipStream_t s = stream_;
        cb_.reserve((size_t)M_ * KS_ * sdm_ * 4);
        cbn_.reserve((size_t)M_ * KS_ * 4);
        codes_.reserve(n * M_ + 4);
        HM_HIP(hipMemcpyAsync(cb_.get(), codebook, (size_t)M_ * KS_ * sdm_ * 4,
                              hipMemcpyHostToDevice, s));
        if (n) HM_HIP(hipMemcpyAsync(codes_.get(), codes, n * M_, hipMemcpyHostToDevice, s));
        metric_ = metric;
        if (metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR)
            codeword_norms(s);
        HM_HIP(hipStreamSynchronize(s));
        trained_ = true;
        sample_for_ = 0;
        n_trained_ = n;
    }
    const char *name() const override { return "PQIndex"; }
    uint64_t nodes_size() const override { return n_; }

    void clear() override
    {
        base_.release();
        n_ = cap_ = 0;
        trained_ = false;
        n_trained_ = 0;
    }

    void get_rows(const uint64_t *idx, uint64_t n, float *out) const override
    {
        for (uint64_t i = 0; i < n; ++i) {
            if (idx[i] >= n_) fail(HM_ERR_INVALID_ARG, "row index out of range");
            HM_HIP(hipMemcpyAsync(out + i * dim_, base_.as<float>() + idx[i] * dpad_, dim_ * 4,
                                  hipMemcpyDeviceToHost, stream_));
        }
        HM_HIP(hipStreamSynchronize(stream_));
    }

    void info(uint32_t *out) const
    {
        out[0] = M_;
        out[1] = KS_;
        out[2] = sdm_;
        out[3] = (uint32_t)n_trained_;
        out[4] = trained_ ? 1 : 0;
        out[5] = dpad_;
        out[6] = (uint32_t)metric_;
        out[7] = 0;
    }

    void export_trained(float *codebook, uint8_t *codes) const
    {
        if (!trained_) fail(HM_ERR_NOT_BUILT, "index not built");
        if (codebook)
            HM_HIP(hipMemcpyAsync(codebook, cb_.get(), (size_t)M_ * KS_ * sdm_ * 4,
                                  hipMemcpyDeviceToHost, stream_));
        if (codes && n_trained_)
            HM_HIP(hipMemcpyAsync(codes, codes_.get(), n_trained_ * M_, hipMemcpyDeviceToHost,
                                  stream_));
        HM_HIP(hipStreamSynchronize(stream_));
    }

    void search_device(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx, float *d_dist,
                       uint32_t *d_count, hipStream_t s) override
    {
        if (!s) s = stream_;
        if (!trained_) fail(HM_ERR_NOT_BUILT, "index not built");
        if (nq == 0) return;
        if (!dq || !d_count || (k && (!d_idx || !d_dist))) fail(HM_ERR_INVALID_ARG, "null pointer");
        const uint32_t n = (uint32_t)n_trained_;
        if (k == 0 || n == 0) {
            launch_fill_results(d_idx, d_dist, d_count, nq, k, s);
            HM_HIP(hipStreamSynchronize(s));
            return;
        }
        const uint32_t K = std::min<uint32_t>(k + 1, n);
        if (K > SELECT_MAX_K) fail(HM_ERR_INVALID_ARG, "k too large (max 8191)");
        err_.reserve(4);
        HM_HIP(hipMemsetAsync(err_.get(), 0, 4, s));
        launch_check_finite(dq, dim_, nq, (uint32_t)dim_, err_.as<uint32_t>(), s);
        last_replayed_ = 0;
        last_filtered_ = 0;
        last_uncert_ = 0;
        if (lut_qb() != 0 && K <= PQ_FILTER_MAX_K && n >= PQ_FILTER_MIN_N)
            search_filter(dq, nq, k, d_idx, d_dist, d_count, s);
        else
            search_full(dq, nq, k, d_idx, d_dist, d_count, s);
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, err_.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
        if (err & 1u) fail(HM_ERR_NAN_DISTANCE, "NaN distance (partial_cmp of NaN)");
    }

    // Whole distance rows: scan, row select of the k+1 smallest, the
    // tie-free rows' results, and the heap replay of tied rows.
    void search_full(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx, float *d_dist,
                     uint32_t *d_count, hipStream_t s)
    {
        const uint32_t n = (uint32_t)n_trained_;
        const uint32_t K = std::min<uint32_t>(k + 1, n);
        const uint64_t ldd = round_up(n, 4);
        // distance rows of up to 4 GiB per chunk: the replay of a chunk's
        // tied rows fills the GPU best when the chunk is large
        uint64_t qc = std::min<uint64_t>(nq, 65535);
        qc = std::min<uint64_t>(qc, std::max<uint64_t>(1, ((size_t)1 << 32) / (ldd * 4)));
        const int qb = lut_qb();
        qpad_.reserve(qc * dpad_ * 4);
        D_.reserve(qc * ldd * 4);
        cand_d_.reserve(qc * K * 4);
        cand_i_.reserve(qc * K * 4);
        flags_.reserve(qc * 4);
        if (qb == 0) lutg_.reserve(qc * M_ * KS_ * 4);
        std::vector<uint32_t> fl(qc);
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            launch_pad_rows(dq + q0 * dim_, dim_, qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                            dpad_, s);
            {
                ProfScope ps("pq_scan", s);
                scan(rows, qb, codes_.as<uint8_t>(), n, D_.as<float>(), ldd, nullptr, s);
            }
            launch_row_select(D_.as<float>(), ldd, rows, n, K, cand_d_.as<float>(),
                              cand_i_.as<uint32_t>(), K, flags_.as<uint32_t>(), s);
            finalize(rows, n, k, K, d_idx + q0 * k, d_dist + q0 * k, d_count + q0, s);
            HM_HIP(hipMemcpyAsync(fl.data(), flags_.get(), rows * 4, hipMemcpyDeviceToHost, s));
            HM_HIP(hipStreamSynchronize(s));
            std::vector<uint32_t> tied;
            for (uint32_t i = 0; i < rows; ++i)
                if ((fl[i] & ROW_TIE) && !(fl[i] & ROW_NAN)) tied.push_back(i);
            last_replayed_ += tied.size();
            if (!tied.empty()) replay(tied, n, k, ldd, d_idx + q0 * k, d_dist + q0 * k, s);
        }
    }

    void finalize(uint32_t rows, uint32_t n, uint32_t k, uint32_t K, uint64_t *d_idx,
                  float *d_dist, uint32_t *d_count, hipStream_t s)
    {
        ProfScope ps("pq_finalize", s);
        pq_finalize_kernel<<<rows, 64, 0, s>>>(cand_d_.as<float>(), cand_i_.as<uint32_t>(), K,
                                               flags_.as<uint32_t>(), rows, n, k, d_idx, d_dist,
                                               d_count, err_.as<uint32_t>());
        HM_CHECK_LAUNCH();
    }

    // Filter form (k + 1 <= 32, the LUTs in LDS): tau[q] = the (k+1)-th
    // smallest ADC distance over a strided sample of PQ_SAMPLE items, an
    // upper bound of the row's (k+1)-th smallest; the scan keeps only the
    // items at or below it, so the first k+1 of the row by (distance, item)
    // are among them.  Rows whose list overflows, and tied rows (the heap
    // replay needs the whole row), are computed again by search_full.
    void search_filter(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx,
                       float *d_dist, uint32_t *d_count, hipStream_t s)
    {
        const uint32_t n = (uint32_t)n_trained_;
        const uint32_t K = std::min<uint32_t>(k + 1, n);
        const uint32_t S = PQ_SAMPLE;
        if (sample_for_ != n) {
            codes_s_.reserve((size_t)S * M_);
            pq_sample_codes_kernel<<<std::min<uint32_t>(ceil_div(S * M_, 256), 1024), 256, 0, s>>>(
                codes_.as<uint8_t>(), M_, n, S, codes_s_.as<uint8_t>());
            HM_CHECK_LAUNCH();
            sample_for_ = n;
        }
        // list capacity: 8x the expected K * n / S entries, a power of two
        uint32_t cap = 1024;
        while (cap < 8ull * K * n / S && cap < 8192) cap <<= 1;
        const uint64_t qc = std::min<uint64_t>(nq, 65535);
        qpad_.reserve(qc * dpad_ * 4);
        Ds_.reserve(qc * S * 4);
        tau_.reserve(qc * 4);
        lists_.reserve(qc * cap * 8);
        lcnt_.reserve(qc * 2 * 4);
        cand_d_.reserve(qc * K * 4);
        cand_i_.reserve(qc * K * 4);
        flags_.reserve(qc * 4);
        std::vector<uint32_t> fl(qc);
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            launch_pad_rows(dq + q0 * dim_, dim_, qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                            dpad_, s);
            uint32_t *cnt = lcnt_.as<uint32_t>(), *nanq = cnt + rows;
            {
                ProfScope ps("pq_scan", s);
                scan(rows, 1, codes_s_.as<uint8_t>(), S, Ds_.as<float>(), S, nullptr, s);
                launch_kth_smallest(Ds_.as<float>(), S, rows, S, K, tau_.as<float>(), s);
                launch_fill_u32(cnt, 2 * rows, 0u, s);
                const PqFilter f{tau_.as<float>(), lists_.as<unsigned long long>(), cap, cnt, nanq};
                scan(rows, 1, codes_.as<uint8_t>(), n, nullptr, 0, &f, s);
            }
            {
                ProfScope ps("pq_list_select", s);
                uint32_t P = 1;
                while (P < cap) P <<= 1;
                pq_list_select_kernel<<<rows, 256, (size_t)P * 8, s>>>(
                    lists_.as<unsigned long long>(), cap, cnt, nanq, K, cand_d_.as<float>(),
                    cand_i_.as<uint32_t>(), flags_.as<uint32_t>());
                HM_CHECK_LAUNCH();
            }
            finalize(rows, n, k, K, d_idx + q0 * k, d_dist + q0 * k, d_count + q0, s);
            HM_HIP(hipMemcpyAsync(fl.data(), flags_.get(), rows * 4, hipMemcpyDeviceToHost, s));
            HM_HIP(hipStreamSynchronize(s));
            std::vector<uint32_t> redo;
            for (uint32_t i = 0; i < rows; ++i)
                if ((fl[i] & ROW_UNCERT) || ((fl[i] & ROW_TIE) && !(fl[i] & ROW_NAN)))
                    redo.push_back((uint32_t)(q0 + i));
            last_filtered_ += rows - redo.size();
            for (uint32_t i = 0; i < rows; ++i) last_uncert_ += (fl[i] & ROW_UNCERT) ? 1u : 0u;
            if (redo.empty()) continue;
            // the rows again with whole distance rows, into compact outputs
            const uint32_t m = (uint32_t)redo.size();
            redo_rows_.reserve(m * 4);
            redo_q_.reserve((size_t)m * dim_ * 4);
            redo_i_.reserve((size_t)m * k * 8);
            redo_d_.reserve((size_t)m * k * 4);
            redo_c_.reserve((size_t)m * 4);
            HM_HIP(hipMemcpyAsync(redo_rows_.get(), redo.data(), m * 4, hipMemcpyHostToDevice, s));
            launch_gather_rows(dq, dim_, redo_rows_.as<uint32_t>(), m, (uint32_t)dim_,
                               redo_q_.as<float>(), dim_, s);
            search_full(redo_q_.as<float>(), m, k, redo_i_.as<uint64_t>(), redo_d_.as<float>(),
                        redo_c_.as<uint32_t>(), s);
            launch_scatter_results(redo_i_.as<uint64_t>(), redo_d_.as<float>(),
                                   redo_c_.as<uint32_t>(), redo_rows_.as<uint32_t>(), m, k, d_idx,
                                   d_dist, d_count, s);
            HM_HIP(hipStreamSynchronize(s));
        }
    }

    void search_stats(uint64_t *fallback, uint64_t *uncertified) const override
    {
        *fallback = last_replayed_;
        *uncertified = last_uncert_;
    }

    // Rows of the last search answered by the filter form.
    uint64_t last_filtered() const { return last_filtered_; }

  private:
    void grow(uint64_t n_rows)
    {
        if (n_ + n_rows > 0xFFFFFFF0ull) fail(HM_ERR_CAPACITY, "more than 2^32-16 rows");
        const uint64_t need = n_ + n_rows;
        if (need > cap_) {
            uint64_t nc = std::max<uint64_t>(need, cap_ + cap_ / 2);
            nc = std::max<uint64_t>(nc, 1024);
            base_.grow_keep(nc * dpad_ * 4, n_ * dpad_ * 4, stream_);
            cap_ = nc;
        }
    }

    // 1 when the LUTs of PS_Q queries fit in LDS (the scan kernel), else 0
    // (LUTs in global memory).
    int lut_qb() const
    {
        const size_t per = (size_t)M_ * KS_ * 4, lim = std::min<size_t>(pq_lds_limit(), 131072);
        return per * PS_Q <= lim ? 1 : 0;
    }

    // ADC of `rows` padded queries (qpad_) against n coded items: the
    // distance block D (stride ldd), or with flt the filter lists.
    template <int MET>
    void scan_m(uint32_t rows, int qb, const uint8_t *codes, uint32_t n, float *D, uint64_t ldd,
                const PqFilter *flt, hipStream_t s)
    {
        if (qb == 0) {
            const uint32_t nsplit = std::max<uint32_t>(1, std::min<uint32_t>(
                ceil_div(1024, rows), std::max<uint32_t>(1, n / 2048)));
            const uint32_t per_split = (uint32_t)round_up(ceil_div(n, nsplit), 64);
            pq_lut_global_kernel<MET><<<rows, PS_THREADS, 0, s>>>(
                qpad_.as<float>(), dpad_, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_, KS_,
                sdm_, cb_.as<float>(), cbn_.as<float>(), lutg_.as<float>());
            HM_CHECK_LAUNCH();
            pq_scan_glut_kernel<<<dim3(ceil_div(n, per_split), rows), PS_THREADS, 0, s>>>(
                rows, M_, KS_, codes, n, per_split, lutg_.as<float>(), D, ldd);
            HM_CHECK_LAUNCH();
            return;
        }
        // about 2048 workgroups; each range a multiple of 2 x 256 items
        const uint32_t groups = ceil_div(rows, PS_Q);
        const uint32_t nsplit = std::max<uint32_t>(1, std::min<uint32_t>(
            ceil_div(2048, groups), std::max<uint32_t>(1, n / 4096)));
        const uint32_t per_split = (uint32_t)round_up(ceil_div(n, nsplit), 2 * PS_THREADS);
        const uint32_t ns = ceil_div(n, per_split);
        const size_t lds = (size_t)M_ * KS_ * PS_Q * 4;
        if (flt)
            pq_scan_kernel<MET, true><<<dim3(ns, groups), PS_THREADS, lds, s>>>(
                qpad_.as<float>(), dpad_, rows, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_,
                KS_, sdm_, cb_.as<float>(), cbn_.as<float>(), codes, n, per_split, nullptr, 0,
                *flt);
        else
            pq_scan_kernel<MET, false><<<dim3(ns, groups), PS_THREADS, lds, s>>>(
                qpad_.as<float>(), dpad_, rows, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_,
                KS_, sdm_, cb_.as<float>(), cbn_.as<float>(), codes, n, per_split, D, ldd,
                PqFilter{});
        HM_CHECK_LAUNCH();
    }

    void scan(uint32_t rows, int qb, const uint8_t *codes, uint32_t n, float *D, uint64_t ldd,
              const PqFilter *flt, hipStream_t s)
    {
        switch (metric_) {
        case HM_METRIC_EUCLIDEAN: scan_m<PM_EUCLID>(rows, qb, codes, n, D, ldd, flt, s); break;
        case HM_METRIC_MANHATTAN: scan_m<PM_MANHATTAN>(rows, qb, codes, n, D, ldd, flt, s); break;
        case HM_METRIC_DOT_PRODUCT: scan_m<PM_DOT>(rows, qb, codes, n, D, ldd, flt, s); break;
        case HM_METRIC_COSINE_SIMILARITY:
            scan_m<PM_COSINE>(rows, qb, codes, n, D, ldd, flt, s);
            break;
        default: scan_m<PM_ANGULAR>(rows, qb, codes, n, D, ldd, flt, s); break;
        }
    }

    void replay(const std::vector<uint32_t> &tied, uint32_t n, uint32_t k, uint64_t ldd,
                uint64_t *d_idx, float *d_dist, hipStream_t s)
    {
        ProfScope ps("pq_replay", s);
        const uint32_t m = (uint32_t)tied.size();
        rp_rows_.reserve(m * 4);
        HM_HIP(hipMemcpyAsync(rp_rows_.get(), tied.data(), m * 4, hipMemcpyHostToDevice, s));
        // rows up to 8,192 items keep the whole heap in LDS (<= 64 KiB);
        // longer rows keep the top 11 levels (16 KiB, 8 rows per CU) and
        // the rest in HBM
        const bool hyb = n > 8192;
        const uint32_t H = hyb ? 2047u : n;
        const uint32_t grid = std::min<uint32_t>(m, hyb ? 2048u : 512u);
        const size_t gslice = hyb ? (size_t)(n - H) : 0;
        if (hyb) rp_g_.reserve((size_t)grid * gslice * 8);
        const size_t use = (size_t)H * 8;
        if (hyb)
            pq_replay_kernel<true><<<grid, PR_THREADS, use, s>>>(
                rp_rows_.as<uint32_t>(), m, D_.as<float>(), ldd, n, k, H, rp_g_.as<uint2>(), d_idx,
                d_dist);
        else
            pq_replay_kernel<false><<<grid, PR_THREADS, use, s>>>(
                rp_rows_.as<uint32_t>(), m, D_.as<float>(), ldd, n, k, H, nullptr, d_idx, d_dist);
        HM_CHECK_LAUNCH();
        HM_HIP(hipStreamSynchronize(s));
    }

    template <int MET> void assign_t(uint8_t *out, uint64_t si, uint64_t ss, hipStream_t s)
    {
        const uint32_t n = (uint32_t)n_;
        const size_t staged_bytes = ((size_t)KS_ + (size_t)KS_ * (sdm_)) * 4;
        dim3 grid(ceil_div(n, PA_THREADS), M_);
        if (staged_bytes <= 65536)
            pq_assign_kernel<MET, true><<<grid, PA_THREADS, staged_bytes, s>>>(
                base_.as<float>(), dpad_, n, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), KS_,
                sdm_, cb_.as<float>(), out, si, ss);
        else
            pq_assign_kernel<MET, false><<<grid, PA_THREADS, (size_t)KS_ * 4, s>>>(
                base_.as<float>(), dpad_, n, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), KS_,
                sdm_, cb_.as<float>(), out, si, ss);
        HM_CHECK_LAUNCH();
    }

    void assign(uint8_t *out, uint64_t si, uint64_t ss, hipStream_t s)
    {
        ProfScope ps("pq_assign", s);
        switch (metric_) {
        case HM_METRIC_EUCLIDEAN: assign_t<PM_EUCLID>(out, si, ss, s); break;
        case HM_METRIC_MANHATTAN: assign_t<PM_MANHATTAN>(out, si, ss, s); break;
        case HM_METRIC_DOT_PRODUCT: assign_t<PM_DOT>(out, si, ss, s); break;
        case HM_METRIC_COSINE_SIMILARITY: assign_t<PM_COSINE>(out, si, ss, s); break;
        default: assign_t<PM_ANGULAR>(out, si, ss, s); break;
        }
    }

    // PQIndex::train_center (pq_idx.rs:122-149) -> Kmeans::train
    // (kmeans.rs:195-207) per sub-space, then search_data for the codes.
    void train()
    {
        hipStream_t s = stream_;
        const uint64_t n = n_;
        // kmeans.rs:158-160 + 203-204: a batch of 0 is an Err that train unwraps
        if (n == 0) fail(HM_ERR_EMPTY, "None to assigned impossible split center");
        const bool fp64 = p_.kmeans_accumulate != 0;
        const uint32_t epochs = p_.train_epoch;
        cb_.reserve((size_t)M_ * KS_ * sdm_ * 4);
        cbn_.reserve((size_t)M_ * KS_ * 4);
        asg_.reserve((size_t)M_ * n);
        counts_.reserve((size_t)M_ * KS_ * 4);
        {
            const size_t nch = ceil_div(n, PU_CH);
            ccnt_.reserve((size_t)M_ * nch * KS_ * 4);
            coff_.reserve((size_t)M_ * nch * KS_ * 4);
            cbase_.reserve((size_t)M_ * (KS_ + 1) * 4);
            members_.reserve((size_t)M_ * n * 4);
        }
        codes_.reserve(n * M_ + 4);
        rng_.reserve((size_t)M_ * 8);
        std::vector<uint64_t> rs(M_);
        for (uint32_t i = 0; i < M_; ++i) rs[i] = host_stream_seed(p_.seed, 1 + i);
        HM_HIP(hipMemcpyAsync(rng_.get(), rs.data(), M_ * 8, hipMemcpyHostToDevice, s));
        HM_HIP(hipMemsetAsync(cb_.get(), 0, (size_t)M_ * KS_ * sdm_ * 4, s));
        {
            ProfScope ps("pq_init", s);
            const uint32_t used = (uint32_t)dim_;
            if (fp64)
                pq_init_kernel<double><<<ceil_div(used, 64), 64, 0, s>>>(
                    base_.as<float>(), dpad_, n, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_,
                    KS_, sdm_, cb_.as<float>());
            else
                pq_init_kernel<float><<<ceil_div(used, 64), 64, 0, s>>>(
                    base_.as<float>(), dpad_, n, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_,
                    KS_, sdm_, cb_.as<float>());
            HM_CHECK_LAUNCH();
        }
        for (uint32_t ep = 0; ep < epochs; ++ep) {
            assign(asg_.as<uint8_t>(), 1, n, s);
            {
                ProfScope ps("pq_update", s);
                const uint32_t nn = (uint32_t)n, nch = ceil_div(nn, PU_CH);
                pq_chunk_hist_kernel<<<dim3(nch, M_), PU_CH, 0, s>>>(asg_.as<uint8_t>(), nn, KS_,
                                                                     nch, ccnt_.as<uint32_t>());
                pq_chunk_scan_kernel<<<M_, 256, 0, s>>>(ccnt_.as<uint32_t>(), nch, KS_,
                                                        coff_.as<uint32_t>(), cbase_.as<uint32_t>(),
                                                        counts_.as<uint32_t>());
                pq_scatter_kernel<<<dim3(nch, M_), PU_CH, 0, s>>>(
                    asg_.as<uint8_t>(), nn, KS_, nch, coff_.as<uint32_t>(), cbase_.as<uint32_t>(),
                    members_.as<uint32_t>());
                const unsigned fb = ceil_div((size_t)M_ * KS_ * sdm_, 256);
                if (fp64)
                    pq_fold_kernel<double><<<fb, 256, 0, s>>>(
                        base_.as<float>(), dpad_, nn, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(),
                        M_, KS_, sdm_, cbase_.as<uint32_t>(), members_.as<uint32_t>(),
                        cb_.as<float>());
                else
                    pq_fold_kernel<float><<<fb, 256, 0, s>>>(
                        base_.as<float>(), dpad_, nn, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(),
                        M_, KS_, sdm_, cbase_.as<uint32_t>(), members_.as<uint32_t>(),
                        cb_.as<float>());
                HM_CHECK_LAUNCH();
            }
            if (ep + 1 < epochs) {
                pq_split_kernel<<<ceil_div(M_, 64), 64, 0, s>>>(
                    cb_.as<float>(), M_, KS_, sdm_, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(),
                    counts_.as<uint32_t>(), n, rng_.as<uint64_t>());
                HM_CHECK_LAUNCH();
            }
        }
        assign(codes_.as<uint8_t>(), M_, 1, s);
        if (metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR)
            codeword_norms(s);
        HM_HIP(hipStreamSynchronize(s));
        trained_ = true;
        sample_for_ = 0;
        n_trained_ = n;
    }

    // cbn[s][c] = the s_yy of the LUT's Cosine / Angular entries.
    void codeword_norms(hipStream_t s)
    {
        pq_norms_kernel<<<ceil_div((size_t)M_ * KS_, 256), 256, 0, s>>>(
            cb_.as<float>(), M_, KS_, sdm_, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(),
            cbn_.as<float>());
        HM_CHECK_LAUNCH();
    }

    hm_params p_;
    uint32_t dpad_ = 0, M_ = 0, KS_ = 0, sdm_ = 0;
    uint64_t n_ = 0, cap_ = 0, n_trained_ = 0, last_replayed_ = 0;
    int32_t metric_ = HM_METRIC_UNKNOWN;
    bool trained_ = false;
    DeviceBuffer base_, sub_b_, sub_e_, cb_, cbn_, asg_, counts_, codes_, rng_;
    DeviceBuffer ccnt_, coff_, cbase_, members_;  // update: items grouped by centre
    DeviceBuffer qpad_, D_, cand_d_, cand_i_, flags_, err_, lutg_, rp_rows_, rp_g_;
    // filter form: sampled codes, sample distances, thresholds, lists, and
    // the rows computed again with whole distance rows
    DeviceBuffer codes_s_, Ds_, tau_, lists_, lcnt_, redo_rows_, redo_q_, redo_i_, redo_d_,
        redo_c_;
    uint32_t sample_for_ = 0;  // item count the sampled codes were taken for
    uint64_t last_filtered_ = 0, last_uncert_ = 0;
};

}  // namespace

std::unique_ptr<Index> make_pq(uint64_t dim, const hm_params &p)
{
    return std::make_unique<PQ>(dim, p);
}

int32_t pq_info(Index *ix, uint32_t *out)
{
    auto *p = dynamic_cast<PQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not a PQIndex");
    p->info(out);
    return HM_OK;
}

const hm_params &pq_params(Index *ix)
{
    auto *p = dynamic_cast<PQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not a PQIndex");
    return p->params();
}

void pq_import(Index *ix, int32_t metric, const float *codebook, const uint8_t *codes,
               uint64_t n_trained)
{
    auto *p = dynamic_cast<PQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not a PQIndex");
    p->import_trained(metric, codebook, codes, n_trained);
}

int32_t pq_export(Index *ix, float *codebook, uint8_t *codes)
{
    auto *p = dynamic_cast<PQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not a PQIndex");
    p->export_trained(codebook, codes);
    return HM_OK;
}

}  // namespace hm
