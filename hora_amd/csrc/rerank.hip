// rerank.hip -- exact FP32 rerank of an MFMA shortlist, with certification.
//
// For each query the shortlist's candidates are rescored with the
// reference-order arithmetic (the same left folds as exact_dist.hip, so each
// value equals the reference's f32 result), sorted by (distance, position),
// and the first min(k, n) are the answer when two conditions hold:
//  * certified: every row outside the shortlist is provably farther than the
//    (k+1)-th exact distance inside it.  Outside rows have approximate key
//    >= key_s (the shortlist threshold), and |approx - exact| <= eps_q, an
//    a-priori bound on the BF16/FP32 rounding (documented per metric below),
//    so exact >= LB(key_s) - eps_q for all of them.
//  * tie-free: the k+1 smallest exact distances are pairwise distinct (the
//    reference's heap then returns them sorted, bruteforce_idx.rs:47-73).
// Rows that fail either test are flagged and the caller reruns them on the
// full exact path (which replays the heap for ties).
#include "kernels.h"
#include "profiler.h"

#pragma clang fp contract(off)

namespace hm {

namespace {

constexpr int RR_THREADS = 256;
constexpr uint32_t RR_CAP = 8192;
constexpr uint32_t RR_MAX_DIM = 4096;

__device__ __forceinline__ uint32_t fkey(float f)
{
    f = f + 0.0f;
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Candidate rows come from Bp (f32) or, U8, from Bp8: a u8 copy of the base
// whose values convert to the same f32 bits (checked when it is made), a
// quarter of the bytes to gather.
template <int ACC, int EPI, bool U8>
__global__ __launch_bounds__(RR_THREADS) void rerank_kernel(
    const float *__restrict__ Qp, uint32_t dpad, const float *__restrict__ Bp,
    const uint32_t *__restrict__ cand_i, const float *__restrict__ cand_key, uint32_t ldc,
    uint32_t s, const float *__restrict__ qnorm2, const float *__restrict__ bnorm2, uint32_t n,
    uint32_t k, float bmax, float gamma, float c_in_exact, float c_in_inexact,
    const uint32_t *__restrict__ q_inexact, int certify, uint64_t *__restrict__ out_idx,
    float *__restrict__ out_dist, uint32_t *__restrict__ out_count, uint32_t *__restrict__ flags,
    const uint32_t *__restrict__ dev_rows, const uint32_t *__restrict__ kq,
    const uint64_t *__restrict__ out_off, const uint32_t *__restrict__ s_q,
    const float *__restrict__ key_s_q, const uint32_t *__restrict__ pre_flags, int by_id,
    const uint8_t *__restrict__ Bp8)
{
    if (dev_rows && blockIdx.x >= *dev_rows) return;  // graph replays with fewer rows
    // dynamic LDS: keys[P] (u64) | vals[P] | sq[dpad]; P covers the largest s
    extern __shared__ __align__(16) unsigned char rr_smem[];
    __shared__ uint32_t s_tie, s_bad, s_nan;
    uint32_t P = 1;
    while (P < s) P <<= 1;
    if (s_q) s = min(s, s_q[blockIdx.x]);  // this query's shortlist length
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(rr_smem);
    float *vals = reinterpret_cast<float *>(rr_smem + (size_t)P * 8);
    float *sq = reinterpret_cast<float *>(rr_smem + (size_t)P * 12);
    float *sb = sq + dpad;  // [RR_THREADS][17]: a 16-dimension slab of the candidate rows
    __shared__ uint32_t s_rows[RR_THREADS];

    const int tid = threadIdx.x;
    const uint32_t r = blockIdx.x;
    if (tid == 0) {
        s_tie = 0;
        s_bad = 0;
        s_nan = 0;
    }
    for (uint32_t t = tid; t < dpad; t += RR_THREADS) sq[t] = Qp[(size_t)r * dpad + t];
    __syncthreads();

    const uint32_t *ci = cand_i + (size_t)r * ldc;
    const float qn2 = qnorm2[r];  // sum q_t^2 (left fold), needed by every bound
    // Candidates in groups of RR_THREADS, one per thread; each folds its own
    // row in reference order while the group's rows pass through LDS 16
    // dimensions at a time, loaded as coalesced 64-byte segments.
    for (uint32_t g0 = 0; g0 < s; g0 += RR_THREADS) {
        const uint32_t R = min((uint32_t)RR_THREADS, s - g0);
        const uint32_t j = g0 + tid;
        uint32_t idx = 0;
        if ((uint32_t)tid < R) idx = ci[j];
        const bool bad = (uint32_t)tid < R && idx >= n;
        s_rows[tid] = ((uint32_t)tid < R && !bad) ? idx : 0u;  // rows the loads may read
        __syncthreads();
        float acc = -0.0f;
        if (U8 && R <= 64) {
            // at most 64 rows of u8: 256-dimension slabs of 256 bytes per
            // row at a stride of 260 bytes (a lane's word w of its row is in
            // bank (lane + w) % 64: conflict-free), four 16-byte loads per
            // thread in flight per slab, so a query makes dpad / 256 round
            // trips; the same LDS as the [256][17] layout
            uint32_t *sw = reinterpret_cast<uint32_t *>(sb);
            for (uint32_t t0 = 0; t0 < dpad; t0 += 256) {
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    const uint32_t i = (uint32_t)tid + RR_THREADS * u;
                    const uint32_t row = i / 16, c = (i % 16) * 16;
                    if (row < R && t0 + c < dpad) {
                        const uint4 v = *reinterpret_cast<const uint4 *>(
                            Bp8 + (size_t)s_rows[row] * dpad + t0 + c);
                        uint32_t *d = sw + row * 65 + c / 4;
                        d[0] = v.x;
                        d[1] = v.y;
                        d[2] = v.z;
                        d[3] = v.w;
                    }
                }
                __syncthreads();
                if ((uint32_t)tid < R) {
                    const uint32_t tn = min(256u, dpad - t0);  // a multiple of 16
                    for (uint32_t w0 = 0; w0 < tn / 4; w0 += 4)
#pragma unroll
                    for (uint32_t w = w0; w < w0 + 4; ++w) {
                        const uint32_t word = sw[tid * 65 + w];
#pragma unroll
                        for (uint32_t b = 0; b < 4; ++b) {
                            const float q = sq[t0 + 4 * w + b];
                            const float bq = (float)((word >> (8 * b)) & 0xFFu);
                            if constexpr (ACC == ACC_EUCLID) {
                                float df = q - bq;
                                float sqr = df * df;
                                acc = acc + sqr;
                            } else if constexpr (ACC == ACC_MANHATTAN) {
                                float df = q - bq;
                                acc = acc + fabsf(df);
                            } else {
                                float p = q * bq;
                                acc = acc + p;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        } else if (R <= 64) {
            // at most 64 rows: 64-dimension slabs ([64][65] floats, the same
            // LDS as the [256][17] layout), four 16-byte loads per thread in
            // flight per slab, so a query makes dpad / 64 round trips
            float *sw = sb;
            for (uint32_t t0 = 0; t0 < dpad; t0 += 64) {
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    const uint32_t i = (uint32_t)tid + RR_THREADS * u;
                    const uint32_t row = i / 16, c = (i % 16) * 4;
                    if (row < R && t0 + c < dpad) {
                        const float4 v = *reinterpret_cast<const float4 *>(
                            Bp + (size_t)s_rows[row] * dpad + t0 + c);
                        float *d = sw + row * 65 + c;
                        d[0] = v.x;
                        d[1] = v.y;
                        d[2] = v.z;
                        d[3] = v.w;
                    }
                }
                __syncthreads();
                if ((uint32_t)tid < R) {
                    const uint32_t tn = min(64u, dpad - t0);  // a multiple of 16
                    for (uint32_t e0 = 0; e0 < tn; e0 += 16)
#pragma unroll
                    for (uint32_t e = e0; e < e0 + 16; ++e) {
                        const float q = sq[t0 + e], bq = sw[tid * 65 + e];
                        if constexpr (ACC == ACC_EUCLID) {
                            float df = q - bq;
                            float sqr = df * df;
                            acc = acc + sqr;
                        } else if constexpr (ACC == ACC_MANHATTAN) {
                            float df = q - bq;
                            acc = acc + fabsf(df);
                        } else {
                            float p = q * bq;
                            acc = acc + p;
                        }
                    }
                }
                __syncthreads();
            }
        } else if ((uint32_t)tid < R) {
            // more than 64 rows: each thread streams its own row in 16-byte
            // loads, 16 dimensions a step with the next steps' loads in
            // flight; no LDS staging and no barriers, so the gathers of a
            // deep shortlist keep many loads outstanding per CU
            const float *row = Bp + (size_t)s_rows[tid] * dpad;
            const uint8_t *row8 = U8 ? Bp8 + (size_t)s_rows[tid] * dpad : nullptr;
#pragma unroll 2
            for (uint32_t t0 = 0; t0 < dpad; t0 += 16) {
                float4 v[4];
                if constexpr (U8) {
                    const uint4 w = *reinterpret_cast<const uint4 *>(row8 + t0);
                    const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                        v[u] = make_float4((float)(ww[u] & 0xFFu), (float)((ww[u] >> 8) & 0xFFu),
                                           (float)((ww[u] >> 16) & 0xFFu), (float)(ww[u] >> 24));
                } else {
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                        v[u] = *reinterpret_cast<const float4 *>(row + t0 + 4 * u);
                }
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    const float bv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                    for (uint32_t e = 0; e < 4; ++e) {
                        const float q = sq[t0 + 4 * u + e], bq = bv[e];
                        if constexpr (ACC == ACC_EUCLID) {
                            float df = q - bq;
                            float sqr = df * df;
                            acc = acc + sqr;
                        } else if constexpr (ACC == ACC_MANHATTAN) {
                            float df = q - bq;
                            acc = acc + fabsf(df);
                        } else {
                            float p = q * bq;
                            acc = acc + p;
                        }
                    }
                }
            }
        }
        if ((uint32_t)tid >= R) continue;
        if (bad) {  // never expected: the read stayed in range; flag the row
            s_bad = 1;
            vals[j] = __int_as_float(0x7f800000);
            keys[j] = ~0ull;
            continue;
        }
        float E = acc;
        if constexpr (EPI == EPI_COSINE) {
            float den = sqrtf(qn2) * sqrtf(bnorm2[idx]);
            E = (-acc) / den;
        } else if constexpr (EPI == EPI_ANGULAR) {
            float rlmul = (-qn2) * (-bnorm2[idx]);
            if (rlmul > 0.0f) {
                float t = 1.0f * (-acc);
                t = t / sqrtf(rlmul);
                E = 1.0f - t;
            } else {
                E = 1.0f;
            }
        }
        if (E != E) s_nan = 1;  // the exact path reports it as the reference fails
        vals[j] = E;
        keys[j] = ((unsigned long long)fkey(E) << 32) | j;
    }
    for (uint32_t j = s + tid; j < P; j += RR_THREADS) keys[j] = ~0ull;
    __syncthreads();
    {
        uint32_t Pq = 1;  // this row's sort length
        while (Pq < s) Pq <<= 1;
        P = Pq;
    }
    for (uint32_t size = 2; size <= P; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < P / 2; t += RR_THREADS) {
                uint32_t lo = bitonic_lo(t, stride);
                uint32_t hi = lo + stride;
                bool asc = (lo & size) == 0;
                unsigned long long a = keys[lo], b = keys[hi];
                if ((a > b) == asc) {
                    keys[lo] = b;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }

    // ragged batches: this query's own k and output offset
    if (kq) k = kq[r];
    const size_t obase = out_off ? out_off[r] : (size_t)r * k;
    const uint32_t m = k < n ? k : n;
    const uint32_t K1 = (k + 1 < s) ? k + 1 : s;  // entries checked for ties
    uint64_t *oi = out_idx + obase;
    float *od = out_dist + obase;
    const float inf = __int_as_float(0x7f800000);
    if (by_id) {
        // order by (distance, row index): each sorted entry moves to its
        // rank by index inside its run of equal distances
        const uint32_t mm = m < s ? m : s;
        for (uint32_t j = tid; j < s; j += RR_THREADS) {
            const uint32_t kj = (uint32_t)(keys[j] >> 32);
            uint32_t a = j, b = j + 1;
            while (a > 0 && (uint32_t)(keys[a - 1] >> 32) == kj) --a;
            while (b < s && (uint32_t)(keys[b] >> 32) == kj) ++b;
            const uint32_t pos = (uint32_t)keys[j];
            if (a >= mm || pos >= s) continue;  // past the result, or a flagged bad row
            const uint32_t id = ci[pos];
            uint32_t slot = a;
            for (uint32_t t = a; t < b; ++t) {
                const uint32_t pt = (uint32_t)keys[t];
                slot += pt < s && ci[pt] < id;
            }
            if (slot < mm) {
                oi[slot] = id;
                od[slot] = vals[pos];
            }
        }
        for (uint32_t j = mm + tid; j < k; j += RR_THREADS) {
            oi[j] = ~0ull;
            od[j] = inf;
        }
    }
    for (uint32_t j = tid; j < k && !by_id; j += RR_THREADS) {
        if (j < m && j < s) {
            const uint32_t pos = (uint32_t)keys[j];
            oi[j] = ci[pos];
            od[j] = vals[pos];
        } else {
            oi[j] = ~0ull;
            od[j] = inf;
        }
        if (j + 1 < K1 && (uint32_t)(keys[j] >> 32) == (uint32_t)(keys[j + 1] >> 32)) s_tie = 1;
    }
    __syncthreads();
    if (tid != 0) return;
    out_count[r] = m < s ? m : s;
    uint32_t f = (s_tie ? ROW_TIE : 0u) | (s_bad ? ROW_BAD : 0u) | (s_nan ? ROW_NAN : 0u) |
                 (pre_flags ? pre_flags[r] : 0u);
    if (s < n && k + 1 > s) {
        f |= ROW_UNCERT;  // too few candidates to hold the k+1 nearest
    } else if (s < n && certify) {
        bool ok = false;
        {
            const float ek1 = vals[(uint32_t)keys[k]];
            const float key_s = key_s_q ? key_s_q[r] : cand_key[(size_t)r * ldc + s - 1];
            const double c_in = q_inexact && *q_inexact ? c_in_inexact : c_in_exact;
            const double g = gamma, u = 5.9604644775390625e-8;  // 2^-24
            const double nq = sqrt((double)qn2);
            double lb;
            if constexpr (ACC == ACC_EUCLID) {
                // |-2(s~ - s)| <= 2(c_in + g)|q||b|; norms: g(|q|^2+|b|^2); the
                // reference's own fold: g * E; the combination: 4u(|q|+|b|)^2.
                double tot = (nq + bmax) * (nq + bmax);
                double eps = 2.0 * (c_in + g) * nq * bmax + g * ((double)qn2 + (double)bmax * bmax) +
                             (g + 4.0 * u) * tot;
                lb = (double)qn2 + (double)key_s - 1.01 * eps - 1e-30;
            } else if constexpr (EPI == EPI_NONE) {  // DotProduct
                double eps = (c_in + 2.0 * g + 2.0 * u) * nq * bmax;
                lb = (double)key_s - 1.01 * eps - 1e-30;
            } else {  // Cosine / Angular in units of cos
                double eps = c_in + 6.0 * g + 16.0 * u;
                double base = (double)key_s / nq;
                lb = (EPI == EPI_COSINE ? base : 1.0 + base) - 1.01 * eps;
            }
            ok = (double)ek1 < lb;
        }
        if (!ok) f |= ROW_UNCERT;
    }
    flags[r] = f;
}

__device__ __forceinline__ float unfkey(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

constexpr int LP_THREADS = 256;

// One workgroup per query: sorts the query's appended (key, row) entries
// and emits the first min(cnt, S).
__global__ __launch_bounds__(LP_THREADS) void list_prepare_kernel(
    const uint64_t *__restrict__ list, uint32_t cap, const uint32_t *__restrict__ cnt,
    const float *__restrict__ tau, uint32_t tau_ld, uint32_t S, uint32_t *__restrict__ cand_i,
    float *__restrict__ cand_key, uint32_t ldc, uint32_t *__restrict__ s_q,
    float *__restrict__ key_s_q, uint32_t *__restrict__ flags, const uint32_t *__restrict__ dev_rows,
    const uint32_t *__restrict__ kq)
{
    const uint32_t r = blockIdx.x;
    if (dev_rows && r >= *dev_rows) return;
    if (kq) {  // ragged batches: this query's own default depth max(64, 4(k+1)), at most S
        const uint32_t kr = kq[r];
        const uint64_t own = kr >= 15 ? 4ull * (kr + 1) : 64ull;
        S = (uint32_t)(own < S ? own : S);
    }
    extern __shared__ __align__(16) unsigned long long lp_keys[];
    const int tid = threadIdx.x;
    const uint32_t raw = cnt[r], c = raw < cap ? raw : cap;
    uint32_t P = 1;
    while (P < c) P <<= 1;
    const uint64_t *src = list + (size_t)r * cap;
    for (uint32_t j = tid; j < P; j += LP_THREADS) lp_keys[j] = j < c ? src[j] : ~0ull;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < P / 2; t += LP_THREADS) {
                const uint32_t lo = bitonic_lo(t, stride), hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const unsigned long long a = lp_keys[lo], b = lp_keys[hi];
                if ((a > b) == asc) {
                    lp_keys[lo] = b;
                    lp_keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    const uint32_t s = c < S ? c : S;
    for (uint32_t j = tid; j < s; j += LP_THREADS) {
        cand_i[(size_t)r * ldc + j] = (uint32_t)lp_keys[j];
        cand_key[(size_t)r * ldc + j] = unfkey((uint32_t)(lp_keys[j] >> 32));
    }
    if (tid == 0) {
        s_q[r] = s;
        // every row left out has key >= the smallest key left out of the
        // list (if any) >= ... and every row never appended has key >= tau
        const float t = tau[(size_t)r * tau_ld];
        const float left = c > S ? unfkey((uint32_t)(lp_keys[S] >> 32)) : t;
        key_s_q[r] = left < t ? left : t;
        flags[r] = raw > cap ? ROW_UNCERT : 0u;
    }
}

__global__ void gather_rows_kernel(const float *__restrict__ src, uint64_t lds,
                                   const uint32_t *__restrict__ rows, uint32_t nrows, uint32_t d,
                                   float *__restrict__ dst, uint64_t ldd)
{
    uint64_t total = (uint64_t)nrows * d;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = e / d, c = e % d;
        const uint32_t row = rows[r];
        if (row != ~0u) dst[r * ldd + c] = src[(uint64_t)row * lds + c];
    }
}

__global__ void scatter_results_kernel(const uint64_t *__restrict__ si, const float *__restrict__ sd,
                                       const uint32_t *__restrict__ sc,
                                       const uint32_t *__restrict__ rows, uint32_t nrows,
                                       uint32_t k, uint64_t *__restrict__ di,
                                       float *__restrict__ dd, uint32_t *__restrict__ dc)
{
    uint64_t total = (uint64_t)nrows * k;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = e / k, j = e % k;
        const uint32_t row = rows[r];
        if (row == ~0u) continue;
        di[(uint64_t)row * k + j] = si[e];
        dd[(uint64_t)row * k + j] = sd[e];
    }
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nrows;
         r += (uint64_t)gridDim.x * blockDim.x)
        if (rows[r] != ~0u) dc[rows[r]] = sc[r];
}

// Exact-path results ([m][kmax]) of rows[] copied into a ragged output:
// query q keeps its first kq[q] results at out_off[q].
__global__ void scatter_ragged_kernel(const uint64_t *__restrict__ si, const float *__restrict__ sd,
                                      const uint32_t *__restrict__ sc,
                                      const uint32_t *__restrict__ rows, uint32_t nrows,
                                      uint32_t kmax, const uint32_t *__restrict__ kq,
                                      const uint64_t *__restrict__ out_off,
                                      uint64_t *__restrict__ di, float *__restrict__ dd,
                                      uint32_t *__restrict__ dc)
{
    const uint32_t r = blockIdx.x;
    if (r >= nrows) return;
    const uint32_t q = rows[r], kr = kq[q];
    for (uint32_t j = threadIdx.x; j < kr; j += blockDim.x) {
        di[out_off[q] + j] = si[(size_t)r * kmax + j];
        dd[out_off[q] + j] = sd[(size_t)r * kmax + j];
    }
    if (threadIdx.x == 0) dc[q] = sc[r] < kr ? sc[r] : kr;
}

__global__ void u64_to_u32_kernel(const uint64_t *__restrict__ src, uint32_t *__restrict__ dst,
                                  uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = (uint32_t)src[i];
}

// one workgroup of 64 threads per chunk row
__global__ __launch_bounds__(64) void ragged_chunk_begin_kernel(
    const RaggedCall *__restrict__ call, uint32_t c, uint32_t QC, uint32_t dim, uint32_t dpad,
    const uint32_t *__restrict__ kq, const uint64_t *__restrict__ off, uint32_t *__restrict__ rows,
    uint32_t *__restrict__ ck, uint64_t *__restrict__ coff, float *__restrict__ qpad,
    uint32_t *err)
{
    const uint64_t Q = call->Q, q0 = (uint64_t)c * QC;
    const uint32_t nr = Q > q0 ? (uint32_t)min<uint64_t>(QC, Q - q0) : 0u;
    const uint32_t r = blockIdx.x, tid = threadIdx.x;
    if (r == 0 && tid == 0) *rows = nr;
    if (r >= nr) return;
    const float *src = call->dq + (q0 + r) * dim;
    float *dst = qpad + (size_t)r * dpad;
    bool bad = false;
    for (uint32_t t = tid; t < dpad; t += 64) {
        const float v = t < dim ? src[t] : 0.0f;
        bad |= !isfinite(v);
        dst[t] = v;
    }
    if (bad) atomicOr(err, 2u);
    if (tid == 0) {
        ck[r] = kq[q0 + r];
        coff[r] = off[q0 + r] - off[q0];
    }
}

__global__ void ragged_chunk_end_kernel(const RaggedCall *__restrict__ call, uint32_t c,
                                        uint32_t QC, const uint32_t *__restrict__ rows,
                                        const uint64_t *__restrict__ off,
                                        const uint64_t *__restrict__ oi,
                                        const float *__restrict__ od,
                                        const uint32_t *__restrict__ oc,
                                        const uint32_t *__restrict__ rf,
                                        uint32_t *__restrict__ flags_all)
{
    const uint32_t nr = *rows;
    const uint64_t q0 = (uint64_t)c * QC;
    const uint64_t o0 = nr ? off[q0] : 0, tot = nr ? off[q0 + nr] - o0 : 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint64_t e = t0; e < tot; e += stride) {
        call->d_idx[o0 + e] = oi[e];
        call->d_dist[o0 + e] = od[e];
    }
    for (uint64_t r = t0; r < QC; r += stride) {
        if (r < nr) call->d_count[q0 + r] = oc[r];
        flags_all[q0 + r] = r < nr ? rf[r] : 0u;
    }
}

__global__ void ragged_rerun_pad_kernel(const RaggedCall *__restrict__ call,
                                        const uint32_t *__restrict__ plan, uint32_t dim,
                                        uint32_t dpad, float *__restrict__ dst)
{
    const uint32_t cnt = plan[2 * RERUN_SLOTS];
    const uint32_t slot = blockIdx.x;
    if (slot >= cnt) return;
    const float *src = call->dq + (uint64_t)plan[slot] * dim;
    for (uint32_t t = threadIdx.x; t < dpad; t += blockDim.x)
        dst[(size_t)slot * dpad + t] = t < dim ? src[t] : 0.0f;
}

__global__ void ragged_rerun_scatter_kernel(const RaggedCall *__restrict__ call,
                                            const uint32_t *__restrict__ plan,
                                            const uint64_t *__restrict__ si,
                                            const float *__restrict__ sd,
                                            const uint32_t *__restrict__ sc, uint32_t kmax,
                                            const uint32_t *__restrict__ kq,
                                            const uint64_t *__restrict__ off)
{
    const uint32_t cnt = plan[2 * RERUN_SLOTS];
    const uint32_t slot = blockIdx.x;
    if (slot >= cnt) return;
    const uint32_t q = plan[slot], kr = kq[q];
    for (uint32_t j = threadIdx.x; j < kr; j += blockDim.x) {
        call->d_idx[off[q] + j] = si[(size_t)slot * kmax + j];
        call->d_dist[off[q] + j] = sd[(size_t)slot * kmax + j];
    }
    if (threadIdx.x == 0) call->d_count[q] = sc[slot] < kr ? sc[slot] : kr;
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t n, uint32_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

void launch_scatter_ragged(const uint64_t *si, const float *sd, const uint32_t *sc,
                           const uint32_t *rows, uint32_t nrows, uint32_t kmax, const uint32_t *kq,
                           const uint64_t *out_off, uint64_t *di, float *dd, uint32_t *dc,
                           hipStream_t s)
{
    if (nrows == 0) return;
    scatter_ragged_kernel<<<nrows, 64, 0, s>>>(si, sd, sc, rows, nrows, kmax, kq, out_off, di, dd,
                                               dc);
    HM_CHECK_LAUNCH();
}

void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s)
{
    if (n == 0) return;
    fill_u32_kernel<<<std::min<uint32_t>(ceil_div(n, 256), 1024), 256, 0, s>>>(p, n, v);
    HM_CHECK_LAUNCH();
}

void launch_rerank(int metric, const float *Qp, uint32_t dpad, const float *Bp,
                   const uint32_t *cand_i, const float *cand_key, uint32_t ldc, uint32_t s,
                   const float *qnorm2, const float *bnorm2, uint32_t rows, uint32_t n,
                   uint32_t k, float bmax, float gamma, float c_in_exact, float c_in_inexact,
                   const uint32_t *q_inexact, int certify, uint64_t *out_idx, float *out_dist,
                   uint32_t *out_count, uint32_t *flags, hipStream_t st,
                   const uint32_t *dev_rows, const uint32_t *kq, const uint64_t *out_off,
                   const uint32_t *s_q, const float *key_s_q, const uint32_t *pre_flags, int by_id,
                   const uint8_t *Bp8)
{
    if (rows == 0) return;
    if (s == 0 || s > RR_CAP) fail(HM_ERR_INVALID_ARG, "rerank: shortlist size out of range");
    if (dpad > RR_MAX_DIM) fail(HM_ERR_INVALID_ARG, "rerank: dimension above 4096");
    ProfScope prof("rerank", st);
    uint32_t P = 1;
    while (P < s) P <<= 1;
    if (dpad % 16 || ((uintptr_t)Bp & 15) || ((uintptr_t)Bp8 & 15))
        fail(HM_ERR_INVALID_ARG, "rerank: rows must be 16-byte aligned and padded to 16");
    const size_t lds = (size_t)P * 12 + (size_t)dpad * 4 + (size_t)RR_THREADS * 17 * 4;
#define HM_RR2(ACC, EPI, U8)                                                                   \
    rerank_kernel<ACC, EPI, U8><<<rows, RR_THREADS, lds, st>>>(                                 \
        Qp, dpad, Bp, cand_i, cand_key, ldc, s, qnorm2, bnorm2, n, k, bmax, gamma, c_in_exact, \
        c_in_inexact, q_inexact, certify, out_idx, out_dist, out_count, flags, dev_rows, kq,    \
        out_off, s_q, key_s_q, pre_flags, by_id, Bp8)
#define HM_RR(ACC, EPI)                                                                        \
    if (Bp8) HM_RR2(ACC, EPI, true);                                                           \
    else HM_RR2(ACC, EPI, false)
    switch (metric) {
    case HM_METRIC_MANHATTAN: HM_RR(ACC_MANHATTAN, EPI_NONE); break;  // no certification bound
    case HM_METRIC_EUCLIDEAN: HM_RR(ACC_EUCLID, EPI_NONE); break;
    case HM_METRIC_DOT_PRODUCT: HM_RR(ACC_DOT, EPI_NONE); break;
    case HM_METRIC_COSINE_SIMILARITY: HM_RR(ACC_DOT, EPI_COSINE); break;
    case HM_METRIC_ANGULAR: HM_RR(ACC_DOT, EPI_ANGULAR); break;
    default: fail(HM_ERR_INVALID_ARG, "rerank: metric has no shortlist form");
    }
#undef HM_RR
#undef HM_RR2
    HM_CHECK_LAUNCH();
}

void launch_ragged_chunk_begin(const RaggedCall *call, uint32_t c, uint32_t QC, uint32_t dim,
                               uint32_t dpad, const uint32_t *kq, const uint64_t *off,
                               uint32_t *rows, uint32_t *ck, uint64_t *coff, float *qpad,
                               uint32_t *err, hipStream_t s)
{
    ragged_chunk_begin_kernel<<<QC, 64, 0, s>>>(call, c, QC, dim, dpad, kq, off, rows, ck, coff,
                                               qpad, err);
    HM_CHECK_LAUNCH();
}

void launch_ragged_chunk_end(const RaggedCall *call, uint32_t c, uint32_t QC,
                             const uint32_t *rows, const uint64_t *off, const uint64_t *oi,
                             const float *od, const uint32_t *oc, const uint32_t *rf,
                             uint32_t *flags_all, hipStream_t s)
{
    ragged_chunk_end_kernel<<<1024, 256, 0, s>>>(call, c, QC, rows, off, oi, od, oc, rf,
                                                 flags_all);
    HM_CHECK_LAUNCH();
}

void launch_ragged_rerun_pad(const RaggedCall *call, const uint32_t *plan, uint32_t dim,
                             uint32_t dpad, float *dst, hipStream_t s)
{
    ragged_rerun_pad_kernel<<<RERUN_SLOTS, 256, 0, s>>>(call, plan, dim, dpad, dst);
    HM_CHECK_LAUNCH();
}

void launch_ragged_rerun_scatter(const RaggedCall *call, const uint32_t *plan, const uint64_t *si,
                                 const float *sd, const uint32_t *sc, uint32_t kmax,
                                 const uint32_t *kq, const uint64_t *off, hipStream_t s)
{
    ragged_rerun_scatter_kernel<<<RERUN_SLOTS, 256, 0, s>>>(call, plan, si, sd, sc, kmax, kq, off);
    HM_CHECK_LAUNCH();
}

void launch_u64_to_u32(const uint64_t *src, uint32_t *dst, uint64_t n, hipStream_t s)
{
    if (n == 0) return;
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    u64_to_u32_kernel<<<blocks, 256, 0, s>>>(src, dst, n);
    HM_CHECK_LAUNCH();
}

void launch_list_prepare(const uint64_t *list, uint32_t cap, const uint32_t *cnt,
                         const float *tau, uint32_t tau_ld, uint32_t rows, uint32_t S,
                         uint32_t *cand_i, float *cand_key, uint32_t ldc, uint32_t *s_q,
                         float *key_s_q, uint32_t *flags, hipStream_t st,
                         const uint32_t *dev_rows, const uint32_t *kq)
{
    if (rows == 0) return;
    if (cap == 0 || cap > RR_CAP || S == 0 || S > ldc)
        fail(HM_ERR_INVALID_ARG, "list_prepare: capacity out of range");
    ProfScope prof("list_prepare", st);
    uint32_t P = 1;
    while (P < cap) P <<= 1;
    list_prepare_kernel<<<rows, LP_THREADS, (size_t)P * 8, st>>>(
        list, cap, cnt, tau, tau_ld, S, cand_i, cand_key, ldc, s_q, key_s_q, flags, dev_rows, kq);
    HM_CHECK_LAUNCH();
}

void launch_gather_rows(const float *src, uint64_t lds, const uint32_t *rows, uint32_t nrows,
                        uint32_t d, float *dst, uint64_t ldd, hipStream_t s)
{
    if (nrows == 0) return;
    unsigned blocks = (unsigned)std::min<uint64_t>(((uint64_t)nrows * d + 255) / 256, 4096);
    gather_rows_kernel<<<blocks, 256, 0, s>>>(src, lds, rows, nrows, d, dst, ldd);
    HM_CHECK_LAUNCH();
}

void launch_scatter_results(const uint64_t *si, const float *sd, const uint32_t *sc,
                            const uint32_t *rows, uint32_t nrows, uint32_t k, uint64_t *di,
                            float *dd, uint32_t *dc, hipStream_t s)
{
    if (nrows == 0) return;
    unsigned blocks = (unsigned)std::min<uint64_t>(((uint64_t)nrows * std::max(k, 1u) + 255) / 256, 4096);
    scatter_results_kernel<<<blocks, 256, 0, s>>>(si, sd, sc, rows, nrows, k, di, dd, dc);
    HM_CHECK_LAUNCH();
}

}  // namespace hm
