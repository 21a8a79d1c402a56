// exact_dist.hip -- reference-order FP32 distance tiles for gfx950.
//
// Every (query, base) pair is accumulated by one lane, strictly in dimension
// order, with no contraction into FMA (`fp contract(off)` below and the
// build's -ffp-contract=off), so each value equals the reference's scalar
// left fold bit for bit:
//   Euclidean  acc = acc + (q_t - b_t)*(q_t - b_t)   simd_metrics.rs:96-102
//   Manhattan  acc = acc + |q_t - b_t|                simd_metrics.rs:60-66
//   Dot        acc = acc + q_t*b_t                    simd_metrics.rs:30-33
// Cosine and Angular are formed in the epilogue from the dot sum and the
// per-vector sequential sums s_qq, s_bb (metrics.rs:73-103).
//
// Tiling: a 256-thread workgroup owns a 128-query x 128-base tile; each lane
// holds an 8x8 register micro-tile, and dimension slabs of 16 are staged
// through LDS (transposed, double buffered) so every LDS read is a 16-byte
// broadcast-friendly load.  The work is VALU-bound (no MFMA: the products
// would be contracted and reordered), so the tile is sized for register
// reuse: 16 LDS bytes feed 24-48 VALU operations.
#include "kernels.h"
#include "profiler.h"

#pragma clang fp contract(off)

namespace hm {

namespace {

constexpr int TQ = 128;   // queries per workgroup tile
constexpr int TN = 128;   // base rows per workgroup tile
constexpr int TK = 16;    // dimensions per LDS slab
constexpr int LDS_PAD = 4;

template <int ACC>
__device__ __forceinline__ float acc_step(float acc, float q, float b)
{
    if constexpr (ACC == ACC_EUCLID) {
        float d = q - b;
        float s = d * d;
        return acc + s;
    } else if constexpr (ACC == ACC_MANHATTAN) {
        float d = q - b;
        return acc + fabsf(d);
    } else {
        float p = q * b;
        return acc + p;
    }
}

// metrics.rs:73-83 in the build's definition: -s_qb / (sqrt(s_qq)*sqrt(s_bb)).
__device__ __forceinline__ float epi_cosine(float sqb, float sqq, float sbb)
{
    float den = sqrtf(sqq) * sqrtf(sbb);
    return (-sqb) / den;
}

// metrics.rs:88-103 with float_two() == 1.0 (node.rs:54-56).
__device__ __forceinline__ float epi_angular(float sqb, float sqq, float sbb)
{
    float rhd = -sqq, lhd = -sbb, rldot = -sqb;
    float rlmul = rhd * lhd;
    if (rlmul > 0.0f) {
        float t = 1.0f * rldot;
        t = t / sqrtf(rlmul);
        return 1.0f - t;
    }
    return 1.0f;
}

__device__ __forceinline__ uint32_t fkey(float f)
{
    f = f + 0.0f;
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// ARGMIN: instead of writing D, fold each row's (distance, column) minimum
// into best[q] with a 64-bit atomicMin (lowest column among equal distances,
// the strict '<' scan of kmeans.rs:136-145).
template <int ACC, int EPI, bool ARGMIN>
__global__ __launch_bounds__(256, 2) void exact_tile_kernel(
    const float *__restrict__ Q, uint32_t ldq, uint32_t nq, const float *__restrict__ B,
    uint32_t ldb, uint32_t n, uint32_t dpad, const float *__restrict__ qnorm,
    const float *__restrict__ bnorm, float *__restrict__ D, uint64_t ldd,
    unsigned long long *__restrict__ best)
{
    __shared__ __align__(16) float sQ[2][TK][TQ + LDS_PAD];
    __shared__ __align__(16) float sB[2][TK][TN + LDS_PAD];

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const uint32_t q0 = blockIdx.y * TQ, n0 = blockIdx.x * TN;

    float acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = -0.0f;  // Rust's f32 Sum starts at -0.0

    // Register staging of the next 16-dim slab: two float4 per operand.  Rows
    // past the end load from a clamped (valid) row and are zeroed afterwards.
    float4 rq0, rq1, rb0, rb1;
    const int row0 = tid >> 2, row1 = (tid + 256) >> 2, c4 = (tid & 3) * 4;
    const bool qa0 = q0 + row0 < nq, qa1 = q0 + row1 < nq;
    const bool ba0 = n0 + row0 < n, ba1 = n0 + row1 < n;
    const float *qp0 = Q + (size_t)(qa0 ? q0 + row0 : nq - 1) * ldq + c4;
    const float *qp1 = Q + (size_t)(qa1 ? q0 + row1 : nq - 1) * ldq + c4;
    const float *bp0 = B + (size_t)(ba0 ? n0 + row0 : n - 1) * ldb + c4;
    const float *bp1 = B + (size_t)(ba1 ? n0 + row1 : n - 1) * ldb + c4;
    auto zero_if = [](float4 v, bool keep) {
        return keep ? v : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    auto gload = [&](uint32_t k0) {
        rq0 = zero_if(*reinterpret_cast<const float4 *>(qp0 + k0), qa0);
        rq1 = zero_if(*reinterpret_cast<const float4 *>(qp1 + k0), qa1);
        rb0 = zero_if(*reinterpret_cast<const float4 *>(bp0 + k0), ba0);
        rb1 = zero_if(*reinterpret_cast<const float4 *>(bp1 + k0), ba1);
    };
    auto put = [&](float (*S)[TQ + LDS_PAD], int row, float4 v) {
        S[c4 + 0][row] = v.x;
        S[c4 + 1][row] = v.y;
        S[c4 + 2][row] = v.z;
        S[c4 + 3][row] = v.w;
    };
    auto sstore = [&](int buf) {
        put(sQ[buf], row0, rq0);
        put(sQ[buf], row1, rq1);
        put(sB[buf], row0, rb0);
        put(sB[buf], row1, rb1);
    };

    const uint32_t nk = dpad / TK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (uint32_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * TK);
#pragma unroll 2
        for (int k = 0; k < TK; ++k) {
            float4 qa = *reinterpret_cast<const float4 *>(&sQ[cur][k][ty * 8]);
            float4 qb = *reinterpret_cast<const float4 *>(&sQ[cur][k][ty * 8 + 4]);
            float4 ba = *reinterpret_cast<const float4 *>(&sB[cur][k][tx * 8]);
            float4 bb = *reinterpret_cast<const float4 *>(&sB[cur][k][tx * 8 + 4]);
            const float qv[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
            const float bv[8] = {ba.x, ba.y, ba.z, ba.w, bb.x, bb.y, bb.z, bb.w};
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = acc_step<ACC>(acc[i][j], qv[i], bv[j]);
        }
        if (kt + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }

    const uint32_t bj0 = n0 + tx * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t qi = q0 + ty * 8 + i;
        const uint32_t qc = qi < nq ? qi : nq - 1;
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float a = acc[i][j];
            if constexpr (EPI == EPI_COSINE) {
                uint32_t bj = bj0 + j < n ? bj0 + j : n - 1;
                a = epi_cosine(a, qnorm[qc], bnorm[bj]);
            } else if constexpr (EPI == EPI_ANGULAR) {
                uint32_t bj = bj0 + j < n ? bj0 + j : n - 1;
                a = epi_angular(a, qnorm[qc], bnorm[bj]);
            }
            v[j] = a;
        }
        if constexpr (ARGMIN) {
            unsigned long long m = ~0ull;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                unsigned long long key = ((unsigned long long)fkey(v[j]) << 32) | (bj0 + j);
                if (bj0 + j < n && key < m) m = key;
            }
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                unsigned long long other = __shfl_xor(m, o, 64);
                m = other < m ? other : m;
            }
            if (tx == 0 && qi < nq) atomicMin(best + qi, m);
        } else {
            if (qi >= nq) continue;
            float *drow = D + (size_t)qi * ldd;
            if (bj0 + 7 < n) {
                *reinterpret_cast<float4 *>(drow + bj0) = make_float4(v[0], v[1], v[2], v[3]);
                *reinterpret_cast<float4 *>(drow + bj0 + 4) = make_float4(v[4], v[5], v[6], v[7]);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (bj0 + j < n) drow[bj0 + j] = v[j];
            }
        }
    }
}

// Few queries (a fallback batch): one lane per base row, the RQ query rows
// in LDS (broadcast reads), the base row streamed with 16-byte loads.  The
// same left folds and epilogues as exact_tile_kernel, so the values are the
// same bits; the 128-query tile would leave most of its work idle here.
template <int ACC, int EPI, int RQ>
__device__ __forceinline__ void exact_rows_body(
    const float *__restrict__ Q, uint32_t ldq, uint32_t nq, const float *__restrict__ B,
    uint32_t ldb, uint32_t n, uint32_t dpad, const float *__restrict__ qnorm,
    const float *__restrict__ bnorm, float *__restrict__ D, uint64_t ldd)
{
    extern __shared__ __align__(16) float xr_q[];  // [RQ][dpad]
    const uint32_t tid = threadIdx.x;
    for (uint32_t e = tid; e < RQ * dpad; e += 256) {
        const uint32_t r = e / dpad, t = e % dpad;
        xr_q[e] = r < nq ? Q[(size_t)r * ldq + t] : 0.0f;
    }
    __syncthreads();
    const uint32_t j = blockIdx.x * 256 + tid;
    const float *b = B + (size_t)(j < n ? j : n - 1) * ldb;
    float acc[RQ];
#pragma unroll
    for (int r = 0; r < RQ; ++r) acc[r] = -0.0f;  // Rust's f32 Sum starts at -0.0
#pragma unroll 4
    for (uint32_t t = 0; t < dpad; t += 4) {
        const float4 bv = *reinterpret_cast<const float4 *>(b + t);
#pragma unroll
        for (int r = 0; r < RQ; ++r) {
            const float4 qv = *reinterpret_cast<const float4 *>(xr_q + r * dpad + t);
            acc[r] = acc_step<ACC>(acc[r], qv.x, bv.x);
            acc[r] = acc_step<ACC>(acc[r], qv.y, bv.y);
            acc[r] = acc_step<ACC>(acc[r], qv.z, bv.z);
            acc[r] = acc_step<ACC>(acc[r], qv.w, bv.w);
        }
    }
    if (j >= n) return;
#pragma unroll
    for (int r = 0; r < RQ; ++r) {
        if ((uint32_t)r >= nq) break;
        float a = acc[r];
        if constexpr (EPI == EPI_COSINE) a = epi_cosine(a, qnorm[r], bnorm[j]);
        else if constexpr (EPI == EPI_ANGULAR) a = epi_angular(a, qnorm[r], bnorm[j]);
        D[(size_t)r * ldd + j] = a;
    }
}

// The same rows from the u8 copy of an integer base (every value is the f32
// of its byte, so the converted values, the folds and the bits are those of
// the f32 rows): a quarter of the bytes, 16 dimensions per 16-byte load.
template <int ACC, int EPI, int RQ>
__device__ __forceinline__ void exact_rows_body_u8(
    const float *__restrict__ Q, uint32_t ldq, uint32_t nq, const uint8_t *__restrict__ B8,
    uint32_t ld8, uint32_t n, uint32_t dpad, const float *__restrict__ qnorm,
    const float *__restrict__ bnorm, float *__restrict__ D, uint64_t ldd)
{
    extern __shared__ __align__(16) float xr_q[];  // [RQ][dpad]
    const uint32_t tid = threadIdx.x;
    for (uint32_t e = tid; e < RQ * dpad; e += 256) {
        const uint32_t r = e / dpad, t = e % dpad;
        xr_q[e] = r < nq ? Q[(size_t)r * ldq + t] : 0.0f;
    }
    __syncthreads();
    const uint32_t j = blockIdx.x * 256 + tid;
    const uint8_t *b = B8 + (size_t)(j < n ? j : n - 1) * ld8;
    float acc[RQ];
#pragma unroll
    for (int r = 0; r < RQ; ++r) acc[r] = -0.0f;
#pragma unroll 2
    for (uint32_t t = 0; t < dpad; t += 16) {
        const uint4 raw = *reinterpret_cast<const uint4 *>(b + t);
        const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float b0 = (float)(w[u] & 0xFFu), b1 = (float)((w[u] >> 8) & 0xFFu);
            const float b2 = (float)((w[u] >> 16) & 0xFFu), b3 = (float)(w[u] >> 24);
#pragma unroll
            for (int r = 0; r < RQ; ++r) {
                const float4 qv = *reinterpret_cast<const float4 *>(xr_q + r * dpad + t + 4 * u);
                acc[r] = acc_step<ACC>(acc[r], qv.x, b0);
                acc[r] = acc_step<ACC>(acc[r], qv.y, b1);
                acc[r] = acc_step<ACC>(acc[r], qv.z, b2);
                acc[r] = acc_step<ACC>(acc[r], qv.w, b3);
            }
        }
    }
    if (j >= n) return;
#pragma unroll
    for (int r = 0; r < RQ; ++r) {
        if ((uint32_t)r >= nq) break;
        float a = acc[r];
        if constexpr (EPI == EPI_COSINE) a = epi_cosine(a, qnorm[r], bnorm[j]);
        else if constexpr (EPI == EPI_ANGULAR) a = epi_angular(a, qnorm[r], bnorm[j]);
        D[(size_t)r * ldd + j] = a;
    }
}

template <int ACC, int EPI, int RQ>
__global__ __launch_bounds__(256) void exact_rows_kernel(
    const float *__restrict__ Q, uint32_t ldq, uint32_t nq, const float *__restrict__ B,
    uint32_t ldb, uint32_t n, uint32_t dpad, const float *__restrict__ qnorm,
    const float *__restrict__ bnorm, float *__restrict__ D, uint64_t ldd)
{
    exact_rows_body<ACC, EPI, RQ>(Q, ldq, nq, B, ldb, n, dpad, qnorm, bnorm, D, ldd);
}

// The first *dev_nq (<= RERUN_SLOTS) rows, eight per workgroup row of the
// grid (blockIdx.y): nothing past the count, the one-query body for a group
// of 1, the eight-query body otherwise (LDS for eight query rows); from the
// u8 copy B8 (rows dpad bytes apart) when it is given.
template <int ACC, int EPI>
__global__ __launch_bounds__(256) void exact_rows_dev_kernel(
    const float *__restrict__ Q, uint32_t ldq, const float *__restrict__ B, uint32_t ldb,
    uint32_t n, uint32_t dpad, const float *__restrict__ qnorm, const float *__restrict__ bnorm,
    float *__restrict__ D, uint64_t ldd, const uint32_t *__restrict__ dev_nq,
    const uint8_t *__restrict__ B8)
{
    const uint32_t g0 = blockIdx.y * 8u, all = min(*dev_nq, RERUN_SLOTS);
    if (g0 >= all) return;
    const uint32_t c = min(all - g0, 8u);
    Q += (size_t)g0 * ldq;
    if (qnorm) qnorm += g0;
    D += (size_t)g0 * ldd;
    if (B8) {
        if (c == 1) exact_rows_body_u8<ACC, EPI, 1>(Q, ldq, 1, B8, dpad, n, dpad, qnorm, bnorm, D, ldd);
        else exact_rows_body_u8<ACC, EPI, 8>(Q, ldq, c, B8, dpad, n, dpad, qnorm, bnorm, D, ldd);
        return;
    }
    if (c == 1) exact_rows_body<ACC, EPI, 1>(Q, ldq, 1, B, ldb, n, dpad, qnorm, bnorm, D, ldd);
    else exact_rows_body<ACC, EPI, 8>(Q, ldq, c, B, ldb, n, dpad, qnorm, bnorm, D, ldd);
}

// s = sum_t x_t*x_t as a left fold from -0.0 (the reference's dot(x,x) up to
// sign, calc.rs:13-18), one lane per row, for rows of any alignment and
// stride (self_dot_wide_kernel takes the aligned ones).  The workgroup's 64
// rows pass through LDS 16 dimensions at a time.
__global__ __launch_bounds__(64) void self_dot_kernel(const float *__restrict__ X, uint32_t ldx,
                                                      uint32_t rows, uint32_t d,
                                                      float *__restrict__ out,
                                                      const uint32_t *__restrict__ dev_rows)
{
    __shared__ float sl[64][17];
    const uint32_t nrows = dev_rows ? min(rows, *dev_rows) : rows;
    const uint32_t r0 = blockIdx.x * 64, tid = threadIdx.x;
    if (r0 >= nrows) return;
    float acc = -0.0f;
    for (uint32_t t0 = 0; t0 < d; t0 += 16) {
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t rl = tid / 4 + 16 * j, c = (tid % 4) * 4;
            const float *src = X + (size_t)min(r0 + rl, nrows - 1) * ldx + t0 + c;
            float4 v;
            v.x = t0 + c < d ? src[0] : 0.0f;
            v.y = t0 + c + 1 < d ? src[1] : 0.0f;
            v.z = t0 + c + 2 < d ? src[2] : 0.0f;
            v.w = t0 + c + 3 < d ? src[3] : 0.0f;
            sl[rl][c] = v.x;
            sl[rl][c + 1] = v.y;
            sl[rl][c + 2] = v.z;
            sl[rl][c + 3] = v.w;
        }
        __syncthreads();
        const uint32_t tn = min(16u, d - t0);
        for (uint32_t t = 0; t < tn; ++t) {
            const float x = sl[tid][t];
            const float p = x * x;
            acc = acc + p;
        }
        __syncthreads();
    }
    if (r0 + tid < nrows) out[r0 + tid] = acc;
}

// The same fold for rows 16-byte aligned with a stride that is a multiple of
// 16 floats: 256 threads per 64 rows load slabs of 128 dimensions (clamped
// addresses, values past d zeroed) into LDS, the next slab in flight while
// the first wave folds the current one, one lane per row in order.
__global__ __launch_bounds__(256) void self_dot_wide_kernel(const float *__restrict__ X,
                                                            uint32_t ldx, uint32_t rows, uint32_t d,
                                                            float *__restrict__ out,
                                                            const uint32_t *__restrict__ dev_rows)
{
    constexpr uint32_t SL = 128;
    __shared__ float sl[2][64][SL + 1];
    const uint32_t nrows = dev_rows ? min(rows, *dev_rows) : rows;
    const uint32_t r0 = blockIdx.x * 64, tid = threadIdx.x;
    if (r0 >= nrows) return;
    // thread tid loads float4 number tid + 256 j (j < 8) of the 64 x 128 slab:
    // row (tid + 256 j) / 32, column 4 ((tid + 256 j) % 32)
    const uint32_t cb = 4 * (tid % 32);
    const float *rp[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
        rp[j] = X + (size_t)min(r0 + (tid + 256 * j) / 32, nrows - 1) * ldx;
    float4 v[8];
    auto load = [&](uint32_t t0) {
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t c = t0 + cb;
            const float4 x = *reinterpret_cast<const float4 *>(rp[j] + min(c, ldx - 4));
            const bool ok = c < d;  // d <= ldx, both multiples of 4 past c
            v[j] = make_float4(ok ? x.x : 0.0f, c + 1 < d ? x.y : 0.0f, c + 2 < d ? x.z : 0.0f,
                               c + 3 < d ? x.w : 0.0f);
        }
    };
    load(0);
    float acc = -0.0f;
    for (uint32_t t0 = 0, b = 0; t0 < d; t0 += SL, b ^= 1) {
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            float *dst = &sl[b][(tid + 256 * j) / 32][cb];
            dst[0] = v[j].x;
            dst[1] = v[j].y;
            dst[2] = v[j].z;
            dst[3] = v[j].w;
        }
        __syncthreads();
        if (t0 + SL < d) load(t0 + SL);
        if (tid < 64) {
            const uint32_t tn = min(SL, d - t0);
            for (uint32_t t = 0; t < tn; ++t) {
                const float x = sl[b][tid][t];
                const float p = x * x;
                acc = acc + p;
            }
        }
    }
    if (tid < 64 && r0 + tid < nrows) out[r0 + tid] = acc;
}

// Batched distance service of the graph searches: one lane per (query,
// node) pair, metric(node, query) in the reference's fold order -- the
// operand order of hnsw_idx.rs:243-254 and ssg_idx.rs:402-456.  Pairs p in
// [off[q], off[q+1]) belong to query q.  A node index out of range writes
// NaN and sets *err bit 2.
template <int ACC, int EPI>
__global__ __launch_bounds__(64) void pair_dist_kernel(
    const float *__restrict__ Qp, uint32_t dpad, uint32_t nq, const float *__restrict__ B,
    uint32_t ldb, uint32_t n, const uint32_t *__restrict__ ids, const uint64_t *__restrict__ off,
    uint64_t total, const float *__restrict__ qnorm, const float *__restrict__ bnorm,
    float *__restrict__ out, uint32_t *__restrict__ err)
{
    // One wave of 64 pairs.  Each lane folds its own pair in order; the 64
    // node rows and query rows pass through LDS 16 dimensions at a time,
    // loaded as coalesced 64-byte row segments (rows padded to 16 floats).
    __shared__ float sb[64][17], sq[64][17];
    __shared__ uint32_t s_id[64], s_q[64];
    const uint32_t lane = threadIdx.x;
    const uint64_t p = (uint64_t)blockIdx.x * 64 + lane;
    const bool live = p < total;
    if (n == 0) {  // no rows: every index is out of range, nothing to load
        if (live) {
            out[p] = __int_as_float(0x7fc00000);
            atomicOr(err, 4u);
        }
        return;
    }
    uint32_t q = 0, id = 0;
    if (live) {
        uint32_t lo = 0, hi = nq;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (off[mid] <= p) lo = mid;
            else hi = mid;
        }
        q = lo;
        id = ids[p];
    }
    const bool bad = live && id >= n;
    s_id[lane] = (live && !bad) ? id : 0u;  // rows the loads may read
    s_q[lane] = q;
    __syncthreads();
    float acc = -0.0f;
    for (uint32_t t0 = 0; t0 < dpad; t0 += 16) {
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t r = lane / 4 + 16 * j, c = (lane % 4) * 4;
            const float4 bv = *reinterpret_cast<const float4 *>(B + (size_t)s_id[r] * ldb + t0 + c);
            const float4 qv = *reinterpret_cast<const float4 *>(Qp + (size_t)s_q[r] * dpad + t0 + c);
            sb[r][c] = bv.x;
            sb[r][c + 1] = bv.y;
            sb[r][c + 2] = bv.z;
            sb[r][c + 3] = bv.w;
            sq[r][c] = qv.x;
            sq[r][c + 1] = qv.y;
            sq[r][c + 2] = qv.z;
            sq[r][c + 3] = qv.w;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t t = 0; t < 16; ++t) acc = acc_step<ACC>(acc, sb[lane][t], sq[lane][t]);
        __syncthreads();
    }
    if (!live) return;
    if (bad) {
        out[p] = __int_as_float(0x7fc00000);
        atomicOr(err, 4u);
        return;
    }
    if constexpr (EPI == EPI_COSINE) acc = epi_cosine(acc, bnorm[id], qnorm[q]);
    else if constexpr (EPI == EPI_ANGULAR) acc = epi_angular(acc, bnorm[id], qnorm[q]);
    out[p] = acc;
}

}  // namespace

template <bool ARGMIN>
static void launch_exact_tile_impl(int metric, const float *Q, uint32_t ldq, uint32_t nq,
                                   const float *B, uint32_t ldb, uint32_t n, uint32_t dpad,
                                   const float *qnorm, const float *bnorm, float *D, uint64_t ldd,
                                   unsigned long long *best, hipStream_t s)
{
    if (nq == 0 || n == 0) return;
    // gridDim.y is at most 65535: longer query ranges run in slices
    constexpr uint32_t QMAX = 65535u * TQ;
    if (nq > QMAX) {
        for (uint32_t q0 = 0; q0 < nq; q0 += QMAX) {
            const uint32_t m = std::min(QMAX, nq - q0);
            launch_exact_tile_impl<ARGMIN>(metric, Q + (size_t)q0 * ldq, ldq, m, B, ldb, n, dpad,
                                           qnorm ? qnorm + q0 : nullptr, bnorm,
                                           D ? D + (size_t)q0 * ldd : nullptr, ldd,
                                           best ? best + q0 : nullptr, s);
        }
        return;
    }
    if (!ARGMIN && nq <= 8 && dpad % 4 == 0 && ldb % 4 == 0 && (size_t)8 * dpad * 4 <= 65536 &&
        ((uintptr_t)B & 15) == 0) {
        ProfScope prof("exact_rows", s);
        const unsigned blocks = ceil_div(n, 256);
#define HM_ROWS(ACC, EPI)                                                                      \
    do {                                                                                       \
        if (nq == 1)                                                                           \
            exact_rows_kernel<ACC, EPI, 1><<<blocks, 256, (size_t)dpad * 4, s>>>(              \
                Q, ldq, nq, B, ldb, n, dpad, qnorm, bnorm, D, ldd);                            \
        else if (nq <= 2)                                                                      \
            exact_rows_kernel<ACC, EPI, 2><<<blocks, 256, (size_t)2 * dpad * 4, s>>>(          \
                Q, ldq, nq, B, ldb, n, dpad, qnorm, bnorm, D, ldd);                            \
        else if (nq <= 4)                                                                      \
            exact_rows_kernel<ACC, EPI, 4><<<blocks, 256, (size_t)4 * dpad * 4, s>>>(          \
                Q, ldq, nq, B, ldb, n, dpad, qnorm, bnorm, D, ldd);                            \
        else                                                                                   \
            exact_rows_kernel<ACC, EPI, 8><<<blocks, 256, (size_t)8 * dpad * 4, s>>>(          \
                Q, ldq, nq, B, ldb, n, dpad, qnorm, bnorm, D, ldd);                            \
    } while (0)
        switch (metric) {
        case HM_METRIC_EUCLIDEAN: HM_ROWS(ACC_EUCLID, EPI_NONE); break;
        case HM_METRIC_MANHATTAN: HM_ROWS(ACC_MANHATTAN, EPI_NONE); break;
        case HM_METRIC_DOT_PRODUCT: HM_ROWS(ACC_DOT, EPI_NONE); break;
        case HM_METRIC_COSINE_SIMILARITY: HM_ROWS(ACC_DOT, EPI_COSINE); break;
        case HM_METRIC_ANGULAR: HM_ROWS(ACC_DOT, EPI_ANGULAR); break;
        default: fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        }
#undef HM_ROWS
        HM_CHECK_LAUNCH();
        return;
    }
    dim3 grid(ceil_div(n, TN), ceil_div(nq, TQ));
    dim3 block(256);
    ProfScope prof(ARGMIN ? "exact_argmin" : "exact_tile", s);
#define HM_EXACT(ACC, EPI)                                                                \
    exact_tile_kernel<ACC, EPI, ARGMIN><<<grid, block, 0, s>>>(Q, ldq, nq, B, ldb, n, dpad, \
                                                              qnorm, bnorm, D, ldd, best)
    switch (metric) {
    case HM_METRIC_EUCLIDEAN: HM_EXACT(ACC_EUCLID, EPI_NONE); break;
    case HM_METRIC_MANHATTAN: HM_EXACT(ACC_MANHATTAN, EPI_NONE); break;
    case HM_METRIC_DOT_PRODUCT: HM_EXACT(ACC_DOT, EPI_NONE); break;
    case HM_METRIC_COSINE_SIMILARITY: HM_EXACT(ACC_DOT, EPI_COSINE); break;
    case HM_METRIC_ANGULAR: HM_EXACT(ACC_DOT, EPI_ANGULAR); break;
    default: fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
    }
#undef HM_EXACT
    HM_CHECK_LAUNCH();
}

void launch_exact_tile(int metric, const float *Q, uint32_t ldq, uint32_t nq, const float *B,
                       uint32_t ldb, uint32_t n, uint32_t dpad, const float *qnorm,
                       const float *bnorm, float *D, uint64_t ldd, hipStream_t s)
{
    launch_exact_tile_impl<false>(metric, Q, ldq, nq, B, ldb, n, dpad, qnorm, bnorm, D, ldd,
                                  nullptr, s);
}

bool exact_rows_dev_ok(uint32_t dpad, uint32_t ldb, const float *B)
{
    return dpad % 4 == 0 && ldb % 4 == 0 && (size_t)8 * dpad * 4 <= 65536 &&
           ((uintptr_t)B & 15) == 0;
}

void launch_exact_rows_dev(int metric, const float *Q, uint32_t ldq, const float *B, uint32_t ldb,
                           uint32_t n, uint32_t dpad, const float *qnorm, const float *bnorm,
                           float *D, uint64_t ldd, const uint32_t *dev_nq, hipStream_t s,
                           const uint8_t *B8)
{
    if (n == 0) return;
    if (!exact_rows_dev_ok(dpad, ldb, B)) fail(HM_ERR_INVALID_ARG, "exact_rows_dev: shape");
    // the u8 rows need 16-byte aligned rows of dpad bytes
    if (B8 && (dpad % 16 != 0 || ((uintptr_t)B8 & 15) != 0)) B8 = nullptr;
    ProfScope prof("exact_rows", s);
    const unsigned blocks = ceil_div(n, 256);
#define HM_ROWS_DEV(ACC, EPI)                                                                  \
    exact_rows_dev_kernel<ACC, EPI><<<dim3(blocks, RERUN_SLOTS / 8), 256, (size_t)8 * dpad * 4, s>>>( \
        Q, ldq, B, ldb, n, dpad, qnorm, bnorm, D, ldd, dev_nq, B8)
    switch (metric) {
    case HM_METRIC_EUCLIDEAN: HM_ROWS_DEV(ACC_EUCLID, EPI_NONE); break;
    case HM_METRIC_MANHATTAN: HM_ROWS_DEV(ACC_MANHATTAN, EPI_NONE); break;
    case HM_METRIC_DOT_PRODUCT: HM_ROWS_DEV(ACC_DOT, EPI_NONE); break;
    case HM_METRIC_COSINE_SIMILARITY: HM_ROWS_DEV(ACC_DOT, EPI_COSINE); break;
    case HM_METRIC_ANGULAR: HM_ROWS_DEV(ACC_DOT, EPI_ANGULAR); break;
    default: fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
    }
#undef HM_ROWS_DEV
    HM_CHECK_LAUNCH();
}

void launch_exact_argmin(int metric, const float *Q, uint32_t ldq, uint32_t nq, const float *B,
                         uint32_t ldb, uint32_t n, uint32_t dpad, const float *qnorm,
                         const float *bnorm, unsigned long long *best, hipStream_t s)
{
    launch_exact_tile_impl<true>(metric, Q, ldq, nq, B, ldb, n, dpad, qnorm, bnorm, nullptr, 0,
                                 best, s);
}

void launch_pair_dist(int metric, const float *Qp, uint32_t dpad, uint32_t nq, const float *B,
                      uint32_t ldb, uint32_t n, const uint32_t *ids, const uint64_t *off,
                      uint64_t total, const float *qnorm, const float *bnorm, float *out,
                      uint32_t *err, hipStream_t s)
{
    if (total == 0 || nq == 0) return;
    if (dpad % 16 || ldb % 16 || ldb < dpad || ((uintptr_t)B & 15) || ((uintptr_t)Qp & 15))
        fail(HM_ERR_INVALID_ARG, "pair_dist: rows must be 16-byte aligned and padded to 16");
    ProfScope prof("pair_dist", s);
    const uint64_t nblk = (total + 63) / 64;
    if (nblk > 0x7FFFFFFFull) fail(HM_ERR_INVALID_ARG, "pair_dist: too many pairs");
    const unsigned blocks = (unsigned)nblk;
#define HM_PD(ACC, EPI)                                                                        \
    pair_dist_kernel<ACC, EPI><<<blocks, 64, 0, s>>>(Qp, dpad, nq, B, ldb, n, ids, off, total,   \
                                                     qnorm, bnorm, out, err)
    switch (metric) {
    case HM_METRIC_EUCLIDEAN: HM_PD(ACC_EUCLID, EPI_NONE); break;
    case HM_METRIC_MANHATTAN: HM_PD(ACC_MANHATTAN, EPI_NONE); break;
    case HM_METRIC_DOT_PRODUCT: HM_PD(ACC_DOT, EPI_NONE); break;
    case HM_METRIC_COSINE_SIMILARITY: HM_PD(ACC_DOT, EPI_COSINE); break;
    case HM_METRIC_ANGULAR: HM_PD(ACC_DOT, EPI_ANGULAR); break;
    default: fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
    }
#undef HM_PD
    HM_CHECK_LAUNCH();
}

void launch_self_dot(const float *X, uint32_t ldx, uint32_t rows, uint32_t d, float *out,
                     hipStream_t s, const uint32_t *dev_rows)
{
    if (rows == 0) return;
    if (ldx % 16 == 0 && ((uintptr_t)X & 15) == 0 && d <= ldx)
        self_dot_wide_kernel<<<ceil_div(rows, 64), 256, 0, s>>>(X, ldx, rows, d, out, dev_rows);
    else
        self_dot_kernel<<<ceil_div(rows, 64), 64, 0, s>>>(X, ldx, rows, d, out, dev_rows);
    HM_CHECK_LAUNCH();
}

}  // namespace hm
