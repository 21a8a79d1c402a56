// mfma_scores.hip -- approximate query x base scores on the CDNA4 matrix cores.
//
// The shortlist forms of BruteForceIndex rank the base by an approximate key
// computed from one GEMM S = Q . B^T on MFMA with FP32 accumulation:
//   Euclidean   key = ||b||^2 - 2 s          (||q||^2 is constant per query)
//   DotProduct  key = s                      (the reference ranks +sum(q*b) ascending)
//   Cosine      key = -s / ||b||             (E ~ key / ||q||)
//   Angular     key =  s / ||b||             (E ~ 1 + key / ||q||)
// and the shortlist is then reranked with the exact reference-order FP32
// arithmetic (rerank.hip).  Operands are BF16 (v_mfma_f32_16x16x32_bf16) or
// FP8-e4m3 with a per-vector scale (v_mfma_f32_16x16x32_fp8_fp8).
//
// Tiling: 256 threads = 4 waves in a 2x2 arrangement; the workgroup owns a
// 128-query x 128-base tile, each wave 64x64 = 4x4 MFMA 16x16 tiles (64 fp32
// accumulators per lane).  Each K step stages 64 bytes per row of both
// operands in LDS (double buffered through registers).
#include "kernels.h"
#include "profiler.h"

namespace hm {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int GM = 128, GN = 128;
constexpr int ROWB = 64;            // bytes per row per K step
constexpr int LDS_ROW = ROWB + 16;  // padded LDS row stride (bytes)

template <int ELEM>  // 2 = bf16, 1 = fp8
__global__ __launch_bounds__(256, 2) void score_gemm_kernel(
    const uint8_t *__restrict__ Q, uint32_t nq, const uint8_t *__restrict__ B, uint32_t n,
    uint32_t row_bytes, const float *__restrict__ qscale, const float *__restrict__ bscale,
    const float *__restrict__ bnorm2, const float *__restrict__ inv_bnorm, int metric,
    float *__restrict__ Kmat, uint64_t ldk, const uint32_t *__restrict__ dev_rows)
{
    if (dev_rows && *dev_rows < nq) nq = *dev_rows;
    if (blockIdx.y * GM >= nq) return;
    __shared__ __align__(16) uint8_t sA[2][GM * LDS_ROW];
    __shared__ __align__(16) uint8_t sB[2][GN * LDS_ROW];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const uint32_t q0 = blockIdx.y * GM, n0 = blockIdx.x * GN;

    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Register staging of the next K step: two 16-byte chunks per operand.
    // Rows past the end load from a clamped (valid) row and are zeroed
    // afterwards, so every load is a plain global load of a real row.
    uint4 ra0, ra1, rb0, rb1;
    const int row0 = tid >> 2, row1 = (tid + 256) >> 2, cc = (tid & 3) * 16;
    const bool qa0 = q0 + row0 < nq, qa1 = q0 + row1 < nq;
    const bool ba0 = n0 + row0 < n, ba1 = n0 + row1 < n;
    const uint8_t *qp0 = Q + (size_t)(qa0 ? q0 + row0 : nq - 1) * row_bytes + cc;
    const uint8_t *qp1 = Q + (size_t)(qa1 ? q0 + row1 : nq - 1) * row_bytes + cc;
    const uint8_t *bp0 = B + (size_t)(ba0 ? n0 + row0 : n - 1) * row_bytes + cc;
    const uint8_t *bp1 = B + (size_t)(ba1 ? n0 + row1 : n - 1) * row_bytes + cc;
    const uint32_t mq0 = qa0 ? ~0u : 0u, mq1 = qa1 ? ~0u : 0u;
    const uint32_t mb0 = ba0 ? ~0u : 0u, mb1 = ba1 ? ~0u : 0u;
#define HM_MASK4(v, m) make_uint4((v).x & (m), (v).y & (m), (v).z & (m), (v).w & (m))
#define HM_GLOAD(kb)                                                                   \
    do {                                                                               \
        const uint4 t0 = *reinterpret_cast<const uint4 *>(qp0 + (kb));                 \
        const uint4 t1 = *reinterpret_cast<const uint4 *>(qp1 + (kb));                 \
        const uint4 t2 = *reinterpret_cast<const uint4 *>(bp0 + (kb));                 \
        const uint4 t3 = *reinterpret_cast<const uint4 *>(bp1 + (kb));                 \
        ra0 = HM_MASK4(t0, mq0);                                                       \
        ra1 = HM_MASK4(t1, mq1);                                                       \
        rb0 = HM_MASK4(t2, mb0);                                                       \
        rb1 = HM_MASK4(t3, mb1);                                                       \
    } while (0)
#define HM_SSTORE(buf)                                                           \
    do {                                                                         \
        *reinterpret_cast<uint4 *>(&sA[buf][row0 * LDS_ROW + cc]) = ra0;         \
        *reinterpret_cast<uint4 *>(&sA[buf][row1 * LDS_ROW + cc]) = ra1;         \
        *reinterpret_cast<uint4 *>(&sB[buf][row0 * LDS_ROW + cc]) = rb0;         \
        *reinterpret_cast<uint4 *>(&sB[buf][row1 * LDS_ROW + cc]) = rb1;         \
    } while (0)

    const uint32_t nk = row_bytes / ROWB;
    HM_GLOAD(0u);
    HM_SSTORE(0);
    __syncthreads();
    const int lr = lane & 15, lk = lane >> 4;
    for (uint32_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) HM_GLOAD((kt + 1) * ROWB);
        const uint8_t *As = sA[cur];
        const uint8_t *Bs = sB[cur];
        if constexpr (ELEM == 2) {
            // one MFMA K=32 (64 bytes) per step; lane holds 8 bf16 at k = 8*lk
            bf16x8 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(As + (wr * 64 + i * 16 + lr) * LDS_ROW + lk * 16);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                b[j] = *reinterpret_cast<const bf16x8 *>(Bs + (wc * 64 + j * 16 + lr) * LDS_ROW + lk * 16);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        } else {
            // two MFMA K=32 (32 bytes each) per step; lane holds 8 fp8 at k = 8*lk
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                long a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a[i] = *reinterpret_cast<const long *>(As + (wr * 64 + i * 16 + lr) * LDS_ROW + kk * 32 + lk * 8);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    b[j] = *reinterpret_cast<const long *>(Bs + (wc * 64 + j * 16 + lr) * LDS_ROW + kk * 32 + lk * 8);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
        if (kt + 1 < nk) HM_SSTORE(cur ^ 1);
        __syncthreads();
    }
#undef HM_GLOAD
#undef HM_SSTORE
#undef HM_MASK4

    // Epilogue: lane holds D[4*lk + e][lr] of each 16x16 tile.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t q = q0 + wr * 64 + i * 16 + lk * 4 + e;
            if (q >= nq) continue;
            const float qs = (ELEM == 1) ? qscale[q] : 1.0f;
            float *krow = Kmat + (size_t)q * ldk;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t c = n0 + wc * 64 + j * 16 + lr;
                if (c >= n) continue;
                float s = acc[i][j][e];
                if (ELEM == 1) s = s * (qs * bscale[c]);
                float key;
                if (metric == HM_METRIC_EUCLIDEAN) key = bnorm2[c] - 2.0f * s;
                else if (metric == HM_METRIC_DOT_PRODUCT) key = s;
                else if (metric == HM_METRIC_COSINE_SIMILARITY) key = -s * inv_bnorm[c];
                else key = s * inv_bnorm[c];
                krow[c] = key;
            }
        }
    }
}

// Round-to-nearest-even f32 -> bf16; flags a value that does not survive.
__global__ void to_bf16_kernel(const float *__restrict__ src, uint64_t lds, uint64_t rows,
                               uint32_t d, uint32_t dk, uint16_t *__restrict__ dst,
                               uint32_t *inexact, const uint32_t *__restrict__ dev_rows)
{
    if (dev_rows && *dev_rows < rows) rows = *dev_rows;
    uint64_t total = rows * dk;
    uint32_t bad = 0;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = e / dk, c = e % dk;
        float x = c < d ? src[r * lds + c] : 0.0f;
        uint32_t u = __float_as_uint(x);
        uint32_t h = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
        if (__uint_as_float(h << 16) != x) bad = 1;
        dst[r * dk + c] = (uint16_t)h;
    }
    if (bad && inexact) atomicOr(inexact, 1u);
}

// Per-row scaled FP8-e4m3: x * (224 / max|x|) rounded by the hardware
// converter; dequant[r] = max|x| / 224.  One wave per row.
__global__ void to_fp8_kernel(const float *__restrict__ src, uint64_t lds, uint32_t rows,
                              uint32_t d, uint32_t dk, uint8_t *__restrict__ dst,
                              float *__restrict__ dequant, const uint32_t *__restrict__ dev_rows)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows || (dev_rows && r >= *dev_rows)) return;
    const float *x = src + (size_t)r * lds;
    float m = 0.0f;
    for (uint32_t c = lane; c < d; c += 64) m = fmaxf(m, fabsf(x[c]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    const float scale = m > 0.0f ? 224.0f / m : 1.0f;
    if (lane == 0) dequant[r] = m > 0.0f ? m / 224.0f : 0.0f;
    uint8_t *o = dst + (size_t)r * dk;
    for (uint32_t c = 2 * lane; c < dk; c += 128) {
        float a = c < d ? x[c] * scale : 0.0f;
        float b = c + 1 < d ? x[c + 1] * scale : 0.0f;
        int v = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
        o[c] = (uint8_t)(v & 0xFF);
        if (c + 1 < dk) o[c + 1] = (uint8_t)((v >> 8) & 0xFF);
    }
}

// inv[r] = 1 / sqrt(norm2[r]); maxnorm = max_r sqrt(norm2[r]) (as float bits).
__global__ void norm_stats_kernel(const float *__restrict__ norm2, uint32_t rows,
                                  float *__restrict__ inv, uint32_t *maxbits)
{
    float m = 0.0f;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows;
         r += gridDim.x * blockDim.x) {
        float s = sqrtf(norm2[r]);
        if (inv) inv[r] = 1.0f / s;
        m = fmaxf(m, s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, __float_as_uint(m));
}

}  // namespace

void launch_score_gemm(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                       uint32_t row_bytes, const float *qscale, const float *bscale,
                       const float *bnorm2, const float *inv_bnorm, int metric, float *Kmat,
                       uint64_t ldk, hipStream_t s, const uint32_t *dev_rows)
{
    if (nq == 0 || n == 0) return;
    if (row_bytes % ROWB) fail(HM_ERR_INVALID_ARG, "score_gemm: row bytes not a multiple of 64");
    dim3 grid(ceil_div(n, GN), ceil_div(nq, GM));
    ProfScope prof("score_gemm", s);
    if (elem_bytes == 2)
        score_gemm_kernel<2><<<grid, 256, 0, s>>>((const uint8_t *)Q, nq, (const uint8_t *)B, n,
                                                  row_bytes, qscale, bscale, bnorm2, inv_bnorm,
                                                  metric, Kmat, ldk, dev_rows);
    else
        score_gemm_kernel<1><<<grid, 256, 0, s>>>((const uint8_t *)Q, nq, (const uint8_t *)B, n,
                                                  row_bytes, qscale, bscale, bnorm2, inv_bnorm,
                                                  metric, Kmat, ldk, dev_rows);
    HM_CHECK_LAUNCH();
}

void launch_to_bf16(const float *src, uint64_t lds, uint64_t rows, uint32_t d, uint32_t dk,
                    uint16_t *dst, uint32_t *inexact, hipStream_t s, const uint32_t *dev_rows)
{
    if (rows == 0) return;
    unsigned blocks = (unsigned)std::min<uint64_t>((rows * dk + 255) / 256, 8192);
    to_bf16_kernel<<<blocks, 256, 0, s>>>(src, lds, rows, d, dk, dst, inexact, dev_rows);
    HM_CHECK_LAUNCH();
}

void launch_to_fp8(const float *src, uint64_t lds, uint32_t rows, uint32_t d, uint32_t dk,
                   uint8_t *dst, float *dequant, hipStream_t s, const uint32_t *dev_rows)
{
    if (rows == 0) return;
    to_fp8_kernel<<<ceil_div(rows, 4), 256, 0, s>>>(src, lds, rows, d, dk, dst, dequant, dev_rows);
    HM_CHECK_LAUNCH();
}

void launch_norm_stats(const float *norm2, uint32_t rows, float *inv, uint32_t *maxbits,
                       hipStream_t s)
{
    if (rows == 0) return;
    unsigned blocks = std::min<unsigned>(ceil_div(rows, 256), 1024);
    norm_stats_kernel<<<blocks, 256, 0, s>>>(norm2, rows, inv, maxbits);
    HM_CHECK_LAUNCH();
}

}  // namespace hm
