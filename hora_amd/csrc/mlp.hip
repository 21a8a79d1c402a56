// mlp.hip -- the probe predictor's network: BF16 MFMA GEMMs with fused
// bias/ReLU epilogues, softmax cross-entropy, bias gradients and Adam.
//
// The network maps a query (d floats) to a score per inverted list:
//   H = relu(X W1 + b1),  Z = H W2 + b2,  p = softmax(Z)
// trained with cross-entropy against the fraction of the query's exact
// top-k that each list holds.  Weights are FP32 masters with BF16 copies;
// every product runs on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
#include "mlp.h"
#include "profiler.h"

#include <cmath>
#include <vector>

namespace hm {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

// C(m,n) = sum_k A(m,k) B(k,n) with A(m,k) = A[m*sam + k*sak] and
// B(k,n) = B[k*sbk + n*sbn] (bf16).  Epilogue: + bias[n], ReLU, then f32
// and/or bf16 stores.  256 threads, 64x64 tile, 4 waves of 32x32.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(
    const uint16_t *__restrict__ A, long sam, long sak, const uint16_t *__restrict__ B, long sbk,
    long sbn, int M, int N, int K, const float *__restrict__ bias, int relu,
    float *__restrict__ C32, long ldc32, uint16_t *__restrict__ C16, long ldc16)
{
    __shared__ __align__(16) uint16_t sA[BM * LDK];
    __shared__ __align__(16) uint16_t sB[BN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int lr = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += BK) {
        // stage: 2048 elements of each operand, 8 per thread
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int idx = tid * 8 + e;
            int r = idx / BK, c = idx % BK;
            int gm = m0 + r, gk = k0 + c;
            sA[r * LDK + c] = (gm < M && gk < K) ? A[(long)gm * sam + (long)gk * sak] : (uint16_t)0;
            int gn = n0 + r;
            sB[r * LDK + c] = (gn < N && gk < K) ? B[(long)gk * sbk + (long)gn * sbn] : (uint16_t)0;
        }
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const bf16x8 *>(&sA[(wr * 32 + i * 16 + lr) * LDK + lk * 8]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            b[j] = *reinterpret_cast<const bf16x8 *>(&sB[(wc * 32 + j * 16 + lr) * LDK + lk * 8]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int m = m0 + wr * 32 + i * 16 + lk * 4 + e;
                int n = n0 + wc * 32 + j * 16 + lr;
                if (m >= M || n >= N) continue;
                float v = acc[i][j][e];
                if (bias) v += bias[n];
                if (relu) v = fmaxf(v, 0.0f);
                if (C32) C32[(long)m * ldc32 + n] = v;
                if (C16) C16[(long)m * ldc16 + n] = f2bf(v);
            }
}

// C(m,n) = sum_k A[m][k] B[n][k] for row-major bf16 A [M][lda] and B
// [N][ldb] (k contiguous in both; K, lda, ldb multiples of 32, 8, 8).  The
// inference GEMMs of the network (B is a transposed weight copy).  256
// threads, a 128x128 tile of four 64x64 wave tiles (4x4 MFMA 16x16x32
// blocks), K stages of 64 (two MFMA steps) staged through two LDS buffers
// with the next stage's 16-byte global loads in flight during the current
// stage's 32 MFMAs per wave.  Rows of the LDS tiles are 72 elements (144
// bytes) apart.  Epilogue: + bias[n], optional ReLU, optional negation, f32
// and/or bf16 stores.
constexpr int NT_B = 128, NT_BK = 64, NT_LDK = NT_BK + 8;

__global__ __launch_bounds__(256) void gemm_nt_bf16_kernel(
    const uint16_t *__restrict__ A, long lda, const uint16_t *__restrict__ B, long ldb, int M,
    int N, int K, const float *__restrict__ bias, int relu, int negate, float *__restrict__ C32,
    long ldc32, uint16_t *__restrict__ C16, long ldc16)
{
    extern __shared__ __align__(16) uint16_t nt_smem[];
    uint16_t *sA = nt_smem, *sB = nt_smem + 2 * NT_B * NT_LDK;  // [2][128][72] each
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int m0 = blockIdx.y * NT_B, n0 = blockIdx.x * NT_B;
    const int lr = lane & 15, lk = lane >> 4;
    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // chunk c of a 128 x 64 operand tile: row c / 8, k offset (c % 8) * 8;
    // a stage of K = 32 (the last one when K % 64 == 32) loads chunks with
    // (c % 8) < 4 only
    uint4 ra[4], rb[4];
    const uint4 zero = make_uint4(0, 0, 0, 0);
    auto load = [&](int k0) {
        const int kw = min(NT_BK, K - k0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = tid + 256 * u, r = c >> 3, kc = (c & 7) * 8;
            const int gm = m0 + r, gn = n0 + r;
            const bool kin = kc < kw;
            ra[u] = (gm < M && kin) ? *reinterpret_cast<const uint4 *>(A + (long)gm * lda + k0 + kc)
                                    : zero;
            rb[u] = (gn < N && kin) ? *reinterpret_cast<const uint4 *>(B + (long)gn * ldb + k0 + kc)
                                    : zero;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = tid + 256 * u, r = c >> 3, kc = (c & 7) * 8;
            *reinterpret_cast<uint4 *>(&sA[(buf * NT_B + r) * NT_LDK + kc]) = ra[u];
            *reinterpret_cast<uint4 *>(&sB[(buf * NT_B + r) * NT_LDK + kc]) = rb[u];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += NT_BK) {
        const bool more = k0 + NT_BK < K;
        if (more) load(k0 + NT_BK);
        const int steps = min(NT_BK, K - k0) / 32;
        for (int ks = 0; ks < steps; ++ks) {
            bf16x8 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(
                    &sA[(buf * NT_B + wr * 64 + i * 16 + lr) * NT_LDK + ks * 32 + lk * 8]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                b[j] = *reinterpret_cast<const bf16x8 *>(
                    &sB[(buf * NT_B + wc * 64 + j * 16 + lr) * NT_LDK + ks * 32 + lk * 8]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wr * 64 + i * 16 + lk * 4 + e;
                const int n = n0 + wc * 64 + j * 16 + lr;
                if (m >= M || n >= N) continue;
                float v = acc[i][j][e];
                if (bias) v += bias[n];
                if (relu) v = fmaxf(v, 0.0f);
                if (negate) v = -v;
                if (C32) C32[(long)m * ldc32 + n] = v;
                if (C16) C16[(long)m * ldc16 + n] = f2bf(v);
            }
}

// dst[c][r] = src[r][c] (bf16 [rows][cols] -> [cols][ldd]; columns of dst
// past `rows` are zero).
__global__ void transpose_bf16_kernel(const uint16_t *__restrict__ src, int rows, int cols,
                                      uint16_t *__restrict__ dst, int ldd)
{
    const long total = (long)cols * ldd;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int c = (int)(e / ldd), r = (int)(e % ldd);
        dst[e] = r < rows ? src[(long)r * cols + c] : (uint16_t)0;
    }
}

// bf16 rows [rows][dpad] from f32 rows [rows][ld] (d valid, zero padded).
__global__ void rows_to_bf16_kernel(const float *__restrict__ x, long ld, long rows, int d,
                                    uint16_t *__restrict__ y, int dpad)
{
    const long total = rows * dpad;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const long r = e / dpad;
        const int c = (int)(e % dpad);
        y[e] = c < d ? f2bf(x[r * ld + c]) : (uint16_t)0;
    }
}

// lse[r] = log sum_j exp(-negz[r][j]) (the log-partition of the logits
// z = -negz).  One workgroup per row.
__global__ __launch_bounds__(256) void row_lse_kernel(const float *__restrict__ negz, long ld,
                                                      int cols, float *__restrict__ lse)
{
    __shared__ float red[256];
    const int r = blockIdx.x, tid = threadIdx.x;
    const float *z = negz + (long)r * ld;
    float mx = -3.0e38f;
    for (int j = tid; j < cols; j += 256) mx = fmaxf(mx, -z[j]);
    red[tid] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float sum = 0.f;
    for (int j = tid; j < cols; j += 256) sum += expf(-z[j] - mx);
    red[tid] = sum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) lse[r] = logf(red[0]) + mx;
}

__global__ void f32_to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ y, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

// Row softmax of Z (rows x cols); G = (p - T) * scale (bf16 and optional
// f32), loss[row] = -sum_j T_j log p_j.  One workgroup per row.
__global__ __launch_bounds__(256) void softmax_xent_kernel(
    const float *__restrict__ Z, const float *__restrict__ T, int cols, float scale,
    uint16_t *__restrict__ G16, float *__restrict__ G32, float *__restrict__ loss,
    float *__restrict__ P)
{
    __shared__ float red[256];
    const int r = blockIdx.x, tid = threadIdx.x;
    const float *z = Z + (long)r * cols;
    float mx = -3.0e38f;
    for (int j = tid; j < cols; j += 256) mx = fmaxf(mx, z[j]);
    red[tid] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float sum = 0.f;
    for (int j = tid; j < cols; j += 256) sum += expf(z[j] - mx);
    red[tid] = sum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const float lse = logf(red[0]) + mx;
    __syncthreads();
    float l = 0.f;
    for (int j = tid; j < cols; j += 256) {
        float p = expf(z[j] - lse);
        if (P) P[(long)r * cols + j] = p;
        if (T) {
            float t = T[(long)r * cols + j];
            float g = (p - t) * scale;
            if (G16) G16[(long)r * cols + j] = f2bf(g);
            if (G32) G32[(long)r * cols + j] = g;
            if (t > 0.f) l -= t * (z[j] - lse);
        }
    }
    if (T && loss) {
        red[tid] = l;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) loss[r] = red[0];
    }
}

// bias gradient: out[j] = sum_r G[r][j] (bf16 rows) or masked by H > 0.
__global__ void colsum_bf16_kernel(const uint16_t *__restrict__ G, int rows, int cols,
                                   float *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += bf2f(G[(long)r * cols + j]);
    out[j] = s;
}

// dH = dH_pre * (H > 0), in place on the bf16 buffer.
__global__ void relu_mask_kernel(uint16_t *__restrict__ dH, const uint16_t *__restrict__ H, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (!(bf2f(H[i]) > 0.f)) dH[i] = 0;
}

// Adam (Kingma & Ba 2015) on FP32 masters, refreshing the bf16 copy.
__global__ void adam_kernel(float *__restrict__ w, uint16_t *__restrict__ wb,
                            const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long n, float lr, float b1, float b2,
                            float eps, float c1, float c2, float wd)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float gi = g[i] + wd * w[i];
        float mi = b1 * m[i] + (1.f - b1) * gi;
        float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

inline unsigned grid1(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

}  // namespace

void gemm_bf16(const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn,
               int M, int N, int K, const float *bias, bool relu, float *C32, long ldc32,
               uint16_t *C16, long ldc16, hipStream_t s)
{
    if (M <= 0 || N <= 0) return;
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
    ProfScope prof("mlp_gemm", s);
    gemm_bf16_kernel<<<grid, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, M, N, K, bias, relu ? 1 : 0,
                                          C32, ldc32, C16, ldc16);
    HM_CHECK_LAUNCH();
}

void gemm_nt_bf16(const uint16_t *A, long lda, const uint16_t *B, long ldb, int M, int N, int K,
                  const float *bias, bool relu, bool negate, float *C32, long ldc32,
                  uint16_t *C16, long ldc16, hipStream_t s)
{
    if (M <= 0 || N <= 0) return;
    if (K % 32 || lda % 8 || ldb % 8 || ((uintptr_t)A & 15) || ((uintptr_t)B & 15))
        fail(HM_ERR_INVALID_ARG, "gemm_nt_bf16: K, strides or alignment");
    dim3 grid((N + NT_B - 1) / NT_B, (M + NT_B - 1) / NT_B);
    ProfScope prof("mlp_gemm", s);
    const size_t lds = (size_t)4 * NT_B * NT_LDK * sizeof(uint16_t);  // 72 KiB
    gemm_nt_bf16_kernel<<<grid, 256, lds, s>>>(A, lda, B, ldb, M, N, K, bias, relu ? 1 : 0,
                                               negate ? 1 : 0, C32, ldc32, C16, ldc16);
    HM_CHECK_LAUNCH();
}

void row_lse(const float *negz, long ld, int rows, int cols, float *lse, hipStream_t s)
{
    if (rows <= 0) return;
    row_lse_kernel<<<rows, 256, 0, s>>>(negz, ld, cols, lse);
    HM_CHECK_LAUNCH();
}

void rows_to_bf16(const float *x, long ld, long rows, int d, uint16_t *y, int dpad, hipStream_t s)
{
    if (rows <= 0) return;
    rows_to_bf16_kernel<<<grid1(rows * dpad), 256, 0, s>>>(x, ld, rows, d, y, dpad);
    HM_CHECK_LAUNCH();
}

void f32_to_bf16(const float *x, uint16_t *y, long n, hipStream_t s)
{
    if (n <= 0) return;
    f32_to_bf16_kernel<<<grid1(n), 256, 0, s>>>(x, y, n);
    HM_CHECK_LAUNCH();
}

void softmax_xent(const float *Z, const float *T, int rows, int cols, float scale, uint16_t *G16,
                  float *G32, float *loss, float *P, hipStream_t s)
{
    if (rows <= 0) return;
    softmax_xent_kernel<<<rows, 256, 0, s>>>(Z, T, cols, scale, G16, G32, loss, P);
    HM_CHECK_LAUNCH();
}

void colsum_bf16(const uint16_t *G, int rows, int cols, float *out, hipStream_t s)
{
    colsum_bf16_kernel<<<(cols + 255) / 256, 256, 0, s>>>(G, rows, cols, out);
    HM_CHECK_LAUNCH();
}

void relu_mask(uint16_t *dH, const uint16_t *H, long n, hipStream_t s)
{
    relu_mask_kernel<<<grid1(n), 256, 0, s>>>(dH, H, n);
    HM_CHECK_LAUNCH();
}

void adam_step(float *w, uint16_t *wb, const float *g, float *m, float *v, long n, float lr,
               int step, float wd, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    adam_kernel<<<grid1(n), 256, 0, s>>>(w, wb, g, m, v, n, lr, b1, b2, eps, c1, c2, wd);
    HM_CHECK_LAUNCH();
}

}  // namespace hm

// ---------------------------------------------------------------- host model
namespace hm {

namespace {
uint64_t mix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

ProbeMLP::ProbeMLP(int din, int hidden, int dout, uint64_t seed, hipStream_t s)
    : din_(din), hid_(hidden), dout_(dout), s_(s)
{
    if (din <= 0 || hidden <= 0 || dout <= 0) fail(HM_ERR_INVALID_ARG, "MLP shape");
    const size_t n1 = (size_t)din * hidden, n2 = (size_t)hidden * dout;
    for (DeviceBuffer *b : {&W1_, &gW1_, &mW1_, &vW1_}) b->reserve(n1 * 4);
    for (DeviceBuffer *b : {&W2_, &gW2_, &mW2_, &vW2_}) b->reserve(n2 * 4);
    for (DeviceBuffer *b : {&b1_, &gb1_, &mb1_, &vb1_}) b->reserve((size_t)hidden * 4);
    for (DeviceBuffer *b : {&b2_, &gb2_, &mb2_, &vb2_}) b->reserve((size_t)dout * 4);
    W1b_.reserve(n1 * 2);
    W2b_.reserve(n2 * 2);
    for (DeviceBuffer *b : {&mW1_, &vW1_, &mW2_, &vW2_, &mb1_, &vb1_, &mb2_, &vb2_})
        HM_HIP(hipMemsetAsync(b->get(), 0, b->bytes(), s_));
    // Glorot-uniform weights, zero biases (seeded SplitMix64).
    std::vector<float> w1(n1), w2(n2), z1(hidden, 0.f), z2(dout, 0.f);
    uint64_t st = seed ^ 0xA5A5A5A5DEADBEEFull;
    const float a1 = std::sqrt(6.0f / (float)(din + hidden)), a2 = std::sqrt(6.0f / (float)(hidden + dout));
    for (auto &x : w1) x = a1 * (2.0f * (float)((mix64(&st) >> 40) * (1.0 / 16777216.0)) - 1.0f);
    for (auto &x : w2) x = a2 * (2.0f * (float)((mix64(&st) >> 40) * (1.0 / 16777216.0)) - 1.0f);
    set_weights(w1.data(), z1.data(), w2.data(), z2.data());
}

void ProbeMLP::set_weights(const float *W1, const float *b1, const float *W2, const float *b2)
{
    const size_t n1 = (size_t)din_ * hid_, n2 = (size_t)hid_ * dout_;
    HM_HIP(hipMemcpyAsync(W1_.get(), W1, n1 * 4, hipMemcpyHostToDevice, s_));
    HM_HIP(hipMemcpyAsync(b1_.get(), b1, (size_t)hid_ * 4, hipMemcpyHostToDevice, s_));
    HM_HIP(hipMemcpyAsync(W2_.get(), W2, n2 * 4, hipMemcpyHostToDevice, s_));
    HM_HIP(hipMemcpyAsync(b2_.get(), b2, (size_t)dout_ * 4, hipMemcpyHostToDevice, s_));
    f32_to_bf16(W1_.as<float>(), W1b_.as<uint16_t>(), (long)n1, s_);
    f32_to_bf16(W2_.as<float>(), W2b_.as<uint16_t>(), (long)n2, s_);
    HM_HIP(hipStreamSynchronize(s_));
    inf_valid_ = false;
}

void ProbeMLP::scores(const uint16_t *Xp, int n, float *negz, long ldz, float *lse)
{
    const int dinp = din_padded(), hidp = (hid_ + 31) / 32 * 32;
    if (!inf_valid_) {
        // k-contiguous bf16 copies of the weights, zero padded to multiples
        // of 32: W1t [hidp][dinp], W2t [dout][hidp]; the bias b1 padded with
        // zeros, so the padding units of H are relu(0) = 0
        W1t_.reserve((size_t)hidp * dinp * 2);
        W2t_.reserve((size_t)dout_ * hidp * 2);
        b1p_.reserve((size_t)hidp * 4);
        HM_HIP(hipMemsetAsync(W1t_.get(), 0, (size_t)hidp * dinp * 2, s_));
        HM_HIP(hipMemsetAsync(b1p_.get(), 0, (size_t)hidp * 4, s_));
        HM_HIP(hipMemcpyAsync(b1p_.get(), b1_.get(), (size_t)hid_ * 4, hipMemcpyDeviceToDevice, s_));
        transpose_bf16_kernel<<<grid1((long)hid_ * dinp), 256, 0, s_>>>(
            W1b_.as<uint16_t>(), din_, hid_, W1t_.as<uint16_t>(), dinp);
        HM_CHECK_LAUNCH();
        transpose_bf16_kernel<<<grid1((long)dout_ * hidp), 256, 0, s_>>>(
            W2b_.as<uint16_t>(), hid_, dout_, W2t_.as<uint16_t>(), hidp);
        HM_CHECK_LAUNCH();
        inf_valid_ = true;
    }
    if (n > hinf_cap_) {
        Hinf_.reserve((size_t)n * hidp * 2);
        hinf_cap_ = n;
    }
    // H = relu(X W1 + b1) (bf16), then -(H W2 + b2) and the rows' log-partition
    gemm_nt_bf16(Xp, dinp, W1t_.as<uint16_t>(), dinp, n, hidp, dinp, b1p_.as<float>(), true,
                 false, nullptr, 0, Hinf_.as<uint16_t>(), hidp, s_);
    gemm_nt_bf16(Hinf_.as<uint16_t>(), hidp, W2t_.as<uint16_t>(), hidp, n, dout_, hidp,
                 b2_.as<float>(), false, true, negz, ldz, nullptr, 0, s_);
    row_lse(negz, ldz, n, dout_, lse, s_);
}

void ProbeMLP::get_weights(float *W1, float *b1, float *W2, float *b2) const
{
    const size_t n1 = (size_t)din_ * hid_, n2 = (size_t)hid_ * dout_;
    if (W1) HM_HIP(hipMemcpyAsync(W1, W1_.get(), n1 * 4, hipMemcpyDeviceToHost, s_));
    if (b1) HM_HIP(hipMemcpyAsync(b1, b1_.get(), (size_t)hid_ * 4, hipMemcpyDeviceToHost, s_));
    if (W2) HM_HIP(hipMemcpyAsync(W2, W2_.get(), n2 * 4, hipMemcpyDeviceToHost, s_));
    if (b2) HM_HIP(hipMemcpyAsync(b2, b2_.get(), (size_t)dout_ * 4, hipMemcpyDeviceToHost, s_));
    HM_HIP(hipStreamSynchronize(s_));
}

void ProbeMLP::get_grads(float *gW1, float *gb1, float *gW2, float *gb2) const
{
    const size_t n1 = (size_t)din_ * hid_, n2 = (size_t)hid_ * dout_;
    if (gW1) HM_HIP(hipMemcpyAsync(gW1, gW1_.get(), n1 * 4, hipMemcpyDeviceToHost, s_));
    if (gb1) HM_HIP(hipMemcpyAsync(gb1, gb1_.get(), (size_t)hid_ * 4, hipMemcpyDeviceToHost, s_));
    if (gW2) HM_HIP(hipMemcpyAsync(gW2, gW2_.get(), n2 * 4, hipMemcpyDeviceToHost, s_));
    if (gb2) HM_HIP(hipMemcpyAsync(gb2, gb2_.get(), (size_t)dout_ * 4, hipMemcpyDeviceToHost, s_));
    HM_HIP(hipStreamSynchronize(s_));
}

void ProbeMLP::reserve_batch(int n)
{
    if (n <= cap_) return;
    Hb_.reserve((size_t)n * hid_ * 2);
    Z_.reserve((size_t)n * dout_ * 4);
    G16_.reserve((size_t)n * dout_ * 2);
    dH16_.reserve((size_t)n * hid_ * 2);
    loss_.reserve((size_t)n * 4);
    cap_ = n;
}

void ProbeMLP::forward(const uint16_t *Xb, int n, float *Z)
{
    reserve_batch(n);
    // H = relu(X W1 + b1) (bf16), Z = H W2 + b2 (f32)
    gemm_bf16(Xb, din_, 1, W1b_.as<uint16_t>(), hid_, 1, n, hid_, din_, b1_.as<float>(), true,
              nullptr, 0, Hb_.as<uint16_t>(), hid_, s_);
    gemm_bf16(Hb_.as<uint16_t>(), hid_, 1, W2b_.as<uint16_t>(), dout_, 1, n, dout_, hid_,
              b2_.as<float>(), false, Z ? Z : Z_.as<float>(), dout_, nullptr, 0, s_);
}

void ProbeMLP::loss_grad(const uint16_t *Xb, const float *T, int n)
{
    reserve_batch(n);
    forward(Xb, n, Z_.as<float>());
    // G = (softmax(Z) - T) / n
    softmax_xent(Z_.as<float>(), T, n, dout_, 1.0f / (float)n, G16_.as<uint16_t>(), nullptr,
                 loss_.as<float>(), nullptr, s_);
    // gW2 = H^T G ; gb2 = colsum(G)
    gemm_bf16(Hb_.as<uint16_t>(), 1, hid_, G16_.as<uint16_t>(), dout_, 1, hid_, dout_, n, nullptr,
              false, gW2_.as<float>(), dout_, nullptr, 0, s_);
    colsum_bf16(G16_.as<uint16_t>(), n, dout_, gb2_.as<float>(), s_);
    // dH = (G W2^T) * (H > 0)
    gemm_bf16(G16_.as<uint16_t>(), dout_, 1, W2b_.as<uint16_t>(), 1, dout_, n, hid_, dout_,
              nullptr, false, nullptr, 0, dH16_.as<uint16_t>(), hid_, s_);
    relu_mask(dH16_.as<uint16_t>(), Hb_.as<uint16_t>(), (long)n * hid_, s_);
    // gW1 = X^T dH ; gb1 = colsum(dH)
    gemm_bf16(Xb, 1, din_, dH16_.as<uint16_t>(), hid_, 1, din_, hid_, n, nullptr, false,
              gW1_.as<float>(), hid_, nullptr, 0, s_);
    colsum_bf16(dH16_.as<uint16_t>(), n, hid_, gb1_.as<float>(), s_);
}

void ProbeMLP::adam(float lr, float wd)
{
    ++step_;
    inf_valid_ = false;
    const long n1 = (long)din_ * hid_, n2 = (long)hid_ * dout_;
    adam_step(W1_.as<float>(), W1b_.as<uint16_t>(), gW1_.as<float>(), mW1_.as<float>(),
              vW1_.as<float>(), n1, lr, step_, wd, s_);
    adam_step(b1_.as<float>(), nullptr, gb1_.as<float>(), mb1_.as<float>(), vb1_.as<float>(),
              hid_, lr, step_, 0.f, s_);
    adam_step(W2_.as<float>(), W2b_.as<uint16_t>(), gW2_.as<float>(), mW2_.as<float>(),
              vW2_.as<float>(), n2, lr, step_, wd, s_);
    adam_step(b2_.as<float>(), nullptr, gb2_.as<float>(), mb2_.as<float>(), vb2_.as<float>(),
              dout_, lr, step_, 0.f, s_);
}

}  // namespace hm
