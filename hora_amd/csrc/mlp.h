// mlp.h -- probe-predictor network kernels (mlp.hip) and the host model.
#pragma once

#include "hm_common.h"

namespace hm {

// C(m,n) = sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]
// (bf16); epilogue + bias[n], optional ReLU, f32 and/or bf16 stores.
void gemm_bf16(const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn,
               int M, int N, int K, const float *bias, bool relu, float *C32, long ldc32,
               uint16_t *C16, long ldc16, hipStream_t s);
// C(m,n) = sum_k A[m][k] B[n][k] (both k-contiguous bf16; K % 32 == 0,
// lda and ldb multiples of 8, 16-byte aligned); epilogue + bias[n],
// optional ReLU, optional negation, f32 and/or bf16 stores.
void gemm_nt_bf16(const uint16_t *A, long lda, const uint16_t *B, long ldb, int M, int N, int K,
                  const float *bias, bool relu, bool negate, float *C32, long ldc32,
                  uint16_t *C16, long ldc16, hipStream_t s);
// lse[r] = log sum_j exp(-negz[r][j]).
void row_lse(const float *negz, long ld, int rows, int cols, float *lse, hipStream_t s);
// bf16 rows [rows][dpad] from f32 rows [rows][ld], zero padded past d.
void rows_to_bf16(const float *x, long ld, long rows, int d, uint16_t *y, int dpad, hipStream_t s);
void f32_to_bf16(const float *x, uint16_t *y, long n, hipStream_t s);
// Row softmax; with T: G = (p - T) * scale and loss[row]; P optional.
void softmax_xent(const float *Z, const float *T, int rows, int cols, float scale, uint16_t *G16,
                  float *G32, float *loss, float *P, hipStream_t s);
void colsum_bf16(const uint16_t *G, int rows, int cols, float *out, hipStream_t s);
void relu_mask(uint16_t *dH, const uint16_t *H, long n, hipStream_t s);
void adam_step(float *w, uint16_t *wb, const float *g, float *m, float *v, long n, float lr,
               int step, float wd, hipStream_t s);

// Two-layer network din -> hidden (ReLU) -> dout with FP32 masters, BF16
// copies and Adam state, all in device memory.
class ProbeMLP {
  public:
    ProbeMLP(int din, int hidden, int dout, uint64_t seed, hipStream_t s);
    int din() const { return din_; }
    int hidden() const { return hid_; }
    int dout() const { return dout_; }

    // Loads / stores the FP32 masters (host arrays, row-major W1 [din][hid],
    // W2 [hid][dout]).
    void set_weights(const float *W1, const float *b1, const float *W2, const float *b2);
    void get_weights(float *W1, float *b1, float *W2, float *b2) const;

    // Z[n][dout] logits for bf16 inputs Xb[n][din].
    void forward(const uint16_t *Xb, int n, float *Z);
    // Inference: negated logits negz[n][ldz] and each row's log-partition
    // lse[n] for bf16 inputs Xp[n][din_padded()] (zero padded).
    void scores(const uint16_t *Xp, int n, float *negz, long ldz, float *lse);
    int din_padded() const { return (din_ + 31) / 32 * 32; }
    // Loss and gradients (into the internal gradient buffers) for a batch;
    // T[n][dout] are target distributions.  Returns nothing: the loss per
    // row is left in loss_buf().
    void loss_grad(const uint16_t *Xb, const float *T, int n);
    void adam(float lr, float wd);
    void get_grads(float *gW1, float *gb1, float *gW2, float *gb2) const;
    const float *loss_buf() const { return loss_.as<float>(); }

  private:
    void reserve_batch(int n);
    int din_, hid_, dout_;
    int step_ = 0, cap_ = 0;
    hipStream_t s_;
    DeviceBuffer W1_, b1_, W2_, b2_, W1b_, W2b_;
    DeviceBuffer gW1_, gb1_, gW2_, gb2_;
    DeviceBuffer mW1_, vW1_, mb1_, vb1_, mW2_, vW2_, mb2_, vb2_;
    DeviceBuffer Hb_, Z_, G16_, dH16_, loss_;
    DeviceBuffer W1t_, W2t_, b1p_, Hinf_;  // inference copies, k contiguous, padded
    bool inf_valid_ = false;
    int hinf_cap_ = 0;
};

}  // namespace hm
