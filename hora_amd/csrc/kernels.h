// kernels.h -- launch wrappers of the gfx950 kernels (host-callable).
#pragma once

#include "hm_common.h"

namespace hm {

enum AccumKind { ACC_EUCLID = 0, ACC_MANHATTAN = 1, ACC_DOT = 2 };
enum EpilogueKind { EPI_NONE = 0, EPI_COSINE = 1, EPI_ANGULAR = 2 };

// Row flags written by the selection kernels.
constexpr uint32_t ROW_TIE = 1u;   // equal distances among the first min(k+1, n)
constexpr uint32_t ROW_NAN = 2u;   // a NaN distance in the row

// ---- exact_dist.hip
// D[q][j] = metric(Q[q], B[j]) in the reference's summation order.
// Q, B rows are zero padded to dpad (a multiple of 16) floats.
void launch_exact_tile(int metric, const float *Q, uint32_t ldq, uint32_t nq, const float *B,
                       uint32_t ldb, uint32_t n, uint32_t dpad, const float *qnorm,
                       const float *bnorm, float *D, uint64_t ldd, hipStream_t s);
// out[r] = sum_t X[r][t]^2 (left fold), t < d.
void launch_self_dot(const float *X, uint32_t ldx, uint32_t rows, uint32_t d, float *out,
                     hipStream_t s);

// ---- select.hip
// For each of `rows` rows of D (n values, row stride ldd): the K smallest by
// (distance, index), ascending, into cand_d/cand_i (row stride ldc), and
// ROW_TIE / ROW_NAN into flags[row].  K <= SELECT_MAX_K.
constexpr uint32_t SELECT_MAX_K = 8192;
void launch_row_select(const float *D, uint64_t ldd, uint32_t rows, uint32_t n, uint32_t K,
                       float *cand_d, uint32_t *cand_i, uint32_t ldc, uint32_t *flags,
                       hipStream_t s);

// Final result of node_search_k per row: rows without ROW_TIE copy the
// first min(k, n) candidates; rows with ROW_TIE replay the reference's
// std-BinaryHeap push/pop stream over the whole distance row.  A NaN row
// sets bit 0 of *err (when k >= 1 and n >= 2).  Indices are offset by
// idx_base.
void launch_bf_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                        const uint32_t *flags, const float *D, uint64_t ldd, uint32_t rows,
                        uint32_t n, uint32_t k, uint64_t *out_idx, float *out_dist,
                        uint32_t *out_count, uint32_t *err, float *heap_d, uint32_t *heap_i,
                        hipStream_t s);

// Non-finite check of a device matrix (rows x d, stride ld): sets *err bit 1.
void launch_check_finite(const float *X, uint64_t ld, uint64_t rows, uint32_t d, uint32_t *err,
                         hipStream_t s);

// dst[r][0..d) = src[r][0..d), dst[r][d..dpad) = 0.
void launch_pad_rows(const float *src, uint64_t lds, float *dst, uint64_t ldd, uint64_t rows,
                     uint32_t d, uint32_t dpad, hipStream_t s);

}  // namespace hm
