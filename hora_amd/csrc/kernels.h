// kernels.h -- launch wrappers of the gfx950 kernels (host-callable).
#pragma once

#include "hm_common.h"

namespace hm {

enum AccumKind { ACC_EUCLID = 0, ACC_MANHATTAN = 1, ACC_DOT = 2 };

// Bitonic sort step: the lower index of compare-exchange pair t for a
// power-of-two stride (2 * stride * (t / stride) + t % stride, by masks).
__device__ __forceinline__ uint32_t bitonic_lo(uint32_t t, uint32_t stride)
{
    return ((t & ~(stride - 1)) << 1) | (t & (stride - 1));
}
enum EpilogueKind { EPI_NONE = 0, EPI_COSINE = 1, EPI_ANGULAR = 2 };

// Row flags written by the selection kernels.
constexpr uint32_t ROW_TIE = 1u;   // equal distances among the first min(k+1, n)
constexpr uint32_t ROW_NAN = 2u;   // a NaN distance in the row
constexpr uint32_t ROW_UNCERT = 4u;  // shortlist could not be certified
constexpr uint32_t ROW_BAD = 8u;     // internal consistency check failed
constexpr uint32_t ROW_TIE_EDGE = 16u;  // the k-th and (k+1)-th distances are equal (IVFPQ)
constexpr uint32_t ROW_SKIP = 0x80000000u;  // bf_finalize: leave the row's outputs alone

// Rows kept on the device for the exact rerun: the plan block of
// RERUN_PLAN_WORDS words holds the rows [0, RERUN_SLOTS) (~0u when unused),
// their finalize flags [RERUN_SLOTS, 2 RERUN_SLOTS), the count, a word that
// is 1 when the host must decide (more tied rows than slots, or rows flagged
// for another reason; the count is then 0), the number of flagged rows and
// the number of uncertified rows.
constexpr uint32_t RERUN_SLOTS = 32, RERUN_PLAN_WORDS = 2 * RERUN_SLOTS + 4;

// ---- exact_dist.hip
// D[q][j] = metric(Q[q], B[j]) in the reference's summation order.
// Q, B rows are zero padded to dpad (a multiple of 16) floats.
void launch_exact_tile(int metric, const float *Q, uint32_t ldq, uint32_t nq, const float *B,
                       uint32_t ldb, uint32_t n, uint32_t dpad, const float *qnorm,
                       const float *bnorm, float *D, uint64_t ldd, hipStream_t s);
// The exact rows of the first *dev_nq (<= RERUN_SLOTS) of RERUN_SLOTS query
// rows, eight per workgroup row of the grid; nothing when *dev_nq is 0.
// Needs dpad <= 2048 and 16-byte aligned B rows (exact_rows_dev_ok).  B8, when given: a u8 copy of B ([n][dpad] bytes)
// holding every value exactly, read instead of B (same bits).
bool exact_rows_dev_ok(uint32_t dpad, uint32_t ldb, const float *B);
void launch_exact_rows_dev(int metric, const float *Q, uint32_t ldq, const float *B, uint32_t ldb,
                           uint32_t n, uint32_t dpad, const float *qnorm, const float *bnorm,
                           float *D, uint64_t ldd, const uint32_t *dev_nq, hipStream_t s,
                           const uint8_t *B8 = nullptr);
// best[q] = min over j of (fkey(metric(Q[q], B[j])) << 32 | j); best must be
// initialised to ~0 by the caller.
void launch_exact_argmin(int metric, const float *Q, uint32_t ldq, uint32_t nq, const float *B,
                         uint32_t ldb, uint32_t n, uint32_t dpad, const float *qnorm,
                         const float *bnorm, unsigned long long *best, hipStream_t s);
// out[r] = sum_t X[r][t]^2 (left fold), t < d.
// dev_rows (optional, device memory): effective row count <= rows, read by
// the kernels, so a captured graph can be replayed for fewer rows.
void launch_self_dot(const float *X, uint32_t ldx, uint32_t rows, uint32_t d, float *out,
                     hipStream_t s, const uint32_t *dev_rows = nullptr);

// out[p] = metric(B[ids[p]], Q[q]) for the pairs p in [off[q], off[q+1])
// (total pairs); ids >= n give NaN and set *err bit 2.
void launch_pair_dist(int metric, const float *Qp, uint32_t dpad, uint32_t nq, const float *B,
                      uint32_t ldb, uint32_t n, const uint32_t *ids, const uint64_t *off,
                      uint64_t total, const float *qnorm, const float *bnorm, float *out,
                      uint32_t *err, hipStream_t s);

// ---- select.hip
// For each of `rows` rows of D (n values, row stride ldd): the K smallest by
// (distance, index), ascending, into cand_d/cand_i (row stride ldc), and
// ROW_TIE / ROW_NAN into flags[row].  K <= SELECT_MAX_K.  row_n (optional,
// device): per-row lengths <= n; a row shorter than K gives row_n[r]
// candidates and fills the rest with (+inf, ~0).
constexpr uint32_t SELECT_MAX_K = 8192;
void launch_row_select(const float *D, uint64_t ldd, uint32_t rows, uint32_t n, uint32_t K,
                       float *cand_d, uint32_t *cand_i, uint32_t ldc, uint32_t *flags,
                       hipStream_t s, const uint32_t *dev_rows = nullptr,
                       const uint32_t *row_n = nullptr);

// out[r] = the m-th smallest value of D's row r (1 <= m <= 32; NaN skipped).
// zero (optional): zero[r] = 0 for every row handled (a list counter the
// next kernel appends to).
void launch_kth_smallest(const float *D, uint64_t ldd, uint32_t rows, uint32_t n, uint32_t m,
                         float *out, hipStream_t s, const uint32_t *dev_rows = nullptr,
                         uint32_t *zero = nullptr);

// Final result of node_search_k per row: rows without ROW_TIE copy the
// first min(k, n) candidates; rows with ROW_TIE replay the reference's
// std-BinaryHeap push/pop stream over the whole distance row.  A NaN row
// sets bit 0 of *err (when k >= 1 and n >= 2).  flags may be null: every
// row is then tied (and free of NaN), and the candidates are not read.
// k is the output stride; row_k (optional) gives row r its own k, read at
// row_k[row_k_map[r]] (or row_k[r]) and capped at the stride.
void launch_bf_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                        const uint32_t *flags, const float *D, uint64_t ldd, uint32_t rows,
                        uint32_t n, uint32_t k, uint64_t *out_idx, float *out_dist,
                        uint32_t *out_count, uint32_t *err, hipStream_t s,
                        const uint32_t *row_k = nullptr, const uint32_t *row_k_map = nullptr);

// Non-finite check of a device matrix (rows x d, stride ld): sets *err bit 1.
void launch_check_finite(const float *X, uint64_t ld, uint64_t rows, uint32_t d, uint32_t *err,
                         hipStream_t s);

// Host rows [n][d] -> device rows of stride dpad (zero padded), and back:
// contiguous copies through a device staging buffer in chunks of at most
// 256 MiB, the padding done on the device (a pitched copy of small rows
// from pageable memory runs row by row).
void upload_rows(const float *h_rows, uint64_t n, uint32_t d, float *d_dst, uint32_t dpad,
                 DeviceBuffer &stage, hipStream_t s);
void download_rows(const float *d_src, uint32_t dpad, uint64_t n, uint32_t d, float *h_out,
                   DeviceBuffer &stage, hipStream_t s);

// dst[r][0..d) = src[r][0..d), dst[r][d..dpad) = 0.
// err (optional): a non-finite source element sets bit 1 of *err, as
// launch_check_finite does.  src_rows (optional): dst row r is source row
// src_rows[r] (rows past *dev_rows are not read).
void launch_pad_rows(const float *src, uint64_t lds, float *dst, uint64_t ldd, uint64_t rows,
                     uint32_t d, uint32_t dpad, hipStream_t s, const uint32_t *dev_rows = nullptr,
                     uint32_t *err = nullptr, const uint32_t *src_rows = nullptr);

// ---- mfma_scores.hip
// Kmat[q][j] = approximate ranking key of base row j for query q (see file
// header); Q, B are bf16 (elem_bytes 2) or fp8 (1) rows of row_bytes bytes
// (a multiple of 64, zero padded).
void launch_score_gemm(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                       uint32_t row_bytes, const float *qscale, const float *bscale,
                       const float *bnorm2, const float *inv_bnorm, int metric, float *Kmat,
                       uint64_t ldk, hipStream_t s, const uint32_t *dev_rows = nullptr,
                       uint32_t group_kib = 0);
// As launch_score_gemm, with each group of four keys of a row (one lane's
// columns) reduced to its minimum: Kmat[q][ceil(n / 128) * 32], ldk at
// least that.  The m-th smallest of these minima is at least the m-th
// smallest key (they are m distinct keys): a threshold estimate from a
// quarter of the data.
void launch_score_min4(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                       uint32_t row_bytes, const float *qscale, const float *bscale,
                       const float *bnorm2, const float *inv_bnorm, int metric, float *Kmat,
                       uint64_t ldk, hipStream_t s, const uint32_t *dev_rows = nullptr,
                       uint32_t group_kib = 0);
// Filter form: keys below tau[q * tau_ld] are appended to list[q][0..cap)
// as fkey(key) << 32 | row, counted in cnt[q] (zeroed by the caller; the
// count may exceed cap, the entries past cap are dropped).  A persistent
// scan writes the rows with keys below their threshold to per-wave record
// regions in `scratch` (at least score_filter_scratch_bytes(nq, cap) bytes
// for a full-size region; a smaller one only lets more lists overflow), and
// a gather appends their keys to the lists.
size_t score_filter_scratch_bytes(uint32_t nq, uint32_t cap);
void launch_score_filter(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                         uint32_t row_bytes, const float *qscale, const float *bscale,
                         const float *bnorm2, const float *inv_bnorm, int metric,
                         const float *tau, uint32_t tau_ld, uint64_t *list, uint32_t cap,
                         uint32_t *cnt, void *scratch, size_t scratch_bytes, hipStream_t s,
                         const uint32_t *dev_rows = nullptr, uint32_t group_kib = 0);
// Score tile order: base tiles are taken in groups of about group_kib KiB of
// base rows per XCD's L2, the query tiles inside a group (0: 2,048 KiB).
// Strided sample of the base: dst row i = src row floor(i * n / ns).
void launch_sample_rows(const void *src, uint32_t row_bytes, uint32_t n, uint32_t ns, void *dst,
                        hipStream_t s);
void launch_sample_f32(const float *src, uint32_t n, uint32_t ns, float *dst, hipStream_t s);
void launch_to_bf16(const float *src, uint64_t lds, uint64_t rows, uint32_t d, uint32_t dk,
                    uint16_t *dst, uint32_t *inexact, hipStream_t s,
                    const uint32_t *dev_rows = nullptr);
void launch_to_fp8(const float *src, uint64_t lds, uint32_t rows, uint32_t d, uint32_t dk,
                   uint8_t *dst, float *dequant, hipStream_t s, const uint32_t *dev_rows = nullptr);
// dst[r][j] = (uint8_t)src[r][j] over [rows][dpad]; *inexact |= 1 when some
// value is not bit-identical to the f32 of its u8 (non-integers, values
// outside 0..255, -0.0): with *inexact == 0 the copy holds the same values.
// The query rows of a shortlist form in one pass (one wave per row): the
// non-finite check (err |= 2), the left-fold norms, the FP8 (elem 1, with
// dequant) or BF16 (elem 2, with inexact) rows of dk columns and, when pad
// is given, the rows zero-padded to dpad floats.  query_prep_ok tells
// whether the shape fits (d % 4 == 0, d <= 1024, 16-byte aligned rows).
bool query_prep_ok(const float *src, uint64_t lds, uint32_t d, const float *pad, uint32_t dpad);
void launch_query_prep(int elem, const float *src, uint64_t lds, uint32_t rows, uint32_t d,
                       uint32_t dk, void *dst, float *dequant, uint32_t *inexact, float *norm2,
                       float *pad, uint32_t dpad, uint32_t *err, hipStream_t s,
                       const uint32_t *dev_rows = nullptr);
void launch_to_u8_exact(const float *src, uint32_t dpad, uint64_t rows, uint8_t *dst,
                        uint32_t *inexact, hipStream_t s);
// inv[r] = 1/sqrt(norm2[r]); stats[0] |= max norm (float bits, atomicMax),
// stats[2] |= 1 when a norm is zero.  stats[0..2] zeroed by the caller.
void launch_norm_stats(const float *norm2, uint32_t rows, float *inv, uint32_t *stats,
                       hipStream_t s);

// ---- rerank.hip
void launch_rerank(int metric, const float *Qp, uint32_t dpad, const float *Bp,
                   const uint32_t *cand_i, const float *cand_key, uint32_t ldc, uint32_t s,
                   const float *qnorm2, const float *bnorm2, uint32_t rows, uint32_t n,
                   uint32_t k, float bmax, float gamma, float c_in_exact, float c_in_inexact,
                   const uint32_t *q_inexact, int certify, uint64_t *out_idx, float *out_dist,
                   uint32_t *out_count, uint32_t *flags, hipStream_t st,
                   const uint32_t *dev_rows = nullptr, const uint32_t *kq = nullptr,
                   const uint64_t *out_off = nullptr, const uint32_t *s_q = nullptr,
                   const float *key_s_q = nullptr, const uint32_t *pre_flags = nullptr,
                   int by_id = 0, const uint8_t *Bp8 = nullptr);
// Candidate lists of the filter form -> the shortlist format of rerank:
// per query the first min(cnt, S) entries by (key, row) into cand_i /
// cand_key (stride ldc), s_q = that count, key_s_q = a lower bound of the
// key of every row left out (the smallest list key left out, else tau),
// flags = ROW_UNCERT when the list overflowed.  kq (ragged batches, when the
// depth is the automatic one): query r keeps min(S, max(64, 4(kq[r]+1))).
void launch_list_prepare(const uint64_t *list, uint32_t cap, const uint32_t *cnt,
                         const float *tau, uint32_t tau_ld, uint32_t rows, uint32_t S,
                         uint32_t *cand_i, float *cand_key, uint32_t ldc, uint32_t *s_q,
                         float *key_s_q, uint32_t *flags, hipStream_t st,
                         const uint32_t *dev_rows = nullptr, const uint32_t *kq = nullptr);
void launch_scatter_ragged(const uint64_t *si, const float *sd, const uint32_t *sc,
                           const uint32_t *rows, uint32_t nrows, uint32_t kmax, const uint32_t *kq,
                           const uint64_t *out_off, uint64_t *di, float *dd, uint32_t *dc,
                           hipStream_t s);
// dst[i] = (uint32_t)src[i] (row indices below 2^32).
void launch_u64_to_u32(const uint64_t *src, uint32_t *dst, uint64_t n, hipStream_t s);
// One ragged-batch call, in device memory: the captured graph's kernels
// read the call's query and output pointers from it, so one graph serves
// every call with the same shape.
struct RaggedCall {
    const float *dq;  // [Q][dim] queries
    uint64_t *d_idx;  // packed outputs, k_b per query
    float *d_dist;
    uint32_t *d_count;  // [Q]
    uint64_t Q;
};
// Chunk c (rows [c QC, min((c+1) QC, Q))): *rows = its row count; the rows
// padded to dpad into qpad (non-finite values set *err bit 1), ck[r] =
// kq[c QC + r], coff[r] = off[c QC + r] - off[c QC].
void launch_ragged_chunk_begin(const RaggedCall *call, uint32_t c, uint32_t QC, uint32_t dim,
                               uint32_t dpad, const uint32_t *kq, const uint64_t *off,
                               uint32_t *rows, uint32_t *ck, uint64_t *coff, float *qpad,
                               uint32_t *err, hipStream_t s);
// The chunk's packed results (oi / od / oc, rf row flags) into the call's
// outputs; flags_all[c QC + r] = rf[r] (0 past the chunk's rows).
void launch_ragged_chunk_end(const RaggedCall *call, uint32_t c, uint32_t QC,
                             const uint32_t *rows, const uint64_t *off, const uint64_t *oi,
                             const float *od, const uint32_t *oc, const uint32_t *rf,
                             uint32_t *flags_all, hipStream_t s);
// The rerun plan's rows (plan[0 .. count)) padded from the call's queries.
void launch_ragged_rerun_pad(const RaggedCall *call, const uint32_t *plan, uint32_t dim,
                             uint32_t dpad, float *dst, hipStream_t s);
// The rerun rows' results ([RERUN_SLOTS][kmax]) into the call's packed
// outputs, query q keeping its first kq[q].
void launch_ragged_rerun_scatter(const RaggedCall *call, const uint32_t *plan, const uint64_t *si,
                                 const float *sd, const uint32_t *sc, uint32_t kmax,
                                 const uint32_t *kq, const uint64_t *off, hipStream_t s);
// Fill kernel (graph-safe replacement of hipMemsetAsync).
void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s);
void launch_rerun_plan(const uint32_t *flags, uint32_t nq, uint32_t *plan, hipStream_t s);
// rows[i] == ~0u: the row is skipped (gather and scatter).
void launch_gather_rows(const float *src, uint64_t lds, const uint32_t *rows, uint32_t nrows,
                        uint32_t d, float *dst, uint64_t ldd, hipStream_t s);
void launch_scatter_results(const uint64_t *si, const float *sd, const uint32_t *sc,
                            const uint32_t *rows, uint32_t nrows, uint32_t k, uint64_t *di,
                            float *dd, uint32_t *dc, hipStream_t s);

// ---- knn_graph.hip (naive_build_knn_graph, knn.rs:14-41)
// D[r][r0 + r] = +inf for the rows of a chunk (the node's own column).
void launch_graph_diag(float *D, uint64_t ldd, uint32_t rows, uint32_t r0, hipStream_t s);
// Graph rows r0.. from the Ksel smallest candidates: rows whose k+1
// nearest others are distinct get those others largest first; the others
// get tied[r] = 1.  NaN rows set *err bit 0.
void launch_graph_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                           uint32_t Ksel, const uint32_t *flags, uint32_t rows, uint32_t r0,
                           uint32_t n, uint32_t k, uint64_t *out_idx, float *out_dist,
                           uint32_t *out_count, uint32_t *tied, uint32_t *err, hipStream_t s);
// C[t] = row rows[t] of D without the node's own column (n - 1 values).
void launch_graph_compact(const float *D, uint64_t ldd, const uint32_t *rows, uint32_t nt,
                          uint32_t r0, uint32_t n, float *C, uint64_t ldc, hipStream_t s);
// Replay output (ascending, compacted columns) -> graph rows, largest first.
void launch_graph_remap(const uint64_t *ti, const float *td, const uint32_t *tc,
                        const uint32_t *rows, uint32_t nt, uint32_t r0, uint32_t k,
                        uint64_t *out_idx, float *out_dist, hipStream_t s);

}  // namespace hm
