// ivfpq.h -- launch wrappers of the IVFPQIndex kernels (ivfpq.hip).
#pragma once

#include "kernels.h"

namespace hm {

// Working memory of the fp64 centre steps (grown on demand).
struct SegScratch {
    DeviceBuffer part, nch, cstart, tmp;
};
// kmeans.rs:50-86 for the coarse quantizer: column means of X (f32 left fold
// in row order, or fp64 sums over fixed row chunks) and the +-1 initial
// centres, written [nc][ldc].
void launch_coarse_init(bool fp64, const float *X, uint32_t ldx, uint64_t n, uint32_t d,
                        uint32_t nc, float *C, uint32_t ldc, SegScratch &ws, hipStream_t s);
void launch_unpack_assign(const unsigned long long *best, uint64_t n, uint32_t *assign,
                          hipStream_t s);
// Stable grouping of rows by list: ids = row indices sorted by (assign, row),
// sorted_keys = assign in that order, off[0..nlist] = list offsets.
void build_lists(const uint32_t *assign, uint64_t n, uint32_t nlist, uint32_t *sorted_keys,
                 uint32_t *ids, uint32_t *off, DeviceBuffer &iota, DeviceBuffer &tmp,
                 hipStream_t s);
// kmeans.rs:88-125: centre = mean of members, empty -> 0 (f32: point-order
// fold; fp64: sums over fixed chunks of SEG_ROWS members).  n = rows in ids.
void launch_update(bool fp64, const float *X, uint32_t ldx, const uint32_t *off,
                   const uint32_t *ids, uint32_t nlist, uint32_t d, float *C, uint32_t ldc,
                   uint64_t n, SegScratch &ws, hipStream_t s);
// R[r] = X[ids[r]] - C[lists[r]] for rows in list order.
void launch_residuals(const float *X, uint32_t ldx, const uint32_t *ids, const uint32_t *lists,
                      const float *C, uint64_t n, uint32_t dpad, float *R, hipStream_t s);
// Per (list, sub-space) k-means on the residuals: codebooks
// [nlist][M][KS][sdm] and codes [n][M] (rows in list order).  lorder[nlist]
// is the order in which workgroups take the lists (longest first);
// asg_scratch [M][n] u16 and ord_scratch [M][n] u32 hold the assignments and
// the member order of lists too long to keep in LDS.
void launch_pq_train(bool fp64, const float *R, uint32_t ldr, const uint32_t *off,
                     const uint32_t *lorder, uint32_t nlist, uint32_t M, const uint32_t *sub_b,
                     const uint32_t *sub_e, uint32_t KS, uint32_t sd_max, uint32_t sdm,
                     uint32_t epochs, uint64_t seed, uint32_t max_list, uint64_t n,
                     uint16_t *asg_scratch, uint32_t *ord_scratch, float *codebooks,
                     uint8_t *codes, hipStream_t s);
// Candidate slots per (query, probe) workgroup.
uint32_t ivf_scan_pool(uint32_t k1);
// out[(q*nprobe + p)*k1 ..] = the k1 smallest (fkey(adc) << 32 | id) of
// the p-th probed list of query q.
void launch_ivf_scan(const float *Qp, uint32_t dpad, const float *C, const uint32_t *probes,
                     uint32_t ldp, uint32_t nprobe, uint32_t rows, const uint32_t *off,
                     const uint32_t *ids, const uint8_t *codes, const float *codebooks,
                     const uint32_t *sub_b, const uint32_t *sub_e, uint32_t M, uint32_t KS,
                     uint32_t sdm, uint32_t k1, const uint32_t *nprobe_q,
                     unsigned long long *out, hipStream_t s);
// List-major form of launch_ivf_scan (same output): (query, probe) pairs
// bucketed by list, one workgroup per list with its codebook in LDS.
// Returns the candidate pool, or 0 when the codebook does not fit in LDS.
uint32_t ivf_scan_lists_pool(uint32_t M, uint32_t KS, uint32_t sdm, uint32_t dpad, uint32_t k1);
// Scratch: pair_cnt, pair_fill [nlist]; pair_off, unit_off [nlist + 1];
// pairs [rows * nprobe].
void launch_ivf_scan_lists(const float *Qp, uint32_t dpad, const float *C,
                           const uint32_t *probes, uint32_t ldp, uint32_t nprobe, uint32_t rows,
                           uint32_t nlist, const uint32_t *off, const uint32_t *ids,
                           const uint8_t *codes, const float *codebooks, const uint32_t *sub_b,
                           const uint32_t *sub_e, uint32_t M, uint32_t KS, uint32_t sdm,
                           uint32_t k1, const uint32_t *nprobe_q, uint32_t *pair_cnt,
                           uint32_t *pair_off, uint32_t *pair_fill, uint32_t *unit_off,
                           uint32_t *pairs, unsigned long long *out, hipStream_t s);
// Row form of the scan: D[q][seg[q*nprobe + p] + i] = ADC of item i of the
// p-th probed list of query q (pq_idx.rs:151-194 arithmetic), row_n[q] =
// the row's length.  Needs a codebook that fits in registers:
// ivf_adc_rows_variant != 0.  D has stride ldd >= nprobe * (longest list).
int ivf_adc_rows_variant(uint32_t M, uint32_t KS, uint32_t sdm);
void launch_ivf_adc_rows(const float *Qp, uint32_t dpad, uint32_t dim, const float *C,
                         const uint32_t *probes, uint32_t ldp, uint32_t nprobe, uint32_t rows,
                         uint32_t nlist,
                         const uint32_t *off, const uint8_t *codes, const float *codebooks,
                         const uint32_t *sub_b, const uint32_t *sub_e, uint32_t M, uint32_t KS,
                         uint32_t sdm, const uint32_t *nprobe_q, uint32_t *pair_cnt,
                         uint32_t *pair_off, uint32_t *pair_fill, uint32_t *unit_off,
                         uint32_t *pairs, uint32_t *seg, uint32_t *row_n, float *D, uint64_t ldd,
                         hipStream_t s);
// The row select's candidates (columns) -> ids, counts and ROW_TIE flags,
// as launch_ivf_merge.
// Narrows the coarse select's ROW_TIE to ties that can change the result
// (ties among empty lists cannot) and marks edge ties ROW_TIE_EDGE.
void launch_coarse_ties(const float *D, uint64_t ldd, uint32_t nlist, const float *pd,
                        const uint32_t *pi, uint32_t Kc, uint32_t nprobe, const uint32_t *off,
                        uint32_t *flags, uint32_t rows, hipStream_t s);
// Row form: makes the selected first k the k smallest by (distance, row
// index) when the k-th and (k+1)-th are equal (the select broke ties by
// column); cand_i holds columns, rows of row_n[q] values in Dr.
void launch_ivf_edge_fixup(const float *Dr, uint64_t ldr, const uint32_t *row_n,
                           const float *cand_d, uint32_t *cand_i, uint32_t ldc, uint32_t k,
                           const uint32_t *probes, uint32_t ldp, uint32_t nprobe,
                           const uint32_t *nprobe_q, const uint32_t *seg, const uint32_t *off,
                           const uint32_t *ids, uint32_t rows, hipStream_t s);
void launch_ivf_rows_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                              const uint32_t *sel_flags, const uint32_t *coarse_flags,
                              const uint32_t *probes, uint32_t ldp, uint32_t nprobe,
                              const uint32_t *nprobe_q, const uint32_t *seg,
                              const uint32_t *row_n, const uint32_t *off, const uint32_t *ids,
                              uint32_t rows, uint32_t k, uint64_t *out_idx, float *out_dist,
                              uint32_t *out_count, uint32_t *flags, hipStream_t s);
void launch_ivf_merge(const unsigned long long *cand, uint32_t per_query, uint32_t rows,
                      uint32_t k, const uint32_t *coarse_flags, uint64_t *out_idx,
                      float *out_dist, uint32_t *out_count, uint32_t *flags, hipStream_t s);
void launch_ivf_replay(const uint32_t *rows, uint32_t nrows, const float *Dc, uint32_t ldd,
                       uint32_t nlist, uint32_t nprobe, const float *Qp, uint32_t dpad,
                       const float *C, const uint32_t *off, const uint32_t *ids,
                       const uint8_t *codes, const float *codebooks, const uint32_t *sub_b,
                       const uint32_t *sub_e, uint32_t M, uint32_t KS, uint32_t sdm, uint32_t k,
                       uint32_t lmax, float *scratch_f, uint32_t *scratch_u, size_t per_thread,
                       const uint32_t *probes, uint32_t ldp, const uint32_t *nprobe_q,
                       uint64_t *out_idx, float *out_dist, uint32_t *out_count, hipStream_t s);
// The same replay with the heaps in LDS and each tied query's probed lists
// simulated in parallel: order [nrows][nprobe] and np [nrows] hold the probe
// order, seq [nrows][nprobe][k] / seq_cnt [nrows][nprobe] each list's
// drained survivors.  Returns false (launching nothing) when the lists are
// too long for LDS.  probes [rows][ldp]: the probe order of every row (the
// learned policy, with nprobe_q), or, with coarse_flags, the coarse
// select's sorted candidates, used for the rows without ROW_TIE there (the
// others replay the coarse heap).
bool launch_ivf_replay_lds(const uint32_t *rows, uint32_t nrows, const float *Dc, uint32_t ldd,
                           uint32_t nlist, uint32_t nprobe, const float *Qp, uint32_t dpad,
                           const float *C, const uint32_t *off, const uint32_t *ids,
                           const uint8_t *codes, const float *codebooks, const uint32_t *sub_b,
                           const uint32_t *sub_e, uint32_t M, uint32_t KS, uint32_t sdm,
                           uint32_t k, uint32_t lmax, const uint32_t *probes, uint32_t ldp,
                           const uint32_t *nprobe_q, uint32_t *order, uint32_t *np,
                           uint2 *seq, uint32_t *seq_cnt, uint64_t *out_idx, float *out_dist,
                           uint32_t *out_count, hipStream_t s,
                           const uint32_t *coarse_flags = nullptr);
// Learned-probe helpers.
void launch_negate(const float *x, float *y, uint64_t n, hipStream_t s);
void launch_probe_count(const float *negp, uint32_t ld, uint32_t rows, uint32_t pmax, float mass,
                        uint32_t pmin, uint32_t *nprobe_q, hipStream_t s, const float *lse = nullptr);
void launch_dense_targets(const uint32_t *lab, uint32_t k, uint32_t rows, uint32_t nlist,
                          float *T, hipStream_t s);
void launch_ids_to_lists(const uint64_t *ids, uint64_t n, const uint32_t *assign, uint64_t n_items,
                         uint32_t *out, hipStream_t s);
// kmeans.rs:150-193 on the device: counts from the list offsets, the
// seeded walk for each empty centre in order, the 1 +- eps copy; *rng (device)
// carries the generator between epochs.  cnt_ws [nc] u32, pk_ws [nc] f64.
void launch_coarse_split(float *C, uint32_t ldc, uint32_t nc, uint32_t dim, const uint32_t *off,
                         uint64_t n, uint64_t *rng, uint32_t *cnt_ws, double *pk_ws, hipStream_t s);
uint64_t host_stream_seed(uint64_t seed, uint64_t instance);

}  // namespace hm
