// ivfpq_index.hip -- IVFPQIndex on the device (pq_idx.rs:267-523).
//
// HBM layout after build:
//   base      [cap][dpad] f32, rows in insertion order (as BruteForce)
//   centers   [nlist][dpad] f32 coarse centroids
//   off/ids   CSR inverted lists: ids[off[l]..off[l+1]) ascending row ids
//             of list l (the order _ivflist[l] is filled in, pq_idx.rs:372-375)
//   codebooks [nlist][M][KS][sdm] f32 per-list residual codebooks
//   codes     [n][M] u8, row r belongs to ids[r] (list order)
// Build runs the reference's k-means (init, epochs of assign/update/split,
// final assign) for the coarse quantizer on the device (split on the host,
// it is O(nlist*d)), then one workgroup per (list, sub-space) trains the
// residual PQ.  Search: exact coarse distances, top-nprobe (radix select),
// one workgroup per (query, probe) builds the LUT and scans the list, a
// merge keeps the k smallest; tied queries replay the reference's heaps.
#include <algorithm>
#include <cstring>
#include <vector>

#include "index.h"
#include "ivfpq.h"
#include "mlp.h"

namespace hm {

namespace {

class IVFPQ final : public Index {
  public:
    IVFPQ(uint64_t dim, const hm_params &p) : Index(HM_INDEX_IVFPQ, dim), p_(p)
    {
        if (dim == 0 || dim > 4096) fail(HM_ERR_INVALID_ARG, "dimension out of range (1..4096)");
        if (p.n_sub == 0 || p.n_sub > 256 || dim / p.n_sub == 0)
            fail(HM_ERR_INVALID_ARG, "n_sub must be in 1..min(256, dimension)");
        if (p.sub_bits == 0 || p.sub_bits > 32) fail(HM_ERR_INVALID_ARG, "sub_bits out of range");
        if (p.sub_bits > 8)
            fail(HM_ERR_NOT_IMPLEMENTED, "sub_bits above 8 (codes are stored as u8)");
        if (p.n_kmeans_center == 0) fail(HM_ERR_INVALID_ARG, "n_kmeans_center must be >= 1");
        dpad_ = (uint32_t)round_up(dim, 16);
        M_ = p.n_sub;
        KS_ = 1u << p.sub_bits;
        nlist_ = p.n_kmeans_center;
        nprobe_ = p.search_n_center;
        // pq_idx.rs:75-88
        const uint32_t sd = (uint32_t)(dim / M_), r = (uint32_t)(dim % sd);
        sub_b_h_.resize(M_);
        sub_e_h_.resize(M_);
        for (uint32_t i = 0; i < M_; ++i) {
            if (i < r) {
                sub_b_h_[i] = i * (sd + 1);
                sub_e_h_[i] = (i + 1) * (sd + 1);
            } else {
                sub_b_h_[i] = r * (sd + 1) + (i - r) * sd;
                sub_e_h_[i] = r * (sd + 1) + (i + 1 - r) * sd;
            }
        }
        sd_max_ = sd + (r ? 1 : 0);
        sdm_ = sd_max_;
        HM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        sub_b_.reserve(M_ * 4);
        sub_e_.reserve(M_ * 4);
        HM_HIP(hipMemcpyAsync(sub_b_.get(), sub_b_h_.data(), M_ * 4, hipMemcpyHostToDevice, stream_));
        HM_HIP(hipMemcpyAsync(sub_e_.get(), sub_e_h_.data(), M_ * 4, hipMemcpyHostToDevice, stream_));
        HM_HIP(hipStreamSynchronize(stream_));
    }
    ~IVFPQ() override
    {
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    void add(const float *rows, uint64_t n_rows) override
    {
        if (n_rows == 0) return;
        if (!rows) fail(HM_ERR_INVALID_ARG, "rows is null");
        check_finite_host(rows, n_rows * dim_);
        if (n_ + n_rows > 0xFFFFFFF0ull) fail(HM_ERR_CAPACITY, "more than 2^32-16 rows");
        const uint64_t need = n_ + n_rows;
        if (need > cap_) {
            uint64_t nc = std::max<uint64_t>(need, cap_ + cap_ / 2);
            nc = std::max<uint64_t>(nc, 1024);
            base_.grow_keep(nc * dpad_ * 4, n_ * dpad_ * 4, stream_);
            cap_ = nc;
        }
        DeviceBuffer stage;
        upload_rows(rows, n_rows, (uint32_t)dim_, base_.as<float>() + n_ * dpad_, dpad_, stage,
                    stream_);
        HM_HIP(hipStreamSynchronize(stream_));
        n_ = need;
    }

    void add_device(const float *d_rows, uint64_t n_rows, uint64_t ld) override
    {
        if (n_rows == 0) return;
        if (n_ + n_rows > 0xFFFFFFF0ull) fail(HM_ERR_CAPACITY, "more than 2^32-16 rows");
        const uint64_t need = n_ + n_rows;
        if (need > cap_) {
            base_.grow_keep(need * dpad_ * 4, n_ * dpad_ * 4, stream_);
            cap_ = need;
        }
        launch_pad_rows(d_rows, ld, base_.as<float>() + n_ * dpad_, dpad_, n_rows, (uint32_t)dim_,
                        dpad_, stream_);
        HM_HIP(hipStreamSynchronize(stream_));
        n_ = need;
    }

    void build(int32_t metric) override
    {
        if (metric < HM_METRIC_UNKNOWN || metric > HM_METRIC_ANGULAR)
            fail(HM_ERR_INVALID_ARG, "metric out of range");
        if (metric == HM_METRIC_UNKNOWN) fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        metric_ = metric;
        train();
    }

    bool built() const override { return true; }  // pq_idx.rs:454-456
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

    // Trained state from arrays (deserialization); the lists must cover
    // rows 0..n-1 of this index exactly once.
    void import_trained(int32_t metric, const float *centers, const uint64_t *offsets,
                        const uint64_t *ids, const float *codebooks, const uint8_t *codes)
    {
        if (metric <= HM_METRIC_UNKNOWN || metric > HM_METRIC_ANGULAR)
            fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        const uint64_t n = offsets[nlist_];
        if (n != n_) fail(HM_ERR_INVALID_ARG, "inverted lists do not cover the rows");
        if (n > 0 && (!ids || !codes)) fail(HM_ERR_INVALID_ARG, "list ids or codes are null");
        std::vector<uint32_t> off(nlist_ + 1), id(n), asg(n, 0xFFFFFFFFu), lst(n);
        uint32_t maxl = 0;
        for (uint32_t l = 0; l < nlist_; ++l) {
            if (offsets[l + 1] < offsets[l]) fail(HM_ERR_INVALID_ARG, "list offsets decrease");
            maxl = std::max<uint32_t>(maxl, (uint32_t)(offsets[l + 1] - offsets[l]));
            for (uint64_t r = offsets[l]; r < offsets[l + 1]; ++r) {
                if (ids[r] >= n || asg[ids[r]] != 0xFFFFFFFFu)
                    fail(HM_ERR_INVALID_ARG, "inverted list ids are not a permutation");
                asg[ids[r]] = l;
                id[r] = (uint32_t)ids[r];
                lst[r] = l;
            }
        }
        for (uint32_t l = 0; l <= nlist_; ++l) off[l] = (uint32_t)offsets[l];
        for (uint64_t i = 0; i < n * M_; ++i)
            if (codes[i] >= KS_) fail(HM_ERR_INVALID_ARG, "code out of range");
        hipStream_t s = stream_;
        centers_.reserve((size_t)nlist_ * dpad_ * 4);
        cnorm_.reserve((size_t)nlist_ * 4);
        assign_.reserve(n * 4 + 4);
        lists_.reserve(n * 4 + 4);
        ids_.reserve(n * 4 + 4);
        off_.reserve((nlist_ + 1) * 4);
        codebooks_.reserve((size_t)nlist_ * M_ * KS_ * sdm_ * 4);
        codes_.reserve(n * M_ + 4);
        HM_HIP(hipMemsetAsync(centers_.get(), 0, (size_t)nlist_ * dpad_ * 4, s));
        HM_HIP(hipMemcpy2DAsync(centers_.get(), dpad_ * 4, centers, dim_ * 4, dim_ * 4, nlist_,
                                hipMemcpyHostToDevice, s));
        HM_HIP(hipMemcpyAsync(off_.get(), off.data(), (nlist_ + 1) * 4, hipMemcpyHostToDevice, s));
        if (n) {
            HM_HIP(hipMemcpyAsync(ids_.get(), id.data(), n * 4, hipMemcpyHostToDevice, s));
            HM_HIP(hipMemcpyAsync(assign_.get(), asg.data(), n * 4, hipMemcpyHostToDevice, s));
            HM_HIP(hipMemcpyAsync(lists_.get(), lst.data(), n * 4, hipMemcpyHostToDevice, s));
            HM_HIP(hipMemcpyAsync(codes_.get(), codes, n * M_, hipMemcpyHostToDevice, s));
        }
        HM_HIP(hipMemcpyAsync(codebooks_.get(), codebooks, (size_t)nlist_ * M_ * KS_ * sdm_ * 4,
                              hipMemcpyHostToDevice, s));
        metric_ = metric;
        if (metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR)
            launch_self_dot(centers_.as<float>(), dpad_, nlist_, (uint32_t)dim_, cnorm_.as<float>(),
                            s);
        HM_HIP(hipStreamSynchronize(s));
        max_list_ = maxl;
        trained_ = true;
        n_trained_ = n;
    }
    const char *name() const override { return "IVFPQIndex"; }
    uint64_t nodes_size() const override { return n_; }

    void clear() override
    {
        base_.release();
        n_ = cap_ = 0;
        rr_bnorm_n_ = 0;
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

    void set_nprobe(uint32_t p) { nprobe_ = p; }

    void info(uint32_t *out) const
    {
        out[0] = nlist_;
        out[1] = M_;
        out[2] = KS_;
        out[3] = sdm_;
        out[4] = (uint32_t)n_trained_;
        out[5] = trained_ ? 1 : 0;
        out[6] = nprobe_;
        out[7] = dpad_;
    }

    void export_trained(float *centers, uint64_t *offsets, uint64_t *ids, float *codebooks,
                        uint8_t *codes) const
    {
        if (!trained_) fail(HM_ERR_NOT_BUILT, "index not built");
        if (centers)
            HM_HIP(hipMemcpy2DAsync(centers, dim_ * 4, centers_.get(), dpad_ * 4, dim_ * 4, nlist_,
                                    hipMemcpyDeviceToHost, stream_));
        std::vector<uint32_t> off(nlist_ + 1), id(n_trained_);
        HM_HIP(hipMemcpyAsync(off.data(), off_.get(), (nlist_ + 1) * 4, hipMemcpyDeviceToHost,
                              stream_));
        if (n_trained_)
            HM_HIP(hipMemcpyAsync(id.data(), ids_.get(), n_trained_ * 4, hipMemcpyDeviceToHost,
                                  stream_));
        if (codebooks)
            HM_HIP(hipMemcpyAsync(codebooks, codebooks_.get(),
                                  (size_t)nlist_ * M_ * KS_ * sdm_ * 4, hipMemcpyDeviceToHost,
                                  stream_));
        if (codes && n_trained_)
            HM_HIP(hipMemcpyAsync(codes, codes_.get(), n_trained_ * M_, hipMemcpyDeviceToHost,
                                  stream_));
        HM_HIP(hipStreamSynchronize(stream_));
        if (offsets)
            for (uint32_t i = 0; i <= nlist_; ++i) offsets[i] = off[i];
        if (ids)
            for (uint64_t i = 0; i < n_trained_; ++i) ids[i] = id[i];
    }

    void search_device(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx, float *d_dist,
                       uint32_t *d_count, hipStream_t s) override
    {
        if (!s) s = stream_;
        if (!trained_) fail(HM_ERR_NOT_BUILT, "index not built");
        if (nq == 0) return;
        if (!dq || !d_count || (k && (!d_idx || !d_dist)))
            fail(HM_ERR_INVALID_ARG, "null pointer");
        if (rerank_ == 0 || k == 0 || nprobe_ == 0 || n_trained_ == 0) {
            search_adc(dq, nq, k, d_idx, d_dist, d_count, s);
            return;
        }
        // Exact rerank (no reference counterpart; the reference returns ADC
        // distances, pq_idx.rs:181-191): the nprobe lists nearest by (coarse
        // distance, list), or the probe network's lists; among their items
        // the R = max(rerank, k) smallest ADC distances by (distance, row
        // index); each candidate's exact distance (the reference's metric
        // fold), sorted by (exact distance, row index); the first k with
        // their exact distances are the result.  Every step is a total
        // order, so no heap is replayed.
        const uint32_t R = std::max(rerank_, k);
        if (dpad_ > 4096) fail(HM_ERR_INVALID_ARG, "exact rerank: dimension above 4096");
        rr_i_.reserve(nq * R * 8);
        rr_d_.reserve(nq * R * 4);
        rr_c_.reserve(nq * 4);
        search_adc(dq, nq, R, rr_i_.as<uint64_t>(), rr_d_.as<float>(), rr_c_.as<uint32_t>(), s,
                   true);
        const bool norms = metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR;
        if (norms && rr_bnorm_n_ != n_) {
            rr_bnorm_.reserve(n_ * 4);
            launch_self_dot(base_.as<float>(), dpad_, (uint32_t)n_, (uint32_t)dim_,
                            rr_bnorm_.as<float>(), s);
            rr_bnorm_n_ = n_;
        }
        rr_ci_.reserve(nq * R * 4);
        rr_qpad_.reserve(nq * dpad_ * 4);
        rr_qn_.reserve(nq * 4);
        rr_f_.reserve(nq * 4);
        launch_u64_to_u32(rr_i_.as<uint64_t>(), rr_ci_.as<uint32_t>(), nq * R, s);
        launch_pad_rows(dq, dim_, rr_qpad_.as<float>(), dpad_, nq, (uint32_t)dim_, dpad_, s);
        launch_self_dot(rr_qpad_.as<float>(), dpad_, (uint32_t)nq, (uint32_t)dim_,
                        rr_qn_.as<float>(), s);
        launch_rerank(metric_, rr_qpad_.as<float>(), dpad_, base_.as<float>(),
                      rr_ci_.as<uint32_t>(), rr_d_.as<float>(), R, R, rr_qn_.as<float>(),
                      rr_bnorm_.as<float>(), (uint32_t)nq, (uint32_t)n_, k, 0.0f, 0.0f, 0.0f, 0.0f,
                      nullptr, 0, d_idx, d_dist, d_count, rr_f_.as<uint32_t>(), s, nullptr,
                      nullptr, nullptr, rr_c_.as<uint32_t>(), nullptr, nullptr, true);
        std::vector<uint32_t> fl(nq);
        HM_HIP(hipMemcpyAsync(fl.data(), rr_f_.get(), nq * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        for (uint64_t q = 0; q < nq; ++q) {
            if (fl[q] & ROW_BAD) fail(HM_ERR_HIP, "internal: rerank candidate out of range");
            if (fl[q] & ROW_NAN) fail(HM_ERR_NAN_DISTANCE, "NaN distance (partial_cmp of NaN)");
        }
    }

    // The top-R depth of the exact rerank (0: the reference's ADC result).
    void set_rerank(uint32_t r)
    {
        if (r > 4095) fail(HM_ERR_INVALID_ARG, "rerank depth above 4095");
        rerank_ = r;
    }

    // IVFPQIndex::search_knn_adc + node_search_k (pq_idx.rs:408-477).
    // canonical: no heap replay; the probed lists are the nprobe nearest by
    // (distance, list) and the result is the set of the k smallest ADC
    // distances by (distance, row index), in no particular order among
    // equal distances (the exact rerank re-sorts it).  Otherwise the
    // reference's search (pq_idx.rs:408-477), tied rows replayed.
    void search_adc(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx, float *d_dist,
                    uint32_t *d_count, hipStream_t s, bool canonical = false)
    {
        // pq_idx.rs:424-425: popping more centres than exist unwraps None
        if (nprobe_ > nlist_) fail(HM_ERR_INVALID_ARG, "search_n_center larger than n_kmeans_center");
        if (k == 0 || nprobe_ == 0 || n_trained_ == 0) {
            launch_fill_results(d_idx, d_dist, d_count, nq, k, s);
            HM_HIP(hipStreamSynchronize(s));
            return;
        }
        err_.reserve(4);
        HM_HIP(hipMemsetAsync(err_.get(), 0, 4, s));
        launch_check_finite(dq, dim_, nq, (uint32_t)dim_, err_.as<uint32_t>(), s);
        const uint32_t k1 = k + 1;
        if (2 * k1 > ivf_scan_pool(k1)) fail(HM_ERR_INVALID_ARG, "k too large (max 4095)");
        const bool learned = probe_mode_ == 1 && mlp_;
        const uint32_t np = learned ? pmax_ : nprobe_;
        const uint32_t per_query = np * k1;
        if (per_query > 16384) fail(HM_ERR_INVALID_ARG, "nprobe*(k+1) above 16384");
        const uint32_t Kc = learned ? pmax_ : std::min<uint32_t>(nprobe_ + 1, nlist_);
        const bool norms = metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR;
        const uint64_t ldd = round_up(nlist_, 4);
        uint64_t qc = std::min<uint64_t>(nq, 65535);
        qc = std::min<uint64_t>(qc, std::max<uint64_t>(1, ((size_t)1 << 31) / (ldd * 4)));
        qpad_.reserve(qc * dpad_ * 4);
        qnorm_.reserve(qc * 4);
        D_.reserve(qc * ldd * 4);
        probe_d_.reserve(qc * Kc * 4);
        probe_i_.reserve(qc * Kc * 4);
        cflags_.reserve(qc * 4);
        mflags_.reserve(qc * 4);

        // Scan forms, fastest first: rows (list-major, codebook in registers,
        // ADC distances into D rows, radix row select), list-major with the
        // codebook in LDS and a per-list top-k1, and one workgroup per
        // (query, probe).
        const bool rows_ok =
            ivf_adc_rows_variant(M_, KS_, sdm_) != 0 && k1 <= SELECT_MAX_K && dpad_ < 65536;
        const bool lists_ok = ivf_scan_lists_pool(M_, KS_, sdm_, dpad_, k1) != 0;
        if ((scan_mode_ == 3 && !rows_ok) || (scan_mode_ == 2 && !lists_ok))
            fail(HM_ERR_INVALID_ARG, "the chosen scan form does not fit this index");
        const bool rows_form = (scan_mode_ == 0 || scan_mode_ == 3) && rows_ok;
        const bool list_major = !rows_form && (scan_mode_ == 0 || scan_mode_ == 2) && lists_ok;
        const uint64_t ldr = round_up(std::max<uint64_t>((uint64_t)np * max_list_, 1), 4);
        if (rows_form) {
            // rows of up to nprobe x (longest list) distances: at most 16 GiB
            qc = std::min<uint64_t>(qc, std::max<uint64_t>(1, ((size_t)1 << 34) / (ldr * 4)));
            Dr_.reserve(qc * ldr * 4);
            seg_.reserve(qc * np * 4);
            row_n_.reserve(qc * 4);
            rcand_d_.reserve(qc * k1 * 4);
            rcand_i_.reserve(qc * k1 * 4);
            rflags_.reserve(qc * 4);
        }
        if (!rows_form) cand_.reserve(qc * per_query * 8);
        if (list_major || rows_form) {
            pair_cnt_.reserve((size_t)nlist_ * 4);
            pair_fill_.reserve((size_t)nlist_ * 4);
            pair_off_.reserve(((size_t)nlist_ + 1) * 4);
            unit_off_.reserve(((size_t)nlist_ + 1) * 4);
            pairs_.reserve(qc * np * 4);
        }
        if (learned) {
            xq16_.reserve(qc * (size_t)mlp_->din_padded() * 2);
            Pq_.reserve(qc * 4);  // the rows' log-partition of the logits
            nprobe_q_.reserve(qc * 4);
        }
        std::vector<uint32_t> fl(qc);
        last_replayed_ = 0;
        uint64_t probes_used = 0;
        std::vector<uint32_t> npq(learned ? qc : 0);
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            launch_pad_rows(dq + q0 * dim_, dim_, qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                            dpad_, s);
            if (learned) {
                // scores of every list from the probe network (negated
                // logits, by (score, list) in the select), then the smallest
                // prefix of the best lists holding `mass` of the softmax
                const int dinp = mlp_->din_padded();
                rows_to_bf16(dq + q0 * dim_, (long)dim_, rows, (int)dim_, xq16_.as<uint16_t>(),
                             dinp, s);
                mlp_->scores(xq16_.as<uint16_t>(), rows, D_.as<float>(), (long)ldd,
                             Pq_.as<float>());
                launch_row_select(D_.as<float>(), ldd, rows, nlist_, Kc, probe_d_.as<float>(),
                                  probe_i_.as<uint32_t>(), Kc, cflags_.as<uint32_t>(), s);
                launch_probe_count(probe_d_.as<float>(), Kc, rows, pmax_, mass_, pmin_,
                                   nprobe_q_.as<uint32_t>(), s, Pq_.as<float>());
                HM_HIP(hipMemsetAsync(cflags_.get(), 0, rows * 4, s));
            } else {
                if (norms)
                    launch_self_dot(qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                                    qnorm_.as<float>(), s);
                // pq_idx.rs:416-421: metric(q, c_i) with the build metric
                launch_exact_tile(metric_, qpad_.as<float>(), dpad_, rows, centers_.as<float>(),
                                  dpad_, nlist_, dpad_, qnorm_.as<float>(), cnorm_.as<float>(),
                                  D_.as<float>(), ldd, s);
                launch_row_select(D_.as<float>(), ldd, rows, nlist_, Kc, probe_d_.as<float>(),
                                  probe_i_.as<uint32_t>(), Kc, cflags_.as<uint32_t>(), s);
                launch_coarse_ties(D_.as<float>(), ldd, nlist_, probe_d_.as<float>(),
                                   probe_i_.as<uint32_t>(), Kc, nprobe_, off_.as<uint32_t>(),
                                   cflags_.as<uint32_t>(), rows, s);
            }
            const uint32_t *npq_dev = learned ? nprobe_q_.as<uint32_t>() : nullptr;
            if (rows_form) {
                launch_ivf_adc_rows(qpad_.as<float>(), dpad_, (uint32_t)dim_, centers_.as<float>(),
                                    probe_i_.as<uint32_t>(), Kc, np, rows, nlist_,
                                    off_.as<uint32_t>(), codes_.as<uint8_t>(),
                                    codebooks_.as<float>(), sub_b_.as<uint32_t>(),
                                    sub_e_.as<uint32_t>(), M_, KS_, sdm_, npq_dev,
                                    pair_cnt_.as<uint32_t>(), pair_off_.as<uint32_t>(),
                                    pair_fill_.as<uint32_t>(), unit_off_.as<uint32_t>(),
                                    pairs_.as<uint32_t>(), seg_.as<uint32_t>(),
                                    row_n_.as<uint32_t>(), Dr_.as<float>(), ldr, s);
                launch_row_select(Dr_.as<float>(), ldr, rows, (uint32_t)ldr, k1,
                                  rcand_d_.as<float>(), rcand_i_.as<uint32_t>(), k1,
                                  rflags_.as<uint32_t>(), s, nullptr, row_n_.as<uint32_t>());
                if (canonical)
                    launch_ivf_edge_fixup(Dr_.as<float>(), ldr, row_n_.as<uint32_t>(),
                                          rcand_d_.as<float>(), rcand_i_.as<uint32_t>(), k1, k,
                                          probe_i_.as<uint32_t>(), Kc, np, npq_dev,
                                          seg_.as<uint32_t>(), off_.as<uint32_t>(),
                                          ids_.as<uint32_t>(), rows, s);
                launch_ivf_rows_finalize(rcand_d_.as<float>(), rcand_i_.as<uint32_t>(), k1,
                                         rflags_.as<uint32_t>(), cflags_.as<uint32_t>(),
                                         probe_i_.as<uint32_t>(), Kc, np, npq_dev,
                                         seg_.as<uint32_t>(), row_n_.as<uint32_t>(),
                                         off_.as<uint32_t>(), ids_.as<uint32_t>(), rows, k,
                                         d_idx + q0 * k, d_dist + q0 * k, d_count + q0,
                                         mflags_.as<uint32_t>(), s);
            } else if (list_major) {
                launch_ivf_scan_lists(qpad_.as<float>(), dpad_, centers_.as<float>(),
                                      probe_i_.as<uint32_t>(), Kc, np, rows, nlist_,
                                      off_.as<uint32_t>(), ids_.as<uint32_t>(),
                                      codes_.as<uint8_t>(), codebooks_.as<float>(),
                                      sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_, KS_, sdm_,
                                      k1, npq_dev, pair_cnt_.as<uint32_t>(),
                                      pair_off_.as<uint32_t>(), pair_fill_.as<uint32_t>(),
                                      unit_off_.as<uint32_t>(), pairs_.as<uint32_t>(),
                                      cand_.as<unsigned long long>(), s);
            } else {
                launch_ivf_scan(qpad_.as<float>(), dpad_, centers_.as<float>(),
                                probe_i_.as<uint32_t>(), Kc, np, rows, off_.as<uint32_t>(),
                                ids_.as<uint32_t>(), codes_.as<uint8_t>(), codebooks_.as<float>(),
                                sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_, KS_, sdm_, k1,
                                npq_dev, cand_.as<unsigned long long>(), s);
            }
            if (!rows_form)
                launch_ivf_merge(cand_.as<unsigned long long>(), per_query, rows, k,
                                 cflags_.as<uint32_t>(), d_idx + q0 * k, d_dist + q0 * k,
                                 d_count + q0, mflags_.as<uint32_t>(), s);
            HM_HIP(hipMemcpyAsync(fl.data(), mflags_.get(), rows * 4, hipMemcpyDeviceToHost, s));
            if (learned)
                HM_HIP(hipMemcpyAsync(npq.data(), nprobe_q_.get(), rows * 4,
                                      hipMemcpyDeviceToHost, s));
            HM_HIP(hipStreamSynchronize(s));
            if (learned)
                for (uint32_t i = 0; i < rows; ++i) probes_used += npq[i];
            else
                probes_used += (uint64_t)rows * np;
            std::vector<uint32_t> tied;
            for (uint32_t i = 0; i < rows && !canonical; ++i)
                if (fl[i] & ROW_TIE) tied.push_back(i);
            last_replayed_ += tied.size();
            if (!tied.empty())
                replay(tied, k, ldd, learned, Kc, d_idx + q0 * k, d_dist + q0 * k, d_count + q0,
                       s);
        }
        last_mean_nprobe_ = (double)probes_used / (double)nq;
        last_rows_ = nq;
        last_chunks_ = (uint32_t)((nq + qc - 1) / qc);
        last_kc_ = Kc;
        last_learned_ = learned;
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, err_.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
    }

    double last_mean_nprobe() const { return last_mean_nprobe_; }

    // Lists scanned per query by the last search (one chunk of queries):
    // probes [nq][ldp] in scan order, counts [nq].
    void last_probes(uint64_t nq, uint32_t ldp, uint32_t *probes, uint32_t *counts) const
    {
        if (nq != last_rows_ || last_chunks_ != 1)
            fail(HM_ERR_INVALID_ARG, "last search did not run as one chunk of this size");
        std::vector<uint32_t> pr((size_t)nq * last_kc_), cn(nq);
        HM_HIP(hipMemcpyAsync(pr.data(), probe_i_.get(), pr.size() * 4, hipMemcpyDeviceToHost,
                              stream_));
        if (last_learned_)
            HM_HIP(hipMemcpyAsync(cn.data(), nprobe_q_.get(), nq * 4, hipMemcpyDeviceToHost,
                                  stream_));
        HM_HIP(hipStreamSynchronize(stream_));
        for (uint64_t q = 0; q < nq; ++q) {
            counts[q] = last_learned_ ? cn[q] : nprobe_;
            for (uint32_t j = 0; j < ldp; ++j)
                probes[q * ldp + j] = j < last_kc_ ? pr[q * last_kc_ + j] : 0xFFFFFFFFu;
        }
    }

    void set_scan_form(int32_t form)
    {
        if (form < 0 || form > 3) fail(HM_ERR_INVALID_ARG, "scan form must be 0..3");
        scan_mode_ = form;
    }

    void set_probe_policy(int32_t mode, float mass, uint32_t pmin, uint32_t pmax)
    {
        if (mode != 0 && mode != 1) fail(HM_ERR_INVALID_ARG, "probe mode must be 0 or 1");
        if (mode == 1 && !mlp_) fail(HM_ERR_NOT_BUILT, "probe network not trained");
        if (pmin == 0 || pmin > pmax || pmax > nlist_)
            fail(HM_ERR_INVALID_ARG, "need 1 <= pmin <= pmax <= n_kmeans_center");
        probe_mode_ = mode;
        mass_ = mass;
        pmin_ = pmin;
        pmax_ = pmax;
    }

    // Trains the probe network: labels are the inverted lists holding each
    // training query's exact top-label_k (GPU brute force over the indexed
    // rows); the target of a query is the fraction of its top-k per list.
    double train_probe(const float *h_queries, uint64_t n, uint32_t label_k, uint32_t hidden,
                       uint32_t epochs, uint32_t batch, float lr, uint64_t seed)
    {
        if (!trained_) fail(HM_ERR_NOT_BUILT, "index not built");
        if (n == 0 || label_k == 0 || hidden == 0 || batch == 0)
            fail(HM_ERR_INVALID_ARG, "empty probe training set or shape");
        check_finite_host(h_queries, n * dim_);
        hipStream_t s = stream_;
        DeviceBuffer Xtr, X16, lab, ids, dist, cnt, T;
        Xtr.reserve(n * dim_ * 4);
        HM_HIP(hipMemcpyAsync(Xtr.get(), h_queries, n * dim_ * 4, hipMemcpyHostToDevice, s));
        X16.reserve(n * dim_ * 2);
        f32_to_bf16(Xtr.as<float>(), X16.as<uint16_t>(), (long)(n * dim_), s);
        HM_HIP(hipStreamSynchronize(s));
        {
            auto bf = make_bruteforce(dim_);
            bf->add_device(base_.as<float>(), n_trained_, dpad_);
            bf->build(metric_);
            if (metric_ != HM_METRIC_MANHATTAN) bf->set_search_mode(HM_SEARCH_BF16_RERANK, 0);
            const uint64_t chunk = 4096;
            ids.reserve(chunk * label_k * 8);
            dist.reserve(chunk * label_k * 4);
            cnt.reserve(chunk * 4);
            lab.reserve(n * label_k * 4);
            for (uint64_t c0 = 0; c0 < n; c0 += chunk) {
                const uint64_t rows = std::min<uint64_t>(chunk, n - c0);
                bf->search_device(Xtr.as<float>() + c0 * dim_, rows, label_k, ids.as<uint64_t>(),
                                  dist.as<float>(), cnt.as<uint32_t>(), s);
                launch_ids_to_lists(ids.as<uint64_t>(), rows * label_k, assign_.as<uint32_t>(),
                                    n_trained_, lab.as<uint32_t>() + c0 * label_k, s);
            }
            HM_HIP(hipStreamSynchronize(s));
        }
        mlp_ = std::make_unique<ProbeMLP>((int)dim_, (int)hidden, (int)nlist_, seed, s);
        T.reserve((size_t)batch * nlist_ * 4);
        std::vector<float> lossv(batch);
        double last = 0.0;
        for (uint32_t ep = 0; ep < epochs; ++ep) {
            double tot = 0.0;
            for (uint64_t b0 = 0; b0 < n; b0 += batch) {
                const int rows = (int)std::min<uint64_t>(batch, n - b0);
                launch_dense_targets(lab.as<uint32_t>() + b0 * label_k, label_k, rows, nlist_,
                                     T.as<float>(), s);
                mlp_->loss_grad(X16.as<uint16_t>() + b0 * dim_, T.as<float>(), rows);
                mlp_->adam(lr, 0.0f);
                if (ep + 1 == epochs) {
                    HM_HIP(hipMemcpyAsync(lossv.data(), mlp_->loss_buf(), rows * 4,
                                          hipMemcpyDeviceToHost, s));
                    HM_HIP(hipStreamSynchronize(s));
                    for (int i = 0; i < rows; ++i) tot += lossv[i];
                }
            }
            if (ep + 1 == epochs) last = tot / (double)n;
        }
        HM_HIP(hipStreamSynchronize(s));
        return last;
    }

    void search_stats(uint64_t *fallback, uint64_t *uncertified) const override
    {
        *fallback = last_replayed_;
        *uncertified = 0;
    }

  private:
    void replay(const std::vector<uint32_t> &tied, uint32_t k, uint64_t ldd, bool learned,
                uint32_t ldp, uint64_t *d_idx, float *d_dist, uint32_t *d_count, hipStream_t s)
    {
        const uint32_t m = (uint32_t)tied.size();
        const size_t per = (size_t)nlist_ + max_list_ + (k + 1) + (size_t)M_ * KS_;
        rp_rows_.reserve(m * 4);
        HM_HIP(hipMemcpyAsync(rp_rows_.get(), tied.data(), m * 4, hipMemcpyHostToDevice, s));
        const uint32_t np_max = learned ? ldp : nprobe_;
        rp_order_.reserve((size_t)m * np_max * 4);
        rp_np_.reserve((size_t)m * 4);
        rp_seq_.reserve((size_t)m * np_max * k * 8);
        rp_scnt_.reserve((size_t)m * np_max * 4);
        if (launch_ivf_replay_lds(rp_rows_.as<uint32_t>(), m, D_.as<float>(), (uint32_t)ldd,
                                  nlist_, np_max, qpad_.as<float>(), dpad_, centers_.as<float>(),
                                  off_.as<uint32_t>(), ids_.as<uint32_t>(), codes_.as<uint8_t>(),
                                  codebooks_.as<float>(), sub_b_.as<uint32_t>(),
                                  sub_e_.as<uint32_t>(), M_, KS_, sdm_, k, max_list_,
                                  probe_i_.as<uint32_t>(), ldp,
                                  learned ? nprobe_q_.as<uint32_t>() : nullptr,
                                  rp_order_.as<uint32_t>(), rp_np_.as<uint32_t>(),
                                  rp_seq_.as<uint2>(), rp_scnt_.as<uint32_t>(), d_idx, d_dist,
                                  d_count, s, learned ? nullptr : cflags_.as<uint32_t>())) {
            HM_HIP(hipStreamSynchronize(s));
            return;
        }
        rp_f_.reserve(m * per * 4);
        rp_u_.reserve(m * per * 4);
        launch_ivf_replay(rp_rows_.as<uint32_t>(), m, D_.as<float>(), (uint32_t)ldd, nlist_,
                          nprobe_, qpad_.as<float>(), dpad_, centers_.as<float>(),
                          off_.as<uint32_t>(), ids_.as<uint32_t>(), codes_.as<uint8_t>(),
                          codebooks_.as<float>(), sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_,
                          KS_, sdm_, k, max_list_, rp_f_.as<float>(), rp_u_.as<uint32_t>(), per,
                          learned ? probe_i_.as<uint32_t>() : nullptr, ldp,
                          learned ? nprobe_q_.as<uint32_t>() : nullptr, d_idx, d_dist, d_count, s);
        HM_HIP(hipStreamSynchronize(s));
    }

    // kmeans.rs:127-148 over all rows against the current centres.
    void assign_all(hipStream_t s)
    {
        const bool norms = metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR;
        if (norms)
            launch_self_dot(centers_.as<float>(), dpad_, nlist_, (uint32_t)dim_, cnorm_.as<float>(),
                            s);
        HM_HIP(hipMemsetAsync(best_.get(), 0xFF, n_ * 8, s));
        launch_exact_argmin(metric_, base_.as<float>(), dpad_, (uint32_t)n_, centers_.as<float>(),
                            dpad_, nlist_, dpad_, xnorm_.as<float>(), cnorm_.as<float>(),
                            best_.as<unsigned long long>(), s);
        launch_unpack_assign(best_.as<unsigned long long>(), n_, assign_.as<uint32_t>(), s);
        build_lists(assign_.as<uint32_t>(), n_, nlist_, lists_.as<uint32_t>(),
                    ids_.as<uint32_t>(), off_.as<uint32_t>(), iota_, sort_tmp_, s);
    }

    // IVFPQIndex::train (pq_idx.rs:357-396).
    void train()
    {
        hipStream_t s = stream_;
        const uint64_t n = n_;
        // An empty index: the reference's split_center returns Err for a
        // batch of 0 and train unwraps it (kmeans.rs:158-160, 203-204).
        if (n == 0) fail(HM_ERR_EMPTY, "None to assigned impossible split center");
        const bool fp64 = p_.kmeans_accumulate != 0;
        const uint32_t epochs = p_.train_epoch;
        centers_.reserve((size_t)nlist_ * dpad_ * 4);
        cnorm_.reserve((size_t)nlist_ * 4);
        best_.reserve(n * 8);
        assign_.reserve(n * 4);
        lists_.reserve(n * 4);
        ids_.reserve(n * 4);
        off_.reserve((nlist_ + 1) * 4);
        const bool norms = metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR;
        if (norms) {
            xnorm_.reserve(n * 4);
            launch_self_dot(base_.as<float>(), dpad_, (uint32_t)n, (uint32_t)dim_,
                            xnorm_.as<float>(), s);
        }
        // Kmeans::train (kmeans.rs:195-207)
        launch_coarse_init(fp64, base_.as<float>(), dpad_, n, (uint32_t)dim_, nlist_,
                           centers_.as<float>(), dpad_, seg_ws_, s);
        // the split's generator lives on the device across epochs
        const uint64_t rng0 = host_stream_seed(p_.seed, 0);
        km_rng_.reserve(8);
        km_cnt_.reserve((size_t)nlist_ * 4);
        km_pk_.reserve((size_t)nlist_ * 8);
        HM_HIP(hipMemcpyAsync(km_rng_.get(), &rng0, 8, hipMemcpyHostToDevice, s));
        for (uint32_t ep = 0; ep < epochs; ++ep) {
            assign_all(s);
            launch_update(fp64, base_.as<float>(), dpad_, off_.as<uint32_t>(), ids_.as<uint32_t>(),
                          nlist_, (uint32_t)dim_, centers_.as<float>(), dpad_, n, seg_ws_, s);
            if (ep + 1 < epochs)
                launch_coarse_split(centers_.as<float>(), dpad_, nlist_, (uint32_t)dim_,
                                    off_.as<uint32_t>(), n, km_rng_.as<uint64_t>(),
                                    km_cnt_.as<uint32_t>(), km_pk_.as<double>(), s);
        }
        std::vector<uint32_t> hoff(nlist_ + 1);
        // search_data after train -> _ivflist (pq_idx.rs:369-375)
        assign_all(s);
        HM_HIP(hipMemcpyAsync(hoff.data(), off_.get(), (nlist_ + 1) * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        max_list_ = 0;
        for (uint32_t c = 0; c < nlist_; ++c) max_list_ = std::max(max_list_, hoff[c + 1] - hoff[c]);
        // per-list residual PQ (pq_idx.rs:376-393); workgroups take the
        // lists longest first
        std::vector<uint32_t> lorder(nlist_);
        for (uint32_t c = 0; c < nlist_; ++c) lorder[c] = c;
        std::stable_sort(lorder.begin(), lorder.end(), [&](uint32_t a, uint32_t b) {
            return hoff[a + 1] - hoff[a] > hoff[b + 1] - hoff[b];
        });
        DeviceBuffer R, asg, ord, dlorder;
        R.reserve(n * dpad_ * 4);
        launch_residuals(base_.as<float>(), dpad_, ids_.as<uint32_t>(), lists_.as<uint32_t>(),
                         centers_.as<float>(), n, dpad_, R.as<float>(), s);
        codebooks_.reserve((size_t)nlist_ * M_ * KS_ * sdm_ * 4);
        codes_.reserve(n * M_);
        asg.reserve((size_t)M_ * n * 2);
        ord.reserve((size_t)M_ * n * 4);
        dlorder.reserve((size_t)nlist_ * 4);
        HM_HIP(hipMemcpyAsync(dlorder.get(), lorder.data(), (size_t)nlist_ * 4,
                              hipMemcpyHostToDevice, s));
        launch_pq_train(fp64, R.as<float>(), dpad_, off_.as<uint32_t>(), dlorder.as<uint32_t>(),
                        nlist_, M_, sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), KS_, sd_max_,
                        sdm_, epochs, p_.seed, max_list_, n, asg.as<uint16_t>(),
                        ord.as<uint32_t>(), codebooks_.as<float>(), codes_.as<uint8_t>(), s);
        HM_HIP(hipStreamSynchronize(s));
        if (norms)
            launch_self_dot(centers_.as<float>(), dpad_, nlist_, (uint32_t)dim_, cnorm_.as<float>(),
                            s);
        HM_HIP(hipStreamSynchronize(s));
        trained_ = true;
        n_trained_ = n;
    }

    hm_params p_;
    uint32_t dpad_ = 0, M_ = 0, KS_ = 0, nlist_ = 0, nprobe_ = 0, sd_max_ = 0, sdm_ = 0;
    uint32_t max_list_ = 0;
    std::vector<uint32_t> sub_b_h_, sub_e_h_;
    uint64_t n_ = 0, cap_ = 0, n_trained_ = 0, last_replayed_ = 0;
    int32_t metric_ = HM_METRIC_UNKNOWN;
    bool trained_ = false;
    DeviceBuffer base_, sub_b_, sub_e_, centers_, cnorm_, xnorm_, best_, assign_, lists_, ids_,
        off_, codebooks_, codes_, iota_, sort_tmp_;
    DeviceBuffer qpad_, qnorm_, D_, probe_d_, probe_i_, cflags_, mflags_, cand_, err_;
    DeviceBuffer pair_cnt_, pair_fill_, pair_off_, unit_off_, pairs_;
    // 0: fastest form that fits; 1: per (query, probe); 2: list-major, top-k1
    // per list; 3: rows (ADC distances into D rows, row select)
    int scan_mode_ = 0;
    DeviceBuffer Dr_, seg_, row_n_, rcand_d_, rcand_i_, rflags_;  // row-form scan
    DeviceBuffer rp_rows_, rp_f_, rp_u_, rp_order_, rp_np_, rp_seq_, rp_scnt_;
    // exact rerank of the ADC top-R
    uint32_t rerank_ = 0;
    uint64_t rr_bnorm_n_ = 0;
    DeviceBuffer rr_i_, rr_d_, rr_c_, rr_ci_, rr_qpad_, rr_qn_, rr_f_, rr_bnorm_;
    // learned probe
    std::unique_ptr<ProbeMLP> mlp_;
    int32_t probe_mode_ = 0;
    float mass_ = 0.9f;
    uint32_t pmin_ = 1, pmax_ = 1;
    double last_mean_nprobe_ = 0.0;
    uint64_t last_rows_ = 0;
    uint32_t last_chunks_ = 0, last_kc_ = 0;
    bool last_learned_ = false;
    DeviceBuffer xq16_, Zq_, Pq_, nprobe_q_;
    SegScratch seg_ws_;                 // fp64 centre steps
    DeviceBuffer km_rng_, km_cnt_, km_pk_;  // coarse split: generator, counts, picks
};

}  // namespace

std::unique_ptr<Index> make_ivfpq(uint64_t dim, const hm_params &p)
{
    return std::make_unique<IVFPQ>(dim, p);
}

const hm_params &ivfpq_params(Index *ix)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    return p->params();
}

void ivfpq_import(Index *ix, int32_t metric, const float *centers, const uint64_t *offsets,
                  const uint64_t *ids, const float *codebooks, const uint8_t *codes)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->import_trained(metric, centers, offsets, ids, codebooks, codes);
}

int32_t ivfpq_info(Index *ix, uint32_t *out)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->info(out);
    return HM_OK;
}

int32_t ivfpq_export(Index *ix, float *centers, uint64_t *offsets, uint64_t *ids,
                     float *codebooks, uint8_t *codes)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->export_trained(centers, offsets, ids, codebooks, codes);
    return HM_OK;
}

int32_t ivfpq_set_rerank(Index *ix, uint32_t depth)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->set_rerank(depth);
    return HM_OK;
}

int32_t ivfpq_set_scan_form(Index *ix, int32_t form)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->set_scan_form(form);
    return HM_OK;
}

int32_t ivfpq_set_probe_policy(Index *ix, int32_t mode, float mass, uint32_t pmin, uint32_t pmax)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->set_probe_policy(mode, mass, pmin, pmax);
    return HM_OK;
}

int32_t ivfpq_train_probe(Index *ix, const float *q, uint64_t n, uint32_t label_k,
                          uint32_t hidden, uint32_t epochs, uint32_t batch, float lr,
                          uint64_t seed, double *loss)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    double l = p->train_probe(q, n, label_k, hidden, epochs, batch, lr, seed);
    if (loss) *loss = l;
    return HM_OK;
}

int32_t ivfpq_last_probes(Index *ix, uint64_t nq, uint32_t ldp, uint32_t *probes,
                          uint32_t *counts)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->last_probes(nq, ldp, probes, counts);
    return HM_OK;
}

double ivfpq_mean_nprobe(Index *ix)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    return p->last_mean_nprobe();
}

int32_t ivfpq_set_nprobe(Index *ix, uint32_t nprobe)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->set_nprobe(nprobe);
    return HM_OK;
}

}  // namespace hm
