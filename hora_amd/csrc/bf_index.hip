// bf_index.hip -- BruteForceIndex on the device (bruteforce_idx.rs:15-106).
//
// HBM layout: the base is one row-major matrix of `cap` rows x dpad floats
// (dpad = dimension rounded up to 16, zero padded), rows in insertion
// order, so row i is the reference's internal index i.  Derived data (the
// per-row s_bb sums for Cosine / Angular) is rebuilt lazily after an add.
//
// EXACT search: for a chunk of queries the exact tile kernel writes the
// full distance block D[chunk][n] (reference-order FP32), the radix select
// picks the min(k+1, n) smallest by (distance, index) and flags ties, and
// finalize emits the first k or replays the heap for a tied row.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "index.h"
#include "kernels.h"

namespace hm {

void check_finite_host(const float *x, uint64_t count)
{
    for (uint64_t i = 0; i < count; ++i)
        if (!std::isfinite(x[i])) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
}

namespace {

__global__ void fill_results_kernel(uint64_t *idx, float *dist, uint32_t *count, uint64_t nq,
                                    uint32_t k)
{
    uint64_t total = nq * k;
    const float inf = __int_as_float(0x7f800000);
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        idx[e] = ~0ull;
        dist[e] = inf;
    }
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nq;
         q += (uint64_t)gridDim.x * blockDim.x)
        count[q] = 0;
}

class BruteForce final : public Index {
  public:
    explicit BruteForce(uint64_t dim) : Index(HM_INDEX_BRUTE_FORCE, dim)
    {
        if (dim == 0 || dim > (1u << 20)) fail(HM_ERR_INVALID_ARG, "dimension out of range");
        dpad_ = (uint32_t)round_up(dim, 16);
        HM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    }
    ~BruteForce() override
    {
        drop_ragged_graphs();
        if (rg_stream_) (void)hipStreamDestroy(rg_stream_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    void add(const float *rows, uint64_t n_rows) override
    {
        if (n_rows == 0) return;
        if (!rows) fail(HM_ERR_INVALID_ARG, "rows is null");
        check_finite_host(rows, n_rows * dim_);
        if (n_ + n_rows > 0xFFFFFFF0ull) fail(HM_ERR_CAPACITY, "more than 2^32-16 rows");
        uint64_t need = n_ + n_rows;
        if (need > cap_) {
            uint64_t nc = std::max<uint64_t>(need, cap_ + cap_ / 2);
            nc = std::max<uint64_t>(nc, 1024);
            base_.grow_keep(nc * dpad_ * sizeof(float), n_ * dpad_ * sizeof(float), stream_);
            cap_ = nc;
        }
        DeviceBuffer stage;
        upload_rows(rows, n_rows, (uint32_t)dim_, base_.as<float>() + n_ * dpad_, dpad_, stage,
                    stream_);
        HM_HIP(hipStreamSynchronize(stream_));
        n_ = need;
        bnorm_valid_ = false;
        sl_valid_ = false;
        samp_ns_ = 0;
        prepare_norms();
    }

    void add_device(const float *d_rows, uint64_t n_rows, uint64_t ld) override
    {
        if (n_rows == 0) return;
        if (n_ + n_rows > 0xFFFFFFF0ull) fail(HM_ERR_CAPACITY, "more than 2^32-16 rows");
        const uint64_t need = n_ + n_rows;
        if (need > cap_) {
            base_.grow_keep(need * dpad_ * sizeof(float), n_ * dpad_ * sizeof(float), stream_);
            cap_ = need;
        }
        float *dst = base_.as<float>() + n_ * dpad_;
        launch_pad_rows(d_rows, ld, dst, dpad_, n_rows, (uint32_t)dim_, dpad_, stream_);
        HM_HIP(hipStreamSynchronize(stream_));
        n_ = need;
        bnorm_valid_ = false;
        sl_valid_ = false;
        samp_ns_ = 0;
        prepare_norms();
    }

    int32_t metric() const override { return metric_; }

    void copy_rows(uint64_t first, uint64_t count, float *out) const override
    {
        if (first + count > n_) fail(HM_ERR_INVALID_ARG, "row range out of bounds");
        if (count == 0) return;
        DeviceBuffer stage;
        download_rows(base_.as<float>() + first * dpad_, dpad_, count, (uint32_t)dim_, out, stage,
                      stream_);
        HM_HIP(hipStreamSynchronize(stream_));
    }

    void build(int32_t metric) override
    {
        if (metric < HM_METRIC_UNKNOWN || metric > HM_METRIC_ANGULAR)
            fail(HM_ERR_INVALID_ARG, "metric out of range");
        metric_ = metric;  // bruteforce_idx.rs:36-39 stores mt and nothing else
        prepare_norms();
    }

    // Cosine / Angular read the base rows' self-dots; computing them here
    // (and after every add) keeps the no-allocation search free of it.
    void prepare_norms()
    {
        if (!need_norm() || n_ == 0 || bnorm_valid_) return;
        ensure_bnorm(stream_);
        HM_HIP(hipStreamSynchronize(stream_));
    }

    bool built() const override { return true; }  // bruteforce_idx.rs:44-46

    const char *name() const override { return "BruteForceIndex"; }
    uint64_t nodes_size() const override { return n_; }

    void clear() override
    {
        base_.release();
        bnorm_.release();
        n_ = cap_ = 0;
        bnorm_valid_ = false;
        sl_valid_ = false;
        b8_valid_ = false;
        b8_.release();
        samp_ns_ = 0;
    }

    void get_rows(const uint64_t *idx, uint64_t n, float *out) const override
    {
        for (uint64_t i = 0; i < n; ++i) {
            if (idx[i] >= n_) fail(HM_ERR_INVALID_ARG, "row index out of range");
            HM_HIP(hipMemcpyAsync(out + i * dim_, base_.as<float>() + idx[i] * dpad_,
                                  dim_ * sizeof(float), hipMemcpyDeviceToHost, stream_));
        }
        HM_HIP(hipStreamSynchronize(stream_));
    }

    void set_search_mode(int32_t mode, uint32_t shortlist) override
    {
        if (mode < HM_SEARCH_EXACT || mode > HM_SEARCH_FP8_RERANK)
            fail(HM_ERR_INVALID_ARG, "unknown search mode");
        if (shortlist > SELECT_MAX_K) fail(HM_ERR_INVALID_ARG, "shortlist above 8192");
        if (mode != mode_) {
            sl_valid_ = false;
            samp_ns_ = 0;
        }
        mode_ = mode;
        shortlist_ = shortlist;
    }

    void search_device(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx, float *d_dist,
                       uint32_t *d_count, hipStream_t s) override
    {
        if (!s) s = stream_;
        if (metric_ == HM_METRIC_UNKNOWN) fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        if (nq == 0) return;
        if (!dq || !d_count || (k && (!d_idx || !d_dist)))
            fail(HM_ERR_INVALID_ARG, "null pointer");
        if (nq > 0x7FFFFFFFull) fail(HM_ERR_INVALID_ARG, "too many queries");
        if (k == 0 || n_ == 0) {
            launch_fill_results(d_idx, d_dist, d_count, nq, k, s);
            HM_HIP(hipStreamSynchronize(s));
            return;
        }
        err_.reserve(sizeof(uint32_t));
        HM_HIP(hipMemsetAsync(err_.get(), 0, sizeof(uint32_t), s));
        const bool shortlist_form = shortlist_usable(s);
        // the shortlist form checks the queries while it pads them
        if (!shortlist_form) launch_check_finite(dq, dim_, nq, (uint32_t)dim_, err_.as<uint32_t>(), s);
        last_fallback_ = last_uncertified_ = 0;
        herr_.reserve(sizeof(uint32_t));
        bool err_read = false;  // the shortlist form may read the error word itself
        if (shortlist_form)
            err_read = search_shortlist(dq, nq, k, d_idx, d_dist, d_count, s);
        else
            search_exact(dq, dim_, nq, k, d_idx, d_dist, d_count, s);
        if (!err_read) {
            HM_HIP(hipMemcpyAsync(herr_.as<uint32_t>(), err_.get(), sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, s));
            HM_HIP(hipStreamSynchronize(s));
        }
        const uint32_t err = *herr_.as<uint32_t>();
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
        if (err & 4u) fail(HM_ERR_HIP, "internal: selection consistency check failed");
        if (err & 1u) fail(HM_ERR_NAN_DISTANCE, "NaN distance (partial_cmp of NaN)");
    }

  private:
    // Metrics with an MFMA shortlist form.  Manhattan has no GEMM form.
    bool shortlist_metric() const
    {
        return metric_ == HM_METRIC_EUCLIDEAN || metric_ == HM_METRIC_DOT_PRODUCT ||
               metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR;
    }

    // A zero-norm base row has no Cosine/Angular key (0 * inf), so the
    // shortlist could not bound it; such bases search on the exact form
    // (where Cosine reports the NaN distance as the reference would fail).
    bool shortlist_usable(hipStream_t s)
    {
        if (mode_ == HM_SEARCH_EXACT || !shortlist_metric() || dpad_ > 4096 || n_ == 0) return false;
        ensure_shortlist_data(s);
        return !(zero_norm_ && need_norm());
    }

    bool need_norm() const
    {
        return metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR;
    }

    void ensure_bnorm(hipStream_t s)
    {
        if (bnorm_valid_) return;
        bnorm_.reserve(n_ * sizeof(float));
        launch_self_dot(base_.as<float>(), dpad_, (uint32_t)n_, (uint32_t)dim_, bnorm_.as<float>(),
                        s);
        bnorm_valid_ = true;
    }

    // Full exact path: every pair in reference order, radix top-(k+1), heap
    // replay for tied rows.  Queries: nq rows of dim floats, row stride ldq.
    // all_tied: every row is known to hold a tie among its k+1 smallest
    // distances (and no NaN), so the select is skipped and each row goes
    // straight to the heap replay over its exact distance row.
    void search_exact(const float *dq, uint64_t ldq, uint64_t nq, uint32_t k, uint64_t *d_idx,
                      float *d_dist, uint32_t *d_count, hipStream_t s, bool all_tied = false)
    {
        const uint32_t n = (uint32_t)n_;
        const uint32_t K = (uint32_t)std::min<uint64_t>((uint64_t)k + 1, n);
        if (K > SELECT_MAX_K) fail(HM_ERR_INVALID_ARG, "k too large (max 8191)");
        if (need_norm()) ensure_bnorm(s);
        const uint64_t ldd = round_up(n, 4);
        uint64_t qc = std::max<uint64_t>(1, d_budget_ / (ldd * sizeof(float)));
        qc = std::min<uint64_t>(qc, nq);
        qc = std::min<uint64_t>(qc, 65535ull * 128);
        qpad_.reserve(qc * dpad_ * sizeof(float));
        D_.reserve(qc * ldd * sizeof(float));
        cand_d_.reserve(qc * K * sizeof(float));
        cand_i_.reserve(qc * K * sizeof(uint32_t));
        flags_.reserve(qc * sizeof(uint32_t));
        qnorm_.reserve(qc * sizeof(float));
        const ExactBufs b{qpad_.as<float>(), qnorm_.as<float>(), D_.as<float>(),
                          cand_d_.as<float>(), cand_i_.as<uint32_t>(), flags_.as<uint32_t>(),
                          err_.as<uint32_t>()};
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            exact_pass(dq + q0 * ldq, ldq, rows, k, K, ldd, b, d_idx + q0 * k, d_dist + q0 * k,
                       d_count + q0, s, all_tied);
        }
    }

    // Working buffers of one exact pass over `rows` queries: padded queries
    // [rows][dpad], their self-dots, the distance block [rows][ldd], the
    // top-K candidates [rows][K] and row flags, and the error word.
    struct ExactBufs {
        float *qpad, *qnorm, *D, *cand_d;
        uint32_t *cand_i, *flags, *err;
    };

    void exact_pass(const float *dq, uint64_t ldq, uint32_t rows, uint32_t k, uint32_t K,
                    uint64_t ldd, const ExactBufs &b, uint64_t *d_idx, float *d_dist,
                    uint32_t *d_count, hipStream_t s, bool all_tied)
    {
        const uint32_t n = (uint32_t)n_;
        launch_pad_rows(dq, ldq, b.qpad, dpad_, rows, (uint32_t)dim_, dpad_, s);
        if (need_norm()) launch_self_dot(b.qpad, dpad_, rows, (uint32_t)dim_, b.qnorm, s);
        launch_exact_tile(metric_, b.qpad, dpad_, rows, base_.as<float>(), dpad_, n, dpad_,
                          b.qnorm, bnorm_.as<float>(), b.D, ldd, s);
        if (!all_tied)
            launch_row_select(b.D, ldd, rows, n, K, b.cand_d, b.cand_i, K, b.flags, s);
        launch_bf_finalize(b.cand_d, b.cand_i, K, all_tied ? nullptr : b.flags, b.D, ldd, rows, n,
                           k, d_idx, d_dist, d_count, b.err, s);
    }

  public:
    // search_device in the exact form with every working buffer carved
    // from `scratch` (no allocation; see the header of noalloc.hip).
    void search_device_scratch(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx,
                               float *d_dist, uint32_t *d_count, hipStream_t s, void *scratch,
                               size_t scratch_bytes) override
    {
        if (!s) s = stream_;
        if (metric_ == HM_METRIC_UNKNOWN) fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        if (nq == 0) return;
        if (!dq || !d_count || (k && (!d_idx || !d_dist)))
            fail(HM_ERR_INVALID_ARG, "null pointer");
        if (nq > 0x7FFFFFFFull) fail(HM_ERR_INVALID_ARG, "too many queries");
        if (k == 0 || n_ == 0) {
            launch_fill_results(d_idx, d_dist, d_count, nq, k, s);
            HM_HIP(hipStreamSynchronize(s));
            return;
        }
        const uint32_t n = (uint32_t)n_;
        const uint32_t K = (uint32_t)std::min<uint64_t>((uint64_t)k + 1, n);
        if (K > SELECT_MAX_K) fail(HM_ERR_INVALID_ARG, "k too large (max 8191)");
        if (need_norm() && !bnorm_valid_)
            fail(HM_ERR_INVALID_ARG, "base norms not prepared: build the index after adding rows");
        const uint64_t ldd = round_up(n, 4);
        // per query: padded row, self-dot, distance row, K candidates, flag;
        // each array of the carve starts on a 256-byte boundary
        const uint64_t per_q = (uint64_t)dpad_ * 4 + 4 + ldd * 4 + (uint64_t)K * 8 + 4;
        const uint64_t slack = 256 * 7;
        if (scratch_bytes < slack + per_q)
            fail(HM_ERR_CAPACITY, "index too large for the no-allocation scratch (" +
                                      std::to_string(scratch_bytes >> 20) + " MiB holds no " +
                                      std::to_string(per_q) + "-byte query)");
        const uint64_t qc = std::min<uint64_t>(nq, (scratch_bytes - slack) / per_q);
        unsigned char *p = static_cast<unsigned char *>(scratch);
        auto carve = [&](uint64_t bytes) {
            void *r = p;
            p += round_up(bytes, 256);
            return r;
        };
        ExactBufs b;
        b.err = static_cast<uint32_t *>(carve(4));
        b.qpad = static_cast<float *>(carve(qc * dpad_ * 4));
        b.qnorm = static_cast<float *>(carve(qc * 4));
        b.D = static_cast<float *>(carve(qc * ldd * 4));
        b.cand_d = static_cast<float *>(carve(qc * K * 4));
        b.cand_i = static_cast<uint32_t *>(carve(qc * K * 4));
        b.flags = static_cast<uint32_t *>(carve(qc * 4));
        // the array keeps the previous call's contents: clear the error word
        // with a kernel (no memset node, no allocation)
        launch_fill_u32(b.err, 1, 0, s);
        launch_check_finite(dq, dim_, nq, (uint32_t)dim_, b.err, s);
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            exact_pass(dq + q0 * dim_, dim_, rows, k, K, ldd, b, d_idx + q0 * k, d_dist + q0 * k,
                       d_count + q0, s, false);
        }
        uint32_t err = 0;  // pageable stack word: no pinned allocation
        HM_HIP(hipMemcpyAsync(&err, b.err, 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
        if (err & 4u) fail(HM_ERR_HIP, "internal: selection consistency check failed");
        if (err & 1u) fail(HM_ERR_NAN_DISTANCE, "NaN distance (partial_cmp of NaN)");
    }

  private:

    // naive_build_knn_graph (knn.rs:14-41): exact distances of a chunk of
    // rows against all rows, the row's own column excluded.
    void knn_graph(uint32_t k, uint64_t *h_idx, float *h_dist, uint32_t *h_count) override
    {
        if (metric_ == HM_METRIC_UNKNOWN) fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        if (!h_count || (k && (!h_idx || !h_dist))) fail(HM_ERR_INVALID_ARG, "null pointer");
        const uint32_t n = (uint32_t)n_;
        if (n == 0) return;
        if (k == 0 || n == 1) {
            for (uint32_t i = 0; i < n; ++i) h_count[i] = 0;
            for (uint64_t j = 0; j < (uint64_t)n * k; ++j) {
                h_idx[j] = ~0ull;
                h_dist[j] = __builtin_inff();
            }
            return;
        }
        if (k + 2 > SELECT_MAX_K) fail(HM_ERR_INVALID_ARG, "k too large (max 8190)");
        hipStream_t s = stream_;
        const uint32_t Ksel = std::min<uint32_t>(k + 2, n);
        if (need_norm()) ensure_bnorm(s);
        const uint64_t ldd = round_up(n, 4), ldc = round_up(n - 1, 4);
        uint64_t qc = std::max<uint64_t>(1, d_budget_ / (ldd * sizeof(float) * 2));
        qc = std::min<uint64_t>(std::min<uint64_t>(qc, n), 65535);
        DeviceBuffer gi, gd, gc, tied, rl, tf, C, ti, td, tc;
        gi.reserve((size_t)n * k * 8);
        gd.reserve((size_t)n * k * 4);
        gc.reserve((size_t)n * 4);
        tied.reserve(qc * 4);
        D_.reserve(qc * ldd * sizeof(float));
        cand_d_.reserve(qc * Ksel * sizeof(float));
        cand_i_.reserve(qc * Ksel * sizeof(uint32_t));
        flags_.reserve(qc * sizeof(uint32_t));
        err_.reserve(4);
        HM_HIP(hipMemsetAsync(err_.get(), 0, 4, s));
        std::vector<uint32_t> th(qc);
        uint64_t replayed = 0;
        for (uint32_t r0 = 0; r0 < n; r0 += (uint32_t)qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, n - r0);
            launch_exact_tile(metric_, base_.as<float>() + (size_t)r0 * dpad_, dpad_, rows,
                              base_.as<float>(), dpad_, n, dpad_,
                              need_norm() ? bnorm_.as<float>() + r0 : nullptr, bnorm_.as<float>(),
                              D_.as<float>(), ldd, s);
            launch_graph_diag(D_.as<float>(), ldd, rows, r0, s);
            launch_row_select(D_.as<float>(), ldd, rows, n, Ksel, cand_d_.as<float>(),
                              cand_i_.as<uint32_t>(), Ksel, flags_.as<uint32_t>(), s);
            launch_graph_finalize(cand_d_.as<float>(), cand_i_.as<uint32_t>(), Ksel, Ksel,
                                  flags_.as<uint32_t>(), rows, r0, n, k, gi.as<uint64_t>(),
                                  gd.as<float>(), gc.as<uint32_t>(), tied.as<uint32_t>(),
                                  err_.as<uint32_t>(), s);
            HM_HIP(hipMemcpyAsync(th.data(), tied.get(), rows * 4, hipMemcpyDeviceToHost, s));
            HM_HIP(hipStreamSynchronize(s));
            std::vector<uint32_t> list;
            for (uint32_t r = 0; r < rows; ++r)
                if (th[r]) list.push_back(r);
            if (list.empty()) continue;
            const uint32_t nt = (uint32_t)list.size();
            replayed += nt;
            rl.reserve(nt * 4);
            tf.reserve(nt * 4);
            C.reserve((size_t)nt * ldc * 4);
            ti.reserve((size_t)nt * k * 8);
            td.reserve((size_t)nt * k * 4);
            tc.reserve((size_t)nt * 4);
            HM_HIP(hipMemcpyAsync(rl.get(), list.data(), nt * 4, hipMemcpyHostToDevice, s));
            launch_graph_compact(D_.as<float>(), ldd, rl.as<uint32_t>(), nt, r0, n, C.as<float>(),
                                 ldc, s);
            launch_fill_u32(tf.as<uint32_t>(), nt, ROW_TIE, s);
            launch_bf_finalize(cand_d_.as<float>(), cand_i_.as<uint32_t>(), Ksel, tf.as<uint32_t>(),
                               C.as<float>(), ldc, nt, n - 1, k, ti.as<uint64_t>(), td.as<float>(),
                               tc.as<uint32_t>(), err_.as<uint32_t>(), s);
            launch_graph_remap(ti.as<uint64_t>(), td.as<float>(), tc.as<uint32_t>(),
                               rl.as<uint32_t>(), nt, r0, k, gi.as<uint64_t>(), gd.as<float>(), s);
            HM_HIP(hipStreamSynchronize(s));
        }
        last_fallback_ = replayed;
        last_uncertified_ = 0;
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, err_.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipMemcpyAsync(h_idx, gi.get(), (size_t)n * k * 8, hipMemcpyDeviceToHost, s));
        HM_HIP(hipMemcpyAsync(h_dist, gd.get(), (size_t)n * k * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipMemcpyAsync(h_count, gc.get(), (size_t)n * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 1u) fail(HM_ERR_NAN_DISTANCE, "NaN distance (partial_cmp of NaN)");
    }

    // The distance evaluations of hnsw_idx.rs:243-254 / ssg_idx.rs:402-456,
    // batched: metric(node, query) for every (query, node) pair.
    void distances_device(const float *dq, uint64_t nq, const uint32_t *ids, const uint64_t *off,
                          uint64_t total, float *out, hipStream_t s) override
    {
        if (!s) s = stream_;
        if (metric_ == HM_METRIC_UNKNOWN) fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        if (nq == 0 || total == 0) return;
        if (!dq || !ids || !off || !out) fail(HM_ERR_INVALID_ARG, "null pointer");
        if (nq > 0xFFFFFFFFull) fail(HM_ERR_INVALID_ARG, "too many queries");
        if (need_norm()) ensure_bnorm(s);
        err_.reserve(4);
        HM_HIP(hipMemsetAsync(err_.get(), 0, 4, s));
        launch_check_finite(dq, dim_, nq, (uint32_t)dim_, err_.as<uint32_t>(), s);
        sv_q_.reserve(nq * dpad_ * 4);
        sv_qn_.reserve(nq * 4);
        launch_pad_rows(dq, dim_, sv_q_.as<float>(), dpad_, nq, (uint32_t)dim_, dpad_, s);
        if (need_norm())
            launch_self_dot(sv_q_.as<float>(), dpad_, (uint32_t)nq, (uint32_t)dim_,
                            sv_qn_.as<float>(), s);
        launch_pair_dist(metric_, sv_q_.as<float>(), dpad_, (uint32_t)nq, base_.as<float>(), dpad_,
                         (uint32_t)n_, ids, off, total, sv_qn_.as<float>(), bnorm_.as<float>(), out,
                         err_.as<uint32_t>(), s);
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, err_.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
        if (err & 4u) fail(HM_ERR_INVALID_ARG, "node index out of range");
    }

    static constexpr uint32_t RAG_CHUNK = 4096;

    // The captured ragged graphs, by call shape: nch chunks of RAG_CHUNK
    // queries each run the shortlist pipeline, then one tail reruns up to
    // RERUN_SLOTS rows that are only tied, with their own k, on the device.
    struct RaggedGraph {
        uint32_t nch, kmax, S, ns;
        int32_t mode, metric;
        uint64_t n;
        std::array<const void *, 10> data;  // the index buffers it reads
        hipGraphExec_t exec;
    };
    static constexpr size_t RAG_GRAPHS = 4;

    void drop_ragged_graphs()
    {
        for (auto &g : rg_graphs_) (void)hipGraphExecDestroy(g.exec);
        rg_graphs_.clear();
    }

    // Grows a buffer the captured graphs use; a growth moves it, so every
    // cached graph is dropped.
    void rg_reserve(DeviceBuffer &b, size_t bytes)
    {
        if (bytes <= b.bytes()) return;
        drop_ragged_graphs();
        b.reserve(bytes);
    }

    hipGraphExec_t ragged_graph(uint32_t nch, uint32_t kmax, uint32_t S, bool dev_rerun)
    {
        const uint32_t QC = RAG_CHUNK;
        const FilterPlan fp = filter_plan(S);
        if (fp.ns) ensure_sample(fp.ns, stream_);
        const uint32_t n = (uint32_t)n_;
        const bool bf16 = mode_ == HM_SEARCH_BF16_RERANK;
        const bool filter = fp.ns != 0;
        const uint32_t row_bytes = bf16 ? dk_ * 2 : dk_;
        const uint64_t ldk = round_up(filter ? fp.ns : n, 4);
        const uint64_t ldd = round_up(n, 4);
        const size_t nq_all = (size_t)nch * QC;
        rg_reserve(rg_call_, sizeof(RaggedCall));
        rg_reserve(rg_kq_, nq_all * 4);
        rg_reserve(rg_off_, (nq_all + 1) * 8);
        rg_reserve(rg_flags_all_, nq_all * 4);
        rg_reserve(rg_rows_, 4);
        rg_reserve(rg_err_, 4);
        rg_reserve(rg_ck_, (size_t)QC * 4);
        rg_reserve(rg_coff_, (size_t)QC * 8);
        rg_reserve(rg_oi_, (size_t)QC * kmax * 8);
        rg_reserve(rg_od_, (size_t)QC * kmax * 4);
        rg_reserve(rg_oc_, (size_t)QC * 4);
        rg_reserve(rg_qpad_, (size_t)QC * dpad_ * 4);
        rg_reserve(rg_qnorm_, (size_t)QC * 4);
        rg_reserve(rg_qlow_, (size_t)QC * row_bytes);
        rg_reserve(rg_qscale_, (size_t)QC * 4);
        rg_reserve(rg_D_, (size_t)QC * ldk * 4);
        rg_reserve(rg_cd_, (size_t)QC * S * 4);
        rg_reserve(rg_ci_, (size_t)QC * S * 4);
        rg_reserve(rg_sf_, (size_t)QC * 4);
        rg_reserve(rg_rf_, (size_t)QC * 4);
        rg_reserve(rg_qinex_, 4);
        rg_reserve(rg_plan_, RERUN_PLAN_WORDS * 4);
        if (filter) {
            rg_reserve(rg_seld_, (size_t)QC * fp.m * 4);
            rg_reserve(rg_seli_, (size_t)QC * fp.m * 4);
            rg_reserve(rg_list_, (size_t)QC * fp.cap * 8);
            rg_reserve(rg_lcnt_, (size_t)QC * 4);
            rg_reserve(rg_fscr_, score_filter_scratch_bytes(QC, fp.cap));
            rg_reserve(rg_sq_, (size_t)QC * 4);
            rg_reserve(rg_ksq_, (size_t)QC * 4);
            rg_reserve(rg_pref_, (size_t)QC * 4);
        }
        if (dev_rerun) {
            rg_reserve(rg_r8p_, (size_t)RERUN_SLOTS * dpad_ * 4);
            rg_reserve(rg_r8n_, RERUN_SLOTS * 4);
            rg_reserve(rg_r8D_, (size_t)RERUN_SLOTS * ldd * 4);
            rg_reserve(rg_r8i_, (size_t)RERUN_SLOTS * kmax * 8);
            rg_reserve(rg_r8d_, (size_t)RERUN_SLOTS * kmax * 4);
            rg_reserve(rg_r8c_, RERUN_SLOTS * 4);
        }
        // a graph reads the index's buffers where they were at capture: one
        // captured before any of them moved (add, clear, mode or sample
        // change) is dropped, never launched
        const std::array<const void *, 10> data{
            base_.get(),     blow_.get(),   bscale_.get(),   bnorm_.get(),
            inv_bnorm_.get(), bsamp_.get(), bs_norm2_.get(), bs_inv_.get(),
            bs_scale_.get(), b8_valid_ ? b8_.get() : nullptr};
        for (size_t i = 0; i < rg_graphs_.size();) {
            if (rg_graphs_[i].data != data) {
                (void)hipGraphExecDestroy(rg_graphs_[i].exec);
                rg_graphs_.erase(rg_graphs_.begin() + i);
            } else {
                ++i;
            }
        }
        for (auto &g : rg_graphs_)
            if (g.nch == nch && g.kmax == kmax && g.S == S && g.ns == fp.ns && g.mode == mode_ &&
                g.metric == metric_ && g.n == n_)
                return g.exec;
        if (rg_graphs_.size() >= RAG_GRAPHS) {
            (void)hipGraphExecDestroy(rg_graphs_.front().exec);
            rg_graphs_.erase(rg_graphs_.begin());
        }
        HM_HIP(hipStreamSynchronize(stream_));  // sample data ready before capture
        if (!rg_stream_) HM_HIP(hipStreamCreateWithFlags(&rg_stream_, hipStreamNonBlocking));
        const float u = 5.9604644775390625e-8f;
        const float gamma = 2.0f * (float)(std::max<uint32_t>(dk_, dpad_) + 8) * u;
        const float c_bf16 = 1.0f / 128.0f + 1.0f / 65536.0f;
        const float c_exact = base_inexact_ ? c_bf16 : 0.0f;
        const uint32_t *dr = rg_rows_.as<uint32_t>();
        const RaggedCall *call = rg_call_.as<RaggedCall>();
        uint32_t *err = rg_err_.as<uint32_t>();
        hipStream_t g = rg_stream_;
        HM_HIP(hipStreamBeginCapture(g, hipStreamCaptureModeThreadLocal));
        launch_fill_u32(err, 1, 0, g);
        for (uint32_t c = 0; c < nch; ++c) {
            launch_ragged_chunk_begin(call, c, QC, (uint32_t)dim_, dpad_, rg_kq_.as<uint32_t>(),
                                      rg_off_.as<uint64_t>(), rg_rows_.as<uint32_t>(),
                                      rg_ck_.as<uint32_t>(), rg_coff_.as<uint64_t>(),
                                      rg_qpad_.as<float>(), err, g);
            launch_fill_u32(rg_qinex_.as<uint32_t>(), 1, 0, g);
            if (query_prep_ok(rg_qpad_.as<float>(), dpad_, (uint32_t)dim_, nullptr, 0)) {
                // the padded rows were checked for finiteness as they were copied
                launch_query_prep(bf16 ? 2 : 1, rg_qpad_.as<float>(), dpad_, QC, (uint32_t)dim_,
                                  dk_, rg_qlow_.get(), rg_qscale_.as<float>(),
                                  rg_qinex_.as<uint32_t>(), rg_qnorm_.as<float>(), nullptr, 0,
                                  nullptr, g, dr);
            } else {
                launch_self_dot(rg_qpad_.as<float>(), dpad_, QC, (uint32_t)dim_,
                                rg_qnorm_.as<float>(), g, dr);
                if (bf16)
                    launch_to_bf16(rg_qpad_.as<float>(), dpad_, QC, (uint32_t)dim_, dk_,
                                   rg_qlow_.as<uint16_t>(), rg_qinex_.as<uint32_t>(), g, dr);
                else
                    launch_to_fp8(rg_qpad_.as<float>(), dpad_, QC, (uint32_t)dim_, dk_,
                                  rg_qlow_.as<uint8_t>(), rg_qscale_.as<float>(), g, dr);
            }
            const uint32_t *s_q = nullptr, *pre = nullptr;
            const float *key_s = nullptr;
            if (filter) {
                if (fp.m <= 32)  // from the minima of the sample keys in fours
                    launch_score_min4(bf16 ? 2 : 1, rg_qlow_.get(), QC, bsamp_.get(), fp.ns,
                                      row_bytes, rg_qscale_.as<float>(), bs_scale_.as<float>(),
                                      bs_norm2_.as<float>(), bs_inv_.as<float>(), metric_,
                                      rg_D_.as<float>(), ldk, g, dr, group_kib_);
                else
                    launch_score_gemm(bf16 ? 2 : 1, rg_qlow_.get(), QC, bsamp_.get(), fp.ns,
                                      row_bytes, rg_qscale_.as<float>(), bs_scale_.as<float>(),
                                      bs_norm2_.as<float>(), bs_inv_.as<float>(), metric_,
                                      rg_D_.as<float>(), ldk, g, dr, group_kib_);
                const float *tau;
                uint32_t tau_ld;
                if (fp.m <= 32) {
                    launch_kth_smallest(rg_D_.as<float>(), ldk, QC, fp.ns / 4, fp.m,
                                        rg_seld_.as<float>(), g, dr);
                    tau = rg_seld_.as<float>();
                    tau_ld = 1;
                } else {
                    launch_row_select(rg_D_.as<float>(), ldk, QC, fp.ns, fp.m,
                                      rg_seld_.as<float>(), rg_seli_.as<uint32_t>(), fp.m,
                                      rg_sf_.as<uint32_t>(), g, dr);
                    tau = rg_seld_.as<float>() + (fp.m - 1);
                    tau_ld = fp.m;
                }
                launch_fill_u32(rg_lcnt_.as<uint32_t>(), QC, 0, g);
                launch_score_filter(bf16 ? 2 : 1, rg_qlow_.get(), QC, blow_.get(), n, row_bytes,
                                    rg_qscale_.as<float>(), bscale_.as<float>(),
                                    bnorm_.as<float>(), inv_bnorm_.as<float>(), metric_, tau,
                                    tau_ld, rg_list_.as<uint64_t>(), fp.cap,
                                    rg_lcnt_.as<uint32_t>(), rg_fscr_.get(), rg_fscr_.bytes(), g,
                                    dr, group_kib_);
                launch_list_prepare(rg_list_.as<uint64_t>(), fp.cap, rg_lcnt_.as<uint32_t>(), tau,
                                    tau_ld, QC, S, rg_ci_.as<uint32_t>(), rg_cd_.as<float>(), S,
                                    rg_sq_.as<uint32_t>(), rg_ksq_.as<float>(),
                                    rg_pref_.as<uint32_t>(), g, dr,
                                    shortlist_ ? nullptr : rg_ck_.as<uint32_t>());
                s_q = rg_sq_.as<uint32_t>();
                key_s = rg_ksq_.as<float>();
                pre = rg_pref_.as<uint32_t>();
            } else {
                launch_score_gemm(bf16 ? 2 : 1, rg_qlow_.get(), QC, blow_.get(), n, row_bytes,
                                  rg_qscale_.as<float>(), bscale_.as<float>(), bnorm_.as<float>(),
                                  inv_bnorm_.as<float>(), metric_, rg_D_.as<float>(), ldk, g, dr,
                                  group_kib_);
                launch_row_select(rg_D_.as<float>(), ldk, QC, n, S, rg_cd_.as<float>(),
                                  rg_ci_.as<uint32_t>(), S, rg_sf_.as<uint32_t>(), g, dr);
            }
            launch_rerank(metric_, rg_qpad_.as<float>(), dpad_, base_.as<float>(),
                          rg_ci_.as<uint32_t>(), rg_cd_.as<float>(), S, S, rg_qnorm_.as<float>(),
                          bnorm_.as<float>(), QC, n, kmax, bmax_, gamma, c_exact, c_bf16,
                          rg_qinex_.as<uint32_t>(), bf16 ? certify_ : 0, rg_oi_.as<uint64_t>(),
                          rg_od_.as<float>(), rg_oc_.as<uint32_t>(), rg_rf_.as<uint32_t>(), g, dr,
                          rg_ck_.as<uint32_t>(), rg_coff_.as<uint64_t>(), s_q, key_s, pre, 0,
                          b8_valid_ ? b8_.as<uint8_t>() : nullptr);
            launch_ragged_chunk_end(call, c, QC, rg_rows_.as<uint32_t>(), rg_off_.as<uint64_t>(),
                                    rg_oi_.as<uint64_t>(), rg_od_.as<float>(),
                                    rg_oc_.as<uint32_t>(), rg_rf_.as<uint32_t>(),
                                    rg_flags_all_.as<uint32_t>(), g);
        }
        // the rows that are only tied: their heap replay over the exact row,
        // each with its own k (the heap's capacity)
        uint32_t *pl = rg_plan_.as<uint32_t>();
        launch_rerun_plan(rg_flags_all_.as<uint32_t>(), (uint32_t)nq_all, pl, g);
        if (dev_rerun) {
            const uint32_t *cnt = pl + 2 * RERUN_SLOTS;
            launch_ragged_rerun_pad(call, pl, (uint32_t)dim_, dpad_, rg_r8p_.as<float>(), g);
            if (need_norm())
                launch_self_dot(rg_r8p_.as<float>(), dpad_, RERUN_SLOTS, (uint32_t)dim_,
                                rg_r8n_.as<float>(), g, cnt);
            launch_exact_rows_dev(metric_, rg_r8p_.as<float>(), dpad_, base_.as<float>(), dpad_, n,
                                  dpad_, rg_r8n_.as<float>(), bnorm_.as<float>(),
                                  rg_r8D_.as<float>(), ldd, cnt, g,
                                  b8_valid_ ? b8_.as<uint8_t>() : nullptr);
            launch_bf_finalize(rg_cd_.as<float>(), rg_ci_.as<uint32_t>(), 1, pl + RERUN_SLOTS,
                               rg_r8D_.as<float>(), ldd, RERUN_SLOTS, n, kmax,
                               rg_r8i_.as<uint64_t>(), rg_r8d_.as<float>(),
                               rg_r8c_.as<uint32_t>(), err, g, rg_kq_.as<uint32_t>(), pl);
            launch_ragged_rerun_scatter(call, pl, rg_r8i_.as<uint64_t>(), rg_r8d_.as<float>(),
                                        rg_r8c_.as<uint32_t>(), kmax, rg_kq_.as<uint32_t>(),
                                        rg_off_.as<uint64_t>(), g);
        }
        hipGraph_t graph = nullptr;
        HM_HIP(hipStreamEndCapture(g, &graph));
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HM_HIP(e);
        rg_graphs_.push_back({nch, kmax, S, fp.ns, mode_, metric_, n_, data, exec});
        return exec;
    }

    // Exact path for the listed queries with kmax, packed into the ragged
    // layout (rg_kq_ / rg_off_ already on the device).
    void ragged_exact(const float *dq, const std::vector<uint32_t> &rows, uint32_t kmax,
                      uint64_t *d_idx, float *d_dist, uint32_t *d_count, hipStream_t s,
                      bool all_tied = false)
    {
        const uint32_t m = (uint32_t)rows.size();
        if (m == 0) return;
        if (kmax == 0 || n_ == 0) {
            for (uint32_t r : rows) HM_HIP(hipMemsetAsync(d_count + r, 0, 4, s));
            HM_HIP(hipStreamSynchronize(s));
            return;
        }
        redo_rows_.reserve(m * 4);
        redo_q_.reserve((size_t)m * dim_ * 4);
        redo_i_.reserve((size_t)m * kmax * 8);
        redo_d_.reserve((size_t)m * kmax * 4);
        redo_c_.reserve((size_t)m * 4);
        HM_HIP(hipMemcpyAsync(redo_rows_.get(), rows.data(), m * 4, hipMemcpyHostToDevice, s));
        launch_gather_rows(dq, dim_, redo_rows_.as<uint32_t>(), m, (uint32_t)dim_,
                           redo_q_.as<float>(), dim_, s);
        search_exact(redo_q_.as<float>(), dim_, m, kmax, redo_i_.as<uint64_t>(),
                     redo_d_.as<float>(), redo_c_.as<uint32_t>(), s, all_tied);
        launch_scatter_ragged(redo_i_.as<uint64_t>(), redo_d_.as<float>(), redo_c_.as<uint32_t>(),
                              redo_rows_.as<uint32_t>(), m, kmax, rg_kq_.as<uint32_t>(),
                              rg_off_.as<uint64_t>(), d_idx, d_dist, d_count, s);
        HM_HIP(hipStreamSynchronize(s));
    }

    // Derived base data of the shortlist forms (rebuilt after add/mode change).
    void ensure_shortlist_data(hipStream_t s)
    {
        ensure_bnorm(s);
        if (sl_valid_) return;
        const uint32_t n = (uint32_t)n_;
        inv_bnorm_.reserve(n_ * sizeof(float));
        stat_.reserve(3 * sizeof(uint32_t));
        b8_stat_.reserve(sizeof(uint32_t));
        HM_HIP(hipMemsetAsync(stat_.get(), 0, 3 * sizeof(uint32_t), s));
        launch_norm_stats(bnorm_.as<float>(), n, inv_bnorm_.as<float>(), stat_.as<uint32_t>(), s);
        if (mode_ == HM_SEARCH_BF16_RERANK) {
            dk_ = (uint32_t)round_up(dim_, 64);  // 128-byte rows
            blow_.reserve(n_ * dk_ * 2);
            launch_to_bf16(base_.as<float>(), dpad_, n_, (uint32_t)dim_, dk_, blow_.as<uint16_t>(),
                           stat_.as<uint32_t>() + 1, s);
        } else {
            dk_ = (uint32_t)round_up(dim_, 128);  // 128-byte rows
            blow_.reserve(n_ * dk_);
            bscale_.reserve(n_ * sizeof(float));
            launch_to_fp8(base_.as<float>(), dpad_, n, (uint32_t)dim_, dk_, blow_.as<uint8_t>(),
                          bscale_.as<float>(), s);
        }
        // a u8 copy of the base when every value is one (integer data such
        // as 0..255 pixels): the rerank then gathers a quarter of the bytes
        // and converts them back to the same f32 values
        b8_.reserve(n_ * dpad_ + 16);
        HM_HIP(hipMemsetAsync(b8_stat_.get(), 0, sizeof(uint32_t), s));
        launch_to_u8_exact(base_.as<float>(), dpad_, n_, b8_.as<uint8_t>(),
                           b8_stat_.as<uint32_t>(), s);
        uint32_t st[4];
        HM_HIP(hipMemcpyAsync(st, stat_.get(), 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HM_HIP(hipMemcpyAsync(st + 3, b8_stat_.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        float bm;
        std::memcpy(&bm, &st[0], 4);
        bmax_ = bm;
        base_inexact_ = st[1] != 0;
        zero_norm_ = st[2] != 0;
        b8_valid_ = st[3] == 0;
        if (!b8_valid_) b8_.release();
        sl_valid_ = true;
    }

    uint32_t shortlist_size(uint32_t k) const
    {
        uint64_t S = shortlist_ ? shortlist_ : std::max<uint64_t>(64, 4ull * (k + 1));
        S = std::max<uint64_t>(S, (uint64_t)k + 1);
        S = std::min<uint64_t>(S, SELECT_MAX_K);
        return (uint32_t)std::min<uint64_t>(S, n_);
    }

    // Filter form of the shortlist: the S-deep shortlist is cut from a
    // candidate list of every row whose key is below a per-query threshold
    // tau, so no [queries x n] key matrix is written or scanned.  tau is the
    // m-th smallest key over a strided sample of ns base rows, sized so that
    // about T = min(2S, cap / 2) rows fall below it (m = clamp(T / 128, 8,
    // 32) sample rows, a relative spread of about 1/sqrt(m) in the list
    // length).  A list shorter than k+1 or
    // longer than its capacity cannot be certified and its query reruns on
    // the exact path, so the estimate affects speed only.
    struct FilterPlan {
        uint32_t ns = 0, m = 0, cap = 0;
    };
    FilterPlan filter_plan(uint32_t S) const
    {
        FilterPlan p;
        if (!use_filter_ || n_ == 0) return p;
        const uint32_t cap =
            (uint32_t)std::min<uint64_t>(8192, std::max<uint64_t>(1024, round_up(8ull * S, 256)));
        // about 2S hits, but at most half the capacity, so that the spread of
        // the estimate rarely overflows a list (a list needs the top k+1
        // under tau for certification, not S entries)
        const uint64_t T = std::min<uint64_t>(2ull * S, cap / 2);
        // m sample rows below tau: the list length spreads by about 1/sqrt(m);
        // the sample costs m/T of the scan, so a long list affords more
        const uint64_t mt = std::min<uint64_t>(32, std::max<uint64_t>(8, T / 128));
        uint64_t ns = round_up((mt * n_ + T - 1) / T, 128);
        ns = std::max<uint64_t>(ns, 1024);
        if (ns * 4 > n_) return p;  // a sample this large does not pay
        p.ns = (uint32_t)ns;
        // at most 32, so that the threshold can come from the minima of the
        // sample keys in fours
        p.m = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((T * ns + n_ - 1) / n_, 1),
                                           std::min<uint64_t>(ns, 32));
        p.cap = cap;
        return p;
    }

    void ensure_sample(uint32_t ns, hipStream_t s)
    {
        if (samp_ns_ == ns) return;
        const uint32_t n = (uint32_t)n_;
        const uint32_t row_bytes = mode_ == HM_SEARCH_BF16_RERANK ? dk_ * 2 : dk_;
        bsamp_.reserve((size_t)ns * row_bytes);
        bs_norm2_.reserve((size_t)ns * 4);
        bs_inv_.reserve((size_t)ns * 4);
        launch_sample_rows(blow_.get(), row_bytes, n, ns, bsamp_.get(), s);
        launch_sample_f32(bnorm_.as<float>(), n, ns, bs_norm2_.as<float>(), s);
        launch_sample_f32(inv_bnorm_.as<float>(), n, ns, bs_inv_.as<float>(), s);
        if (mode_ == HM_SEARCH_FP8_RERANK) {
            bs_scale_.reserve((size_t)ns * 4);
            launch_sample_f32(bscale_.as<float>(), n, ns, bs_scale_.as<float>(), s);
        }
        samp_ns_ = ns;
    }

    // Largest query chunk whose score grid fits the launch limit.
    uint64_t grid_rows_limit(uint64_t ncols) const
    {
        const uint64_t tx = (ncols + 127) / 128;
        return std::max<uint64_t>(1, 0x7FFFFFull / tx) * 128;
    }

    // MFMA shortlist (BF16 or FP8) + exact rerank; rows that are tied or not
    // certified are rerun on the exact path.
    // Returns true when no row needed the exact path: the error word was
    // then read (into herr_) together with the row flags.
    bool search_shortlist(const float *dq, uint64_t nq, uint32_t k, uint64_t *d_idx,
                          float *d_dist, uint32_t *d_count, hipStream_t s)
    {
        ensure_shortlist_data(s);
        const uint32_t n = (uint32_t)n_;
        const uint32_t S = shortlist_size(k);
        const bool bf16 = mode_ == HM_SEARCH_BF16_RERANK;
        const uint32_t row_bytes = bf16 ? dk_ * 2 : dk_;
        const FilterPlan fp = filter_plan(S);
        const bool filter = fp.ns != 0;
        if (filter) ensure_sample(fp.ns, s);
        const uint64_t ldk = round_up(filter ? fp.ns : n, 4);
        // per query: its key row, and (filter form) its list and scan records
        const uint64_t per_q = ldk * sizeof(float) + (filter ? (uint64_t)fp.cap * (8 + 24) : 0);
        uint64_t qc = std::max<uint64_t>(1, d_budget_ / per_q);
        qc = std::min<uint64_t>(qc, nq);
        qc = std::min<uint64_t>(qc, 65535ull * 128);
        qc = std::min<uint64_t>(qc, grid_rows_limit(n));
        qpad_.reserve(qc * dpad_ * sizeof(float));
        qnorm_.reserve(qc * sizeof(float));
        qlow_.reserve(qc * row_bytes);
        qscale_.reserve(qc * sizeof(float));
        D_.reserve(qc * ldk * sizeof(float));
        cand_d_.reserve(qc * S * sizeof(float));
        cand_i_.reserve(qc * S * sizeof(uint32_t));
        flags_.reserve(qc * sizeof(uint32_t));
        rflags_.reserve(nq * sizeof(uint32_t));
        qinexact_.reserve(sizeof(uint32_t));
        if (filter) {
            sel_d_.reserve(qc * fp.m * sizeof(float));
            sel_i_.reserve(qc * fp.m * sizeof(uint32_t));
            sel_f_.reserve(qc * sizeof(uint32_t));
            list_.reserve(qc * fp.cap * sizeof(uint64_t));
            lcnt_.reserve(qc * sizeof(uint32_t));
            fscr_.reserve(score_filter_scratch_bytes((uint32_t)qc, fp.cap));
            sq_.reserve(qc * sizeof(uint32_t));
            ksq_.reserve(qc * sizeof(float));
            pref_.reserve(qc * sizeof(uint32_t));
        }
        // the queries' BF16-inexact flag: read only by the certification (BF16 form)
        if (bf16) HM_HIP(hipMemsetAsync(qinexact_.get(), 0, sizeof(uint32_t), s));

        const float u = 5.9604644775390625e-8f;
        const float gamma = 2.0f * (float)(std::max<uint32_t>(dk_, dpad_) + 8) * u;
        const float c_bf16 = 1.0f / 128.0f + 1.0f / 65536.0f;
        const float c_exact = base_inexact_ ? c_bf16 : 0.0f;
        // Rows that fit the one-pass preparation are read once; when they
        // are already dpad floats apart and aligned, the rerank reads them
        // in place instead of a padded copy.
        const bool prep1 = query_prep_ok(dq, dim_, (uint32_t)dim_, qpad_.as<float>(), dpad_);
        const bool in_place = prep1 && dim_ == dpad_;
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            const float *qp = in_place ? dq + q0 * dim_ : qpad_.as<float>();
            if (prep1) {
                launch_query_prep(bf16 ? 2 : 1, dq + q0 * dim_, dim_, rows, (uint32_t)dim_, dk_,
                                  qlow_.get(), qscale_.as<float>(), qinexact_.as<uint32_t>(),
                                  qnorm_.as<float>(), in_place ? nullptr : qpad_.as<float>(),
                                  dpad_, err_.as<uint32_t>(), s);
            } else {
                launch_pad_rows(dq + q0 * dim_, dim_, qpad_.as<float>(), dpad_, rows,
                                (uint32_t)dim_, dpad_, s, nullptr, err_.as<uint32_t>());
                launch_self_dot(qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                                qnorm_.as<float>(), s);
                if (bf16)
                    launch_to_bf16(qpad_.as<float>(), dpad_, rows, (uint32_t)dim_, dk_,
                                   qlow_.as<uint16_t>(), qinexact_.as<uint32_t>(), s);
                else
                    launch_to_fp8(qpad_.as<float>(), dpad_, rows, (uint32_t)dim_, dk_,
                                  qlow_.as<uint8_t>(), qscale_.as<float>(), s);
            }
            const uint32_t *s_q = nullptr, *pre = nullptr;
            const float *key_s = nullptr;
            if (filter) {
                // threshold from the sample, then the filtered full scan
                // (m <= 32: from the minima of the sample keys in fours,
                // a quarter of the key matrix; see launch_score_min4)
                if (fp.m <= 32)
                    launch_score_min4(bf16 ? 2 : 1, qlow_.get(), rows, bsamp_.get(), fp.ns,
                                      row_bytes, qscale_.as<float>(), bs_scale_.as<float>(),
                                      bs_norm2_.as<float>(), bs_inv_.as<float>(), metric_,
                                      D_.as<float>(), ldk, s, nullptr, group_kib_);
                else
                    launch_score_gemm(bf16 ? 2 : 1, qlow_.get(), rows, bsamp_.get(), fp.ns,
                                      row_bytes, qscale_.as<float>(), bs_scale_.as<float>(),
                                      bs_norm2_.as<float>(), bs_inv_.as<float>(), metric_,
                                      D_.as<float>(), ldk, s, nullptr, group_kib_);
                const float *tau;
                uint32_t tau_ld;
                if (fp.m <= 32) {  // the select also zeroes the list counters
                    launch_kth_smallest(D_.as<float>(), ldk, rows, fp.ns / 4, fp.m,
                                        sel_d_.as<float>(), s, nullptr, lcnt_.as<uint32_t>());
                    tau = sel_d_.as<float>();
                    tau_ld = 1;
                } else {
                    launch_row_select(D_.as<float>(), ldk, rows, fp.ns, fp.m, sel_d_.as<float>(),
                                      sel_i_.as<uint32_t>(), fp.m, sel_f_.as<uint32_t>(), s);
                    tau = sel_d_.as<float>() + (fp.m - 1);
                    tau_ld = fp.m;
                    launch_fill_u32(lcnt_.as<uint32_t>(), rows, 0, s);
                }
                launch_score_filter(bf16 ? 2 : 1, qlow_.get(), rows, blow_.get(), n, row_bytes,
                                    qscale_.as<float>(), bscale_.as<float>(), bnorm_.as<float>(),
                                    inv_bnorm_.as<float>(), metric_, tau, tau_ld,
                                    list_.as<uint64_t>(), fp.cap, lcnt_.as<uint32_t>(),
                                    fscr_.get(), fscr_.bytes(), s, nullptr, group_kib_);
                launch_list_prepare(list_.as<uint64_t>(), fp.cap, lcnt_.as<uint32_t>(), tau, tau_ld,
                                    rows, S, cand_i_.as<uint32_t>(), cand_d_.as<float>(), S,
                                    sq_.as<uint32_t>(), ksq_.as<float>(), pref_.as<uint32_t>(), s);
                s_q = sq_.as<uint32_t>();
                key_s = ksq_.as<float>();
                pre = pref_.as<uint32_t>();
            } else {
                launch_score_gemm(bf16 ? 2 : 1, qlow_.get(), rows, blow_.get(), n, row_bytes,
                                  qscale_.as<float>(), bscale_.as<float>(), bnorm_.as<float>(),
                                  inv_bnorm_.as<float>(), metric_, D_.as<float>(), ldk, s, nullptr,
                                  group_kib_);
                launch_row_select(D_.as<float>(), ldk, rows, n, S, cand_d_.as<float>(),
                                  cand_i_.as<uint32_t>(), S, flags_.as<uint32_t>(), s);
            }
            launch_rerank(metric_, qp, dpad_, base_.as<float>(),
                          cand_i_.as<uint32_t>(), cand_d_.as<float>(), S, S, qnorm_.as<float>(),
                          bnorm_.as<float>(), rows, n, k, bmax_, gamma, c_exact, c_bf16,
                          qinexact_.as<uint32_t>(), bf16 ? certify_ : 0, d_idx + q0 * k,
                          d_dist + q0 * k, d_count + q0, rflags_.as<uint32_t>() + q0, s,
                          nullptr, nullptr, nullptr, s_q, key_s, pre, 0,
                          b8_valid_ ? b8_.as<uint8_t>() : nullptr);
        }
        // Rows needing the exact path (ties: heap replay; uncertified).  Up
        // to RERUN_SLOTS rows that are only tied are rerun on the device
        // without a host round trip: a plan kernel lists them and the exact
        // row and heap replay kernels take their count from device memory.
        // The flags, the plan and the error word then come back in one
        // pinned copy; the host reruns the rows only when the plan left them.
        const bool dev_rerun = exact_rows_dev_ok(dpad_, dpad_, base_.as<float>());
        if (dev_rerun) {
            const uint64_t ldd = round_up(n, 4);
            plan_.reserve(RERUN_PLAN_WORDS * sizeof(uint32_t));
            r8p_.reserve((size_t)RERUN_SLOTS * dpad_ * sizeof(float));
            r8n_.reserve(RERUN_SLOTS * sizeof(float));
            r8D_.reserve((size_t)RERUN_SLOTS * ldd * sizeof(float));
            r8i_.reserve((size_t)RERUN_SLOTS * k * sizeof(uint64_t));
            r8d_.reserve((size_t)RERUN_SLOTS * k * sizeof(float));
            r8c_.reserve(RERUN_SLOTS * sizeof(uint32_t));
            uint32_t *pl = plan_.as<uint32_t>();
            const uint32_t *cnt = pl + 2 * RERUN_SLOTS;
            launch_rerun_plan(rflags_.as<uint32_t>(), (uint32_t)nq, pl, s);
            launch_pad_rows(dq, dim_, r8p_.as<float>(), dpad_, RERUN_SLOTS, (uint32_t)dim_, dpad_,
                            s, cnt, nullptr, pl);
            if (need_norm())
                launch_self_dot(r8p_.as<float>(), dpad_, RERUN_SLOTS, (uint32_t)dim_,
                                r8n_.as<float>(), s, cnt);
            launch_exact_rows_dev(metric_, r8p_.as<float>(), dpad_, base_.as<float>(), dpad_, n,
                                  dpad_, r8n_.as<float>(), bnorm_.as<float>(), r8D_.as<float>(),
                                  ldd, cnt, s, b8_valid_ ? b8_.as<uint8_t>() : nullptr);
            launch_bf_finalize(cand_d_.as<float>(), cand_i_.as<uint32_t>(), 1, pl + RERUN_SLOTS,
                               r8D_.as<float>(), ldd, RERUN_SLOTS, n, k, r8i_.as<uint64_t>(),
                               r8d_.as<float>(), r8c_.as<uint32_t>(), err_.as<uint32_t>(), s);
            launch_scatter_results(r8i_.as<uint64_t>(), r8d_.as<float>(), r8c_.as<uint32_t>(), pl,
                                   RERUN_SLOTS, k, d_idx, d_dist, d_count, s);
        }
        hflags_.reserve(nq * sizeof(uint32_t));
        hrows_.reserve(nq * sizeof(uint32_t));
        hplan_.reserve(RERUN_PLAN_WORDS * sizeof(uint32_t));
        HM_HIP(hipMemcpyAsync(herr_.as<uint32_t>(), err_.get(), sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
        if (dev_rerun) {
            // the plan and the error word; the flags only if the host must act
            HM_HIP(hipMemcpyAsync(hplan_.as<uint32_t>(), plan_.get(),
                                  RERUN_PLAN_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HM_HIP(hipStreamSynchronize(s));
            const uint32_t *hp = hplan_.as<uint32_t>();
            last_fallback_ = hp[2 * RERUN_SLOTS + 2];
            last_uncertified_ = hp[2 * RERUN_SLOTS + 3];
            if (hp[2 * RERUN_SLOTS + 1] == 0) return true;  // tied rows rerun on the device
        }
        HM_HIP(hipMemcpyAsync(hflags_.as<uint32_t>(), rflags_.get(), nq * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        const uint32_t *fl = hflags_.as<uint32_t>();
        // Certified rows with a tie only need the heap replay over their
        // exact row; the others take the whole exact path.  Both lists go
        // to the device from pinned memory: tied rows from the front, the
        // others from the back.
        uint32_t *rl = hrows_.as<uint32_t>();
        uint32_t nt = 0, nr = 0;
        last_uncertified_ = 0;
        for (uint64_t i = 0; i < nq; ++i) {
            if (fl[i] == ROW_TIE)
                rl[nt++] = (uint32_t)i;
            else if (fl[i] & (ROW_TIE | ROW_NAN | ROW_UNCERT | ROW_BAD))
                rl[nq - 1 - nr++] = (uint32_t)i;
            last_uncertified_ += (fl[i] & ROW_UNCERT) ? 1 : 0;
        }
        last_fallback_ = nt + nr;
        if (nt) rerun_rows(dq, rl, nt, k, d_idx, d_dist, d_count, true, s);
        if (nr) rerun_rows(dq, rl + (nq - nr), nr, k, d_idx, d_dist, d_count, false, s);
        return nt + nr == 0;
    }

    // Exact path for the listed query rows (m entries of `rows`, in pinned
    // host memory that stays unchanged until the stream has run the copy);
    // results scattered back.  No synchronisation: the caller reads the
    // error word after the last launch.
    void rerun_rows(const float *dq, const uint32_t *rows, uint32_t m, uint32_t k,
                    uint64_t *d_idx, float *d_dist, uint32_t *d_count, bool all_tied,
                    hipStream_t s)
    {
        redo_rows_.reserve(m * sizeof(uint32_t));
        redo_q_.reserve((size_t)m * dim_ * sizeof(float));
        redo_i_.reserve((size_t)m * k * sizeof(uint64_t));
        redo_d_.reserve((size_t)m * k * sizeof(float));
        redo_c_.reserve((size_t)m * sizeof(uint32_t));
        HM_HIP(hipMemcpyAsync(redo_rows_.get(), rows, m * sizeof(uint32_t), hipMemcpyHostToDevice,
                              s));
        launch_gather_rows(dq, dim_, redo_rows_.as<uint32_t>(), m, (uint32_t)dim_,
                           redo_q_.as<float>(), dim_, s);
        search_exact(redo_q_.as<float>(), dim_, m, k, redo_i_.as<uint64_t>(), redo_d_.as<float>(),
                     redo_c_.as<uint32_t>(), s, all_tied);
        launch_scatter_results(redo_i_.as<uint64_t>(), redo_d_.as<float>(),
                               redo_c_.as<uint32_t>(), redo_rows_.as<uint32_t>(), m, k, d_idx,
                               d_dist, d_count, s);
    }

  public:
    // Ragged batches (batch b = queries [off[b], off[b+1]) with its own k):
    // the shortlist pipeline for a chunk of up to RAG_CHUNK queries is one
    // captured hipGraph whose kernels read the chunk's row count, per-query
    // k and output offsets from device memory, so one graph serves every
    // chunk and every call.  Results are packed batch after batch, k_b per
    // query.  Tied / uncertified rows are rerun on the exact path.
    void search_ragged_device(const float *dq, uint32_t nb, const uint32_t *h_off,
                              const uint32_t *h_k, uint64_t *d_idx, float *d_dist,
                              uint32_t *d_count, hipStream_t s) override
    {
        if (!s) s = stream_;
        if (metric_ == HM_METRIC_UNKNOWN) fail(HM_ERR_UNKNOWN_METRIC, "unknown method");
        if (nb == 0) return;
        if (h_off[0] != 0) fail(HM_ERR_INVALID_ARG, "batch_offsets[0] must be 0");
        uint64_t total = 0;
        uint32_t kmax = 0;
        for (uint32_t b = 0; b < nb; ++b) {
            if (h_off[b + 1] < h_off[b]) fail(HM_ERR_INVALID_ARG, "batch offsets not ascending");
            total += (uint64_t)(h_off[b + 1] - h_off[b]) * h_k[b];
            kmax = std::max(kmax, h_k[b]);
        }
        const uint64_t Q = h_off[nb];
        if (Q == 0) return;
        if (!dq || !d_count || (total && (!d_idx || !d_dist)))
            fail(HM_ERR_INVALID_ARG, "null pointer");
        if (Q > 0x7FFFFFFFull) fail(HM_ERR_INVALID_ARG, "too many queries");
        err_.reserve(4);
        HM_HIP(hipMemsetAsync(err_.get(), 0, 4, s));
        launch_check_finite(dq, dim_, Q, (uint32_t)dim_, err_.as<uint32_t>(), s);
        const bool shortlist_form = kmax > 0 && shortlist_usable(s);
        if (!shortlist_form) {
            // exact form: each batch is an exact search with its own k,
            // written where the packed layout puts it
            uint64_t o = 0;
            last_fallback_ = last_uncertified_ = 0;
            for (uint32_t b = 0; b < nb; ++b) {
                const uint64_t q0 = h_off[b], m = h_off[b + 1] - q0;
                if (m == 0) continue;
                if (h_k[b] == 0 || n_ == 0)
                    launch_fill_results(d_idx + o, d_dist + o, d_count + q0, m, h_k[b], s);
                else
                    search_exact(dq + q0 * dim_, dim_, m, h_k[b], d_idx + o, d_dist + o,
                                 d_count + q0, s);
                o += m * h_k[b];
            }
            read_error(s);
            return;
        }
        std::vector<uint32_t> kq(Q);
        std::vector<uint64_t> off(Q + 1);
        for (uint32_t b = 0; b < nb; ++b)
            for (uint64_t q = h_off[b]; q < h_off[b + 1]; ++q) kq[q] = h_k[b];
        off[0] = 0;
        for (uint64_t q = 0; q < Q; ++q) off[q + 1] = off[q] + kq[q];
        ensure_shortlist_data(s);
        const uint32_t S = shortlist_size(kmax);
        const uint32_t QC = RAG_CHUNK;
        const uint32_t nch = (uint32_t)((Q + QC - 1) / QC);
        const bool dev_rerun = exact_rows_dev_ok(dpad_, dpad_, base_.as<float>());
        hipGraphExec_t exec = ragged_graph(nch, kmax, S, dev_rerun);
        // the call's inputs, staged once: per-query k and output offsets,
        // and the pointers the graph reads
        rg_host_.reserve(sizeof(RaggedCall) + (RERUN_PLAN_WORDS + 1) * 4);
        RaggedCall *hc = rg_host_.as<RaggedCall>();
        *hc = RaggedCall{dq, d_idx, d_dist, d_count, Q};
        HM_HIP(hipMemcpyAsync(rg_kq_.get(), kq.data(), Q * 4, hipMemcpyHostToDevice, s));
        HM_HIP(hipMemcpyAsync(rg_off_.get(), off.data(), (Q + 1) * 8, hipMemcpyHostToDevice, s));
        HM_HIP(hipMemcpyAsync(rg_call_.get(), hc, sizeof(RaggedCall), hipMemcpyHostToDevice, s));
        HM_HIP(hipGraphLaunch(exec, s));
        ++graph_launches_;
        uint32_t *hp = reinterpret_cast<uint32_t *>(rg_host_.as<unsigned char>() + sizeof(RaggedCall));
        HM_HIP(hipMemcpyAsync(hp, rg_plan_.get(), RERUN_PLAN_WORDS * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipMemcpyAsync(hp + RERUN_PLAN_WORDS, rg_err_.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        HM_HIP(hipMemcpyAsync(err_.get(), rg_err_.get(), 4, hipMemcpyDeviceToDevice, s));
        last_fallback_ = hp[2 * RERUN_SLOTS + 2];
        last_uncertified_ = hp[2 * RERUN_SLOTS + 3];
        const bool host_acts = hp[2 * RERUN_SLOTS + 1] != 0 || (!dev_rerun && last_fallback_ > 0);
        if (!host_acts) {
            read_error(s);
            return;
        }
        // rows the device plan left (more than RERUN_SLOTS tied rows, or
        // uncertified / NaN rows): the host reruns them
        std::vector<uint32_t> fl(Q);
        HM_HIP(hipMemcpyAsync(fl.data(), rg_flags_all_.get(), Q * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        std::vector<uint32_t> redo, tied;
        last_uncertified_ = 0;
        for (uint64_t q = 0; q < Q; ++q) {
            if (fl[q] == ROW_TIE)
                tied.push_back((uint32_t)q);
            else if (fl[q] & (ROW_TIE | ROW_NAN | ROW_UNCERT | ROW_BAD))
                redo.push_back((uint32_t)q);
            last_uncertified_ += (fl[q] & ROW_UNCERT) ? 1 : 0;
        }
        last_fallback_ = redo.size() + tied.size();
        // a tied row's result depends on its own k (the heap's capacity),
        // so the rows rerun in groups of equal k
        auto rerun_by_k = [&](const std::vector<uint32_t> &rows, bool all_tied) {
            std::vector<uint32_t> order(rows);
            std::stable_sort(order.begin(), order.end(),
                             [&](uint32_t a, uint32_t b) { return kq[a] < kq[b]; });
            for (size_t i = 0; i < order.size();) {
                size_t j = i;
                while (j < order.size() && kq[order[j]] == kq[order[i]]) ++j;
                ragged_exact(dq, std::vector<uint32_t>(order.begin() + i, order.begin() + j),
                             kq[order[i]], d_idx, d_dist, d_count, s, all_tied);
                i = j;
            }
        };
        if (!tied.empty()) rerun_by_k(tied, true);
        if (!redo.empty()) rerun_by_k(redo, false);
        read_error(s);
    }

    // The error word of the last search: non-finite input, a NaN distance,
    // an internal consistency failure.
    void read_error(hipStream_t s)
    {
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, err_.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
        if (err & 4u) fail(HM_ERR_HIP, "internal: selection consistency check failed");
        if (err & 1u) fail(HM_ERR_NAN_DISTANCE, "NaN distance (partial_cmp of NaN)");
    }

    // Picks the shortlist depth that minimises the search time of a query
    // sample (k fixed), trying 64..4096; returns the choice.
    uint32_t autotune(const float *h_queries, uint64_t nq, uint32_t k) override
    {
        if (mode_ == HM_SEARCH_EXACT || !shortlist_metric() || n_ == 0 || nq == 0) return 0;
        hipStream_t s = stream_;
        DeviceBuffer q, oi, od, oc;
        q.reserve(nq * dim_ * 4);
        oi.reserve(nq * std::max(k, 1u) * 8);
        od.reserve(nq * std::max(k, 1u) * 4);
        oc.reserve(nq * 4);
        HM_HIP(hipMemcpyAsync(q.get(), h_queries, nq * dim_ * 4, hipMemcpyHostToDevice, s));
        // events destroyed and the caller's depth restored on every exit,
        // also when a trial search throws (NaN distance, out of memory)
        struct Sweep {
            uint32_t &depth;
            const uint32_t saved;
            uint32_t &group;
            const uint32_t saved_group;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            bool done = false;
            ~Sweep()
            {
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                if (!done) {
                    depth = saved;
                    group = saved_group;
                }
            }
        } sw{shortlist_, shortlist_, group_kib_, group_kib_};
        HM_HIP(hipEventCreate(&sw.e0));
        HM_HIP(hipEventCreate(&sw.e1));
        hipEvent_t e0 = sw.e0, e1 = sw.e1;
        float best_ms = 1e30f;
        uint32_t best = 0;
        const uint32_t saved = shortlist_;
        for (uint32_t S = 64; S <= 4096; S *= 2) {
            if (S < k + 1 || S > n_) continue;
            shortlist_ = S;
            search_device(q.as<float>(), nq, k, oi.as<uint64_t>(), od.as<float>(),
                          oc.as<uint32_t>(), s);  // warm
            HM_HIP(hipEventRecord(e0, s));
            search_device(q.as<float>(), nq, k, oi.as<uint64_t>(), od.as<float>(),
                          oc.as<uint32_t>(), s);
            HM_HIP(hipEventRecord(e1, s));
            HM_HIP(hipEventSynchronize(e1));
            float ms = 0;
            HM_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = S;
            }
        }
        shortlist_ = best ? best : saved;
        // then the score tile order at that depth: base tiles per group as
        // KiB of base rows per XCD's L2 (each timed three times, the best kept)
        if (best) {
            const uint32_t saved_g = group_kib_;
            uint32_t best_g = saved_g;
            float best_g_ms = 1e30f;
            for (uint32_t gk : {512u, 1024u, 2048u, 3072u}) {
                group_kib_ = gk;
                float t = 1e30f;
                for (int rep = 0; rep < 3; ++rep) {
                    HM_HIP(hipEventRecord(e0, s));
                    search_device(q.as<float>(), nq, k, oi.as<uint64_t>(), od.as<float>(),
                                  oc.as<uint32_t>(), s);
                    HM_HIP(hipEventRecord(e1, s));
                    HM_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    HM_HIP(hipEventElapsedTime(&ms, e0, e1));
                    t = std::min(t, ms);
                }
                if (t < best_g_ms) {
                    best_g_ms = t;
                    best_g = gk;
                }
            }
            group_kib_ = best_g;
            if (group_kib_ != saved_g) drop_ragged_graphs();  // captured with the old order
        }
        sw.done = true;
        return shortlist_;
    }

    void tuning(uint32_t *shortlist, uint32_t *group_kib) const override
    {
        *shortlist = shortlist_;
        *group_kib = group_kib_ ? group_kib_ : 2048u;
    }

    void search_stats(uint64_t *fallback, uint64_t *uncertified) const override
    {
        *fallback = last_fallback_;
        *uncertified = last_uncertified_;
    }

  private:
    uint64_t n_ = 0, cap_ = 0;
    uint32_t dpad_ = 0;
    uint32_t group_kib_ = 0;  // score tile order (0: the default 2,048 KiB)
    int32_t metric_ = HM_METRIC_UNKNOWN;
    DeviceBuffer sv_q_, sv_qn_;  // distance service: padded queries, their norms
    int32_t mode_ = HM_SEARCH_EXACT;
    uint32_t shortlist_ = 0;
    size_t d_budget_ = (size_t)4 << 30;
    DeviceBuffer base_, bnorm_;
    bool bnorm_valid_ = false;
    DeviceBuffer qpad_, qnorm_, D_, cand_d_, cand_i_, flags_, err_;
    // shortlist forms
    bool sl_valid_ = false, base_inexact_ = true, zero_norm_ = false;
    bool b8_valid_ = false;              // b8_ holds the base exactly
    DeviceBuffer b8_, b8_stat_;          // u8 copy of the base rows [n][dpad]
    int certify_ = 1;
    uint32_t dk_ = 0;
    float bmax_ = 0;
    DeviceBuffer inv_bnorm_, stat_, blow_, bscale_, qlow_, qscale_, qinexact_, rflags_;
    // filter form: strided base sample (low-precision rows, norms, scales)
    // and the per-chunk threshold / candidate-list buffers
    uint32_t samp_ns_ = 0;
    bool use_filter_ = true;
    DeviceBuffer bsamp_, bs_norm2_, bs_inv_, bs_scale_;
    DeviceBuffer Ds_, sel_d_, sel_i_, sel_f_, list_, lcnt_, fscr_, sq_, ksq_, pref_;
    DeviceBuffer redo_rows_, redo_q_, redo_i_, redo_d_, redo_c_;
    PinnedBuffer hflags_, hrows_, herr_, hplan_;  // row flags, rerun rows, error word, rerun plan
    DeviceBuffer plan_, r8p_, r8n_, r8D_, r8i_, r8d_, r8c_;  // device rerun of tied rows
    uint64_t last_fallback_ = 0, last_uncertified_ = 0;
    // ragged graphs
    hipStream_t rg_stream_ = nullptr;
    std::vector<RaggedGraph> rg_graphs_;
    uint64_t graph_launches_ = 0;
    PinnedBuffer rg_host_;  // the call block and the plan / error read-back
    DeviceBuffer rg_call_, rg_err_, rg_plan_, rg_r8p_, rg_r8n_, rg_r8D_, rg_r8i_, rg_r8d_, rg_r8c_;
    DeviceBuffer rg_rows_, rg_ck_, rg_coff_, rg_oi_, rg_od_, rg_oc_, rg_kq_, rg_off_,
        rg_flags_all_;
    // buffers owned by the captured graph (never reallocated by other paths)
    DeviceBuffer rg_qpad_, rg_qnorm_, rg_qlow_, rg_qscale_, rg_D_, rg_cd_, rg_ci_, rg_sf_, rg_rf_,
        rg_qinex_;
    DeviceBuffer rg_seld_, rg_seli_, rg_list_, rg_lcnt_, rg_fscr_, rg_sq_, rg_ksq_, rg_pref_;
};

}  // namespace

void launch_fill_results(uint64_t *idx, float *dist, uint32_t *count, uint64_t nq, uint32_t k,
                         hipStream_t s)
{
    if (nq == 0) return;
    unsigned blocks = (unsigned)std::min<uint64_t>((nq * std::max<uint32_t>(k, 1) + 255) / 256, 4096);
    fill_results_kernel<<<blocks, 256, 0, s>>>(idx, dist, count, nq, k);
    HM_CHECK_LAUNCH();
}

std::unique_ptr<Index> make_bruteforce(uint64_t dim) { return std::make_unique<BruteForce>(dim); }

}  // namespace hm
