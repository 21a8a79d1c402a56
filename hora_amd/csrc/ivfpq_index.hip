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
        float *dst = base_.as<float>() + n_ * dpad_;
        if (dpad_ != dim_) HM_HIP(hipMemsetAsync(dst, 0, n_rows * dpad_ * 4, stream_));
        HM_HIP(hipMemcpy2DAsync(dst, dpad_ * 4, rows, dim_ * 4, dim_ * 4, n_rows,
                                hipMemcpyHostToDevice, stream_));
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
    const char *name() const override { return "IVFPQIndex"; }
    uint64_t nodes_size() const override { return n_; }

    void clear() override
    {
        base_.release();
        n_ = cap_ = 0;
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
        const uint32_t per_query = nprobe_ * k1;
        if (per_query > 16384) fail(HM_ERR_INVALID_ARG, "nprobe*(k+1) above 16384");
        const uint32_t Kc = std::min<uint32_t>(nprobe_ + 1, nlist_);
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
        cand_.reserve(qc * per_query * 8);
        std::vector<uint32_t> fl(qc);
        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            launch_pad_rows(dq + q0 * dim_, dim_, qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                            dpad_, s);
            if (norms)
                launch_self_dot(qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                                qnorm_.as<float>(), s);
            // pq_idx.rs:416-421: metric(q, c_i) with the build metric
            launch_exact_tile(metric_, qpad_.as<float>(), dpad_, rows, centers_.as<float>(), dpad_,
                              nlist_, dpad_, qnorm_.as<float>(), cnorm_.as<float>(),
                              D_.as<float>(), ldd, s);
            launch_row_select(D_.as<float>(), ldd, rows, nlist_, Kc, probe_d_.as<float>(),
                              probe_i_.as<uint32_t>(), Kc, cflags_.as<uint32_t>(), s);
            launch_ivf_scan(qpad_.as<float>(), dpad_, centers_.as<float>(),
                            probe_i_.as<uint32_t>(), Kc, nprobe_, rows, off_.as<uint32_t>(),
                            ids_.as<uint32_t>(), codes_.as<uint8_t>(), codebooks_.as<float>(),
                            sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_, KS_, sdm_, k1,
                            cand_.as<unsigned long long>(), s);
            launch_ivf_merge(cand_.as<unsigned long long>(), per_query, rows, k,
                             cflags_.as<uint32_t>(), d_idx + q0 * k, d_dist + q0 * k,
                             d_count + q0, mflags_.as<uint32_t>(), s);
            HM_HIP(hipMemcpyAsync(fl.data(), mflags_.get(), rows * 4, hipMemcpyDeviceToHost, s));
            HM_HIP(hipStreamSynchronize(s));
            std::vector<uint32_t> tied;
            for (uint32_t i = 0; i < rows; ++i)
                if (fl[i] & ROW_TIE) tied.push_back(i);
            last_replayed_ += tied.size();
            if (!tied.empty()) replay(tied, rows, k, ldd, d_idx + q0 * k, d_dist + q0 * k,
                                      d_count + q0, s);
        }
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, err_.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
    }

    void search_stats(uint64_t *fallback, uint64_t *uncertified) const override
    {
        *fallback = last_replayed_;
        *uncertified = 0;
    }

  private:
    void replay(const std::vector<uint32_t> &tied, uint32_t rows, uint32_t k, uint64_t ldd,
                uint64_t *d_idx, float *d_dist, uint32_t *d_count, hipStream_t s)
    {
        (void)rows;
        const uint32_t m = (uint32_t)tied.size();
        const size_t per = (size_t)nlist_ + max_list_ + (k + 1) + (size_t)M_ * KS_;
        rp_rows_.reserve(m * 4);
        rp_f_.reserve(m * per * 4);
        rp_u_.reserve(m * per * 4);
        HM_HIP(hipMemcpyAsync(rp_rows_.get(), tied.data(), m * 4, hipMemcpyHostToDevice, s));
        launch_ivf_replay(rp_rows_.as<uint32_t>(), m, D_.as<float>(), (uint32_t)ldd, nlist_,
                          nprobe_, qpad_.as<float>(), dpad_, centers_.as<float>(),
                          off_.as<uint32_t>(), ids_.as<uint32_t>(), codes_.as<uint8_t>(),
                          codebooks_.as<float>(), sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), M_,
                          KS_, sdm_, k, max_list_, rp_f_.as<float>(), rp_u_.as<uint32_t>(), per,
                          d_idx, d_dist, d_count, s);
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
                           centers_.as<float>(), dpad_, s);
        uint64_t rng = host_stream_seed(p_.seed, 0);
        std::vector<float> hc((size_t)nlist_ * dpad_);
        std::vector<uint32_t> hoff(nlist_ + 1);
        std::vector<uint64_t> cnt(nlist_);
        for (uint32_t ep = 0; ep < epochs; ++ep) {
            assign_all(s);
            launch_update(fp64, base_.as<float>(), dpad_, off_.as<uint32_t>(), ids_.as<uint32_t>(),
                          nlist_, (uint32_t)dim_, centers_.as<float>(), dpad_, s);
            if (ep + 1 < epochs) {
                HM_HIP(hipMemcpyAsync(hoff.data(), off_.get(), (nlist_ + 1) * 4,
                                      hipMemcpyDeviceToHost, s));
                HM_HIP(hipMemcpyAsync(hc.data(), centers_.get(), hc.size() * 4,
                                      hipMemcpyDeviceToHost, s));
                HM_HIP(hipStreamSynchronize(s));
                for (uint32_t c = 0; c < nlist_; ++c) cnt[c] = hoff[c + 1] - hoff[c];
                split_centers_host(hc.data(), dpad_, nlist_, (uint32_t)dim_, n, cnt.data(), &rng);
                HM_HIP(hipMemcpyAsync(centers_.get(), hc.data(), hc.size() * 4,
                                      hipMemcpyHostToDevice, s));
            }
        }
        // search_data after train -> _ivflist (pq_idx.rs:369-375)
        assign_all(s);
        HM_HIP(hipMemcpyAsync(hoff.data(), off_.get(), (nlist_ + 1) * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        max_list_ = 0;
        for (uint32_t c = 0; c < nlist_; ++c) max_list_ = std::max(max_list_, hoff[c + 1] - hoff[c]);
        // per-list residual PQ (pq_idx.rs:376-393)
        DeviceBuffer R, asg;
        R.reserve(n * dpad_ * 4);
        launch_residuals(base_.as<float>(), dpad_, ids_.as<uint32_t>(), lists_.as<uint32_t>(),
                         centers_.as<float>(), n, dpad_, R.as<float>(), s);
        codebooks_.reserve((size_t)nlist_ * M_ * KS_ * sdm_ * 4);
        codes_.reserve(n * M_);
        asg.reserve((size_t)M_ * n * 2);
        launch_pq_train(fp64, R.as<float>(), dpad_, off_.as<uint32_t>(), nlist_, M_,
                        sub_b_.as<uint32_t>(), sub_e_.as<uint32_t>(), KS_, sd_max_, sdm_, epochs,
                        p_.seed, max_list_, n, asg.as<uint16_t>(), codebooks_.as<float>(),
                        codes_.as<uint8_t>(), s);
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
    DeviceBuffer rp_rows_, rp_f_, rp_u_;
};

}  // namespace

std::unique_ptr<Index> make_ivfpq(uint64_t dim, const hm_params &p)
{
    return std::make_unique<IVFPQ>(dim, p);
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

int32_t ivfpq_set_nprobe(Index *ix, uint32_t nprobe)
{
    auto *p = dynamic_cast<IVFPQ *>(ix);
    if (!p) fail(HM_ERR_INVALID_ARG, "not an IVFPQIndex");
    p->set_nprobe(nprobe);
    return HM_OK;
}

}  // namespace hm
