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
#include <cmath>
#include <cstring>

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
        float *dst = base_.as<float>() + n_ * dpad_;
        if (dpad_ != dim_) HM_HIP(hipMemsetAsync(dst, 0, n_rows * dpad_ * sizeof(float), stream_));
        HM_HIP(hipMemcpy2DAsync(dst, dpad_ * sizeof(float), rows, dim_ * sizeof(float),
                                dim_ * sizeof(float), n_rows, hipMemcpyHostToDevice, stream_));
        HM_HIP(hipStreamSynchronize(stream_));
        n_ = need;
        bnorm_valid_ = false;
    }

    void build(int32_t metric) override
    {
        if (metric < HM_METRIC_UNKNOWN || metric > HM_METRIC_ANGULAR)
            fail(HM_ERR_INVALID_ARG, "metric out of range");
        metric_ = metric;  // bruteforce_idx.rs:36-39 stores mt and nothing else
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
        if (mode != HM_SEARCH_EXACT)
            fail(HM_ERR_NOT_IMPLEMENTED, "search mode not available in this build");
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
        if (k == 0 || n_ == 0) {
            if (k) {
                unsigned blocks = (unsigned)std::min<uint64_t>((nq * k + 255) / 256, 4096);
                fill_results_kernel<<<blocks, 256, 0, s>>>(d_idx, d_dist, d_count, nq, k);
            } else {
                HM_HIP(hipMemsetAsync(d_count, 0, nq * sizeof(uint32_t), s));
            }
            HM_HIP(hipStreamSynchronize(s));
            return;
        }
        const uint32_t n = (uint32_t)n_;
        const uint32_t K = (uint32_t)std::min<uint64_t>((uint64_t)k + 1, n);
        if (K > SELECT_MAX_K) fail(HM_ERR_INVALID_ARG, "k too large (max 8191)");
        if (nq > 0x7FFFFFFFull) fail(HM_ERR_INVALID_ARG, "too many queries");

        err_.reserve(sizeof(uint32_t));
        HM_HIP(hipMemsetAsync(err_.get(), 0, sizeof(uint32_t), s));
        launch_check_finite(dq, dim_, nq, (uint32_t)dim_, err_.as<uint32_t>(), s);

        const bool need_norm =
            metric_ == HM_METRIC_COSINE_SIMILARITY || metric_ == HM_METRIC_ANGULAR;
        if (need_norm && !bnorm_valid_) {
            bnorm_.reserve(n_ * sizeof(float));
            launch_self_dot(base_.as<float>(), dpad_, n, (uint32_t)dim_, bnorm_.as<float>(), s);
            bnorm_valid_ = true;
        }

        const uint64_t ldd = round_up(n, 4);
        uint64_t qc = std::max<uint64_t>(1, d_budget_ / (ldd * sizeof(float)));
        qc = std::min<uint64_t>(qc, nq);
        qc = std::min<uint64_t>(qc, 65535ull * 128);
        qpad_.reserve(qc * dpad_ * sizeof(float));
        D_.reserve(qc * ldd * sizeof(float));
        cand_d_.reserve(qc * K * sizeof(float));
        cand_i_.reserve(qc * K * sizeof(uint32_t));
        flags_.reserve(qc * sizeof(uint32_t));
        if (need_norm) qnorm_.reserve(qc * sizeof(float));
        if ((size_t)64 * (k + 1) * 8 > 64 * 1024) {
            heap_d_.reserve(qc * (k + 1) * sizeof(float));
            heap_i_.reserve(qc * (k + 1) * sizeof(uint32_t));
        }

        for (uint64_t q0 = 0; q0 < nq; q0 += qc) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(qc, nq - q0);
            launch_pad_rows(dq + q0 * dim_, dim_, qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                            dpad_, s);
            if (need_norm)
                launch_self_dot(qpad_.as<float>(), dpad_, rows, (uint32_t)dim_,
                                qnorm_.as<float>(), s);
            launch_exact_tile(metric_, qpad_.as<float>(), dpad_, rows, base_.as<float>(), dpad_,
                              n, dpad_, qnorm_.as<float>(), bnorm_.as<float>(), D_.as<float>(),
                              ldd, s);
            launch_row_select(D_.as<float>(), ldd, rows, n, K, cand_d_.as<float>(),
                              cand_i_.as<uint32_t>(), K, flags_.as<uint32_t>(), s);
            launch_bf_finalize(cand_d_.as<float>(), cand_i_.as<uint32_t>(), K,
                               flags_.as<uint32_t>(), D_.as<float>(), ldd, rows, n, k,
                               d_idx + q0 * k, d_dist + q0 * k, d_count + q0,
                               err_.as<uint32_t>(), heap_d_.as<float>(), heap_i_.as<uint32_t>(),
                               s);
        }
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, err_.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err & 2u) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
        if (err & 1u) fail(HM_ERR_NAN_DISTANCE, "NaN distance (partial_cmp of NaN)");
    }

  private:
    uint64_t n_ = 0, cap_ = 0;
    uint32_t dpad_ = 0;
    int32_t metric_ = HM_METRIC_UNKNOWN;
    int32_t mode_ = HM_SEARCH_EXACT;
    uint32_t shortlist_ = 0;
    size_t d_budget_ = (size_t)4 << 30;
    DeviceBuffer base_, bnorm_;
    bool bnorm_valid_ = false;
    DeviceBuffer qpad_, qnorm_, D_, cand_d_, cand_i_, flags_, err_, heap_d_, heap_i_;
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
