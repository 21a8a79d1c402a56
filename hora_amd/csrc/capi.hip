// capi.hip -- the extern "C" boundary (include/hora_mi355x.h).
//
// Each entry translates C++ exceptions into an hm_status plus a
// thread-local message, so no error crosses the ABI as an exception or a
// process abort (the reference panics in the same situations).
#include <cstring>
#include <vector>

#include "index.h"
#include "kernels.h"
#include "mlp.h"

using namespace hm;

namespace {

template <class F> int32_t guard(F &&f)
{
    try {
        f();
        return HM_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return HM_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return HM_ERR_INVALID_ARG;
    }
}

void need(const void *p, const char *what)
{
    if (!p) fail(HM_ERR_INVALID_ARG, std::string(what) + " is null");
}

// Binds the calling thread to the index's device for the duration of a call.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) HM_HIP(hipSetDevice(dev));
    }
    ~DeviceScope()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

HM_API int32_t hm_params_default(int32_t kind, hm_params *out)
{
    return guard([&] {
        need(out, "out");
        std::memset(out, 0, sizeof(*out));
        if (kind == HM_INDEX_IVFPQ) {
            // IVFPQParams::default (pq_params.rs:80-91)
            out->n_sub = 25;
            out->sub_bits = 4;
            out->n_kmeans_center = 256;
            out->search_n_center = 8;
            out->train_epoch = 100;
            out->kmeans_accumulate = 0;
            out->seed = 0x5EEDull;
        } else if (kind == HM_INDEX_PQ) {
            // PQParams::default (pq_params.rs:31-40)
            out->n_sub = 4;
            out->sub_bits = 4;
            out->train_epoch = 100;
            out->kmeans_accumulate = 0;
            out->seed = 0x5EEDull;
        } else if (kind != HM_INDEX_BRUTE_FORCE) {
            fail(HM_ERR_INVALID_ARG, "unknown index kind");
        }
    });
}

HM_API int32_t hm_index_create(int32_t kind, uint64_t dimension, const hm_params *params,
                               hm_index **out)
{
    return guard([&] {
        need(out, "out");
        *out = nullptr;
        int dev = 0;
        HM_HIP(hipGetDevice(&dev));
        auto h = std::make_unique<hm_index>();
        h->device = dev;
        if (kind == HM_INDEX_BRUTE_FORCE) {
            h->impl = make_bruteforce(dimension);
        } else if (kind == HM_INDEX_IVFPQ) {
            hm_params p;
            if (params) {
                p = *params;
            } else {
                hm_params_default(HM_INDEX_IVFPQ, &p);
            }
            h->impl = make_ivfpq(dimension, p);
        } else if (kind == HM_INDEX_PQ) {
            hm_params p;
            if (params) {
                p = *params;
            } else {
                hm_params_default(HM_INDEX_PQ, &p);
            }
            h->impl = make_pq(dimension, p);
        } else {
            fail(HM_ERR_INVALID_ARG, "unknown index kind");
        }
        *out = h.release();
    });
}

HM_API int32_t hm_index_destroy(hm_index *h)
{
    return guard([&] {
        if (!h) return;
        {
            DeviceScope ds(h->device);
            delete h;
        }
    });
}

HM_API int32_t hm_index_add(hm_index *h, const float *rows, uint64_t n_rows)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        h->impl->add(rows, n_rows);
    });
}

HM_API int32_t hm_index_add_device(hm_index *h, const float *d_rows, uint64_t n_rows)
{
    return guard([&] {
        need(h, "index");
        if (n_rows == 0) return;
        need(d_rows, "rows");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        hipStream_t s = h->impl->stream();
        const uint64_t d = h->impl->dimension();
        h->err_stage.reserve(4);
        HM_HIP(hipMemsetAsync(h->err_stage.get(), 0, 4, s));
        launch_check_finite(d_rows, d, n_rows, (uint32_t)d, h->err_stage.as<uint32_t>(), s);
        uint32_t err = 0;
        HM_HIP(hipMemcpyAsync(&err, h->err_stage.get(), 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        if (err) fail(HM_ERR_NONFINITE_INPUT, "invalid float element");
        h->impl->add_device(d_rows, n_rows, d);
    });
}

HM_API int32_t hm_index_build(hm_index *h, int32_t metric)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        h->impl->build(metric);
    });
}

HM_API int32_t hm_index_built(const hm_index *h, int32_t *built)
{
    return guard([&] {
        need(h, "index");
        need(built, "built");
        *built = h->impl->built() ? 1 : 0;
    });
}

HM_API int32_t hm_index_search(hm_index *h, const float *queries, uint64_t nq, uint32_t k,
                               uint64_t *idx, float *dist, uint32_t *count)
{
    return guard([&] {
        need(h, "index");
        if (nq == 0) return;
        need(queries, "queries");
        need(count, "count");
        if (k) {
            need(idx, "idx");
            need(dist, "dist");
        }
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        const uint64_t d = h->impl->dimension();
        check_finite_host(queries, nq * d);
        hipStream_t s = h->impl->stream();
        h->q_stage.reserve(nq * d * sizeof(float));
        h->idx_stage.reserve(nq * std::max<uint32_t>(k, 1) * sizeof(uint64_t));
        h->dist_stage.reserve(nq * std::max<uint32_t>(k, 1) * sizeof(float));
        h->count_stage.reserve(nq * sizeof(uint32_t));
        HM_HIP(hipMemcpyAsync(h->q_stage.get(), queries, nq * d * sizeof(float),
                              hipMemcpyHostToDevice, s));
        h->impl->search_device(h->q_stage.as<float>(), nq, k, h->idx_stage.as<uint64_t>(),
                               h->dist_stage.as<float>(), h->count_stage.as<uint32_t>(), s);
        if (k) {
            HM_HIP(hipMemcpyAsync(idx, h->idx_stage.get(), nq * k * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, s));
            HM_HIP(hipMemcpyAsync(dist, h->dist_stage.get(), nq * k * sizeof(float),
                                  hipMemcpyDeviceToHost, s));
        }
        HM_HIP(hipMemcpyAsync(count, h->count_stage.get(), nq * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
    });
}

HM_API int32_t hm_index_search_device(hm_index *h, const float *d_queries, uint64_t nq,
                                      uint32_t k, uint64_t *d_idx, float *d_dist,
                                      uint32_t *d_count, void *stream)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        h->impl->search_device(d_queries, nq, k, d_idx, d_dist, d_count,
                               static_cast<hipStream_t>(stream));
    });
}

HM_API int32_t hm_index_search_ragged_device(hm_index *h, const float *d_queries,
                                             uint32_t n_batches, const uint32_t *batch_offsets,
                                             const uint32_t *batch_k, uint64_t *d_idx,
                                             float *d_dist, uint32_t *d_count, void *stream)
{
    return guard([&] {
        need(h, "index");
        if (n_batches == 0) return;
        need(batch_offsets, "batch_offsets");
        need(batch_k, "batch_k");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        h->impl->search_ragged_device(d_queries, n_batches, batch_offsets, batch_k, d_idx, d_dist,
                                      d_count, static_cast<hipStream_t>(stream));
    });
}

HM_API int32_t hm_index_autotune(hm_index *h, const float *queries, uint64_t nq, uint32_t k,
                                 uint32_t *shortlist)
{
    return guard([&] {
        need(h, "index");
        need(queries, "queries");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        check_finite_host(queries, nq * h->impl->dimension());
        uint32_t sl = h->impl->autotune(queries, nq, k);
        if (shortlist) *shortlist = sl;
    });
}

HM_API int32_t hm_index_tuning(const hm_index *h, uint32_t *shortlist, uint32_t *score_group_kib)
{
    return guard([&] {
        need(h, "index");
        uint32_t sl = 0, g = 0;
        {
            std::lock_guard<std::mutex> lk(h->mu);
            h->impl->tuning(&sl, &g);
        }
        if (shortlist) *shortlist = sl;
        if (score_group_kib) *score_group_kib = g;
    });
}

HM_API int32_t hm_index_set_search_mode(hm_index *h, int32_t mode, uint32_t shortlist)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        h->impl->set_search_mode(mode, shortlist);
    });
}

HM_API int32_t hm_index_search_stats(const hm_index *h, uint64_t *fallback,
                                     uint64_t *uncertified)
{
    return guard([&] {
        need(h, "index");
        need(fallback, "fallback");
        need(uncertified, "uncertified");
        h->impl->search_stats(fallback, uncertified);
    });
}

HM_API int32_t hm_index_dump(hm_index *h, const char *path, const uint8_t *id_bytes,
                             const uint64_t *id_offsets)
{
    return guard([&] {
        need(h, "index");
        need(path, "path");
        if (id_bytes) need(id_offsets, "id_offsets");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        dump_index(h->impl.get(), path, id_bytes, id_offsets);
    });
}

HM_API int32_t hm_index_load(int32_t kind, const char *path, int32_t id_width, hm_index **out)
{
    return guard([&] {
        need(out, "out");
        need(path, "path");
        *out = nullptr;
        int dev = 0;
        HM_HIP(hipGetDevice(&dev));
        auto h = std::make_unique<hm_index>();
        h->device = dev;
        h->impl = load_index(kind, path, id_width, h->loaded_ids, h->loaded_id_off);
        *out = h.release();
    });
}

HM_API int32_t hm_index_loaded_ids(const hm_index *h, uint8_t *id_bytes, uint64_t *id_offsets,
                                   uint64_t *total_bytes)
{
    return guard([&] {
        need(h, "index");
        need(total_bytes, "total_bytes");
        *total_bytes = h->loaded_ids.size();
        if (id_bytes && !h->loaded_ids.empty())
            std::memcpy(id_bytes, h->loaded_ids.data(), h->loaded_ids.size());
        if (id_offsets)
            std::memcpy(id_offsets, h->loaded_id_off.data(), h->loaded_id_off.size() * 8);
    });
}

HM_API int32_t hm_index_knn_graph(hm_index *h, uint32_t k, uint64_t *idx, float *dist,
                                  uint32_t *count)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        h->impl->knn_graph(k, idx, dist, count);
    });
}

HM_API int32_t hm_index_distances(hm_index *h, const float *queries, uint64_t nq,
                                  const uint32_t *ids, const uint64_t *offsets, float *out)
{
    return guard([&] {
        need(h, "index");
        if (nq == 0) return;
        need(queries, "queries");
        need(offsets, "offsets");
        const uint64_t total = offsets[nq];
        for (uint64_t q = 0; q < nq; ++q)
            if (offsets[q + 1] < offsets[q]) fail(HM_ERR_INVALID_ARG, "offsets decrease");
        if (total == 0) return;
        need(ids, "ids");
        need(out, "out");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        hipStream_t s = h->impl->stream();
        const uint64_t d = h->impl->dimension();
        hm::DeviceBuffer dq, di, doff, dout;
        dq.reserve(nq * d * 4);
        di.reserve(total * 4);
        doff.reserve((nq + 1) * 8);
        dout.reserve(total * 4);
        HM_HIP(hipMemcpyAsync(dq.get(), queries, nq * d * 4, hipMemcpyHostToDevice, s));
        HM_HIP(hipMemcpyAsync(di.get(), ids, total * 4, hipMemcpyHostToDevice, s));
        HM_HIP(hipMemcpyAsync(doff.get(), offsets, (nq + 1) * 8, hipMemcpyHostToDevice, s));
        h->impl->distances_device(dq.as<float>(), nq, di.as<uint32_t>(), doff.as<uint64_t>(),
                                  total, dout.as<float>(), s);
        HM_HIP(hipMemcpyAsync(out, dout.get(), total * 4, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
    });
}

HM_API int32_t hm_index_distances_device(hm_index *h, const float *d_queries, uint64_t nq,
                                         const uint32_t *d_ids, const uint64_t *d_offsets,
                                         uint64_t total, float *d_out, void *stream)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        h->impl->distances_device(d_queries, nq, d_ids, d_offsets, total, d_out,
                                  static_cast<hipStream_t>(stream));
    });
}

HM_API int32_t hm_index_nodes_size(const hm_index *h, uint64_t *out)
{
    return guard([&] {
        need(h, "index");
        need(out, "out");
        *out = h->impl->nodes_size();
    });
}

HM_API int32_t hm_index_dimension(const hm_index *h, uint64_t *out)
{
    return guard([&] {
        need(h, "index");
        need(out, "out");
        *out = h->impl->dimension();
    });
}

HM_API const char *hm_index_name(const hm_index *h) { return h ? h->impl->name() : ""; }

HM_API int32_t hm_index_clear(hm_index *h)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        h->impl->clear();
    });
}

HM_API int32_t hm_index_rebuild(hm_index *h, int32_t metric)
{
    (void)metric;
    return guard([&] {
        need(h, "index");
        fail(HM_ERR_NOT_IMPLEMENTED, "not implement");  // ann_index.rs:70-72
    });
}

HM_API int32_t hm_index_idx_info(const hm_index *h, char *buf, uint64_t cap, uint64_t *len)
{
    return guard([&] {
        need(h, "index");
        if (cap > 0) need(buf, "buf");
        static const char text[] = "not implement";  // ann_index.rs:113-115
        const uint64_t n = sizeof(text) - 1;
        if (len) *len = n;
        if (cap > 0) {
            const uint64_t c = n < cap - 1 ? n : cap - 1;
            std::memcpy(buf, text, c);
            buf[c] = '\0';
        }
    });
}

HM_API int32_t hm_index_get_rows(const hm_index *h, const uint64_t *idx, uint64_t n, float *out)
{
    return guard([&] {
        need(h, "index");
        if (n == 0) return;
        need(idx, "idx");
        need(out, "out");
        DeviceScope ds(h->device);
        h->impl->get_rows(idx, n, out);
    });
}

HM_API int32_t hm_index_set_nprobe(hm_index *h, uint32_t nprobe)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        ivfpq_set_nprobe(h->impl.get(), nprobe);
    });
}

HM_API int32_t hm_ivfpq_info(hm_index *h, uint32_t *out8)
{
    return guard([&] {
        need(h, "index");
        need(out8, "out");
        std::lock_guard<std::mutex> lk(h->mu);
        ivfpq_info(h->impl.get(), out8);
    });
}

HM_API int32_t hm_ivfpq_export(hm_index *h, float *centers, uint64_t *list_offsets,
                               uint64_t *list_ids, float *codebooks, uint8_t *codes)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        ivfpq_export(h->impl.get(), centers, list_offsets, list_ids, codebooks, codes);
    });
}

HM_API int32_t hm_ivfpq_import(hm_index *h, int32_t metric, const float *centers,
                               const uint64_t *list_offsets, const uint64_t *list_ids,
                               const float *codebooks, const uint8_t *codes)
{
    return guard([&] {
        need(h, "index");
        need(centers, "centers");
        need(list_offsets, "list_offsets");
        need(codebooks, "codebooks");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        ivfpq_import(h->impl.get(), metric, centers, list_offsets, list_ids, codebooks, codes);
    });
}

HM_API int32_t hm_pq_info(hm_index *h, uint32_t *out8)
{
    return guard([&] {
        need(h, "index");
        need(out8, "out8");
        std::lock_guard<std::mutex> lk(h->mu);
        pq_info(h->impl.get(), out8);
    });
}

HM_API int32_t hm_pq_export(hm_index *h, float *codebook, uint8_t *codes)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        pq_export(h->impl.get(), codebook, codes);
    });
}

HM_API int32_t hm_pq_import(hm_index *h, int32_t metric, const float *codebook,
                            const uint8_t *codes, uint64_t n_coded)
{
    return guard([&] {
        need(h, "index");
        need(codebook, "codebook");
        if (n_coded > 0) need(codes, "codes");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        pq_import(h->impl.get(), metric, codebook, codes, n_coded);
    });
}

HM_API int32_t hm_ivfpq_train_probe(hm_index *h, const float *train_queries, uint64_t n,
                                    uint32_t label_k, uint32_t hidden, uint32_t epochs,
                                    uint32_t batch, float lr, uint64_t seed, double *final_loss)
{
    return guard([&] {
        need(h, "index");
        need(train_queries, "train_queries");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        ivfpq_train_probe(h->impl.get(), train_queries, n, label_k, hidden, epochs, batch, lr,
                          seed, final_loss);
    });
}

HM_API int32_t hm_ivfpq_set_probe_policy(hm_index *h, int32_t mode, float mass, uint32_t pmin,
                                         uint32_t pmax)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        ivfpq_set_probe_policy(h->impl.get(), mode, mass, pmin, pmax);
    });
}

HM_API int32_t hm_ivfpq_set_scan_form(hm_index *h, int32_t form)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        ivfpq_set_scan_form(h->impl.get(), form);
    });
}

HM_API int32_t hm_ivfpq_set_rerank(hm_index *h, uint32_t depth)
{
    return guard([&] {
        need(h, "index");
        std::lock_guard<std::mutex> lk(h->mu);
        ivfpq_set_rerank(h->impl.get(), depth);
    });
}

HM_API int32_t hm_ivfpq_mean_nprobe(hm_index *h, double *out)
{
    return guard([&] {
        need(h, "index");
        need(out, "out");
        std::lock_guard<std::mutex> lk(h->mu);
        *out = ivfpq_mean_nprobe(h->impl.get());
    });
}

HM_API int32_t hm_ivfpq_last_probes(hm_index *h, uint64_t nq, uint32_t ldp, uint32_t *probes,
                                    uint32_t *counts)
{
    return guard([&] {
        need(h, "index");
        need(probes, "probes");
        need(counts, "counts");
        std::lock_guard<std::mutex> lk(h->mu);
        DeviceScope ds(h->device);
        ivfpq_last_probes(h->impl.get(), nq, ldp, probes, counts);
    });
}

HM_API int32_t hm_mlp_loss_grad(uint32_t din, uint32_t hidden, uint32_t dout, uint32_t n,
                                const float *W1, const float *b1, const float *W2,
                                const float *b2, const float *X, const float *T, float *loss,
                                float *gW1, float *gb1, float *gW2, float *gb2, float *logits)
{
    return guard([&] {
        for (const void *p : {(const void *)W1, (const void *)b1, (const void *)W2,
                              (const void *)b2, (const void *)X})
            need(p, "input");
        hipStream_t s = nullptr;
        HM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};
        ProbeMLP net((int)din, (int)hidden, (int)dout, 1, s);
        net.set_weights(W1, b1, W2, b2);
        DeviceBuffer Xd, X16, Td, Zd;
        Xd.reserve((size_t)n * din * 4);
        X16.reserve((size_t)n * din * 2);
        HM_HIP(hipMemcpyAsync(Xd.get(), X, (size_t)n * din * 4, hipMemcpyHostToDevice, s));
        f32_to_bf16(Xd.as<float>(), X16.as<uint16_t>(), (long)n * din, s);
        if (logits) {
            Zd.reserve((size_t)n * dout * 4);
            net.forward(X16.as<uint16_t>(), (int)n, Zd.as<float>());
            HM_HIP(hipMemcpyAsync(logits, Zd.get(), (size_t)n * dout * 4, hipMemcpyDeviceToHost, s));
        }
        if (T) {
            Td.reserve((size_t)n * dout * 4);
            HM_HIP(hipMemcpyAsync(Td.get(), T, (size_t)n * dout * 4, hipMemcpyHostToDevice, s));
            net.loss_grad(X16.as<uint16_t>(), Td.as<float>(), (int)n);
            if (loss)
                HM_HIP(hipMemcpyAsync(loss, net.loss_buf(), (size_t)n * 4, hipMemcpyDeviceToHost, s));
            net.get_grads(gW1, gb1, gW2, gb2);
        }
        HM_HIP(hipStreamSynchronize(s));
    });
}

HM_API const char *hm_last_error(void) { return last_error(); }

}  // extern "C"
