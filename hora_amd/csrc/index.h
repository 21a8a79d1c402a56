// index.h -- host runtime objects behind the C ABI.
#pragma once

#include "hm_common.h"

#include <memory>
#include <mutex>
#include <vector>

namespace hm {

// One ANNIndex implementation on one device (ann_index.rs:25-121).
class Index {
  public:
    Index(int32_t kind, uint64_t dim) : kind_(kind), dim_(dim) {}
    virtual ~Index() = default;

    int32_t kind() const { return kind_; }
    uint64_t dimension() const { return dim_; }

    virtual void add(const float *rows, uint64_t n_rows) = 0;
    virtual void build(int32_t metric) = 0;
    virtual bool built() const = 0;
    // Queries and outputs in device memory; outputs are [nq][k] / [nq].
    virtual void search_device(const float *d_queries, uint64_t nq, uint32_t k, uint64_t *d_idx,
                               float *d_dist, uint32_t *d_count, hipStream_t s) = 0;
    // The exact form of search_device with every working buffer carved from
    // the caller's scratch (scratch_bytes at a 256-byte aligned device
    // address): no device allocation.  Queries run in chunks sized to fit.
    virtual void search_device_scratch(const float *d_queries, uint64_t nq, uint32_t k,
                                       uint64_t *d_idx, float *d_dist, uint32_t *d_count,
                                       hipStream_t s, void *scratch, size_t scratch_bytes)
    {
        (void)d_queries, (void)nq, (void)k, (void)d_idx, (void)d_dist, (void)d_count, (void)s,
            (void)scratch, (void)scratch_bytes;
        fail(HM_ERR_NOT_IMPLEMENTED, "no-allocation search not supported by this index");
    }
    virtual const char *name() const = 0;
    virtual uint64_t nodes_size() const = 0;
    // The metric of the last build (HM_METRIC_UNKNOWN before one).
    virtual int32_t metric() const = 0;
    // Rows [first, first + count) into out (count x dimension floats).
    virtual void copy_rows(uint64_t first, uint64_t count, float *out) const = 0;
    virtual void clear() = 0;
    virtual void get_rows(const uint64_t *idx, uint64_t n, float *out) const = 0;
    // Ragged batches: batch b = queries [off[b], off[b+1]) with k = ks[b];
    // outputs packed batch after batch (k_b per query), counts per query.
    virtual void search_ragged_device(const float *dq, uint32_t nb, const uint32_t *h_off,
                                      const uint32_t *h_k, uint64_t *d_idx, float *d_dist,
                                      uint32_t *d_count, hipStream_t s)
    {
        (void)dq, (void)nb, (void)h_off, (void)h_k, (void)d_idx, (void)d_dist, (void)d_count,
            (void)s;
        fail(HM_ERR_NOT_IMPLEMENTED, "ragged search not supported by this index");
    }
    // naive_build_knn_graph (knn.rs:14-41) over the indexed rows: row i of
    // [n][k] holds its k nearest other rows, largest distance first.
    virtual void knn_graph(uint32_t k, uint64_t *h_idx, float *h_dist, uint32_t *h_count)
    {
        (void)k, (void)h_idx, (void)h_dist, (void)h_count;
        fail(HM_ERR_NOT_IMPLEMENTED, "kNN graph build not supported by this index");
    }
    // Batched distance service of the graph searches: out[p] =
    // metric(row ids[p], query q) for p in [off[q], off[q+1]), total pairs;
    // everything in device memory.
    virtual void distances_device(const float *dq, uint64_t nq, const uint32_t *ids,
                                  const uint64_t *off, uint64_t total, float *out, hipStream_t s)
    {
        (void)dq, (void)nq, (void)ids, (void)off, (void)total, (void)out, (void)s;
        fail(HM_ERR_NOT_IMPLEMENTED, "distance service not supported by this index");
    }
    // Chooses tuning parameters on a query sample; returns the shortlist.
    virtual uint32_t autotune(const float *h_queries, uint64_t nq, uint32_t k)
    {
        (void)h_queries, (void)nq, (void)k;
        return 0;
    }
    // The shortlist depth and score tile group (KiB) searches use; 0, 0 for
    // an index without shortlist forms.
    virtual void tuning(uint32_t *shortlist, uint32_t *group_kib) const
    {
        *shortlist = 0;
        *group_kib = 0;
    }
    // Appends rows already in device memory (n rows of dimension floats,
    // row stride ld floats).
    virtual void add_device(const float *d_rows, uint64_t n_rows, uint64_t ld)
    {
        (void)d_rows;
        (void)n_rows;
        (void)ld;
        fail(HM_ERR_NOT_IMPLEMENTED, "add_device not supported by this index");
    }
    virtual void set_search_mode(int32_t mode, uint32_t shortlist)
    {
        (void)mode;
        (void)shortlist;
        fail(HM_ERR_INVALID_ARG, "search mode not supported by this index");
    }

    // Rows of the last search that were rerun on the exact path, and how
    // many of those failed shortlist certification (shortlist forms only).
    virtual void search_stats(uint64_t *fallback, uint64_t *uncertified) const
    {
        *fallback = 0;
        *uncertified = 0;
    }

    hipStream_t stream() const { return stream_; }

  protected:
    int32_t kind_;
    uint64_t dim_;
    hipStream_t stream_ = nullptr;
};

std::unique_ptr<Index> make_bruteforce(uint64_t dim);
std::unique_ptr<Index> make_ivfpq(uint64_t dim, const hm_params &p);
std::unique_ptr<Index> make_pq(uint64_t dim, const hm_params &p);

// Standalone PQIndex extras (fail with INVALID_ARG on another index kind).
int32_t pq_info(Index *ix, uint32_t *out8);
int32_t pq_export(Index *ix, float *codebook, uint8_t *codes);
const hm_params &pq_params(Index *ix);
// Trained state from arrays (deserialization): codebook [M][KS][sdm],
// codes [n_trained][M] for the first n_trained rows.
void pq_import(Index *ix, int32_t metric, const float *codebook, const uint8_t *codes,
               uint64_t n_trained);
const hm_params &ivfpq_params(Index *ix);
// Trained state from arrays: centers [nlist][dim], list offsets [nlist+1]
// and ids (list order), codebooks [nlist][M][KS][sdm], codes [n][M] in list
// order.
void ivfpq_import(Index *ix, int32_t metric, const float *centers, const uint64_t *offsets,
                  const uint64_t *ids, const float *codebooks, const uint8_t *codes);

// IVFPQIndex extras (fail with INVALID_ARG on another index kind).
int32_t ivfpq_info(Index *ix, uint32_t *out8);
int32_t ivfpq_export(Index *ix, float *centers, uint64_t *offsets, uint64_t *ids,
                     float *codebooks, uint8_t *codes);
int32_t ivfpq_set_nprobe(Index *ix, uint32_t nprobe);
int32_t ivfpq_set_probe_policy(Index *ix, int32_t mode, float mass, uint32_t pmin, uint32_t pmax);
int32_t ivfpq_set_scan_form(Index *ix, int32_t form);
int32_t ivfpq_set_rerank(Index *ix, uint32_t depth);
int32_t ivfpq_train_probe(Index *ix, const float *q, uint64_t n, uint32_t label_k,
                          uint32_t hidden, uint32_t epochs, uint32_t batch, float lr,
                          uint64_t seed, double *loss);
double ivfpq_mean_nprobe(Index *ix);
int32_t ivfpq_last_probes(Index *ix, uint64_t nq, uint32_t ldp, uint32_t *probes,
                          uint32_t *counts);

// SerializableIndex (serialize.hip): the reference's bincode layout.
// id_bytes/id_off: node i's Option<T> encoding is id_bytes[id_off[i] ..
// id_off[i+1]); null writes Some(i as usize).  Load: id_width > 0 reads a
// fixed-width T, 0 a u64-length-prefixed T; the encodings come back in
// id_bytes / id_off (n + 1 offsets).
void dump_index(Index *ix, const char *path, const uint8_t *id_bytes, const uint64_t *id_off);
std::unique_ptr<Index> load_index(int32_t kind, const char *path, int32_t id_width,
                                  std::vector<uint8_t> &id_bytes, std::vector<uint64_t> &id_off);

// Host-side validation shared by the index types (node.rs:158-166).
void check_finite_host(const float *x, uint64_t count);

// Fills unused result slots (index UINT64_MAX, distance +inf, count 0).
void launch_fill_results(uint64_t *idx, float *dist, uint32_t *count, uint64_t nq, uint32_t k,
                         hipStream_t s);

}  // namespace hm

struct hm_index {
    std::unique_ptr<hm::Index> impl;
    mutable std::mutex mu;  // every entry that touches the index state holds it
    int device = 0;
    // Host-entry staging buffers.
    hm::DeviceBuffer q_stage, idx_stage, dist_stage, count_stage, err_stage;
    // Option<T> encodings of the nodes read by hm_index_load.
    std::vector<uint8_t> loaded_ids;
    std::vector<uint64_t> loaded_id_off;
};
