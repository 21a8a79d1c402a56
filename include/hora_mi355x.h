/*
 * hora_mi355x.h -- C ABI of the MI355X (gfx950) distance-and-scan engine.
 *
 * This is the drop-in boundary for hora's hot path: the entry points below
 * are what a hora FFI crate binds in place of the Rust implementations of
 * the ANNIndex trait (src/core/ann_index.rs:25-121) for BruteForceIndex
 * (src/index/bruteforce_idx.rs) and IVFPQIndex / PQIVFIndex
 * (src/index/pq_idx.rs).  Each declaration cites the reference item it
 * replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Host pointers unless the name ends in
 *    _device (then device pointers, valid on the index's device).
 *  - Every entry returns an int32 status (hm_status); on failure
 *    hm_last_error() returns a thread-local message that reuses the
 *    reference's error strings where the reference has one.
 *  - Item indices are positional: the i-th row ever added has index i
 *    (the reference's internal usize index, bruteforce_idx.rs:50-53).  The
 *    caller keeps its own id table to map indices to its id type T, as the
 *    reference's Node::idx does (node.rs:96).
 *  - Result rows are k wide; row q holds count[q] results in ascending
 *    distance order (node_search_k's drain+reverse, bruteforce_idx.rs:63-71);
 *    unused slots hold index UINT64_MAX and distance +inf.
 *  - Reference panics become error statuses (the reference has no status
 *    for them); the reference's 100,001-item caps of IVFPQIndex and PQIndex
 *    (pq_idx.rs:326, 66) are not reproduced.
 */
#ifndef HORA_MI355X_H
#define HORA_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HM_API __attribute__((visibility("default")))

typedef struct hm_index hm_index;

/* Variant order of `enum Metric` (src/core/metrics.rs:5-13); also the
 * bincode variant index. */
typedef enum {
    HM_METRIC_UNKNOWN = 0,
    HM_METRIC_MANHATTAN = 1,
    HM_METRIC_DOT_PRODUCT = 2,
    HM_METRIC_EUCLIDEAN = 3,
    HM_METRIC_COSINE_SIMILARITY = 4,
    HM_METRIC_ANGULAR = 5
} hm_metric;

typedef enum {
    HM_OK = 0,
    HM_ERR_INVALID_ARG = 1,
    HM_ERR_DIM_MISMATCH = 2,     /* "different dimensions" (calc.rs:25) / "dimension is different" (pq_idx.rs:342) */
    HM_ERR_NONFINITE_INPUT = 3,  /* reference: panic!("invalid float element") (node.rs:162) */
    HM_ERR_UNKNOWN_METRIC = 4,   /* "unknown method" (metrics.rs:32) */
    HM_ERR_NOT_BUILT = 5,
    HM_ERR_CAPACITY = 6,
    HM_ERR_HIP = 7,
    HM_ERR_NAN_DISTANCE = 8,     /* reference: partial_cmp().unwrap() panic (neighbor.rs:32) */
    HM_ERR_EMPTY = 9,            /* "None to assigned impossible split center" (kmeans.rs:159) */
    HM_ERR_NOT_IMPLEMENTED = 10, /* "not implement" (ann_index.rs:70) */
    HM_ERR_OUT_OF_MEMORY = 11
} hm_status;

typedef enum {
    HM_INDEX_BRUTE_FORCE = 1, /* BruteForceIndex (bruteforce_idx.rs:15-22) */
    HM_INDEX_IVFPQ = 2,       /* IVFPQIndex, exported as PQIVFIndex too (pq_idx.rs:267-292) */
    HM_INDEX_PQ = 3           /* standalone PQIndex (pq_idx.rs:21-45) */
} hm_index_kind;

/* BruteForceParams is empty (bruteforce_params.rs:5-14); IVFPQParams fields
 * (pq_params.rs:43-91); PQParams (pq_params.rs:7-41) uses n_sub, sub_bits
 * and train_epoch.  `seed` seeds the k-means split step, which the
 * reference draws from an unseeded thread_rng (kmeans.rs:168). */
typedef struct {
    uint32_t n_sub;
    uint32_t sub_bits;
    uint32_t n_kmeans_center;
    uint32_t search_n_center;
    uint32_t train_epoch;
    uint32_t kmeans_accumulate; /* 0: f32 in point order (reference kmeans.rs:102-113), 1: fp64 */
    uint64_t seed;
} hm_params;

/* Search forms of BruteForceIndex.  EXACT runs every pair through the
 * reference-order FP32 arithmetic; the RERANK forms build a shortlist with
 * MFMA (BF16 or FP8-e4m3 with per-vector scale, FP32 accumulate) and rerank
 * it with the exact arithmetic. */
typedef enum {
    HM_SEARCH_EXACT = 0,
    HM_SEARCH_BF16_RERANK = 1,
    HM_SEARCH_FP8_RERANK = 2
} hm_search_mode;

/* Fills the reference defaults: IVFPQParams::default() (pq_params.rs:80-91)
 * for HM_INDEX_IVFPQ, PQParams::default() (pq_params.rs:31-40) for
 * HM_INDEX_PQ; zeros for HM_INDEX_BRUTE_FORCE. */
HM_API int32_t hm_params_default(int32_t kind, hm_params *out);

/* BruteForceIndex::new (bruteforce_idx.rs:25-32), IVFPQIndex::new
 * (pq_idx.rs:295-330), PQIndex::new (pq_idx.rs:48-90).  Binds to the
 * current HIP device. */
HM_API int32_t hm_index_create(int32_t kind, uint64_t dimension, const hm_params *params,
                               hm_index **out);

/* Drop of the Rust index value. */
HM_API int32_t hm_index_destroy(hm_index *h);

/* ANNIndex::add / madd / add_node (ann_index.rs:36-59; bruteforce_idx.rs:40-43;
 * pq_idx.rs:340-355, 448-453): appends n_rows rows of `dimension` floats.
 * Non-finite elements are rejected (node.rs:158-166) and nothing is added. */
HM_API int32_t hm_index_add(hm_index *h, const float *rows, uint64_t n_rows);

/* Same as hm_index_add with the rows already in device memory (n_rows x
 * dimension floats, contiguous); rejected if any element is non-finite. */
HM_API int32_t hm_index_add_device(hm_index *h, const float *d_rows, uint64_t n_rows);

/* ANNIndex::build (ann_index.rs:30; bruteforce_idx.rs:36-39; pq_idx.rs:443-447). */
HM_API int32_t hm_index_build(hm_index *h, int32_t metric);

/* ANNIndex::built (ann_index.rs:64). */
HM_API int32_t hm_index_built(const hm_index *h, int32_t *built);

/* ANNIndex::node_search_k / search for a batch of nq queries
 * (ann_index.rs:74-97; bruteforce_idx.rs:47-73; pq_idx.rs:408-477). */
HM_API int32_t hm_index_search(hm_index *h, const float *queries, uint64_t nq, uint32_t k,
                               uint64_t *idx, float *dist, uint32_t *count);

/* Same, with queries and outputs in device memory, enqueued on `stream`
 * (a hipStream_t; NULL = the index's own stream).  Returns after the work is
 * enqueued and the stream has been synchronised. */
HM_API int32_t hm_index_search_device(hm_index *h, const float *d_queries, uint64_t nq,
                                      uint32_t k, uint64_t *d_idx, float *d_dist,
                                      uint32_t *d_count, void *stream);

/* Size of the device array that hm_index_search_device_noalloc works in. */
#define HM_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)

/* hm_index_search_device with no allocation of any kind: no hipMalloc,
 * hipFree or allocating library call.  No counterpart in the reference; it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  BruteForceIndex only.
 * It runs the exact form whatever hm_index_set_search_mode chose (ids,
 * distances and counts equal hm_index_search_device's in HM_SEARCH_EXACT
 * mode) in a file-scope device array of HM_NOALLOC_SCRATCH_BYTES per GPU,
 * queries in chunks that fit it.  HM_ERR_CAPACITY when one query's buffers
 * (padded row, distance row of nodes_size floats, min(k+1, n) candidates)
 * exceed it, i.e. an index of more than about 8.3 million rows.  Cosine and
 * Angular read base norms that build and add compute.  Calls on one GPU
 * share the array: they are serialised, and each returns after its stream
 * has drained. */
HM_API int32_t hm_index_search_device_noalloc(hm_index *h, const float *d_queries, uint64_t nq,
                                              uint32_t k, uint64_t *d_idx, float *d_dist,
                                              uint32_t *d_count, void *stream);

/* A set of query batches of uneven size, each with its own k: batch b is
 * queries [batch_offsets[b], batch_offsets[b+1]) (host array, nb+1 entries)
 * searched with k = batch_k[b] (host array).  d_queries holds all queries
 * back to back; results are packed batch after batch, k_b entries per query
 * (d_idx / d_dist hold sum_b n_b*k_b entries), d_count one per query.  In
 * the shortlist forms one captured hipGraph per call runs every chunk of
 * <= 4096 queries and the device rerun of up to 8 tied rows (each with its
 * own k); the host reruns rows only when that plan leaves some.  The exact
 * form searches each batch with its own k.  No reference counterpart (the
 * reference searches one query at a time, ann_index.rs:91-97). */
HM_API int32_t hm_index_search_ragged_device(hm_index *h, const float *d_queries,
                                             uint32_t n_batches, const uint32_t *batch_offsets,
                                             const uint32_t *batch_k, uint64_t *d_idx,
                                             float *d_dist, uint32_t *d_count, void *stream);

/* Times the shortlist forms on a host query sample for shortlist depths
 * 64..4096 and keeps the fastest (written to *shortlist); then, at that
 * depth, the score GEMM's tile order: base tiles are taken in groups of
 * 512, 1,024, 2,048 or 3,072 KiB of base rows per XCD's L2 (the query tiles
 * inside a group), each timed three times, and the fastest is kept (read it
 * with hm_index_tuning).  No reference counterpart. */
HM_API int32_t hm_index_autotune(hm_index *h, const float *queries, uint64_t nq, uint32_t k,
                                 uint32_t *shortlist);

/* The shortlist depth (0: automatic) and the score tile group size in KiB
 * (2,048 by default) that the index's shortlist searches use; 0 and 0 for an
 * index without shortlist forms.  Either pointer may be NULL. */
HM_API int32_t hm_index_tuning(const hm_index *h, uint32_t *shortlist, uint32_t *score_group_kib);

/* Selects the BruteForceIndex search form and its shortlist depth
 * (0 = automatic).  No reference counterpart (the reference has one form). */
HM_API int32_t hm_index_set_search_mode(hm_index *h, int32_t mode, uint32_t shortlist);

/* Rows of the last search rerun on the exact path (tied or uncertified
 * shortlist) and how many were uncertified.  No reference counterpart. */
HM_API int32_t hm_index_search_stats(const hm_index *h, uint64_t *fallback,
                                     uint64_t *uncertified);

/* SerializableIndex::dump (ann_index.rs:152-157; bruteforce_idx.rs:98-105,
 * pq_idx.rs:253-264, 507-522): writes the index to `path` in the reference's
 * on-disk layout (serde derive over the index struct, bincode 1.x default
 * encoding: little-endian, usize and lengths as u64, enum variants as u32,
 * bool and Option tags as one byte).  Node i's `idx: Option<T>` is written
 * from the caller's bincode bytes id_bytes[id_offsets[i] .. id_offsets[i+1])
 * (n + 1 offsets); id_bytes NULL writes Some(i as usize). */
HM_API int32_t hm_index_dump(hm_index *h, const char *path, const uint8_t *id_bytes,
                             const uint64_t *id_offsets);

/* SerializableIndex::load (ann_index.rs:145-150; bruteforce_idx.rs:87-96,
 * pq_idx.rs:243-251, 490-505) for an index of `kind`: rows, metric and
 * trained state come back without retraining.  id_width: the byte width of
 * a fixed-width id type T (8 for usize), or 0 for a u64-length-prefixed T
 * (String).  The ids are kept for hm_index_loaded_ids. */
HM_API int32_t hm_index_load(int32_t kind, const char *path, int32_t id_width, hm_index **out);

/* The Option<T> encodings of the nodes read by hm_index_load, node after
 * node: *total_bytes is always set; id_bytes (total_bytes) and id_offsets
 * (n + 1) are filled when not NULL. */
HM_API int32_t hm_index_loaded_ids(const hm_index *h, uint8_t *id_bytes, uint64_t *id_offsets,
                                   uint64_t *total_bytes);

/* naive_build_knn_graph (knn.rs:14-41; the same loop as SSGIndex's
 * build_knn_graph, ssg_idx.rs:69-98) over the rows of a built
 * BruteForceIndex, with its metric: row i of idx/dist ([n][k], host memory)
 * holds the k nearest other rows in the order the reference's heap drains
 * them -- largest distance first -- and count[i] = min(k, n - 1).  Unused
 * slots hold UINT64_MAX / +inf. */
HM_API int32_t hm_index_knn_graph(hm_index *h, uint32_t k, uint64_t *idx, float *dist,
                                  uint32_t *count);

/* Batched distance service for graph searches: the distance evaluations of
 * HNSWIndex::search_layer (get_distance_from_vec, hnsw_idx.rs:243-254) and
 * SSGIndex::search (ssg_idx.rs:402-456), metric(node, query) with the
 * index's metric and the reference's arithmetic, for many (query, node)
 * pairs per call: out[p] for p in [offsets[q], offsets[q+1]) is the distance
 * from node ids[p] to query q (offsets has nq + 1 entries, host memory in
 * hm_index_distances).  A BruteForceIndex holds the nodes. */
HM_API int32_t hm_index_distances(hm_index *h, const float *queries, uint64_t nq,
                                  const uint32_t *ids, const uint64_t *offsets, float *out);

/* Same with every buffer in device memory; `total` = offsets[nq]. */
HM_API int32_t hm_index_distances_device(hm_index *h, const float *d_queries, uint64_t nq,
                                         const uint32_t *d_ids, const uint64_t *d_offsets,
                                         uint64_t total, float *d_out, void *stream);

/* ANNIndex::nodes_size (ann_index.rs:105) -- number of rows added. */
HM_API int32_t hm_index_nodes_size(const hm_index *h, uint64_t *out);

/* ANNIndex::dimension (ann_index.rs:118). */
HM_API int32_t hm_index_dimension(const hm_index *h, uint64_t *out);

/* ANNIndex::name (ann_index.rs:102): "BruteForceIndex" / "IVFPQIndex" /
 * "PQIndex". */
HM_API const char *hm_index_name(const hm_index *h);

/* ANNIndex::clear (ann_index.rs:110): drops all rows and built state. */
HM_API int32_t hm_index_clear(hm_index *h);

/* ANNIndex::rebuild (ann_index.rs:66-72).  None of the three index types
 * overrides the trait default, so this returns HM_ERR_NOT_IMPLEMENTED with
 * the default's message "not implement" and leaves the index unchanged. */
HM_API int32_t hm_index_rebuild(hm_index *h, int32_t metric);

/* ANNIndex::idx_info (ann_index.rs:113-115).  None of the three index
 * types overrides the trait default, so the text is the default's "not
 * implement".  Writes at most cap bytes including the terminating NUL into
 * buf (buf may be NULL when cap is 0); *len (optional) = the text's length
 * without the NUL. */
HM_API int32_t hm_index_idx_info(const hm_index *h, char *buf, uint64_t cap, uint64_t *len);

/* Copies rows by index into `out` (n x dimension floats): what
 * search_nodes returns with each Node (ann_index.rs:81-84). */
HM_API int32_t hm_index_get_rows(const hm_index *h, const uint64_t *idx, uint64_t n,
                                 float *out);

/* IVFPQIndex: number of lists probed per query, the search_n_center of
 * IVFPQParams (pq_params.rs:53-55, 592-595) changed after construction. */
HM_API int32_t hm_index_set_nprobe(hm_index *h, uint32_t nprobe);

/* IVFPQIndex shape after build: out8 = {n_kmeans_center, n_sub, 2^sub_bits,
 * sub_dim_max, rows trained, trained flag, nprobe, padded dimension}.  The
 * reference keeps these in private fields (pq_idx.rs:268-292). */
HM_API int32_t hm_ivfpq_info(hm_index *h, uint32_t *out8);

/* IVFPQIndex trained state (_centers, _ivflist, per-list _centers and
 * _assigned_center of pq_idx.rs:280-282, 33, 41) as arrays:
 * centers [nlist][dim] f32; list_offsets [nlist+1]; list_ids [n] ascending
 * within each list; codebooks [nlist][n_sub][2^sub_bits][sub_dim_max] f32
 * (a sub-space with fewer dims is zero padded); codes [n][n_sub] u8 with row
 * r belonging to list_ids[r].  Any output pointer may be NULL. */
HM_API int32_t hm_ivfpq_export(hm_index *h, float *centers, uint64_t *list_offsets,
                               uint64_t *list_ids, float *codebooks, uint8_t *codes);

/* The inverse of hm_ivfpq_export: installs a trained state (the same arrays
 * and layouts; codes in list order) on an index that holds the rows, and
 * marks it built for `metric`.  The lists must hold rows 0..n-1 of the
 * index exactly once.  The reference takes such state only through load
 * (pq_idx.rs:488-523); this entry lets a caller bring centroids and
 * codebooks trained elsewhere, as conditional parity does. */
HM_API int32_t hm_ivfpq_import(hm_index *h, int32_t metric, const float *centers,
                               const uint64_t *list_offsets, const uint64_t *list_ids,
                               const float *codebooks, const uint8_t *codes);

/* Standalone PQIndex shape after build: out8 = {n_sub, 2^sub_bits,
 * sub_dim_max, rows trained, trained flag, padded dimension, metric, 0}
 * (private fields of pq_idx.rs:22-45). */
HM_API int32_t hm_pq_info(hm_index *h, uint32_t *out8);

/* Standalone PQIndex trained state (_centers and _assigned_center of
 * pq_idx.rs:33, 41): codebook [n_sub][2^sub_bits][sub_dim_max] f32 (a
 * sub-space with fewer dims is zero padded), codes [n][n_sub] u8 (row i is
 * item i).  Either pointer may be NULL. */
HM_API int32_t hm_pq_export(hm_index *h, float *codebook, uint8_t *codes);

/* The inverse of hm_pq_export: installs the codebook and the codes of the
 * first n_coded rows (n_coded <= rows added) and marks the index built for
 * `metric` (the reference: load, pq_idx.rs:243-265). */
HM_API int32_t hm_pq_import(hm_index *h, int32_t metric, const float *codebook,
                            const uint8_t *codes, uint64_t n_coded);

/* Learned probe (no reference counterpart: the reference probes a fixed
 * search_n_center lists, pq_idx.rs:424).  Trains, on the GPU, a network
 * dimension -> hidden (ReLU) -> n_kmeans_center whose softmax predicts
 * which lists hold a query's exact top-label_k (labels from a GPU brute
 * force over the indexed rows).  BF16 weights, FP32 accumulation, Adam.
 * final_loss: mean cross-entropy of the last epoch. */
HM_API int32_t hm_ivfpq_train_probe(hm_index *h, const float *train_queries, uint64_t n,
                                    uint32_t label_k, uint32_t hidden, uint32_t epochs,
                                    uint32_t batch, float lr, uint64_t seed, double *final_loss);

/* mode 0: fixed search_n_center (the reference behaviour, and the
 * fallback); mode 1: learned -- each query probes the smallest set of its
 * highest-scored lists whose predicted mass reaches `mass`, between pmin and
 * pmax lists, scanned in descending score with the reference's per-list
 * scan and merge. */
HM_API int32_t hm_ivfpq_set_probe_policy(hm_index *h, int32_t mode, float mass, uint32_t pmin,
                                         uint32_t pmax);

/* How IVFPQ search scans the probed lists (the results are the same):
 * 0 (default) the fastest form that fits; 1 one workgroup per (query,
 * probed list); 2 list-major, the list's codebook in LDS and a top-(k+1)
 * per list; 3 list-major with the codebook in registers, every ADC
 * distance into the query's row and one radix select per row.  Forms 2 and 3
 * fail at search time when the index's PQ shape does not fit them.  No
 * counterpart in the reference (pq_idx.rs:408-455 has one scan loop); for
 * tuning and for testing every form. */
HM_API int32_t hm_ivfpq_set_scan_form(hm_index *h, int32_t form);

/* Exact rerank of the IVFPQ result (0, the default, keeps the reference's
 * ADC result, pq_idx.rs:181-191, 429-436).  With depth > 0: the nprobe
 * lists nearest by (coarse distance, list index) -- or the probe network's
 * lists -- are scanned; the R = max(depth, k) smallest ADC distances by
 * (distance, row index) are rescored with the exact metric (the reference's
 * f32 fold, metrics.rs:22-34) and sorted by (exact distance, row index); the
 * first k come back with their exact distances.  Every step is a total
 * order, so no heap is replayed.  No counterpart in the reference: it wins
 * back the accuracy the PQ codes lose.  depth <= 4095;
 * nprobe * (R + 1) <= 16384. */
HM_API int32_t hm_ivfpq_set_rerank(hm_index *h, uint32_t depth);

/* Mean number of lists probed per query by the last search. */
HM_API int32_t hm_ivfpq_mean_nprobe(hm_index *h, double *out);

/* The lists the last search scanned, in scan order: probes [nq][ldp]
 * (unused slots UINT32_MAX), counts [nq].  Valid when the last search ran
 * as a single chunk of nq queries.  For parity checks of the learned form. */
HM_API int32_t hm_ivfpq_last_probes(hm_index *h, uint64_t nq, uint32_t ldp, uint32_t *probes,
                                    uint32_t *counts);

/* Probe network forward/backward on given weights, for gradient checks:
 * X [n][din], T [n][dout] targets (NULL: forward only), W1 [din][hidden],
 * W2 [hidden][dout]; outputs (any may be NULL) per-row loss, gradients of
 * the mean loss, logits [n][dout].  Inputs are rounded to bf16 as in
 * training. */
HM_API int32_t hm_mlp_loss_grad(uint32_t din, uint32_t hidden, uint32_t dout, uint32_t n,
                                const float *W1, const float *b1, const float *W2,
                                const float *b2, const float *X, const float *T, float *loss,
                                float *gW1, float *gb1, float *gW2, float *gb2, float *logits);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * each launch (off by default; instrumentation, no reference counterpart).
 * hm_profile_query synchronises and returns the accumulated milliseconds
 * and launch count of a kernel family ("exact_tile", "row_select", ...). */
HM_API int32_t hm_profile_enable(int32_t on);
/* Restricts the timing to one family (NULL or "" times every family), so a
 * timed region carries only the events of the kernel it measures. */
HM_API int32_t hm_profile_only(const char *kernel);
HM_API int32_t hm_profile_reset(void);
HM_API int32_t hm_profile_query(const char *kernel, double *total_ms, uint64_t *launches);

/* Thread-local message of the last failed call on this thread. */
HM_API const char *hm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HORA_MI355X_H */
