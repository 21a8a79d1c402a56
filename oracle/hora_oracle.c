/*
 * hora_oracle.c -- CPU restatement of hora's distance-and-scan path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (hora_amd/libhora_mi355x.so) never links or calls it.
 *
 * Every function states the reference file:line it restates (paths relative
 * to the hora source tree).  Compile with -O2 -ffp-contract=off
 * -fno-fast-math so that every float operation is a single IEEE-754 binary32
 * operation in source order, as the reference's Rust code is (Rust never
 * contracts a*b+c into an FMA).
 *
 * Parity pinning (what checks each function; see tests/test_oracle.py):
 *   metrics          -- calc.rs:49 (dot == -14), node.rs:185 (Manhattan == 2),
 *                       closed forms, FP64 bound, sequential-order property.
 *   heap             -- hand-traced push/pop layouts (tests/golden/heap_*.json),
 *                       heap invariants, tie-free reduction to a sort.
 *   bf_search        -- brute force by full sort on tie-free inputs; heap
 *                       fixtures for ties.
 *   kmeans init      -- closed form (two distinct centres, kmeans.rs:72-84).
 *   kmeans assign    -- brute-force argmin with lowest-index rule.
 *   kmeans update    -- closed form on hand-made clusters; empty -> zero.
 *   kmeans split     -- parity unpinned against the reference: it draws from
 *                       an unseeded thread_rng (kmeans.rs:168).  The oracle
 *                       uses SplitMix64 (seeded) instead; checked only for
 *                       its invariants (counts conserved, epsilon pattern).
 *   pq ranges        -- closed form fixtures (d=10,M=4 / d=10,M=3).
 *   pq lut / adc     -- closed form on tiny codebooks, brute force.
 *   ivfpq search     -- brute force over probed lists on tie-free inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum {
    ORC_UNKNOWN = 0,
    ORC_MANHATTAN = 1,
    ORC_DOT = 2,
    ORC_EUCLIDEAN = 3,
    ORC_COSINE = 4,
    ORC_ANGULAR = 5
};

enum {
    ORC_OK = 0,
    ORC_ERR_UNKNOWN_METRIC = 4,
    ORC_ERR_NAN = 8,
    ORC_ERR_EMPTY = 9,
    ORC_ERR_ARG = 1
};

/* Rust's `impl Sum for f32` folds from -0.0 (core/src/iter/traits/accum.rs
 * in current std).  Only the sign of an all-zero sum depends on this, and the
 * sign never changes a comparison. */
#define ORC_SUM_INIT (-0.0f)

/* ---------------------------------------------------------------- metrics */

/* simd_metrics.rs:30-33 (scalar branch): -(sum_t a_t*b_t), left fold. */
float orc_dot_neg(const float *a, const float *b, size_t d)
{
    float acc = ORC_SUM_INIT;
    for (size_t t = 0; t < d; ++t) {
        float p = a[t] * b[t];
        acc = acc + p;
    }
    return -acc;
}

/* simd_metrics.rs:60-66: sum_t |a_t-b_t|, left fold. */
float orc_manhattan(const float *a, const float *b, size_t d)
{
    float acc = ORC_SUM_INIT;
    for (size_t t = 0; t < d; ++t) {
        float df = a[t] - b[t];
        acc = acc + fabsf(df);
    }
    return acc;
}

/* simd_metrics.rs:96-102: sum_t (a_t-b_t)^2 (powi(2) is one multiply),
 * left fold, no sqrt. */
float orc_euclidean(const float *a, const float *b, size_t d)
{
    float acc = ORC_SUM_INIT;
    for (size_t t = 0; t < d; ++t) {
        float df = a[t] - b[t];
        float sq = df * df;
        acc = acc + sq;
    }
    return acc;
}

/* metrics.rs:73-83 as written: -dot(a,b)/(sqrt(dot(a,a))*sqrt(dot(b,b)))
 * with dot already negated, so both square roots take values <= 0 and the
 * result is NaN for every pair with a nonzero vector. */
float orc_cosine_reference(const float *a, const float *b, size_t d)
{
    float num = -orc_dot_neg(a, b, d);
    float na = sqrtf(orc_dot_neg(a, a, d));
    float nb = sqrtf(orc_dot_neg(b, b, d));
    return num / (na * nb);
}

/* Cosine as this build defines it (the reference value is always NaN):
 * -s_ab / (sqrt(s_aa) * sqrt(s_bb)), each s a left-fold f32 sum of products.
 * Smaller means more similar. */
float orc_cosine_build(const float *a, const float *b, size_t d)
{
    float sab = -orc_dot_neg(a, b, d);
    float saa = -orc_dot_neg(a, a, d);
    float sbb = -orc_dot_neg(b, b, d);
    float den = sqrtf(saa) * sqrtf(sbb);
    return (-sab) / den;
}

/* metrics.rs:88-103 with float_two() == 1.0 (node.rs:54-56):
 * rhd = dot(a,a), lhd = dot(b,b), rldot = dot(a,b) (all negated),
 * r = rhd*lhd; r > 0 ? two - two*rldot/sqrt(r) : two. */
float orc_angular(const float *a, const float *b, size_t d)
{
    float rhd = orc_dot_neg(a, a, d);
    float lhd = orc_dot_neg(b, b, d);
    float rldot = orc_dot_neg(a, b, d);
    float rlmul = rhd * lhd;
    const float two = 1.0f;
    if (rlmul > 0.0f) {
        float t = two * rldot;
        t = t / sqrtf(rlmul);
        return two - t;
    }
    return two;
}

/* metrics.rs:22-34 (metric dispatch).  Cosine uses the build definition
 * (orc_cosine_build); the reference's own value is orc_cosine_reference. */
int orc_metric(const float *a, const float *b, size_t d, int mt, float *out)
{
    switch (mt) {
    case ORC_EUCLIDEAN: *out = orc_euclidean(a, b, d); return ORC_OK;
    case ORC_MANHATTAN: *out = orc_manhattan(a, b, d); return ORC_OK;
    case ORC_DOT: *out = -orc_dot_neg(a, b, d); return ORC_OK; /* metrics.rs:56 */
    case ORC_COSINE: *out = orc_cosine_build(a, b, d); return ORC_OK;
    case ORC_ANGULAR: *out = orc_angular(a, b, d); return ORC_OK;
    default: return ORC_ERR_UNKNOWN_METRIC;
    }
}

/* FP64 versions of the mathematical quantities, for error statements. */
double orc_metric_f64(const float *a, const float *b, size_t d, int mt)
{
    double s = 0.0, saa = 0.0, sbb = 0.0;
    for (size_t t = 0; t < d; ++t) {
        double x = a[t], y = b[t];
        switch (mt) {
        case ORC_EUCLIDEAN: s += (x - y) * (x - y); break;
        case ORC_MANHATTAN: s += fabs(x - y); break;
        default: s += x * y; saa += x * x; sbb += y * y; break;
        }
    }
    if (mt == ORC_COSINE) return -s / (sqrt(saa) * sqrt(sbb));
    if (mt == ORC_ANGULAR) return (saa * sbb > 0) ? 1.0 + s / sqrt(saa * sbb) : 1.0;
    return s;
}

/* ------------------------------------------------------- std BinaryHeap  */
/* Neighbor ordered by distance only (neighbor.rs:30-34).  The algorithm is
 * the std BinaryHeap one, whose text the reference carries in heap.rs:
 * push = append + sift_up (heap.rs:189-193, 212-226); pop = take the last
 * element, swap it with the root, sift_down_to_bottom then sift_up
 * (heap.rs:179-187, 253-270). */
typedef struct {
    float dist;
    uint64_t idx;
} orc_nb;

typedef struct {
    orc_nb *data;
    size_t len, cap;
} orc_heap;

static void heap_sift_up(orc_heap *h, size_t start, size_t pos)
{
    orc_nb elem = h->data[pos];
    while (pos > start) {
        size_t parent = (pos - 1) / 2;
        if (elem.dist <= h->data[parent].dist) break;
        h->data[pos] = h->data[parent];
        pos = parent;
    }
    h->data[pos] = elem;
}

static void heap_sift_down_to_bottom(orc_heap *h, size_t pos)
{
    size_t end = h->len;
    size_t start = pos;
    orc_nb elem = h->data[pos];
    size_t child = 2 * pos + 1;
    while (end >= 2 && child < end - 1) {
        child += (h->data[child].dist <= h->data[child + 1].dist) ? 1 : 0;
        h->data[pos] = h->data[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (end >= 1 && child == end - 1) {
        h->data[pos] = h->data[child];
        pos = child;
    }
    h->data[pos] = elem;
    heap_sift_up(h, start, pos);
}

static int heap_reserve(orc_heap *h, size_t need)
{
    if (need <= h->cap) return 0;
    size_t nc = h->cap ? h->cap * 2 : 16;
    while (nc < need) nc *= 2;
    orc_nb *p = (orc_nb *)realloc(h->data, nc * sizeof(orc_nb));
    if (!p) return -1;
    h->data = p;
    h->cap = nc;
    return 0;
}

static void heap_push(orc_heap *h, orc_nb x)
{
    heap_reserve(h, h->len + 1);
    h->data[h->len] = x;
    h->len += 1;
    heap_sift_up(h, 0, h->len - 1);
}

static orc_nb heap_pop(orc_heap *h)
{
    orc_nb item = h->data[h->len - 1];
    h->len -= 1;
    if (h->len > 0) {
        orc_nb root = h->data[0];
        h->data[0] = item;
        item = root;
        heap_sift_down_to_bottom(h, 0);
    }
    return item;
}

/* Exported for the hand-traced fixtures: run a script of operations
 * (op[i] = 0 push (dist[i], id[i]); op[i] = 1 pop) and return the final
 * layout and the popped sequence. */
int orc_heap_script(const int *ops, const float *dists, const uint64_t *ids, size_t nops,
                    float *out_layout_d, uint64_t *out_layout_i, size_t *out_len,
                    float *popped_d, uint64_t *popped_i, size_t *n_popped)
{
    orc_heap h = {0};
    size_t np = 0;
    for (size_t i = 0; i < nops; ++i) {
        if (ops[i] == 0) {
            orc_nb x = {dists[i], ids[i]};
            heap_push(&h, x);
        } else if (h.len > 0) {
            orc_nb x = heap_pop(&h);
            popped_d[np] = x.dist;
            popped_i[np] = x.idx;
            ++np;
        }
    }
    for (size_t i = 0; i < h.len; ++i) {
        out_layout_d[i] = h.data[i].dist;
        out_layout_i[i] = h.data[i].idx;
    }
    *out_len = h.len;
    *n_popped = np;
    free(h.data);
    return ORC_OK;
}

/* ------------------------------------------------------- BruteForceIndex */
/* bruteforce_idx.rs:47-73: scan all nodes in insertion order, push
 * Neighbor(i, item.metric(node)), pop when len > k, drain by pop and reverse.
 * Query is the first metric operand (bruteforce_idx.rs:56).  Returns the
 * number of results (min(k, n)).  A NaN distance makes the reference panic
 * (neighbor.rs:32): reported here as ORC_ERR_NAN. */
static int bf_one(const float *base, size_t n, size_t d, const float *q, size_t k, int mt,
                  orc_heap *h, uint64_t *out_idx, float *out_dist, uint32_t *out_count)
{
    h->len = 0;
    heap_reserve(h, k + 2);
    int nan_seen = 0;
    for (size_t i = 0; i < n; ++i) {
        float dist;
        int st = orc_metric(q, base + i * d, d, mt, &dist);
        if (st) return st;
        if (dist != dist) nan_seen = 1;
        orc_nb x = {dist, (uint64_t)i};
        heap_push(h, x);
        if (h->len > k) (void)heap_pop(h);
    }
    if (nan_seen && n >= 2) return ORC_ERR_NAN;
    size_t m = h->len;
    for (size_t r = m; r > 0; --r) {
        orc_nb x = heap_pop(h);
        out_idx[r - 1] = x.idx;
        out_dist[r - 1] = x.dist;
    }
    *out_count = (uint32_t)m;
    return ORC_OK;
}

/* Batch driver: queries are independent, so they run in parallel over
 * nthreads host threads (<=0: all); each query's scan is the sequential
 * reference loop above.  Output rows are k wide. */
int orc_bf_search(const float *base, size_t n, size_t d, const float *queries, size_t nq,
                  size_t k, int mt, int nthreads, uint64_t *out_idx, float *out_dist,
                  uint32_t *out_count)
{
    if (mt == ORC_UNKNOWN || mt > ORC_ANGULAR || mt < 0) return ORC_ERR_UNKNOWN_METRIC;
    int status = ORC_OK;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        orc_heap h = {0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (long qi = 0; qi < (long)nq; ++qi) {
            int st = bf_one(base, n, d, queries + (size_t)qi * d, k, mt, &h,
                            out_idx + (size_t)qi * k, out_dist + (size_t)qi * k,
                            out_count + qi);
            if (st) {
#ifdef _OPENMP
#pragma omp critical
#endif
                status = st;
            }
        }
        free(h.data);
    }
    return status;
}

/* Full distance rows in reference order (used to build golden rows). */
int orc_distance_rows(const float *base, size_t n, size_t d, const float *queries, size_t nq,
                      int mt, int nthreads, float *out)
{
    if (mt == ORC_UNKNOWN || mt > ORC_ANGULAR || mt < 0) return ORC_ERR_UNKNOWN_METRIC;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (long qi = 0; qi < (long)nq; ++qi)
        for (size_t i = 0; i < n; ++i)
            orc_metric(queries + (size_t)qi * d, base + i * d, d, mt, out + (size_t)qi * n + i);
    return ORC_OK;
}

/* Heap replay over a given distance row (the tie-exact reference order). */
int orc_replay_row(const float *row, size_t n, size_t k, uint64_t *out_idx, float *out_dist,
                   uint32_t *out_count)
{
    orc_heap h = {0};
    heap_reserve(&h, k + 2);
    for (size_t i = 0; i < n; ++i) {
        orc_nb x = {row[i], (uint64_t)i};
        heap_push(&h, x);
        if (h.len > k) (void)heap_pop(&h);
    }
    size_t m = h.len;
    for (size_t r = m; r > 0; --r) {
        orc_nb x = heap_pop(&h);
        out_idx[r - 1] = x.idx;
        out_dist[r - 1] = x.dist;
    }
    *out_count = (uint32_t)m;
    free(h.data);
    return ORC_OK;
}

/* ------------------------------------------------------------ RNG        */
/* SplitMix64 (Steele, Lea & Flood 2014).  Seeded stand-in for the
 * reference's unseeded rand::thread_rng() in Kmeans::split_center
 * (kmeans.rs:168-171); uniform double in [0,1) from the top 53 bits. */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double splitmix_unit(uint64_t *s)
{
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

uint64_t orc_splitmix64_next(uint64_t *state) { return splitmix64(state); }

/* Per-k-means-instance RNG stream: instance i of a run seeded with `seed`
 * starts from this state (the coarse k-means is instance 0, the PQ
 * sub-quantizer of list l, sub s is instance 1 + l*n_sub + s). */
uint64_t orc_stream_seed(uint64_t seed, uint64_t instance)
{
    uint64_t s = seed ^ (instance * 0xD1B54A32D192ED03ull);
    return splitmix64(&s);
}

/* naive_build_knn_graph (knn.rs:14-41): for node n, every other node i in
 * index order is pushed as Neighbor(i, metric(node_n, node_i)) and the heap
 * popped while len > k; the heap is drained by pop into the row (largest
 * distance first, no reverse).  Node n itself is skipped, not pushed.
 * out_idx/out_dist [n][k] (unused slots UINT64_MAX / +inf), out_count [n].
 * A NaN compared in the heap (k >= 1 and at least 2 other nodes) is the
 * reference's partial_cmp().unwrap() panic: ORC_ERR_NAN. */
int orc_knn_graph(const float *base, size_t n, size_t d, size_t k, int mt, int nthreads,
                  uint64_t *out_idx, float *out_dist, uint32_t *out_count)
{
    if (mt == ORC_UNKNOWN || mt > ORC_ANGULAR || mt < 0) return ORC_ERR_UNKNOWN_METRIC;
    int status = ORC_OK;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        orc_heap h = {0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (long r = 0; r < (long)n; ++r) {
            int nan_seen = 0;
            h.len = 0;
            for (size_t i = 0; i < n; ++i) {
                if (i == (size_t)r) continue;
                float dist;
                orc_metric(base + (size_t)r * d, base + i * d, d, mt, &dist);
                if (dist != dist) nan_seen = 1;
                orc_nb x = {dist, (uint64_t)i};
                heap_push(&h, x);
                if (h.len > k) (void)heap_pop(&h);
            }
            if (nan_seen && k >= 1 && n >= 3) {
#ifdef _OPENMP
#pragma omp critical
#endif
                status = ORC_ERR_NAN;
            }
            size_t m = h.len;
            for (size_t j = 0; j < m; ++j) {
                orc_nb x = heap_pop(&h);
                out_idx[(size_t)r * k + j] = x.idx;
                out_dist[(size_t)r * k + j] = x.dist;
            }
            for (size_t j = m; j < k; ++j) {
                out_idx[(size_t)r * k + j] = UINT64_MAX;
                out_dist[(size_t)r * k + j] = INFINITY;
            }
            out_count[r] = (uint32_t)m;
        }
        free(h.data);
    }
    return status;
}

/* ------------------------------------------------------------ Kmeans     */
/* kmeans.rs:8-214.  Data is addressed as rows of `stride` floats; the
 * k-means works on columns [begin, begin+dim) with an optional residual of
 * full length (kmeans.rs:36-43). */
typedef struct {
    size_t dim, n_center, begin;
    const float *residual; /* full-length vector or NULL */
    int mt;
    float *centers; /* n_center x dim */
} orc_kmeans;

/* kmeans.rs:36-43: z = x[begin..end] (- residual[begin..end]) then metric. */
static float km_distance(const orc_kmeans *km, const float *x, const float *c, float *z)
{
    for (size_t i = 0; i < km->dim; ++i) {
        float v = x[km->begin + i];
        if (km->residual) v = v - km->residual[km->begin + i];
        z[i] = v;
    }
    float out = 0.0f;
    orc_metric(z, c, km->dim, km->mt, &out);
    return out;
}

/* kmeans.rs:50-86.  mean_center[j] += x (or x - residual) in point order,
 * divided by batch_size; centre i gets +1 on dim j when
 * (i & (1 << j)) == 1, which (shift amount masked to 6 bits, release build
 * with overflow-checks = false, Cargo.toml) holds only for odd i and
 * j % 64 == 0; -1 everywhere else. */
static void km_init(orc_kmeans *km, size_t n, const float *data, size_t stride)
{
    size_t dim = km->dim;
    float *mean = (float *)calloc(dim, sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        const float *x = data + i * stride;
        for (size_t j = 0; j < dim; ++j) {
            float v = x[km->begin + j];
            if (km->residual) v = v - km->residual[km->begin + j];
            mean[j] = mean[j] + v;
        }
    }
    for (size_t j = 0; j < dim; ++j) mean[j] = mean[j] / (float)n;
    for (size_t i = 0; i < km->n_center; ++i)
        for (size_t j = 0; j < dim; ++j) {
            uint64_t bit = (uint64_t)1 << (j & 63);
            float v = mean[j];
            if (((uint64_t)i & bit) == 1) v = v + 1.0f;
            else v = v - 1.0f;
            km->centers[i * dim + j] = v;
        }
    free(mean);
}

/* kmeans.rs:127-148: argmin with strict '<', starting at centre 0. */
static void km_search(const orc_kmeans *km, size_t n, const float *data, size_t stride,
                      uint32_t *assign)
{
    float *z = (float *)malloc(km->dim * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        const float *x = data + i * stride;
        size_t best = 0;
        for (size_t j = 1; j < km->n_center; ++j) {
            float dj = km_distance(km, x, km->centers + j * km->dim, z);
            float db = km_distance(km, x, km->centers + best * km->dim, z);
            if (dj < db) best = j;
        }
        assign[i] = (uint32_t)best;
    }
    free(z);
}

/* kmeans.rs:88-125: per-centre f32 sums in point order, divided by the
 * count; an empty centre becomes all zeros. */
static void km_update(orc_kmeans *km, size_t n, const float *data, size_t stride,
                      const uint32_t *assign, uint64_t *counts)
{
    size_t dim = km->dim;
    memset(km->centers, 0, km->n_center * dim * sizeof(float));
    memset(counts, 0, km->n_center * sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) {
        const float *x = data + i * stride;
        uint32_t c = assign[i];
        counts[c] += 1;
        float *ctr = km->centers + (size_t)c * dim;
        for (size_t j = 0; j < dim; ++j) {
            float v = x[km->begin + j];
            if (km->residual) v = v - km->residual[km->begin + j];
            ctr[j] = ctr[j] + v;
        }
    }
    for (size_t c = 0; c < km->n_center; ++c) {
        if (counts[c] == 0) continue;
        for (size_t j = 0; j < dim; ++j)
            km->centers[c * dim + j] = km->centers[c * dim + j] / (float)counts[c];
    }
}

/* kmeans.rs:150-193 with the seeded RNG.  EPS = 1/1024; even dims scale the
 * new centre by (1-EPS) and the source by (1+EPS), odd dims the reverse. */
static int km_split(orc_kmeans *km, size_t n, uint64_t *counts, uint64_t *rng)
{
    if (n == 0) return ORC_ERR_EMPTY;
    const float EPS = 1.0f / 1024.0f;
    size_t nc = km->n_center, dim = km->dim;
    for (size_t i = 0; i < nc; ++i) {
        if (counts[i] != 0) continue;
        size_t s = (i + 1) % nc;
        for (;;) {
            double pick = (double)counts[s] / (double)n;
            if (splitmix_unit(rng) < pick) break;
            s = (s + 1) % nc;
        }
        for (size_t j = 0; j < dim; ++j) {
            float src = km->centers[s * dim + j];
            if (j % 2 == 0) {
                km->centers[i * dim + j] = src * (1.0f - EPS);
                km->centers[s * dim + j] = src * (1.0f + EPS);
            } else {
                km->centers[i * dim + j] = src * (1.0f + EPS);
                km->centers[s * dim + j] = src * (1.0f - EPS);
            }
        }
        counts[i] = counts[s] / 2;
        counts[s] -= counts[i];
    }
    return ORC_OK;
}

/* kmeans.rs:195-207: init, then per epoch search + update (+ split except
 * after the last epoch).  `assign` receives the final search_data result
 * that the callers (pq_idx.rs:144, 370) compute after train. */
static int km_train(orc_kmeans *km, size_t n, const float *data, size_t stride, size_t epochs,
                    uint64_t *rng, uint32_t *assign)
{
    uint64_t *counts = (uint64_t *)malloc(km->n_center * sizeof(uint64_t));
    int st = ORC_OK;
    km_init(km, n, data, stride);
    for (size_t e = 0; e < epochs; ++e) {
        km_search(km, n, data, stride, assign);
        km_update(km, n, data, stride, assign, counts);
        if (e + 1 < epochs) {
            st = km_split(km, n, counts, rng);
            if (st) break;
        }
    }
    if (!st) km_search(km, n, data, stride, assign);
    free(counts);
    return st;
}

/* Exported k-means entry: columns [begin, begin+dim) of rows of `stride`
 * floats, optional residual (length >= begin+dim). */
int orc_kmeans_train(const float *data, size_t n, size_t stride, size_t begin, size_t dim,
                     const float *residual, size_t n_center, int mt, size_t epochs,
                     uint64_t seed, float *centers_out, uint32_t *assign_out)
{
    orc_kmeans km = {dim, n_center, begin, residual, mt, centers_out};
    uint64_t rng = orc_stream_seed(seed, 0);
    return km_train(&km, n, data, stride, epochs, &rng, assign_out);
}

int orc_kmeans_init(const float *data, size_t n, size_t stride, size_t begin, size_t dim,
                    const float *residual, size_t n_center, float *centers_out)
{
    orc_kmeans km = {dim, n_center, begin, residual, ORC_EUCLIDEAN, centers_out};
    km_init(&km, n, data, stride);
    return ORC_OK;
}

int orc_kmeans_assign(const float *data, size_t n, size_t stride, size_t begin, size_t dim,
                      const float *residual, const float *centers, size_t n_center, int mt,
                      int nthreads, uint32_t *assign_out)
{
    orc_kmeans km = {dim, n_center, begin, residual, mt, (float *)centers};
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static, 256)
#endif
    for (long i = 0; i < (long)n; ++i)
        km_search(&km, 1, data + (size_t)i * stride, stride, assign_out + i);
    return ORC_OK;
}

int orc_kmeans_update(const float *data, size_t n, size_t stride, size_t begin, size_t dim,
                      const float *residual, const uint32_t *assign, size_t n_center,
                      float *centers_out, uint64_t *counts_out)
{
    orc_kmeans km = {dim, n_center, begin, residual, ORC_EUCLIDEAN, centers_out};
    km_update(&km, n, data, stride, assign, counts_out);
    return ORC_OK;
}

int orc_kmeans_split(float *centers, size_t n_center, size_t dim, size_t n, uint64_t *counts,
                     uint64_t *rng_state)
{
    orc_kmeans km = {dim, n_center, 0, NULL, ORC_EUCLIDEAN, centers};
    return km_split(&km, n, counts, rng_state);
}

/* ------------------------------------------------------------ PQ         */
/* pq_idx.rs:75-88: sub_dim = d / n_sub; the first d % sub_dim subs get
 * sub_dim+1 dims, the rest sub_dim (trailing dims may be dropped). */
int orc_pq_ranges(size_t d, size_t n_sub, uint32_t *begin, uint32_t *end)
{
    if (n_sub == 0 || d / n_sub == 0) return ORC_ERR_ARG;
    size_t sd = d / n_sub, r = d % sd;
    for (size_t i = 0; i < n_sub; ++i) {
        if (i < r) {
            begin[i] = (uint32_t)(i * (sd + 1));
            end[i] = (uint32_t)((i + 1) * (sd + 1));
        } else {
            begin[i] = (uint32_t)(r * (sd + 1) + (i - r) * sd);
            end[i] = (uint32_t)(r * (sd + 1) + (i + 1 - r) * sd);
        }
    }
    return ORC_OK;
}

/* pq_idx.rs:122-149 (train_center): for each sub-space a Kmeans over the
 * list's items restricted to [begin,end) with the list residual, trained
 * then searched; codebook layout [n_sub][n_center][sub_dim_max] (a sub with
 * fewer dims leaves its tail zero), codes [item][n_sub].  The nested PQ
 * always uses Euclidean (pq_idx.rs:71).  The RNG state runs through the subs
 * in order. */
static int pq_train_mt(const float *items, size_t n, size_t d, size_t n_sub, size_t n_center,
                       const float *residual, int mt, size_t epochs, uint64_t seed,
                       uint64_t first_instance, float *codebook, size_t sub_dim_max,
                       uint32_t *codes)
{
    uint32_t b[256], e[256];
    if (n_sub > 256) return ORC_ERR_ARG;
    int st = orc_pq_ranges(d, n_sub, b, e);
    if (st) return st;
    if (n == 0) {
        /* An empty inverted list makes the reference panic (split_center
         * returns Err for batch_size 0, unwrapped at kmeans.rs:203-204).
         * This build leaves the list's codebook zero and skips it. */
        memset(codebook, 0, n_sub * n_center * sub_dim_max * sizeof(float));
        return ORC_OK;
    }
    uint32_t *assign = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    for (size_t s = 0; s < n_sub; ++s) {
        size_t sdim = e[s] - b[s];
        float *cent = (float *)calloc(n_center * sdim, sizeof(float));
        orc_kmeans km = {sdim, n_center, b[s], residual, mt, cent};
        uint64_t rng = orc_stream_seed(seed, first_instance + s);
        st = km_train(&km, n, items, d, epochs, &rng, assign);
        if (st) {
            free(cent);
            break;
        }
        for (size_t c = 0; c < n_center; ++c)
            for (size_t j = 0; j < sub_dim_max; ++j)
                codebook[(s * n_center + c) * sub_dim_max + j] = j < sdim ? cent[c * sdim + j] : 0.0f;
        for (size_t i = 0; i < n; ++i) codes[i * n_sub + s] = assign[i];
        free(cent);
    }
    free(assign);
    return st;
}

int orc_pq_train(const float *items, size_t n, size_t d, size_t n_sub, size_t n_center,
                 const float *residual, size_t epochs, uint64_t seed, uint64_t first_instance,
                 float *codebook, size_t sub_dim_max, uint32_t *codes)
{
    return pq_train_mt(items, n, d, n_sub, n_center, residual, ORC_EUCLIDEAN, epochs, seed,
                       first_instance, codebook, sub_dim_max, codes);
}

/* Standalone PQIndex::train_center (pq_idx.rs:122-149) after build(mt)
 * (pq_idx.rs:198-202): no residual, and the sub-space k-means use the build
 * metric (Kmeans::new(.., self.mt), pq_idx.rs:133).  The RNG streams of the
 * subs are instances 1..n_sub of the seed. */
int orc_pqindex_train(const float *items, size_t n, size_t d, size_t n_sub, size_t n_center,
                      int mt, size_t epochs, uint64_t seed, float *codebook, size_t sub_dim_max,
                      uint32_t *codes)
{
    if (mt == ORC_UNKNOWN || mt > ORC_ANGULAR || mt < 0) return ORC_ERR_UNKNOWN_METRIC;
    if (n == 0) return ORC_ERR_EMPTY;
    return pq_train_mt(items, n, d, n_sub, n_center, NULL, mt, epochs, seed, 1, codebook,
                       sub_dim_max, codes);
}

/* pq_idx.rs:151-163, 170-179: LUT[s*K + c] = Euclid(fl(q[b..e] - res[b..e]),
 * codeword[s][c]); the residual difference is rounded to f32 first. */
static void pq_lut_mt(const float *q, const float *residual, size_t n_sub, const uint32_t *b,
                      const uint32_t *e, const float *codebook, size_t n_center,
                      size_t sub_dim_max, int mt, float *lut)
{
    float z[4096];
    for (size_t s = 0; s < n_sub; ++s) {
        size_t sdim = e[s] - b[s];
        for (size_t j = 0; j < sdim; ++j) {
            float v = q[b[s] + j];
            if (residual) v = v - residual[b[s] + j];
            z[j] = v;
        }
        for (size_t c = 0; c < n_center; ++c)
            orc_metric(z, codebook + (s * n_center + c) * sub_dim_max, sdim, mt,
                       &lut[s * n_center + c]);
    }
}

static void pq_lut(const float *q, const float *residual, size_t d, size_t n_sub,
                   const uint32_t *b, const uint32_t *e, const float *codebook, size_t n_center,
                   size_t sub_dim_max, float *lut)
{
    float z[1024];
    (void)d;
    for (size_t s = 0; s < n_sub; ++s) {
        size_t sdim = e[s] - b[s];
        for (size_t j = 0; j < sdim; ++j) {
            float v = q[b[s] + j];
            if (residual) v = v - residual[b[s] + j];
            z[j] = v;
        }
        for (size_t c = 0; c < n_center; ++c)
            lut[s * n_center + c] =
                orc_euclidean(z, codebook + (s * n_center + c) * sub_dim_max, sdim);
    }
}

int orc_pq_lut(const float *q, const float *residual, size_t d, size_t n_sub,
               const float *codebook, size_t n_center, size_t sub_dim_max, float *lut)
{
    uint32_t b[256], e[256];
    if (n_sub > 256) return ORC_ERR_ARG;
    int st = orc_pq_ranges(d, n_sub, b, e);
    if (st) return st;
    pq_lut(q, residual, d, n_sub, b, e, codebook, n_center, sub_dim_max, lut);
    return ORC_OK;
}

/* pq_idx.rs:181-191: dist = 0; dist += LUT[j*K + code[j]] for j in order;
 * push every item into an unbounded heap, pop until len <= k.  The heap is
 * returned (layout matters for the merge). */
static void pq_adc_heap(const float *lut, const uint32_t *codes, size_t n, size_t n_sub,
                        size_t n_center, size_t k, orc_heap *h)
{
    h->len = 0;
    for (size_t i = 0; i < n; ++i) {
        float dist = 0.0f;
        for (size_t j = 0; j < n_sub; ++j) dist = dist + lut[j * n_center + codes[i * n_sub + j]];
        orc_nb x = {dist, (uint64_t)i};
        heap_push(h, x);
    }
    while (h->len > k) (void)heap_pop(h);
}

/* Standalone PQIndex::search_knn_adc (pq_idx.rs:165-194) + node_search_k
 * (212-233): LUT with the build metric on the raw query sub-vectors (no
 * residual), ADC over every item in insertion order, all pushed into an
 * unbounded heap, popped down to k, then drained (descending) and reversed.
 * codebook [n_sub][n_center][sub_dim_max], codes [n][n_sub].  A NaN
 * distance among n >= 2 items is the reference's partial_cmp().unwrap()
 * panic (neighbor.rs:32): ORC_ERR_NAN. */
int orc_pqindex_search(const float *queries, size_t nq, size_t d, int mt, const float *codebook,
                       size_t n_sub, size_t n_center, size_t sub_dim_max, const uint32_t *codes,
                       size_t n, size_t k, int nthreads, uint64_t *out_idx, float *out_dist,
                       uint32_t *out_count)
{
    uint32_t b[256], e[256];
    if (mt == ORC_UNKNOWN || mt > ORC_ANGULAR || mt < 0) return ORC_ERR_UNKNOWN_METRIC;
    if (n_sub > 256) return ORC_ERR_ARG;
    int st = orc_pq_ranges(d, n_sub, b, e);
    if (st) return st;
    int status = ORC_OK;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        orc_heap h = {0};
        float *lut = (float *)malloc(n_sub * n_center * sizeof(float));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (long qi = 0; qi < (long)nq; ++qi) {
            pq_lut_mt(queries + (size_t)qi * d, NULL, n_sub, b, e, codebook, n_center,
                      sub_dim_max, mt, lut);
            int nan_seen = 0;
            h.len = 0;
            for (size_t i = 0; i < n; ++i) {
                float dist = 0.0f;
                for (size_t j = 0; j < n_sub; ++j)
                    dist = dist + lut[j * n_center + codes[i * n_sub + j]];
                if (dist != dist) nan_seen = 1;
                orc_nb x = {dist, (uint64_t)i};
                heap_push(&h, x);
            }
            if (nan_seen && n >= 2) {
#ifdef _OPENMP
#pragma omp critical
#endif
                status = ORC_ERR_NAN;
            }
            while (h.len > k) (void)heap_pop(&h);
            size_t m = h.len;
            for (size_t r = m; r > 0; --r) {
                orc_nb x = heap_pop(&h);
                out_idx[(size_t)qi * k + r - 1] = x.idx;
                out_dist[(size_t)qi * k + r - 1] = x.dist;
            }
            for (size_t r = m; r < k; ++r) {
                out_idx[(size_t)qi * k + r] = UINT64_MAX;
                out_dist[(size_t)qi * k + r] = INFINITY;
            }
            out_count[qi] = (uint32_t)m;
        }
        free(lut);
        free(h.data);
    }
    return status;
}

/* LUT of the standalone PQIndex (pq_idx.rs:170-179) with the build metric. */
int orc_pqindex_lut(const float *q, size_t d, size_t n_sub, const float *codebook,
                    size_t n_center, size_t sub_dim_max, int mt, float *lut)
{
    uint32_t b[256], e[256];
    if (n_sub > 256) return ORC_ERR_ARG;
    int st = orc_pq_ranges(d, n_sub, b, e);
    if (st) return st;
    pq_lut_mt(q, NULL, n_sub, b, e, codebook, n_center, sub_dim_max, mt, lut);
    return ORC_OK;
}

/* IVFPQIndex::search_knn_adc (pq_idx.rs:408-439) + node_search_k (458-477).
 * With given_probes ([nq][nprobe], given_counts[nq] used per query) the
 * coarse heap is skipped and the given lists are scanned in that order
 * (the learned-probe form, which has no reference counterpart; the per-list
 * scan and the merge are the reference's).  The index is given as arrays:
 *   centers [nlist][d], codebooks [nlist][n_sub][n_center][sub_dim_max],
 *   list_offsets [nlist+1], list_ids [n] (global ids, ascending per list),
 *   codes [n][n_sub] in list order (row r belongs to list_ids[r]).
 * Coarse: push Neighbor(i, -metric(q, c_i)) for all i, pop nprobe.
 * Per probed list: LUT + ADC heap trimmed to k, popped descending, remapped
 * and pushed into the global heap trimmed to k.  Final drain ascending. */
int orc_ivfpq_search(const float *queries, size_t nq, size_t d, int mt, const float *centers,
                     size_t nlist, const float *codebooks, size_t n_sub, size_t n_center,
                     size_t sub_dim_max, const uint64_t *list_offsets, const uint64_t *list_ids,
                     const uint32_t *codes, size_t nprobe, size_t k, int nthreads,
                     uint64_t *out_idx, float *out_dist, uint32_t *out_count,
                     uint32_t *out_probes, const uint32_t *given_probes,
                     const uint32_t *given_counts)
{
    uint32_t b[256], e[256];
    if (n_sub > 256) return ORC_ERR_ARG;
    if (nprobe > nlist) return ORC_ERR_ARG; /* reference: unwrap panic, pq_idx.rs:425 */
    int st = orc_pq_ranges(d, n_sub, b, e);
    if (st) return st;
    int status = ORC_OK;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        orc_heap coarse = {0}, top = {0}, local = {0};
        float *lut = (float *)malloc(n_sub * n_center * sizeof(float));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (long qi = 0; qi < (long)nq; ++qi) {
            const float *q = queries + (size_t)qi * d;
            coarse.len = 0;
            top.len = 0;
            for (size_t i = 0; i < nlist && !given_probes; ++i) {
                float dist;
                int s2 = orc_metric(q, centers + i * d, d, mt, &dist);
                if (s2) {
#ifdef _OPENMP
#pragma omp critical
#endif
                    status = s2;
                    break;
                }
                orc_nb x = {-dist, (uint64_t)i};
                heap_push(&coarse, x);
            }
            size_t np_q = given_probes ? given_counts[qi] : nprobe;
            for (size_t p = 0; p < np_q && (given_probes || coarse.len > 0); ++p) {
                size_t c = given_probes ? given_probes[(size_t)qi * nprobe + p]
                                        : (size_t)heap_pop(&coarse).idx;
                if (out_probes) out_probes[(size_t)qi * nprobe + p] = (uint32_t)c;
                size_t off = list_offsets[c], len = list_offsets[c + 1] - off;
                pq_lut(q, centers + c * d, d, n_sub, b, e,
                       codebooks + c * n_sub * n_center * sub_dim_max, n_center, sub_dim_max, lut);
                pq_adc_heap(lut, codes + off * n_sub, len, n_sub, n_center, k, &local);
                while (local.len > 0) {
                    orc_nb x = heap_pop(&local);
                    x.idx = list_ids[off + x.idx];
                    heap_push(&top, x);
                    if (top.len > k) (void)heap_pop(&top);
                }
            }
            size_t m = top.len;
            for (size_t r = m; r > 0; --r) {
                orc_nb x = heap_pop(&top);
                out_idx[(size_t)qi * k + r - 1] = x.idx;
                out_dist[(size_t)qi * k + r - 1] = x.dist;
            }
            out_count[qi] = (uint32_t)m;
        }
        free(lut);
        free(coarse.data);
        free(top.data);
        free(local.data);
    }
    return status;
}

/* PQ encode given codebooks: per sub-space the Kmeans search_data of
 * pq_idx.rs:143-144 (argmin, strict '<', residual subtracted first). */
int orc_pq_encode(const float *items, size_t n, size_t d, size_t n_sub, size_t n_center,
                  const float *residual, const float *codebook, size_t sub_dim_max,
                  uint32_t *codes)
{
    uint32_t b[256], e[256];
    if (n_sub > 256) return ORC_ERR_ARG;
    int st = orc_pq_ranges(d, n_sub, b, e);
    if (st) return st;
    uint32_t *assign = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    float *cent = (float *)malloc(n_center * sub_dim_max * sizeof(float));
    for (size_t s = 0; s < n_sub; ++s) {
        size_t sdim = e[s] - b[s];
        for (size_t c = 0; c < n_center; ++c)
            for (size_t j = 0; j < sdim; ++j)
                cent[c * sdim + j] = codebook[(s * n_center + c) * sub_dim_max + j];
        orc_kmeans km = {sdim, n_center, b[s], residual, ORC_EUCLIDEAN, cent};
        km_search(&km, n, items, d, assign);
        for (size_t i = 0; i < n; ++i) codes[i * n_sub + s] = assign[i];
    }
    free(assign);
    free(cent);
    return ORC_OK;
}
