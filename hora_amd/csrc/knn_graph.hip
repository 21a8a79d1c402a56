// knn_graph.hip -- the all-pairs kNN graph of naive_build_knn_graph
// (knn.rs:14-41) and SSGIndex::build_knn_graph (ssg_idx.rs:69-98).
//
// Row n of the graph streams every other node i != n, in index order,
// through a std BinaryHeap (push, pop while len > k) and drains it by pop,
// so the row holds the k nearest in DESCENDING distance order.  Node n itself
// is never pushed.  On the device: the exact distance block of a chunk of
// nodes against all nodes, the node's own column set to +inf for the radix
// select only, and then
//   - rows whose k+1 nearest others are pairwise distinct: those others,
//     written largest first (graph_finalize_kernel);
//   - the other rows: a copy of the row with the node's own column removed
//     (graph_compact_kernel) goes through the heap replay of the
//     BruteForceIndex path, whose ascending output is remapped to node ids
//     and reversed (graph_remap_kernel).
#include "kernels.h"
#include "profiler.h"

namespace hm {

namespace {

__global__ void graph_diag_kernel(float *__restrict__ D, uint64_t ldd, uint32_t rows, uint32_t r0)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) D[(size_t)r * ldd + r0 + r] = __int_as_float(0x7f800000);
}

// One lane per row.  cand: the Ksel smallest of the row by (distance,
// node), the node's own entry among them or not.
__global__ void graph_finalize_kernel(const float *__restrict__ cand_d,
                                      const uint32_t *__restrict__ cand_i, uint32_t ldc,
                                      uint32_t Ksel, const uint32_t *__restrict__ flags,
                                      uint32_t rows, uint32_t r0, uint32_t n, uint32_t k,
                                      uint64_t *__restrict__ out_idx, float *__restrict__ out_dist,
                                      uint32_t *__restrict__ out_count, uint32_t *__restrict__ tied,
                                      uint32_t *__restrict__ err)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t self = r0 + r, others = n - 1;
    const uint32_t m1 = min(k + 1, others), m = min(k, others);
    const float *cd = cand_d + (size_t)r * ldc;
    const uint32_t *ci = cand_i + (size_t)r * ldc;
    uint64_t *oi = out_idx + (size_t)self * k;
    float *od = out_dist + (size_t)self * k;
    // the j-th other candidate: skip the node's own entry
    auto pos = [&](uint32_t j) -> uint32_t {
        uint32_t p = 0, seen = 0;
        for (; p < Ksel; ++p) {
            if (ci[p] == self) continue;
            if (seen == j) break;
            ++seen;
        }
        return p;
    };
    bool tie = false;
    float prev = 0.0f;
    for (uint32_t j = 0, p = 0; j < m1; ++j, ++p) {
        while (ci[p] == self) ++p;
        const float v = cd[p];
        if (j > 0 && v == prev) tie = true;
        prev = v;
    }
    const bool nan = (flags[r] & ROW_NAN) && k >= 1 && others >= 2;
    if (nan) atomicOr(err, 1u);
    tied[r] = (tie && !nan) ? 1u : 0u;
    out_count[self] = m;
    const float inf = __int_as_float(0x7f800000);
    for (uint32_t j = m; j < k; ++j) {
        oi[j] = ~0ull;
        od[j] = inf;
    }
    if (tie) return;
    for (uint32_t j = 0; j < m; ++j) {
        const uint32_t p = pos(m - 1 - j);
        oi[j] = ci[p];
        od[j] = cd[p];
    }
}

// C[t][j] = D[rows[t]][j < self ? j : j + 1] for j < n - 1.
__global__ void graph_compact_kernel(const float *__restrict__ D, uint64_t ldd,
                                     const uint32_t *__restrict__ rows, uint32_t r0, uint32_t n,
                                     float *__restrict__ C, uint64_t ldc)
{
    const uint32_t t = blockIdx.y;
    const uint32_t r = rows[t], self = r0 + r;
    const float *src = D + (size_t)r * ldd;
    float *dst = C + (size_t)t * ldc;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j + 1 < n; j += gridDim.x * blockDim.x)
        dst[j] = src[j < self ? j : j + 1];
}

// The replay's ascending output (compacted columns) -> node ids, largest
// first.
__global__ void graph_remap_kernel(const uint64_t *__restrict__ ti, const float *__restrict__ td,
                                   const uint32_t *__restrict__ tc, const uint32_t *__restrict__ rows,
                                   uint32_t nt, uint32_t r0, uint32_t k, uint64_t *__restrict__ out_idx,
                                   float *__restrict__ out_dist)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t self = r0 + rows[t], m = tc[t];
    for (uint32_t j = 0; j < m; ++j) {
        const uint64_t c = ti[(size_t)t * k + m - 1 - j];
        out_idx[(size_t)self * k + j] = c < self ? c : c + 1;
        out_dist[(size_t)self * k + j] = td[(size_t)t * k + m - 1 - j];
    }
}

}  // namespace

void launch_graph_diag(float *D, uint64_t ldd, uint32_t rows, uint32_t r0, hipStream_t s)
{
    graph_diag_kernel<<<ceil_div(rows, 256), 256, 0, s>>>(D, ldd, rows, r0);
    HM_CHECK_LAUNCH();
}

void launch_graph_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                           uint32_t Ksel, const uint32_t *flags, uint32_t rows, uint32_t r0,
                           uint32_t n, uint32_t k, uint64_t *out_idx, float *out_dist,
                           uint32_t *out_count, uint32_t *tied, uint32_t *err, hipStream_t s)
{
    ProfScope ps("graph_finalize", s);
    graph_finalize_kernel<<<ceil_div(rows, 64), 64, 0, s>>>(cand_d, cand_i, ldc, Ksel, flags, rows,
                                                            r0, n, k, out_idx, out_dist, out_count,
                                                            tied, err);
    HM_CHECK_LAUNCH();
}

void launch_graph_compact(const float *D, uint64_t ldd, const uint32_t *rows, uint32_t nt,
                          uint32_t r0, uint32_t n, float *C, uint64_t ldc, hipStream_t s)
{
    if (!nt || n < 2) return;
    dim3 grid(std::min<unsigned>(ceil_div(n - 1, 256), 64), nt);
    graph_compact_kernel<<<grid, 256, 0, s>>>(D, ldd, rows, r0, n, C, ldc);
    HM_CHECK_LAUNCH();
}

void launch_graph_remap(const uint64_t *ti, const float *td, const uint32_t *tc,
                        const uint32_t *rows, uint32_t nt, uint32_t r0, uint32_t k,
                        uint64_t *out_idx, float *out_dist, hipStream_t s)
{
    if (!nt) return;
    graph_remap_kernel<<<ceil_div(nt, 64), 64, 0, s>>>(ti, td, tc, rows, nt, r0, k, out_idx,
                                                       out_dist);
    HM_CHECK_LAUNCH();
}

}  // namespace hm
