// select.hip -- per-row top-K selection, tie detection and the std-heap replay.
//
// The reference keeps its top-k in std::collections::BinaryHeap<Neighbor>
// ordered by distance only (bruteforce_idx.rs:47-73, neighbor.rs:30-34).
// When the k+1 smallest distances of a row are pairwise distinct, the
// reference output is exactly the k smallest in ascending order, whatever
// the heap layout; this file computes that with a radix select over the
// (distance, index) order.  When they are not distinct, the reference's
// ids and order depend on the whole push/pop stream, so the row is flagged
// and bf_finalize replays the heap algorithm (the std BinaryHeap text the
// reference carries in heap.rs:179-270) over the row in index order.
#include "kernels.h"
#include "profiler.h"

namespace hm {

namespace {

constexpr int SEL_THREADS = 256;
constexpr uint32_t SEL_CAP = 8192;  // candidate slots in LDS, power of two

// Order-preserving map of a float to uint32 (ascending float order ==
// ascending key).  -0.0 is canonicalised to +0.0 first: the reference
// compares with partial_cmp, for which they are equal.
__device__ __forceinline__ uint32_t fkey(float f)
{
    f = f + 0.0f;
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Exclusive scan of one value per thread over a 256-thread block (4 waves
// of 64).  Every thread must call it; *total gets the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_w, uint32_t *total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SEL_THREADS / 64; ++i) {
        uint32_t t = s_w[i];
        if (i < w) before += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

__global__ __launch_bounds__(SEL_THREADS) void row_select_kernel(
    const float *__restrict__ D, uint64_t ldd, uint32_t n, uint32_t K, float *__restrict__ cand_d,
    uint32_t *__restrict__ cand_i, uint32_t ldc, uint32_t *__restrict__ flags,
    const uint32_t *__restrict__ dev_rows)
{
    if (dev_rows && blockIdx.x >= *dev_rows) return;  // graph replays with fewer rows
    __shared__ uint32_t hist[2048];
    __shared__ unsigned long long cand[SEL_CAP];
    __shared__ uint32_t s_w[SEL_THREADS / 64];
    __shared__ uint32_t s_prefix, s_remaining, s_bucket, s_cnt, s_nan, s_tie;

    const int tid = threadIdx.x;
    const float *row = D + (size_t)blockIdx.x * ldd;
    if (tid == 0) {
        s_prefix = 0;
        s_remaining = K;
        s_bucket = 0;
        s_nan = 0;
        s_tie = 0;
        s_cnt = 0;
    }

    // Radix select of the K-th smallest key: digits of 11, 11 and 10 bits.
    // After each pass, when every key up to the target bucket fits in LDS,
    // they are collected and sorted directly (usually after the first pass).
    uint32_t mask = 0;
    bool early = false;
#pragma unroll 1
    for (int pass = 0; pass < 3; ++pass) {
        const int shift = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
        const uint32_t nb = pass < 2 ? 2048u : 1024u;
        for (int i = tid; i < 2048; i += SEL_THREADS) hist[i] = 0;
        __syncthreads();
        const uint32_t prefix = s_prefix, remaining = s_remaining;
        uint32_t nan_local = 0;
        for (uint32_t i = tid; i < n; i += SEL_THREADS) {
            float f = row[i];
            if (pass == 0 && f != f) nan_local = 1;
            uint32_t key = fkey(f);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & (nb - 1)], 1u);
        }
        if (nan_local) s_nan = 1;
        __syncthreads();
        const uint32_t per = nb / SEL_THREADS;
        uint32_t local = 0;
        for (uint32_t b = 0; b < per; ++b) local += hist[tid * per + b];
        uint32_t tot;
        uint32_t excl = block_excl_scan(local, s_w, &tot);
        if (local > 0 && excl < remaining && remaining <= excl + local) {
            uint32_t c = excl;
            for (uint32_t b = 0; b < per; ++b) {
                uint32_t h = hist[tid * per + b];
                if (c + h >= remaining) {
                    s_prefix = prefix | ((tid * per + b) << shift);
                    s_remaining = remaining - c;
                    s_bucket = h;
                    break;
                }
                c += h;
            }
        }
        __syncthreads();
        mask |= (nb - 1) << shift;
        if ((K - s_remaining) + s_bucket <= SEL_CAP) {
            early = true;
            break;
        }
    }

    uint32_t total;  // candidates in LDS, sorted below; the first K are kept
    bool bad = false;
    const uint32_t T = s_prefix, rem = s_remaining, ceq = s_bucket;
    if (early) {
        // every key whose top bits are <= the target prefix
        total = (K - rem) + ceq;
        for (uint32_t i = tid; i < n; i += SEL_THREADS) {
            uint32_t key = fkey(row[i]);
            if ((key & mask) <= T) {
                uint32_t pos = atomicAdd(&s_cnt, 1u);
                if (pos < SEL_CAP) cand[pos] = ((unsigned long long)key << 32) | i;
            }
        }
        __syncthreads();
        bad = s_cnt != total;
    } else {
        // Collect: every key < T, then the `rem` lowest indices with key == T.
        total = K;
        const uint32_t nlt = K - rem;
        const bool all_eq = (ceq == rem);
        for (uint32_t i = tid; i < n; i += SEL_THREADS) {
            uint32_t key = fkey(row[i]);
            if (key < T || (all_eq && key == T)) {
                uint32_t pos = atomicAdd(&s_cnt, 1u);
                if (pos < SEL_CAP) cand[pos] = ((unsigned long long)key << 32) | i;
            }
        }
        __syncthreads();
        bad = s_cnt != (all_eq ? K : nlt);
        if (!all_eq && !bad) {
            uint32_t taken = 0;  // uniform over the block
            for (uint32_t base = 0; base < n && taken < rem; base += SEL_THREADS) {
                uint32_t i = base + tid;
                uint32_t p = (i < n && fkey(row[i]) == T) ? 1u : 0u;
                uint32_t tot;
                uint32_t ex = block_excl_scan(p, s_w, &tot);
                if (p && taken + ex < rem) cand[nlt + taken + ex] = ((unsigned long long)T << 32) | i;
                taken += tot;
            }
        }
    }
    if (bad) {  // never expected: keep every index in range and flag the row
        total = K;
        for (uint32_t j = tid; j < K; j += SEL_THREADS) cand[j] = ~0ull << 32;
    }

    // Bitonic sort of the candidates (padded to a power of two).
    uint32_t P = 1;
    while (P < total) P <<= 1;
    for (uint32_t j = total + tid; j < P; j += SEL_THREADS) cand[j] = ~0ull;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < P / 2; t += SEL_THREADS) {
                uint32_t lo = 2 * stride * (t / stride) + (t % stride);
                uint32_t hi = lo + stride;
                bool asc = (lo & size) == 0;
                unsigned long long a = cand[lo], b = cand[hi];
                if ((a > b) == asc) {
                    cand[lo] = b;
                    cand[hi] = a;
                }
            }
            __syncthreads();
        }
    }

    float *od = cand_d + (size_t)blockIdx.x * ldc;
    uint32_t *oi = cand_i + (size_t)blockIdx.x * ldc;
    uint32_t tie = 0;
    for (uint32_t j = tid; j < K; j += SEL_THREADS) {
        unsigned long long c = cand[j];
        uint32_t idx = (uint32_t)c;
        oi[j] = idx;
        od[j] = idx < n ? row[idx] : __int_as_float(0x7f800000);
        if (j + 1 < K && (uint32_t)(c >> 32) == (uint32_t)(cand[j + 1] >> 32)) tie = 1;
    }
    if (tie) s_tie = 1;
    __syncthreads();
    if (tid == 0)
        flags[blockIdx.x] =
            (s_tie ? ROW_TIE : 0u) | (s_nan ? ROW_NAN : 0u) | (bad ? ROW_BAD : 0u);
}

// ---- std BinaryHeap model (max-heap on distance only) on a per-lane array.
struct LaneHeap {
    float *d;
    uint32_t *ix;
    uint32_t len;

    __device__ void sift_up(uint32_t start, uint32_t pos)
    {
        float ed = d[pos];
        uint32_t ei = ix[pos];
        while (pos > start) {
            uint32_t parent = (pos - 1) / 2;
            if (ed <= d[parent]) break;
            d[pos] = d[parent];
            ix[pos] = ix[parent];
            pos = parent;
        }
        d[pos] = ed;
        ix[pos] = ei;
    }
    __device__ void sift_down_to_bottom(uint32_t pos)
    {
        uint32_t end = len, start = pos;
        float ed = d[pos];
        uint32_t ei = ix[pos];
        uint32_t child = 2 * pos + 1;
        while (end >= 2 && child < end - 1) {
            child += (d[child] <= d[child + 1]) ? 1u : 0u;
            d[pos] = d[child];
            ix[pos] = ix[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (end >= 1 && child == end - 1) {
            d[pos] = d[child];
            ix[pos] = ix[child];
            pos = child;
        }
        d[pos] = ed;
        ix[pos] = ei;
        sift_up(start, pos);
    }
    __device__ void push(float x, uint32_t i)
    {
        d[len] = x;
        ix[len] = i;
        len += 1;
        sift_up(0, len - 1);
    }
    __device__ void pop(float *od, uint32_t *oi)
    {
        float ld = d[len - 1];
        uint32_t li = ix[len - 1];
        len -= 1;
        if (len > 0) {
            *od = d[0];
            *oi = ix[0];
            d[0] = ld;
            ix[0] = li;
            sift_down_to_bottom(0);
        } else {
            *od = ld;
            *oi = li;
        }
    }
};

// ---- std BinaryHeap held across the lanes of one wave: lane j owns slot j.
// Every operation runs with uniform control flow; a slot is read with
// v_readlane (uniform index) and written by the owning lane.
struct WaveHeap {
    float d;     // this lane's slot
    uint32_t i;  // this lane's id
    uint32_t len;
    int lane;

    __device__ float gd(uint32_t p) const
    {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), (int)p));
    }
    __device__ uint32_t gi(uint32_t p) const
    {
        return (uint32_t)__builtin_amdgcn_readlane((int)i, (int)p);
    }
    __device__ void put(uint32_t p, float v, uint32_t id)
    {
        if ((uint32_t)lane == p) {
            d = v;
            i = id;
        }
    }
    __device__ void sift_up(uint32_t start, uint32_t pos)
    {
        const float ed = gd(pos);
        const uint32_t ei = gi(pos);
        while (pos > start) {
            const uint32_t parent = (pos - 1) / 2;
            const float pd = gd(parent);
            if (ed <= pd) break;
            put(pos, pd, gi(parent));
            pos = parent;
        }
        put(pos, ed, ei);
    }
    __device__ void sift_down_to_bottom(uint32_t pos)
    {
        const uint32_t end = len, start = pos;
        const float ed = gd(pos);
        const uint32_t ei = gi(pos);
        uint32_t child = 2 * pos + 1;
        while (end >= 2 && child < end - 1) {
            child += (gd(child) <= gd(child + 1)) ? 1u : 0u;
            put(pos, gd(child), gi(child));
            pos = child;
            child = 2 * pos + 1;
        }
        if (end >= 1 && child == end - 1) {
            put(pos, gd(child), gi(child));
            pos = child;
        }
        put(pos, ed, ei);
        sift_up(start, pos);
    }
    __device__ void push(float x, uint32_t id)
    {
        put(len, x, id);
        len += 1;
        sift_up(0, len - 1);
    }
    __device__ void pop(float *od, uint32_t *oi)
    {
        const float ld = gd(len - 1);
        const uint32_t li = gi(len - 1);
        len -= 1;
        if (len > 0) {
            *od = gd(0);
            *oi = gi(0);
            put(0, ld, li);
            sift_down_to_bottom(0);
        } else {
            *od = ld;
            *oi = li;
        }
    }
};

// Replays bruteforce_idx.rs:47-73 over a tied row with the heap in the
// lanes of one wave (k + 1 <= 64).  An element x greater than the heap's
// maximum is pushed and popped again; that round trip permutes the layout
// in a way that depends on the layout only, so when it is known to leave
// the current layout unchanged (checked once after every change) such
// elements are skipped.  One workgroup of 64 lanes per row; rows without a
// tie copy their candidates.
__global__ __launch_bounds__(64) void bf_finalize_wave_kernel(
    const float *__restrict__ cand_d, const uint32_t *__restrict__ cand_i, uint32_t ldc,
    const uint32_t *__restrict__ flags, const float *__restrict__ D, uint64_t ldd, uint32_t n,
    uint32_t k, uint64_t *__restrict__ out_idx, float *__restrict__ out_dist,
    uint32_t *__restrict__ out_count, uint32_t *err)
{
    const uint32_t r = blockIdx.x;
    const int lane = threadIdx.x;
    const float inf = __int_as_float(0x7f800000);
    uint64_t *oi = out_idx + (size_t)r * k;
    float *od = out_dist + (size_t)r * k;
    const uint32_t m = k < n ? k : n;
    const uint32_t f = flags[r];
    if ((f & (ROW_NAN | ROW_BAD)) && n >= 2) {
        if (lane == 0) atomicOr(err, (f & ROW_BAD) ? 4u : 1u);
        for (uint32_t j = lane; j < k; j += 64) {
            oi[j] = ~0ull;
            od[j] = inf;
        }
        if (lane == 0) out_count[r] = 0;
        return;
    }
    if (!(f & ROW_TIE)) {
        const float *cd = cand_d + (size_t)r * ldc;
        const uint32_t *ci = cand_i + (size_t)r * ldc;
        for (uint32_t j = lane; j < k; j += 64) {
            oi[j] = j < m ? (uint64_t)ci[j] : ~0ull;
            od[j] = j < m ? cd[j] : inf;
        }
        if (lane == 0) out_count[r] = m;
        return;
    }
    WaveHeap h{0.0f, 0u, 0u, lane};
    const float *row = D + (size_t)r * ldd;
    bool stable = false;
    float root = inf;
    // The row streams through LDS in batches of FB x 64 values; the next
    // batch's loads are in flight while the current one is replayed.  While
    // the heap is full and its layout is stable, a ballot finds the next
    // element that is not greater than the root (the only ones that change
    // anything), so the rejected majority costs one ballot per 64 values.
    constexpr uint32_t FB = 16;
    __shared__ float sv[FB * 64];
    float nx[FB];
#pragma unroll
    for (uint32_t c = 0; c < FB; ++c) nx[c] = row[min((uint32_t)(c * 64 + lane), n - 1)];
    for (uint32_t b0 = 0; b0 < n; b0 += FB * 64) {
#pragma unroll
        for (uint32_t c = 0; c < FB; ++c) sv[c * 64 + lane] = nx[c];
        __syncthreads();
        const uint32_t b1 = b0 + FB * 64;
        if (b1 < n) {
#pragma unroll
            for (uint32_t c = 0; c < FB; ++c) nx[c] = row[min(b1 + c * 64 + lane, n - 1)];
        }
        for (uint32_t c = 0; c < FB; ++c) {
            const uint32_t base = b0 + c * 64;
            if (base >= n) break;
            const float v = sv[c * 64 + lane];
            const uint32_t cnt = n - base < 64 ? n - base : 64;
            uint32_t j = 0;
            while (j < cnt) {
                if (h.len == k && stable) {
                    const unsigned long long live =
                        __ballot((uint32_t)lane >= j && (uint32_t)lane < cnt && !(v > root));
                    if (live == 0) break;
                    j = (uint32_t)__builtin_ctzll(live);
                }
                const float x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)j));
                h.push(x, base + j);
                if (h.len > k) {
                    float a;
                    uint32_t b;
                    h.pop(&a, &b);
                }
                if (h.len == k) {
                    root = h.gd(0);
                    // would a larger element's push + pop leave this layout as is?
                    WaveHeap t = h;
                    t.push(inf, 0xFFFFFFFFu);
                    float a;
                    uint32_t b;
                    t.pop(&a, &b);
                    const bool same = (uint32_t)lane >= k || (t.d == h.d && t.i == h.i);
                    stable = root < inf && __all(same);
                }
                ++j;
            }
        }
        __syncthreads();
    }
    const uint32_t cntk = h.len;
    for (uint32_t p = cntk; p > 0; --p) {
        float a;
        uint32_t b;
        h.pop(&a, &b);
        if (lane == 0) {
            oi[p - 1] = b;
            od[p - 1] = a;
        }
    }
    for (uint32_t j = cntk + lane; j < k; j += 64) {
        oi[j] = ~0ull;
        od[j] = inf;
    }
    if (lane == 0) out_count[r] = cntk;
}

template <bool LDS_HEAP>
__global__ __launch_bounds__(64) void bf_finalize_kernel(
    const float *__restrict__ cand_d, const uint32_t *__restrict__ cand_i, uint32_t ldc,
    const uint32_t *__restrict__ flags, const float *__restrict__ D, uint64_t ldd, uint32_t rows,
    uint32_t n, uint32_t k, uint64_t *__restrict__ out_idx, float *__restrict__ out_dist,
    uint32_t *__restrict__ out_count, uint32_t *err, float *heap_d, uint32_t *heap_i)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    if (r >= rows) return;
    const float inf = __int_as_float(0x7f800000);
    uint64_t *oi = out_idx + (size_t)r * k;
    float *od = out_dist + (size_t)r * k;
    const uint32_t m = k < n ? k : n;
    const uint32_t f = flags[r];
    if (k == 0) {
        out_count[r] = 0;
        return;
    }
    if ((f & (ROW_NAN | ROW_BAD)) && n >= 2) {
        atomicOr(err, (f & ROW_BAD) ? 4u : 1u);
        for (uint32_t j = 0; j < k; ++j) {
            oi[j] = ~0ull;
            od[j] = inf;
        }
        out_count[r] = 0;
        return;
    }
    uint32_t cnt = m;
    if (!(f & ROW_TIE)) {
        const float *cd = cand_d + (size_t)r * ldc;
        const uint32_t *ci = cand_i + (size_t)r * ldc;
        for (uint32_t j = 0; j < m; ++j) {
            oi[j] = ci[j];
            od[j] = cd[j];
        }
    } else {
        LaneHeap h;
        if constexpr (LDS_HEAP) {
            float *base_d = reinterpret_cast<float *>(smem);
            uint32_t *base_i = reinterpret_cast<uint32_t *>(smem + (size_t)64 * (k + 1) * 4);
            h.d = base_d + threadIdx.x * (k + 1);
            h.ix = base_i + threadIdx.x * (k + 1);
        } else {
            h.d = heap_d + (size_t)r * (k + 1);
            h.ix = heap_i + (size_t)r * (k + 1);
        }
        h.len = 0;
        const float *row = D + (size_t)r * ldd;
        for (uint32_t i = 0; i < n; ++i) {
            h.push(row[i], i);
            if (h.len > k) {
                float xd;
                uint32_t xi;
                h.pop(&xd, &xi);
            }
        }
        cnt = h.len;
        for (uint32_t p = cnt; p > 0; --p) {
            float xd;
            uint32_t xi;
            h.pop(&xd, &xi);
            oi[p - 1] = xi;
            od[p - 1] = xd;
        }
    }
    for (uint32_t j = cnt; j < k; ++j) {
        oi[j] = ~0ull;
        od[j] = inf;
    }
    out_count[r] = cnt;
}

__global__ void check_finite_kernel(const float *__restrict__ X, uint64_t ld, uint64_t rows,
                                    uint32_t d, uint32_t *err)
{
    uint64_t total = rows * d;
    uint32_t bad = 0;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = e / d, c = e % d;
        float v = X[r * ld + c];
        if (!isfinite(v)) bad = 1;
    }
    if (bad) atomicOr(err, 2u);
}

__global__ void pad_rows_kernel(const float *__restrict__ src, uint64_t lds, float *__restrict__ dst,
                                uint64_t ldd, uint64_t rows, uint32_t d, uint32_t dpad,
                                const uint32_t *__restrict__ dev_rows)
{
    if (dev_rows && *dev_rows < rows) rows = *dev_rows;
    uint64_t total = rows * dpad;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = e / dpad, c = e % dpad;
        dst[r * ldd + c] = c < d ? src[r * lds + c] : 0.0f;
    }
}

}  // namespace

void launch_row_select(const float *D, uint64_t ldd, uint32_t rows, uint32_t n, uint32_t K,
                       float *cand_d, uint32_t *cand_i, uint32_t ldc, uint32_t *flags,
                       hipStream_t s, const uint32_t *dev_rows)
{
    if (rows == 0 || n == 0 || K == 0) return;
    if (K > SEL_CAP || K > n) fail(HM_ERR_INVALID_ARG, "row_select: K out of range");
    ProfScope prof("row_select", s);
    row_select_kernel<<<rows, SEL_THREADS, 0, s>>>(D, ldd, n, K, cand_d, cand_i, ldc, flags,
                                                   dev_rows);
    HM_CHECK_LAUNCH();
}

void launch_bf_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                        const uint32_t *flags, const float *D, uint64_t ldd, uint32_t rows,
                        uint32_t n, uint32_t k, uint64_t *out_idx, float *out_dist,
                        uint32_t *out_count, uint32_t *err, float *heap_d, uint32_t *heap_i,
                        hipStream_t s)
{
    if (rows == 0) return;
    ProfScope prof("bf_finalize", s);
    if (k >= 1 && k + 1 <= 64) {
        bf_finalize_wave_kernel<<<rows, 64, 0, s>>>(cand_d, cand_i, ldc, flags, D, ldd, n, k,
                                                     out_idx, out_dist, out_count, err);
        HM_CHECK_LAUNCH();
        return;
    }
    size_t lds = (size_t)64 * (k + 1) * 8;
    if (lds <= 64 * 1024) {
        bf_finalize_kernel<true><<<ceil_div(rows, 64), 64, lds, s>>>(
            cand_d, cand_i, ldc, flags, D, ldd, rows, n, k, out_idx, out_dist, out_count, err,
            heap_d, heap_i);
    } else {
        if (!heap_d || !heap_i) fail(HM_ERR_INVALID_ARG, "bf_finalize: heap scratch missing");
        bf_finalize_kernel<false><<<ceil_div(rows, 64), 64, 0, s>>>(
            cand_d, cand_i, ldc, flags, D, ldd, rows, n, k, out_idx, out_dist, out_count, err,
            heap_d, heap_i);
    }
    HM_CHECK_LAUNCH();
}

void launch_check_finite(const float *X, uint64_t ld, uint64_t rows, uint32_t d, uint32_t *err,
                         hipStream_t s)
{
    if (rows == 0 || d == 0) return;
    uint64_t total = rows * d;
    unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 4096);
    check_finite_kernel<<<blocks, 256, 0, s>>>(X, ld, rows, d, err);
    HM_CHECK_LAUNCH();
}

void launch_pad_rows(const float *src, uint64_t lds, float *dst, uint64_t ldd, uint64_t rows,
                     uint32_t d, uint32_t dpad, hipStream_t s, const uint32_t *dev_rows)
{
    if (rows == 0) return;
    uint64_t total = rows * dpad;
    unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 8192);
    pad_rows_kernel<<<blocks, 256, 0, s>>>(src, lds, dst, ldd, rows, d, dpad, dev_rows);
    HM_CHECK_LAUNCH();
}

}  // namespace hm
