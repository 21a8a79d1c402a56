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
#include "wave_heap.h"

namespace hm {

namespace {

constexpr int SEL_THREADS = 256;
constexpr uint32_t SEL_CAP = 8192;  // most candidate slots in LDS, power of two

// Keys up to the target bucket collected for the final sort at most.
__host__ __device__ inline uint32_t sel_early_cap(uint32_t K)
{
    const uint32_t c = 2 * K > 1024u ? 2 * K : 1024u;
    return c < SEL_CAP ? c : SEL_CAP;
}

// Candidate slots for K_max: a power of two >= max(K_max, early cap).
inline uint32_t sel_slots(uint32_t K_max)
{
    uint32_t need = sel_early_cap(K_max) > K_max ? sel_early_cap(K_max) : K_max, c = 1;
    while (c < need) c <<= 1;
    return c;
}

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

// Visits every element of a row once, in no particular order: f(value,
// index).  Several loads per thread are in flight before any is used; VEC
// uses 16-byte loads (row 16-byte aligned, ldd a multiple of 4 and >= n, so
// a load clamped to the row's last full vector stays inside the row).
template <bool VEC, class F>
__device__ __forceinline__ void row_for_each(const float *__restrict__ row, uint32_t n, uint64_t ldd,
                                             int tid, F &&f)
{
    if constexpr (VEC) {
        constexpr uint32_t U = 4, STEP = SEL_THREADS * 4 * U;
        for (uint64_t b = 0; b < n; b += STEP) {
            float4 v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t e = min(b + u * SEL_THREADS * 4 + 4 * (uint32_t)tid, ldd - 4);
                v[u] = *reinterpret_cast<const float4 *>(row + e);
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t e = b + u * SEL_THREADS * 4 + 4 * (uint32_t)tid;
                if (e < n) f(v[u].x, (uint32_t)e);
                if (e + 1 < n) f(v[u].y, (uint32_t)e + 1);
                if (e + 2 < n) f(v[u].z, (uint32_t)e + 2);
                if (e + 3 < n) f(v[u].w, (uint32_t)e + 3);
            }
        }
    } else {
        constexpr uint32_t U = 8;
        for (uint64_t b = tid; b < n; b += SEL_THREADS * U) {
            float v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) v[u] = row[min(b + u * SEL_THREADS, (uint64_t)n - 1)];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                if (b + u * SEL_THREADS < n) f(v[u], (uint32_t)(b + u * SEL_THREADS));
        }
    }
}

template <bool VEC>
__global__ __launch_bounds__(SEL_THREADS) void row_select_kernel(
    const float *__restrict__ D, uint64_t ldd, uint32_t n_max, uint32_t K_max,
    float *__restrict__ cand_d, uint32_t *__restrict__ cand_i, uint32_t ldc,
    uint32_t *__restrict__ flags, const uint32_t *__restrict__ dev_rows,
    const uint32_t *__restrict__ row_n, uint32_t cap)
{
    if (dev_rows && blockIdx.x >= *dev_rows) return;  // graph replays with fewer rows
    // ragged rows: this row's length and the candidates it can give
    const uint32_t n = row_n ? min(row_n[blockIdx.x], n_max) : n_max;
    const uint32_t K = min(K_max, n);
    __shared__ uint32_t hist[2048];
    // candidate slots: cap (a power of two >= max(K_max, early_cap)) in
    // dynamic LDS, so small K leaves room for many rows per CU
    extern __shared__ unsigned long long cand[];
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
    // After each pass, when the keys up to the target bucket are few (at
    // most max(2K, 1024)), they are collected and sorted directly; a larger
    // target bucket (a narrow value range) is refined by the next digit
    // first, so the sort stays small.
    const uint32_t early_cap = sel_early_cap(K);
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
        row_for_each<VEC>(row, n, ldd, tid, [&](float f, uint32_t) {
            if (pass == 0 && f != f) nan_local = 1;
            const uint32_t key = fkey(f);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & (nb - 1)], 1u);
        });
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
        if ((K - s_remaining) + s_bucket <= early_cap) {
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
        row_for_each<VEC>(row, n, ldd, tid, [&](float f, uint32_t i) {
            const uint32_t key = fkey(f);
            if ((key & mask) <= T) {
                const uint32_t pos = atomicAdd(&s_cnt, 1u);
                if (pos < cap) cand[pos] = ((unsigned long long)key << 32) | i;
            }
        });
        __syncthreads();
        bad = s_cnt != total;
    } else {
        // Collect: every key < T, then the `rem` lowest indices with key == T.
        total = K;
        const uint32_t nlt = K - rem;
        const bool all_eq = (ceq == rem);
        row_for_each<VEC>(row, n, ldd, tid, [&](float f, uint32_t i) {
            const uint32_t key = fkey(f);
            if (key < T || (all_eq && key == T)) {
                const uint32_t pos = atomicAdd(&s_cnt, 1u);
                if (pos < cap) cand[pos] = ((unsigned long long)key << 32) | i;
            }
        });
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
                uint32_t lo = bitonic_lo(t, stride);
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
    for (uint32_t j = K + tid; j < K_max; j += SEL_THREADS) {  // short ragged rows
        oi[j] = ~0u;
        od[j] = __int_as_float(0x7f800000);
    }
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

// Streams a distance row in batches of 4096 values (16 float4 loads per
// lane, the next batch in flight while the current one is processed) and
// hands over, in index order, the 64-value chunks that may hold a value not
// above threshold(): on_chunk(first_index, this_lane's_value,
// values_in_lds, count); the values passed over go to on_skip(count), in
// the same order.  Chunks are filtered by register ballots per group
// of 256 values (rows with a NaN never reach a replay).  Rows are 16-byte
// aligned with a stride ldd >= n that is a multiple of 4, so the vector
// loads (clamped to the last full vector of the row) stay inside the row.
template <class Threshold, class OnChunk, class OnSkip>
__device__ void stream_row(const float *__restrict__ row, uint32_t n, uint64_t ldd, int lane,
                           float *sv, Threshold &&threshold, OnChunk &&on_chunk, OnSkip &&on_skip)
{
    constexpr uint32_t FB4 = 16, BATCH = FB4 * 256;
    const float inf = __int_as_float(0x7f800000);
    float nx[4 * FB4];  // scalar array: promoted to registers once the loops unroll
#define HM_ROW_LOAD(b)                                                                     \
    _Pragma("clang loop unroll(full)") for (uint32_t c4 = 0; c4 < FB4; ++c4)               \
    {                                                                                      \
        const uint64_t e = min((uint64_t)(b) + c4 * 256 + 4 * (uint32_t)lane, ldd - 4);    \
        const float4 t = *reinterpret_cast<const float4 *>(row + e);                       \
        nx[4 * c4] = t.x;                                                                  \
        nx[4 * c4 + 1] = t.y;                                                              \
        nx[4 * c4 + 2] = t.z;                                                              \
        nx[4 * c4 + 3] = t.w;                                                              \
    }
    HM_ROW_LOAD(0u)
    for (uint32_t b0 = 0; b0 < n; b0 += BATCH) {
        // Which 256-value groups hold a value not above the threshold (the
        // root of a full heap; +inf otherwise)?  The root only falls while
        // the heap is full, so a group with none can be passed over later.
        const float thr = threshold();
        uint32_t gmask = 0;
#pragma clang loop unroll(full)
        for (uint32_t c4 = 0; c4 < FB4; ++c4) {
            bool any = false;
#pragma clang loop unroll(full)
            for (uint32_t jj = 0; jj < 4; ++jj) {
                const uint64_t e = (uint64_t)b0 + c4 * 256 + 4 * (uint32_t)lane + jj;
                any = any || (e < n && !(nx[4 * c4 + jj] > thr));
            }
            if (__ballot(any)) gmask |= 1u << c4;
        }
        __syncthreads();
        if (gmask) {
#pragma clang loop unroll(full)
            for (uint32_t c4 = 0; c4 < FB4; ++c4)
                reinterpret_cast<float4 *>(sv)[c4 * 64 + lane] =
                    make_float4(nx[4 * c4], nx[4 * c4 + 1], nx[4 * c4 + 2], nx[4 * c4 + 3]);
        }
        __syncthreads();
        if (n - b0 > BATCH) {
            HM_ROW_LOAD(b0 + BATCH)
        }
        if (!gmask) {
            on_skip(n - b0 < BATCH ? n - b0 : BATCH);
            continue;
        }
        for (uint32_t c = 0; c < BATCH / 64; ++c) {
            const uint32_t base = b0 + c * 64;
            if (base >= n) break;
            const uint32_t cnt = n - base < 64 ? n - base : 64u;
            if (!((gmask >> (c >> 2)) & 1u) && threshold() < inf) {
                on_skip(cnt);
                continue;
            }
            on_chunk(base, sv[c * 64 + lane], sv + c * 64, cnt);
        }
    }
#undef HM_ROW_LOAD
}


// Replays bruteforce_idx.rs:47-73 over a tied row (k + 1 <= 64) with the
// heap in LDS, run by one wave (wave_heap.h).  An element x greater than the heap's
// maximum is pushed and popped again; that round trip permutes the layout
// in a way that depends on the layout only, so when it is known to leave
// the current layout unchanged (heap_noop_roundtrip, checked after every
// change) such elements are skipped.  One workgroup of 256 threads per
// row (four waves find the groups worth visiting, one wave runs the heap);
// rows without a tie copy their candidates.
constexpr uint32_t BF_BLOCK = 256;       // values per lane-wide load (a float4 per lane)
constexpr uint32_t BF_THREADS = 512;  // bf_finalize_wave_kernel: the group minima by 8 waves, the heap by one
constexpr uint32_t BF_GMIN_MAX = 16384;  // group minima held in LDS
constexpr uint32_t BF_RT_LEN = 512;      // bf_replay_lds_kernel: heaps whose layouts are compared

// Wave-wide minimum (every lane gets it).
__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

// bpg: blocks of 256 values per group, so that the row has at most
// BF_GMIN_MAX groups.
__global__ __launch_bounds__(BF_THREADS) void bf_finalize_wave_kernel(
    const float *__restrict__ cand_d, const uint32_t *__restrict__ cand_i, uint32_t ldc,
    const uint32_t *__restrict__ flags, const float *__restrict__ D, uint64_t ldd, uint32_t n,
    uint32_t k, uint32_t bpg, uint64_t *__restrict__ out_idx, float *__restrict__ out_dist,
    uint32_t *__restrict__ out_count, uint32_t *err, const uint32_t *__restrict__ row_k,
    const uint32_t *__restrict__ row_k_map)
{
    const uint32_t r = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float inf = __int_as_float(0x7f800000);
    uint64_t *oi = out_idx + (size_t)r * k;  // k: the output stride
    float *od = out_dist + (size_t)r * k;
    const uint32_t f = flags ? flags[r] : ROW_TIE;  // no flags: every row tied
    if (f == ROW_SKIP) return;  // (a skipped row's map entry may be unset)
    // this row's k (at most the stride), and the heap's capacity: a heap
    // over n <= k elements never pops, so min(k_r, n) replays the same steps
    const uint32_t kr = row_k ? min(row_k[row_k_map ? row_k_map[r] : r], k) : k;
    const uint32_t m = kr < n ? kr : n;
    if ((f & (ROW_NAN | ROW_BAD)) && n >= 2) {
        if (tid == 0) atomicOr(err, (f & ROW_BAD) ? 4u : 1u);
        for (uint32_t j = tid; j < k; j += BF_THREADS) {
            oi[j] = ~0ull;
            od[j] = inf;
        }
        if (tid == 0) out_count[r] = 0;
        return;
    }
    if (!(f & ROW_TIE)) {
        const float *cd = cand_d + (size_t)r * ldc;
        const uint32_t *ci = cand_i + (size_t)r * ldc;
        for (uint32_t j = tid; j < k; j += BF_THREADS) {
            oi[j] = j < m ? (uint64_t)ci[j] : ~0ull;
            od[j] = j < m ? cd[j] : inf;
        }
        if (tid == 0) out_count[r] = m;
        return;
    }
    const float *row = D + (size_t)r * ldd;
    // A tied row.  Every element that is not greater than the root of a
    // full heap with a stable layout is pushed; the others change nothing.
    // The eight waves first take the minimum key of every group of bpg
    // blocks of 256 values (16 blocks per load round, all in flight
    // together), so the heap wave visits only the groups that can hold such
    // an element, with the next block loading while it works on the current
    // one.  The row's NaN-free (tied rows only), so fkey order is value order.
    const uint32_t nb = (n + BF_BLOCK - 1) / BF_BLOCK, ng = (nb + bpg - 1) / bpg;
    extern __shared__ __align__(16) uint2 bf_smem[];
    uint32_t *gkey = reinterpret_cast<uint32_t *>(bf_smem);  // [ng] group minima (fkey)
    for (uint32_t g = tid; g < ng; g += BF_THREADS) gkey[g] = ~0u;
    __syncthreads();
    for (uint32_t b0 = (uint32_t)wave * 16; b0 < nb; b0 += (BF_THREADS / 64) * 16) {
        float4 v[16];
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u) {
            const uint64_t e = min((uint64_t)(b0 + u) * BF_BLOCK + 4 * (uint32_t)lane, ldd - 4);
            v[u] = *reinterpret_cast<const float4 *>(row + e);
        }
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u) {
            const uint64_t e = (uint64_t)(b0 + u) * BF_BLOCK + 4 * (uint32_t)lane;
            float mn = inf;
            if (e < n) mn = v[u].x;
            if (e + 1 < n) mn = fminf(mn, v[u].y);
            if (e + 2 < n) mn = fminf(mn, v[u].z);
            if (e + 3 < n) mn = fminf(mn, v[u].w);
            mn = wave_min(mn);
            if (lane == 0 && b0 + u < nb) atomicMin(&gkey[(b0 + u) / bpg], fkey(mn));
        }
    }
    __syncthreads();
    if (wave != 0) return;
    // the reference's heap (wave_heap.h), k + 1 <= 64 entries in the wave's
    // registers: (item, distance bits)
    WaveHeapReg h((uint32_t)lane);
    bool stable = false;
    float root = inf;
    auto val = [&](uint32_t pos) { return h.val_lane(pos); };
    // heap_noop_roundtrip on the heap, all ancestor steps at once: lane
    // j checks step j of the walk from position L = len to the root
    auto noop = [&]() -> bool {
        const uint32_t L = h.len;
        if (L == 0) return true;
        const uint32_t steps = 32u - (uint32_t)__clz(L + 1);  // ancestors of L, plus L
        const bool act = (uint32_t)lane < steps;
        const uint32_t a = act ? ((L + 1) >> lane) - 1 : 0u;
        const uint32_t p = a > 0 ? (a - 1) / 2 : 0u;
        // the cross-lane reads are made by every lane, outside the branches
        const float vp = val(p), vq = val(lane == 0 ? L - 1 : a + 1);
        bool bad = false;
        if (act && a > 0) {
            if (lane == 0) bad = a == 2 * p + 2 && !(vp > vq);
            else bad = a == 2 * p + 1 && a + 1 < L && !(vp > vq);
        }
        return __ballot(bad) == 0;
    };
    // Once the heap is full, an element above the root climbs to the root
    // and is popped again: that round trip is a map F of the heap's layout
    // that does not depend on the element's value (it beats every entry).
    // A run of j such elements applies F^j; F is the identity while the
    // layout is stable (noop), and otherwise the layouts F, F^2, ... either
    // become stable or repeat with some period p, so F^j is F^(j mod p)
    // after the first repeat.  Up to RT_HIST layouts are compared.
    constexpr uint32_t RT_HIST = 8;
    auto round_trip = [&]() {  // F: push an element above every entry, pop it
        h.push(make_uint2(~0u, 0x7f800000u));  // +inf beats every (finite) entry
        (void)h.pop();
    };
    auto apply_runs = [&](uint64_t j) {
        if (j == 0 || stable) return;
        if (j <= 2 * RT_HIST) {  // short runs: the round trips cost less than the layouts
            for (uint64_t t = 0; t < j; ++t) {
                round_trip();
                if (noop()) {
                    stable = true;
                    return;
                }
            }
            return;
        }
        uint32_t hist[RT_HIST];
        hist[0] = h.it;
        uint64_t t = 0;
        while (t < j) {
            round_trip();
            ++t;
            if (noop()) {  // stable from here on: the rest are no-ops
                stable = true;
                return;
            }
            const uint32_t seen = t < RT_HIST ? (uint32_t)t : RT_HIST;
            uint32_t prev = RT_HIST;  // an earlier layout equal to this one
            for (uint32_t q = 0; q < seen; ++q)
                if (prev == RT_HIST && __ballot((uint32_t)lane < h.len && hist[q] != h.it) == 0)
                    prev = q;
            if (prev != RT_HIST) {  // period t - prev
                const uint64_t left = (j - t) % (t - prev);
                for (uint64_t u = 0; u < left; ++u) round_trip();
                stable = noop();
                return;
            }
            if (t < RT_HIST) hist[t] = h.it;
        }
    };
    auto skip = [&]() { return h.len == m; };
    // first block at or after b that can change the heap: inside a group
    // being visited the next block; else the first block of the next group
    // whose minimum is not above the root (any group while the heap is not
    // full); the caller applies the skipped elements' round trips
    auto next_block = [&](uint32_t b) -> uint32_t {
        if (b >= nb) return nb;
        if (b % bpg != 0 || !skip()) return b;
        const uint32_t rk = fkey(root);
        for (uint32_t g0 = b / bpg; g0 < ng; g0 += 64) {
            const uint32_t g = g0 + (uint32_t)lane;
            const uint64_t live = __ballot(g < ng && !(gkey[g] > rk));
            if (live) return min((g0 + (uint32_t)__builtin_ctzll(live)) * bpg, nb);
        }
        return nb;
    };
    auto load_block = [&](uint32_t b) -> float4 {
        const uint64_t e = min((uint64_t)b * BF_BLOCK + 4 * (uint32_t)lane, ldd - 4);
        return *reinterpret_cast<const float4 *>(row + e);
    };
    // past the end: the last block again (an in-bounds load, never used)
    auto ld = [&](uint32_t b) { return load_block(min(b, nb - 1)); };
    uint32_t b = next_block(0);
    float4 v = ld(b);
    while (b < nb) {
        const uint32_t nx = next_block(b + 1);  // under the current root
        const float4 pf = ld(nx);
        const uint32_t base = b * BF_BLOCK, cnt = min(BF_BLOCK, n - base);
        const float vc[4] = {v.x, v.y, v.z, v.w};
        uint32_t e = 0;  // element 4 * lane + c of the block is vc[c] of that lane
        while (e < cnt) {
            if (skip()) {
                uint32_t best = BF_BLOCK;
#pragma unroll
                for (uint32_t c = 0; c < 4; ++c) {
                    const uint32_t idx = 4 * (uint32_t)lane + c;
                    const uint64_t live = __ballot(idx >= e && idx < cnt && !(vc[c] > root));
                    if (live) best = min(best, 4 * (uint32_t)__builtin_ctzll(live) + c);
                }
                apply_runs((best < cnt ? best : cnt) - e);  // the elements above the root
                if (best >= cnt) break;
                e = best;
            }
            const uint32_t c = e & 3;
            const float sel = c == 0 ? vc[0] : (c == 1 ? vc[1] : (c == 2 ? vc[2] : vc[3]));
            const float x =
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sel), (int)(e >> 2)));
            h.push(make_uint2(base + e, __float_as_uint(x)));
            if (h.len > m) (void)h.pop();
            if (h.len == m) {
                root = h.val_at(0);
                stable = root < inf && noop();
            }
            ++e;
        }
        // The root only falls while the heap is full, so the next block is
        // at or after nx; the whole blocks skipped hold elements above it.
        const uint32_t b2 = next_block(b + 1);
        if (b2 > b + 1) apply_runs((uint64_t)min(b2 * BF_BLOCK, n) - (uint64_t)(b + 1) * BF_BLOCK);
        v = b2 == nx ? pf : ld(b2);
        b = b2;
    }
    const uint32_t cntk = h.len;
    for (uint32_t p = cntk; p > 0; --p) {
        const uint2 x = h.pop();
        if (lane == 0) {
            oi[p - 1] = x.x;
            od[p - 1] = __uint_as_float(x.y);
        }
    }
    for (uint32_t j = cntk + lane; j < k; j += 64) {
        oi[j] = ~0ull;
        od[j] = inf;
    }
    if (lane == 0) out_count[r] = cntk;
}

// Would pushing an element larger than every entry and popping it again
// leave this (full) heap's layout as it is?  The push shifts the path from
// the new leaf L to the root down one level; the pop's sift-down retraces
// that path exactly when, at every level where the path runs through a
// left child that has a right sibling, the parent is strictly greater than
// the sibling (on equality the sift-down takes the right child), and, at
// the bottom, when L is a right child, its parent is strictly greater than
// the left sibling L-1.  Checked against a direct simulation of the two
// operations on 200,000 random heaps with heavy ties.
__device__ bool heap_noop_roundtrip(const float *d, uint32_t L)
{
    if (L == 0) return true;
    uint32_t a = L, p = (a - 1) / 2;
    if (a == 2 * p + 2 && !(d[p] > d[L - 1])) return false;
    a = p;
    while (a > 0) {
        p = (a - 1) / 2;
        if (a == 2 * p + 1 && a + 1 < L && !(d[p] > d[a + 1])) return false;
        a = p;
    }
    return true;
}

// Replays bruteforce_idx.rs:47-73 over a tied row for k + 1 > 64, with the
// heap in LDS.  Every lane of the single wave executes the same heap
// operations on the same LDS words (uniform control flow, broadcast reads),
// and the lanes cooperate only to scan: while the heap is full and the
// round trip of a larger element is a no-op, a ballot over 64 values finds
// the next element that is not greater than the root.
__global__ __launch_bounds__(64) void bf_replay_lds_kernel(
    const float *__restrict__ cand_d, const uint32_t *__restrict__ cand_i, uint32_t ldc,
    const uint32_t *__restrict__ flags, const float *__restrict__ D, uint64_t ldd, uint32_t n,
    uint32_t k, uint64_t *__restrict__ out_idx, float *__restrict__ out_dist,
    uint32_t *__restrict__ out_count, uint32_t *err, const uint32_t *__restrict__ row_k,
    const uint32_t *__restrict__ row_k_map)
{
    extern __shared__ __align__(16) unsigned char rp_smem[];
    const uint32_t r = blockIdx.x;
    const int lane = threadIdx.x;
    const float inf = __int_as_float(0x7f800000);
    uint64_t *oi = out_idx + (size_t)r * k;  // k: the output stride
    float *od = out_dist + (size_t)r * k;
    const uint32_t f = flags ? flags[r] : ROW_TIE;  // no flags: every row tied
    if (f == ROW_SKIP) return;
    const uint32_t kr = row_k ? min(row_k[row_k_map ? row_k_map[r] : r], k) : k;
    const uint32_t m = kr < n ? kr : n;  // heap capacity (see the wave kernel)
    const uint32_t cap = k < n ? k : n;  // the LDS layout's capacity, every row
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
    LaneHeap h;
    h.d = reinterpret_cast<float *>(rp_smem);
    h.ix = reinterpret_cast<uint32_t *>(rp_smem + (size_t)(cap + 1) * 4);
    h.len = 0;
    const float *row = D + (size_t)r * ldd;
    bool stable = false;
    float root = inf;
    __shared__ __align__(16) float sv[4096];
    // Runs of elements above the root of the full heap: one layout map F
    // applied j times (see bf_finalize_wave_kernel), until the layout is
    // stable or repeats one of the last RT_HIST (kept for heaps of at most
    // RT_LEN entries), then by the period.
    constexpr uint32_t RT_HIST = 8;
    const uint32_t RT_LEN = cap < BF_RT_LEN ? cap : BF_RT_LEN;
    // [RT_HIST][RT_LEN] after the heap's arrays
    uint32_t *rt = reinterpret_cast<uint32_t *>(rp_smem + (size_t)(cap + 1) * 8);
    auto round_trip = [&]() {
        h.push(inf, ~0u);  // above every (finite) entry
        float a;
        uint32_t b;
        h.pop(&a, &b);
    };
    auto save = [&](uint32_t q) {
        for (uint32_t p = lane; p < h.len; p += 64) rt[q * RT_LEN + p] = h.ix[p];
    };
    auto same_as = [&](uint32_t q) {
        bool diff = false;
        for (uint32_t p = lane; p < h.len; p += 64) diff = diff || rt[q * RT_LEN + p] != h.ix[p];
        return __ballot(diff) == 0;
    };
    auto apply_runs = [&](uint64_t j) {
        if (j == 0 || stable) return;
        // short runs: the round trips themselves cost less than the layouts
        const bool track = h.len <= RT_LEN && j > 2 * RT_HIST;
        if (track) save(0);
        for (uint64_t t = 1; t <= j; ++t) {
            round_trip();
            if (heap_noop_roundtrip(h.d, h.len)) {
                stable = true;
                return;
            }
            if (!track) continue;
            const uint32_t seen = t < RT_HIST ? (uint32_t)t : RT_HIST;
            for (uint32_t q = 0; q < seen; ++q)
                if (same_as(q)) {
                    const uint64_t left = (j - t) % (t - q);
                    for (uint64_t u = 0; u < left; ++u) round_trip();
                    return;
                }
            if (t < RT_HIST) save((uint32_t)t);
        }
    };
    auto threshold = [&]() { return h.len == m ? root : inf; };
    stream_row(row, n, ldd, lane, sv, threshold,
               [&](uint32_t base, float v, const float *cv, uint32_t cnt) {
        uint32_t j = 0;
        while (j < cnt) {
            if (h.len == m) {
                const unsigned long long live =
                    __ballot((uint32_t)lane >= j && (uint32_t)lane < cnt && !(v > root));
                const uint32_t nxt = live ? (uint32_t)__builtin_ctzll(live) : cnt;
                apply_runs(nxt - j);
                if (live == 0) break;
                j = nxt;
            }
            const float x = cv[j];
            h.push(x, base + j);
            if (h.len > m) {
                float a;
                uint32_t b;
                h.pop(&a, &b);
            }
            if (h.len == m) {
                root = h.d[0];
                stable = root < inf && heap_noop_roundtrip(h.d, h.len);
            }
            ++j;
        }
    }, [&](uint32_t cnt) { apply_runs(cnt); });
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

// out[r] = the m-th smallest value of row r (m <= KTH_MAX), one wave per
// row: each lane keeps its m smallest in registers (insertion into a
// sorted run), then m rounds of a wave minimum take the m-th overall.
// NaN values are never taken.
constexpr uint32_t KTH_MAX = 32;

// One wave per row: each lane keeps the KT smallest of its values (KT >= m)
// as a sorted run; the row is read 4,096 values per round with every load
// of the round in flight (16-byte loads when the row allows them), then m
// rounds of a wave minimum over the lanes' heads give the m-th smallest.
template <uint32_t KT, bool VEC, uint32_t U = 16>  // U float4 loads per lane per round
__global__ __launch_bounds__(256) void kth_smallest_kernel(const float *__restrict__ D,
                                                           uint64_t ldd, uint32_t rows, uint32_t n,
                                                           uint32_t m, float *__restrict__ out,
                                                           const uint32_t *__restrict__ dev_rows,
                                                           uint32_t *__restrict__ zero)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows || (dev_rows && r >= *dev_rows)) return;
    if (zero && lane == 0) zero[r] = 0;
    const float inf = __int_as_float(0x7f800000);
    const float *row = D + (size_t)r * ldd;
    float best[KT];
#pragma unroll
    for (uint32_t i = 0; i < KT; ++i) best[i] = inf;
    auto insert = [&](float x) {
        if (!(x < best[KT - 1])) return;  // also drops NaN
#pragma unroll
        for (uint32_t t = 0; t < KT; ++t) {
            const float b = best[t];
            const bool lt = x < b;
            best[t] = lt ? x : b;
            x = lt ? b : x;
        }
    };
    // cut: an upper bound of the row's m-th smallest value, the m-th smallest
    // of the lanes' minima over the first round (m distinct elements); only
    // values not above it can be among the m smallest, so only they are
    // inserted.  Until it is known (and in the scalar path) it is +inf.
    float cut = inf;
    for (uint64_t b0 = 0; b0 < n; b0 += (uint64_t)U * 256) {
        if constexpr (VEC) {
            float4 v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t e = min(b0 + u * 256 + 4 * (uint32_t)lane, ldd - 4);
                v[u] = *reinterpret_cast<const float4 *>(row + e);
            }
            if (b0 == 0) {
                float lm = inf;
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint64_t e = u * 256 + 4 * (uint32_t)lane;
                    if (e < n) lm = fminf(lm, v[u].x);
                    if (e + 1 < n) lm = fminf(lm, v[u].y);
                    if (e + 2 < n) lm = fminf(lm, v[u].z);
                    if (e + 3 < n) lm = fminf(lm, v[u].w);
                }
                // m rounds of the wave minimum over the lanes' minima; the
                // owner (lowest lane holding it) drops out each round
                float mine = lm;
                for (uint32_t round = 0; round < m; ++round) {
                    float w = mine;
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) w = fminf(w, __shfl_xor(w, o, 64));
                    cut = w;
                    const unsigned long long owner = __ballot(mine == w);
                    if ((uint32_t)lane == (uint32_t)__builtin_ctzll(owner)) mine = inf;
                    if (owner == 0) break;  // fewer than m finite minima: cut stays +inf
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t e = b0 + u * 256 + 4 * (uint32_t)lane;
                if (e < n && v[u].x <= cut) insert(v[u].x);
                if (e + 1 < n && v[u].y <= cut) insert(v[u].y);
                if (e + 2 < n && v[u].z <= cut) insert(v[u].z);
                if (e + 3 < n && v[u].w <= cut) insert(v[u].w);
            }
        } else {
            // scalar loads: the round in two halves of 32 per lane
#pragma unroll
            for (uint32_t hh = 0; hh < 2; ++hh) {
                float v[2 * U];
#pragma unroll
                for (uint32_t u = 0; u < 2 * U; ++u)
                    v[u] = row[min(b0 + (hh * 2 * U + u) * 64 + (uint32_t)lane, (uint64_t)n - 1)];
#pragma unroll
                for (uint32_t u = 0; u < 2 * U; ++u)
                    if (b0 + (hh * 2 * U + u) * 64 + (uint32_t)lane < n) insert(v[u]);
            }
        }
    }
    // m rounds: the wave minimum of the lanes' heads; its lane drops its head
    float head = best[0];
    uint32_t used = 0;
    float kth = inf;
    for (uint32_t round = 0; round < m; ++round) {
        float w = head;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) w = fminf(w, __shfl_xor(w, o, 64));
        kth = w;
        const unsigned long long owner = __ballot(head == w);
        if ((uint32_t)lane == (uint32_t)__builtin_ctzll(owner)) {
            ++used;
            // shift: the lane's next smallest becomes its head
            float nxt = inf;
#pragma unroll
            for (uint32_t t = 1; t < KT; ++t)
                if (t == used) nxt = best[t];
            head = used < KT ? nxt : inf;
        }
    }
    if (lane == 0) out[r] = kth;
}

__global__ void check_finite_kernel(const float *__restrict__ X, uint64_t ld, uint64_t rows,
                                    uint32_t d, uint32_t *err)
{
    bool bad = false;
    if (ld == d && ld % 4 == 0 && ((uintptr_t)X & 15) == 0) {
        // contiguous rows: the whole block as float4s
        const float4 *X4 = reinterpret_cast<const float4 *>(X);
        const uint64_t total = rows * d / 4;
        for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
             e += (uint64_t)gridDim.x * blockDim.x) {
            const float4 v = X4[e];
            bad |= !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w));
        }
    } else {
        for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x)
            for (uint32_t c = threadIdx.x; c < d; c += blockDim.x) bad |= !isfinite(X[r * ld + c]);
    }
    if (bad) atomicOr(err, 2u);
}

__global__ void pad_rows_kernel(const float *__restrict__ src, uint64_t lds, float *__restrict__ dst,
                                uint64_t ldd, uint64_t rows, uint32_t d, uint32_t dpad,
                                const uint32_t *__restrict__ dev_rows, uint32_t *err,
                                const uint32_t *__restrict__ src_rows)
{
    if (dev_rows && *dev_rows < rows) rows = *dev_rows;
    bool bad = false;
    for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const uint64_t sr = src_rows ? src_rows[r] : r;
        const float *x = src + sr * lds;
        float *y = dst + r * ldd;
        for (uint32_t c = threadIdx.x; c < dpad; c += blockDim.x) {
            const float v = c < d ? x[c] : 0.0f;
            bad |= !isfinite(v);
            y[c] = v;
        }
    }
    if (err && bad) atomicOr(err, 2u);
}

}  // namespace

void launch_row_select(const float *D, uint64_t ldd, uint32_t rows, uint32_t n, uint32_t K,
                       float *cand_d, uint32_t *cand_i, uint32_t ldc, uint32_t *flags,
                       hipStream_t s, const uint32_t *dev_rows, const uint32_t *row_n)
{
    if (rows == 0 || n == 0 || K == 0) return;
    if (K > SEL_CAP || (!row_n && K > n)) fail(HM_ERR_INVALID_ARG, "row_select: K out of range");
    if (ldd < n) fail(HM_ERR_INVALID_ARG, "row_select: row stride below row length");
    ProfScope prof("row_select", s);
    const uint32_t cap = sel_slots(K);
    const size_t lds = (size_t)cap * 8;
    if (ldd % 4 == 0 && ((uintptr_t)D & 15) == 0)
        row_select_kernel<true><<<rows, SEL_THREADS, lds, s>>>(D, ldd, n, K, cand_d, cand_i, ldc,
                                                               flags, dev_rows, row_n, cap);
    else
        row_select_kernel<false><<<rows, SEL_THREADS, lds, s>>>(D, ldd, n, K, cand_d, cand_i, ldc,
                                                                flags, dev_rows, row_n, cap);
    HM_CHECK_LAUNCH();
}

void launch_bf_finalize(const float *cand_d, const uint32_t *cand_i, uint32_t ldc,
                        const uint32_t *flags, const float *D, uint64_t ldd, uint32_t rows,
                        uint32_t n, uint32_t k, uint64_t *out_idx, float *out_dist,
                        uint32_t *out_count, uint32_t *err, hipStream_t s,
                        const uint32_t *row_k, const uint32_t *row_k_map)
{
    if (rows == 0) return;
    if (k == 0) fail(HM_ERR_INVALID_ARG, "bf_finalize: k is zero");
    // the heap never holds more than min(k, n) + 1 entries
    const uint32_t cap = k < n ? k : n;
    if (cap + 1 > SELECT_MAX_K) fail(HM_ERR_INVALID_ARG, "bf_finalize: k too large");
    if (ldd % 4 || ldd < n || ((uintptr_t)D & 15))
        fail(HM_ERR_INVALID_ARG, "bf_finalize: rows must be 16-byte aligned, stride a multiple of 4");
    ProfScope prof("bf_finalize", s);
    if (cap + 1 <= 64) {
        const uint32_t nb = (n + BF_BLOCK - 1) / BF_BLOCK;
        uint32_t bpg = 1;
        while ((nb + bpg - 1) / bpg > BF_GMIN_MAX) bpg <<= 1;
        const size_t lds = (size_t)((nb + bpg - 1) / bpg) * 4;  // the group minima
        bf_finalize_wave_kernel<<<rows, BF_THREADS, lds, s>>>(cand_d, cand_i, ldc, flags, D, ldd, n, k,
                                                       bpg, out_idx, out_dist, out_count, err,
                                                       row_k, row_k_map);
    } else {
        // the heap (<= 64 KiB) and the layouts its runs compare
        const size_t lds = (size_t)(cap + 1) * 8 + (size_t)8 * std::min(cap, BF_RT_LEN) * 4;
        bf_replay_lds_kernel<<<rows, 64, lds, s>>>(cand_d, cand_i, ldc, flags, D, ldd, n, k,
                                                   out_idx, out_dist, out_count, err, row_k,
                                                   row_k_map);
    }
    HM_CHECK_LAUNCH();
}

void launch_kth_smallest(const float *D, uint64_t ldd, uint32_t rows, uint32_t n, uint32_t m,
                         float *out, hipStream_t s, const uint32_t *dev_rows, uint32_t *zero)
{
    if (rows == 0) return;
    if (m == 0 || m > KTH_MAX || m > n) fail(HM_ERR_INVALID_ARG, "kth_smallest: m out of range");
    ProfScope prof("kth_smallest", s);
    const bool vec = ldd % 4 == 0 && ((uintptr_t)D & 15) == 0 && ldd >= 4;
#define HM_KTH(KT)                                                                             \
    if (vec && n <= 1024)                                                                      \
        kth_smallest_kernel<KT, true, 4><<<ceil_div(rows, 4), 256, 0, s>>>(D, ldd, rows, n, m,   \
                                                                           out, dev_rows, zero); \
    else if (vec)                                                                              \
        kth_smallest_kernel<KT, true><<<ceil_div(rows, 4), 256, 0, s>>>(D, ldd, rows, n, m, out, \
                                                                        dev_rows, zero);       \
    else                                                                                       \
        kth_smallest_kernel<KT, false><<<ceil_div(rows, 4), 256, 0, s>>>(D, ldd, rows, n, m,     \
                                                                         out, dev_rows, zero)
    if (m <= 8) {
        HM_KTH(8);
    } else if (m <= 16) {
        HM_KTH(16);
    } else {
        HM_KTH(32);
    }
#undef HM_KTH
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

namespace {
// One workgroup: the rows flagged ROW_TIE alone go to the plan's slots when
// there are at most RERUN_SLOTS of them and no row has another flag;
// otherwise the host takes over (the count stays 0).  Four flags per 16-byte
// load when the array allows it.
__global__ __launch_bounds__(1024) void rerun_plan_kernel(const uint32_t *__restrict__ flags,
                                                          uint32_t nq, uint32_t *__restrict__ plan)
{
    __shared__ uint32_t s_tie, s_other, s_unc, s_rows[RERUN_SLOTS];
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_tie = s_other = s_unc = 0;
    __syncthreads();
    auto visit = [&](uint32_t r, uint32_t f) {
        if (f == ROW_TIE) {
            const uint32_t p = atomicAdd(&s_tie, 1u);
            if (p < RERUN_SLOTS) s_rows[p] = r;
        } else if (f & (ROW_TIE | ROW_NAN | ROW_UNCERT | ROW_BAD)) {
            atomicAdd(&s_other, 1u);
        }
        if (f & ROW_UNCERT) atomicAdd(&s_unc, 1u);
    };
    const uint32_t nv = ((uintptr_t)flags & 15) == 0 ? nq / 4 : 0;
    for (uint32_t v = tid; v < nv; v += 1024) {
        const uint4 f4 = reinterpret_cast<const uint4 *>(flags)[v];
        if (f4.x | f4.y | f4.z | f4.w) {
            visit(4 * v, f4.x);
            visit(4 * v + 1, f4.y);
            visit(4 * v + 2, f4.z);
            visit(4 * v + 3, f4.w);
        }
    }
    for (uint32_t r = 4 * nv + tid; r < nq; r += 1024) visit(r, flags[r]);
    __syncthreads();
    const bool dev = s_other == 0 && s_tie <= RERUN_SLOTS;
    const uint32_t c = dev ? s_tie : 0u;
    if (tid < RERUN_SLOTS) {
        plan[tid] = tid < c ? s_rows[tid] : ~0u;
        plan[RERUN_SLOTS + tid] = tid < c ? ROW_TIE : ROW_SKIP;
    }
    if (tid == 0) {
        plan[2 * RERUN_SLOTS] = c;
        plan[2 * RERUN_SLOTS + 1] = dev ? 0u : 1u;
        plan[2 * RERUN_SLOTS + 2] = s_tie + s_other;
        plan[2 * RERUN_SLOTS + 3] = s_unc;
    }
}
}  // namespace

void launch_rerun_plan(const uint32_t *flags, uint32_t nq, uint32_t *plan, hipStream_t s)
{
    rerun_plan_kernel<<<1, 1024, 0, s>>>(flags, nq, plan);
    HM_CHECK_LAUNCH();
}

void upload_rows(const float *h_rows, uint64_t n, uint32_t d, float *d_dst, uint32_t dpad,
                 DeviceBuffer &stage, hipStream_t s)
{
    if (n == 0) return;
    if (dpad == d) {
        HM_HIP(hipMemcpyAsync(d_dst, h_rows, n * d * sizeof(float), hipMemcpyHostToDevice, s));
        return;
    }
    const uint64_t chunk = std::max<uint64_t>(1, ((uint64_t)64 << 20) / d);  // 256 MiB of rows
    stage.reserve(std::min(n, chunk) * d * sizeof(float));
    for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
        const uint64_t m = std::min(chunk, n - r0);
        HM_HIP(hipMemcpyAsync(stage.get(), h_rows + r0 * d, m * d * sizeof(float),
                              hipMemcpyHostToDevice, s));
        launch_pad_rows(stage.as<float>(), d, d_dst + r0 * dpad, dpad, m, d, dpad, s);
    }
}

void download_rows(const float *d_src, uint32_t dpad, uint64_t n, uint32_t d, float *h_out,
                   DeviceBuffer &stage, hipStream_t s)
{
    if (n == 0) return;
    if (dpad == d) {
        HM_HIP(hipMemcpyAsync(h_out, d_src, n * d * sizeof(float), hipMemcpyDeviceToHost, s));
        return;
    }
    const uint64_t chunk = std::max<uint64_t>(1, ((uint64_t)64 << 20) / d);
    stage.reserve(std::min(n, chunk) * d * sizeof(float));
    for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
        const uint64_t m = std::min(chunk, n - r0);
        launch_pad_rows(d_src + r0 * dpad, dpad, stage.as<float>(), d, m, d, d, s);
        HM_HIP(hipMemcpyAsync(h_out + r0 * d, stage.get(), m * d * sizeof(float),
                              hipMemcpyDeviceToHost, s));
    }
}

void launch_pad_rows(const float *src, uint64_t lds, float *dst, uint64_t ldd, uint64_t rows,
                     uint32_t d, uint32_t dpad, hipStream_t s, const uint32_t *dev_rows,
                     uint32_t *err, const uint32_t *src_rows)
{
    if (rows == 0) return;
    const unsigned blocks = (unsigned)std::min<uint64_t>(rows, 8192);
    pad_rows_kernel<<<blocks, 256, 0, s>>>(src, lds, dst, ldd, rows, d, dpad, dev_rows, err,
                                           src_rows);
    HM_CHECK_LAUNCH();
}

}  // namespace hm
