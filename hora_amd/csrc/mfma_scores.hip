// mfma_scores.hip -- approximate query x base scores on the CDNA4 matrix cores.
//
// The shortlist forms of BruteForceIndex rank the base by an approximate key
// computed from one GEMM S = Q . B^T on MFMA with FP32 accumulation:
//   Euclidean   key = ||b||^2 - 2 s          (||q||^2 is constant per query)
//   DotProduct  key = s                      (the reference ranks +sum(q*b) ascending)
//   Cosine      key = -s / ||b||             (E ~ key / ||q||)
//   Angular     key =  s / ||b||             (E ~ 1 + key / ||q||)
// and the shortlist is then reranked with the exact reference-order FP32
// arithmetic (rerank.hip).  Operands are BF16 (v_mfma_f32_16x16x32_bf16) or
// FP8-e4m3 with a per-vector scale (v_mfma_f32_16x16x128_f8f6f4).
//
// Tiling: 256 threads = 4 waves in a 2x2 arrangement; the workgroup owns a
// 128-query x 128-base tile, each wave 64x64 = 4x4 MFMA 16x16 tiles (64 fp32
// accumulators per lane).  Each K step brings 128 bytes per row of both
// operands into LDS with direct global->LDS loads (two LDS buffers; no
// register staging).  The filter form writes no key matrix: keys below a
// per-query threshold are appended to per-query candidate lists in the
// epilogue.
#include <type_traits>

#include "kernels.h"
#include "profiler.h"

namespace hm {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int GM = 128, GN = 128;
constexpr int ROWB = 128;  // bytes per row per K step (64 bf16 / 128 fp8)

// LDS layout of one K step of one operand: rows of 128 bytes, unpadded; the
// 16-byte chunk c of row r sits at chunk position c ^ swz(r).  A wave's
// fragment read touches 16 rows at one chunk index, and the swizzle spreads
// those rows over all eight chunk positions of both row parities, so the
// 16-byte reads hit distinct banks with 32 or 64 banks.
__device__ __forceinline__ uint32_t swz(uint32_t r) { return (r ^ (r >> 3)) & 7u; }

// f(integral_constant<int, I>) for I = B .. E-1, in order
template <int I, int E, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < E) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, E>(f);
    }
}

__device__ __forceinline__ uint32_t fkey(float f)
{
    f = f + 0.0f;
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Output of a key-matrix launch.
struct ScoreOut {
    float *Kmat;  // [nq][ldk]
    uint64_t ldk;
};

// Tile order.  The workgroups are dispatched round robin over the 8 XCDs
// (workgroup L to XCD L % 8), each XCD with its own L2.  The linear id is
// remapped so that every XCD walks a contiguous range of tiles, and tiles
// are ordered in groups of G base tiles with the query tiles inside a
// group: the group's base tiles (about 2 MiB) stay in that XCD's L2 while
// the query matrix streams past them, so the base is read once and the
// queries once per group.
__device__ __forceinline__ void tile_of(uint32_t L, uint32_t TX, uint32_t TY, uint32_t G,
                                        uint32_t *bt, uint32_t *qt)
{
    const uint32_t T = TX * TY, q = T / 8, r = T % 8, x = L % 8, i = L / 8;
    const uint32_t t = x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    const uint32_t per = G * TY, g = t / per, w = t % per;
    const uint32_t gb = min(G, TX - g * G);
    *qt = w / gb;
    *bt = g * G + w % gb;
}

// Software pipeline: step kt reads its fragments from LDS buffer kt & 1,
// then issues the direct loads of step kt+1 into the other buffer, then
// runs its MFMAs, so each load has one step of matrix work to land.  A
// global_load_lds_dwordx4 writes 16 bytes per lane at (wave-uniform LDS
// base + 16 * lane): wave w's instruction c fills rows 8 (4w + c) .. +7,
// lane l taking row 8 (4w + c) + l / 8 and chunk position l % 8, that is
// global chunk (l % 8) ^ swz(row).  Rows past the end load a clamped valid
// row; their products are finite and the epilogue masks them.
// The metric is a template parameter: the epilogue has no run-time branch
// on it (a uniform branch chain there compiled to a store whose address
// register was left unset on the last arm).
// MIN4 (key matrix form only): each lane writes the minimum of its four keys
// of a row (columns lr + 16 j of the wave's 64), at column
// btile * 32 + wc * 16 + lr of a [nq][ceil(n / 128) * 32] matrix.
template <int ELEM, int MET, bool MIN4 = false, bool OFF32 = true>  // ELEM: 2 = bf16, 1 = fp8
__global__ __launch_bounds__(256, 2) void score_gemm_kernel(
    const uint8_t *__restrict__ Q, uint32_t nq, uint32_t TY, uint32_t G,
    const uint8_t *__restrict__ B, uint32_t n, uint32_t row_bytes, const float *__restrict__ qscale,
    const float *__restrict__ bscale, const float *__restrict__ bnorm2,
    const float *__restrict__ inv_bnorm, ScoreOut out, const uint32_t *__restrict__ dev_rows)
{
    uint32_t btile, qtile;
    tile_of(blockIdx.x, (n + GN - 1) / GN, TY, G, &btile, &qtile);
    if (dev_rows && *dev_rows < nq) nq = *dev_rows;
    if (qtile * GM >= nq) return;
    __shared__ __align__(1024) uint8_t sA[2][GM * ROWB];
    __shared__ __align__(1024) uint8_t sB[2][GN * ROWB];
    // the epilogue's per-query scale and per-row (norm term, scale)
    // operands, loaded before the K loop so their latency is hidden
    __shared__ float s_rs[GM], s_cv[GN], s_cs[GN];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const uint32_t q0 = qtile * GM, n0 = btile * GN;
    if (tid < GM) {
        const uint32_t q = q0 + tid, qc = q < nq ? q : nq - 1;
        s_rs[tid] = (ELEM == 1) ? qscale[qc] : 1.0f;
    } else {
        const uint32_t c = n0 + (tid - GM), cc = c < n ? c : n - 1;
        if constexpr (MET == HM_METRIC_EUCLIDEAN) s_cv[tid - GM] = bnorm2[cc];
        else if constexpr (MET == HM_METRIC_DOT_PRODUCT) s_cv[tid - GM] = 0.0f;
        else s_cv[tid - GM] = inv_bnorm[cc];
        s_cs[tid - GM] = (ELEM == 1) ? bscale[cc] : 1.0f;
    }

    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Loop-invariant addressing, so that a K step spends no VALU on it:
    //  - each lane's four load rows as 32-bit byte offsets from the operand
    //    base (OFF32: the launcher checks that both operands span < 4 GiB),
    //    the step's k offset added to the wave-uniform base;
    //  - the LDS destination of each load from the wave index in an SGPR;
    //  - each lane's fragment read offsets within a buffer; the K loop is
    //    unrolled by two, so the buffer is a constant in each half.
    const uint32_t wave_s = (uint32_t)__builtin_amdgcn_readfirstlane(wave);
    uint32_t qo[4], bo[4];
    const uint8_t *qp[4], *bp[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t r = 8 * (4 * wave + c) + lane / 8;
        const uint32_t g = (uint32_t)(lane % 8) ^ swz(r);
        const uint32_t qa = q0 + r < nq ? q0 + r : nq - 1, ba = n0 + r < n ? n0 + r : n - 1;
        if constexpr (OFF32) {
            qo[c] = qa * row_bytes + 16 * g;
            bo[c] = ba * row_bytes + 16 * g;
        } else {
            qp[c] = Q + (size_t)qa * row_bytes + 16 * g;
            bp[c] = B + (size_t)ba * row_bytes + 16 * g;
        }
    }
    auto issue = [&](int buf, uint32_t kb) {
        const uint8_t *Qk = Q + kb, *Bk = B + kb;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint8_t *src_q = OFF32 ? Qk + qo[c] : qp[c] + kb;
            const uint8_t *src_b = OFF32 ? Bk + bo[c] : bp[c] + kb;
            __builtin_amdgcn_global_load_lds(src_q, &sA[buf][(4 * wave_s + c) * 1024], 16, 0, 0);
            __builtin_amdgcn_global_load_lds(src_b, &sB[buf][(4 * wave_s + c) * 1024], 16, 0, 0);
        }
    };
    const int lr = lane & 15, lk = lane >> 4;
    // fragment offsets: h = 0, 1 are the lane's two 16-byte chunks of a row
    // (FP8: chunks 2 lk, 2 lk + 1; BF16: chunks lk and 4 + lk)
    uint32_t fa[4][2], fb[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t ch = ELEM == 1 ? 2 * lk + h : 4 * h + lk;
            const uint32_t ra = wr * 64 + i * 16 + lr, rb = wc * 64 + i * 16 + lr;
            fa[i][h] = ra * ROWB + 16 * (ch ^ swz(ra));
            fb[i][h] = rb * ROWB + 16 * (ch ^ swz(rb));
        }

    const uint32_t nk = row_bytes / ROWB;
    // Software pipeline: step kt reads its fragments from LDS buffer kt & 1,
    // then issues the direct loads of step kt+1 into the other buffer, then
    // runs its MFMAs, so each load has one step of matrix work to land.
    auto step = [&](auto bufc, uint32_t kt) {
        constexpr int buf = decltype(bufc)::value;
        __builtin_amdgcn_s_waitcnt(0);  // this wave's loads of step kt have landed
        __syncthreads();                // ... and every wave's; buffer buf ^ 1 is free
        const uint8_t *As = sA[buf], *Bs = sB[buf];
        uint4 av[4][2], bv[4][2];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                av[i][h] = *reinterpret_cast<const uint4 *>(As + fa[i][h]);
                bv[i][h] = *reinterpret_cast<const uint4 *>(Bs + fb[i][h]);
            }
        if (kt + 1 < nk) issue(buf ^ 1, (kt + 1) * ROWB);
        if constexpr (ELEM == 2) {
            // two MFMA K=32 per step; lane holds 8 bf16 at k = 32*kk + 8*lk
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, av[i][kk]),
                            __builtin_bit_cast(bf16x8, bv[j][kk]), acc[i][j], 0, 0, 0);
        } else {
            // one MFMA K=128 per step (v_mfma_f32_16x16x128_f8f6f4, e4m3 x
            // e4m3, no block scale); lane holds 32 fp8 at k = 32*lk.  A and B
            // use the same k packing, so the products pair up for any order
            // the hardware assigns within a lane.
            i32x8 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint4 lo = av[i][0], hi = av[i][1];
                a[i] = i32x8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w,
                             (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
                const uint4 lb = bv[i][0], hb = bv[i][1];
                b[i] = i32x8{(int)lb.x, (int)lb.y, (int)lb.z, (int)lb.w,
                             (int)hb.x, (int)hb.y, (int)hb.z, (int)hb.w};
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        a[i], b[j], acc[i][j], 0, 0, 0, 0, 0, 0);
        }
    };
    issue(0, 0u);
    for (uint32_t kt = 0; kt < nk; kt += 2) {
        step(std::integral_constant<int, 0>{}, kt);
        if (kt + 1 < nk) step(std::integral_constant<int, 1>{}, kt + 1);
    }

    // Epilogue: lane holds D[4*lk + e][lr] of each 16x16 tile.  Every
    // per-column and per-row operand is loaded up front (clamped indices,
    // no branches), so the loads overlap instead of each waiting inside a
    // branch; rows/columns past the end are masked afterwards.
    float colv[4], cols[4];
    uint32_t colc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t cl = wc * 64 + j * 16 + lr;
        colc[j] = n0 + cl;
        colv[j] = s_cv[cl];
        cols[j] = s_cs[cl];
    }
    float rows_[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) rows_[i][e] = s_rs[wr * 64 + i * 16 + lk * 4 + e];
    if constexpr (MIN4) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t q = q0 + wr * 64 + i * 16 + lk * 4 + e;
                float mn = __int_as_float(0x7f800000);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float s = acc[i][j][e];
                    if (ELEM == 1) s = s * (rows_[i][e] * cols[j]);
                    float key;
                    if constexpr (MET == HM_METRIC_EUCLIDEAN) key = colv[j] - 2.0f * s;
                    else if constexpr (MET == HM_METRIC_DOT_PRODUCT) key = s;
                    else if constexpr (MET == HM_METRIC_COSINE_SIMILARITY) key = -s * colv[j];
                    else key = s * colv[j];
                    if (colc[j] < n) mn = fminf(mn, key);
                }
                if (q < nq) out.Kmat[(size_t)q * out.ldk + btile * 32 + wc * 16 + lr] = mn;
            }
        }
        return;
    }
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t q = q0 + wr * 64 + i * 16 + lk * 4 + e;
                const bool qok = q < nq;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float s = acc[i][j][e];
                    if (ELEM == 1) s = s * (rows_[i][e] * cols[j]);
                    float key;
                    if constexpr (MET == HM_METRIC_EUCLIDEAN) key = colv[j] - 2.0f * s;
                    else if constexpr (MET == HM_METRIC_DOT_PRODUCT) key = s;
                    else if constexpr (MET == HM_METRIC_COSINE_SIMILARITY) key = -s * colv[j];
                    else key = s * colv[j];
                    const uint32_t c = colc[j];
                    if (qok && c < n) out.Kmat[(size_t)q * out.ldk + c] = key;
                }
            }
        }
    }
}

// ---- The filtered scan --------------------------------------------------
//
// A persistent grid of one 512-thread workgroup per CU walks 256-query x
// 256-row tiles: 8 waves in a 2 x 4 arrangement, each wave a 128 x 64 tile
// of 8 x 4 MFMA 16x16 blocks (128 fp32 accumulators per lane).  A K step
// brings 128 bytes of each of the 512 rows into LDS (64 KiB; two buffers):
// half the bytes per MFMA of a 128 x 128 tile, within what the direct loads
// sustain from L2.  Workgroup w takes tiles w, w + P, w + 2P, ... of the
// XCD-aware order (P a multiple of 8: each workgroup stays on its XCD's
// tile range), and the direct loads run one step ahead across tile
// boundaries: the next tile's first step and its epilogue operands are in
// flight while the current tile finishes.  The step barrier is a bare
// s_barrier after an explicit wait: __syncthreads' fence would also wait
// for the record stores of the epilogue.
//
// The epilogue writes no list: a lane whose four keys of a row (columns
// col0 + 16 j) include one below the row's threshold writes a record (the
// four keys and {query row, col0}) to its wave's own region at a position
// counted by ballots, with no atomic and no LDS staging.  A second kernel
// (scan_gather_kernel) appends the keys below the threshold to the lists.
// A record past the region's end marks its query's list as overflowed
// (count past cap: the list is not trusted).
constexpr int SM = 256, SN = 256, SCAN_WAVES = 8;

struct ScanRecords {
    float4 *keys;     // [waves][R]
    uint2 *meta;      // [waves][R]: {query row, base row of key 0}
    uint32_t *count;  // [waves]: records of the wave (past R: dropped)
    uint32_t R;
};

__device__ __forceinline__ void scan_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int ELEM, int MET>
__global__ __launch_bounds__(512, 1) void score_scan_kernel(
    const uint8_t *__restrict__ Q, uint32_t nq, uint32_t TX, uint32_t TY, uint32_t G,
    const uint8_t *__restrict__ B, uint32_t n, uint32_t row_bytes, const float *__restrict__ qscale,
    const float *__restrict__ bscale, const float *__restrict__ bnorm2,
    const float *__restrict__ inv_bnorm, const float *__restrict__ tau, uint32_t tau_ld,
    uint32_t *__restrict__ cnt, uint32_t cap, ScanRecords rec, const uint32_t *__restrict__ dev_rows)
{
    if (dev_rows && *dev_rows < nq) nq = *dev_rows;
    const uint32_t TYe = (nq + SM - 1) / SM;  // query tiles holding rows
    const uint32_t T = TX * TY, P = gridDim.x;
    // the two buffers are separate objects: the compiler then knows that a
    // direct load into one does not write the other, and reads of the
    // current buffer need not wait for the loads in flight
    __shared__ __align__(1024) uint8_t sA0[SM * ROWB], sA1[SM * ROWB];
    __shared__ __align__(1024) uint8_t sB0[SN * ROWB], sB1[SN * ROWB];
    // epilogue operands of the current and the next tile: per query row
    // scale and threshold, per base row norm term and scale
    __shared__ __align__(16) float s_rs[2][SM], s_rt[2][SM];
    __shared__ float s_cv[2][SN], s_cs[2][SN];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    const int lr = lane & 15, lk = lane >> 4;
    const uint32_t wave_s = (uint32_t)__builtin_amdgcn_readfirstlane(wave);
    const uint32_t nk = row_bytes / ROWB;
    const uint32_t W = blockIdx.x * SCAN_WAVES + wave_s;
    float4 *rk = rec.keys + (size_t)W * rec.R;
    uint2 *rm = rec.meta + (size_t)W * rec.R;
    uint32_t rec_n = 0;

    auto next_tile = [&](uint32_t L, uint32_t &bt, uint32_t &qt) __attribute__((always_inline)) -> uint32_t {
        for (; L < T; L += P) {
            tile_of(L, TX, TY, G, &bt, &qt);
            if (qt < TYe) return L;
        }
        return T;
    };
    uint32_t bt, qt;
    uint32_t L = next_tile(blockIdx.x, bt, qt);
    if (L >= T) {
        if (lane == 0) rec.count[W] = 0;
        return;
    }

    // each wave's direct loads fill rows 8 (4 w + c) + l / 8 of both
    // operands: per-lane 32-bit byte offsets from the tile's first rows,
    // whose addresses are wave-uniform (rows past the end load the last row)
    uint32_t qo[4], bo[4];
    const uint8_t *Qt = Q, *Bt = B;
    auto set_rows = [&](uint32_t bt_, uint32_t qt_) __attribute__((always_inline)) {
        const uint32_t q0 = qt_ * SM, n0 = bt_ * SN;
        Qt = Q + (size_t)q0 * row_bytes;
        Bt = B + (size_t)n0 * row_bytes;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t r = 8 * (4 * wave + c) + lane / 8;
            const uint32_t g = (uint32_t)(lane % 8) ^ swz(r);
            const uint32_t qa = q0 + r < nq ? r : nq - 1 - q0, ba = n0 + r < n ? r : n - 1 - n0;
            qo[c] = qa * row_bytes + 16 * g;
            bo[c] = ba * row_bytes + 16 * g;
        }
    };
    auto issue = [&](auto bufc, uint32_t kb) __attribute__((always_inline)) {
        constexpr int buf = decltype(bufc)::value;
        const uint8_t *Qk = Qt + kb, *Bk = Bt + kb;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            __builtin_amdgcn_global_load_lds(Qk + qo[c], &(buf ? sA1 : sA0)[(4 * wave_s + c) * 1024], 16, 0, 0);
            __builtin_amdgcn_global_load_lds(Bk + bo[c], &(buf ? sB1 : sB0)[(4 * wave_s + c) * 1024], 16, 0, 0);
        }
    };
    // tid through an empty asm: the per-tile helpers below recompute their
    // addresses where they run instead of keeping them live across the K loop
    auto opaque_tid = [&]() __attribute__((always_inline)) {
        uint32_t t = (uint32_t)tid;
        asm volatile("" : "+v"(t));
        return t;
    };
    // FP8 Euclidean keys are tested as halves (below): their threshold is
    // stored as -rt/2.  A row past nq gets a threshold no key passes.
    constexpr bool HALFKEY = ELEM == 1 && MET == HM_METRIC_EUCLIDEAN;
    auto load_ops = [&](uint32_t bt_, uint32_t qt_, float &o0, float &o1) __attribute__((always_inline)) {
        const uint32_t tid = opaque_tid();
        if (tid < SM) {
            const uint32_t q = qt_ * SM + tid, qc = q < nq ? q : nq - 1;
            o0 = ELEM == 1 ? qscale[qc] : 1.0f;
            o1 = tau[(size_t)qc * tau_ld];  // transformed when stored: its load is not waited on here
        } else {
            const uint32_t c = bt_ * SN + (tid - SM), cc = c < n ? c : n - 1;
            if constexpr (MET == HM_METRIC_EUCLIDEAN) o0 = bnorm2[cc];
            else if constexpr (MET == HM_METRIC_DOT_PRODUCT) o0 = 0.0f;
            else o0 = inv_bnorm[cc];
            o1 = ELEM == 1 ? bscale[cc] : 1.0f;
        }
    };
    auto store_ops = [&](int par, uint32_t qt_, float o0, float o1) __attribute__((always_inline)) {
        const uint32_t tid = opaque_tid();
        if (tid < SM) {
            constexpr float PINF = __builtin_huge_valf();
            const bool qok = qt_ * SM + tid < nq;
            s_rs[par][tid] = o0;
            if constexpr (HALFKEY) s_rt[par][tid] = qok ? -0.5f * o1 : PINF;
            else s_rt[par][tid] = qok ? o1 : -PINF;
        } else {
            s_cv[par][tid - SM] = o0;
            s_cs[par][tid - SM] = o1;
        }
    };
    // fragment offsets within a buffer: h = 0, 1 are the lane's two 16-byte
    // chunks of a row (FP8: chunks 2 lk, 2 lk + 1; BF16: chunks lk, 4 + lk).
    // Row blocks i and i + 4 are 64 rows apart, which leaves the swizzle
    // unchanged: block i + 4 reads at fa[i] + 64 rows (an immediate).
    uint32_t fa[4][2], fb[4][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t ch_ = ELEM == 1 ? 2 * lk + h : 4 * h + lk;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t ra = wr * 128 + i * 16 + lr;
            fa[i][h] = ra * ROWB + 16 * (ch_ ^ swz(ra));
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t rb = wc * 64 + j * 16 + lr;
            fb[j][h] = rb * ROWB + 16 * (ch_ ^ swz(rb));
        }
    }

    // FP8: the MFMA applies the rows' power-of-two scales, so acc holds the
    // dot product s itself.  Euclidean keys cv - 2 s are tested as halves:
    // t = s - cv/2 (cv/2 = +inf for a column past n), key = -2 t exactly,
    // and key < rt becomes t > -rt/2 (scaling by -2 and halving are exact),
    // one max over a lane's four columns and one compare per row.
    floatx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // E8M0 scale bytes of the lane's rows: query blocks i (byte i & 3 of
    // sa_pk[i >> 2]) and base blocks j (byte j of sb_pk)
    uint32_t sa_pk[2] = {0u, 0u}, sb_pk = 0u;
    auto tile_start = [&](int par) __attribute__((always_inline)) {
        const uint32_t lane = opaque_tid() & 63u;
        const uint32_t lr = lane & 15u;
        if constexpr (ELEM == 1) {
            sa_pk[0] = sa_pk[1] = sb_pk = 0u;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                sa_pk[i >> 2] |= ((__float_as_uint(s_rs[par][wr * 128 + i * 16 + lr]) >> 23) & 0xFFu)
                                 << (8 * (i & 3));
#pragma unroll
            for (int j = 0; j < 4; ++j)
                sb_pk |= ((__float_as_uint(s_cs[par][wc * 64 + j * 16 + lr]) >> 23) & 0xFFu) << (8 * j);
        }
    };

    // The epilogue runs inside the tile's last K step: the rows of query
    // block i are tested while the MFMAs of block i + 1 are in the matrix
    // core, so its VALU work overlaps the matrix work of both waves of the
    // SIMD instead of leaving the matrix core idle.
    // epi_cols: the column operands of the lane's four columns
    auto epi_cols = [&](uint32_t bt_, int par, float (&cvv)[4], bool (&cok)[4]) __attribute__((always_inline)) {
        constexpr float PINF = __builtin_huge_valf();
        const uint32_t lane = opaque_tid() & 63u;
        const uint32_t lr = lane & 15u;
        const uint32_t n0 = bt_ * SN;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t cl = wc * 64 + j * 16 + lr;
            cok[j] = n0 + cl < n;
            cvv[j] = (MET == HM_METRIC_EUCLIDEAN && !cok[j]) ? PINF : s_cv[par][cl];
            if constexpr (HALFKEY) cvv[j] = 0.5f * cvv[j];
        }
    };
    // row_keys: the lane's four keys of row e of query block i against the
    // row's stored threshold rt; true when one passes
    auto row_keys = [&](auto ic, int e, float rt, const float (&cvv)[4], const bool (&cok)[4],
                        float (&key)[4]) __attribute__((always_inline)) -> bool {
        constexpr int i = decltype(ic)::value;
        constexpr float PINF = __builtin_huge_valf();
        if constexpr (HALFKEY) {
            // key < rt  <=>  t > -rt/2 (the stored threshold); fmaxf drops a
            // NaN, whose key never compares below rt anyway
#pragma unroll
            for (int j = 0; j < 4; ++j) key[j] = acc[i][j][e] - cvv[j];  // t (cvv = cv/2)
            return fmaxf(fmaxf(key[0], key[1]), fmaxf(key[2], key[3])) > rt;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float a = acc[i][j][e];
                if constexpr (MET == HM_METRIC_EUCLIDEAN) {  // BF16: cv - 2 s
                    key[j] = cvv[j] - 2.0f * a;
                } else {  // FP8: acc is s itself (scales applied by the MFMA)
                    float kv;
                    if constexpr (MET == HM_METRIC_DOT_PRODUCT) kv = a;
                    else if constexpr (MET == HM_METRIC_COSINE_SIMILARITY) kv = -a * cvv[j];
                    else kv = a * cvv[j];
                    key[j] = cok[j] ? kv : PINF;
                }
            }
            // fminf drops a NaN key, which never compares below rt anyway
            return fminf(fminf(key[0], key[1]), fminf(key[2], key[3])) < rt;
        }
    };
    // the four rows' thresholds of query block i held by the lane (rows past
    // nq: no key passes), one 16-byte read
    auto block_rt = [&](auto ic, int par, uint32_t lk) __attribute__((always_inline)) -> float4 {
        constexpr int i = decltype(ic)::value;
        return *reinterpret_cast<const float4 *>(&s_rt[par][wr * 128 + i * 16 + lk * 4]);
    };
    // epi_block: the records of the four rows of query block i held by the
    // lane
    auto epi_block = [&](auto ic, uint32_t bt_, uint32_t qt_, int par, const float (&cvv)[4],
                         const bool (&cok)[4]) __attribute__((always_inline)) {
        constexpr int i = decltype(ic)::value;
        const uint32_t lane = opaque_tid() & 63u;
        const uint32_t lr = lane & 15u, lk = lane >> 4;
        const uint32_t col0 = bt_ * SN + wc * 64 + lr;
        const float4 t4 = block_rt(ic, par, lk);
        const float rtv[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t q = qt_ * SM + wr * 128 + i * 16 + lk * 4 + e;
            float key[4];
            const bool hit = row_keys(ic, e, rtv[e], cvv, cok, key);
            const uint64_t bal = __ballot(hit);
            if (bal) {
                if constexpr (HALFKEY) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) key[j] = -2.0f * key[j];  // cv - 2 s, exactly
                }
                if (hit) {
                    const uint32_t pos = rec_n + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    if (pos < rec.R) {
                        rk[pos] = make_float4(key[0], key[1], key[2], key[3]);
                        rm[pos] = make_uint2(q, col0);
                    } else {
                        atomicAdd(cnt + q, cap + 1);  // mark the list overflowed
                    }
                }
                rec_n += (uint32_t)__popcll(bal);
            }
        }
    };

    set_rows(bt, qt);
    issue(std::integral_constant<int, 0>{}, 0u);
    {
        float o0, o1;
        load_ops(bt, qt, o0, o1);
        store_ops(0, qt, o0, o1);
    }
    uint32_t nbt = 0, nqt = 0;
    uint32_t NL = next_tile(L + P, nbt, nqt);
    float p0 = 0.0f, p1 = 0.0f;  // the next tile's epilogue operands
    int par = 0;
    uint32_t k = 0;
    // one K step from buffer buf; true when the workgroup has no tile left
    auto step = [&](auto bufc) __attribute__((always_inline)) -> bool {
        constexpr int buf = decltype(bufc)::value;
        using Other = std::integral_constant<int, buf ^ 1>;
        __builtin_amdgcn_s_waitcnt(0);  // this wave's loads of the step have landed
        scan_barrier();                 // ... and every wave's; buffer buf ^ 1 is free
        if (k + 1 < nk) {
            issue(Other{}, (k + 1) * ROWB);
        } else if (NL < T) {
            set_rows(nbt, nqt);
            issue(Other{}, 0u);
        }
        if (k == 0) {
            tile_start(par);
            if (NL < T) load_ops(nbt, nqt, p0, p1);
        }
        const uint8_t *As = buf ? sA1 : sA0, *Bs = buf ? sB1 : sB0;
        uint4 bv[4][2];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int h = 0; h < 2; ++h) bv[j][h] = *reinterpret_cast<const uint4 *>(Bs + fb[j][h]);
        // query fragments: block i + 1 is read while block i's MFMAs issue
        // (i and j are compile-time indices: the MFMA's scale-byte selects
        // must be constants)
        uint4 avb[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) avb[0][h] = *reinterpret_cast<const uint4 *>(As + fa[0][h]);
        // the MFMAs of the step, block by block; in the tile's last step each
        // block's rows are tested after the next block's MFMAs
        const bool last = k + 1 == nk;
        float cvv[4] = {0.f, 0.f, 0.f, 0.f};
        bool cok[4] = {false, false, false, false};
        if (last) epi_cols(bt, par, cvv, cok);
        {
            static_for<0, 8>([&](auto ic) __attribute__((always_inline)) {
                constexpr int i = decltype(ic)::value;
                if (i + 1 < 8) {
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        avb[(i + 1) & 1][h] = *reinterpret_cast<const uint4 *>(
                            As + fa[(i + 1) & 3][h] + ((i + 1) >> 2) * 64 * ROWB);
                }
                const uint4 *av = avb[i & 1];
                if constexpr (ELEM == 2) {
                    // two MFMA K=32 per step; lane holds 8 bf16 at k = 32*kk + 8*lk
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(bf16x8, av[kk]), __builtin_bit_cast(bf16x8, bv[j][kk]),
                                acc[i][j], 0, 0, 0);
                } else {
                    // one MFMA K=128 (e4m3 x e4m3); lane holds 32 fp8 at k = 32*lk
                    // in A and B alike
                    const i32x8 a = i32x8{(int)av[0].x, (int)av[0].y, (int)av[0].z, (int)av[0].w,
                                          (int)av[1].x, (int)av[1].y, (int)av[1].z, (int)av[1].w};
                    static_for<0, 4>([&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        const i32x8 b = i32x8{(int)bv[j][0].x, (int)bv[j][0].y, (int)bv[j][0].z,
                                              (int)bv[j][0].w, (int)bv[j][1].x, (int)bv[j][1].y,
                                              (int)bv[j][1].z, (int)bv[j][1].w};
                        // the rows' scales applied by the matrix core (E8M0 bytes
                        // i & 3 of sa_pk[i >> 2] and j of sb_pk)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                            a, b, acc[i][j], 0, 0, i & 3, (int)sa_pk[i >> 2], j, (int)sb_pk);
                    });
                }
                // the records of the previous block, whose MFMAs are done
                if constexpr (i > 0)
                    if (last) epi_block(std::integral_constant<int, i - 1>{}, bt, qt, par, cvv, cok);
            });
            if (last) epi_block(std::integral_constant<int, 7>{}, bt, qt, par, cvv, cok);
        }
        if (++k < nk) return false;
        if (NL >= T) return true;
        // the next tile: its operands go to the other half of the operand
        // arrays (the previous tile's epilogue, which read them, is done:
        // every step of this tile had a barrier)
        store_ops(par ^ 1, nqt, p0, p1);
        par ^= 1;
        bt = nbt;
        qt = nqt;
        L = NL;
        NL = next_tile(L + P, nbt, nqt);
        k = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        return false;
    };
    for (;;) {
        if (step(std::integral_constant<int, 0>{})) break;
        if (step(std::integral_constant<int, 1>{})) break;
    }
    if (lane == 0) rec.count[W] = rec_n;
}

// Appends the keys below the threshold of every record of scan region
// blockIdx.x to its query's list: one atomic per record reserves its hits.
// 1,024 threads per region: a region's records are usually one round, all
// of their atomics in flight together.
constexpr int GATHER_THREADS = 1024;
__global__ __launch_bounds__(GATHER_THREADS) void scan_gather_kernel(ScanRecords rec,
                                                          const float *__restrict__ tau,
                                                          uint32_t tau_ld, uint64_t *__restrict__ list,
                                                          uint32_t cap, uint32_t *__restrict__ cnt)
{
    const uint32_t W = blockIdx.x;
    const uint32_t c0 = rec.count[W], c = c0 < rec.R ? c0 : rec.R;
    const float4 *rk = rec.keys + (size_t)W * rec.R;
    const uint2 *rm = rec.meta + (size_t)W * rec.R;
    for (uint32_t r = threadIdx.x; r < c; r += GATHER_THREADS) {
        const float4 kv = rk[r];
        const uint2 m = rm[r];
        const float t = tau[(size_t)m.x * tau_ld];
        const float kk[4] = {kv.x, kv.y, kv.z, kv.w};
        uint32_t hits = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) hits |= (kk[j] < t ? 1u : 0u) << j;
        if (!hits) continue;
        uint32_t pos = atomicAdd(cnt + m.x, (uint32_t)__builtin_popcount(hits));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if ((hits >> j) & 1u) {
                if (pos < cap)
                    list[(size_t)m.x * cap + pos] = ((uint64_t)fkey(kk[j]) << 32) | (m.y + 16u * j);
                ++pos;
            }
        }
    }
}

// Rows of a strided sample of the base: dst row i = src row floor(i*n/ns).
__global__ void sample_rows_kernel(const uint8_t *__restrict__ src, uint32_t row_bytes, uint32_t n,
                                   uint32_t ns, uint8_t *__restrict__ dst)
{
    const uint32_t words = row_bytes / 4;
    const uint64_t total = (uint64_t)ns * words;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / words, w = e % words;
        const uint64_t r = i * n / ns;
        reinterpret_cast<uint32_t *>(dst)[e] =
            reinterpret_cast<const uint32_t *>(src)[r * words + w];
    }
}

__global__ void sample_f32_kernel(const float *__restrict__ src, uint32_t n, uint32_t ns,
                                  float *__restrict__ dst)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x)
        dst[i] = src[(uint64_t)i * n / ns];
}

// Round-to-nearest-even f32 -> bf16; flags a value that does not survive.
__global__ void to_bf16_kernel(const float *__restrict__ src, uint64_t lds, uint64_t rows,
                               uint32_t d, uint32_t dk, uint16_t *__restrict__ dst,
                               uint32_t *inexact, const uint32_t *__restrict__ dev_rows)
{
    if (dev_rows && *dev_rows < rows) rows = *dev_rows;
    uint64_t total = rows * dk;
    uint32_t bad = 0;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = e / dk, c = e % dk;
        float x = c < d ? src[r * lds + c] : 0.0f;
        uint32_t u = __float_as_uint(x);
        uint32_t h = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
        if (__uint_as_float(h << 16) != x) bad = 1;
        dst[r * dk + c] = (uint16_t)h;
    }
    if (bad && inexact) atomicOr(inexact, 1u);
}

// Per-row scaled FP8-e4m3 with a power-of-two scale: x * mult rounded by
// the hardware converter, mult = 2^(8 - e) for max|x| = f 2^e, f in
// [0.5, 1), so that max|x| * mult lies in [128, 256) (below e4m3's 448);
// dequant[r] = 1 / mult exactly (0 for a zero row).  The scan applies it
// as the row's E8M0 block scale (fp8_scale_bits).  One wave per row.
__device__ __forceinline__ float fp8_mult(float m) { int e = 0; frexpf(m, &e); return m > 0.0f ? ldexpf(1.0f, 8 - e) : 1.0f; }
__device__ __forceinline__ float fp8_dequant(float m) { int e = 0; frexpf(m, &e); return m > 0.0f ? ldexpf(1.0f, e - 8) : 0.0f; }
__global__ void to_fp8_kernel(const float *__restrict__ src, uint64_t lds, uint32_t rows,
                              uint32_t d, uint32_t dk, uint8_t *__restrict__ dst,
                              float *__restrict__ dequant, const uint32_t *__restrict__ dev_rows)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows || (dev_rows && r >= *dev_rows)) return;
    const float *x = src + (size_t)r * lds;
    float m = 0.0f;
    for (uint32_t c = lane; c < d; c += 64) m = fmaxf(m, fabsf(x[c]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    const float scale = fp8_mult(m);
    if (lane == 0) dequant[r] = fp8_dequant(m);
    uint8_t *o = dst + (size_t)r * dk;
    for (uint32_t c = 2 * lane; c < dk; c += 128) {
        float a = c < d ? x[c] * scale : 0.0f;
        float b = c + 1 < d ? x[c + 1] * scale : 0.0f;
        int v = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
        o[c] = (uint8_t)(v & 0xFF);
        if (c + 1 < dk) o[c + 1] = (uint8_t)((v >> 8) & 0xFF);
    }
}

// The query side of the shortlist forms in one pass, one wave per row, for
// rows 16-byte aligned with d % 4 == 0 and d <= 1024: lane l holds the
// float4s at columns 4 (l + 64 j), j < 4, and from them
//  - err |= 2 if a value is NaN or infinite;
//  - norm2[r] = sum_t x_t*x_t as a left fold from -0.0 (lane 0 of the wave
//    over the row staged in LDS; the same bits as self_dot_kernel);
//  - MODE 1: the scaled FP8-e4m3 row of to_fp8_kernel (power-of-two scale,
//    dequant[r] its inverse), four bytes per lane store;
//    MODE 2: the RNE BF16 row of to_bf16_kernel, *inexact set if a value
//    does not survive;
//  - pad (when given): the row zero-padded to dpad f32 columns.
// Columns past d are zero in every output, up to dk (dpad for pad).
template <int MODE>
__global__ __launch_bounds__(256) void query_prep_kernel(
    const float *__restrict__ src, uint64_t lds, uint32_t rows, uint32_t d, uint32_t dk,
    void *__restrict__ dst, float *__restrict__ dequant, uint32_t *__restrict__ inexact,
    float *__restrict__ norm2, float *__restrict__ pad, uint32_t dpad, uint32_t *__restrict__ err,
    const uint32_t *__restrict__ dev_rows)
{
    __shared__ float4 srow[4][256];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (dev_rows && *dev_rows < rows) rows = *dev_rows;
    const uint32_t r = blockIdx.x * 4 + wave;
    const bool live = r < rows;  // no early return: the workgroup barrier below
    const float *x = src + (size_t)(live ? r : 0) * lds;
    float4 v[4];
    float m = 0.0f;
    bool bad = false;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t c = 4 * (lane + 64 * j);
        v[j] = live && c < d ? *reinterpret_cast<const float4 *>(x + c) : make_float4(0, 0, 0, 0);
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v[j].x), fabsf(v[j].y)), fmaxf(fabsf(v[j].z), fabsf(v[j].w))));
        bad |= !(isfinite(v[j].x) && isfinite(v[j].y) && isfinite(v[j].z) && isfinite(v[j].w));
        srow[wave][lane + 64 * j] = v[j];
    }
    if (err && __ballot(bad) && lane == 0) atomicOr(err, 2u);
    __syncthreads();
    if (!live) return;
    if (lane == 0) {
        float acc = -0.0f;
        const float4 *p = srow[wave];
        for (uint32_t i = 0; i < d / 4; ++i) {
            const float4 t = p[i];
            float q = t.x * t.x;
            acc = acc + q;
            q = t.y * t.y;
            acc = acc + q;
            q = t.z * t.z;
            acc = acc + q;
            q = t.w * t.w;
            acc = acc + q;
        }
        norm2[r] = acc;
    }
    if (pad) {
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t c = 4 * (lane + 64 * j);
            if (c < dpad) *reinterpret_cast<float4 *>(pad + (size_t)r * dpad + c) = v[j];
        }
    }
    if constexpr (MODE == 1) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        const float scale = fp8_mult(m);
        if (lane == 0) dequant[r] = fp8_dequant(m);
        uint32_t *o = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(dst) + (size_t)r * dk);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t c = 4 * (lane + 64 * j);
            if (c < dk) {
                const int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[j].x * scale, v[j].y * scale, 0,
                                                               false);
                o[c / 4] = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(v[j].z * scale,
                                                                     v[j].w * scale, lo, true);
            }
        }
    } else {
        uint32_t ib = 0;
        uint16_t *o = static_cast<uint16_t *>(dst) + (size_t)r * dk;
        auto rne = [&](float f) {
            const uint32_t u = __float_as_uint(f);
            const uint32_t h = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
            ib |= __uint_as_float(h << 16) != f ? 1u : 0u;
            return h;
        };
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t c = 4 * (lane + 64 * j);
            if (c < dk) {
                uint2 w;
                w.x = rne(v[j].x) | (rne(v[j].y) << 16);
                w.y = rne(v[j].z) | (rne(v[j].w) << 16);
                *reinterpret_cast<uint2 *>(o + c) = w;
            }
        }
        if (inexact && __ballot(ib != 0) && lane == 0) atomicOr(inexact, 1u);
    }
}

// dst = the u8 of each value; *inexact |= 1 if some value's bits differ
// from those of the f32 of its u8 (a fraction, a value outside 0..255,
// -0.0 or NaN): then the copy does not hold the base.
__global__ void to_u8_exact_kernel(const float *__restrict__ src, uint64_t total,
                                   uint8_t *__restrict__ dst, uint32_t *__restrict__ inexact)
{
    bool bad = false;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float v = src[i];
        const uint8_t u = (v >= 0.0f && v <= 255.0f) ? (uint8_t)v : (uint8_t)0;
        bad |= __float_as_uint(v) != __float_as_uint((float)u);
        dst[i] = u;
    }
    if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(inexact, 1u);
}

// inv[r] = 1 / sqrt(norm2[r]); stats[0] = max_r sqrt(norm2[r]) (as float
// bits); stats[2] = 1 if some row has norm zero.
__global__ void norm_stats_kernel(const float *__restrict__ norm2, uint32_t rows,
                                  float *__restrict__ inv, uint32_t *stats)
{
    float m = 0.0f;
    uint32_t zero = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows;
         r += gridDim.x * blockDim.x) {
        float s = sqrtf(norm2[r]);
        if (inv) inv[r] = 1.0f / s;
        m = fmaxf(m, s);
        zero |= s == 0.0f ? 1u : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(stats, __float_as_uint(m));
    if (zero) atomicOr(stats + 2, 1u);
}

}  // namespace

template <bool MIN4>
static void score_launch(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                         uint32_t row_bytes, const float *qscale, const float *bscale,
                         const float *bnorm2, const float *inv_bnorm, int metric,
                         const ScoreOut &out, hipStream_t s, const uint32_t *dev_rows,
                         uint32_t group_kib)
{
    if (nq == 0 || n == 0) return;
    if (row_bytes == 0 || row_bytes % ROWB)
        fail(HM_ERR_INVALID_ARG, "score_gemm: row bytes not a multiple of 128");
    if (elem_bytes != 1 && elem_bytes != 2) fail(HM_ERR_INVALID_ARG, "score_gemm: element size");
    const uint64_t TX = ceil_div(n, GN), TY = ceil_div(nq, GM);
    if (TX * TY > 0x7FFFFFull) fail(HM_ERR_INVALID_ARG, "score_gemm: too many tiles");
    const dim3 grid((unsigned)(TX * TY));
    // base tiles per group: about 2 MiB of base rows per XCD's L2
    const uint64_t group_bytes = (uint64_t)(group_kib ? group_kib : 2048) << 10;
    const uint32_t G = (uint32_t)std::min<uint64_t>(
        32, std::max<uint64_t>(2, group_bytes / ((uint64_t)GN * row_bytes)));
    // 32-bit per-lane load offsets when both operands span less than 4 GiB
    const bool off32 = (uint64_t)nq * row_bytes <= 0xFFFFFFFFull &&
                       (uint64_t)n * row_bytes <= 0xFFFFFFFFull;
#define HM_SG2(E, M, O)                                                                       \
    score_gemm_kernel<E, M, MIN4, O><<<grid, 256, 0, s>>>(                            \
        (const uint8_t *)Q, nq, (uint32_t)TY, G, (const uint8_t *)B, n, row_bytes, qscale,     \
        bscale, bnorm2, inv_bnorm, out, dev_rows)
#define HM_SG(E, M)                                                                           \
    if (off32) HM_SG2(E, M, true);                                                            \
    else HM_SG2(E, M, false)
#define HM_SG_METRIC(E)                                                                       \
    switch (metric) {                                                                         \
    case HM_METRIC_EUCLIDEAN: HM_SG(E, HM_METRIC_EUCLIDEAN); break;                           \
    case HM_METRIC_DOT_PRODUCT: HM_SG(E, HM_METRIC_DOT_PRODUCT); break;                       \
    case HM_METRIC_COSINE_SIMILARITY: HM_SG(E, HM_METRIC_COSINE_SIMILARITY); break;           \
    case HM_METRIC_ANGULAR: HM_SG(E, HM_METRIC_ANGULAR); break;                               \
    default: fail(HM_ERR_INVALID_ARG, "score_gemm: metric has no shortlist form");            \
    }
    if (elem_bytes == 2) {
        HM_SG_METRIC(2)
    } else {
        HM_SG_METRIC(1)
    }
#undef HM_SG_METRIC
#undef HM_SG
#undef HM_SG2
    HM_CHECK_LAUNCH();
}

void launch_score_gemm(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                       uint32_t row_bytes, const float *qscale, const float *bscale,
                       const float *bnorm2, const float *inv_bnorm, int metric, float *Kmat,
                       uint64_t ldk, hipStream_t s, const uint32_t *dev_rows,
                       uint32_t group_kib)
{
    ProfScope prof("score_gemm", s);
    ScoreOut out{Kmat, ldk};
    score_launch<false>(elem_bytes, Q, nq, B, n, row_bytes, qscale, bscale, bnorm2, inv_bnorm,
                        metric, out, s, dev_rows, group_kib);
}

void launch_score_min4(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                       uint32_t row_bytes, const float *qscale, const float *bscale,
                       const float *bnorm2, const float *inv_bnorm, int metric, float *Kmat,
                       uint64_t ldk, hipStream_t s, const uint32_t *dev_rows,
                       uint32_t group_kib)
{
    if (ldk < (uint64_t)ceil_div(n, GN) * 32) fail(HM_ERR_INVALID_ARG, "score_min4: ldk too small");
    ProfScope prof("score_gemm", s);
    ScoreOut out{Kmat, ldk};
    score_launch<true>(elem_bytes, Q, nq, B, n, row_bytes, qscale, bscale, bnorm2,
                              inv_bnorm, metric, out, s, dev_rows, group_kib);
}

// Workgroups of the persistent scan: one per CU, a multiple of 8 (the XCDs).
static uint32_t scan_workgroups()
{
    int dev = 0, cus = 0;
    HM_HIP(hipGetDevice(&dev));
    HM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return (uint32_t)std::max(8, cus / 8 * 8);
}

// Scratch of the filtered scan: the record count of each of its waves, then
// their record regions (16-byte keys and 8-byte meta each) of
// max(1024, nq * cap / waves) records.
static uint64_t scan_count_bytes(uint64_t waves) { return (waves * 4 + 255) / 256 * 256; }

size_t score_filter_scratch_bytes(uint32_t nq, uint32_t cap)
{
    const uint64_t waves = (uint64_t)scan_workgroups() * SCAN_WAVES;
    const uint64_t R = std::max<uint64_t>(1024, ceil_div((uint64_t)nq * cap, waves));
    return (size_t)(scan_count_bytes(waves) + waves * R * 24);
}

void launch_score_filter(int elem_bytes, const void *Q, uint32_t nq, const void *B, uint32_t n,
                         uint32_t row_bytes, const float *qscale, const float *bscale,
                         const float *bnorm2, const float *inv_bnorm, int metric,
                         const float *tau, uint32_t tau_ld, uint64_t *list, uint32_t cap,
                         uint32_t *cnt, void *scratch, size_t scratch_bytes, hipStream_t s,
                         const uint32_t *dev_rows, uint32_t group_kib)
{
    if (!tau || !list || !cnt || cap == 0 || !scratch)
        fail(HM_ERR_INVALID_ARG, "score_filter: null output");
    if (nq == 0 || n == 0) return;
    if (row_bytes == 0 || row_bytes % ROWB)
        fail(HM_ERR_INVALID_ARG, "score_filter: row bytes not a multiple of 128");
    if (elem_bytes != 1 && elem_bytes != 2) fail(HM_ERR_INVALID_ARG, "score_filter: element size");
    const uint64_t TX = ceil_div(n, SN), TY = ceil_div(nq, SM);
    if (TX * TY > 0xFFFFFFFFull - 4096) fail(HM_ERR_INVALID_ARG, "score_filter: too many tiles");
    const uint32_t P = scan_workgroups();
    const uint64_t waves = (uint64_t)P * SCAN_WAVES, cb = scan_count_bytes(waves);
    if (scratch_bytes < cb + waves * 24) fail(HM_ERR_INVALID_ARG, "score_filter: scratch too small");
    const uint64_t R = std::min<uint64_t>((scratch_bytes - cb) / (waves * 24), 0xFFFFFFFFull);
    uint8_t *base = static_cast<uint8_t *>(scratch);
    ScanRecords rec{reinterpret_cast<float4 *>(base + cb),
                    reinterpret_cast<uint2 *>(base + cb + waves * R * 16),
                    reinterpret_cast<uint32_t *>(base), (uint32_t)R};
    // base tiles per group: about 2 MiB of base rows per XCD's L2
    const uint64_t group_bytes = (uint64_t)(group_kib ? group_kib : 2048) << 10;
    const uint32_t G = (uint32_t)std::min<uint64_t>(
        32, std::max<uint64_t>(2, group_bytes / ((uint64_t)SN * row_bytes)));
    {
        ProfScope prof("score_filter", s);
#define HM_SC(E, M)                                                                           \
    score_scan_kernel<E, M><<<P, 64 * SCAN_WAVES, 0, s>>>(                                    \
        (const uint8_t *)Q, nq, (uint32_t)TX, (uint32_t)TY, G, (const uint8_t *)B, n,         \
        row_bytes, qscale, bscale, bnorm2, inv_bnorm, tau, tau_ld, cnt, cap, rec, dev_rows)
#define HM_SC_METRIC(E)                                                                       \
    switch (metric) {                                                                         \
    case HM_METRIC_EUCLIDEAN: HM_SC(E, HM_METRIC_EUCLIDEAN); break;                           \
    case HM_METRIC_DOT_PRODUCT: HM_SC(E, HM_METRIC_DOT_PRODUCT); break;                       \
    case HM_METRIC_COSINE_SIMILARITY: HM_SC(E, HM_METRIC_COSINE_SIMILARITY); break;           \
    case HM_METRIC_ANGULAR: HM_SC(E, HM_METRIC_ANGULAR); break;                               \
    default: fail(HM_ERR_INVALID_ARG, "score_filter: metric has no shortlist form");          \
    }
        if (elem_bytes == 2) {
            HM_SC_METRIC(2)
        } else {
            HM_SC_METRIC(1)
        }
#undef HM_SC_METRIC
#undef HM_SC
        HM_CHECK_LAUNCH();
    }
    {
        ProfScope prof("list_gather", s);
        scan_gather_kernel<<<(unsigned)waves, GATHER_THREADS, 0, s>>>(rec, tau, tau_ld, list, cap, cnt);
        HM_CHECK_LAUNCH();
    }
}

void launch_sample_rows(const void *src, uint32_t row_bytes, uint32_t n, uint32_t ns, void *dst,
                        hipStream_t s)
{
    if (ns == 0) return;
    if (row_bytes % 4 || ns > n) fail(HM_ERR_INVALID_ARG, "sample_rows: bad shape");
    const unsigned blocks = (unsigned)std::min<uint64_t>(((uint64_t)ns * row_bytes / 4 + 255) / 256, 4096);
    sample_rows_kernel<<<blocks, 256, 0, s>>>((const uint8_t *)src, row_bytes, n, ns,
                                              (uint8_t *)dst);
    HM_CHECK_LAUNCH();
}

void launch_sample_f32(const float *src, uint32_t n, uint32_t ns, float *dst, hipStream_t s)
{
    if (ns == 0) return;
    if (ns > n) fail(HM_ERR_INVALID_ARG, "sample_f32: bad shape");
    sample_f32_kernel<<<std::min<unsigned>(ceil_div(ns, 256), 1024), 256, 0, s>>>(src, n, ns, dst);
    HM_CHECK_LAUNCH();
}

void launch_to_bf16(const float *src, uint64_t lds, uint64_t rows, uint32_t d, uint32_t dk,
                    uint16_t *dst, uint32_t *inexact, hipStream_t s, const uint32_t *dev_rows)
{
    if (rows == 0) return;
    unsigned blocks = (unsigned)std::min<uint64_t>((rows * dk + 255) / 256, 8192);
    to_bf16_kernel<<<blocks, 256, 0, s>>>(src, lds, rows, d, dk, dst, inexact, dev_rows);
    HM_CHECK_LAUNCH();
}

void launch_to_fp8(const float *src, uint64_t lds, uint32_t rows, uint32_t d, uint32_t dk,
                   uint8_t *dst, float *dequant, hipStream_t s, const uint32_t *dev_rows)
{
    if (rows == 0) return;
    to_fp8_kernel<<<ceil_div(rows, 4), 256, 0, s>>>(src, lds, rows, d, dk, dst, dequant, dev_rows);
    HM_CHECK_LAUNCH();
}

bool query_prep_ok(const float *src, uint64_t lds, uint32_t d, const float *pad, uint32_t dpad)
{
    return d > 0 && d % 4 == 0 && d <= 1024 && lds % 4 == 0 && ((uintptr_t)src & 15) == 0 &&
           (!pad || (dpad >= d && dpad % 4 == 0 && dpad <= 1024 && ((uintptr_t)pad & 15) == 0));
}

void launch_query_prep(int elem, const float *src, uint64_t lds, uint32_t rows, uint32_t d,
                       uint32_t dk, void *dst, float *dequant, uint32_t *inexact, float *norm2,
                       float *pad, uint32_t dpad, uint32_t *err, hipStream_t s,
                       const uint32_t *dev_rows)
{
    if (rows == 0) return;
    if (!query_prep_ok(src, lds, d, pad, dpad) || dk < d || dk > 1024 || dk % 4 != 0 ||
        ((uintptr_t)dst & 7) != 0)
        fail(HM_ERR_INVALID_ARG, "query_prep: unsupported row shape or alignment");
    ProfScope prof("query_prep", s);
    if (elem == 1)
        query_prep_kernel<1><<<ceil_div(rows, 4), 256, 0, s>>>(src, lds, rows, d, dk, dst, dequant,
                                                               inexact, norm2, pad, dpad, err,
                                                               dev_rows);
    else
        query_prep_kernel<2><<<ceil_div(rows, 4), 256, 0, s>>>(src, lds, rows, d, dk, dst, dequant,
                                                               inexact, norm2, pad, dpad, err,
                                                               dev_rows);
    HM_CHECK_LAUNCH();
}

void launch_to_u8_exact(const float *src, uint32_t dpad, uint64_t rows, uint8_t *dst,
                        uint32_t *inexact, hipStream_t s)
{
    const uint64_t total = rows * dpad;
    if (total == 0) return;
    const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 8192);
    to_u8_exact_kernel<<<blocks, 256, 0, s>>>(src, total, dst, inexact);
    HM_CHECK_LAUNCH();
}

void launch_norm_stats(const float *norm2, uint32_t rows, float *inv, uint32_t *stats,
                       hipStream_t s)
{
    if (rows == 0) return;
    unsigned blocks = std::min<unsigned>(ceil_div(rows, 256), 1024);
    norm_stats_kernel<<<blocks, 256, 0, s>>>(norm2, rows, inv, stats);
    HM_CHECK_LAUNCH();
}

}  // namespace hm
