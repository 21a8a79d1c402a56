// pq_common.h -- k-means helpers shared by the IVFPQ and PQ kernels.
#pragma once

#include "hm_common.h"

namespace hm {

// SplitMix64, the seeded stand-in for the reference's thread_rng
// (kmeans.rs:168-171); the oracle uses the same generator.
__device__ __host__ __forceinline__ uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __host__ __forceinline__ double splitmix_unit(uint64_t *s)
{
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __host__ __forceinline__ uint64_t stream_seed(uint64_t seed, uint64_t instance)
{
    uint64_t s = seed ^ (instance * 0xD1B54A32D192ED03ull);
    return splitmix64(&s);
}

// Advances a SplitMix64 state by k draws: the state is a Weyl sequence, so
// skipping is one multiply-add.
__device__ __host__ __forceinline__ void splitmix_skip(uint64_t *s, uint64_t k)
{
    *s += k * 0x9E3779B97F4A7C15ull;
}

// Bit set of the centres with a nonzero count (words w[0 .. ceil(n/64))).
struct NonzeroSet {
    uint64_t *w;
    uint32_t n;
    __device__ __host__ void put(uint32_t i, bool on)
    {
        const uint64_t b = 1ull << (i & 63);
        w[i >> 6] = on ? (w[i >> 6] | b) : (w[i >> 6] & ~b);
    }
    // The first member at or after `from`, cyclically (the set is not empty).
    __device__ __host__ uint32_t next(uint32_t from) const
    {
        const uint32_t nw = (n + 63) / 64;
        uint32_t wi = from >> 6;
        uint64_t m = w[wi] & (~0ull << (from & 63));
        for (uint32_t t = 0; t <= nw; ++t) {
            if (m) return (wi << 6) + (uint32_t)__builtin_ctzll(m);
            wi = wi + 1 == nw ? 0 : wi + 1;
            m = w[wi];
        }
        return from;
    }
};

// split_center's pick walk (kmeans.rs:166-175): from sc = (i + 1) % nc, one
// draw u per step, stop at the first centre with u < count / n.  A centre
// whose count is 0 fails every draw (u >= 0), so a run of them is passed by
// advancing the generator over the run: the same draws are consumed and the
// same centre is picked as by the step-by-step walk, in work proportional to
// the nonzero centres visited.  The counts sum to n > 0, so the set is never
// empty.
template <class COUNT>
__device__ __host__ inline uint32_t split_pick(uint32_t i, uint32_t nc, uint64_t n, COUNT count,
                                               const NonzeroSet &nz, uint64_t *rng)
{
    uint32_t sc = (i + 1) % nc;
    for (;;) {
        const uint32_t nx = nz.next(sc);
        splitmix_skip(rng, (uint64_t)((nx + nc - sc) % nc));
        sc = nx;
        const double pick = (double)count(sc) / (double)n;
        if (splitmix_unit(rng) < pick) return sc;
        sc = (sc + 1) % nc;
    }
}

// The draw that SplitMix64 state s yields after t + 1 more draws, as a unit
// double (the state itself is not advanced).
__device__ __host__ __forceinline__ double splitmix_unit_at(uint64_t s, uint64_t t)
{
    uint64_t z = s + (t + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// kmeans.rs:76: (i & (1 << j)) == 1, shift amount masked to 6 bits.
__device__ __host__ __forceinline__ float init_offset(uint64_t i, uint32_t j)
{
    return ((i & (1ull << (j & 63))) == 1ull) ? 1.0f : -1.0f;
}

}  // namespace hm
