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

// kmeans.rs:76: (i & (1 << j)) == 1, shift amount masked to 6 bits.
__device__ __host__ __forceinline__ float init_offset(uint64_t i, uint32_t j)
{
    return ((i & (1ull << (j & 63))) == 1ull) ? 1.0f : -1.0f;
}

}  // namespace hm
