// Counter-based input generator shared by host and device code.
//
// SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the i-th output of a generator with state `key`
// is mix64(key + (i + 1) * GOLDEN).  Because output i depends only on (key, i)
// any element of a 128 GiB matrix can be generated in place, and the host
// produces exactly the same integers.  The mapping to doubles uses the top 53
// bits and a single correctly rounded fma, so host and device doubles are
// bitwise identical as well.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MPLA_HD __host__ __device__ __forceinline__
#else
#define MPLA_HD static inline
#endif

namespace mplak {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

MPLA_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Generator state for (seed, stream): independent streams per input array.
MPLA_HD uint64_t rng_key(uint64_t seed, uint64_t stream) { return mix64(seed * kGolden + stream); }

MPLA_HD uint64_t rng_u64(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * kGolden); }

// Uniform in [0, 1) with 53 random bits.
MPLA_HD double rng_unit(uint64_t key, uint64_t i) {
  return (double)(rng_u64(key, i) >> 11) * 0x1.0p-53;
}

// lo + span * U, rounded once.
MPLA_HD double rng_uniform_ab(uint64_t key, uint64_t i, double lo, double span) {
  return __builtin_fma(span, rng_unit(key, i), lo);
}

}  // namespace mplak
