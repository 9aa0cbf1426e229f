// pdr_rng.hpp -- counter-based random numbers shared by host and device code.
//
// The reference draws from gfortran's random_number stream (src/pg_random.f90),
// which cannot be reproduced outside gfortran.  This library instead derives
// every draw from (seed, draw index) with SplitMix64 (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014), so a draw is
// the same on any thread, device or host and no generator state is carried.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pdr {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// uniform in [0, 1): top 53 bits of the mixed (seed, index) pair
__host__ __device__ inline double u01(uint64_t seed, uint64_t idx) {
  const uint64_t h = splitmix64(splitmix64(seed) ^ (idx * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
  return double(h >> 11) * (1.0 / 9007199254740992.0);
}

// Pair of independent N(0, 1) deviates by the Marsaglia polar form of the
// Box-Muller method, as g_RandomNormal (src/pg_random.f90:112-146) does with
// mean 0 and variance 1.  The rejection loop reads uniforms 2a, 2a+1 of the
// block of 64 that belongs to `key`; after 32 rejections (probability below
// 1e-21) both deviates are 0.
__host__ __device__ inline void normal_pair(uint64_t seed, uint64_t key, double* r1, double* r2) {
  *r1 = 0.0;
  *r2 = 0.0;
  for (int a = 0; a < 32; ++a) {
    const double x1 = 2.0 * u01(seed, key * 64u + uint64_t(2 * a)) - 1.0;
    const double x2 = 2.0 * u01(seed, key * 64u + uint64_t(2 * a + 1)) - 1.0;
    const double w = x1 * x1 + x2 * x2;
    if (w < 1.0 && w > 0.0) {
      const double f = sqrt((-2.0 * log(w)) / w);
      *r1 = x1 * f;
      *r2 = x2 * f;
      return;
    }
  }
}

}  // namespace pdr
