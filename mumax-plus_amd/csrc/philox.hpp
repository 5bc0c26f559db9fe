// Counter-based random numbers for the thermal field.
//
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC11): ten rounds of two 32x32->64 multiplies with the
// multipliers 0xD2511F53, 0xCD9E8D57, the key bumped by the Weyl constants
// 0x9E3779B9, 0xBB67AE85 between rounds.  A (counter, key) pair maps to four
// independent 32-bit words, so every cell draws its own numbers without any
// generator state: the same counter always gives the same numbers, whichever
// thread or launch computes them.  (The reference draws from a cuRAND
// generator seeded by the clock, thermalnoise.cu:55-61, ferromagnet.cpp:79-81.)
//
// Normal deviates: Box-Muller on the top 24 bits of each word, u = (x >> 8 +
// 0.5) / 2^24 in (0, 1): z0, z1 = sqrt(-2 ln u0) (cos, sin)(2 pi u1),
// z2 = sqrt(-2 ln u2) cos(2 pi u3).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mpa {

constexpr uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;
// Second key word of the thermal field (the first is the magnet's seed).
constexpr uint32_t NOISE_KEY1 = 0x6d756d78u;

__host__ __device__ inline void philox4x32_10(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c0, p1 = (uint64_t)PHILOX_M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Three standard normal deviates for cell `cell` of draw `draw` of stream
// `purpose` under `seed`.
__device__ inline void normal3(uint32_t seed, uint32_t cell, uint32_t purpose, uint64_t draw, float z[3]) {
  const uint32_t ctr[4] = {cell, purpose, (uint32_t)draw, (uint32_t)(draw >> 32)};
  const uint32_t key[2] = {seed, NOISE_KEY1};
  uint32_t x[4];
  philox4x32_10(ctr, key, x);
  float u[4];
  for (int i = 0; i < 4; ++i) u[i] = ((float)(x[i] >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float r1 = sqrtf(-2.0f * logf(u[0])), r2 = sqrtf(-2.0f * logf(u[2]));
  float s1, c1, s2, c2;
  sincospif(2.0f * u[1], &s1, &c1);
  sincospif(2.0f * u[3], &s2, &c2);
  (void)s2;
  z[0] = r1 * c1;
  z[1] = r1 * s1;
  z[2] = r2 * c2;
}

}  // namespace mpa
