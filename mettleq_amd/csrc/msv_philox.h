// Counter-based Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): ten rounds of two
// 32x32->64 multiplications with the published multipliers and Weyl key
// increments. Shot k of a sampling call uses counter (k, stream_id) and key
// seed, so every shot is independent of batch order and thread count.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace msv {

struct Philox4 {
  uint32_t v[4];
};

__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                             uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(M0) * c0;
    const uint64_t p1 = static_cast<uint64_t>(M1) * c2;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  return Philox4{{c0, c1, c2, c3}};
}

// Uniform double in [0, 1) with 53 random bits from the first two words.
__host__ __device__ inline double philox_u01(uint64_t seed, uint64_t stream_id, uint64_t counter) {
  const Philox4 r = philox4x32_10(static_cast<uint32_t>(counter), static_cast<uint32_t>(counter >> 32),
                                  static_cast<uint32_t>(stream_id),
                                  static_cast<uint32_t>(stream_id >> 32),
                                  static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
  const uint64_t a = r.v[0] >> 5;  // 27 bits
  const uint64_t b = r.v[1] >> 6;  // 26 bits
  return (static_cast<double>(a) * 67108864.0 + static_cast<double>(b)) *
         (1.0 / 9007199254740992.0);
}

}  // namespace msv
