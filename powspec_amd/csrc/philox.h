// philox.h -- Philox4x32-10 counter-based random numbers (Salmon, Moraes,
// Dror & Shaw 2011, "Parallel random numbers: as easy as 1, 2, 3"), written
// from the paper for host and device, so that CPU and GPU generators draw
// identical streams.
//
// Stream layout used by the synthetic catalogues: counter = (index_lo,
// index_hi, stream, draw), key = (seed_lo, seed_hi); every counter yields
// four 32-bit words.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define PSH_HD __host__ __device__
#else
#define PSH_HD
#endif

namespace pshrng {

struct W4 {
  uint32_t v[4];
};

PSH_HD inline W4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                               uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;   // multipliers of the paper
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;   // Weyl key increments (golden ratio, sqrt(3) - 1)
    k1 += 0xBB67AE85u;
  }
  W4 w;
  w.v[0] = c0;
  w.v[1] = c1;
  w.v[2] = c2;
  w.v[3] = c3;
  return w;
}

PSH_HD inline W4 draw(uint64_t seed, uint64_t index, uint32_t stream, uint32_t d) {
  return philox4x32_10((uint32_t)index, (uint32_t)(index >> 32), stream, d, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}

// 53-bit uniform double in [0, 1) from two words (27 + 26 bits).
PSH_HD inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

enum Stream : uint32_t {
  kUniformPos = 1,   // uniform box positions (draws 0, 1)
  kNoise = 3,        // white noise of the Gaussian field, one normal per cell
  kPoisson = 4,      // Poisson count per cell
  kPlace = 5,        // position inside the cell, index (cell << 16 | slot) (draws 0, 1)
  kSky = 6,          // survey angles and redshift per object (draws 0, 1)
};

}  // namespace pshrng
