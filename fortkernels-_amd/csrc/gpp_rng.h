// gpp_rng.h — counter-based random numbers shared by the host generators
// (gpp_inputs.cpp, ff_gen.hip) and the device generators (gpp_gen.hip,
// ff_gen.hip).  Every draw is
// a pure function of (seed, stream, index): SplitMix64 finaliser over a
// keyed counter; integer arithmetic only, so host and device agree bit for
// bit on every uniform draw.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define GPP_HD __host__ __device__
#else
#define GPP_HD
#endif

namespace gpp {
namespace rng {

enum Stream : uint64_t {
  kAqsm = 1,
  kAqsn = 2,
  kIeps = 3,
  kWtilde = 4,
  kEnergy = 5,
  kResPick = 6,
  kResWhere = 7,
  // full-frequency (FF) inputs
  kFFAqsm = 11,
  kFFAqsn = 12,
  kFFEpsR = 13,
  kFFEpsA = 14,
  kFFEkq = 15,
  kFFGrid = 16,
  kFFBrd = 17,
  kFFSchDt = 18,
};

GPP_HD inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

GPP_HD inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t key = mix64(seed * 0xD1B54A32D192ED03ull + stream);
  return mix64(key ^ mix64(i));
}

// [0,1) with 53 random bits.
GPP_HD inline double u01(uint64_t x) {
  return (double)(x >> 11) * 0x1.0p-53;
}
// (0,1] with 53 random bits (Box-Muller radius argument).
GPP_HD inline double u01_open0(uint64_t x) {
  return (double)((x >> 11) + 1) * 0x1.0p-53;
}

}  // namespace rng
}  // namespace gpp
