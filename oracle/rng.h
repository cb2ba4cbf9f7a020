/* Counter-based input generator used by the oracle and by the reference-side
 * fixture drivers (test infrastructure, not product code).
 *
 * The reference fills dats with std::mt19937 + libstdc++ distributions over
 * the padded allocation (ops/c/src/core/ops_lib_core.cpp:2394-2494), which is
 * not reproducible across backends.  Every input of this project is instead a
 * pure function of (seed, stream, index):
 *
 *   bits(seed, stream, i) = mix(mix(seed ^ stream*G) + (i+1)*G)
 *   u01                   = (bits >> 11) * 2^-53                    in [0,1)
 *   normal (Box-Muller)   = sqrt(-2 ln(1-u01(2i))) * cos(2 pi u01(2i+1))
 *
 * where mix is the SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014) and
 * G = 0x9e3779b97f4a7c15.  The uniform part is integer arithmetic plus one
 * exact scaling, so any implementation reproduces it bit for bit.  The normal
 * part goes through libm log/cos, so two libms may differ in the last ulp.
 */
#ifndef OPS_ORACLE_RNG_H
#define OPS_ORACLE_RNG_H

#include <stdint.h>
#include <math.h>

#define ORC_RNG_GOLDEN 0x9e3779b97f4a7c15ULL

static inline uint64_t orc_rng_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static inline uint64_t orc_rng_bits(uint64_t seed, uint64_t stream, uint64_t i) {
  return orc_rng_mix(orc_rng_mix(seed ^ (stream * ORC_RNG_GOLDEN)) + (i + 1) * ORC_RNG_GOLDEN);
}

static inline double orc_rng_u01(uint64_t seed, uint64_t stream, uint64_t i) {
  return (double)(orc_rng_bits(seed, stream, i) >> 11) * (1.0 / 9007199254740992.0);
}

static inline double orc_rng_normal(uint64_t seed, uint64_t stream, uint64_t i) {
  const double two_pi = 6.283185307179586476925286766559;
  double u1 = orc_rng_u01(seed, stream, 2 * i);
  double u2 = orc_rng_u01(seed, stream, 2 * i + 1);
  return sqrt(-2.0 * log(1.0 - u1)) * cos(two_pi * u2);
}

#endif
