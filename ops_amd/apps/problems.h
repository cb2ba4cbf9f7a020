// Input definitions of the benchmark problems, as pure functions of grid
// index and seed (so every run and every backend sees the same inputs).
//   Heat (TeaLeaf-form implicit diffusion): conductivity K = exp(N(0,1)) over
//     the halo-extended box (stream 1), solver coefficient w = 1/K, rhs
//     u0 = U[0,1) (stream 2) + 4 exp(-r^2 / (2 s^2)), centre of the block,
//     s = (sum of sides)/30; rx = ry = rz = 1.
//   Poisson: rhs 2 U[0,1) - 1 (stream 3), zero Dirichlet halo.
//   Tridiagonal: per solve direction s, a = U - 1 (stream 10+3s),
//     c = U - 1 (stream 11+3s), b = (|a| + |c|) + (0.1 + 0.9 U) (stream 12+3s);
//     rhs d = 2U - 1 (stream 20).
// U(stream, i) = (bits >> 11) 2^-53 with bits = mix(mix(seed ^ stream G) +
// (i+1) G), mix the SplitMix64 finaliser, G = 0x9e3779b97f4a7c15; normals by
// Box-Muller on counters 2i, 2i+1.  Index i is x + nx (y + ny z) over the
// interior, or over the (n+2)-sided extended box for K.
#pragma once

#include <math.h>
#include <stdint.h>

// the same definitions generate inputs on the host (apps) and on the device
// (the batched solve engine)
#if defined(__HIPCC__)
#define OPSAPP_HD __host__ __device__
#else
#define OPSAPP_HD
#endif

namespace opsapp {

OPSAPP_HD static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
OPSAPP_HD static inline double u01(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t G = 0x9e3779b97f4a7c15ULL;
  return (double)(mix64(mix64(seed ^ (stream * G)) + (i + 1) * G) >> 11) * (1.0 / 9007199254740992.0);
}
OPSAPP_HD static inline double normal(uint64_t seed, uint64_t stream, uint64_t i) {
  const double two_pi = 6.283185307179586476925286766559;
  double a = u01(seed, stream, 2 * i), b = u01(seed, stream, 2 * i + 1);
  return sqrt(-2.0 * log(1.0 - a)) * cos(two_pi * b);
}
OPSAPP_HD static inline uint64_t lin(int i, int j, int k, int nx, int ny) {
  return (uint64_t)i + (uint64_t)nx * ((uint64_t)j + (uint64_t)ny * (uint64_t)k);
}
OPSAPP_HD static inline double heat_K(uint64_t seed, int i, int j, int k, int nx, int ny) {
  return exp(normal(seed, 1, lin(i + 1, j + 1, k + 1, nx + 2, ny + 2)));
}
OPSAPP_HD static inline double heat_rhs(uint64_t seed, int i, int j, int k, int nx, int ny, int nz, int ndim) {
  double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1), cz = 0.5 * (nz - 1);
  double sig = (ndim == 3 ? (double)(nx + ny + nz) : (double)(nx + ny)) / 30.0;
  double dx = i - cx, dy = j - cy, dz = ndim == 3 ? (k - cz) : 0.0;
  double r2 = dx * dx + dy * dy + dz * dz;
  return u01(seed, 2, lin(i, j, k, nx, ny)) + 4.0 * exp(-r2 / (2.0 * sig * sig));
}
OPSAPP_HD static inline double poisson_rhs(uint64_t seed, int i, int j, int k, int nx, int ny) {
  return 2.0 * u01(seed, 3, lin(i, j, k, nx, ny)) - 1.0;
}
OPSAPP_HD static inline void trid_coefs(uint64_t seed, int s, int i, int j, int k, int nx, int ny, double *a, double *b,
                              double *c) {
  uint64_t l = lin(i, j, k, nx, ny);
  double av = u01(seed, 10 + 3 * s, l) - 1.0, cv = u01(seed, 11 + 3 * s, l) - 1.0;
  *a = av;
  *c = cv;
  *b = (fabs(av) + fabs(cv)) + (0.1 + 0.9 * u01(seed, 12 + 3 * s, l));
}
OPSAPP_HD static inline double trid_rhs(uint64_t seed, int i, int j, int k, int nx, int ny) {
  return 2.0 * u01(seed, 20, lin(i, j, k, nx, ny)) - 1.0;
}

}  // namespace opsapp
