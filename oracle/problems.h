/* Test-problem definitions shared by the oracle and the reference-side
 * fixture drivers (test infrastructure, not product code).  Each input value
 * is a pure function of its grid index and a seed, so any implementation can
 * regenerate it.  Index conventions follow OPS: x fastest, interior indices
 * 0..n-1, halo at -1 and n (ops/c/include/ops_lib_core.h:1356-2038).
 *
 * Problems
 *  - Laplace (2D Jacobi):  top halo row u(i,ny) = sin(pi (i+1)/(nx+1)),
 *    i in [-1,nx]; everything else 0.  The reference tutorial app instead
 *    sets left/right halo columns (apps/c/laplace2d_tutorial/step7/
 *    laplace2d.cpp:60-105, laplace_kernels.h:12-18); both are provided.
 *  - Heat (TeaLeaf-form implicit diffusion, 2D/3D): conductivity
 *    K = exp(N(0,1)) on the full range incl. halo (stream 1); the solver
 *    coefficient w = 1/K (TeaLeaf RECIP_CONDUCTIVITY, apps/c/TeaLeaf/
 *    tea_leaf_common.cpp:86-90); rhs u0 = U[0,1) (stream 2) + a Gaussian hot
 *    spot of amplitude 4 centred in the block, sigma = (sum of sides)/30.
 *    rx = ry = rz = 1.
 *  - Poisson (constant-coefficient 5/7-point, zero Dirichlet halo):
 *    rhs f = 2 U[0,1) - 1 (stream 3).
 *  - Tridiagonal (ADI sweeps): for solve direction s in {0,1,2}:
 *    a = U - 1 (stream 10+3s), c = U - 1 (stream 11+3s),
 *    b = (|a| + |c|) + (0.1 + 0.9 U) (stream 12+3s); rhs d = 2U - 1 (stream 20).
 */
#ifndef OPS_ORACLE_PROBLEMS_H
#define OPS_ORACLE_PROBLEMS_H

#include "rng.h"

#define ORC_PI 3.14159265358979323846

/* linear index of interior point (i,j,k) with sides n[] */
static inline uint64_t orc_lin3(int i, int j, int k, int nx, int ny) {
  return (uint64_t)i + (uint64_t)nx * ((uint64_t)j + (uint64_t)ny * (uint64_t)k);
}
/* linear index over the halo-extended range [-1,n] in every dimension */
static inline uint64_t orc_lin3_ext(int i, int j, int k, int nx, int ny) {
  return orc_lin3(i + 1, j + 1, k + 1, nx + 2, ny + 2);
}

static inline double orc_laplace_top(int i, int nx) {
  return sin(ORC_PI * (i + 1) / (nx + 1));
}
static inline double orc_laplace_left(int j, int ny) {
  return sin(ORC_PI * (j + 1) / (ny + 1));
}
static inline double orc_laplace_right(int j, int ny) {
  return sin(ORC_PI * (j + 1) / (ny + 1)) * exp(-ORC_PI);
}

/* heat: conductivity over the halo-extended range (2D: pass k = 0, nz = 1) */
static inline double orc_heat_conductivity(uint64_t seed, int i, int j, int k, int nx, int ny) {
  return exp(orc_rng_normal(seed, 1, orc_lin3_ext(i, j, k, nx, ny)));
}
static inline double orc_heat_rhs(uint64_t seed, int i, int j, int k, int nx, int ny, int nz, int ndim) {
  double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1), cz = 0.5 * (nz - 1);
  double sig = (ndim == 3 ? (double)(nx + ny + nz) : (double)(nx + ny)) / 30.0;
  double dx = i - cx, dy = j - cy, dz = (ndim == 3) ? (k - cz) : 0.0;
  double r2 = dx * dx + dy * dy + dz * dz;
  return orc_rng_u01(seed, 2, orc_lin3(i, j, k, nx, ny)) + 4.0 * exp(-r2 / (2.0 * sig * sig));
}

static inline double orc_poisson_rhs(uint64_t seed, int i, int j, int k, int nx, int ny) {
  return 2.0 * orc_rng_u01(seed, 3, orc_lin3(i, j, k, nx, ny)) - 1.0;
}

/* tridiagonal coefficients for solve direction s at interior point */
static inline void orc_trid_coefs(uint64_t seed, int s, int i, int j, int k, int nx, int ny,
                                  double *a, double *b, double *c) {
  uint64_t l = orc_lin3(i, j, k, nx, ny);
  double av = orc_rng_u01(seed, 10 + 3 * s, l) - 1.0;
  double cv = orc_rng_u01(seed, 11 + 3 * s, l) - 1.0;
  *a = av;
  *c = cv;
  *b = (fabs(av) + fabs(cv)) + (0.1 + 0.9 * orc_rng_u01(seed, 12 + 3 * s, l));
}
static inline double orc_trid_rhs(uint64_t seed, int i, int j, int k, int nx, int ny) {
  return 2.0 * orc_rng_u01(seed, 20, orc_lin3(i, j, k, nx, ny)) - 1.0;
}

#endif
