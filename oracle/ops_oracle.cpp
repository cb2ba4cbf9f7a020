/* ORACLE — test infrastructure only (see ops_oracle.h).  Never linked into
 * or called by the product library. */
#include <chrono>
#include "ops_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "problems.h"

namespace {

/* ext-array geometry: 1-cell halo in each of ndim dimensions */
struct Geo {
  int ndim, nx, ny, nz;
  long sx, sy;
  Geo(int d, int x, int y, int z) : ndim(d), nx(x), ny(y), nz(d == 3 ? z : 1), sx(x + 2), sy(y + 2) {}
  long at(int i, int j, int k) const {
    return (i + 1) + sx * ((j + 1) + (ndim == 3 ? sy * (long)(k + 1) : 0L));
  }
  long next() const { return sx * sy * (ndim == 3 ? (long)(nz + 2) : 1L); }
  long dz() const { return sx * sy; }
  long nint() const { return (long)nx * ny * nz; }
  long lin(int i, int j, int k) const { return i + (long)nx * (j + (long)ny * k); }
};

/* loop over a box [i0,i1)x[j0,j1)x[k0,k1) in lexicographic order; threaded
 * over (k,j) when nt > 1, with an OpenMP sum reduction */
template <class F>
double sweep_sum(int nt, int i0, int i1, int j0, int j1, int k0, int k1, F f) {
  double s = 0.0;
#pragma omp parallel for collapse(2) num_threads(nt) reduction(+ : s) schedule(static)
  for (int k = k0; k < k1; k++)
    for (int j = j0; j < j1; j++)
      for (int i = i0; i < i1; i++) f(i, j, k, s);
  return s;
}
template <class F>
void sweep(int nt, int i0, int i1, int j0, int j1, int k0, int k1, F f) {
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
  for (int k = k0; k < k1; k++)
    for (int j = j0; j < j1; j++)
      for (int i = i0; i < i1; i++) f(i, j, k);
}

}  // namespace

/* definitions take C linkage from the declarations in ops_oracle.h */

/* ------------------------------------------------------------------ inputs */
void orc_gen_u01(uint64_t seed, uint64_t stream, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) out[i] = orc_rng_u01(seed, stream, (uint64_t)i);
}
void orc_gen_normal(uint64_t seed, uint64_t stream, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) out[i] = orc_rng_normal(seed, stream, (uint64_t)i);
}
void orc_gen_laplace_init(int nx, int ny, int mode, double *u) {
  Geo g(2, nx, ny, 1);
  memset(u, 0, sizeof(double) * g.next());
  if (mode == 0) {
    for (int i = -1; i <= nx; i++) u[g.at(i, ny, 0)] = orc_laplace_top(i, nx);
  } else {
    /* set_zero everywhere, then left then right columns incl. corners
     * (apps/c/laplace2d_tutorial/step7/laplace2d.cpp:63-82) */
    for (int j = -1; j <= ny; j++) u[g.at(-1, j, 0)] = orc_laplace_left(j, ny);
    for (int j = -1; j <= ny; j++) u[g.at(nx, j, 0)] = orc_laplace_right(j, ny);
  }
}
void orc_gen_heat(uint64_t seed, int ndim, int nx, int ny, int nz, double *K, double *rhs) {
  Geo g(ndim, nx, ny, nz);
  int kl = ndim == 3 ? -1 : 0, kh = ndim == 3 ? nz : 0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = kl; k <= kh; k++)
    for (int j = -1; j <= ny; j++)
      for (int i = -1; i <= nx; i++) K[g.at(i, j, k)] = orc_heat_conductivity(seed, i, j, k, nx, ny);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < g.nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) rhs[g.lin(i, j, k)] = orc_heat_rhs(seed, i, j, k, nx, ny, g.nz, ndim);
}
void orc_gen_poisson_rhs(uint64_t seed, int ndim, int nx, int ny, int nz, double *f) {
  Geo g(ndim, nx, ny, nz);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < g.nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) f[g.lin(i, j, k)] = orc_poisson_rhs(seed, i, j, k, nx, ny);
}
void orc_gen_trid(uint64_t seed, int s, int nx, int ny, int nz, double *a, double *b, double *c) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long l = i + (long)nx * (j + (long)ny * k);
        orc_trid_coefs(seed, s, i, j, k, nx, ny, &a[l], &b[l], &c[l]);
      }
}
void orc_gen_trid_rhs(uint64_t seed, int nx, int ny, int nz, double *d) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) d[i + (long)nx * (j + (long)ny * k)] = orc_trid_rhs(seed, i, j, k, nx, ny);
}

/* ----------------------------------------------------------------- Laplace */
template <class T>
static T laplace_run(T *A, T *An, int nx, int ny, int iters, int nt, T *hist) {
  Geo g(2, nx, ny, 1);
  const long sx = g.sx;
  T err = 0;
  for (int it = 0; it < iters; it++) {
    /* OPS_MAX handle starts at -DBL_MAX / -FLT_MAX (ops_lib_core.cpp:840-846) */
    T e = (sizeof(T) == 8) ? (T)-DBL_MAX : (T)-FLT_MAX;
#pragma omp parallel for num_threads(nt) reduction(max : e) schedule(static)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long q = g.at(i, j, 0);
        /* laplace_kernels.h:20-23: 0.25f * (A(1,0)+A(-1,0)+A(0,-1)+A(0,1)) */
        T v = 0.25f * (A[q + 1] + A[q - 1] + A[q - sx] + A[q + sx]);
        An[q] = v;
        e = fmax(e, fabs(An[q] - A[q]));
      }
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long q = g.at(i, j, 0);
        A[q] = An[q];
      }
    err = e;
    if (hist) hist[it] = e;
  }
  return err;
}
double orc_laplace_run(double *A, double *An, int nx, int ny, int iters, int nt, double *hist) {
  return laplace_run<double>(A, An, nx, ny, iters, nt, hist);
}
float orc_laplace_run_f32(float *A, float *An, int nx, int ny, int iters, int nt, float *hist) {
  return laplace_run<float>(A, An, nx, ny, iters, nt, hist);
}

/* ---------------------------------------------------------------- heat CG */
void orc_heat_coefs(int ndim, int nx, int ny, int nz, const double *K, double *Kx, double *Ky, double *Kz) {
  Geo g(ndim, nx, ny, nz);
  long N = g.next();
  std::vector<double> w(N);
  /* tea_leaf_recip_kernel over the ext range: w = 1/K */
  for (long q = 0; q < N; q++) w[q] = 1.0 / K[q];
  memset(Kx, 0, sizeof(double) * N);
  memset(Ky, 0, sizeof(double) * N);
  if (ndim == 3) memset(Kz, 0, sizeof(double) * N);
  const long sx = g.sx, sz = g.dz();
  int kh = ndim == 3 ? nz + 1 : 1;
  /* tea_leaf_common_init_Kx_Ky_kernel over [0,n] in every dimension */
  for (int k = 0; k < kh; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        long q = g.at(i, j, k);
        Kx[q] = (w[q - 1] + w[q]) / (2.0 * w[q - 1] * w[q]);
        Ky[q] = (w[q - sx] + w[q]) / (2.0 * w[q - sx] * w[q]);
        if (ndim == 3) Kz[q] = (w[q - sz] + w[q]) / (2.0 * w[q - sz] * w[q]);
      }
  /* zeroed external faces (tea_leaf_common.cpp:104-127) */
  int kl = ndim == 3 ? -1 : 0, kx = ndim == 3 ? nz : 0;
  for (int k = kl; k <= kx; k++)
    for (int j = -1; j <= ny; j++) {
      Kx[g.at(-1, j, k)] = 0.0;
      Kx[g.at(0, j, k)] = 0.0;
      Kx[g.at(nx, j, k)] = 0.0;
    }
  for (int k = kl; k <= kx; k++)
    for (int i = -1; i <= nx; i++) {
      Ky[g.at(i, -1, k)] = 0.0;
      Ky[g.at(i, 0, k)] = 0.0;
      Ky[g.at(i, ny, k)] = 0.0;
    }
  if (ndim == 3)
    for (int j = -1; j <= ny; j++)
      for (int i = -1; i <= nx; i++) {
        Kz[g.at(i, j, -1)] = 0.0;
        Kz[g.at(i, j, 0)] = 0.0;
        Kz[g.at(i, j, nz)] = 0.0;
      }
}

/* operator of tea_leaf_cg_kernels.h:8-15 (+ z terms appended in 3D), in
 * element type T (double; float for the FP32 inner sweeps of the mixed
 * heat solve) */
template <class T>
static inline T heat_op(const Geo &g, const T *Kx, const T *Ky, const T *Kz, const T *p, long q, T rx, T ry, T rz) {
  const long sx = g.sx, sz = g.dz();
  if (g.ndim == 2)
    return ((T)1.0 + ry * (Ky[q + sx] + Ky[q]) + rx * (Kx[q + 1] + Kx[q])) * p[q] -
           ry * (Ky[q + sx] * p[q + sx] + Ky[q] * p[q - sx]) - rx * (Kx[q + 1] * p[q + 1] + Kx[q] * p[q - 1]);
  return ((T)1.0 + ry * (Ky[q + sx] + Ky[q]) + rx * (Kx[q + 1] + Kx[q]) + rz * (Kz[q + sz] + Kz[q])) * p[q] -
         ry * (Ky[q + sx] * p[q + sx] + Ky[q] * p[q - sx]) - rx * (Kx[q + 1] * p[q + 1] + Kx[q] * p[q - 1]) -
         rz * (Kz[q + sz] * p[q + sz] + Kz[q] * p[q - sz]);
}

double orc_heat_apply(int ndim, int nx, int ny, int nz, const double *Kx, const double *Ky, const double *Kz,
                      const double *p, double rx, double ry, double rz, double *w, int nt) {
  Geo g(ndim, nx, ny, nz);
  return sweep_sum(nt, 0, nx, 0, ny, 0, g.nz, [&](int i, int j, int k, double &s) {
    long q = g.at(i, j, k);
    w[q] = heat_op(g, Kx, Ky, Kz, p, q, rx, ry, rz);
    s = s + w[q] * p[q];
  });
}

/* wall time of the iteration loop of the last orc_heat_cg call (set-up and
 * output copies excluded): the CPU baseline's timer */
static double g_loop_seconds = 0.0;

int orc_heat_cg(int ndim, int nx, int ny, int nz, const double *K, const double *rhs, double rx, double ry,
                double rz, int max_iters, double tol, int nt, double *u_int, double *pw_h, double *rrn_h,
                double *alpha_h, double *beta_h) {
  Geo g(ndim, nx, ny, nz);
  long N = g.next();
  std::vector<double> Kx(N), Ky(N), Kz(ndim == 3 ? N : 1), u(N, 0.0), w(N, 0.0), r(N, 0.0), p(N, 0.0);
  orc_heat_coefs(ndim, nx, ny, nz, K, Kx.data(), Ky.data(), Kz.data());
  const double *kz = ndim == 3 ? Kz.data() : nullptr;
  int NZ = g.nz;
  sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) { u[g.at(i, j, k)] = rhs[g.lin(i, j, k)]; });
  /* tea_leaf_common_init_kernel: w = A u, r = u - w */
  sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
    long q = g.at(i, j, k);
    w[q] = heat_op(g, Kx.data(), Ky.data(), kz, u.data(), q, rx, ry, rz);
    r[q] = u[q] - w[q];
  });
  /* tea_leaf_cg_init: p = r, rro = r.p */
  sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) { p[g.at(i, j, k)] = r[g.at(i, j, k)]; });
  double rro = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
    long q = g.at(i, j, k);
    s = s + r[q] * p[q];
  });
  if (rrn_h) rrn_h[0] = rro;
  const double stop = tol > 0 ? tol * sqrt(rro) : -1.0;
  int n = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (n < max_iters) {
    double pw = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
      long q = g.at(i, j, k);
      w[q] = heat_op(g, Kx.data(), Ky.data(), kz, p.data(), q, rx, ry, rz);
      s = s + w[q] * p[q];
    });
    double alpha = rro / pw;
    sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
      long q = g.at(i, j, k);
      u[q] = u[q] + alpha * p[q];
    });
    double rrn = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
      long q = g.at(i, j, k);
      r[q] = r[q] - alpha * w[q];
      s = s + r[q] * r[q];
    });
    double beta = rrn / rro;
    /* tea_leaf_axpby_kernel(p, r, beta, 1.0): p = beta*p + 1.0*r */
    sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
      long q = g.at(i, j, k);
      p[q] = beta * p[q] + 1.0 * r[q];
    });
    if (pw_h) pw_h[n] = pw;
    if (alpha_h) alpha_h[n] = alpha;
    if (beta_h) beta_h[n] = beta;
    if (rrn_h) rrn_h[n + 1] = rrn;
    rro = rrn;
    n++;
    if (sqrt(rrn) <= stop) break;
  }
  g_loop_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (u_int) sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) { u_int[g.lin(i, j, k)] = u[g.at(i, j, k)]; });
  return n;
}

double orc_last_loop_seconds(void) { return g_loop_seconds; }

/* FP64 iterative refinement around FP32 CG for the 3D heat problem (no
 * reference counterpart: the batched engine's FP32-inner heat mode).  The
 * outer loop starts from TeaLeaf's x = u0 = rhs, computes r = f - A x in
 * FP64 (f = rhs) and stops at |r| <= tol |r_0|.  Each correction solves
 * A e = r by unpreconditioned CG in FP32 storage and arithmetic, with the
 * face coefficients rounded to float, every dot product accumulated in
 * FP64, until |rr| <= max(inner_tol, tol / (2 relres)) |rr_0|; then
 * x += (double) e.  Returns the outer iterations; relres_h[k] = |r_k|/|r_0|. */
int orc_heat_mixed(int nx, int ny, int nz, const double *K, const double *rhs, double tol, int max_outer,
                   int inner_iters, double inner_tol, int nt, double *x_int, double *relres_h, int *inner_total) {
  Geo g(3, nx, ny, nz);
  long N = g.next();
  std::vector<double> Kx(N), Ky(N), Kz(N), x(N, 0.0), f(N, 0.0), r(N, 0.0);
  orc_heat_coefs(3, nx, ny, nz, K, Kx.data(), Ky.data(), Kz.data());
  std::vector<float> kx(N), ky(N), kz(N), e(N, 0.0f), rr(N, 0.0f), z(N, 0.0f), p(N, 0.0f), w(N, 0.0f);
  for (long q = 0; q < N; q++) {
    kx[q] = (float)Kx[q];
    ky[q] = (float)Ky[q];
    kz[q] = (float)Kz[q];
  }
  sweep(nt, 0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
    const long q = g.at(i, j, k);
    x[q] = f[q] = rhs[g.lin(i, j, k)];
  });
  double fn = 0.0;
  int outer = 0, total = 0;
  while (true) {
    const double rrd = sweep_sum(nt, 0, nx, 0, ny, 0, nz, [&](int i, int j, int k, double &s) {
      const long q = g.at(i, j, k);
      const double ax = heat_op<double>(g, Kx.data(), Ky.data(), Kz.data(), x.data(), q, 1.0, 1.0, 1.0);
      r[q] = f[q] - ax;
      s = s + r[q] * r[q];
    });
    if (outer == 0) fn = sqrt(rrd);
    const double rel = sqrt(rrd) / fn;
    if (relres_h) relres_h[outer] = rel;
    if (rel <= tol || outer >= max_outer) break;
    const double rz = sweep_sum(nt, 0, nx, 0, ny, 0, nz, [&](int i, int j, int k, double &s) {
      const long q = g.at(i, j, k);
      e[q] = 0.0f;
      rr[q] = (float)r[q];
      z[q] = rr[q];
      p[q] = z[q];
      s = s + (double)rr[q] * (double)z[q];
    });
    double rzc = rz;
    const double istop = fmax(inner_tol, 0.5 * tol / rel) * sqrt(rz);
    for (int it = 0; it < inner_iters; it++) {
      const double pw = sweep_sum(nt, 0, nx, 0, ny, 0, nz, [&](int i, int j, int k, double &s) {
        const long q = g.at(i, j, k);
        w[q] = heat_op<float>(g, kx.data(), ky.data(), kz.data(), p.data(), q, 1.0f, 1.0f, 1.0f);
        s = s + (double)w[q] * (double)p[q];
      });
      const float a = (float)(rzc / pw);
      double rzn = 0.0, rrn = 0.0;
#pragma omp parallel for collapse(2) num_threads(nt) reduction(+ : rzn, rrn) schedule(static)
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++) {
            const long q = g.at(i, j, k);
            e[q] = e[q] + a * p[q];
            rr[q] = rr[q] - a * w[q];
            z[q] = rr[q];
            rzn = rzn + (double)rr[q] * (double)z[q];
            rrn = rrn + (double)rr[q] * (double)rr[q];
          }
      total++;
      if (sqrt(rrn) <= istop) break;
      const float b = (float)(rzn / rzc);
      sweep(nt, 0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
        const long q = g.at(i, j, k);
        p[q] = z[q] + b * p[q];
      });
      rzc = rzn;
    }
    sweep(nt, 0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
      const long q = g.at(i, j, k);
      x[q] = x[q] + (double)e[q];
    });
    outer++;
  }
  if (inner_total) *inner_total = total;
  if (x_int) sweep(nt, 0, nx, 0, ny, 0, nz, [&](int i, int j, int k) { x_int[g.lin(i, j, k)] = x[g.at(i, j, k)]; });
  return outer;
}

/* ------------------------------------------------------------- Poisson CG */
template <class T>
static inline T poisson_op(const Geo &g, const T *p, long q) {
  const long sx = g.sx, sz = g.dz();
  if (g.ndim == 2) return (T)4.0 * p[q] - (p[q - 1] + p[q + 1]) - (p[q - sx] + p[q + sx]);
  return (T)6.0 * p[q] - (p[q - 1] + p[q + 1]) - (p[q - sx] + p[q + sx]) - (p[q - sz] + p[q + sz]);
}

int orc_poisson_cg(int ndim, int nx, int ny, int nz, const double *f, int max_iters, double tol, int nt,
                   double *x_int, double *rrn_h, double *alpha_h, double *beta_h) {
  Geo g(ndim, nx, ny, nz);
  long N = g.next();
  int NZ = g.nz;
  std::vector<double> x(N, 0.0), r(N, 0.0), p(N, 0.0), w(N, 0.0);
  sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
    long q = g.at(i, j, k);
    r[q] = f[g.lin(i, j, k)];
    p[q] = r[q];
  });
  double rro = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
    long q = g.at(i, j, k);
    s = s + r[q] * p[q];
  });
  if (rrn_h) rrn_h[0] = rro;
  const double stop = tol > 0 ? tol * sqrt(rro) : -1.0;
  int n = 0;
  while (n < max_iters) {
    double pw = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
      long q = g.at(i, j, k);
      w[q] = poisson_op<double>(g, p.data(), q);
      s = s + w[q] * p[q];
    });
    double alpha = rro / pw;
    sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
      long q = g.at(i, j, k);
      x[q] = x[q] + alpha * p[q];
    });
    double rrn = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
      long q = g.at(i, j, k);
      r[q] = r[q] - alpha * w[q];
      s = s + r[q] * r[q];
    });
    double beta = rrn / rro;
    sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
      long q = g.at(i, j, k);
      p[q] = beta * p[q] + 1.0 * r[q];
    });
    if (alpha_h) alpha_h[n] = alpha;
    if (beta_h) beta_h[n] = beta;
    if (rrn_h) rrn_h[n + 1] = rrn;
    rro = rrn;
    n++;
    if (sqrt(rrn) <= stop) break;
  }
  if (x_int) sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) { x_int[g.lin(i, j, k)] = x[g.at(i, j, k)]; });
  return n;
}

double orc_poisson_apply_f32(int ndim, int nx, int ny, int nz, const float *p, float *w) {
  Geo g(ndim, nx, ny, nz);
  return sweep_sum(1, 0, nx, 0, ny, 0, g.nz, [&](int i, int j, int k, double &s) {
    long q = g.at(i, j, k);
    w[q] = poisson_op<float>(g, p, q);
    s = s + (double)w[q] * (double)p[q];
  });
}

double orc_poisson_residual(int ndim, int nx, int ny, int nz, const double *f, const double *x, double *r,
                            int nt) {
  Geo g(ndim, nx, ny, nz);
  return sweep_sum(nt, 0, nx, 0, ny, 0, g.nz, [&](int i, int j, int k, double &s) {
    long q = g.at(i, j, k);
    r[q] = f[g.lin(i, j, k)] - poisson_op<double>(g, x, q);
    s = s + r[q] * r[q];
  });
}

/* Mixed precision.  Outer (FP64):  r = f - A x;  relres = |r|/|f|.
 * Inner (FP32 storage and arithmetic, FP64 dot accumulation), Jacobi PCG on
 * A e = (float)r with Mi = (float)(1/diag):
 *   e = 0; rr = (float)r; z = Mi*rr; p = z; rz = rr.z
 *   loop: w = A p; pw = p.w; a = (float)(rz/pw)
 *         e = e + a*p; rr = rr - a*w; z = Mi*rr; rzn = rr.z; rrn = rr.rr
 *         stop if sqrt(rrn) <= t*sqrt(rr0.rr0) or it == inner_iters, with
 *         t = max(inner_tol, tol/(2 relres)): the last correction only
 *         reduces the residual as far as the outer target needs
 *         b = (float)(rzn/rz); p = z + b*p; rz = rzn
 * Correction (FP64): x = x + (double)e.  Dots take products of floats in
 * FP64 (double(p)*double(w)). */
int orc_poisson_mixed(int ndim, int nx, int ny, int nz, const double *f, double tol, int max_outer,
                      int inner_iters, double inner_tol, int nt, double *x_int, double *relres_h,
                      int *inner_total) {
  Geo g(ndim, nx, ny, nz);
  long N = g.next();
  int NZ = g.nz;
  std::vector<double> x(N, 0.0), r(N, 0.0);
  std::vector<float> e(N, 0.0f), rr(N, 0.0f), z(N, 0.0f), p(N, 0.0f), w(N, 0.0f);
  const float Mi = (float)(1.0 / (ndim == 3 ? 6.0 : 4.0));
  double ff = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
    double v = f[g.lin(i, j, k)];
    s = s + v * v;
  });
  double fn = sqrt(ff);
  int outer = 0, total = 0;
  while (true) {
    double rrd = orc_poisson_residual(ndim, nx, ny, nz, f, x.data(), r.data(), nt);
    double rel = sqrt(rrd) / fn;
    if (relres_h) relres_h[outer] = rel;
    if (rel <= tol || outer >= max_outer) break;
    double rz = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
      long q = g.at(i, j, k);
      e[q] = 0.0f;
      rr[q] = (float)r[q];
      z[q] = Mi * rr[q];
      p[q] = z[q];
      s = s + (double)rr[q] * (double)z[q];
    });
    double rr0 = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
      long q = g.at(i, j, k);
      s = s + (double)rr[q] * (double)rr[q];
    });
    double istop = fmax(inner_tol, 0.5 * tol / rel) * sqrt(rr0);
    for (int it = 0; it < inner_iters; it++) {
      double pw = sweep_sum(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k, double &s) {
        long q = g.at(i, j, k);
        w[q] = poisson_op<float>(g, p.data(), q);
        s = s + (double)p[q] * (double)w[q];
      });
      float a = (float)(rz / pw);
      double rzn = 0.0, rrn = 0.0;
#pragma omp parallel for collapse(2) num_threads(nt) reduction(+ : rzn, rrn) schedule(static)
      for (int k = 0; k < NZ; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++) {
            long q = g.at(i, j, k);
            e[q] = e[q] + a * p[q];
            rr[q] = rr[q] - a * w[q];
            z[q] = Mi * rr[q];
            rzn = rzn + (double)rr[q] * (double)z[q];
            rrn = rrn + (double)rr[q] * (double)rr[q];
          }
      total++;
      if (sqrt(rrn) <= istop) break;
      float b = (float)(rzn / rz);
      sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
        long q = g.at(i, j, k);
        p[q] = z[q] + b * p[q];
      });
      rz = rzn;
    }
    sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) {
      long q = g.at(i, j, k);
      x[q] = x[q] + (double)e[q];
    });
    outer++;
  }
  if (inner_total) *inner_total = total;
  if (x_int) sweep(nt, 0, nx, 0, ny, 0, NZ, [&](int i, int j, int k) { x_int[g.lin(i, j, k)] = x[g.at(i, j, k)]; });
  return outer;
}

/* --------------------------------------------------------- ADI / 27-point */
void orc_adi_preproc(int nx, int ny, int nz, double lambda, const double *u, double *du, double *ax, double *bx,
                     double *cx) {
  Geo g(3, nx, ny, nz);
  const long sx = g.sx, sz = g.dz();
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long q = g.at(i, j, k), l = g.lin(i, j, k);
        double a, b, c, d;
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1) {
          d = 0.0f; a = 0.0f; b = 1.0f; c = 0.0f;
        } else {
          d = lambda * (u[q - 1] + u[q + 1] + u[q - sx] + u[q + sx] + u[q - sz] + u[q + sz] - 6.0f * u[q]);
          a = -0.5f * lambda;
          b = 1.0f + lambda;
          c = -0.5f * lambda;
        }
        du[l] = d; ax[l] = a; bx[l] = b; cx[l] = c;
      }
}

void orc_stencil27(int nx, int ny, int nz, const double *u, double *out) {
  static const double W[4] = {0.125, 0.0625, 0.03125, 0.015625};
  Geo g(3, nx, ny, nz);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double acc = 0.0;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++)
              acc = acc + W[abs(dx) + abs(dy) + abs(dz)] * u[g.at(i + dx, j + dy, k + dz)];
        out[g.lin(i, j, k)] = acc;
      }
}

/* ------------------------------------------------------------ tridiagonal */
namespace {
struct Lines {
  long n, stride, count;
  long other_stride[2];
  int other_dim[2];
  int other_ext[2];
};
static Lines make_lines(int ndim, int solvedim, const int *dims, const int *pads) {
  Lines L;
  long st[3] = {1, pads[0], (long)pads[0] * (ndim > 1 ? pads[1] : 1)};
  L.n = dims[solvedim];
  L.stride = st[solvedim];
  int q = 0;
  L.count = 1;
  for (int d = 0; d < 3; d++) {
    if (d == solvedim) continue;
    L.other_dim[q] = d;
    L.other_ext[q] = d < ndim ? dims[d] : 1;
    L.other_stride[q] = st[d];
    L.count *= L.other_ext[q];
    q++;
  }
  return L;
}
static inline long line_base(const Lines &L, long id) {
  long i0 = id % L.other_ext[0], i1 = id / L.other_ext[0];
  return i0 * L.other_stride[0] + i1 * L.other_stride[1];
}
template <class T>
static void thomas_line(const T *a, const T *b, const T *c, T *d, T *u, long n, long s, T *c2, T *d2) {
  T bb = (T)1.0 / b[0];
  c2[0] = bb * c[0];
  d2[0] = bb * d[0];
  for (long i = 1; i < n; i++) {
    bb = (T)1.0 / (b[i * s] - a[i * s] * c2[i - 1]);
    d2[i] = bb * (d[i * s] - a[i * s] * d2[i - 1]);
    c2[i] = bb * c[i * s];
  }
  /* back substitution in place of d2 */
  for (long i = n - 2; i >= 0; i--) d2[i] = d2[i] - c2[i] * d2[i + 1];
  if (u) {
    for (long i = 0; i < n; i++) u[i * s] = u[i * s] + d2[i];
  } else {
    for (long i = 0; i < n; i++) d[i * s] = d2[i];
  }
}
template <class T>
static void thomas_batch(int ndim, int solvedim, const int *dims, const int *pads, const T *a, const T *b,
                         const T *c, T *d, T *u, int nt) {
  Lines L = make_lines(ndim, solvedim, dims, pads);
#pragma omp parallel num_threads(nt)
  {
    std::vector<T> c2(L.n), d2(L.n);
#pragma omp for schedule(static)
    for (long id = 0; id < L.count; id++) {
      long o = line_base(L, id);
      thomas_line<T>(a + o, b + o, c + o, d + o, u ? u + o : nullptr, L.n, L.stride, c2.data(), d2.data());
    }
  }
}
}  // namespace

void orc_thomas_batch(int ndim, int solvedim, const int *dims, const int *pads, const double *a, const double *b,
                      const double *c, double *d, double *u, int nt) {
  thomas_batch<double>(ndim, solvedim, dims, pads, a, b, c, d, u, nt);
}
void orc_thomas_batch_f32(int ndim, int solvedim, const int *dims, const int *pads, const float *a,
                          const float *b, const float *c, float *d, float *u, int nt) {
  thomas_batch<float>(ndim, solvedim, dims, pads, a, b, c, d, u, nt);
}

/* PCR step with stride s over arrays of length n; out-of-range neighbours
 * contribute nothing, and a[0], c[n-1] are taken as 0. */
void orc_pcr_batch(int ndim, int solvedim, const int *dims, const int *pads, const double *a, const double *b,
                   const double *c, double *d, double *u, int nt) {
  Lines L = make_lines(ndim, solvedim, dims, pads);
#pragma omp parallel num_threads(nt)
  {
    long n = L.n;
    std::vector<double> A(n), B(n), C(n), D(n), A2(n), B2(n), C2(n), D2(n);
#pragma omp for schedule(static)
    for (long id = 0; id < L.count; id++) {
      long o = line_base(L, id);
      for (long i = 0; i < n; i++) {
        A[i] = i == 0 ? 0.0 : a[o + i * L.stride];
        B[i] = b[o + i * L.stride];
        C[i] = i == n - 1 ? 0.0 : c[o + i * L.stride];
        D[i] = d[o + i * L.stride];
      }
      for (long s = 1; s < n; s *= 2) {
        for (long i = 0; i < n; i++) {
          bool lo = i - s >= 0, hi = i + s < n;
          double k1 = lo ? A[i] / B[i - s] : 0.0;
          double k2 = hi ? C[i] / B[i + s] : 0.0;
          A2[i] = lo ? -A[i - s] * k1 : 0.0;
          C2[i] = hi ? -C[i + s] * k2 : 0.0;
          double bv = B[i], dv = D[i];
          if (lo) { bv = bv - C[i - s] * k1; dv = dv - D[i - s] * k1; }
          if (hi) { bv = bv - A[i + s] * k2; dv = dv - D[i + s] * k2; }
          B2[i] = bv;
          D2[i] = dv;
        }
        A.swap(A2); B.swap(B2); C.swap(C2); D.swap(D2);
      }
      for (long i = 0; i < n; i++) {
        double x = D[i] / B[i];
        if (u) u[o + i * L.stride] = u[o + i * L.stride] + x;
        else d[o + i * L.stride] = x;
      }
    }
  }
}


/* --------------------------------------------- multi-block Poisson test app
 * apps/c/poisson/poisson.cpp:191-288 with one block; kernels of
 * apps/c/poisson/poisson_kernel.h:4-38 in their operation order. */
double orc_mb_poisson(int nx, int ny, int iters, double *u_int) {
  const double dx = 0.01, dy = 0.01, pi = 3.14159265358979323846;
  const long W = nx + 2, H = ny + 2;
  std::vector<double> u(W * H), u2(W * H), f(W * H), ref(W * H);
  auto at = [&](int i, int j) { return (long)(i + 1) + W * (long)(j + 1); };
  for (int j = -1; j <= ny; j++)
    for (int i = -1; i <= nx; i++) {
      const double x = dx * (double)(i + 0), y = dy * (double)(j + 0);
      const long q = at(i, j);
      u[q] = (sin(pi * x) * cos(2.0 * pi * y) + 1.0) - 1.0;
      f[q] = -5.0 * pi * pi * sin(pi * x) * cos(2.0 * pi * y);
      ref[q] = sin(pi * x) * cos(2.0 * pi * y);
      u2[q] = u[q];
    }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) u[at(i, j)] = 0.0;
  for (int it = 0; it < iters; it++) {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        u2[at(i, j)] = (((u[at(i - 1, j)] + u[at(i + 1, j)]) * dx * dx) + ((u[at(i, j - 1)] + u[at(i, j + 1)]) * dy * dy) -
                        (f[at(i, j)] * dx * dx * dy * dy)) /
                       (2.0 * (dx * dx + dy * dy));
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) u[at(i, j)] = u2[at(i, j)];
  }
  double err = 0.0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const double d = u[at(i, j)] - ref[at(i, j)];
      err = err + d * d;
      if (u_int) u_int[(long)i + (long)nx * j] = u[at(i, j)];
    }
  return err;
}
