// Application drivers written only against the public C ABI
// (include/ops_mi355x.h), the way the reference apps drive OPS:
//   Laplace 2D Jacobi   apps/c/laplace2d_tutorial/step7/laplace2d.cpp:37-139
//   TeaLeaf-form CG     apps/c/TeaLeaf/tea_leaf_common.cpp:76-146,
//                       tea_leaf_cg.cpp:54-214, tea_solve.cpp:300-325 (2D; 3D
//                       generalisation with the z pair appended)
//   Poisson CG          constant-coefficient 5/7-point operator
//   ADI tridiagonal     apps/c/adi/adi.cpp:344-424 (random coefficients)
//   mixed-precision Poisson and the request stream on the batched engine
//   (include/ops_mi355x_serve.h) have no reference app
// Exported through include/ops_mi355x_apps.h for the benchmark and tests.
#include "../../include/ops_mi355x_apps.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/ops_mi355x.h"
#include "../../include/ops_mi355x_serve.h"
#include "problems.h"

using namespace opsapp;

namespace {

int S000[] = {0, 0, 0};
int S7[] = {0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1};
int S5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
int S00[] = {0, 0};

struct Geo {
  int ndim, nx, ny, nz;
  int full[6], inner[6], ext[6];
  Geo(int d, int x, int y, int z) : ndim(d), nx(x), ny(y), nz(d == 3 ? z : 1) {
    int n[3] = {nx, ny, nz};
    for (int q = 0; q < 3; q++) {
      full[2 * q] = -1;
      full[2 * q + 1] = n[q] + 1;
      inner[2 * q] = 0;
      inner[2 * q + 1] = n[q];
      ext[2 * q] = 0;
      ext[2 * q + 1] = n[q] + 1;
    }
  }
  size_t n_ext() const { return (size_t)(nx + 2) * (ny + 2) * (ndim == 3 ? nz + 2 : 1); }
  size_t n_int() const { return (size_t)nx * ny * nz; }
};

ops_dat decl(ops_block b, const Geo &g, const char *name, const char *type = "double", int ts = 8) {
  int size[3] = {g.nx, g.ny, g.nz}, base[3] = {0, 0, 0}, dm[3] = {-1, -1, -1}, dp[3] = {1, 1, 1};
  int stride[3] = {1, 1, 1};
  return ops_decl_dat_char(b, 1, size, base, dm, dp, stride, nullptr, ts, type, name);
}

ops_arg gbl(double *v) { return ops_arg_gbl_char((char *)v, 1, sizeof(double), OPS_READ); }

double wtime() {
  double cpu, et;
  ops_timers(&cpu, &et);
  return et;
}

// ---------------------------------------------------------------- heat CG
struct Heat {
  Geo g;
  double rx, ry, rz;
  ops_block blk;
  ops_dat density, u, u0, w, Kx, Ky, Kz, r, p;
  ops_stencil s0, sp, sx, sy, sz, sm;
  ops_reduction red;
  double rro = 0.0, rro0 = 0.0;
  double one = 1.0;
  Heat(int d, int x, int y, int z) : g(d, x, y, z) {}
};

}  // namespace

extern "C" {

void *opsapp_heat_create(int ndim, int nx, int ny, int nz, uint64_t seed, double rx, double ry, double rz,
                         const double *K_ext, const double *rhs) {
  Heat *h = new Heat(ndim, nx, ny, nz);
  Geo &g = h->g;
  h->rx = rx;
  h->ry = ry;
  h->rz = rz;
  h->blk = ops_decl_block(ndim, "tea_grid");
  h->density = decl(h->blk, g, "density");
  h->u = decl(h->blk, g, "u");
  h->u0 = decl(h->blk, g, "u0");
  h->w = decl(h->blk, g, "w");
  h->Kx = decl(h->blk, g, "Kx");
  h->Ky = decl(h->blk, g, "Ky");
  h->Kz = ndim == 3 ? decl(h->blk, g, "Kz") : nullptr;
  h->r = decl(h->blk, g, "r");
  h->p = decl(h->blk, g, "p");
  if (ndim == 3) {
    int sx[] = {0, 0, 0, 1, 0, 0}, sy[] = {0, 0, 0, 0, 1, 0}, sz[] = {0, 0, 0, 0, 0, 1};
    int sm[] = {0, 0, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};
    h->s0 = ops_decl_stencil(3, 1, S000, "S3D_000");
    h->sp = ops_decl_stencil(3, 7, S7, "S3D_7PT");
    h->sx = ops_decl_stencil(3, 2, sx, "S3D_000_P100");
    h->sy = ops_decl_stencil(3, 2, sy, "S3D_000_0P10");
    h->sz = ops_decl_stencil(3, 2, sz, "S3D_000_00P1");
    h->sm = ops_decl_stencil(3, 4, sm, "S3D_000_M100_0M10_00M1");
  } else {
    int sx[] = {0, 0, 1, 0}, sy[] = {0, 0, 0, 1}, sm[] = {0, 0, -1, 0, 0, -1};
    h->s0 = ops_decl_stencil(2, 1, S00, "S2D_00");
    h->sp = ops_decl_stencil(2, 5, S5, "S2D_00_0M1_M10_P10_0P1");
    h->sx = ops_decl_stencil(2, 2, sx, "S2D_00_P10");
    h->sy = ops_decl_stencil(2, 2, sy, "S2D_00_0P1");
    h->sz = nullptr;
    h->sm = ops_decl_stencil(2, 3, sm, "S2D_00_M10_0M1");
  }
  h->red = ops_decl_reduction_handle(sizeof(double), "double", "red_temp");

  // inputs: given, or generated from the problem definition
  std::vector<double> Kv, fv;
  if (!K_ext || !rhs) {
    Kv.resize(g.n_ext());
    fv.resize(g.n_int());
    int kl = ndim == 3 ? -1 : 0, kh = ndim == 3 ? nz : 0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = kl; k <= kh; k++)
      for (int j = -1; j <= ny; j++)
        for (int i = -1; i <= nx; i++)
          Kv[(size_t)(i + 1) + (size_t)(nx + 2) * ((j + 1) + (size_t)(ny + 2) * (k - kl))] =
              heat_K(seed, i, j, k, nx, ny);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) fv[lin(i, j, k, nx, ny)] = heat_rhs(seed, i, j, k, nx, ny, g.nz, ndim);
    K_ext = Kv.data();
    rhs = fv.data();
  }
  ops_dat_set_data_slab_memspace(h->density, 0, (char *)K_ext, g.full, OPS_HOST);
  ops_dat_set_data(h->u0, 0, (char *)rhs);
  Kv.clear();
  Kv.shrink_to_fit();
  fv.clear();
  fv.shrink_to_fit();

  ops_block b = h->blk;
  int d = ndim;
  ops_par_loop_tea_leaf_yeqx_kernel("tea_leaf_yeqx_kernel", b, d, g.inner, ops_arg_dat(h->u, 1, h->s0, "double", OPS_WRITE),
                                    ops_arg_dat(h->u0, 1, h->s0, "double", OPS_READ));
  // tea_leaf_common_init, RECIP_CONDUCTIVITY, halo depth 1
  ops_par_loop_tea_leaf_recip_kernel("tea_leaf_recip_kernel", b, d, g.full, ops_arg_dat(h->w, 1, h->s0, "double", OPS_WRITE),
                                     ops_arg_dat(h->density, 1, h->s0, "double", OPS_READ));
  if (ndim == 3)
    ops_par_loop_tea_leaf_common_init_Kx_Ky_Kz_kernel(
        "tea_leaf_common_init_Kx_Ky_Kz_kernel", b, d, g.ext, ops_arg_dat(h->Kx, 1, h->s0, "double", OPS_WRITE),
        ops_arg_dat(h->Ky, 1, h->s0, "double", OPS_WRITE), ops_arg_dat(h->Kz, 1, h->s0, "double", OPS_WRITE),
        ops_arg_dat(h->w, 1, h->sm, "double", OPS_READ));
  else
    ops_par_loop_tea_leaf_common_init_Kx_Ky_kernel(
        "tea_leaf_common_init_Kx_Ky_kernel", b, d, g.ext, ops_arg_dat(h->Kx, 1, h->s0, "double", OPS_WRITE),
        ops_arg_dat(h->Ky, 1, h->s0, "double", OPS_WRITE), ops_arg_dat(h->w, 1, h->sm, "double", OPS_READ));
  // zeroed external faces
  ops_dat K3[3] = {h->Kx, h->Ky, h->Kz};
  for (int q = 0; q < ndim; q++) {
    int lo[6], hi[6];
    memcpy(lo, g.full, sizeof(lo));
    memcpy(hi, g.full, sizeof(hi));
    lo[2 * q] = -1;
    lo[2 * q + 1] = 1;
    int n = q == 0 ? nx : (q == 1 ? ny : nz);
    hi[2 * q] = n;
    hi[2 * q + 1] = n + 1;
    ops_par_loop_tea_leaf_init_zero_kernel("tea_leaf_init_zero_kernel", b, d, lo, ops_arg_dat(K3[q], 1, h->s0, "double", OPS_WRITE));
    ops_par_loop_tea_leaf_init_zero_kernel("tea_leaf_init_zero_kernel", b, d, hi, ops_arg_dat(K3[q], 1, h->s0, "double", OPS_WRITE));
  }
  if (ndim == 3)
    ops_par_loop_tea_leaf_common_init_kernel_3d(
        "tea_leaf_common_init_kernel", b, d, g.inner, ops_arg_dat(h->w, 1, h->s0, "double", OPS_WRITE),
        ops_arg_dat(h->r, 1, h->s0, "double", OPS_WRITE), ops_arg_dat(h->Kx, 1, h->sx, "double", OPS_READ),
        ops_arg_dat(h->Ky, 1, h->sy, "double", OPS_READ), ops_arg_dat(h->Kz, 1, h->sz, "double", OPS_READ),
        ops_arg_dat(h->u, 1, h->sp, "double", OPS_READ), gbl(&h->rx), gbl(&h->ry), gbl(&h->rz));
  else
    ops_par_loop_tea_leaf_common_init_kernel(
        "tea_leaf_common_init_kernel", b, d, g.inner, ops_arg_dat(h->w, 1, h->s0, "double", OPS_WRITE),
        ops_arg_dat(h->r, 1, h->s0, "double", OPS_WRITE), ops_arg_dat(h->Kx, 1, h->sx, "double", OPS_READ),
        ops_arg_dat(h->Ky, 1, h->sy, "double", OPS_READ), ops_arg_dat(h->u, 1, h->sp, "double", OPS_READ),
        gbl(&h->rx), gbl(&h->ry));
  // tea_leaf_cg_init without preconditioner
  ops_par_loop_tea_leaf_yeqx_kernel("tea_leaf_yeqx_kernel", b, d, g.inner, ops_arg_dat(h->p, 1, h->s0, "double", OPS_WRITE),
                                    ops_arg_dat(h->r, 1, h->s0, "double", OPS_READ));
  ops_par_loop_tea_leaf_dot_kernel("tea_leaf_dot_kernel", b, d, g.inner, ops_arg_dat(h->r, 1, h->s0, "double", OPS_READ),
                                   ops_arg_dat(h->p, 1, h->s0, "double", OPS_READ),
                                   ops_arg_reduce(h->red, 1, "double", OPS_INC));
  ops_reduction_result_char(h->red, sizeof(double), (char *)&h->rro);
  h->rro0 = h->rro;
  return h;
}

int opsapp_heat_iterate(void *ctx, int max_iters, double tol, double *pw_h, double *rrn_h, double *alpha_h,
                        double *beta_h) {
  Heat *h = (Heat *)ctx;
  Geo &g = h->g;
  ops_block b = h->blk;
  int d = g.ndim;
  const double stop = tol > 0 ? tol * sqrt(h->rro0) : -1.0;
  int n = 0;
  while (n < max_iters) {
    double pw, rrn;
    if (d == 3)
      ops_par_loop_tea_leaf_cg_calc_w_reduce_kernel_3d(
          "tea_leaf_cg_calc_w_reduce_kernel", b, d, g.inner, ops_arg_dat(h->w, 1, h->s0, "double", OPS_WRITE),
          ops_arg_dat(h->Kx, 1, h->sx, "double", OPS_READ), ops_arg_dat(h->Ky, 1, h->sy, "double", OPS_READ),
          ops_arg_dat(h->Kz, 1, h->sz, "double", OPS_READ), ops_arg_dat(h->p, 1, h->sp, "double", OPS_READ),
          gbl(&h->rx), gbl(&h->ry), gbl(&h->rz), ops_arg_reduce(h->red, 1, "double", OPS_INC));
    else
      ops_par_loop_tea_leaf_cg_calc_w_reduce_kernel(
          "tea_leaf_cg_calc_w_reduce_kernel", b, d, g.inner, ops_arg_dat(h->w, 1, h->s0, "double", OPS_WRITE),
          ops_arg_dat(h->Kx, 1, h->sx, "double", OPS_READ), ops_arg_dat(h->Ky, 1, h->sy, "double", OPS_READ),
          ops_arg_dat(h->p, 1, h->sp, "double", OPS_READ), gbl(&h->rx), gbl(&h->ry),
          ops_arg_reduce(h->red, 1, "double", OPS_INC));
    ops_reduction_result_char(h->red, sizeof(double), (char *)&pw);
    double alpha = h->rro / pw;
    ops_par_loop_tea_leaf_axpy_kernel("tea_leaf_axpy_kernel", b, d, g.inner, ops_arg_dat(h->u, 1, h->s0, "double", OPS_INC),
                                      ops_arg_dat(h->p, 1, h->s0, "double", OPS_READ), gbl(&alpha));
    ops_par_loop_tea_leaf_cg_calc_ur_r_reduce_kernel(
        "tea_leaf_cg_calc_ur_r_reduce_kernel", b, d, g.inner, ops_arg_dat(h->r, 1, h->s0, "double", OPS_INC),
        ops_arg_dat(h->w, 1, h->s0, "double", OPS_READ), gbl(&alpha), ops_arg_reduce(h->red, 1, "double", OPS_INC));
    ops_reduction_result_char(h->red, sizeof(double), (char *)&rrn);
    double beta = rrn / h->rro;
    ops_par_loop_tea_leaf_axpby_kernel("tea_leaf_axpby_kernel", b, d, g.inner, ops_arg_dat(h->p, 1, h->s0, "double", OPS_RW),
                                       ops_arg_dat(h->r, 1, h->s0, "double", OPS_READ), gbl(&beta), gbl(&h->one));
    if (pw_h) pw_h[n] = pw;
    if (alpha_h) alpha_h[n] = alpha;
    if (beta_h) beta_h[n] = beta;
    if (rrn_h) rrn_h[n] = rrn;
    h->rro = rrn;
    n++;
    if (sqrt(rrn) <= stop) break;
  }
  return n;
}

double opsapp_heat_rr(void *ctx, int initial) { return initial ? ((Heat *)ctx)->rro0 : ((Heat *)ctx)->rro; }

void opsapp_heat_fetch_u(void *ctx, double *out) { ops_dat_fetch_data(((Heat *)ctx)->u, 0, (char *)out); }

void opsapp_heat_destroy(void *ctx) {
  Heat *h = (Heat *)ctx;
  ops_dat all[9] = {h->density, h->u, h->u0, h->w, h->Kx, h->Ky, h->Kz, h->r, h->p};
  for (ops_dat q : all)
    if (q) ops_free_dat(q);
  delete h;
}

// ---------------------------------------------------------------- Poisson CG
struct Poisson {
  Geo g;
  ops_block blk;
  ops_dat x, r, p, w;
  ops_stencil s0, sp;
  ops_reduction red;
  double rro = 0, rro0 = 0, one = 1.0;
  Poisson(int d, int a, int b, int c) : g(d, a, b, c) {}
};

void *opsapp_poisson_create(int ndim, int nx, int ny, int nz, uint64_t seed, const double *f) {
  Poisson *P = new Poisson(ndim, nx, ny, nz);
  Geo &g = P->g;
  P->blk = ops_decl_block(ndim, "poisson");
  P->x = decl(P->blk, g, "x");
  P->r = decl(P->blk, g, "r");
  P->p = decl(P->blk, g, "p");
  P->w = decl(P->blk, g, "w");
  P->s0 = ndim == 3 ? ops_decl_stencil(3, 1, S000, "000") : ops_decl_stencil(2, 1, S00, "00");
  P->sp = ndim == 3 ? ops_decl_stencil(3, 7, S7, "7pt") : ops_decl_stencil(2, 5, S5, "5pt");
  P->red = ops_decl_reduction_handle(sizeof(double), "double", "red");
  std::vector<double> fv;
  if (!f) {
    fv.resize(g.n_int());
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) fv[lin(i, j, k, nx, ny)] = poisson_rhs(seed, i, j, k, nx, ny);
    f = fv.data();
  }
  ops_dat_set_data(P->r, 0, (char *)f);
  int d = ndim;
  ops_par_loop_tea_leaf_yeqx_kernel("tea_leaf_yeqx_kernel", P->blk, d, g.inner, ops_arg_dat(P->p, 1, P->s0, "double", OPS_WRITE),
                                    ops_arg_dat(P->r, 1, P->s0, "double", OPS_READ));
  ops_par_loop_tea_leaf_dot_kernel("tea_leaf_dot_kernel", P->blk, d, g.inner, ops_arg_dat(P->r, 1, P->s0, "double", OPS_READ),
                                   ops_arg_dat(P->p, 1, P->s0, "double", OPS_READ),
                                   ops_arg_reduce(P->red, 1, "double", OPS_INC));
  ops_reduction_result_char(P->red, sizeof(double), (char *)&P->rro);
  P->rro0 = P->rro;
  return P;
}

int opsapp_poisson_iterate(void *ctx, int max_iters, double tol, double *rrn_h, double *alpha_h, double *beta_h) {
  Poisson *P = (Poisson *)ctx;
  Geo &g = P->g;
  int d = g.ndim;
  const double stop = tol > 0 ? tol * sqrt(P->rro0) : -1.0;
  int n = 0;
  while (n < max_iters) {
    double pw, rrn;
    ops_par_loop_poisson_calc_w_reduce_kernel("poisson_calc_w_reduce_kernel", P->blk, d, g.inner,
                                              ops_arg_dat(P->w, 1, P->s0, "double", OPS_WRITE),
                                              ops_arg_dat(P->p, 1, P->sp, "double", OPS_READ),
                                              ops_arg_reduce(P->red, 1, "double", OPS_INC));
    ops_reduction_result_char(P->red, sizeof(double), (char *)&pw);
    double alpha = P->rro / pw;
    ops_par_loop_tea_leaf_axpy_kernel("tea_leaf_axpy_kernel", P->blk, d, g.inner, ops_arg_dat(P->x, 1, P->s0, "double", OPS_INC),
                                      ops_arg_dat(P->p, 1, P->s0, "double", OPS_READ), gbl(&alpha));
    ops_par_loop_tea_leaf_cg_calc_ur_r_reduce_kernel(
        "tea_leaf_cg_calc_ur_r_reduce_kernel", P->blk, d, g.inner, ops_arg_dat(P->r, 1, P->s0, "double", OPS_INC),
        ops_arg_dat(P->w, 1, P->s0, "double", OPS_READ), gbl(&alpha), ops_arg_reduce(P->red, 1, "double", OPS_INC));
    ops_reduction_result_char(P->red, sizeof(double), (char *)&rrn);
    double beta = rrn / P->rro;
    ops_par_loop_tea_leaf_axpby_kernel("tea_leaf_axpby_kernel", P->blk, d, g.inner, ops_arg_dat(P->p, 1, P->s0, "double", OPS_RW),
                                       ops_arg_dat(P->r, 1, P->s0, "double", OPS_READ), gbl(&beta), gbl(&P->one));
    if (alpha_h) alpha_h[n] = alpha;
    if (beta_h) beta_h[n] = beta;
    if (rrn_h) rrn_h[n] = rrn;
    P->rro = rrn;
    n++;
    if (sqrt(rrn) <= stop) break;
  }
  return n;
}

void opsapp_poisson_fetch_x(void *ctx, double *out) { ops_dat_fetch_data(((Poisson *)ctx)->x, 0, (char *)out); }

void opsapp_poisson_destroy(void *ctx) {
  Poisson *P = (Poisson *)ctx;
  ops_free_dat(P->x);
  ops_free_dat(P->r);
  ops_free_dat(P->p);
  ops_free_dat(P->w);
  delete P;
}

// ------------------------------------------ mixed-precision Poisson (IR + FP32 PCG)
struct PoissonMixed {
  Geo g;
  ops_block blk;
  ops_dat f, x, r;      // double
  ops_dat e, rr, p, w;  // float
  ops_stencil s0, sp;
  ops_reduction red, dots;
  double fn = 0;
  PoissonMixed(int d, int a, int b, int c) : g(d, a, b, c) {}
};

void *opsapp_pmixed_create(int ndim, int nx, int ny, int nz, uint64_t seed, const double *f) {
  PoissonMixed *M = new PoissonMixed(ndim, nx, ny, nz);
  Geo &g = M->g;
  M->blk = ops_decl_block(ndim, "poisson_mixed");
  M->f = decl(M->blk, g, "f");
  M->x = decl(M->blk, g, "x");
  M->r = decl(M->blk, g, "r");
  M->e = decl(M->blk, g, "e", "float", 4);
  M->rr = decl(M->blk, g, "rr", "float", 4);
  M->p = decl(M->blk, g, "p32", "float", 4);
  M->w = decl(M->blk, g, "w32", "float", 4);
  M->s0 = ndim == 3 ? ops_decl_stencil(3, 1, S000, "000") : ops_decl_stencil(2, 1, S00, "00");
  M->sp = ndim == 3 ? ops_decl_stencil(3, 7, S7, "7pt") : ops_decl_stencil(2, 5, S5, "5pt");
  M->red = ops_decl_reduction_handle(sizeof(double), "double", "red");
  M->dots = ops_decl_reduction_handle(2 * sizeof(double), "double", "dots");
  std::vector<double> fv;
  if (!f) {
    fv.resize(g.n_int());
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) fv[lin(i, j, k, nx, ny)] = poisson_rhs(seed, i, j, k, nx, ny);
    f = fv.data();
  }
  ops_dat_set_data(M->f, 0, (char *)f);
  double ff = 0;
  ops_par_loop_red_sumsq("red_sumsq", M->blk, ndim, g.inner, ops_arg_dat(M->f, 1, M->s0, "double", OPS_READ),
                         ops_arg_reduce(M->red, 1, "double", OPS_INC));
  ops_reduction_result_char(M->red, sizeof(double), (char *)&ff);
  M->fn = sqrt(ff);
  return M;
}

// Iterative refinement from the current x: the outer loop recomputes the
// FP64 residual and stops at |f - A x| <= tol |f|; each correction solves
// A e = r in FP32 with Jacobi PCG to inner_tol (relative, relaxed on the last
// correction) or inner_iters.  Same steps as the oracle's orc_poisson_mixed.
int opsapp_pmixed_solve(void *ctx, double tol, int max_outer, int inner_iters, double inner_tol,
                        double *relres_h, int *inner_total) {
  PoissonMixed *M = (PoissonMixed *)ctx;
  Geo &g = M->g;
  const int d = g.ndim;
  float mi = (float)(1.0 / (d == 3 ? 6.0 : 4.0));
  int outer = 0, total = 0;
  while (true) {
    double rrd = 0;
    ops_par_loop_poisson_residual_kernel("poisson_residual_kernel", M->blk, d, g.inner,
                                         ops_arg_dat(M->r, 1, M->s0, "double", OPS_WRITE),
                                         ops_arg_dat(M->f, 1, M->s0, "double", OPS_READ),
                                         ops_arg_dat(M->x, 1, M->sp, "double", OPS_READ),
                                         ops_arg_reduce(M->red, 1, "double", OPS_INC));
    ops_reduction_result_char(M->red, sizeof(double), (char *)&rrd);
    const double rel = sqrt(rrd) / M->fn;
    if (relres_h) relres_h[outer] = rel;
    if (rel <= tol || outer >= max_outer) break;
    double dots[2];
    ops_par_loop_mp_init_kernel("mp_init_kernel", M->blk, d, g.inner, ops_arg_dat(M->e, 1, M->s0, "float", OPS_WRITE),
                                ops_arg_dat(M->rr, 1, M->s0, "float", OPS_WRITE),
                                ops_arg_dat(M->p, 1, M->s0, "float", OPS_WRITE),
                                ops_arg_dat(M->r, 1, M->s0, "double", OPS_READ),
                                ops_arg_gbl_char((char *)&mi, 1, sizeof(float), OPS_READ),
                                ops_arg_reduce(M->dots, 2, "double", OPS_INC));
    ops_reduction_result_char(M->dots, 2 * sizeof(double), (char *)dots);
    double rz = dots[0];
    // the last correction needs only reach the outer target: stop the inner
    // solve at max(inner_tol, tol / (2 relres)) of its initial residual
    const double istop = fmax(inner_tol, 0.5 * tol / rel) * sqrt(dots[1]);
    for (int it = 0; it < inner_iters; it++) {
      double pw = 0;
      ops_par_loop_poisson_calc_w_reduce_kernel("poisson_calc_w_reduce_kernel", M->blk, d, g.inner,
                                                ops_arg_dat(M->w, 1, M->s0, "float", OPS_WRITE),
                                                ops_arg_dat(M->p, 1, M->sp, "float", OPS_READ),
                                                ops_arg_reduce(M->red, 1, "double", OPS_INC));
      ops_reduction_result_char(M->red, sizeof(double), (char *)&pw);
      float a = (float)(rz / pw);
      // rr -= a w with its dots, then e += a p: the dots result does not need
      // the e update, so the queue holds it until the next p update and
      // operator sweep and runs the three as one fused pass
      ops_par_loop_mp_rr_update_kernel("mp_rr_update_kernel", M->blk, d, g.inner,
                                       ops_arg_dat(M->rr, 1, M->s0, "float", OPS_RW),
                                       ops_arg_dat(M->w, 1, M->s0, "float", OPS_READ),
                                       ops_arg_gbl_char((char *)&a, 1, sizeof(float), OPS_READ),
                                       ops_arg_gbl_char((char *)&mi, 1, sizeof(float), OPS_READ),
                                       ops_arg_reduce(M->dots, 2, "double", OPS_INC));
      ops_par_loop_mp_e_update_kernel("mp_e_update_kernel", M->blk, d, g.inner,
                                      ops_arg_dat(M->e, 1, M->s0, "float", OPS_RW),
                                      ops_arg_dat(M->p, 1, M->s0, "float", OPS_READ),
                                      ops_arg_gbl_char((char *)&a, 1, sizeof(float), OPS_READ));
      ops_reduction_result_char(M->dots, 2 * sizeof(double), (char *)dots);
      total++;
      if (sqrt(dots[1]) <= istop) break;
      float b = (float)(dots[0] / rz);
      ops_par_loop_mp_p_kernel("mp_p_kernel", M->blk, d, g.inner, ops_arg_dat(M->p, 1, M->s0, "float", OPS_RW),
                               ops_arg_dat(M->rr, 1, M->s0, "float", OPS_READ),
                               ops_arg_gbl_char((char *)&mi, 1, sizeof(float), OPS_READ),
                               ops_arg_gbl_char((char *)&b, 1, sizeof(float), OPS_READ));
      rz = dots[0];
    }
    ops_par_loop_mp_correct_kernel("mp_correct_kernel", M->blk, d, g.inner, ops_arg_dat(M->x, 1, M->s0, "double", OPS_RW),
                                   ops_arg_dat(M->e, 1, M->s0, "float", OPS_READ));
    outer++;
  }
  if (inner_total) *inner_total = total;
  return outer;
}

void opsapp_pmixed_fetch_x(void *ctx, double *out) { ops_dat_fetch_data(((PoissonMixed *)ctx)->x, 0, (char *)out); }

void opsapp_pmixed_destroy(void *ctx) {
  PoissonMixed *M = (PoissonMixed *)ctx;
  for (ops_dat q : {M->f, M->x, M->r, M->e, M->rr, M->p, M->w}) ops_free_dat(q);
  delete M;
}

// ---------------------------------------------------------------- Laplace
double opsapp_laplace2d(int nx, int ny, int iters, int mode, const double *init_ext, double *err_hist,
                        double *A_out) {
  Geo g(2, nx, ny, 1);
  ops_block blk = ops_decl_block(2, "my_grid");
  ops_dat A = decl(blk, g, "A"), An = decl(blk, g, "Anew");
  ops_stencil s00 = ops_decl_stencil(2, 1, S00, "00");
  ops_stencil s5 = ops_decl_stencil(2, 5, S5, "5pt");
  ops_reduction h_err = ops_decl_reduction_handle(sizeof(double), "double", "error");
  double pi = 2.0 * asin(1.0);
  ops_decl_const_char(nullptr, 1, "int", sizeof(int), (char *)&nx, "imax");
  ops_decl_const_char(nullptr, 1, "int", sizeof(int), (char *)&ny, "jmax");
  ops_decl_const_char(nullptr, 1, "double", sizeof(double), (char *)&pi, "pi");
  ops_dat both[2] = {A, An};
  if (init_ext) {
    for (ops_dat q : both) ops_dat_set_data_slab_memspace(q, 0, (char *)init_ext, g.full, OPS_HOST);
  } else {
    int full[] = {-1, nx + 1, -1, ny + 1}, top[] = {-1, nx + 1, ny, ny + 1};
    int left[] = {-1, 0, -1, ny + 1}, right[] = {nx, nx + 1, -1, ny + 1};
    for (ops_dat q : both) {
      ops_par_loop_set_zero("set zero", blk, 2, full, ops_arg_dat(q, 1, s00, "double", OPS_WRITE));
      if (mode == 0) {
        ops_par_loop_top_bndcon("top bndcon", blk, 2, top, ops_arg_dat(q, 1, s00, "double", OPS_WRITE), ops_arg_idx());
      } else {
        ops_par_loop_left_bndcon("left bndcon", blk, 2, left, ops_arg_dat(q, 1, s00, "double", OPS_WRITE), ops_arg_idx());
        ops_par_loop_right_bndcon("right bndcon", blk, 2, right, ops_arg_dat(q, 1, s00, "double", OPS_WRITE), ops_arg_idx());
      }
    }
  }
  int interior[] = {0, nx, 0, ny};
  double error = 1.0;
  for (int it = 0; it < iters; it++) {
    ops_par_loop_apply_stencil("apply stencil", blk, 2, interior, ops_arg_dat(A, 1, s5, "double", OPS_READ),
                               ops_arg_dat(An, 1, s00, "double", OPS_WRITE),
                               ops_arg_reduce(h_err, 1, "double", OPS_MAX));
    ops_reduction_result_char(h_err, sizeof(double), (char *)&error);
    ops_par_loop_copy("copy", blk, 2, interior, ops_arg_dat(A, 1, s00, "double", OPS_WRITE),
                      ops_arg_dat(An, 1, s00, "double", OPS_READ));
    if (err_hist) err_hist[it] = error;
  }
  if (A_out) ops_dat_fetch_data(A, 0, (char *)A_out);
  ops_free_dat(A);
  ops_free_dat(An);
  return error;
}

// ---------------------------------------------------------------- tridiagonal
struct Trid {
  Geo g;
  ops_block blk;
  ops_dat a[3], b[3], c[3], d, u;
  ops_tridsolver_params_c *ctx;
  Trid(int x, int y, int z) : g(3, x, y, z) {}
};

void *opsapp_trid_create(int nx, int ny, int nz, uint64_t seed, int strategy) {
  Trid *T = new Trid(nx, ny, nz);
  Geo &g = T->g;
  T->blk = ops_decl_block(3, "adi");
  const char *an[3] = {"ax", "ay", "az"}, *bn[3] = {"bx", "by", "bz"}, *cn[3] = {"cx", "cy", "cz"};
  size_t n = g.n_int();
  std::vector<double> va(n), vb(n), vc(n);
  for (int s = 0; s < 3; s++) {
    T->a[s] = decl(T->blk, g, an[s]);
    T->b[s] = decl(T->blk, g, bn[s]);
    T->c[s] = decl(T->blk, g, cn[s]);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          size_t l = lin(i, j, k, nx, ny);
          trid_coefs(seed, s, i, j, k, nx, ny, &va[l], &vb[l], &vc[l]);
        }
    ops_dat_set_data(T->a[s], 0, (char *)va.data());
    ops_dat_set_data(T->b[s], 0, (char *)vb.data());
    ops_dat_set_data(T->c[s], 0, (char *)vc.data());
  }
  T->d = decl(T->blk, g, "du");
  T->u = decl(T->blk, g, "u");
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) va[lin(i, j, k, nx, ny)] = trid_rhs(seed, i, j, k, nx, ny);
  ops_dat_set_data(T->d, 0, (char *)va.data());
  T->ctx = ops_tridsolver_params_create(T->blk, strategy);
  return T;
}

// one ADI sweep: solves along x, then y, then z (z accumulating into u when inc)
void opsapp_trid_sweep(void *ctx, int inc) {
  Trid *T = (Trid *)ctx;
  int *r = T->g.inner;
  ops_tridMultiDimBatch_c(3, 0, r, T->a[0], T->b[0], T->c[0], T->d, T->ctx);
  ops_tridMultiDimBatch_c(3, 1, r, T->a[1], T->b[1], T->c[1], T->d, T->ctx);
  if (inc) ops_tridMultiDimBatch_Inc_c(3, 2, r, T->a[2], T->b[2], T->c[2], T->d, T->u, T->ctx);
  else ops_tridMultiDimBatch_c(3, 2, r, T->a[2], T->b[2], T->c[2], T->d, T->ctx);
}

void opsapp_trid_solve(void *ctx, int solvedim, int inc) {
  Trid *T = (Trid *)ctx;
  int s = solvedim;
  if (inc) ops_tridMultiDimBatch_Inc_c(3, s, T->g.inner, T->a[s], T->b[s], T->c[s], T->d, T->u, T->ctx);
  else ops_tridMultiDimBatch_c(3, s, T->g.inner, T->a[s], T->b[s], T->c[s], T->d, T->ctx);
}

void opsapp_trid_fetch(void *ctx, double *d_out, double *u_out) {
  Trid *T = (Trid *)ctx;
  if (d_out) ops_dat_fetch_data(T->d, 0, (char *)d_out);
  if (u_out) ops_dat_fetch_data(T->u, 0, (char *)u_out);
}

void opsapp_trid_destroy(void *ctx) {
  Trid *T = (Trid *)ctx;
  for (int s = 0; s < 3; s++) {
    ops_free_dat(T->a[s]);
    ops_free_dat(T->b[s]);
    ops_free_dat(T->c[s]);
  }
  ops_free_dat(T->d);
  ops_free_dat(T->u);
  ops_tridsolver_params_destroy(T->ctx);
  delete T;
}

// ---------------------------------------------------------------- multi-block Poisson
// The reference's multi-block test app (apps/c/poisson/poisson.cpp:56-307)
// with its ngrid_x x ngrid_y partition exposed: blocks of uniform size
// ((n-1)/ngrid + 1, the last ones shorter), dats u, u2, f, ref with depth-1
// halos, halos between neighbouring blocks exchanged each iteration by
// ops_halo_transfer, then Jacobi sweeps (poisson_kernel_stencil) and copies.
// Returns the total squared error against the exact solution; u_out (may
// be NULL) gets the whole interior, x fastest.
double opsapp_mb_poisson(int nx, int ny, int ngx, int ngy, int iters, double *u_out) {
  double dx = 0.01, dy = 0.01;
  ops_decl_const_char(nullptr, 1, "double", sizeof(double), (char *)&dx, "dx");
  ops_decl_const_char(nullptr, 1, "double", sizeof(double), (char *)&dy, "dy");
  const int nb = ngx * ngy;
  std::vector<ops_block> blocks(nb);
  std::vector<ops_dat> u(nb), u2(nb), f(nb), ref(nb);
  std::vector<int> sizes(2 * nb), disps(2 * nb);
  int s00[] = {0, 0}, s5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
  ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
  ops_stencil S5 = ops_decl_stencil(2, 5, s5, "00:10:-10:01:0-1");
  ops_reduction red_err = ops_decl_reduction_handle(sizeof(double), "double", "err");
  const int usz[2] = {(nx - 1) / ngx + 1, (ny - 1) / ngy + 1};
  int dm[2] = {-1, -1}, dp[2] = {1, 1}, base[2] = {0, 0}, stride[2] = {1, 1};
  for (int j = 0; j < ngy; j++)
    for (int i = 0; i < ngx; i++) {
      const int b = i + ngx * j;
      blocks[b] = ops_decl_block(2, "block");
      int size[2] = {usz[0], usz[1]};
      if ((i + 1) * size[0] > nx) size[0] = nx - i * size[0];
      if ((j + 1) * size[1] > ny) size[1] = ny - j * size[1];
      u[b] = ops_decl_dat_char(blocks[b], 1, size, base, dm, dp, stride, nullptr, 8, "double", "u");
      u2[b] = ops_decl_dat_char(blocks[b], 1, size, base, dm, dp, stride, nullptr, 8, "double", "u2");
      f[b] = ops_decl_dat_char(blocks[b], 1, size, base, dm, dp, stride, nullptr, 8, "double", "f");
      ref[b] = ops_decl_dat_char(blocks[b], 1, size, base, dm, dp, stride, nullptr, 8, "double", "ref");
      sizes[2 * b] = size[0];
      sizes[2 * b + 1] = size[1];
      disps[2 * b] = i * usz[0];
      disps[2 * b + 1] = j * usz[1];
    }
  std::vector<ops_halo> halos;
  for (int j = 0; j < ngy; j++)
    for (int i = 0; i < ngx; i++) {
      const int b = i + ngx * j;
      int dir[] = {1, 2};
      // the reverse halo lands in the neighbour's far halo, at the
      // neighbour's size (the reference app uses the sending block's size,
      // poisson.cpp:158,167, which is the same only for equal blocks)
      if (i > 0) {
        int iter[] = {1, sizes[2 * b + 1]}, from[] = {sizes[2 * (b - 1)] - 1, 0}, to[] = {-1, 0};
        halos.push_back(ops_decl_halo(u[b - 1], u[b], iter, from, to, dir, dir));
        from[0] = 0;
        to[0] = sizes[2 * (b - 1)];
        halos.push_back(ops_decl_halo(u[b], u[b - 1], iter, from, to, dir, dir));
      }
      if (j > 0) {
        int iter[] = {sizes[2 * b], 1}, from[] = {0, sizes[2 * (b - ngx) + 1] - 1}, to[] = {0, -1};
        halos.push_back(ops_decl_halo(u[b - ngx], u[b], iter, from, to, dir, dir));
        from[1] = 0;
        to[1] = sizes[2 * (b - ngx) + 1];
        halos.push_back(ops_decl_halo(u[b], u[b - ngx], iter, from, to, dir, dir));
      }
    }
  ops_halo_group grp = halos.empty() ? nullptr : ops_decl_halo_group((int)halos.size(), halos.data());
  for (int b = 0; b < nb; b++) {
    int rng[] = {-1, sizes[2 * b] + 1, -1, sizes[2 * b + 1] + 1};
    ops_par_loop_poisson_kernel_populate(
        "poisson_kernel_populate", blocks[b], 2, rng, ops_arg_gbl_char((char *)&disps[2 * b], 1, 4, OPS_READ),
        ops_arg_gbl_char((char *)&disps[2 * b + 1], 1, 4, OPS_READ), ops_arg_idx(),
        ops_arg_dat(u[b], 1, S00, "double", OPS_WRITE), ops_arg_dat(f[b], 1, S00, "double", OPS_WRITE),
        ops_arg_dat(ref[b], 1, S00, "double", OPS_WRITE));
    ops_par_loop_poisson_kernel_update("poisson_kernel_update", blocks[b], 2, rng,
                                       ops_arg_dat(u[b], 1, S00, "double", OPS_READ),
                                       ops_arg_dat(u2[b], 1, S00, "double", OPS_WRITE));
  }
  for (int b = 0; b < nb; b++) {
    int rng[] = {0, sizes[2 * b], 0, sizes[2 * b + 1]};
    ops_par_loop_poisson_kernel_initialguess("poisson_kernel_initialguess", blocks[b], 2, rng,
                                             ops_arg_dat(u[b], 1, S00, "double", OPS_WRITE));
  }
  for (int it = 0; it < iters; it++) {
    if (grp) ops_halo_transfer(grp);
    for (int b = 0; b < nb; b++) {
      int rng[] = {0, sizes[2 * b], 0, sizes[2 * b + 1]};
      ops_par_loop_poisson_kernel_stencil("poisson_kernel_stencil", blocks[b], 2, rng,
                                          ops_arg_dat(u[b], 1, S5, "double", OPS_READ),
                                          ops_arg_dat(f[b], 1, S00, "double", OPS_READ),
                                          ops_arg_dat(u2[b], 1, S00, "double", OPS_WRITE));
    }
    for (int b = 0; b < nb; b++) {
      int rng[] = {0, sizes[2 * b], 0, sizes[2 * b + 1]};
      ops_par_loop_poisson_kernel_update("poisson_kernel_update", blocks[b], 2, rng,
                                         ops_arg_dat(u2[b], 1, S00, "double", OPS_READ),
                                         ops_arg_dat(u[b], 1, S00, "double", OPS_WRITE));
    }
  }
  double err = 0.0;
  for (int b = 0; b < nb; b++) {
    int rng[] = {0, sizes[2 * b], 0, sizes[2 * b + 1]};
    ops_par_loop_poisson_kernel_error("poisson_kernel_error", blocks[b], 2, rng,
                                      ops_arg_dat(u[b], 1, S00, "double", OPS_READ),
                                      ops_arg_dat(ref[b], 1, S00, "double", OPS_READ),
                                      ops_arg_reduce(red_err, 1, "double", OPS_INC));
  }
  ops_reduction_result_char(red_err, sizeof(double), (char *)&err);
  if (u_out) {
    std::vector<double> tmp;
    for (int b = 0; b < nb; b++) {
      const int sx = sizes[2 * b], sy = sizes[2 * b + 1];
      tmp.resize((size_t)sx * sy);
      ops_dat_fetch_data(u[b], 0, (char *)tmp.data());
      for (int y = 0; y < sy; y++)
        memcpy(u_out + (size_t)(disps[2 * b + 1] + y) * nx + disps[2 * b], tmp.data() + (size_t)y * sx,
               (size_t)sx * sizeof(double));
    }
  }
  for (int b = 0; b < nb; b++) {
    ops_free_dat(u[b]);
    ops_free_dat(u2[b]);
    ops_free_dat(f[b]);
    ops_free_dat(ref[b]);
  }
  return err;
}

// ---------------------------------------------------------------- serving
// Request stream on the batched engine, on the wall clock: request i arrives
// arrival_s[i] seconds after the start (ascending), is submitted once it has
// arrived, and completes when a scheduler step reports it converged (its
// memory is then released).  Between steps nothing else runs on the host.
double opsapp_serve_stream(int nreq, const int *kind, const int *n, const uint64_t *seed, const double *arrival_s,
                           double tol, size_t budget, int max_slots, int iters_per_launch, int autotune, int mixed,
                           double *done_s, int *iters) {
  opsmi_server S = opsmi_serve_create(budget, max_slots, iters_per_launch, 1);
  if (!S) return -1.0;
  if (autotune && opsmi_serve_autotune(S, 10) != 0) {
    opsmi_serve_destroy(S);
    return -1.0;
  }
  std::vector<int> idx_of;
  std::vector<opsmi_solve_done> buf(256);
  const double t0 = wtime();
  int next = 0, done = 0;
  while (done < nreq) {
    double t = wtime() - t0;
    while (next < nreq && arrival_s[next] <= t) {
      opsmi_solve_req q{kind[next], n[next], seed[next], tol, 1000000, mixed ? 1 : 0};
      const int id = opsmi_serve_submit(S, &q);
      if (id < 0) {
        opsmi_serve_destroy(S);
        return -1.0;
      }
      if ((int)idx_of.size() <= id) idx_of.resize(id + 1, -1);
      idx_of[id] = next++;
    }
    if (opsmi_serve_active(S) + opsmi_serve_queued(S) == 0) {
      while (next < nreq && wtime() - t0 < arrival_s[next]) {
      }
      continue;
    }
    if (opsmi_serve_step(S) < 0) {
      opsmi_serve_destroy(S);
      return -1.0;
    }
    const double tc = wtime() - t0;
    int k;
    while ((k = opsmi_serve_completed(S, buf.data(), (int)buf.size())) > 0)
      for (int j = 0; j < k; j++) {
        const int i = idx_of[buf[j].id];
        done_s[i] = tc;
        iters[i] = buf[j].status == 1 ? buf[j].iters : -buf[j].iters;
        opsmi_serve_release(S, buf[j].id);
        done++;
      }
  }
  const double span = wtime() - t0;
  opsmi_serve_destroy(S);
  return span;
}

}  // extern "C"
