/* 3D fixture driver: runs this project's 3D test problems through the
 * reference's sequential OPS library (ops/c/include/ops_seq_v2.h:212-335).
 *
 * Problems:
 *  - heat3d: TeaLeaf-form CG (apps/c/TeaLeaf/tea_leaf_cg.cpp:54-214,
 *    tea_solve.cpp:300-325) generalised to 3D.  The 3D operator appends the
 *    z terms after the x terms of apps/c/TeaLeaf/tea_leaf_cg_kernels.h:8-15:
 *      w = (1 + ry(Ky(0,1,0)+Ky) + rx(Kx(1,0,0)+Kx) + rz(Kz(0,0,1)+Kz)) p
 *          - ry(Ky(0,1,0) p(0,1,0) + Ky p(0,-1,0))
 *          - rx(Kx(1,0,0) p(1,0,0) + Kx p(-1,0,0))
 *          - rz(Kz(0,0,1) p(0,0,1) + Kz p(0,0,-1))
 *  - poisson3d: constant-coefficient 7-point CG,
 *      w = 6 p - (p(-1,0,0)+p(1,0,0)) - (p(0,-1,0)+p(0,1,0)) - (p(0,0,-1)+p(0,0,1))
 *    plus one FP32 sweep of the same operator.
 *  - adi preproc: apps/c/adi/preproc_kernel.h:4-39 with lambda = 1.
 *  - stencil27: a 27-point weighted box sweep (weights 1/8,1/16,1/32,1/64 by
 *    Manhattan distance 0..3, summed z-outer, y, x-inner).
 *
 * usage: ref_3d OUTDIR
 */
#define OPS_3D
#include <ops_seq_v2.h>

#include "problems.h"
#include "refdrv_common.h"

static const uint64_t SEED = 20240611ULL;

static const double *g_src = nullptr;
static int g_x0, g_y0, g_z0, g_w, g_h;
static void load_kernel(ACC<double> &u, const int *idx) {
  u(0, 0, 0) = g_src[(idx[0] - g_x0) + (size_t)g_w * ((idx[1] - g_y0) + (size_t)g_h * (idx[2] - g_z0))];
}
static void set_src(const std::vector<double> &v, int x0, int y0, int z0, int w, int h) {
  g_src = v.data(); g_x0 = x0; g_y0 = y0; g_z0 = z0; g_w = w; g_h = h;
}

static int S000[] = {0, 0, 0};
static int S7[] = {0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1};
static int S_P100[] = {0, 0, 0, 1, 0, 0};
static int S_0P10[] = {0, 0, 0, 0, 1, 0};
static int S_00P1[] = {0, 0, 0, 0, 0, 1};
static int S_M[] = {0, 0, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};

static void init_zero(ACC<double> &p) { p(0, 0, 0) = 0.0; }
static void yeqx(ACC<double> &p, const ACC<double> &x) { p(0, 0, 0) = x(0, 0, 0); }
static void recip(ACC<double> &u, const ACC<double> &p) { u(0, 0, 0) = 1.0 / p(0, 0, 0); }
static void dot_k(const ACC<double> &r, const ACC<double> &p, double *rro) {
  *rro = *rro + r(0, 0, 0) * p(0, 0, 0);
}
static void axpy(ACC<double> &u, const ACC<double> &p, const double *alpha) {
  u(0, 0, 0) = u(0, 0, 0) + (*alpha) * p(0, 0, 0);
}
static void axpby(ACC<double> &u, const ACC<double> &p, const double *alpha, const double *beta) {
  u(0, 0, 0) = (*alpha) * u(0, 0, 0) + (*beta) * p(0, 0, 0);
}
static void init_K3(ACC<double> &Kx, ACC<double> &Ky, ACC<double> &Kz, const ACC<double> &w) {
  Kx(0, 0, 0) = (w(-1, 0, 0) + w(0, 0, 0)) / (2.0 * w(-1, 0, 0) * w(0, 0, 0));
  Ky(0, 0, 0) = (w(0, -1, 0) + w(0, 0, 0)) / (2.0 * w(0, -1, 0) * w(0, 0, 0));
  Kz(0, 0, 0) = (w(0, 0, -1) + w(0, 0, 0)) / (2.0 * w(0, 0, -1) * w(0, 0, 0));
}
static inline double heat_op(const ACC<double> &Kx, const ACC<double> &Ky, const ACC<double> &Kz,
                             const ACC<double> &p, double rx, double ry, double rz) {
  return (1.0 + ry * (Ky(0, 1, 0) + Ky(0, 0, 0)) + rx * (Kx(1, 0, 0) + Kx(0, 0, 0)) +
          rz * (Kz(0, 0, 1) + Kz(0, 0, 0))) * p(0, 0, 0) -
         ry * (Ky(0, 1, 0) * p(0, 1, 0) + Ky(0, 0, 0) * p(0, -1, 0)) -
         rx * (Kx(1, 0, 0) * p(1, 0, 0) + Kx(0, 0, 0) * p(-1, 0, 0)) -
         rz * (Kz(0, 0, 1) * p(0, 0, 1) + Kz(0, 0, 0) * p(0, 0, -1));
}
static void common_init3(ACC<double> &w, ACC<double> &r, const ACC<double> &Kx, const ACC<double> &Ky,
                         const ACC<double> &Kz, const ACC<double> &u, const double *rx, const double *ry,
                         const double *rz) {
  w(0, 0, 0) = heat_op(Kx, Ky, Kz, u, *rx, *ry, *rz);
  r(0, 0, 0) = u(0, 0, 0) - w(0, 0, 0);
}
static void cg_calc_w3(ACC<double> &w, const ACC<double> &Kx, const ACC<double> &Ky, const ACC<double> &Kz,
                       const ACC<double> &p, const double *rx, const double *ry, const double *rz, double *pw) {
  w(0, 0, 0) = heat_op(Kx, Ky, Kz, p, *rx, *ry, *rz);
  *pw = *pw + w(0, 0, 0) * p(0, 0, 0);
}
static void cg_calc_ur_r(ACC<double> &r, const ACC<double> &w, const double *alpha, double *rnn) {
  r(0, 0, 0) = r(0, 0, 0) - (*alpha) * w(0, 0, 0);
  *rnn = *rnn + r(0, 0, 0) * r(0, 0, 0);
}
static void poisson_calc_w3(ACC<double> &w, const ACC<double> &p, double *pw) {
  w(0, 0, 0) = 6.0 * p(0, 0, 0) - (p(-1, 0, 0) + p(1, 0, 0)) - (p(0, -1, 0) + p(0, 1, 0)) -
               (p(0, 0, -1) + p(0, 0, 1));
  *pw = *pw + w(0, 0, 0) * p(0, 0, 0);
}
static void poisson_calc_w3_f(ACC<float> &w, const ACC<float> &p, double *pw) {
  w(0, 0, 0) = 6.0f * p(0, 0, 0) - (p(-1, 0, 0) + p(1, 0, 0)) - (p(0, -1, 0) + p(0, 1, 0)) -
               (p(0, 0, -1) + p(0, 0, 1));
  *pw = *pw + (double)w(0, 0, 0) * (double)p(0, 0, 0);
}
static const float *g_srcf = nullptr;
static void load_kernel_f(ACC<float> &u, const int *idx) {
  u(0, 0, 0) = g_srcf[(idx[0] - g_x0) + (size_t)g_w * ((idx[1] - g_y0) + (size_t)g_h * (idx[2] - g_z0))];
}

static void run_heat3d(int nx, int ny, int nz, int iters) {
  ops_block g = ops_decl_block(3, "heat");
  int size[] = {nx, ny, nz}, base[] = {0, 0, 0}, d_m[] = {-1, -1, -1}, d_p[] = {1, 1, 1};
  double *tmp = NULL;
  auto mk = [&](const char *n) { return ops_decl_dat(g, 1, size, base, d_m, d_p, tmp, "double", n); };
  ops_dat density = mk("density"), u = mk("u"), u0 = mk("u0"), w = mk("w"), Kx = mk("Kx"), Ky = mk("Ky"),
          Kz = mk("Kz"), r = mk("r"), p = mk("p");
  ops_stencil s000 = ops_decl_stencil(3, 1, S000, "000");
  ops_stencil s7 = ops_decl_stencil(3, 7, S7, "7pt");
  ops_stencil sx = ops_decl_stencil(3, 2, S_P100, "P100");
  ops_stencil sy = ops_decl_stencil(3, 2, S_0P10, "0P10");
  ops_stencil sz = ops_decl_stencil(3, 2, S_00P1, "00P1");
  ops_stencil sm = ops_decl_stencil(3, 4, S_M, "M");
  ops_reduction red = ops_decl_reduction_handle(sizeof(double), "double", "red");

  size_t ex = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  std::vector<double> K(ex), rhs((size_t)nx * ny * nz);
  for (int k = -1; k <= nz; k++)
    for (int j = -1; j <= ny; j++)
      for (int i = -1; i <= nx; i++)
        K[orc_lin3_ext(i, j, k, nx, ny)] = orc_heat_conductivity(SEED, i, j, k, nx, ny);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) rhs[orc_lin3(i, j, k, nx, ny)] = orc_heat_rhs(SEED, i, j, k, nx, ny, nz, 3);
  int full[] = {-1, nx + 1, -1, ny + 1, -1, nz + 1};
  int inner[] = {0, nx, 0, ny, 0, nz};
  set_src(K, -1, -1, -1, nx + 2, ny + 2);
  ops_par_loop(load_kernel, "load", g, 3, full, ops_arg_dat(density, 1, s000, "double", OPS_WRITE), ops_arg_idx());
  set_src(rhs, 0, 0, 0, nx, ny);
  ops_par_loop(load_kernel, "load", g, 3, inner, ops_arg_dat(u0, 1, s000, "double", OPS_WRITE), ops_arg_idx());
  ops_par_loop(yeqx, "yeqx", g, 3, inner, ops_arg_dat(u, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(u0, 1, s000, "double", OPS_READ));
  double rx = 1.0, ry = 1.0, rz = 1.0;
  ops_par_loop(recip, "recip", g, 3, full, ops_arg_dat(w, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(density, 1, s000, "double", OPS_READ));
  int ext[] = {0, nx + 1, 0, ny + 1, 0, nz + 1};
  ops_par_loop(init_K3, "init_K", g, 3, ext, ops_arg_dat(Kx, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(Ky, 1, s000, "double", OPS_WRITE), ops_arg_dat(Kz, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(w, 1, sm, "double", OPS_READ));
  int r0[] = {-1, 1, -1, ny + 1, -1, nz + 1}, r1[] = {nx, nx + 1, -1, ny + 1, -1, nz + 1};
  int r2[] = {-1, nx + 1, -1, 1, -1, nz + 1}, r3[] = {-1, nx + 1, ny, ny + 1, -1, nz + 1};
  int r4[] = {-1, nx + 1, -1, ny + 1, -1, 1}, r5[] = {-1, nx + 1, -1, ny + 1, nz, nz + 1};
  ops_par_loop(init_zero, "zero", g, 3, r0, ops_arg_dat(Kx, 1, s000, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 3, r1, ops_arg_dat(Kx, 1, s000, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 3, r2, ops_arg_dat(Ky, 1, s000, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 3, r3, ops_arg_dat(Ky, 1, s000, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 3, r4, ops_arg_dat(Kz, 1, s000, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 3, r5, ops_arg_dat(Kz, 1, s000, "double", OPS_WRITE));
  ops_par_loop(common_init3, "common_init", g, 3, inner, ops_arg_dat(w, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(r, 1, s000, "double", OPS_WRITE), ops_arg_dat(Kx, 1, sx, "double", OPS_READ),
               ops_arg_dat(Ky, 1, sy, "double", OPS_READ), ops_arg_dat(Kz, 1, sz, "double", OPS_READ),
               ops_arg_dat(u, 1, s7, "double", OPS_READ), ops_arg_gbl(&rx, 1, "double", OPS_READ),
               ops_arg_gbl(&ry, 1, "double", OPS_READ), ops_arg_gbl(&rz, 1, "double", OPS_READ));
  double rro, one = 1.0;
  ops_par_loop(yeqx, "yeqx", g, 3, inner, ops_arg_dat(p, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(r, 1, s000, "double", OPS_READ));
  ops_par_loop(dot_k, "dot", g, 3, inner, ops_arg_dat(r, 1, s000, "double", OPS_READ),
               ops_arg_dat(p, 1, s000, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
  ops_reduction_result(red, &rro);
  std::vector<double> h_pw, h_rrn, h_alpha, h_beta;
  h_rrn.push_back(rro);
  for (int n = 0; n < iters; n++) {
    double pw, rrn;
    ops_par_loop(cg_calc_w3, "cg_calc_w", g, 3, inner, ops_arg_dat(w, 1, s000, "double", OPS_WRITE),
                 ops_arg_dat(Kx, 1, sx, "double", OPS_READ), ops_arg_dat(Ky, 1, sy, "double", OPS_READ),
                 ops_arg_dat(Kz, 1, sz, "double", OPS_READ), ops_arg_dat(p, 1, s7, "double", OPS_READ),
                 ops_arg_gbl(&rx, 1, "double", OPS_READ), ops_arg_gbl(&ry, 1, "double", OPS_READ),
                 ops_arg_gbl(&rz, 1, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &pw);
    double alpha = rro / pw;
    ops_par_loop(axpy, "axpy", g, 3, inner, ops_arg_dat(u, 1, s000, "double", OPS_INC),
                 ops_arg_dat(p, 1, s000, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ));
    ops_par_loop(cg_calc_ur_r, "cg_calc_ur_r", g, 3, inner, ops_arg_dat(r, 1, s000, "double", OPS_INC),
                 ops_arg_dat(w, 1, s000, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ),
                 ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &rrn);
    double beta = rrn / rro;
    ops_par_loop(axpby, "axpby", g, 3, inner, ops_arg_dat(p, 1, s000, "double", OPS_RW),
                 ops_arg_dat(r, 1, s000, "double", OPS_READ), ops_arg_gbl(&beta, 1, "double", OPS_READ),
                 ops_arg_gbl(&one, 1, "double", OPS_READ));
    rro = rrn;
    h_pw.push_back(pw); h_rrn.push_back(rrn); h_alpha.push_back(alpha); h_beta.push_back(beta);
  }
  size_t n3 = (size_t)nx * ny * nz;
  drv_write_vec("heat3d_pw", h_pw);
  drv_write_vec("heat3d_rrn", h_rrn);
  drv_write_vec("heat3d_alpha", h_alpha);
  drv_write_vec("heat3d_beta", h_beta);
  drv_write_vec("heat3d_u", drv_fetch<double>(u, n3));
  drv_write_vec("heat3d_Kx", drv_fetch<double>(Kx, n3));
  drv_write_vec("heat3d_Kz", drv_fetch<double>(Kz, n3));
}

static void run_poisson3d(int nx, int ny, int nz, int iters) {
  ops_block g = ops_decl_block(3, "pois");
  int size[] = {nx, ny, nz}, base[] = {0, 0, 0}, d_m[] = {-1, -1, -1}, d_p[] = {1, 1, 1};
  double *tmp = NULL;
  auto mk = [&](const char *n) { return ops_decl_dat(g, 1, size, base, d_m, d_p, tmp, "double", n); };
  ops_dat x = mk("x"), r = mk("r"), p = mk("p"), w = mk("w");
  ops_stencil s000 = ops_decl_stencil(3, 1, S000, "000");
  ops_stencil s7 = ops_decl_stencil(3, 7, S7, "7pt");
  ops_reduction red = ops_decl_reduction_handle(sizeof(double), "double", "red");
  std::vector<double> f((size_t)nx * ny * nz);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) f[orc_lin3(i, j, k, nx, ny)] = orc_poisson_rhs(SEED, i, j, k, nx, ny);
  int inner[] = {0, nx, 0, ny, 0, nz};
  set_src(f, 0, 0, 0, nx, ny);
  ops_par_loop(load_kernel, "load", g, 3, inner, ops_arg_dat(r, 1, s000, "double", OPS_WRITE), ops_arg_idx());
  ops_par_loop(yeqx, "yeqx", g, 3, inner, ops_arg_dat(p, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(r, 1, s000, "double", OPS_READ));
  double rro, one = 1.0;
  ops_par_loop(dot_k, "dot", g, 3, inner, ops_arg_dat(r, 1, s000, "double", OPS_READ),
               ops_arg_dat(p, 1, s000, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
  ops_reduction_result(red, &rro);
  std::vector<double> h_alpha, h_beta, h_rrn;
  h_rrn.push_back(rro);
  for (int n = 0; n < iters; n++) {
    double pw, rrn;
    ops_par_loop(poisson_calc_w3, "poisson_calc_w", g, 3, inner, ops_arg_dat(w, 1, s000, "double", OPS_WRITE),
                 ops_arg_dat(p, 1, s7, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &pw);
    double alpha = rro / pw;
    ops_par_loop(axpy, "axpy", g, 3, inner, ops_arg_dat(x, 1, s000, "double", OPS_INC),
                 ops_arg_dat(p, 1, s000, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ));
    ops_par_loop(cg_calc_ur_r, "cg_calc_ur_r", g, 3, inner, ops_arg_dat(r, 1, s000, "double", OPS_INC),
                 ops_arg_dat(w, 1, s000, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ),
                 ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &rrn);
    double beta = rrn / rro;
    ops_par_loop(axpby, "axpby", g, 3, inner, ops_arg_dat(p, 1, s000, "double", OPS_RW),
                 ops_arg_dat(r, 1, s000, "double", OPS_READ), ops_arg_gbl(&beta, 1, "double", OPS_READ),
                 ops_arg_gbl(&one, 1, "double", OPS_READ));
    rro = rrn;
    h_alpha.push_back(alpha); h_beta.push_back(beta); h_rrn.push_back(rrn);
  }
  drv_write_vec("poisson3d_alpha", h_alpha);
  drv_write_vec("poisson3d_beta", h_beta);
  drv_write_vec("poisson3d_rrn", h_rrn);
  drv_write_vec("poisson3d_x", drv_fetch<double>(x, (size_t)nx * ny * nz));

  /* one FP32 sweep of the same operator on p rounded to float */
  float *tmpf = NULL;
  ops_dat pf = ops_decl_dat(g, 1, size, base, d_m, d_p, tmpf, "float", "pf");
  ops_dat wf = ops_decl_dat(g, 1, size, base, d_m, d_p, tmpf, "float", "wf");
  std::vector<float> fp(f.size());
  for (size_t q = 0; q < f.size(); q++) fp[q] = (float)f[q];
  g_srcf = fp.data(); g_x0 = 0; g_y0 = 0; g_z0 = 0; g_w = nx; g_h = ny;
  ops_par_loop(load_kernel_f, "load", g, 3, inner, ops_arg_dat(pf, 1, s000, "float", OPS_WRITE), ops_arg_idx());
  double pwf;
  ops_par_loop(poisson_calc_w3_f, "poisson_calc_w_f", g, 3, inner, ops_arg_dat(wf, 1, s000, "float", OPS_WRITE),
               ops_arg_dat(pf, 1, s7, "float", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
  ops_reduction_result(red, &pwf);
  drv_write_vec("poisson3d_f32_w", drv_fetch<float>(wf, (size_t)nx * ny * nz));
  std::vector<double> pv(1, pwf);
  drv_write_vec("poisson3d_f32_pw", pv);
}

/* apps/c/adi/preproc_kernel.h:4-39, lambda = 1 */
static int g_anx, g_any, g_anz;
static double g_lambda = 1.0;
static void preproc_kernel(const ACC<double> &u, ACC<double> &du, ACC<double> &ax, ACC<double> &bx,
                           ACC<double> &cx, ACC<double> &ay, ACC<double> &by, ACC<double> &cy, ACC<double> &az,
                           ACC<double> &bz, ACC<double> &cz, const int *idx) {
  double a, b, c, d;
  if (idx[0] == 0 || idx[0] == g_anx - 1 || idx[1] == 0 || idx[1] == g_any - 1 || idx[2] == 0 ||
      idx[2] == g_anz - 1) {
    d = 0.0f; a = 0.0f; b = 1.0f; c = 0.0f;
  } else {
    d = g_lambda * (u(-1, 0, 0) + u(1, 0, 0) + u(0, -1, 0) + u(0, 1, 0) + u(0, 0, -1) + u(0, 0, 1) -
                    6.0f * u(0, 0, 0));
    a = -0.5f * g_lambda;
    b = 1.0f + g_lambda;
    c = -0.5f * g_lambda;
  }
  du(0, 0, 0) = d;
  ax(0, 0, 0) = a; bx(0, 0, 0) = b; cx(0, 0, 0) = c;
  ay(0, 0, 0) = a; by(0, 0, 0) = b; cy(0, 0, 0) = c;
  az(0, 0, 0) = a; bz(0, 0, 0) = b; cz(0, 0, 0) = c;
}

static const double W27[4] = {0.125, 0.0625, 0.03125, 0.015625};
static void stencil27(ACC<double> &out, const ACC<double> &u) {
  double acc = 0.0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++)
        acc = acc + W27[abs(dx) + abs(dy) + abs(dz)] * u(dx, dy, dz);
  out(0, 0, 0) = acc;
}

static void run_adi_preproc_and_27(int nx, int ny, int nz) {
  g_anx = nx; g_any = ny; g_anz = nz;
  ops_block g = ops_decl_block(3, "adi");
  int size[] = {nx, ny, nz}, base[] = {0, 0, 0}, d_m[] = {-1, -1, -1}, d_p[] = {1, 1, 1};
  double *tmp = NULL;
  auto mk = [&](const char *n) { return ops_decl_dat(g, 1, size, base, d_m, d_p, tmp, "double", n); };
  ops_dat u = mk("u"), du = mk("du"), ax = mk("ax"), bx = mk("bx"), cx = mk("cx"), ay = mk("ay"),
          by = mk("by"), cy = mk("cy"), az = mk("az"), bz = mk("bz"), cz = mk("cz"), o = mk("o");
  ops_stencil s000 = ops_decl_stencil(3, 1, S000, "000");
  ops_stencil s7 = ops_decl_stencil(3, 7, S7, "7pt");
  int s27pts[81];
  int q = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) { s27pts[q++] = dx; s27pts[q++] = dy; s27pts[q++] = dz; }
  ops_stencil s27 = ops_decl_stencil(3, 27, s27pts, "27pt");
  /* u over the full range incl. halo: U[0,1) (stream 6) */
  std::vector<double> uv((size_t)(nx + 2) * (ny + 2) * (nz + 2));
  for (int k = -1; k <= nz; k++)
    for (int j = -1; j <= ny; j++)
      for (int i = -1; i <= nx; i++)
        uv[orc_lin3_ext(i, j, k, nx, ny)] = orc_rng_u01(SEED, 6, orc_lin3_ext(i, j, k, nx, ny));
  drv_write_vec("adi_u_ext", uv);
  int full[] = {-1, nx + 1, -1, ny + 1, -1, nz + 1};
  int inner[] = {0, nx, 0, ny, 0, nz};
  set_src(uv, -1, -1, -1, nx + 2, ny + 2);
  ops_par_loop(load_kernel, "load", g, 3, full, ops_arg_dat(u, 1, s000, "double", OPS_WRITE), ops_arg_idx());
  ops_par_loop(preproc_kernel, "preproc_kernel", g, 3, inner, ops_arg_dat(u, 1, s7, "double", OPS_READ),
               ops_arg_dat(du, 1, s000, "double", OPS_WRITE), ops_arg_dat(ax, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(bx, 1, s000, "double", OPS_WRITE), ops_arg_dat(cx, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(ay, 1, s000, "double", OPS_WRITE), ops_arg_dat(by, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(cy, 1, s000, "double", OPS_WRITE), ops_arg_dat(az, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(bz, 1, s000, "double", OPS_WRITE), ops_arg_dat(cz, 1, s000, "double", OPS_WRITE),
               ops_arg_idx());
  size_t n3 = (size_t)nx * ny * nz;
  drv_write_vec("adi_du", drv_fetch<double>(du, n3));
  drv_write_vec("adi_ax", drv_fetch<double>(ax, n3));
  drv_write_vec("adi_bx", drv_fetch<double>(bx, n3));
  ops_par_loop(stencil27, "stencil27", g, 3, inner, ops_arg_dat(o, 1, s000, "double", OPS_WRITE),
               ops_arg_dat(u, 1, s27, "double", OPS_READ));
  drv_write_vec("st27_out", drv_fetch<double>(o, n3));
}

int main(int argc, const char **argv) {
  if (argc < 2) { fprintf(stderr, "usage: %s OUTDIR\n", argv[0]); return 2; }
  g_outdir = argv[1];
  ops_init(1, argv, 1);
  run_heat3d(24, 20, 18, 50);
  run_poisson3d(20, 18, 16, 40);
  run_adi_preproc_and_27(16, 12, 10);
  ops_exit();
  return 0;
}
