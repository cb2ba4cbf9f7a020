/* 2D fixture driver: runs this project's 2D test problems through the
 * reference's sequential OPS library (dev path ops/c/include/ops_seq_v2.h:
 * 212-335) and dumps reduction histories and interior fields.
 *
 * The user kernels below are written for this project.  Where a problem is a
 * reference app, each kernel keeps the reference kernel's arithmetic and
 * association order so the reference library's loop machinery (ranges,
 * addressing, reduction init/re-arm) is what the fixture pins:
 *   Laplace:  apps/c/laplace2d_tutorial/step7/laplace_kernels.h:4-24
 *   TeaLeaf:  apps/c/TeaLeaf/tea_leaf_cg_kernels.h:8-21,
 *             apps/c/TeaLeaf/tea_leaf_kernels.h:8-62,
 *             apps/c/TeaLeaf/tea_leaf_common_kernels.h:12-25
 *   CG loop:  apps/c/TeaLeaf/tea_leaf_cg.cpp:54-214, tea_solve.cpp:300-325
 *
 * usage: ref_2d OUTDIR [kat]
 */
#define OPS_2D
#include <ops_seq_v2.h>

#include "problems.h"
#include "refdrv_common.h"

static const uint64_t SEED = 20240611ULL;

/* ------------------------------------------------------------------ loaders */
static const double *g_src = nullptr;
static int g_src_x0, g_src_y0, g_src_w;
static void load_kernel(ACC<double> &u, const int *idx) {
  u(0, 0) = g_src[(idx[0] - g_src_x0) + (size_t)(idx[1] - g_src_y0) * g_src_w];
}
static const float *g_srcf = nullptr;
static void load_kernel_f(ACC<float> &u, const int *idx) {
  u(0, 0) = g_srcf[(idx[0] - g_src_x0) + (size_t)(idx[1] - g_src_y0) * g_src_w];
}

/* --------------------------------------------------------------- Laplace */
static int g_nx, g_ny;
static void set_zero(ACC<double> &A) { A(0, 0) = 0.0; }
static void copy_k(ACC<double> &A, const ACC<double> &Anew) { A(0, 0) = Anew(0, 0); }
static void top_bndcon(ACC<double> &A, const int *idx) { A(0, 0) = orc_laplace_top(idx[0], g_nx); }
static void left_bndcon(ACC<double> &A, const int *idx) { A(0, 0) = orc_laplace_left(idx[1], g_ny); }
static void right_bndcon(ACC<double> &A, const int *idx) { A(0, 0) = orc_laplace_right(idx[1], g_ny); }
static void apply_stencil(const ACC<double> &A, ACC<double> &Anew, double *error) {
  Anew(0, 0) = 0.25f * (A(1, 0) + A(-1, 0) + A(0, -1) + A(0, 1));
  *error = fmax(*error, fabs(Anew(0, 0) - A(0, 0)));
}

static void set_zero_f(ACC<float> &A) { A(0, 0) = 0.0f; }
static void copy_f(ACC<float> &A, const ACC<float> &Anew) { A(0, 0) = Anew(0, 0); }
static void apply_stencil_f(const ACC<float> &A, ACC<float> &Anew, float *error) {
  Anew(0, 0) = 0.25f * (A(1, 0) + A(-1, 0) + A(0, -1) + A(0, 1));
  *error = fmax(*error, fabs(Anew(0, 0) - A(0, 0)));
}

static int S00[] = {0, 0};
static int S5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};

/* mode 0: top-edge boundary (project Laplace problem);
 * mode 1: left/right boundary (reference tutorial app) */
static void run_laplace(const char *tag, int nx, int ny, int iters, int mode, bool dump_field) {
  g_nx = nx;
  g_ny = ny;
  ops_block block = ops_decl_block(2, "grid");
  int size[] = {nx, ny}, base[] = {0, 0}, d_m[] = {-1, -1}, d_p[] = {1, 1};
  double *tmp = NULL;
  ops_dat A = ops_decl_dat(block, 1, size, base, d_m, d_p, tmp, "double", "A");
  ops_dat Anew = ops_decl_dat(block, 1, size, base, d_m, d_p, tmp, "double", "Anew");
  ops_stencil s00 = ops_decl_stencil(2, 1, S00, "00");
  ops_stencil s5 = ops_decl_stencil(2, 5, S5, "5pt");
  ops_reduction h_err = ops_decl_reduction_handle(sizeof(double), "double", "error");

  int full[] = {-1, nx + 1, -1, ny + 1};
  int top[] = {-1, nx + 1, ny, ny + 1};
  int left[] = {-1, 0, -1, ny + 1};
  int right[] = {nx, nx + 1, -1, ny + 1};
  ops_dat both[2] = {A, Anew};
  for (ops_dat d : both) {
    ops_par_loop(set_zero, "set_zero", block, 2, full, ops_arg_dat(d, 1, s00, "double", OPS_WRITE));
    if (mode == 0) {
      ops_par_loop(top_bndcon, "top_bndcon", block, 2, top,
                   ops_arg_dat(d, 1, s00, "double", OPS_WRITE), ops_arg_idx());
    } else {
      ops_par_loop(left_bndcon, "left_bndcon", block, 2, left,
                   ops_arg_dat(d, 1, s00, "double", OPS_WRITE), ops_arg_idx());
      ops_par_loop(right_bndcon, "right_bndcon", block, 2, right,
                   ops_arg_dat(d, 1, s00, "double", OPS_WRITE), ops_arg_idx());
    }
  }
  int interior[] = {0, nx, 0, ny};
  std::vector<double> hist;
  double error = 0.0;
  for (int it = 0; it < iters; it++) {
    ops_par_loop(apply_stencil, "apply_stencil", block, 2, interior,
                 ops_arg_dat(A, 1, s5, "double", OPS_READ),
                 ops_arg_dat(Anew, 1, s00, "double", OPS_WRITE),
                 ops_arg_reduce(h_err, 1, "double", OPS_MAX));
    ops_reduction_result(h_err, &error);
    ops_par_loop(copy_k, "copy", block, 2, interior,
                 ops_arg_dat(A, 1, s00, "double", OPS_WRITE),
                 ops_arg_dat(Anew, 1, s00, "double", OPS_READ));
    hist.push_back(error);
  }
  std::string t(tag);
  drv_write_vec(t + "_err", hist);
  if (dump_field) drv_write_vec(t + "_A", drv_fetch<double>(A, (size_t)nx * ny));
  printf("%s final error %.17g\n", tag, error);
}

/* float storage Laplace with a prescribed initial field (incl. halo) */
static void run_laplace_f32(int nx, int ny, int iters) {
  ops_block block = ops_decl_block(2, "gridf");
  int size[] = {nx, ny}, base[] = {0, 0}, d_m[] = {-1, -1}, d_p[] = {1, 1};
  float *tmp = NULL;
  ops_dat A = ops_decl_dat(block, 1, size, base, d_m, d_p, tmp, "float", "A");
  ops_dat Anew = ops_decl_dat(block, 1, size, base, d_m, d_p, tmp, "float", "Anew");
  ops_stencil s00 = ops_decl_stencil(2, 1, S00, "00");
  ops_stencil s5 = ops_decl_stencil(2, 5, S5, "5pt");
  ops_reduction h_err = ops_decl_reduction_handle(sizeof(float), "float", "error");
  /* initial field: top halo row sin(), interior U[0,1) rounded to float */
  std::vector<float> init((size_t)(nx + 2) * (ny + 2), 0.0f);
  for (int j = -1; j <= ny; j++)
    for (int i = -1; i <= nx; i++) {
      float v = 0.0f;
      if (j == ny) v = (float)orc_laplace_top(i, nx);
      else if (i >= 0 && i < nx && j >= 0 && j < ny) v = (float)orc_rng_u01(SEED, 4, orc_lin3(i, j, 0, nx, ny));
      init[(i + 1) + (size_t)(j + 1) * (nx + 2)] = v;
    }
  drv_write_vec("laplace_f32_init", init);
  int full[] = {-1, nx + 1, -1, ny + 1};
  g_srcf = init.data();
  g_src_x0 = -1;
  g_src_y0 = -1;
  g_src_w = nx + 2;
  ops_par_loop(load_kernel_f, "load", block, 2, full, ops_arg_dat(A, 1, s00, "float", OPS_WRITE), ops_arg_idx());
  ops_par_loop(set_zero_f, "set_zero", block, 2, full, ops_arg_dat(Anew, 1, s00, "float", OPS_WRITE));
  int interior[] = {0, nx, 0, ny};
  std::vector<float> hist;
  float error = 0.0f;
  for (int it = 0; it < iters; it++) {
    ops_par_loop(apply_stencil_f, "apply_stencil", block, 2, interior,
                 ops_arg_dat(A, 1, s5, "float", OPS_READ),
                 ops_arg_dat(Anew, 1, s00, "float", OPS_WRITE),
                 ops_arg_reduce(h_err, 1, "float", OPS_MAX));
    ops_reduction_result(h_err, &error);
    ops_par_loop(copy_f, "copy", block, 2, interior,
                 ops_arg_dat(A, 1, s00, "float", OPS_WRITE),
                 ops_arg_dat(Anew, 1, s00, "float", OPS_READ));
    hist.push_back(error);
  }
  drv_write_vec("laplace_f32_err", hist);
  drv_write_vec("laplace_f32_A", drv_fetch<float>(A, (size_t)nx * ny));
}

/* ------------------------------------------------------------- TeaLeaf CG */
static void init_zero(ACC<double> &p) { p(0, 0) = 0.0; }
static void init_zero2(ACC<double> &p, ACC<double> &z) { p(0, 0) = 0.0; z(0, 0) = 0.0; }
static void yeqx(ACC<double> &p, const ACC<double> &x) { p(0, 0) = x(0, 0); }
static void recip(ACC<double> &u, const ACC<double> &p) { u(0, 0) = 1.0 / p(0, 0); }
static void dot_k(const ACC<double> &r, const ACC<double> &p, double *rro) { *rro = *rro + r(0, 0) * p(0, 0); }
static void axpy(ACC<double> &u, const ACC<double> &p, const double *alpha) {
  u(0, 0) = u(0, 0) + (*alpha) * p(0, 0);
}
static void axpby(ACC<double> &u, const ACC<double> &p, const double *alpha, const double *beta) {
  u(0, 0) = (*alpha) * u(0, 0) + (*beta) * p(0, 0);
}
static void init_Kx_Ky(ACC<double> &Kx, ACC<double> &Ky, const ACC<double> &w) {
  Kx(0, 0) = (w(-1, 0) + w(0, 0)) / (2.0 * w(-1, 0) * w(0, 0));
  Ky(0, 0) = (w(0, -1) + w(0, 0)) / (2.0 * w(0, -1) * w(0, 0));
}
static void common_init(ACC<double> &w, ACC<double> &r, const ACC<double> &Kx, const ACC<double> &Ky,
                        const ACC<double> &u, const double *rx, const double *ry) {
  w(0, 0) = (1.0 + (*ry) * (Ky(0, 1) + Ky(0, 0)) + (*rx) * (Kx(1, 0) + Kx(0, 0))) * u(0, 0) -
            (*ry) * (Ky(0, 1) * u(0, 1) + Ky(0, 0) * u(0, -1)) -
            (*rx) * (Kx(1, 0) * u(1, 0) + Kx(0, 0) * u(-1, 0));
  r(0, 0) = u(0, 0) - w(0, 0);
}
static void cg_calc_w(ACC<double> &w, const ACC<double> &Kx, const ACC<double> &Ky, const ACC<double> &p,
                      const double *rx, const double *ry, double *pw) {
  w(0, 0) = (1.0 + (*ry) * (Ky(0, 1) + Ky(0, 0)) + (*rx) * (Kx(1, 0) + Kx(0, 0))) * p(0, 0) -
            (*ry) * (Ky(0, 1) * p(0, 1) + Ky(0, 0) * p(0, -1)) -
            (*rx) * (Kx(1, 0) * p(1, 0) + Kx(0, 0) * p(-1, 0));
  *pw = *pw + w(0, 0) * p(0, 0);
}
static void cg_calc_ur_r(ACC<double> &r, const ACC<double> &w, const double *alpha, double *rnn) {
  r(0, 0) = r(0, 0) - (*alpha) * w(0, 0);
  *rnn = *rnn + r(0, 0) * r(0, 0);
}
/* constant-coefficient 5-point Poisson operator (this project's problem) */
static void poisson_calc_w(ACC<double> &w, const ACC<double> &p, double *pw) {
  w(0, 0) = 4.0 * p(0, 0) - (p(-1, 0) + p(1, 0)) - (p(0, -1) + p(0, 1));
  *pw = *pw + w(0, 0) * p(0, 0);
}

static int S_P10[] = {0, 0, 1, 0};
static int S_0P1[] = {0, 0, 0, 1};
static int S_M10_0M1[] = {0, 0, -1, 0, 0, -1};

static void run_tealeaf_cg(int nx, int ny, int iters) {
  ops_block g = ops_decl_block(2, "tea");
  int size[] = {nx, ny}, base[] = {0, 0}, d_m[] = {-1, -1}, d_p[] = {1, 1};
  double *tmp = NULL;
  auto mk = [&](const char *n) { return ops_decl_dat(g, 1, size, base, d_m, d_p, tmp, "double", n); };
  ops_dat density = mk("density"), u = mk("u"), u0 = mk("u0"), w = mk("w"), Kx = mk("Kx"), Ky = mk("Ky"),
          r = mk("r"), p = mk("p"), z = mk("z");
  ops_stencil s00 = ops_decl_stencil(2, 1, S00, "00");
  ops_stencil s5 = ops_decl_stencil(2, 5, S5, "5pt");
  ops_stencil sp10 = ops_decl_stencil(2, 2, S_P10, "P10");
  ops_stencil s0p1 = ops_decl_stencil(2, 2, S_0P1, "0P1");
  ops_stencil sm = ops_decl_stencil(2, 3, S_M10_0M1, "M10_0M1");
  ops_reduction red = ops_decl_reduction_handle(sizeof(double), "double", "red");

  std::vector<double> K((size_t)(nx + 2) * (ny + 2)), rhs((size_t)nx * ny);
  for (int j = -1; j <= ny; j++)
    for (int i = -1; i <= nx; i++)
      K[(i + 1) + (size_t)(j + 1) * (nx + 2)] = orc_heat_conductivity(SEED, i, j, 0, nx, ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) rhs[i + (size_t)j * nx] = orc_heat_rhs(SEED, i, j, 0, nx, ny, 1, 2);
  int full[] = {-1, nx + 1, -1, ny + 1};
  int inner[] = {0, nx, 0, ny};
  g_src = K.data(); g_src_x0 = -1; g_src_y0 = -1; g_src_w = nx + 2;
  ops_par_loop(load_kernel, "load", g, 2, full, ops_arg_dat(density, 1, s00, "double", OPS_WRITE), ops_arg_idx());
  g_src = rhs.data(); g_src_x0 = 0; g_src_y0 = 0; g_src_w = nx;
  ops_par_loop(load_kernel, "load", g, 2, inner, ops_arg_dat(u0, 1, s00, "double", OPS_WRITE), ops_arg_idx());
  ops_par_loop(yeqx, "yeqx", g, 2, inner, ops_arg_dat(u, 1, s00, "double", OPS_WRITE),
               ops_arg_dat(u0, 1, s00, "double", OPS_READ));

  double rx = 1.0, ry = 1.0;
  /* tea_leaf_common_init (apps/c/TeaLeaf/tea_leaf_common.cpp:76-146), halo depth 1 */
  ops_par_loop(recip, "recip", g, 2, full, ops_arg_dat(w, 1, s00, "double", OPS_WRITE),
               ops_arg_dat(density, 1, s00, "double", OPS_READ));
  int ext[] = {0, nx + 1, 0, ny + 1};
  ops_par_loop(init_Kx_Ky, "init_Kx_Ky", g, 2, ext, ops_arg_dat(Kx, 1, s00, "double", OPS_WRITE),
               ops_arg_dat(Ky, 1, s00, "double", OPS_WRITE), ops_arg_dat(w, 1, sm, "double", OPS_READ));
  int rl[] = {-1, 1, -1, ny + 1}, rr[] = {nx, nx + 1, -1, ny + 1};
  int rb[] = {-1, nx + 1, -1, 1}, rt[] = {-1, nx + 1, ny, ny + 1};
  ops_par_loop(init_zero, "zero", g, 2, rl, ops_arg_dat(Kx, 1, s00, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 2, rr, ops_arg_dat(Kx, 1, s00, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 2, rb, ops_arg_dat(Ky, 1, s00, "double", OPS_WRITE));
  ops_par_loop(init_zero, "zero", g, 2, rt, ops_arg_dat(Ky, 1, s00, "double", OPS_WRITE));
  ops_par_loop(common_init, "common_init", g, 2, inner, ops_arg_dat(w, 1, s00, "double", OPS_WRITE),
               ops_arg_dat(r, 1, s00, "double", OPS_WRITE), ops_arg_dat(Kx, 1, sp10, "double", OPS_READ),
               ops_arg_dat(Ky, 1, s0p1, "double", OPS_READ), ops_arg_dat(u, 1, s5, "double", OPS_READ),
               ops_arg_gbl(&rx, 1, "double", OPS_READ), ops_arg_gbl(&ry, 1, "double", OPS_READ));
  /* tea_leaf_cg_init, no preconditioner */
  double rro;
  ops_par_loop(init_zero2, "zero2", g, 2, inner, ops_arg_dat(p, 1, s00, "double", OPS_WRITE),
               ops_arg_dat(z, 1, s00, "double", OPS_WRITE));
  ops_par_loop(yeqx, "yeqx", g, 2, inner, ops_arg_dat(p, 1, s00, "double", OPS_WRITE),
               ops_arg_dat(r, 1, s00, "double", OPS_READ));
  ops_par_loop(dot_k, "dot", g, 2, inner, ops_arg_dat(r, 1, s00, "double", OPS_READ),
               ops_arg_dat(p, 1, s00, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
  ops_reduction_result(red, &rro);
  std::vector<double> h_pw, h_rrn, h_alpha, h_beta;
  h_rrn.push_back(rro);
  double one = 1.0;
  for (int n = 0; n < iters; n++) {
    double pw, rrn;
    ops_par_loop(cg_calc_w, "cg_calc_w", g, 2, inner, ops_arg_dat(w, 1, s00, "double", OPS_WRITE),
                 ops_arg_dat(Kx, 1, sp10, "double", OPS_READ), ops_arg_dat(Ky, 1, s0p1, "double", OPS_READ),
                 ops_arg_dat(p, 1, s5, "double", OPS_READ), ops_arg_gbl(&rx, 1, "double", OPS_READ),
                 ops_arg_gbl(&ry, 1, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &pw);
    double alpha = rro / pw;
    ops_par_loop(axpy, "axpy", g, 2, inner, ops_arg_dat(u, 1, s00, "double", OPS_INC),
                 ops_arg_dat(p, 1, s00, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ));
    ops_par_loop(cg_calc_ur_r, "cg_calc_ur_r", g, 2, inner, ops_arg_dat(r, 1, s00, "double", OPS_INC),
                 ops_arg_dat(w, 1, s00, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ),
                 ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &rrn);
    double beta = rrn / rro;
    ops_par_loop(axpby, "axpby", g, 2, inner, ops_arg_dat(p, 1, s00, "double", OPS_RW),
                 ops_arg_dat(r, 1, s00, "double", OPS_READ), ops_arg_gbl(&beta, 1, "double", OPS_READ),
                 ops_arg_gbl(&one, 1, "double", OPS_READ));
    rro = rrn;
    h_pw.push_back(pw); h_rrn.push_back(rrn); h_alpha.push_back(alpha); h_beta.push_back(beta);
  }
  drv_write_vec("cg2d_pw", h_pw);
  drv_write_vec("cg2d_rrn", h_rrn);
  drv_write_vec("cg2d_alpha", h_alpha);
  drv_write_vec("cg2d_beta", h_beta);
  drv_write_vec("cg2d_u", drv_fetch<double>(u, (size_t)nx * ny));
  drv_write_vec("cg2d_Kx", drv_fetch<double>(Kx, (size_t)nx * ny));
}

static void run_poisson2d_cg(int nx, int ny, int iters) {
  ops_block g = ops_decl_block(2, "pois");
  int size[] = {nx, ny}, base[] = {0, 0}, d_m[] = {-1, -1}, d_p[] = {1, 1};
  double *tmp = NULL;
  auto mk = [&](const char *n) { return ops_decl_dat(g, 1, size, base, d_m, d_p, tmp, "double", n); };
  ops_dat x = mk("x"), r = mk("r"), p = mk("p"), w = mk("w");
  ops_stencil s00 = ops_decl_stencil(2, 1, S00, "00");
  ops_stencil s5 = ops_decl_stencil(2, 5, S5, "5pt");
  ops_reduction red = ops_decl_reduction_handle(sizeof(double), "double", "red");
  std::vector<double> f((size_t)nx * ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) f[i + (size_t)j * nx] = orc_poisson_rhs(SEED, i, j, 0, nx, ny);
  int inner[] = {0, nx, 0, ny};
  g_src = f.data(); g_src_x0 = 0; g_src_y0 = 0; g_src_w = nx;
  ops_par_loop(load_kernel, "load", g, 2, inner, ops_arg_dat(r, 1, s00, "double", OPS_WRITE), ops_arg_idx());
  ops_par_loop(yeqx, "yeqx", g, 2, inner, ops_arg_dat(p, 1, s00, "double", OPS_WRITE),
               ops_arg_dat(r, 1, s00, "double", OPS_READ));
  double rro, one = 1.0;
  ops_par_loop(dot_k, "dot", g, 2, inner, ops_arg_dat(r, 1, s00, "double", OPS_READ),
               ops_arg_dat(p, 1, s00, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
  ops_reduction_result(red, &rro);
  std::vector<double> h_alpha, h_beta, h_rrn;
  h_rrn.push_back(rro);
  for (int n = 0; n < iters; n++) {
    double pw, rrn;
    ops_par_loop(poisson_calc_w, "poisson_calc_w", g, 2, inner, ops_arg_dat(w, 1, s00, "double", OPS_WRITE),
                 ops_arg_dat(p, 1, s5, "double", OPS_READ), ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &pw);
    double alpha = rro / pw;
    ops_par_loop(axpy, "axpy", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_INC),
                 ops_arg_dat(p, 1, s00, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ));
    ops_par_loop(cg_calc_ur_r, "cg_calc_ur_r", g, 2, inner, ops_arg_dat(r, 1, s00, "double", OPS_INC),
                 ops_arg_dat(w, 1, s00, "double", OPS_READ), ops_arg_gbl(&alpha, 1, "double", OPS_READ),
                 ops_arg_reduce(red, 1, "double", OPS_INC));
    ops_reduction_result(red, &rrn);
    double beta = rrn / rro;
    ops_par_loop(axpby, "axpby", g, 2, inner, ops_arg_dat(p, 1, s00, "double", OPS_RW),
                 ops_arg_dat(r, 1, s00, "double", OPS_READ), ops_arg_gbl(&beta, 1, "double", OPS_READ),
                 ops_arg_gbl(&one, 1, "double", OPS_READ));
    rro = rrn;
    h_alpha.push_back(alpha); h_beta.push_back(beta); h_rrn.push_back(rrn);
  }
  drv_write_vec("poisson2d_alpha", h_alpha);
  drv_write_vec("poisson2d_beta", h_beta);
  drv_write_vec("poisson2d_rrn", h_rrn);
  drv_write_vec("poisson2d_x", drv_fetch<double>(x, (size_t)nx * ny));
}

/* ------------------------------------------------------------ reductions */
static void red_sum(const ACC<double> &x, double *s) { *s = *s + x(0, 0); }
static void red_sumsq(const ACC<double> &x, double *s) { *s = *s + x(0, 0) * x(0, 0); }
static void red_abs(const ACC<double> &x, double *s) { *s = *s + fabs(x(0, 0)); }
static void red_min(const ACC<double> &x, double *m) { *m = fmin(*m, x(0, 0)); }
static void red_max(const ACC<double> &x, double *m) { *m = fmax(*m, x(0, 0)); }
static void red_two(const ACC<double> &x, double *s) { s[0] = s[0] + x(0, 0); s[1] = s[1] + 1.0; }
static void red_imax(const int *idx, int *m) { *m = MAX(*m, idx[0] * idx[1]); }
static void red_scaled(const ACC<double> &x, const double *scale, double *s) { *s = *s + (*scale) * x(0, 0); }

static void run_reductions(int nx, int ny) {
  ops_block g = ops_decl_block(2, "red");
  int size[] = {nx, ny}, base[] = {0, 0}, d_m[] = {-1, -1}, d_p[] = {1, 1};
  double *tmp = NULL;
  ops_dat x = ops_decl_dat(g, 1, size, base, d_m, d_p, tmp, "double", "x");
  ops_stencil s00 = ops_decl_stencil(2, 1, S00, "00");
  std::vector<double> v((size_t)nx * ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) v[i + (size_t)j * nx] = 2.0 * orc_rng_u01(SEED, 5, orc_lin3(i, j, 0, nx, ny)) - 1.0;
  drv_write_vec("red_x", v);
  int inner[] = {0, nx, 0, ny};
  int sub[] = {2, nx - 3, 1, ny - 2};
  g_src = v.data(); g_src_x0 = 0; g_src_y0 = 0; g_src_w = nx;
  ops_par_loop(load_kernel, "load", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_WRITE), ops_arg_idx());
  ops_reduction h_sum = ops_decl_reduction_handle(sizeof(double), "double", "sum");
  ops_reduction h_min = ops_decl_reduction_handle(sizeof(double), "double", "min");
  ops_reduction h_max = ops_decl_reduction_handle(sizeof(double), "double", "max");
  ops_reduction h_two = ops_decl_reduction_handle(2 * sizeof(double), "double", "two");
  ops_reduction h_imax = ops_decl_reduction_handle(sizeof(int), "int", "imax");
  std::vector<double> out;
  double t;
  ops_par_loop(red_sum, "sum", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_sum, 1, "double", OPS_INC));
  ops_reduction_result(h_sum, &t); out.push_back(t);                    /* 0: sum */
  ops_par_loop(red_min, "min", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_min, 1, "double", OPS_MIN));
  ops_reduction_result(h_min, &t); out.push_back(t);                    /* 1: min */
  ops_par_loop(red_max, "max", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_max, 1, "double", OPS_MAX));
  ops_reduction_result(h_max, &t); out.push_back(t);                    /* 2: max */
  /* two loops accumulate into one handle before the result is read */
  ops_par_loop(red_sum, "sum", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_sum, 1, "double", OPS_INC));
  ops_par_loop(red_sumsq, "sumsq", g, 2, sub, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_sum, 1, "double", OPS_INC));
  ops_reduction_result(h_sum, &t); out.push_back(t);                    /* 3: sum + sumsq(sub) */
  ops_par_loop(red_abs, "abs", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_sum, 1, "double", OPS_INC));
  ops_reduction_result(h_sum, &t); out.push_back(t);                    /* 4: re-armed sum|x| */
  /* min/max handles re-armed, over a sub-range */
  ops_par_loop(red_min, "min", g, 2, sub, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_min, 1, "double", OPS_MIN));
  ops_reduction_result(h_min, &t); out.push_back(t);                    /* 5: min(sub) */
  ops_par_loop(red_max, "max", g, 2, sub, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_max, 1, "double", OPS_MAX));
  ops_reduction_result(h_max, &t); out.push_back(t);                    /* 6: max(sub) */
  double two[2];
  ops_par_loop(red_two, "two", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_two, 2, "double", OPS_INC));
  ops_reduction_result(h_two, two); out.push_back(two[0]); out.push_back(two[1]); /* 7,8 */
  int im;
  ops_par_loop(red_imax, "imax", g, 2, inner, ops_arg_idx(), ops_arg_reduce(h_imax, 1, "int", OPS_MAX));
  ops_reduction_result(h_imax, &im); out.push_back((double)im);        /* 9: max i*j */
  double scale = -2.5;
  ops_par_loop(red_scaled, "scaled", g, 2, inner, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_gbl(&scale, 1, "double", OPS_READ), ops_arg_reduce(h_sum, 1, "double", OPS_INC));
  ops_reduction_result(h_sum, &t); out.push_back(t);                    /* 10: sum(-2.5 x) */
  /* an empty range leaves the freshly initialised value */
  int empty[] = {3, 3, 0, ny};
  ops_par_loop(red_min, "min", g, 2, empty, ops_arg_dat(x, 1, s00, "double", OPS_READ),
               ops_arg_reduce(h_min, 1, "double", OPS_MIN));
  ops_reduction_result(h_min, &t); out.push_back(t);                    /* 11: DBL_MAX */
  drv_write_vec("red_out", out);
}

int main(int argc, const char **argv) {
  if (argc < 2) { fprintf(stderr, "usage: %s OUTDIR [kat]\n", argv[0]); return 2; }
  g_outdir = argv[1];
  ops_init(1, argv, 1);
  run_laplace("laplace_top", 62, 46, 60, 0, true);
  run_laplace("laplace_app", 38, 30, 25, 1, true);
  run_laplace_f32(50, 40, 30);
  run_tealeaf_cg(30, 26, 40);
  run_poisson2d_cg(40, 33, 40);
  run_reductions(17, 13);
  if (argc > 2 && strcmp(argv[2], "kat") == 0) {
    /* the tutorial app's own known answer: 4094^2, 100 iterations
     * (apps/c/laplace2d_tutorial/step7/laplace2d.cpp:19-24,134) */
    run_laplace("laplace_kat", 4094, 4094, 100, 1, false);
  }
  ops_exit();
  return 0;
}
