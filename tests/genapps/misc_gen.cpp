// More dev-API loops for the translator: a 1D block with float dats and a
// float MAX reduction, a 3D three-component dat with a three-component INC
// and an int MIN reduction, an optional dat (ops_arg_dat_opt), and a 2D
// fill whose result the caller reads through a hand-written loop (the two
// kinds of loop share one queue).
#include "ops_seq_v2.h"

void ramp1d(ACC<float> &f, const int *idx, const float *c) {
  f(0) = c[0] * idx[0] + c[1];
}

void smooth1d(const ACC<float> &f, ACC<float> &g, float *mx) {
  g(0) = 0.25f * f(-1) + 0.5f * f(0) + 0.25f * f(1);
  *mx = fmaxf(*mx, g(0));
}

void fill_vec3(ACC<double> &v, const int *idx) {
  v(0, 0, 0, 0) = idx[0];
  v(1, 0, 0, 0) = 2.0 * idx[1];
  v(2, 0, 0, 0) = 3.0 * idx[2] + 1.0;
}

void div_vec3(const ACC<double> &v, ACC<double> &d, double *sums, int *imin) {
  d(0, 0, 0) = (v(0, 1, 0, 0) - v(0, -1, 0, 0)) + (v(1, 0, 1, 0) - v(1, 0, -1, 0)) + (v(2, 0, 0, 1) - v(2, 0, 0, -1));
  sums[0] = sums[0] + v(0, 0, 0, 0);
  sums[1] = sums[1] + v(1, 0, 0, 0);
  sums[2] = sums[2] + v(2, 0, 0, 0);
  const int q = (int)d(0, 0, 0) - (int)v(0, 0, 0, 0);
  *imin = q < *imin ? q : *imin;
}

void opt_add(ACC<double> &d, const ACC<double> &e, const int *use) {
  if (*use) d(0, 0, 0) = d(0, 0, 0) + e(0, 0, 0);
  else d(0, 0, 0) = 2.0 * d(0, 0, 0);
}

void fill2d(ACC<double> &a, const int *idx, const double *val) {
  a(0, 0) = *val + idx[0] - 0.5 * idx[1];
}

// f, g: 1D float dats of n points with a one-point halo
extern "C" int gen_ramp1d(ops_block block, ops_dat f, ops_dat g, int n, float c0, float c1, float *mx_out) {
  int s0[] = {0};
  ops_stencil S0 = ops_decl_stencil(1, 1, s0, "m0");
  int s3[] = {0, -1, 1};
  ops_stencil S3 = ops_decl_stencil(1, 3, s3, "m3");
  ops_reduction h = ops_decl_reduction_handle(sizeof(float), "float", "mx");
  int whole[] = {-1, n + 1};
  int inner[] = {0, n};
  float c[2] = {c0, c1};
  ops_par_loop(ramp1d, "ramp1d", block, 1, whole, ops_arg_dat(f, 1, S0, "float", OPS_WRITE), ops_arg_idx(),
               ops_arg_gbl(c, 2, "float", OPS_READ));
  ops_par_loop(smooth1d, "smooth1d", block, 1, inner, ops_arg_dat(f, 1, S3, "float", OPS_READ),
               ops_arg_dat(g, 1, S0, "float", OPS_WRITE), ops_arg_reduce(h, 1, "float", OPS_MAX));
  ops_reduction_result(h, mx_out);
  return 0;
}

// v: 3D three-component dat, d: one component; both with a one-point halo
extern "C" int gen_vec3(ops_block block, ops_dat v, ops_dat d, int nx, int ny, int nz, double *sums, int *imin) {
  int s0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, s0, "v0");
  int s7[] = {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
  ops_stencil S7 = ops_decl_stencil(3, 7, s7, "v7");
  ops_reduction hs = ops_decl_reduction_handle(3 * sizeof(double), "double", "sums3");
  ops_reduction hm = ops_decl_reduction_handle(sizeof(int), "int", "imin");
  int whole[] = {-1, nx + 1, -1, ny + 1, -1, nz + 1};
  int inner[] = {0, nx, 0, ny, 0, nz};
  ops_par_loop(fill_vec3, "fill_vec3", block, 3, whole, ops_arg_dat(v, 3, S0, "double", OPS_WRITE), ops_arg_idx());
  ops_par_loop(div_vec3, "div_vec3", block, 3, inner, ops_arg_dat(v, 3, S7, "double", OPS_READ),
               ops_arg_dat(d, 1, S0, "double", OPS_WRITE), ops_arg_reduce(hs, 3, "double", OPS_INC),
               ops_arg_reduce(hm, 1, "int", OPS_MIN));
  ops_reduction_result(hs, sums);
  ops_reduction_result(hm, imin);
  return 0;
}

// d += e where use != 0 (e passed through ops_arg_dat_opt), else d *= 2
extern "C" int gen_opt_add(ops_block block, ops_dat d, ops_dat e, int nx, int ny, int nz, int use) {
  int s0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, s0, "o0");
  int inner[] = {0, nx, 0, ny, 0, nz};
  ops_par_loop(opt_add, "opt_add", block, 3, inner, ops_arg_dat(d, 1, S0, "double", OPS_RW),
               ops_arg_dat_opt(e, 1, S0, "double", OPS_READ, use), ops_arg_gbl(&use, 1, "int", OPS_READ));
  return 0;
}

// queues a fill of a over the interior and returns without a result
extern "C" int gen_fill2d(ops_block block, ops_dat a, int nx, int ny, double val) {
  int s0[] = {0, 0};
  ops_stencil S0 = ops_decl_stencil(2, 1, s0, "f0");
  int inner[] = {0, nx, 0, ny};
  ops_par_loop(fill2d, "fill2d", block, 2, inner, ops_arg_dat(a, 1, S0, "double", OPS_WRITE), ops_arg_idx(),
               ops_arg_gbl(&val, 1, "double", OPS_READ));
  return 0;
}

// reps fills of a in a row, then the queue runs: the per-loop cost of a
// small generated loop (host work and launch) when the grid is small
extern "C" int gen_fill2d_many(ops_block block, ops_dat a, int nx, int ny, int reps) {
  int s0[] = {0, 0};
  ops_stencil S0 = ops_decl_stencil(2, 1, s0, "fm0");
  int inner[] = {0, nx, 0, ny};
  for (int k = 0; k < reps; k++) {
    double val = k;
    ops_par_loop(fill2d, "fill2d many", block, 2, inner, ops_arg_dat(a, 1, S0, "double", OPS_WRITE), ops_arg_idx(),
                 ops_arg_gbl(&val, 1, "double", OPS_READ));
  }
  ops_execute();
  return 0;
}

void double_it(ACC<double> &d, const ACC<double> &s) {
  d(0, 0, 0) = 2.0 * s(0, 0, 0);
}

// d = 2 d with d passed as both arguments (the kernel without restrict-
// qualified dat pointers runs)
extern "C" int gen_double_in_place(ops_block block, ops_dat d, int nx, int ny, int nz) {
  int s0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, s0, "d0");
  int inner[] = {0, nx, 0, ny, 0, nz};
  ops_par_loop(double_it, "double_it", block, 3, inner, ops_arg_dat(d, 1, S0, "double", OPS_WRITE),
               ops_arg_dat(d, 1, S0, "double", OPS_READ));
  return 0;
}

void cg_axpy(ACC<double> &u, const ACC<double> &p, const double *a) {
  u(0, 0, 0) = u(0, 0, 0) + (*a) * p(0, 0, 0);
}

void cg_axpby_dot(ACC<double> &p, const ACC<double> &r, const double *b, double *pp) {
  p(0, 0, 0) = r(0, 0, 0) + (*b) * p(0, 0, 0);
  *pp = *pp + p(0, 0, 0) * p(0, 0, 0);
}

// The tail of a CG iteration, `iters` times over the interior: u = u + a p,
// then p = r + b p with p.p into pp[it] (two loops on shared dats that touch
// them only at the point itself, so the queue may run them as one pass)
extern "C" int gen_cg_tail(ops_block block, ops_dat u, ops_dat p, ops_dat r, int nx, int ny, int nz, double a,
                           double b, int iters, double *pp) {
  int s0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, s0, "t0");
  ops_reduction h = ops_decl_reduction_handle(sizeof(double), "double", "pp");
  int inner[] = {0, nx, 0, ny, 0, nz};
  for (int it = 0; it < iters; it++) {
    ops_par_loop(cg_axpy, "cg_axpy", block, 3, inner, ops_arg_dat(u, 1, S0, "double", OPS_RW),
                 ops_arg_dat(p, 1, S0, "double", OPS_READ), ops_arg_gbl(&a, 1, "double", OPS_READ));
    ops_par_loop(cg_axpby_dot, "cg_axpby_dot", block, 3, inner, ops_arg_dat(p, 1, S0, "double", OPS_RW),
                 ops_arg_dat(r, 1, S0, "double", OPS_READ), ops_arg_gbl(&b, 1, "double", OPS_READ),
                 ops_arg_reduce(h, 1, "double", OPS_INC));
    ops_reduction_result(h, &pp[it]);
  }
  return 0;
}

void cg3_u(ACC<double> &u, const ACC<double> &p, const double *a) {
  u(0, 0, 0) = u(0, 0, 0) + (*a) * p(0, 0, 0);
}

void cg3_r(ACC<double> &r, const ACC<double> &w, const double *a, double *rr) {
  r(0, 0, 0) = r(0, 0, 0) - (*a) * w(0, 0, 0);
  *rr = *rr + r(0, 0, 0) * r(0, 0, 0);
}

void cg3_p(ACC<double> &p, const ACC<double> &r, const double *b, double *pp) {
  p(0, 0, 0) = r(0, 0, 0) + (*b) * p(0, 0, 0);
  *pp = *pp + p(0, 0, 0) * p(0, 0, 0);
}

// Three loops on shared dats that touch them only at the point, `iters`
// times: u = u + a p; r = r - a w with r.r into rr[it]; p = r + b p with p.p
// into pp[it].  short_p != 0 runs the third loop one z plane short, so the
// three cannot run as one pass (the first two still can).
extern "C" int gen_cg_three(ops_block block, ops_dat u, ops_dat p, ops_dat r, ops_dat w, int nx, int ny, int nz,
                            double a, double b, int iters, int short_p, double *rr, double *pp) {
  int s0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, s0, "c0");
  ops_reduction hr = ops_decl_reduction_handle(sizeof(double), "double", "rr");
  ops_reduction hp = ops_decl_reduction_handle(sizeof(double), "double", "pp");
  int inner[] = {0, nx, 0, ny, 0, nz};
  int pr[] = {0, nx, 0, ny, 0, short_p ? nz - 1 : nz};
  for (int it = 0; it < iters; it++) {
    ops_par_loop(cg3_u, "cg3_u", block, 3, inner, ops_arg_dat(u, 1, S0, "double", OPS_RW),
                 ops_arg_dat(p, 1, S0, "double", OPS_READ), ops_arg_gbl(&a, 1, "double", OPS_READ));
    ops_par_loop(cg3_r, "cg3_r", block, 3, inner, ops_arg_dat(r, 1, S0, "double", OPS_RW),
                 ops_arg_dat(w, 1, S0, "double", OPS_READ), ops_arg_gbl(&a, 1, "double", OPS_READ),
                 ops_arg_reduce(hr, 1, "double", OPS_INC));
    ops_par_loop(cg3_p, "cg3_p", block, 3, pr, ops_arg_dat(p, 1, S0, "double", OPS_RW),
                 ops_arg_dat(r, 1, S0, "double", OPS_READ), ops_arg_gbl(&b, 1, "double", OPS_READ),
                 ops_arg_reduce(hp, 1, "double", OPS_INC));
    // p.p first: its flush needs all three loops (r.r alone would need only
    // the second)
    ops_reduction_result(hp, &pp[it]);
    ops_reduction_result(hr, &rr[it]);
  }
  return 0;
}
