// Loop pairs whose second loop reads, at stencil offsets, what the first
// writes: the queue runs them as one pass (the first loop's output at a
// neighbour recomputed there) when that gives the results of running them in
// turn, and apart otherwise.
#include "ops_seq_v2.h"

void ch_src(ACC<double> &u, const ACC<double> &a) { u(0, 0, 0) = 2.0 * a(0, 0, 0) + 1.0; }

void ch_lap7(const ACC<double> &u, ACC<double> &w, double *s) {
  w(0, 0, 0) = 6.0 * u(0, 0, 0) - (u(-1, 0, 0) + u(1, 0, 0)) - (u(0, -1, 0) + u(0, 1, 0)) - (u(0, 0, -1) + u(0, 0, 1));
  *s = *s + w(0, 0, 0) * w(0, 0, 0);
}

void ch_mix(ACC<double> &u, const ACC<double> &a, const ACC<double> &b) { u(0, 0) = a(0, 0) - 0.5 * b(0, 0); }

void ch_avg(const ACC<double> &u, ACC<double> &a, double *mx) {
  a(0, 0) = 0.25 * (u(1, 0) + u(-1, 0) + u(0, 1) + u(0, -1));
  *mx = fmax(*mx, a(0, 0));
}

void ch_scale(ACC<double> &u, const ACC<double> &v) { u(0, 0) = 3.0 * v(0, 0); }

void ch_diff(const ACC<double> &u, ACC<double> &w) { w(0, 0) = u(1, 0) - u(-1, 0) + u(0, 1) - u(0, -1); }

// u = 2 a + 1 over range r1, then w = A u (7-point) with sum of w^2 over
// range r2 (r2 == r1: one pass)
extern "C" int gen_chain_src_lap(ops_block block, ops_dat u, ops_dat a, ops_dat w, const int *r1, const int *r2,
                                 double *s_out) {
  int p0[] = {0, 0, 0};
  int p7[] = {0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1};
  ops_stencil S0 = ops_decl_stencil(3, 1, p0, "ch0");
  ops_stencil S7 = ops_decl_stencil(3, 7, p7, "ch7");
  ops_reduction h = ops_decl_reduction_handle(sizeof(double), "double", "ch_s");
  int ra[6], rb[6];
  for (int q = 0; q < 6; q++) {
    ra[q] = r1[q];
    rb[q] = r2[q];
  }
  ops_par_loop(ch_src, "ch_src", block, 3, ra, ops_arg_dat(u, 1, S0, "double", OPS_WRITE),
               ops_arg_dat(a, 1, S0, "double", OPS_READ));
  ops_par_loop(ch_lap7, "ch_lap7", block, 3, rb, ops_arg_dat(u, 1, S7, "double", OPS_READ),
               ops_arg_dat(w, 1, S0, "double", OPS_WRITE), ops_arg_reduce(h, 1, "double", OPS_INC));
  ops_reduction_result(h, s_out);
  return 0;
}

// iters of u = a - b/2 then a = average of u's four neighbours, max(a):
// the second loop overwrites a, which the first reads, so the pass writes a
// into a second buffer
extern "C" int gen_chain_mix_avg(ops_block block, ops_dat u, ops_dat a, ops_dat b, int nx, int ny, int iters,
                                 double *mx_out) {
  int p0[] = {0, 0};
  int p5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
  ops_stencil S0 = ops_decl_stencil(2, 1, p0, "ch2_0");
  ops_stencil S5 = ops_decl_stencil(2, 5, p5, "ch2_5");
  ops_reduction h = ops_decl_reduction_handle(sizeof(double), "double", "ch_mx");
  int r[] = {0, nx, 0, ny};
  for (int it = 0; it < iters; it++) {
    ops_par_loop(ch_mix, "ch_mix", block, 2, r, ops_arg_dat(u, 1, S0, "double", OPS_WRITE),
                 ops_arg_dat(a, 1, S0, "double", OPS_READ), ops_arg_dat(b, 1, S0, "double", OPS_READ));
    ops_par_loop(ch_avg, "ch_avg", block, 2, r, ops_arg_dat(u, 1, S5, "double", OPS_READ),
                 ops_arg_dat(a, 1, S0, "double", OPS_WRITE), ops_arg_reduce(h, 1, "double", OPS_MAX));
    ops_reduction_result(h, mx_out + it);
  }
  return 0;
}

// u = 3 v then w = differences of u's neighbours, where the dat the text
// calls p is swapped for another between the two loops (mode 1) or the first
// loop writes the dat it also reads (mode 2): the pair runs apart
extern "C" int gen_chain_refused(ops_block block, ops_dat u, ops_dat v, ops_dat w, ops_dat other, int nx, int ny,
                                 int mode) {
  int p0[] = {0, 0};
  int p5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
  ops_stencil S0 = ops_decl_stencil(2, 1, p0, "chr_0");
  ops_stencil S5 = ops_decl_stencil(2, 5, p5, "chr_5");
  int r[] = {0, nx, 0, ny};
  ops_dat p = u;
  ops_dat src = mode == 2 ? u : v;
  ops_par_loop(ch_scale, "ch_scale", block, 2, r, ops_arg_dat(p, 1, S0, "double", OPS_WRITE),
               ops_arg_dat(src, 1, S0, "double", OPS_READ));
  if (mode == 1) p = other;
  ops_par_loop(ch_diff, "ch_diff", block, 2, r, ops_arg_dat(p, 1, S5, "double", OPS_READ),
               ops_arg_dat(w, 1, S0, "double", OPS_WRITE));
  ops_execute();
  return 0;
}
