// A 2D application whose kernels are defined in this file, covering the
// argument kinds of ops_par_loop: a two-component (AoS) dat read with
// ACC(d, dx, dy), an int dat, ops_arg_idx, ops_arg_gbl values, constants
// from ops_decl_const (scalar, int and array), and INC (two components),
// MIN, MAX and int INC reductions.
#define OPS_2D
#include "ops_seq_v2.h"

double shift;
int mult;
double weights[3];

void fill_fields(ACC<double> &v, ACC<int> &m, const int *idx, const double *scale) {
  v(0, 0, 0) = scale[0] * idx[0] + shift;
  v(1, 0, 0) = scale[1] * idx[1] - 0.5 * idx[0];
  m(0, 0) = mult * idx[0] - idx[1];
}

void mix_fields(const ACC<double> &v, ACC<double> &w) {
  w(0, 0) = weights[0] * v(0, 1, 0) + weights[1] * v(1, 0, -1) + weights[2] * (v(0, 0, 1) - v(1, -1, 0));
}

void summary(const ACC<double> &v, const ACC<double> &w, const ACC<int> &m, double *sums, double *lo, double *hi,
             int *isum) {
  sums[0] = sums[0] + v(0, 0, 0);
  sums[1] = sums[1] + w(0, 0) * w(0, 0);
  *lo = fmin(*lo, w(0, 0));
  *hi = fmax(*hi, v(1, 0, 0));
  *isum = *isum + m(0, 0);
}

// v: 2 components, m: int, w: double; all with a one-point halo.  out[0..4]
// = sums[0], sums[1], lo, hi, isum
extern "C" int gen_fields(ops_block block, ops_dat v, ops_dat m, ops_dat w, int nx, int ny, double sx, double sy,
                          double sh, int mu, const double *wts, double *out) {
  shift = sh;
  mult = mu;
  for (int i = 0; i < 3; i++) weights[i] = wts[i];
  ops_decl_const("shift", 1, "double", &shift);
  ops_decl_const("mult", 1, "int", &mult);
  ops_decl_const("weights", 3, "double", weights);
  int s00[] = {0, 0};
  ops_stencil S00 = ops_decl_stencil(2, 1, s00, "f00");
  int s5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
  ops_stencil S5 = ops_decl_stencil(2, 5, s5, "f5");
  ops_reduction h_sums = ops_decl_reduction_handle(2 * sizeof(double), "double", "sums");
  ops_reduction h_lo = ops_decl_reduction_handle(sizeof(double), "double", "lo");
  ops_reduction h_hi = ops_decl_reduction_handle(sizeof(double), "double", "hi");
  ops_reduction h_is = ops_decl_reduction_handle(sizeof(int), "int", "isum");
  int whole[] = {-1, nx + 1, -1, ny + 1};
  int inner[] = {0, nx, 0, ny};
  double scale[2] = {sx, sy};
  ops_par_loop(fill_fields, "fill_fields", block, 2, whole,
               ops_arg_dat(v, 2, S00, "double", OPS_WRITE),
               ops_arg_dat(m, 1, S00, "int", OPS_WRITE),
               ops_arg_idx(),
               ops_arg_gbl(scale, 2, "double", OPS_READ));
  ops_par_loop(mix_fields, "mix_fields", block, 2, inner,
               ops_arg_dat(v, 2, S5, "double", OPS_READ),
               ops_arg_dat(w, 1, S00, "double", OPS_WRITE));
  ops_par_loop(summary, "summary", block, 2, inner,
               ops_arg_dat(v, 2, S00, "double", OPS_READ),
               ops_arg_dat(w, 1, S00, "double", OPS_READ),
               ops_arg_dat(m, 1, S00, "int", OPS_READ),
               ops_arg_reduce(h_sums, 2, "double", OPS_INC),
               ops_arg_reduce(h_lo, 1, "double", OPS_MIN),
               ops_arg_reduce(h_hi, 1, "double", OPS_MAX),
               ops_arg_reduce(h_is, 1, "int", OPS_INC));
  double sums[2], lo, hi;
  int is;
  ops_reduction_result(h_sums, sums);
  ops_reduction_result(h_lo, &lo);
  ops_reduction_result(h_hi, &hi);
  ops_reduction_result(h_is, &is);
  out[0] = sums[0];
  out[1] = sums[1];
  out[2] = lo;
  out[3] = hi;
  out[4] = is;
  return 0;
}
