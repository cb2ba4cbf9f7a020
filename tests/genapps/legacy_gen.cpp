// Kernels in the legacy pointer form (T * parameters indexed with the
// OPS_ACCn / OPS_ACC_MDn macros, ops/c/include/ops_macros.h), each beside the
// same kernel written with ACC<T> references.  The exported functions run
// both forms on the same dats so a test can compare them bit for bit:
//   gen_legacy_lap2d:  5-point average in 2D, then (legacy) copy back and a
//                      second average, a pair the queue fuses
//   gen_legacy_md3d:   a two-component 3D dat read through OPS_ACC_MDn at
//                      neighbours, with a sum reduction weighted by the index
#include "ops_seq.h"

void leg_lap(const double *a, double *b) {
  b[OPS_ACC1(0, 0)] = 0.25 * (a[OPS_ACC0(1, 0)] + a[OPS_ACC0(-1, 0)] + a[OPS_ACC0(0, 1)] + a[OPS_ACC0(0, -1)]);
}

void acc_lap(const ACC<double> &a, ACC<double> &b) {
  b(0, 0) = 0.25 * (a(1, 0) + a(-1, 0) + a(0, 1) + a(0, -1));
}

void leg_copy(double *a, const double *b) {
  a[OPS_ACC0(0, 0)] = b[OPS_ACC1(0, 0)];
}

void acc_copy(ACC<double> &a, const ACC<double> &b) {
  a(0, 0) = b(0, 0);
}

void leg_md(const double *v, double *w, double *s, const int *idx) {
  w[OPS_ACC1(0, 0, 0)] = v[OPS_ACC_MD0(0, 1, 0, 0)] - v[OPS_ACC_MD0(1, 0, 0, -1)] +
                         0.5 * (v[OPS_ACC_MD0(1, 0, 1, 0)] - v[OPS_ACC_MD0(0, 0, -1, 0)]);
  *s = *s + w[OPS_ACC1(0, 0, 0)] * (idx[0] + 2 * idx[1] + 3 * idx[2]);
}

void acc_md(const ACC<double> &v, ACC<double> &w, double *s, const int *idx) {
  w(0, 0, 0) = v(0, 1, 0, 0) - v(1, 0, 0, -1) + 0.5 * (v(1, 0, 1, 0) - v(0, 0, -1, 0));
  *s = *s + w(0, 0, 0) * (idx[0] + 2 * idx[1] + 3 * idx[2]);
}

// a, b: 2D dats of nx * ny points with a one-point halo.  mode 0: legacy
// kernels, mode 1: ACC kernels.  Each runs b = lap(a); a = b; b = lap(a).
extern "C" int gen_legacy_lap2d(ops_block block, ops_dat a, ops_dat b, int nx, int ny, int mode) {
  int s0[] = {0, 0};
  ops_stencil S0 = ops_decl_stencil(2, 1, s0, "l00");
  int s5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
  ops_stencil S5 = ops_decl_stencil(2, 5, s5, "l5pt");
  int r[] = {0, nx, 0, ny};
  if (mode == 0) {
    ops_par_loop(leg_lap, "leg_lap", block, 2, r, ops_arg_dat(a, 1, S5, "double", OPS_READ),
                 ops_arg_dat(b, 1, S0, "double", OPS_WRITE));
    ops_par_loop(leg_copy, "leg_copy", block, 2, r, ops_arg_dat(a, 1, S0, "double", OPS_WRITE),
                 ops_arg_dat(b, 1, S0, "double", OPS_READ));
    ops_par_loop(leg_lap, "leg_lap", block, 2, r, ops_arg_dat(a, 1, S5, "double", OPS_READ),
                 ops_arg_dat(b, 1, S0, "double", OPS_WRITE));
  } else {
    ops_par_loop(acc_lap, "acc_lap", block, 2, r, ops_arg_dat(a, 1, S5, "double", OPS_READ),
                 ops_arg_dat(b, 1, S0, "double", OPS_WRITE));
    ops_par_loop(acc_copy, "acc_copy", block, 2, r, ops_arg_dat(a, 1, S0, "double", OPS_WRITE),
                 ops_arg_dat(b, 1, S0, "double", OPS_READ));
    ops_par_loop(acc_lap, "acc_lap", block, 2, r, ops_arg_dat(a, 1, S5, "double", OPS_READ),
                 ops_arg_dat(b, 1, S0, "double", OPS_WRITE));
  }
  ops_execute();
  return 0;
}

// v: 3D two-component dat, w: one component, both nx * ny * nz with a
// one-point halo; *s receives the reduction
extern "C" int gen_legacy_md3d(ops_block block, ops_dat v, ops_dat w, int nx, int ny, int nz, int mode,
                               double *s) {
  int s0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, s0, "m000");
  int s4[] = {1, 0, 0, 0, 0, -1, 0, 1, 0, 0, -1, 0};
  ops_stencil S4 = ops_decl_stencil(3, 4, s4, "m4");
  ops_reduction h = ops_decl_reduction_handle(sizeof(double), "double", "legsum");
  int r[] = {0, nx, 0, ny, 0, nz};
  if (mode == 0)
    ops_par_loop(leg_md, "leg_md", block, 3, r, ops_arg_dat(v, 2, S4, "double", OPS_READ),
                 ops_arg_dat(w, 1, S0, "double", OPS_WRITE), ops_arg_reduce(h, 1, "double", OPS_INC),
                 ops_arg_idx());
  else
    ops_par_loop(acc_md, "acc_md", block, 3, r, ops_arg_dat(v, 2, S4, "double", OPS_READ),
                 ops_arg_dat(w, 1, S0, "double", OPS_WRITE), ops_arg_reduce(h, 1, "double", OPS_INC),
                 ops_arg_idx());
  ops_reduction_result(h, s);
  return 0;
}
