// Low-dimensional (edge) dats through the translator: dats declared with
// size 1 in some dimensions, read and accumulated into by loops over the
// full block through strided stencils (stride 0 where the dat has size 1),
// with the reference's single-node rule for accumulations
// (ops/c/src/core/ops_dummy_singlenode.cpp:204-391): an OPS_INC/MIN/MAX
// loop after a different access to the dat first resets it to the
// identity; consecutive accumulations of the same kind add up.
#include "ops_seq_v2.h"

void ld_fill(ACC<double> &f, const int *idx) {
  f(0, 0, 0) = (double)((idx[0] * 7 + idx[1] * 3 + idx[2] * 11) % 17) - 5.0;
}

void ld_colsum(const ACC<double> &f, ACC<double> &c) { c.combine_inc(0, 0, 0, f(0, 0, 0)); }

void ld_halve(ACC<double> &c) { c(0, 0, 0) = 0.5 * c(0, 0, 0); }

void ld_extrema(const ACC<double> &f, ACC<double> &rmax, ACC<double> &zmin, ACC<int> &cnt, ACC<float> &ymax) {
  rmax.combine_max(0, 0, 0, f(0, 0, 0));
  zmin.combine_min(0, 0, 0, f(0, 0, 0));
  cnt.combine_inc(0, 0, 0, 1);
  ymax.combine_max(0, 0, 0, (float)f(0, 0, 0) + 0.5f);
}

void ld_gather(const ACC<double> &f, const ACC<double> &c, const ACC<double> &rmax, const ACC<double> &zmin,
               const ACC<int> &cnt, ACC<double> &out) {
  out(0, 0, 0) = f(0, 0, 0) + 2.0 * c(0, 0, 0) + 4.0 * rmax(0, 0, 0) + 8.0 * zmin(0, 0, 0) + 16.0 * cnt(0, 0, 0);
}

void ld_col2d(const ACC<double> &g, ACC<double> &m) { m.combine_min(0, 0, g(0, 0)); }

void ld_fill2d(ACC<double> &g, const int *idx) { g(0, 0) = (double)((idx[0] * 5 + idx[1] * 9) % 13) - 6.0; }

void ld_write_edge(ACC<double> &c, const double *v) { c(0, 0, 0) = *v; }

static ops_stencil strided(int sx, int sy, int sz, const char *name) {
  int pt[] = {0, 0, 0};
  int st[] = {sx, sy, sz};
  return ops_decl_strided_stencil(3, 1, pt, st, name);
}

// 3D block of nx x ny x nz.  colsum_steps: four nx*ny snapshots of the
// column sums dat (size nx x ny x 1) after INC, INC again, RW (halve), INC
// after the RW; rowmax: nx (size nx x 1 x 1, MAX over y, z); zmin: nz (size
// 1 x 1 x nz, MIN over x, y); cnt: ny ints (1 x ny x 1, INC of 1); ymax: ny
// floats (1 x ny x 1, MAX); out: nx*ny*nz, a 3D loop reading every edge dat.
extern "C" int gen_lowdim3d(ops_block block, int nx, int ny, int nz, double *colsum_steps, double *rowmax,
                            double *zmin, int *cnt, float *ymax, double *out) {
  int d_m[] = {-1, -1, -1}, d_p[] = {1, 1, 1}, base[] = {0, 0, 0};
  int size[] = {nx, ny, nz};
  double *none = NULL;
  ops_dat f = ops_decl_dat(block, 1, size, base, d_m, d_p, none, "double", "ld_f");
  ops_dat o = ops_decl_dat(block, 1, size, base, d_m, d_p, none, "double", "ld_out");
  // column sums, declared with a non-zero initial value: the first
  // accumulation must start from 0 regardless
  int sc[] = {nx, ny, 1}, mc[] = {-1, -1, 0}, pc[] = {1, 1, 0};
  double *init = (double *)malloc(sizeof(double) * (nx + 2) * (ny + 2));
  for (int i = 0; i < (nx + 2) * (ny + 2); i++) init[i] = 1000.0;
  ops_dat c = ops_decl_dat(block, 1, sc, base, mc, pc, init, "double", "ld_colsum");
  free(init);
  int sr[] = {nx, 1, 1}, mr[] = {-1, 0, 0}, pr[] = {1, 0, 0};
  ops_dat rmax = ops_decl_dat(block, 1, sr, base, mr, pr, none, "double", "ld_rmax");
  int sz_[] = {1, 1, nz}, mz[] = {0, 0, -1}, pz[] = {0, 0, 1};
  ops_dat zm = ops_decl_dat(block, 1, sz_, base, mz, pz, none, "double", "ld_zmin");
  int sy[] = {1, ny, 1}, my[] = {0, -1, 0}, py[] = {0, 1, 0};
  int *inone = NULL;
  float *fnone = NULL;
  ops_dat cn = ops_decl_dat(block, 1, sy, base, my, py, inone, "int", "ld_cnt");
  ops_dat ym = ops_decl_dat(block, 1, sy, base, my, py, fnone, "float", "ld_ymax");

  int p0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, p0, "ld0");
  ops_stencil Sxy = strided(1, 1, 0, "ld_xy"), Sx = strided(1, 0, 0, "ld_x"), Sz = strided(0, 0, 1, "ld_z"),
              Sy = strided(0, 1, 0, "ld_y");
  int all[] = {0, nx, 0, ny, 0, nz};
  int plane[] = {0, nx, 0, ny, 0, 1};
  ops_par_loop(ld_fill, "ld_fill", block, 3, all, ops_arg_dat(f, 1, S0, "double", OPS_WRITE), ops_arg_idx());
  for (int step = 0; step < 4; step++) {
    if (step == 2)
      ops_par_loop(ld_halve, "ld_halve", block, 3, plane, ops_arg_dat(c, 1, Sxy, "double", OPS_RW));
    else
      ops_par_loop(ld_colsum, "ld_colsum", block, 3, all, ops_arg_dat(f, 1, S0, "double", OPS_READ),
                   ops_arg_dat(c, 1, Sxy, "double", OPS_INC));
    ops_dat_fetch_data(c, 0, (char *)(colsum_steps + (size_t)step * nx * ny));
  }
  ops_par_loop(ld_extrema, "ld_extrema", block, 3, all, ops_arg_dat(f, 1, S0, "double", OPS_READ),
               ops_arg_dat(rmax, 1, Sx, "double", OPS_MAX), ops_arg_dat(zm, 1, Sz, "double", OPS_MIN),
               ops_arg_dat(cn, 1, Sy, "int", OPS_INC), ops_arg_dat(ym, 1, Sy, "float", OPS_MAX));
  ops_par_loop(ld_gather, "ld_gather", block, 3, all, ops_arg_dat(f, 1, S0, "double", OPS_READ),
               ops_arg_dat(c, 1, Sxy, "double", OPS_READ), ops_arg_dat(rmax, 1, Sx, "double", OPS_READ),
               ops_arg_dat(zm, 1, Sz, "double", OPS_READ), ops_arg_dat(cn, 1, Sy, "int", OPS_READ),
               ops_arg_dat(o, 1, S0, "double", OPS_WRITE));
  ops_dat_fetch_data(rmax, 0, (char *)rowmax);
  ops_dat_fetch_data(zm, 0, (char *)zmin);
  ops_dat_fetch_data(cn, 0, (char *)cnt);
  ops_dat_fetch_data(ym, 0, (char *)ymax);
  ops_dat_fetch_data(o, 0, (char *)out);
  for (ops_dat d : {f, o, c, rmax, zm, cn, ym}) ops_free_dat(d);
  return 0;
}

// 2D block nx x ny: g filled from the indices, then, for each x, the MIN
// over y into a dat of size nx x 1
extern "C" int gen_lowdim2d(ops_block block, int nx, int ny, double *colmin) {
  int d_m[] = {-1, -1}, d_p[] = {1, 1}, base[] = {0, 0};
  int size[] = {nx, ny};
  double *none = NULL;
  ops_dat g = ops_decl_dat(block, 1, size, base, d_m, d_p, none, "double", "ld2_g");
  int sm[] = {nx, 1}, mm[] = {-1, 0}, pm[] = {1, 0};
  ops_dat m = ops_decl_dat(block, 1, sm, base, mm, pm, none, "double", "ld2_m");
  int p0[] = {0, 0}, st[] = {1, 0};
  ops_stencil S0 = ops_decl_stencil(2, 1, p0, "ld2_0");
  ops_stencil Sx = ops_decl_strided_stencil(2, 1, p0, st, "ld2_x");
  int all[] = {0, nx, 0, ny};
  ops_par_loop(ld_fill2d, "ld_fill2d", block, 2, all, ops_arg_dat(g, 1, S0, "double", OPS_WRITE), ops_arg_idx());
  ops_par_loop(ld_col2d, "ld_col2d", block, 2, all, ops_arg_dat(g, 1, S0, "double", OPS_READ),
               ops_arg_dat(m, 1, Sx, "double", OPS_MIN));
  ops_dat_fetch_data(m, 0, (char *)colmin);
  ops_free_dat(g);
  ops_free_dat(m);
  return 0;
}

// An OPS_WRITE through a stride-0 stencil over a range that puts several
// points on one element: refused (the caller runs with error mode 1 and
// reads the error).  mode 1 instead uses a stride of 2, also refused.
extern "C" int gen_lowdim_bad(ops_block block, int nx, int ny, int nz, int mode) {
  int sc[] = {nx, ny, 1}, mc[] = {-1, -1, 0}, pc[] = {1, 1, 0}, base[] = {0, 0, 0};
  double *none = NULL;
  ops_dat c = ops_decl_dat(block, 1, sc, base, mc, pc, none, "double", "ld_bad");
  ops_stencil S = mode == 0 ? strided(1, 1, 0, "ld_bad_xy") : strided(1, 1, 2, "ld_bad_2");
  int all[] = {0, nx, 0, ny, 0, nz};
  double v = 3.0;
  ops_par_loop(ld_write_edge, "ld_write_edge", block, 3, all, ops_arg_dat(c, 1, S, "double", OPS_WRITE),
               ops_arg_gbl(&v, 1, "double", OPS_READ));
  ops_free_dat(c);
  return 0;
}
