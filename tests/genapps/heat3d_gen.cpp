// The TeaLeaf CG operator and residual update in 3D (the 3D form of
// apps/c/TeaLeaf/tea_leaf_cg_kernels.h:8-21, z pair after the x pair),
// written in the OPS dev API and run through generated kernels.
#define OPS_3D
#include "ops_seq_v2.h"

void heat_calc_w(ACC<double> &w, const ACC<double> &Kx, const ACC<double> &Ky, const ACC<double> &Kz,
                 const ACC<double> &p, const double *rx, const double *ry, const double *rz, double *pw) {
  w(0, 0, 0) = (1.0 + (*ry) * (Ky(0, 1, 0) + Ky(0, 0, 0)) + (*rx) * (Kx(1, 0, 0) + Kx(0, 0, 0)) +
                (*rz) * (Kz(0, 0, 1) + Kz(0, 0, 0))) * p(0, 0, 0) -
               (*ry) * (Ky(0, 1, 0) * p(0, 1, 0) + Ky(0, 0, 0) * p(0, -1, 0)) -
               (*rx) * (Kx(1, 0, 0) * p(1, 0, 0) + Kx(0, 0, 0) * p(-1, 0, 0)) -
               (*rz) * (Kz(0, 0, 1) * p(0, 0, 1) + Kz(0, 0, 0) * p(0, 0, -1));
  *pw = *pw + w(0, 0, 0) * p(0, 0, 0);
}

void heat_ur_r(ACC<double> &r, const ACC<double> &w, const double *alpha, double *rrn) {
  r(0, 0, 0) = r(0, 0, 0) - (*alpha) * w(0, 0, 0);
  *rrn = *rrn + r(0, 0, 0) * r(0, 0, 0);
}

// one operator application and one residual update over the interior;
// out[0] = p.w, out[1] = r.r after the update
extern "C" int gen_heat_step(ops_block block, ops_dat w, ops_dat Kx, ops_dat Ky, ops_dat Kz, ops_dat p, ops_dat r,
                             int nx, int ny, int nz, double rx, double ry, double rz, double alpha, double *out) {
  int s0[] = {0, 0, 0};
  ops_stencil S0 = ops_decl_stencil(3, 1, s0, "h0");
  int s7[] = {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
  ops_stencil S7 = ops_decl_stencil(3, 7, s7, "h7");
  int sx[] = {0, 0, 0, 1, 0, 0};
  ops_stencil SX = ops_decl_stencil(3, 2, sx, "hx");
  int sy[] = {0, 0, 0, 0, 1, 0};
  ops_stencil SY = ops_decl_stencil(3, 2, sy, "hy");
  int sz[] = {0, 0, 0, 0, 0, 1};
  ops_stencil SZ = ops_decl_stencil(3, 2, sz, "hz");
  ops_reduction h_pw = ops_decl_reduction_handle(sizeof(double), "double", "pw");
  ops_reduction h_rr = ops_decl_reduction_handle(sizeof(double), "double", "rrn");
  int inner[] = {0, nx, 0, ny, 0, nz};
  ops_par_loop(heat_calc_w, "gen heat_calc_w", block, 3, inner,
               ops_arg_dat(w, 1, S0, "double", OPS_WRITE),
               ops_arg_dat(Kx, 1, SX, "double", OPS_READ),
               ops_arg_dat(Ky, 1, SY, "double", OPS_READ),
               ops_arg_dat(Kz, 1, SZ, "double", OPS_READ),
               ops_arg_dat(p, 1, S7, "double", OPS_READ),
               ops_arg_gbl(&rx, 1, "double", OPS_READ),
               ops_arg_gbl(&ry, 1, "double", OPS_READ),
               ops_arg_gbl(&rz, 1, "double", OPS_READ),
               ops_arg_reduce(h_pw, 1, "double", OPS_INC));
  ops_par_loop(heat_ur_r, "gen heat_ur_r", block, 3, inner,
               ops_arg_dat(r, 1, S0, "double", OPS_RW),
               ops_arg_dat(w, 1, S0, "double", OPS_READ),
               ops_arg_gbl(&alpha, 1, "double", OPS_READ),
               ops_arg_reduce(h_rr, 1, "double", OPS_INC));
  ops_reduction_result(h_pw, &out[0]);
  ops_reduction_result(h_rr, &out[1]);
  return 0;
}
