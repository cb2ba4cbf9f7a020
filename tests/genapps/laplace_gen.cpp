// 2D Laplace Jacobi written against the OPS dev API: ops_par_loop(kernel,
// ...) with the kernels in laplace_gen_kernels.h.  The translator turns it
// into laplace_gen_ops.cpp + generated kernels; gen_laplace runs `iters`
// iterations on dats the caller declared and returns the per-iteration
// max-change history.
#define OPS_2D
#include "ops_seq_v2.h"
#include "laplace_gen_kernels.h"

extern "C" int gen_laplace(ops_block block, ops_dat A, ops_dat Anew, int imax, int jmax, int iters, double *hist) {
  int s2d_00[] = {0, 0};
  ops_stencil S2D_00 = ops_decl_stencil(2, 1, s2d_00, "gen_00");
  int s2d_5pt[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
  ops_stencil S2D_5pt = ops_decl_stencil(2, 5, s2d_5pt, "gen_5pt");
  ops_reduction h_err = ops_decl_reduction_handle(sizeof(double), "double", "gen_error");
  int interior[] = {0, imax, 0, jmax};
  for (int it = 0; it < iters; it++) {
    ops_par_loop(apply_stencil, "gen apply_stencil", block, 2, interior,
                 ops_arg_dat(A, 1, S2D_5pt, "double", OPS_READ),
                 ops_arg_dat(Anew, 1, S2D_00, "double", OPS_WRITE),
                 ops_arg_reduce(h_err, 1, "double", OPS_MAX));
    ops_par_loop(copy, "gen copy", block, 2, interior,
                 ops_arg_dat(A, 1, S2D_00, "double", OPS_WRITE),
                 ops_arg_dat(Anew, 1, S2D_00, "double", OPS_READ));
    double err = 0.0;
    ops_reduction_result(h_err, &err);
    hist[it] = err;
  }
  return 0;
}

// apply_stencil with a one-point stencil on A: the loop is refused before
// it runs (the kernel reads four neighbours the stencil does not hold)
extern "C" int gen_laplace_bad_stencil(ops_block block, ops_dat A, ops_dat Anew, int imax, int jmax) {
  int s2d_00[] = {0, 0};
  ops_stencil S = ops_decl_stencil(2, 1, s2d_00, "gen_bad_00");
  ops_reduction h = ops_decl_reduction_handle(sizeof(double), "double", "gen_bad_error");
  int interior[] = {0, imax, 0, jmax};
  ops_par_loop(apply_stencil, "gen bad apply_stencil", block, 2, interior,
               ops_arg_dat(A, 1, S, "double", OPS_READ),
               ops_arg_dat(Anew, 1, S, "double", OPS_WRITE),
               ops_arg_reduce(h, 1, "double", OPS_MAX));
  return 0;
}
