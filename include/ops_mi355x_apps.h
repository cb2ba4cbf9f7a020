/*
 * ops_mi355x_apps.h — application drivers built on the OPS C ABI
 * (libopsmi_apps.so).  They play the role of the reference's example apps and
 * have no counterpart in the OPS library itself; each cites the reference app
 * it follows.  Inputs are generated from the problem definitions in
 * ops_amd/apps/problems.h when the caller passes NULL arrays.
 */
#ifndef OPS_MI355X_APPS_H
#define OPS_MI355X_APPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TeaLeaf-form implicit heat CG (apps/c/TeaLeaf/tea_leaf_common.cpp:76-146,
 * tea_leaf_cg.cpp:54-214, tea_solve.cpp:300-325), 2D or 3D.  K_ext holds the
 * conductivity over the halo-extended box ((n+2) points per side), rhs the
 * interior right-hand side.  create runs the setup loops and r0.p0. */
void *opsapp_heat_create(int ndim, int nx, int ny, int nz, uint64_t seed, double rx, double ry, double rz,
                         const double *K_ext, const double *rhs);
/* up to max_iters CG iterations; stops once sqrt(r.r) <= tol sqrt(r0.r0)
 * (tol <= 0: never).  Histories (may be NULL) get one entry per iteration;
 * rrn_h[i] is r.r after iteration i.  Returns iterations run. */
int opsapp_heat_iterate(void *ctx, int max_iters, double tol, double *pw_h, double *rrn_h, double *alpha_h,
                        double *beta_h);
double opsapp_heat_rr(void *ctx, int initial);
void opsapp_heat_fetch_u(void *ctx, double *out);
void opsapp_heat_destroy(void *ctx);

/* constant-coefficient Poisson CG (5-point 2D / 7-point 3D, zero Dirichlet) */
void *opsapp_poisson_create(int ndim, int nx, int ny, int nz, uint64_t seed, const double *f);
int opsapp_poisson_iterate(void *ctx, int max_iters, double tol, double *rrn_h, double *alpha_h, double *beta_h);
void opsapp_poisson_fetch_x(void *ctx, double *out);
void opsapp_poisson_destroy(void *ctx);

/* mixed-precision Poisson: FP64 iterative refinement (residual recomputed in
 * FP64, correction x += e in FP64) around an FP32 Jacobi-preconditioned CG
 * with FP64 dot accumulation.  solve stops once |f - A x| <= tol |f| or after
 * max_outer corrections; each correction runs the inner CG to
 * max(inner_tol, tol / (2 relres)) relative, or inner_iters.  relres_h gets |f - A x|/|f| before each
 * correction and after the last (outer + 1 entries).  Returns outer
 * iterations; *inner_total gets the inner iterations summed. */
void *opsapp_pmixed_create(int ndim, int nx, int ny, int nz, uint64_t seed, const double *f);
int opsapp_pmixed_solve(void *ctx, double tol, int max_outer, int inner_iters, double inner_tol,
                        double *relres_h, int *inner_total);
void opsapp_pmixed_fetch_x(void *ctx, double *out);
void opsapp_pmixed_destroy(void *ctx);

/* Laplace Jacobi (apps/c/laplace2d_tutorial/step7/laplace2d.cpp:37-139):
 * mode 0 top-edge boundary, 1 the tutorial's left/right boundary; init_ext
 * (halo-including) replaces the boundary loops when given.  Returns the last
 * max-abs-diff. */
double opsapp_laplace2d(int nx, int ny, int iters, int mode, const double *init_ext, double *err_hist,
                        double *A_out);

/* batched tridiagonal ADI sweeps (apps/c/adi/adi.cpp:344-424 with random
 * coefficients); strategy 3 selects PCR */
void *opsapp_trid_create(int nx, int ny, int nz, uint64_t seed, int strategy);
void opsapp_trid_sweep(void *ctx, int inc);
void opsapp_trid_solve(void *ctx, int solvedim, int inc);
void opsapp_trid_fetch(void *ctx, double *d_out, double *u_out);
void opsapp_trid_destroy(void *ctx);

/* the reference's multi-block Poisson test app (apps/c/poisson/poisson.cpp)
 * on an nx x ny grid split into ngx x ngy blocks joined by halo transfers;
 * returns the total squared error (the app's printed "Total error"); u_out
 * (may be NULL) gets the assembled interior, x fastest */
double opsapp_mb_poisson(int nx, int ny, int ngx, int ngy, int iters, double *u_out);

/* request stream on the batched engine (include/ops_mi355x_serve.h), on the
 * wall clock: request i (kind[i], n[i], seed[i]) arrives arrival_s[i]
 * seconds after the start (ascending) and is submitted once it has arrived;
 * done_s[i] gets the time of the scheduler step that reported it converged,
 * iters[i] its iteration count (negative: stopped without converging).
 * autotune != 0 runs the tile-shape autotuner before the clock starts;
 * mixed != 0 solves every request in the FP32-inner refinement mode.
 * Returns the time from start to the last completion, or -1 on error. */
double opsapp_serve_stream(int nreq, const int *kind, const int *n, const uint64_t *seed, const double *arrival_s,
                           double tol, size_t budget, int max_slots, int iters_per_launch, int autotune, int mixed,
                           double *done_s, int *iters);

#ifdef __cplusplus
}
#endif

#endif
