/* ORACLE — test infrastructure only.  A CPU restatement of the reference OPS
 * sequential/OpenMP path for this project's hot-path loops.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the CPU baseline, never as the measured product.
 *
 * Array convention ("ext" arrays): every field carries a 1-cell halo on both
 * sides of each of its ndim dimensions, x fastest:
 *   2D  (i,j)   -> (i+1) + (nx+2)*(j+1)
 *   3D  (i,j,k) -> (i+1) + (nx+2)*((j+1) + (ny+2)*(k+1))
 * "int" arrays are interior-only, x fastest.
 *
 * nthreads == 1 reproduces the reference dev-path order exactly
 * (ops/c/include/ops_seq_v2.h:272-293: lexicographic, x innermost, sums
 * accumulated into the handle from its initial value).  nthreads > 1 follows
 * the shape of the generated OpenMP loop (ops_translator/resources/templates/
 * cpp/mpi_openmp/loop_host.cpp.j2:124-341: parallel over the outer
 * dimensions, OpenMP reduction clauses), so sums change order.
 *
 * Parity status of each function (what pins it) is listed beside it.
 */
#ifndef OPS_ORACLE_H
#define OPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- inputs (pinned: uniform part is integer arithmetic; fixtures) ---- */
void orc_gen_u01(uint64_t seed, uint64_t stream, int64_t n, double *out);
void orc_gen_normal(uint64_t seed, uint64_t stream, int64_t n, double *out);
/* mode 0: top-edge sin(); mode 1: left/right columns (reference tutorial) */
void orc_gen_laplace_init(int nx, int ny, int mode, double *u_ext);
/* K over the ext range, rhs over the interior; ndim 2 ignores nz */
void orc_gen_heat(uint64_t seed, int ndim, int nx, int ny, int nz, double *K_ext, double *rhs_int);
void orc_gen_poisson_rhs(uint64_t seed, int ndim, int nx, int ny, int nz, double *f_int);
void orc_gen_trid(uint64_t seed, int s, int nx, int ny, int nz, double *a, double *b, double *c);
void orc_gen_trid_rhs(uint64_t seed, int nx, int ny, int nz, double *d);

/* ---- Laplace Jacobi (apps/c/laplace2d_tutorial/step7/laplace2d.cpp:112-130,
 *      laplace_kernels.h:8-24) — pinned: bit-exact vs libops_seq fixtures
 *      and the tutorial app's known answer.  Returns the last error. ---- */
double orc_laplace_run(double *A_ext, double *Anew_ext, int nx, int ny, int iters, int nthreads,
                       double *err_hist);
float orc_laplace_run_f32(float *A_ext, float *Anew_ext, int nx, int ny, int iters, int nthreads,
                          float *err_hist);

/* ---- TeaLeaf-form heat CG (apps/c/TeaLeaf/tea_leaf_common.cpp:76-146,
 *      tea_leaf_cg.cpp:54-214, tea_solve.cpp:300-325; 3D operator appends
 *      z after x) — pinned: libops_seq fixtures (2D and 3D).
 *      Runs at most max_iters iterations, stopping when
 *      sqrt(rrn) <= tol * sqrt(rro_0) (tol <= 0: never).  Histories hold
 *      one entry per iteration; rrn_hist has iters+1 (entry 0 = r0.r0). ---- */
int orc_heat_cg(int ndim, int nx, int ny, int nz, const double *K_ext, const double *rhs_int, double rx,
                double ry, double rz, int max_iters, double tol, int nthreads, double *u_int, double *pw_hist,
                double *rrn_hist, double *alpha_hist, double *beta_hist);
/* seconds spent in the iteration loop of the last orc_heat_cg call */
double orc_last_loop_seconds(void);
/* face coefficients only (tea_leaf_common_init_Kx_Ky_kernel + zeroed
 * boundary faces), interior-sized outputs */
void orc_heat_coefs(int ndim, int nx, int ny, int nz, const double *K_ext, double *Kx_ext, double *Ky_ext,
                    double *Kz_ext);
/* one application w = A p (+ p.w) of the heat operator on ext arrays */
double orc_heat_apply(int ndim, int nx, int ny, int nz, const double *Kx_ext, const double *Ky_ext,
                      const double *Kz_ext, const double *p_ext, double rx, double ry, double rz, double *w_ext,
                      int nthreads);

/* ---- constant-coefficient Poisson CG, 5/7-point, zero Dirichlet halo —
 *      pinned: libops_seq fixtures (2D and 3D). ---- */
int orc_poisson_cg(int ndim, int nx, int ny, int nz, const double *f_int, int max_iters, double tol,
                   int nthreads, double *x_int, double *rrn_hist, double *alpha_hist, double *beta_hist);
/* one FP32 sweep w = A p with an FP64 p.w accumulation — pinned: fixture */
double orc_poisson_apply_f32(int ndim, int nx, int ny, int nz, const float *p_ext, float *w_ext);
/* FP64 residual r = f - A x, returns r.r — pinned: closed form on the
 * operator eigenvectors (tests) */
double orc_poisson_residual(int ndim, int nx, int ny, int nz, const double *f_int, const double *x_ext,
                            double *r_ext, int nthreads);

/* ---- mixed precision: FP64 iterative refinement around an FP32
 *      Jacobi-preconditioned CG (this project's algorithm; no reference
 *      counterpart) — pinned: converged solution vs orc_poisson_cg's and the
 *      FP64 residual recomputed independently in tests.
 *      Returns outer iterations; *inner_total = total inner iterations. ---- */
int orc_poisson_mixed(int ndim, int nx, int ny, int nz, const double *f_int, double tol, int max_outer,
                      int inner_iters, double inner_tol, int nthreads, double *x_int, double *relres_hist,
                      int *inner_total);
/* ---- FP64 refinement around FP32 CG for the 3D heat problem (the batched
 *      engine's FP32-inner heat mode; no reference counterpart): x0 = rhs
 *      (TeaLeaf's start), f = rhs, FP64 residual, stop at |r| <= tol |r_0|;
 *      inner unpreconditioned CG in FP32 with float face coefficients, FP64
 *      dots, inner stop at max(inner_tol, tol/(2 relres)) |rr_0|.  Pinned:
 *      converged solution vs orc_heat_cg's and the FP64 residual recomputed
 *      independently in tests.  Returns outer iterations. ---- */
int orc_heat_mixed(int nx, int ny, int nz, const double *K_ext, const double *rhs_int, double tol, int max_outer,
                   int inner_iters, double inner_tol, int nthreads, double *x_int, double *relres_hist,
                   int *inner_total);

/* ---- ADI preproc (apps/c/adi/preproc_kernel.h:4-39, lambda) — pinned:
 *      fixture.  Outputs interior-sized. ---- */
void orc_adi_preproc(int nx, int ny, int nz, double lambda, const double *u_ext, double *du, double *ax,
                     double *bx, double *cx);
/* 27-point weighted box sweep — pinned: fixture */
void orc_stencil27(int nx, int ny, int nz, const double *u_ext, double *out_int);

/* ---- the reference's multi-block Poisson test app
 *      (apps/c/poisson/poisson.cpp:56-307, poisson_kernel.h:4-38) on one
 *      block: populate over the halo-extended box, zero initial guess,
 *      `iters` Jacobi sweeps + copies, returns the total squared error
 *      against sin(pi x) cos(2 pi y) — pinned: the app's own known answer
 *      (16.5173570176708 at 20 x 20, 10 iterations) and runs of the
 *      reference app built from its sources at other sizes (fixtures).
 *      u_int (may be NULL) gets the interior. ---- */
double orc_mb_poisson(int nx, int ny, int iters, double *u_int);

/* ---- batched tridiagonal, semantics of tridDmtsvStridedBatch as called at
 *      ops/c/src/tridiag/ops_tridiag.cpp:126-129: systems along solvedim of
 *      extents dims[], element strides from pads[] (the dat's padded sizes),
 *      d overwritten with the solution; _inc adds it into u and leaves d.
 *      Thomas in reciprocal form (Conte & de Boor 4.3).  libtrid is not
 *      vendored: parity unpinned at that boundary; pinned here by a dense
 *      long-double elimination and residual checks in tests. ---- */
void orc_thomas_batch(int ndim, int solvedim, const int *dims, const int *pads, const double *a,
                      const double *b, const double *c, double *d, double *u, int nthreads);
void orc_thomas_batch_f32(int ndim, int solvedim, const int *dims, const int *pads, const float *a,
                          const float *b, const float *c, float *d, float *u, int nthreads);
/* parallel cyclic reduction (Hockney 1965), same interface — pinned vs
 * Thomas (1e-13 relative) and the dense solve */
void orc_pcr_batch(int ndim, int solvedim, const int *dims, const int *pads, const double *a, const double *b,
                   const double *c, double *d, double *u, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
