/*
 * ops_mi355x_serve.h — batched multi-block solve engine (part of libopsmi.so).
 *
 * Many independent blocks of uneven sizes (the OPS multi-block model: one
 * ops_decl_block per problem, ops/c/src/core/ops_lib_core.cpp:376-434) are
 * solved together with unpreconditioned CG (the loop body of
 * apps/c/TeaLeaf/tea_leaf_cg.cpp:129-214: calc_w + p.w, u += alpha p and
 * r -= alpha w + r.r, p = beta p + r).  Every loop of the iteration is one
 * persistent launch over the tiles of all active blocks; the CG scalars stay
 * on the device, so iterations run without host round trips and are captured
 * in a HIP graph.  A continuous-batching scheduler admits queued solves
 * whenever a slot and HBM budget are free, between graph launches, while
 * the others keep iterating.  No counterpart in the reference library (its
 * multi-block apps run blocks one loop at a time); the interface follows the
 * library's C conventions (errors through ops_mi_last_error, handles as
 * opaque pointers).
 *
 * Problems (inputs generated on the device from ops_amd/apps/problems.h):
 *   OPSMI_SOLVE_POISSON2D  2D 5-point constant-coefficient Poisson, n x n,
 *                          rhs 2U-1 (stream 3), zero Dirichlet halo, x0 = 0
 *   OPSMI_SOLVE_HEAT3D     3D TeaLeaf-form heat, n^3, K = exp(N(0,1)),
 *                          u0 = rhs, rx = ry = rz = 1 (as opsapp_heat_create)
 * Stop: sqrt(r.r) <= tol sqrt(r0.r0), or max_iters.  Results do not depend
 * on what else is in the batch: per-tile partial sums are folded per block
 * in a fixed order.
 */
#ifndef OPS_MI355X_SERVE_H
#define OPS_MI355X_SERVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPSMI_SOLVE_POISSON2D 0
#define OPSMI_SOLVE_HEAT3D 1

typedef struct opsmi_server_s *opsmi_server;

typedef struct {
  int kind;      /* OPSMI_SOLVE_* */
  int n;         /* grid side */
  uint64_t seed; /* input generator seed */
  double tol;    /* relative residual target */
  int max_iters; /* CG iterations (mixed: inner iterations in total) */
  int mixed;     /* 1: FP32-inner refinement: FP64 x, f and residual, CG on
                    FP32 fields with FP64 dots (2D Poisson: Jacobi
                    preconditioned; 3D heat: unpreconditioned, with the face
                    coefficients rounded to float), inner stop
                    max(1e-6, tol / (2 relres)) of the inner start, at most
                    30 corrections; stop at |r| <= tol |r_0| (the 3D heat
                    solve starts from x = rhs, as TeaLeaf's) */
} opsmi_solve_req;

typedef struct {
  int id;
  int status; /* 1 converged, 2 stopped at max_iters (or 30 corrections) */
  int iters;  /* CG iterations (mixed: inner iterations) */
  int outer;  /* mixed: corrections applied */
  double rr0, rr; /* r.r at start and at the end (mixed: true FP64 residual) */
} opsmi_solve_done;

/* hbm_budget: bytes of device memory for solve fields (one arena);
 * max_slots: solves iterating at once; iters_per_launch: CG iterations per
 * scheduler step (one graph launch); use_graph 0 launches the kernels
 * directly (same results). */
opsmi_server opsmi_serve_create(size_t hbm_budget, int max_slots, int iters_per_launch, int use_graph);
/* queues a solve; returns its id (>= 0), or -1 on error (too large for the
 * budget, bad kind/size) */
int opsmi_serve_submit(opsmi_server s, const opsmi_solve_req *req);
/* admits queued solves (FIFO) that fit, runs iters_per_launch iterations of
 * every active solve, retires converged ones; returns active + queued */
int opsmi_serve_step(opsmi_server s);
/* moves up to cap completions (oldest first) into out; returns the count */
int opsmi_serve_completed(opsmi_server s, opsmi_solve_done *out, int cap);
/* copies a completed solve's solution (interior, x fastest) to out and
 * frees its memory; release frees without copying */
int opsmi_serve_fetch(opsmi_server s, int id, double *out);
int opsmi_serve_release(opsmi_server s, int id);
/* bytes of the budget in use, active solves, queued solves */
size_t opsmi_serve_bytes_in_use(opsmi_server s);
int opsmi_serve_active(opsmi_server s);
int opsmi_serve_queued(opsmi_server s);
/* Tile-shape autotuner (server idle): for each kind and size class (3D
 * heat n < 64, < 128, >= 128; 2D Poisson n < 256, < 1024, >= 1024) times
 * `iters` CG iterations of a representative solve with every candidate
 * (3D: 4, 8, 16, 32 z planes per 128 x 4 column; 2D: 1, 2, 4, 8 rows per
 * wave of a 128-wide tile) and keeps the fastest.  Results of later solves depend on
 * the chosen shape only through the order of their partial sums.  2D
 * FP32-inner (mixed) solves are tuned separately from FP64 ones.
 * tile_param returns the parameter an FP64 solve of that kind and side gets. */
int opsmi_serve_autotune(opsmi_server s, int iters);
int opsmi_serve_tile_param(opsmi_server s, int kind, int n);
/* bytes a request needs (fields with halos and row padding) */
size_t opsmi_serve_request_bytes(const opsmi_solve_req *req);
void opsmi_serve_destroy(opsmi_server s);

#ifdef __cplusplus
}
#endif

#endif
