/* Run-time support for translator-generated loops (MI355X OPS backend).
 *
 * ops_amd/translator/opsmi_translate.py turns an application's
 * ops_par_loop(kernel, name, block, dim, range, args...) calls into calls of
 * generated host stubs ops_par_loop_<kernel>(name, block, dim, range,
 * args...), the shape the reference translator emits
 * (ops_translator/resources/templates/cpp/loop_host.cpp.j2:12-33), and
 * writes one HIP kernel per user kernel (device side:
 * include/ops_mi355x_device.h).  A stub packs what the loop needs by value
 * (ops_arg_gbl values and ops_decl_const constants, taken when the loop is
 * called, as the reference's lazy queue copies them,
 * ops/c/src/core/ops_lazy.cpp:1433-1438) and hands the loop to
 * ops_mi_gen_par_loop, which validates it, queues it with the hand-written
 * loops (same dependency tracking and reduction handling) and calls the
 * stub's launch function when the loop runs.
 *
 * This is the backend's counterpart of the reference HIP/CUDA run-time
 * support used by generated code (ops/c/include/ops_hip_rt_support.h,
 * ops/c/src/hip/ops_hip_common.cpp): plain C, no HIP types.
 */
#ifndef OPS_MI355X_GEN_H
#define OPS_MI355X_GEN_H

#include <stddef.h>

#include "ops_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OPSMI_GEN_MAX_ARGS 32
/* grid points per lane along x: a workgroup row covers 64 * OPSMI_GEN_XP
 * consecutive points, lane l taking points l, l + 64, ... */
#define OPSMI_GEN_XP 2

/* element type codes of a generated loop's arguments */
enum { OPSMI_GEN_F64 = 0, OPSMI_GEN_F32 = 1, OPSMI_GEN_I32 = 2, OPSMI_GEN_NONE = 3, OPSMI_GEN_F16 = 4 };

/* one dat argument as the kernel sees it: base = device address of
 * component 0 of point (0, 0, 0) of the dat's current buffer; strides in
 * elements (md = components per point, the x stride of an AoS dat).  A
 * dimension the argument's stencil does not advance in (stride 0, a dat of
 * size 1 there: ops/c/include/ops_lib_core.h:824-852) has stride 0 here
 * (md = 0 for x). */
typedef struct opsmi_gen_dat {
  void *base;
  long md, sy, sz;
  long cs; /* elements between components of a point: 1 (AoS), or the
              points of the whole array (SoA, ops_set_soa; md is then 1) */
} opsmi_gen_dat;

/* what a launch function receives when its loop runs */
typedef struct opsmi_gen_launch {
  int range[6];      /* x0, x1, y0, y1, z0, z1 (1, 1 for missing dimensions) */
  int chunk;         /* 3D: z planes per workgroup; 2D/1D: groups of 4 rows */
  int three_d;       /* 1: grid[2] runs over z chunks */
  unsigned grid[3];  /* workgroups of 64 x 4 threads */
  int alias;         /* 1: the general kernel (plain pointers, x steps from md): one dat is
                        passed in two arguments, or a stencil does not advance in a dimension */
  int spread;        /* 1: a lane's points 64 apart also where they may be adjacent
                        (OPSMI_GEN_ADJ=0) */
  opsmi_gen_dat dat[OPSMI_GEN_MAX_ARGS];    /* dat arguments, by argument index */
  void *partials[OPSMI_GEN_MAX_ARGS];       /* reduction arguments: dim values per workgroup */
} opsmi_gen_launch;

/* launches the loop's kernel on `stream` (a hipStream_t); blob is the
 * stub's by-value data */
typedef void (*opsmi_gen_launch_fn)(const opsmi_gen_launch *L, const void *blob, void *stream);

/* Queues a generated loop.  types[i]: element type the kernel declares for
 * argument i (dat, gbl or reduce; OPSMI_GEN_NONE for ops_arg_idx).
 * npts[i]: number of points the kernel reads or writes through dat argument
 * i, listed in pts as (component or -1, dx, dy, dz) quadruples, argument
 * after argument; -1 when the translator could not list them (non-literal
 * indices); for an OPS_READ ops_arg_gbl the number of values the kernel's
 * copy holds (a call passing more is refused); 0 for other arguments.  Checks, as the hand-written entry
 * points do (ops/c/src/core/ops_lib_core.cpp:1704-1836 under OPS_DEBUG):
 * argument kinds and access modes, element types against the declared ones,
 * component counts, that each listed point is in the argument's stencil and
 * each listed component below the dat's count, and that range plus the
 * stencil's reach stays inside the dat.  Reductions: up to 4 components
 * each; the kernel stores one partial per workgroup and component, which
 * the library folds in a fixed order into the handle. */
void ops_mi_gen_par_loop(const char *name, ops_block block, int dim, const int *range, const ops_arg *args,
                         int nargs, const int *types, const int *npts, const int *pts, opsmi_gen_launch_fn fn,
                         const void *blob, size_t blob_bytes);

/* Copies the current value of constant `name` (declared with
 * ops_decl_const) into dst; an OPS error if it is missing or its size is
 * not dim values of `type`. */
void ops_mi_gen_const(const char *name, const char *type, int dim, void *dst);
/* The same for a constant of any type (user structs included) or of a count
 * that is not a literal: elem_size is sizeof the type; with exact != 0 the
 * constant must hold max_dim values, otherwise at most max_dim (the kernel's
 * copy holds OPSMI_CONST_CAP(T) of them, OPSMI_CONST_BYTES bytes); an OPS
 * error if the type name differs from the declaration's or the count does
 * not fit. */
void ops_mi_gen_const_sized(const char *name, const char *type, size_t elem_size, int max_dim, int exact,
                            void *dst);
#define OPSMI_CONST_BYTES 2048
#define OPSMI_CONST_CAP(T) ((OPSMI_CONST_BYTES + sizeof(T) - 1) / sizeof(T))

/* launches one kernel that runs, at each point, the first loop's kernel and
 * then the second's (L1, L2 share range and tiling): the neighbour form
 * below */
typedef void (*opsmi_gen_fused_fn)(const opsmi_gen_launch *L1, const void *blob1, const opsmi_gen_launch *L2,
                                   const void *blob2, void *stream);


/* Registers `fused` as the one-pass form of a loop launched by `first`
 * directly followed by a loop launched by `second` that reads, at stencil
 * offsets, dats the first one writes (the GPU form of the reference's
 * run-time tiling of loop chains, ops/c/src/core/ops_lazy.cpp:380-1187:
 * the reference skews tiles so a tile of the second loop runs after the
 * tiles of the first it depends on; here each point of the second loop
 * gets the first loop's output at its neighbours by running the first
 * loop's kernel there into registers).  links holds nlinks pairs
 * (i, j): argument i of the second loop reads what argument j of the first
 * writes.  The queue launches it at fusion level 2 when: same block, range
 * and dimension; no shared reduction handle; for each link the two
 * arguments name the same dat, the first loop writes it with OPS_WRITE and
 * the second reads it with OPS_READ at listed points; the first loop reads
 * no dat it writes; the second loop touches every other dat the first
 * writes only at the point; and no dat is written by the second loop and
 * linked.  Dats the second loop writes and the first reads get a second
 * buffer for the pass (the first loop's recomputation still reads the old
 * values), which then becomes theirs.  Fields are the same bits as the two
 * loops run one after the other; reductions of the second loop too, those
 * of the first are accumulated at each point once. */
void ops_mi_gen_fused_nb(opsmi_gen_launch_fn first, opsmi_gen_launch_fn second, opsmi_gen_fused_fn fused,
                         const int *links, int nlinks);

/* launches one kernel that runs, at each point, the kernels of n loops in
 * turn (L[k], blob[k]: the k-th loop's launch data; they share range and
 * tiling) */
typedef void (*opsmi_gen_chain_fn)(const opsmi_gen_launch *const *L, const void *const *blob, void *stream);
#define OPSMI_GEN_MAX_CHAIN 4

/* Registers `fused` as the one-pass form of n (2 to OPSMI_GEN_MAX_CHAIN)
 * loops launched by loops[0], ..., loops[n - 1] in that order (no
 * counterpart in the reference, whose lazy queue,
 * ops/c/src/core/ops_lazy.cpp:380-1187, launches each loop of a chain
 * separately).  The translator registers one per pair of ops_par_loop calls
 * that follow each other in an application's source, and one per run of
 * three or four such calls that touch what they write only at the point.
 * The queue launches the longest registered form that matches the loops it
 * runs back to back when running them point by point gives what running
 * them one after the other gives: the same block, range and dimension; no
 * reduction handle shared by two of them; every dat any of them writes
 * accessed by all of them only at the point itself (listed points, offset
 * 0); pairs at fusion level >= 1 and
 * longer runs at level 3 (ops_mi_set_fusion), and only when the arguments
 * agree with `groups`: one entry per argument of each loop in turn
 * (ngroups in all), the index of the fused kernel's pointer a dat argument
 * uses, -1 for other arguments; arguments with one index must name one
 * dat, arguments with different indices different dats (the kernel's
 * pointers are restrict-qualified), and every dat argument must be
 * present.  Fields and reductions are the same bits as the loops run one
 * after the other. */
void ops_mi_gen_fused_chain(const opsmi_gen_launch_fn *loops, int n, const int *groups, int ngroups,
                            opsmi_gen_chain_fn fused);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* OPS_MI355X_GEN_H */
