/*
 * ops_mi355x.h — C ABI of the MI355X (gfx950) execution backend for the OPS
 * structured-mesh API.
 *
 * Every entry point is a plain C function over plain pointers and sizes.  The
 * names, argument meaning and error behaviour follow the reference OPS API;
 * each declaration cites the reference declaration it replaces
 * (paths relative to the OPS source tree).  Object structs keep the public
 * fields of the reference types so application code that reads e.g.
 * dat->size[0] keeps compiling; backend state hangs off the trailing impl
 * pointer.
 *
 * Per-loop entry points have the shape of the reference translator's
 * generated host stubs (ops_translator/resources/templates/cpp/
 * loop_host.cpp.j2:12-33):
 *   void ops_par_loop_<kernel>(const char *name, ops_block block, int dim,
 *                              int *range, ops_arg arg0, ...);
 * Only the hand-written kernels declared below exist; each validates its
 * arguments (kind, access mode, type, stencil coverage and bounds) before it
 * touches the GPU.
 *
 * Errors: by default an error prints "OPS error <code>: <msg>" and aborts,
 * the outcome of the reference's uncaught OPSException
 * (ops/c/include/ops_exceptions.h:46-53).  ops_mi_set_error_mode(1) makes the
 * failing call return instead; ops_mi_last_error() then reports it.
 *
 * Threading: one host thread drives the API (ops/c/src/core/
 * ops_instance.cpp:13,37-40).  All device work goes to one in-order HIP
 * stream owned by the library (ops_mi_stream()).
 */
#ifndef OPS_MI355X_H
#define OPS_MI355X_H

#include <stddef.h>
#include <stdio.h>

/* ---- utility macros applications use (ops/c/include/ops_macros.h:43-57) ---- */
#ifndef MIN
#define MIN(a, b) ((a < b) ? (a) : (b))
#endif
#ifndef MIN3
#define MIN3(a, b, c) MIN(a, MIN(b, c))
#endif
#ifndef MAX
#define MAX(a, b) ((a > b) ? (a) : (b))
#endif
#ifndef MAX3
#define MAX3(a, b, c) MAX(a, MAX(b, c))
#endif
#ifndef SIGN
#define SIGN(a, b) ((b < 0.0) ? (a * (-1)) : (a))
#endif

/* The legacy index macros OPS_ACCn / OPS_ACC_MDn (ops/c/include/
 * ops_macros.h:723-1133) mean something only inside a kernel that the
 * translator generates a loop from (include/ops_mi355x_device.h defines them
 * there); a legacy kernel run any other way stops with an error. */
#include <stdlib.h>
static inline long ops_mi_legacy_acc_host(int arg) {
  fprintf(stderr, "OPS error %d: OPS_ACC%d indexing outside a translator-generated loop\n", 1, arg);
  abort();
  return 0;
}
#ifndef OPSMI_DEVICE_LEGACY_ACC
#define OPS_ACC0(...) ops_mi_legacy_acc_host(0)
#define OPS_ACC1(...) ops_mi_legacy_acc_host(1)
#define OPS_ACC2(...) ops_mi_legacy_acc_host(2)
#define OPS_ACC3(...) ops_mi_legacy_acc_host(3)
#define OPS_ACC4(...) ops_mi_legacy_acc_host(4)
#define OPS_ACC5(...) ops_mi_legacy_acc_host(5)
#define OPS_ACC6(...) ops_mi_legacy_acc_host(6)
#define OPS_ACC7(...) ops_mi_legacy_acc_host(7)
#define OPS_ACC8(...) ops_mi_legacy_acc_host(8)
#define OPS_ACC9(...) ops_mi_legacy_acc_host(9)
#define OPS_ACC_MD0(...) ops_mi_legacy_acc_host(0)
#define OPS_ACC_MD1(...) ops_mi_legacy_acc_host(1)
#define OPS_ACC_MD2(...) ops_mi_legacy_acc_host(2)
#define OPS_ACC_MD3(...) ops_mi_legacy_acc_host(3)
#define OPS_ACC_MD4(...) ops_mi_legacy_acc_host(4)
#define OPS_ACC_MD5(...) ops_mi_legacy_acc_host(5)
#endif

/* ---- constants: ops/c/include/ops_lib_core.h:57-92,163-165 ---- */
#define OPS_MAX_DIM 5
#define OPS_READ 0
#define OPS_WRITE 1
#define OPS_RW 2
#define OPS_INC 3
#define OPS_MIN 4
#define OPS_MAX 5
#define OPS_ARG_GBL 0
#define OPS_ARG_DAT 1
#define OPS_ARG_IDX 2
#define OPS_HOST 1
#define OPS_DEVICE 2
/* error codes: ops/c/include/ops_exceptions.h:46-53 */
#define OPS_NOT_IMPLEMENTED 1
#define OPS_RUNTIME_CONFIGURATION_ERROR 2
#define OPS_INTERNAL_ERROR 3
#define OPS_RUNTIME_ERROR 4
#define OPS_INVALID_ARGUMENT 6

typedef int ops_access;
typedef int ops_arg_type;
typedef int ops_memspace;

/* IEEE binary16 element type of "half" dats (ops/c/include/ops_lib_core.h:
 * 98-119): the compiler's _Float16 where it has one (clang and hipcc, as the
 * translator builds applications; GCC 12 and later), otherwise its bits */
#include <float.h>
#if defined(__clang__) || defined(FLT16_MIN) || defined(__STDCPP_FLOAT16_T__)
typedef _Float16 half;
#else
#include <stdint.h>
typedef uint16_t half;
#endif

/* ---- object types: ops/c/include/ops_lib_core.h:181-552.  Under C++ the
 * structs also carry the reference's member functions (its OPS_CPP_API,
 * defined after the C declarations below); the data layout is the same. ---- */
struct ops_dat_core;
struct ops_stencil_core;
#ifdef __cplusplus
/* the objects' instance member (ops/c/include/ops_lib_core.h: an
 * OPS_instance *): one pointer wide, as in C, and under C++ it reaches the
 * library's single instance, e.g. block->instance->OPS_diags */
class OPS_instance;
struct opsmi_instance_ref {
  void *unused_;
  inline OPS_instance *operator->() const;
  inline operator OPS_instance *() const;
};
#define OPSMI_INSTANCE_MEMBER opsmi_instance_ref instance
#else
#define OPSMI_INSTANCE_MEMBER void *instance
#endif
typedef struct ops_block_core {
  int index;
  int dims;
  const char *name;
  OPSMI_INSTANCE_MEMBER;
#ifdef __cplusplus
  template <class T>
  struct ops_dat_core *decl_dat(int data_size, int *block_size, int *base, int *d_m, int *d_p, int *stride, T *data,
                                const char *type, const char *name);
  template <class T>
  struct ops_dat_core *decl_dat(int data_size, int *block_size, int *base, int *d_m, int *d_p, T *data,
                                const char *type, const char *name);
#endif
} ops_block_core;
typedef ops_block_core *ops_block;

typedef struct ops_dat_core {
  int index;
  ops_block block;
  int dim;                  /* components per grid point (AoS) */
  int type_size;            /* bytes per component */
  int elem_size;            /* bytes per grid point */
  int size[OPS_MAX_DIM];    /* allocated points per dimension, incl. halo and x padding */
  int base[OPS_MAX_DIM];
  int d_m[OPS_MAX_DIM];     /* negative halo depths (<= 0) */
  int d_p[OPS_MAX_DIM];     /* positive halo depths, x includes x_pad */
  int x_pad;
  char *data;               /* host mirror (allocated on demand) */
  char *data_d;             /* HBM allocation start of the array */
  const char *name;
  const char *type;
  int dirty_hd;             /* 0 in sync, 1 host newer, 2 device newer */
  int locked_hd;
  int user_managed;
  int is_hdf5;
  const char *hdf5_file;
  int e_dat;
  size_t mem;               /* bytes of the array */
  size_t base_offset;       /* bytes from data(_d) to point (0,...,0) */
  int stride[OPS_MAX_DIM];
  void *impl;
#ifdef __cplusplus
  void get_data();
  void print_to_txtfile(const char *file_name);
  int get_local_npartitions();
  void get_extents(int part, int *disp, int *size);
  void get_raw_metadata(int part, int *disp, int *size, int *stride, int *d_m, int *d_p);
  char *get_raw_pointer(int part, struct ops_stencil_core *stencil, int *memspace);
  void release_raw_data(int part, int acc);
  void release_raw_data(int part, int acc, int *memspace);
  void fetch_data(int part, char *data, int memspace = OPS_HOST);
  void fetch_data_slab(int part, char *data, int *range, int memspace = OPS_HOST);
  void set_data(int part, char *data, int memspace = OPS_HOST);
  void set_data_slab(int part, char *data, int *range, int memspace = OPS_HOST);
#endif
} ops_dat_core;
typedef ops_dat_core *ops_dat;

typedef struct ops_stencil_core {
  int index;
  int dims;
  int points;
  const char *name;
  int *stencil;
  int *stride;
  int *mgrid_stride;
  int type;
} ops_stencil_core;
typedef ops_stencil_core *ops_stencil;

typedef struct ops_reduction_core {
  char *data;               /* host value (size bytes) */
  int size;
  int initialized;
  int index;
  ops_access acc;
  char *type;
  char *name;
  OPSMI_INSTANCE_MEMBER;
  void *impl;
#ifdef __cplusplus
  template <class T>
  void get_result(T *ptr);
#endif
} ops_reduction_core;
typedef ops_reduction_core *ops_reduction;

typedef struct ops_halo_core {
  ops_dat from;
  ops_dat to;
  int iter_size[OPS_MAX_DIM];
  int from_base[OPS_MAX_DIM];
  int to_base[OPS_MAX_DIM];
  int from_dir[OPS_MAX_DIM];
  int to_dir[OPS_MAX_DIM];
  int index;
} ops_halo_core;
typedef ops_halo_core *ops_halo;

typedef struct ops_halo_group_core {
  int nhalos;
  ops_halo *halos;
  int index;
  OPSMI_INSTANCE_MEMBER;
#ifdef __cplusplus
  void halo_transfer();
#endif
} ops_halo_group_core;
typedef ops_halo_group_core *ops_halo_group;

/* ops/c/include/ops_lib_core.h:509-520 */
typedef struct ops_arg {
  ops_dat dat;
  ops_stencil stencil;
  int dim;
  int elem_size;
  char *data;
  char *data_d;
  ops_access acc;
  ops_arg_type argtype;
  int opt;
} ops_arg;

/* types above, functions below: the structs carry C++ member templates
 * under C++, which must not have C linkage */
#ifdef __cplusplus
extern "C" {
#endif

/* ---- lifecycle ---- */
/* ops/c/include/ops_lib_core.h:601-608.  argv tokens (as parsed by
 * ops/c/src/core/ops_lib_core.cpp:100-249):
 *   OPS_DIAGS=n, -OPS_DIAGS=n   diagnostics level
 *   OPS_DEVICE=n                HIP device
 *   OPS_TILING                  run-time loop-chain fusion on (fusion level 2,
 *                               the default unless OPSMI_FUSION lowers it)
 *   OPS_TILING_MAXDEPTH=n       recorded; it bounds the halo depth tiles
 *                               exchange between MPI processes, and one GPU
 *                               exchanges none
 *   OPS_BLOCK_SIZE_Y=n          translator-generated 2D loops: rows per
 *                               workgroup (rounded up to a multiple of 4, the
 *                               workgroup's waves)
 *   OPS_BLOCK_SIZE_Z=n          translator-generated 3D loops: z planes per
 *                               workgroup
 *   OPS_BLOCK_SIZE_X=n          recorded; a workgroup is one 64-lane wave
 *                               wide in x on this backend
 * Hand-written kernels keep their tuned tile shapes. */
void ops_init(const int argc, const char *const argv[], const int diags_level);
void ops_exit(void);
/* ops/c/include/ops_lib_core.h:1076-1077 (single GPU: no decomposition) */
void ops_partition(const char *routine);
/* ops/c/include/ops_lib_core.h:611: multi-component dats declared after
 * ops_set_soa(1) keep each component in its own array (structure of
 * arrays); fetch/set buffers keep a point's components together either way */
void ops_set_soa(const int soa_val);

/* ---- declarations ---- */
/* ops/c/include/ops_lib_core.h:620-621 */
ops_block ops_decl_block(int dims, const char *name);
/* ops/c/include/ops_internal1.h (behind the ops_decl_dat<T> templates,
 * ops/c/include/ops_lib_core.h:647-694); data may be NULL (zero-filled) or a
 * host array laid out with halos, size[d] - d_m[d] + d_p[d] points per dim */
ops_dat ops_decl_dat_char(ops_block block, int dim, int *size, int *base, int *d_m, int *d_p, int *stride,
                          char *data, int type_size, const char *type, const char *name);
/* ops/c/include/ops_lib_core.h:702-703 */
void ops_free_dat(ops_dat dat);
/* ops/c/include/ops_lib_core.h:824-852 */
ops_stencil ops_decl_stencil(int dims, int points, int *stencil, const char *name);
ops_stencil ops_decl_strided_stencil(int dims, int points, int *sten, int *stride, const char *name);
/* ops/c/include/ops_lib_core.h:810-811 */
ops_reduction ops_decl_reduction_handle(int size, const char *type, const char *name);
/* ops/c/include/ops_lib_core.h:888-902 */
ops_halo ops_decl_halo(ops_dat from, ops_dat to, int *iter_size, int *from_base, int *to_base, int *from_dir,
                       int *to_dir);
ops_halo_group ops_decl_halo_group(int nhalos, ops_halo *halos);
/* ops/c/include/ops_internal1.h:99; instance may be NULL (the one
 * instance this backend keeps) */
void ops_decl_const_char(void *instance, int dim, const char *type, int type_size, char *data, const char *name);

/* ---- loop arguments: ops/c/include/ops_lib_core.h:732-796 ---- */
ops_arg ops_arg_dat(ops_dat dat, int dim, ops_stencil stencil, const char *type, ops_access acc);
ops_arg ops_arg_dat_opt(ops_dat dat, int dim, ops_stencil stencil, const char *type, ops_access acc, int flag);
ops_arg ops_arg_idx(void);
ops_arg ops_arg_reduce(ops_reduction handle, int dim, const char *type, ops_access acc);
/* behind ops_arg_gbl<T> (ops/c/include/ops_lib_core.h:793-796); READ data is
 * copied when the loop is queued (ops/c/src/core/ops_lazy.cpp:1433-1438) */
ops_arg ops_arg_gbl_char(char *data, int dim, int size, ops_access acc);

/* ---- synchronisation and data access ---- */
/* ops/c/include/ops_lib_core.h:910-911 */
void ops_halo_transfer(ops_halo_group group);
/* ops/c/include/ops_lib_core.h:922-934, ops/c/src/core/ops_lib_core.cpp:1143-1149 */
void ops_reduction_result_char(ops_reduction handle, int type_size, char *ptr);
/* flushes queued loops (ops/c/src/core/ops_lazy.cpp:1248) */
void ops_execute(void);
/* ops/c/include/ops_lib_core.h:1104-1296 */
void ops_dat_get_extents(ops_dat dat, int part, int *disp, int *size);
void ops_dat_get_raw_metadata(ops_dat dat, int part, int *disp, int *size, int *stride, int *d_m, int *d_p);
char *ops_dat_get_raw_pointer(ops_dat dat, int part, ops_stencil stencil, ops_memspace *memspace);
void ops_dat_release_raw_data(ops_dat dat, int part, ops_access acc);
void ops_dat_release_raw_data_memspace(ops_dat dat, int part, ops_access acc, ops_memspace *memspace);
void ops_dat_fetch_data(ops_dat dat, int part, char *data);
void ops_dat_fetch_data_memspace(ops_dat dat, int part, char *data, ops_memspace memspace);
void ops_dat_fetch_data_slab_memspace(ops_dat dat, int part, char *data, int *range, ops_memspace memspace);
/* interior-sized buffer, as the reference device path and the documentation
 * specify (ops/c/src/core/ops_device_singlenode_common.cpp:227-255) */
void ops_dat_set_data(ops_dat dat, int part, char *data);
void ops_dat_set_data_memspace(ops_dat dat, int part, char *data, ops_memspace memspace);
void ops_dat_set_data_slab_memspace(ops_dat dat, int part, char *data, int *range, ops_memspace memspace);
int ops_dat_get_local_npartitions(ops_dat dat);
int ops_dat_get_global_npartitions(ops_dat dat);

/* ---- random fills: ops/c/include/ops_lib_core.h:1302-1309.  Counter-based
 * (SplitMix64 finaliser) so values depend only on (seed, dat index, point),
 * unlike the reference's std::mt19937 stream; fills the whole allocation ---- */
void ops_randomgen_init(unsigned int seed, int options);
void ops_fill_random_uniform(ops_dat dat);
void ops_fill_random_normal(ops_dat dat);
void ops_randomgen_exit(void);

/* ---- utilities: ops/c/include/ops_lib_core.h:990-1061 ---- */
void ops_timers(double *cpu, double *et);
void ops_timers_core(double *cpu, double *et);
void ops_printf(const char *format, ...);
void ops_fprintf(FILE *stream, const char *format, ...);
void ops_timing_output_stdout(void);
void ops_diagnostic_output(void);
int ops_is_root(void);

/* ---- tridiagonal: ops/c/include/ops_tridiag.h:45-83.  The reference's
 * ops_tridsolver_params is a C++ class whose one data member,
 * tridsolver_params, points at the solver's settings; the C++ class below
 * keeps that layout, and C code fills the struct itself from
 * ops_tridsolver_params_create. ---- */
typedef struct ops_tridsolver_params_c ops_tridsolver_params_c;
/* strategy: the reference SolveStrategy values; PCR (3) selects parallel
 * cyclic reduction, every other value the Thomas algorithm */
ops_tridsolver_params_c *ops_tridsolver_params_create(ops_block block, int strategy);
void ops_tridsolver_params_destroy(ops_tridsolver_params_c *p);
#ifdef __cplusplus
class ops_tridsolver_params;
#else
typedef struct ops_tridsolver_params {
  void *tridsolver_params; /* an ops_tridsolver_params_c */
} ops_tridsolver_params;
#endif
/* ops/c/include/ops_tridiag.h:73-76 (OPS_FTN_INTEROP: C linkage),
 * ops/c/src/tridiag/ops_tridiag.cpp:78-160 */
void ops_tridMultiDimBatch(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                           ops_tridsolver_params *tridsolver_ctx);
/* ops/c/include/ops_tridiag.h:78-81, ops/c/src/tridiag/ops_tridiag.cpp:168-249 */
void ops_tridMultiDimBatch_Inc(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                               ops_dat u, ops_tridsolver_params *tridsolver_ctx);
/* the same solves over the settings object itself */
void ops_tridMultiDimBatch_c(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                             ops_tridsolver_params_c *ctx);
void ops_tridMultiDimBatch_Inc_c(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c,
                                 ops_dat d, ops_dat u, ops_tridsolver_params_c *ctx);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ops_tridMultiDimBatch (u NULL) or
 * ops_tridMultiDimBatch_Inc (u given) on the same kernels, with the c'/d'
 * scratch in a fixed 64 MiB device array of the library instead of
 * library-allocated memory.  Limit: 2 * rows * systems * sizeof(element)
 * <= 64 MiB (4,194,304 points in double, 8,388,608 in float); larger
 * solves are refused with OPS_INVALID_ARGUMENT.  It calls no hipMalloc or
 * hipFree and records no statistics or timing events; it refuses to run
 * while loops are queued (call ops_execute() first) or while a dat's host
 * copy is newer than its device copy.  Calls on one GPU share the scratch:
 * calls that overlap on one GPU (several host threads) race. */
void ops_tridMultiDimBatch_noalloc(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c,
                                   ops_dat d, ops_dat u, ops_tridsolver_params *tridsolver_ctx);

/* ops/c/include/ops_internal2.h:219: reports the first NaN in the dat
 * (halos included) and exits with status 2 */
void ops_NaNcheck(ops_dat dat);
/* ops/c/include/ops_lib_core.h:1050 (ops_print_dat_to_txtfile,
 * ops/c/src/core/ops_lib_core.cpp:1277-1462): appends the dat's whole array,
 * halos included, to a text file in the reference's format */
void ops_print_dat_to_txtfile(ops_dat dat, const char *file_name);

/* ---- utilities applications call: ops/c/include/ops_util.h:41-57 (the
 * allocators report an OPS error instead of returning NULL) ---- */
void *xmalloc(size_t size);
void *xrealloc(void *ptr, size_t size);
void *xcalloc(size_t number, size_t size);
int file_exist(const char *filename);

/* ---- checkpointing: ops/c/include/ops_checkpointing.h:59-180.  This
 * backend has no HDF5, so it behaves as the reference built without
 * checkpointing (ops/c/src/externlib/ops_checkpointing.cpp:1734-1779):
 * ops_checkpointing_init returns false (no restart), the rest do nothing
 * and return false. ---- */
#define OPS_CHECKPOINT_INITPHASE 1
#define OPS_CHECKPOINT_MANUAL_DATLIST 2
#define OPS_CHECKPOINT_FASTFW 4
#define OPS_CHECKPOINT_MANUAL 8
#ifndef __cplusplus
#include <stdbool.h>
#endif
bool ops_checkpointing_init(const char *filename, double interval, int options);
void ops_checkpointing_initphase_done(void);
void ops_checkpointing_manual_datlist(int ndats, ops_dat *datlist);
bool ops_checkpointing_fastfw(int nbytes, char *payload);
bool ops_checkpointing_manual_datlist_fastfw(int ndats, ops_dat *datlist, int nbytes, char *payload);
bool ops_checkpointing_manual_datlist_fastfw_trigger(int ndats, ops_dat *datlist, int nbytes, char *payload);

/* ---- HDF5 output: ops/c/include/ops_hdf5.h:130-166.  This backend has no
 * HDF5 library; each call reports OPS_NOT_IMPLEMENTED (error modes below) ---- */
void ops_fetch_block_hdf5_file(ops_block block, const char *file_name);
void ops_fetch_dat_hdf5_file(ops_dat dat, const char *file_name);
void ops_fetch_stencil_hdf5_file(ops_stencil stencil, const char *file_name);
void ops_fetch_halo_hdf5_file(ops_halo halo, const char *file_name);

/* ==== hand-written per-loop entry points ============================== */

/* Laplace tutorial kernels: apps/c/laplace2d_tutorial/step7/laplace_kernels.h
 * (double or float dats, 2D) */
/* :4-6  A(0,0) = 0 (any dims) */
void ops_par_loop_set_zero(const char *name, ops_block block, int dim, int *range, ops_arg arg0);
/* :8-10 A(0,0) = Anew(0,0) (any dims) */
void ops_par_loop_copy(const char *name, ops_block block, int dim, int *range, ops_arg arg0, ops_arg arg1);
/* :12-14, constants pi, jmax (ops_decl_const) */
void ops_par_loop_left_bndcon(const char *name, ops_block block, int dim, int *range, ops_arg arg0, ops_arg arg1);
/* :16-18 */
void ops_par_loop_right_bndcon(const char *name, ops_block block, int dim, int *range, ops_arg arg0, ops_arg arg1);
/* top edge u = sin(pi (i+1)/(imax+1)) (this project's Laplace problem) */
void ops_par_loop_top_bndcon(const char *name, ops_block block, int dim, int *range, ops_arg arg0, ops_arg arg1);
/* :20-24 Anew = 0.25 (A(1,0)+A(-1,0)+A(0,-1)+A(0,1)); err = max(err,|Anew-A|) */
void ops_par_loop_apply_stencil(const char *name, ops_block block, int dim, int *range, ops_arg arg0,
                                ops_arg arg1, ops_arg arg2);

/* TeaLeaf kernels: apps/c/TeaLeaf/tea_leaf_kernels.h, tea_leaf_cg_kernels.h,
 * tea_leaf_common_kernels.h (2D as in the reference; 3D appends z terms) */
void ops_par_loop_tea_leaf_init_zero_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p);
void ops_par_loop_tea_leaf_init_zero2_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p,
                                             ops_arg z);
void ops_par_loop_tea_leaf_yeqx_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p,
                                       ops_arg x);
void ops_par_loop_tea_leaf_recip_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                        ops_arg p);
void ops_par_loop_tea_leaf_dot_kernel(const char *name, ops_block block, int dim, int *range, ops_arg r,
                                      ops_arg p, ops_arg rro);
void ops_par_loop_tea_leaf_norm2_kernel(const char *name, ops_block block, int dim, int *range, ops_arg x,
                                        ops_arg norm);
void ops_par_loop_tea_leaf_axpy_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                       ops_arg p, ops_arg alpha);
void ops_par_loop_tea_leaf_axpby_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                        ops_arg p, ops_arg alpha, ops_arg beta);
void ops_par_loop_tea_leaf_zeqxty_kernel(const char *name, ops_block block, int dim, int *range, ops_arg z,
                                         ops_arg x, ops_arg y);
void ops_par_loop_tea_leaf_common_init_Kx_Ky_kernel(const char *name, ops_block block, int dim, int *range,
                                                    ops_arg Kx, ops_arg Ky, ops_arg w);
void ops_par_loop_tea_leaf_common_init_Kx_Ky_Kz_kernel(const char *name, ops_block block, int dim, int *range,
                                                       ops_arg Kx, ops_arg Ky, ops_arg Kz, ops_arg w);
void ops_par_loop_tea_leaf_common_init_kernel(const char *name, ops_block block, int dim, int *range, ops_arg w,
                                              ops_arg r, ops_arg Kx, ops_arg Ky, ops_arg u, ops_arg rx,
                                              ops_arg ry);
void ops_par_loop_tea_leaf_common_init_kernel_3d(const char *name, ops_block block, int dim, int *range,
                                                 ops_arg w, ops_arg r, ops_arg Kx, ops_arg Ky, ops_arg Kz,
                                                 ops_arg u, ops_arg rx, ops_arg ry, ops_arg rz);
void ops_par_loop_tea_leaf_cg_calc_w_reduce_kernel(const char *name, ops_block block, int dim, int *range,
                                                   ops_arg w, ops_arg Kx, ops_arg Ky, ops_arg p, ops_arg rx,
                                                   ops_arg ry, ops_arg pw);
void ops_par_loop_tea_leaf_cg_calc_w_reduce_kernel_3d(const char *name, ops_block block, int dim, int *range,
                                                      ops_arg w, ops_arg Kx, ops_arg Ky, ops_arg Kz, ops_arg p,
                                                      ops_arg rx, ops_arg ry, ops_arg rz, ops_arg pw);
void ops_par_loop_tea_leaf_cg_calc_ur_r_reduce_kernel(const char *name, ops_block block, int dim, int *range,
                                                      ops_arg r, ops_arg w, ops_arg alpha, ops_arg rnn);

/* constant-coefficient Poisson (this project's problem): 2D 5-point
 * w = 4p - (p(-1,0)+p(1,0)) - (p(0,-1)+p(0,1)), 3D 7-point with 6p and z
 * pair appended; double or float dats, pw always accumulated in double */
void ops_par_loop_poisson_calc_w_reduce_kernel(const char *name, ops_block block, int dim, int *range,
                                               ops_arg w, ops_arg p, ops_arg pw);

/* Mixed-precision Poisson (this project's solver; the reference has no
 * mixed-precision path): FP64 iterative refinement around a
 * Jacobi-preconditioned CG with FP32 storage and arithmetic, every dot
 * product accumulated in FP64.  Same per-loop stub shape as above.
 *   poisson_residual_kernel  r = f - A x (FP64 dats), rr += r*r (double)
 *   mp_init_kernel           e = 0; rr = (float)r; p = Mi*rr; float dats
 *                            e, rr, p; r double; Mi float gbl;
 *                            dots[0] += rr*(Mi*rr), dots[1] += rr*rr
 *   mp_update_kernel         e += a*p; rr -= a*w; same two dots (a, Mi
 *                            float gbls; e, rr RW; p, w READ)
 *   mp_rr_update_kernel      rr -= a*w; the two dots (the rr half of
 *                            mp_update_kernel)
 *   mp_e_update_kernel       e += a*p (the e half; the lazy queue holds it
 *                            past the dots result and fuses it with the next
 *                            mp_p_kernel and operator sweep)
 *   mp_p_kernel              p = Mi*rr + b*p (Mi, b float gbls)
 *   mp_correct_kernel        x = x + (double)e
 * The FP32 operator sweep is poisson_calc_w_reduce_kernel on float dats. */
void ops_par_loop_poisson_residual_kernel(const char *name, ops_block block, int dim, int *range, ops_arg r,
                                          ops_arg f, ops_arg x, ops_arg rr);
void ops_par_loop_mp_init_kernel(const char *name, ops_block block, int dim, int *range, ops_arg e, ops_arg rr,
                                 ops_arg p, ops_arg r, ops_arg mi, ops_arg dots);
void ops_par_loop_mp_update_kernel(const char *name, ops_block block, int dim, int *range, ops_arg e, ops_arg rr,
                                   ops_arg p, ops_arg w, ops_arg a, ops_arg mi, ops_arg dots);
void ops_par_loop_mp_rr_update_kernel(const char *name, ops_block block, int dim, int *range, ops_arg rr,
                                      ops_arg w, ops_arg a, ops_arg mi, ops_arg dots);
void ops_par_loop_mp_e_update_kernel(const char *name, ops_block block, int dim, int *range, ops_arg e, ops_arg p,
                                     ops_arg a);
void ops_par_loop_mp_p_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p, ops_arg rr,
                              ops_arg mi, ops_arg b);
void ops_par_loop_mp_correct_kernel(const char *name, ops_block block, int dim, int *range, ops_arg x,
                                    ops_arg e);

/* Multi-block Poisson test app kernels: apps/c/poisson/poisson_kernel.h:4-38
 * (constants dx, dy from ops_decl_const; double dats, 2D)
 *   populate     (gbl int dispx, gbl int dispy, idx, u W, f W, ref W)
 *   initialguess (u W)
 *   stencil      (u R 5-point, f R, u2 W)
 *   update       (in R, out W)
 *   error        (u R, ref R, err INC) */
void ops_par_loop_poisson_kernel_populate(const char *name, ops_block block, int dim, int *range, ops_arg dispx,
                                          ops_arg dispy, ops_arg idx, ops_arg u, ops_arg f, ops_arg ref);
void ops_par_loop_poisson_kernel_initialguess(const char *name, ops_block block, int dim, int *range, ops_arg u);
void ops_par_loop_poisson_kernel_stencil(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                         ops_arg f, ops_arg u2);
void ops_par_loop_poisson_kernel_update(const char *name, ops_block block, int dim, int *range, ops_arg in,
                                        ops_arg out);
void ops_par_loop_poisson_kernel_error(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                       ops_arg ref, ops_arg err);

/* ADI right-hand side and coefficients: apps/c/adi/preproc_kernel.h:4-39
 * (constants nx, ny, nz, lambda) */
void ops_par_loop_preproc_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u, ops_arg du,
                                 ops_arg ax, ops_arg bx, ops_arg cx, ops_arg ay, ops_arg by, ops_arg cy,
                                 ops_arg az, ops_arg bz, ops_arg cz, ops_arg idx);
/* 27-point weighted box sweep (weights 1/8,1/16,1/32,1/64 by Manhattan
 * distance; summed z-outer, x-inner) */
void ops_par_loop_stencil27_kernel(const char *name, ops_block block, int dim, int *range, ops_arg out,
                                   ops_arg u);

/* small reductions used by reduction-semantics tests (sum, sum of squares,
 * sum of |x|, min, max, (sum, count) pair, max of i*j, sum of scale*x) */
void ops_par_loop_red_sum(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s);
void ops_par_loop_red_sumsq(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s);
void ops_par_loop_red_abs(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s);
void ops_par_loop_red_min(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg m);
void ops_par_loop_red_max(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg m);
void ops_par_loop_red_two(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s2);
void ops_par_loop_red_imax(const char *name, ops_block block, int dim, int *range, ops_arg idx, ops_arg m);
void ops_par_loop_red_scaled(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg scale,
                             ops_arg s);

/* ==== backend extensions (no reference counterpart) =================== */
/* 0: print and abort on error (default); 1: record and return; 2: print,
 * record and return.  Initial mode from OPSMI_ERROR_MODE. */
void ops_mi_set_error_mode(int mode);
/* diagnostics level after ops_init (the argument of ops_init, OPS_DIAGS):
 * above 1 every launch is timed with HIP events for ops_mi_loop_stats and
 * the timing table; queued loops run first */
void ops_mi_set_diags(int level);
/* the current diagnostics level */
int ops_mi_get_diags(void);
/* 1 (default): a result read right after the one loop reducing into it
 * (FP64, first use since the last read, at most 4096 partials) is folded by
 * the host from partials the loop writes to pinned memory, with the same
 * bits as the fold kernel; 0: always the fold kernel.  Initial value from
 * OPSMI_HOST_FOLD (read by ops_init) */
void ops_mi_set_host_fold(int on);
/* last error code (0 = none) and message; reading clears it */
int ops_mi_last_error(char *msg, int msg_len);
/* the library's HIP stream (hipStream_t) */
void *ops_mi_stream(void);
/* blocks until all queued device work is done */
void ops_mi_sync(void);
/* per-loop statistics gathered when diags > 1 (the reference's
 * ops_timing_output table, ops/c/src/core/ops_lib_core.cpp:1468-1624):
 * launches made under this name, their summed kernel milliseconds (HIP events
 * on the library stream) and algorithmic bytes by ops_compute_transfer's rule
 * (:1687-1697; a loop the queue fused into another kernel adds its bytes here
 * but its launch and time to the fused kernel's name).  Returns 0 if the name
 * is unknown. */
int ops_mi_loop_stats(const char *name, long *count, double *kernel_ms, double *bytes);
void ops_mi_reset_stats(void);
/* memory-stream probe: a pointwise kernel reading nread and writing nwrite
 * arrays of `points` doubles (16-byte lanes; nt: non-temporal loads and
 * stores), launched reps times after one untimed launch; *avg_us = average
 * launch time.  The read/write mixes of this path's kernels: (1,0) (4,0)
 * (0,1) (1,1) (1,2) (2,1) (3,2) (4,1) (6,3) (4,2) (5,1).  Allocates its
 * arrays for the call.  Returns 0, or -1 on error. */
int ops_mi_stream_probe(int nread, int nwrite, int nt, long points, int reps, double *avg_us);
/* loop-chain fusion in the queue: 0 off, 1 the pointwise CG axpy+axpby pair,
 * 2 also the column chains (TeaLeaf and Poisson CG axpy+axpby+calc_w, the
 * mixed solver's e update + p update + FP32 sweep, the Laplace tutorial's
 * copy + next apply_stencil), each one pass; 3 also runs of three or four
 * generated point-wise loops (ops_mi_gen_fused_chain).  Every level gives
 * the same bits.  Initial level from OPSMI_FUSION (default 2). */
void ops_mi_set_fusion(int level);

#ifdef __cplusplus
} /* extern "C" */

/* ---- C++ conveniences with the reference's template signatures
 *      (ops/c/include/ops_lib_core.h:647-694,793-796,922-934,947-975) ---- */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

/* The accessor type user kernels are written with
 * (ops/c/include/ops_lib_core.h:1356-2038), so kernel definitions an
 * application includes compile as host code.  They are never run here: the
 * loops run as hand-written or translator-generated GPU kernels, whose
 * device accessor is opsmi_gen::Acc (include/ops_mi355x_device.h). */
template <class T>
class ACC {
 public:
  ACC(T *p = nullptr, long sx = 0, long sy = 0, long sz = 0) : p_(p), sx_(sx), sy_(sy), sz_(sz) {}
  T &operator()(int x, int y = 0, int z = 0, int w = 0) const {
    return p_[x + sx_ * (y + sy_ * ((long)z + sz_ * (long)w))];
  }

 private:
  T *p_;
  long sx_, sy_, sz_;
};
/* applies the OPS_soa flag of the instance object, if there is one (below) */
static inline void opsmi_instance_layout(void);
template <class T>
static inline ops_dat ops_decl_dat(ops_block block, int data_size, int *block_size, int *base, int *d_m,
                                   int *d_p, T *data, const char *type, const char *name) {
  int stride[OPS_MAX_DIM] = {1, 1, 1, 1, 1};
  opsmi_instance_layout();
  return ops_decl_dat_char(block, data_size, block_size, base, d_m, d_p, stride, (char *)data, sizeof(T), type,
                           name);
}
/* ops/c/include/ops_lib_core.h:647-660: with the dat's stride per dimension */
template <class T>
static inline ops_dat ops_decl_dat(ops_block block, int data_size, int *block_size, int *base, int *d_m,
                                   int *d_p, int *stride, T *data, const char *type, const char *name) {
  opsmi_instance_layout();
  return ops_decl_dat_char(block, data_size, block_size, base, d_m, d_p, stride, (char *)data, sizeof(T), type,
                           name);
}
template <class T>
static inline ops_arg ops_arg_gbl(T *data, int dim, const char *type, ops_access acc) {
  (void)type;
  return ops_arg_gbl_char((char *)data, dim, sizeof(T), acc);
}
template <class T>
static inline void ops_reduction_result(ops_reduction handle, T *ptr) {
  ops_reduction_result_char(handle, sizeof(T), (char *)ptr);
}
template <class T>
static inline void ops_decl_const(const char *name, int dim, const char *type, T *data) {
  ops_decl_const_char(nullptr, dim, type, sizeof(T), (char *)data, name);
}
/* ops/c/include/ops_lib_core.h:946-956: a new value for a declared constant */
template <class T>
static inline void ops_update_const(const char *name, int dim, const char *type, T *data) {
  ops_decl_const_char(nullptr, dim, type, sizeof(T), (char *)data, name);
}
/* ops/c/include/ops_tridiag.h:45-71: same single data member */
class ops_tridsolver_params {
 public:
  enum SolveStrategy { GATHER_SCATTER = 0, ALLGATHER, JACOBI, PCR, LATENCY_HIDING_INTERLEAVED,
                       LATENCY_HIDING_TWO_STEP };
  explicit ops_tridsolver_params(ops_block block) : tridsolver_params(ops_tridsolver_params_create(block, 0)) {}
  ops_tridsolver_params(ops_block block, SolveStrategy s)
      : tridsolver_params(ops_tridsolver_params_create(block, (int)s)) {}
  ~ops_tridsolver_params() { ops_tridsolver_params_destroy((ops_tridsolver_params_c *)tridsolver_params); }
  ops_tridsolver_params(const ops_tridsolver_params &) = delete;
  ops_tridsolver_params &operator=(const ops_tridsolver_params &) = delete;
  /* settings of the distributed and CUDA solvers: no effect on one GPU */
  void set_jacobi_params(double, double, int) {}
  void set_batch_size(int) {}
  void set_cuda_opts(int, int, int) {}
  void set_cuda_sync(int) {}
  void *tridsolver_params;
};
/* the reference's C++ signatures over the single instance
 * (ops/c/include/ops_lib_core.h: ops_execute(OPS_instance *),
 * ops_timing_output(std::ostream &)) */
#include <iostream>
static inline void ops_execute(void *instance) {
  (void)instance;
  ops_execute();
}
static inline void ops_timing_output(std::ostream &) { ops_timing_output_stdout(); }

/* ---- the reference's C++ object API (OPS_CPP_API: ops/c/include/
 *      ops_instance.h:62-380, ops/c/include/ops_lib_core.h:183-552,
 *      ops/c/include/ops_exceptions.h:46-53) over the C entry points.  This
 *      backend keeps one instance per process: an OPS_instance is it (its
 *      constructor runs ops_init, its destructor ops_exit), and a second one
 *      while the first exists is refused. ---- */
#include <exception>
#include <sstream>

struct OPSException : public std::exception {
  int code;
  std::stringstream msg;
  mutable std::string text;
  explicit OPSException(int c) : code(c) {}
  OPSException(int c, const char *m) : code(c) { msg << m; }
  OPSException(const OPSException &o) : std::exception(o), code(o.code) { msg << o.msg.str(); }
  template <class T>
  OPSException &operator<<(const T &v) {
    msg << v;
    return *this;
  }
  const char *what() const throw() override {
    text = msg.str();
    return text.c_str();
  }
};

class OPS_instance {
 public:
  OPS_instance(const int argc, const char *const argv[], const int diags_level, std::ostream &os = std::cout)
      : OPS_soa(0), os_(&os) {
    if (current()) throw OPSException(OPS_RUNTIME_CONFIGURATION_ERROR, "this backend keeps one OPS instance at a time");
    ops_init(argc, argv, diags_level);
    current() = this;
  }
  ~OPS_instance() { exit(); }
  OPS_instance(const OPS_instance &) = delete;
  OPS_instance &operator=(const OPS_instance &) = delete;
  /* the live instance, or NULL */
  static OPS_instance *&current() {
    static OPS_instance *p = nullptr;
    return p;
  }
  /* ops/c/include/ops_instance.h:67: the instance of the C API.  After
   * ops_init there is no instance object; this returns one that stands for
   * the library's single instance (created on first use, never calling
   * ops_init or ops_exit), whose OPS_soa the C-API declarations follow. */
  static OPS_instance *getOPSInstance() {
    if (current()) return current();
    if (!view()) {
      static OPS_instance v{ViewTag{}};
      view() = &v;
    }
    return view();
  }
  static OPS_instance *&view() {
    static OPS_instance *p = nullptr;
    return p;
  }
  static int numInstances() { return current() ? 1 : 0; }
  std::ostream &ostream() { return *os_; }
  void set_ostream(std::ostream &s) { os_ = &s; }
  void exit() {
    if (current() == this) {
      ops_exit();
      current() = nullptr;
    }
  }
  /* declarations follow the current OPS_soa setting */
  ops_block decl_block(int dims, const char *name) {
    ops_set_soa(OPS_soa);
    return ops_decl_block(dims, name);
  }
  ops_reduction decl_reduction_handle(int size, const char *type, const char *name) {
    return ops_decl_reduction_handle(size, type, name);
  }
  ops_stencil decl_stencil(int dims, int points, int *stencil, const char *name) {
    return ops_decl_stencil(dims, points, stencil, name);
  }
  ops_stencil decl_strided_stencil(int dims, int points, int *sten, int *stride, const char *name) {
    return ops_decl_strided_stencil(dims, points, sten, stride, name);
  }
  ops_halo decl_halo(ops_dat from, ops_dat to, int *iter_size, int *from_base, int *to_base, int *from_dir,
                     int *to_dir) {
    return ops_decl_halo(from, to, iter_size, from_base, to_base, from_dir, to_dir);
  }
  ops_halo_group decl_halo_group(int nhalos, ops_halo *halos) { return ops_decl_halo_group(nhalos, halos); }
  template <class T>
  void decl_const(const char *name, int dim, const char *type, T *data) {
    ops_decl_const_char(nullptr, dim, type, sizeof(T), (char *)data, name);
  }
  template <class T>
  void update_const(const char *name, int dim, const char *type, T *data) {
    decl_const(name, dim, type, data);
  }
  void partition(const char *routine) { ops_partition(routine); }
  void diagnostic_output() { ops_diagnostic_output(); }
  void timing_output(std::ostream &) { ops_timing_output_stdout(); }
  void timing_output_stdout() { ops_timing_output_stdout(); }
  int is_root() { return 1; }
  void execute() { ops_execute(); }
  int OPS_soa; /* 1: multi-component dats declared from now on are SoA */
  /* ops/c/include/ops_instance.h: OPS_diags, the diagnostics level (read
   * from the library) */
  struct DiagsLevel {
    operator int() const { return ops_mi_get_diags(); }
  } OPS_diags;
  /* ops/c/include/ops_instance.h:246: restarts the energy counters the
   * reference reads through rocm_smi/NVML when built with them; no counters
   * here */
  void reset_power_counters() {}

 private:
  struct ViewTag {};
  explicit OPS_instance(ViewTag) : OPS_soa(0), os_(&std::cout) {}
  std::ostream *os_;
};

inline OPS_instance *opsmi_instance_ref::operator->() const { return OPS_instance::getOPSInstance(); }
inline opsmi_instance_ref::operator OPS_instance *() const { return OPS_instance::getOPSInstance(); }

static inline void opsmi_instance_layout(void) {
  OPS_instance *i = OPS_instance::current() ? OPS_instance::current() : OPS_instance::view();
  if (i) ops_set_soa(i->OPS_soa);
}

template <class T>
inline ops_dat ops_block_core::decl_dat(int data_size, int *block_size, int *base, int *d_m, int *d_p, int *stride,
                                        T *data, const char *type, const char *name) {
  opsmi_instance_layout();
  return ops_decl_dat_char(this, data_size, block_size, base, d_m, d_p, stride, (char *)data, sizeof(T), type, name);
}
template <class T>
inline ops_dat ops_block_core::decl_dat(int data_size, int *block_size, int *base, int *d_m, int *d_p, T *data,
                                        const char *type, const char *name) {
  int stride[OPS_MAX_DIM] = {1, 1, 1, 1, 1};
  return decl_dat(data_size, block_size, base, d_m, d_p, stride, data, type, name);
}
template <class T>
inline void ops_reduction_core::get_result(T *ptr) {
  ops_reduction_result_char(this, sizeof(T), (char *)ptr);
}
inline void ops_halo_group_core::halo_transfer() { ops_halo_transfer(this); }
inline void ops_dat_core::get_data() { ops_execute(); }
inline void ops_dat_core::print_to_txtfile(const char *file_name) { ops_print_dat_to_txtfile(this, file_name); }
inline int ops_dat_core::get_local_npartitions() { return ops_dat_get_local_npartitions(this); }
inline void ops_dat_core::get_extents(int part, int *disp, int *size) { ops_dat_get_extents(this, part, disp, size); }
inline void ops_dat_core::get_raw_metadata(int part, int *disp, int *size, int *stride, int *d_m, int *d_p) {
  ops_dat_get_raw_metadata(this, part, disp, size, stride, d_m, d_p);
}
inline char *ops_dat_core::get_raw_pointer(int part, ops_stencil_core *stencil, int *memspace) {
  return ops_dat_get_raw_pointer(this, part, stencil, memspace);
}
inline void ops_dat_core::release_raw_data(int part, int acc) { ops_dat_release_raw_data(this, part, acc); }
inline void ops_dat_core::release_raw_data(int part, int acc, int *memspace) {
  ops_dat_release_raw_data_memspace(this, part, acc, memspace);
}
inline void ops_dat_core::fetch_data(int part, char *data, int memspace) {
  ops_dat_fetch_data_memspace(this, part, data, memspace);
}
inline void ops_dat_core::fetch_data_slab(int part, char *data, int *range, int memspace) {
  ops_dat_fetch_data_slab_memspace(this, part, data, range, memspace);
}
inline void ops_dat_core::set_data(int part, char *data, int memspace) {
  ops_dat_set_data_memspace(this, part, data, memspace);
}
inline void ops_dat_core::set_data_slab(int part, char *data, int *range, int memspace) {
  ops_dat_set_data_slab_memspace(this, part, data, range, memspace);
}
#endif

#endif /* OPS_MI355X_H */
