// Internal runtime of the MI355X OPS backend: object tables, device memory,
// argument validation, the loop queue and reduction bookkeeping.  Not part of
// the public ABI (include/ops_mi355x.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ops_mi355x.h"
#include "../../include/ops_mi355x_gen.h"

namespace opsmi {

// ---------------------------------------------------------------- errors
struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const std::string &msg);
void report(const Error &e);

#define OPSMI_HIP(call)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      ::opsmi::fail(OPS_RUNTIME_ERROR, std::string("HIP error in " #call ": ") + hipGetErrorString(e_)); \
  } while (0)

// every extern "C" entry point runs inside API_BEGIN/API_END so no C++
// exception crosses the ABI
#define OPSMI_API_BEGIN try {
#define OPSMI_API_END(ret)              \
  }                                     \
  catch (const ::opsmi::Error &e_) {    \
    ::opsmi::report(e_);                \
    return ret;                         \
  }                                     \
  catch (const std::exception &e_) {    \
    ::opsmi::report({OPS_INTERNAL_ERROR, e_.what()}); \
    return ret;                         \
  }

template <class... A>
std::string cat(const A &...a) {
  std::ostringstream s;
  (s << ... << a);
  return s.str();
}

// ---------------------------------------------------------------- types
enum DType { DT_F64 = 0, DT_F32 = 1, DT_I32 = 2, DT_OTHER = 3, DT_F16 = 4 };
DType dtype_of(const char *type, int type_size);

// device-side view of one dat argument: pointer to component 0 of point
// (0,0,0) and strides in elements: sx along x (the component count for an
// AoS dat, 1 for SoA), sy, sz, and cs between components (1 for AoS, the
// points of the whole array for SoA)
struct DatView {
  char *base;
  long sy, sz;
  long sx = 1, cs = 1;
};

struct RedImpl {
  void *d_slot = nullptr;    // device value of the handle (size bytes)
  void *h_mirror = nullptr;  // pinned, device-mapped host copy the final fold also writes
  void *d_mirror = nullptr;  // its device address
  // sequence word in the same pinned allocation: the final fold of the
  // n-th launch reducing into the handle stores n after its values
  volatile unsigned long long *h_seq = nullptr;
  unsigned long long *d_seq = nullptr;
  unsigned long long seq_issued = 0;  // launches prepared so far
  bool dev_valid = false;    // device slot holds the running value since re-arm
  DType dt = DT_OTHER;
};

struct DatImpl {
  char *alloc = nullptr;  // hipMalloc'd pointer (data_d may be offset for alignment)
  DType dt = DT_OTHER;
  size_t host_bytes = 0;
  char *spare = nullptr;      // second buffer for double-buffered fused chains
  bool spare_synced = false;  // spare equals the current buffer outside spare_range
  bool spare_ranged = false;  // spare_range is set (else: equal everywhere)
  int spare_range[6] = {0, 0, 0, 0, 0, 0};
  // edge dats: access of the last loop that wrote the dat, in program
  // order; an OPS_INC/MIN/MAX loop after a different access first resets the
  // dat to the identity (ops/c/src/core/ops_dummy_singlenode.cpp:204-391)
  int edat_prev_acc = OPS_READ;
  // structure-of-arrays layout (declared with more than one component while
  // the instance's SoA flag was set): component d of every point is one
  // array, d arrays apart
  bool soa = false;
};

struct LoopStat {
  long count = 0;     // calls of the loop (fused ones included)
  long launches = 0;  // launches under this name (ms is their kernel time)
  double ms = 0.0;
  double bytes = 0.0;
};

struct Const {
  std::string type;
  std::vector<char> bytes;
};

struct Instance {
  bool initialised = false;
  int diags = 1;
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev_wait = nullptr;  // wait_stream's completion event
  int error_mode = 0;
  int last_error = 0;
  std::string last_msg;
  int fusion = 2;  // 0 off, 1 the pointwise pair only, 2 also the column chains
  int soa = 0;     // ops_set_soa: multi-component dats declared from now on are SoA
  // OPS_BLOCK_SIZE_{X,Y,Z} (0: not given) and OPS_TILING_MAXDEPTH
  int block_x = 0, block_y = 0, block_z = 0;
  int tiling_maxdepth = 0;
  bool block_y_set() const { return block_y > 0; }
  bool block_z_set() const { return block_z > 0; }
  // rows of a generated 2D loop's workgroup: OPS_BLOCK_SIZE_Y rounded up to
  // a multiple of its 4 waves (default 16); planes of a generated 3D loop's
  // workgroup: OPS_BLOCK_SIZE_Z (default OPSMI_GEN_ZCHUNK, else 1)
  int gen_rows_2d() const;
  int gen_planes_3d() const;

  std::vector<ops_block> blocks;
  std::vector<ops_dat> dats;  // index -> dat (nullptr after free)
  std::vector<ops_stencil> stencils;
  std::vector<ops_reduction> reductions;
  std::vector<ops_halo> halos;
  std::vector<ops_halo_group> groups;
  std::map<std::string, Const> consts;
  std::map<std::string, LoopStat> stats;

  // reduction scratch for the one in-order stream: block partials of the
  // loop kernel and of the second reduction stage
  void *d_partials = nullptr;
  size_t partials_bytes = 0;
  void *d_partials2 = nullptr;

  // host staging for reduction results
  void *h_pinned = nullptr;

  // Host fold of a result read right away: the one loop of a flush that
  // reduces into the handle being read (FP64, first use since the last
  // read) writes its workgroup partials into this pinned, device-mapped
  // buffer (kHostFoldMax values, each preset to a sentinel NaN), and the
  // host folds them in the device fold's order; no fold kernel runs
  double *h_hfold = nullptr;
  void *d_hfold = nullptr;
  const void *hf_slot = nullptr;  // device slot of the armed handle (null: none armed)
  bool hf_used = false;           // the armed handle's loop took the host fold
  unsigned hf_n = 0;              // partials of that launch
  int hf_dim = 1;
  double hf_init[4] = {0, 0, 0, 0};  // the handle's initial values
  bool host_fold = true;             // ops_mi_set_host_fold
  bool gen_spread = false;           // OPSMI_GEN_ADJ=0: generated loops keep lane points 64 apart

  // random fills
  uint64_t rng_seed = 0;
  bool rng_init = false;

  // event pool for diags > 1 timing
  std::vector<hipEvent_t> ev_free;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
};

Instance &inst();
void require_init();

// ---------------------------------------------------------------- dats
DatView view_of(ops_dat d);
void ensure_host(ops_dat d);      // host mirror allocated and current
void mark_device_dirty(ops_dat d);  // device now holds the newest data
void sync_to_device(ops_dat d);   // host newer -> copy up
// Waits until the work queued on s has completed, polling an event (a
// blocking wait adds tens of microseconds of wake-up latency to every
// reduction result); after 50 ms of polling it blocks instead.
void wait_stream(hipStream_t s);

// ---------------------------------------------------------------- args
// Per-argument contract of a hand-written kernel.
enum ArgKind { AK_DAT, AK_GBL, AK_RED, AK_IDX };
struct ArgSpec {
  ArgKind kind;
  bool write = false;          // dat written by the kernel
  std::vector<int> offsets;    // dat: flattened (x,y,z) offsets the kernel reads/writes
  int red_op = OPS_INC;        // reduce: required access
  int dim = 1;                 // gbl/reduce components
  DType dt = DT_F64;           // expected element type (DT_OTHER: any of f64/f32)
};
std::vector<int> pts_point();
std::vector<int> pts_5pt();
std::vector<int> pts_7pt();
std::vector<int> pts_27pt();

struct Range3 {
  int x0, x1, y0, y1, z0, z1;
  long npts() const {
    long n = 1;
    n *= (x1 > x0 ? x1 - x0 : 0);
    n *= (y1 > y0 ? y1 - y0 : 0);
    n *= (z1 > z0 ? z1 - z0 : 0);
    return n;
  }
};

// Validates args against specs (kind, access, type, stencil coverage, bounds)
// and returns the loop range.  dat_dt receives the common dat element type
// for specs with dt == DT_OTHER.
Range3 check_loop(const char *name, ops_block block, int dim, const int *range, const ops_arg *args,
                  const std::vector<ArgSpec> &specs, DType *dat_dt);

// Validates the args of a translator-generated loop against the element
// types its kernel declares (types[i], DType codes; DT_OTHER for idx args):
// kinds, access modes, types, component counts, unit-stride stencils, the
// points the kernel touches (npts/pts as in ops_mi_gen_par_loop) inside
// each stencil, and range plus stencil reach inside every dat.  Returns the
// loop range.
Range3 check_gen_loop(const char *name, ops_block block, int dim, const int *range, const ops_arg *args, int nargs,
                      const int *types, const int *npts, const int *pts);

// value of a READ gbl arg
template <class T>
T gbl(const ops_arg &a) {
  T v;
  memcpy(&v, a.data, sizeof(T));
  return v;
}
const Const &get_const(const char *name);
double const_as_double(const char *name);
int const_as_int(const char *name);

// ---------------------------------------------------------------- reductions
// Device-side control block passed to reducing kernels.
// partials one host fold takes (workgroups x components), and the bit
// pattern the host presets each to (a signalling NaN no arithmetic
// produces): a partial has arrived when it differs
constexpr unsigned kHostFoldMax = 4096;
constexpr unsigned long long kHostFoldEmpty = 0x7ff4dead5eed0001ULL;

struct RedCtl {
  void *slot;          // handle's device value
  void *mirror;        // host-mapped copy of it (written by the final fold)
  void *partials;      // per-block partials of the loop kernel (dim values per block)
  void *partials2;     // second-stage partials
  int use_init;        // 1: start from init[] (first use since re-arm)
  double init[4];      // initial values, stored as the reduction type
  int init_i[4];
  unsigned long long *seq;     // host-mapped sequence word (null: none)
  unsigned long long seq_val;  // stored there after the values
  int hostfold;                // 1: partials go to host memory and the host folds them
};
// Prepares a handle for a launch (device slot, init values).  Marks the
// handle's device value as live; the result is copied back at result time.
RedCtl red_prepare(const ops_arg &a, int dim, DType dt, int nblocks);
void ensure_partials(size_t bytes);

// ---------------------------------------------------------------- loops
// Records one executed loop for statistics; with diags > 1 the launches made
// by fn are bracketed by events on the library stream.
void run_loop(const char *name, double bytes, const std::function<void(hipStream_t)> &fn);
double transfer_bytes(const ops_arg *args, int nargs, const Range3 &r);
int grid_blocks(long rows, int rows_per_block);

// ---------------------------------------------------------------- lazy queue
// Loops are queued (the reference's OPS_LAZY mode, ops/c/src/core/
// ops_lazy.cpp:186-276) and run when a result is needed.  A reduction result
// runs only the queued loops it depends on: the loops reducing into the
// handle plus, transitively, every earlier loop that conflicts with one of
// them (read-after-write, write-after-read, write-after-write on a dat, or
// the same reduction handle); the rest stay queued in program order.  Data
// access (fetch/set/raw pointers/halo transfers/frees) runs everything.
// Device pointers are resolved when a loop runs, so a dat's buffer may be
// swapped by a fused loop chain in between.
enum LoopKind {
  LK_OPAQUE = 0,
  LK_AXPY,       // tea_leaf_axpy_kernel: u = u + a p
  LK_AXPBY,      // tea_leaf_axpby_kernel: u = a u + b p
  LK_HEAT_W3,    // tea_leaf_cg_calc_w_reduce_kernel (3D): w = A p, pw
  LK_POISSON_W3, // poisson_calc_w_reduce_kernel (3D column form): w = A p, pw; g[0] = element bytes
  LK_MP_E,       // mp_e_update_kernel: e = e + a p (float)
  LK_MP_P,       // mp_p_kernel: p = mi rr + b p (float)
  LK_COPY,       // copy: d[0] = d[1]; g[0] = element bytes
  LK_JACOBI5,    // apply_stencil: d[1] = 5-point average of d[0], MAX |change|; g[0] = element bytes
};
struct GenLoop;  // a translator-generated loop's launch data (genloop.hip)
struct LoopRec {
  std::string name;
  std::shared_ptr<const GenLoop> gen;  // set for a generated loop
  LoopKind kind = LK_OPAQUE;
  Range3 range{};
  int dims = 0;
  std::vector<ops_dat> reads, writes;  // dats read (READ/RW/INC), written (WRITE/RW/INC)
  ops_reduction red = nullptr;         // handle this loop reduces into
  std::vector<ops_reduction> more_reds;  // further handles (loops with several reduction args)
  ops_dat d[8] = {};                   // kind-specific dat operands
  double g[4] = {0, 0, 0, 0};          // kind-specific scalar operands
  RedCtl ctl{};                        // reduction control of a reducing loop
  double bytes = 0.0;
  std::function<void(hipStream_t)> launch;
};
// a fusion rule looks at the front of a run of loops about to execute; when
// it matches it launches one fused kernel for the first n loops and returns n
typedef int (*FusionRule)(std::vector<LoopRec> &run, size_t at, hipStream_t s);
void register_fusion(FusionRule rule);
void enqueue(LoopRec &&rec);
void flush_all();
bool queue_empty();
void flush_for(ops_reduction h);
void deps_from_args(LoopRec &rec, const ops_arg *args, int n);
// Double buffering for fused chains that rewrite a dat its own stencil reads
// over a range r: spare_base(d, r) is point (0,...,0) of a second allocation
// of d that equals d's current buffer everywhere outside r (copied in full
// when that is not known to hold); the fused kernel then writes all of r
// into it and swap_spare(d, r) makes it d's buffer.  The old buffer, now the
// spare, differs from the new one only inside r, so a following chain over
// the same r needs no copy.  Any other write to d marks the spare out of date.
char *spare_base(ops_dat d, const Range3 &r);
void swap_spare(ops_dat d, const Range3 &r);
void note_external_write(ops_dat d);
void trid_noalloc_setup(int dev);  // trid.hip: per-device address of the no-allocation scratch
// counts a loop that ran inside a fused kernel (its time is the fused entry's)
void note_fused(const char *name, double bytes);
// z planes per workgroup of the z-marching 7-point kernels (16, or
// OPSMI_HEAT_ZCHUNK)
int heat_zchunk_default();

}  // namespace opsmi
