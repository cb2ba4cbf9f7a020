// Host runtime of the MI355X OPS backend: lifecycle, declarations, arguments,
// HBM-resident dats with lazy host mirrors, reductions, statistics and the
// C ABI wrappers for all of them.
#include "runtime.h"

#include <float.h>
#include <limits.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>

namespace opsmi {

static Instance g_inst;
Instance &inst() { return g_inst; }

void fail(int code, const std::string &msg) { throw Error{code, msg}; }

void report(const Error &e) {
  Instance &I = inst();
  if (I.error_mode != 1) {
    fprintf(stderr, "OPS error %d: %s\n", e.code, e.msg.c_str());
    fflush(stderr);
    if (I.error_mode == 0) abort();
  }
  I.last_error = e.code;
  I.last_msg = e.msg;
}

// initial error mode from OPSMI_ERROR_MODE (read when the library loads)
static int g_error_mode_env = [] {
  const char *e = getenv("OPSMI_ERROR_MODE");
  const int m = e ? atoi(e) : 0;
  inst().error_mode = (m >= 0 && m <= 2) ? m : 0;
  return m;
}();

void require_init() {
  if (!inst().initialised) fail(OPS_RUNTIME_CONFIGURATION_ERROR, "ops_init has not been called");
}

static char *copy_str(const char *s) {
  if (!s) s = "";
  size_t n = strlen(s) + 1;
  char *c = (char *)malloc(n);
  memcpy(c, s, n);
  return c;
}

DType dtype_of(const char *type, int type_size) {
  if (!type) return DT_OTHER;
  if ((!strcmp(type, "double") || !strcmp(type, "real(8)") || !strcmp(type, "real(kind=8)") ||
       !strcmp(type, "double precision")) && type_size == 8)
    return DT_F64;
  if ((!strcmp(type, "float") || !strcmp(type, "real") || !strcmp(type, "real(4)") ||
       !strcmp(type, "real(kind=4)")) && type_size == 4)
    return DT_F32;
  if ((!strcmp(type, "int") || !strcmp(type, "int(4)") || !strcmp(type, "integer") ||
       !strcmp(type, "integer(4)") || !strcmp(type, "integer(kind=4)")) && type_size == 4)
    return DT_I32;
  if (!strcmp(type, "half") && type_size == 2) return DT_F16;
  return DT_OTHER;
}

static const char *dt_name(DType d) {
  switch (d) {
    case DT_F64: return "double";
    case DT_F32: return "float";
    case DT_I32: return "int";
    case DT_F16: return "half";
    default: return "other";
  }
}

// ------------------------------------------------------------------ dats
DatView view_of(ops_dat d) {
  DatView v;
  v.base = d->data_d + d->base_offset;
  const bool soa = ((DatImpl *)d->impl)->soa;
  const long cx = soa ? 1 : d->dim;  // AoS: components of a point are adjacent
  v.sx = cx;
  v.sy = cx * d->size[0];
  v.sz = v.sy * d->size[1];
  v.cs = 1;
  if (soa) {
    v.cs = 1;
    for (int n = 0; n < d->block->dims; n++) v.cs *= d->size[n];
  }
  return v;
}

static DatImpl *impl(ops_dat d) { return (DatImpl *)d->impl; }

static size_t dat_lead(ops_dat d) { return (size_t)(d->data_d - impl(d)->alloc); }

void ensure_host(ops_dat d) {
  Instance &I = inst();
  DatImpl *m = impl(d);
  if (!d->data) {
    d->data = (char *)malloc(d->mem);
    if (!d->data) fail(OPS_RUNTIME_ERROR, cat("host allocation failed for dat ", d->name));
    m->host_bytes = d->mem;
    if (d->dirty_hd == 1) d->dirty_hd = 2;  // nothing on host yet
    if (d->dirty_hd == 0) d->dirty_hd = 2;
  }
  if (d->dirty_hd == 2) {
    OPSMI_HIP(hipMemcpyAsync(d->data, d->data_d, d->mem, hipMemcpyDeviceToHost, I.stream));
    OPSMI_HIP(hipStreamSynchronize(I.stream));
    d->dirty_hd = 0;
  }
}

void wait_stream(hipStream_t s) {
  Instance &I = inst();
  if (!I.ev_wait) OPSMI_HIP(hipEventCreateWithFlags(&I.ev_wait, hipEventDisableTiming));
  OPSMI_HIP(hipEventRecord(I.ev_wait, s));
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipEventQuery(I.ev_wait);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) OPSMI_HIP(e);
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      OPSMI_HIP(hipEventSynchronize(I.ev_wait));
      return;
    }
  }
}

void sync_to_device(ops_dat d) {
  if (d->dirty_hd == 1) {
    Instance &I = inst();
    OPSMI_HIP(hipMemcpyAsync(d->data_d, d->data, d->mem, hipMemcpyHostToDevice, I.stream));
    d->dirty_hd = 0;
  }
}

void mark_device_dirty(ops_dat d) { d->dirty_hd = 2; }

// ------------------------------------------------------------------ stencils
std::vector<int> pts_point() { return {0, 0, 0}; }
std::vector<int> pts_5pt() { return {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0}; }
std::vector<int> pts_7pt() {
  return {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
}
std::vector<int> pts_27pt() {
  std::vector<int> v;
  for (int z = -1; z <= 1; z++)
    for (int y = -1; y <= 1; y++)
      for (int x = -1; x <= 1; x++) {
        v.push_back(x);
        v.push_back(y);
        v.push_back(z);
      }
  return v;
}

static bool stencil_has(ops_stencil s, const int *off) {
  for (int p = 0; p < s->points; p++) {
    bool eq = true;
    for (int d = 0; d < 3; d++) {
      int sv = d < s->dims ? s->stencil[p * s->dims + d] : 0;
      if (sv != off[d]) eq = false;
    }
    if (eq) return true;
  }
  return false;
}

// ------------------------------------------------------------------ checks
Range3 check_loop(const char *name, ops_block block, int dim, const int *range, const ops_arg *args,
                  const std::vector<ArgSpec> &specs, DType *dat_dt) {
  require_init();
  if (!block) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": NULL block"));
  if (dim != block->dims || dim < 1 || dim > 3)
    fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dim ", dim, " does not match block ", block->name, " (",
                                   block->dims, "D); this backend supports 1-3D"));
  if (!range) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": NULL range"));
  Range3 r{0, 1, 0, 1, 0, 1};
  r.x0 = range[0];
  r.x1 = range[1];
  if (dim > 1) { r.y0 = range[2]; r.y1 = range[3]; }
  if (dim > 2) { r.z0 = range[4]; r.z1 = range[5]; }
  DType common = DT_OTHER;
  for (size_t i = 0; i < specs.size(); i++) {
    const ops_arg &a = args[i];
    const ArgSpec &s = specs[i];
    switch (s.kind) {
      case AK_IDX:
        if (a.argtype != OPS_ARG_IDX) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg ", i, " must be ops_arg_idx()"));
        break;
      case AK_GBL: {
        if (a.argtype != OPS_ARG_GBL || a.acc != OPS_READ || a.dat != nullptr)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg ", i, " must be an OPS_READ ops_arg_gbl"));
        int want = s.dt == DT_F32 ? 4 : (s.dt == DT_I32 ? 4 : 8);
        if (a.elem_size != want || a.dim < s.dim || !a.data)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": gbl arg ", i, " must hold ", s.dim, " ", dt_name(s.dt)));
        break;
      }
      case AK_RED: {
        if (a.argtype != OPS_ARG_GBL || a.acc != s.red_op || a.dat != nullptr || !a.data)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg ", i, " must be an ops_arg_reduce with access ",
                                         s.red_op == OPS_INC ? "OPS_INC" : (s.red_op == OPS_MIN ? "OPS_MIN" : "OPS_MAX")));
        ops_reduction h = (ops_reduction)a.data;
        RedImpl *ri = (RedImpl *)h->impl;
        DType want = s.dt == DT_OTHER ? common : s.dt;
        if (ri->dt != want)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": reduction handle ", h->name, " has type ", dt_name(ri->dt),
                                         ", kernel reduces ", dt_name(want)));
        int es = want == DT_F64 ? 8 : 4;
        if (a.dim != s.dim || h->size < s.dim * es)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": reduction handle ", h->name, " must hold ", s.dim, " values"));
        break;
      }
      case AK_DAT: {
        if (a.argtype != OPS_ARG_DAT || !a.dat)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg ", i, " must be an ops_arg_dat"));
        ops_dat d = a.dat;
        if (d->block != block)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " is not declared on block ", block->name));
        if (d->dim != 1 || a.dim != 1)
          fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": dat ", d->name, " must have one component"));
        if (d->locked_hd)
          fail(OPS_RUNTIME_ERROR, cat("loop ", name, ": dat ", d->name,
                                      " is locked by ops_dat_get_raw_pointer (release it first)"));
        DType dt = impl(d)->dt;
        DType want = s.dt == DT_OTHER ? (common == DT_OTHER ? dt : common) : s.dt;
        if (dt != want || (dt != DT_F64 && dt != DT_F32))
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " has type ", dt_name(dt), ", expected ",
                                         dt_name(want)));
        if (s.dt == DT_OTHER) common = dt;
        if (s.write && a.acc == OPS_READ)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " is written but passed as OPS_READ"));
        if (!a.stencil) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " has no stencil"));
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (size_t p = 0; p + 2 < s.offsets.size(); p += 3) {
          const int *off = &s.offsets[p];
          if (!stencil_has(a.stencil, off))
            fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": stencil ", a.stencil->name, " of dat ", d->name,
                                           " lacks point (", off[0], ",", off[1], ",", off[2], ") read by the kernel"));
          for (int q = 0; q < 3; q++) {
            lo[q] = std::min(lo[q], off[q]);
            hi[q] = std::max(hi[q], off[q]);
          }
        }
        // bounds: every touched point lies in [d_m, size + d_m) per dimension
        int rl[3] = {r.x0, r.y0, r.z0}, rh[3] = {r.x1, r.y1, r.z1};
        if (r.npts() > 0)
          for (int q = 0; q < dim; q++) {
            int first = d->d_m[q] + d->base[q];
            int last = d->size[q] + d->d_m[q] + d->base[q];
            if (rl[q] + lo[q] < first || rh[q] - 1 + hi[q] >= last)
              fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": range [", rl[q], ",", rh[q], ") in dimension ", q,
                                             " with stencil reach [", lo[q], ",", hi[q], "] leaves dat ", d->name,
                                             " (valid indices [", first, ",", last, "))"));
          }
        break;
      }
    }
  }
  if (dat_dt) *dat_dt = common == DT_OTHER ? DT_F64 : common;
  // host-side newer data goes up before the kernel runs
  for (size_t i = 0; i < specs.size(); i++)
    if (specs[i].kind == AK_DAT) sync_to_device(args[i].dat);
  return r;
}

Range3 check_gen_loop(const char *name, ops_block block, int dim, const int *range, const ops_arg *args, int nargs,
                      const int *types, const int *npts, const int *pts) {
  require_init();
  if (!block) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": NULL block"));
  if (dim != block->dims || dim < 1 || dim > 3)
    fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dim ", dim, " does not match block ", block->name, " (",
                                   block->dims, "D); this backend supports 1-3D"));
  if (!range) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": NULL range"));
  if (nargs < 0 || nargs > OPSMI_GEN_MAX_ARGS || (nargs > 0 && (!args || !types)))
    fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": ", nargs, " arguments (at most ", OPSMI_GEN_MAX_ARGS, ")"));
  Range3 r{0, 1, 0, 1, 0, 1};
  r.x0 = range[0];
  r.x1 = range[1];
  if (dim > 1) { r.y0 = range[2]; r.y1 = range[3]; }
  if (dim > 2) { r.z0 = range[4]; r.z1 = range[5]; }
  if (nargs > 0 && (!npts || !pts)) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": no point lists"));
  const int *pt = pts;
  for (int i = 0; i < nargs; i++) {
    const ops_arg &a = args[i];
    const DType want = (DType)types[i];
    const int es = want == DT_F64 ? 8 : (want == DT_F16 ? 2 : 4);
    const int np = npts[i];
    const int *mine = pt;
    if (np > 0 && a.argtype == OPS_ARG_DAT) pt += 4 * np;
    if (a.argtype == OPS_ARG_IDX) continue;
    if (want != DT_F64 && want != DT_F32 && want != DT_I32 && want != DT_F16)
      fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg ", i, " has no element type (only double, float, int, half)"));
    if (want == DT_F16 && a.argtype != OPS_ARG_DAT)
      fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": arg ", i, ": half values are supported in dats only"));
    if (a.argtype == OPS_ARG_GBL) {
      if (a.acc == OPS_READ) {
        if (!a.data || a.dim < 1 || a.elem_size != es)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": gbl arg ", i, " must hold ", dt_name(want), " values"));
        if (np > 0 && a.dim > np)
          fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": gbl arg ", i, " passes ", a.dim,
                                        " values, the generated kernel holds ", np));
      } else if (a.acc == OPS_INC || a.acc == OPS_MIN || a.acc == OPS_MAX) {
        if (!a.data) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg ", i, " has no reduction handle"));
        ops_reduction h = (ops_reduction)a.data;
        RedImpl *ri = (RedImpl *)h->impl;
        if (ri->dt != want)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": reduction handle ", h->name, " has type ", dt_name(ri->dt),
                                         ", kernel reduces ", dt_name(want)));
        if (a.dim < 1 || a.dim > 4)
          fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": reduction ", h->name, " has ", a.dim,
                                        " components (this backend folds 1 to 4)"));
        if (h->size < a.dim * es)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": reduction handle ", h->name, " must hold ", a.dim, " values"));
      } else {
        fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": gbl arg ", i, " must be OPS_READ or a reduction"));
      }
      continue;
    }
    if (a.argtype != OPS_ARG_DAT) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg ", i, " has unknown kind"));
    if (!a.dat) {
      if (a.opt) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat arg ", i, " is NULL"));
      continue;  // ops_arg_dat_opt switched off: the kernel must not touch it
    }
    ops_dat d = a.dat;
    if (d->block != block)
      fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " is not declared on block ", block->name));
    if (d->locked_hd)
      fail(OPS_RUNTIME_ERROR, cat("loop ", name, ": dat ", d->name, " is locked by ops_dat_get_raw_pointer (release it first)"));
    if (impl(d)->dt != want)
      fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " has type ", dt_name(impl(d)->dt),
                                     ", the kernel declares ", dt_name(want)));
    if (a.dim != d->dim)
      fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " has ", d->dim, " components, the argument says ",
                                     a.dim));
    ops_stencil s = a.stencil;
    if (!s) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": dat ", d->name, " has no stencil"));
    if (s->mgrid_stride)
      fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": stencil ", s->name, " is a multigrid (restrict/prolong) stencil"));
    // strides: 1, or 0 for a dimension the stencil does not advance in
    // (the point's index there is 0: a dat of size 1, lowdim)
    int adv[3] = {1, 1, 1};
    if (s->stride)
      for (int q = 0; q < s->dims && q < 3; q++) {
        if (s->stride[q] != 0 && s->stride[q] != 1)
          fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": stencil ", s->name, " has stride ", s->stride[q],
                                        " in dimension ", q, " (this backend takes strides 0 and 1)"));
        adv[q] = s->stride[q];
      }
    for (int k = 0; k < np; k++) {
      const int *q4 = mine + 4 * k;
      if (q4[0] >= d->dim)
        fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": the kernel reads component ", q4[0], " of dat ", d->name,
                                       ", which has ", d->dim));
      if (!stencil_has(s, q4 + 1))
        fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": stencil ", s->name, " of dat ", d->name, " lacks point (",
                                       q4[1], ",", q4[2], ",", q4[3], ") read by the kernel"));
    }
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int p = 0; p < s->points; p++)
      for (int q = 0; q < 3 && q < s->dims; q++) {
        lo[q] = std::min(lo[q], s->stencil[p * s->dims + q]);
        hi[q] = std::max(hi[q], s->stencil[p * s->dims + q]);
      }
    int rl[3] = {r.x0, r.y0, r.z0}, rh[3] = {r.x1, r.y1, r.z1};
    if (r.npts() > 0)
      for (int q = 0; q < dim; q++) {
        // a dimension of size 1 (an edge dat's) holds its one point at
        // index 0 (ops/c/src/core/ops_lib_core.cpp:466-478)
        int first = d->size[q] == 1 ? 0 : d->d_m[q] + d->base[q];
        int last = d->size[q] == 1 ? 1 : d->size[q] + d->d_m[q] + d->base[q];
        const int a0 = adv[q] ? rl[q] : 0, a1 = adv[q] ? rh[q] - 1 : 0;
        if (a0 + lo[q] < first || a1 + hi[q] >= last)
          fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": range [", rl[q], ",", rh[q], ") in dimension ", q,
                                         adv[q] ? "" : " (stride 0)", " with stencil reach [", lo[q], ",", hi[q],
                                         "] leaves dat ", d->name, " (valid indices [", first, ",", last, "))"));
        // several points on one element: OPS_INC/MIN/MAX combine into it
        // (atomics here); OPS_WRITE lets every point store into it, as the
        // reference's GPU backends do (the reference's apps store one value
        // from every point, e.g. TeaLeaf's initialise_chunk_kernel_zero_x;
        // points storing different values leave any one of them); OPS_RW
        // would be a read-modify-write race and is refused
        if (!adv[q] && rh[q] - rl[q] > 1 && a.acc == OPS_RW)
          fail(OPS_NOT_IMPLEMENTED, cat("loop ", name, ": dat ", d->name, " is read and written with OPS_RW through stencil ",
                                        s->name, ", which maps ", rh[q] - rl[q], " points of dimension ", q,
                                        " to one element (use OPS_INC, OPS_MIN or OPS_MAX)"));
      }
  }
  for (int i = 0; i < nargs; i++)
    if (args[i].argtype == OPS_ARG_DAT && args[i].dat) sync_to_device(args[i].dat);
  return r;
}

const Const &get_const(const char *name) {
  auto it = inst().consts.find(name);
  if (it == inst().consts.end()) fail(OPS_RUNTIME_CONFIGURATION_ERROR, cat("constant ", name, " was not declared with ops_decl_const"));
  return it->second;
}
double const_as_double(const char *name) {
  const Const &c = get_const(name);
  if (c.bytes.size() >= 8 && (c.type == "double" || c.type == "real(8)")) {
    double v;
    memcpy(&v, c.bytes.data(), 8);
    return v;
  }
  if (c.bytes.size() >= 4 && c.type == "float") {
    float v;
    memcpy(&v, c.bytes.data(), 4);
    return v;
  }
  if (c.bytes.size() >= 4 && c.type == "int") {
    int v;
    memcpy(&v, c.bytes.data(), 4);
    return v;
  }
  fail(OPS_INVALID_ARGUMENT, cat("constant ", name, " has unsupported type ", c.type));
}
int const_as_int(const char *name) {
  const Const &c = get_const(name);
  if (c.type != "int" || c.bytes.size() < 4) fail(OPS_INVALID_ARGUMENT, cat("constant ", name, " must be an int"));
  int v;
  memcpy(&v, c.bytes.data(), 4);
  return v;
}

// ------------------------------------------------------------------ reductions
void ensure_partials(size_t bytes) {
  Instance &I = inst();
  if (bytes <= I.partials_bytes) return;
  if (I.d_partials) {
    OPSMI_HIP(hipStreamSynchronize(I.stream));
    OPSMI_HIP(hipFree(I.d_partials));
  }
  size_t nb = std::max(bytes, (size_t)1 << 20);
  OPSMI_HIP(hipMalloc(&I.d_partials, nb));
  I.partials_bytes = nb;
}

RedCtl red_prepare(const ops_arg &a, int dim, DType dt, int nblocks) {
  Instance &I = inst();
  ops_reduction h = (ops_reduction)a.data;
  RedImpl *ri = (RedImpl *)h->impl;
  ensure_partials((size_t)nblocks * dim * 8);
  RedCtl c{};
  c.slot = ri->d_slot;
  c.mirror = ri->d_mirror;
  c.partials = I.d_partials;
  c.partials2 = I.d_partials2;
  c.use_init = ri->dev_valid ? 0 : 1;
  for (int q = 0; q < dim && q < 4; q++) {
    if (dt == DT_F64) memcpy(&c.init[q], h->data + 8 * q, 8);
    else if (dt == DT_F32) { float f; memcpy(&f, h->data + 4 * q, 4); c.init[q] = f; }
    else if (dt == DT_I32) { memcpy(&c.init_i[q], h->data + 4 * q, 4); }
  }
  c.seq = ri->d_seq;
  c.seq_val = ++ri->seq_issued;
  ri->dev_valid = true;
  return c;
}

// ------------------------------------------------------------------ loops
double transfer_bytes(const ops_arg *args, int nargs, const Range3 &r) {
  // ops_compute_transfer (ops/c/src/core/ops_lib_core.cpp:1687-1697)
  double tot = 0.0;
  for (int i = 0; i < nargs; i++) {
    const ops_arg &a = args[i];
    if (a.argtype != OPS_ARG_DAT || !a.dat) continue;
    double n = (double)r.npts() * a.dat->elem_size;
    tot += n * ((a.acc == OPS_READ || a.acc == OPS_WRITE) ? 1.0 : 2.0);
  }
  return tot;
}

int grid_blocks(long rows, int rows_per_block) {
  long need = (rows + rows_per_block - 1) / rows_per_block;
  long cap = (long)inst().num_cus * 8;
  if (need < 1) need = 1;
  return (int)std::min(need, cap);
}

static void drain_events() {
  Instance &I = inst();
  if (I.pending.empty()) return;
  OPSMI_HIP(hipStreamSynchronize(I.stream));
  for (auto &p : I.pending) {
    float ms = 0.f;
    OPSMI_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    I.stats[p.name].ms += ms;
    I.ev_free.push_back(p.a);
    I.ev_free.push_back(p.b);
  }
  I.pending.clear();
}

// timing events (diags > 1): only their timestamps are read, so they carry
// no system-scope fence (recording one with the fence cost about 3.4 us of
// a 2048^2 Laplace iteration)
static hipEvent_t get_event() {
  Instance &I = inst();
  if (I.ev_free.empty()) {
    hipEvent_t e;
    OPSMI_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return e;
  }
  hipEvent_t e = I.ev_free.back();
  I.ev_free.pop_back();
  return e;
}

void run_loop(const char *name, double bytes, const std::function<void(hipStream_t)> &fn) {
  Instance &I = inst();
  LoopStat &s = I.stats[name ? name : "?"];
  s.count++;
  s.launches++;
  s.bytes += bytes;
  if (I.diags > 1) {
    hipEvent_t a = get_event(), b = get_event();
    OPSMI_HIP(hipEventRecord(a, I.stream));
    fn(I.stream);
    OPSMI_HIP(hipEventRecord(b, I.stream));
    I.pending.push_back({name ? name : "?", a, b});
    if (I.pending.size() > 4096) drain_events();
  } else {
    fn(I.stream);
  }
  OPSMI_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ lazy queue
static std::vector<FusionRule> &rules() {
  static std::vector<FusionRule> r;
  return r;
}
void register_fusion(FusionRule rule) { rules().push_back(rule); }

static std::vector<LoopRec> g_queue;

void deps_from_args(LoopRec &rec, const ops_arg *args, int n) {
  for (int i = 0; i < n; i++) {
    const ops_arg &a = args[i];
    if (a.argtype == OPS_ARG_DAT && a.dat) {
      if (a.acc != OPS_WRITE) rec.reads.push_back(a.dat);
      if (a.acc != OPS_READ) rec.writes.push_back(a.dat);
    } else if (a.argtype == OPS_ARG_GBL && a.data && (a.acc == OPS_INC || a.acc == OPS_MIN || a.acc == OPS_MAX)) {
      if (!rec.red) rec.red = (ops_reduction)a.data;
      else rec.more_reds.push_back((ops_reduction)a.data);
    }
  }
}

static bool meet(const std::vector<ops_dat> &x, const std::vector<ops_dat> &y) {
  for (ops_dat p : x)
    for (ops_dat q : y)
      if (p == q) return true;
  return false;
}
static bool reduces_into(const LoopRec &L, ops_reduction h) {
  if (!h) return false;
  if (L.red == h) return true;
  for (ops_reduction q : L.more_reds)
    if (q == h) return true;
  return false;
}
static bool share_red(const LoopRec &a, const LoopRec &b) {
  if (reduces_into(b, a.red)) return true;
  for (ops_reduction q : a.more_reds)
    if (reduces_into(b, q)) return true;
  return false;
}
// a (earlier) must run before b
static bool conflict(const LoopRec &a, const LoopRec &b) {
  return meet(a.writes, b.reads) || meet(a.writes, b.writes) || meet(a.reads, b.writes) || share_red(a, b);
}

static void invalidate_spare(ops_dat d) {
  DatImpl *m = (DatImpl *)d->impl;
  if (m) m->spare_synced = false;
}

static void execute(std::vector<LoopRec> &run) {
  Instance &I = inst();
  size_t i = 0;
  while (i < run.size()) {
    int n = 0;
    if (I.fusion)
      for (FusionRule r : rules()) {
        n = r(run, i, I.stream);
        if (n > 0) break;
      }
    if (n > 0) {
      i += n;
      continue;
    }
    LoopRec &L = run[i];
    run_loop(L.name.c_str(), L.bytes, L.launch);
    for (ops_dat d : L.writes) invalidate_spare(d);
    i++;
  }
}

void enqueue(LoopRec &&rec) {
  g_queue.push_back(std::move(rec));
  if (g_queue.size() > 64) flush_all();
}

bool queue_empty() { return g_queue.empty(); }

void flush_all() {
  if (g_queue.empty()) return;
  std::vector<LoopRec> run;
  run.swap(g_queue);
  execute(run);
}

void flush_for(ops_reduction h) {
  const long n = (long)g_queue.size();
  if (n == 0) return;
  std::vector<char> need(n, 0);
  bool any = false;
  for (long i = n - 1; i >= 0; i--) {
    if (reduces_into(g_queue[i], h)) {
      need[i] = 1;
      any = true;
      continue;
    }
    for (long j = i + 1; j < n; j++)
      if (need[j] && conflict(g_queue[i], g_queue[j])) {
        need[i] = 1;
        break;
      }
  }
  if (!any) return;
  std::vector<LoopRec> run, rest;
  for (long i = 0; i < n; i++) (need[i] ? run : rest).push_back(std::move(g_queue[i]));
  g_queue.swap(rest);
  // the handle's result is read right after this flush: when one loop of
  // the run reduces into it, with no other handle, starting from the OPS
  // initial value, its partials may be folded by the host
  // (with_host_fold, host_fold_result)
  Instance &I = inst();
  const LoopRec *only = nullptr;
  int count = 0;
  for (const LoopRec &L : run)
    if (reduces_into(L, h)) {
      count++;
      only = &L;
    }
  RedImpl *ri = (RedImpl *)h->impl;
  I.hf_used = false;
  I.hf_slot = nullptr;
  if (I.host_fold && count == 1 && only->red == h && only->more_reds.empty() && only->ctl.use_init && ri->dt == DT_F64 &&
      only->ctl.slot == ri->d_slot) {
    I.hf_slot = ri->d_slot;
    memcpy(I.hf_init, only->ctl.init, sizeof(I.hf_init));
  }
  // disarmed however execute() ends (an error thrown out of a launch
  // included), so no later flush takes the host fold for this handle
  struct Disarm {
    Instance &I;
    ~Disarm() { I.hf_slot = nullptr; }
  } disarm{I};
  try {
    execute(run);
  } catch (...) {
    I.hf_used = false;
    throw;
  }
}

static bool same_range6(const int *a, const Range3 &r) {
  return a[0] == r.x0 && a[1] == r.x1 && a[2] == r.y0 && a[3] == r.y1 && a[4] == r.z0 && a[5] == r.z1;
}

char *spare_base(ops_dat d, const Range3 &r) {
  Instance &I = inst();
  DatImpl *m = (DatImpl *)d->impl;
  size_t lead = (size_t)(d->data_d - m->alloc);
  if (!m->spare) {
    OPSMI_HIP(hipMalloc(&m->spare, d->mem + lead + 256));
    m->spare_synced = false;
  }
  // the spare may differ from the buffer inside the last swapped range only
  if (!m->spare_synced || (m->spare_ranged && !same_range6(m->spare_range, r))) {
    OPSMI_HIP(hipMemcpyAsync(m->spare + lead, d->data_d, d->mem, hipMemcpyDeviceToDevice, I.stream));
    m->spare_synced = true;
    m->spare_ranged = false;
  }
  return m->spare + lead + d->base_offset;
}

void swap_spare(ops_dat d, const Range3 &r) {
  DatImpl *m = (DatImpl *)d->impl;
  size_t lead = (size_t)(d->data_d - m->alloc);
  std::swap(m->alloc, m->spare);
  d->data_d = m->alloc + lead;
  const int rr[6] = {r.x0, r.x1, r.y0, r.y1, r.z0, r.z1};
  memcpy(m->spare_range, rr, sizeof(rr));
  m->spare_ranged = true;
}

void note_external_write(ops_dat d) { invalidate_spare(d); }

void note_fused(const char *name, double bytes) {
  LoopStat &s = inst().stats[name ? name : "?"];
  s.count++;
  s.bytes += bytes;
}

int heat_zchunk_default() {
  const char *env = getenv("OPSMI_HEAT_ZCHUNK");
  return env ? std::max(1, atoi(env)) : 16;
}

static double wall_now() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

}  // namespace opsmi

using namespace opsmi;

// =================================================================== C ABI
extern "C" {

void ops_init(const int argc, const char *const argv[], const int diags_level) {
  OPSMI_API_BEGIN
  Instance &I = inst();
  if (I.initialised) return;
  I.diags = diags_level;
  int dev = 0;
  bool tiling = false;
  for (int n = 1; n < argc && argv; n++) {
    const char *a = argv[n];
    if (!a) continue;
    const char *p;
    if ((p = strstr(a, "OPS_DIAGS="))) I.diags = atoi(p + 10);
    if ((p = strstr(a, "OPS_DEVICE="))) dev = atoi(p + 11);
    if ((p = strstr(a, "OPS_BLOCK_SIZE_X="))) I.block_x = atoi(p + 17);
    if ((p = strstr(a, "OPS_BLOCK_SIZE_Y="))) I.block_y = atoi(p + 17);
    if ((p = strstr(a, "OPS_BLOCK_SIZE_Z="))) I.block_z = atoi(p + 17);
    if (strstr(a, "OPS_TILING")) tiling = true;
    if ((p = strstr(a, "OPS_TILING_MAXDEPTH="))) I.tiling_maxdepth = atoi(p + 20);
  }
  if (I.block_x < 0 || I.block_y < 0 || I.block_z < 0 || I.block_y > 1024 || I.block_z > 1024)
    fail(OPS_RUNTIME_CONFIGURATION_ERROR, cat("OPS_BLOCK_SIZE_{X,Y,Z}: ", I.block_x, ", ", I.block_y, ", ",
                                              I.block_z, " out of range"));
  int ndev = 0;
  OPSMI_HIP(hipGetDeviceCount(&ndev));
  if (ndev < 1) fail(OPS_RUNTIME_CONFIGURATION_ERROR, "no HIP device available");
  if (dev < 0 || dev >= ndev) fail(OPS_RUNTIME_CONFIGURATION_ERROR, cat("OPS_DEVICE=", dev, " out of range"));
  OPSMI_HIP(hipSetDevice(dev));
  I.device = dev;
  hipDeviceProp_t prop;
  OPSMI_HIP(hipGetDeviceProperties(&prop, dev));
  I.num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  const char *fz = getenv("OPSMI_FUSION");
  I.fusion = fz ? atoi(fz) : 2;
  const char *hf = getenv("OPSMI_HOST_FOLD");
  I.host_fold = hf ? atoi(hf) != 0 : true;
  const char *ga = getenv("OPSMI_GEN_ADJ");
  I.gen_spread = ga && atoi(ga) == 0;
  if (tiling) I.fusion = 2;
  if (I.diags > 0 && (tiling || I.block_y_set() || I.block_z_set()))
    printf(" OPS (MI355X): loop-chain fusion level %d, generated-loop workgroups %d rows (2D) / %d planes (3D)\n",
           I.fusion, I.gen_rows_2d(), I.gen_planes_3d());
  OPSMI_HIP(hipStreamCreateWithFlags(&I.stream, hipStreamNonBlocking));
  trid_noalloc_setup(dev);
  ensure_partials((size_t)1 << 20);
  OPSMI_HIP(hipMalloc(&I.d_partials2, (size_t)1 << 20));
  OPSMI_HIP(hipHostMalloc(&I.h_pinned, 4096));
  OPSMI_HIP(hipHostMalloc((void **)&I.h_hfold, 8 * 4096, hipHostMallocMapped | hipHostMallocCoherent));
  OPSMI_HIP(hipHostGetDevicePointer(&I.d_hfold, I.h_hfold, 0));
  I.initialised = true;
  OPSMI_API_END()
}

void ops_exit(void) {
  OPSMI_API_BEGIN
  Instance &I = inst();
  if (!I.initialised) return;
  flush_all();
  OPSMI_HIP(hipStreamSynchronize(I.stream));
  for (ops_dat d : I.dats) {
    if (!d) continue;
    ops_free_dat(d);
  }
  I.dats.clear();
  for (ops_reduction r : I.reductions) {
    RedImpl *ri = (RedImpl *)r->impl;
    (void)hipFree(ri->d_slot);
    (void)hipHostFree(ri->h_mirror);
    delete ri;
    free(r->data);
    free(r->name);
    free(r->type);
    delete r;
  }
  I.reductions.clear();
  for (ops_stencil s : I.stencils) {
    free(s->stencil);
    free(s->stride);
    free((char *)s->name);
    delete s;
  }
  I.stencils.clear();
  for (ops_block b : I.blocks) {
    free((char *)b->name);
    delete b;
  }
  I.blocks.clear();
  for (ops_halo h : I.halos) delete h;
  I.halos.clear();
  for (ops_halo_group g : I.groups) {
    free(g->halos);
    delete g;
  }
  I.groups.clear();
  for (auto &p : I.pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  I.pending.clear();
  for (hipEvent_t e : I.ev_free) (void)hipEventDestroy(e);
  I.ev_free.clear();
  (void)hipFree(I.d_partials);
  I.d_partials = nullptr;
  I.partials_bytes = 0;
  (void)hipFree(I.d_partials2);
  I.d_partials2 = nullptr;
  (void)hipHostFree(I.h_pinned);
  (void)hipHostFree(I.h_hfold);
  I.h_hfold = nullptr;
  I.d_hfold = nullptr;
  I.hf_slot = nullptr;
  I.hf_used = false;
  if (I.ev_wait) (void)hipEventDestroy(I.ev_wait);
  I.ev_wait = nullptr;
  (void)hipStreamDestroy(I.stream);
  I.consts.clear();
  I.stats.clear();
  I.initialised = false;
  OPSMI_API_END()
}

void ops_partition(const char *routine) {
  (void)routine;
}

ops_block ops_decl_block(int dims, const char *name) {
  OPSMI_API_BEGIN
  require_init();
  if (dims < 1 || dims > 3) fail(OPS_NOT_IMPLEMENTED, cat("ops_decl_block: ", dims, "D blocks are not supported (1-3D)"));
  ops_block b = new ops_block_core;
  b->index = (int)inst().blocks.size();
  b->dims = dims;
  b->name = copy_str(name);
  b->instance.unused_ = &inst();
  inst().blocks.push_back(b);
  return b;
  OPSMI_API_END(nullptr)
}

ops_dat ops_decl_dat_char(ops_block block, int dim, int *dsize, int *base, int *d_m, int *d_p, int *stride,
                          char *data, int type_size, const char *type, const char *name) {
  OPSMI_API_BEGIN
  require_init();
  if (!block) fail(OPS_INVALID_ARGUMENT, "ops_decl_dat: NULL block");
  if (dim <= 0)
    fail(OPS_INVALID_ARGUMENT, cat("Error: ops_decl_dat -- negative/zero number of items per grid point in dataset: ", name));
  // a stride other than 1 makes a multigrid dat, read through restrict /
  // prolong stencils, which this backend does not run
  for (int n = 0; stride && n < block->dims; n++)
    if (stride[n] != 1)
      fail(OPS_NOT_IMPLEMENTED, cat("ops_decl_dat: dat ", name ? name : "?", " has stride ", stride[n],
                                    " in dimension ", n, " (multigrid dats are not supported)"));
  Instance &I = inst();
  ops_dat d = new ops_dat_core;
  memset(d, 0, sizeof(*d));
  DatImpl *m = new DatImpl;
  d->impl = m;
  d->index = (int)I.dats.size();
  d->block = block;
  d->dim = dim;
  d->type_size = type_size;
  d->elem_size = type_size * dim;
  d->name = copy_str(name);
  d->type = copy_str(type);
  m->dt = dtype_of(type, type_size);
  int nd = block->dims;
  // metadata as ops_dat_init_metadata_core (ops/c/src/core/ops_lib_core.cpp:445-546)
  for (int n = 0; n < OPS_MAX_DIM; n++) {
    d->size[n] = 1;
    d->stride[n] = 1;
  }
  for (int n = 0; n < nd; n++) {
    if (d_m[n] > 0)
      fail(OPS_INVALID_ARGUMENT, cat("Error: ops_decl_dat -- Non negative d_m during declaration of: ", name));
    if (d_p[n] < 0)
      fail(OPS_INVALID_ARGUMENT, cat("Error: ops_decl_dat -- Non positive d_p during declaration of: ", name));
    // a dimension of size 1 is an edge (lowdim) dimension: one point at
    // index 0, no halo
    const bool edge = dsize[n] == 1;
    if (!edge) {
      d->size[n] = dsize[n] - d_m[n] + d_p[n];
    } else {
      d->size[n] = 1;
      d->e_dat = 1;
    }
    d->base[n] = d->size[n] == 1 ? 0 : base[n];
    d->d_m[n] = edge ? 0 : d_m[n];
    d->d_p[n] = edge ? 0 : d_p[n];
    if (stride) d->stride[n] = stride[n];
  }
  // multi-component dats follow the instance's layout (ops_set_soa,
  // ops/c/src/sequential/ops_host_singlenode.cpp:97): SoA keeps each
  // component in its own array
  m->soa = I.soa && dim > 1;
  const size_t unit = m->soa ? (size_t)d->type_size : (size_t)d->elem_size;  // bytes per point of one array
  // x padding so each row starts on a 256-byte boundary (rows at least 1 KiB)
  const size_t A = 256;
  size_t row = (size_t)d->size[0] * unit;
  int user_x = d->size[0];
  if (d->size[0] > 1 && row >= 1024 && A % unit == 0) {
    size_t padded = (row + A - 1) / A * A;
    d->x_pad = (int)((padded - row) / unit);
    d->size[0] += d->x_pad;
    d->d_p[0] += d->x_pad;
  }
  size_t bytes = (size_t)d->elem_size;
  for (int n = 0; n < nd; n++) bytes *= d->size[n];
  d->mem = bytes;
  d->base_offset = 0;
  size_t cum = 1;
  for (int n = 0; n < nd; n++) {
    d->base_offset += unit * cum * (size_t)(-d->base[n] - d->d_m[n]);
    cum *= d->size[n];
  }
  // leading pad so that point x = 0 is 256-byte aligned
  size_t lead = (A - ((size_t)(-d->d_m[0]) * unit) % A) % A;
  OPSMI_HIP(hipMalloc(&m->alloc, bytes + lead + A));
  d->data_d = m->alloc + lead;
  OPSMI_HIP(hipMemset(m->alloc, 0, bytes + lead + A));
  if (data) {
    // user array: same layout without the x padding (SoA: one such array
    // per component, one after the other)
    size_t upitch = (size_t)user_x * unit;
    size_t rows = 1;
    for (int n = 1; n < nd; n++) rows *= d->size[n];
    const int ncomp = m->soa ? dim : 1;
    for (int c = 0; c < ncomp; c++)
      OPSMI_HIP(hipMemcpy2D(d->data_d + (size_t)c * row * rows + (size_t)c * d->x_pad * unit * rows,
                            row + (size_t)d->x_pad * unit, data + (size_t)c * upitch * rows, upitch, upitch, rows,
                            hipMemcpyHostToDevice));
  }
  d->data = nullptr;
  d->user_managed = 0;
  d->dirty_hd = 2;
  d->hdf5_file = "none";
  I.dats.push_back(d);
  return d;
  OPSMI_API_END(nullptr)
}

void ops_free_dat(ops_dat d) {
  OPSMI_API_BEGIN
  if (!d) return;
  Instance &I = inst();
  flush_all();
  OPSMI_HIP(hipStreamSynchronize(I.stream));
  DatImpl *m = (DatImpl *)d->impl;
  if (m) {
    (void)hipFree(m->alloc);
    if (m->spare) (void)hipFree(m->spare);
    delete m;
  }
  free(d->data);
  if (d->index >= 0 && d->index < (int)I.dats.size()) I.dats[d->index] = nullptr;
  free((char *)d->name);
  free((char *)d->type);
  delete d;
  OPSMI_API_END()
}

ops_stencil ops_decl_strided_stencil(int dims, int points, int *sten, int *stride, const char *name) {
  OPSMI_API_BEGIN
  require_init();
  if (dims < 1 || points < 1 || !sten) fail(OPS_INVALID_ARGUMENT, cat("ops_decl_stencil: invalid stencil ", name));
  ops_stencil s = new ops_stencil_core;
  s->index = (int)inst().stencils.size();
  s->dims = dims;
  s->points = points;
  s->name = copy_str(name);
  s->stencil = (int *)malloc(sizeof(int) * dims * points);
  memcpy(s->stencil, sten, sizeof(int) * dims * points);
  s->stride = (int *)malloc(sizeof(int) * dims);
  for (int d = 0; d < dims; d++) s->stride[d] = stride ? stride[d] : 1;
  s->mgrid_stride = nullptr;
  s->type = 0;
  inst().stencils.push_back(s);
  return s;
  OPSMI_API_END(nullptr)
}

ops_stencil ops_decl_stencil(int dims, int points, int *stencil, const char *name) {
  return ops_decl_strided_stencil(dims, points, stencil, nullptr, name);
}

ops_reduction ops_decl_reduction_handle(int size, const char *type, const char *name) {
  OPSMI_API_BEGIN
  require_init();
  if (size <= 0) fail(OPS_INVALID_ARGUMENT, cat("ops_decl_reduction_handle: bad size for ", name));
  ops_reduction r = new ops_reduction_core;
  memset(r, 0, sizeof(*r));
  r->size = size;
  r->data = (char *)calloc(size, 1);
  r->initialized = 0;
  r->index = (int)inst().reductions.size();
  r->name = copy_str(name);
  r->type = copy_str(type);
  r->instance.unused_ = &inst();
  RedImpl *ri = new RedImpl;
  int ts = 8;
  if (!strcmp(type, "float") || !strcmp(type, "int") || !strcmp(type, "real") || !strcmp(type, "integer")) ts = 4;
  ri->dt = dtype_of(type, ts);
  OPSMI_HIP(hipMalloc(&ri->d_slot, std::max(size, 64)));
  // the result is read from here after the stream completes: no copy kernel
  // values first, the sequence word on its own 64-byte line after them
  const size_t vbytes = ((size_t)std::max(size, 64) + 63) / 64 * 64;
  OPSMI_HIP(hipHostMalloc(&ri->h_mirror, vbytes + 64, hipHostMallocMapped | hipHostMallocCoherent));
  OPSMI_HIP(hipHostGetDevicePointer(&ri->d_mirror, ri->h_mirror, 0));
  ri->h_seq = (volatile unsigned long long *)((char *)ri->h_mirror + vbytes);
  ri->d_seq = (unsigned long long *)((char *)ri->d_mirror + vbytes);
  *ri->h_seq = 0;
  r->impl = ri;
  inst().reductions.push_back(r);
  return r;
  OPSMI_API_END(nullptr)
}

ops_halo ops_decl_halo(ops_dat from, ops_dat to, int *iter_size, int *from_base, int *to_base, int *from_dir,
                       int *to_dir) {
  OPSMI_API_BEGIN
  require_init();
  if (!from || !to) fail(OPS_INVALID_ARGUMENT, "ops_decl_halo: NULL dat");
  ops_halo h = new ops_halo_core;
  memset(h, 0, sizeof(*h));
  int nd = from->block->dims;
  for (int d = 0; d < nd; d++) {
    h->iter_size[d] = iter_size[d];
    h->from_base[d] = from_base[d];
    h->to_base[d] = to_base[d];
    h->from_dir[d] = from_dir[d];
    h->to_dir[d] = to_dir[d];
  }
  for (int d = nd; d < OPS_MAX_DIM; d++) {
    h->iter_size[d] = 1;
    h->from_dir[d] = h->to_dir[d] = d + 1;
  }
  h->from = from;
  h->to = to;
  h->index = (int)inst().halos.size();
  inst().halos.push_back(h);
  return h;
  OPSMI_API_END(nullptr)
}

ops_halo_group ops_decl_halo_group(int nhalos, ops_halo *halos) {
  OPSMI_API_BEGIN
  require_init();
  ops_halo_group g = new ops_halo_group_core;
  g->nhalos = nhalos;
  g->halos = (ops_halo *)malloc(sizeof(ops_halo) * std::max(nhalos, 1));
  for (int i = 0; i < nhalos; i++) g->halos[i] = halos[i];
  g->index = (int)inst().groups.size();
  g->instance.unused_ = &inst();
  inst().groups.push_back(g);
  return g;
  OPSMI_API_END(nullptr)
}

void ops_decl_const_char(void *instance, int dim, const char *type, int type_size, char *data, const char *name) {
  OPSMI_API_BEGIN
  require_init();
  if (instance && instance != (void *)&inst()) fail(OPS_INVALID_ARGUMENT, "ops_decl_const: unknown OPS instance");
  if (!name || dim < 0 || type_size <= 0 || (dim > 0 && !data)) fail(OPS_INVALID_ARGUMENT, "ops_decl_const: bad arguments");
  Const c;
  c.type = type ? type : "";
  c.bytes.assign(data, data + (size_t)dim * type_size);
  inst().consts[name] = c;
  OPSMI_API_END()
}

// ------------------------------------------------------------------ args
ops_arg ops_arg_dat(ops_dat dat, int dim, ops_stencil stencil, const char *type, ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof(a));
  OPSMI_API_BEGIN
  (void)type;
  a.argtype = OPS_ARG_DAT;
  a.dat = dat;
  a.stencil = stencil;
  if (acc == OPS_WRITE && stencil && stencil->points != 1)
    fail(OPS_INVALID_ARGUMENT,
         "Error: OPS does not support OPS_WRITE arguments with a non (0,0,0) stencil due to potential race conditions");
  if (dat) {
    a.data = dat->data;
    a.data_d = dat->data_d;
    a.elem_size = dat->elem_size;
  }
  a.acc = acc;
  a.opt = 1;
  a.dim = dim;
  return a;
  OPSMI_API_END(a)
}

ops_arg ops_arg_dat_opt(ops_dat dat, int dim, ops_stencil stencil, const char *type, ops_access acc, int flag) {
  ops_arg a = ops_arg_dat(dat, dim, stencil, type, acc);
  a.opt = flag;
  return a;
}

ops_arg ops_arg_idx(void) {
  ops_arg a;
  memset(&a, 0, sizeof(a));
  a.argtype = OPS_ARG_IDX;
  return a;
}

ops_arg ops_arg_gbl_char(char *data, int dim, int size, ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof(a));
  a.argtype = OPS_ARG_GBL;
  a.dim = dim;
  a.data = data;
  a.acc = acc;
  a.elem_size = size;
  return a;
}

// ops_arg_reduce_core (ops/c/src/core/ops_lib_core.cpp:856-965)
ops_arg ops_arg_reduce(ops_reduction handle, int dim, const char *type, ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof(a));
  OPSMI_API_BEGIN
  if (!handle) fail(OPS_INVALID_ARGUMENT, "ops_arg_reduce: NULL handle");
  a.argtype = OPS_ARG_GBL;
  a.dim = dim;
  a.data = (char *)handle;
  a.acc = acc;
  if (handle->initialized == 0) {
    handle->initialized = 1;
    handle->acc = acc;
    if (acc == OPS_INC) memset(handle->data, 0, handle->size);
    DType dt = dtype_of(type, (!strcmp(type, "float") || !strcmp(type, "int")) ? 4 : 8);
    if (dt == DT_F64) {
      for (int i = 0; i < handle->size / 8; i++) {
        double v = acc == OPS_MIN ? DBL_MAX : -DBL_MAX;
        if (acc == OPS_MIN || acc == OPS_MAX) memcpy(handle->data + 8 * i, &v, 8);
      }
    } else if (dt == DT_F32) {
      for (int i = 0; i < handle->size / 4; i++) {
        float v = acc == OPS_MIN ? FLT_MAX : -FLT_MAX;
        if (acc == OPS_MIN || acc == OPS_MAX) memcpy(handle->data + 4 * i, &v, 4);
      }
    } else if (dt == DT_I32) {
      for (int i = 0; i < handle->size / 4; i++) {
        int v = acc == OPS_MIN ? INT_MAX : INT_MIN;
        if (acc == OPS_MIN || acc == OPS_MAX) memcpy(handle->data + 4 * i, &v, 4);
      }
    } else {
      fail(OPS_NOT_IMPLEMENTED, "Error, reduction type not recognised, please add in ops_lib_core.cpp");
    }
    ((RedImpl *)handle->impl)->dev_valid = false;
  } else if (handle->acc != acc) {
    fail(OPS_INVALID_ARGUMENT, cat("Error: ops_reduction handle ", handle->name,
                                   " was already used with a different access type"));
  }
  return a;
  OPSMI_API_END(a)
}

// ------------------------------------------------------------------ sync
void ops_execute(void) {
  OPSMI_API_BEGIN
  if (inst().initialised) flush_all();
  OPSMI_API_END()
}

// The device's final fold (k_red_final over n partials of dim components,
// engine.h) restated on the host in the same order, so the result has the
// same bits: for each component, "thread" t < 256 combines partials t,
// t + 256, ... in that order, each wave of 64 such values folds by xor
// butterflies (offsets 32 .. 1), the four wave values fold in order, and the
// result combines with the initial value.  The loops run across the
// independent threads and lanes, which leaves every value's order of
// operations as it is.
extern "C++" {
template <int OP>
static inline double fold_op(double a, double b) {
  return OP == OPS_MIN ? std::fmin(a, b) : (OP == OPS_MAX ? std::fmax(a, b) : a + b);
}
template <int OP>
static double host_fold(const double *in, unsigned n, int dim, int q, double init) {
  const double ident = OP == OPS_MIN ? HUGE_VAL : (OP == OPS_MAX ? -HUGE_VAL : 0.0);
  double v[256];
  for (int t = 0; t < 256; t++) v[t] = ident;
  unsigned i0 = 0;
  for (; i0 + 256 <= n; i0 += 256)
    for (int t = 0; t < 256; t++) v[t] = fold_op<OP>(v[t], in[(size_t)(i0 + t) * dim + q]);
  for (unsigned t = 0; i0 + t < n; t++) v[t] = fold_op<OP>(v[t], in[(size_t)(i0 + t) * dim + q]);
  double sh[4];
  for (int w = 0; w < 4; w++) {
    double a[64], b[64];
    for (int l = 0; l < 64; l++) a[l] = v[64 * w + l];
    for (int off = 32; off > 0; off >>= 1) {
      for (int l = 0; l < 64; l++) b[l] = fold_op<OP>(a[l], a[l ^ off]);
      for (int l = 0; l < 64; l++) a[l] = b[l];
    }
    sh[w] = a[0];
  }
  double sum = sh[0];
  for (int w = 1; w < 4; w++) sum = fold_op<OP>(sum, sh[w]);
  return fold_op<OP>(init, sum);
}
}  // extern "C++"
static void host_fold_result(ops_reduction handle) {
  Instance &I = inst();
  const unsigned n = I.hf_n;
  const int dim = I.hf_dim;
  const size_t total = (size_t)n * dim;
  // the partials arrive as the workgroups finish; the stream is waited for
  // after 50 ms, and a partial still missing then is an error
  const volatile unsigned long long *raw = (const volatile unsigned long long *)I.h_hfold;
  const auto t0 = std::chrono::steady_clock::now();
  size_t k = 0;
  bool waited = false;
  while (k < total) {
    if (raw[k] != kHostFoldEmpty) {
      k++;
      continue;
    }
    if (waited) fail(OPS_RUNTIME_ERROR, cat("reduction ", handle->name, ": a workgroup partial never arrived"));
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      wait_stream(I.stream);
      waited = true;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  const double *in = (const double *)I.h_hfold;
  double res[4];
  for (int q = 0; q < dim; q++) {
    if (handle->acc == OPS_MIN) res[q] = host_fold<OPS_MIN>(in, n, dim, q, I.hf_init[q]);
    else if (handle->acc == OPS_MAX) res[q] = host_fold<OPS_MAX>(in, n, dim, q, I.hf_init[q]);
    else res[q] = host_fold<OPS_INC>(in, n, dim, q, I.hf_init[q]);
  }
  memcpy(handle->data, res, sizeof(double) * dim);
}

void ops_reduction_result_char(ops_reduction handle, int type_size, char *ptr) {
  OPSMI_API_BEGIN
  (void)type_size;
  require_init();
  Instance &I = inst();
  flush_for(handle);
  RedImpl *ri = (RedImpl *)handle->impl;
  if (I.hf_used) {
    I.hf_used = false;
    host_fold_result(handle);
  } else if (ri->dev_valid) {
    // the last fold into the handle stores its sequence number after the
    // values: poll it (no event round trip); after 50 ms, or if it never
    // comes, wait for the stream instead
    const unsigned long long want = ri->seq_issued;
    const auto t0 = std::chrono::steady_clock::now();
    bool seen = false;
    for (long k = 0;; k++) {
      if (__atomic_load_n(ri->h_seq, __ATOMIC_ACQUIRE) >= want) {
        seen = true;
        break;
      }
      if ((k & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
    }
    if (!seen) wait_stream(I.stream);
    memcpy(handle->data, ri->h_mirror, handle->size);
  }
  memcpy(ptr, handle->data, handle->size);
  handle->initialized = 0;
  ri->dev_valid = false;
  OPSMI_API_END()
}

// ------------------------------------------------------------------ data access
void ops_dat_get_extents(ops_dat dat, int part, int *disp, int *size) {
  (void)part;
  if (disp)
    for (int d = 0; d < dat->block->dims; d++) disp[d] = 0;
  if (size)
    for (int d = 0; d < dat->block->dims; d++) size[d] = dat->size[d] + dat->d_m[d] - dat->d_p[d];
}

void ops_dat_get_raw_metadata(ops_dat dat, int part, int *disp, int *size, int *stride, int *d_m, int *d_p) {
  (void)part;
  ops_dat_get_extents(dat, part, disp, size);
  if (stride)
    for (int d = 0; d < dat->block->dims; d++) stride[d] = dat->size[d];
  if (d_m)
    for (int d = 0; d < dat->block->dims; d++) d_m[d] = dat->d_m[d];
  if (d_p)
    for (int d = 0; d < dat->block->dims; d++) d_p[d] = dat->d_p[d];
}

int ops_dat_get_local_npartitions(ops_dat dat) {
  (void)dat;
  return 1;
}
int ops_dat_get_global_npartitions(ops_dat dat) {
  (void)dat;
  return 1;
}

// ops/c/src/core/ops_dummy_singlenode.cpp:662-735
char *ops_dat_get_raw_pointer(ops_dat dat, int part, ops_stencil stencil, ops_memspace *memspace) {
  OPSMI_API_BEGIN
  (void)part;
  (void)stencil;
  require_init();
  flush_all();
  if (!memspace) fail(OPS_INVALID_ARGUMENT, "ops_dat_get_raw_pointer: NULL memspace");
  if (*memspace == OPS_HOST) {
    ensure_host(dat);
  } else if (*memspace == OPS_DEVICE) {
    sync_to_device(dat);
  } else {
    if (dat->dirty_hd == 1) *memspace = OPS_HOST;
    else *memspace = OPS_DEVICE;
  }
  if (*memspace == OPS_DEVICE) OPSMI_HIP(hipStreamSynchronize(inst().stream));
  dat->locked_hd = *memspace;
  return (*memspace == OPS_HOST ? dat->data : dat->data_d) + dat->base_offset;
  OPSMI_API_END(nullptr)
}

void ops_dat_release_raw_data(ops_dat dat, int part, ops_access acc) {
  OPSMI_API_BEGIN
  (void)part;
  if (dat->locked_hd == 0)
    fail(OPS_RUNTIME_ERROR, cat("Error: ops_dat_release_raw_data() called, but with no matching "
                                "ops_dat_get_raw_pointer() beforehand: ", dat->name));
  if (acc != OPS_READ) {
    dat->dirty_hd = dat->locked_hd;
    note_external_write(dat);
  }
  dat->locked_hd = 0;
  OPSMI_API_END()
}

void ops_dat_release_raw_data_memspace(ops_dat dat, int part, ops_access acc, ops_memspace *memspace) {
  OPSMI_API_BEGIN
  (void)part;
  if (dat->locked_hd == 0)
    fail(OPS_RUNTIME_ERROR, cat("Error: ops_dat_release_raw_data_memspace() called, but with no matching "
                                "ops_dat_get_raw_pointer() beforehand: ", dat->name));
  if (acc != OPS_READ) {
    dat->dirty_hd = *memspace;
    note_external_write(dat);
  }
  dat->locked_hd = 0;
  OPSMI_API_END()
}

// copies a box [lo,hi) of points between the dat's device array and a dense
// buffer (x fastest), one 2D copy per z-plane
static void slab_copy(ops_dat dat, char *buf, const int *lo, const int *hi, bool to_dat) {
  Instance &I = inst();
  int nd = dat->block->dims;
  int l[3] = {0, 0, 0}, h[3] = {1, 1, 1};
  for (int d = 0; d < nd; d++) {
    l[d] = lo[d];
    h[d] = hi[d];
    int first = dat->d_m[d] + dat->base[d], last = dat->size[d] + dat->d_m[d] + dat->base[d];
    if (l[d] < first || h[d] > last || h[d] < l[d])
      fail(OPS_INVALID_ARGUMENT, cat("slab [", l[d], ",", h[d], ") outside dat ", dat->name, " in dimension ", d));
  }
  size_t es = dat->elem_size;
  size_t w = (size_t)(h[0] - l[0]) * es, rows = (size_t)(h[1] - l[1]);
  if (w == 0 || rows == 0 || h[2] <= l[2]) return;
  DatView v = view_of(dat);
  size_t pitch = (size_t)v.sy * dat->type_size;
  if (impl(dat)->soa) {
    // the buffer holds the points' components together, as the reference's
    // slab copy presents a SoA dat (ops/c/src/externlib/ops_util.cpp:353-500):
    // each component's box goes through a component-major staging copy
    const size_t ts = dat->type_size, nc = dat->dim;
    const size_t wx = (size_t)(h[0] - l[0]), planes = (size_t)(h[2] - l[2]), np = wx * rows * planes;
    std::vector<char> tmp(np * es);
    if (to_dat)
      for (size_t p = 0; p < np; p++)
        for (size_t c = 0; c < nc; c++) memcpy(&tmp[(c * np + p) * ts], buf + (p * nc + c) * ts, ts);
    for (size_t c = 0; c < nc; c++)
      for (int z = l[2]; z < h[2]; z++) {
        char *dp = v.base + ((long)c * v.cs + (long)l[0] + (long)l[1] * v.sy + (long)z * v.sz) * (long)ts;
        char *bp = &tmp[(c * np + (size_t)(z - l[2]) * wx * rows) * ts];
        if (to_dat)
          OPSMI_HIP(hipMemcpy2DAsync(dp, pitch, bp, wx * ts, wx * ts, rows, hipMemcpyDefault, I.stream));
        else
          OPSMI_HIP(hipMemcpy2DAsync(bp, wx * ts, dp, pitch, wx * ts, rows, hipMemcpyDefault, I.stream));
      }
    OPSMI_HIP(hipStreamSynchronize(I.stream));
    if (!to_dat)
      for (size_t p = 0; p < np; p++)
        for (size_t c = 0; c < nc; c++) memcpy(buf + (p * nc + c) * ts, &tmp[(c * np + p) * ts], ts);
    return;
  }
  for (int z = l[2]; z < h[2]; z++) {
    char *dp = v.base + ((long)l[0] * dat->dim + (long)l[1] * v.sy + (long)z * v.sz) * dat->type_size;
    char *bp = buf + (size_t)(z - l[2]) * w * rows;
    if (to_dat)
      OPSMI_HIP(hipMemcpy2DAsync(dp, pitch, bp, w, w, rows, hipMemcpyDefault, I.stream));
    else
      OPSMI_HIP(hipMemcpy2DAsync(bp, w, dp, pitch, w, rows, hipMemcpyDefault, I.stream));
  }
  OPSMI_HIP(hipStreamSynchronize(I.stream));
}

static void interior_box(ops_dat dat, int *lo, int *hi) {
  int sz[OPS_MAX_DIM];
  ops_dat_get_extents(dat, 0, nullptr, sz);
  for (int d = 0; d < dat->block->dims; d++) {
    lo[d] = dat->base[d];
    hi[d] = lo[d] + sz[d];
  }
}

void ops_dat_fetch_data_memspace(ops_dat dat, int part, char *data, ops_memspace memspace) {
  OPSMI_API_BEGIN
  (void)part;
  (void)memspace;
  require_init();
  flush_all();
  sync_to_device(dat);
  int lo[3], hi[3];
  interior_box(dat, lo, hi);
  slab_copy(dat, data, lo, hi, false);
  OPSMI_API_END()
}

void ops_dat_fetch_data(ops_dat dat, int part, char *data) { ops_dat_fetch_data_memspace(dat, part, data, OPS_HOST); }

void ops_dat_fetch_data_slab_memspace(ops_dat dat, int part, char *data, int *range, ops_memspace memspace) {
  OPSMI_API_BEGIN
  (void)part;
  (void)memspace;
  require_init();
  flush_all();
  sync_to_device(dat);
  int lo[3], hi[3];
  for (int d = 0; d < dat->block->dims; d++) {
    lo[d] = range[2 * d];
    hi[d] = range[2 * d + 1];
  }
  slab_copy(dat, data, lo, hi, false);
  OPSMI_API_END()
}

void ops_dat_set_data_memspace(ops_dat dat, int part, char *data, ops_memspace memspace) {
  OPSMI_API_BEGIN
  (void)part;
  (void)memspace;
  require_init();
  flush_all();
  sync_to_device(dat);
  int lo[3], hi[3];
  interior_box(dat, lo, hi);
  slab_copy(dat, data, lo, hi, true);
  mark_device_dirty(dat);
  note_external_write(dat);
  OPSMI_API_END()
}

void ops_dat_set_data(ops_dat dat, int part, char *data) { ops_dat_set_data_memspace(dat, part, data, OPS_HOST); }

void ops_dat_set_data_slab_memspace(ops_dat dat, int part, char *data, int *range, ops_memspace memspace) {
  OPSMI_API_BEGIN
  (void)part;
  (void)memspace;
  require_init();
  flush_all();
  sync_to_device(dat);
  int lo[3], hi[3];
  for (int d = 0; d < dat->block->dims; d++) {
    lo[d] = range[2 * d];
    hi[d] = range[2 * d + 1];
  }
  slab_copy(dat, data, lo, hi, true);
  mark_device_dirty(dat);
  note_external_write(dat);
  OPSMI_API_END()
}

// ------------------------------------------------------------------ utilities
void ops_timers_core(double *cpu, double *et) {
  *cpu = (double)clock() / CLOCKS_PER_SEC;
  *et = wall_now();
}

void ops_timers(double *cpu, double *et) {
  OPSMI_API_BEGIN
  if (inst().initialised) {
    flush_all();
    OPSMI_HIP(hipStreamSynchronize(inst().stream));
  }
  ops_timers_core(cpu, et);
  OPSMI_API_END()
}

void ops_printf(const char *format, ...) {
  va_list ap;
  va_start(ap, format);
  vprintf(format, ap);
  va_end(ap);
}

void ops_fprintf(FILE *stream, const char *format, ...) {
  va_list ap;
  va_start(ap, format);
  vfprintf(stream, format, ap);
  va_end(ap);
}

int ops_is_root(void) { return 1; }

// HDF5 output (ops/c/include/ops_hdf5.h:130-166): this backend has no HDF5
// library; the calls are recognised and refused
static void no_hdf5(const char *what, const char *file) {
  fail(OPS_NOT_IMPLEMENTED, cat(what, "(\"", file ? file : "", "\"): HDF5 output is not available in this backend"));
}
void ops_fetch_block_hdf5_file(ops_block block, const char *file_name) {
  OPSMI_API_BEGIN
  (void)block;
  no_hdf5("ops_fetch_block_hdf5_file", file_name);
  OPSMI_API_END()
}
void ops_fetch_dat_hdf5_file(ops_dat dat, const char *file_name) {
  OPSMI_API_BEGIN
  (void)dat;
  no_hdf5("ops_fetch_dat_hdf5_file", file_name);
  OPSMI_API_END()
}
void ops_fetch_stencil_hdf5_file(ops_stencil stencil, const char *file_name) {
  OPSMI_API_BEGIN
  (void)stencil;
  no_hdf5("ops_fetch_stencil_hdf5_file", file_name);
  OPSMI_API_END()
}
void ops_fetch_halo_hdf5_file(ops_halo halo, const char *file_name) {
  OPSMI_API_BEGIN
  (void)halo;
  no_hdf5("ops_fetch_halo_hdf5_file", file_name);
  OPSMI_API_END()
}

// ops/c/src/core/ops_dummy_singlenode.cpp:871-878, ops/c/src/core/
// ops_lib_core.cpp:1845-1960: scans the whole dat (halos included) on the
// host; the first NaN is reported and the process exits with status 2
// allocation helpers applications use (ops/c/include/ops_util.h:41-45)
void *xmalloc(size_t size) {
  void *p = malloc(size);
  if (!p && size) {
    OPSMI_API_BEGIN
    fail(OPS_RUNTIME_ERROR, cat("xmalloc: out of host memory (", size, " bytes)"));
    OPSMI_API_END(nullptr)
  }
  return p;
}
void *xrealloc(void *ptr, size_t size) {
  void *p = realloc(ptr, size);
  if (!p && size) {
    OPSMI_API_BEGIN
    fail(OPS_RUNTIME_ERROR, cat("xrealloc: out of host memory (", size, " bytes)"));
    OPSMI_API_END(nullptr)
  }
  return p;
}
void *xcalloc(size_t number, size_t size) {
  void *p = calloc(number, size);
  if (!p && number && size) {
    OPSMI_API_BEGIN
    fail(OPS_RUNTIME_ERROR, cat("xcalloc: out of host memory (", number, " x ", size, " bytes)"));
    OPSMI_API_END(nullptr)
  }
  return p;
}
int file_exist(const char *filename) {
  FILE *f = filename ? fopen(filename, "r") : nullptr;
  if (!f) return 0;
  fclose(f);
  return 1;
}

// checkpointing as the reference built without it
// (ops/c/src/externlib/ops_checkpointing.cpp:1734-1779)
bool ops_checkpointing_init(const char *filename, double interval, int options) {
  (void)filename;
  (void)interval;
  (void)options;
  return false;
}
void ops_checkpointing_initphase_done(void) {}
void ops_checkpointing_manual_datlist(int ndats, ops_dat *datlist) {
  (void)ndats;
  (void)datlist;
}
bool ops_checkpointing_fastfw(int nbytes, char *payload) {
  (void)nbytes;
  (void)payload;
  return false;
}
bool ops_checkpointing_manual_datlist_fastfw(int ndats, ops_dat *datlist, int nbytes, char *payload) {
  (void)ndats;
  (void)datlist;
  (void)nbytes;
  (void)payload;
  return false;
}
bool ops_checkpointing_manual_datlist_fastfw_trigger(int ndats, ops_dat *datlist, int nbytes, char *payload) {
  return ops_checkpointing_manual_datlist_fastfw(ndats, datlist, nbytes, payload);
}

void ops_NaNcheck(ops_dat dat) {
  OPSMI_API_BEGIN
  require_init();
  if (!dat) fail(OPS_INVALID_ARGUMENT, "ops_NaNcheck: NULL dat");
  flush_all();
  ensure_host(dat);
  printf("ops_NaNcheck_core called for %s \n", dat->name);
  const DType dt = impl(dat)->dt;
  const int nd = dat->block->dims;
  long n[3] = {1, 1, 1};
  for (int q = 0; q < nd && q < 3; q++) n[q] = dat->size[q];
  for (long k = 0; k < n[2]; k++)
    for (long j = 0; j < n[1]; j++)
      for (long i = 0; i < n[0]; i++)
        for (int d = 0; d < dat->dim; d++) {
          const size_t off = (size_t)((k * n[1] + j) * n[0] + i) * dat->dim + d;
          bool nan = false;
          if (dt == DT_F64) nan = std::isnan(((const double *)dat->data)[off]);
          else if (dt == DT_F32) nan = std::isnan(((const float *)dat->data)[off]);
          else if (dt == DT_F16) {
            const uint16_t h = ((const uint16_t *)dat->data)[off];
            nan = (h & 0x7c00u) == 0x7c00u && (h & 0x03ffu) != 0;
          }
          if (nan) {
            const long idx[3] = {i + dat->d_m[0], j + (nd > 1 ? dat->d_m[1] : 0), k + (nd > 2 ? dat->d_m[2] : 0)};
            printf("Error: NaN detected at element dim:%d,index:(%ld", d, idx[0]);
            for (int q = 1; q < nd && q < 3; q++) printf(",%ld", idx[q]);
            printf(")\n");
            exit(2);
          }
        }
  OPSMI_API_END()
}

// ops/c/src/core/ops_lib_core.cpp:1277-1462 (ops_print_dat_to_txtfile_core):
// appends the dat's whole array (halos and padding included) to the file:
// a header, then for each row every point's components, rows ending in a
// newline, planes in a blank line
void ops_print_dat_to_txtfile(ops_dat dat, const char *file_name) {
  OPSMI_API_BEGIN
  require_init();
  if (!dat || !file_name) fail(OPS_INVALID_ARGUMENT, "ops_print_dat_to_txtfile: NULL argument");
  const DType dt = impl(dat)->dt;
  const bool cplx = strcmp(dat->type, "complexd") == 0;
  if (dt != DT_F64 && dt != DT_F32 && dt != DT_I32 && !cplx)
    fail(OPS_NOT_IMPLEMENTED, cat("Error: Unknown type ", dat->type, " cannot be written to file ", file_name));
  flush_all();
  ensure_host(dat);
  FILE *fp = fopen(file_name, "a");
  if (!fp) fail(OPS_RUNTIME_ERROR, cat("Error: can't open file ", file_name));
  const int nd = dat->block->dims;
  long n[3] = {1, 1, 1};
  for (int q = 0; q < nd && q < 3; q++) n[q] = dat->size[q];
  fprintf(fp, "ops_dat:  %s \n", dat->name);
  fprintf(fp, "ops_dat dim:  %d \n", dat->dim);
  fprintf(fp, "block Dims : %d ", nd);
  for (int q = 0; q < nd; q++) fprintf(fp, "[%d]", dat->size[q]);
  fprintf(fp, "\n");
  fprintf(fp, "elem size %d \n", dat->elem_size);
  const size_t plane = (size_t)(n[0] * n[1] * n[2]);
  const bool soa = impl(dat)->soa;
  for (long k = 0; k < n[2]; k++) {
    for (long j = 0; j < n[1]; j++) {
      for (long i = 0; i < n[0]; i++)
        for (int d = 0; d < dat->dim; d++) {
          const size_t pt = (size_t)((k * n[1] + j) * n[0] + i);
          const size_t off = soa ? pt + d * plane : pt * dat->dim + d;
          if (cplx) {
            const double *c = (const double *)dat->data;
            fprintf(fp, "%3.10lf+%3.10lfi ", c[2 * off], c[2 * off + 1]);
          } else if (dt == DT_F64) {
            fprintf(fp, " %3.10lf", ((const double *)dat->data)[off]);
          } else if (dt == DT_F32) {
            fprintf(fp, "%e ", ((const float *)dat->data)[off]);
          } else {
            fprintf(fp, "%d ", ((const int *)dat->data)[off]);
          }
        }
      if (n[0] > 1) fprintf(fp, "\n");
    }
    if (n[1] > 1) fprintf(fp, "\n");
  }
  if (n[2] > 1) fprintf(fp, "\n");
  const bool bad = ferror(fp) != 0;
  fclose(fp);
  if (bad) fail(OPS_RUNTIME_ERROR, cat("Error: error writing to file ", file_name));
  OPSMI_API_END()
}

void ops_diagnostic_output(void) {
  Instance &I = inst();
  printf("\n OPS diagnostic output (MI355X backend, device %d, %d CUs)\n", I.device, I.num_cus);
  printf(" blocks %zu, dats %zu, stencils %zu, reductions %zu\n", I.blocks.size(), I.dats.size(),
         I.stencils.size(), I.reductions.size());
  printf(" fusion level %d, tiling max depth %d, OPS_BLOCK_SIZE %d x %d x %d (generated loops: %d rows 2D, "
         "%d planes 3D per workgroup)\n", I.fusion, I.tiling_maxdepth, I.block_x, I.block_y, I.block_z,
         I.gen_rows_2d(), I.gen_planes_3d());
  for (ops_dat d : I.dats)
    if (d)
      printf("   %-16s block %-10s dim %d type %-6s size [%d %d %d] mem %.3f MB\n", d->name, d->block->name, d->dim,
             d->type, d->size[0], d->size[1], d->size[2], d->mem / 1e6);
}

void ops_timing_output_stdout(void) {
  OPSMI_API_BEGIN
  flush_all();
  drain_events();
  // ops_timing_output layout (ops/c/src/core/ops_lib_core.cpp:1468-1624),
  // bandwidth in decimal GB/s of kernel time
  printf("\n%-40s %8s %12s %14s\n", "Loop", "count", "time(ms)", "GB/s");
  for (auto &kv : inst().stats) {
    const LoopStat &s = kv.second;
    double gbs = s.ms > 0 ? s.bytes / (s.ms * 1e-3) / 1e9 : 0.0;
    printf("%-40s %8ld %12.4f %14.2f\n", kv.first.c_str(), s.count, s.ms, gbs);
  }
  OPSMI_API_END()
}

// ------------------------------------------------------------------ extensions
void ops_mi_set_error_mode(int mode) { inst().error_mode = mode; }

// ops/c/src/core/ops_lib_core.cpp:372-374
void ops_set_soa(const int soa_val) { inst().soa = soa_val ? 1 : 0; }

int ops_mi_get_diags(void) { return inst().diags; }

void ops_mi_set_host_fold(int on) { inst().host_fold = on != 0; }

void ops_mi_set_diags(int level) {
  OPSMI_API_BEGIN
  flush_all();
  drain_events();
  inst().diags = level;
  OPSMI_API_END()
}

int ops_mi_last_error(char *msg, int msg_len) {
  Instance &I = inst();
  int c = I.last_error;
  if (msg && msg_len > 0) {
    strncpy(msg, I.last_msg.c_str(), msg_len - 1);
    msg[msg_len - 1] = 0;
  }
  I.last_error = 0;
  I.last_msg.clear();
  return c;
}

void *ops_mi_stream(void) { return (void *)inst().stream; }

void ops_mi_sync(void) {
  OPSMI_API_BEGIN
  require_init();
  flush_all();
  OPSMI_HIP(hipStreamSynchronize(inst().stream));
  OPSMI_API_END()
}

int ops_mi_loop_stats(const char *name, long *count, double *kernel_ms, double *bytes) {
  OPSMI_API_BEGIN
  flush_all();
  drain_events();
  auto it = inst().stats.find(name);
  if (it == inst().stats.end()) return 0;
  if (count) *count = it->second.launches;
  if (kernel_ms) *kernel_ms = it->second.ms;
  if (bytes) *bytes = it->second.bytes;
  return 1;
  OPSMI_API_END(0)
}

void ops_mi_reset_stats(void) {
  OPSMI_API_BEGIN
  flush_all();
  drain_events();
  inst().stats.clear();
  OPSMI_API_END()
}

void ops_mi_set_fusion(int level) { inst().fusion = level < 0 ? 0 : level; }

}  // extern "C"
