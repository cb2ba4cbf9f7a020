// Execution of translator-generated loops (include/ops_mi355x_gen.h).
//
// A generated stub hands over its arguments, the element types its kernel
// declares, a launch function and its by-value data.  The loop goes into the
// same lazy queue as the hand-written ones (dependencies from its dat and
// reduction arguments); when it runs, the launch function gets the dats'
// current device views, the tiling the kernel maps its points with, and one
// partial buffer per reduction argument, and the library then folds each
// argument's partials into its handle in a fixed order (the same two-stage
// fold as the hand-written loops).
#include <stdlib.h>
#include <string.h>

#include "engine.h"

namespace opsmi {
namespace {

template <class T, int OP>
void finish_dim(const RedCtl &c, int dim, unsigned parts, hipStream_t s) {
  switch (dim) {
    case 1: launch_red_finish<T, OP, 1>(c, parts, s); break;
    case 2: launch_red_finish<T, OP, 2>(c, parts, s); break;
    case 3: launch_red_finish<T, OP, 3>(c, parts, s); break;
    default: launch_red_finish<T, OP, 4>(c, parts, s); break;
  }
}
template <class T>
void finish_op(const RedCtl &c, int op, int dim, unsigned parts, hipStream_t s) {
  if (op == OPS_INC) finish_dim<T, OPS_INC>(c, dim, parts, s);
  else if (op == OPS_MIN) finish_dim<T, OPS_MIN>(c, dim, parts, s);
  else finish_dim<T, OPS_MAX>(c, dim, parts, s);
}
void finish(DType dt, const RedCtl &c, int op, int dim, unsigned parts, hipStream_t s) {
  if (dt == DT_F64) finish_op<double>(c, op, dim, parts, s);
  else if (dt == DT_F32) finish_op<float>(c, op, dim, parts, s);
  else finish_op<int>(c, op, dim, parts, s);
}

// z planes per workgroup of a generated 3D loop (OPSMI_GEN_ZCHUNK)
int gen_zchunk() {
  static int v = [] {
    const char *e = getenv("OPSMI_GEN_ZCHUNK");
    const int n = e ? atoi(e) : 1;
    return n >= 1 && n <= 256 ? n : 1;
  }();
  return v;
}

struct GenRed {
  int arg;
  DType dt;
  int op, dim;
  size_t offset;  // bytes into the partial buffer
  RedCtl ctl;
};

}  // namespace
}  // namespace opsmi

using namespace opsmi;

extern "C" {

void ops_mi_gen_par_loop(const char *name, ops_block block, int dim, const int *range, const ops_arg *args, int nargs,
                         const int *types, const int *npts, const int *pts, opsmi_gen_launch_fn fn, const void *blob,
                         size_t blob_bytes) {
  OPSMI_API_BEGIN
  if (!fn) fail(OPS_INVALID_ARGUMENT, cat("loop ", name ? name : "?", ": no launch function"));
  if (blob_bytes > 0 && !blob) fail(OPS_INVALID_ARGUMENT, cat("loop ", name ? name : "?", ": NULL data block"));
  const Range3 r = check_gen_loop(name ? name : "?", block, dim, range, args, nargs, types, npts, pts);
  LoopRec rec;
  rec.name = name ? name : "?";
  rec.range = r;
  rec.dims = dim;
  rec.bytes = transfer_bytes(args, nargs, r);
  deps_from_args(rec, args, nargs);
  if (r.npts() <= 0) {
    rec.bytes = 0.0;
    rec.launch = [](hipStream_t) {};
  } else {
    // the hand-written engine's tiling with rows of 64 * OPSMI_GEN_XP
    // points: 3D one plane of 4 rows per workgroup, 2D four groups of four
    // rows
    Tiling t = make_tiling(r, r.z1 - r.z0 > 1 ? gen_zchunk() : 4);
    t.grid.x = (unsigned)(((long)r.x1 - r.x0 + kWave * OPSMI_GEN_XP - 1) / (kWave * OPSMI_GEN_XP));
    const unsigned parts = t.blocks();
    std::vector<GenRed> reds;
    size_t off = 0;
    for (int i = 0; i < nargs; i++) {
      const ops_arg &a = args[i];
      if (a.argtype != OPS_ARG_GBL || a.acc == OPS_READ) continue;
      GenRed g;
      g.arg = i;
      g.dt = (DType)types[i];
      g.op = a.acc;
      g.dim = a.dim;
      g.offset = off;
      // handles are armed in argument order, when the loop is called
      g.ctl = red_prepare(a, a.dim, g.dt, 0);
      off += ((size_t)parts * a.dim * 8 + 255) / 256 * 256;
      reds.push_back(g);
    }
    ensure_partials(off);
    if (!reds.empty()) rec.ctl = reds[0].ctl;
    std::vector<char> data((const char *)blob, (const char *)blob + blob_bytes);
    std::vector<ops_dat> dats((size_t)nargs, nullptr);
    for (int i = 0; i < nargs; i++)
      if (args[i].argtype == OPS_ARG_DAT) dats[i] = args[i].dat;
    int alias = 0;
    for (int i = 0; i < nargs; i++)
      for (int j = i + 1; j < nargs; j++)
        if (dats[i] && dats[i] == dats[j]) alias = 1;
    rec.launch = [=](hipStream_t s) {
      opsmi_gen_launch L;
      memset(&L, 0, sizeof(L));
      const int rr[6] = {r.x0, r.x1, r.y0, r.y1, r.z0, r.z1};
      memcpy(L.range, rr, sizeof(rr));
      L.chunk = t.chunk;
      L.three_d = t.three_d ? 1 : 0;
      L.grid[0] = t.grid.x;
      L.grid[1] = t.grid.y;
      L.grid[2] = t.grid.z;
      L.alias = alias;
      for (size_t i = 0; i < dats.size(); i++)
        if (dats[i]) {
          const DatView v = view_of(dats[i]);
          L.dat[i].base = v.base;
          L.dat[i].md = dats[i]->dim;
          L.dat[i].sy = v.sy;
          L.dat[i].sz = v.sz;
        }
      char *pbase = (char *)inst().d_partials;
      for (const GenRed &g : reds) L.partials[g.arg] = pbase + g.offset;
      fn(&L, data.empty() ? nullptr : data.data(), (void *)s);
      OPSMI_HIP(hipGetLastError());
      for (const GenRed &g : reds) {
        RedCtl c = with_scratch(g.ctl);
        c.partials = pbase + g.offset;
        finish(g.dt, c, g.op, g.dim, parts, s);
      }
    };
  }
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_mi_gen_const(const char *name, const char *type, int dim, void *dst) {
  OPSMI_API_BEGIN
  require_init();
  if (!name || !type || !dst) fail(OPS_INVALID_ARGUMENT, "ops_mi_gen_const: NULL argument");
  const Const &c = get_const(name);
  const size_t es = !strcmp(type, "double") ? 8 : ((!strcmp(type, "float") || !strcmp(type, "int")) ? 4 : 0);
  if (es == 0) fail(OPS_INVALID_ARGUMENT, cat("constant ", name, ": unsupported type ", type));
  if (c.type != type || c.bytes.size() != es * (size_t)dim)
    fail(OPS_INVALID_ARGUMENT, cat("constant ", name, " was declared as ", c.bytes.size(), " bytes of ", c.type,
                                   ", the kernel uses ", dim, " ", type));
  memcpy(dst, c.bytes.data(), c.bytes.size());
  OPSMI_API_END()
}

}  // extern "C"
