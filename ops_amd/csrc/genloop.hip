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
#include <float.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>

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

// z planes per workgroup of a generated 3D loop without OPS_BLOCK_SIZE_Z
// (OPSMI_GEN_ZCHUNK, else 1: more planes per workgroup measured slower)
int gen_zchunk_env() {
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
  size_t offset;  // bytes into the loop's share of the partial buffer
  RedCtl ctl;
};

}  // namespace

int Instance::gen_rows_2d() const {
  return block_y > 0 ? (block_y + kWavesPerBlock - 1) / kWavesPerBlock * kWavesPerBlock : 4 * kWavesPerBlock;
}
int Instance::gen_planes_3d() const { return block_z > 0 ? block_z : gen_zchunk_env(); }

// A generated loop as queued: what its launch function is given when it
// runs, alone or as the second half of a fused pair.
struct GenLoop {
  opsmi_gen_launch_fn fn = nullptr;
  ops_block block = nullptr;
  Range3 r{};
  int dims = 0;
  Tiling t{};
  int alias = 0;
  std::vector<char> data;
  std::vector<ops_dat> dats;  // by argument (nullptr for other kinds)
  std::vector<int> acc;       // access mode by argument
  std::vector<char> at_point; // dat argument touched only at offset 0 (listed points)
  std::vector<unsigned char> adv;  // dat argument: bit q set when its stencil advances in dimension q
  bool strided = false;       // some stencil does not advance in a dimension (stride 0)
  // edge dats reset to the identity of their accumulation before the kernel
  std::vector<std::pair<ops_dat, int>> resets;
  std::vector<GenRed> reds;
  size_t pbytes = 0;          // partial buffer bytes the reductions take

  unsigned parts() const { return t.blocks(); }
  // the launch descriptor with reductions at pbase
  void fill(opsmi_gen_launch &L, char *pbase) const {
    memset(&L, 0, sizeof(L));
    const int rr[6] = {r.x0, r.x1, r.y0, r.y1, r.z0, r.z1};
    memcpy(L.range, rr, sizeof(rr));
    L.chunk = t.chunk;
    L.three_d = t.three_d ? 1 : 0;
    L.grid[0] = t.grid.x;
    L.grid[1] = t.grid.y;
    L.grid[2] = t.grid.z;
    L.alias = alias || strided;
    L.spread = inst().gen_spread ? 1 : 0;
    for (size_t i = 0; i < dats.size(); i++)
      if (dats[i]) {
        const DatView v = view_of(dats[i]);
        L.dat[i].base = v.base;
        L.dat[i].md = (adv[i] & 1) ? v.sx : 0;
        L.dat[i].sy = (adv[i] & 2) ? v.sy : 0;
        L.dat[i].sz = (adv[i] & 4) ? v.sz : 0;
        L.dat[i].cs = v.cs;
      }
    for (const GenRed &g : reds) L.partials[g.arg] = pbase + g.offset;
    // a loop with one reduction whose result is read right after this
    // flush: the host folds its partials (with_host_fold in engine.h)
    if (reds.size() == 1 && host_fold_claim(reds[0].ctl, parts(), reds[0].dim, reds[0].dt))
      L.partials[reds[0].arg] = (char *)inst().d_hfold;
  }
  const void *blob() const { return data.empty() ? nullptr : data.data(); }
  // the identity fills of edge dats this loop accumulates into afresh
  void run_resets(hipStream_t s) const;
  // folds each reduction argument's partials into its handle
  void finish_reds(char *pbase, hipStream_t s) const {
    for (const GenRed &g : reds) {
      if (reds.size() == 1 && inst().hf_used && g.ctl.slot == inst().hf_slot) continue;  // host fold
      RedCtl c = with_scratch(g.ctl);
      c.partials = pbase + g.offset;
      finish(g.dt, c, g.op, g.dim, parts(), s);
    }
  }
  bool writes(ops_dat d) const {
    for (size_t i = 0; i < dats.size(); i++)
      if (dats[i] == d && acc[i] != OPS_READ) return true;
    return false;
  }
  // every argument on d touches it only at the point itself
  bool only_at_point(ops_dat d) const {
    for (size_t i = 0; i < dats.size(); i++)
      if (dats[i] == d && !at_point[i]) return false;
    return true;
  }
};

namespace {

template <class T>
__global__ void k_fill_value(T *p, size_t n, T v) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <class T>
void fill_value(ops_dat d, T v, hipStream_t s) {
  const size_t n = d->mem / sizeof(T);
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  if (n) hipLaunchKernelGGL(k_fill_value<T>, dim3(blocks), dim3(256), 0, s, (T *)d->data_d, n, v);
}

}  // namespace

// identity of an accumulation into an edge dat: 0, the type's largest value
// (MIN) or lowest value (MAX), as ops/c/src/core/ops_dummy_singlenode.cpp:
// 252-378 fills them
void GenLoop::run_resets(hipStream_t s) const {
  for (const auto &rs : resets) {
    const ops_dat d = rs.first;
    const int op = rs.second;
    switch (((DatImpl *)d->impl)->dt) {
      case DT_F64:
        fill_value<double>(d, op == OPS_INC ? 0.0 : (op == OPS_MIN ? DBL_MAX : -DBL_MAX), s);
        break;
      case DT_F32:
        fill_value<float>(d, op == OPS_INC ? 0.0f : (op == OPS_MIN ? FLT_MAX : -FLT_MAX), s);
        break;
      case DT_F16:  // 0, +65504, -65504 as binary16 bits
        fill_value<unsigned short>(d, op == OPS_INC ? 0x0000 : (op == OPS_MIN ? 0x7bff : 0xfbff), s);
        break;
      default:
        fill_value<int>(d, op == OPS_INC ? 0 : (op == OPS_MIN ? INT_MAX : INT_MIN), s);
        break;
    }
    OPSMI_HIP(hipGetLastError());
  }
}

namespace {

struct NbFused {
  opsmi_gen_fused_fn fn = nullptr;
  std::vector<std::pair<int, int>> links;  // (argument of the second loop, argument of the first)
};
std::map<std::pair<opsmi_gen_launch_fn, opsmi_gen_launch_fn>, NbFused> &nb_table() {
  static std::map<std::pair<opsmi_gen_launch_fn, opsmi_gen_launch_fn>, NbFused> m;
  return m;
}

bool same_range(const Range3 &a, const Range3 &b) {
  return a.x0 == b.x0 && a.x1 == b.x1 && a.y0 == b.y0 && a.y1 == b.y1 && a.z0 == b.z0 && a.z1 == b.z1;
}

bool share_handle(const GenLoop &a, const GenLoop &b) {
  for (const GenRed &x : a.reds)
    for (const GenRed &y : b.reds)
      if (x.ctl.slot == y.ctl.slot) return true;
  return false;
}

struct ChainFused {
  opsmi_gen_chain_fn fn = nullptr;
  std::vector<int> groups;  // per argument of each loop in turn: pointer index, -1 for non-dats
};
std::map<std::vector<opsmi_gen_launch_fn>, ChainFused> &chain_table() {
  static std::map<std::vector<opsmi_gen_launch_fn>, ChainFused> m;
  return m;
}

// the run's dat arguments agree with the kernel's pointers: one dat per
// pointer, distinct dats for distinct pointers, none absent
bool groups_match(const GenLoop *const *G, int n, const std::vector<int> &groups) {
  std::vector<ops_dat> d;
  std::vector<int> g;
  size_t pos = 0;
  for (int k = 0; k < n; k++)
    for (size_t i = 0; i < G[k]->dats.size(); i++, pos++) {
      if (pos >= groups.size()) return false;
      const int j = groups[pos];
      if (j < 0) continue;
      if (!G[k]->dats[i]) return false;
      d.push_back(G[k]->dats[i]);
      g.push_back(j);
    }
  if (pos != groups.size()) return false;
  for (size_t a = 0; a < d.size(); a++)
    for (size_t b = 0; b < a; b++)
      if ((g[a] == g[b]) != (d[a] == d[b])) return false;
  return true;
}
int g_longest_chain = 0;

// n generated loops back to back as one pass (ops_mi_gen_fused_chain) when
// running them point by point equals running them in turn: every dat any of
// them writes is touched by all of them only at the point itself, so no
// point reads what another point's later loop writes, or reads before
// another point's earlier loop writes
int gen_chain_rule(std::vector<LoopRec> &run, size_t at, hipStream_t s) {
  (void)s;
  const int avail = (int)std::min<size_t>((size_t)g_longest_chain, run.size() - at);
  // pairs at level >= 1, longer runs at level 3
  for (int n = avail; n >= 2; n--) {
    if (n >= 3 && inst().fusion < 3) continue;
    std::vector<opsmi_gen_launch_fn> key;
    const GenLoop *G[OPSMI_GEN_MAX_CHAIN];
    bool ok = true;
    for (int k = 0; k < n && ok; k++) {
      ok = run[at + k].gen != nullptr;
      if (ok) {
        G[k] = run[at + k].gen.get();
        key.push_back(G[k]->fn);
      }
    }
    if (!ok) continue;
    const auto it = chain_table().find(key);
    if (it == chain_table().end()) continue;
    for (int k = 0; k < n && ok; k++) {
      const GenLoop &A = *G[k];
      ok = A.block == G[0]->block && A.dims == G[0]->dims && same_range(A.r, G[0]->r) && A.r.npts() > 0 &&
           !A.alias && !A.strided && A.resets.empty();
      for (int j = 0; j < k && ok; j++) ok = !share_handle(*G[j], A);
      for (size_t i = 0; i < A.dats.size() && ok; i++) {
        const ops_dat d = A.dats[i];
        if (d && A.acc[i] != OPS_READ)
          for (int j = 0; j < n && ok; j++) ok = G[j]->only_at_point(d);
      }
    }
    if (!ok || !groups_match(G, n, it->second.groups)) continue;
    const opsmi_gen_chain_fn fused = it->second.fn;
    size_t off[OPSMI_GEN_MAX_CHAIN], total = 0, bytes = 0;
    std::string name = "fused(";
    for (int k = 0; k < n; k++) {
      off[k] = total;
      total += G[k]->pbytes;
      bytes += run[at + k].bytes;
      name += (k ? "+" : "") + run[at + k].name;
    }
    name += ")";
    ensure_partials(total);
    run_loop(name.c_str(), bytes, [&](hipStream_t st) {
      char *pbase = (char *)inst().d_partials;
      opsmi_gen_launch L[OPSMI_GEN_MAX_CHAIN];
      const opsmi_gen_launch *Lp[OPSMI_GEN_MAX_CHAIN];
      const void *blobs[OPSMI_GEN_MAX_CHAIN];
      for (int k = 0; k < n; k++) {
        G[k]->fill(L[k], pbase + off[k]);
        Lp[k] = &L[k];
        blobs[k] = G[k]->blob();
      }
      fused(Lp, blobs, (void *)st);
      OPSMI_HIP(hipGetLastError());
      for (int k = 0; k < n; k++) G[k]->finish_reds(pbase + off[k], st);
    });
    for (int k = 0; k < n; k++) {
      note_fused(run[at + k].name.c_str(), run[at + k].bytes);
      for (ops_dat d : run[at + k].writes) note_external_write(d);
    }
    return n;
  }
  return 0;
}

bool reads(const GenLoop &G, ops_dat d) {
  for (size_t i = 0; i < G.dats.size(); i++)
    if (G.dats[i] == d && G.acc[i] != OPS_WRITE) return true;
  return false;
}

// The neighbour form (ops_mi_gen_fused_nb): the second loop reads, at
// stencil offsets, what the first writes.  Conditions in ops_mi355x_gen.h.
int gen_nb_rule(std::vector<LoopRec> &run, size_t at, hipStream_t s) {
  if (inst().fusion < 2) return 0;
  if (at + 1 >= run.size() || !run[at].gen || !run[at + 1].gen) return 0;
  const GenLoop &A = *run[at].gen, &B = *run[at + 1].gen;
  const auto it = nb_table().find({A.fn, B.fn});
  if (it == nb_table().end()) return 0;
  const NbFused &nb = it->second;
  if (A.block != B.block || A.dims != B.dims || !same_range(A.r, B.r) || A.r.npts() <= 0) return 0;
  if (share_handle(A, B)) return 0;
  if (A.alias || B.alias || A.strided || B.strided || !A.resets.empty() || !B.resets.empty()) return 0;
  std::vector<char> linked_b(B.dats.size(), 0);
  std::vector<ops_dat> linked;
  for (const auto &l : nb.links) {
    const int i = l.first, j = l.second;
    if (i < 0 || j < 0 || (size_t)i >= B.dats.size() || (size_t)j >= A.dats.size()) return 0;
    if (!B.dats[i] || B.dats[i] != A.dats[j] || A.acc[j] != OPS_WRITE || B.acc[i] != OPS_READ) return 0;
    linked_b[i] = 1;
    linked.push_back(B.dats[i]);
  }
  // the first loop reads no dat it writes (its recomputation at a neighbour
  // must see the values from before the pair)
  for (size_t j = 0; j < A.dats.size(); j++)
    if (A.dats[j] && A.acc[j] != OPS_READ && reads(A, A.dats[j])) return 0;
  // the second loop: other dats the first writes only at the point; no
  // write to a linked dat
  for (size_t i = 0; i < B.dats.size(); i++) {
    const ops_dat d = B.dats[i];
    if (!d || linked_b[i]) continue;
    if (A.writes(d) && !B.at_point[i]) return 0;
    if (B.acc[i] != OPS_READ && std::find(linked.begin(), linked.end(), d) != linked.end()) return 0;
  }
  // dats the second loop writes and the first reads: second buffer
  std::vector<ops_dat> spare;
  for (size_t i = 0; i < B.dats.size(); i++) {
    const ops_dat d = B.dats[i];
    if (d && B.acc[i] != OPS_READ && reads(A, d) && std::find(spare.begin(), spare.end(), d) == spare.end())
      spare.push_back(d);
  }
  const size_t off_b = A.pbytes;
  ensure_partials(off_b + B.pbytes);
  std::vector<char *> spare_base_of(spare.size());
  for (size_t k = 0; k < spare.size(); k++) spare_base_of[k] = spare_base(spare[k], B.r);
  LoopRec &ra = run[at], &rb = run[at + 1];
  const std::string name = "fused(" + ra.name + "+" + rb.name + ")";
  run_loop(name.c_str(), ra.bytes + rb.bytes, [&](hipStream_t st) {
    char *pbase = (char *)inst().d_partials;
    opsmi_gen_launch La, Lb;
    A.fill(La, pbase);
    B.fill(Lb, pbase + off_b);
    for (size_t i = 0; i < B.dats.size(); i++)
      for (size_t k = 0; k < spare.size(); k++)
        if (B.dats[i] == spare[k]) Lb.dat[i].base = spare_base_of[k];
    nb.fn(&La, A.blob(), &Lb, B.blob(), (void *)st);
    OPSMI_HIP(hipGetLastError());
    A.finish_reds(pbase, st);
    B.finish_reds(pbase + off_b, st);
  });
  (void)s;
  for (ops_dat d : spare) swap_spare(d, B.r);
  note_fused(ra.name.c_str(), ra.bytes);
  note_fused(rb.name.c_str(), rb.bytes);
  for (const LoopRec *L : {&ra, &rb})
    for (ops_dat d : L->writes)
      if (std::find(spare.begin(), spare.end(), d) == spare.end()) note_external_write(d);
  return 2;
}

struct RegisterGen {
  RegisterGen() {
    register_fusion(gen_chain_rule);
    register_fusion(gen_nb_rule);
  }
} g_register_gen;

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
  // edge dats (ops/c/src/core/ops_dummy_singlenode.cpp:204-391): an
  // accumulation after a different access starts from the identity; the
  // access of every write is remembered, in program order
  std::vector<std::pair<ops_dat, int>> resets;
  for (int i = 0; i < nargs; i++) {
    const ops_arg &a = args[i];
    if (a.argtype != OPS_ARG_DAT || !a.dat || !a.dat->e_dat) continue;
    DatImpl *m = (DatImpl *)a.dat->impl;
    if ((a.acc == OPS_INC || a.acc == OPS_MIN || a.acc == OPS_MAX) && m->edat_prev_acc != a.acc) {
      m->edat_prev_acc = a.acc;
      resets.push_back({a.dat, a.acc});
    }
  }
  for (int i = 0; i < nargs; i++) {
    const ops_arg &a = args[i];
    if (a.argtype == OPS_ARG_DAT && a.dat && a.dat->e_dat && a.acc != OPS_READ)
      ((DatImpl *)a.dat->impl)->edat_prev_acc = a.acc;
  }
  if (r.npts() <= 0) {
    rec.bytes = 0.0;
    if (resets.empty()) {
      rec.launch = [](hipStream_t) {};
    } else {
      auto G = std::make_shared<GenLoop>();
      G->resets = resets;
      rec.launch = [G](hipStream_t s) { G->run_resets(s); };
    }
  } else {
    auto G = std::make_shared<GenLoop>();
    G->fn = fn;
    G->block = block;
    G->r = r;
    G->dims = dim;
    // the hand-written engine's tiling with rows of 64 * OPSMI_GEN_XP
    // points: 3D one plane of 4 rows per workgroup, 2D four groups of four
    // rows
    G->t = make_tiling(r, r.z1 - r.z0 > 1 ? inst().gen_planes_3d() : inst().gen_rows_2d() / kWavesPerBlock);
    G->t.grid.x = (unsigned)(((long)r.x1 - r.x0 + kWave * OPSMI_GEN_XP - 1) / (kWave * OPSMI_GEN_XP));
    const unsigned parts = G->parts();
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
      G->reds.push_back(g);
    }
    G->pbytes = off;
    ensure_partials(off);
    if (!G->reds.empty()) rec.ctl = G->reds[0].ctl;
    G->data.assign((const char *)blob, (const char *)blob + blob_bytes);
    G->dats.assign((size_t)nargs, nullptr);
    G->acc.assign((size_t)nargs, 0);
    G->at_point.assign((size_t)nargs, 0);
    G->adv.assign((size_t)nargs, 7);
    const int *pt = pts;
    for (int i = 0; i < nargs; i++) {
      const int np = npts[i];
      if (args[i].argtype == OPS_ARG_DAT) {
        G->dats[i] = args[i].dat;
        G->acc[i] = args[i].acc;
        // only listed points count: a kernel that reaches the dat another
        // way (np < 0) or not at all (np == 0) is not known to stay at the point
        bool zero = np > 0;
        for (int q = 0; q < np; q++)
          if (pt[4 * q + 1] != 0 || pt[4 * q + 2] != 0 || pt[4 * q + 3] != 0) zero = false;
        G->at_point[i] = zero;
        if (np > 0) pt += 4 * np;
        const ops_stencil st = args[i].stencil;
        if (args[i].dat && st && st->stride) {
          unsigned char m = 0;
          for (int q = 0; q < 3; q++)
            if (q >= st->dims || st->stride[q] != 0) m |= (unsigned char)(1u << q);
          G->adv[i] = m;
          if (m != 7) G->strided = true;
        }
      }
    }
    G->resets = resets;
    for (int i = 0; i < nargs; i++)
      for (int j = i + 1; j < nargs; j++)
        if (G->dats[i] && G->dats[i] == G->dats[j]) G->alias = 1;
    rec.gen = G;
    const GenLoop *g = G.get();
    rec.launch = [G, g](hipStream_t s) {
      char *pbase = (char *)inst().d_partials;
      g->run_resets(s);
      opsmi_gen_launch L;
      g->fill(L, pbase);
      g->fn(&L, g->blob(), (void *)s);
      OPSMI_HIP(hipGetLastError());
      g->finish_reds(pbase, s);
    };
  }
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_mi_gen_fused_chain(const opsmi_gen_launch_fn *loops, int n, const int *groups, int ngroups,
                            opsmi_gen_chain_fn fused) {
  OPSMI_API_BEGIN
  if (!loops || !fused || n < 2 || n > OPSMI_GEN_MAX_CHAIN || ngroups < 0 ||
      ngroups > n * OPSMI_GEN_MAX_ARGS || (ngroups > 0 && !groups))
    fail(OPS_INVALID_ARGUMENT, "ops_mi_gen_fused_chain: bad arguments");
  std::vector<opsmi_gen_launch_fn> key(loops, loops + n);
  for (opsmi_gen_launch_fn f : key)
    if (!f) fail(OPS_INVALID_ARGUMENT, "ops_mi_gen_fused_chain: NULL function");
  ChainFused c;
  c.fn = fused;
  c.groups.assign(groups, groups + ngroups);
  chain_table()[key] = c;
  g_longest_chain = std::max(g_longest_chain, n);
  OPSMI_API_END()
}

void ops_mi_gen_fused_nb(opsmi_gen_launch_fn first, opsmi_gen_launch_fn second, opsmi_gen_fused_fn fused,
                         const int *links, int nlinks) {
  OPSMI_API_BEGIN
  if (!first || !second || !fused || (nlinks > 0 && !links) || nlinks < 1 || nlinks > OPSMI_GEN_MAX_ARGS)
    fail(OPS_INVALID_ARGUMENT, "ops_mi_gen_fused_nb: bad arguments");
  NbFused nb;
  nb.fn = fused;
  for (int k = 0; k < nlinks; k++) nb.links.push_back({links[2 * k], links[2 * k + 1]});
  nb_table()[{first, second}] = nb;
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

void ops_mi_gen_const_sized(const char *name, const char *type, size_t elem_size, int max_dim, int exact,
                            void *dst) {
  OPSMI_API_BEGIN
  require_init();
  if (!name || !type || !dst || elem_size == 0 || max_dim < 1)
    fail(OPS_INVALID_ARGUMENT, "ops_mi_gen_const_sized: bad arguments");
  const Const &c = get_const(name);
  const size_t n = c.bytes.size() / elem_size;
  if (c.type != type || c.bytes.size() % elem_size != 0 || (exact ? n != (size_t)max_dim : n > (size_t)max_dim))
    fail(OPS_INVALID_ARGUMENT, cat("constant ", name, " was declared as ", c.bytes.size(), " bytes of ", c.type,
                                   "; the kernel holds ", exact ? "exactly " : "at most ", max_dim, " ", type,
                                   " of ", elem_size, " bytes"));
  memcpy(dst, c.bytes.data(), c.bytes.size());
  OPSMI_API_END()
}

}  // extern "C"
