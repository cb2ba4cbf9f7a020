// Run-time fused loop chains (the GPU form of OPS lazy execution and loop
// tiling, ops/c/src/core/ops_lazy.cpp:380-1187).  When the queue runs a
// sequence of loops, these rules look for chains they can execute as one
// kernel with the same results loop by loop.
//
// CG chain (TeaLeaf CG, apps/c/TeaLeaf/tea_leaf_cg.cpp:129-214):
//     [tea_leaf_axpy_kernel        u = u + alpha p          (deferred from the
//                                                            previous iteration)]
//      tea_leaf_axpby_kernel       p = a p + b r
//      tea_leaf_cg_calc_w_reduce   w = A p ; pw += w p      (3D, 7-point)
// The lazy queue defers the axpy past the r.r result (nothing it depends on
// changes) until the axpby that overwrites p needs it gone, so the three
// loops meet here.  One z-marching column kernel (heat3_chain_column_v)
// reads u, p_old, r, Kx, Ky, Kz and writes u, p_new, w: 72 bytes per point
// instead of 40 (the fused pair below) + 40 (fusion level 2, the default).
// p_new is needed at stencil neighbours before every column has written it,
// so each column forms it
// from p_old and r for its planes and rims and writes it into p's spare
// buffer, which then becomes p's buffer (double buffering).  Outside the
// loop range (the halo) p_new equals p_old, as the axpby leaves it.  Every
// value is computed with the same operations, in the same order, as the
// separate loops.
#include "column.h"
#include "engine.h"

using namespace opsmi;

namespace {

// Same tiles (128 x kHeatRows, heat_column_grid) as the calc_w column, so
// the per-workgroup sums of w p, and the result, are the same bits.
template <bool AXPY>
__global__ __launch_bounds__(kWave * kHeatRows) void k_cg_chain(ChainPtrs c, Range3 R, double alpha_u, double a,
                                                                double b, double rx, double ry, double rz, int chunk,
                                                                RedCtl ctl) {
  constexpr int V = 2;
  __shared__ HeatLdsV<V, kHeatRows> lds;
  const TileId t = tile_id();
  const int z0 = R.z0 + (int)t.z * chunk;
  double acc[1] = {heat3_chain_column_v<V, AXPY, kHeatRows>(lds, c, R, alpha_u, a, b, rx, ry, rz,
                                                            R.x0 + (int)t.x * V * kWave, R.y0 + (int)t.y * kHeatRows,
                                                            z0, min(z0 + chunk, R.z1))};
  reduce_block_store<double, OPS_INC, 1>(acc, ctl);
}

bool same_range(const Range3 &a, const Range3 &b) {
  return a.x0 == b.x0 && a.x1 == b.x1 && a.y0 == b.y0 && a.y1 == b.y1 && a.z0 == b.z0 && a.z1 == b.z1;
}

int cg_chain_rule(std::vector<LoopRec> &run, size_t at, hipStream_t s) {
  size_t i = at;
  const LoopRec *ax = nullptr;
  if (i < run.size() && run[i].kind == LK_AXPY) ax = &run[i++];
  if (i + 1 >= run.size()) return 0;
  const LoopRec &bp = run[i], &cw = run[i + 1];
  if (bp.kind != LK_AXPBY || cw.kind != LK_HEAT_W3 || cw.dims != 3 || bp.dims != 3) return 0;
  ops_dat P = bp.d[0], Rr = bp.d[1];
  ops_dat W = cw.d[0], Kx = cw.d[1], Ky = cw.d[2], Kz = cw.d[3];
  if (cw.d[4] != P || P == Rr || W == P || W == Rr) return 0;
  if (Kx == P || Ky == P || Kz == P || Kx == W || Ky == W || Kz == W) return 0;
  if (!same_range(bp.range, cw.range) || cw.range.npts() <= 0) return 0;
  ops_dat U = nullptr;
  if (ax) {
    U = ax->d[0];
    // the deferred axpy must read this p_old and touch nothing else in the chain
    if (ax->d[1] != P || !same_range(ax->range, cw.range) || U == P || U == Rr || U == W || U == Kx ||
        U == Ky || U == Kz || ax->dims != 3)
      ax = nullptr, U = nullptr;
  }
  const Range3 R = cw.range;
  // the column kernel takes 16-byte loads from every array
  for (ops_dat d : {P, Rr, W, Kx, Ky, Kz})
    if (!vec16_ok(d, R.x0)) return 0;
  if (U && !vec16_ok(U, R.x0)) ax = nullptr, U = nullptr;
  if (!use_vec() || cw.ctl.slot == nullptr) return 0;
  constexpr int V = 2;
  const ColGrid g = heat_column_grid(R, V * kWave, kHeatRows);
  ensure_partials((size_t)g.blocks() * 8);
  ChainPtrs c;
  DatView pv = view_of(P);
  c.pnew = Ptr<double>{(double *)spare_base(P, R), pv.sy, pv.sz};
  c.pold = ptr_of<const double>(P);
  c.r = ptr_of<const double>(Rr);
  c.w = ptr_of<double>(W);
  c.Kx = ptr_of<const double>(Kx);
  c.Ky = ptr_of<const double>(Ky);
  c.Kz = ptr_of<const double>(Kz);
  c.u = U ? ptr_of<double>(U) : Ptr<double>{nullptr, 0, 0};
  const double alpha_u = ax ? ax->g[0] : 0.0, a = bp.g[0], b = bp.g[1];
  const double rx = cw.g[0], ry = cw.g[1], rz = cw.g[2];
  const RedCtl ctl = with_host_fold(cw.ctl, g.blocks(), 1, DT_F64);
  const double bytes = (ax ? ax->bytes : 0.0) + bp.bytes + cw.bytes;
  run_loop(ax ? "cg_chain(axpy+axpby+calc_w)" : "cg_chain(axpby+calc_w)", bytes, [&](hipStream_t st) {
    hipLaunchKernelGGL(ax ? k_cg_chain<true> : k_cg_chain<false>, dim3(g.gx, g.gy, g.gz), dim3(kWave, kHeatRows), 0,
                       st, c, R, alpha_u, a, b, rx, ry, rz, g.chunk, ctl);
    launch_red_finish<double, OPS_INC, 1>(ctl, g.blocks(), st);
  });
  (void)s;
  swap_spare(P, R);
  if (U) note_external_write(U);
  note_external_write(W);
  if (ax) note_fused(ax->name.c_str(), ax->bytes);
  note_fused(bp.name.c_str(), bp.bytes);
  note_fused(cw.name.c_str(), cw.bytes);
  return (int)(i + 2 - at);
}

// Pointwise pair: the deferred axpy and the axpby that overwrites its p,
//     u = u + alpha p_old ;  p = a p_old + b r
// in one pass: reads u, p, r and writes u, p (40 bytes per point instead of
// 24 + 24).  Pointwise, so p is updated in place.  One point per lane: at
// 512^3 it measured 955 us against 983 us (16-byte loads) and 990 us
// (16-byte loads and non-temporal stores).
struct PairF {
  OPSMI_NO_RED
  Ptr<double> u, p;
  Ptr<const double> r;
  double alpha_u, a, b;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    const double po = p(x, y, z);
    u(x, y, z) = u(x, y, z) + alpha_u * po;
    p(x, y, z) = a * po + b * r(x, y, z);
  }
};

int axpy_axpby_rule(std::vector<LoopRec> &run, size_t at, hipStream_t s) {
  if (at + 1 >= run.size()) return 0;
  const LoopRec &ax = run[at], &bp = run[at + 1];
  if (ax.kind != LK_AXPY || bp.kind != LK_AXPBY) return 0;
  ops_dat U = ax.d[0], P = ax.d[1], Rr = bp.d[1];
  if (bp.d[0] != P || U == P || U == Rr || P == Rr) return 0;
  if (!same_range(ax.range, bp.range) || ax.range.npts() <= 0) return 0;
  const Range3 R = ax.range;
  const Tiling t = make_tiling(R, 1);
  PairF f{ptr_of<double>(U), ptr_of<double>(P), ptr_of<const double>(Rr), ax.g[0], bp.g[0], bp.g[1]};
  run_loop("cg_pair(axpy+axpby)", ax.bytes + bp.bytes, [&](hipStream_t st) {
    hipLaunchKernelGGL(k_loop<PairF>, t.grid, dim3(kWave, kWavesPerBlock), 0, st, R, f, RedCtl{}, t.chunk,
                       t.three_d ? 1 : 0);
  });
  (void)s;
  note_external_write(U);
  note_external_write(P);
  note_fused(ax.name.c_str(), ax.bytes);
  note_fused(bp.name.c_str(), bp.bytes);
  return 2;
}

// The three-loop chain moves 72 instead of 80 bytes per point: at 512^3 it
// measured 1853 us against 1025 + 934 us for calc_w and the pair (2093 us
// when it formed p_new as each load arrived, which made every load wait).
// It runs at fusion level 2 (the default); at level 1 the pair rule takes the
// axpy + axpby and calc_w runs alone.
int cg_chain_rule_level2(std::vector<LoopRec> &run, size_t at, hipStream_t s) {
  return inst().fusion >= 2 ? cg_chain_rule(run, at, s) : 0;
}

// Constant-coefficient Poisson chains (poisson3_chain_column_v), on the
// tiles of the Poisson calc_w column (k_poisson3_march) so that the sums of
// w p are the same bits:
//   MODE 0, FP64 CG:   [tea_leaf_axpy] + tea_leaf_axpby + poisson_calc_w
//                      (48 instead of 24 + 40 - ... = 56 bytes per point
//                      for the pair and the sweep)
//   MODE 1, FP32 inner iterations of the mixed solver:
//                      [mp_e_update] + mp_p + poisson_calc_w (float)
//                      (24 instead of 12 + 12 + 8 bytes per point)
template <class T, bool AXPY, int MODE>
__global__ __launch_bounds__(kWave * kHeatRows) void k_pois_chain(PoisChainPtrs<T> c, Range3 R, T alpha_u, T a, T b,
                                                                  int chunk, RedCtl ctl) {
  constexpr int V = 16 / (int)sizeof(T);
  __shared__ PoisLdsV<T, V, kHeatRows> lds;
  const TileId t = tile_id();
  const int z0 = R.z0 + (int)t.z * chunk;
  double acc[1] = {poisson3_chain_column_v<T, V, AXPY, MODE, kHeatRows>(
      lds, c, R, alpha_u, a, b, R.x0 + (int)t.x * V * kWave, R.y0 + (int)t.y * kHeatRows, z0, min(z0 + chunk, R.z1))};
  reduce_block_store<double, OPS_INC, 1>(acc, ctl);
}

template <class T, int MODE>
int pois_chain_rule_t(std::vector<LoopRec> &run, size_t at) {
  const LoopKind ka = MODE == 0 ? LK_AXPY : LK_MP_E, kb = MODE == 0 ? LK_AXPBY : LK_MP_P;
  size_t i = at;
  const LoopRec *ax = nullptr;
  if (i < run.size() && run[i].kind == ka) ax = &run[i++];
  if (i + 1 >= run.size()) return 0;
  const LoopRec &bp = run[i], &cw = run[i + 1];
  if (bp.kind != kb || cw.kind != LK_POISSON_W3 || cw.g[0] != (double)sizeof(T) || bp.dims != 3) return 0;
  ops_dat P = bp.d[0], Rr = bp.d[1], W = cw.d[0];
  if (cw.d[1] != P || P == Rr || W == P || W == Rr) return 0;
  if (!same_range(bp.range, cw.range) || cw.range.npts() <= 0) return 0;
  ops_dat U = nullptr;
  if (ax) {
    U = ax->d[0];
    if (ax->d[1] != P || !same_range(ax->range, cw.range) || U == P || U == Rr || U == W || ax->dims != 3)
      ax = nullptr, U = nullptr;
  }
  const Range3 R = cw.range;
  for (ops_dat d : {P, Rr, W})
    if (!vec16_ok(d, R.x0)) return 0;
  if (U && !vec16_ok(U, R.x0)) ax = nullptr, U = nullptr;
  if (!use_vec() || cw.ctl.slot == nullptr) return 0;
  constexpr int V = 16 / (int)sizeof(T);
  const ColGrid g = heat_column_grid(R, V * kWave, kHeatRows);
  ensure_partials((size_t)g.blocks() * 8);
  PoisChainPtrs<T> c;
  DatView pv = view_of(P);
  c.pnew = Ptr<T>{(T *)spare_base(P, R), pv.sy, pv.sz};
  c.pold = ptr_of<const T>(P);
  c.r = ptr_of<const T>(Rr);
  c.w = ptr_of<T>(W);
  c.u = U ? ptr_of<T>(U) : Ptr<T>{nullptr, 0, 0};
  const T alpha_u = ax ? (T)ax->g[0] : (T)0, a = (T)bp.g[0], b = (T)bp.g[1];
  const RedCtl ctl = with_host_fold(cw.ctl, g.blocks(), 1, DT_F64);
  const double bytes = (ax ? ax->bytes : 0.0) + bp.bytes + cw.bytes;
  const char *nm = MODE == 0 ? (ax ? "poisson_chain(axpy+axpby+calc_w)" : "poisson_chain(axpby+calc_w)")
                             : (ax ? "mp_chain(e+p+calc_w)" : "mp_chain(p+calc_w)");
  run_loop(nm, bytes, [&](hipStream_t st) {
    auto kern = ax ? k_pois_chain<T, true, MODE> : k_pois_chain<T, false, MODE>;
    hipLaunchKernelGGL(kern, dim3(g.gx, g.gy, g.gz), dim3(kWave, kHeatRows), 0, st, c, R, alpha_u, a, b, g.chunk,
                       ctl);
    launch_red_finish<double, OPS_INC, 1>(ctl, g.blocks(), st);
  });
  swap_spare(P, R);
  if (U) note_external_write(U);
  note_external_write(W);
  if (ax) note_fused(ax->name.c_str(), ax->bytes);
  note_fused(bp.name.c_str(), bp.bytes);
  note_fused(cw.name.c_str(), cw.bytes);
  return (int)(i + 2 - at);
}

int pois_chain_rule_level2(std::vector<LoopRec> &run, size_t at, hipStream_t s) {
  (void)s;
  if (inst().fusion < 2) return 0;
  if (int k = pois_chain_rule_t<double, 0>(run, at)) return k;
  return pois_chain_rule_t<float, 1>(run, at);
}

// Jacobi step of the Laplace tutorial (apps/c/laplace2d_tutorial/step7/
// laplace2d.cpp: apply_stencil, its error result, copy).  copy(k), A = Anew,
// is still queued when apply_stencil(k + 1) arrives, and the two run as one
// pass over R:
//   A = Anew_old                                    (the copy)
//   Anew' = 0.25 (A'(x+1) + A'(x-1) + A'(y-1) + A'(y+1)),  err = max(err, |Anew' - A'|)
// where A' is A after the copy (Anew_old inside R, A's own values outside).
// Anew' goes to Anew's spare buffer (neighbouring waves still read
// Anew_old), which then becomes Anew's buffer.  24 bytes per point instead of
// 16 + 16, and one launch.  Each wave marches kJacRows rows of a (64 V)-wide
// strip keeping three rows in registers; x neighbours come from the adjacent
// lanes (shuffles) and, at the strip ends, one scalar load.
constexpr int kJacRows = 16;

template <class T>
struct JacPtrs {
  Ptr<T> A, out;
  Ptr<const T> An;
};

// FULL: every tile is 64 V points wide inside the range (every lane has V
// points): the element counts are literals, the edge values are loaded by
// every lane (one broadcast address) and kept by the edge lane, so a row has
// no lane-divergent branches (a branch join makes the compiler wait for
// every load in flight), and each row is loaded one row ahead of its use
template <class T, int V, bool FULL>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_jacobi_chain(JacPtrs<T> c, Range3 R, RedCtl ctl) {
  T err[1] = {RedOp<T, OPS_MAX>::identity()};
  const int lx = (int)threadIdx.x, wv = (int)threadIdx.y;
  const TileId t = tile_id();
  const int x0 = R.x0 + (int)t.x * kWave * V;
  const int ya = R.y0 + ((int)t.y * kWavesPerBlock + wv) * kJacRows;
  const int yb = min(ya + kJacRows, R.y1);
  const int z = R.z0;
  const int cnt = FULL ? kWave * V : min(kWave * V, R.x1 - x0);
  const int x = x0 + V * lx;
  const int n = FULL ? V : max(0, min(V, cnt - V * lx));
  const bool hxh = lx == (cnt - 1) / V;
  const bool in_xl = x0 - 1 >= R.x0, in_xh = x0 + cnt < R.x1;
  auto row = [&](int yy, T (&v)[V]) {
    const bool in = yy >= R.y0 && yy < R.y1;
    ldv<T, V>(in ? c.An.at(x, yy, z) : (const T *)c.A.at(x, yy, z), n, v);
  };
  auto edge = [&](int xx, int yy, bool inx) -> T {
    return inx && yy >= R.y0 && yy < R.y1 ? c.An(xx, yy, z) : c.A(xx, yy, z);
  };
  if (FULL && ya < yb) {
    // rows y + 1 and y + 2 are in registers while row y is computed (the
    // load of row y + 2 is issued one row ahead; rows past yb read row yb,
    // a row of the range or its halo)
    T up[V], cur[V], dn[V], dn2[V];
    row(ya - 1, up);
    row(ya, cur);
    row(ya + 1, dn);
    T cl = edge(x0 - 1, ya, in_xl), cr = edge(x0 + cnt, ya, in_xh);
    for (int y = ya; y < yb; y++) {
      row(min(y + 2, yb), dn2);
      T ncl = 0, ncr = 0;
      if (y + 1 < yb) {
        ncl = edge(x0 - 1, y + 1, in_xl);
        ncr = edge(x0 + cnt, y + 1, in_xh);
      }
      const T fromleft = __shfl_up(cur[V - 1], 1, kWave), fromright = __shfl_down(cur[0], 1, kWave);
      T v[V];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const T xm = j > 0 ? cur[j > 0 ? j - 1 : 0] : (lx == 0 ? cl : fromleft);
        const T xp = j + 1 < V ? cur[j + 1 < V ? j + 1 : j] : (hxh ? cr : fromright);
        v[j] = (T)0.25f * (xp + xm + up[j] + dn[j]);
        err[0] = RedOp<T, OPS_MAX>::combine(err[0], fabs(v[j] - cur[j]));
      }
      stv<T, V, false>(c.A.at(x, y, z), V, cur);
      stv<T, V, false>(c.out.at(x, y, z), V, v);
#pragma unroll
      for (int j = 0; j < V; j++) {
        up[j] = cur[j];
        cur[j] = dn[j];
        dn[j] = dn2[j];
      }
      cl = ncl;
      cr = ncr;
    }
  } else if (ya < yb) {
    T up[V], cur[V], dn[V];
    row(ya - 1, up);
    row(ya, cur);
    T cl = lx == 0 ? edge(x0 - 1, ya, in_xl) : (T)0, cr = hxh ? edge(x0 + cnt, ya, in_xh) : (T)0;
    for (int y = ya; y < yb; y++) {
      row(y + 1, dn);
      T ncl = 0, ncr = 0;
      if (y + 1 < yb) {
        if (lx == 0) ncl = edge(x0 - 1, y + 1, in_xl);
        if (hxh) ncr = edge(x0 + cnt, y + 1, in_xh);
      }
      const T fromleft = __shfl_up(cur[V - 1], 1, kWave), fromright = __shfl_down(cur[0], 1, kWave);
      T v[V];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const T xm = j > 0 ? cur[j > 0 ? j - 1 : 0] : (lx == 0 ? cl : fromleft);
        const T xp = j + 1 < n ? cur[j + 1 < V ? j + 1 : j] : (hxh ? cr : fromright);
        v[j] = (T)0.25f * (xp + xm + up[j] + dn[j]);
        if (j < n) err[0] = RedOp<T, OPS_MAX>::combine(err[0], fabs(v[j] - cur[j]));
      }
      if (n > 0) {
        stv<T, V, false>(c.A.at(x, y, z), n, cur);
        stv<T, V, false>(c.out.at(x, y, z), n, v);
      }
#pragma unroll
      for (int j = 0; j < V; j++) {
        up[j] = cur[j];
        cur[j] = dn[j];
      }
      cl = ncl;
      cr = ncr;
    }
  }
  reduce_block_store<T, OPS_MAX, 1>(err, ctl);
}

template <class T>
void launch_jacobi_chain(const LoopRec &cp, const LoopRec &jac, const Range3 &R) {
  constexpr int V = 16 / (int)sizeof(T);
  ops_dat A = jac.d[0], An = jac.d[1];
  const unsigned gx = (unsigned)((R.x1 - R.x0 + kWave * V - 1) / (kWave * V));
  const unsigned gy = (unsigned)((R.y1 - R.y0 + kWavesPerBlock * kJacRows - 1) / (kWavesPerBlock * kJacRows));
  ensure_partials((size_t)gx * gy * sizeof(T));
  JacPtrs<T> c;
  DatView v = view_of(An);
  c.out = Ptr<T>{(T *)spare_base(An, R), v.sy, v.sz};
  c.A = ptr_of<T>(A);
  c.An = ptr_of<const T>(An);
  const RedCtl ctl = with_host_fold(jac.ctl, gx * gy, 1, sizeof(T) == 8 ? DT_F64 : DT_F32);
  run_loop("jacobi_chain(copy+apply_stencil)", cp.bytes + jac.bytes, [&](hipStream_t st) {
    // every tile whole: the branch-free form
    auto kern = (R.x1 - R.x0) % (kWave * V) == 0 ? k_jacobi_chain<T, V, true> : k_jacobi_chain<T, V, false>;
    hipLaunchKernelGGL(kern, dim3(gx, gy, 1), dim3(kWave, kWavesPerBlock), 0, st, c, R, ctl);
    launch_red_finish<T, OPS_MAX, 1>(ctl, gx * gy, st);
  });
  swap_spare(An, R);
  note_external_write(A);
}

int jacobi_chain_rule(std::vector<LoopRec> &run, size_t at, hipStream_t s) {
  (void)s;
  if (inst().fusion < 2 || at + 1 >= run.size()) return 0;
  const LoopRec &cp = run[at], &jac = run[at + 1];
  if (cp.kind != LK_COPY || jac.kind != LK_JACOBI5 || cp.dims != 2 || jac.dims != 2) return 0;
  ops_dat A = jac.d[0], An = jac.d[1];
  if (cp.d[0] != A || cp.d[1] != An || A == An || cp.g[0] != jac.g[0]) return 0;
  if (!same_range(cp.range, jac.range) || jac.range.npts() <= 0 || jac.ctl.slot == nullptr) return 0;
  const Range3 R = jac.range;
  if (!use_vec() || !vec16_ok(A, R.x0) || !vec16_ok(An, R.x0)) return 0;
  if (jac.g[0] == 8) launch_jacobi_chain<double>(cp, jac, R);
  else launch_jacobi_chain<float>(cp, jac, R);
  note_fused(cp.name.c_str(), cp.bytes);
  note_fused(jac.name.c_str(), jac.bytes);
  return 2;
}

struct Register {
  Register() {
    register_fusion(cg_chain_rule_level2);
    register_fusion(pois_chain_rule_level2);
    register_fusion(jacobi_chain_rule);
    register_fusion(axpy_axpby_rule);
  }
} g_register;

}  // namespace
