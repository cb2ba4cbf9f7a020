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
// loops meet here.  One z-marching kernel reads u, p_old, r, Kx, Ky, Kz and
// writes u, p_new, w: 72 bytes per point instead of 24 + 24 + 40 = 88.
// p_new is needed at stencil neighbours before every thread has written it,
// so it is recomputed there from p_old and r and written into p's spare
// buffer, which then becomes p's buffer (double buffering).  Outside the
// loop range (the halo) p_new equals p_old, as the axpby leaves it.  Every
// value is computed with the same operations, in the same order, as the
// three separate loops.
#include "engine.h"

using namespace opsmi;

namespace {

struct ChainArgs {
  Ptr<double> u, pnew, w;
  Ptr<const double> pold, r, Kx, Ky, Kz;
  double alpha_u, a, b, rx, ry, rz;
};

template <bool AXPY>
__global__ __launch_bounds__(256) void k_cg_chain3(ChainArgs c, Range3 R, int chunk, RedCtl ctl) {
  double acc[1] = {0.0};
  const TileId t = tile_id();
  const int x = R.x0 + (int)t.x * kWave + (int)threadIdx.x;
  const int y = R.y0 + (int)t.y * kWavesPerBlock + (int)threadIdx.y;
  if (x < R.x1 && y < R.y1) {
    const int z0 = R.z0 + (int)t.z * chunk;
    const int z1 = min(z0 + chunk, R.z1);
    const double a = c.a, b = c.b;
    // p_new at (x',y',z'): the axpby value inside R, p_old outside
    auto pv = [&](int xx, int yy, int zz) -> double {
      const double po = c.pold(xx, yy, zz);
      if (xx < R.x0 || xx >= R.x1 || yy < R.y0 || yy >= R.y1 || zz < R.z0 || zz >= R.z1) return po;
      return a * po + b * c.r(xx, yy, zz);
    };
    double pm = pv(x, y, z0 - 1);
    double poc = c.pold(x, y, z0);
    double pc = a * poc + b * c.r(x, y, z0);
    double pn = pv(x, y, z0 + 1);
    double kzc = c.Kz(x, y, z0), kzn = c.Kz(x, y, z0 + 1);
    double kxc = c.Kx(x, y, z0), kxn = c.Kx(x + 1, y, z0), kyc = c.Ky(x, y, z0), kyn = c.Ky(x, y + 1, z0);
    double pxm = pv(x - 1, y, z0), pxp = pv(x + 1, y, z0), pym = pv(x, y - 1, z0), pyp = pv(x, y + 1, z0);
    double uc = AXPY ? c.u(x, y, z0) : 0.0;
    for (int z = z0; z < z1; z++) {
      const double v = (1.0 + c.ry * (kyn + kyc) + c.rx * (kxn + kxc) + c.rz * (kzn + kzc)) * pc -
                       c.ry * (kyn * pyp + kyc * pym) - c.rx * (kxn * pxp + kxc * pxm) -
                       c.rz * (kzn * pn + kzc * pm);
      const double pc_v = pc, poc_v = poc, uc_v = uc;
      if (z + 1 < z1) {
        const int zn = z + 1;
        const double pnn = pv(x, y, zn + 1), kznn = c.Kz(x, y, zn + 1);
        poc = c.pold(x, y, zn);
        if (AXPY) uc = c.u(x, y, zn);
        kxc = c.Kx(x, y, zn);
        kxn = c.Kx(x + 1, y, zn);
        kyc = c.Ky(x, y, zn);
        kyn = c.Ky(x, y + 1, zn);
        pxm = pv(x - 1, y, zn);
        pxp = pv(x + 1, y, zn);
        pym = pv(x, y - 1, zn);
        pyp = pv(x, y + 1, zn);
        pm = pc;
        pc = pn;
        pn = pnn;
        kzc = kzn;
        kzn = kznn;
      }
      if (AXPY) c.u(x, y, z) = uc_v + c.alpha_u * poc_v;
      c.pnew(x, y, z) = pc_v;
      c.w(x, y, z) = v;
      acc[0] = acc[0] + v * pc_v;
    }
  }
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
  const int nx = R.x1 - R.x0, ny = R.y1 - R.y0, nz = R.z1 - R.z0;
  const int chunk = heat_zchunk_default();
  const unsigned gx = (nx + kWave - 1) / kWave, gy = (ny + kWavesPerBlock - 1) / kWavesPerBlock;
  const unsigned gz = (nz + chunk - 1) / chunk;
  if (cw.ctl.slot == nullptr) return 0;
  ChainArgs c;
  DatView pv = view_of(P);
  c.pnew = Ptr<double>{(double *)spare_base(P), pv.sy, pv.sz};
  c.pold = ptr_of<const double>(P);
  c.r = ptr_of<const double>(Rr);
  c.w = ptr_of<double>(W);
  c.Kx = ptr_of<const double>(Kx);
  c.Ky = ptr_of<const double>(Ky);
  c.Kz = ptr_of<const double>(Kz);
  c.u = U ? ptr_of<double>(U) : Ptr<double>{nullptr, 0, 0};
  c.alpha_u = ax ? ax->g[0] : 0.0;
  c.a = bp.g[0];
  c.b = bp.g[1];
  c.rx = cw.g[0];
  c.ry = cw.g[1];
  c.rz = cw.g[2];
  const RedCtl ctl = with_scratch(cw.ctl);
  const double bytes = (ax ? ax->bytes : 0.0) + bp.bytes + cw.bytes;
  run_loop(ax ? "cg_chain(axpy+axpby+calc_w)" : "cg_chain(axpby+calc_w)", bytes, [&](hipStream_t st) {
    if (ax) hipLaunchKernelGGL(k_cg_chain3<true>, dim3(gx, gy, gz), dim3(kWave, kWavesPerBlock), 0, st, c, R, chunk, ctl);
    else hipLaunchKernelGGL(k_cg_chain3<false>, dim3(gx, gy, gz), dim3(kWave, kWavesPerBlock), 0, st, c, R, chunk, ctl);
    launch_red_finish<double, OPS_INC, 1>(ctl, gx * gy * gz, st);
  });
  (void)s;
  swap_spare(P);
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

// the three-loop chain recomputes p_new at every stencil neighbour, which
// costs more load instructions than it saves traffic on this GPU (measured
// 2363 us against 2345 us for the loops one by one at 512^3), so it is only
// tried when OPSMI_FUSE_CHAIN3=1; the pointwise pair is the default
bool chain3_enabled() {
  const char *e = getenv("OPSMI_FUSE_CHAIN3");
  return e && e[0] == '1';
}

struct Register {
  Register() {
    if (chain3_enabled()) register_fusion(cg_chain_rule);
    register_fusion(axpy_axpby_rule);
  }
} g_register;

}  // namespace
