// One 64 x 4 x (z1 - z0) column of the 3D TeaLeaf operator (the 7-point
// variable-coefficient heat operator of tea_leaf_cg_calc_w_reduce_kernel,
// apps/c/TeaLeaf/tea_leaf_cg_kernels.h:8-21, with the z pair appended after
// the x pair), computed by a 64 x 4 workgroup marching in z.
//
// Each thread keeps its column's p(z-1), p(z), p(z+1), Kz(z), Kz(z+1) in
// registers, so every value is loaded from memory once per plane.  The
// plane's p, Kx and Ky go to LDS with the rim the stencil needs (p on all
// four sides, Kx at x+1, Ky at y+1), double-buffered with one barrier per
// plane; the in-plane neighbours are read from there (4 global loads per
// point instead of 10).  Next-plane loads are issued before the barrier, so
// they are in flight while the plane is computed.  Lanes and rows past xe /
// ye skip loads and stores but keep to the barriers; the last active lane
// and row load the far rim.  Arithmetic in the order of the reference kernel.
//
// INIT: w = A p, r = p - w (tea_leaf_common_init_kernel); otherwise w = A p
// and the return value is this thread's sum of w p over its points.
#pragma once

#include "engine.h"

namespace opsmi {

struct HeatLds {
  double sp[2][kWavesPerBlock + 2][kWave + 2];
  double skx[2][kWavesPerBlock][kWave + 1];
  double sky[2][kWavesPerBlock + 1][kWave];
};

template <bool INIT>
__device__ __forceinline__ double heat3_column(HeatLds &L, const Ptr<double> &w, const Ptr<double> &r,
                                               const Ptr<const double> &Kx, const Ptr<const double> &Ky,
                                               const Ptr<const double> &Kz, const Ptr<const double> &p, double rx,
                                               double ry, double rz, int x0, int y0, int xe, int ye, int z0, int z1) {
  double acc = 0.0;
  const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
  const int lxe = min(kWave, xe - x0) - 1, lye = min(kWavesPerBlock, ye - y0) - 1;
  const bool act = lx <= lxe && ly <= lye;
  const bool hxl = lx == 0, hxh = lx == lxe, hyl = ly == 0, hyh = ly == lye;
  const int x = x0 + lx, y = y0 + ly;
  double pm = 0, pc = 0, pn = 0, kzc = 0, kzn = 0, kxc = 0, kyc = 0;
  double rxl = 0, rxh = 0, rkx = 0, ryl = 0, ryh = 0, rky = 0;  // rim values of this thread
  if (act) {
    pm = p(x, y, z0 - 1);
    pc = p(x, y, z0);
    pn = p(x, y, z0 + 1);
    kzc = Kz(x, y, z0);
    kzn = Kz(x, y, z0 + 1);
    kxc = Kx(x, y, z0);
    kyc = Ky(x, y, z0);
    if (hxl) rxl = p(x - 1, y, z0);
    if (hxh) {
      rxh = p(x + 1, y, z0);
      rkx = Kx(x + 1, y, z0);
    }
    if (hyl) ryl = p(x, y - 1, z0);
    if (hyh) {
      ryh = p(x, y + 1, z0);
      rky = Ky(x, y + 1, z0);
    }
  }
  for (int z = z0; z < z1; z++) {
    const int b = (z - z0) & 1;
    if (act) {
      L.sp[b][ly + 1][lx + 1] = pc;
      L.skx[b][ly][lx] = kxc;
      L.sky[b][ly][lx] = kyc;
      if (hxl) L.sp[b][ly + 1][0] = rxl;
      if (hxh) {
        L.sp[b][ly + 1][lx + 2] = rxh;
        L.skx[b][ly][lx + 1] = rkx;
      }
      if (hyl) L.sp[b][0][lx + 1] = ryl;
      if (hyh) {
        L.sp[b][ly + 2][lx + 1] = ryh;
        L.sky[b][ly + 1][lx] = rky;
      }
    }
    double pnn = 0, kznn = 0, kxn_c = 0, kyn_c = 0;
    if (act && z + 1 < z1) {
      const int zn = z + 1;
      pnn = p(x, y, zn + 1);
      kznn = Kz(x, y, zn + 1);
      kxn_c = Kx(x, y, zn);
      kyn_c = Ky(x, y, zn);
      if (hxl) rxl = p(x - 1, y, zn);
      if (hxh) {
        rxh = p(x + 1, y, zn);
        rkx = Kx(x + 1, y, zn);
      }
      if (hyl) ryl = p(x, y - 1, zn);
      if (hyh) {
        ryh = p(x, y + 1, zn);
        rky = Ky(x, y + 1, zn);
      }
    }
    __syncthreads();
    if (act) {
      const double pxm = L.sp[b][ly + 1][lx], pxp = L.sp[b][ly + 1][lx + 2];
      const double pym = L.sp[b][ly][lx + 1], pyp = L.sp[b][ly + 2][lx + 1];
      const double kxn = L.skx[b][ly][lx + 1], kyn = L.sky[b][ly + 1][lx];
      const double v = (1.0 + ry * (kyn + kyc) + rx * (kxn + kxc) + rz * (kzn + kzc)) * pc -
                       ry * (kyn * pyp + kyc * pym) - rx * (kxn * pxp + kxc * pxm) - rz * (kzn * pn + kzc * pm);
      w(x, y, z) = v;
      if constexpr (INIT) r(x, y, z) = pc - v;
      else acc = acc + v * pc;
    }
    pm = pc;
    pc = pn;
    pn = pnn;
    kzc = kzn;
    kzn = kznn;
    kxc = kxn_c;
    kyc = kyn_c;
  }
  return acc;
}

}  // namespace opsmi
