// z-marching stencil columns with V points per lane (16-byte vector loads).
//
// A 64 x WY workgroup (WY = 4 rows, or kHeatRows for the heat columns) owns
// a (64 V) x WY column of the grid over planes [z0, z1); lane l of row r
// holds points x0 + V l ... x0 + V l + V - 1 of row y0 + r.  Each thread keeps its points' p(z-1), p(z), p(z+1) (and, for the
// heat operator, Kz(z), Kz(z+1)) in registers, so every value is loaded from
// memory once per plane, as one 16-byte load per lane.  The plane's p (and
// Kx, Ky) go to LDS with the rim the stencil needs, double-buffered with one
// barrier per plane; x neighbours inside a lane come from registers, the
// two at a lane's ends and the y neighbours from LDS.  Loads of the next
// plane (or the next PF planes) are issued before the barrier.  Outputs are written with non-temporal stores
// when NT (the next loop reads them from HBM anyway at the sizes this path
// is for).  Lanes past the range end keep to the barriers; a lane with fewer
// than V points in range loads and stores them one by one, and the lane
// holding the last point loads the far x rim.
//
// Requirements (checked by the launchers): every array's point (x0, y, z)
// is 16-byte aligned for all rows and planes (row and plane pitches are
// multiples of 16 bytes and x0 is a multiple of V).
#pragma once

#include "engine.h"

namespace opsmi {

// LDS row slice (16-byte aligned when n == V)
template <class T, int V>
__device__ __forceinline__ void lds_st(T *p, int n, const T (&v)[V]) {
  if (n >= V) {
    vec_t<T, V> q;
#pragma unroll
    for (int j = 0; j < V; j++) q[j] = v[j];
    *reinterpret_cast<vec_t<T, V> *>(p) = q;
  } else {
#pragma unroll
    for (int j = 0; j < V; j++)
      if (j < n) p[j] = v[j];
  }
}
template <class T, int V>
__device__ __forceinline__ void lds_ld(const T *p, T (&o)[V]) {
  const vec_t<T, V> q = *reinterpret_cast<const vec_t<T, V> *>(p);
#pragma unroll
  for (int j = 0; j < V; j++) o[j] = q[j];
}

// lane geometry of a column tile
struct ColLane {
  int n;     // points of this lane in range (0..V)
  int cnt;   // points of the tile row in range
  int lye;   // last row in range
  bool hxl, hxh, hyl, hyh;
};
template <int V, int WY = kWavesPerBlock>
__device__ __forceinline__ ColLane col_lane(int x0, int y0, int xe, int ye) {
  ColLane c;
  const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
  c.cnt = min(kWave * V, xe - x0);
  c.lye = min(WY, ye - y0) - 1;
  c.n = ly <= c.lye ? max(0, min(V, c.cnt - V * lx)) : 0;
  c.hxl = lx == 0;
  c.hxh = lx == (c.cnt - 1) / V;
  c.hyl = ly == 0;
  c.hyh = ly == c.lye;
  return c;
}

// ------------------------------------------------------------ heat (7-point, variable coefficient, double)
// The heat columns are kHeatRows rows tall (512-thread workgroups): against
// 4-row tiles the y rims are half as many re-read rows per point, and at
// 512^3 the kernel measured 987 us against 1060 us.
constexpr int kHeatRows = 8;

template <int V, int WY = kWavesPerBlock>
struct HeatLdsV {
  __attribute__((aligned(16))) double sp[2][WY + 2][kWave * V + 2 * V];
  __attribute__((aligned(16))) double skx[2][WY][kWave * V + 2 * V];
  __attribute__((aligned(16))) double sky[2][WY + 1][kWave * V];
};

// Launch grid of a heat column kernel over r, tiles of tile_x x rows points.
// Planes per workgroup: OPSMI_HEAT_ZCHUNK when set, otherwise the longest
// power of two in [8, 64] that leaves at least 1024 workgroups (4 per CU);
// each chunk re-reads two planes of p and one of Kz, so longer is cheaper
// (512^3: 64 planes, 987 us against 1003 us with 32 and 1034 us with 16).
struct ColGrid {
  unsigned gx, gy, gz;
  int chunk;
  unsigned blocks() const { return gx * gy * gz; }
};
inline ColGrid heat_column_grid(const Range3 &r, int tile_x, int rows) {
  ColGrid g;
  const long nx = r.x1 - r.x0, ny = r.y1 - r.y0, nz = r.z1 - r.z0;
  g.gx = (unsigned)((nx + tile_x - 1) / tile_x);
  g.gy = (unsigned)((ny + rows - 1) / rows);
  if (getenv("OPSMI_HEAT_ZCHUNK")) {
    g.chunk = heat_zchunk_default();
  } else {
    g.chunk = 64;
    while (g.chunk > 8 && (long)g.gx * g.gy * ((nz + g.chunk - 1) / g.chunk) < 1024) g.chunk >>= 1;
  }
  g.gz = (unsigned)((nz + g.chunk - 1) / g.chunk);
  return g;
}

// one plane of a heat column as a lane loads it: its V points of p, Kz, Kx,
// Ky and the rims it owns
template <int V>
struct HeatPlane {
  double p[V], kz[V], kx[V], ky[V], ryl[V], ryh[V], rky[V];
  double rxl, rxh, rkx;
};

// TeaLeaf 3D operator (apps/c/TeaLeaf/tea_leaf_cg_kernels.h:8-21 with the z
// pair after the x pair), arithmetic in the reference order, over one column
// tile.  INIT: w = A p, r = p - w; otherwise w = A p and run() returns this
// thread's sum of w p.  Loads run PF planes ahead of the plane being
// computed, so a lane has PF planes of loads in flight while it waits at the
// barrier; the PF + 2 planes held in registers form a ring indexed by the
// step number modulo PF + 2, and the z loop is unrolled by the ring length
// so every index is a compile-time constant (no register moves).  Only p
// and Kz are read at plane z1 (the loop's stencils guarantee those, not the
// rest).  FULL: the tile lies inside the range (every lane has V points, all
// rows active), which drops the per-lane partial-load paths.
template <int V, bool INIT, bool NT, int WY, int PF, bool FULL>
struct HeatColumn {
  static constexpr int RING = PF + 2;
  HeatLdsV<V, WY> &L;
  const Ptr<double> &w, &r;
  const Ptr<const double> &Kx, &Ky, &Kz, &p;
  double rx, ry, rz;
  int x0, z0, z1;
  ColLane c;
  int x, y, o, lx, ly;
  double acc;
  double pm[V];
  // (one rim slot shared by the low and high rims of a full tile saves 10
  // VGPRs but makes the compiler wait for all loads before refilling it)
  using Plane = HeatPlane<V>;
  Plane R[RING];

  __device__ __forceinline__ HeatColumn(HeatLdsV<V, WY> &L_, const Ptr<double> &w_, const Ptr<double> &r_,
                                        const Ptr<const double> &Kx_, const Ptr<const double> &Ky_,
                                        const Ptr<const double> &Kz_, const Ptr<const double> &p_, double rx_,
                                        double ry_, double rz_, int x0_, int y0, int xe, int ye, int z0_, int z1_)
      : L(L_), w(w_), r(r_), Kx(Kx_), Ky(Ky_), Kz(Kz_), p(p_), rx(rx_), ry(ry_), rz(rz_), x0(x0_), z0(z0_), z1(z1_) {
    lx = (int)threadIdx.x;
    ly = (int)threadIdx.y;
    c = col_lane<V, WY>(x0, y0, xe, ye);
    if (FULL) c.n = V;
    x = x0 + V * lx;
    y = y0 + ly;
    o = V + V * lx;
    acc = 0.0;
  }
  __device__ __forceinline__ int n() const { return FULL ? V : c.n; }
  __device__ __forceinline__ bool act() const { return FULL || c.n > 0; }

  // plane q <= z1; the plane-z1 values of Kx, Ky and the rims are never
  // used and are read from plane z1 - 1 instead, so every step issues the
  // same loads (the compiler then keeps them in flight across the barrier)
  __device__ __forceinline__ void load(int q, Plane &P) const {
    const int qk = min(q, z1 - 1);
    ldv<double, V>(p.at(x, y, q), n(), P.p);
    ldv<double, V>(Kz.at(x, y, q), n(), P.kz);
    ldv<double, V>(Kx.at(x, y, qk), n(), P.kx);
    ldv<double, V>(Ky.at(x, y, qk), n(), P.ky);
    if (c.hxl) P.rxl = p(x0 - 1, y, qk);
    if (c.hxh) {
      P.rxh = p(x0 + c.cnt, y, qk);
      P.rkx = Kx(x0 + c.cnt, y, qk);
    }
    if (c.hyl) ldv<double, V>(p.at(x, y - 1, qk), n(), P.ryl);
    if (c.hyh) {
      ldv<double, V>(p.at(x, y + 1, qk), n(), P.ryh);
      ldv<double, V>(Ky.at(x, y + 1, qk), n(), P.rky);
    }
  }

  // plane z = z0 + s with s % RING == K
  template <int K>
  __device__ __forceinline__ void step(int z) {
    const Plane &C = R[K % RING], &N = R[(K + 1) % RING];
    const int b = (z - z0) & 1, nn = n();
    if (act()) {
      lds_st<double, V>(&L.sp[b][ly + 1][o], nn, C.p);
      lds_st<double, V>(&L.skx[b][ly][o], nn, C.kx);
      lds_st<double, V>(&L.sky[b][ly][V * lx], nn, C.ky);
      if (c.hxl) L.sp[b][ly + 1][V - 1] = C.rxl;
      if (c.hxh) {
        L.sp[b][ly + 1][V + c.cnt] = C.rxh;
        L.skx[b][ly][V + c.cnt] = C.rkx;
      }
      if (c.hyl) lds_st<double, V>(&L.sp[b][0][o], nn, C.ryl);
      if (c.hyh) {
        lds_st<double, V>(&L.sp[b][ly + 2][o], nn, C.ryh);
        lds_st<double, V>(&L.sky[b][ly + 1][V * lx], nn, C.rky);
      }
      // the slot takes plane z + PF + 1 (or, near the end, plane z1 again:
      // the slot held plane z - 1, which is no longer needed)
      load(min(z + PF + 1, z1), R[(K + PF + 1) % RING]);
    }
    __syncthreads();
    if (act()) {
      double pym[V], pyp[V], kyn[V], v[V], rv[V];
      lds_ld<double, V>(&L.sp[b][ly][o], pym);
      lds_ld<double, V>(&L.sp[b][ly + 2][o], pyp);
      lds_ld<double, V>(&L.sky[b][ly + 1][V * lx], kyn);
      const double left = L.sp[b][ly + 1][o - 1];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const double pxm = j == 0 ? left : C.p[j == 0 ? 0 : j - 1];
        const double pxp = j + 1 < nn ? C.p[j + 1 < V ? j + 1 : j] : L.sp[b][ly + 1][o + j + 1];
        const double kxn = j + 1 < nn ? C.kx[j + 1 < V ? j + 1 : j] : L.skx[b][ly][o + j + 1];
        v[j] = (1.0 + ry * (kyn[j] + C.ky[j]) + rx * (kxn + C.kx[j]) + rz * (N.kz[j] + C.kz[j])) * C.p[j] -
               ry * (kyn[j] * pyp[j] + C.ky[j] * pym[j]) - rx * (kxn * pxp + C.kx[j] * pxm) -
               rz * (N.kz[j] * N.p[j] + C.kz[j] * pm[j]);
        rv[j] = C.p[j] - v[j];
        if (!INIT && j < nn) acc = acc + v[j] * C.p[j];
      }
      stv<double, V, NT>(w.at(x, y, z), nn, v);
      if constexpr (INIT) stv<double, V, NT>(r.at(x, y, z), nn, rv);
    }
#pragma unroll
    for (int j = 0; j < V; j++) pm[j] = C.p[j];
  }
  template <int K>
  __device__ __forceinline__ void steps_all(int z) {
    if constexpr (K < RING) {
      step<K>(z);
      steps_all<K + 1>(z + 1);
    }
  }
  template <int K>
  __device__ __forceinline__ void steps_tail(int z) {
    if constexpr (K < RING - 1) {
      if (z < z1) {
        step<K>(z);
        steps_tail<K + 1>(z + 1);
      }
    }
  }

  __device__ __forceinline__ double run() {
#pragma unroll
    for (int i = 0; i < RING; i++) {
#pragma unroll
      for (int j = 0; j < V; j++)
        R[i].p[j] = R[i].kz[j] = R[i].kx[j] = R[i].ky[j] = R[i].ryl[j] = R[i].ryh[j] = R[i].rky[j] = 0.0;
      R[i].rxl = R[i].rxh = R[i].rkx = 0.0;
    }
#pragma unroll
    for (int j = 0; j < V; j++) pm[j] = 0.0;
    if (act()) {
      ldv<double, V>(p.at(x, y, z0 - 1), n(), pm);
#pragma unroll
      for (int i = 0; i < PF + 1; i++) load(min(z0 + i, z1), R[i]);
    }
    int z = z0;
    for (; z + RING <= z1; z += RING) steps_all<0>(z);
    steps_tail<0>(z);
    return acc;
  }
};

template <int V, bool INIT, bool NT, int WY = kWavesPerBlock, int PF = 1>
__device__ __forceinline__ double heat3_column_v(HeatLdsV<V, WY> &L, const Ptr<double> &w, const Ptr<double> &r,
                                                 const Ptr<const double> &Kx, const Ptr<const double> &Ky,
                                                 const Ptr<const double> &Kz, const Ptr<const double> &p, double rx,
                                                 double ry, double rz, int x0, int y0, int xe, int ye, int z0,
                                                 int z1) {
  if (x0 + kWave * V <= xe && y0 + WY <= ye)
    return HeatColumn<V, INIT, NT, WY, PF, true>(L, w, r, Kx, Ky, Kz, p, rx, ry, rz, x0, y0, xe, ye, z0, z1).run();
  return HeatColumn<V, INIT, NT, WY, PF, false>(L, w, r, Kx, Ky, Kz, p, rx, ry, rz, x0, y0, xe, ye, z0, z1).run();
}

// ------------------------------------------------------------ fused CG chain
// The TeaLeaf CG loops between two r.r results, as one column pass
// (apps/c/TeaLeaf/tea_leaf_cg.cpp:129-214):
//   [AXPY] u = u + alpha_u p_old          (tea_leaf_axpy_kernel, deferred)
//          p_new = a p_old + b r          (tea_leaf_axpby_kernel, inside R only)
//          w = A p_new ; return sum w p_new  (tea_leaf_cg_calc_w_reduce_kernel)
// p_new is formed from p_old and r wherever the stencil needs it (the
// column's planes and rims; p_old itself outside R, where the axpby does not
// write) with the same operations as the separate loops, and written to a
// second buffer (pnew) because neighbouring columns still read p_old.
struct ChainPtrs {
  Ptr<double> u, pnew, w;
  Ptr<const double> pold, r, Kx, Ky, Kz;
};

template <int V, bool AXPY, int WY = kWavesPerBlock>
__device__ __forceinline__ double heat3_chain_column_v(HeatLdsV<V, WY> &L, const ChainPtrs &c, const Range3 &R,
                                                       double alpha_u, double a, double b, double rx, double ry,
                                                       double rz, int x0, int y0, int z0, int z1) {
  double acc = 0.0;
  const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
  const ColLane cl = col_lane<V, WY>(x0, y0, R.x1, R.y1);
  const int n = cl.n;
  const bool act = n > 0;
  const int x = x0 + V * lx, y = y0 + ly;
  const int o = V + V * lx;
  // rims inside R take the axpby value
  const bool in_xl = x0 - 1 >= R.x0, in_xh = x0 + cl.cnt < R.x1;
  const bool in_yl = y - 1 >= R.y0, in_yh = y + 1 < R.y1;
  auto pnew1 = [&](int xx, int yy, int zz, bool in) -> double {
    const double po = c.pold(xx, yy, zz);
    return in ? a * po + b * c.r(xx, yy, zz) : po;
  };
  // p_new of V points of row yy, plane zz into o_ (p_old into po)
  auto pnewv = [&](int yy, int zz, bool in, double (&o_)[V], double (&po)[V]) {
    double rv[V];
    ldv<double, V>(c.pold.at(x, yy, zz), n, po);
    if (in) {
      ldv<double, V>(c.r.at(x, yy, zz), n, rv);
#pragma unroll
      for (int j = 0; j < V; j++) o_[j] = a * po[j] + b * rv[j];
    } else {
#pragma unroll
      for (int j = 0; j < V; j++) o_[j] = po[j];
    }
  };
  double pm[V], pc[V], pn[V], poc[V], pon[V], kzc[V], kzn[V], kxc[V], kyc[V], uc[V];
  double ryl[V], ryh[V], rky[V], scratch[V];
  double rxl = 0, rxh = 0, rkx = 0;
#pragma unroll
  for (int j = 0; j < V; j++)
    pm[j] = pc[j] = pn[j] = poc[j] = pon[j] = kzc[j] = kzn[j] = kxc[j] = kyc[j] = uc[j] = ryl[j] = ryh[j] =
        rky[j] = 0;
  if (act) {
    pnewv(y, z0 - 1, z0 - 1 >= R.z0, pm, scratch);
    pnewv(y, z0, true, pc, poc);
    pnewv(y, z0 + 1, z0 + 1 < R.z1, pn, pon);
    ldv<double, V>(c.Kz.at(x, y, z0), n, kzc);
    ldv<double, V>(c.Kz.at(x, y, z0 + 1), n, kzn);
    ldv<double, V>(c.Kx.at(x, y, z0), n, kxc);
    ldv<double, V>(c.Ky.at(x, y, z0), n, kyc);
    if (AXPY) ldv<double, V>(c.u.at(x, y, z0), n, uc);
    if (cl.hxl) rxl = pnew1(x0 - 1, y, z0, in_xl);
    if (cl.hxh) {
      rxh = pnew1(x0 + cl.cnt, y, z0, in_xh);
      rkx = c.Kx(x0 + cl.cnt, y, z0);
    }
    if (cl.hyl) pnewv(y - 1, z0, in_yl, ryl, scratch);
    if (cl.hyh) {
      pnewv(y + 1, z0, in_yh, ryh, scratch);
      ldv<double, V>(c.Ky.at(x, y + 1, z0), n, rky);
    }
  }
  for (int z = z0; z < z1; z++) {
    const int bb = (z - z0) & 1;
    if (act) {
      lds_st<double, V>(&L.sp[bb][ly + 1][o], n, pc);
      lds_st<double, V>(&L.skx[bb][ly][o], n, kxc);
      lds_st<double, V>(&L.sky[bb][ly][V * lx], n, kyc);
      if (cl.hxl) L.sp[bb][ly + 1][V - 1] = rxl;
      if (cl.hxh) {
        L.sp[bb][ly + 1][V + cl.cnt] = rxh;
        L.skx[bb][ly][V + cl.cnt] = rkx;
      }
      if (cl.hyl) lds_st<double, V>(&L.sp[bb][0][o], n, ryl);
      if (cl.hyh) {
        lds_st<double, V>(&L.sp[bb][ly + 2][o], n, ryh);
        lds_st<double, V>(&L.sky[bb][ly + 1][V * lx], n, rky);
      }
    }
    // next plane: p_new and p_old of z + 2 for the register window, u(z + 1)
    // for the axpy, coefficients and rims of plane z + 1
    double pnn[V], ponn[V], kznn[V], kxn_c[V], kyn_c[V], un[V];
#pragma unroll
    for (int j = 0; j < V; j++) pnn[j] = ponn[j] = kznn[j] = kxn_c[j] = kyn_c[j] = un[j] = 0;
    if (act && z + 1 < z1) {
      const int zn = z + 1;
      pnewv(y, zn + 1, zn + 1 < R.z1, pnn, ponn);
      ldv<double, V>(c.Kz.at(x, y, zn + 1), n, kznn);
      ldv<double, V>(c.Kx.at(x, y, zn), n, kxn_c);
      ldv<double, V>(c.Ky.at(x, y, zn), n, kyn_c);
      if (AXPY) ldv<double, V>(c.u.at(x, y, zn), n, un);
      if (cl.hxl) rxl = pnew1(x0 - 1, y, zn, in_xl);
      if (cl.hxh) {
        rxh = pnew1(x0 + cl.cnt, y, zn, in_xh);
        rkx = c.Kx(x0 + cl.cnt, y, zn);
      }
      if (cl.hyl) pnewv(y - 1, zn, in_yl, ryl, scratch);
      if (cl.hyh) {
        pnewv(y + 1, zn, in_yh, ryh, scratch);
        ldv<double, V>(c.Ky.at(x, y + 1, zn), n, rky);
      }
    }
    __syncthreads();
    if (act) {
      double pym[V], pyp[V], kyn[V], v[V], unew[V];
      lds_ld<double, V>(&L.sp[bb][ly][o], pym);
      lds_ld<double, V>(&L.sp[bb][ly + 2][o], pyp);
      lds_ld<double, V>(&L.sky[bb][ly + 1][V * lx], kyn);
      const double left = L.sp[bb][ly + 1][o - 1];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const double pxm = j == 0 ? left : pc[j == 0 ? 0 : j - 1];
        const double pxp = j + 1 < n ? pc[j + 1 < V ? j + 1 : j] : L.sp[bb][ly + 1][o + j + 1];
        const double kxn = j + 1 < n ? kxc[j + 1 < V ? j + 1 : j] : L.skx[bb][ly][o + j + 1];
        v[j] = (1.0 + ry * (kyn[j] + kyc[j]) + rx * (kxn + kxc[j]) + rz * (kzn[j] + kzc[j])) * pc[j] -
               ry * (kyn[j] * pyp[j] + kyc[j] * pym[j]) - rx * (kxn * pxp + kxc[j] * pxm) -
               rz * (kzn[j] * pn[j] + kzc[j] * pm[j]);
        unew[j] = uc[j] + alpha_u * poc[j];
        if (j < n) acc = acc + v[j] * pc[j];
      }
      stv<double, V, true>(c.w.at(x, y, z), n, v);
      stv<double, V, true>(c.pnew.at(x, y, z), n, pc);
      if (AXPY) stv<double, V, true>(c.u.at(x, y, z), n, unew);
    }
#pragma unroll
    for (int j = 0; j < V; j++) {
      pm[j] = pc[j];
      pc[j] = pn[j];
      pn[j] = pnn[j];
      poc[j] = pon[j];
      pon[j] = ponn[j];
      kzc[j] = kzn[j];
      kzn[j] = kznn[j];
      kxc[j] = kxn_c[j];
      kyc[j] = kyn_c[j];
      uc[j] = un[j];
    }
  }
  return acc;
}

// ------------------------------------------------------------ Poisson (7-point, constant coefficient)
template <class T, int V>
struct PoisLdsV {
  __attribute__((aligned(16))) T sp[2][kWavesPerBlock + 2][kWave * V + 2 * V];
};

// w = 6p - (p(-1,0,0) + p(1,0,0)) - (p(0,-1,0) + p(0,1,0)) - (p(0,0,-1) +
// p(0,0,1)) in T; returns this thread's sum of (double) w (double) p.
template <class T, int V, bool NT>
__device__ __forceinline__ double poisson3_column_v(PoisLdsV<T, V> &L, const Ptr<T> &w, const Ptr<const T> &p, int x0,
                                                    int y0, int xe, int ye, int z0, int z1) {
  double acc = 0.0;
  const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
  const ColLane c = col_lane<V>(x0, y0, xe, ye);
  const int n = c.n;
  const bool act = n > 0;
  const int x = x0 + V * lx, y = y0 + ly;
  const int o = V + V * lx;
  T pm[V], pc[V], pn[V], ryl[V], ryh[V];
  T rxl = 0, rxh = 0;
#pragma unroll
  for (int j = 0; j < V; j++) pm[j] = pc[j] = pn[j] = ryl[j] = ryh[j] = 0;
  if (act) {
    ldv<T, V>(p.at(x, y, z0 - 1), n, pm);
    ldv<T, V>(p.at(x, y, z0), n, pc);
    ldv<T, V>(p.at(x, y, z0 + 1), n, pn);
    if (c.hxl) rxl = p(x0 - 1, y, z0);
    if (c.hxh) rxh = p(x0 + c.cnt, y, z0);
    if (c.hyl) ldv<T, V>(p.at(x, y - 1, z0), n, ryl);
    if (c.hyh) ldv<T, V>(p.at(x, y + 1, z0), n, ryh);
  }
  for (int z = z0; z < z1; z++) {
    const int b = (z - z0) & 1;
    if (act) {
      lds_st<T, V>(&L.sp[b][ly + 1][o], n, pc);
      if (c.hxl) L.sp[b][ly + 1][V - 1] = rxl;
      if (c.hxh) L.sp[b][ly + 1][V + c.cnt] = rxh;
      if (c.hyl) lds_st<T, V>(&L.sp[b][0][o], n, ryl);
      if (c.hyh) lds_st<T, V>(&L.sp[b][ly + 2][o], n, ryh);
    }
    T pnn[V];
#pragma unroll
    for (int j = 0; j < V; j++) pnn[j] = 0;
    if (act && z + 1 < z1) {
      const int zn = z + 1;
      ldv<T, V>(p.at(x, y, zn + 1), n, pnn);
      if (c.hxl) rxl = p(x0 - 1, y, zn);
      if (c.hxh) rxh = p(x0 + c.cnt, y, zn);
      if (c.hyl) ldv<T, V>(p.at(x, y - 1, zn), n, ryl);
      if (c.hyh) ldv<T, V>(p.at(x, y + 1, zn), n, ryh);
    }
    __syncthreads();
    if (act) {
      T pym[V], pyp[V], v[V];
      lds_ld<T, V>(&L.sp[b][ly][o], pym);
      lds_ld<T, V>(&L.sp[b][ly + 2][o], pyp);
      const T left = L.sp[b][ly + 1][o - 1];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const T pxm = j == 0 ? left : pc[j == 0 ? 0 : j - 1];
        const T pxp = j + 1 < n ? pc[j + 1 < V ? j + 1 : j] : L.sp[b][ly + 1][o + j + 1];
        v[j] = (T)6.0 * pc[j] - (pxm + pxp) - (pym[j] + pyp[j]) - (pm[j] + pn[j]);
        if (j < n) acc = acc + (double)v[j] * (double)pc[j];
      }
      stv<T, V, NT>(w.at(x, y, z), n, v);
    }
#pragma unroll
    for (int j = 0; j < V; j++) {
      pm[j] = pc[j];
      pc[j] = pn[j];
      pn[j] = pnn[j];
    }
  }
  return acc;
}

}  // namespace opsmi
