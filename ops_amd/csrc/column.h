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
// is for).  The coefficient planes Kx, Ky, Kz and the axpy target u, which no
// other workgroup reads, load with the non-temporal hint; p and r, whose
// edge rows are neighbouring tiles' rims, load without it (TeaLeaf chain at
// 512^3: 1788 us, against 1820 us with no hints and 1824 us with hints on
// every centre load).  Lanes past the range end keep to the barriers; a lane with fewer
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

template <int V, int WY = kWavesPerBlock, class T = double>
struct HeatLdsV {
  __attribute__((aligned(16))) T sp[2][WY + 2][kWave * V + 2 * V];
  __attribute__((aligned(16))) T skx[2][WY][kWave * V + 2 * V];
  __attribute__((aligned(16))) T sky[2][WY + 1][kWave * V];
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
template <int V, class T = double>
struct HeatPlane {
  T p[V], kz[V], kx[V], ky[V], ryl[V], ryh[V], rky[V];
  T rxl, rxh, rkx;
};

// Output stages of the heat column: called per lane and plane with the
// operator's values v = (A p) at the lane's nn points and the centre values
// pc of p.
//   HeatOutW:    w = v; acc = sum of v p (the calc_w reduction)
//   HeatOutInit: w = v, r = p - v (tea_leaf_common_init's residual)
template <class T, int V, bool NT>
struct HeatOutW {
  const Ptr<T> &w;
  double acc = 0.0;
  __device__ __forceinline__ void operator()(int x, int y, int z, int nn, const T (&v)[V], const T (&pc)[V]) {
#pragma unroll
    for (int j = 0; j < V; j++)
      if (j < nn) acc = acc + (double)v[j] * (double)pc[j];
    stv<T, V, NT>(w.at(x, y, z), nn, v);
  }
};
template <class T, int V, bool NT>
struct HeatOutInit {
  const Ptr<T> &w, &r;
  double acc = 0.0;
  __device__ __forceinline__ void operator()(int x, int y, int z, int nn, const T (&v)[V], const T (&pc)[V]) {
    T rv[V];
#pragma unroll
    for (int j = 0; j < V; j++) rv[j] = pc[j] - v[j];
    stv<T, V, NT>(w.at(x, y, z), nn, v);
    stv<T, V, NT>(r.at(x, y, z), nn, rv);
  }
};

// TeaLeaf 3D operator (apps/c/TeaLeaf/tea_leaf_cg_kernels.h:8-21 with the z
// pair after the x pair), arithmetic in the reference order, over one column
// tile, in element type T (double; float for the FP32 inner sweeps of the
// batched engine), the result of each plane handed to the output stage OUT
// (above).  Loads run PF planes ahead of the plane being
// computed, so a lane has PF planes of loads in flight while it waits at the
// barrier; the PF + 2 planes held in registers form a ring indexed by the
// step number modulo PF + 2, and the z loop is unrolled by the ring length
// so every index is a compile-time constant (no register moves).  Only p
// and Kz are read at plane z1 (the loop's stencils guarantee those, not the
// rest).  FULL: the tile lies inside the range (every lane has V points, all
// rows active), which drops the per-lane partial-load paths.
template <class T, int V, int WY, int PF, bool FULL, class OUT>
struct HeatColumn {
  static constexpr int RING = PF + 2;
  HeatLdsV<V, WY, T> &L;
  OUT &out;
  const Ptr<const T> &Kx, &Ky, &Kz, &p;
  T rx, ry, rz;
  int x0, z0, z1;
  ColLane c;
  int x, y, o, lx, ly;
  T pm[V];
  // (one rim slot shared by the low and high rims of a full tile saves 10
  // VGPRs but makes the compiler wait for all loads before refilling it)
  using Plane = HeatPlane<V, T>;
  Plane R[RING];

  __device__ __forceinline__ HeatColumn(HeatLdsV<V, WY, T> &L_, OUT &out_, const Ptr<const T> &Kx_,
                                        const Ptr<const T> &Ky_, const Ptr<const T> &Kz_, const Ptr<const T> &p_,
                                        T rx_, T ry_, T rz_, int x0_, int y0, int xe, int ye, int z0_, int z1_)
      : L(L_), out(out_), Kx(Kx_), Ky(Ky_), Kz(Kz_), p(p_), rx(rx_), ry(ry_), rz(rz_), x0(x0_), z0(z0_), z1(z1_) {
    lx = (int)threadIdx.x;
    ly = (int)threadIdx.y;
    c = col_lane<V, WY>(x0, y0, xe, ye);
    if (FULL) c.n = V;
    x = x0 + V * lx;
    y = y0 + ly;
    o = V + V * lx;
  }
  __device__ __forceinline__ int n() const { return FULL ? V : c.n; }
  __device__ __forceinline__ bool act() const { return FULL || c.n > 0; }

  // plane q <= z1; the plane-z1 values of Kx, Ky and the rims are never
  // used and are read from plane z1 - 1 instead, so every step issues the
  // same loads (the compiler then keeps them in flight across the barrier)
  __device__ __forceinline__ void load(int q, Plane &P) const {
    const int qk = min(q, z1 - 1);
    ldv<T, V>(p.at(x, y, q), n(), P.p);
    ldv<T, V, true>(Kz.at(x, y, q), n(), P.kz);
    ldv<T, V, true>(Kx.at(x, y, qk), n(), P.kx);
    ldv<T, V, true>(Ky.at(x, y, qk), n(), P.ky);
    if (c.hxl) P.rxl = p(x0 - 1, y, qk);
    if (c.hxh) {
      P.rxh = p(x0 + c.cnt, y, qk);
      P.rkx = Kx(x0 + c.cnt, y, qk);
    }
    if (c.hyl) ldv<T, V>(p.at(x, y - 1, qk), n(), P.ryl);
    if (c.hyh) {
      ldv<T, V>(p.at(x, y + 1, qk), n(), P.ryh);
      ldv<T, V>(Ky.at(x, y + 1, qk), n(), P.rky);
    }
  }

  // plane z = z0 + s with s % RING == K
  template <int K>
  __device__ __forceinline__ void step(int z) {
    const Plane &C = R[K % RING], &N = R[(K + 1) % RING];
    const int b = (z - z0) & 1, nn = n();
    if (act()) {
      lds_st<T, V>(&L.sp[b][ly + 1][o], nn, C.p);
      lds_st<T, V>(&L.skx[b][ly][o], nn, C.kx);
      lds_st<T, V>(&L.sky[b][ly][V * lx], nn, C.ky);
      if (c.hxl) L.sp[b][ly + 1][V - 1] = C.rxl;
      if (c.hxh) {
        L.sp[b][ly + 1][V + c.cnt] = C.rxh;
        L.skx[b][ly][V + c.cnt] = C.rkx;
      }
      if (c.hyl) lds_st<T, V>(&L.sp[b][0][o], nn, C.ryl);
      if (c.hyh) {
        lds_st<T, V>(&L.sp[b][ly + 2][o], nn, C.ryh);
        lds_st<T, V>(&L.sky[b][ly + 1][V * lx], nn, C.rky);
      }
      // the slot takes plane z + PF + 1 (or, near the end, plane z1 again:
      // the slot held plane z - 1, which is no longer needed)
      load(min(z + PF + 1, z1), R[(K + PF + 1) % RING]);
    }
    __syncthreads();
    if (act()) {
      T pym[V], pyp[V], kyn[V], v[V];
      lds_ld<T, V>(&L.sp[b][ly][o], pym);
      lds_ld<T, V>(&L.sp[b][ly + 2][o], pyp);
      lds_ld<T, V>(&L.sky[b][ly + 1][V * lx], kyn);
      const T left = L.sp[b][ly + 1][o - 1];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const T pxm = j == 0 ? left : C.p[j == 0 ? 0 : j - 1];
        const T pxp = j + 1 < nn ? C.p[j + 1 < V ? j + 1 : j] : L.sp[b][ly + 1][o + j + 1];
        const T kxn = j + 1 < nn ? C.kx[j + 1 < V ? j + 1 : j] : L.skx[b][ly][o + j + 1];
        v[j] = ((T)1.0 + ry * (kyn[j] + C.ky[j]) + rx * (kxn + C.kx[j]) + rz * (N.kz[j] + C.kz[j])) * C.p[j] -
               ry * (kyn[j] * pyp[j] + C.ky[j] * pym[j]) - rx * (kxn * pxp + C.kx[j] * pxm) -
               rz * (N.kz[j] * N.p[j] + C.kz[j] * pm[j]);
      }
      out(x, y, z, nn, v, C.p);
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

  __device__ __forceinline__ void run() {
#pragma unroll
    for (int i = 0; i < RING; i++) {
#pragma unroll
      for (int j = 0; j < V; j++)
        R[i].p[j] = R[i].kz[j] = R[i].kx[j] = R[i].ky[j] = R[i].ryl[j] = R[i].ryh[j] = R[i].rky[j] = (T)0;
      R[i].rxl = R[i].rxh = R[i].rkx = (T)0;
    }
#pragma unroll
    for (int j = 0; j < V; j++) pm[j] = (T)0;
    if (act()) {
      ldv<T, V>(p.at(x, y, z0 - 1), n(), pm);
#pragma unroll
      for (int i = 0; i < PF + 1; i++) load(min(z0 + i, z1), R[i]);
    }
    int z = z0;
    for (; z + RING <= z1; z += RING) steps_all<0>(z);
    steps_tail<0>(z);
  }
};

// the column over one tile with output stage out: the branch-free form when
// the tile lies inside the range
template <class T, int V, int WY, int PF, class OUT>
__device__ __forceinline__ void heat3_column_out(HeatLdsV<V, WY, T> &L, OUT &out, const Ptr<const T> &Kx,
                                                 const Ptr<const T> &Ky, const Ptr<const T> &Kz, const Ptr<const T> &p,
                                                 T rx, T ry, T rz, int x0, int y0, int xe, int ye, int z0, int z1) {
  if (x0 + kWave * V <= xe && y0 + WY <= ye)
    HeatColumn<T, V, WY, PF, true, OUT>(L, out, Kx, Ky, Kz, p, rx, ry, rz, x0, y0, xe, ye, z0, z1).run();
  else
    HeatColumn<T, V, WY, PF, false, OUT>(L, out, Kx, Ky, Kz, p, rx, ry, rz, x0, y0, xe, ye, z0, z1).run();
}

// INIT: w = A p, r = p - w (returns 0); otherwise w = A p, returns this
// thread's sum of w p
template <int V, bool INIT, bool NT, int WY = kWavesPerBlock, int PF = 1>
__device__ __forceinline__ double heat3_column_v(HeatLdsV<V, WY> &L, const Ptr<double> &w, const Ptr<double> &r,
                                                 const Ptr<const double> &Kx, const Ptr<const double> &Ky,
                                                 const Ptr<const double> &Kz, const Ptr<const double> &p, double rx,
                                                 double ry, double rz, int x0, int y0, int xe, int ye, int z0,
                                                 int z1) {
  if constexpr (INIT) {
    HeatOutInit<double, V, NT> out{w, r};
    heat3_column_out<double, V, WY, PF>(L, out, Kx, Ky, Kz, p, rx, ry, rz, x0, y0, xe, ye, z0, z1);
    return 0.0;
  } else {
    HeatOutW<double, V, NT> out{w};
    heat3_column_out<double, V, WY, PF>(L, out, Kx, Ky, Kz, p, rx, ry, rz, x0, y0, xe, ye, z0, z1);
    return out.acc;
  }
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

// one plane of a chain column as a lane loads it, raw: p_old and r of its
// points (p_new is formed when the plane is used), Kz, Kx, Ky, u and the
// rims' p_old, r (and Kx, Ky) values
template <int V>
struct ChainPlane {
  double po[V], rr[V], kz[V], kx[V], ky[V], u[V];
  double yl_po[V], yl_r[V], yh_po[V], yh_r[V], yh_ky[V];
  double xl_po, xl_r, xh_po, xh_r, xh_kx;
};

// The chain over one column tile.  Loads run one plane ahead (the three
// planes in registers form a ring, the z loop is unrolled by three), every
// step issues the same loads, and p_new is formed from the raw values only
// when the plane is used, so loads stay in flight across the barrier.  r is
// read only inside R (elsewhere from a clamped in-range point whose value is
// not used), Kx, Ky, u and the rims only up to plane z1 - 1.
template <int V, bool AXPY, int WY, bool FULL>
struct ChainColumn {
  static constexpr int RING = 3;
  HeatLdsV<V, WY> &L;
  const ChainPtrs &c;
  const Range3 &R;
  double alpha_u, a, b, rx, ry, rz;
  int x0, z0, z1;
  ColLane cl;
  int x, y, o, lx, ly;
  bool in_xl, in_xh, in_yl, in_yh;
  int xl_r, xh_r, yl_r, yh_r;  // where the rims read r (inside R)
  double acc;
  double pm[V], pc[V];
  ChainPlane<V> P[RING];

  __device__ __forceinline__ ChainColumn(HeatLdsV<V, WY> &L_, const ChainPtrs &c_, const Range3 &R_, double alpha_u_,
                                         double a_, double b_, double rx_, double ry_, double rz_, int x0_, int y0,
                                         int z0_, int z1_)
      : L(L_), c(c_), R(R_), alpha_u(alpha_u_), a(a_), b(b_), rx(rx_), ry(ry_), rz(rz_), x0(x0_), z0(z0_), z1(z1_) {
    lx = (int)threadIdx.x;
    ly = (int)threadIdx.y;
    cl = col_lane<V, WY>(x0, y0, R.x1, R.y1);
    if (FULL) cl.n = V;
    x = x0 + V * lx;
    y = y0 + ly;
    o = V + V * lx;
    in_xl = x0 - 1 >= R.x0;
    in_xh = x0 + cl.cnt < R.x1;
    in_yl = y - 1 >= R.y0;
    in_yh = y + 1 < R.y1;
    xl_r = in_xl ? x0 - 1 : x0;
    xh_r = in_xh ? x0 + cl.cnt : x0 + cl.cnt - 1;
    yl_r = in_yl ? y - 1 : y;
    yh_r = in_yh ? y + 1 : y;
    acc = 0.0;
  }
  __device__ __forceinline__ int n() const { return FULL ? V : cl.n; }
  __device__ __forceinline__ bool act() const { return FULL || cl.n > 0; }
  // p_new of one value: the axpby inside R, p_old outside
  __device__ __forceinline__ double pnew(double po, double r, bool in) const { return in ? a * po + b * r : po; }

  __device__ __forceinline__ void load(int q, ChainPlane<V> &Pl) const {
    const int qk = min(q, z1 - 1);
    const int qr = min(max(q, R.z0), R.z1 - 1);
    const int nn = n();
    ldv<double, V>(c.pold.at(x, y, q), nn, Pl.po);
    ldv<double, V>(c.r.at(x, y, qr), nn, Pl.rr);
    ldv<double, V, true>(c.Kz.at(x, y, q), nn, Pl.kz);
    ldv<double, V, true>(c.Kx.at(x, y, qk), nn, Pl.kx);
    ldv<double, V, true>(c.Ky.at(x, y, qk), nn, Pl.ky);
    if (AXPY) ldv<double, V, true>(c.u.at(x, y, qk), nn, Pl.u);
    if (cl.hxl) {
      Pl.xl_po = c.pold(x0 - 1, y, qk);
      Pl.xl_r = c.r(xl_r, y, qk);
    }
    if (cl.hxh) {
      Pl.xh_po = c.pold(x0 + cl.cnt, y, qk);
      Pl.xh_r = c.r(xh_r, y, qk);
      Pl.xh_kx = c.Kx(x0 + cl.cnt, y, qk);
    }
    if (cl.hyl) {
      ldv<double, V>(c.pold.at(x, y - 1, qk), nn, Pl.yl_po);
      ldv<double, V>(c.r.at(x, yl_r, qk), nn, Pl.yl_r);
    }
    if (cl.hyh) {
      ldv<double, V>(c.pold.at(x, y + 1, qk), nn, Pl.yh_po);
      ldv<double, V>(c.r.at(x, yh_r, qk), nn, Pl.yh_r);
      ldv<double, V>(c.Ky.at(x, y + 1, qk), nn, Pl.yh_ky);
    }
  }

  template <int K>
  __device__ __forceinline__ void step(int z) {
    const ChainPlane<V> &C = P[K % RING], &N = P[(K + 1) % RING];
    const int bb = (z - z0) & 1, nn = n();
    double pn[V];
    if (act()) {
      lds_st<double, V>(&L.sp[bb][ly + 1][o], nn, pc);
      lds_st<double, V>(&L.skx[bb][ly][o], nn, C.kx);
      lds_st<double, V>(&L.sky[bb][ly][V * lx], nn, C.ky);
      if (cl.hxl) L.sp[bb][ly + 1][V - 1] = pnew(C.xl_po, C.xl_r, in_xl);
      if (cl.hxh) {
        L.sp[bb][ly + 1][V + cl.cnt] = pnew(C.xh_po, C.xh_r, in_xh);
        L.skx[bb][ly][V + cl.cnt] = C.xh_kx;
      }
      if (cl.hyl) {
        double t[V];
#pragma unroll
        for (int j = 0; j < V; j++) t[j] = pnew(C.yl_po[j], C.yl_r[j], in_yl);
        lds_st<double, V>(&L.sp[bb][0][o], nn, t);
      }
      if (cl.hyh) {
        double t[V];
#pragma unroll
        for (int j = 0; j < V; j++) t[j] = pnew(C.yh_po[j], C.yh_r[j], in_yh);
        lds_st<double, V>(&L.sp[bb][ly + 2][o], nn, t);
        lds_st<double, V>(&L.sky[bb][ly + 1][V * lx], nn, C.yh_ky);
      }
      load(min(z + 2, z1), P[(K + 2) % RING]);
    }
    __syncthreads();
    const bool in_n = z + 1 < R.z1;
#pragma unroll
    for (int j = 0; j < V; j++) pn[j] = pnew(N.po[j], N.rr[j], in_n);
    if (act()) {
      double pym[V], pyp[V], kyn[V], v[V], unew[V];
      lds_ld<double, V>(&L.sp[bb][ly][o], pym);
      lds_ld<double, V>(&L.sp[bb][ly + 2][o], pyp);
      lds_ld<double, V>(&L.sky[bb][ly + 1][V * lx], kyn);
      const double left = L.sp[bb][ly + 1][o - 1];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const double pxm = j == 0 ? left : pc[j == 0 ? 0 : j - 1];
        const double pxp = j + 1 < nn ? pc[j + 1 < V ? j + 1 : j] : L.sp[bb][ly + 1][o + j + 1];
        const double kxn = j + 1 < nn ? C.kx[j + 1 < V ? j + 1 : j] : L.skx[bb][ly][o + j + 1];
        v[j] = (1.0 + ry * (kyn[j] + C.ky[j]) + rx * (kxn + C.kx[j]) + rz * (N.kz[j] + C.kz[j])) * pc[j] -
               ry * (kyn[j] * pyp[j] + C.ky[j] * pym[j]) - rx * (kxn * pxp + C.kx[j] * pxm) -
               rz * (N.kz[j] * pn[j] + C.kz[j] * pm[j]);
        unew[j] = C.u[j] + alpha_u * C.po[j];
        if (j < nn) acc = acc + v[j] * pc[j];
      }
      stv<double, V, true>(c.w.at(x, y, z), nn, v);
      stv<double, V, true>(c.pnew.at(x, y, z), nn, pc);
      if (AXPY) stv<double, V, true>(c.u.at(x, y, z), nn, unew);
    }
#pragma unroll
    for (int j = 0; j < V; j++) {
      pm[j] = pc[j];
      pc[j] = pn[j];
    }
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
        P[i].po[j] = P[i].rr[j] = P[i].kz[j] = P[i].kx[j] = P[i].ky[j] = P[i].u[j] = P[i].yl_po[j] = P[i].yl_r[j] =
            P[i].yh_po[j] = P[i].yh_r[j] = P[i].yh_ky[j] = 0.0;
      P[i].xl_po = P[i].xl_r = P[i].xh_po = P[i].xh_r = P[i].xh_kx = 0.0;
    }
#pragma unroll
    for (int j = 0; j < V; j++) pm[j] = pc[j] = 0.0;
    if (act()) {
      // p_new of plane z0 - 1 (inside R when z0 > R.z0)
      const int qm = z0 - 1;
      double po[V], rv[V];
      ldv<double, V>(c.pold.at(x, y, qm), n(), po);
      ldv<double, V>(c.r.at(x, y, max(qm, R.z0)), n(), rv);
      load(z0, P[0]);
      load(min(z0 + 1, z1), P[1]);
#pragma unroll
      for (int j = 0; j < V; j++) {
        pm[j] = pnew(po[j], rv[j], qm >= R.z0);
        pc[j] = pnew(P[0].po[j], P[0].rr[j], true);
      }
    }
    int z = z0;
    for (; z + RING <= z1; z += RING) steps_all<0>(z);
    steps_tail<0>(z);
    return acc;
  }
};

template <int V, bool AXPY, int WY = kWavesPerBlock>
__device__ __forceinline__ double heat3_chain_column_v(HeatLdsV<V, WY> &L, const ChainPtrs &c, const Range3 &R,
                                                       double alpha_u, double a, double b, double rx, double ry,
                                                       double rz, int x0, int y0, int z0, int z1) {
  if (x0 + kWave * V <= R.x1 && y0 + WY <= R.y1)
    return ChainColumn<V, AXPY, WY, true>(L, c, R, alpha_u, a, b, rx, ry, rz, x0, y0, z0, z1).run();
  return ChainColumn<V, AXPY, WY, false>(L, c, R, alpha_u, a, b, rx, ry, rz, x0, y0, z0, z1).run();
}

// ------------------------------------------------------------ Poisson (7-point, constant coefficient)
template <class T, int V, int WY = kWavesPerBlock>
struct PoisLdsV {
  __attribute__((aligned(16))) T sp[2][WY + 2][kWave * V + 2 * V];
};

// one plane of a Poisson column as a lane loads it: its V points of p and
// the rims it owns
template <class T, int V>
struct PoisPlane {
  T p[V], ryl[V], ryh[V];
  T rxl, rxh;
};

// w = 6p - (p(-1,0,0) + p(1,0,0)) - (p(0,-1,0) + p(0,1,0)) - (p(0,0,-1) +
// p(0,0,1)) in T over one column tile; run() returns this thread's sum of
// (double) w (double) p.  Same structure as HeatColumn: PF planes of loads
// ahead in an unrolled register ring, the same loads every step (the rims of
// plane z1, never used, come from plane z1 - 1), and a branch-free form for
// tiles inside the range.
template <class T, int V, bool NT, int WY, int PF, bool FULL>
struct PoisColumn {
  static constexpr int RING = PF + 2;
  PoisLdsV<T, V, WY> &L;
  const Ptr<T> &w;
  const Ptr<const T> &p;
  int x0, z0, z1;
  ColLane c;
  int x, y, o, lx, ly;
  double acc;
  T pm[V];
  PoisPlane<T, V> R[RING];

  __device__ __forceinline__ PoisColumn(PoisLdsV<T, V, WY> &L_, const Ptr<T> &w_, const Ptr<const T> &p_, int x0_,
                                        int y0, int xe, int ye, int z0_, int z1_)
      : L(L_), w(w_), p(p_), x0(x0_), z0(z0_), z1(z1_) {
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

  __device__ __forceinline__ void load(int q, PoisPlane<T, V> &P) const {
    const int qk = min(q, z1 - 1);
    ldv<T, V>(p.at(x, y, q), n(), P.p);
    if (c.hxl) P.rxl = p(x0 - 1, y, qk);
    if (c.hxh) P.rxh = p(x0 + c.cnt, y, qk);
    if (c.hyl) ldv<T, V>(p.at(x, y - 1, qk), n(), P.ryl);
    if (c.hyh) ldv<T, V>(p.at(x, y + 1, qk), n(), P.ryh);
  }

  template <int K>
  __device__ __forceinline__ void step(int z) {
    const PoisPlane<T, V> &C = R[K % RING], &N = R[(K + 1) % RING];
    const int b = (z - z0) & 1, nn = n();
    if (act()) {
      lds_st<T, V>(&L.sp[b][ly + 1][o], nn, C.p);
      if (c.hxl) L.sp[b][ly + 1][V - 1] = C.rxl;
      if (c.hxh) L.sp[b][ly + 1][V + c.cnt] = C.rxh;
      if (c.hyl) lds_st<T, V>(&L.sp[b][0][o], nn, C.ryl);
      if (c.hyh) lds_st<T, V>(&L.sp[b][ly + 2][o], nn, C.ryh);
      load(min(z + PF + 1, z1), R[(K + PF + 1) % RING]);
    }
    __syncthreads();
    if (act()) {
      T pym[V], pyp[V], v[V];
      lds_ld<T, V>(&L.sp[b][ly][o], pym);
      lds_ld<T, V>(&L.sp[b][ly + 2][o], pyp);
      const T left = L.sp[b][ly + 1][o - 1];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const T pxm = j == 0 ? left : C.p[j == 0 ? 0 : j - 1];
        const T pxp = j + 1 < nn ? C.p[j + 1 < V ? j + 1 : j] : L.sp[b][ly + 1][o + j + 1];
        v[j] = (T)6.0 * C.p[j] - (pxm + pxp) - (pym[j] + pyp[j]) - (pm[j] + N.p[j]);
        if (j < nn) acc = acc + (double)v[j] * (double)C.p[j];
      }
      stv<T, V, NT>(w.at(x, y, z), nn, v);
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
      for (int j = 0; j < V; j++) R[i].p[j] = R[i].ryl[j] = R[i].ryh[j] = (T)0;
      R[i].rxl = R[i].rxh = (T)0;
    }
#pragma unroll
    for (int j = 0; j < V; j++) pm[j] = (T)0;
    if (act()) {
      ldv<T, V>(p.at(x, y, z0 - 1), n(), pm);
#pragma unroll
      for (int i = 0; i < PF + 1; i++) load(min(z0 + i, z1), R[i]);
    }
    int z = z0;
    for (; z + RING <= z1; z += RING) steps_all<0>(z);
    steps_tail<0>(z);
    return acc;
  }
};

template <class T, int V, bool NT, int WY = kWavesPerBlock, int PF = 1>
__device__ __forceinline__ double poisson3_column_v(PoisLdsV<T, V, WY> &L, const Ptr<T> &w, const Ptr<const T> &p,
                                                    int x0, int y0, int xe, int ye, int z0, int z1) {
  if (x0 + kWave * V <= xe && y0 + WY <= ye)
    return PoisColumn<T, V, NT, WY, PF, true>(L, w, p, x0, y0, xe, ye, z0, z1).run();
  return PoisColumn<T, V, NT, WY, PF, false>(L, w, p, x0, y0, xe, ye, z0, z1).run();
}

// ------------------------------------------------------------ fused Poisson CG chains
// The constant-coefficient operator's CG loops between two results, as one
// column pass over the calc_w column's tiles (so the sums of w p are the
// same bits):
//   MODE 0 (FP64 CG, tea_leaf kernels):
//     [AXPY] u = u + alpha_u p_old ; p_new = a p_old + b r ; w = A p_new
//   MODE 1 (FP32 inner iterations of the mixed solver):
//     [AXPY] e = e + alpha_u p_old ; p_new = mi rr + b p_old ; w = A p_new
// with the separate loops' operations in their order; p_new is p_old outside
// the loop range R.  Same ring structure as ChainColumn.
template <class T>
struct PoisChainPtrs {
  Ptr<T> u, pnew, w;
  Ptr<const T> pold, r;
};

template <class T, int V>
struct PoisChainPlane {
  T po[V], rr[V], u[V];
  T yl_po[V], yl_r[V], yh_po[V], yh_r[V];
  T xl_po, xl_r, xh_po, xh_r;
};

template <class T, int V, bool AXPY, int MODE, int WY, bool FULL>
struct PoisChainColumn {
  static constexpr int RING = 3;
  PoisLdsV<T, V, WY> &L;
  Ptr<T> u, pnew, w;
  Ptr<const T> pold, r;
  const Range3 &R;
  T alpha_u, a, b;
  int x0, z0, z1;
  ColLane cl;
  int x, y, o, lx, ly;
  bool in_xl, in_xh, in_yl, in_yh;
  int xl_r, xh_r, yl_r, yh_r;
  double acc;
  T pm[V], pc[V];
  PoisChainPlane<T, V> P[RING];

  __device__ __forceinline__ PoisChainColumn(PoisLdsV<T, V, WY> &L_, const PoisChainPtrs<T> &c, const Range3 &R_,
                                             T alpha_u_, T a_, T b_, int x0_, int y0, int z0_, int z1_)
      : L(L_), u(c.u), pnew(c.pnew), w(c.w), pold(c.pold), r(c.r), R(R_), alpha_u(alpha_u_), a(a_), b(b_), x0(x0_),
        z0(z0_), z1(z1_) {
    lx = (int)threadIdx.x;
    ly = (int)threadIdx.y;
    cl = col_lane<V, WY>(x0, y0, R.x1, R.y1);
    if (FULL) cl.n = V;
    x = x0 + V * lx;
    y = y0 + ly;
    o = V + V * lx;
    in_xl = x0 - 1 >= R.x0;
    in_xh = x0 + cl.cnt < R.x1;
    in_yl = y - 1 >= R.y0;
    in_yh = y + 1 < R.y1;
    xl_r = in_xl ? x0 - 1 : x0;
    xh_r = in_xh ? x0 + cl.cnt : x0 + cl.cnt - 1;
    yl_r = in_yl ? y - 1 : y;
    yh_r = in_yh ? y + 1 : y;
    acc = 0.0;
  }
  __device__ __forceinline__ int n() const { return FULL ? V : cl.n; }
  __device__ __forceinline__ bool act() const { return FULL || cl.n > 0; }
  __device__ __forceinline__ T pnewv(T po, T rv, bool in) const {
    if (!in) return po;
    if constexpr (MODE == 0) return a * po + b * rv;
    else return a * rv + b * po;  // a = mi
  }

  __device__ __forceinline__ void load(int q, PoisChainPlane<T, V> &Pl) const {
    const int qk = min(q, z1 - 1);
    const int qr = min(max(q, R.z0), R.z1 - 1);
    const int nn = n();
    ldv<T, V>(pold.at(x, y, q), nn, Pl.po);
    ldv<T, V>(r.at(x, y, qr), nn, Pl.rr);
    if (AXPY) ldv<T, V, true>(u.at(x, y, qk), nn, Pl.u);
    if (cl.hxl) {
      Pl.xl_po = pold(x0 - 1, y, qk);
      Pl.xl_r = r(xl_r, y, qk);
    }
    if (cl.hxh) {
      Pl.xh_po = pold(x0 + cl.cnt, y, qk);
      Pl.xh_r = r(xh_r, y, qk);
    }
    if (cl.hyl) {
      ldv<T, V>(pold.at(x, y - 1, qk), nn, Pl.yl_po);
      ldv<T, V>(r.at(x, yl_r, qk), nn, Pl.yl_r);
    }
    if (cl.hyh) {
      ldv<T, V>(pold.at(x, y + 1, qk), nn, Pl.yh_po);
      ldv<T, V>(r.at(x, yh_r, qk), nn, Pl.yh_r);
    }
  }

  template <int K>
  __device__ __forceinline__ void step(int z) {
    const PoisChainPlane<T, V> &C = P[K % RING], &N = P[(K + 1) % RING];
    const int bb = (z - z0) & 1, nn = n();
    T pn[V];
    if (act()) {
      lds_st<T, V>(&L.sp[bb][ly + 1][o], nn, pc);
      if (cl.hxl) L.sp[bb][ly + 1][V - 1] = pnewv(C.xl_po, C.xl_r, in_xl);
      if (cl.hxh) L.sp[bb][ly + 1][V + cl.cnt] = pnewv(C.xh_po, C.xh_r, in_xh);
      if (cl.hyl) {
        T t[V];
#pragma unroll
        for (int j = 0; j < V; j++) t[j] = pnewv(C.yl_po[j], C.yl_r[j], in_yl);
        lds_st<T, V>(&L.sp[bb][0][o], nn, t);
      }
      if (cl.hyh) {
        T t[V];
#pragma unroll
        for (int j = 0; j < V; j++) t[j] = pnewv(C.yh_po[j], C.yh_r[j], in_yh);
        lds_st<T, V>(&L.sp[bb][ly + 2][o], nn, t);
      }
      load(min(z + 2, z1), P[(K + 2) % RING]);
    }
    __syncthreads();
    const bool in_n = z + 1 < R.z1;
#pragma unroll
    for (int j = 0; j < V; j++) pn[j] = pnewv(N.po[j], N.rr[j], in_n);
    if (act()) {
      T pym[V], pyp[V], v[V], unew[V];
      lds_ld<T, V>(&L.sp[bb][ly][o], pym);
      lds_ld<T, V>(&L.sp[bb][ly + 2][o], pyp);
      const T left = L.sp[bb][ly + 1][o - 1];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const T pxm = j == 0 ? left : pc[j == 0 ? 0 : j - 1];
        const T pxp = j + 1 < nn ? pc[j + 1 < V ? j + 1 : j] : L.sp[bb][ly + 1][o + j + 1];
        v[j] = (T)6.0 * pc[j] - (pxm + pxp) - (pym[j] + pyp[j]) - (pm[j] + pn[j]);
        unew[j] = C.u[j] + alpha_u * C.po[j];
        if (j < nn) acc = acc + (double)v[j] * (double)pc[j];
      }
      stv<T, V, true>(w.at(x, y, z), nn, v);
      stv<T, V, true>(pnew.at(x, y, z), nn, pc);
      if (AXPY) stv<T, V, true>(u.at(x, y, z), nn, unew);
    }
#pragma unroll
    for (int j = 0; j < V; j++) {
      pm[j] = pc[j];
      pc[j] = pn[j];
    }
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
        P[i].po[j] = P[i].rr[j] = P[i].u[j] = P[i].yl_po[j] = P[i].yl_r[j] = P[i].yh_po[j] = P[i].yh_r[j] = (T)0;
      P[i].xl_po = P[i].xl_r = P[i].xh_po = P[i].xh_r = (T)0;
    }
#pragma unroll
    for (int j = 0; j < V; j++) pm[j] = pc[j] = (T)0;
    if (act()) {
      const int qm = z0 - 1;
      T po[V], rv[V];
      ldv<T, V>(pold.at(x, y, qm), n(), po);
      ldv<T, V>(r.at(x, y, max(qm, R.z0)), n(), rv);
      load(z0, P[0]);
      load(min(z0 + 1, z1), P[1]);
#pragma unroll
      for (int j = 0; j < V; j++) {
        pm[j] = pnewv(po[j], rv[j], qm >= R.z0);
        pc[j] = pnewv(P[0].po[j], P[0].rr[j], true);
      }
    }
    int z = z0;
    for (; z + RING <= z1; z += RING) steps_all<0>(z);
    steps_tail<0>(z);
    return acc;
  }
};

template <class T, int V, bool AXPY, int MODE, int WY>
__device__ __forceinline__ double poisson3_chain_column_v(PoisLdsV<T, V, WY> &L, const PoisChainPtrs<T> &c,
                                                          const Range3 &R, T alpha_u, T a, T b, int x0, int y0, int z0,
                                                          int z1) {
  if (x0 + kWave * V <= R.x1 && y0 + WY <= R.y1)
    return PoisChainColumn<T, V, AXPY, MODE, WY, true>(L, c, R, alpha_u, a, b, x0, y0, z0, z1).run();
  return PoisChainColumn<T, V, AXPY, MODE, WY, false>(L, c, R, alpha_u, a, b, x0, y0, z0, z1).run();
}

// ------------------------------------------------------------ 27-point (weights by distance)
// out = sum over dz, dy, dx in -1..1 (that order, dx fastest) of
// W[|dx| + |dy| + |dz|] u(x + dx, y + dy, z + dz), W = 1/8, 1/16, 1/32, 1/64.
// Every neighbour comes from LDS: a (128 + 2) x (WY + 2) plane with its
// corners per z, five plane buffers in rotation so one barrier per plane
// suffices (the buffer written at step z last held plane z - 3, read at step
// z - 2).  A lane reads each (dy, dz) row of its two points as one 16-byte
// and two 8-byte LDS loads; the plane three steps ahead is loaded into
// registers while the current one is computed.
constexpr double kW27[4] = {0.125, 0.0625, 0.03125, 0.015625};
constexpr int kS27Bufs = 5;

template <int WY>
struct S27Lds {
  __attribute__((aligned(16))) double sp[kS27Bufs][WY + 2][kWave * 2 + 4];
};

template <int WY, bool FULL>
struct S27Column {
  static constexpr int V = 2;
  S27Lds<WY> &L;
  const Ptr<double> &out;
  const Ptr<const double> &u;
  int x0, z0, z1;
  ColLane c;
  int x, y, o, lx, ly;
  // the plane in flight: own row, x rims, rim rows with their corners
  double p[V], rl[V], rh[V];
  double xl, xh, rl_xl, rl_xh, rh_xl, rh_xh;

  __device__ __forceinline__ S27Column(S27Lds<WY> &L_, const Ptr<double> &out_, const Ptr<const double> &u_, int x0_,
                                       int y0, int xe, int ye, int z0_, int z1_)
      : L(L_), out(out_), u(u_), x0(x0_), z0(z0_), z1(z1_) {
    lx = (int)threadIdx.x;
    ly = (int)threadIdx.y;
    c = col_lane<V, WY>(x0, y0, xe, ye);
    if (FULL) c.n = V;
    x = x0 + V * lx;
    y = y0 + ly;
    o = V + V * lx;
  }
  __device__ __forceinline__ int n() const { return FULL ? V : c.n; }
  __device__ __forceinline__ bool act() const { return FULL || c.n > 0; }
  __device__ __forceinline__ int buf(int q) const { return (q - z0 + kS27Bufs) % kS27Bufs; }

  __device__ __forceinline__ void load(int q) {
    const int nn = n();
    ldv<double, V>(u.at(x, y, q), nn, p);
    if (c.hxl) xl = u(x0 - 1, y, q);
    if (c.hxh) xh = u(x0 + c.cnt, y, q);
    if (c.hyl) {
      ldv<double, V>(u.at(x, y - 1, q), nn, rl);
      if (c.hxl) rl_xl = u(x0 - 1, y - 1, q);
      if (c.hxh) rl_xh = u(x0 + c.cnt, y - 1, q);
    }
    if (c.hyh) {
      ldv<double, V>(u.at(x, y + 1, q), nn, rh);
      if (c.hxl) rh_xl = u(x0 - 1, y + 1, q);
      if (c.hxh) rh_xh = u(x0 + c.cnt, y + 1, q);
    }
  }
  template <int B>
  __device__ __forceinline__ void stage_b() {
    const int nn = n();
    lds_st<double, V>(&L.sp[B][ly + 1][o], nn, p);
    if (c.hxl) L.sp[B][ly + 1][V - 1] = xl;
    if (c.hxh) L.sp[B][ly + 1][V + c.cnt] = xh;
    if (c.hyl) {
      lds_st<double, V>(&L.sp[B][0][o], nn, rl);
      if (c.hxl) L.sp[B][0][V - 1] = rl_xl;
      if (c.hxh) L.sp[B][0][V + c.cnt] = rl_xh;
    }
    if (c.hyh) {
      lds_st<double, V>(&L.sp[B][ly + 2][o], nn, rh);
      if (c.hxl) L.sp[B][ly + 2][V - 1] = rh_xl;
      if (c.hxh) L.sp[B][ly + 2][V + c.cnt] = rh_xh;
    }
  }
  __device__ __forceinline__ void stage(int q) {
    switch (buf(q)) {
      case 0: stage_b<0>(); break;
      case 1: stage_b<1>(); break;
      case 2: stage_b<2>(); break;
      case 3: stage_b<3>(); break;
      default: stage_b<4>(); break;
    }
  }

  // plane z = z0 + s with s % kS27Bufs == K: planes z - 1, z, z + 1 sit in
  // buffers (K + 4) % 5, K, (K + 1) % 5 and plane z + 2 is staged into
  // (K + 2) % 5 (the z loop is unrolled by the ring length, so every buffer
  // index is a compile-time constant)
  template <int K>
  __device__ __forceinline__ void step(int z) {
    if (act()) {
      if (z + 2 <= z1) {
        stage_b<(K + 2) % kS27Bufs>();
        load(min(z + 3, z1));
      }
    }
    __syncthreads();
    if (act()) {
      double acc[V] = {0.0, 0.0};
#pragma unroll
      for (int dz = -1; dz <= 1; dz++) {
        const int b = (K + dz + kS27Bufs) % kS27Bufs;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
          const double *row = &L.sp[b][ly + 1 + dy][0];
          const vec_t<double, 2> mid = *reinterpret_cast<const vec_t<double, 2> *>(row + o);
          const double lft = row[o - 1], rgt = row[o + 2];
          const int ad = (dy < 0 ? -dy : dy) + (dz < 0 ? -dz : dz);
          acc[0] = acc[0] + kW27[ad + 1] * lft;
          acc[0] = acc[0] + kW27[ad] * mid[0];
          acc[0] = acc[0] + kW27[ad + 1] * mid[1];
          acc[1] = acc[1] + kW27[ad + 1] * mid[0];
          acc[1] = acc[1] + kW27[ad] * mid[1];
          acc[1] = acc[1] + kW27[ad + 1] * rgt;
        }
      }
      stv<double, V, true>(out.at(x, y, z), n(), acc);
    }
  }

  __device__ __forceinline__ void run() {
    p[0] = p[1] = rl[0] = rl[1] = rh[0] = rh[1] = 0.0;
    xl = xh = rl_xl = rl_xh = rh_xl = rh_xh = 0.0;
    if (act()) {
      // planes z0 - 1 .. z0 + 1 into LDS, z0 + 2 into registers
      for (int q = z0 - 1; q <= z0 + 1 && q <= z1; q++) {
        load(q);
        stage(q);
      }
      if (z0 + 2 <= z1) load(z0 + 2);
    }
    int z = z0;
    for (; z + kS27Bufs <= z1; z += kS27Bufs) {
      step<0>(z);
      step<1>(z + 1);
      step<2>(z + 2);
      step<3>(z + 3);
      step<4>(z + 4);
    }
    const int rem = z1 - z;
    if (rem > 0) step<0>(z);
    if (rem > 1) step<1>(z + 1);
    if (rem > 2) step<2>(z + 2);
    if (rem > 3) step<3>(z + 3);
  }
};

template <int WY>
__device__ __forceinline__ void stencil27_column(S27Lds<WY> &L, const Ptr<double> &out, const Ptr<const double> &u,
                                                 int x0, int y0, int xe, int ye, int z0, int z1) {
  if (x0 + kWave * 2 <= xe && y0 + WY <= ye) S27Column<WY, true>(L, out, u, x0, y0, xe, ye, z0, z1).run();
  else S27Column<WY, false>(L, out, u, x0, y0, xe, ye, z0, z1).run();
}

}  // namespace opsmi
