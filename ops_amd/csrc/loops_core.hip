// Hand-written ops_par_loop kernels: Laplace tutorial, TeaLeaf CG family,
// constant-coefficient Poisson, ADI preproc, 27-point sweep and the small
// reduction loops.  Each functor restates the arithmetic of the reference
// kernel it names, in the same association order (compiled with FP
// contraction off), so results match the reference sequential backend bit
// for bit; sums differ only by reduction order.
#include "column.h"
#include "engine.h"
#include "heat_tile.h"

using namespace opsmi;

namespace {

// ------------------------------------------------------------ Laplace family
// apps/c/laplace2d_tutorial/step7/laplace_kernels.h:4-24
template <class T>
struct SetZero {
  OPSMI_NO_RED
  Ptr<T> a;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const { a(x, y, z) = T(0.0); }
};
template <class T>
struct Copy {
  OPSMI_NO_RED
  static constexpr int kVec = 16 / (int)sizeof(T);
  Ptr<T> a;
  Ptr<const T> b;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const { a(x, y, z) = b(x, y, z); }
  template <int V, bool NTL = false>
  __device__ void vop(int x, int y, int z, int n, double (&)[1]) const {
    T v[V];
    ldv<T, V, NTL>(b.at(x, y, z), n, v);
    stv<T, V, true>(a.at(x, y, z), n, v);
  }
};
// mode 0 left: sin(pi (y+1)/(jmax+1)); 1 right: same * exp(-pi);
// 2 top: sin(pi (x+1)/(imax+1))
template <class T>
struct Bndcon {
  OPSMI_NO_RED
  Ptr<T> a;
  double pi;
  int n;
  int mode;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    double v;
    if (mode == 2) v = sin(pi * (x + 1) / (n + 1));
    else {
      v = sin(pi * (y + 1) / (n + 1));
      if (mode == 1) v = v * exp(-pi);
    }
    a(x, y, z) = (T)v;
  }
};
template <class T>
struct ApplyStencil {
  using red_t = T;
  static constexpr int red_op = OPS_MAX;
  static constexpr int red_dim = 1;
  Ptr<const T> A;
  Ptr<T> An;
  __device__ void operator()(int x, int y, int z, T (&err)[1]) const {
    const T v = (T)0.25f * (A(x + 1, y, z) + A(x - 1, y, z) + A(x, y - 1, z) + A(x, y + 1, z));
    An(x, y, z) = v;
    err[0] = RedOp<T, OPS_MAX>::combine(err[0], fabs(v - A(x, y, z)));
  }
};

// ------------------------------------------------------------ TeaLeaf family
// apps/c/TeaLeaf/tea_leaf_kernels.h:8-62, tea_leaf_cg_kernels.h:8-21,
// tea_leaf_common_kernels.h:12-25
struct InitZero2 {
  OPSMI_NO_RED
  Ptr<double> p, z2;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    p(x, y, z) = 0.0;
    z2(x, y, z) = 0.0;
  }
};
struct Recip {
  OPSMI_NO_RED
  Ptr<double> u;
  Ptr<const double> p;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const { u(x, y, z) = 1.0 / p(x, y, z); }
};
struct Dot {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 1;
  Ptr<const double> r, p;
  __device__ void operator()(int x, int y, int z, double (&s)[1]) const { s[0] = s[0] + r(x, y, z) * p(x, y, z); }
};
struct Norm2 {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 1;
  Ptr<const double> a;
  __device__ void operator()(int x, int y, int z, double (&s)[1]) const {
    s[0] = s[0] + a(x, y, z) * a(x, y, z);
  }
};
struct Axpy {
  OPSMI_NO_RED
  static constexpr int kVec = 2;
  Ptr<double> u;
  Ptr<const double> p;
  double alpha;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const { u(x, y, z) = u(x, y, z) + alpha * p(x, y, z); }
  template <int V, bool NTL = false>
  __device__ void vop(int x, int y, int z, int n, double (&)[1]) const {
    double a[V], b[V];
    ldv<double, V, NTL>(u.at(x, y, z), n, a);
    ldv<double, V, NTL>(p.at(x, y, z), n, b);
#pragma unroll
    for (int j = 0; j < V; j++) a[j] = a[j] + alpha * b[j];
    stv<double, V, true>(u.at(x, y, z), n, a);
  }
};
struct Axpby {
  OPSMI_NO_RED
  static constexpr int kVec = 2;
  Ptr<double> u;
  Ptr<const double> p;
  double alpha, beta;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    u(x, y, z) = alpha * u(x, y, z) + beta * p(x, y, z);
  }
  template <int V, bool NTL = false>
  __device__ void vop(int x, int y, int z, int n, double (&)[1]) const {
    double a[V], b[V];
    ldv<double, V, NTL>(u.at(x, y, z), n, a);
    ldv<double, V, NTL>(p.at(x, y, z), n, b);
#pragma unroll
    for (int j = 0; j < V; j++) a[j] = alpha * a[j] + beta * b[j];
    stv<double, V, true>(u.at(x, y, z), n, a);
  }
};
struct Zeqxty {
  OPSMI_NO_RED
  Ptr<double> zz;
  Ptr<const double> a, b;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const { zz(x, y, z) = a(x, y, z) * b(x, y, z); }
};
template <int ND>
struct InitK {
  OPSMI_NO_RED
  Ptr<double> Kx, Ky, Kz;
  Ptr<const double> w;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    const double c = w(x, y, z);
    Kx(x, y, z) = (w(x - 1, y, z) + c) / (2.0 * w(x - 1, y, z) * c);
    Ky(x, y, z) = (w(x, y - 1, z) + c) / (2.0 * w(x, y - 1, z) * c);
    if constexpr (ND == 3) Kz(x, y, z) = (w(x, y, z - 1) + c) / (2.0 * w(x, y, z - 1) * c);
  }
};
// the TeaLeaf operator; 3D appends the z pair after the x pair
template <int ND>
__device__ __forceinline__ double heat_op(const Ptr<const double> &Kx, const Ptr<const double> &Ky,
                                          const Ptr<const double> &Kz, const Ptr<const double> &p, double rx,
                                          double ry, double rz, int x, int y, int z) {
  if constexpr (ND == 2) {
    return (1.0 + ry * (Ky(x, y + 1, z) + Ky(x, y, z)) + rx * (Kx(x + 1, y, z) + Kx(x, y, z))) * p(x, y, z) -
           ry * (Ky(x, y + 1, z) * p(x, y + 1, z) + Ky(x, y, z) * p(x, y - 1, z)) -
           rx * (Kx(x + 1, y, z) * p(x + 1, y, z) + Kx(x, y, z) * p(x - 1, y, z));
  } else {
    return (1.0 + ry * (Ky(x, y + 1, z) + Ky(x, y, z)) + rx * (Kx(x + 1, y, z) + Kx(x, y, z)) +
            rz * (Kz(x, y, z + 1) + Kz(x, y, z))) * p(x, y, z) -
           ry * (Ky(x, y + 1, z) * p(x, y + 1, z) + Ky(x, y, z) * p(x, y - 1, z)) -
           rx * (Kx(x + 1, y, z) * p(x + 1, y, z) + Kx(x, y, z) * p(x - 1, y, z)) -
           rz * (Kz(x, y, z + 1) * p(x, y, z + 1) + Kz(x, y, z) * p(x, y, z - 1));
  }
}
template <int ND>
struct CommonInit {
  OPSMI_NO_RED
  Ptr<double> w, r;
  Ptr<const double> Kx, Ky, Kz, u;
  double rx, ry, rz;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    const double v = heat_op<ND>(Kx, Ky, Kz, u, rx, ry, rz, x, y, z);
    w(x, y, z) = v;
    r(x, y, z) = u(x, y, z) - v;
  }
};
template <int ND>
struct CalcW {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 1;
  Ptr<double> w;
  Ptr<const double> Kx, Ky, Kz, p;
  double rx, ry, rz;
  __device__ void operator()(int x, int y, int z, double (&s)[1]) const {
    const double v = heat_op<ND>(Kx, Ky, Kz, p, rx, ry, rz, x, y, z);
    w(x, y, z) = v;
    s[0] = s[0] + v * p(x, y, z);
  }
};
// 3D TeaLeaf operator, z-marching form (heat3_column): a 64 x 4 workgroup
// owns a 64 x 4 column of the grid over `chunk` z planes.
// INIT = tea_leaf_common_init_kernel (w = A u, r = u - w); otherwise
// tea_leaf_cg_calc_w_reduce_kernel (w = A p, pw += w p).
template <bool INIT>
__global__ __launch_bounds__(256) void k_heat3_march(Ptr<double> w, Ptr<double> r, Ptr<const double> Kx,
                                                     Ptr<const double> Ky, Ptr<const double> Kz,
                                                     Ptr<const double> p, double rx, double ry, double rz, Range3 rg,
                                                     int chunk, RedCtl ctl) {
  __shared__ HeatLds lds;
  const TileId t = tile_id();
  const int z0 = rg.z0 + (int)t.z * chunk;
  double acc[1] = {heat3_column<INIT>(lds, w, r, Kx, Ky, Kz, p, rx, ry, rz, rg.x0 + (int)t.x * kWave,
                                      rg.y0 + (int)t.y * kWavesPerBlock, rg.x1, rg.y1, z0, min(z0 + chunk, rg.z1))};
  if constexpr (!INIT) reduce_block_store<double, OPS_INC, 1>(acc, ctl);
}

// Same with two points per lane (16-byte loads), 128 x kHeatRows columns,
// loads two planes ahead and non-temporal stores of w (and r):
// heat3_column_v.
template <bool INIT>
__global__ __launch_bounds__(kWave * kHeatRows) void k_heat3_march_v(Ptr<double> w, Ptr<double> r,
                                                                     Ptr<const double> Kx, Ptr<const double> Ky,
                                                                     Ptr<const double> Kz, Ptr<const double> p,
                                                                     double rx, double ry, double rz, Range3 rg,
                                                                     int chunk, RedCtl ctl) {
  __shared__ HeatLdsV<2, kHeatRows> lds;
  const TileId t = tile_id();
  const int z0 = rg.z0 + (int)t.z * chunk;
  double acc[1] = {heat3_column_v<2, INIT, true, kHeatRows, 2>(
      lds, w, r, Kx, Ky, Kz, p, rx, ry, rz, rg.x0 + (int)t.x * 2 * kWave, rg.y0 + (int)t.y * kHeatRows, rg.x1, rg.y1,
      z0, min(z0 + chunk, rg.z1))};
  if constexpr (!INIT) reduce_block_store<double, OPS_INC, 1>(acc, ctl);
}

// Queue record for the z-marching 3D heat kernel (16 z planes per workgroup
// by default, measured best among 1-16 on 512^3; OPSMI_HEAT_ZCHUNK
// overrides).  INIT args: w, r, Kx, Ky, Kz, u; otherwise w, Kx, Ky, Kz, p.
// The two-points-per-lane form runs when every array allows 16-byte loads
// at the range's x start (OPSMI_VEC=0 keeps one point per lane).
inline int heat_zchunk() { return heat_zchunk_default(); }

template <bool INIT>
LoopRec make_heat3_loop(const char *name, const ops_arg *args, int nargs, const Range3 &r, ops_arg w, ops_arg rr,
                        ops_arg Kx, ops_arg Ky, ops_arg Kz, ops_arg p, double rx, double ry, double rz,
                        const ops_arg *red) {
  LoopRec rec;
  rec.name = name;
  rec.range = r;
  rec.bytes = transfer_bytes(args, nargs, r);
  deps_from_args(rec, args, nargs);
  if (r.npts() <= 0) {
    rec.bytes = 0.0;
    rec.launch = [](hipStream_t) {};
    return rec;
  }
  bool vec = use_vec();
  for (ops_dat d : {w.dat, Kx.dat, Ky.dat, Kz.dat, p.dat}) vec = vec && vec16_ok(d, r.x0);
  if (INIT) vec = vec && vec16_ok(rr.dat, r.x0);
  ColGrid g;
  if (vec) {
    g = heat_column_grid(r, 2 * kWave, kHeatRows);
  } else {
    g.chunk = heat_zchunk();
    g.gx = (unsigned)((r.x1 - r.x0 + kWave - 1) / kWave);
    g.gy = (unsigned)((r.y1 - r.y0 + kWavesPerBlock - 1) / kWavesPerBlock);
    g.gz = (unsigned)((r.z1 - r.z0 + g.chunk - 1) / g.chunk);
  }
  RedCtl ctl{};
  if (!INIT) ctl = red_prepare(*red, 1, DT_F64, g.blocks());
  rec.ctl = ctl;
  rec.launch = [=](hipStream_t s) {
    const RedCtl c = INIT ? with_scratch(ctl) : with_host_fold(ctl, g.blocks(), 1, DT_F64);
    Ptr<double> rp = INIT ? ptr_of<double>(rr.dat) : Ptr<double>{nullptr, 0, 0};
    hipLaunchKernelGGL(vec ? k_heat3_march_v<INIT> : k_heat3_march<INIT>, dim3(g.gx, g.gy, g.gz),
                       dim3(kWave, vec ? kHeatRows : kWavesPerBlock), 0, s, ptr_of<double>(w.dat), rp,
                       ptr_of<const double>(Kx.dat), ptr_of<const double>(Ky.dat), ptr_of<const double>(Kz.dat),
                       ptr_of<const double>(p.dat), rx, ry, rz, r, g.chunk, c);
    if (!INIT) launch_red_finish<double, OPS_INC, 1>(c, g.blocks(), s);
  };
  return rec;
}

bool use_march() {
  const char *e = getenv("OPSMI_HEAT_MARCH");
  return !(e && e[0] == '0');
}

struct CalcUrR {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 1;
  static constexpr int kVec = 2;
  Ptr<double> r;
  Ptr<const double> w;
  double alpha;
  __device__ void operator()(int x, int y, int z, double (&s)[1]) const {
    const double v = r(x, y, z) - alpha * w(x, y, z);
    r(x, y, z) = v;
    s[0] = s[0] + v * v;
  }
  template <int V, bool NTL = false>
  __device__ void vop(int x, int y, int z, int n, double (&s)[1]) const {
    double a[V], b[V];
    ldv<double, V, NTL>(r.at(x, y, z), n, a);
    ldv<double, V, NTL>(w.at(x, y, z), n, b);
#pragma unroll
    for (int j = 0; j < V; j++) {
      a[j] = a[j] - alpha * b[j];
      if (j < n) s[0] = s[0] + a[j] * a[j];
    }
    stv<double, V, true>(r.at(x, y, z), n, a);
  }
};

// ------------------------------------------------------------ Poisson
template <class T, int ND>
struct PoissonW {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 1;
  Ptr<T> w;
  Ptr<const T> p;
  __device__ void operator()(int x, int y, int z, double (&s)[1]) const {
    T v;
    if constexpr (ND == 2)
      v = (T)4.0 * p(x, y, z) - (p(x - 1, y, z) + p(x + 1, y, z)) - (p(x, y - 1, z) + p(x, y + 1, z));
    else
      v = (T)6.0 * p(x, y, z) - (p(x - 1, y, z) + p(x + 1, y, z)) - (p(x, y - 1, z) + p(x, y + 1, z)) -
          (p(x, y, z - 1) + p(x, y, z + 1));
    w(x, y, z) = v;
    s[0] = s[0] + (double)v * (double)p(x, y, z);
  }
};

// 3D form as z-marching columns with 16 bytes per lane (2 doubles or 4
// floats), (64 V) x kHeatRows tiles, p planes in LDS: poisson3_column_v, same
// arithmetic.  Planes of loads in flight: at 768^3 FP64 1464 / 1478 / 1476 /
// 1432 us and FP32 759 / 769 / 782 / 767 us for 1 / 2 / 3 / 4.
template <class T>
constexpr int pois_pf() {
  return sizeof(T) == 8 ? 4 : 1;
}
template <class T, int kPoisPF>
__global__ __launch_bounds__(kWave * kHeatRows) void k_poisson3_march(Ptr<T> w, Ptr<const T> p, Range3 rg, int chunk,
                                                                      RedCtl ctl) {
  constexpr int V = 16 / (int)sizeof(T);
  __shared__ PoisLdsV<T, V, kHeatRows> lds;
  const TileId t = tile_id();
  const int z0 = rg.z0 + (int)t.z * chunk;
  double acc[1] = {poisson3_column_v<T, V, true, kHeatRows, kPoisPF>(lds, w, p, rg.x0 + (int)t.x * kWave * V,
                                                                     rg.y0 + (int)t.y * kHeatRows, rg.x1, rg.y1, z0,
                                                                     min(z0 + chunk, rg.z1))};
  reduce_block_store<double, OPS_INC, 1>(acc, ctl);
}

template <class T>
LoopRec make_poisson3_loop(const char *name, const ops_arg *args, int nargs, const Range3 &r, ops_arg w, ops_arg p,
                           const ops_arg *red) {
  LoopRec rec;
  rec.name = name;
  rec.range = r;
  rec.bytes = transfer_bytes(args, nargs, r);
  deps_from_args(rec, args, nargs);
  if (r.npts() <= 0) {
    rec.bytes = 0.0;
    rec.launch = [](hipStream_t) {};
    return rec;
  }
  constexpr int V = 16 / (int)sizeof(T);
  const ColGrid g = heat_column_grid(r, kWave * V, kHeatRows);
  const RedCtl ctl = red_prepare(*red, 1, DT_F64, g.blocks());
  rec.ctl = ctl;
  rec.launch = [=](hipStream_t s) {
    const RedCtl c = with_host_fold(ctl, g.blocks(), 1, DT_F64);
    hipLaunchKernelGGL((k_poisson3_march<T, pois_pf<T>()>), dim3(g.gx, g.gy, g.gz), dim3(kWave, kHeatRows), 0, s,
                       ptr_of<T>(w.dat), ptr_of<const T>(p.dat), r, g.chunk, c);
    launch_red_finish<double, OPS_INC, 1>(c, g.blocks(), s);
  };
  return rec;
}

// ------------------------------------------------------------ ADI preproc
// apps/c/adi/preproc_kernel.h:4-39
struct Preproc {
  OPSMI_NO_RED
  Ptr<const double> u;
  Ptr<double> du, ax, bx, cx, ay, by, cy, az, bz, cz;
  int nx, ny, nz;
  double lambda;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    double a, b, c, d;
    if (x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1) {
      d = 0.0f;
      a = 0.0f;
      b = 1.0f;
      c = 0.0f;
    } else {
      d = lambda * (u(x - 1, y, z) + u(x + 1, y, z) + u(x, y - 1, z) + u(x, y + 1, z) + u(x, y, z - 1) +
                    u(x, y, z + 1) - 6.0f * u(x, y, z));
      a = -0.5f * lambda;
      b = 1.0f + lambda;
      c = -0.5f * lambda;
    }
    du(x, y, z) = d;
    ax(x, y, z) = a; bx(x, y, z) = b; cx(x, y, z) = c;
    ay(x, y, z) = a; by(x, y, z) = b; cy(x, y, z) = c;
    az(x, y, z) = a; bz(x, y, z) = b; cz(x, y, z) = c;
  }
};

struct Stencil27 {
  OPSMI_NO_RED
  Ptr<double> out;
  Ptr<const double> u;
  __device__ void operator()(int x, int y, int z, double (&)[1]) const {
    const double W[4] = {0.125, 0.0625, 0.03125, 0.015625};
    double acc = 0.0;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
      for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++)
          acc = acc + W[(dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) + (dz < 0 ? -dz : dz)] * u(x + dx, y + dy, z + dz);
    out(x, y, z) = acc;
  }
};

// 3D form as an LDS z-column (stencil27_column): 128 x kHeatRows tiles, the
// same sum in the same order.
__global__ __launch_bounds__(kWave * kHeatRows) void k_stencil27_march(Ptr<double> out, Ptr<const double> u,
                                                                       Range3 rg, int chunk) {
  __shared__ S27Lds<kHeatRows> lds;
  const TileId t = tile_id();
  const int z0 = rg.z0 + (int)t.z * chunk;
  stencil27_column<kHeatRows>(lds, out, u, rg.x0 + (int)t.x * 2 * kWave, rg.y0 + (int)t.y * kHeatRows, rg.x1, rg.y1,
                              z0, min(z0 + chunk, rg.z1));
}

LoopRec make_stencil27_loop(const char *name, const ops_arg *args, int nargs, const Range3 &r, ops_arg out,
                            ops_arg u) {
  LoopRec rec;
  rec.name = name;
  rec.range = r;
  rec.bytes = transfer_bytes(args, nargs, r);
  deps_from_args(rec, args, nargs);
  if (r.npts() <= 0) {
    rec.bytes = 0.0;
    rec.launch = [](hipStream_t) {};
    return rec;
  }
  const ColGrid g = heat_column_grid(r, 2 * kWave, kHeatRows);
  rec.launch = [=](hipStream_t s) {
    hipLaunchKernelGGL(k_stencil27_march, dim3(g.gx, g.gy, g.gz), dim3(kWave, kHeatRows), 0, s,
                       ptr_of<double>(out.dat), ptr_of<const double>(u.dat), r, g.chunk);
  };
  return rec;
}

// ------------------------------------------------------------ small reductions
template <int KIND, int OP>
struct RedX {
  using red_t = double;
  static constexpr int red_op = OP;
  static constexpr int red_dim = 1;
  Ptr<const double> a;
  double scale;
  __device__ void operator()(int x, int y, int z, double (&s)[1]) const {
    const double v = a(x, y, z);
    if constexpr (KIND == 0) s[0] = s[0] + v;
    if constexpr (KIND == 1) s[0] = s[0] + v * v;
    if constexpr (KIND == 2) s[0] = s[0] + fabs(v);
    if constexpr (KIND == 3) s[0] = fmin(s[0], v);
    if constexpr (KIND == 4) s[0] = fmax(s[0], v);
    if constexpr (KIND == 5) s[0] = s[0] + scale * v;
  }
};
struct RedTwo {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 2;
  Ptr<const double> a;
  __device__ void operator()(int x, int y, int z, double (&s)[2]) const {
    s[0] = s[0] + a(x, y, z);
    s[1] = s[1] + 1.0;
  }
};
struct RedImax {
  using red_t = int;
  static constexpr int red_op = OPS_MAX;
  static constexpr int red_dim = 1;
  __device__ void operator()(int x, int y, int z, int (&s)[1]) const {
    const int v = x * y;
    s[0] = s[0] > v ? s[0] : v;
  }
};

}  // namespace

// small-reduction entry helper
template <int KIND, int OP>
static void red_x(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s, double scale,
                  const ops_arg *gb) {
  if (gb) {
    ops_arg args[] = {x, *gb, s};
    Range3 r = check_loop(name, block, dim, range, args,
                          {dat_r(pts_point(), DT_F64), gbl_r(), red_a(OP, DT_F64)}, nullptr);
    launch_loop(name, args, 3, r, [=] { return RedX<KIND, OP>{ptr_of<const double>(x.dat), scale}; }, &s);
  } else {
    ops_arg args[] = {x, s};
    Range3 r = check_loop(name, block, dim, range, args, {dat_r(pts_point(), DT_F64), red_a(OP, DT_F64)}, nullptr);
    launch_loop(name, args, 2, r, [=] { return RedX<KIND, OP>{ptr_of<const double>(x.dat), scale}; }, &s);
  }
}

// ============================================================ entry points
#define ARGS(...) \
  ops_arg args[] = {__VA_ARGS__}; \
  const int nargs = (int)(sizeof(args) / sizeof(args[0]));
#define PD(a) ptr_of<double>((a).dat)
#define PCD(a) ptr_of<const double>((a).dat)
#define PF(a) ptr_of<float>((a).dat)
#define PCF(a) ptr_of<const float>((a).dat)
#define LAUNCH(r, ...) launch_loop(name, args, nargs, r, [=] { return __VA_ARGS__; })
#define LAUNCH_RED(r, red, ...) launch_loop(name, args, nargs, r, [=] { return __VA_ARGS__; }, &(red))

static void need_dim(const char *name, int dim, int want) {
  if (dim != want) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": ", want, "D kernel on a ", dim, "D block"));
}

extern "C" {

void ops_par_loop_set_zero(const char *name, ops_block block, int dim, int *range, ops_arg a0) {
  OPSMI_API_BEGIN
  ARGS(a0)
  DType dt;
  Range3 r = check_loop(name, block, dim, range, args, {dat_w()}, &dt);
  if (dt == DT_F64) LAUNCH(r, SetZero<double>{PD(a0)});
  else LAUNCH(r, SetZero<float>{PF(a0)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_copy(const char *name, ops_block block, int dim, int *range, ops_arg a0, ops_arg a1) {
  OPSMI_API_BEGIN
  ARGS(a0, a1)
  DType dt;
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(), dat_r()}, &dt);
  LoopRec rec = dt == DT_F64 ? make_loop(name, args, nargs, r, [=] { return Copy<double>{PD(a0), PCD(a1)}; })
                             : make_loop(name, args, nargs, r, [=] { return Copy<float>{PF(a0), PCF(a1)}; });
  rec.kind = LK_COPY;
  rec.dims = dim;
  rec.d[0] = a0.dat;
  rec.d[1] = a1.dat;
  rec.g[0] = dt == DT_F64 ? 8 : 4;
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}

static void bndcon(const char *name, ops_block block, int dim, int *range, ops_arg a0, ops_arg a1, int mode) {
  ARGS(a0, a1)
  DType dt;
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(), idx_a()}, &dt);
  need_dim(name, dim, 2);
  const double pi = const_as_double("pi");
  const int n = const_as_int(mode == 2 ? "imax" : "jmax");
  if (dt == DT_F64) LAUNCH(r, Bndcon<double>{PD(a0), pi, n, mode});
  else LAUNCH(r, Bndcon<float>{PF(a0), pi, n, mode});
  finish_loop(args, nargs);
}

void ops_par_loop_left_bndcon(const char *name, ops_block block, int dim, int *range, ops_arg a0, ops_arg a1) {
  OPSMI_API_BEGIN
  bndcon(name, block, dim, range, a0, a1, 0);
  OPSMI_API_END()
}
void ops_par_loop_right_bndcon(const char *name, ops_block block, int dim, int *range, ops_arg a0, ops_arg a1) {
  OPSMI_API_BEGIN
  bndcon(name, block, dim, range, a0, a1, 1);
  OPSMI_API_END()
}
void ops_par_loop_top_bndcon(const char *name, ops_block block, int dim, int *range, ops_arg a0, ops_arg a1) {
  OPSMI_API_BEGIN
  bndcon(name, block, dim, range, a0, a1, 2);
  OPSMI_API_END()
}

void ops_par_loop_apply_stencil(const char *name, ops_block block, int dim, int *range, ops_arg a0, ops_arg a1,
                                ops_arg a2) {
  OPSMI_API_BEGIN
  ARGS(a0, a1, a2)
  DType dt;
  Range3 r = check_loop(name, block, dim, range, args, {dat_r(pts_5pt()), dat_w(), red_a(OPS_MAX)}, &dt);
  need_dim(name, dim, 2);
  LoopRec rec =
      dt == DT_F64
          ? make_loop(name, args, nargs, r, [=] { return ApplyStencil<double>{PCD(a0), PD(a1)}; }, &a2)
          : make_loop(name, args, nargs, r, [=] { return ApplyStencil<float>{PCF(a0), PF(a1)}; }, &a2);
  rec.kind = LK_JACOBI5;
  rec.dims = dim;
  rec.d[0] = a0.dat;
  rec.d[1] = a1.dat;
  rec.g[0] = dt == DT_F64 ? 8 : 4;
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}

// ---- TeaLeaf
void ops_par_loop_tea_leaf_init_zero_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p) {
  OPSMI_API_BEGIN
  ARGS(p)
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64)}, nullptr);
  LAUNCH(r, SetZero<double>{PD(p)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_init_zero2_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p,
                                             ops_arg z) {
  OPSMI_API_BEGIN
  ARGS(p, z)
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64), dat_w(DT_F64)}, nullptr);
  LAUNCH(r, InitZero2{PD(p), PD(z)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_yeqx_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p,
                                       ops_arg x) {
  OPSMI_API_BEGIN
  ARGS(p, x)
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64), dat_r(pts_point(), DT_F64)}, nullptr);
  LAUNCH(r, Copy<double>{PD(p), PCD(x)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_recip_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                        ops_arg p) {
  OPSMI_API_BEGIN
  ARGS(u, p)
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64), dat_r(pts_point(), DT_F64)}, nullptr);
  LAUNCH(r, Recip{PD(u), PCD(p)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_dot_kernel(const char *name, ops_block block, int dim, int *range, ops_arg rr,
                                      ops_arg p, ops_arg rro) {
  OPSMI_API_BEGIN
  ARGS(rr, p, rro)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_r(pts_point(), DT_F64), dat_r(pts_point(), DT_F64), red_a(OPS_INC, DT_F64)}, nullptr);
  LAUNCH_RED(r, rro, Dot{PCD(rr), PCD(p)});
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_norm2_kernel(const char *name, ops_block block, int dim, int *range, ops_arg x,
                                        ops_arg norm) {
  OPSMI_API_BEGIN
  ARGS(x, norm)
  Range3 r = check_loop(name, block, dim, range, args, {dat_r(pts_point(), DT_F64), red_a(OPS_INC, DT_F64)}, nullptr);
  LAUNCH_RED(r, norm, Norm2{PCD(x)});
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_axpy_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                       ops_arg p, ops_arg alpha) {
  OPSMI_API_BEGIN
  ARGS(u, p, alpha)
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64), dat_r(pts_point(), DT_F64), gbl_r()}, nullptr);
  const double a = gbl<double>(alpha);
  LoopRec rec = make_loop(name, args, nargs, r, [=] { return Axpy{PD(u), PCD(p), a}; });
  rec.kind = LK_AXPY;
  rec.dims = dim;
  rec.d[0] = u.dat;
  rec.d[1] = p.dat;
  rec.g[0] = a;
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_axpby_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                        ops_arg p, ops_arg alpha, ops_arg beta) {
  OPSMI_API_BEGIN
  ARGS(u, p, alpha, beta)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_r(pts_point(), DT_F64), gbl_r(), gbl_r()}, nullptr);
  const double a = gbl<double>(alpha), b = gbl<double>(beta);
  LoopRec rec = make_loop(name, args, nargs, r, [=] { return Axpby{PD(u), PCD(p), a, b}; });
  rec.kind = LK_AXPBY;
  rec.dims = dim;
  rec.d[0] = u.dat;
  rec.d[1] = p.dat;
  rec.g[0] = a;
  rec.g[1] = b;
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_zeqxty_kernel(const char *name, ops_block block, int dim, int *range, ops_arg z,
                                         ops_arg x, ops_arg y) {
  OPSMI_API_BEGIN
  ARGS(z, x, y)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_r(pts_point(), DT_F64), dat_r(pts_point(), DT_F64)}, nullptr);
  LAUNCH(r, Zeqxty{PD(z), PCD(x), PCD(y)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_common_init_Kx_Ky_kernel(const char *name, ops_block block, int dim, int *range,
                                                    ops_arg Kx, ops_arg Ky, ops_arg w) {
  OPSMI_API_BEGIN
  ARGS(Kx, Ky, w)
  std::vector<int> m = {0, 0, 0, -1, 0, 0, 0, -1, 0};
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64), dat_w(DT_F64), dat_r(m, DT_F64)}, nullptr);
  need_dim(name, dim, 2);
  LAUNCH(r, InitK<2>{PD(Kx), PD(Ky), Ptr<double>{nullptr, 0, 0}, PCD(w)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_common_init_Kx_Ky_Kz_kernel(const char *name, ops_block block, int dim, int *range,
                                                       ops_arg Kx, ops_arg Ky, ops_arg Kz, ops_arg w) {
  OPSMI_API_BEGIN
  ARGS(Kx, Ky, Kz, w)
  std::vector<int> m = {0, 0, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_w(DT_F64), dat_w(DT_F64), dat_r(m, DT_F64)}, nullptr);
  need_dim(name, dim, 3);
  LAUNCH(r, InitK<3>{PD(Kx), PD(Ky), PD(Kz), PCD(w)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}

static std::vector<int> kx_pts() { return {0, 0, 0, 1, 0, 0}; }
static std::vector<int> ky_pts() { return {0, 0, 0, 0, 1, 0}; }
static std::vector<int> kz_pts() { return {0, 0, 0, 0, 0, 1}; }

void ops_par_loop_tea_leaf_common_init_kernel(const char *name, ops_block block, int dim, int *range, ops_arg w,
                                              ops_arg rr, ops_arg Kx, ops_arg Ky, ops_arg u, ops_arg rx,
                                              ops_arg ry) {
  OPSMI_API_BEGIN
  ARGS(w, rr, Kx, Ky, u, rx, ry)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_w(DT_F64), dat_r(kx_pts(), DT_F64), dat_r(ky_pts(), DT_F64),
                         dat_r(pts_5pt(), DT_F64), gbl_r(), gbl_r()},
                        nullptr);
  need_dim(name, dim, 2);
  const double vx = gbl<double>(rx), vy = gbl<double>(ry);
  LAUNCH(r, CommonInit<2>{PD(w), PD(rr), PCD(Kx), PCD(Ky), Ptr<const double>{nullptr, 0, 0}, PCD(u), vx, vy, 0.0});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_common_init_kernel_3d(const char *name, ops_block block, int dim, int *range,
                                                 ops_arg w, ops_arg rr, ops_arg Kx, ops_arg Ky, ops_arg Kz,
                                                 ops_arg u, ops_arg rx, ops_arg ry, ops_arg rz) {
  OPSMI_API_BEGIN
  ARGS(w, rr, Kx, Ky, Kz, u, rx, ry, rz)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_w(DT_F64), dat_r(kx_pts(), DT_F64), dat_r(ky_pts(), DT_F64),
                         dat_r(kz_pts(), DT_F64), dat_r(pts_7pt(), DT_F64), gbl_r(), gbl_r(), gbl_r()},
                        nullptr);
  need_dim(name, dim, 3);
  const double vx = gbl<double>(rx), vy = gbl<double>(ry), vz = gbl<double>(rz);
  if (use_march())
    enqueue(make_heat3_loop<true>(name, args, nargs, r, w, rr, Kx, Ky, Kz, u, vx, vy, vz, nullptr));
  else
    LAUNCH(r, CommonInit<3>{PD(w), PD(rr), PCD(Kx), PCD(Ky), PCD(Kz), PCD(u), vx, vy, vz});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_cg_calc_w_reduce_kernel(const char *name, ops_block block, int dim, int *range,
                                                   ops_arg w, ops_arg Kx, ops_arg Ky, ops_arg p, ops_arg rx,
                                                   ops_arg ry, ops_arg pw) {
  OPSMI_API_BEGIN
  ARGS(w, Kx, Ky, p, rx, ry, pw)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_r(kx_pts(), DT_F64), dat_r(ky_pts(), DT_F64), dat_r(pts_5pt(), DT_F64),
                         gbl_r(), gbl_r(), red_a(OPS_INC, DT_F64)},
                        nullptr);
  need_dim(name, dim, 2);
  const double vx = gbl<double>(rx), vy = gbl<double>(ry);
  LAUNCH_RED(r, pw, CalcW<2>{PD(w), PCD(Kx), PCD(Ky), Ptr<const double>{nullptr, 0, 0}, PCD(p), vx, vy, 0.0});
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_cg_calc_w_reduce_kernel_3d(const char *name, ops_block block, int dim, int *range,
                                                      ops_arg w, ops_arg Kx, ops_arg Ky, ops_arg Kz, ops_arg p,
                                                      ops_arg rx, ops_arg ry, ops_arg rz, ops_arg pw) {
  OPSMI_API_BEGIN
  ARGS(w, Kx, Ky, Kz, p, rx, ry, rz, pw)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_r(kx_pts(), DT_F64), dat_r(ky_pts(), DT_F64), dat_r(kz_pts(), DT_F64),
                         dat_r(pts_7pt(), DT_F64), gbl_r(), gbl_r(), gbl_r(), red_a(OPS_INC, DT_F64)},
                        nullptr);
  need_dim(name, dim, 3);
  const double vx = gbl<double>(rx), vy = gbl<double>(ry), vz = gbl<double>(rz);
  LoopRec rec = use_march()
                    ? make_heat3_loop<false>(name, args, nargs, r, w, w, Kx, Ky, Kz, p, vx, vy, vz, &pw)
                    : make_loop(name, args, nargs, r,
                                [=] { return CalcW<3>{PD(w), PCD(Kx), PCD(Ky), PCD(Kz), PCD(p), vx, vy, vz}; }, &pw);
  rec.kind = LK_HEAT_W3;
  rec.dims = dim;
  rec.d[0] = w.dat;
  rec.d[1] = Kx.dat;
  rec.d[2] = Ky.dat;
  rec.d[3] = Kz.dat;
  rec.d[4] = p.dat;
  rec.g[0] = vx;
  rec.g[1] = vy;
  rec.g[2] = vz;
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}
void ops_par_loop_tea_leaf_cg_calc_ur_r_reduce_kernel(const char *name, ops_block block, int dim, int *range,
                                                      ops_arg rr, ops_arg w, ops_arg alpha, ops_arg rnn) {
  OPSMI_API_BEGIN
  ARGS(rr, w, alpha, rnn)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(DT_F64), dat_r(pts_point(), DT_F64), gbl_r(), red_a(OPS_INC, DT_F64)}, nullptr);
  const double a = gbl<double>(alpha);
  LAUNCH_RED(r, rnn, CalcUrR{PD(rr), PCD(w), a});
  finish_loop(args, nargs);
  OPSMI_API_END()
}

// ---- Poisson
void ops_par_loop_poisson_calc_w_reduce_kernel(const char *name, ops_block block, int dim, int *range, ops_arg w,
                                               ops_arg p, ops_arg pw) {
  OPSMI_API_BEGIN
  ARGS(w, p, pw)
  DType dt;
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_w(), dat_r(dim == 3 ? pts_7pt() : pts_5pt()), red_a(OPS_INC, DT_F64)}, &dt);
  if (dim < 2) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": needs a 2D or 3D block"));
  const bool col = dim == 3 && use_march() && use_vec() && vec16_ok(w.dat, r.x0) && vec16_ok(p.dat, r.x0);
  if (col) {
    LoopRec rec = dt == DT_F64 ? make_poisson3_loop<double>(name, args, nargs, r, w, p, &pw)
                               : make_poisson3_loop<float>(name, args, nargs, r, w, p, &pw);
    rec.kind = LK_POISSON_W3;
    rec.dims = 3;
    rec.d[0] = w.dat;
    rec.d[1] = p.dat;
    rec.g[0] = dt == DT_F64 ? 8 : 4;
    enqueue(std::move(rec));
  } else if (dt == DT_F64) {
    if (dim == 2) LAUNCH_RED(r, pw, PoissonW<double, 2>{PD(w), PCD(p)});
    else LAUNCH_RED(r, pw, PoissonW<double, 3>{PD(w), PCD(p)});
  } else {
    if (dim == 2) LAUNCH_RED(r, pw, PoissonW<float, 2>{PF(w), PCF(p)});
    else LAUNCH_RED(r, pw, PoissonW<float, 3>{PF(w), PCF(p)});
  }
  finish_loop(args, nargs);
  OPSMI_API_END()
}

// ---- ADI preproc, 27-point
void ops_par_loop_preproc_kernel(const char *name, ops_block block, int dim, int *range, ops_arg u, ops_arg du,
                                 ops_arg ax, ops_arg bx, ops_arg cx, ops_arg ay, ops_arg by, ops_arg cy,
                                 ops_arg az, ops_arg bz, ops_arg cz, ops_arg idx) {
  OPSMI_API_BEGIN
  ARGS(u, du, ax, bx, cx, ay, by, cy, az, bz, cz, idx)
  std::vector<ArgSpec> specs = {dat_r(pts_7pt(), DT_F64)};
  for (int i = 0; i < 10; i++) specs.push_back(dat_w(DT_F64));
  specs.push_back(idx_a());
  Range3 r = check_loop(name, block, dim, range, args, specs, nullptr);
  need_dim(name, dim, 3);
  const int nx = const_as_int("nx"), ny = const_as_int("ny"), nz = const_as_int("nz");
  const double lambda = const_as_double("lambda");
  LAUNCH(r, Preproc{PCD(u), PD(du), PD(ax), PD(bx), PD(cx), PD(ay), PD(by), PD(cy), PD(az), PD(bz), PD(cz), nx, ny,
                    nz, lambda});
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_stencil27_kernel(const char *name, ops_block block, int dim, int *range, ops_arg out, ops_arg u) {
  OPSMI_API_BEGIN
  ARGS(out, u)
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64), dat_r(pts_27pt(), DT_F64)}, nullptr);
  need_dim(name, dim, 3);
  if (use_march() && use_vec() && vec16_ok(out.dat, r.x0) && vec16_ok(u.dat, r.x0))
    enqueue(make_stencil27_loop(name, args, nargs, r, out, u));
  else
    LAUNCH(r, Stencil27{PD(out), PCD(u)});
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_red_sum(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s) {
  OPSMI_API_BEGIN
  red_x<0, OPS_INC>(name, block, dim, range, x, s, 0.0, nullptr);
  OPSMI_API_END()
}
void ops_par_loop_red_sumsq(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s) {
  OPSMI_API_BEGIN
  red_x<1, OPS_INC>(name, block, dim, range, x, s, 0.0, nullptr);
  OPSMI_API_END()
}
void ops_par_loop_red_abs(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s) {
  OPSMI_API_BEGIN
  red_x<2, OPS_INC>(name, block, dim, range, x, s, 0.0, nullptr);
  OPSMI_API_END()
}
void ops_par_loop_red_min(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg m) {
  OPSMI_API_BEGIN
  red_x<3, OPS_MIN>(name, block, dim, range, x, m, 0.0, nullptr);
  OPSMI_API_END()
}
void ops_par_loop_red_max(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg m) {
  OPSMI_API_BEGIN
  red_x<4, OPS_MAX>(name, block, dim, range, x, m, 0.0, nullptr);
  OPSMI_API_END()
}
void ops_par_loop_red_scaled(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg scale,
                             ops_arg s) {
  OPSMI_API_BEGIN
  if (scale.argtype != OPS_ARG_GBL || !scale.data || scale.elem_size != 8)
    fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": arg 1 must be an OPS_READ ops_arg_gbl of one double"));
  red_x<5, OPS_INC>(name, block, dim, range, x, s, gbl<double>(scale), &scale);
  OPSMI_API_END()
}
void ops_par_loop_red_two(const char *name, ops_block block, int dim, int *range, ops_arg x, ops_arg s2) {
  OPSMI_API_BEGIN
  ARGS(x, s2)
  Range3 r = check_loop(name, block, dim, range, args, {dat_r(pts_point(), DT_F64), red_a(OPS_INC, DT_F64, 2)}, nullptr);
  LAUNCH_RED(r, s2, RedTwo{PCD(x)});
  OPSMI_API_END()
}
void ops_par_loop_red_imax(const char *name, ops_block block, int dim, int *range, ops_arg idx, ops_arg m) {
  OPSMI_API_BEGIN
  ARGS(idx, m)
  Range3 r = check_loop(name, block, dim, range, args, {idx_a(), red_a(OPS_MAX, DT_I32)}, nullptr);
  LAUNCH_RED(r, m, RedImax{});
  OPSMI_API_END()
}

}  // extern "C"
