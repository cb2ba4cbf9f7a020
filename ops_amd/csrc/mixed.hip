// Mixed-precision Poisson loops: an FP64 iterative-refinement outer loop
// (residual, correction) around a Jacobi-preconditioned CG whose sweeps keep
// FP32 storage and FP32 arithmetic while every dot product accumulates the
// FP32 values in FP64.  The reference has no mixed-precision solver; the
// loops take the per-loop host stub form of the reference's generated code
// (ops_translator/resources/templates/cpp/loop_host.cpp.j2:12-21) and the
// constant-coefficient 5/7-point operator of the Poisson loops.
//
// Outer step (FP64):   r = f - A x ; |r|^2              poisson_residual_kernel
// Inner start:         e = 0 ; rr = (float) r ; p = Mi rr ;
//                      (rr.Mi rr, rr.rr)                mp_init_kernel
// Inner iteration:     w = A p ; p.w  (FP32 sweep)      poisson_calc_w_reduce_kernel
//                      e += a p ; rr -= a w ;
//                      (rr.Mi rr, rr.rr)                mp_update_kernel
//                      p = Mi rr + b p                  mp_p_kernel
// Correction (FP64):   x = x + (double) e               mp_correct_kernel
// Mi = 1/diag is a scalar for the constant-coefficient operator.  The
// preconditioned residual z = Mi rr is recomputed where it is needed instead
// of being stored: 44 bytes per point per inner iteration (4 + 4, 8 + 8 + 4 +
// 4, 8 + 4) against 88 for an FP64 CG iteration.
#include "engine.h"

using namespace opsmi;

namespace {

template <int ND>
struct PoissonRes {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 1;
  Ptr<double> r;
  Ptr<const double> f, x;
  __device__ void operator()(int i, int j, int k, double (&s)[1]) const {
    double ax;
    if constexpr (ND == 2)
      ax = 4.0 * x(i, j, k) - (x(i - 1, j, k) + x(i + 1, j, k)) - (x(i, j - 1, k) + x(i, j + 1, k));
    else
      ax = 6.0 * x(i, j, k) - (x(i - 1, j, k) + x(i + 1, j, k)) - (x(i, j - 1, k) + x(i, j + 1, k)) -
           (x(i, j, k - 1) + x(i, j, k + 1));
    const double v = f(i, j, k) - ax;
    r(i, j, k) = v;
    s[0] = s[0] + v * v;
  }
};

struct MpInit {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 2;
  static constexpr int kVec = 2;
  Ptr<float> e, rr, p;
  Ptr<const double> r;
  float mi;
  __device__ void operator()(int i, int j, int k, double (&s)[2]) const {
    const float q = (float)r(i, j, k);
    const float z = mi * q;
    e(i, j, k) = 0.0f;
    rr(i, j, k) = q;
    p(i, j, k) = z;
    s[0] = s[0] + (double)q * (double)z;
    s[1] = s[1] + (double)q * (double)q;
  }
  template <int V, bool NTL = false>
  __device__ void vop(int i, int j, int k, int n, double (&s)[2]) const {
    double rv[V];
    float q[V], z[V], zero[V];
    ldv<double, V, NTL>(r.at(i, j, k), n, rv);
#pragma unroll
    for (int c = 0; c < V; c++) {
      q[c] = (float)rv[c];
      z[c] = mi * q[c];
      zero[c] = 0.0f;
      if (c < n) {
        s[0] = s[0] + (double)q[c] * (double)z[c];
        s[1] = s[1] + (double)q[c] * (double)q[c];
      }
    }
    stv<float, V, true>(e.at(i, j, k), n, zero);
    stv<float, V, true>(rr.at(i, j, k), n, q);
    stv<float, V, true>(p.at(i, j, k), n, z);
  }
};

struct MpUpdate {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 2;
  static constexpr int kVec = 4;
  Ptr<float> e, rr;
  Ptr<const float> p, w;
  float a, mi;
  __device__ void operator()(int i, int j, int k, double (&s)[2]) const {
    e(i, j, k) = e(i, j, k) + a * p(i, j, k);
    const float q = rr(i, j, k) - a * w(i, j, k);
    rr(i, j, k) = q;
    const float z = mi * q;
    s[0] = s[0] + (double)q * (double)z;
    s[1] = s[1] + (double)q * (double)q;
  }
  template <int V, bool NTL = false>
  __device__ void vop(int i, int j, int k, int n, double (&s)[2]) const {
    float ev[V], rv[V], pv[V], wv[V];
    ldv<float, V, NTL>(e.at(i, j, k), n, ev);
    ldv<float, V, NTL>(rr.at(i, j, k), n, rv);
    ldv<float, V, NTL>(p.at(i, j, k), n, pv);
    ldv<float, V, NTL>(w.at(i, j, k), n, wv);
#pragma unroll
    for (int c = 0; c < V; c++) {
      ev[c] = ev[c] + a * pv[c];
      rv[c] = rv[c] - a * wv[c];
      const float z = mi * rv[c];
      if (c < n) {
        s[0] = s[0] + (double)rv[c] * (double)z;
        s[1] = s[1] + (double)rv[c] * (double)rv[c];
      }
    }
    stv<float, V, true>(e.at(i, j, k), n, ev);
    stv<float, V, true>(rr.at(i, j, k), n, rv);
  }
};

// the rr half of MpUpdate (same arithmetic and dots)
struct MpRrUpdate {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 2;
  static constexpr int kVec = 4;
  Ptr<float> rr;
  Ptr<const float> w;
  float a, mi;
  __device__ void operator()(int i, int j, int k, double (&s)[2]) const {
    const float q = rr(i, j, k) - a * w(i, j, k);
    rr(i, j, k) = q;
    const float z = mi * q;
    s[0] = s[0] + (double)q * (double)z;
    s[1] = s[1] + (double)q * (double)q;
  }
  template <int V, bool NTL = false>
  __device__ void vop(int i, int j, int k, int n, double (&s)[2]) const {
    float rv[V], wv[V];
    ldv<float, V, NTL>(rr.at(i, j, k), n, rv);
    ldv<float, V, NTL>(w.at(i, j, k), n, wv);
#pragma unroll
    for (int c = 0; c < V; c++) {
      rv[c] = rv[c] - a * wv[c];
      const float z = mi * rv[c];
      if (c < n) {
        s[0] = s[0] + (double)rv[c] * (double)z;
        s[1] = s[1] + (double)rv[c] * (double)rv[c];
      }
    }
    stv<float, V, true>(rr.at(i, j, k), n, rv);
  }
};

// the e half of MpUpdate
struct MpEUpdate {
  OPSMI_NO_RED
  static constexpr int kVec = 4;
  Ptr<float> e;
  Ptr<const float> p;
  float a;
  __device__ void operator()(int i, int j, int k, double (&)[1]) const { e(i, j, k) = e(i, j, k) + a * p(i, j, k); }
  template <int V, bool NTL = false>
  __device__ void vop(int i, int j, int k, int n, double (&)[1]) const {
    float ev[V], pv[V];
    ldv<float, V, NTL>(e.at(i, j, k), n, ev);
    ldv<float, V, NTL>(p.at(i, j, k), n, pv);
#pragma unroll
    for (int c = 0; c < V; c++) ev[c] = ev[c] + a * pv[c];
    stv<float, V, true>(e.at(i, j, k), n, ev);
  }
};

struct MpP {
  OPSMI_NO_RED
  static constexpr int kVec = 4;
  Ptr<float> p;
  Ptr<const float> rr;
  float mi, b;
  __device__ void operator()(int i, int j, int k, double (&)[1]) const {
    p(i, j, k) = mi * rr(i, j, k) + b * p(i, j, k);
  }
  template <int V, bool NTL = false>
  __device__ void vop(int i, int j, int k, int n, double (&)[1]) const {
    float pv[V], rv[V];
    ldv<float, V, NTL>(p.at(i, j, k), n, pv);
    ldv<float, V, NTL>(rr.at(i, j, k), n, rv);
#pragma unroll
    for (int c = 0; c < V; c++) pv[c] = mi * rv[c] + b * pv[c];
    stv<float, V, true>(p.at(i, j, k), n, pv);
  }
};

struct MpCorrect {
  OPSMI_NO_RED
  static constexpr int kVec = 2;
  Ptr<double> x;
  Ptr<const float> e;
  __device__ void operator()(int i, int j, int k, double (&)[1]) const {
    x(i, j, k) = x(i, j, k) + (double)e(i, j, k);
  }
  template <int V, bool NTL = false>
  __device__ void vop(int i, int j, int k, int n, double (&)[1]) const {
    double xv[V];
    float ev[V];
    ldv<double, V, NTL>(x.at(i, j, k), n, xv);
    ldv<float, V, NTL>(e.at(i, j, k), n, ev);
#pragma unroll
    for (int c = 0; c < V; c++) xv[c] = xv[c] + (double)ev[c];
    stv<double, V, true>(x.at(i, j, k), n, xv);
  }
};

}  // namespace

#define ARGS(...) \
  ops_arg args[] = {__VA_ARGS__}; \
  const int nargs = (int)(sizeof(args) / sizeof(args[0]));

extern "C" {

void ops_par_loop_poisson_residual_kernel(const char *name, ops_block block, int dim, int *range, ops_arg r,
                                          ops_arg f, ops_arg x, ops_arg rr) {
  OPSMI_API_BEGIN
  ARGS(r, f, x, rr)
  if (dim < 2) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": needs a 2D or 3D block"));
  Range3 rg = check_loop(name, block, dim, range, args,
                         {dat_w(DT_F64), dat_r(pts_point(), DT_F64), dat_r(dim == 3 ? pts_7pt() : pts_5pt(), DT_F64),
                          red_a(OPS_INC, DT_F64)},
                         nullptr);
  if (dim == 2)
    launch_loop(name, args, nargs, rg, [=] {
      return PoissonRes<2>{ptr_of<double>(r.dat), ptr_of<const double>(f.dat), ptr_of<const double>(x.dat)};
    }, &rr);
  else
    launch_loop(name, args, nargs, rg, [=] {
      return PoissonRes<3>{ptr_of<double>(r.dat), ptr_of<const double>(f.dat), ptr_of<const double>(x.dat)};
    }, &rr);
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_mp_init_kernel(const char *name, ops_block block, int dim, int *range, ops_arg e, ops_arg rr,
                                 ops_arg p, ops_arg r, ops_arg mi, ops_arg dots) {
  OPSMI_API_BEGIN
  ARGS(e, rr, p, r, mi, dots)
  Range3 rg = check_loop(name, block, dim, range, args,
                         {dat_w(DT_F32), dat_w(DT_F32), dat_w(DT_F32), dat_r(pts_point(), DT_F64), gbl_r(DT_F32),
                          red_a(OPS_INC, DT_F64, 2)},
                         nullptr);
  const float m = gbl<float>(mi);
  launch_loop(name, args, nargs, rg, [=] {
    return MpInit{ptr_of<float>(e.dat), ptr_of<float>(rr.dat), ptr_of<float>(p.dat), ptr_of<const double>(r.dat), m};
  }, &dots);
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_mp_update_kernel(const char *name, ops_block block, int dim, int *range, ops_arg e, ops_arg rr,
                                   ops_arg p, ops_arg w, ops_arg a, ops_arg mi, ops_arg dots) {
  OPSMI_API_BEGIN
  ARGS(e, rr, p, w, a, mi, dots)
  Range3 rg = check_loop(name, block, dim, range, args,
                         {dat_w(DT_F32), dat_w(DT_F32), dat_r(pts_point(), DT_F32), dat_r(pts_point(), DT_F32), gbl_r(DT_F32),
                          gbl_r(DT_F32), red_a(OPS_INC, DT_F64, 2)},
                         nullptr);
  const float av = gbl<float>(a), m = gbl<float>(mi);
  launch_loop(name, args, nargs, rg, [=] {
    return MpUpdate{ptr_of<float>(e.dat), ptr_of<float>(rr.dat), ptr_of<const float>(p.dat),
                    ptr_of<const float>(w.dat), av, m};
  }, &dots);
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_mp_rr_update_kernel(const char *name, ops_block block, int dim, int *range, ops_arg rr,
                                      ops_arg w, ops_arg a, ops_arg mi, ops_arg dots) {
  OPSMI_API_BEGIN
  ARGS(rr, w, a, mi, dots)
  Range3 rg = check_loop(name, block, dim, range, args,
                         {dat_w(DT_F32), dat_r(pts_point(), DT_F32), gbl_r(DT_F32), gbl_r(DT_F32),
                          red_a(OPS_INC, DT_F64, 2)},
                         nullptr);
  const float av = gbl<float>(a), m = gbl<float>(mi);
  launch_loop(name, args, nargs, rg,
              [=] { return MpRrUpdate{ptr_of<float>(rr.dat), ptr_of<const float>(w.dat), av, m}; }, &dots);
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_mp_e_update_kernel(const char *name, ops_block block, int dim, int *range, ops_arg e, ops_arg p,
                                     ops_arg a) {
  OPSMI_API_BEGIN
  ARGS(e, p, a)
  Range3 rg = check_loop(name, block, dim, range, args, {dat_w(DT_F32), dat_r(pts_point(), DT_F32), gbl_r(DT_F32)},
                         nullptr);
  const float av = gbl<float>(a);
  LoopRec rec =
      make_loop(name, args, nargs, rg, [=] { return MpEUpdate{ptr_of<float>(e.dat), ptr_of<const float>(p.dat), av}; });
  rec.kind = LK_MP_E;
  rec.dims = dim;
  rec.d[0] = e.dat;
  rec.d[1] = p.dat;
  rec.g[0] = av;
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_mp_p_kernel(const char *name, ops_block block, int dim, int *range, ops_arg p, ops_arg rr,
                              ops_arg mi, ops_arg b) {
  OPSMI_API_BEGIN
  ARGS(p, rr, mi, b)
  Range3 rg = check_loop(name, block, dim, range, args,
                         {dat_w(DT_F32), dat_r(pts_point(), DT_F32), gbl_r(DT_F32), gbl_r(DT_F32)}, nullptr);
  const float m = gbl<float>(mi), bv = gbl<float>(b);
  LoopRec rec =
      make_loop(name, args, nargs, rg, [=] { return MpP{ptr_of<float>(p.dat), ptr_of<const float>(rr.dat), m, bv}; });
  rec.kind = LK_MP_P;
  rec.dims = dim;
  rec.d[0] = p.dat;
  rec.d[1] = rr.dat;
  rec.g[0] = m;
  rec.g[1] = bv;
  enqueue(std::move(rec));
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_mp_correct_kernel(const char *name, ops_block block, int dim, int *range, ops_arg x,
                                    ops_arg e) {
  OPSMI_API_BEGIN
  ARGS(x, e)
  Range3 rg = check_loop(name, block, dim, range, args, {dat_w(DT_F64), dat_r(pts_point(), DT_F32)}, nullptr);
  launch_loop(name, args, nargs, rg,
              [=] { return MpCorrect{ptr_of<double>(x.dat), ptr_of<const float>(e.dat)}; });
  finish_loop(args, nargs);
  OPSMI_API_END()
}

}  // extern "C"
