// Multi-block Poisson (Jacobi) loops of the reference's multi-block test app
// (apps/c/poisson/poisson_kernel.h:4-38, driven by apps/c/poisson/poisson.cpp),
// in the per-loop stub form.  Constants dx, dy come from ops_decl_const; the
// arithmetic keeps the reference kernels' order (contraction off), so a
// partition into blocks joined by ops_halo_transfer gives the same bits as
// one block.
#include "engine.h"

using namespace opsmi;

namespace {

// poisson_kernel_populate: boundary/initial u, forcing f, exact solution ref
// at x = dx (i + dispx), y = dy (j + dispy); myfun(a, b) = a b + 1
// (apps/c/poisson/user_types.h)
struct MbPopulate {
  OPSMI_NO_RED
  Ptr<double> u, f, ref;
  int dispx, dispy;
  double dx, dy;
  __device__ void operator()(int i, int j, int k, double (&)[1]) const {
    const double pi = 3.14159265358979323846;
    const double x = dx * (double)(i + dispx);
    const double y = dy * (double)(j + dispy);
    const double a = sin(pi * x), b = cos(2.0 * pi * y);
    u(i, j, k) = (a * b + 1.0) - 1.0;
    f(i, j, k) = -5.0 * pi * pi * sin(pi * x) * cos(2.0 * pi * y);
    ref(i, j, k) = sin(pi * x) * cos(2.0 * pi * y);
  }
};

struct MbZero {
  OPSMI_NO_RED
  Ptr<double> u;
  __device__ void operator()(int i, int j, int k, double (&)[1]) const { u(i, j, k) = 0.0; }
};

struct MbStencil {
  OPSMI_NO_RED
  Ptr<const double> u, f;
  Ptr<double> u2;
  double dx, dy;
  __device__ void operator()(int i, int j, int k, double (&)[1]) const {
    u2(i, j, k) = (((u(i - 1, j, k) + u(i + 1, j, k)) * dx * dx) + ((u(i, j - 1, k) + u(i, j + 1, k)) * dy * dy) -
                   (f(i, j, k) * dx * dx * dy * dy)) /
                  (2.0 * (dx * dx + dy * dy));
  }
};

struct MbCopy {
  OPSMI_NO_RED
  Ptr<const double> in;
  Ptr<double> out;
  __device__ void operator()(int i, int j, int k, double (&)[1]) const { out(i, j, k) = in(i, j, k); }
};

struct MbError {
  using red_t = double;
  static constexpr int red_op = OPS_INC;
  static constexpr int red_dim = 1;
  Ptr<const double> u, ref;
  __device__ void operator()(int i, int j, int k, double (&s)[1]) const {
    const double d = u(i, j, k) - ref(i, j, k);
    s[0] = s[0] + d * d;
  }
};

void need2d(const char *name, int dim) {
  if (dim != 2) fail(OPS_INVALID_ARGUMENT, cat("loop ", name, ": 2D kernel on a ", dim, "D block"));
}

}  // namespace

#define ARGS(...) \
  ops_arg args[] = {__VA_ARGS__}; \
  const int nargs = (int)(sizeof(args) / sizeof(args[0]));

extern "C" {

void ops_par_loop_poisson_kernel_populate(const char *name, ops_block block, int dim, int *range, ops_arg dispx,
                                          ops_arg dispy, ops_arg idx, ops_arg u, ops_arg f, ops_arg ref) {
  OPSMI_API_BEGIN
  ARGS(dispx, dispy, idx, u, f, ref)
  Range3 r = check_loop(name, block, dim, range, args,
                        {gbl_r(DT_I32), gbl_r(DT_I32), idx_a(), dat_w(DT_F64), dat_w(DT_F64), dat_w(DT_F64)}, nullptr);
  need2d(name, dim);
  const int dx0 = gbl<int>(dispx), dy0 = gbl<int>(dispy);
  const double cdx = const_as_double("dx"), cdy = const_as_double("dy");
  launch_loop(name, args, nargs, r, [=] {
    return MbPopulate{ptr_of<double>(u.dat), ptr_of<double>(f.dat), ptr_of<double>(ref.dat), dx0, dy0, cdx, cdy};
  });
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_poisson_kernel_initialguess(const char *name, ops_block block, int dim, int *range, ops_arg u) {
  OPSMI_API_BEGIN
  ARGS(u)
  Range3 r = check_loop(name, block, dim, range, args, {dat_w(DT_F64)}, nullptr);
  need2d(name, dim);
  launch_loop(name, args, nargs, r, [=] { return MbZero{ptr_of<double>(u.dat)}; });
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_poisson_kernel_stencil(const char *name, ops_block block, int dim, int *range, ops_arg u, ops_arg f,
                                         ops_arg u2) {
  OPSMI_API_BEGIN
  ARGS(u, f, u2)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_r(pts_5pt(), DT_F64), dat_r(pts_point(), DT_F64), dat_w(DT_F64)}, nullptr);
  need2d(name, dim);
  const double cdx = const_as_double("dx"), cdy = const_as_double("dy");
  launch_loop(name, args, nargs, r, [=] {
    return MbStencil{ptr_of<const double>(u.dat), ptr_of<const double>(f.dat), ptr_of<double>(u2.dat), cdx, cdy};
  });
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_poisson_kernel_update(const char *name, ops_block block, int dim, int *range, ops_arg in,
                                        ops_arg out) {
  OPSMI_API_BEGIN
  ARGS(in, out)
  Range3 r = check_loop(name, block, dim, range, args, {dat_r(pts_point(), DT_F64), dat_w(DT_F64)}, nullptr);
  need2d(name, dim);
  launch_loop(name, args, nargs, r,
              [=] { return MbCopy{ptr_of<const double>(in.dat), ptr_of<double>(out.dat)}; });
  finish_loop(args, nargs);
  OPSMI_API_END()
}

void ops_par_loop_poisson_kernel_error(const char *name, ops_block block, int dim, int *range, ops_arg u,
                                       ops_arg ref, ops_arg err) {
  OPSMI_API_BEGIN
  ARGS(u, ref, err)
  Range3 r = check_loop(name, block, dim, range, args,
                        {dat_r(pts_point(), DT_F64), dat_r(pts_point(), DT_F64), red_a(OPS_INC, DT_F64)}, nullptr);
  need2d(name, dim);
  launch_loop(name, args, nargs, r,
              [=] { return MbError{ptr_of<const double>(u.dat), ptr_of<const double>(ref.dat)}; }, &err);
  finish_loop(args, nargs);
  OPSMI_API_END()
}

}  // extern "C"
