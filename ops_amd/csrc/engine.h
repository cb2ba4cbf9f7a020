// Device loop engine shared by the hand-written ops_par_loop kernels.
//
// Work mapping: a workgroup is 64 x 4 threads (four wave64s).  Each wave owns
// one (y,z) row of the loop range at a time and its 64 lanes stride along x,
// so every global access of a wave is one contiguous 512-byte (f64) segment.
// Rows are dealt to waves grid-stride; the grid is capped at 8 workgroups per
// CU, so every thread visits many points and the number of reduction partials
// stays small.
//
// Reductions: each thread folds its points into registers in its own order,
// then the block reduces with wave64 butterflies (__shfl_xor) and LDS, writes
// one partial, and takes a ticket.  The last block to finish combines all
// partials in a fixed order (so results do not depend on scheduling), folds
// in the handle's previous value (or the OPS initial value on first use),
// writes the handle's device slot and resets the ticket.  Partials and the
// ticket go through device-scope atomics so blocks on different XCDs (each
// with its own L2) see each other's writes.
#pragma once

#include <float.h>
#include <limits.h>

#include <type_traits>

#include "runtime.h"

namespace opsmi {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

template <class T>
struct Ptr {
  T *p;
  long sy, sz;
  __device__ __forceinline__ T &operator()(int x, int y, int z) const { return p[x + y * sy + z * sz]; }
  __device__ __forceinline__ T *at(int x, int y, int z) const { return p + (x + y * sy + z * sz); }
};

template <class T>
inline Ptr<T> ptr_of(ops_dat d) {
  DatView v = view_of(d);
  return Ptr<T>{(T *)v.base, v.sy, v.sz};
}

// members of a non-reducing loop functor
#define OPSMI_NO_RED                    \
  using red_t = double;                 \
  static constexpr int red_op = OPS_INC; \
  static constexpr int red_dim = 0;

template <class T, int OP>
struct RedOp;
template <class T>
struct RedOp<T, OPS_INC> {
  __device__ static T identity() { return T(0); }
  __device__ static T combine(T a, T b) { return a + b; }
};
template <>
struct RedOp<double, OPS_MIN> {
  __device__ static double identity() { return __builtin_huge_val(); }
  __device__ static double combine(double a, double b) { return fmin(a, b); }
};
template <>
struct RedOp<double, OPS_MAX> {
  __device__ static double identity() { return -__builtin_huge_val(); }
  __device__ static double combine(double a, double b) { return fmax(a, b); }
};
template <>
struct RedOp<float, OPS_MIN> {
  __device__ static float identity() { return __builtin_huge_valf(); }
  __device__ static float combine(float a, float b) { return fminf(a, b); }
};
template <>
struct RedOp<float, OPS_MAX> {
  __device__ static float identity() { return -__builtin_huge_valf(); }
  __device__ static float combine(float a, float b) { return fmaxf(a, b); }
};
template <>
struct RedOp<int, OPS_MIN> {
  __device__ static int identity() { return INT_MAX; }
  __device__ static int combine(int a, int b) { return a < b ? a : b; }
};
template <>
struct RedOp<int, OPS_MAX> {
  __device__ static int identity() { return INT_MIN; }
  __device__ static int combine(int a, int b) { return a > b ? a : b; }
};

template <class T>
__device__ __forceinline__ T init_value(const RedCtl &c, int q);
template <>
__device__ __forceinline__ double init_value<double>(const RedCtl &c, int q) { return c.init[q]; }
template <>
__device__ __forceinline__ float init_value<float>(const RedCtl &c, int q) { return (float)c.init[q]; }
template <>
__device__ __forceinline__ int init_value<int>(const RedCtl &c, int q) { return c.init_i[q]; }

template <class T>
__device__ __forceinline__ void store_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T load_agent(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T, int OP>
__device__ __forceinline__ T wave_reduce(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = RedOp<T, OP>::combine(v, __shfl_xor(v, off, kWave));
  return v;
}

// block-wide fold of D values; result valid in thread 0
template <class T, int OP, int D>
__device__ __forceinline__ void block_reduce(T (&v)[D], T (*sh)[D]) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nw = (blockDim.x * blockDim.y) / kWave;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
#pragma unroll
  for (int q = 0; q < D; q++) {
    T w = wave_reduce<T, OP>(v[q]);
    if (lane == 0) sh[wid][q] = w;
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < D; q++) {
      T s = sh[0][q];
      for (int w = 1; w < nw; w++) s = RedOp<T, OP>::combine(s, sh[w][q]);
      v[q] = s;
    }
  }
}

template <class T, int OP, int D>
__device__ void reduce_finish(T (&acc)[D], const RedCtl &c) {
  __shared__ T sh[16][D];
  __shared__ int is_last;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  T *part = (T *)c.partials;
  block_reduce<T, OP, D>(acc, sh);
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < D; q++) store_agent(&part[(size_t)blockIdx.x * D + q], acc[q]);
    __threadfence();
    unsigned t = atomicAdd(c.ticket, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  T v[D];
#pragma unroll
  for (int q = 0; q < D; q++) {
    v[q] = RedOp<T, OP>::identity();
    for (int b = tid; b < (int)gridDim.x; b += nthr) v[q] = RedOp<T, OP>::combine(v[q], load_agent(&part[(size_t)b * D + q]));
  }
  __syncthreads();
  block_reduce<T, OP, D>(v, sh);
  if (tid == 0) {
    T *slot = (T *)c.slot;
#pragma unroll
    for (int q = 0; q < D; q++) {
      T base = c.use_init ? init_value<T>(c, q) : load_agent(&slot[q]);
      store_agent(&slot[q], RedOp<T, OP>::combine(base, v[q]));
    }
    store_agent(c.ticket, 0u);
  }
}

// F provides: using red_t; static constexpr int red_op, red_dim;
//   __device__ void operator()(int x, int y, int z, red_t (&acc)[red_dim or 1]) const
template <class F>
__global__ __launch_bounds__(256) void k_loop(Range3 r, F f, RedCtl ctl) {
  using R = typename F::red_t;
  constexpr int D = F::red_dim > 0 ? F::red_dim : 1;
  constexpr int OP = F::red_op;
  R acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = RedOp<R, OP>::identity();
  const int lane = threadIdx.x;
  const int ny = r.y1 - r.y0;
  const long nrows = (long)ny * (r.z1 - r.z0);
  for (long row = (long)blockIdx.x * blockDim.y + threadIdx.y; row < nrows; row += (long)gridDim.x * blockDim.y) {
    const int y = r.y0 + (int)(row % ny), z = r.z0 + (int)(row / ny);
    for (int x = r.x0 + lane; x < r.x1; x += kWave) f(x, y, z, acc);
  }
  if constexpr (F::red_dim > 0) reduce_finish<R, OP, D>(acc, ctl);
}

template <class T>
constexpr DType dtype_v() {
  return std::is_same<T, double>::value ? DT_F64 : (std::is_same<T, float>::value ? DT_F32 : DT_I32);
}

// launches F over r; red_arg is the ops_arg_reduce of a reducing F
template <class F>
void launch_loop(const char *name, double bytes, const Range3 &r, const F &f, const ops_arg *red_arg = nullptr) {
  if (r.npts() <= 0) {
    run_loop(name, 0.0, [](hipStream_t) {});
    return;
  }
  long rows = (long)(r.y1 - r.y0) * (r.z1 - r.z0);
  int nb = grid_blocks(rows, kWavesPerBlock);
  RedCtl ctl{};
  if constexpr (F::red_dim > 0) ctl = red_prepare(*red_arg, F::red_dim, dtype_v<typename F::red_t>(), nb);
  run_loop(name, bytes, [&](hipStream_t s) {
    hipLaunchKernelGGL(k_loop<F>, dim3(nb), dim3(kWave, kWavesPerBlock), 0, s, r, f, ctl);
  });
}

// marks every written dat as newest on the device
inline void finish_loop(const ops_arg *args, int n) {
  for (int i = 0; i < n; i++)
    if (args[i].argtype == OPS_ARG_DAT && args[i].dat && args[i].acc != OPS_READ) mark_device_dirty(args[i].dat);
}

// spec builders
inline ArgSpec dat_r(std::vector<int> pts = pts_point(), DType dt = DT_OTHER) {
  ArgSpec s{AK_DAT};
  s.offsets = std::move(pts);
  s.dt = dt;
  return s;
}
inline ArgSpec dat_w(DType dt = DT_OTHER) {
  ArgSpec s{AK_DAT};
  s.write = true;
  s.offsets = pts_point();
  s.dt = dt;
  return s;
}
inline ArgSpec gbl_r(DType dt = DT_F64, int dim = 1) {
  ArgSpec s{AK_GBL};
  s.dt = dt;
  s.dim = dim;
  return s;
}
inline ArgSpec red_a(int op, DType dt = DT_OTHER, int dim = 1) {
  ArgSpec s{AK_RED};
  s.red_op = op;
  s.dt = dt;
  s.dim = dim;
  return s;
}
inline ArgSpec idx_a() { return ArgSpec{AK_IDX}; }

}  // namespace opsmi
