// Device loop engine shared by the hand-written ops_par_loop kernels.
//
// Work mapping: a workgroup is 64 x 4 threads (four wave64s); lanes run
// along x, so every global access of a wave is one contiguous 512-byte (f64)
// segment.  Non-reducing loops give each thread one point: short-lived waves
// keep the HBM pipeline full.  Workgroups are dealt to the 8 XCDs
// round-robin by the hardware, so the engine renumbers them: XCD k works
// through the k-th contiguous eighth of the tiles, and stencil neighbours
// (adjacent rows and planes) are mostly found in the same XCD's L2.
//
// Reductions are two-stage and fence-free: each thread folds its points into
// registers, the block folds them with wave64 butterflies (__shfl_xor) and
// LDS and stores one partial; a second kernel folds the partials in a fixed
// order (spans of 2048, then one block) and combines the result with the
// handle's running value (or the OPS initial value on first use).  The order
// depends only on the launch shape, so results are reproducible run to run.
#pragma once

#include <float.h>
#include <limits.h>

#include <type_traits>

#include "runtime.h"

namespace opsmi {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kSpan = 2048;  // partials folded by one second-stage block

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

__device__ __forceinline__ unsigned flat_block() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ __forceinline__ unsigned num_blocks() { return gridDim.x * gridDim.y * gridDim.z; }

// XCD-aware renumbering: hardware block b runs on XCD b % 8; it takes logical
// tile start(b % 8) + b / 8, where XCD k owns a contiguous logical range of
// its floor/ceil share of the nb tiles.  A bijection on [0, nb).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  const unsigned q8 = nb >> 3, r8 = nb & 7u, k = b & 7u;
  return k * q8 + (k < r8 ? k : r8) + (b >> 3);
}

// stage 1: block fold + one partial per block
template <class T, int OP, int D>
__device__ __forceinline__ void reduce_block_store(T (&acc)[D], const RedCtl &c) {
  __shared__ T sh[16][D];
  block_reduce<T, OP, D>(acc, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    T *part = (T *)c.partials;
#pragma unroll
    for (int q = 0; q < D; q++) part[(size_t)flat_block() * D + q] = acc[q];
  }
}

// stage 2: block b folds in[b*kSpan, (b+1)*kSpan) (eight loads per thread)
template <class T, int OP, int D>
__global__ __launch_bounds__(256) void k_red_stage2(const T *__restrict__ in, unsigned n, T *__restrict__ out) {
  __shared__ T sh[16][D];
  const unsigned tid = threadIdx.x, base = blockIdx.x * kSpan;
  T v[D];
#pragma unroll
  for (int q = 0; q < D; q++) {
    T t8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const unsigned i = base + tid + u * 256u;
      t8[u] = i < n ? in[(size_t)i * D + q] : RedOp<T, OP>::identity();
    }
    v[q] = t8[0];
#pragma unroll
    for (int u = 1; u < 8; u++) v[q] = RedOp<T, OP>::combine(v[q], t8[u]);
  }
  block_reduce<T, OP, D>(v, sh);
  if (tid == 0)
#pragma unroll
    for (int q = 0; q < D; q++) out[(size_t)blockIdx.x * D + q] = v[q];
}

// final: one block folds n values and updates the handle's device slot
template <class T, int OP, int D>
__global__ __launch_bounds__(256) void k_red_final(const T *__restrict__ in, unsigned n, RedCtl c) {
  __shared__ T sh[16][D];
  const unsigned tid = threadIdx.x;
  T v[D];
#pragma unroll
  for (int q = 0; q < D; q++) {
    v[q] = RedOp<T, OP>::identity();
    unsigned i = tid;
    for (; i + 7 * 256u < n; i += 8 * 256u) {
      T t8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) t8[u] = in[(size_t)(i + u * 256u) * D + q];
#pragma unroll
      for (int u = 0; u < 8; u++) v[q] = RedOp<T, OP>::combine(v[q], t8[u]);
    }
    for (; i < n; i += 256u) v[q] = RedOp<T, OP>::combine(v[q], in[(size_t)i * D + q]);
  }
  block_reduce<T, OP, D>(v, sh);
  if (tid == 0) {
    T *slot = (T *)c.slot;
    T *mirror = (T *)c.mirror;
#pragma unroll
    for (int q = 0; q < D; q++) {
      const T base = c.use_init ? init_value<T>(c, q) : slot[q];
      const T r = RedOp<T, OP>::combine(base, v[q]);
      slot[q] = r;
      mirror[q] = r;
    }
    // the values reach host memory before the sequence word the host polls
    if (c.seq) {
      __threadfence_system();
      *(volatile unsigned long long *)c.seq = c.seq_val;
    }
  }
}

// launches the fold of `parts` stage-1 partials into the handle (none when
// the host folds them, with_host_fold)
template <class T, int OP, int D>
void launch_red_finish(const RedCtl &c, unsigned parts, hipStream_t s) {
  if (c.hostfold) return;
  if (parts <= 4u * kSpan) {
    hipLaunchKernelGGL((k_red_final<T, OP, D>), dim3(1), dim3(256), 0, s, (const T *)c.partials, parts, c);
  } else {
    const unsigned g2 = (parts + kSpan - 1) / kSpan;
    hipLaunchKernelGGL((k_red_stage2<T, OP, D>), dim3(g2), dim3(256), 0, s, (const T *)c.partials, parts,
                       (T *)c.partials2);
    hipLaunchKernelGGL((k_red_final<T, OP, D>), dim3(1), dim3(256), 0, s, (const T *)c.partials2, g2, c);
  }
}

// Tiling of a loop range into 64 x 4 workgroups.  3D ranges: tile (tx, ty,
// tz) covers x in 64 tx + [0,64), y in 4 ty + [0,4) and `chunk` z planes
// from chunk*tz.  2D/1D ranges: tile (tx, ty) covers 64 x-points and `chunk`
// groups of 4 rows.
struct Tiling {
  dim3 grid;
  int chunk;
  bool three_d;
  unsigned blocks() const { return grid.x * grid.y * grid.z; }
};

inline Tiling make_tiling(const Range3 &r, int chunk = 1) {
  Tiling t;
  const long nx = r.x1 - r.x0, ny = r.y1 - r.y0, nz = r.z1 - r.z0;
  t.three_d = nz > 1;
  const long gx = (nx + kWave - 1) / kWave, gy = (ny + kWavesPerBlock - 1) / kWavesPerBlock;
  const long planes = t.three_d ? nz : gy;
  const long gc = (planes + chunk - 1) / chunk;
  if (t.three_d) t.grid = dim3((unsigned)gx, (unsigned)gy, (unsigned)gc);
  else t.grid = dim3((unsigned)gx, (unsigned)gc, 1);
  t.chunk = chunk;
  return t;
}

// logical tile coordinates of this workgroup (XCD-aware)
struct TileId {
  unsigned x, y, z;
};
__device__ __forceinline__ TileId tile_id() {
  const unsigned l = xcd_tile(flat_block(), num_blocks());
  TileId t;
  t.x = l % gridDim.x;
  const unsigned rest = l / gridDim.x;
  t.y = rest % gridDim.y;
  t.z = rest / gridDim.y;
  return t;
}

// ------------------------------------------------------------ vector access
template <class T, int V>
using vec_t = T __attribute__((ext_vector_type(V)));

// host: 16-byte loads are possible at (x0, y, z) of every row and plane
inline bool vec16_ok(ops_dat d, int x0) {
  if (!d) return false;
  const DatView v = view_of(d);
  const long es = d->type_size;
  return (uintptr_t)(v.base + x0 * es) % 16 == 0 && (v.sy * es) % 16 == 0 && (v.sz * es) % 16 == 0;
}
// host: OPSMI_VEC=0 keeps the one-point-per-lane kernels
inline bool use_vec() {
  const char *e = getenv("OPSMI_VEC");
  return !(e && e[0] == '0');
}

// NTL: loads carry the non-temporal hint.  The pointwise engine uses it for
// loops whose arrays are larger than the Infinity Cache: a pointwise 4-read
// 1-write sweep over 1 GiB arrays measured 6.24 TB/s with it against
// 5.77 TB/s without (6 reads 3 writes: 5.96 against 5.25).  The stencil
// columns load without it (their rims are re-read by neighbouring
// workgroups from L2), as do loops over arrays that stay in the cache.
template <class T, bool NTL>
__device__ __forceinline__ T ldg(const T *p) {
  if constexpr (NTL) return __builtin_nontemporal_load(p);
  else return *p;
}

// n <= V consecutive values from p (one vector load when all V are valid)
template <class T, int V, bool NTL = false>
__device__ __forceinline__ void ldv(const T *p, int n, T (&o)[V]) {
  if (n >= V) {
    const vec_t<T, V> q = ldg<vec_t<T, V>, NTL>(reinterpret_cast<const vec_t<T, V> *>(p));
#pragma unroll
    for (int j = 0; j < V; j++) o[j] = q[j];
  } else {
#pragma unroll
    for (int j = 0; j < V; j++) o[j] = j < n ? ldg<T, NTL>(p + j) : T(0);
  }
}

// n <= V consecutive values to p; NT: non-temporal (streaming) stores
template <class T, int V, bool NT>
__device__ __forceinline__ void stv(T *p, int n, const T (&v)[V]) {
  if (n >= V) {
    vec_t<T, V> q;
#pragma unroll
    for (int j = 0; j < V; j++) q[j] = v[j];
    if constexpr (NT) __builtin_nontemporal_store(q, reinterpret_cast<vec_t<T, V> *>(p));
    else *reinterpret_cast<vec_t<T, V> *>(p) = q;
  } else {
#pragma unroll
    for (int j = 0; j < V; j++)
      if (j < n) {
        if constexpr (NT) __builtin_nontemporal_store(v[j], p + j);
        else p[j] = v[j];
      }
  }
}

// F provides: using red_t; static constexpr int red_op, red_dim;
//   __device__ void operator()(int x, int y, int z, red_t (&acc)[red_dim or 1]) const
// and optionally, for the vector engine (k_loop_v), kVec > 1 and
//   template <int V, bool NTL> __device__ void vop(int x, int y, int z, int n, red_t (&acc)[...]) const
// which handles the n <= V consecutive points x .. x + n - 1 (NTL: its loads
// take the non-temporal hint).
template <class F>
__global__ __launch_bounds__(256) void k_loop(Range3 r, F f, RedCtl ctl, int chunk, int three_d) {
  using R = typename F::red_t;
  constexpr int D = F::red_dim > 0 ? F::red_dim : 1;
  constexpr int OP = F::red_op;
  R acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = RedOp<R, OP>::identity();
  const TileId t = tile_id();
  const int x = r.x0 + (int)t.x * kWave + (int)threadIdx.x;
  if (x < r.x1) {
    if (three_d) {
      const int y = r.y0 + (int)t.y * kWavesPerBlock + (int)threadIdx.y;
      if (y < r.y1) {
        const int z0 = r.z0 + (int)t.z * chunk;
        const int z1 = min(z0 + chunk, r.z1);
        for (int z = z0; z < z1; z++) f(x, y, z, acc);
      }
    } else {
      const int y0 = r.y0 + (int)t.y * kWavesPerBlock * chunk;
      const int y1 = min(y0 + kWavesPerBlock * chunk, r.y1);
      for (int y = y0 + (int)threadIdx.y; y < y1; y += kWavesPerBlock) f(x, y, r.z0, acc);
    }
  }
  if constexpr (F::red_dim > 0) reduce_block_store<R, OP, D>(acc, ctl);
}

// Vector form of k_loop: each lane takes V consecutive x points (one 16-byte
// load per array when the functor's element is 16 / V bytes), so a wave
// covers 64 V points of a row.
template <class F, int V, bool NTL>
__global__ __launch_bounds__(256) void k_loop_v(Range3 r, F f, RedCtl ctl, int chunk, int three_d) {
  using R = typename F::red_t;
  constexpr int D = F::red_dim > 0 ? F::red_dim : 1;
  constexpr int OP = F::red_op;
  R acc[D];
#pragma unroll
  for (int q = 0; q < D; q++) acc[q] = RedOp<R, OP>::identity();
  const TileId t = tile_id();
  const int x = r.x0 + (int)t.x * kWave * V + V * (int)threadIdx.x;
  const int n = min(V, r.x1 - x);
  // a lane with all V points calls vop with the literal V, so the
  // element-count branches of its loads and stores fold away (a branch join
  // makes the compiler wait for every load in flight)
  auto pt = [&](int y, int z) {
    if (n == V) f.template vop<V, NTL>(x, y, z, V, acc);
    else f.template vop<V, NTL>(x, y, z, n, acc);
  };
  if (n > 0) {
    if (three_d) {
      const int y = r.y0 + (int)t.y * kWavesPerBlock + (int)threadIdx.y;
      if (y < r.y1) {
        const int z0 = r.z0 + (int)t.z * chunk;
        const int z1 = min(z0 + chunk, r.z1);
        for (int z = z0; z < z1; z++) pt(y, z);
      }
    } else {
      const int y0 = r.y0 + (int)t.y * kWavesPerBlock * chunk;
      const int y1 = min(y0 + kWavesPerBlock * chunk, r.y1);
      for (int y = y0 + (int)threadIdx.y; y < y1; y += kWavesPerBlock) pt(y, r.z0);
    }
  }
  if constexpr (F::red_dim > 0) reduce_block_store<R, OP, D>(acc, ctl);
}

// bytes (OPS transfer rule) from which a pointwise loop loads non-temporally:
// the 256 MB Infinity Cache
constexpr double kNtLoadBytes = 256.0 * (1 << 20);

template <class F, class = void>
struct vec_width {
  static constexpr int value = 1;
};
template <class F>
struct vec_width<F, std::void_t<decltype(F::kVec)>> {
  static constexpr int value = F::kVec;
};

template <class T>
constexpr DType dtype_v() {
  return std::is_same<T, double>::value ? DT_F64 : (std::is_same<T, float>::value ? DT_F32 : DT_I32);
}

// reduction control with the current scratch buffers (taken at launch time)
inline RedCtl with_scratch(RedCtl c) {
  c.partials = inst().d_partials;
  c.partials2 = inst().d_partials2;
  c.hostfold = 0;
  return c;
}

// with_scratch for a kernel of nblocks workgroups that ends in
// reduce_block_store<T, OP, dim> in every workgroup: when the handle's result
// is read right after this flush and the launch qualifies (FP64, at most
// kHostFoldMax partials), its partials go to the pinned host buffer instead,
// preset to kHostFoldEmpty, and launch_red_finish launches nothing; the
// reader folds them (host_fold_result in runtime.cpp)
// claims the host fold for a launch reducing through ctl (true: its
// partials go to I.d_hfold, preset here)
inline bool host_fold_claim(const RedCtl &c, unsigned nblocks, int dim, DType dt) {
  Instance &I = inst();
  if (!I.hf_slot || I.hf_used || c.slot != I.hf_slot || dt != DT_F64 || dim < 1 || dim > 4 ||
      (size_t)nblocks * dim > kHostFoldMax)
    return false;
  unsigned long long *h = (unsigned long long *)I.h_hfold;
  for (size_t i = 0; i < (size_t)nblocks * dim; i++) h[i] = kHostFoldEmpty;
  I.hf_used = true;
  I.hf_n = nblocks;
  I.hf_dim = dim;
  return true;
}
inline RedCtl with_host_fold(RedCtl c, unsigned nblocks, int dim, DType dt) {
  c = with_scratch(c);
  if (host_fold_claim(c, nblocks, dim, dt)) {
    c.partials = inst().d_hfold;
    c.hostfold = 1;
  }
  return c;
}

// Builds the queue record of a loop over r.  make() constructs the kernel
// functor; it runs when the loop is launched, so device pointers are those
// of the dats at that moment.  red_arg is the ops_arg_reduce of a reducing
// functor.
template <class Make>
LoopRec make_loop(const char *name, const ops_arg *args, int nargs, const Range3 &r, Make make,
                  const ops_arg *red_arg = nullptr) {
  using F = decltype(make());
  LoopRec rec;
  rec.name = name ? name : "?";
  rec.range = r;
  rec.bytes = transfer_bytes(args, nargs, r);
  deps_from_args(rec, args, nargs);
  if (r.npts() <= 0) {
    rec.bytes = 0.0;
    rec.launch = [](hipStream_t) {};
    return rec;
  }
  // 2D ranges: four groups of four rows per workgroup (64 x 16 points), so a
  // 2048^2 sweep is 4096 workgroups and its reduction one fold kernel; 3D
  // ranges: one plane of 64 x 4 per workgroup
  Tiling t = make_tiling(r, r.z1 - r.z0 > 1 ? 1 : 4);
  // vector engine when the functor has one and every dat allows 16-byte
  // loads at the range's x start
  constexpr int V = vec_width<F>::value;
  bool vec = V > 1 && use_vec();
  for (int i = 0; i < nargs && vec; i++)
    if (args[i].argtype == OPS_ARG_DAT) vec = vec16_ok(args[i].dat, r.x0);
  if (vec) t.grid.x = (unsigned)((r.x1 - r.x0 + kWave * V - 1) / (kWave * V));
  // non-temporal loads once the loop's arrays outgrow the Infinity Cache
  const bool ntl = rec.bytes >= kNtLoadBytes;
  RedCtl ctl{};
  if constexpr (F::red_dim > 0) ctl = red_prepare(*red_arg, F::red_dim, dtype_v<typename F::red_t>(), t.blocks());
  rec.ctl = ctl;
  rec.launch = [=](hipStream_t s) {
    const F f = make();
    const RedCtl c = F::red_dim > 0 ? with_host_fold(ctl, t.blocks(), F::red_dim, dtype_v<typename F::red_t>())
                                    : with_scratch(ctl);
    if (vec) {
      if constexpr (V > 1) {
        auto kern = ntl ? k_loop_v<F, V, true> : k_loop_v<F, V, false>;
        hipLaunchKernelGGL(kern, t.grid, dim3(kWave, kWavesPerBlock), 0, s, r, f, c, t.chunk, t.three_d ? 1 : 0);
      }
    } else {
      hipLaunchKernelGGL(k_loop<F>, t.grid, dim3(kWave, kWavesPerBlock), 0, s, r, f, c, t.chunk, t.three_d ? 1 : 0);
    }
    if constexpr (F::red_dim > 0) launch_red_finish<typename F::red_t, F::red_op, F::red_dim>(c, t.blocks(), s);
  };
  return rec;
}

template <class Make>
void launch_loop(const char *name, const ops_arg *args, int nargs, const Range3 &r, Make make,
                 const ops_arg *red_arg = nullptr) {
  enqueue(make_loop(name, args, nargs, r, make, red_arg));
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
