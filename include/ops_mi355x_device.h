// Device side of translator-generated loops (MI355X OPS backend).
//
// A generated kernel runs a 64 x 4 workgroup per tile, OPSMI_GEN_XP grid
// points per thread along x (each wave reads contiguous rows); tiles are
// renumbered so each XCD works through a contiguous share of them (its L2
// then holds the rows its stencils share).  3D loops: a tile is
// (64 OPSMI_GEN_XP) x 4 points over `chunk` z planes; 2D/1D loops: 64
// OPSMI_GEN_XP points over `chunk` groups of 4 rows.  Reductions keep one accumulator per thread and component,
// initialised as the reference's generated code does (0, +-DBL_MAX,
// +-FLT_MAX, +-INT_MAX: ops_translator/resources/templates/cpp/cuda/
// loop_host.cpp.j2:116-127, ops/c/include/ops_internal2.h:68-80), fold
// them across the workgroup with wave64 shuffles, and store one partial per
// workgroup; the library folds the partials in a fixed order.
//
// ACC<T> (ops/c/include/ops_lib_core.h:1356-2038, AoS layout): A(dx[, dy[,
// dz]]) is the point at that offset from the current one, A(d, dx[, ...])
// component d of it; combine_max/min/inc update a point atomically.
#pragma once

#include <float.h>
#include <hip/hip_runtime.h>
#include <limits.h>

#include <type_traits>

// the legacy index macros are this header's (below), not the host
// fallbacks of ops_mi355x.h
#define OPSMI_DEVICE_LEGACY_ACC
#include "ops_mi355x_gen.h"

namespace opsmi_gen {

constexpr int kWave = 64;
constexpr int kRows = 4;

struct Loop {
  int x0, x1, y0, y1, z0, z1;
  int chunk, three_d, spread;
};

inline Loop loop_of(const opsmi_gen_launch *L) {
  Loop l;
  l.x0 = L->range[0];
  l.x1 = L->range[1];
  l.y0 = L->range[2];
  l.y1 = L->range[3];
  l.z0 = L->range[4];
  l.z1 = L->range[5];
  l.chunk = L->chunk;
  l.three_d = L->three_d;
  l.spread = L->spread;
  return l;
}

// Legacy kernels (the OPS_ACC macro API, ops/c/include/ops_macros.h:
// 723-1133, 1345-1656) take T * parameters and index them with
// OPS_ACCn(dx[, dy[, dz]]) or OPS_ACC_MDn(d, dx[, dy[, dz]]), where n is the
// argument's position and the macro turns the offsets into an element index
// with that argument's sizes.  Here the parameters are accessors and the
// macros build the offsets themselves, which each accessor resolves with its
// own strides (operator[] below), so n plays no part.
struct Off {
  int x = 0, y = 0, z = 0;
};
struct OffMD {
  int d = 0, x = 0, y = 0, z = 0;
};

// NDIM: block dimensions; MULTI: the dat may have more than one component
// (otherwise the component count is the constant 1); GEN: the x step comes
// from the launch (0 when the argument's stencil does not advance in x, a
// strided stencil over a dat of size 1 in x) instead of being the constant
// component count.  Strides along y and z always come from the launch (0
// for a dimension the stencil does not advance in).
template <class T, int NDIM, bool MULTI, bool GEN = false>
class Acc {
  T *p_;
  long md_, sy_, sz_, cs_;

  __device__ __forceinline__ long pt(int dx, int dy, int dz) const {
    long o = (long)dx * ((MULTI || GEN) ? md_ : 1);
    if (NDIM > 1) o += (long)dy * sy_;
    if (NDIM > 2) o += (long)dz * sz_;
    return o;
  }

 public:
  __device__ __forceinline__ Acc(const opsmi_gen_dat &d, int x, int y, int z)
      : md_(d.md), sy_(d.sy), sz_(d.sz), cs_(d.cs) {
    p_ = (T *)d.base + pt(x, y, z);
  }
  // base: the dat's (restrict-qualified) kernel argument, d.base's value
  __device__ __forceinline__ Acc(const T *base, const opsmi_gen_dat &d, int x, int y, int z)
      : md_(d.md), sy_(d.sy), sz_(d.sz), cs_(d.cs) {
    p_ = const_cast<T *>(base) + pt(x, y, z);
  }
  // NDIM indices: a point; NDIM + 1: component, then the point
  template <class... I>
  __device__ __forceinline__ T &operator()(I... i) const {
    constexpr int n = sizeof...(I);
    static_assert(n == NDIM || n == NDIM + 1, "ACC: give NDIM offsets, or a component and NDIM offsets");
    const int v[n + 3] = {int(i)..., 0, 0, 0};
    if constexpr (n == NDIM) return p_[pt(v[0], v[1], v[2])];
    else return p_[(MULTI ? (long)v[0] * cs_ : (long)v[0]) + pt(v[1], v[2], v[3])];
  }
  // legacy indexing: p[OPS_ACCn(dx, dy, dz)], p[OPS_ACC_MDn(d, dx, dy, dz)]
  __device__ __forceinline__ T &operator[](Off o) const { return p_[pt(o.x, o.y, o.z)]; }
  __device__ __forceinline__ T &operator[](OffMD o) const {
    return p_[(MULTI ? (long)o.d * cs_ : (long)o.d) + pt(o.x, o.y, o.z)];
  }
  __device__ void combine_max(int dx, T val) const { amax(&p_[pt(dx, 0, 0)], val); }
  __device__ void combine_max(int dx, int dy, T val) const { amax(&p_[pt(dx, dy, 0)], val); }
  __device__ void combine_max(int dx, int dy, int dz, T val) const { amax(&p_[pt(dx, dy, dz)], val); }
  __device__ void combine_min(int dx, T val) const { amin(&p_[pt(dx, 0, 0)], val); }
  __device__ void combine_min(int dx, int dy, T val) const { amin(&p_[pt(dx, dy, 0)], val); }
  __device__ void combine_min(int dx, int dy, int dz, T val) const { amin(&p_[pt(dx, dy, dz)], val); }
  __device__ void combine_inc(int dx, T val) const { atomicAdd(&p_[pt(dx, 0, 0)], val); }
  __device__ void combine_inc(int dx, int dy, T val) const { atomicAdd(&p_[pt(dx, dy, 0)], val); }
  __device__ void combine_inc(int dx, int dy, int dz, T val) const { atomicAdd(&p_[pt(dx, dy, dz)], val); }

 private:
  // compare-and-swap loops on the value's bits (integers use the native
  // atomics)
  template <bool MAX>
  __device__ static void acas(T *a, T val) {
    if constexpr (sizeof(T) == 8) {
      unsigned long long *u = (unsigned long long *)a;
      unsigned long long old = *u;
      while (true) {
        const T cur = __longlong_as_double(old);
        if (!(MAX ? val > cur : val < cur)) break;
        const unsigned long long prev = atomicCAS(u, old, (unsigned long long)__double_as_longlong(val));
        if (prev == old) break;
        old = prev;
      }
    } else {
      unsigned *u = (unsigned *)a;
      unsigned old = *u;
      while (true) {
        const T cur = __uint_as_float(old);
        if (!(MAX ? val > cur : val < cur)) break;
        const unsigned prev = atomicCAS(u, old, __float_as_uint(val));
        if (prev == old) break;
        old = prev;
      }
    }
  }
  __device__ static void amax(T *a, T val) {
    if constexpr (std::is_integral_v<T>) atomicMax(a, val);
    else acas<true>(a, val);
  }
  __device__ static void amin(T *a, T val) {
    if constexpr (std::is_integral_v<T>) atomicMin(a, val);
    else acas<false>(a, val);
  }
};

// ------------------------------------------------------------ fused pairs
// A written argument of a kernel run at a neighbouring point only for the
// values it would write there (a fused pair whose second loop reads, at
// stencil offsets, what the first loop writes): the writes go to registers,
// nothing to memory.  The first loop writes such a dat only at the point
// itself, so every access names that point; with C > 1 components the first
// index is the component when NDIM + 1 indices are given.
template <class T, int NDIM, int C>
struct Cap {
  mutable T v[C];
  // legacy (OPS_ACCn) kernels: the point itself, or component o.d of it
  __device__ __forceinline__ T &operator[](Off) const { return v[0]; }
  __device__ __forceinline__ T &operator[](OffMD o) const { return v[C == 1 ? 0 : o.d]; }
  template <class... I>
  __device__ __forceinline__ T &operator()(I... i) const {
    constexpr int n = sizeof...(I);
    static_assert(n == NDIM || n == NDIM + 1, "ACC: give NDIM offsets, or a component and NDIM offsets");
    if constexpr (n == NDIM || C == 1) {
      return v[0];
    } else {
      const int q[n] = {int(i)...};
      return v[q[0]];
    }
  }
};

// ------------------------------------------------------------ reductions
template <class T, int OP>
__device__ __forceinline__ T ident() {
  if constexpr (OP == OPS_INC) return T(0);
  else if constexpr (OP == OPS_MIN) {
    if constexpr (std::is_integral_v<T>) return (T)INT_MAX;
    else if constexpr (sizeof(T) == 8) return (T)DBL_MAX;
    else return (T)FLT_MAX;
  } else {
    if constexpr (std::is_integral_v<T>) return (T)-INT_MAX;
    else if constexpr (sizeof(T) == 8) return (T)-DBL_MAX;
    else return (T)-FLT_MAX;
  }
}

template <class T, int OP>
__device__ __forceinline__ T combine(T a, T b) {
  if constexpr (OP == OPS_INC) return a + b;
  else if constexpr (!std::is_integral_v<T>) return OP == OPS_MIN ? (T)fmin(a, b) : (T)fmax(a, b);
  else return OP == OPS_MIN ? (a < b ? a : b) : (a > b ? a : b);
}

// folds v over the workgroup and stores part[workgroup * D + q]; every
// thread of the workgroup calls it
template <class T, int OP, int D>
__device__ __forceinline__ void store_partials(T (&v)[D], T *part) {
  __shared__ T sh[kRows][D];
  const int lane = threadIdx.x, w = threadIdx.y;
#pragma unroll
  for (int q = 0; q < D; q++) {
    T s = v[q];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) s = combine<T, OP>(s, __shfl_xor(s, off, kWave));
    if (lane == 0) sh[w][q] = s;
  }
  __syncthreads();
  if (lane == 0 && w == 0) {
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int q = 0; q < D; q++) {
      T s = sh[0][q];
      for (int k = 1; k < kRows; k++) s = combine<T, OP>(s, sh[k][q]);
      part[(size_t)b * D + q] = s;
    }
  }
}

// ------------------------------------------------------------ point mapping
// workgroup b runs on XCD b % 8 and takes logical tile start(b % 8) + b / 8,
// XCD k owning a contiguous share of the nb tiles (a bijection on [0, nb))
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  const unsigned q8 = nb >> 3, r8 = nb & 7u, k = b & 7u;
  return k * q8 + (k < r8 ? k : r8) + (b >> 3);
}

// calls f(x, y, z) for this thread's points: OPSMI_GEN_XP points along x
// per row, 64 apart (ADJ false) or adjacent (ADJ true: with restrict-
// qualified dat arguments the compiler merges a lane's accesses of one array
// into one 16-byte access; a loop without reductions uses it, since the
// order of a reduction's per-thread sums follows the point layout and must
// not differ from that of the fused pair kernels, which keep the 64-apart
// form).  A loop writes a dat only at the point itself, so the points are
// independent and, with restrict-qualified arguments, the compiler overlaps
// their loads.
template <bool ADJ = false, class F>
__device__ __forceinline__ void for_each_point(const Loop &L, F &&f);
template <bool ADJ, class F>
__device__ __forceinline__ void for_each_point_as(const Loop &L, F &&f);

template <bool ADJ, class F>
__device__ __forceinline__ void for_each_point(const Loop &L, F &&f) {
  if (ADJ && L.spread) for_each_point_as<false>(L, f);
  else for_each_point_as<ADJ>(L, f);
}

template <bool ADJ, class F>
__device__ __forceinline__ void for_each_point_as(const Loop &L, F &&f) {
  constexpr int XP = OPSMI_GEN_XP;
  constexpr int STEP = ADJ ? 1 : kWave;  // x distance between a lane's points
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned l = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
  const unsigned tx = l % gridDim.x, rest = l / gridDim.x;
  const unsigned ty = rest % gridDim.y, tz = rest / gridDim.y;
  const int x = L.x0 + (int)tx * kWave * XP + (int)threadIdx.x * (ADJ ? XP : 1);
  if (x >= L.x1) return;
  // a lane whose points all lie in the range calls f for them with no branch
  // in between (a branch join makes the compiler wait for every load in
  // flight, which would serialise the points)
  const bool whole = x + (XP - 1) * STEP < L.x1;
  auto row = [&](int y, int z) {
    if (whole) {
#pragma unroll
      for (int k = 0; k < XP; k++) f(x + k * STEP, y, z);
    } else {
#pragma unroll
      for (int k = 0; k < XP; k++)
        if (x + k * STEP < L.x1) f(x + k * STEP, y, z);
    }
  };
  if (L.three_d) {
    const int y = L.y0 + (int)ty * kRows + (int)threadIdx.y;
    if (y >= L.y1) return;
    const int z0 = L.z0 + (int)tz * L.chunk;
    const int z1 = min(z0 + L.chunk, L.z1);
    for (int z = z0; z < z1; z++) row(y, z);
  } else {
    const int y0 = L.y0 + (int)ty * kRows * L.chunk;
    const int y1 = min(y0 + kRows * L.chunk, L.y1);
    for (int y = y0 + (int)threadIdx.y; y < y1; y += kRows) row(y, L.z0);
  }
}

}  // namespace opsmi_gen

// the legacy index macros (argument positions 0-31): offsets for the accessors
#define OPS_ACC0(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD0(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC1(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD1(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC2(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD2(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC3(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD3(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC4(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD4(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC5(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD5(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC6(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD6(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC7(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD7(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC8(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD8(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC9(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD9(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC10(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD10(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC11(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD11(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC12(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD12(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC13(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD13(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC14(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD14(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC15(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD15(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC16(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD16(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC17(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD17(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC18(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD18(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC19(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD19(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC20(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD20(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC21(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD21(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC22(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD22(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC23(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD23(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC24(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD24(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC25(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD25(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC26(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD26(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC27(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD27(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC28(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD28(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC29(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD29(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC30(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD30(...) (::opsmi_gen::OffMD{__VA_ARGS__})
#define OPS_ACC31(...) (::opsmi_gen::Off{__VA_ARGS__})
#define OPS_ACC_MD31(...) (::opsmi_gen::OffMD{__VA_ARGS__})
