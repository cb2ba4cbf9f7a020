// Random fills behind ops_randomgen_init / ops_fill_random_uniform /
// ops_fill_random_normal (ops/c/include/ops_lib_core.h:1302-1309).
//
// The reference draws from one std::mt19937 stream on the host
// (ops/c/src/core/ops_lib_core.cpp:2394-2494), so values depend on the
// backend and on call order.  Here every value is a pure function of
// (seed, fill number, point): the k-th fill after ops_randomgen_init uses
// stream k, and point (x,y,z) of the dat's extended (halo-including,
// unpadded) box maps to counter x' + W (y' + H z') with primed coordinates
// counted from the negative halo edge; component c of a dat with D
// components takes counter D i + c, in either layout (AoS or SoA), so a
// one-component dat draws the same values.  bits = mix(mix(seed ^ k G) + (i+1) G)
// with the SplitMix64 finaliser mix and G = 0x9e3779b97f4a7c15;
// uniform = (bits >> 11) 2^-53 in [0,1); normal = Box-Muller on the uniforms
// of counters 2i and 2i+1.  Padding columns are set to 0.
#include "engine.h"

using namespace opsmi;

namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(unsigned long long seed, unsigned long long stream, unsigned long long i) {
  const unsigned long long G = 0x9e3779b97f4a7c15ULL;
  unsigned long long b = mix64(mix64(seed ^ (stream * G)) + (i + 1) * G);
  return (double)(b >> 11) * (1.0 / 9007199254740992.0);
}

template <class T, bool NORMAL>
__global__ __launch_bounds__(256) void k_fill(T *base, long sx, long sy, long sz, long cs, int dim, int w, int wpad,
                                              int h, int dz, unsigned long long seed, unsigned long long stream) {
  const long n = (long)wpad * h * dz * dim;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int c = (int)(q % dim);
    const long pt = q / dim;
    const int xi = (int)(pt % wpad);
    const long rest = pt / wpad;
    const int yi = (int)(rest % h), zi = (int)(rest / h);
    T *p = base + (long)xi * sx + yi * sy + zi * sz + c * cs;
    if (xi >= w) {
      *p = T(0);
      continue;
    }
    const unsigned long long i =
        ((unsigned long long)xi + (unsigned long long)w * (yi + (unsigned long long)h * zi)) * dim + c;
    double v;
    if (NORMAL) {
      const double two_pi = 6.283185307179586476925286766559;
      double a = u01(seed, stream, 2 * i), b = u01(seed, stream, 2 * i + 1);
      v = sqrt(-2.0 * log(1.0 - a)) * cos(two_pi * b);
    } else {
      v = u01(seed, stream, i);
    }
    *p = (T)v;
  }
}

unsigned long long g_fill_count = 0;

void fill(ops_dat dat, bool normal) {
  require_init();
  flush_all();
  Instance &I = inst();
  if (!I.rng_init) fail(OPS_RUNTIME_CONFIGURATION_ERROR, "ops_randomgen_init has not been called");
  if (!dat) fail(OPS_INVALID_ARGUMENT, "ops_fill_random: NULL dat");
  DType dt = dtype_of(dat->type, dat->type_size);
  if (dt != DT_F64 && dt != DT_F32) fail(OPS_INVALID_ARGUMENT, "ops_fill_random: dat must be double or float");
  sync_to_device(dat);
  DatView v = view_of(dat);
  int nd = dat->block->dims;
  int wpad = dat->size[0], w = dat->size[0] - dat->x_pad;
  int h = nd > 1 ? dat->size[1] : 1, dz = nd > 2 ? dat->size[2] : 1;
  // start of the extended box (negative halo corner)
  char *start = v.base + ((long)dat->d_m[0] * v.sx + (nd > 1 ? dat->d_m[1] * v.sy : 0) +
                          (nd > 2 ? dat->d_m[2] * v.sz : 0)) *
                             dat->type_size;
  unsigned long long stream = g_fill_count++;
  const int dim = dat->dim;
  long n = (long)wpad * h * dz * dim;
  int nb = (int)std::min<long>((n + 255) / 256, (long)I.num_cus * 16);
  run_loop(normal ? "ops_fill_random_normal" : "ops_fill_random_uniform", (double)n * dat->type_size,
           [&](hipStream_t s) {
             if (dt == DT_F64) {
               if (normal) hipLaunchKernelGGL((k_fill<double, true>), dim3(nb), dim3(256), 0, s, (double *)start, v.sx, v.sy, v.sz, v.cs, dim, w, wpad, h, dz, (unsigned long long)I.rng_seed, stream);
               else hipLaunchKernelGGL((k_fill<double, false>), dim3(nb), dim3(256), 0, s, (double *)start, v.sx, v.sy, v.sz, v.cs, dim, w, wpad, h, dz, (unsigned long long)I.rng_seed, stream);
             } else {
               if (normal) hipLaunchKernelGGL((k_fill<float, true>), dim3(nb), dim3(256), 0, s, (float *)start, v.sx, v.sy, v.sz, v.cs, dim, w, wpad, h, dz, (unsigned long long)I.rng_seed, stream);
               else hipLaunchKernelGGL((k_fill<float, false>), dim3(nb), dim3(256), 0, s, (float *)start, v.sx, v.sy, v.sz, v.cs, dim, w, wpad, h, dz, (unsigned long long)I.rng_seed, stream);
             }
           });
  mark_device_dirty(dat);
  note_external_write(dat);
}

}  // namespace

extern "C" {

void ops_randomgen_init(unsigned int seed, int options) {
  OPSMI_API_BEGIN
  (void)options;
  require_init();
  inst().rng_seed = seed;
  inst().rng_init = true;
  g_fill_count = 0;
  OPSMI_API_END()
}

void ops_fill_random_uniform(ops_dat dat) {
  OPSMI_API_BEGIN
  fill(dat, false);
  OPSMI_API_END()
}

void ops_fill_random_normal(ops_dat dat) {
  OPSMI_API_BEGIN
  fill(dat, true);
  OPSMI_API_END()
}

void ops_randomgen_exit(void) { inst().rng_init = false; }

}  // extern "C"
