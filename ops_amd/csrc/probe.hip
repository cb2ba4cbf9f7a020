// Memory-stream probes: the read/write mixes of this path's kernels without
// their arithmetic or neighbour reads.  Each point reads NR arrays and
// writes NW arrays (one 16-byte vector per lane per array, one pair of
// points per thread, short-lived waves in the engine's XCD-aware order), so
// the time of a launch is the practical HBM ceiling for that mix on this
// card; the stencil kernels are measured against the mix they move.
#include "engine.h"

using namespace opsmi;

namespace {

constexpr int kMaxR = 8, kMaxW = 4;

struct StreamArgs {
  const double *in[kMaxR];
  double *out[kMaxW];
  long n;        // points per array (a multiple of 2)
  double never;  // NW == 0: the sum is stored only where it equals this (no point does)
};

template <int NR, int NW, bool NTL, bool NTS>
__global__ __launch_bounds__(256) void k_stream(StreamArgs a) {
  using V2 = vec_t<double, 2>;
  const long i = ((long)xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x) * 2;
  if (i >= a.n) return;
  V2 s = {1.0, 2.0};
#pragma unroll
  for (int r = 0; r < NR; r++) s += ldg<V2, NTL>(reinterpret_cast<const V2 *>(a.in[r] + i));
#pragma unroll
  for (int w = 0; w < NW; w++) {
    const V2 v = s + (double)w;
    if constexpr (NTS) __builtin_nontemporal_store(v, reinterpret_cast<V2 *>(a.out[w] + i));
    else *reinterpret_cast<V2 *>(a.out[w] + i) = v;
  }
  // a comparison of the sum keeps every load (a store under a flag alone let
  // the compiler move the loads into the never-taken branch)
  if constexpr (NW == 0)
    if (s.x == a.never || s.y == a.never) *reinterpret_cast<V2 *>(a.out[0] + i) = s;
}

template <int NR, int NW>
void launch_mix(bool nt, unsigned grid, hipStream_t st, const StreamArgs &a) {
  if (nt) hipLaunchKernelGGL((k_stream<NR, NW, true, true>), dim3(grid), dim3(256), 0, st, a);
  else hipLaunchKernelGGL((k_stream<NR, NW, false, false>), dim3(grid), dim3(256), 0, st, a);
}

// the mixes of this path's kernels
bool launch(int nr, int nw, bool nt, unsigned grid, hipStream_t st, const StreamArgs &a) {
#define OPSMI_MIX(R, W)                    \
  if (nr == R && nw == W) {                \
    launch_mix<R, W>(nt, grid, st, a);     \
    return true;                           \
  }
  OPSMI_MIX(1, 0) OPSMI_MIX(4, 0) OPSMI_MIX(0, 1) OPSMI_MIX(1, 1) OPSMI_MIX(1, 2) OPSMI_MIX(2, 1)
  OPSMI_MIX(3, 2) OPSMI_MIX(4, 1) OPSMI_MIX(6, 3) OPSMI_MIX(4, 2) OPSMI_MIX(5, 1)
#undef OPSMI_MIX
  return false;
}

}  // namespace

extern "C" {

// Reads nread and writes nwrite arrays of `points` doubles, `reps` times
// after one untimed launch; *avg_us = the average launch time (HIP events on
// the library's stream).  nt: non-temporal loads and stores.  Mixes:
// (1,0) (4,0) (0,1) (1,1) (1,2) (2,1) (3,2) (4,1) (6,3) (4,2) (5,1).
// Allocates its arrays for the call and frees them after.
int ops_mi_stream_probe(int nread, int nwrite, int nt, long points, int reps, double *avg_us) {
  OPSMI_API_BEGIN
  require_init();
  flush_all();
  if (!avg_us || reps < 1 || points < 2 || nread < 0 || nread > kMaxR || nwrite < 0 || nwrite > kMaxW ||
      nread + nwrite == 0)
    fail(OPS_INVALID_ARGUMENT, "ops_mi_stream_probe: bad arguments");
  const long n = points / 2 * 2;
  const size_t bytes = (size_t)n * sizeof(double);
  StreamArgs a{};
  a.n = n;
  a.never = -1.0;  // the arrays hold zeros, so every sum is (1, 2)
  std::vector<void *> bufs;
  auto alloc = [&]() {
    void *p = nullptr;
    OPSMI_HIP(hipMalloc(&p, bytes));
    bufs.push_back(p);
    OPSMI_HIP(hipMemsetAsync(p, 0, bytes, inst().stream));
    return p;
  };
  struct Free {
    std::vector<void *> &b;
    ~Free() {
      (void)hipDeviceSynchronize();
      for (void *p : b) (void)hipFree(p);
    }
  } guard{bufs};
  for (int r = 0; r < nread; r++) a.in[r] = (const double *)alloc();
  for (int w = 0; w < std::max(nwrite, 1); w++) a.out[w] = (double *)alloc();
  const unsigned grid = (unsigned)((n / 2 + 255) / 256);
  hipStream_t st = inst().stream;
  if (!launch(nread, nwrite, nt != 0, grid, st, a))
    fail(OPS_INVALID_ARGUMENT, cat("ops_mi_stream_probe: no kernel for ", nread, " reads and ", nwrite, " writes"));
  OPSMI_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  OPSMI_HIP(hipEventCreate(&e0));
  OPSMI_HIP(hipEventCreate(&e1));
  OPSMI_HIP(hipEventRecord(e0, st));
  for (int k = 0; k < reps; k++) launch(nread, nwrite, nt != 0, grid, st, a);
  OPSMI_HIP(hipEventRecord(e1, st));
  OPSMI_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  OPSMI_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_us = (double)ms * 1e3 / reps;
  return 0;
  OPSMI_API_END(-1)
}

}  // extern "C"
