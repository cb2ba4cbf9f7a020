// Batched tridiagonal solves behind ops_tridMultiDimBatch[_Inc]
// (ops/c/include/ops_tridiag.h:45-83, ops/c/src/tridiag/ops_tridiag.cpp:
// 78-249).  The reference hands raw pointers to the external libtrid
// (tridDmtsvStridedBatch); this file is the backend's own solver.
//
// Thomas: one thread per system.  Lines are numbered so that consecutive
// threads take systems that are neighbours in x (for solves along y and z the
// loads of a wave are then contiguous).  The forward sweep keeps the modified
// super-diagonal c' and right-hand side d' in a step-major scratch array
// ([step][line]), so scratch traffic is coalesced for every direction.
// Arithmetic is the reciprocal form
//   r = 1/(b_i - a_i c'_{i-1});  d'_i = r (d_i - a_i d'_{i-1});  c'_i = r c_i
//   x_i = d'_i - c'_i x_{i+1}
// with contraction off, so results equal the CPU restatement bit for bit.
//
// PCR (strategy PCR): one workgroup per system held in LDS, log2(n) steps of
// parallel cyclic reduction (Hockney 1965); a_0 and c_{n-1} are taken as 0.
#include "engine.h"

using namespace opsmi;

struct ops_tridsolver_params_c {
  int strategy;
};

namespace {

struct LineGeo {
  long n;         // system length
  long stride;    // element stride along the solve dimension
  long count;     // number of systems
  int e0;         // extent of the first other dimension (fastest)
  long s0, s1;    // element strides of the two other dimensions
};

__device__ __forceinline__ long line_offset(const LineGeo &g, long id) {
  long i0 = id % g.e0, i1 = id / g.e0;
  return i0 * g.s0 + i1 * g.s1;
}

// The recurrence is serial per system but its loads are not: the passes
// run in blocks of P steps, and while a block's recurrence runs the next
// block's loads are already in flight (two register sets, the loop unrolled
// by two so each set has fixed registers), so a thread waits on memory once
// per pass rather than once per block.  Loads of the block past the end are
// clamped to the last step (in bounds, never used), so the loads and the
// recurrence of full blocks carry no branches (a branch join would make the
// compiler wait for every load in flight); the last, partial block runs with
// per-step conditions.  The arithmetic and its order are those of the plain
// loop.  Without _Inc, d' is written in place into d (the backward pass
// overwrites it with x), so only c' goes through the scratch; _Inc leaves d
// untouched and keeps d' in scratch too.  The forward pass reads a, b, c, d
// once, with the non-temporal hint.
constexpr int kPre = 8;

template <class T, int P>
struct FwdBuf {
  T a[P], b[P], c[P], d[P];
};
template <class T, int P>
struct BwdBuf {
  T c[P], d[P], u[P];
};

template <class T, bool INC, int P>
__global__ __launch_bounds__(256) void k_thomas(const T *__restrict__ a, const T *__restrict__ b,
                                                const T *__restrict__ c, T *__restrict__ d, T *__restrict__ u,
                                                T *__restrict__ cs, T *__restrict__ ds, LineGeo g) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= g.count) return;
  const long o = line_offset(g, id);
  const long s = g.stride, L = g.count, n = g.n;
  const T *ap = a + o, *bp = b + o, *cp = c + o;
  T *dp = d + o;
  T bb = (T)1.0 / bp[0];
  T cprev = bb * cp[0];
  T dprev = bb * dp[0];
  cs[id] = cprev;
  if (INC) ds[id] = dprev;
  else dp[0] = dprev;
  // ---- forward: steps 1 .. n-1
  auto fload = [&](long i0, FwdBuf<T, P> &f) {
#pragma unroll
    for (int k = 0; k < P; k++) {
      const long i = min(i0 + k, n - 1);
      f.a[k] = ldg<T, true>(ap + i * s);
      f.b[k] = ldg<T, true>(bp + i * s);
      f.c[k] = ldg<T, true>(cp + i * s);
      f.d[k] = ldg<T, true>(dp + i * s);
    }
  };
  auto fstep = [&](const FwdBuf<T, P> &f, int k, long i) {
    bb = (T)1.0 / (f.b[k] - f.a[k] * cprev);
    dprev = bb * (f.d[k] - f.a[k] * dprev);
    cprev = bb * f.c[k];
    cs[i * L + id] = cprev;
    if (INC) ds[i * L + id] = dprev;
    else dp[i * s] = dprev;
  };
  auto ffull = [&](const FwdBuf<T, P> &f, long i0) {
#pragma unroll
    for (int k = 0; k < P; k++) fstep(f, k, i0 + k);
  };
  auto ftail = [&](const FwdBuf<T, P> &f, long i0) {
#pragma unroll
    for (int k = 0; k < P; k++)
      if (i0 + k < n) fstep(f, k, i0 + k);
  };
  if (n > 1) {
    FwdBuf<T, P> X, Y;
    long i0 = 1;
    fload(i0, X);
    while (true) {
      if (i0 + P > n) {
        ftail(X, i0);
        break;
      }
      fload(i0 + P, Y);
      ffull(X, i0);
      i0 += P;
      if (i0 >= n) break;
      if (i0 + P > n) {
        ftail(Y, i0);
        break;
      }
      fload(i0 + P, X);
      ffull(Y, i0);
      i0 += P;
      if (i0 >= n) break;
    }
  }
  // ---- backward: x_{n-1} = d'_{n-1}, then steps n-2 .. 0
  T x = dprev;
  if (INC) u[o + (n - 1) * s] = u[o + (n - 1) * s] + x;
  else dp[(n - 1) * s] = x;
  auto bload = [&](long i1, BwdBuf<T, P> &w) {
#pragma unroll
    for (int k = 0; k < P; k++) {
      const long i = max(i1 - k, 0L);
      w.c[k] = cs[i * L + id];
      w.d[k] = INC ? ds[i * L + id] : dp[i * s];
      if (INC) w.u[k] = u[o + i * s];
    }
  };
  auto bstep = [&](const BwdBuf<T, P> &w, int k, long i) {
    x = w.d[k] - w.c[k] * x;
    if (INC) u[o + i * s] = w.u[k] + x;
    else dp[i * s] = x;
  };
  auto bfull = [&](const BwdBuf<T, P> &w, long i1) {
#pragma unroll
    for (int k = 0; k < P; k++) bstep(w, k, i1 - k);
  };
  auto btail = [&](const BwdBuf<T, P> &w, long i1) {
#pragma unroll
    for (int k = 0; k < P; k++)
      if (i1 - k >= 0) bstep(w, k, i1 - k);
  };
  if (n > 1) {
    BwdBuf<T, P> X, Y;
    long i1 = n - 2;
    bload(i1, X);
    while (true) {
      if (i1 - P + 1 < 0) {
        btail(X, i1);
        break;
      }
      bload(i1 - P, Y);
      bfull(X, i1);
      i1 -= P;
      if (i1 < 0) break;
      if (i1 - P + 1 < 0) {
        btail(Y, i1);
        break;
      }
      bload(i1 - P, X);
      bfull(Y, i1);
      i1 -= P;
      if (i1 < 0) break;
    }
  }
}

// Solves along x (element stride 1).  One thread per system would make each
// load of a wave touch 64 rows (64 cache lines); here a one-wave workgroup
// takes 64 systems and moves them through LDS in chunks of kChunk steps: the
// wave loads a chunk of a, b, c, d with row-contiguous V-wide loads (kChunk/V
// lanes per system), stores it transposed ([step][system]) in LDS, and each
// lane runs the same recurrence as k_thomas on its system from LDS, writing
// c' and d' to the step-major scratch.  The next chunk's loads are issued
// before the recurrence runs.  The backward pass reads c', d' per step
// (coalesced), puts a chunk of x in LDS and writes it (or adds it to u) with
// row-contiguous stores.  Same operations in the same order as k_thomas.
constexpr int kChunk = 16;

template <class T, int V>
struct XChunk {
  static constexpr int kLanesPerSys = kChunk / V;      // lanes covering one system's chunk
  static constexpr int kSysPerLoad = 64 / kLanesPerSys;  // systems per wave-wide load
  static constexpr int kLoads = 64 / kSysPerLoad;        // loads per array per chunk
};

template <class T, int V, bool INC>
__global__ __launch_bounds__(64) void k_thomas_x(const T *__restrict__ a, const T *__restrict__ b,
                                                 const T *__restrict__ c, T *__restrict__ d, T *__restrict__ u,
                                                 T *__restrict__ cs, T *__restrict__ ds, LineGeo g) {
  using XC = XChunk<T, V>;
  __shared__ T LA[kChunk][65], LB[kChunk][65], LC[kChunk][65], LD[kChunk][65];
  const int lane = (int)threadIdx.x;
  const long first = (long)blockIdx.x * 64;
  const long id = first + lane;  // this lane's system in the recurrence
  const long L = g.count, n = g.n;
  // the systems and chunk positions this lane covers in row-contiguous loads
  const int sub = lane % XC::kLanesPerSys, sys0 = lane / XC::kLanesPerSys;
  long rowoff[XC::kLoads];
  bool rowok[XC::kLoads];
#pragma unroll
  for (int k = 0; k < XC::kLoads; k++) {
    const long sid = first + sys0 + k * XC::kSysPerLoad;
    rowok[k] = sid < L;
    rowoff[k] = rowok[k] ? line_offset(g, sid) : 0;
  }
  T ra[XC::kLoads][V], rb[XC::kLoads][V], rc[XC::kLoads][V], rd[XC::kLoads][V];
  // every system of the workgroup exists (workgroup-uniform)
  const bool all_rows = first + 64 <= L;
  auto load_chunk = [&](long i0) {
    if (all_rows && i0 + kChunk <= n) {
      // a full chunk: no per-row branches, so the loads of all rows are in
      // flight together (a branch join makes the compiler wait for every
      // outstanding load, which serialised the rows: 241 us along x at 256^3)
#pragma unroll
      for (int k = 0; k < XC::kLoads; k++) {
        using VT = vec_t<T, V>;
        const long i = i0 + sub * V;
        const VT va = ldg<VT, true>((const VT *)(a + rowoff[k] + i));
        const VT vb = ldg<VT, true>((const VT *)(b + rowoff[k] + i));
        const VT vc = ldg<VT, true>((const VT *)(c + rowoff[k] + i));
        const VT vd = ldg<VT, true>((const VT *)(d + rowoff[k] + i));
#pragma unroll
        for (int j = 0; j < V; j++) {
          ra[k][j] = va[j];
          rb[k][j] = vb[j];
          rc[k][j] = vc[j];
          rd[k][j] = vd[j];
        }
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < XC::kLoads; k++) {
      const long i = i0 + sub * V;
      if (rowok[k] && i + V <= n) {
        using VT = vec_t<T, V>;
        const VT va = ldg<VT, true>((const VT *)(a + rowoff[k] + i));
        const VT vb = ldg<VT, true>((const VT *)(b + rowoff[k] + i));
        const VT vc = ldg<VT, true>((const VT *)(c + rowoff[k] + i));
        const VT vd = ldg<VT, true>((const VT *)(d + rowoff[k] + i));
#pragma unroll
        for (int j = 0; j < V; j++) {
          ra[k][j] = va[j];
          rb[k][j] = vb[j];
          rc[k][j] = vc[j];
          rd[k][j] = vd[j];
        }
      } else {
#pragma unroll
        for (int j = 0; j < V; j++) {
          const bool ok = rowok[k] && i + j < n;
          ra[k][j] = ok ? a[rowoff[k] + i + j] : (T)0;
          rb[k][j] = ok ? b[rowoff[k] + i + j] : (T)1;
          rc[k][j] = ok ? c[rowoff[k] + i + j] : (T)0;
          rd[k][j] = ok ? d[rowoff[k] + i + j] : (T)0;
        }
      }
    }
  };
  auto stage_chunk = [&]() {
#pragma unroll
    for (int k = 0; k < XC::kLoads; k++) {
      const int sy = sys0 + k * XC::kSysPerLoad;
#pragma unroll
      for (int j = 0; j < V; j++) {
        LA[sub * V + j][sy] = ra[k][j];
        LB[sub * V + j][sy] = rb[k][j];
        LC[sub * V + j][sy] = rc[k][j];
        LD[sub * V + j][sy] = rd[k][j];
      }
    }
  };
  const bool live = id < L;
  T cprev = 0, dprev = 0;
  load_chunk(0);
  for (long i0 = 0; i0 < n; i0 += kChunk) {
    stage_chunk();
    __syncthreads();
    if (i0 + kChunk < n) load_chunk(i0 + kChunk);
    if (live && i0 > 0 && i0 + kChunk <= n) {
      // an inner full chunk: no per-step branches (their joins would make
      // the compiler wait for the next chunk's loads already in flight)
#pragma unroll
      for (int k = 0; k < kChunk; k++) {
        const long i = i0 + k;
        const T av = LA[k][lane], bv = LB[k][lane], cv = LC[k][lane], dv = LD[k][lane];
        const T bb = (T)1.0 / (bv - av * cprev);
        dprev = bb * (dv - av * dprev);
        cprev = bb * cv;
        cs[i * L + id] = cprev;
        ds[i * L + id] = dprev;
      }
    } else if (live) {
#pragma unroll
      for (int k = 0; k < kChunk; k++) {
        const long i = i0 + k;
        if (i < n) {
          const T av = LA[k][lane], bv = LB[k][lane], cv = LC[k][lane], dv = LD[k][lane];
          if (i == 0) {
            const T bb = (T)1.0 / bv;
            cprev = bb * cv;
            dprev = bb * dv;
          } else {
            const T bb = (T)1.0 / (bv - av * cprev);
            dprev = bb * (dv - av * dprev);
            cprev = bb * cv;
          }
          cs[i * L + id] = cprev;
          ds[i * L + id] = dprev;
        }
      }
    }
    __syncthreads();
  }
  // backward: chunks from the last one down; x of a chunk goes to LA.  The
  // c', d' of the next chunk are loaded while this one is solved (two
  // register sets used in turn).
  T x = dprev;
  const long last0 = (n - 1) / kChunk * kChunk;
  T cv[2][kChunk], dv[2][kChunk];
  auto load_back = [&](int q, long i0) {
    if (!live || i0 < 0) return;
#pragma unroll
    for (int k = 0; k < kChunk; k++) {
      const long i = i0 + k;
      if (i < n - 1) {
        cv[q][k] = cs[i * L + id];
        dv[q][k] = ds[i * L + id];
      }
    }
  };
  load_back(0, last0);
  int q = 0;
  for (long i0 = last0; i0 >= 0; i0 -= kChunk, q ^= 1) {
    if (q == 0) load_back(1, i0 - kChunk);
    else load_back(0, i0 - kChunk);
    if (live) {
#pragma unroll
      for (int k = kChunk - 1; k >= 0; k--) {
        const long i = i0 + k;
        const T cq = q == 0 ? cv[0][k] : cv[1][k];
        const T dq = q == 0 ? dv[0][k] : dv[1][k];
        if (i == n - 1) LA[k][lane] = x;
        else if (i < n - 1) {
          x = dq - cq * x;
          LA[k][lane] = x;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < XC::kLoads; k++) {
      const long i = i0 + sub * V;
      const int sy = sys0 + k * XC::kSysPerLoad;
      if (!rowok[k]) continue;
      if (i + V <= n) {
        using VT = vec_t<T, V>;
        VT v;
#pragma unroll
        for (int j = 0; j < V; j++) v[j] = LA[sub * V + j][sy];
        if (INC) {
          const VT uo = *(const VT *)(u + rowoff[k] + i);
#pragma unroll
          for (int j = 0; j < V; j++) v[j] = uo[j] + v[j];
          *(VT *)(u + rowoff[k] + i) = v;
        } else {
          *(VT *)(d + rowoff[k] + i) = v;
        }
      } else {
#pragma unroll
        for (int j = 0; j < V; j++)
          if (i + j < n) {
            const T xv = LA[sub * V + j][sy];
            if (INC) u[rowoff[k] + i + j] = u[rowoff[k] + i + j] + xv;
            else d[rowoff[k] + i + j] = xv;
          }
      }
    }
    __syncthreads();
  }
}

template <class T, bool INC>
__global__ __launch_bounds__(1024) void k_pcr(const T *__restrict__ a, const T *__restrict__ b,
                                              const T *__restrict__ c, T *__restrict__ d, T *__restrict__ u,
                                              LineGeo g) {
  extern __shared__ unsigned char smem[];
  T *A = (T *)smem;
  T *B = A + 2 * g.n;
  T *C = B + 2 * g.n;
  T *D = C + 2 * g.n;
  const long o = line_offset(g, blockIdx.x);
  const int n = (int)g.n;
  const int i = threadIdx.x;
  const long s = g.stride;
  if (i < n) {
    A[i] = i == 0 ? (T)0 : a[o + i * s];
    B[i] = b[o + i * s];
    C[i] = i == n - 1 ? (T)0 : c[o + i * s];
    D[i] = d[o + i * s];
  }
  __syncthreads();
  int cur = 0;
  for (int st = 1; st < n; st *= 2) {
    const int nx = cur ^ n;  // offset of the other buffer half (0 or n)
    if (i < n) {
      const T *Ac = A + cur, *Bc = B + cur, *Cc = C + cur, *Dc = D + cur;
      const bool lo = i - st >= 0, hi = i + st < n;
      const T k1 = lo ? Ac[i] / Bc[i - st] : (T)0;
      const T k2 = hi ? Cc[i] / Bc[i + st] : (T)0;
      T bv = Bc[i], dv = Dc[i];
      if (lo) {
        bv = bv - Cc[i - st] * k1;
        dv = dv - Dc[i - st] * k1;
      }
      if (hi) {
        bv = bv - Ac[i + st] * k2;
        dv = dv - Dc[i + st] * k2;
      }
      A[nx + i] = lo ? -Ac[i - st] * k1 : (T)0;
      C[nx + i] = hi ? -Cc[i + st] * k2 : (T)0;
      B[nx + i] = bv;
      D[nx + i] = dv;
    }
    __syncthreads();
    cur = nx;
  }
  if (i < n) {
    const T x = D[cur + i] / B[cur + i];
    if (INC) u[o + i * s] = u[o + i * s] + x;
    else d[o + i * s] = x;
  }
}

struct TridScratch {
  void *p = nullptr;
  size_t bytes = 0;
};
TridScratch g_scratch;

void *trid_scratch(size_t bytes) {
  if (bytes > g_scratch.bytes) {
    if (g_scratch.p) {
      OPSMI_HIP(hipStreamSynchronize(inst().stream));
      OPSMI_HIP(hipFree(g_scratch.p));
    }
    OPSMI_HIP(hipMalloc(&g_scratch.p, bytes));
    g_scratch.bytes = bytes;
  }
  return g_scratch.p;
}

// Scratch of the entry point that allocates nothing
// (ops_tridMultiDimBatch_noalloc): c' and d' of every system, step-major, as
// the library scratch above.  Each GPU holds its own copy for the life of the
// process; host code reaches it only through the address that
// hipGetSymbolAddress returned on that device (the array's name in host code
// is not a device address).  Calls on one GPU share it, so calls that overlap
// on one GPU race.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
__device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
char *g_noalloc_ptr[kMaxDevices];

// trid_run with fixed_scr set: no flush, no library scratch, no timing events
// or statistics records, so nothing is allocated on the host or the device.
void trid_run(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d, ops_dat u,
              ops_tridsolver_params_c *ctx, char *fixed_scr = nullptr) {
  require_init();
  if (!fixed_scr) flush_all();
  else if (!queue_empty())
    fail(OPS_RUNTIME_ERROR, "ops_tridMultiDimBatch_noalloc: loops are queued; call ops_execute() first "
                            "(running them may allocate)");
  if (!a || !b || !c || !d || !range || !ctx) fail(OPS_INVALID_ARGUMENT, "Tridsolver error: NULL argument");
  if (ndim < 1 || ndim > 3 || ndim != a->block->dims)
    fail(OPS_INVALID_ARGUMENT, "Tridsolver error: ndim must match the block (1-3D)");
  if (solvedim < 0 || solvedim >= ndim) fail(OPS_INVALID_ARGUMENT, "Tridsolver error: solvedim out of range");
  ops_dat all[5] = {a, b, c, d, u};
  int nd = u ? 5 : 4;
  // ops/c/src/tridiag/ops_tridiag.cpp:91-101
  for (int i = 0; i < ndim; i++)
    for (int k = 1; k < nd; k++)
      if (all[k]->size[i] != a->size[i])
        fail(OPS_RUNTIME_ERROR, "Tridsolver error: the a,b,c,d datasets all need to be the same size");
  for (int k = 1; k < nd; k++)
    if (strcmp(all[k]->type, a->type) != 0)
      fail(OPS_RUNTIME_ERROR, "Tridsolver error: the a,b,c,d datasets must all be of the same type");
  DType dt = dtype_of(a->type, a->type_size);
  if (dt != DT_F64 && dt != DT_F32)
    fail(OPS_RUNTIME_ERROR, "Tridsolver error: unsupported type. Dataset type must either be \"double\" or \"float\"");
  for (int k = 0; k < nd; k++) {
    if (all[k]->dim != 1) fail(OPS_NOT_IMPLEMENTED, "Tridsolver error: multi-component dats are not supported");
    if (all[k]->locked_hd) fail(OPS_RUNTIME_ERROR, cat("Tridsolver error: dat ", all[k]->name, " is locked"));
  }
  int dims[3] = {1, 1, 1};
  long offset = 0;
  long st[3] = {1, a->size[0], (long)a->size[0] * a->size[1]};
  for (int i = 0; i < ndim; i++) {
    dims[i] = range[2 * i + 1] - range[2 * i];
    int first = a->d_m[i] + a->base[i], last = a->size[i] + a->d_m[i] + a->base[i];
    if (dims[i] < 0 || range[2 * i] < first || range[2 * i + 1] > last)
      fail(OPS_INVALID_ARGUMENT, cat("Tridsolver error: range outside the dats in dimension ", i));
    offset += (long)range[2 * i] * st[i];
  }
  for (int k = 0; k < nd; k++) {
    if (fixed_scr && all[k]->dirty_hd == 1)
      fail(OPS_RUNTIME_ERROR, cat("ops_tridMultiDimBatch_noalloc: dat ", all[k]->name,
                                  " has newer data on the host; upload it first (ops_execute, or any loop)"));
    sync_to_device(all[k]);
  }
  LineGeo g;
  g.n = dims[solvedim];
  g.stride = st[solvedim];
  int od[2], q = 0;
  for (int i = 0; i < 3; i++)
    if (i != solvedim) od[q++] = i;
  g.e0 = dims[od[0]];
  g.s0 = st[od[0]];
  g.s1 = st[od[1]];
  g.count = (long)dims[od[0]] * dims[od[1]];
  // statistics per solve direction (the timing table's loop name)
  static const char *names[2][3] = {{"ops_tridMultiDimBatch x", "ops_tridMultiDimBatch y", "ops_tridMultiDimBatch z"},
                                    {"ops_tridMultiDimBatch_Inc x", "ops_tridMultiDimBatch_Inc y",
                                     "ops_tridMultiDimBatch_Inc z"}};
  const char *nm = names[u ? 1 : 0][solvedim];
  double bytes = (double)g.n * g.count * a->type_size * (u ? 6.0 : 5.0);
  if (g.n <= 0 || g.count <= 0) {
    if (!fixed_scr) run_loop(nm, 0.0, [](hipStream_t) {});
    return;
  }
  if (fixed_scr && (double)g.n * g.count * a->type_size * 2.0 > (double)kNoAllocScratchBytes)
    fail(OPS_INVALID_ARGUMENT,
         cat("ops_tridMultiDimBatch_noalloc: ", g.count, " systems of ", g.n, " rows need ",
             2 * g.n * g.count * (long)a->type_size, " bytes of scratch; the maximum is ", kNoAllocScratchBytes,
             " (", kNoAllocScratchBytes / 16, " points in double, ", kNoAllocScratchBytes / 8, " in float)"));
  // the launch: through run_loop (statistics, timing events) or, for the
  // no-allocation entry point, straight onto the library's stream
  auto launch = [&](const std::function<void(hipStream_t)> &fn) {
    if (fixed_scr) {
      fn(inst().stream);
      OPSMI_HIP(hipGetLastError());
    } else {
      run_loop(nm, bytes, fn);
    }
  };
  size_t es = a->type_size;
  char *pa = a->data_d + a->base_offset + offset * es, *pb = b->data_d + b->base_offset + offset * es;
  char *pc = c->data_d + c->base_offset + offset * es, *pd = d->data_d + d->base_offset + offset * es;
  char *pu = u ? u->data_d + u->base_offset + offset * es : nullptr;
  bool pcr = ctx && ctx->strategy == 3;
  if (pcr) {
    if (g.n > 1024) fail(OPS_NOT_IMPLEMENTED, "Tridsolver error: PCR supports systems of at most 1024 rows");
    int threads = (int)((g.n + 63) / 64 * 64);
    size_t sm = 8 * g.n * es;
    launch([&](hipStream_t s) {
      if (dt == DT_F64) {
        if (u) hipLaunchKernelGGL((k_pcr<double, true>), dim3(g.count), dim3(threads), sm, s, (double *)pa, (double *)pb, (double *)pc, (double *)pd, (double *)pu, g);
        else hipLaunchKernelGGL((k_pcr<double, false>), dim3(g.count), dim3(threads), sm, s, (double *)pa, (double *)pb, (double *)pc, (double *)pd, (double *)pu, g);
      } else {
        if (u) hipLaunchKernelGGL((k_pcr<float, true>), dim3(g.count), dim3(threads), sm, s, (float *)pa, (float *)pb, (float *)pc, (float *)pd, (float *)pu, g);
        else hipLaunchKernelGGL((k_pcr<float, false>), dim3(g.count), dim3(threads), sm, s, (float *)pa, (float *)pb, (float *)pc, (float *)pd, (float *)pu, g);
      }
    });
  } else {
    size_t sb = (size_t)g.n * g.count * es;
    char *scr = fixed_scr ? fixed_scr : (char *)trid_scratch(2 * sb);
    int nblk = (int)((g.count + 255) / 256);
    if (solvedim == 0) {
      // x solves: LDS-staged chunks, 16-byte row loads when every array's
      // rows start 16-byte aligned
      bool v16 = g.s0 % 2 == 0 && g.s1 % 2 == 0;
      for (char *p : {pa, pb, pc, pd, pu})
        if (p && (uintptr_t)p % 16 != 0) v16 = false;
      if (dt == DT_F32) v16 = false;
      const int nw = (int)((g.count + 63) / 64);
      launch([&](hipStream_t s) {
        if (dt == DT_F64) {
          double *A = (double *)pa, *B = (double *)pb, *C = (double *)pc, *D = (double *)pd, *U = (double *)pu;
          double *CS = (double *)scr, *DS = (double *)(scr + sb);
          if (v16) {
            if (u) hipLaunchKernelGGL((k_thomas_x<double, 2, true>), dim3(nw), dim3(64), 0, s, A, B, C, D, U, CS, DS, g);
            else hipLaunchKernelGGL((k_thomas_x<double, 2, false>), dim3(nw), dim3(64), 0, s, A, B, C, D, U, CS, DS, g);
          } else {
            if (u) hipLaunchKernelGGL((k_thomas_x<double, 1, true>), dim3(nw), dim3(64), 0, s, A, B, C, D, U, CS, DS, g);
            else hipLaunchKernelGGL((k_thomas_x<double, 1, false>), dim3(nw), dim3(64), 0, s, A, B, C, D, U, CS, DS, g);
          }
        } else {
          float *A = (float *)pa, *B = (float *)pb, *C = (float *)pc, *D = (float *)pd, *U = (float *)pu;
          float *CS = (float *)scr, *DS = (float *)(scr + sb);
          if (u) hipLaunchKernelGGL((k_thomas_x<float, 1, true>), dim3(nw), dim3(64), 0, s, A, B, C, D, U, CS, DS, g);
          else hipLaunchKernelGGL((k_thomas_x<float, 1, false>), dim3(nw), dim3(64), 0, s, A, B, C, D, U, CS, DS, g);
        }
      });
      mark_device_dirty(u ? u : d);
      note_external_write(u ? u : d);
      return;
    }
    launch([&](hipStream_t s) {
      if (dt == DT_F64) {
        if (u) hipLaunchKernelGGL((k_thomas<double, true, kPre>), dim3(nblk), dim3(256), 0, s, (double *)pa, (double *)pb, (double *)pc, (double *)pd, (double *)pu, (double *)scr, (double *)(scr + sb), g);
        else hipLaunchKernelGGL((k_thomas<double, false, kPre>), dim3(nblk), dim3(256), 0, s, (double *)pa, (double *)pb, (double *)pc, (double *)pd, (double *)pu, (double *)scr, (double *)(scr + sb), g);
      } else {
        if (u) hipLaunchKernelGGL((k_thomas<float, true, kPre>), dim3(nblk), dim3(256), 0, s, (float *)pa, (float *)pb, (float *)pc, (float *)pd, (float *)pu, (float *)scr, (float *)(scr + sb), g);
        else hipLaunchKernelGGL((k_thomas<float, false, kPre>), dim3(nblk), dim3(256), 0, s, (float *)pa, (float *)pb, (float *)pc, (float *)pd, (float *)pu, (float *)scr, (float *)(scr + sb), g);
      }
    });
  }
  mark_device_dirty(u ? u : d);
  note_external_write(u ? u : d);
}

}  // namespace

namespace opsmi {
// Called by ops_init with the library's device current: the device address
// of this device's copy of the no-allocation scratch.
void trid_noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) fail(OPS_RUNTIME_CONFIGURATION_ERROR, cat("device ", dev, " out of range"));
  if (g_noalloc_ptr[dev]) return;
  void *ptr = nullptr;
  OPSMI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(OPS_RUNTIME_ERROR, "hipGetSymbolAddress returned no address for the tridiagonal scratch");
  g_noalloc_ptr[dev] = (char *)ptr;
}
}  // namespace opsmi

extern "C" {

ops_tridsolver_params_c *ops_tridsolver_params_create(ops_block block, int strategy) {
  (void)block;
  ops_tridsolver_params_c *p = new ops_tridsolver_params_c;
  p->strategy = strategy;
  return p;
}

void ops_tridsolver_params_destroy(ops_tridsolver_params_c *p) { delete p; }

void ops_tridMultiDimBatch_c(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                             ops_tridsolver_params_c *ctx) {
  OPSMI_API_BEGIN
  trid_run(ndim, solvedim, range, a, b, c, d, nullptr, ctx);
  OPSMI_API_END()
}

void ops_tridMultiDimBatch_Inc_c(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                                 ops_dat u, ops_tridsolver_params_c *ctx) {
  OPSMI_API_BEGIN
  if (!u) fail(OPS_INVALID_ARGUMENT, "Tridsolver error: NULL u");
  trid_run(ndim, solvedim, range, a, b, c, d, u, ctx);
  OPSMI_API_END()
}

// The reference's C-linkage entry points: the params object's one data
// member points at the solver settings.
void ops_tridMultiDimBatch(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                           ops_tridsolver_params *tridsolver_ctx) {
  OPSMI_API_BEGIN
  if (!tridsolver_ctx) fail(OPS_INVALID_ARGUMENT, "Tridsolver error: NULL ops_tridsolver_params");
  trid_run(ndim, solvedim, range, a, b, c, d, nullptr, *(ops_tridsolver_params_c **)tridsolver_ctx);
  OPSMI_API_END()
}

void ops_tridMultiDimBatch_Inc(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                               ops_dat u, ops_tridsolver_params *tridsolver_ctx) {
  OPSMI_API_BEGIN
  if (!tridsolver_ctx) fail(OPS_INVALID_ARGUMENT, "Tridsolver error: NULL ops_tridsolver_params");
  if (!u) fail(OPS_INVALID_ARGUMENT, "Tridsolver error: NULL u");
  trid_run(ndim, solvedim, range, a, b, c, d, u, *(ops_tridsolver_params_c **)tridsolver_ctx);
  OPSMI_API_END()
}

// ops_tridMultiDimBatch (u NULL) or ops_tridMultiDimBatch_Inc (u given) on
// the same kernels, with the c'/d' scratch in the library's fixed device
// array instead of library-allocated memory: at most 64 MiB of scratch,
// i.e. 2 * rows * systems * sizeof(element) <= 64 MiB.
void ops_tridMultiDimBatch_noalloc(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
                                   ops_dat u, ops_tridsolver_params *tridsolver_ctx) {
  OPSMI_API_BEGIN
  if (!tridsolver_ctx) fail(OPS_INVALID_ARGUMENT, "Tridsolver error: NULL ops_tridsolver_params");
  require_init();
  const int dev = inst().device;
  char *ptr = dev >= 0 && dev < kMaxDevices ? g_noalloc_ptr[dev] : nullptr;
  if (!ptr) fail(OPS_RUNTIME_ERROR, "ops_tridMultiDimBatch_noalloc: scratch address not set up (ops_init)");
  trid_run(ndim, solvedim, range, a, b, c, d, u, *(ops_tridsolver_params_c **)tridsolver_ctx, ptr);
  OPSMI_API_END()
}

}  // extern "C"
