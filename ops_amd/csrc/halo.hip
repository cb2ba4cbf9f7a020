// Multi-block halo exchange behind ops_halo_transfer
// (ops/c/include/ops_lib_core.h:888-911; semantics of the host path
// ops/c/src/sequential/ops_host_singlenode.cpp:172-446).
//
// A halo copies an iter_size box from one dat to another through a linear
// buffer.  Buffer dimension q has extent iter_size[q]; dat dimension i walks
// buffer dimension |dir[i]|-1, forwards when dir[i] > 0 and backwards when
// dir[i] < 0, so blocks may meet with permuted or flipped axes.  Values pass
// through the narrower of the two element types when a float and a double
// dat meet.  Each halo is one pack kernel (source -> buffer) and one unpack
// kernel (buffer -> target), so overlapping source and target regions behave
// as in the reference.
#include "engine.h"

using namespace opsmi;

namespace {

struct HaloSide {
  char *data;          // start of the dat's array (raw index 0)
  long size0, size1;   // raw extents of dims 0, 1
  long pstride;        // bytes between neighbouring points of one component
  long cstride;        // bytes between the components of a point
  int tsize;           // bytes per component
  int dim;             // components
  int first[3];        // raw index of the first box point per dat dimension
  int q[3];            // buffer dimension walked by each dat dimension
  int sgn[3];          // +1 forward, -1 backward
};

struct HaloBox {
  int ext[3];  // buffer extents
};

__device__ __forceinline__ long side_offset(const HaloSide &s, const int *b, const HaloBox &box, int d) {
  long r[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const int pos = b[s.q[i]];
    r[i] = s.sgn[i] > 0 ? s.first[i] + pos : s.first[i] + (box.ext[s.q[i]] - 1 - pos);
  }
  return (r[0] + r[1] * s.size0 + r[2] * s.size0 * s.size1) * s.pstride + (long)d * s.cstride;
}

// storage: 8 or 4 bytes per buffer value
template <bool PACK>
__global__ __launch_bounds__(256) void k_halo(HaloSide s, HaloBox box, char *buf, int storage) {
  const long n = (long)box.ext[0] * box.ext[1] * box.ext[2] * s.dim;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int d = (int)(t % s.dim);
    long rest = t / s.dim;
    int b[3];
    b[0] = (int)(rest % box.ext[0]);
    rest /= box.ext[0];
    b[1] = (int)(rest % box.ext[1]);
    b[2] = (int)(rest / box.ext[1]);
    char *p = s.data + side_offset(s, b, box, d);
    char *q = buf + t * storage;
    if (PACK) {
      if (storage == 8) {
        *(double *)q = s.tsize == 8 ? *(double *)p : (double)*(float *)p;
      } else {
        *(float *)q = s.tsize == 4 ? *(float *)p : (float)*(double *)p;
      }
    } else {
      if (s.tsize == 8) *(double *)p = storage == 8 ? *(double *)q : (double)*(float *)q;
      else *(float *)p = storage == 4 ? *(float *)q : (float)*(double *)q;
    }
  }
}

HaloSide make_side(ops_dat dat, const int *base, const int *dir, const int *iter) {
  HaloSide s;
  s.data = dat->data_d;
  s.size0 = dat->size[0];
  s.size1 = dat->size[1];
  // AoS: a point's components are adjacent; SoA (ops_set_soa): each
  // component is one array (ops/c/src/sequential/ops_host_singlenode.cpp:124-131)
  const DatView v = view_of(dat);
  s.pstride = v.sx * dat->type_size;
  s.cstride = v.cs * dat->type_size;
  s.tsize = dat->type_size;
  s.dim = dat->dim;
  const int nd = dat->block->dims;
  for (int i = 0; i < 3; i++) {
    if (i >= nd) {
      s.first[i] = 0;
      s.q[i] = i;
      s.sgn[i] = 1;
      continue;
    }
    int q = abs(dir[i]) - 1;
    if (q < 0 || q >= nd) fail(OPS_INVALID_ARGUMENT, cat("ops_halo_transfer: bad direction ", dir[i]));
    s.q[i] = q;
    s.sgn[i] = dir[i] > 0 ? 1 : -1;
    s.first[i] = base[i] - dat->d_m[i] - dat->base[i];
    if (s.first[i] < 0 || s.first[i] + iter[q] > dat->size[i])
      fail(OPS_INVALID_ARGUMENT, cat("ops_halo_transfer: halo box leaves dat ", dat->name, " in dimension ", i));
  }
  return s;
}

void *g_buf = nullptr;
size_t g_buf_bytes = 0;

}  // namespace

extern "C" void ops_halo_transfer(ops_halo_group group) {
  OPSMI_API_BEGIN
  require_init();
  flush_all();
  if (!group) fail(OPS_INVALID_ARGUMENT, "ops_halo_transfer: NULL group");
  Instance &I = inst();
  for (int h = 0; h < group->nhalos; h++) {
    ops_halo halo = group->halos[h];
    ops_dat from = halo->from, to = halo->to;
    if (from->dim != to->dim) fail(OPS_INVALID_ARGUMENT, "ops_halo_transfer: dats differ in components");
    DType df = dtype_of(from->type, from->type_size), dt = dtype_of(to->type, to->type_size);
    if (from->type_size != to->type_size &&
        !((df == DT_F64 || df == DT_F32) && (dt == DT_F64 || dt == DT_F32)))
      fail(OPS_INVALID_ARGUMENT, "ops_halo_transfer: mixed exchange needs float/double dats");
    if ((from->type_size != 4 && from->type_size != 8) || (to->type_size != 4 && to->type_size != 8))
      fail(OPS_NOT_IMPLEMENTED, "ops_halo_transfer: element types of 4 or 8 bytes are supported");
    HaloBox box;
    int nd = from->block->dims;
    for (int i = 0; i < 3; i++) box.ext[i] = i < nd ? halo->iter_size[i] : 1;
    long n = (long)box.ext[0] * box.ext[1] * box.ext[2] * from->dim;
    if (n <= 0) continue;
    HaloSide sf = make_side(from, halo->from_base, halo->from_dir, box.ext);
    HaloSide st = make_side(to, halo->to_base, halo->to_dir, box.ext);
    int storage = std::min(from->type_size, to->type_size);
    size_t need = (size_t)n * storage;
    if (need > g_buf_bytes) {
      if (g_buf) {
        OPSMI_HIP(hipStreamSynchronize(I.stream));
        OPSMI_HIP(hipFree(g_buf));
      }
      OPSMI_HIP(hipMalloc(&g_buf, need));
      g_buf_bytes = need;
    }
    sync_to_device(from);
    sync_to_device(to);
    int nb = (int)std::min<long>((n + 255) / 256, (long)I.num_cus * 8);
    char *buf = (char *)g_buf;
    run_loop("ops_halo_transfer", 2.0 * n * storage, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_halo<true>, dim3(nb), dim3(256), 0, s, sf, box, buf, storage);
      hipLaunchKernelGGL(k_halo<false>, dim3(nb), dim3(256), 0, s, st, box, buf, storage);
    });
    mark_device_dirty(to);
    note_external_write(to);
  }
  OPSMI_API_END()
}
