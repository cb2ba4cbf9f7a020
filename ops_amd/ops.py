"""Python front end of libopsmi.so (include/ops_mi355x.h) over ctypes.

Function names, argument order and meaning follow the OPS C API
(ops/c/include/ops_lib_core.h:581-1297) so tests and drivers read like the
reference's own apps; numpy arrays stand in for host buffers.  Errors raise
OPSError (the library runs in its "record and return" error mode), mirroring
the reference's OPSException codes (ops/c/include/ops_exceptions.h:46-53).

The library is loaded from ops_amd/libopsmi.so; importing this module fails
loudly when it has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopsmi.so")
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "ops_mi355x.h")

OPS_READ, OPS_WRITE, OPS_RW, OPS_INC, OPS_MIN, OPS_MAX = range(6)
OPS_ARG_GBL, OPS_ARG_DAT, OPS_ARG_IDX = range(3)
OPS_HOST, OPS_DEVICE = 1, 2
OPS_MAX_DIM = 5


class OPSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"OPS error {code}: {msg}")
        self.code = code
        self.msg = msg


class ops_block_core(C.Structure):
    _fields_ = [("index", C.c_int), ("dims", C.c_int), ("name", C.c_char_p), ("instance", C.c_void_p)]


class ops_dat_core(C.Structure):
    _fields_ = [
        ("index", C.c_int), ("block", C.POINTER(ops_block_core)), ("dim", C.c_int), ("type_size", C.c_int),
        ("elem_size", C.c_int), ("size", C.c_int * OPS_MAX_DIM), ("base", C.c_int * OPS_MAX_DIM),
        ("d_m", C.c_int * OPS_MAX_DIM), ("d_p", C.c_int * OPS_MAX_DIM), ("x_pad", C.c_int),
        ("data", C.c_void_p), ("data_d", C.c_void_p), ("name", C.c_char_p), ("type", C.c_char_p),
        ("dirty_hd", C.c_int), ("locked_hd", C.c_int), ("user_managed", C.c_int), ("is_hdf5", C.c_int),
        ("hdf5_file", C.c_char_p), ("e_dat", C.c_int), ("mem", C.c_size_t), ("base_offset", C.c_size_t),
        ("stride", C.c_int * OPS_MAX_DIM), ("impl", C.c_void_p),
    ]


class ops_reduction_core(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_int), ("initialized", C.c_int), ("index", C.c_int),
                ("acc", C.c_int), ("type", C.c_char_p), ("name", C.c_char_p), ("instance", C.c_void_p),
                ("impl", C.c_void_p)]


class ops_arg(C.Structure):
    _fields_ = [("dat", C.c_void_p), ("stencil", C.c_void_p), ("dim", C.c_int), ("elem_size", C.c_int),
                ("data", C.c_void_p), ("data_d", C.c_void_p), ("acc", C.c_int), ("argtype", C.c_int),
                ("opt", C.c_int)]


_lib = None
_keep = []  # host buffers referenced by queued ops_arg_gbl values


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp, ci, cd, cs = C.c_void_p, C.c_int, C.c_double, C.c_char_p
        ip = C.POINTER(C.c_int)
        sig = {
            "ops_init": (None, [ci, C.POINTER(cs), ci]),
            "ops_exit": (None, []),
            "ops_execute": (None, []),
            "ops_decl_block": (vp, [ci, cs]),
            "ops_decl_dat_char": (vp, [vp, ci, ip, ip, ip, ip, ip, vp, ci, cs, cs]),
            "ops_free_dat": (None, [vp]),
            "ops_decl_stencil": (vp, [ci, ci, ip, cs]),
            "ops_decl_strided_stencil": (vp, [ci, ci, ip, ip, cs]),
            "ops_decl_reduction_handle": (vp, [ci, cs, cs]),
            "ops_decl_const_char": (None, [vp, ci, cs, ci, vp, cs]),
            "ops_decl_halo": (vp, [vp, vp, ip, ip, ip, ip, ip]),
            "ops_decl_halo_group": (vp, [ci, C.POINTER(vp)]),
            "ops_halo_transfer": (None, [vp]),
            "ops_arg_dat": (ops_arg, [vp, ci, vp, cs, ci]),
            "ops_arg_idx": (ops_arg, []),
            "ops_arg_reduce": (ops_arg, [vp, ci, cs, ci]),
            "ops_arg_gbl_char": (ops_arg, [vp, ci, ci, ci]),
            "ops_reduction_result_char": (None, [vp, ci, vp]),
            "ops_dat_fetch_data": (None, [vp, ci, vp]),
            "ops_dat_fetch_data_slab_memspace": (None, [vp, ci, vp, ip, ci]),
            "ops_dat_set_data": (None, [vp, ci, vp]),
            "ops_dat_set_data_slab_memspace": (None, [vp, ci, vp, ip, ci]),
            "ops_dat_get_extents": (None, [vp, ci, ip, ip]),
            "ops_dat_get_raw_pointer": (vp, [vp, ci, vp, ip]),
            "ops_dat_release_raw_data": (None, [vp, ci, ci]),
            "ops_randomgen_init": (None, [C.c_uint, ci]),
            "ops_fill_random_uniform": (None, [vp]),
            "ops_fill_random_normal": (None, [vp]),
            "ops_timing_output_stdout": (None, []),
            "ops_tridsolver_params_create": (vp, [vp, ci]),
            "ops_tridsolver_params_destroy": (None, [vp]),
            "ops_tridMultiDimBatch_c": (None, [ci, ci, ip, vp, vp, vp, vp, vp]),
            "ops_tridMultiDimBatch_Inc_c": (None, [ci, ci, ip, vp, vp, vp, vp, vp, vp]),
            "ops_tridMultiDimBatch": (None, [ci, ci, ip, vp, vp, vp, vp, vp]),
            "ops_tridMultiDimBatch_Inc": (None, [ci, ci, ip, vp, vp, vp, vp, vp, vp]),
            "ops_tridMultiDimBatch_noalloc": (None, [ci, ci, ip, vp, vp, vp, vp, vp, vp]),
            "ops_mi_set_error_mode": (None, [ci]),
            "ops_mi_last_error": (ci, [cs, ci]),
            "ops_mi_stream": (vp, []),
            "ops_mi_sync": (None, []),
            "ops_mi_loop_stats": (ci, [cs, C.POINTER(C.c_long), C.POINTER(cd), C.POINTER(cd)]),
            "ops_mi_reset_stats": (None, []),
            "ops_mi_set_fusion": (None, [ci]),
            "ops_mi_set_diags": (None, [ci]),
            "ops_mi_set_host_fold": (None, [ci]),
            "ops_mi_stream_probe": (ci, [ci, ci, ci, C.c_long, ci, C.POINTER(cd)]),
            "ops_set_soa": (None, [ci]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.ops_mi_set_error_mode(1)
        _lib = L
    return _lib


def _check():
    buf = C.create_string_buffer(1024)
    code = lib().ops_mi_last_error(buf, 1024)
    if code:
        raise OPSError(code, buf.value.decode(errors="replace"))


def _call(name, *a):
    r = getattr(lib(), name)(*a)
    _check()
    return r


def _ints(v):
    v = list(v)
    return (C.c_int * max(len(v), 1))(*v)


def _b(s):
    return s.encode() if isinstance(s, str) else s


# ---------------------------------------------------------------- objects
class Dat:
    """Handle of an ops_dat with its numpy dtype and interior shape."""

    def __init__(self, ptr, block, size, d_m, d_p, dtype, name, dim=1):
        self.ptr = ptr
        self.block = block
        self.size = list(size)
        self.d_m = list(d_m)
        self.d_p = list(d_p)
        self.dtype = np.dtype(dtype)
        self.name = name
        self.dim = dim

    @property
    def core(self):
        return C.cast(self.ptr, C.POINTER(ops_dat_core)).contents

    @property
    def shape(self):
        return tuple(reversed(self.size))


class Block:
    def __init__(self, ptr, dims, name):
        self.ptr = ptr
        self.dims = dims
        self.name = name


_TYPES = {np.dtype(np.float64): "double", np.dtype(np.float32): "float", np.dtype(np.int32): "int"}


def ops_init(argv=(), diags=1):
    args = [b"ops"] + [_b(a) for a in argv]
    arr = (C.c_char_p * len(args))(*args)
    _call("ops_init", len(args), arr, diags)


def ops_exit():
    _call("ops_exit")


def ops_execute():
    _call("ops_execute")


def ops_set_soa(soa):
    """multi-component dats declared from now on: 1 structure of arrays, 0
    array of structures (the default)"""
    _call("ops_set_soa", 1 if soa else 0)


def ops_decl_block(dims, name):
    return Block(_call("ops_decl_block", dims, _b(name)), dims, name)


def ops_decl_dat(block, size, d_m=None, d_p=None, dtype=np.float64, name="dat", data=None, base=None, dim=1):
    nd = block.dims
    d_m = d_m if d_m is not None else [-1] * nd
    d_p = d_p if d_p is not None else [1] * nd
    base = base if base is not None else [0] * nd
    dt = np.dtype(dtype)
    ptr = None
    if data is not None:
        data = np.ascontiguousarray(data, dtype=dt)
        ptr = data.ctypes.data
    p = _call("ops_decl_dat_char", block.ptr, dim, _ints(size), _ints(base), _ints(d_m), _ints(d_p),
              _ints([1] * nd), ptr, dt.itemsize, _b(_TYPES[dt]), _b(name))
    return Dat(p, block, size, d_m, d_p, dt, name, dim)


def ops_free_dat(dat):
    _call("ops_free_dat", dat.ptr)


def ops_decl_stencil(dims, points, offsets, name):
    return _call("ops_decl_stencil", dims, points, _ints(offsets), _b(name))


class Reduction:
    def __init__(self, ptr, dtype, count):
        self.ptr = ptr
        self.dtype = np.dtype(dtype)
        self.count = count


def ops_decl_reduction_handle(size, type_, name):
    dt = {"double": np.float64, "float": np.float32, "int": np.int32}[type_]
    return Reduction(_call("ops_decl_reduction_handle", size, _b(type_), _b(name)), dt,
                     size // np.dtype(dt).itemsize)


def ops_decl_const(name, dim, type_, value):
    dt = {"double": np.float64, "float": np.float32, "int": np.int32}[type_]
    arr = np.ascontiguousarray(np.atleast_1d(np.asarray(value, dtype=dt)))
    _call("ops_decl_const_char", None, dim, _b(type_), arr.itemsize, arr.ctypes.data, _b(name))


def ops_arg_dat(dat, dim, stencil, type_, acc):
    return _call("ops_arg_dat", dat.ptr if dat is not None else None, dim, stencil, _b(type_), acc)


def ops_arg_idx():
    return _call("ops_arg_idx")


def ops_arg_reduce(handle, dim, type_, acc):
    return _call("ops_arg_reduce", handle.ptr, dim, _b(type_), acc)


def ops_arg_gbl(value, dim, type_, acc=OPS_READ):
    dt = {"double": np.float64, "float": np.float32, "int": np.int32}[type_]
    arr = np.ascontiguousarray(np.atleast_1d(np.asarray(value, dtype=dt)))
    _keep.append(arr)
    if len(_keep) > 256:
        del _keep[:128]
    return _call("ops_arg_gbl_char", arr.ctypes.data, dim, arr.itemsize, acc)


def ops_par_loop(kernel, name, block, dim, rng, *args):
    """Calls ops_par_loop_<kernel>(name, block, dim, range, args...)."""
    fn = getattr(lib(), "ops_par_loop_" + kernel)
    fn.restype = None
    fn.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)] + [ops_arg] * len(args)
    fn(_b(name), block.ptr, dim, _ints(rng), *args)
    _check()


def ops_reduction_result(handle):
    out = np.zeros(handle.count, dtype=handle.dtype)
    _call("ops_reduction_result_char", handle.ptr, out.itemsize, out.ctypes.data)
    return out if handle.count > 1 else out[0]


def ops_dat_fetch_data(dat):
    out = np.empty(dat.shape, dtype=dat.dtype)
    _call("ops_dat_fetch_data", dat.ptr, 0, out.ctypes.data)
    return out


def ops_dat_set_data(dat, arr):
    arr = np.ascontiguousarray(arr, dtype=dat.dtype)
    assert arr.size == int(np.prod(dat.size)), "ops_dat_set_data takes the interior"
    _call("ops_dat_set_data", dat.ptr, 0, arr.ctypes.data)


def ops_dat_fetch_data_slab(dat, rng):
    nd = dat.block.dims
    shape = tuple(rng[2 * d + 1] - rng[2 * d] for d in reversed(range(nd))) + ((dat.dim,) if dat.dim > 1 else ())
    out = np.empty(shape, dtype=dat.dtype)
    _call("ops_dat_fetch_data_slab_memspace", dat.ptr, 0, out.ctypes.data, _ints(rng), OPS_HOST)
    return out


def ops_dat_set_data_slab(dat, rng, arr):
    arr = np.ascontiguousarray(arr, dtype=dat.dtype)
    _call("ops_dat_set_data_slab_memspace", dat.ptr, 0, arr.ctypes.data, _ints(rng), OPS_HOST)


def set_ext(dat, arr_ext):
    """Uploads a halo-including array (1-cell halo per dimension)."""
    nd = dat.block.dims
    rng = []
    for d in range(nd):
        rng += [-1, dat.size[d] + 1]
    ops_dat_set_data_slab(dat, rng, arr_ext)


def get_ext(dat):
    nd = dat.block.dims
    rng = []
    for d in range(nd):
        rng += [-1, dat.size[d] + 1]
    return ops_dat_fetch_data_slab(dat, rng)


def _whole_range(dat):
    rng = []
    for d in range(dat.block.dims):
        rng += [dat.d_m[d], dat.size[d] + dat.d_p[d]]
    return rng


def set_whole(dat, arr):
    """Uploads the dat's whole array: halos of depth -d_m / d_p included,
    indexed [z, y, x(, component)]."""
    ops_dat_set_data_slab(dat, _whole_range(dat), arr)


def get_whole(dat):
    """Downloads the dat's whole array (halos included), as set_whole."""
    return ops_dat_fetch_data_slab(dat, _whole_range(dat))


# ---------------------------------------------------------------- halos
def ops_decl_halo(frm, to, iter_size, from_base, to_base, from_dir, to_dir):
    return _call("ops_decl_halo", frm.ptr, to.ptr, _ints(iter_size), _ints(from_base), _ints(to_base),
                 _ints(from_dir), _ints(to_dir))


def ops_decl_halo_group(halos):
    return _call("ops_decl_halo_group", len(halos), (C.c_void_p * len(halos))(*halos))


def ops_halo_transfer(group):
    _call("ops_halo_transfer", group)


def ops_randomgen_init(seed, options=0):
    _call("ops_randomgen_init", seed, options)


def ops_fill_random_uniform(dat):
    _call("ops_fill_random_uniform", dat.ptr)


def ops_fill_random_normal(dat):
    _call("ops_fill_random_normal", dat.ptr)


class TridParams:
    """ops_tridsolver_params: ptr is the settings object, ref the address of
    a struct laid out as the reference class (one void* member pointing at
    it), which the C-linkage ops_tridMultiDimBatch[_Inc] take."""
    GATHER_SCATTER, ALLGATHER, JACOBI, PCR = 0, 1, 2, 3

    def __init__(self, block, strategy=0):
        self.ptr = _call("ops_tridsolver_params_create", block.ptr, strategy)
        self._obj = C.c_void_p(self.ptr)
        self.ref = C.addressof(self._obj)

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.ops_tridsolver_params_destroy(self.ptr)
        except Exception:
            pass


def ops_tridMultiDimBatch(ndim, solvedim, rng, a, b, c, d, ctx):
    _call("ops_tridMultiDimBatch", ndim, solvedim, _ints(rng), a.ptr, b.ptr, c.ptr, d.ptr, ctx.ref)


def ops_tridMultiDimBatch_Inc(ndim, solvedim, rng, a, b, c, d, u, ctx):
    _call("ops_tridMultiDimBatch_Inc", ndim, solvedim, _ints(rng), a.ptr, b.ptr, c.ptr, d.ptr, u.ptr, ctx.ref)


def ops_tridMultiDimBatch_noalloc(ndim, solvedim, rng, a, b, c, d, u, ctx):
    """u None: ops_tridMultiDimBatch; u a dat: ops_tridMultiDimBatch_Inc; the
    c'/d' scratch in the library's fixed device array (no allocation)."""
    _call("ops_tridMultiDimBatch_noalloc", ndim, solvedim, _ints(rng), a.ptr, b.ptr, c.ptr, d.ptr,
          u.ptr if u is not None else None, ctx.ref)


def ops_mi_sync():
    _call("ops_mi_sync")


def ops_mi_stream():
    return lib().ops_mi_stream()


def ops_mi_loop_stats(name):
    cnt, ms, by = C.c_long(), C.c_double(), C.c_double()
    ok = _call("ops_mi_loop_stats", _b(name), C.byref(cnt), C.byref(ms), C.byref(by))
    return (cnt.value, ms.value, by.value) if ok else None


def ops_mi_reset_stats():
    _call("ops_mi_reset_stats")


def ops_mi_set_diags(level):
    """diagnostics level: above 1 every launch is timed with HIP events"""
    _call("ops_mi_set_diags", level)


def ops_mi_set_host_fold(on):
    """1: results read right after their one loop are folded by the host (same bits)"""
    _call("ops_mi_set_host_fold", 1 if on else 0)


def ops_mi_set_fusion(level):
    """loop-chain fusion level of the queue: 0 off, 1 point-wise pairs, 2 chains,
    3 also runs of three or four generated point-wise loops"""
    _call("ops_mi_set_fusion", level)


SERVE_HEADER = os.path.join(ROOT, "include", "ops_mi355x_serve.h")
GEN_HEADER = os.path.join(ROOT, "include", "ops_mi355x_gen.h")


def declared_symbols():
    """C entry points declared in include/ops_mi355x.h (the extern "C" part),
    include/ops_mi355x_serve.h and include/ops_mi355x_gen.h, all exported by
    libopsmi.so."""
    import re
    text = open(HEADER).read()
    c_part = text.split("} /* extern \"C\" */")[0] + open(SERVE_HEADER).read() + open(GEN_HEADER).read()
    c_part = re.sub(r"/\*.*?\*/", "", c_part, flags=re.S)
    names = re.findall(r"\b([a-z_][a-zA-Z0-9_]*)\s*\(", c_part)
    skip = {"if", "for", "while", "sizeof", "return"}
    # header-only helpers are not library symbols
    skip |= set(re.findall(r"static\s+inline\s+[\w\s\*]+?\b([a-z_]\w*)\s*\(", c_part))
    return sorted({n for n in names if (n.startswith("ops_") or n.startswith("opsmi_")) and n not in skip})
