"""ORACLE — test infrastructure only.

ctypes front end of oracle/_build/libops_oracle.so, the CPU restatement of
the reference OPS sequential/OpenMP path.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module, as the checker or the CPU
baseline.  Array shapes follow the oracle's conventions (ops_oracle.h):
"ext" arrays carry a 1-cell halo per dimension and are indexed [k, j, i]
(numpy C order, x fastest); interior arrays likewise without the halo.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libops_oracle.so")

_lib = None

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64, i64, ci, cd = C.c_uint64, C.c_int64, C.c_int, C.c_double
        sig = {
            "orc_gen_u01": (None, [u64, u64, i64, _d]),
            "orc_gen_normal": (None, [u64, u64, i64, _d]),
            "orc_gen_laplace_init": (None, [ci, ci, ci, _d]),
            "orc_gen_heat": (None, [u64, ci, ci, ci, ci, _d, _d]),
            "orc_gen_poisson_rhs": (None, [u64, ci, ci, ci, ci, _d]),
            "orc_gen_trid": (None, [u64, ci, ci, ci, ci, _d, _d, _d]),
            "orc_gen_trid_rhs": (None, [u64, ci, ci, ci, _d]),
            "orc_laplace_run": (cd, [_d, _d, ci, ci, ci, ci, _d]),
            "orc_laplace_run_f32": (C.c_float, [_f, _f, ci, ci, ci, ci, _f]),
            "orc_heat_cg": (ci, [ci, ci, ci, ci, _d, _d, cd, cd, cd, ci, cd, ci, _d, _d, _d, _d, _d]),
            "orc_last_loop_seconds": (cd, []),
            "orc_heat_coefs": (None, [ci, ci, ci, ci, _d, _d, _d, _d]),
            "orc_heat_apply": (cd, [ci, ci, ci, ci, _d, _d, _d, _d, cd, cd, cd, _d, ci]),
            "orc_poisson_cg": (ci, [ci, ci, ci, ci, _d, ci, cd, ci, _d, _d, _d, _d]),
            "orc_poisson_apply_f32": (cd, [ci, ci, ci, ci, _f, _f]),
            "orc_poisson_residual": (cd, [ci, ci, ci, ci, _d, _d, _d, ci]),
            "orc_poisson_mixed": (ci, [ci, ci, ci, ci, _d, cd, ci, ci, cd, ci, _d, _d, _i]),
            "orc_heat_mixed": (ci, [ci, ci, ci, _d, _d, cd, ci, ci, cd, ci, _d, _d, _i]),
            "orc_adi_preproc": (None, [ci, ci, ci, cd, _d, _d, _d, _d, _d]),
            "orc_stencil27": (None, [ci, ci, ci, _d, _d]),
            "orc_mb_poisson": (cd, [ci, ci, ci, _d]),
            "orc_thomas_batch": (None, [ci, ci, _i, _i, _d, _d, _d, _d, _d, ci]),
            "orc_thomas_batch_f32": (None, [ci, ci, _i, _i, _f, _f, _f, _f, _f, ci]),
            "orc_pcr_batch": (None, [ci, ci, _i, _i, _d, _d, _d, _d, _d, ci]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def P(a):
    """ctypes pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(_d)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_f)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_i)
    raise TypeError(a.dtype)


def ext_shape(ndim, nx, ny, nz=1):
    return (ny + 2, nx + 2) if ndim == 2 else (nz + 2, ny + 2, nx + 2)


def int_shape(ndim, nx, ny, nz=1):
    return (ny, nx) if ndim == 2 else (nz, ny, nx)


def interior(a_ext):
    return a_ext[(slice(1, -1),) * a_ext.ndim]


# ------------------------------------------------------------------ inputs
def u01(seed, stream, n):
    out = np.empty(n)
    lib().orc_gen_u01(seed, stream, n, P(out))
    return out


def normal(seed, stream, n):
    out = np.empty(n)
    lib().orc_gen_normal(seed, stream, n, P(out))
    return out


def laplace_init(nx, ny, mode=0):
    u = np.zeros(ext_shape(2, nx, ny))
    lib().orc_gen_laplace_init(nx, ny, mode, P(u))
    return u


def heat_inputs(seed, ndim, nx, ny, nz=1):
    K = np.zeros(ext_shape(ndim, nx, ny, nz))
    rhs = np.zeros(int_shape(ndim, nx, ny, nz))
    lib().orc_gen_heat(seed, ndim, nx, ny, nz, P(K), P(rhs))
    return K, rhs


def poisson_rhs(seed, ndim, nx, ny, nz=1):
    f = np.zeros(int_shape(ndim, nx, ny, nz))
    lib().orc_gen_poisson_rhs(seed, ndim, nx, ny, nz, P(f))
    return f


def trid_inputs(seed, s, nx, ny, nz):
    a, b, c = (np.zeros((nz, ny, nx)) for _ in range(3))
    lib().orc_gen_trid(seed, s, nx, ny, nz, P(a), P(b), P(c))
    return a, b, c


def trid_rhs(seed, nx, ny, nz):
    d = np.zeros((nz, ny, nx))
    lib().orc_gen_trid_rhs(seed, nx, ny, nz, P(d))
    return d


# ----------------------------------------------------------------- solvers
def laplace_run(A_ext, iters, nthreads=1):
    """Runs Jacobi in place on A_ext (float64 or float32); returns (A, hist)."""
    A = np.ascontiguousarray(A_ext).copy()
    An = A.copy()
    ny, nx = A.shape[0] - 2, A.shape[1] - 2
    if A.dtype == np.float32:
        hist = np.zeros(iters, np.float32)
        lib().orc_laplace_run_f32(P(A), P(An), nx, ny, iters, nthreads, P(hist))
    else:
        hist = np.zeros(iters)
        lib().orc_laplace_run(P(A), P(An), nx, ny, iters, nthreads, P(hist))
    return A, hist


def heat_cg(ndim, nx, ny, nz, K_ext, rhs, rx=1.0, ry=1.0, rz=1.0, max_iters=100, tol=0.0, nthreads=1):
    u = np.zeros(int_shape(ndim, nx, ny, nz))
    pw, al, be = (np.zeros(max_iters) for _ in range(3))
    rrn = np.zeros(max_iters + 1)
    n = lib().orc_heat_cg(ndim, nx, ny, nz, P(np.ascontiguousarray(K_ext)), P(np.ascontiguousarray(rhs)),
                          rx, ry, rz, max_iters, tol, nthreads, P(u), P(pw), P(rrn), P(al), P(be))
    return dict(iters=n, u=u, pw=pw[:n], rrn=rrn[: n + 1], alpha=al[:n], beta=be[:n])


def last_loop_seconds():
    """Wall time of the iteration loop of the last heat_cg call."""
    return lib().orc_last_loop_seconds()


def heat_coefs(ndim, nx, ny, nz, K_ext):
    Kx, Ky, Kz = (np.zeros(ext_shape(ndim, nx, ny, nz)) for _ in range(3))
    lib().orc_heat_coefs(ndim, nx, ny, nz, P(np.ascontiguousarray(K_ext)), P(Kx), P(Ky), P(Kz))
    return Kx, Ky, Kz


def heat_apply(ndim, nx, ny, nz, Kx, Ky, Kz, p_ext, rx=1.0, ry=1.0, rz=1.0, nthreads=1):
    w = np.zeros_like(p_ext)
    s = lib().orc_heat_apply(ndim, nx, ny, nz, P(Kx), P(Ky), P(Kz), P(p_ext), rx, ry, rz, P(w), nthreads)
    return w, s


def poisson_cg(ndim, nx, ny, nz, f, max_iters=100, tol=0.0, nthreads=1):
    x = np.zeros(int_shape(ndim, nx, ny, nz))
    al, be = np.zeros(max_iters), np.zeros(max_iters)
    rrn = np.zeros(max_iters + 1)
    n = lib().orc_poisson_cg(ndim, nx, ny, nz, P(np.ascontiguousarray(f)), max_iters, tol, nthreads, P(x),
                             P(rrn), P(al), P(be))
    return dict(iters=n, x=x, rrn=rrn[: n + 1], alpha=al[:n], beta=be[:n])


def poisson_apply_f32(ndim, nx, ny, nz, p_ext):
    w = np.zeros_like(p_ext, dtype=np.float32)
    s = lib().orc_poisson_apply_f32(ndim, nx, ny, nz, P(np.ascontiguousarray(p_ext, np.float32)), P(w))
    return w, s


def poisson_residual(ndim, nx, ny, nz, f, x_ext, nthreads=1):
    r = np.zeros_like(x_ext)
    s = lib().orc_poisson_residual(ndim, nx, ny, nz, P(np.ascontiguousarray(f)),
                                   P(np.ascontiguousarray(x_ext)), P(r), nthreads)
    return r, s


def poisson_mixed(ndim, nx, ny, nz, f, tol=1e-10, max_outer=30, inner_iters=1000, inner_tol=1e-4,
                  nthreads=1):
    x = np.zeros(int_shape(ndim, nx, ny, nz))
    rel = np.zeros(max_outer + 1)
    tot = np.zeros(1, np.int32)
    n = lib().orc_poisson_mixed(ndim, nx, ny, nz, P(np.ascontiguousarray(f)), tol, max_outer, inner_iters,
                                inner_tol, nthreads, P(x), P(rel), P(tot))
    return dict(outer=n, x=x, relres=rel[: n + 1], inner_total=int(tot[0]))


def heat_mixed(nx, ny, nz, K_ext, rhs, tol=1e-10, max_outer=30, inner_iters=100000, inner_tol=1e-6, nthreads=1):
    """3D heat problem by FP64 refinement around FP32 CG (orc_heat_mixed)."""
    x = np.zeros(int_shape(3, nx, ny, nz))
    rel = np.zeros(max_outer + 1)
    tot = np.zeros(1, np.int32)
    n = lib().orc_heat_mixed(nx, ny, nz, P(np.ascontiguousarray(K_ext)), P(np.ascontiguousarray(rhs)), tol,
                             max_outer, inner_iters, inner_tol, nthreads, P(x), P(rel), P(tot))
    return dict(outer=n, x=x, relres=rel[: n + 1], inner_total=int(tot[0]))


def adi_preproc(nx, ny, nz, u_ext, lam=1.0):
    out = [np.zeros((nz, ny, nx)) for _ in range(4)]
    lib().orc_adi_preproc(nx, ny, nz, lam, P(np.ascontiguousarray(u_ext)), *[P(o) for o in out])
    return out


def stencil27(nx, ny, nz, u_ext):
    out = np.zeros((nz, ny, nx))
    lib().orc_stencil27(nx, ny, nz, P(np.ascontiguousarray(u_ext)), P(out))
    return out


def mb_poisson(nx=20, ny=20, iters=10):
    """The reference's multi-block Poisson test app on one block: (total
    squared error, interior u)."""
    u = np.zeros((ny, nx))
    err = lib().orc_mb_poisson(nx, ny, iters, P(u))
    return err, u


def _trid(fn, ndim, solvedim, dims, pads, a, b, c, d, u, nthreads):
    dims = np.ascontiguousarray(dims, np.int32)
    pads = np.ascontiguousarray(pads, np.int32)
    fn(ndim, solvedim, P(dims), P(pads), P(a), P(b), P(c), P(d), P(u), nthreads)


def thomas_batch(ndim, solvedim, dims, pads, a, b, c, d, u=None, nthreads=1):
    """In place on d (or accumulates into u); arrays are flat OPS-layout
    buffers starting at the first solved element."""
    f = lib().orc_thomas_batch_f32 if a.dtype == np.float32 else lib().orc_thomas_batch
    _trid(f, ndim, solvedim, dims, pads, a, b, c, d, u, nthreads)


def pcr_batch(ndim, solvedim, dims, pads, a, b, c, d, u=None, nthreads=1):
    _trid(lib().orc_pcr_batch, ndim, solvedim, dims, pads, a, b, c, d, u, nthreads)


def load_golden():
    """The reference fixtures (tests/golden/ref_seq.npz) and their metadata."""
    import json
    path = os.path.join(os.path.dirname(HERE), "tests", "golden", "ref_seq.npz")
    z = np.load(path, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    meta = json.loads(bytes(g.pop("meta_json")).decode())
    return g, meta


# ------------------------------------------------------------ halo transfer
class HaloDat:
    """A dat for the halo-transfer restatement: its whole array ``a`` (halos
    included) indexed [z, y, x, component] from the lowest halo point, the
    negative halo depths ``d_m`` (<= 0) and the element type."""

    def __init__(self, a, d_m, ndim):
        self.a = a
        self.d_m = list(d_m) + [0] * (3 - len(d_m))
        self.ndim = ndim


def _halo_ranges(dat, base, dirs, iters):
    # ranges/step/buffer strides of ops_host_singlenode.cpp:201-219 (source)
    # and :327-342 (target), in raw coordinates of the dat's array
    lo, hi, step, bstr = [], [], [], []
    for i in range(3):
        q = abs(dirs[i]) - 1
        if dirs[i] > 0:
            r0 = base[i] - dat.d_m[i]
            r1 = r0 + iters[q]
            step.append(1)
        else:
            r1 = base[i] - 1 - dat.d_m[i]
            r0 = r1 + iters[q]
            step.append(-1)
        lo.append(r0)
        hi.append(r1)
        s = 1
        for j in range(q):
            s *= iters[j]
        bstr.append(s)
    return lo, hi, step, bstr


def _halo_points(lo, hi):
    # the reference's loops: min(r0, r1 + 1) <= idx < max(r0 + 1, r1)
    rng = [range(min(lo[i], hi[i] + 1), max(lo[i] + 1, hi[i])) for i in range(3)]
    for k in rng[2]:
        for j in rng[1]:
            for i in rng[0]:
                yield i, j, k


def halo_transfer(frm, to, iter_size, from_base, to_base, from_dir, to_dir):
    """One halo of ops_halo_transfer (ops/c/src/sequential/
    ops_host_singlenode.cpp:172-446), in place on ``to.a``: the source box is
    packed into a linear buffer (buffer dimension q has extent iter_size[q],
    source dimension i walks buffer dimension |from_dir[i]| - 1, backwards
    when from_dir[i] < 0), then the buffer is unpacked into the target box the
    same way.  Entries beyond the block's dimensions are 1 / 0 / i + 1, as
    ops_decl_halo_core sets them (ops/c/src/core/ops_lib_core.cpp:1023-1029).
    When a float and a double dat meet, values pass through the narrower type
    (:225-228).  The reference places component d of buffer point pos at
    pos + d in that mixed case and at pos * dim + d otherwise; the two agree
    for one component, the only mixed case the fixtures exercise."""
    nd = frm.ndim
    pad = lambda v, fill: list(v[:nd]) + [fill(i) for i in range(nd, 3)]  # noqa: E731
    it = pad(iter_size, lambda i: 1)
    fb, tb = pad(from_base, lambda i: 0), pad(to_base, lambda i: 0)
    fd, td = pad(from_dir, lambda i: i + 1), pad(to_dir, lambda i: i + 1)
    dim = frm.a.shape[3]
    storage = np.float32 if np.float32 in (frm.a.dtype.type, to.a.dtype.type) else np.float64
    buf = np.zeros(it[0] * it[1] * it[2] * dim, dtype=storage)
    lo, hi, st, bs = _halo_ranges(frm, fb, fd, it)
    for p in _halo_points(lo, hi):
        pos = sum((p[i] - lo[i]) * st[i] * bs[i] for i in range(3))
        for d in range(dim):
            buf[pos * dim + d] = storage(frm.a[p[2], p[1], p[0], d])
    lo, hi, st, bs = _halo_ranges(to, tb, td, it)
    for p in _halo_points(lo, hi):
        pos = sum((p[i] - lo[i]) * st[i] * bs[i] for i in range(3))
        for d in range(dim):
            to.a[p[2], p[1], p[0], d] = buf[pos * dim + d]


def halo_whole_shape(ndim, size, d_m, d_p, dim):
    """Shape [z, y, x, component] of a dat's whole array (halos included)."""
    e = [size[i] - d_m[i] + d_p[i] if i < ndim else 1 for i in range(3)]
    return (e[2], e[1], e[0], dim)


def halo_fixture(g, case):
    """The reference's arrays for one halo-transfer case of the fixtures:
    {dat name: [whole array before any group, after group 1, ...]}."""
    out = {}
    for name, typ, dim, size, d_m, d_p in case["dats"]:
        shp = halo_whole_shape(case["ndim"], size, d_m, d_p, dim)
        out[name] = [g[f"halo_{case['name']}_{name}_g{k}"].reshape(shp) for k in range(len(case["groups"]) + 1)]
    return out
