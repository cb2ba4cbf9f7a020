"""ctypes front end of libopsmi_apps.so (include/ops_mi355x_apps.h): the
C++ application drivers (Laplace, TeaLeaf-form heat CG, Poisson CG,
mixed-precision Poisson, ADI tridiagonal sweeps) that run on the OPS C ABI."""
import ctypes as C
import os

import numpy as np

from . import ops

LIB_PATH = os.path.join(ops.HERE, "libopsmi_apps.so")
_lib = None
_d = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        ops.lib()  # loads libopsmi.so first, with the error mode set
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp, ci, cd, u64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64
        sig = {
            "opsapp_heat_create": (vp, [ci, ci, ci, ci, u64, cd, cd, cd, _d, _d]),
            "opsapp_heat_iterate": (ci, [vp, ci, cd, _d, _d, _d, _d]),
            "opsapp_heat_rr": (cd, [vp, ci]),
            "opsapp_heat_fetch_u": (None, [vp, _d]),
            "opsapp_heat_destroy": (None, [vp]),
            "opsapp_poisson_create": (vp, [ci, ci, ci, ci, u64, _d]),
            "opsapp_poisson_iterate": (ci, [vp, ci, cd, _d, _d, _d]),
            "opsapp_poisson_fetch_x": (None, [vp, _d]),
            "opsapp_poisson_destroy": (None, [vp]),
            "opsapp_pmixed_create": (vp, [ci, ci, ci, ci, u64, _d]),
            "opsapp_pmixed_solve": (ci, [vp, cd, ci, ci, cd, _d, C.POINTER(C.c_int)]),
            "opsapp_pmixed_fetch_x": (None, [vp, _d]),
            "opsapp_pmixed_destroy": (None, [vp]),
            "opsapp_laplace2d": (cd, [ci, ci, ci, ci, _d, _d, _d]),
            "opsapp_trid_create": (vp, [ci, ci, ci, u64, ci]),
            "opsapp_trid_sweep": (None, [vp, ci]),
            "opsapp_trid_solve": (None, [vp, ci, ci]),
            "opsapp_trid_fetch": (None, [vp, _d, _d]),
            "opsapp_trid_destroy": (None, [vp]),
            "opsapp_serve_stream": (cd, [ci, C.POINTER(ci), C.POINTER(ci), C.POINTER(u64), _d, cd, C.c_size_t, ci,
                                         ci, ci, ci, _d, C.POINTER(ci)]),
            "opsapp_mb_poisson": (cd, [ci, ci, ci, ci, ci, _d]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_d)


def _chk():
    ops._check()


class Heat:
    """TeaLeaf-form heat CG on the GPU (host loop in C++)."""

    def __init__(self, ndim, nx, ny, nz=1, seed=20240611, rx=1.0, ry=1.0, rz=1.0, K_ext=None, rhs=None):
        self.shape = (ny, nx) if ndim == 2 else (nz, ny, nx)
        self._K = None if K_ext is None else np.ascontiguousarray(K_ext, dtype=np.float64)
        self._f = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
        self.ctx = lib().opsapp_heat_create(ndim, nx, ny, nz, seed, rx, ry, rz, _p(self._K), _p(self._f))
        _chk()
        self._K = self._f = None

    def iterate(self, n, tol=0.0, history=False):
        h = [np.zeros(n) for _ in range(4)] if history else [None] * 4
        k = lib().opsapp_heat_iterate(self.ctx, n, tol, *[_p(x) if x is not None else None for x in h])
        _chk()
        if history:
            return k, dict(pw=h[0][:k], rrn=h[1][:k], alpha=h[2][:k], beta=h[3][:k])
        return k

    def rr(self, initial=False):
        return lib().opsapp_heat_rr(self.ctx, 1 if initial else 0)

    def u(self):
        out = np.zeros(self.shape)
        lib().opsapp_heat_fetch_u(self.ctx, _p(out))
        _chk()
        return out

    def close(self):
        if self.ctx:
            lib().opsapp_heat_destroy(self.ctx)
            _chk()
            self.ctx = None


class Poisson:
    def __init__(self, ndim, nx, ny, nz=1, seed=20240611, f=None):
        self.shape = (ny, nx) if ndim == 2 else (nz, ny, nx)
        self._f = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        self.ctx = lib().opsapp_poisson_create(ndim, nx, ny, nz, seed, _p(self._f))
        _chk()

    def iterate(self, n, tol=0.0, history=False):
        h = [np.zeros(n) for _ in range(3)] if history else [None] * 3
        k = lib().opsapp_poisson_iterate(self.ctx, n, tol, *[_p(x) if x is not None else None for x in h])
        _chk()
        if history:
            return k, dict(rrn=h[0][:k], alpha=h[1][:k], beta=h[2][:k])
        return k

    def x(self):
        out = np.zeros(self.shape)
        lib().opsapp_poisson_fetch_x(self.ctx, _p(out))
        _chk()
        return out

    def close(self):
        if self.ctx:
            lib().opsapp_poisson_destroy(self.ctx)
            _chk()
            self.ctx = None


class PoissonMixed:
    """Mixed-precision Poisson: FP64 iterative refinement around an FP32
    Jacobi-preconditioned CG (FP64 dot accumulation)."""

    def __init__(self, ndim, nx, ny, nz=1, seed=20240611, f=None):
        self.shape = (ny, nx) if ndim == 2 else (nz, ny, nx)
        self._f = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        self.ctx = lib().opsapp_pmixed_create(ndim, nx, ny, nz, seed, _p(self._f))
        _chk()
        self._f = None

    def solve(self, tol=1e-10, max_outer=30, inner_iters=100000, inner_tol=1e-4):
        rel = np.zeros(max_outer + 1)
        tot = C.c_int(0)
        n = lib().opsapp_pmixed_solve(self.ctx, tol, max_outer, inner_iters, inner_tol, _p(rel), C.byref(tot))
        _chk()
        return dict(outer=n, relres=rel[: n + 1], inner_total=tot.value)

    def x(self):
        out = np.zeros(self.shape)
        lib().opsapp_pmixed_fetch_x(self.ctx, _p(out))
        _chk()
        return out

    def close(self):
        if self.ctx:
            lib().opsapp_pmixed_destroy(self.ctx)
            _chk()
            self.ctx = None


def laplace2d(nx, ny, iters, mode=0, init_ext=None, want_field=True):
    hist = np.zeros(iters)
    A = np.zeros((ny, nx)) if want_field else None
    init = None if init_ext is None else np.ascontiguousarray(init_ext, dtype=np.float64)
    err = lib().opsapp_laplace2d(nx, ny, iters, mode, _p(init), _p(hist), _p(A))
    _chk()
    return err, hist, A


class Trid:
    def __init__(self, nx, ny, nz, seed=20240611, strategy=0):
        self.shape = (nz, ny, nx)
        self.ctx = lib().opsapp_trid_create(nx, ny, nz, seed, strategy)
        _chk()

    def sweep(self, inc=False):
        lib().opsapp_trid_sweep(self.ctx, 1 if inc else 0)
        _chk()

    def solve(self, solvedim, inc=False):
        lib().opsapp_trid_solve(self.ctx, solvedim, 1 if inc else 0)
        _chk()

    def fetch(self):
        d, u = np.zeros(self.shape), np.zeros(self.shape)
        lib().opsapp_trid_fetch(self.ctx, _p(d), _p(u))
        _chk()
        return d, u

    def close(self):
        if self.ctx:
            lib().opsapp_trid_destroy(self.ctx)
            _chk()
            self.ctx = None


def mb_poisson(nx=20, ny=20, ngx=1, ngy=1, iters=10, want_field=True):
    """The reference's multi-block Poisson test app; returns (total error, u)."""
    u = np.zeros((ny, nx)) if want_field else None
    err = lib().opsapp_mb_poisson(nx, ny, ngx, ngy, iters, _p(u))
    _chk()
    return err, u


def serve_stream(kinds, sizes, seeds, arrival_s, tol=1e-10, budget=200 * 10 ** 9, max_slots=256,
                 iters_per_launch=8, autotune=False, mixed=False):
    """Runs a request stream on the batched engine (opsapp_serve_stream);
    mixed solves every request with FP32-inner refinement.
    Returns (makespan_s, done_s, iters)."""
    n = len(kinds)
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    s = np.ascontiguousarray(sizes, dtype=np.int32)
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
    done = np.zeros(n)
    it = np.zeros(n, dtype=np.int32)
    ci = C.POINTER(C.c_int)
    span = lib().opsapp_serve_stream(n, k.ctypes.data_as(ci), s.ctypes.data_as(ci),
                                     sd.ctypes.data_as(C.POINTER(C.c_uint64)), _p(arr), tol, budget, max_slots,
                                     iters_per_launch, 1 if autotune else 0, 1 if mixed else 0, _p(done),
                                     it.ctypes.data_as(ci))
    _chk()
    if span < 0:
        raise RuntimeError("opsapp_serve_stream failed")
    return span, done, it
