"""ctypes front end of the batched multi-block solve engine
(include/ops_mi355x_serve.h, exported by libopsmi.so)."""
import ctypes as C

import numpy as np

from . import ops

POISSON2D, HEAT3D = 0, 1


class SolveReq(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("seed", C.c_uint64), ("tol", C.c_double),
                ("max_iters", C.c_int), ("mixed", C.c_int)]


class SolveDone(C.Structure):
    _fields_ = [("id", C.c_int), ("status", C.c_int), ("iters", C.c_int), ("outer", C.c_int),
                ("rr0", C.c_double), ("rr", C.c_double)]


_sigs = None


def lib():
    global _sigs
    L = ops.lib()
    if _sigs is None:
        vp, ci, sz = C.c_void_p, C.c_int, C.c_size_t
        sig = {
            "opsmi_serve_create": (vp, [sz, ci, ci, ci]),
            "opsmi_serve_submit": (ci, [vp, C.POINTER(SolveReq)]),
            "opsmi_serve_step": (ci, [vp]),
            "opsmi_serve_completed": (ci, [vp, C.POINTER(SolveDone), ci]),
            "opsmi_serve_fetch": (ci, [vp, ci, C.POINTER(C.c_double)]),
            "opsmi_serve_release": (ci, [vp, ci]),
            "opsmi_serve_bytes_in_use": (sz, [vp]),
            "opsmi_serve_active": (ci, [vp]),
            "opsmi_serve_queued": (ci, [vp]),
            "opsmi_serve_request_bytes": (sz, [C.POINTER(SolveReq)]),
            "opsmi_serve_autotune": (ci, [vp, ci]),
            "opsmi_serve_tile_param": (ci, [vp, ci, ci]),
            "opsmi_serve_destroy": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _sigs = sig
    return L


def request_bytes(kind, n, tol=1e-10, max_iters=100000, mixed=False):
    r = SolveReq(kind, n, 0, tol, max_iters, 1 if mixed else 0)
    v = lib().opsmi_serve_request_bytes(C.byref(r))
    ops._check()
    return v


class Server:
    """Continuous-batching CG engine: submit solves, step, collect results."""

    def __init__(self, budget_bytes=1 << 30, max_slots=64, iters_per_launch=8, use_graph=True):
        self.h = lib().opsmi_serve_create(budget_bytes, max_slots, iters_per_launch, 1 if use_graph else 0)
        ops._check()
        self.shapes = {}

    def submit(self, kind, n, seed=20240611, tol=1e-10, max_iters=100000, mixed=False):
        """mixed: FP32-inner refinement (2D Poisson and 3D heat)."""
        r = SolveReq(kind, n, seed, tol, max_iters, int(mixed))
        i = lib().opsmi_serve_submit(self.h, C.byref(r))
        ops._check()
        self.shapes[i] = (n, n) if kind == POISSON2D else (n, n, n)
        return i

    def step(self):
        v = lib().opsmi_serve_step(self.h)
        ops._check()
        return v

    def completed(self, cap=1024):
        buf = (SolveDone * cap)()
        k = lib().opsmi_serve_completed(self.h, buf, cap)
        ops._check()
        return [dict(id=b.id, status=b.status, iters=b.iters, outer=b.outer, rr0=b.rr0, rr=b.rr) for b in buf[:k]]

    def fetch(self, i):
        out = np.zeros(self.shapes[i])
        lib().opsmi_serve_fetch(self.h, i, out.ctypes.data_as(C.POINTER(C.c_double)))
        ops._check()
        return out

    def release(self, i):
        lib().opsmi_serve_release(self.h, i)
        ops._check()

    def autotune(self, iters=10):
        lib().opsmi_serve_autotune(self.h, iters)
        ops._check()

    def tile_param(self, kind, n):
        v = lib().opsmi_serve_tile_param(self.h, kind, n)
        ops._check()
        return v

    def bytes_in_use(self):
        return lib().opsmi_serve_bytes_in_use(self.h)

    def active(self):
        return lib().opsmi_serve_active(self.h)

    def queued(self):
        return lib().opsmi_serve_queued(self.h)

    def run(self, fetch=True):
        """Steps until nothing is active or queued; returns {id: (done, x)}."""
        out = {}
        while True:
            left = self.step()
            for c in self.completed():
                out[c["id"]] = (c, self.fetch(c["id"]) if fetch else None)
                if not fetch:
                    self.release(c["id"])
            if left == 0:
                break
        return out

    def close(self):
        if self.h:
            lib().opsmi_serve_destroy(self.h)
            self.h = None
