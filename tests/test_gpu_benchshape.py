"""Parity at the launch shapes the benchmarks time.

The column kernels choose their z-chunk from the grid (the longest power of
two in [8, 64] planes that leaves at least 1024 workgroups; OPSMI_HEAT_ZCHUNK
overrides it), so small test grids run 8-plane chunks while the 512^3 heat CG
and the 768^3 Poisson solves run 64-plane chunks over 2048 and more
workgroups.  These tests reach the timed shapes two ways:

* forced chunks of 64 and 32 planes on small grids whose z extent is not a
  multiple of the chunk (a partial last chunk) and whose x/y extents leave
  partial tiles, compared bit-exact with the oracle loop by loop, with the
  fused chains firing (the lockstep is arranged so the lazy queue still holds
  the deferred update when the operator sweep arrives);
* the benchmark sizes themselves: 512^3 heat CG (the chain kernel with
  64-plane chunks, 2048 workgroups), one FP64 and one FP32 Poisson chain
  launch at 768^3, and the 2048^2 Laplace run for its full 1000 sweeps.

Bars: fields bit-exact (FP contraction off, same association as the
reference kernels); sums within 1e-12 relative."""
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from opshelp import heat_setup_py

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)
RTOL = 1e-12
SEED = 20240611  # the benchmarks' default seed
F32 = np.float32


def relerr(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


class zchunk:
    """Forces the column kernels' z-chunk for the loops queued inside."""

    def __init__(self, planes):
        self.planes = planes

    def __enter__(self):
        self.old = os.environ.get("OPSMI_HEAT_ZCHUNK")
        if self.planes:
            os.environ["OPSMI_HEAT_ZCHUNK"] = str(self.planes)
        return self

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("OPSMI_HEAT_ZCHUNK", None)
        else:
            os.environ["OPSMI_HEAT_ZCHUNK"] = self.old


def _stat(ops, name):
    s = ops.ops_mi_loop_stats(name)
    return s[0] if s else 0


# ------------------------------------------------------------------ heat CG chain
def heat_chain_lockstep(ops, n, K, rhs, iters):
    """TeaLeaf CG (apps/c/TeaLeaf/tea_leaf_cg.cpp:129-214) through the Python
    front end, checked after every iteration against the oracle applied to
    the GPU's own inputs.  No data is read between the deferred axpy and the
    next calc_w, so the queue runs axpy + axpby + calc_w as the fused chain;
    its outputs u, p and w are then compared bit-exact and p.w, r.r within
    1e-12.  Returns the number of chain launches."""
    nx, ny, nz = n
    blk, D, st, red, inner, rro = heat_setup_py(ops, 3, n, K, rhs)
    Kx, Ky, Kz = O.heat_coefs(3, nx, ny, nz, K)
    R, W, RW, INC = ops.OPS_READ, ops.OPS_WRITE, ops.OPS_RW, ops.OPS_INC
    A = lambda q, s, acc: ops.ops_arg_dat(D[q], 1, st[s], "double", acc)
    g = lambda v: ops.ops_arg_gbl(v, 1, "double")

    def calc_w():
        ops.ops_par_loop("tea_leaf_cg_calc_w_reduce_kernel_3d", "w", blk, 3, inner, A("w", "s0", W),
                         A("Kx", "sx", R), A("Ky", "sy", R), A("Kz", "sz", R), A("p", "sp", R), g(1.0), g(1.0),
                         g(1.0), ops.ops_arg_reduce(red, 1, "double", INC))
        return ops.ops_reduction_result(red)

    def check_w(p_ext, pw, tag):
        w_ref, pw_ref = O.heat_apply(3, nx, ny, nz, Kx, Ky, Kz, p_ext, nthreads=NT)
        w = ops.ops_dat_fetch_data(D["w"])
        assert np.array_equal(w, O.interior(w_ref)), tag
        assert relerr(pw, pw_ref) < RTOL, tag
        return w

    ops.ops_mi_reset_stats()
    p_ext = ops.get_ext(D["p"])
    u_prev = ops.ops_dat_fetch_data(D["u"])
    r_prev = ops.ops_dat_fetch_data(D["r"])
    pw = calc_w()
    w_prev = check_w(p_ext, pw, "first calc_w")
    for it in range(iters):
        alpha = rro / pw
        ops.ops_par_loop("tea_leaf_axpy_kernel", "axpy", blk, 3, inner, A("u", "s0", INC), A("p", "s0", R), g(alpha))
        ops.ops_par_loop("tea_leaf_cg_calc_ur_r_reduce_kernel", "ur", blk, 3, inner, A("r", "s0", INC),
                         A("w", "s0", R), g(alpha), ops.ops_arg_reduce(red, 1, "double", INC))
        rrn = ops.ops_reduction_result(red)
        beta = rrn / rro
        ops.ops_par_loop("tea_leaf_axpby_kernel", "axpby", blk, 3, inner, A("p", "s0", RW), A("r", "s0", R),
                         g(beta), g(1.0))
        pw = calc_w()  # the queue now runs axpy + axpby + calc_w
        r1 = ops.ops_dat_fetch_data(D["r"])
        assert np.array_equal(r1, r_prev - alpha * w_prev), it
        assert relerr(rrn, np.sum(r1 * r1)) < RTOL, it
        u1 = ops.ops_dat_fetch_data(D["u"])
        assert np.array_equal(u1, u_prev + alpha * O.interior(p_ext)), it
        p_new = ops.get_ext(D["p"])
        p_ref = p_ext.copy()
        O.interior(p_ref)[:] = beta * O.interior(p_ext) + 1.0 * r1
        assert np.array_equal(p_new, p_ref), it
        w_prev = check_w(p_new, pw, it)
        p_ext, u_prev, r_prev, rro = p_new, u1, r1, rrn
        del u1, p_ref
    chains = _stat(ops, "cg_chain(axpy+axpby+calc_w)")
    for d in D.values():
        ops.ops_free_dat(d)
    return chains


@pytest.mark.parametrize("planes,n", [(64, (130, 20, 150)), (32, (260, 9, 97)), (64, (250, 17, 64))])
def test_heat_chain_forced_chunks_lockstep(gpu_ops, planes, n):
    K, rhs = O.heat_inputs(37, 3, *n)
    with zchunk(planes):
        chains = heat_chain_lockstep(gpu_ops, n, K, rhs, iters=4)
    assert chains == 4


def test_heat_chain_512_lockstep(gpu_ops):
    """The benchmark's own problem (512^3, its seed): the chain kernel with
    64-plane chunks over 4 x 64 x 8 = 2048 workgroups, two iterations."""
    n = (512, 512, 512)
    K, rhs = O.heat_inputs(SEED, 3, *n)
    with zchunk(None):
        chains = heat_chain_lockstep(gpu_ops, n, K, rhs, iters=2)
    assert chains == 2


def test_heat_cg_512_trajectory(gpu_ops):
    """The benchmark's C++ driver at 512^3 (inputs generated by the app from
    the seed), six CG iterations at the default fusion level, against the
    oracle's OpenMP CG on the same inputs: every scalar within 1e-12."""
    from ops_amd import apps
    n = 512
    h = apps.Heat(3, n, n, n, seed=SEED)
    gpu_ops.ops_mi_reset_stats()
    rr0 = h.rr(initial=True)
    k, hist = h.iterate(6, history=True)
    chains = _stat(gpu_ops, "cg_chain(axpy+axpby+calc_w)")
    h.close()
    assert k == 6 and chains == 5
    K, rhs = O.heat_inputs(SEED, 3, n, n, n)
    ref = O.heat_cg(3, n, n, n, K, rhs, max_iters=6, nthreads=NT)
    assert relerr(rr0, ref["rrn"][0]) < RTOL
    for name in ("pw", "alpha", "beta"):
        assert relerr(hist[name], ref[name]) < RTOL, name
    assert relerr(hist["rrn"], ref["rrn"][1:]) < RTOL


# ------------------------------------------------------------------ Poisson chains
def _pois_setup(ops, n, dtype, names):
    blk = ops.ops_decl_block(3, f"pc_{np.dtype(dtype).name}_{n[0]}_{n[2]}")
    s0 = ops.ops_decl_stencil(3, 1, [0, 0, 0], "0")
    s7 = ops.ops_decl_stencil(3, 7, [0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1], "7pt")
    D = {q: ops.ops_decl_dat(blk, list(n), dtype=dtype, name=q) for q in names}
    return blk, s0, s7, D, [0, n[0], 0, n[1], 0, n[2]]


def _apply_f64(n, p_ext):
    """w = A p of the constant-coefficient operator: the oracle's FP64
    residual with f = 0 is exactly -A p."""
    r = O.poisson_residual(3, *n, np.zeros(n[::-1]), p_ext, nthreads=NT)[0]
    return -O.interior(r)


def pois64_chain_lockstep(ops, n, seed, iters):
    """FP64 CG on the Poisson operator with the TeaLeaf vector loops; the
    queue runs axpy + axpby + poisson calc_w as one chain.  Checked as the
    heat chain is."""
    blk, s0, s7, D, rg = _pois_setup(ops, n, np.float64, ("x", "r", "p", "w"))
    R, W, RW, INC = ops.OPS_READ, ops.OPS_WRITE, ops.OPS_RW, ops.OPS_INC
    A = lambda q, s, acc: ops.ops_arg_dat(D[q], 1, s, "double", acc)
    g = lambda v: ops.ops_arg_gbl(v, 1, "double")
    red = ops.ops_decl_reduction_handle(8, "double", "red")
    f = O.poisson_rhs(seed, 3, *n)
    ops.ops_dat_set_data(D["r"], f)
    ops.ops_dat_set_data(D["p"], f)
    rro = float(np.sum(f * f))

    def calc_w():
        ops.ops_par_loop("poisson_calc_w_reduce_kernel", "pw", blk, 3, rg, A("w", s0, W), A("p", s7, R),
                         ops.ops_arg_reduce(red, 1, "double", INC))
        return ops.ops_reduction_result(red)

    def check_w(p_ext, pw, tag):
        w_ref = _apply_f64(n, p_ext)
        w = ops.ops_dat_fetch_data(D["w"])
        assert np.array_equal(w, w_ref), tag
        assert relerr(pw, np.sum(w_ref * O.interior(p_ext))) < RTOL, tag
        return w

    ops.ops_mi_reset_stats()
    p_ext = ops.get_ext(D["p"])
    x_prev = ops.ops_dat_fetch_data(D["x"])
    r_prev = f
    pw = calc_w()
    w_prev = check_w(p_ext, pw, "first calc_w")
    for it in range(iters):
        alpha = rro / pw
        ops.ops_par_loop("tea_leaf_axpy_kernel", "axpy", blk, 3, rg, A("x", s0, INC), A("p", s0, R), g(alpha))
        ops.ops_par_loop("tea_leaf_cg_calc_ur_r_reduce_kernel", "ur", blk, 3, rg, A("r", s0, INC), A("w", s0, R),
                         g(alpha), ops.ops_arg_reduce(red, 1, "double", INC))
        rrn = ops.ops_reduction_result(red)
        beta = rrn / rro
        ops.ops_par_loop("tea_leaf_axpby_kernel", "axpby", blk, 3, rg, A("p", s0, RW), A("r", s0, R), g(beta),
                         g(1.0))
        pw = calc_w()
        r1 = ops.ops_dat_fetch_data(D["r"])
        assert np.array_equal(r1, r_prev - alpha * w_prev), it
        assert relerr(rrn, np.sum(r1 * r1)) < RTOL, it
        x1 = ops.ops_dat_fetch_data(D["x"])
        assert np.array_equal(x1, x_prev + alpha * O.interior(p_ext)), it
        del x1
        p_new = ops.get_ext(D["p"])
        p_ref = p_ext.copy()
        O.interior(p_ref)[:] = beta * O.interior(p_ext) + 1.0 * r1
        assert np.array_equal(p_new, p_ref), it
        del p_ref
        w_prev = check_w(p_new, pw, it)
        x_prev = ops.ops_dat_fetch_data(D["x"])
        p_ext, r_prev, rro = p_new, r1, rrn
    chains = _stat(ops, "poisson_chain(axpy+axpby+calc_w)")
    for d in D.values():
        ops.ops_free_dat(d)
    return chains


def pois32_chain_lockstep(ops, n, seed, iters):
    """The mixed solver's FP32 inner iteration (rr update with its FP64 dots,
    e update, p update, FP32 operator sweep); the queue holds the e update
    past the dots result and runs e + p + calc_w as one chain.  FP32 values
    are compared bit-exact with numpy float32 arithmetic and the oracle's
    FP32 sweep, FP64 dots within 1e-12."""
    blk, s0, s7, D, rg = _pois_setup(ops, n, F32, ("e", "rr", "p", "w"))
    R, RW, W, INC = ops.OPS_READ, ops.OPS_RW, ops.OPS_WRITE, ops.OPS_INC
    A = lambda q, s, acc: ops.ops_arg_dat(D[q], 1, s, "float", acc)
    gf = lambda v: ops.ops_arg_gbl(F32(v), 1, "float")
    red = ops.ops_decl_reduction_handle(8, "double", "pw")
    dots = ops.ops_decl_reduction_handle(16, "double", "dots")
    mi = F32(1.0 / 6.0)
    gen = np.random.default_rng(seed)
    shape = n[::-1]
    rr_h = gen.standard_normal(shape).astype(F32)
    p_h = (mi * rr_h).astype(F32)
    ops.ops_dat_set_data(D["rr"], rr_h)
    ops.ops_dat_set_data(D["p"], p_h)
    ops.ops_dat_set_data(D["e"], np.zeros(shape, F32))
    rz = float(np.sum(rr_h.astype(np.float64) * p_h.astype(np.float64)))
    del p_h

    def calc_w():
        ops.ops_par_loop("poisson_calc_w_reduce_kernel", "pw", blk, 3, rg, A("w", s0, W), A("p", s7, R),
                         ops.ops_arg_reduce(red, 1, "double", INC))
        return ops.ops_reduction_result(red)

    def check_w(p_ext, pw, tag):
        w_ref, pw_ref = O.poisson_apply_f32(3, *n, p_ext)
        w = ops.ops_dat_fetch_data(D["w"])
        assert np.array_equal(w, O.interior(w_ref)), tag
        assert relerr(pw, pw_ref) < RTOL, tag
        return w

    ops.ops_mi_reset_stats()
    p_ext = ops.get_ext(D["p"])
    e_prev = ops.ops_dat_fetch_data(D["e"])
    pw = calc_w()
    w_prev = check_w(p_ext, pw, "first calc_w")
    for it in range(iters):
        a = F32(rz / pw)
        ops.ops_par_loop("mp_rr_update_kernel", "rr", blk, 3, rg, A("rr", s0, RW), A("w", s0, R), gf(a), gf(mi),
                         ops.ops_arg_reduce(dots, 2, "double", INC))
        ops.ops_par_loop("mp_e_update_kernel", "e", blk, 3, rg, A("e", s0, RW), A("p", s0, R), gf(a))
        d = ops.ops_reduction_result(dots)
        b = F32(d[0] / rz)
        ops.ops_par_loop("mp_p_kernel", "p", blk, 3, rg, A("p", s0, RW), A("rr", s0, R), gf(mi), gf(b))
        pw = calc_w()  # the queue now runs e + p + calc_w
        rr1 = ops.ops_dat_fetch_data(D["rr"])
        assert np.array_equal(rr1, rr_h - a * w_prev), it
        r64 = rr1.astype(np.float64)
        assert relerr(d[0], np.sum(r64 * (mi * rr1).astype(np.float64))) < RTOL, it
        assert relerr(d[1], np.sum(r64 * r64)) < RTOL, it
        del r64
        e1 = ops.ops_dat_fetch_data(D["e"])
        assert np.array_equal(e1, e_prev + a * O.interior(p_ext)), it
        p_new = ops.get_ext(D["p"])
        p_ref = p_ext.copy()
        O.interior(p_ref)[:] = mi * rr1 + b * O.interior(p_ext)
        assert p_ref.dtype == F32 and np.array_equal(p_new, p_ref), it
        del p_ref
        w_prev = check_w(p_new, pw, it)
        p_ext, e_prev, rr_h, rz = p_new, e1, rr1, d[0]
    chains = _stat(ops, "mp_chain(e+p+calc_w)")
    for q in D.values():
        ops.ops_free_dat(q)
    return chains


@pytest.mark.parametrize("planes,n", [(64, (130, 20, 150)), (32, (260, 9, 97)), (64, (384, 12, 70))])
def test_poisson_chains_forced_chunks_lockstep(gpu_ops, planes, n):
    with zchunk(planes):
        c64 = pois64_chain_lockstep(gpu_ops, n, 41, iters=3)
        c32 = pois32_chain_lockstep(gpu_ops, n, 43, iters=3)
    assert c64 == 3 and c32 == 3


def test_poisson_chain_768_f64(gpu_ops):
    """The 768^3 FP64 CG's shape: one chain launch with 64-plane chunks over
    6 x 96 x 12 workgroups."""
    with zchunk(None):
        assert pois64_chain_lockstep(gpu_ops, (768, 768, 768), SEED, iters=1) == 1


def test_poisson_chain_768_f32(gpu_ops):
    """The 768^3 mixed solver's FP32 inner iteration: one chain launch with
    64-plane chunks over 3 x 96 x 12 workgroups."""
    with zchunk(None):
        assert pois32_chain_lockstep(gpu_ops, (768, 768, 768), SEED, iters=1) == 1


# ------------------------------------------------------------------ Laplace
def test_laplace_2048_1000_bit_exact(gpu_ops):
    """The Laplace workload as timed: 2048^2, 1000 sweeps (each copy fused
    with the next apply_stencil), every max-abs-diff and the final field
    bit-exact against the oracle."""
    from ops_amd import apps
    nx = ny = 2048
    it = 1000
    init = O.laplace_init(nx, ny, 0)
    gpu_ops.ops_mi_reset_stats()
    err, hist, A = apps.laplace2d(nx, ny, it, 0, init)
    assert _stat(gpu_ops, "jacobi_chain(copy+apply_stencil)") == it - 1
    A_ref, h_ref = O.laplace_run(init, it, nthreads=NT)
    assert np.array_equal(hist, h_ref)
    assert np.array_equal(A, O.interior(A_ref))


# ------------------------------------------------------------------ 27-point
@pytest.mark.parametrize("planes,n", [(64, (130, 20, 150)), (32, (256, 9, 97))])
def test_stencil27_forced_chunks(gpu_ops, planes, n):
    ops = gpu_ops
    nx, ny, nz = n
    blk = ops.ops_decl_block(3, f"s27c_{nx}")
    s000 = ops.ops_decl_stencil(3, 1, [0, 0, 0], "000")
    pts = [c for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1) for c in (dx, dy, dz)]
    s27 = ops.ops_decl_stencil(3, 27, pts, "27pt")
    u = ops.ops_decl_dat(blk, [nx, ny, nz], name="u27")
    o = ops.ops_decl_dat(blk, [nx, ny, nz], name="o27")
    ue = O.u01(43, nx, (nx + 2) * (ny + 2) * (nz + 2)).reshape(nz + 2, ny + 2, nx + 2) - 0.5
    ops.set_ext(u, ue)
    with zchunk(planes):
        ops.ops_par_loop("stencil27_kernel", "stencil27", blk, 3, [0, nx, 0, ny, 0, nz],
                         ops.ops_arg_dat(o, 1, s000, "double", ops.OPS_WRITE),
                         ops.ops_arg_dat(u, 1, s27, "double", ops.OPS_READ))
        out = ops.ops_dat_fetch_data(o)
    assert np.array_equal(out, O.stencil27(nx, ny, nz, ue))
    ops.ops_free_dat(u)
    ops.ops_free_dat(o)
