"""Mixed-precision Poisson: FP64 iterative refinement around an FP32
Jacobi-preconditioned CG whose dots accumulate in FP64.

Kernels: every FP32 value is bit-identical to numpy float32 arithmetic (one
rounding per operation, no contraction); FP64 values bit-identical to the
oracle's; dots within 1e-12 relative (summation order differs).  Solver: the
FP32 trajectory only depends on the dots through float casts of their ratios,
so iteration counts and the solution match the oracle restatement exactly
(up to a last-bit flip of a cast, not seen in practice); the final FP64
residual is recomputed on the CPU, and the error against the FP64 CG
solution is bounded."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

F32 = np.float32


def _setup(ops, n):
    nd = len(n)
    blk = ops.ops_decl_block(nd, "mp")
    s0 = ops.ops_decl_stencil(nd, 1, [0] * nd, "0")
    rg = []
    for v in n:
        rg += [0, v]
    return blk, s0, rg


def _rel(a, b):
    return abs(a - b) / abs(b)


def test_mp_update_and_p_bit_exact(gpu_ops):
    ops = gpu_ops
    n = (70, 33, 9)
    blk, s0, rg = _setup(ops, n)
    g = np.random.default_rng(5)
    vals = {k: g.standard_normal(n[::-1]).astype(F32) for k in ("e", "rr", "p", "w")}
    D = {k: ops.ops_decl_dat(blk, list(n), dtype=F32, name=k) for k in vals}
    for k, v in vals.items():
        ops.ops_dat_set_data(D[k], v)
    dots = ops.ops_decl_reduction_handle(16, "double", "dots")
    a, mi, b = F32(0.3711), F32(1.0 / 6.0), F32(0.8123)
    R, RW = ops.OPS_READ, ops.OPS_RW
    ops.ops_par_loop("mp_update_kernel", "mp_update_kernel", blk, 3, rg,
                     ops.ops_arg_dat(D["e"], 1, s0, "float", RW), ops.ops_arg_dat(D["rr"], 1, s0, "float", RW),
                     ops.ops_arg_dat(D["p"], 1, s0, "float", R), ops.ops_arg_dat(D["w"], 1, s0, "float", R),
                     ops.ops_arg_gbl(a, 1, "float"), ops.ops_arg_gbl(mi, 1, "float"),
                     ops.ops_arg_reduce(dots, 2, "double", ops.OPS_INC))
    got = ops.ops_reduction_result(dots)
    e = vals["e"] + a * vals["p"]
    rr = vals["rr"] - a * vals["w"]
    z = mi * rr
    assert e.dtype == F32 and rr.dtype == F32
    assert np.array_equal(ops.ops_dat_fetch_data(D["e"]), e)
    assert np.array_equal(ops.ops_dat_fetch_data(D["rr"]), rr)
    r64 = rr.astype(np.float64)
    assert _rel(got[0], np.sum(r64 * z.astype(np.float64))) <= 1e-12
    assert _rel(got[1], np.sum(r64 * r64)) <= 1e-12
    ops.ops_par_loop("mp_p_kernel", "mp_p_kernel", blk, 3, rg, ops.ops_arg_dat(D["p"], 1, s0, "float", RW),
                     ops.ops_arg_dat(D["rr"], 1, s0, "float", R), ops.ops_arg_gbl(mi, 1, "float"),
                     ops.ops_arg_gbl(b, 1, "float"))
    assert np.array_equal(ops.ops_dat_fetch_data(D["p"]), mi * rr + b * vals["p"])
    for d in D.values():
        ops.ops_free_dat(d)


def test_mp_init_and_correct_bit_exact(gpu_ops):
    ops = gpu_ops
    n = (45, 38)
    blk, s0, rg = _setup(ops, n)
    g = np.random.default_rng(6)
    r = g.standard_normal(n[::-1])
    ev = g.standard_normal(n[::-1]).astype(F32)
    xv = g.standard_normal(n[::-1])
    Dr = ops.ops_decl_dat(blk, list(n), name="r")
    Dx = ops.ops_decl_dat(blk, list(n), name="x")
    De, Drr, Dp = (ops.ops_decl_dat(blk, list(n), dtype=F32, name=k) for k in ("e", "rr", "p"))
    ops.ops_dat_set_data(Dr, r)
    ops.ops_dat_set_data(Dx, xv)
    ops.ops_dat_set_data(De, ev)
    dots = ops.ops_decl_reduction_handle(16, "double", "dots")
    mi = F32(0.25)
    W, R = ops.OPS_WRITE, ops.OPS_READ
    ops.ops_par_loop("mp_init_kernel", "mp_init_kernel", blk, 2, rg, ops.ops_arg_dat(De, 1, s0, "float", W),
                     ops.ops_arg_dat(Drr, 1, s0, "float", W), ops.ops_arg_dat(Dp, 1, s0, "float", W),
                     ops.ops_arg_dat(Dr, 1, s0, "double", R), ops.ops_arg_gbl(mi, 1, "float"),
                     ops.ops_arg_reduce(dots, 2, "double", ops.OPS_INC))
    got = ops.ops_reduction_result(dots)
    rr = r.astype(F32)
    assert np.all(ops.ops_dat_fetch_data(De) == 0)
    assert np.array_equal(ops.ops_dat_fetch_data(Drr), rr)
    assert np.array_equal(ops.ops_dat_fetch_data(Dp), mi * rr)
    r64 = rr.astype(np.float64)
    assert _rel(got[0], np.sum(r64 * (mi * rr).astype(np.float64))) <= 1e-12
    assert _rel(got[1], np.sum(r64 * r64)) <= 1e-12
    ops.ops_dat_set_data(De, ev)
    ops.ops_par_loop("mp_correct_kernel", "mp_correct_kernel", blk, 2, rg, ops.ops_arg_dat(Dx, 1, s0, "double",
                                                                                             ops.OPS_RW),
                     ops.ops_arg_dat(De, 1, s0, "float", R))
    assert np.array_equal(ops.ops_dat_fetch_data(Dx), xv + ev.astype(np.float64))
    for d in (Dr, Dx, De, Drr, Dp):
        ops.ops_free_dat(d)


@pytest.mark.parametrize("n", [(31, 27, 20), (60, 47)])
def test_residual_kernel_matches_oracle(gpu_ops, n):
    ops = gpu_ops
    nd = len(n)
    nx, ny, nz = n[0], n[1], (n[2] if nd == 3 else 1)
    blk, s0, rg = _setup(ops, n)
    sp = ops.ops_decl_stencil(nd, 2 * nd + 1, O_pts(nd), "sp")
    g = np.random.default_rng(7)
    x_ext = g.standard_normal(O.ext_shape(nd, nx, ny, nz))
    f = g.standard_normal(n[::-1])
    Dx, Df, Dr = (ops.ops_decl_dat(blk, list(n), name=k) for k in ("x", "f", "r"))
    ops.set_ext(Dx, x_ext)
    ops.ops_dat_set_data(Df, f)
    red = ops.ops_decl_reduction_handle(8, "double", "rr")
    ops.ops_par_loop("poisson_residual_kernel", "poisson_residual_kernel", blk, nd, rg,
                     ops.ops_arg_dat(Dr, 1, s0, "double", ops.OPS_WRITE),
                     ops.ops_arg_dat(Df, 1, s0, "double", ops.OPS_READ),
                     ops.ops_arg_dat(Dx, 1, sp, "double", ops.OPS_READ),
                     ops.ops_arg_reduce(red, 1, "double", ops.OPS_INC))
    rr = ops.ops_reduction_result(red)
    r_ref, rr_ref = O.poisson_residual(nd, nx, ny, nz, f, x_ext)
    assert np.array_equal(ops.ops_dat_fetch_data(Dr), O.interior(r_ref))
    assert _rel(rr, rr_ref) <= 1e-12
    for d in (Dx, Df, Dr):
        ops.ops_free_dat(d)


def O_pts(nd):
    if nd == 2:
        return [0, 0, 1, 0, -1, 0, 0, 1, 0, -1]
    return [0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1]


@pytest.mark.parametrize("n", [(40, 36, 30), (210, 170)])
def test_mixed_solve_matches_oracle(gpu_ops, n):
    from ops_amd import apps
    nd = len(n)
    nx, ny, nz = n[0], n[1], (n[2] if nd == 3 else 1)
    f = O.poisson_rhs(20240611, nd, nx, ny, nz)
    ref = O.poisson_mixed(nd, nx, ny, nz, f, tol=1e-10, nthreads=8)
    m = apps.PoissonMixed(nd, nx, ny, nz, f=f)
    got = m.solve(tol=1e-10)
    x = m.x()
    m.close()
    assert got["outer"] == ref["outer"] and got["inner_total"] == ref["inner_total"]
    np.testing.assert_allclose(got["relres"], ref["relres"], rtol=1e-10, atol=0)
    assert got["relres"][-1] <= 1e-10
    assert np.max(np.abs(x - ref["x"])) <= 1e-12 * np.max(np.abs(ref["x"]))
    # FP64 residual recomputed on the CPU from the GPU solution
    x_ext = np.zeros(O.ext_shape(nd, nx, ny, nz))
    x_ext[(slice(1, -1),) * nd] = x
    _, rr = O.poisson_residual(nd, nx, ny, nz, f, x_ext)
    assert np.sqrt(rr) / np.linalg.norm(f) <= 1e-10
    # error against FP64 CG converged further
    cg = O.poisson_cg(nd, nx, ny, nz, f, max_iters=5000, tol=1e-13, nthreads=8)
    err = np.max(np.abs(x - cg["x"])) / np.max(np.abs(cg["x"]))
    assert err <= 1e-8
