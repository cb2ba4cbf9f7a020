"""Lazy execution and run-time loop-chain fusion.

A reduction result runs only the queued loops it depends on; data access runs
everything.  The fused CG chain (axpy + axpby + calc_w) and pair (axpy +
axpby) must give the same bits as the loops run one by one."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu


def _setup(ops, n=(37, 29)):
    blk = ops.ops_decl_block(2, "lazy")
    s00 = ops.ops_decl_stencil(2, 1, [0, 0], "00")
    A = ops.ops_decl_dat(blk, list(n), name="A")
    B = ops.ops_decl_dat(blk, list(n), name="B")
    v = O.u01(3, 9, n[0] * n[1]).reshape(n[1], n[0])
    ops.ops_dat_set_data(A, v)
    h = ops.ops_decl_reduction_handle(8, "double", "s")
    rng = [0, n[0], 0, n[1]]
    return blk, s00, A, B, v, h, rng


def test_deferred_loop_not_needed_by_result(gpu_ops):
    ops = gpu_ops
    blk, s00, A, B, v, h, rng = _setup(ops)
    R, W = ops.OPS_READ, ops.OPS_WRITE
    ops.ops_par_loop("copy", "copy", blk, 2, rng, ops.ops_arg_dat(B, 1, s00, "double", W),
                     ops.ops_arg_dat(A, 1, s00, "double", R))
    ops.ops_par_loop("set_zero", "zero", blk, 2, rng, ops.ops_arg_dat(A, 1, s00, "double", W))
    ops.ops_par_loop("red_sum", "sum", blk, 2, rng, ops.ops_arg_dat(B, 1, s00, "double", R),
                     ops.ops_arg_reduce(h, 1, "double", ops.OPS_INC))
    s = ops.ops_reduction_result(h)
    assert abs(s - v.sum()) <= 1e-12 * abs(v.sum())
    assert np.all(ops.ops_dat_fetch_data(A) == 0.0)
    assert np.array_equal(ops.ops_dat_fetch_data(B), v)


def test_result_runs_its_dependencies_in_order(gpu_ops):
    ops = gpu_ops
    blk, s00, A, B, v, h, rng = _setup(ops)
    R, W = ops.OPS_READ, ops.OPS_WRITE
    ops.ops_par_loop("set_zero", "zero", blk, 2, rng, ops.ops_arg_dat(A, 1, s00, "double", W))
    ops.ops_par_loop("copy", "copy", blk, 2, rng, ops.ops_arg_dat(B, 1, s00, "double", W),
                     ops.ops_arg_dat(A, 1, s00, "double", R))
    ops.ops_par_loop("red_sum", "sum", blk, 2, rng, ops.ops_arg_dat(B, 1, s00, "double", R),
                     ops.ops_arg_reduce(h, 1, "double", ops.OPS_INC))
    assert ops.ops_reduction_result(h) == 0.0


def test_write_after_read_kept_in_order(gpu_ops):
    # a later loop overwrites what an earlier queued loop reads: the earlier
    # one must run first even though only the later one is needed
    ops = gpu_ops
    blk, s00, A, B, v, h, rng = _setup(ops)
    R, W = ops.OPS_READ, ops.OPS_WRITE
    h2 = ops.ops_decl_reduction_handle(8, "double", "s2")
    ops.ops_par_loop("red_sum", "sumA", blk, 2, rng, ops.ops_arg_dat(A, 1, s00, "double", R),
                     ops.ops_arg_reduce(h2, 1, "double", ops.OPS_INC))
    ops.ops_par_loop("copy", "copy", blk, 2, rng, ops.ops_arg_dat(B, 1, s00, "double", W),
                     ops.ops_arg_dat(A, 1, s00, "double", R))
    ops.ops_par_loop("set_zero", "zero", blk, 2, rng, ops.ops_arg_dat(A, 1, s00, "double", W))
    ops.ops_par_loop("red_sum", "sumB", blk, 2, rng, ops.ops_arg_dat(B, 1, s00, "double", R),
                     ops.ops_arg_reduce(h, 1, "double", ops.OPS_INC))
    sB = ops.ops_reduction_result(h)
    sA = ops.ops_reduction_result(h2)
    assert abs(sA - v.sum()) <= 1e-12 * abs(v.sum())
    assert abs(sB - v.sum()) <= 1e-12 * abs(v.sum())


@pytest.mark.parametrize("n", [(48, 40, 33), (130, 66, 40), (131, 37, 21)])
def test_fused_cg_chain_bit_identical(gpu_ops, n):
    from ops_amd import apps
    ops = gpu_ops
    K, rhs = O.heat_inputs(23, 3, *n)
    runs = []
    for fusion in (0, 1, 2):
        ops.lib().ops_mi_set_fusion(fusion)
        try:
            h = apps.Heat(3, *n, K_ext=K, rhs=rhs)
            ops.ops_mi_reset_stats()
            k, hist = h.iterate(30, history=True)
            assert k == 30
            st = [ops.ops_mi_loop_stats(name) for name in ("cg_pair(axpy+axpby)", "cg_chain(axpy+axpby+calc_w)")]
            u = h.u()
            h.close()
        finally:
            ops.lib().ops_mi_set_fusion(2)
        runs.append((hist, u, [s[0] if s else 0 for s in st]))
    (h0, u0, st0), (h1, u1, st1), (h2, u2, st2) = runs
    # each iteration's axpy waits in the queue past the r.r result until the
    # axpby that overwrites p; at level 1 the two run as the pointwise pair;
    # at level 2 they join the next calc_w in the three-loop chain where every
    # array takes 16-byte lanes (rows of at least 1 KiB are padded and
    # aligned); the last iteration's axpy + axpby is always a pair
    assert st0 == [0, 0]
    assert st1 == [30, 0]
    assert st2[0] + st2[1] == 30
    if n[0] + 2 >= 128:
        assert st2 == [1, 29]
    for hist, u in ((h1, u1), (h2, u2)):
        for key in ("pw", "alpha", "beta", "rrn"):
            assert np.array_equal(h0[key], hist[key]), key
        assert np.array_equal(u0, u)


@pytest.mark.parametrize("n", [(48, 40, 33), (130, 66, 40), (260, 37, 21)])
def test_fused_poisson_chains_bit_identical(gpu_ops, n):
    """The constant-coefficient chains: FP64 CG (axpy + axpby + Poisson
    calc_w) and the mixed solver's FP32 inner iterations (e update + p update
    + calc_w) give the same bits at fusion levels 0, 1 and 2."""
    from ops_amd import apps
    ops = gpu_ops
    f = O.poisson_rhs(31, 3, *n)
    runs = []
    for fusion in (0, 1, 2):
        ops.lib().ops_mi_set_fusion(fusion)
        try:
            ops.ops_mi_reset_stats()
            P = apps.Poisson(3, *n, f=f)
            k, hist = P.iterate(25, history=True)
            x = P.x()
            P.close()
            M = apps.PoissonMixed(3, *n, f=f)
            res = M.solve(tol=1e-9)
            xm = M.x()
            M.close()
            st = [ops.ops_mi_loop_stats(nm) for nm in ("poisson_chain(axpy+axpby+calc_w)", "mp_chain(e+p+calc_w)")]
        finally:
            ops.lib().ops_mi_set_fusion(2)
        runs.append((hist, x, res, xm, [s[0] if s else 0 for s in st]))
    (h0, x0, r0, m0, s0), _, (h2, x2, r2, m2, s2) = runs
    assert s0 == [0, 0] and runs[1][4] == [0, 0]
    # every array takes 16-byte lanes (rows of at least 1 KiB are padded)
    if n[0] + 2 >= 128:
        assert s2[0] >= 23 and s2[1] > 0, s2
    for _, x, r, xm, _ in runs[1:]:
        assert np.array_equal(x0, x) and np.array_equal(m0, xm)
        assert r["inner_total"] == r0["inner_total"] and np.array_equal(r["relres"], r0["relres"])
    for key in ("rrn", "alpha", "beta"):
        assert np.array_equal(h0[key], h2[key]), key


@pytest.mark.parametrize("n", [(130, 90), (61, 47), (256, 129), (2048, 70)])
def test_fused_jacobi_step_bit_identical(gpu_ops, n):
    """The Laplace tutorial's copy and next apply_stencil run as one pass at
    fusion level 2 (rows that allow 16-byte lanes); field and per-iteration
    max-change history are the same bits at levels 0, 1 and 2."""
    from ops_amd import apps
    ops = gpu_ops
    nx, ny = n
    runs = []
    for fusion in (0, 1, 2):
        ops.lib().ops_mi_set_fusion(fusion)
        try:
            ops.ops_mi_reset_stats()
            err, hist, A = apps.laplace2d(nx, ny, 30, 0, O.laplace_init(nx, ny, 0))
            st = ops.ops_mi_loop_stats("jacobi_chain(copy+apply_stencil)")
        finally:
            ops.lib().ops_mi_set_fusion(2)
        runs.append((hist, A, st[0] if st else 0))
    (h0, A0, s0), (h1, A1, s1), (h2, A2, s2) = runs
    assert s0 == 0 and s1 == 0
    if ((nx + 2) * 8) % 16 == 0 or (nx + 2) * 8 >= 1024:
        assert s2 == 29
    for h, A in ((h1, A1), (h2, A2)):
        assert np.array_equal(h0, h) and np.array_equal(A0, A)
    ref_A, ref_h = O.laplace_run(O.laplace_init(nx, ny, 0), 30)
    assert np.array_equal(h2, ref_h) and np.array_equal(A2, O.interior(ref_A))
