"""ops_tridMultiDimBatch_noalloc: the tridiagonal entry point that allocates
nothing (its c'/d' scratch is a fixed device array of the library), called
through the C ABI and compared bit for bit with ops_tridMultiDimBatch /
ops_tridMultiDimBatch_Inc on the same inputs: two calls in a row, the second
smaller, so the second runs on scratch that still holds the first call's
values; plus the limits it refuses."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

LIMIT_BYTES = 64 << 20


def _dats(ops, blk, n):
    return {k: ops.ops_decl_dat(blk, list(n), name=k) for k in ("a", "b", "c", "d", "u")}


def _solve(ops, D, sd, rng, inc, ctx, noalloc, coeffs, d0, u0):
    for k, v in zip("abc", coeffs):
        ops.ops_dat_set_data(D[k], v)
    ops.ops_dat_set_data(D["d"], d0)
    ops.ops_dat_set_data(D["u"], u0)
    ops.ops_execute()
    if noalloc:
        ops.ops_tridMultiDimBatch_noalloc(3, sd, rng, D["a"], D["b"], D["c"], D["d"], D["u"] if inc else None, ctx)
    elif inc:
        ops.ops_tridMultiDimBatch_Inc(3, sd, rng, D["a"], D["b"], D["c"], D["d"], D["u"], ctx)
    else:
        ops.ops_tridMultiDimBatch(3, sd, rng, D["a"], D["b"], D["c"], D["d"], ctx)
    return ops.ops_dat_fetch_data(D["d"]), ops.ops_dat_fetch_data(D["u"])


@pytest.mark.parametrize("strategy", [0, 3])
@pytest.mark.parametrize("sd", [0, 1, 2])
@pytest.mark.parametrize("inc", [False, True])
def test_noalloc_matches_allocating_entry_points(gpu_ops, strategy, sd, inc):
    ops = gpu_ops
    big, small = (96, 80, 72), (37, 21, 13)
    blk = ops.ops_decl_block(3, f"noalloc_{strategy}_{sd}_{int(inc)}")
    ctx = ops.TridParams(blk, strategy)
    for n, rng_of in ((big, lambda n: [0, n[0], 0, n[1], 0, n[2]]),
                      (small, lambda n: [2, n[0] - 3, 1, n[1] - 2, 3, n[2] - 1])):
        nx, ny, nz = n
        assert 2 * nx * ny * nz * 8 <= LIMIT_BYTES
        coeffs = O.trid_inputs(31 + sd, sd, nx, ny, nz)
        d0 = O.trid_rhs(77 + sd, nx, ny, nz)
        u0 = O.trid_rhs(78 + sd, nx, ny, nz)
        rg = rng_of(n)
        A, B = _dats(ops, blk, n), _dats(ops, blk, n)
        d_ref, u_ref = _solve(ops, A, sd, rg, inc, ctx, False, coeffs, d0, u0)
        d_got, u_got = _solve(ops, B, sd, rg, inc, ctx, True, coeffs, d0, u0)
        assert np.array_equal(d_got, d_ref), (n, rg)
        assert np.array_equal(u_got, u_ref), (n, rg)
        # the solve did something (and, for Thomas, what the oracle does)
        assert not np.array_equal(u_got if inc else d_got, u0 if inc else d0)
        if strategy == 0:
            d_o, u_o = d0.copy(), u0.copy()
            off = rg[0] + nx * (rg[2] + ny * rg[4])
            dims = [rg[1] - rg[0], rg[3] - rg[2], rg[5] - rg[4]]
            flat = lambda q: q.reshape(-1)[off:]
            O.thomas_batch(3, sd, dims, [nx, ny, nz], *(flat(q) for q in coeffs), flat(d_o),
                           flat(u_o) if inc else None, nthreads=4)
            assert np.array_equal(d_got, d_o) and np.array_equal(u_got, u_o)
        for q in list(A.values()) + list(B.values()):
            ops.ops_free_dat(q)


def test_noalloc_refusals(gpu_ops):
    ops = gpu_ops
    blk = ops.ops_decl_block(3, "noalloc_limits")
    ctx = ops.TridParams(blk, 0)
    # one point past the limit: 2 * 8 B * 256 * 256 * 65 > 64 MiB
    n = (256, 256, 65)
    assert 2 * 8 * n[0] * n[1] * n[2] > LIMIT_BYTES
    D = _dats(ops, blk, n)
    full = [0, n[0], 0, n[1], 0, n[2]]
    with pytest.raises(ops.OPSError, match="maximum is 67108864"):
        ops.ops_tridMultiDimBatch_noalloc(3, 2, full, D["a"], D["b"], D["c"], D["d"], None, ctx)
    # a sub-range inside the limit runs
    coeffs = O.trid_inputs(5, 2, *n)
    for k, v in zip("abc", coeffs):
        ops.ops_dat_set_data(D[k], v)
    ops.ops_dat_set_data(D["d"], O.trid_rhs(6, *n))
    ops.ops_execute()
    ops.ops_tridMultiDimBatch_noalloc(3, 2, [0, 256, 0, 256, 0, 64], D["a"], D["b"], D["c"], D["d"], None, ctx)
    # loops still queued are refused (running them might allocate)
    s0 = ops.ops_decl_stencil(3, 1, [0, 0, 0], "s0")
    ops.ops_par_loop("tea_leaf_yeqx_kernel", "yeqx", blk, 3, full,
                     ops.ops_arg_dat(D["u"], 1, s0, "double", ops.OPS_WRITE),
                     ops.ops_arg_dat(D["d"], 1, s0, "double", ops.OPS_READ))
    with pytest.raises(ops.OPSError, match="loops are queued"):
        ops.ops_tridMultiDimBatch_noalloc(3, 0, [0, 64, 0, 64, 0, 8], D["a"], D["b"], D["c"], D["d"], None, ctx)
    ops.ops_execute()
    ops.ops_tridMultiDimBatch_noalloc(3, 0, [0, 64, 0, 64, 0, 8], D["a"], D["b"], D["c"], D["d"], None, ctx)
    ops.ops_mi_sync()
    for q in D.values():
        ops.ops_free_dat(q)
