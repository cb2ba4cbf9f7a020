"""Shared set-up for GPU tests: the TeaLeaf CG set-up loops through the
Python front end (the same loops, in the same order, as the C++ heat app,
apps/c/TeaLeaf/tea_leaf_common.cpp:76-146 and tea_leaf_cg.cpp:54-90)."""
import numpy as np  # noqa: F401


def heat_setup_py(ops, ndim, n, K_ext, rhs):
    """TeaLeaf setup through the Python front end (same loops as the C++ app)."""
    nx, ny, nz = n
    blk = ops.ops_decl_block(ndim, "tea_lockstep")
    size = [nx, ny] if ndim == 2 else [nx, ny, nz]
    D = {q: ops.ops_decl_dat(blk, size, name=q) for q in ("density", "u", "w", "Kx", "Ky", "Kz", "r", "p")}
    if ndim == 3:
        S = dict(s0=[0, 0, 0], sp=[0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1],
                 sx=[0, 0, 0, 1, 0, 0], sy=[0, 0, 0, 0, 1, 0], sz=[0, 0, 0, 0, 0, 1],
                 sm=[0, 0, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1])
    else:
        S = dict(s0=[0, 0], sp=[0, 0, 1, 0, -1, 0, 0, 1, 0, -1], sx=[0, 0, 1, 0], sy=[0, 0, 0, 1],
                 sm=[0, 0, -1, 0, 0, -1])
    st = {k: ops.ops_decl_stencil(ndim, len(v) // ndim, v, k) for k, v in S.items()}
    red = ops.ops_decl_reduction_handle(8, "double", "red")
    ops.set_ext(D["density"], K_ext)
    ops.ops_dat_set_data(D["u"], rhs)
    full = [-1, nx + 1, -1, ny + 1, -1, nz + 1][: 2 * ndim]
    inner = [0, nx, 0, ny, 0, nz][: 2 * ndim]
    ext = [0, nx + 1, 0, ny + 1, 0, nz + 1][: 2 * ndim]
    A = lambda q, s, acc: ops.ops_arg_dat(D[q], 1, st[s], "double", acc)
    R, W, RW, INC = ops.OPS_READ, ops.OPS_WRITE, ops.OPS_RW, ops.OPS_INC
    ops.ops_par_loop("tea_leaf_recip_kernel", "recip", blk, ndim, full, A("w", "s0", W), A("density", "s0", R))
    if ndim == 3:
        ops.ops_par_loop("tea_leaf_common_init_Kx_Ky_Kz_kernel", "initK", blk, 3, ext, A("Kx", "s0", W),
                         A("Ky", "s0", W), A("Kz", "s0", W), A("w", "sm", R))
    else:
        ops.ops_par_loop("tea_leaf_common_init_Kx_Ky_kernel", "initK", blk, 2, ext, A("Kx", "s0", W),
                         A("Ky", "s0", W), A("w", "sm", R))
    for q, name in enumerate(["Kx", "Ky", "Kz"][:ndim]):
        lo, hi = list(full), list(full)
        lo[2 * q:2 * q + 2] = [-1, 1]
        hi[2 * q:2 * q + 2] = [n[q], n[q] + 1]
        for rg in (lo, hi):
            ops.ops_par_loop("tea_leaf_init_zero_kernel", "zero", blk, ndim, rg, A(name, "s0", W))
    one = ops.ops_arg_gbl(1.0, 1, "double")
    if ndim == 3:
        ops.ops_par_loop("tea_leaf_common_init_kernel_3d", "init", blk, 3, inner, A("w", "s0", W), A("r", "s0", W),
                         A("Kx", "sx", R), A("Ky", "sy", R), A("Kz", "sz", R), A("u", "sp", R), one,
                         ops.ops_arg_gbl(1.0, 1, "double"), ops.ops_arg_gbl(1.0, 1, "double"))
    else:
        ops.ops_par_loop("tea_leaf_common_init_kernel", "init", blk, 2, inner, A("w", "s0", W), A("r", "s0", W),
                         A("Kx", "sx", R), A("Ky", "sy", R), A("u", "sp", R), one, ops.ops_arg_gbl(1.0, 1, "double"))
    ops.ops_par_loop("tea_leaf_yeqx_kernel", "yeqx", blk, ndim, inner, A("p", "s0", W), A("r", "s0", R))
    ops.ops_par_loop("tea_leaf_dot_kernel", "dot", blk, ndim, inner, A("r", "s0", R), A("p", "s0", R),
                     ops.ops_arg_reduce(red, 1, "double", INC))
    rro = ops.ops_reduction_result(red)
    return blk, D, st, red, inner, rro
