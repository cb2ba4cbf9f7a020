"""Translator-generated loops on the GPU.

tests/genapps/*.cpp are applications written in the OPS dev API
(ops_par_loop(kernel, ...) with ACC<T> kernels); __graft_entry__.build()
runs them through ops_amd/translator/opsmi_translate.py into
tests/genapps/libgenapps.so.  Their results must match the oracle (Laplace:
bit-exact), closed forms computed with numpy in the kernels' own operation
order (fields), and the hand-written entry points of the same kernels
(TeaLeaf calc_w: field bit-exact, p.w within 1e-12 relative)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

GEN_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "genapps", "libgenapps.so")
vp, ci, cd = C.c_void_p, C.c_int, C.c_double
dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def gen(gpu_ops):
    if not os.path.exists(GEN_LIB):
        raise RuntimeError(f"{GEN_LIB} is not built; run __graft_entry__.build() first")
    L = C.CDLL(GEN_LIB)
    L.gen_laplace.argtypes = [vp, vp, vp, ci, ci, ci, dp]
    L.gen_laplace_bad_stencil.argtypes = [vp, vp, vp, ci, ci]
    L.gen_fields.argtypes = [vp, vp, vp, vp, ci, ci, cd, cd, cd, ci, dp, dp]
    L.gen_heat_step.argtypes = [vp] * 7 + [ci, ci, ci, cd, cd, cd, cd, dp]
    L.gen_ramp1d.argtypes = [vp, vp, vp, ci, C.c_float, C.c_float, C.POINTER(C.c_float)]
    L.gen_vec3.argtypes = [vp, vp, vp, ci, ci, ci, dp, C.POINTER(C.c_int)]
    L.gen_opt_add.argtypes = [vp, vp, vp, ci, ci, ci, ci]
    L.gen_fill2d.argtypes = [vp, vp, ci, ci, cd]
    L.gen_double_in_place.argtypes = [vp, vp, ci, ci, ci]
    L.gen_cg_tail.argtypes = [vp, vp, vp, vp, ci, ci, ci, cd, cd, ci, dp]
    L.gen_cg_three.argtypes = [vp, vp, vp, vp, vp, ci, ci, ci, cd, cd, ci, ci, dp, dp]
    L.gen_lowdim3d.argtypes = [vp, ci, ci, ci, dp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_float), dp]
    L.gen_lowdim2d.argtypes = [vp, ci, ci, dp]
    L.gen_lowdim_bad.argtypes = [vp, ci, ci, ci, ci]
    ip = C.POINTER(C.c_int)
    L.gen_chain_src_lap.argtypes = [vp, vp, vp, vp, ip, ip, dp]
    L.gen_chain_mix_avg.argtypes = [vp, vp, vp, vp, ci, ci, ci, dp]
    L.gen_chain_refused.argtypes = [vp, vp, vp, vp, vp, ci, ci, ci]
    L.gen_legacy_lap2d.argtypes = [vp, vp, vp, ci, ci, ci]
    L.gen_legacy_md3d.argtypes = [vp, vp, vp, ci, ci, ci, ci, dp]
    return L


def _p(a):
    return a.ctypes.data_as(dp)


@pytest.mark.parametrize("n", [(64, 48), (130, 70), (257, 33)])
def test_generated_laplace_bit_exact(gpu_ops, gen, n):
    ops = gpu_ops
    nx, ny = n
    it = 25
    init = O.laplace_init(nx, ny, 0)
    blk = ops.ops_decl_block(2, "genlap")
    A = ops.ops_decl_dat(blk, [nx, ny], name="A")
    An = ops.ops_decl_dat(blk, [nx, ny], name="Anew")
    ops.set_ext(A, init)
    ops.set_ext(An, init)
    hist = np.zeros(it)
    ops.ops_mi_reset_stats()
    gen.gen_laplace(blk.ptr, A.ptr, An.ptr, nx, ny, it, _p(hist))
    ops._check()
    ref_A, ref_h = O.laplace_run(init, it)
    assert np.array_equal(hist, ref_h)
    assert np.array_equal(ops.ops_dat_fetch_data(A), O.interior(ref_A))
    # every apply_stencil ran: on its own, or with the copy before it as one
    # pass (fusion level 2: the first iteration's alone, the rest fused)
    assert _fused_count(ops, "gen apply_stencil") + _fused_count(ops, "fused(gen copy+gen apply_stencil)") == it
    assert _fused_count(ops, "fused(gen copy+gen apply_stencil)") == it - 1


def test_generated_loop_checks_its_stencil(gpu_ops, gen):
    ops = gpu_ops
    blk = ops.ops_decl_block(2, "genbad")
    A = ops.ops_decl_dat(blk, [40, 30], name="A")
    An = ops.ops_decl_dat(blk, [40, 30], name="An")
    gen.gen_laplace_bad_stencil(blk.ptr, A.ptr, An.ptr, 40, 30)
    with pytest.raises(ops.OPSError, match="lacks point"):
        ops._check()
    # element type against the kernel's declaration (checked before the stencil)
    F = ops.ops_decl_dat(blk, [40, 30], dtype=np.float32, name="F")
    gen.gen_laplace_bad_stencil(blk.ptr, F.ptr, An.ptr, 40, 30)
    with pytest.raises(ops.OPSError, match="declares double"):
        ops._check()


@pytest.mark.parametrize("n", [(37, 29), (200, 61)])
def test_generated_fields_all_argument_kinds(gpu_ops, gen, n):
    ops = gpu_ops
    nx, ny = n
    blk = ops.ops_decl_block(2, "genfields")
    v = ops.ops_decl_dat(blk, [nx, ny], name="v", dim=2)
    m = ops.ops_decl_dat(blk, [nx, ny], dtype=np.int32, name="m")
    w = ops.ops_decl_dat(blk, [nx, ny], name="w")
    sx, sy, sh, mu = 0.5, 0.25, 1.5, 3
    wts = np.array([0.5, -0.25, 2.0])
    out = np.zeros(5)
    gen.gen_fields(blk.ptr, v.ptr, m.ptr, w.ptr, nx, ny, sx, sy, sh, mu, _p(wts), _p(out))
    ops._check()
    # closed forms over the whole array (halo included), index [y, x]
    X, Y = np.meshgrid(np.arange(-1, nx + 1), np.arange(-1, ny + 1))
    v0 = sx * X + sh
    v1 = sy * Y - 0.5 * X
    mm = (mu * X - Y).astype(np.int32)
    V = ops.get_whole(v)
    assert np.array_equal(V[..., 0], v0) and np.array_equal(V[..., 1], v1)
    assert np.array_equal(ops.get_whole(m), mm)
    i = (slice(1, -1), slice(1, -1))
    wref = (wts[0] * v0[1:-1, 2:] + wts[1] * v1[:-2, 1:-1]) + wts[2] * (v0[2:, 1:-1] - v1[1:-1, :-2])
    W = ops.ops_dat_fetch_data(w)
    assert np.array_equal(W, wref)
    assert out[0] == v0[i].sum()  # multiples of 1/4: exact in any order
    assert abs(out[1] - np.sum(wref * wref)) <= 1e-12 * np.sum(wref * wref)
    assert out[2] == wref.min() and out[3] == v1[i].max()
    assert out[4] == int(mm[i].sum())


@pytest.mark.parametrize("n", [(40, 33, 20), (130, 17, 9)])
def test_generated_heat_matches_hand_written(gpu_ops, gen, n):
    ops = gpu_ops
    nx, ny, nz = n
    blk = ops.ops_decl_block(3, "genheat")
    names = ["w", "Kx", "Ky", "Kz", "p", "r", "w2"]
    D = {k: ops.ops_decl_dat(blk, [nx, ny, nz], name=k) for k in names}
    ext = (nz + 2, ny + 2, nx + 2)
    for s, k in enumerate(["Kx", "Ky", "Kz", "p", "r"]):
        ops.set_whole(D[k], O.u01(91, s, int(np.prod(ext))).reshape(ext) + (1.0 if k[0] == "K" else 0.0))
    rx, ry, rz, alpha = 0.3, 0.7, 1.1, 0.37
    r0 = ops.ops_dat_fetch_data(D["r"])
    out = np.zeros(2)
    gen.gen_heat_step(blk.ptr, *[D[k].ptr for k in ["w", "Kx", "Ky", "Kz", "p", "r"]], nx, ny, nz, rx, ry, rz,
                      alpha, _p(out))
    ops._check()
    # the hand-written entry point of the same kernel
    R, W = ops.OPS_READ, ops.OPS_WRITE
    s0 = ops.ops_decl_stencil(3, 1, [0, 0, 0], "g0")
    s7 = ops.ops_decl_stencil(3, 7, [0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1], "g7")
    sxs = ops.ops_decl_stencil(3, 2, [0, 0, 0, 1, 0, 0], "gx")
    sys_ = ops.ops_decl_stencil(3, 2, [0, 0, 0, 0, 1, 0], "gy")
    szs = ops.ops_decl_stencil(3, 2, [0, 0, 0, 0, 0, 1], "gz")
    h = ops.ops_decl_reduction_handle(8, "double", "pw_hw")
    ops.ops_par_loop("tea_leaf_cg_calc_w_reduce_kernel_3d", "hw calc_w", blk, 3, [0, nx, 0, ny, 0, nz],
                     ops.ops_arg_dat(D["w2"], 1, s0, "double", W), ops.ops_arg_dat(D["Kx"], 1, sxs, "double", R),
                     ops.ops_arg_dat(D["Ky"], 1, sys_, "double", R), ops.ops_arg_dat(D["Kz"], 1, szs, "double", R),
                     ops.ops_arg_dat(D["p"], 1, s7, "double", R), ops.ops_arg_gbl(rx, 1, "double"),
                     ops.ops_arg_gbl(ry, 1, "double"), ops.ops_arg_gbl(rz, 1, "double"),
                     ops.ops_arg_reduce(h, 1, "double", ops.OPS_INC))
    pw = ops.ops_reduction_result(h)
    w_gen = ops.ops_dat_fetch_data(D["w"])
    assert np.array_equal(w_gen, ops.ops_dat_fetch_data(D["w2"]))
    assert abs(out[0] - pw) <= 1e-12 * abs(pw)
    # and the oracle's operator with the same non-unit rx, ry, rz
    Kx, Ky, Kz, p_ext = (ops.get_ext(D[k]) for k in ("Kx", "Ky", "Kz", "p"))
    w_ref, pw_ref = O.heat_apply(3, nx, ny, nz, Kx, Ky, Kz, p_ext, rx, ry, rz)
    assert np.array_equal(w_gen, O.interior(w_ref))
    assert abs(out[0] - pw_ref) <= 1e-12 * abs(pw_ref)
    r1 = ops.ops_dat_fetch_data(D["r"])
    assert np.array_equal(r1, r0 - alpha * w_gen)
    assert abs(out[1] - np.sum(r1 * r1)) <= 1e-12 * np.sum(r1 * r1)


@pytest.mark.parametrize("n", [2, 63, 1000])  # (a size of 1 declares an OPS edge dat, without halo)
def test_generated_1d_float_loops(gpu_ops, gen, n):
    ops = gpu_ops
    blk = ops.ops_decl_block(1, "gen1d")
    f = ops.ops_decl_dat(blk, [n], d_m=[-1], d_p=[1], dtype=np.float32, name="f")
    g = ops.ops_decl_dat(blk, [n], d_m=[-1], d_p=[1], dtype=np.float32, name="g")
    c0, c1 = np.float32(0.3), np.float32(-1.7)
    mx = C.c_float(0)
    gen.gen_ramp1d(blk.ptr, f.ptr, g.ptr, n, c0, c1, C.byref(mx))
    ops._check()
    X = np.arange(-1, n + 1).astype(np.float32)
    fr = c0 * X + c1
    assert np.array_equal(ops.get_whole(f), fr)
    gr = (np.float32(0.25) * fr[:-2] + np.float32(0.5) * fr[1:-1]) + np.float32(0.25) * fr[2:]
    assert np.array_equal(ops.ops_dat_fetch_data(g), gr)
    assert np.float32(mx.value) == gr.max()


@pytest.mark.parametrize("n", [(9, 7, 5), (70, 12, 6)])
@pytest.mark.parametrize("soa", [0, 1])
def test_generated_3d_three_components(gpu_ops, gen, n, soa):
    """A three-component dat in either layout (ops_set_soa: structure of
    arrays keeps each component in its own array); fetches and slabs present
    a point's components together either way."""
    ops = gpu_ops
    nx, ny, nz = n
    blk = ops.ops_decl_block(3, "genv3")
    ops.ops_set_soa(soa)
    try:
        v = ops.ops_decl_dat(blk, [nx, ny, nz], name="v", dim=3)
    finally:
        ops.ops_set_soa(0)
    d = ops.ops_decl_dat(blk, [nx, ny, nz], name="d")
    sums = np.zeros(3)
    imin = C.c_int(0)
    gen.gen_vec3(blk.ptr, v.ptr, d.ptr, nx, ny, nz, _p(sums), C.byref(imin))
    ops._check()
    Z, Y, X = np.meshgrid(np.arange(-1, nz + 1), np.arange(-1, ny + 1), np.arange(-1, nx + 1), indexing="ij")
    V = ops.get_whole(v)
    assert np.array_equal(V[..., 0], X.astype(float))
    assert np.array_equal(V[..., 1], 2.0 * Y) and np.array_equal(V[..., 2], 3.0 * Z + 1.0)
    assert np.all(ops.ops_dat_fetch_data(d) == 12.0)
    i = (slice(1, -1),) * 3
    assert list(sums) == [X[i].sum(), (2.0 * Y[i]).sum(), (3.0 * Z[i] + 1.0).sum()]
    assert imin.value == 12 - (nx - 1)
    # the array itself, through the host raw pointer: component-major for SoA
    core = v.core
    ms = C.c_int(ops.OPS_HOST)
    addr = ops._call("ops_dat_get_raw_pointer", v.ptr, 0, None, C.byref(ms))
    npts = core.size[0] * core.size[1] * core.size[2]
    raw = np.ctypeslib.as_array((C.c_double * (3 * npts)).from_address(addr - core.base_offset)).copy()
    ops._call("ops_dat_release_raw_data", v.ptr, 0, ops.OPS_READ)
    if soa:
        comp = raw.reshape(3, core.size[2], core.size[1], core.size[0])[:, :, :, :nx + 2]
    else:
        comp = np.moveaxis(raw.reshape(core.size[2], core.size[1], core.size[0], 3), 3, 0)[:, :, :, :nx + 2]
    assert np.array_equal(comp[0], X.astype(float)) and np.array_equal(comp[2], 3.0 * Z + 1.0)


def test_generated_optional_dat(gpu_ops, gen):
    ops = gpu_ops
    nx, ny, nz = 20, 9, 4
    blk = ops.ops_decl_block(3, "genopt")
    d = ops.ops_decl_dat(blk, [nx, ny, nz], name="d")
    e = ops.ops_decl_dat(blk, [nx, ny, nz], name="e")
    d0 = O.u01(5, 1, nx * ny * nz).reshape(nz, ny, nx)
    e0 = O.u01(5, 2, nx * ny * nz).reshape(nz, ny, nx)
    ops.ops_dat_set_data(d, d0)
    ops.ops_dat_set_data(e, e0)
    gen.gen_opt_add(blk.ptr, d.ptr, e.ptr, nx, ny, nz, 1)
    ops._check()
    assert np.array_equal(ops.ops_dat_fetch_data(d), d0 + e0)
    gen.gen_opt_add(blk.ptr, d.ptr, e.ptr, nx, ny, nz, 0)
    ops._check()
    assert np.array_equal(ops.ops_dat_fetch_data(d), 2.0 * (d0 + e0))


def test_generated_loop_with_one_dat_twice(gpu_ops, gen):
    ops = gpu_ops
    nx, ny, nz = 130, 7, 5
    blk = ops.ops_decl_block(3, "gentwice")
    d = ops.ops_decl_dat(blk, [nx, ny, nz], name="d")
    d0 = O.u01(6, 1, nx * ny * nz).reshape(nz, ny, nx)
    ops.ops_dat_set_data(d, d0)
    gen.gen_double_in_place(blk.ptr, d.ptr, nx, ny, nz)
    ops._check()
    assert np.array_equal(ops.ops_dat_fetch_data(d), 2.0 * d0)


def test_generated_and_hand_written_loops_share_the_queue(gpu_ops, gen):
    """A generated fill and hand-written loops in one lazy queue: a result
    that does not need the fill is right, a result that reads the filled dat
    sees it, and a later hand-written write to it and a later generated fill
    keep program order."""
    ops = gpu_ops
    nx, ny = 50, 21
    blk = ops.ops_decl_block(2, "genq")
    a = ops.ops_decl_dat(blk, [nx, ny], name="a")
    b = ops.ops_decl_dat(blk, [nx, ny], name="b")
    ops.ops_dat_set_data(b, np.ones((ny, nx)))
    s00 = ops.ops_decl_stencil(2, 1, [0, 0], "q00")
    rng = [0, nx, 0, ny]
    R, W = ops.OPS_READ, ops.OPS_WRITE

    def total(d, tag):
        h = ops.ops_decl_reduction_handle(8, "double", tag)
        ops.ops_par_loop("red_sum", tag, blk, 2, rng, ops.ops_arg_dat(d, 1, s00, "double", R),
                         ops.ops_arg_reduce(h, 1, "double", ops.OPS_INC))
        return ops.ops_reduction_result(h)

    Y, X = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    gen.gen_fill2d(blk.ptr, a.ptr, nx, ny, 1.25)
    ops._check()
    assert total(b, "sb") == nx * ny
    assert total(a, "sa") == ((1.25 + X) - 0.5 * Y).sum()  # quarter multiples: exact
    # generated fill, then hand-written copy over it, then a generated fill of b
    gen.gen_fill2d(blk.ptr, a.ptr, nx, ny, 7.0)
    ops.ops_par_loop("copy", "copy b to a", blk, 2, rng, ops.ops_arg_dat(a, 1, s00, "double", W),
                     ops.ops_arg_dat(b, 1, s00, "double", R))
    gen.gen_fill2d(blk.ptr, b.ptr, nx, ny, -2.0)
    ops._check()
    assert total(a, "sa2") == nx * ny
    assert np.array_equal(ops.ops_dat_fetch_data(b), (-2.0 + X) - 0.5 * Y)
    assert ops.ops_mi_loop_stats("fill2d")[0] == 3


def _cg_tail(ops, gen, n, iters, fusion):
    nx, ny, nz = n
    blk = ops.ops_decl_block(3, "gentail")
    ext = (nz + 2, ny + 2, nx + 2)
    D = {}
    for s, k in enumerate(["u", "p", "r"]):
        D[k] = ops.ops_decl_dat(blk, [nx, ny, nz], name=k)
        ops.set_whole(D[k], O.u01(57, s, int(np.prod(ext))).reshape(ext) - 0.5)
    start = {k: ops.get_whole(D[k]) for k in D}
    pp = np.zeros(iters)
    ops.lib().ops_mi_set_fusion(fusion)
    ops.ops_mi_reset_stats()
    try:
        gen.gen_cg_tail(blk.ptr, D["u"].ptr, D["p"].ptr, D["r"].ptr, nx, ny, nz, 0.37, -0.61, iters, _p(pp))
        ops._check()
        fused = ops.ops_mi_loop_stats("fused(cg_axpy+cg_axpby_dot)")
    finally:
        ops.lib().ops_mi_set_fusion(2)
    return start, {k: ops.get_whole(D[k]) for k in D}, pp, fused


@pytest.mark.parametrize("n", [(64, 20, 9), (131, 7, 5)])
def test_generated_pair_runs_fused_and_bit_exact(gpu_ops, gen, n):
    """axpy then axpby + dot: one fused pass per iteration, the same bits as
    the two loops run one after the other, and numpy's values (fields
    bit-exact, p.p within 1e-12 relative)."""
    iters = 3
    s0, fused, pp_f, st = _cg_tail(gpu_ops, gen, n, iters, 2)
    _, apart, pp_a, st_a = _cg_tail(gpu_ops, gen, n, iters, 0)
    assert st is not None and st[0] == iters
    assert st_a is None or st_a[0] == 0
    for k in ("u", "p", "r"):
        assert np.array_equal(fused[k], apart[k]), k
    assert np.array_equal(pp_f, pp_a)
    u, p, r = (s0[k].copy() for k in ("u", "p", "r"))
    i = (slice(1, -1),) * 3
    for it in range(iters):
        u[i] = u[i] + 0.37 * p[i]
        p[i] = r[i] + -0.61 * p[i]
        ref = np.sum(p[i] * p[i])
        assert abs(pp_f[it] - ref) <= 1e-12 * abs(ref)
    assert np.array_equal(fused["u"], u) and np.array_equal(fused["p"], p)
    assert np.array_equal(fused["r"], s0["r"])


def _cg_three(ops, gen, n, iters, fusion, short_p, w_is_u=False):
    nx, ny, nz = n
    blk = ops.ops_decl_block(3, "genthree")
    ext = (nz + 2, ny + 2, nx + 2)
    D = {}
    for s, k in enumerate(["u", "p", "r", "w"]):
        if k == "w" and w_is_u:
            D[k] = D["u"]
            continue
        D[k] = ops.ops_decl_dat(blk, [nx, ny, nz], name=k)
        ops.set_whole(D[k], O.u01(71, s, int(np.prod(ext))).reshape(ext) - 0.5)
    start = {k: ops.get_whole(D[k]) for k in D}
    rr, pp = np.zeros(iters), np.zeros(iters)
    ops.lib().ops_mi_set_fusion(fusion)
    ops.ops_mi_reset_stats()
    try:
        gen.gen_cg_three(blk.ptr, D["u"].ptr, D["p"].ptr, D["r"].ptr, D["w"].ptr, nx, ny, nz, 0.29, -0.53, iters,
                         short_p, _p(rr), _p(pp))
        ops._check()
        st = {"three": ops.ops_mi_loop_stats("fused(cg3_u+cg3_r+cg3_p)"),
              "pair": ops.ops_mi_loop_stats("fused(cg3_u+cg3_r)")}
    finally:
        ops.lib().ops_mi_set_fusion(2)
    return start, {k: ops.get_whole(D[k]) for k in D}, rr, pp, st


@pytest.mark.parametrize("n", [(64, 20, 9), (131, 7, 5)])
def test_generated_three_loops_run_as_one_pass(gpu_ops, gen, n):
    """Three point-wise loops on shared dats (u += a p; r -= a w with r.r;
    p = r + b p with p.p): one pass per iteration at fusion level 3, none
    below it, the same bits as the loops one by one (level 0) at every
    level, and numpy's values (fields bit-exact, sums within 1e-12
    relative)."""
    iters = 3
    s0, apart, rr_a, pp_a, st_a = _cg_three(gpu_ops, gen, n, iters, 0, 0)
    assert st_a["three"] is None or st_a["three"][0] == 0
    for level in (1, 2, 3):
        _, fused, rr_f, pp_f, st = _cg_three(gpu_ops, gen, n, iters, level, 0)
        if level == 3:
            assert st["three"] is not None and st["three"][0] == iters
        else:
            assert st["three"] is None or st["three"][0] == 0, level
        for k in ("u", "p", "r", "w"):
            assert np.array_equal(fused[k], apart[k]), (level, k)
        assert np.array_equal(rr_f, rr_a) and np.array_equal(pp_f, pp_a), level
    u, p, r, w = (s0[k].copy() for k in ("u", "p", "r", "w"))
    i = (slice(1, -1),) * 3
    for it in range(iters):
        u[i] = u[i] + 0.29 * p[i]
        r[i] = r[i] - 0.29 * w[i]
        ref_r = np.sum(r[i] * r[i])
        p[i] = r[i] + -0.53 * p[i]
        ref_p = np.sum(p[i] * p[i])
        assert abs(rr_a[it] - ref_r) <= 1e-12 * abs(ref_r)
        assert abs(pp_a[it] - ref_p) <= 1e-12 * abs(ref_p)
    for k, v in (("u", u), ("p", p), ("r", r), ("w", w)):
        assert np.array_equal(apart[k], v), k


def test_generated_three_loops_with_a_shorter_range_stay_apart(gpu_ops, gen):
    """The third loop one plane short: no three-loop pass (the first two
    still run as a pair), and the results equal the unfused run's."""
    n, iters = (64, 20, 9), 2
    _, apart, rr_a, pp_a, _ = _cg_three(gpu_ops, gen, n, iters, 0, 1)
    _, fused, rr_f, pp_f, st = _cg_three(gpu_ops, gen, n, iters, 3, 1)
    assert st["three"] is None or st["three"][0] == 0
    assert st["pair"] is not None and st["pair"][0] == iters
    for k in ("u", "p", "r", "w"):
        assert np.array_equal(fused[k], apart[k]), k
    assert np.array_equal(rr_f, rr_a) and np.array_equal(pp_f, pp_a)


def test_generated_three_loops_with_one_dat_under_two_names_stay_apart(gpu_ops, gen):
    """w and u the same dat: the fused kernels give the two names separate
    restrict-qualified pointers, so the queue launches neither the run nor
    the pair of the first two loops; results equal the unfused run's."""
    n, iters = (64, 20, 9), 2
    _, apart, rr_a, pp_a, _ = _cg_three(gpu_ops, gen, n, iters, 0, 0, w_is_u=True)
    _, fused, rr_f, pp_f, st = _cg_three(gpu_ops, gen, n, iters, 3, 0, w_is_u=True)
    assert st["three"] is None or st["three"][0] == 0
    assert st["pair"] is None or st["pair"][0] == 0
    for k in ("u", "p", "r"):
        assert np.array_equal(fused[k], apart[k]), k
    assert np.array_equal(rr_f, rr_a) and np.array_equal(pp_f, pp_a)


def test_generated_pair_with_a_neighbour_read_is_not_fused(gpu_ops, gen):
    """copy then apply_stencil: the second loop reads neighbours of what the
    first writes, so the point-wise pair rule (fusion level 1) keeps them
    apart; only the neighbour form (level 2) joins them.  The results are the
    Laplace test's, bit-exact against the oracle."""
    ops = gpu_ops
    nx, ny = 64, 48
    init = O.laplace_init(nx, ny, 0)
    blk = ops.ops_decl_block(2, "genlapnf")
    A = ops.ops_decl_dat(blk, [nx, ny], name="A")
    An = ops.ops_decl_dat(blk, [nx, ny], name="Anew")
    ops.set_ext(A, init)
    ops.set_ext(An, init)
    hist = np.zeros(4)
    ops.lib().ops_mi_set_fusion(1)
    try:
        ops.ops_mi_reset_stats()
        gen.gen_laplace(blk.ptr, A.ptr, An.ptr, nx, ny, 4, _p(hist))
        ops._check()
        for name in ("fused(gen copy+gen apply_stencil)", "fused(gen apply_stencil+gen copy)"):
            assert _fused_count(ops, name) == 0, name
        assert ops.ops_mi_loop_stats("gen copy")[0] == 4 and ops.ops_mi_loop_stats("gen apply_stencil")[0] == 4
    finally:
        ops.lib().ops_mi_set_fusion(2)
    ref_A, ref_h = O.laplace_run(init, 4)
    assert np.array_equal(hist, ref_h)
    assert np.array_equal(ops.ops_dat_fetch_data(A), O.interior(ref_A))


# ------------------------------------------------------------------ low-dimensional dats
@pytest.mark.parametrize("n", [(9, 7, 5), (130, 20, 11), (300, 3, 70)])
def test_generated_lowdim_accumulation(gpu_ops, gen, n):
    """Edge dats (size 1 in some dimensions) through strided stencils: MAX,
    MIN and INC accumulation from a 3D field into 2D and 1D dats (double,
    float, int) with the reference's reset rule
    (ops/c/src/core/ops_dummy_singlenode.cpp:204-391): the first INC starts
    from 0 whatever the dat held, a second INC adds to it, an INC after an
    RW loop starts from 0 again; then one 3D loop reads every edge dat.
    Integer-valued data, so every sum is exact in any order."""
    ops = gpu_ops
    nx, ny, nz = n
    blk = ops.ops_decl_block(3, f"lowdim_{nx}")
    steps = np.zeros((4, ny, nx))
    rmax, zmin = np.zeros(nx), np.zeros(nz)
    cnt = np.zeros(ny, np.int32)
    ymax = np.zeros(ny, np.float32)
    out = np.zeros((nz, ny, nx))
    gen.gen_lowdim3d(blk.ptr, nx, ny, nz, _p(steps), _p(rmax), _p(zmin), cnt.ctypes.data_as(C.POINTER(C.c_int)),
                     ymax.ctypes.data_as(C.POINTER(C.c_float)), _p(out))
    ops._check()
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = ((X * 7 + Y * 3 + Z * 11) % 17).astype(np.float64) - 5.0
    col = f.sum(axis=0)
    assert np.array_equal(steps[0], col)
    assert np.array_equal(steps[1], 2 * col)
    assert np.array_equal(steps[2], col)
    assert np.array_equal(steps[3], col)
    assert np.array_equal(rmax, f.max(axis=(0, 1)))
    assert np.array_equal(zmin, f.min(axis=(1, 2)))
    assert np.array_equal(cnt, np.full(ny, nx * nz, np.int32))
    assert np.array_equal(ymax, (f.astype(np.float32) + np.float32(0.5)).max(axis=(0, 2)))
    ref = f + 2.0 * col[None] + 4.0 * rmax[None, None, :] + 8.0 * zmin[:, None, None] + 16.0 * cnt[None, :, None]
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("n", [(33, 17), (200, 64)])
def test_generated_lowdim_2d_min(gpu_ops, gen, n):
    ops = gpu_ops
    nx, ny = n
    blk = ops.ops_decl_block(2, f"lowdim2_{nx}")
    m = np.zeros(nx)
    gen.gen_lowdim2d(blk.ptr, nx, ny, _p(m))
    ops._check()
    Y, X = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    g = ((X * 5 + Y * 9) % 13).astype(np.float64) - 6.0
    assert np.array_equal(m, g.min(axis=0))


@pytest.mark.parametrize("mode,msg", [(0, None), (1, "stride 2")])
def test_generated_lowdim_refused(gpu_ops, gen, mode, msg):
    """A stride other than 0 or 1 is refused before launch; a plain write
    through a stride-0 stencil that puts several points on one element runs
    (every point stores the same value, as the reference's apps do)."""
    ops = gpu_ops
    blk = ops.ops_decl_block(3, f"lowdim_bad{mode}")
    gen.gen_lowdim_bad(blk.ptr, 8, 6, 4, mode)
    if msg is None:
        ops._check()
        ops.ops_mi_sync()
        return
    with pytest.raises(ops.OPSError, match=msg):
        ops._check()


# ------------------------------------------------------------------ neighbour-form fused pairs
def _levels(ops, fn):
    """fn() at fusion levels 0, 1 and 2 (the library's default is 2)"""
    out = []
    for lev in (0, 1, 2):
        ops.lib().ops_mi_set_fusion(lev)
        try:
            ops.ops_mi_reset_stats()
            out.append(fn())
        finally:
            ops.lib().ops_mi_set_fusion(2)
    return out


def _fused_count(ops, name):
    st = ops.ops_mi_loop_stats(name)
    return st[0] if st else 0


@pytest.mark.parametrize("n,shift", [((70, 20, 13), 0), ((130, 9, 40), 0), ((41, 30, 17), 1)])
def test_generated_nb_fusion_src_lap7(gpu_ops, gen, n, shift):
    """u = 2a + 1, then w = 7-point operator of u with sum of w^2: one pass at
    fusion level 2 when both loops share the range (u at neighbours
    recomputed from a inside the range, read from memory in the halo), two
    passes otherwise.  Fields bit-identical across levels and to numpy; the
    sum within 1e-12 and the same bits across levels."""
    ops = gpu_ops
    nx, ny, nz = n
    ext = (nz + 2, ny + 2, nx + 2)
    u0 = O.u01(71, 1, int(np.prod(ext))).reshape(ext) - 0.5
    a0 = O.u01(71, 2, int(np.prod(ext))).reshape(ext)
    r1 = [0, nx, 0, ny, 0, nz]
    r2 = [0, nx, 0, ny, shift, nz]

    def run():
        blk = ops.ops_decl_block(3, f"chn_{nx}")
        u, a, w = (ops.ops_decl_dat(blk, [nx, ny, nz], name=k) for k in ("chu", "cha", "chw"))
        ops.set_ext(u, u0)
        ops.set_ext(a, a0)
        sv = np.zeros(1)
        gen.gen_chain_src_lap(blk.ptr, u.ptr, a.ptr, w.ptr, (C.c_int * 6)(*r1), (C.c_int * 6)(*r2), _p(sv))
        ops._check()
        out = (ops.get_ext(u), ops.get_ext(w), sv[0], _fused_count(ops, "fused(ch_src+ch_lap7)"))
        for d in (u, a, w):
            ops.ops_free_dat(d)
        return out

    runs = _levels(ops, run)
    ref_u = u0.copy()
    ref_u[1:-1, 1:-1, 1:-1] = 2.0 * a0[1:-1, 1:-1, 1:-1] + 1.0
    U = ref_u
    c = U[1:-1, 1:-1, 1:-1]
    wint = 6.0 * c - (U[1:-1, 1:-1, :-2] + U[1:-1, 1:-1, 2:]) - (U[1:-1, :-2, 1:-1] + U[1:-1, 2:, 1:-1]) - \
        (U[:-2, 1:-1, 1:-1] + U[2:, 1:-1, 1:-1])
    z0 = shift
    for lev, (uu, ww, sv, cnt) in enumerate(runs):
        assert np.array_equal(uu, ref_u), lev
        assert np.array_equal(ww[1 + z0:-1, 1:-1, 1:-1], wint[z0:]), lev
        ref_s = np.sum(wint[z0:] * wint[z0:])
        assert abs(sv - ref_s) <= 1e-12 * ref_s, lev
        assert cnt == (1 if lev == 2 and shift == 0 else 0), (lev, cnt)
    assert runs[0][2] == runs[2][2] or shift != 0


@pytest.mark.parametrize("n", [(50, 37), (300, 41)])
def test_generated_nb_fusion_double_buffered(gpu_ops, gen, n):
    """u = a - b/2 then a = average of u's neighbours with max(a), six times:
    the second loop overwrites a, which the first reads, so the fused pass
    writes a into a second buffer.  Bit-identical across levels and to numpy."""
    ops = gpu_ops
    nx, ny = n
    it = 6
    ext = (ny + 2, nx + 2)
    a0 = O.u01(73, 1, int(np.prod(ext))).reshape(ext)
    b0 = O.u01(73, 2, int(np.prod(ext))).reshape(ext) - 0.5
    u0 = O.u01(73, 3, int(np.prod(ext))).reshape(ext)

    def run():
        blk = ops.ops_decl_block(2, f"chm_{nx}")
        u, a, b = (ops.ops_decl_dat(blk, [nx, ny], name=k) for k in ("cmu", "cma", "cmb"))
        ops.set_ext(u, u0)
        ops.set_ext(a, a0)
        ops.set_ext(b, b0)
        mx = np.zeros(it)
        gen.gen_chain_mix_avg(blk.ptr, u.ptr, a.ptr, b.ptr, nx, ny, it, _p(mx))
        ops._check()
        out = (ops.get_ext(u), ops.get_ext(a), mx, _fused_count(ops, "fused(ch_mix+ch_avg)"))
        for d in (u, a, b):
            ops.ops_free_dat(d)
        return out

    runs = _levels(ops, run)
    U, A, B = u0.copy(), a0.copy(), b0.copy()
    mxs = []
    for _ in range(it):
        U[1:-1, 1:-1] = A[1:-1, 1:-1] - 0.5 * B[1:-1, 1:-1]
        A[1:-1, 1:-1] = 0.25 * (U[1:-1, 2:] + U[1:-1, :-2] + U[2:, 1:-1] + U[:-2, 1:-1])
        mxs.append(A[1:-1, 1:-1].max())
    for lev, (uu, aa, mx, cnt) in enumerate(runs):
        assert np.array_equal(uu, U) and np.array_equal(aa, A), lev
        assert np.array_equal(mx, np.array(mxs)), lev
        assert cnt == (it if lev == 2 else 0), (lev, cnt)


@pytest.mark.parametrize("mode", [1, 2])
def test_generated_nb_fusion_refused(gpu_ops, gen, mode):
    """The dat the source links is not the one the second loop reads (the
    variable was reassigned between the calls), or the first loop reads the
    dat it writes: the neighbour form is refused.  In the first case the
    second loop's neighbour reads are of a dat neither loop writes, but the
    point-wise pair form shares one pointer between the arguments written
    alike (p), which here name two dats, so it is refused too; in the second
    the pair runs as two passes.  Either way the results are those of the loops in turn
    (the neighbour form would have given w from 3 v instead of the other
    dat)."""
    ops = gpu_ops
    nx, ny = 67, 23
    ext = (ny + 2, nx + 2)
    blk = ops.ops_decl_block(2, f"chr_{mode}")
    D = {k: ops.ops_decl_dat(blk, [nx, ny], name=f"chr{mode}{k}") for k in ("u", "v", "w", "o")}
    init = {k: O.u01(79, i, int(np.prod(ext))).reshape(ext) for i, k in enumerate(("u", "v", "w", "o"))}
    for k, d in D.items():
        ops.set_ext(d, init[k])
    ops.ops_mi_reset_stats()
    gen.gen_chain_refused(blk.ptr, D["u"].ptr, D["v"].ptr, D["w"].ptr, D["o"].ptr, nx, ny, mode)
    ops._check()
    assert _fused_count(ops, "fused(ch_scale+ch_diff)") == 0
    U = init["u"].copy()
    U[1:-1, 1:-1] = 3.0 * (init["v"] if mode == 1 else init["u"])[1:-1, 1:-1]
    P = init["o"] if mode == 1 else U
    wref = P[1:-1, 2:] - P[1:-1, :-2] + P[2:, 1:-1] - P[:-2, 1:-1]
    assert np.array_equal(ops.get_ext(D["u"]), U)
    assert np.array_equal(ops.ops_dat_fetch_data(D["w"]), wref)
    for d in D.values():
        ops.ops_free_dat(d)


@pytest.mark.parametrize("n", [(64, 48), (257, 33)])
def test_generated_laplace_fused_across_levels(gpu_ops, gen, n):
    """The Laplace loop through the translator: each copy runs with the next
    apply_stencil as one pass at fusion level 2 (apply_stencil reads A, which
    copy writes, at its four neighbours); history and field bit-identical at
    every level and to the oracle."""
    ops = gpu_ops
    nx, ny = n
    it = 20
    init = O.laplace_init(nx, ny, 0)

    def run():
        blk = ops.ops_decl_block(2, "genlapf")
        A = ops.ops_decl_dat(blk, [nx, ny], name="A")
        An = ops.ops_decl_dat(blk, [nx, ny], name="Anew")
        ops.set_ext(A, init)
        ops.set_ext(An, init)
        hist = np.zeros(it)
        gen.gen_laplace(blk.ptr, A.ptr, An.ptr, nx, ny, it, _p(hist))
        ops._check()
        out = (hist, ops.ops_dat_fetch_data(A), _fused_count(ops, "fused(gen copy+gen apply_stencil)"))
        ops.ops_free_dat(A)
        ops.ops_free_dat(An)
        return out

    runs = _levels(ops, run)
    ref_A, ref_h = O.laplace_run(init, it)
    for lev, (h, A, cnt) in enumerate(runs):
        assert np.array_equal(h, ref_h) and np.array_equal(A, O.interior(ref_A)), lev
        assert cnt == (it - 1 if lev == 2 else 0), (lev, cnt)


@pytest.mark.parametrize("n", [(64, 40), (131, 23)])
def test_legacy_pointer_kernels_2d(gpu_ops, gen, n):
    """Kernels in the legacy form (T * parameters indexed with OPS_ACCn) give
    the bits of the same kernels written with ACC<T>, and of numpy in the
    kernels' operation order, at every fusion level; at level 2 the legacy
    copy runs with the next 5-point average as one pass."""
    ops = gpu_ops
    nx, ny = n
    a0 = O.u01(11, 1, (nx + 2) * (ny + 2)).reshape(ny + 2, nx + 2)

    def lap(a):
        b = a.copy()
        b[1:-1, 1:-1] = 0.25 * (((a[1:-1, 2:] + a[1:-1, :-2]) + a[2:, 1:-1]) + a[:-2, 1:-1])
        return b

    b1 = lap(a0)
    a1 = a0.copy()
    a1[1:-1, 1:-1] = b1[1:-1, 1:-1]
    want_b = lap(a1)[1:-1, 1:-1]

    def run(mode):
        blk = ops.ops_decl_block(2, "genleg2")
        a = ops.ops_decl_dat(blk, [nx, ny], name="a")
        b = ops.ops_decl_dat(blk, [nx, ny], name="b")
        ops.set_ext(a, a0)
        assert gen.gen_legacy_lap2d(blk.ptr, a.ptr, b.ptr, nx, ny, mode) == 0
        ops._check()
        out = (ops.ops_dat_fetch_data(a), ops.ops_dat_fetch_data(b),
               _fused_count(ops, "fused(leg_copy+leg_lap)" if mode == 0 else "fused(acc_copy+acc_lap)"))
        ops.ops_free_dat(a)
        ops.ops_free_dat(b)
        return out

    for lev, (leg, acc) in enumerate(zip(_levels(ops, lambda: run(0)), _levels(ops, lambda: run(1)))):
        assert np.array_equal(leg[0], acc[0]) and np.array_equal(leg[1], acc[1]), lev
        assert np.array_equal(leg[0], a1[1:-1, 1:-1]) and np.array_equal(leg[1], want_b), lev
        assert leg[2] == acc[2] == (1 if lev == 2 else 0), (lev, leg[2], acc[2])


@pytest.mark.parametrize("n", [(20, 9, 6), (67, 13, 5)])
@pytest.mark.parametrize("soa", [0, 1])
def test_legacy_multi_component_kernels_3d(gpu_ops, gen, n, soa):
    """OPS_ACC_MDn on a two-component 3D dat (either layout) at neighbours,
    with an index-weighted sum: field bit-identical to the ACC<T> form and to
    numpy in the kernel's order; the sum bit-identical between the forms and
    within 1e-12 of numpy's."""
    ops = gpu_ops
    nx, ny, nz = n
    v0 = O.u01(12, 1, (nz + 2) * (ny + 2) * (nx + 2) * 2).reshape(nz + 2, ny + 2, nx + 2, 2)
    w_ref = (v0[1:-1, 1:-1, 2:, 0] - v0[:-2, 1:-1, 1:-1, 1]) + 0.5 * (v0[1:-1, 2:, 1:-1, 1] - v0[1:-1, :-2, 1:-1, 0])
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    s_ref = float((w_ref * (X + 2 * Y + 3 * Z)).sum())
    out = []
    for mode in (0, 1):
        blk = ops.ops_decl_block(3, "genleg3")
        ops.ops_set_soa(soa)
        try:
            v = ops.ops_decl_dat(blk, [nx, ny, nz], name="v", dim=2)
        finally:
            ops.ops_set_soa(0)
        w = ops.ops_decl_dat(blk, [nx, ny, nz], name="w")
        ops.set_ext(v, v0)
        s = np.zeros(1)
        assert gen.gen_legacy_md3d(blk.ptr, v.ptr, w.ptr, nx, ny, nz, mode, _p(s)) == 0
        ops._check()
        out.append((ops.ops_dat_fetch_data(w), s[0]))
        ops.ops_free_dat(v)
        ops.ops_free_dat(w)
    (w_leg, s_leg), (w_acc, s_acc) = out
    assert np.array_equal(w_leg, w_acc) and s_leg == s_acc
    assert np.array_equal(w_leg, w_ref)
    assert abs(s_leg - s_ref) <= 1e-12 * max(1.0, np.abs(w_ref * (X + 2 * Y + 3 * Z)).sum())


@pytest.mark.parametrize("n", [(40, 30, 20), (300, 50, 40)])
def test_generated_host_fold_keeps_the_bits(gpu_ops, gen, n):
    """The generated calc_w and residual update with each result read after
    its loop: the host fold of the workgroup partials gives the same bits as
    the fold kernel (ops_mi_set_host_fold(0))."""
    ops = gpu_ops
    nx, ny, nz = n
    blk = ops.ops_decl_block(3, "genhf")
    names = ["w", "Kx", "Ky", "Kz", "p", "r"]
    D = {k: ops.ops_decl_dat(blk, [nx, ny, nz], name=k) for k in names}
    ext = (nz + 2, ny + 2, nx + 2)
    res = []
    for on in (1, 0):
        for s, k in enumerate(["Kx", "Ky", "Kz", "p", "r"]):
            ops.set_whole(D[k], O.u01(93, s, int(np.prod(ext))).reshape(ext) + (1.0 if k[0] == "K" else -0.5))
        ops.ops_mi_set_host_fold(on)
        try:
            out = np.zeros(2)
            gen.gen_heat_step(blk.ptr, *[D[k].ptr for k in names], nx, ny, nz, 0.3, 0.7, 1.1, 0.37, _p(out))
            ops._check()
        finally:
            ops.ops_mi_set_host_fold(1)
        res.append((out.copy(), ops.ops_dat_fetch_data(D["r"])))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    for k in names:
        ops.ops_free_dat(D[k])
