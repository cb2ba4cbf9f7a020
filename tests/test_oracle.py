"""Pins the oracle (CPU restatement of the reference path) before it is used
as the checker: bit-exact against fixtures produced by the reference's own
libops_seq, the reference app's known answer, and closed forms for what the
fixtures leave open (Thomas/PCR, the Poisson operator spectrum, CG
invariants, mixed-precision refinement)."""
import numpy as np
import pytest

from oracle import oracle_py as O


@pytest.fixture(scope="module")
def golden():
    return O.load_golden()


def test_laplace_top_bit_exact(golden):
    g, m = golden
    c = m["laplace_top"]
    A, hist = O.laplace_run(O.laplace_init(c["nx"], c["ny"], 0), c["iters"])
    assert np.array_equal(hist, g["laplace_top_err"])
    assert np.array_equal(O.interior(A).ravel(), g["laplace_top_A"])


def test_laplace_app_bit_exact(golden):
    g, m = golden
    c = m["laplace_app"]
    A, hist = O.laplace_run(O.laplace_init(c["nx"], c["ny"], 1), c["iters"])
    assert np.array_equal(hist, g["laplace_app_err"])
    assert np.array_equal(O.interior(A).ravel(), g["laplace_app_A"])


def test_laplace_f32_bit_exact(golden):
    g, m = golden
    c = m["laplace_f32"]
    init = g["laplace_f32_init"].reshape(c["ny"] + 2, c["nx"] + 2)
    A, hist = O.laplace_run(init, c["iters"])
    assert np.array_equal(hist, g["laplace_f32_err"])
    assert np.array_equal(O.interior(A).ravel(), g["laplace_f32_A"])


def test_laplace_threads_match_serial(golden):
    # MAX reductions are order independent: the OpenMP form is bit-exact too
    A1, h1 = O.laplace_run(O.laplace_init(130, 97, 1), 30, nthreads=1)
    A8, h8 = O.laplace_run(O.laplace_init(130, 97, 1), 30, nthreads=8)
    assert np.array_equal(h1, h8) and np.array_equal(A1, A8)


def test_laplace_reference_known_answer(golden):
    # apps/c/laplace2d_tutorial/step7/laplace2d.cpp:134 (4094^2, 100 iterations)
    g, m = golden
    c = m["laplace_kat"]
    kat = g["laplace_kat_err"][-1]
    assert abs(100.0 * kat / c["expected_app"] - 100.0) < 0.001
    A, hist = O.laplace_run(O.laplace_init(c["nx"], c["ny"], 1), c["iters"], nthreads=8)
    assert hist[-1] == kat


@pytest.mark.parametrize("ndim,key", [(2, "cg2d"), (3, "heat3d")])
def test_heat_cg_bit_exact(golden, ndim, key):
    g, m = golden
    c = m[key]
    nz = c.get("nz", 1)
    K, rhs = O.heat_inputs(m["seed"], ndim, c["nx"], c["ny"], nz)
    r = O.heat_cg(ndim, c["nx"], c["ny"], nz, K, rhs, max_iters=c["iters"])
    for k in ("pw", "alpha", "beta", "rrn"):
        assert np.array_equal(r[k], g[f"{key}_{k}"]), k
    assert np.array_equal(r["u"].ravel(), g[f"{key}_u"])
    Kx, Ky, Kz = O.heat_coefs(ndim, c["nx"], c["ny"], nz, K)
    assert np.array_equal(O.interior(Kx).ravel(), g[f"{key}_Kx"])


@pytest.mark.parametrize("ndim,key", [(2, "poisson2d"), (3, "poisson3d")])
def test_poisson_cg_bit_exact(golden, ndim, key):
    g, m = golden
    c = m[key]
    nz = c.get("nz", 1)
    f = O.poisson_rhs(m["seed"], ndim, c["nx"], c["ny"], nz)
    r = O.poisson_cg(ndim, c["nx"], c["ny"], nz, f, max_iters=c["iters"])
    for k in ("alpha", "beta", "rrn"):
        assert np.array_equal(r[k], g[f"{key}_{k}"]), k
    assert np.array_equal(r["x"].ravel(), g[f"{key}_x"])


def test_poisson_f32_sweep_bit_exact(golden):
    g, m = golden
    c = m["poisson3d"]
    f = O.poisson_rhs(m["seed"], 3, c["nx"], c["ny"], c["nz"])
    p = np.zeros(O.ext_shape(3, c["nx"], c["ny"], c["nz"]), np.float32)
    O.interior(p)[:] = f.astype(np.float32)
    w, s = O.poisson_apply_f32(3, c["nx"], c["ny"], c["nz"], p)
    assert np.array_equal(O.interior(w).ravel(), g["poisson3d_f32_w"])
    assert s == g["poisson3d_f32_pw"][0]


def test_adi_preproc_and_27pt_bit_exact(golden):
    g, m = golden
    c = m["adi"]
    u = g["adi_u_ext"].reshape(c["nz"] + 2, c["ny"] + 2, c["nx"] + 2)
    du, ax, bx, cx = O.adi_preproc(c["nx"], c["ny"], c["nz"], u, c["lambda"])
    assert np.array_equal(du.ravel(), g["adi_du"])
    assert np.array_equal(ax.ravel(), g["adi_ax"])
    assert np.array_equal(bx.ravel(), g["adi_bx"])
    assert np.array_equal(O.stencil27(c["nx"], c["ny"], c["nz"], u).ravel(), g["st27_out"])


def test_reductions_fixture_closed_form(golden):
    # the reduction-semantics fixture agrees with numpy on the same data
    g, m = golden
    c = m["red"]
    x = g["red_x"].reshape(c["ny"], c["nx"])
    out = g["red_out"]
    sub = x[1:c["ny"] - 2, 2:c["nx"] - 3]
    assert out[0] == pytest.approx(x.sum(), rel=1e-14)
    assert out[1] == x.min() and out[2] == x.max()
    assert out[3] == pytest.approx(x.sum() + (sub * sub).sum(), rel=1e-14)
    assert out[4] == pytest.approx(np.abs(x).sum(), rel=1e-14)
    assert out[5] == sub.min() and out[6] == sub.max()
    assert out[7] == pytest.approx(x.sum(), rel=1e-14) and out[8] == x.size
    assert out[9] == (c["nx"] - 1) * (c["ny"] - 1)
    assert out[10] == pytest.approx(-2.5 * x.sum(), rel=1e-14)
    assert out[11] == np.finfo(np.float64).max


def test_uniform_generator_is_integer_exact():
    # SplitMix64 finaliser restated independently in numpy
    G = np.uint64(0x9E3779B97F4A7C15)

    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        i = np.arange(1000, dtype=np.uint64)
        seed, stream = np.uint64(12345), np.uint64(7)
        bits = mix(mix(seed ^ (stream * G)) + (i + np.uint64(1)) * G)
    ref = (bits >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    assert np.array_equal(O.u01(12345, 7, 1000), ref)
    n = O.normal(99, 1, 200000)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1.0) < 0.01


# ---------------------------------------------------------------- Thomas / PCR
def _dense_solve(a, b, c, d):
    n = len(b)
    M = np.zeros((n, n), dtype=np.longdouble)
    for i in range(n):
        M[i, i] = b[i]
        if i > 0:
            M[i, i - 1] = a[i]
        if i < n - 1:
            M[i, i + 1] = c[i]
    return np.linalg.solve(M.astype(np.float64), d)


@pytest.mark.parametrize("solvedim", [0, 1, 2])
def test_thomas_matches_dense_solve(solvedim):
    nx, ny, nz = 7, 5, 6
    a, b, c = O.trid_inputs(3, solvedim, nx, ny, nz)
    d = O.trid_rhs(3, nx, ny, nz)
    x = d.copy()
    O.thomas_batch(3, solvedim, [nx, ny, nz], [nx, ny, nz], a, b, c, x)
    ax = [2, 1, 0][solvedim]  # numpy axis of the solve dimension
    A, B, Cc, D, X = (np.moveaxis(v, ax, -1).reshape(-1, v.shape[ax]) for v in (a, b, c, d, x))
    for q in range(A.shape[0]):
        ref = _dense_solve(A[q], B[q], Cc[q], D[q])
        assert np.allclose(X[q], ref, rtol=1e-13, atol=1e-14)
        # residual of the tridiagonal system
        n = len(ref)
        r = B[q] * X[q] - D[q]
        r[1:] += A[q][1:] * X[q][:-1]
        r[:-1] += Cc[q][:-1] * X[q][1:]
        assert np.linalg.norm(r) / np.linalg.norm(D[q]) < 1e-14


def test_thomas_ignores_first_a_and_last_c():
    n = 9
    a, b, c = O.trid_inputs(5, 0, n, 1, 1)
    d = O.trid_rhs(5, n, 1, 1)
    x1 = d.copy()
    O.thomas_batch(3, 0, [n, 1, 1], [n, 1, 1], a, b, c, x1)
    a2, c2 = a.copy(), c.copy()
    a2.ravel()[0] = 123.0
    c2.ravel()[-1] = -77.0
    x2 = d.copy()
    O.thomas_batch(3, 0, [n, 1, 1], [n, 1, 1], a2, b, c2, x2)
    assert np.array_equal(x1, x2)


def test_thomas_strided_subrange_and_inc():
    # a range inside padded arrays (the pads argument of the libtrid call)
    px, py, pz = 12, 9, 8
    nx, ny, nz = 10, 7, 6
    rng = np.random.default_rng(0)
    a = -rng.random((pz, py, px))
    c = -rng.random((pz, py, px))
    b = np.abs(a) + np.abs(c) + 0.5
    d = rng.random((pz, py, px)) * 2 - 1
    u = rng.random((pz, py, px))
    off = 1 + px * (1 + py * 1)
    for sd in range(3):
        dd = d.copy()
        O.thomas_batch(3, sd, [nx, ny, nz], [px, py, pz], a.ravel()[off:], b.ravel()[off:], c.ravel()[off:],
                       dd.ravel()[off:])
        uu = u.copy()
        O.thomas_batch(3, sd, [nx, ny, nz], [px, py, pz], a.ravel()[off:], b.ravel()[off:], c.ravel()[off:],
                       d.ravel()[off:], uu.ravel()[off:])
        inner = (slice(1, 1 + nz), slice(1, 1 + ny), slice(1, 1 + nx))
        assert np.array_equal(uu[inner], u[inner] + dd[inner])
        outside = np.ones_like(d, bool)
        outside[inner] = False
        assert np.array_equal(dd[outside], d[outside])
        ax = [2, 1, 0][sd]
        sub = [v[inner] for v in (a, b, c, d, dd)]
        A, B, Cc, D, X = (np.moveaxis(v, ax, -1).reshape(-1, v.shape[ax]) for v in sub)
        for q in range(0, A.shape[0], 7):
            assert np.allclose(X[q], _dense_solve(A[q], B[q], Cc[q], D[q]), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 256])
def test_pcr_matches_thomas(n):
    a, b, c = O.trid_inputs(8, 0, n, 6, 2)
    d = O.trid_rhs(8, n, 6, 2)
    xt, xp = d.copy(), d.copy()
    O.thomas_batch(3, 0, [n, 6, 2], [n, 6, 2], a, b, c, xt)
    O.pcr_batch(3, 0, [n, 6, 2], [n, 6, 2], a, b, c, xp)
    assert np.max(np.abs(xp - xt)) <= 1e-13 * np.max(np.abs(xt))


# ---------------------------------------------------------------- closed forms
def test_poisson_operator_eigenvector():
    # sin modes are eigenvectors of the zero-Dirichlet 7-point operator with
    # eigenvalue sum_d 2(1 - cos(k_d pi/(n_d+1)))
    nx, ny, nz = 11, 9, 7
    kx, ky, kz = 2, 3, 1
    i, j, k = np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1)
    v = (np.sin(kz * np.pi * k / (nz + 1))[:, None, None] * np.sin(ky * np.pi * j / (ny + 1))[None, :, None]
         * np.sin(kx * np.pi * i / (nx + 1))[None, None, :])
    lam = sum(2 * (1 - np.cos(kk * np.pi / (nn + 1))) for kk, nn in ((kx, nx), (ky, ny), (kz, nz)))
    x = np.zeros(O.ext_shape(3, nx, ny, nz))
    O.interior(x)[:] = v
    r, _ = O.poisson_residual(3, nx, ny, nz, np.zeros((nz, ny, nx)), x)
    assert np.allclose(-O.interior(r), lam * v, atol=1e-13)


def test_cg_residual_invariant():
    # the recursive residual r.r matches |f - A x|^2 recomputed from x
    nx, ny, nz = 16, 12, 10
    f = O.poisson_rhs(7, 3, nx, ny, nz)
    res = O.poisson_cg(3, nx, ny, nz, f, max_iters=30)
    x = np.zeros(O.ext_shape(3, nx, ny, nz))
    O.interior(x)[:] = res["x"]
    _, rr = O.poisson_residual(3, nx, ny, nz, f, x)
    assert rr == pytest.approx(res["rrn"][-1], rel=1e-8)


def test_heat_operator_symmetric_positive():
    nx, ny, nz = 6, 5, 4
    K, _ = O.heat_inputs(11, 3, nx, ny, nz)
    Kx, Ky, Kz = O.heat_coefs(3, nx, ny, nz, K)
    n = nx * ny * nz
    M = np.zeros((n, n))
    for q in range(n):
        p = np.zeros(O.ext_shape(3, nx, ny, nz))
        O.interior(p)[np.unravel_index(q, (nz, ny, nx))] = 1.0
        w, _ = O.heat_apply(3, nx, ny, nz, Kx, Ky, Kz, p)
        M[:, q] = O.interior(w).ravel()
    assert np.allclose(M, M.T, rtol=1e-14, atol=1e-14)
    assert np.linalg.eigvalsh(M).min() >= 1.0 - 1e-12  # I + r L with L >= 0


def test_heat_cg_converges_to_direct_solve():
    nx, ny, nz = 6, 5, 4
    K, rhs = O.heat_inputs(11, 3, nx, ny, nz)
    r = O.heat_cg(3, nx, ny, nz, K, rhs, max_iters=500, tol=1e-12)
    Kx, Ky, Kz = O.heat_coefs(3, nx, ny, nz, K)
    n = nx * ny * nz
    M = np.zeros((n, n))
    for q in range(n):
        p = np.zeros(O.ext_shape(3, nx, ny, nz))
        O.interior(p)[np.unravel_index(q, (nz, ny, nx))] = 1.0
        w, _ = O.heat_apply(3, nx, ny, nz, Kx, Ky, Kz, p)
        M[:, q] = O.interior(w).ravel()
    ref = np.linalg.solve(M, rhs.ravel())
    assert np.allclose(r["u"].ravel(), ref, rtol=1e-10)


def test_mixed_precision_reaches_fp64_solution():
    nx, ny, nz = 24, 20, 18
    f = O.poisson_rhs(5, 3, nx, ny, nz)
    mp = O.poisson_mixed(3, nx, ny, nz, f, tol=1e-10, max_outer=20, inner_iters=500, inner_tol=1e-4)
    assert mp["relres"][-1] <= 1e-10
    cg = O.poisson_cg(3, nx, ny, nz, f, max_iters=2000, tol=1e-12)
    err = np.linalg.norm(mp["x"] - cg["x"]) / np.linalg.norm(cg["x"])
    assert err < 1e-8
    # the reported residual is the FP64 residual recomputed from x
    x = np.zeros(O.ext_shape(3, nx, ny, nz))
    O.interior(x)[:] = mp["x"]
    _, rr = O.poisson_residual(3, nx, ny, nz, f, x)
    assert np.sqrt(rr) / np.linalg.norm(f) == pytest.approx(mp["relres"][-1], rel=1e-6)


def test_mixed_heat_reaches_fp64_solution():
    # FP64 refinement around FP32 CG on the 3D heat problem: reaches the
    # FP64 CG solution, and its reported residual is the FP64 residual of x
    # recomputed here (relative to the residual of TeaLeaf's start x = rhs)
    nx, ny, nz = 22, 19, 17
    K, rhs = O.heat_inputs(9, 3, nx, ny, nz)
    mp = O.heat_mixed(nx, ny, nz, K, rhs, tol=1e-10, inner_tol=1e-6, nthreads=4)
    assert mp["relres"][-1] <= 1e-10 and mp["outer"] >= 2  # FP32 inner solves cannot reach 1e-10 at once
    cg = O.heat_cg(3, nx, ny, nz, K, rhs, max_iters=5000, tol=1e-13, nthreads=4)
    assert np.linalg.norm(mp["x"] - cg["u"]) / np.linalg.norm(cg["u"]) < 1e-8
    Kx, Ky, Kz = O.heat_coefs(3, nx, ny, nz, K)
    x = np.zeros(O.ext_shape(3, nx, ny, nz))
    u0 = np.zeros_like(x)
    O.interior(x)[:] = mp["x"]
    O.interior(u0)[:] = rhs
    ax, _ = O.heat_apply(3, nx, ny, nz, Kx, Ky, Kz, x)
    a0, _ = O.heat_apply(3, nx, ny, nz, Kx, Ky, Kz, u0)
    rel = np.linalg.norm(rhs - O.interior(ax)) / np.linalg.norm(rhs - O.interior(a0))
    assert rel == pytest.approx(mp["relres"][-1], rel=1e-6)


def test_mb_poisson_matches_reference_app(golden):
    # the reference's multi-block Poisson app, built from its sources and run
    # unchanged (its printed total error has 15 significant digits)
    g, m = golden
    c = m["mb_poisson"]
    assert abs(g["mb_poisson_err"][0] - c["expected_app"]) <= 1e-13 * c["expected_app"]
    for (nx, ny, it), ref in zip(c["configs"], g["mb_poisson_err"]):
        err, u = O.mb_poisson(nx, ny, it)
        assert abs(err - ref) <= 1e-13 * ref, (nx, ny, it)
        assert u.shape == (ny, nx)


def test_halo_transfer_matches_reference(golden):
    """The restatement of ops_halo_transfer against the reference library's
    own transfers (oracle/refdrv/ref_halo.cpp): the multi-block app's five
    halo groups, 3D permuted/flipped axes with unequal halo depths, an
    overlapping self-transfer, two components per point and float/double
    dats meeting — bit for bit after every group."""
    g, m = golden
    for case in m["halo"]:
        fx = O.halo_fixture(g, case)
        dats = {name: O.HaloDat(fx[name][0].copy(), d_m, case["ndim"])
                for name, _, _, _, d_m, _ in case["dats"]}
        for k, grp in enumerate(case["groups"], start=1):
            for f, t, it, fb, tb, fd, td in grp:
                O.halo_transfer(dats[f], dats[t], it, fb, tb, fd, td)
            for name in dats:
                assert np.array_equal(dats[name].a, fx[name][k]), (case["name"], k, name)
        # every group moved something
        for k in range(1, len(case["groups"]) + 1):
            assert any(not np.array_equal(fx[n][k], fx[n][k - 1]) for n in dats), (case["name"], k)


def test_halo_transfer_mblock_rotation_closed_form(golden):
    """The rotated group of the multi-block app (x of one block against y of
    the other): data1(x, -2 + c) = data0(18 + c, x) for c in 0..1 — the
    mapping written out by hand, independent of the restatement."""
    g, m = golden
    case = [c for c in m["halo"] if c["name"] == "mblock"][0]
    fx = O.halo_fixture(g, case)
    before0, after1 = fx["data0"][4], fx["data1"][5]
    for c in range(2):
        for x in range(20):
            # whole-array index = logical index + 2 (depth-2 halos)
            assert after1[0, -2 + c + 2, x + 2, 0] == before0[0, x + 2, 18 + c + 2, 0]
