"""The reference's own apps, unchanged, on this backend: the Laplace tutorial
(apps/c/laplace2d_tutorial/step7/laplace2d.cpp) and the multi-block Poisson
test app (apps/c/poisson/poisson.cpp), compiled against include/ops_seq_v2.h
and linked with libopsmi.so (oracle/Makefile.ref, in the build container).
Each prints its own known-answer check; the tests read it.  Skipped where
the binaries were not built (no reference sources at build time)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def _run(name, *args):
    exe = os.path.join(REF, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    out = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120, cwd=REF)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout


def test_laplace_tutorial_app_passes_its_known_answer():
    out = _run("app_laplace2d_opsmi")
    assert "This run is considered PASSED" in out, out[-2000:]
    pct = float(re.search(r"within ([0-9.eE+-]+) %", out).group(1))
    assert pct < 1e-9  # the app accepts 1e-3 %


def test_multiblock_poisson_app_passes_its_known_answer():
    out = _run("app_poisson_opsmi")
    assert "This run is considered PASSED" in out, out[-2000:]
    err = float(re.search(r"Total error: ([0-9.eE+-]+)", out).group(1))
    assert abs(err - 16.5173570176708) <= 1e-12 * 16.5173570176708


def test_multiblock_poisson_app_other_sizes():
    from oracle import oracle_py as O
    g, m = O.load_golden()
    for (nx, ny, it), ref in zip(m["mb_poisson"]["configs"], g["mb_poisson_err"]):
        out = _run("app_poisson_opsmi", f"-sizex={nx}", f"-sizey={ny}", f"-iters={it}")
        err = float(re.search(r"Total error: ([0-9.eE+-]+)", out).group(1))
        assert abs(err - ref) <= 1e-12 * ref, (nx, ny, it)


# The same two apps through this project's translator: their own kernel
# headers become generated HIP kernels (oracle/Makefile.ref, APPS_GEN).
def test_laplace_tutorial_app_through_the_translator():
    """The tutorial's copy and the next apply_stencil run as one generated
    pass (apply_stencil reads, at neighbours, the A that copy writes); the
    printed error is the hand-written entry points' to the last digit, and
    the same with fusion off (OPSMI_FUSION=0)."""
    out = _run("app_laplace2d_gen")
    assert re.search(r"fused\(copy\+apply stencil\)\s+(\d+)", out), out[-3000:]
    assert "This run is considered PASSED" in out, out[-2000:]
    pct = float(re.search(r"within ([0-9.eE+-]+) %", out).group(1))
    assert pct < 1e-9
    hw = _run("app_laplace2d_opsmi")
    assert re.search(r"within ([0-9.eE+-]+) %", out).group(1) == re.search(r"within ([0-9.eE+-]+) %", hw).group(1)
    exe = os.path.join(REF, "app_laplace2d_gen")
    off = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=REF,
                         env=dict(os.environ, OPSMI_FUSION="0"))
    assert off.returncode == 0 and "fused(copy+apply stencil)" not in off.stdout
    assert re.search(r"within ([0-9.eE+-]+) %", off.stdout).group(1) == re.search(r"within ([0-9.eE+-]+) %",
                                                                                  out).group(1)


def test_multiblock_poisson_app_through_the_translator():
    from oracle import oracle_py as O
    out = _run("app_poisson_gen")
    assert "This run is considered PASSED" in out, out[-2000:]
    err = float(re.search(r"Total error: ([0-9.eE+-]+)", out).group(1))
    assert abs(err - 16.5173570176708) <= 1e-12 * 16.5173570176708
    g, m = O.load_golden()
    for (nx, ny, it), ref in zip(m["mb_poisson"]["configs"], g["mb_poisson_err"]):
        out = _run("app_poisson_gen", f"-sizex={nx}", f"-sizey={ny}", f"-iters={it}")
        err = float(re.search(r"Total error: ([0-9.eE+-]+)", out).group(1))
        assert abs(err - ref) <= 1e-12 * ref, (nx, ny, it)


def test_lowdim_app_through_the_translator():
    """The reference's low-dimensional dat test (apps/c/lowdim_test/lowdim.cpp,
    unchanged): MAX, MIN and INC from a 3D dat into 2D and 1D edge dats
    through strided stencils, reads back through them, its own checks after
    each step.  Its closing HDF5 output is not available here: run with
    OPSMI_ERROR_MODE=2 the calls are reported and skipped, and the app goes
    on to its last check."""
    exe = os.path.join(REF, "app_lowdim_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    env = dict(os.environ, OPSMI_ERROR_MODE="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=REF, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for what in ("OPS_MAX on 2D directions OK", "OPS_MIN on 2D directions OK", "OPS_INC on 2D directions OK",
                 "OPS_MAX on 1D directions OK", "OPS_MIN on 1D directions OK", "OPS_INC on 1D directions OK",
                 "Calc done", "All checks PASSED"):
        assert what in r.stdout, (what, r.stdout[-2000:])
    assert "TEST FAILED" not in r.stdout and "errors found" not in r.stdout
    assert r.stderr.count("HDF5 output is not available") == 8, r.stderr[-2000:]


def test_half_precision_app_through_the_translator():
    """The reference's half-precision Laplace app (apps/c/halfprecision/
    laplace2d.cpp, unchanged): "half" dats with binary16 arithmetic in the
    generated kernels (native FP16 instructions, one rounding per operation)
    and a float MAX reduction, 4094^2 for 100 iterations; its own check
    against the known answer 0.002685546875 (laplace2d.cpp:122).  HDF5
    output is reported and skipped (OPSMI_ERROR_MODE=2); the app's NaN check
    of A runs."""
    exe = os.path.join(REF, "app_halfprec_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    env = dict(os.environ, OPSMI_ERROR_MODE="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=REF, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "This run is considered PASSED" in r.stdout, r.stdout[-3000:]
    assert re.search(r"^\s*100, 0\.002686\s*$", r.stdout, re.M), r.stdout[-3000:]
    assert "ops_NaNcheck_core called for A" in r.stdout and "NaN detected" not in r.stdout
    assert r.stderr.count("HDF5 output is not available") == 2, r.stderr[-2000:]


def test_multidim_soa_app_through_the_translator():
    """The reference's multiDim test (apps/c/multiDim/multidim.cpp, unchanged,
    the C++ object API): an OPS_instance with OPS_soa = 1, two-component dats
    in SoA layout, a halo group exchange, a two-component reduction checked
    against its known answer 24 + 24 (multidim.cpp:171), gbl arrays indexed
    by the loop index; HDF5 output reported and skipped."""
    exe = os.path.join(REF, "app_multidim_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    env = dict(os.environ, OPSMI_ERROR_MODE="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=REF, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "This test is considered PASSED" in r.stdout, r.stdout[-3000:]
    assert "Reduction result = 24, 24" in r.stdout, r.stdout[-3000:]
    assert "The test using ops_arg_gbl arguments passed!" in r.stdout


def test_mb_trid_app_through_the_translator():
    """The reference's mb_trid_test (apps/c/mb_trid_test/mb_trid.cpp): an
    OPS_instance with two blocks and a tridiagonal settings object per block."""
    out = _run("app_mbtrid_gen")
    assert "error" not in out.lower()


# The reference's TeaLeaf (apps/c/TeaLeaf, 18 files, legacy OPS_ACCn
# kernels) through the translator as one program, on its own test input
# (tea.in: 10 x 10 cells, 10 steps, tl_eps 1e-15, test_problem 1) with each of
# its four solvers; field_summary.cpp:82-103 compares the final temperature
# with 157.55084183279294 and prints PASSED within 1e-3 %.
TEA_IN = """*tea
 state 1 density=100.0 energy=0.0001
 state 2 density=0.1 energy=25.0 geometry=rectangle xmin=0.0 xmax=1.0 ymin=1.0 ymax=2.0
 state 3 density=0.1 energy=0.1 geometry=rectangle xmin=1.0 xmax=6.0 ymin=1.0 ymax=2.0
 state 4 density=0.1 energy=0.1 geometry=rectangle xmin=5.0 xmax=6.0 ymin=1.0 ymax=8.0
 state 5 density=0.1 energy=0.1 geometry=rectangle xmin=5.0 xmax=10.0 ymin=7.0 ymax=8.0
 x_cells=10
 y_cells=10
 xmin=0.0
 ymin=0.0
 xmax=10.0
 ymax=10.0
 initial_timestep=0.004
 end_step=10
 tl_max_iters=1000
 {solver}
 tl_eps=1e-15
 test_problem 1
*endtea
"""


@pytest.mark.parametrize("solver", ["tl_use_jacobi", "tl_use_cg", "tl_use_chebyshev", "tl_use_ppcg"])
def test_tealeaf_app_through_the_translator(tmp_path, solver):
    exe = os.path.join(REF, "app_tealeaf_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    (tmp_path / "tea.in").write_text(TEA_IN.format(solver=solver))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    txt = out.stdout + (tmp_path / "tea.out").read_text()
    assert "This test is considered PASSED" in txt, txt[-3000:]
    pct = float(re.search(r"within\s+([0-9.eE+-]+)%", txt).group(1))
    assert pct < 1e-3, pct



def _adi_oracle(nx, ny, nz, iters, pcr):
    """The ADI app's computation restated on the oracle: init_kernel (1 on
    the faces of the range, 0 inside, apps/c/adi/init_kernel.h), then per
    iteration preproc_kernel (lambda = 1) and the x, y solves in place on du
    and the z solve added to u (adi.cpp:385-421); the app's a, b, c are the
    same for the three directions.  Returns u with its one-point halo."""
    import numpy as np

    from oracle import oracle_py as O

    u = np.zeros((nz + 2, ny + 2, nx + 2))
    core = u[1:-1, 1:-1, 1:-1]
    core[...] = 1.0
    core[1:-1, 1:-1, 1:-1] = 0.0
    solve = O.pcr_batch if pcr else O.thomas_batch
    dims = [nx, ny, nz]
    for _ in range(iters):
        du, a, b, c = O.adi_preproc(nx, ny, nz, u, 1.0)
        solve(3, 0, dims, dims, a, b, c, du)
        solve(3, 1, dims, dims, a, b, c, du)
        ui = np.ascontiguousarray(u[1:-1, 1:-1, 1:-1])
        solve(3, 2, dims, dims, a, b, c, du, ui)
        u[1:-1, 1:-1, 1:-1] = ui
    return u


# The reference's ADI app (apps/c/adi/adi.cpp, legacy OPS_ACCn kernels,
# ops_tridMultiDimBatch along x and y and _Inc along z) through the
# translator.  Its HDF5 dumps are refused by this backend (no HDF5 library)
# and reported, so the run sets OPSMI_ERROR_MODE=2 (print and continue); it
# then checks its halo ("Halo Test") and writes the first nx*ny*nz doubles of
# u's host array to <argv[0]>.dat (adi.cpp:88-118,491), which the test
# compares with the oracle's u in the same layout.  -m 0 selects Thomas
# (bit-exact), -m 5 PCR (within 1e-12 of the oracle's PCR).
@pytest.mark.parametrize("m", [0, 5])
def test_adi_app_through_the_translator(tmp_path, m):
    import numpy as np

    exe = os.path.join(REF, "app_adi_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    nx, ny, nz, iters = 37, 20, 13, 3
    env = dict(os.environ, OPSMI_ERROR_MODE="2")
    out = subprocess.run(["adi_out", "-nx", str(nx), "-ny", str(ny), "-nz", str(nz), "-iter", str(iters),
                          "-m", str(m)], executable=exe, capture_output=True, text=True, timeout=120,
                         cwd=tmp_path, env=env)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert "Halo Test: PASSED" in out.stdout, out.stdout[-2000:]
    assert "Successful exit from OPS!" in out.stdout
    assert "HDF5 output is not available" in out.stderr
    sx, sy, sz = (int(v) for v in re.search(r"Local dimensions: (\d+) x (\d+) x (\d+)", out.stdout).groups())
    got = np.fromfile(tmp_path / "adi_out.dat", dtype=np.float64)
    assert got.size == nx * ny * nz
    full = np.zeros((sz, sy, sx))
    full[: nz + 2, : ny + 2, : nx + 2] = _adi_oracle(nx, ny, nz, iters, pcr=m == 5)
    want = full.ravel()[: nx * ny * nz]
    if m == 0:
        assert np.array_equal(got, want)
    else:
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_multidim3d_app_through_the_translator():
    """The reference's multiDim3D test (apps/c/multiDim/multidim3D, unchanged):
    the C API with OPS_instance::getOPSInstance()->OPS_soa = 1, so its
    three-component dats are SoA; a 7-point copy, a halo group within one
    dat and a three-component reduction whose sum the app checks against
    3 x 96 to 1e-13 % (multidim.cpp:136-148); HDF5 output reported and
    skipped."""
    exe = os.path.join(REF, "app_multidim3d_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    env = dict(os.environ, OPSMI_ERROR_MODE="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=REF, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "This run is considered PASSED" in r.stdout, r.stdout[-3000:]
    assert "Reduction result = 96.000000, 96.000000, 96.000000" in r.stdout, r.stdout[-3000:]


@pytest.mark.parametrize("step", [5, 6])
def test_laplace_tutorial_steps_through_the_translator(step):
    """Steps 5 and 6 of the reference's Laplace tutorial
    (apps/c/laplace2d_tutorial/step5,6: 4094^2, 100 iterations, every loop an
    ops_par_loop; step 6 takes the error from an OPS_MAX reduction) against
    the tutorial's known answer 2.421354960840227e-03 within 1e-3 %."""
    exe = os.path.join(REF, f"app_lap_step{step}_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=REF)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "This run is considered PASSED" in r.stdout, r.stdout[-3000:]
    pct = float(re.search(r"within ([0-9.eE+-]+) %", r.stdout).group(1))
    assert pct < 1e-9, pct


def test_random_app_through_the_translator(tmp_path):
    """The reference's random test (apps/c/random/random.cpp, unchanged): two
    uniform fills (seed 0) of a two-component 4 x 4 dat with halos (-2, -1)
    and (2, 1), each appended to random.dat with ops_print_dat_to_txtfile.
    The file holds, per fill, the whole array (halos and padding columns) in
    the reference's text format; its values are the generator's (fill k =
    stream k, counter 2 i + c) to the 10 printed decimals, padding 0."""
    import numpy as np

    from oracle import oracle_py as O

    exe = os.path.join(REF, "app_random_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "This run is considered PASSED" in r.stdout
    txt = (tmp_path / "random.dat").read_text()
    blocks = txt.split("ops_dat:  dat0 \n")[1:]
    assert len(blocks) == 2
    w, h, dim = 4 + 2 + 2, 4 + 1 + 1, 2
    for k, b in enumerate(blocks):
        lines = b.split("\n")
        assert lines[0] == "ops_dat dim:  2 "
        sx, sy = (int(v) for v in re.match(r"block Dims : 2 \[(\d+)\]\[(\d+)\]", lines[1]).groups())
        assert sy == h and sx >= w
        assert lines[2] == f"elem size {8 * dim} "
        rows = [ln for ln in lines[3:] if ln.strip()]
        got = np.array([[float(v) for v in ln.split()] for ln in rows])
        assert got.shape == (sy, sx * dim)
        got = got.reshape(sy, sx, dim)
        want = O.u01(0, k, w * h * dim).reshape(h, w, dim)
        assert np.max(np.abs(got[:, :w] - want)) <= 5.1e-11
        assert np.all(got[:, w:] == 0.0)


# The reference's CloverLeaf 2D (apps/c/CloverLeaf: 21 files, legacy OPS_ACCn
# kernels, ops_update_const, the C API's instance object) through the
# translator as one program.  field_summary.cpp:89-103 compares the final
# kinetic energy with the known answer of each test problem and prints
# PASSED within 1e-3 %: problem 1 is clover.in_default (10 x 2 cells, to
# t = 3), problem 2 clover_bm_short.in (960^2 cells, 87 steps).
CLOVER_IN = """*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 geometry=rectangle xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0
 x_cells={n}
 y_cells={m}
 xmin=0.0
 ymin=0.0
 xmax=10.0
 ymax={ymax}
 initial_timestep=0.04
 timestep_rise=1.5
 max_timestep=0.04
 {stop}
 test_problem {tp}
*endclover
"""


@pytest.mark.parametrize("tp", [1, 2, 3])
def test_cloverleaf_app_through_the_translator(tmp_path, tp):
    """Test problem 3 is the benchmark input clover_bm.in (960^2 cells to
    t = 15.5 or 2955 steps); the run's wall clock is printed."""
    exe = os.path.join(REF, "app_cloverleaf_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    cfg = {1: dict(n=10, m=2, ymax="2.0", stop="end_time=3.0", tp=1),
           2: dict(n=960, m=960, ymax="10.0", stop="end_step=87", tp=2),
           3: dict(n=960, m=960, ymax="10.0", stop="end_time=15.5\n end_step=2955", tp=3)}[tp]
    (tmp_path / "clover.in").write_text(CLOVER_IN.format(**cfg))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    txt = out.stdout + (tmp_path / "clover.out").read_text()
    assert "This test is considered PASSED" in txt, txt[-3000:]
    pct = float(re.search(r"within\s+([0-9.eE+-]+)\s*%", txt).group(1))
    assert pct < 1e-3, pct
    wall = re.findall(r"Total Wall time\s+([0-9.]+)", txt)[-1]
    print(f"CloverLeaf 2D test problem {tp}: wall clock {wall} s")


CLOVER3_IN = """*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 geometry=cuboid xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0 zmin=0.0 zmax={zmax2}
 x_cells={n}
 y_cells={m}
 z_cells={m}
 xmin=0.0
 ymin=0.0
 zmin=0.0
 xmax=10.0
 ymax={ymax}
 zmax={ymax}
 initial_timestep={dt}
 max_timestep={dt}
 end_step={steps}
 end_time={end_time}
 test_problem {tp}
 checkpoint_frequency={ckpt}
*endclover
"""


@pytest.mark.parametrize("tp", [1, 2, 3])
def test_cloverleaf3d_app_through_the_translator(tmp_path, tp):
    """CloverLeaf 3D (apps/c/CloverLeaf_3D) the same way: test problem 1
    (clover.in_default: 10 x 2 x 2 cells, 75 steps) and 2
    (clover_bm_short.in: 96^3 cells, 87 steps) and 3 (clover_bm.in: 96^3,
    2955 steps), field_summary.cpp:93-107; the wall clock is printed.
    The input sets checkpoint_frequency as the reference's inputs do: the
    main loop takes step % checkpoint_frequency (clover_leaf.cpp:209)."""
    exe = os.path.join(REF, "app_cloverleaf3d_gen")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    cfg = {1: dict(n=10, m=2, ymax="2.0", zmax2="2.0", dt="0.04", steps=75, end_time="10.0", tp=1, ckpt=100),
           2: dict(n=96, m=96, ymax="10.0", zmax2="4.0", dt="0.1", steps=87, end_time="1000.1", tp=2, ckpt=10),
           3: dict(n=96, m=96, ymax="10.0", zmax2="4.0", dt="0.1", steps=2955, end_time="1000.1", tp=3,
                   ckpt=100)}[tp]
    (tmp_path / "clover.in").write_text(CLOVER3_IN.format(**cfg))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    txt = out.stdout + (tmp_path / "clover.out").read_text()
    assert "This test is considered PASSED" in txt, txt[-3000:]
    pct = float(re.search(r"within\s+([0-9.eE+-]+)\s*%", txt).group(1))
    assert pct < 1e-3, pct
    wall = re.findall(r"Total Wall time\s+([0-9.]+)", txt)[-1]
    print(f"CloverLeaf 3D test problem {tp}: wall clock {wall} s")


def test_shsgc_apps_through_the_translator(tmp_path):
    """The reference's shock-capturing test apps, unchanged: shsgc
    (apps/c/shsgc/shsgc.cpp:409-420: solution RMS against 0.233688543536201
    within 1e-3 %) and its multi-block form
    (apps/c/mb_shsgc/Max_datatransfer/shsgc.cpp:500-511: post-shock error
    against 0.0003206 for 200 points)."""
    exe = os.path.join(REF, "app_shsgc_gen")
    exe2 = os.path.join(REF, "app_mbshsgc_gen")
    if not (os.path.exists(exe) and os.path.exists(exe2)):
        pytest.skip("shsgc apps not built")
    env = dict(os.environ, OPSMI_ERROR_MODE="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=tmp_path, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "This run is considered PASSED" in r.stdout, r.stdout[-3000:]
    pct = float(re.search(r"within ([0-9.eE+-]+) %", r.stdout).group(1))
    assert pct < 1e-3, pct
    r = subprocess.run([exe2], capture_output=True, text=True, timeout=120, cwd=tmp_path, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "PASSED" in r.stdout, r.stdout[-3000:]
