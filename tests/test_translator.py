"""The source-to-source translator (CPU only): what it extracts from an
application and its kernels, what it writes, and what it refuses."""
import os
import re
import subprocess

import pytest

from ops_amd.translator import opsmi_translate as T

HERE = os.path.dirname(os.path.abspath(__file__))
APPS = [os.path.join(HERE, "genapps", a) for a in ("laplace_gen.cpp", "fields_gen.cpp", "heat3d_gen.cpp")]


def test_mask_blanks_comments_and_strings():
    src = 'a(/* x(y) */ "b(c", \'(\') // d(\ne'
    m = T.mask(src)
    assert len(m) == len(src)
    assert m.count("(") == 1 and m.count(")") == 1
    assert m.endswith("\ne") and "d(" not in m and "x" not in m


def test_translates_dev_api_applications(tmp_path):
    out = str(tmp_path)
    ks = [T.translate(a, out) for a in APPS]
    assert ks == [["apply_stencil", "copy"], ["fill_fields", "mix_fields", "summary"], ["heat_calc_w", "heat_ur_r"]]
    lap = open(os.path.join(out, "laplace_gen_ops.cpp")).read()
    assert not re.search(r"\bops_par_loop\s*\(", T.mask(lap))
    assert lap.count("ops_par_loop_apply_stencil(") == 2 and "ops_par_loop_copy(" in lap
    assert '#include "ops_mi355x.h"\n#include "laplace_gen_opsmi_kernels.h"' in lap
    assert '/* #include "laplace_gen_kernels.h"' in lap and '\n#include "laplace_gen_kernels.h"' not in lap
    fields = open(os.path.join(out, "fields_gen_ops.cpp")).read()
    assert "void summary(" not in fields and "/* kernel summary: generated" in fields
    hip = open(os.path.join(out, "fields_gen_opsmi_kernels.hip")).read()
    # constants the body names become members; the two-component dat is MULTI
    assert re.search(r"struct opsk_mix_fields \{\s*double weights\[3\];", hip)
    assert re.search(r"struct opsk_fill_fields \{\s*int mult;\s*double shift;", hip)
    assert re.search(r"template <class OPSMI_A0, class OPSMI_A1> __device__ __forceinline__ void operator\(\)"
                     r"\(const OPSMI_A0 &v, OPSMI_A1 &w", hip)
    assert "::opsmi_gen::Acc<double, 2, true, false> a0(p0, A.d0, x, y, z);" in hip
    # the general kernel (plain pointers, x steps from the launch) beside the restrict-qualified one
    assert "::opsmi_gen::Acc<double, 2, true, true> a0(p0, A.d0, x, y, z);" in hip
    assert "auto kern = L->alias ? opsg_fill_fields_alias : opsg_fill_fields;" in hip
    assert "store_partials<int, OPS_INC, 1>" in hip and "store_partials<double, OPS_INC, 2>" in hip
    # one fused kernel per pair of loops that follow each other in the source
    # adjacent pairs: one restrict pointer per distinct dat expression
    assert "opsg_chain_mix_fields__summary(const opsa_mix_fields A0, const opsa_summary A1, " \
        "const double *__restrict__ g0, double *__restrict__ g1, const int *__restrict__ g2)" in hip
    assert "{&opsl_fill_fields, &opsl_mix_fields}; static const int g[] = {0, 1, -1, -1, 0, 2}; " \
        "ops_mi_gen_fused_chain(c, 2, g, 6, &opsc_fill_fields__mix_fields);" in hip
    assert "ops_mi_gen_fused_chain(c, 2, g, 9, &opsc_mix_fields__summary);" in hip
    assert "opsc_summary__" not in hip
    decl = open(os.path.join(out, "heat3d_gen_opsmi_kernels.h")).read()
    assert 'visibility("hidden")' in decl and "ops_par_loop_heat_calc_w(" in decl


def test_point_lists_follow_the_kernel_body():
    body = "{ w(0,0) = v(0, 1, 0) + v(1,0,-1) - v(0,0,1); }"
    assert T.accessed_points(body, "v", 2) == [(0, 0, 1, 0), (0, 1, 0, 0), (1, 0, -1, 0)]
    assert T.accessed_points(body, "w", 2) == [(-1, 0, 0, 0)]
    assert T.accessed_points("{ p(0, 0, -1) = p(1, 0, 0); }", "p", 3) == [(-1, 0, 0, -1), (-1, 1, 0, 0)]
    # a computed index or an atomic combine: no list (the stencil is not checked)
    assert T.accessed_points("{ a(i, 0) = 1; }", "a", 2) is None
    assert T.accessed_points("{ a.combine_max(0, 0, 1.0); }", "a", 2) is None
    # comments do not count
    assert T.accessed_points("{ a(0, 0) = 1; // a(5, 5)\n }", "a", 2) == [(-1, 0, 0, 0)]
    # an accessor handed to a helper, bound to another name, or never named:
    # the points it reaches are unknown
    assert T.accessed_points("{ w(0, 0) = nb(b); }", "b", 2) is None
    assert T.accessed_points("{ const ACC<double> &q = b; w(0, 0) = q(1, 0); }", "b", 2) is None
    assert T.accessed_points("{ w(0, 0) = b(0, 0) + nb(b); }", "b", 2) is None
    assert T.accessed_points("{ w(0, 0) = 1.0; }", "b", 2) is None


def test_helper_reading_neighbours_is_not_fused(tmp_path):
    """Two adjacent loops where the second reads the first's output through a
    helper: the stub of the second lists no points (-1), so the runtime does
    not take the dat as touched only at the point and keeps the loops apart."""
    p = tmp_path / "helper.cpp"
    p.write_text('#include "ops_seq_v2.h"\n'
                 'inline double nb(const ACC<double> &b) { return b(1, 0) + b(-1, 0); }\n'
                 'void k1(ACC<double> &a) { a(0, 0) = 1.0; }\n'
                 'void k2(const ACC<double> &a, ACC<double> &w) { w(0, 0) = nb(a); }\n'
                 'void run(ops_block b, ops_dat d, ops_dat e, ops_stencil s0, ops_stencil s3, int *r) {\n'
                 '  ops_par_loop(k1, "k1", b, 2, r, ops_arg_dat(d, 1, s0, "double", OPS_WRITE));\n'
                 '  ops_par_loop(k2, "k2", b, 2, r, ops_arg_dat(d, 1, s3, "double", OPS_READ),\n'
                 '               ops_arg_dat(e, 1, s0, "double", OPS_WRITE));\n'
                 '}\n')
    T.translate(str(p), str(tmp_path / "out"))
    hip = open(tmp_path / "out" / "helper_opsmi_kernels.hip").read()
    stub = hip[hip.index("void ops_par_loop_k2("):]
    assert "static const int npts[2] = {-1, 1};" in stub


def _app(tmp_path, loop, kernel="void k(ACC<double> &a) { a(0, 0) = 1.0; }"):
    p = tmp_path / "app.cpp"
    p.write_text(f'#include "ops_seq_v2.h"\n{kernel}\nvoid run(ops_block b, ops_dat d, ops_stencil s, int *r) {{\n'
                 f'  {loop};\n}}\n')
    return str(p)


@pytest.mark.parametrize("loop,kernel,msg", [
    ('ops_par_loop(k, "k", b, nd, r, ops_arg_dat(d, 1, s, "double", OPS_WRITE))', None, "integer literal"),
    ('ops_par_loop(k, "k", b, 2, r, ops_arg_dat(d, 1, s, dtype, OPS_WRITE))', None, "string literal"),
    ('ops_par_loop(k, "k", b, 2, r, ops_arg_dat(d, 1, s, "double", OPS_WRITE))', "void other(ACC<double> &a) {}",
     "not defined"),
    ('ops_par_loop(k, "k", b, 2, r, ops_arg_dat(d, 1, s, "double", OPS_WRITE), ops_arg_idx())', None,
     "takes 1 parameters"),
    ('ops_par_loop(k, "k", b, 2, r, ops_arg_dat(d, 1, s, "double", OPS_WRITE))', "void k(double *a) { *a = 1; }",
     "ACC<T> reference"),
    ('ops_par_loop(k, "k", b, 2, r, ops_arg_reduce(h, n, "double", OPS_INC))', "void k(double *x) {}",
     "integer literal for a reduction"),
    ('ops_par_loop(k, "k", b, 2, r, ops_arg_dat(d, 1, s, "complexd", OPS_WRITE))', None,
     "only double, float, int, half"),
])
def test_refuses_what_it_cannot_translate(tmp_path, loop, kernel, msg):
    app = _app(tmp_path, loop) if kernel is None else _app(tmp_path, loop, kernel)
    with pytest.raises(T.TranslateError, match=msg):
        T.translate(app, str(tmp_path / "out"))


def test_generated_library_built():
    lib = os.path.join(HERE, "genapps", "libgenapps.so")
    assert os.path.exists(lib), "run __graft_entry__.build() first"
    dyn = subprocess.run(["nm", "-D", lib], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)", dyn))
    assert {"gen_laplace", "gen_fields", "gen_heat_step", "gen_laplace_bad_stencil", "gen_ramp1d", "gen_vec3",
            "gen_opt_add", "gen_fill2d", "gen_double_in_place", "gen_cg_tail", "gen_lowdim3d", "gen_lowdim2d",
            "gen_lowdim_bad"} <= exported
    # the generated stubs stay inside the library (no clash with the
    # hand-written ops_par_loop_apply_stencil of libopsmi.so)
    assert not any(n.startswith("ops_par_loop_") for n in exported)
    assert re.search(r" U ops_mi_gen_par_loop\b", dyn)


def test_gbl_with_a_run_time_dim_gets_a_bounded_copy(tmp_path):
    app = _app(tmp_path, 'ops_par_loop(k, "k", b, 2, r, ops_arg_dat(d, 1, s, "double", OPS_WRITE), '
                         'ops_arg_gbl(xs, n, "double", OPS_READ))',
               "void k(ACC<double> &a, const double *xs) { a(0, 0) = xs[1]; }")
    T.translate(app, str(tmp_path / "out"))
    hip = open(tmp_path / "out" / "app_opsmi_kernels.hip").read()
    assert f"double g1[{T.GBL_CAP}];" in hip
    # the runtime learns the capacity (npts of a gbl argument) and refuses more
    assert f"static const int npts[2] = {{1, {T.GBL_CAP}}};" in hip


def test_legacy_pointer_kernels_translate(tmp_path):
    """Kernels in the legacy OPS_ACCn form (T * parameters indexed with
    OPS_ACCn(dx, dy) / OPS_ACC_MDn(d, dx, dy), ops/c/include/ops_macros.h)
    become accessor parameters; the literal offsets are listed as points
    (stencil check, fusion) like ACC<T> calls, and a pointer used any other
    way lists none."""
    p = tmp_path / "legacy.cpp"
    p.write_text('#include "ops_seq.h"\n'
                 'void lk(const double *a, double * restrict b, double *v) {\n'
                 '  b[OPS_ACC1(0,0)] = a[OPS_ACC0(1,0)] + a[OPS_ACC0(-1,0)]; *v = *v + b[OPS_ACC1(0,0)]; }\n'
                 'void lk2(const double *a, double *b) { b[OPS_ACC_MD1(1,0,0)] = a[OPS_ACC0(0,0)]; }\n'
                 'void run(ops_block bl, ops_dat d, ops_dat e, ops_stencil s0, ops_stencil s3, ops_reduction h, int *r) {\n'
                 '  ops_par_loop(lk, "lk", bl, 2, r, ops_arg_dat(d, 1, s3, "double", OPS_READ),\n'
                 '               ops_arg_dat(e, 1, s0, "double", OPS_WRITE), ops_arg_reduce(h, 1, "double", OPS_INC));\n'
                 '  ops_par_loop(lk2, "lk2", bl, 2, r, ops_arg_dat(e, 1, s0, "double", OPS_READ),\n'
                 '               ops_arg_dat(d, 2, s0, "double", OPS_WRITE));\n'
                 '}\n')
    T.translate(str(p), str(tmp_path / "out"))
    hip = open(tmp_path / "out" / "legacy_opsmi_kernels.hip").read()
    assert "const OPSMI_A0 &a, const OPSMI_A1 &b, double *v" in hip
    assert "b[OPS_ACC1(0,0)] = a[OPS_ACC0(1,0)]" in hip  # the body is kept as written
    stub = hip[hip.index("void ops_par_loop_lk("):]
    assert "static const int npts[3] = {2, 1, 0};" in stub
    assert "static const int pts[12] = {-1, -1, 0, 0, -1, 1, 0, 0, -1, 0, 0, 0};" in stub
    stub2 = hip[hip.index("void ops_par_loop_lk2("):]
    assert "static const int npts[2] = {1, 1};" in stub2 and "1, 0, 0, 0" in stub2  # component 1 of d
    assert "ops_mi_gen_fused_chain(" in hip
    assert T.legacy_points("{ b[OPS_ACC0(0,0)] = *(&a[OPS_ACC1(0,0)] + 1); }", "a", 2) is None
    assert T.accessed_points("{ w(0, 0) = *(&p(0, 0) + 1); }", "p", 2) is None
    assert T.legacy_points("{ b[OPS_ACC0(0,0)] = (a[OPS_ACC1(0,0)] > 0) && (a[OPS_ACC1(1,0)] > 0); }", "a", 2) \
        == [(-1, 0, 0, 0), (-1, 1, 0, 0)]
    assert T.legacy_points("{ double *q = a; b[OPS_ACC0(0,0)] = q[0]; }", "a", 2) is None
    assert T.legacy_points("{ b[OPS_ACC0(0,0)] = a[OPS_ACC1(i,0)]; }", "a", 2) is None
    app = open(tmp_path / "out" / "legacy_ops.cpp").read()
    assert "ops_par_loop_lk(" in app and '#include "ops_mi355x.h"' in app


def test_program_constants_and_globals(tmp_path):
    """Files of one program share constants (a user struct of run-time
    count declared in another file); headers that define globals stay out of
    the generated kernel files."""
    (tmp_path / "types.h").write_text("typedef struct { double a; int g; } st_t;\n")
    (tmp_path / "vars.h").write_text('#include "types.h"\nst_t *sts;\nint nst;\n')
    (tmp_path / "main.cpp").write_text('#include "ops_seq.h"\n#include "vars.h"\n'
                                       'void setup() { ops_decl_const("sts", nst, "st_t", sts); }\n')
    (tmp_path / "k.cpp").write_text('#include "ops_seq.h"\n#include "types.h"\n'
                                    'void kk(double *u) { u[OPS_ACC0(0,0)] = sts[1].a; }\n'
                                    'void run(ops_block bl, ops_dat d, ops_stencil s0, int *r) {\n'
                                    '  ops_par_loop(kk, "kk", bl, 2, r, ops_arg_dat(d, 1, s0, "double", OPS_WRITE));\n}\n')
    prog = T.Program([str(tmp_path / "main.cpp"), str(tmp_path / "k.cpp")])
    assert prog.consts["sts"] == (None, "st_t")
    T.translate(str(tmp_path / "k.cpp"), str(tmp_path / "out"), program=prog)
    hip = open(tmp_path / "out" / "k_opsmi_kernels.hip").read()
    assert "st_t sts[OPSMI_CONST_CAP(st_t)];" in hip
    assert 'ops_mi_gen_const_sized("sts", "st_t", sizeof(st_t), (int)OPSMI_CONST_CAP(st_t), 0, &b.k.sts);' in hip
    assert T.defines_globals(open(tmp_path / "vars.h").read())
    assert not T.defines_globals(open(tmp_path / "types.h").read())
    assert not T.defines_globals("extern int q;\nstatic const int k = 3;\nint f(int x);\n")


def test_implementation_files_and_commented_parameters(tmp_path):
    """An included file that defines functions with external linkage (an
    implementation file the program includes, as CloverLeaf's profile.cpp)
    stays out of the generated kernel file, which would define them again;
    inline, static and template definitions and declarations do not count.
    Comments inside a kernel's parameter list are dropped from the
    parameter text."""
    assert T.defines_functions("void process(int n) { n++; }\n")
    assert T.defines_functions("OPS_FUN_PREFIX double g(double x) { return x; }\n")
    assert not T.defines_functions("inline double g(double x) { return x; }\nstatic int h() { return 1; }\n"
                                   "template <class T> T k(T a) { return a; }\nstruct S { int a; void f() {} };\n"
                                   "double decl(double);\n#define M(x) { x }\n")
    (tmp_path / "prof.cpp").write_text("#include <stdio.h>\nvoid report() { printf(\"x\\n\"); }\n")
    (tmp_path / "helpers.h").write_text("static inline OPS_FUN_PREFIX double twice(double v) { return 2.0 * v; }\n")
    p = tmp_path / "app.cpp"
    p.write_text('#include "ops_seq.h"\n#include "helpers.h"\n#include "prof.cpp"\n'
                 'void kc(const double *a, // input\n'
                 '        double *b /* output */) {\n'
                 '  b[OPS_ACC1(0,0)] = twice(a[OPS_ACC0(0,0)]); }\n'
                 'void run(ops_block bl, ops_dat d, ops_dat e, ops_stencil s0, int *r) {\n'
                 '  ops_par_loop(kc, "kc", bl, 2, r, ops_arg_dat(d, 1, s0, "double", OPS_READ),\n'
                 '               ops_arg_dat(e, 1, s0, "double", OPS_WRITE));\n  report();\n}\n')
    T.translate(str(p), str(tmp_path / "out"))
    hip = open(tmp_path / "out" / "app_opsmi_kernels.hip").read()
    assert '#include "helpers.h"' in hip and '#include "prof.cpp"' not in hip
    assert "const OPSMI_A0 &a, const OPSMI_A1 &b) const {" in hip
