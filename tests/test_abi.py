"""The C-ABI library loads and exports every entry point include/ops_mi355x.h
declares, and errors cross the boundary as codes, not exceptions or aborts
(CPU only: nothing here touches a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from ops_amd import ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_built():
    assert os.path.exists(ops.LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(ops.LIB_PATH)
    names = ops.declared_symbols()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_exports_have_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", ops.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    for n in ops.declared_symbols():
        assert n in exported, n


def test_par_loop_entry_points_cover_reference_kernels():
    names = set(ops.declared_symbols())
    for k in ["apply_stencil", "copy", "set_zero", "left_bndcon", "right_bndcon",
              "tea_leaf_cg_calc_w_reduce_kernel", "tea_leaf_cg_calc_ur_r_reduce_kernel", "tea_leaf_axpy_kernel",
              "tea_leaf_axpby_kernel", "tea_leaf_dot_kernel", "tea_leaf_common_init_Kx_Ky_kernel",
              "tea_leaf_common_init_kernel", "preproc_kernel"]:
        assert "ops_par_loop_" + k in names, k
    for k in ["ops_tridMultiDimBatch", "ops_tridMultiDimBatch_Inc", "ops_tridMultiDimBatch_c",
              "ops_tridMultiDimBatch_Inc_c", "ops_reduction_result_char", "ops_decl_const_char",
              "ops_decl_dat_char", "ops_arg_reduce", "ops_arg_gbl_char", "ops_halo_transfer",
              "ops_dat_fetch_data", "ops_dat_set_data", "ops_dat_get_raw_pointer"]:
        assert k in names, k


def test_error_before_init_is_reported_not_raised():
    # runs in a child process so this interpreter never initialises HIP
    code = (
        "from ops_amd import ops\n"
        "try:\n"
        "    ops.ops_decl_block(2, 'b')\n"
        "except ops.OPSError as e:\n"
        "    print('code', e.code)\n"
    )
    r = subprocess.run(["python", "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert "code 2" in r.stdout, r.stdout + r.stderr


def test_header_compiles_as_c_and_cpp(tmp_path):
    src_c = tmp_path / "t.c"
    src_c.write_text('#include "ops_mi355x.h"\nint main(void){ops_arg a = ops_arg_idx(); return a.argtype;}\n')
    subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src_c)],
                   check=True)
    src_cpp = tmp_path / "t.cpp"
    src_cpp.write_text('#include "ops_mi355x.h"\nint main(){double x=1; ops_arg a = ops_arg_gbl(&x, 1, "double", OPS_READ);'
                       ' return a.dim;}\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src_cpp)],
                   check=True)


def test_reference_tridiagonal_call_compiles_as_c_and_cpp(tmp_path):
    """The reference's calls compile unchanged: C++ constructs the
    ops_tridsolver_params class (apps/c/adi/adi.cpp style), C fills the struct
    with the settings object; both reach the C-linkage symbols, whose
    mangled-free names the library exports."""
    inc = os.path.join(ROOT, "include")
    src_cpp = tmp_path / "trid.cpp"
    src_cpp.write_text(
        '#include "ops_mi355x.h"\n'
        'void solve(ops_block b, ops_dat a, ops_dat bb, ops_dat c, ops_dat d, ops_dat u, int *r) {\n'
        '  ops_tridsolver_params p(b);\n'
        '  ops_tridsolver_params q(b, ops_tridsolver_params::PCR);\n'
        '  ops_tridMultiDimBatch(3, 0, r, a, bb, c, d, &p);\n'
        '  ops_tridMultiDimBatch_Inc(3, 2, r, a, bb, c, d, u, &q);\n'
        '}\n'
        'static_assert(sizeof(ops_tridsolver_params) == sizeof(void *), "one data member");\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", inc, str(src_cpp)], check=True)
    src_c = tmp_path / "trid.c"
    src_c.write_text(
        '#include "ops_mi355x.h"\n'
        'void solve(ops_block b, ops_dat a, ops_dat bb, ops_dat c, ops_dat d, int *r) {\n'
        '  ops_tridsolver_params p = {ops_tridsolver_params_create(b, 0)};\n'
        '  ops_tridMultiDimBatch(3, 1, r, a, bb, c, d, &p);\n'
        '  ops_tridsolver_params_destroy(p.tridsolver_params);\n'
        '}\n')
    subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-I", inc, str(src_c)], check=True)
    out = subprocess.run(["nm", "-D", "--defined-only", ops.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    assert {"ops_tridMultiDimBatch", "ops_tridMultiDimBatch_Inc"} <= exported


def test_init_flags_parsed_without_gpu():
    """OPS_TILING / OPS_BLOCK_SIZE_* tokens: an out-of-range block size is a
    configuration error raised before the device is touched (child process:
    this interpreter never initialises HIP)."""
    code = (
        "from ops_amd import ops\n"
        "try:\n"
        "    ops.ops_init(argv=['app', 'OPS_TILING', 'OPS_BLOCK_SIZE_Z=-3'])\n"
        "except ops.OPSError as e:\n"
        "    print('code', e.code, e)\n"
    )
    r = subprocess.run(["python", "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert "code 2" in r.stdout and "OPS_BLOCK_SIZE" in r.stdout, r.stdout + r.stderr
