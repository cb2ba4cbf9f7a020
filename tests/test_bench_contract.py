"""bench.py's contract, checked without a GPU: every ops.* / apps.* name it
calls exists in the bindings, every workload is registered, and --gpus N > 1
exits with a message (the backend runs on one GPU)."""
import ast
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _names(module_alias):
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    out = set()
    for node in ast.walk(tree):
        if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) and node.value.id == module_alias:
            out.add(node.attr)
    return out


def test_bench_calls_only_existing_bindings():
    from ops_amd import apps, ops
    missing = [f"ops.{n}" for n in sorted(_names("ops")) if not hasattr(ops, n)]
    missing += [f"apps.{n}" for n in sorted(_names("apps")) if not hasattr(apps, n)]
    assert not missing, missing


def test_bench_workloads_registered():
    sys.path.insert(0, ROOT)
    import bench
    for w in ("heat3d_cg_512", "poisson_mixed_768", "laplace2d_2048", "adi_trid_256", "serve_stream",
              "gen_cg_tail_512", "stream_mix"):
        assert callable(bench.WORKLOADS[w]), w


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr
