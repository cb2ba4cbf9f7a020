"""Benchmark: one step = one CG iteration of the 3D TeaLeaf-form implicit heat
solve on a 512^3 FP64 grid (7-point variable-coefficient operator, lognormal
conductivity), i.e. the four ops_par_loops of tea_leaf_cg_calc_w / calc_ur /
calc_p with their two FP64 reductions, driven by the C++ app over the C ABI.

value = algorithmic bytes of the step (OPS ops_compute_transfer rule: 1x for
READ/WRITE, 2x for RW/INC, 112 B/point, counted per loop as written even when
the lazy queue fuses the axpy+axpby pair into one 40 B/point pass) / measured
time, decimal GB/s.
roofline = the kernel with the largest share of the step
(tea_leaf_cg_calc_w_reduce_kernel, 40 B/point compulsory traffic) timed with
HIP events recorded on the library's stream around every launch.
cpu_baseline = the oracle's OpenMP restatement of the reference CPU path on
the same 512^3 problem, a few iterations on the host cores.

usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--size 512]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes per point of each kernel: the four TeaLeaf loops, the
# fused pointwise pair (reads u, p, r; writes u, p) and the fused chain
# (reads u, p_old, r, Kx, Ky, Kz; writes u, p_new, w)
BYTES_PER_PT = {"tea_leaf_cg_calc_w_reduce_kernel": 40, "tea_leaf_axpy_kernel": 24,
                "tea_leaf_cg_calc_ur_r_reduce_kernel": 24, "tea_leaf_axpby_kernel": 24,
                "cg_pair(axpy+axpby)": 40, "cg_chain(axpy+axpby+calc_w)": 72}


def cpu_baseline(n, seed, iters=3):
    """Oracle (OpenMP restatement of the reference CPU path) on the same
    problem: times `iters` CG iterations as the difference of two runs."""
    import numpy as np

    from oracle import oracle_py as O

    nt = os.cpu_count() or 1
    K, rhs = O.heat_inputs(seed, 3, n, n, n)
    t = []
    for k in (1, 1 + iters):
        t0 = time.perf_counter()
        O.heat_cg(3, n, n, n, K, rhs, max_iters=k, nthreads=nt)
        t.append(time.perf_counter() - t0)
    dt = max(t[1] - t[0], 1e-9) / iters
    gbs = 112.0 * n ** 3 / dt / 1e9
    del K, rhs
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": nt, "kind": "port",
            "sample": f"{iters} CG iterations of the same {n}^3 heat problem (oracle OpenMP, "
                      f"{dt * 1e3:.1f} ms/iteration)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20240611)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from ops_amd import apps, ops

    n = a.size
    ops.ops_init(diags=2)
    h = apps.Heat(3, n, n, n, seed=a.seed)
    h.iterate(a.warmup)
    ops.ops_mi_sync()
    ops.ops_mi_reset_stats()
    t0 = time.perf_counter()
    k = h.iterate(a.steps)
    ops.ops_mi_sync()
    t1 = time.perf_counter()
    assert k == a.steps
    ms = (t1 - t0) / a.steps * 1e3
    pts = float(n) ** 3
    step_bytes = 112.0 * pts
    value = step_bytes / (ms * 1e-3) / 1e9

    kernels = {}
    for name, bpp in BYTES_PER_PT.items():
        st = ops.ops_mi_loop_stats(name)
        if st and st[0] and st[1] > 0:
            cnt, kms, by = st
            kernels[name] = {"launches": cnt, "avg_us": round(kms / cnt * 1e3, 2),
                             "GB_s": round(bpp * pts / (kms / cnt * 1e-3) / 1e9, 1)}
    # the dominant kernel: largest total time in the step
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    bpp = BYTES_PER_PT[dom]
    avg_s = kernels[dom]["avg_us"] * 1e-6
    achieved = bpp * pts / avg_s / 1e9
    rr_rel = (h.rr() / h.rr(initial=True)) ** 0.5
    h.close()

    out = {
        "metric": "HBM GB/s per ops_par_loop; batched solves/s + p99 latency at 1e-10 residual",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based generator: lognormal conductivity, uniform + Gaussian rhs)",
        "config": {"workload": f"heat3d_cg_{n}", "grid": [n, n, n], "operator": "7-point variable coefficient",
                   "step": "one CG iteration (4 ops_par_loops, 2 FP64 reductions)",
                   "bytes_per_point": 112, "relres_after_steps": rr_rel},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": bpp * pts, "avg_launch_us": round(avg_s * 1e6, 2)},
        "kernels": kernels,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, a.seed)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
