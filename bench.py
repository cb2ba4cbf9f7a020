"""Benchmarks of the OPS hot path on one GPU.  Each run prints one JSON line.

Default workload (heat3d_cg_512): one step = one CG iteration of the 3D
TeaLeaf-form implicit heat solve on a 512^3 FP64 grid (7-point
variable-coefficient operator, lognormal conductivity), i.e. the four
ops_par_loops of tea_leaf_cg_calc_w / calc_ur / calc_p with their two FP64
reductions, driven by the C++ app over the C ABI.
  value = algorithmic bytes of the step (OPS ops_compute_transfer rule: 1x for
  READ/WRITE, 2x for RW/INC, 112 B/point, counted per loop as written even
  when the lazy queue fuses the axpy+axpby pair into one 40 B/point pass) /
  measured time, decimal GB/s.
  roofline = the kernel with the largest share of the step (the fused
  axpy + axpby + calc_w chain, 72 B/point compulsory traffic; calc_w alone,
  40 B/point, with OPSMI_FUSION=1) timed with HIP events recorded on the
  library's stream around every launch.
  cpu_baseline = the oracle's OpenMP restatement of the reference CPU path on
  the same 512^3 problem, a few iterations on the host cores.
  configs = the other benchmark configurations, measured in the same run
  (--no-configs skips them): laplace2d_2048, adi_trid_256,
  poisson_mixed_768 and serve_stream (60 requests), each with its own
  roofline or speed-up figures.

Workloads on their own (--workload):
  poisson_mixed_768  768^3 7-point Poisson to 1e-10 relative residual: FP64
                     CG against FP64 iterative refinement around FP32 PCG
                     (solves/s, iteration counts, error against FP64 CG)
  laplace2d_2048     2048^2 Jacobi with max-abs-diff reduction, 1000 iterations
  adi_trid_256       256^3 batched tridiagonal solves along x, y, z (Thomas)
  serve_stream       request stream (60 % 2D Poisson, 40 % 3D heat, uneven
                     sizes) on the batched engine: solves/s with all
                     requests queued, p50/p99 latency under Poisson arrivals,
                     against the per-solve FP64 path one solve at a time
  gen_cg_tail_512    512^3 CG tail u = u + a p, then p = r + b p with p.p,
                     written as two dev-API loops (tests/genapps/misc_gen.cpp)
                     and run as the translator's fused pair (one pass, 40 B/pt)
  stream_mix         pointwise read/write mixes of the path's kernels over
                     512^3-point FP64 arrays (plain and non-temporal): the
                     practical HBM ceiling each kernel is measured against

usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload heat3d_cg_512]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "HBM GB/s per ops_par_loop; batched solves/s + p99 latency at 1e-10 residual"
# compulsory HBM bytes per point of each kernel: the four TeaLeaf loops, the
# fused pointwise pair (reads u, p, r; writes u, p) and the fused chain
# (reads u, p_old, r, Kx, Ky, Kz; writes u, p_new, w)
BYTES_PER_PT = {"tea_leaf_cg_calc_w_reduce_kernel": 40, "tea_leaf_axpy_kernel": 24,
                "tea_leaf_cg_calc_ur_r_reduce_kernel": 24, "tea_leaf_axpby_kernel": 24,
                "cg_pair(axpy+axpby)": 40, "cg_chain(axpy+axpby+calc_w)": 72}


def kernel_table(ops, bpp_map, pts):
    out = {}
    for name, bpp in bpp_map.items():
        st = ops.ops_mi_loop_stats(name)
        if st and st[0] and st[1] > 0:
            cnt, kms, _ = st
            out[name] = {"launches": cnt, "avg_us": round(kms / cnt * 1e3, 2),
                         "GB_s": round(bpp * pts / (kms / cnt * 1e-3) / 1e9, 1)}
    return out


def profiled_traffic(profile, kernel):
    """HBM bytes per launch measured with rocprofv3 PMC (FETCH_SIZE, WRITE_SIZE,
    calibrated on a copy kernel) for the workload's own configuration, from the
    summary kept under profiles/; None when not recorded."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        f = ln.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def roofline(kernels, bpp_map, pts, traffic=None):
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    avg_s = kernels[dom]["avg_us"] * 1e-6
    achieved = bpp_map[dom] * pts / avg_s / 1e9
    return {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": bpp_map[dom] * pts,
            "avg_launch_us": round(avg_s * 1e6, 2)}


def line(value, unit, ms, a, dtype, config, **extra):
    out = {"metric": METRIC, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype,
           "data": "synthetic (seeded counter-based generator)", "config": config}
    out.update(extra)
    return out


# ------------------------------------------------------------------ heat CG
def host_threads():
    """Threads for the CPU baseline: the cores this process may run on, at
    most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_heat(n, seed, target_s=10.0):
    """Oracle (OpenMP restatement of the reference CPU path) on the same
    problem: CG iterations timed by the oracle around its iteration loop
    (set-up excluded); a 3-iteration probe sizes the sample to about
    target_s seconds (3 to 200 iterations)."""
    from oracle import oracle_py as O

    nt = host_threads()
    K, rhs = O.heat_inputs(seed, 3, n, n, n)
    O.heat_cg(3, n, n, n, K, rhs, max_iters=3, nthreads=nt)
    iters = int(min(200, max(3, target_s / max(O.last_loop_seconds() / 3, 1e-6))))
    O.heat_cg(3, n, n, n, K, rhs, max_iters=iters, nthreads=nt)
    dt = max(O.last_loop_seconds(), 1e-9) / iters
    gbs = 112.0 * n ** 3 / dt / 1e9
    # one thread, as the reference's sequential (dev_seq) build runs
    O.heat_cg(3, n, n, n, K, rhs, max_iters=2, nthreads=1)
    dt1 = max(O.last_loop_seconds(), 1e-9) / 2
    del K, rhs
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": nt, "kind": "port",
            "sample": f"{iters} CG iterations of the same {n}^3 heat problem (oracle OpenMP, "
                      f"{dt * 1e3:.1f} ms/iteration)",
            "one_thread": {"value": round(112.0 * n ** 3 / dt1 / 1e9, 2), "unit": "GB/s",
                           "ms_per_iteration": round(dt1 * 1e3, 1), "sample": "2 CG iterations, 1 thread"}}


def bench_heat(a, ops, apps):
    n = a.size or 512
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
    value = 112.0 * pts / (ms * 1e-3) / 1e9
    kernels = kernel_table(ops, BYTES_PER_PT, pts)
    rr_rel = (h.rr() / h.rr(initial=True)) ** 0.5
    h.close()
    rl = roofline(kernels, BYTES_PER_PT, pts)
    if n == 512:
        rl["traffic"] = profiled_traffic("heat3d_cg_512_hbm_traffic.md", rl["kernel"])
    out = line(value, "GB/s", ms, a, "f64",
               {"workload": f"heat3d_cg_{n}", "grid": [n, n, n], "operator": "7-point variable coefficient",
                "step": "one CG iteration (4 ops_par_loops, 2 FP64 reductions)", "bytes_per_point": 112,
                "relres_after_steps": rr_rel},
               roofline=rl, kernels=kernels,
               value_basis="OPS ops_compute_transfer rule: 112 B/pt for the four loops as written, "
                           "also when the queue fuses them (96 B/pt compulsory after fusion)")
    # HBM traffic actually moved per step, from the PMC profile of the two
    # kernels that make up the step (reduction folds add under 0.1 %)
    if n == 512:
        t_chain = profiled_traffic("heat3d_cg_512_hbm_traffic.md", "cg_chain(axpy+axpby+calc_w)")
        t_urr = profiled_traffic("heat3d_cg_512_hbm_traffic.md", "tea_leaf_cg_calc_ur_r_reduce_kernel")
        if t_chain and t_urr:
            out["hbm_traffic_per_step"] = {"bytes": t_chain + t_urr,
                                           "GB_s": round((t_chain + t_urr) / (ms * 1e-3) / 1e9, 1),
                                           "frac_of_peak": round((t_chain + t_urr) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                           "source": "profiles/heat3d_cg_512_hbm_traffic.md (PMC)"}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_heat(n, a.seed)
    if not a.no_configs:
        # the other configurations, measured in the same run (one GPU)
        out["configs"] = {
            "laplace2d_2048": measure_laplace(ops, apps, 2048),
            "adi_trid_256": measure_trid(ops, apps, 256, a.seed),
            "poisson_mixed_768": measure_mixed(ops, apps, 768, a.seed),
            "serve_stream": measure_serve(ops, apps, 60, a.seed),
        }
    return out


# ------------------------------------------------------ mixed-precision Poisson
# compulsory bytes per point; the fused chains read u/e, p_old, r/rr and
# write u/e, p_new, w (48 B in FP64, 24 B in FP32)
MIXED_BPP = {"poisson_calc_w_reduce_kernel": 8, "mp_update_kernel": 24, "mp_rr_update_kernel": 12,
             "mp_e_update_kernel": 12, "mp_p_kernel": 12, "mp_chain(e+p+calc_w)": 24, "mp_chain(p+calc_w)": 16,
             "poisson_residual_kernel": 24, "mp_init_kernel": 20, "mp_correct_kernel": 20}
FP64_BPP = {"poisson_calc_w_reduce_kernel": 16, "tea_leaf_axpy_kernel": 24,
            "tea_leaf_cg_calc_ur_r_reduce_kernel": 24, "tea_leaf_axpby_kernel": 24, "cg_pair(axpy+axpby)": 40,
            "poisson_chain(axpy+axpby+calc_w)": 48}


def pois_roofline(kernels, bpp, pts, n):
    """roofline of the dominant chain, with its HBM traffic per launch from
    the PMC profile at 768^3"""
    rl = roofline(kernels, bpp, pts)
    if n == 768:
        rl["traffic"] = profiled_traffic("poisson_mixed_768_hbm_traffic.md", rl["kernel"])
    return rl


def measure_mixed(ops, apps, n, seed):
    """768^3 Poisson to 1e-10: FP64 CG, then FP64 refinement around FP32 PCG,
    each timed from resident inputs to the converged solution."""
    import numpy as np

    tol = 1e-10
    pts = float(n) ** 3
    ops.ops_mi_set_diags(2)
    # FP64 CG (x0 = 0, stop at |r|/|f| <= tol on the recursive residual)
    P = apps.Poisson(3, n, n, n, seed=seed)
    ops.ops_mi_sync()
    ops.ops_mi_reset_stats()
    t0 = time.perf_counter()
    it64 = P.iterate(100000, tol=tol)
    ops.ops_mi_sync()
    t64 = time.perf_counter() - t0
    k64 = kernel_table(ops, FP64_BPP, pts)
    x64 = P.x()
    P.close()
    print(f"poisson {n}^3 fp64 CG: {it64} iterations, {t64:.2f} s", file=sys.stderr, flush=True)
    M = apps.PoissonMixed(3, n, n, n, seed=seed)
    ops.ops_mi_sync()
    ops.ops_mi_reset_stats()
    t0 = time.perf_counter()
    res = M.solve(tol=tol)
    ops.ops_mi_sync()
    tmp = time.perf_counter() - t0
    kmp = kernel_table(ops, MIXED_BPP, pts)
    xm = M.x()
    M.close()
    print(f"poisson {n}^3 mixed: {res['outer']} outer, {res['inner_total']} inner, {tmp:.2f} s", file=sys.stderr,
          flush=True)
    err = float(np.max(np.abs(xm - x64)) / np.max(np.abs(x64)))
    del xm, x64
    return {"workload": f"poisson_mixed_{n}", "grid": [n, n, n], "tol": tol,
            "fp64_cg": {"iterations": it64, "seconds": round(t64, 3), "solves_per_s": round(1.0 / t64, 4),
                        "roofline": pois_roofline(k64, FP64_BPP, pts, n), "kernels": k64},
            "mixed": {"outer": res["outer"], "inner_total": res["inner_total"],
                      "relres": [float(v) for v in res["relres"]], "seconds": round(tmp, 3),
                      "solves_per_s": round(1.0 / tmp, 4), "roofline": pois_roofline(kmp, MIXED_BPP, pts, n),
                      "kernels": kmp},
            "speedup_vs_fp64": round(t64 / tmp, 3), "max_rel_diff_vs_fp64_cg": err}


def bench_mixed(a, ops, apps):
    n = a.size or 768
    ops.ops_init(diags=2)
    m = measure_mixed(ops, apps, n, a.seed)
    tmp = m["mixed"]["seconds"]
    return line(1.0 / tmp, "solves/s", tmp * 1e3, a, "f32 sweeps, f64 reductions/residual/correction",
                {"workload": m["workload"], "grid": m["grid"], "operator": "7-point constant coefficient",
                 "tol": m["tol"]},
                fp64_cg=m["fp64_cg"], mixed=m["mixed"], speedup_vs_fp64=m["speedup_vs_fp64"],
                max_rel_diff_vs_fp64_cg=m["max_rel_diff_vs_fp64_cg"], roofline=m["mixed"]["roofline"])


# ------------------------------------------------------------------ Laplace
def measure_laplace(ops, apps, n, iters=1000):
    """2048^2 Jacobi, 1000 iterations, each with its max-abs-diff result
    (the tutorial's loop); per-iteration time from the difference of a
    (10 + iters)- and a 10-iteration run, so set-up cancels.  The timed runs
    record no timing events (each costs about 1.8 us of a 27 us iteration);
    the per-kernel table comes from a third run with events."""
    ops.ops_mi_set_diags(1)
    apps.laplace2d(n, n, 10, want_field=False)
    ops.ops_mi_sync()
    t0 = time.perf_counter()
    apps.laplace2d(n, n, 10, want_field=False)
    t1 = time.perf_counter()
    t2 = time.perf_counter()
    err, _, _ = apps.laplace2d(n, n, 10 + iters, want_field=False)
    t3 = time.perf_counter()
    ms = ((t3 - t2) - (t1 - t0)) / iters * 1e3
    ops.ops_mi_set_diags(2)
    ops.ops_mi_reset_stats()
    apps.laplace2d(n, n, 10 + iters, want_field=False)
    ops.ops_mi_sync()
    pts = float(n) ** 2
    # the queue runs each copy with the next apply_stencil as one pass
    # (reads Anew, writes A and the new Anew: 24 B/pt)
    bpp = {"apply stencil": 16, "copy": 16, "jacobi_chain(copy+apply_stencil)": 24}
    kernels = kernel_table(ops, bpp, pts)
    return {"workload": f"laplace2d_{n}", "grid": [n, n], "iterations": iters, "bytes_per_point": 32,
            "ms_per_iteration": round(ms, 5), "GB_s": round(32.0 * pts / (ms * 1e-3) / 1e9, 1),
            "final_max_abs_diff": err, "roofline": roofline(kernels, bpp, pts), "kernels": kernels}


def bench_laplace(a, ops, apps):
    n = a.size or 2048
    ops.ops_init(diags=2)
    m = measure_laplace(ops, apps, n)
    return line(m["GB_s"], "GB/s", m["ms_per_iteration"], a, "f64",
                {k: m[k] for k in ("workload", "grid", "iterations", "bytes_per_point", "final_max_abs_diff")},
                roofline=m["roofline"], kernels=m["kernels"])


# --------------------------------------------------------------- tridiagonal
# algorithmic bytes per point: R a, b, c and RW d (x, y); R a, b, c, d and RW u (z, _Inc)
TRID_BPP = {"ops_tridMultiDimBatch x": 40, "ops_tridMultiDimBatch y": 40, "ops_tridMultiDimBatch z": 40,
            "ops_tridMultiDimBatch_Inc x": 48, "ops_tridMultiDimBatch_Inc y": 48, "ops_tridMultiDimBatch_Inc z": 48}


def trid_roofline(kernels, pts, n):
    """roofline of the dominant solve, with its HBM traffic per launch from
    the PMC profile at 256^3"""
    rl = roofline(kernels, TRID_BPP, pts)
    if n == 256:
        rl["traffic"] = profiled_traffic("adi_trid_256_hbm_traffic.md", rl["kernel"].replace(" ", "_"))
    return rl


def measure_trid(ops, apps, n, seed, sweeps=20, warmup=3):
    """256^3 Thomas solves along x, y, then z with _Inc (the ADI sweep)."""
    ops.ops_mi_set_diags(2)
    T = apps.Trid(n, n, n, seed=seed, strategy=0)
    for _ in range(warmup):
        T.sweep(inc=True)
    ops.ops_mi_sync()
    ops.ops_mi_reset_stats()
    t0 = time.perf_counter()
    for _ in range(sweeps):
        T.sweep(inc=True)
    ops.ops_mi_sync()
    ms = (time.perf_counter() - t0) / sweeps * 1e3
    T.close()
    pts = float(n) ** 3
    kernels = kernel_table(ops, TRID_BPP, pts)
    return {"workload": f"adi_trid_{n}", "grid": [n, n, n], "step": "Thomas solves along x, y, z (z with _Inc)",
            "bytes_per_point": 128, "ms_per_sweep": round(ms, 4),
            "GB_s": round(128.0 * pts / (ms * 1e-3) / 1e9, 1), "roofline": trid_roofline(kernels, pts, n),
            "kernels": kernels}


def bench_trid(a, ops, apps):
    n = a.size or 256
    ops.ops_init(diags=2)
    m = measure_trid(ops, apps, n, a.seed, a.steps, a.warmup)
    return line(m["GB_s"], "GB/s", m["ms_per_sweep"], a, "f64",
                {k: m[k] for k in ("workload", "grid", "step", "bytes_per_point")}, roofline=m["roofline"],
                kernels=m["kernels"])


# ------------------------------------------------------------------ serving
def serve_requests(nreq, seed, frac2d=0.6):
    """60 % (frac2d) 2D 5-point Poisson with sides log-uniform in [64, 2048],
    the rest 3D 7-point heat with sides log-uniform in [32, 256]."""
    import numpy as np

    g = np.random.default_rng(seed)
    kinds = np.where(g.random(nreq) < frac2d, 0, 1).astype(np.int32)
    s2 = np.exp(g.uniform(np.log(64), np.log(2048), nreq))
    s3 = np.exp(g.uniform(np.log(32), np.log(256), nreq))
    sizes = np.where(kinds == 0, s2, s3).round().astype(np.int32)
    seeds = g.integers(1, 2 ** 62, nreq, dtype=np.uint64)
    return kinds, sizes, seeds, g


def measure_serve(ops, apps, nreq, seed, frac2d=0.6):
    """The request stream on the batched engine (FP64, then FP32-inner
    refinement for the 2D Poisson requests), all queued at t = 0 and under
    Poisson arrivals, against the per-solve FP64 path (the library apps, one
    solve after another).  Launches are not timed with events here (diags 1),
    so neither path pays for them."""
    import numpy as np

    tol, budget = 1e-10, 200 * 10 ** 9
    kinds, sizes, seeds, g = serve_requests(nreq, seed, frac2d)
    ops.ops_mi_set_diags(1)
    # saturation: every request queued at t = 0
    span, done, iters = apps.serve_stream(kinds, sizes, seeds, np.zeros(nreq), tol, budget, autotune=True)
    assert np.all(iters > 0), "a solve stopped without converging"
    sat = nreq / span
    print(f"serve: batched, all queued at t=0: {nreq} solves in {span:.2f} s", file=sys.stderr, flush=True)
    span_m, _, iters_m = apps.serve_stream(kinds, sizes, seeds, np.zeros(nreq), tol, budget, autotune=True,
                                           mixed=True)
    assert np.all(iters_m > 0), "a mixed solve stopped without converging"
    sat_m = nreq / span_m
    print(f"serve: batched FP32-inner, all queued at t=0: {nreq} solves in {span_m:.2f} s", file=sys.stderr,
          flush=True)
    # latency under Poisson arrivals at 70 % of the saturated rate
    rate = 0.7 * sat
    arr = np.cumsum(g.exponential(1.0 / rate, nreq))
    arr -= arr[0]
    span2, done2, iters2 = apps.serve_stream(kinds, sizes, seeds, arr, tol, budget, autotune=True)
    lat = done2 - arr
    print(f"serve: batched, Poisson arrivals at {rate:.1f}/s: span {span2:.2f} s", file=sys.stderr, flush=True)
    # per-solve FP64 path: the library apps, one solve after another
    t_base, it_base, t_each = 0.0, [], []
    for j, (k, n, s) in enumerate(zip(kinds, sizes, seeds)):
        n = int(n)
        P = apps.Poisson(2, n, n, 1, seed=int(s)) if k == 0 else apps.Heat(3, n, n, n, seed=int(s))
        ops.ops_mi_sync()
        t0 = time.perf_counter()
        it_base.append(P.iterate(1000000, tol=tol))
        ops.ops_mi_sync()
        t_each.append(time.perf_counter() - t0)
        t_base += t_each[-1]
        P.close()
        if j % 20 == 19:
            print(f"serve: per-solve path: {j + 1} solves, {t_base:.2f} s", file=sys.stderr, flush=True)
    base = nreq / t_base
    ops.ops_mi_set_diags(2)
    return {"workload": f"serve_stream_{nreq}", "requests": nreq, "tol": tol, "budget_bytes": budget,
            "mix": f"{100 * frac2d:.0f}% 2D 5-pt Poisson side log-U[64,2048], {100 * (1 - frac2d):.0f}% 3D 7-pt heat "
                   "side log-U[32,256]",
            "batched_saturated": {"solves_per_s": round(sat, 3), "makespan_s": round(span, 3)},
            "batched_fp32_inner_saturated": {"solves_per_s": round(sat_m, 3), "makespan_s": round(span_m, 3)},
            "batched_poisson_arrivals": {"rate_per_s": round(rate, 3),
                                         "p50_latency_s": float(np.percentile(lat, 50)),
                                         "p99_latency_s": float(np.percentile(lat, 99)),
                                         "solves_per_s": round(nreq / span2, 3)},
            "per_solve_fp64": {"solves_per_s": round(base, 3), "solve_seconds_total": round(t_base, 3)},
            "speedup_batched_fp64_vs_per_solve": round(sat / base, 3),
            "speedup_batched_fp32_inner_vs_per_solve": round(sat_m / base, 3),
            "iterations": {"batched_total": int(iters.sum()), "batched_fp32_inner_total": int(iters_m.sum()),
                           "per_solve_total": int(sum(it_base))},
            # the requests behind the tail latency: kind, side, CG iterations,
            # latency in the arrival stream and the same solve's time alone on
            # the per-solve path
            "slowest_requests": [{"kind": "poisson2d" if kinds[i] == 0 else "heat3d", "side": int(sizes[i]),
                                  "iterations": int(iters2[i]), "latency_s": round(float(lat[i]), 4),
                                  "alone_per_solve_s": round(t_each[i], 4)}
                                 for i in np.argsort(lat)[::-1][:3]]}


def bench_serve(a, ops, apps):
    ops.ops_init(diags=1)
    m = measure_serve(ops, apps, a.requests, a.seed, a.frac2d)
    sat = m["batched_saturated"]["solves_per_s"]
    return line(sat, "solves/s", 1e3 / sat, a, "f64",
                {k: m[k] for k in ("workload", "requests", "tol", "budget_bytes", "mix")},
                **{k: m[k] for k in m if k not in ("workload", "requests", "tol", "budget_bytes", "mix")})


# ------------------------------------------------------ generated loops
def bench_gen_tail(a, ops, apps):
    """One step = one iteration of the CG tail through translator-generated
    loops (libgenapps.so, built by __graft_entry__.build()): the queue runs
    cg_axpy and cg_axpby_dot as one fused pass, then the p.p result is read."""
    import ctypes as C

    import numpy as np
    lib = os.path.join(ROOT, "tests", "genapps", "libgenapps.so")
    if not os.path.exists(lib):
        raise RuntimeError(f"{lib} is not built; run __graft_entry__.build() first")
    G = C.CDLL(lib)
    dp = C.POINTER(C.c_double)
    G.gen_cg_tail.argtypes = [C.c_void_p] * 4 + [C.c_int] * 3 + [C.c_double] * 2 + [C.c_int, dp]
    n = a.size or 512
    ops.ops_init(diags=2)
    blk = ops.ops_decl_block(3, "gen_tail")
    rng = np.random.default_rng(a.seed)
    D = []
    for k in ("u", "p", "r"):
        d = ops.ops_decl_dat(blk, [n, n, n], name=k)
        ops.set_whole(d, rng.random((n + 2, n + 2, n + 2)) - 0.5)
        D.append(d)
    pp = np.zeros(max(a.steps, a.warmup, 1))
    args = [blk.ptr] + [d.ptr for d in D] + [n, n, n, 0.37, -0.61]
    G.gen_cg_tail(*args, a.warmup, pp.ctypes.data_as(dp))
    ops.ops_mi_sync()
    ops.ops_mi_reset_stats()
    t0 = time.perf_counter()
    G.gen_cg_tail(*args, a.steps, pp.ctypes.data_as(dp))
    ops.ops_mi_sync()
    t1 = time.perf_counter()
    ops._check()
    ms = (t1 - t0) / a.steps * 1e3
    pts = float(n) ** 3
    # fused: reads u, p, r, writes u, p; apart: 24 B/pt each
    bpp = {"fused(cg_axpy+cg_axpby_dot)": 40, "cg_axpy": 24, "cg_axpby_dot": 24}
    kernels = kernel_table(ops, bpp, pts)
    value = 40.0 * pts / (ms * 1e-3) / 1e9
    return line(value, "GB/s", ms, a, "f64",
                {"workload": f"gen_cg_tail_{n}", "grid": [n, n, n],
                 "loops": "cg_axpy + cg_axpby_dot (translator-generated, fused pair)", "bytes_per_point": 40,
                 "last_pp": float(pp[a.steps - 1])},
                roofline=roofline(kernels, bpp, pts), kernels=kernels, data="synthetic (numpy PCG64, seeded)")


def bench_gen_heat(a, ops, apps):
    """One step = the TeaLeaf CG operator (w = A p with p.w, 7-point variable
    coefficient, 40 B/pt) and the residual update (r -= alpha w with r.r,
    24 B/pt) at 512^3 through translator-generated loops
    (tests/genapps/heat3d_gen.cpp): at the default fusion level they run as
    one pass (56 B/pt, w is written and read at the point), with
    OPSMI_FUSION=0 as two.  Both results are read each step."""
    import ctypes as C

    import numpy as np
    lib = os.path.join(ROOT, "tests", "genapps", "libgenapps.so")
    if not os.path.exists(lib):
        raise RuntimeError(f"{lib} is not built; run __graft_entry__.build() first")
    G = C.CDLL(lib)
    dp = C.POINTER(C.c_double)
    G.gen_heat_step.argtypes = [C.c_void_p] * 7 + [C.c_int] * 3 + [C.c_double] * 4 + [dp]
    n = a.size or 512
    ops.ops_init(diags=2)
    blk = ops.ops_decl_block(3, "gen_heat")
    rng = np.random.default_rng(a.seed)
    D = {}
    for k in ("w", "Kx", "Ky", "Kz", "p", "r"):
        d = ops.ops_decl_dat(blk, [n, n, n], name=k)
        v = rng.random((n + 2, n + 2, n + 2))
        ops.set_whole(d, v + 0.5 if k.startswith("K") else v - 0.5)
        D[k] = d
    out = np.zeros(2)
    args = [blk.ptr] + [D[k].ptr for k in ("w", "Kx", "Ky", "Kz", "p", "r")] + [n, n, n, 0.3, 0.3, 0.3, 1e-3,
                                                                                 out.ctypes.data_as(dp)]
    for _ in range(a.warmup):
        G.gen_heat_step(*args)
    ops.ops_mi_sync()
    ops.ops_mi_reset_stats()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        G.gen_heat_step(*args)
    ops.ops_mi_sync()
    t1 = time.perf_counter()
    ops._check()
    ms = (t1 - t0) / a.steps * 1e3
    pts = float(n) ** 3
    bpp = {"fused(gen heat_calc_w+gen heat_ur_r)": 56, "gen heat_calc_w": 40, "gen heat_ur_r": 24}
    kernels = kernel_table(ops, bpp, pts)
    value = 64.0 * pts / (ms * 1e-3) / 1e9
    return line(value, "GB/s", ms, a, "f64",
                {"workload": f"gen_heat_step_{n}", "grid": [n, n, n],
                 "loops": "TeaLeaf 3D calc_w + ur_r (translator-generated)", "bytes_per_point": 64,
                 "pw": float(out[0]), "rr": float(out[1])},
                roofline=roofline(kernels, bpp, pts), kernels=kernels, data="synthetic (numpy PCG64, seeded)",
                value_basis="64 B/pt: the two loops as written (40 + 24)")


def bench_gen_overhead(a, ops, apps):
    """One step = one small generated loop (a 64 x 64 fill): the per-loop
    cost of the generated path when the grid is too small to matter (the
    stub, the queue, the launch), against the kernel's own time from HIP
    events in a second pass."""
    import ctypes as C
    lib = os.path.join(ROOT, "tests", "genapps", "libgenapps.so")
    if not os.path.exists(lib):
        raise RuntimeError(f"{lib} is not built; run __graft_entry__.build() first")
    G = C.CDLL(lib)
    G.gen_fill2d_many.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
    ops.ops_init(diags=1)
    blk = ops.ops_decl_block(2, "gen_ovh")
    d = ops.ops_decl_dat(blk, [64, 64], name="f")
    G.gen_fill2d_many(blk.ptr, d.ptr, 64, 64, max(a.warmup, 1))
    ops.ops_mi_sync()
    reps = max(a.steps, 1) * 1000
    t0 = time.perf_counter()
    G.gen_fill2d_many(blk.ptr, d.ptr, 64, 64, reps)
    ops.ops_mi_sync()
    t1 = time.perf_counter()
    us = (t1 - t0) / reps * 1e6
    ops.ops_mi_set_diags(2)
    ops.ops_mi_reset_stats()
    G.gen_fill2d_many(blk.ptr, d.ptr, 64, 64, 1000)
    ops.ops_mi_sync()
    st = ops.ops_mi_loop_stats("fill2d many")
    kern_us = st[1] / st[0] * 1e3 if st and st[0] else None
    ops._check()
    return line(1e6 / us, "loops/s", us * 1e-3, a, "f64",
                {"workload": "gen_loop_overhead", "grid": [64, 64], "loop": "generated 2D fill, 1000 per step"},
                us_per_loop=round(us, 2), kernel_us_by_events=round(kern_us, 2) if kern_us else None)


# ------------------------------------------------------ stream-mix ceilings
# (reads, writes) of this path's kernels: the practical HBM ceiling each is
# measured against (pointwise, 16-byte lanes, no arithmetic or neighbours)
STREAM_MIXES = {"copy 1R1W": (1, 1), "read-only 4R": (4, 0), "write-only 1W": (0, 1),
                "calc_ur_r / axpy 2R1W": (2, 1), "jacobi_chain 1R2W": (1, 2), "cg pair 3R2W": (3, 2),
                "calc_w column 4R1W": (4, 1), "cg chain 6R3W": (6, 3), "thomas fwd (_Inc) 4R2W": (4, 2),
                "thomas fwd 5R1W": (5, 1)}


def measure_stream_mixes(ops, n, reps=10):
    """Each mix over n^3-point FP64 arrays, plain and non-temporal loads and
    stores: GB/s of the bytes it moves (decimal)."""
    import ctypes as C
    L = ops.lib()
    L.ops_mi_stream_probe.restype = C.c_int
    L.ops_mi_stream_probe.argtypes = [C.c_int, C.c_int, C.c_int, C.c_long, C.c_int, C.POINTER(C.c_double)]
    pts = n ** 3
    out = {}
    for name, (r, w) in STREAM_MIXES.items():
        row = {"reads": r, "writes": w}
        for nt in (0, 1):
            us = C.c_double()
            if L.ops_mi_stream_probe(r, w, nt, pts, reps, C.byref(us)) != 0:
                ops._check()
            gbs = (r + w) * 8.0 * pts / (us.value * 1e-6) / 1e9
            row["nt" if nt else "plain"] = {"avg_us": round(us.value, 2), "GB_s": round(gbs, 1),
                                            "frac": round(gbs / HBM_PEAK_GBS, 4)}
        out[name] = row
        print(f"stream {name}: plain {row['plain']['GB_s']} GB/s, nt {row['nt']['GB_s']} GB/s", file=sys.stderr,
              flush=True)
    return out


def bench_stream(a, ops, apps):
    n = a.size or 512
    ops.ops_init(diags=1)
    m = measure_stream_mixes(ops, n)
    best = max(v["nt"]["GB_s"] for v in m.values())
    chain = m["cg chain 6R3W"]["nt"]
    return line(best, "GB/s", chain["avg_us"] * 1e-3, a, "f64",
                {"workload": f"stream_mix_{n}", "points_per_array": n ** 3,
                 "kernel": "pointwise, 2 doubles per lane (16-byte), XCD-aware order"},
                mixes=m, roofline={"kernel": "stream cg chain 6R3W (nt)", "bound": "hbm", "achieved": chain["GB_s"],
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": chain["frac"], "traffic": None},
                data="zero-filled arrays")


WORKLOADS = {"heat3d_cg_512": bench_heat, "poisson_mixed_768": bench_mixed, "laplace2d_2048": bench_laplace,
             "adi_trid_256": bench_trid, "serve_stream": bench_serve, "gen_cg_tail_512": bench_gen_tail,
             "gen_heat_step_512": bench_gen_heat, "gen_loop_overhead": bench_gen_overhead,
             "stream_mix": bench_stream}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=0, help="grid side (default: the workload's)")
    ap.add_argument("--seed", type=int, default=20240611)
    ap.add_argument("--workload", default="heat3d_cg_512", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="heat3d_cg_512: skip the other configurations")
    ap.add_argument("--requests", type=int, default=100, help="serve_stream: number of requests")
    ap.add_argument("--frac2d", type=float, default=0.6, help="serve_stream: share of 2D Poisson requests")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a 1 GiB copy (the calibration kernel of PMC byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: the backend runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from ops_amd import apps, ops

    if a.calibrate:
        # a copy of known size (1 GiB read, 1 GiB written; one untimed and one
        # timed launch of k_stream<1, 1, false, false>) that scales the PMC
        # FETCH_SIZE / WRITE_SIZE counters of a profiled run
        ops.ops_init(diags=2)  # the level the workloads ask for (a second ops_init is a no-op)
        import ctypes as C
        us = C.c_double()
        ops.lib().ops_mi_stream_probe(1, 1, 0, 1 << 27, 1, C.byref(us))
        ops._check()
    out = WORKLOADS[a.workload](a, ops, apps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
