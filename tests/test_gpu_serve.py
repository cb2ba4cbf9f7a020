"""Batched multi-block CG engine with continuous batching.

Each solve is checked against the oracle CG on the same inputs (iteration
count within one, solution within 1e-9 of the largest value: reduction
orders differ, and the heat conductivity is generated with device exp/log/
cos, which may differ from the host's in the last bit).  A solve's result
must be bit-identical whether it ran alone or in a mix, with or without the
HIP graph; admission must respect the HBM budget."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

SEED = 20240611


@pytest.fixture
def serve(gpu_ops):
    from ops_amd import serve as S
    return S


def _run_alone(S, kind, n, seed, **kw):
    srv = S.Server(budget_bytes=1 << 30, max_slots=4, **kw)
    i = srv.submit(kind, n, seed=seed)
    out = srv.run()
    srv.close()
    return out[i]


def test_poisson_single_matches_oracle(serve):
    n = 80
    done, x = _run_alone(serve, serve.POISSON2D, n, SEED)
    f = O.poisson_rhs(SEED, 2, n, n)
    ref = O.poisson_cg(2, n, n, 1, f, max_iters=100000, tol=1e-10, nthreads=8)
    assert done["status"] == 1
    assert abs(done["iters"] - ref["iters"]) <= 1
    assert abs(done["rr0"] - ref["rrn"][0]) <= 1e-12 * ref["rrn"][0]
    assert np.sqrt(done["rr"]) <= 1e-10 * np.sqrt(done["rr0"])
    assert np.max(np.abs(x - ref["x"])) <= 1e-9 * np.max(np.abs(ref["x"]))


def test_heat_single_matches_oracle(serve):
    n = 21
    done, u = _run_alone(serve, serve.HEAT3D, n, SEED)
    K, rhs = O.heat_inputs(SEED, 3, n, n, n)
    ref = O.heat_cg(3, n, n, n, K, rhs, max_iters=100000, tol=1e-10, nthreads=8)
    assert done["status"] == 1
    assert abs(done["iters"] - ref["iters"]) <= 1
    assert abs(done["rr0"] - ref["rrn"][0]) <= 1e-10 * ref["rrn"][0]
    assert np.max(np.abs(u - ref["u"])) <= 1e-9 * np.max(np.abs(ref["u"]))


MIX = [(0, 40, 1), (1, 17, 2), (0, 130, 3), (1, 24, 4), (0, 70, 5), (0, 257, 6), (1, 33, 7), (0, 64, 8),
       (1, 20, 9), (0, 97, 10)]


def test_mix_equals_alone_bit_for_bit(serve):
    S = serve
    srv = S.Server(budget_bytes=1 << 30, max_slots=16, iters_per_launch=8)
    ids = [srv.submit(k, n, seed=s) for k, n, s in MIX]
    out = srv.run()
    srv.close()
    assert sorted(out) == sorted(ids)
    for i, (k, n, s) in zip(ids, MIX):
        done, x = out[i]
        assert done["status"] == 1
        d1, x1 = _run_alone(S, k, n, s, iters_per_launch=3)
        assert d1["iters"] == done["iters"] and d1["rr"] == done["rr"]
        assert np.array_equal(x1, x), (k, n, s)


def test_graph_and_direct_launches_agree(serve):
    S = serve
    res = []
    for g in (True, False):
        srv = S.Server(budget_bytes=1 << 28, max_slots=8, iters_per_launch=5, use_graph=g)
        ids = [srv.submit(k, n, seed=s) for k, n, s in MIX[:4]]
        out = srv.run()
        srv.close()
        res.append([out[i] for i in ids])
    for (d0, x0), (d1, x1) in zip(*res):
        assert d0["iters"] == d1["iters"]
        assert np.array_equal(x0, x1)


def test_admission_respects_budget(serve):
    S = serve
    reqs = [(0, 257, 11), (1, 33, 12), (0, 200, 13), (1, 30, 14), (0, 150, 15), (0, 230, 16)]
    need = [S.request_bytes(k, n) for k, n, _ in reqs]
    budget = max(need) + min(need) + (1 << 20)
    srv = S.Server(budget_bytes=budget, max_slots=8, iters_per_launch=16)
    ids = [srv.submit(k, n, seed=s) for k, n, s in reqs]
    seen_queue, results, peak = False, {}, 0
    while True:
        left = srv.step()
        peak = max(peak, srv.bytes_in_use())
        assert srv.bytes_in_use() <= budget
        seen_queue |= srv.queued() > 0 and srv.active() > 0
        for c in srv.completed():
            results[c["id"]] = c
            srv.release(c["id"])
        if left == 0:
            break
    srv.close()
    assert seen_queue  # some solves waited while others ran
    assert sorted(results) == sorted(ids)
    for c in results.values():
        assert c["status"] == 1 and np.sqrt(c["rr"]) <= 1e-10 * np.sqrt(c["rr0"])


@pytest.mark.parametrize("n", [60, 150])
def test_mixed_poisson_matches_oracle(serve, n):
    # FP32-inner refinement in the batched engine: the oracle's refinement
    # loop (orc_poisson_mixed), iteration counts and solution
    S = serve
    srv = S.Server(budget_bytes=1 << 28, max_slots=4, iters_per_launch=8)
    i = srv.submit(S.POISSON2D, n, seed=SEED, mixed=True)
    done, x = srv.run()[i]
    srv.close()
    f = O.poisson_rhs(SEED, 2, n, n)
    # the engine's inner FP32 solves run to 1e-6 of their initial residual
    ref = O.poisson_mixed(2, n, n, 1, f, tol=1e-10, inner_tol=1e-6, nthreads=8)
    assert done["status"] == 1
    assert done["outer"] == ref["outer"] and done["iters"] == ref["inner_total"]
    assert np.max(np.abs(x - ref["x"])) <= 1e-12 * np.max(np.abs(ref["x"]))
    assert np.sqrt(done["rr"]) <= 1e-10 * np.sqrt(done["rr0"])


@pytest.mark.parametrize("n", [24, 45])
def test_mixed_heat_matches_oracle(serve, n):
    # FP32-inner refinement of the 3D heat problem in the batched engine
    # against the oracle's (orc_heat_mixed): same corrections, inner
    # iterations within a few (the device generates the conductivity with
    # its own exp/log/cos, last-bit differences from the host's), solution
    # within 1e-9 of the largest value, and the FP64 residual of the GPU's
    # solution, recomputed on the host, at the target
    S = serve
    srv = S.Server(budget_bytes=1 << 28, max_slots=4, iters_per_launch=8)
    i = srv.submit(S.HEAT3D, n, seed=SEED, mixed=True)
    done, x = srv.run()[i]
    srv.close()
    K, rhs = O.heat_inputs(SEED, 3, n, n, n)
    ref = O.heat_mixed(n, n, n, K, rhs, tol=1e-10, inner_tol=1e-6, nthreads=8)
    assert done["status"] == 1
    assert done["outer"] == ref["outer"]
    assert abs(done["iters"] - ref["inner_total"]) <= 2 * ref["outer"]
    assert np.max(np.abs(x - ref["x"])) <= 1e-9 * np.max(np.abs(ref["x"]))
    Kx, Ky, Kz = O.heat_coefs(3, n, n, n, K)
    x_ext = np.zeros((n + 2, n + 2, n + 2))
    x_ext[1:-1, 1:-1, 1:-1] = x
    u0 = np.zeros_like(x_ext)
    u0[1:-1, 1:-1, 1:-1] = rhs
    ax, _ = O.heat_apply(3, n, n, n, Kx, Ky, Kz, x_ext)
    a0, _ = O.heat_apply(3, n, n, n, Kx, Ky, Kz, u0)
    r = rhs - ax[1:-1, 1:-1, 1:-1]
    r0 = rhs - a0[1:-1, 1:-1, 1:-1]
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(r0)
    # the engine's own FP64 residual at the end
    assert np.sqrt(done["rr"]) <= 1e-10 * np.sqrt(done["rr0"])


def test_mixed_heat_takes_fp64_solution(serve):
    # FP64 CG and FP32-inner refinement of the same heat problem reach the
    # same solution (both at 1e-10 relative residual)
    S = serve
    srv = S.Server(budget_bytes=1 << 28, max_slots=4, iters_per_launch=8)
    a = srv.submit(S.HEAT3D, 30, seed=SEED + 1)
    b = srv.submit(S.HEAT3D, 30, seed=SEED + 1, mixed=True)
    out = srv.run()
    srv.close()
    x64, xm = out[a][1], out[b][1]
    assert out[a][0]["status"] == 1 and out[b][0]["status"] == 1
    assert np.max(np.abs(xm - x64)) <= 1e-8 * np.max(np.abs(x64))


def test_mixed_and_fp64_in_one_batch(serve):
    S = serve
    reqs = [(0, 70, 31, True), (1, 21, 32, False), (0, 130, 33, True), (0, 90, 34, False), (1, 26, 35, True),
            (1, 19, 36, True)]
    srv = S.Server(budget_bytes=1 << 28, max_slots=8, iters_per_launch=8)
    ids = [srv.submit(k, n, seed=s, mixed=m) for k, n, s, m in reqs]
    out = srv.run()
    srv.close()
    for i, (k, n, s, m) in zip(ids, reqs):
        a = S.Server(budget_bytes=1 << 28, max_slots=2, iters_per_launch=5)
        j = a.submit(k, n, seed=s, mixed=m)
        alone = a.run()[j]
        a.close()
        assert alone[0]["iters"] == out[i][0]["iters"] and alone[0]["outer"] == out[i][0]["outer"]
        assert np.array_equal(alone[1], out[i][1]), (k, n, s, m)


def test_autotune_picks_candidates_and_solves_stay_exact(serve):
    S = serve
    srv = S.Server(budget_bytes=1 << 30, max_slots=8, iters_per_launch=8)
    srv.autotune(iters=4)
    for n in (40, 100, 200):
        assert srv.tile_param(S.HEAT3D, n) in (4, 8, 16, 32)
    for n in (100, 600, 1500):
        assert srv.tile_param(S.POISSON2D, n) in (1, 2, 4, 8)
    assert srv.bytes_in_use() == 0 and srv.active() == 0
    ids = [srv.submit(k, n, seed=s) for k, n, s in MIX[:5]]
    out = srv.run()
    # the same server, one solve at a time: bit-identical
    for i, (k, n, s) in zip(ids, MIX[:5]):
        j = srv.submit(k, n, seed=s)
        alone = srv.run()[j]
        assert alone[0]["iters"] == out[i][0]["iters"]
        assert np.array_equal(alone[1], out[i][1])
    srv.close()


def test_max_iters_and_errors(serve):
    S = serve
    from ops_amd import ops
    srv = S.Server(budget_bytes=1 << 26, max_slots=2, iters_per_launch=4)
    i = srv.submit(S.POISSON2D, 300, max_iters=6)
    with pytest.raises(ops.OPSError):
        srv.release(i)  # not completed yet
    out = srv.run()
    assert out[i][0]["status"] == 2 and out[i][0]["iters"] == 6
    with pytest.raises(ops.OPSError):
        srv.submit(7, 10)
    with pytest.raises(ops.OPSError):
        srv.submit(S.HEAT3D, 10, mixed=2)  # mixed is 0 or 1
    with pytest.raises(ops.OPSError):
        srv.submit(S.HEAT3D, 600)  # larger than the budget
    with pytest.raises(ops.OPSError):
        srv.release(12345)
    srv.close()


def test_request_stream_driver(gpu_ops):
    from ops_amd import apps
    kinds = [0, 1, 0, 0, 1, 0]
    sizes = [60, 20, 33, 100, 26, 48]
    seeds = [21, 22, 23, 24, 25, 26]
    arrival = [0.0, 0.0, 0.01, 0.02, 0.05, 0.05]
    span, done, iters = apps.serve_stream(kinds, sizes, seeds, arrival, budget=1 << 28, max_slots=4,
                                          iters_per_launch=4)
    assert span > 0 and np.all(iters > 0)
    assert np.all(done >= np.asarray(arrival)) and np.max(done) <= span
    # same iteration counts as the engine gives each solve on its own
    from ops_amd import serve as S
    for k, n, s, it in zip(kinds, sizes, seeds, iters):
        d, _ = _run_alone(S, k, n, s)
        assert d["iters"] == it
