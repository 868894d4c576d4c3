"""Continuous-batching solve server (an addition with no counterpart in the reference).
Every request must reach the requested FP64-level backward error, and each
result must be the same alone and in the mix: |x_mix - x_alone| / |x_alone|
<= 1e-12 (max norm)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 8 * 2.0 ** -53
GB = 1 << 30


def workload():
    rng = np.random.default_rng(17)
    reqs = []
    for i in range(14):
        n = int(np.exp(rng.uniform(np.log(40), np.log(1024))))
        reqs.append(dict(matrix_id=i, n=n, nrhs=int(rng.integers(1, 9)), mu=0.1, seed=100 + i,
                         rhs_seed=500 + i))
    for i in range(14, 18):
        n = int(rng.integers(1500, 4500))
        reqs.append(dict(matrix_id=i, n=n, nrhs=int(rng.integers(1, 17)),
                         mu=[1e-1, 1e-4][i % 2], seed=100 + i, rhs_seed=500 + i))
    # reuse of an earlier large matrix with new right-hand sides (cache hits)
    for k, src in enumerate((14, 15)):
        r = dict(reqs[src])
        r.update(rhs_seed=900 + k, nrhs=5)
        reqs.append(r)
    # a small matrix with more right-hand sides than the batched kernel holds
    reqs.append(dict(matrix_id=40, n=300, nrhs=24, mu=0.1, seed=140, rhs_seed=540))
    return reqs


def run(srv, reqs, prec="fp32"):
    tickets = [srv.submit(precision=prec, tol=TOL, **r) for r in reqs]
    return [srv.wait(t, want_x=True) for t in tickets]


def test_server_mix_equals_alone(gpu):
    ooc = gpu
    reqs = workload()
    srv = ooc.Server(64 * GB, slots=3)
    mix = run(srv, reqs)
    stats = srv.stats()
    srv.close()
    routes = [m[0]["route"] for m in mix]
    assert routes.count(0) >= 10 and routes.count(1) >= 6
    assert max(m[0]["batch_size"] for m in mix) > 1          # continuous batching happened
    assert any(m[0]["cache_hit"] for m in mix)                # cached factor reused
    for (res, x), r in zip(mix, reqs):
        assert res["status"] == 0, (r, res)
        assert res["backward_error"] <= TOL, (r, res)
    # each request alone, in a fresh one-slot server
    for i in (0, 5, 14, 15, 18, 20):
        one = ooc.Server(64 * GB, slots=1)
        (res1, x1), = run(one, [reqs[i]])
        one.close()
        x = mix[i][1]
        assert np.abs(x - x1).max() / np.abs(x1).max() <= 1e-12, (i, reqs[i])
    print(stats)


def test_server_routes_over_budget_to_out_of_core(gpu):
    ooc = gpu
    r = dict(matrix_id=1, n=3000, nrhs=3, mu=0.1, seed=77, rhs_seed=78)
    small_budget = ooc.Server(96 << 20, slots=1)       # 96 MiB: the request does not fit
    (res, x), = run(small_budget, [r], prec="fp64")
    small_budget.close()
    assert res["route"] == 2 and res["status"] == 0
    assert res["backward_error"] <= 1e-15
    big = ooc.Server(16 * GB, slots=1)
    (res2, x2), = run(big, [r], prec="fp64")
    big.close()
    assert res2["route"] == 1
    assert np.abs(x - x2).max() / np.abs(x2).max() <= 1e-10


def gram_dense(ooc, n, mu, seed):
    nb = 256 if n < 4096 else 512
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, mu, seed=seed)
    A = np.zeros((n, n), order="F")
    ooc.tiles_to_host(buf, A, nb)
    buf.free()
    return np.tril(A) + np.tril(A, -1).T


def rhs(ooc, n, nrhs, seed):
    d = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(d, n, nrhs, n, seed=seed, stream_id=2)
    return d.to_array((n, nrhs))


def test_server_out_of_core_route_inside_its_reservation(gpu):
    """n = 6000 (NB = 512): the packed matrix (163 MB) alone exceeds a
    96 MiB budget.  The route generates A tile column by tile column, factors
    it out of core inside its reservation, solves and checks it with streamed
    tile columns: the solution matches numpy's on the same matrix, the
    backward error is FP64-level, the PDLASCHK residual <= 10."""
    ooc = gpu
    r = dict(matrix_id=3, n=6000, nrhs=4, mu=0.1, seed=91, rhs_seed=92)
    est = ooc.Server.estimate(**r)
    budget = 96 << 20
    assert est["miss"] > budget and est["ooc_min"] <= budget
    srv = ooc.Server(budget, slots=2)
    (res, x), = run(srv, [r], prec="fp64")
    st = srv.stats()
    srv.close()
    assert res["route"] == 2 and res["status"] == 0, res
    assert res["backward_error"] <= 4e-16 and res["solve_resid"] <= 10.0, res
    assert st["used_bytes"] == 0
    A = gram_dense(ooc, r["n"], r["mu"], r["seed"])
    B = rhs(ooc, r["n"], r["nrhs"], r["rhs_seed"])
    xref = np.linalg.solve(A, B)
    assert np.abs(x - xref).max() / np.abs(xref).max() <= 1e-10


def test_server_cache_hit_is_not_evicted_by_its_own_admission(gpu):
    """Two slots, a budget that holds matrix 1's cached factor plus a miss
    for matrix 2 but not also a hit on matrix 1: the hit, admitted while the
    miss holds its reservation, must neither evict nor read a freed factor.
    It waits for room and returns the same x as when run alone."""
    ooc = gpu
    r1 = dict(matrix_id=1, n=5000, nrhs=2, mu=0.1, seed=11, rhs_seed=12)
    r2 = dict(matrix_id=2, n=5000, nrhs=2, mu=0.1, seed=21, rhs_seed=22)
    r3 = dict(r1, rhs_seed=13)
    e1, e2, e3 = (ooc.Server.estimate(**r) for r in (r1, r2, r3))
    budget = e1["cached"] + e2["miss"] + e3["hit"] // 2
    assert e1["miss"] <= budget
    srv = ooc.Server(budget, slots=2)
    run(srv, [r1])
    assert srv.stats()["used_bytes"] == e1["cached"]
    (res2, _), (res3, x3) = run(srv, [r2, r3])
    srv.close()
    assert res2["status"] == 0 and res3["status"] == 0
    assert res3["cache_hit"] == 1 and res3["route"] == 1
    alone = ooc.Server(64 * GB, slots=1)
    run(alone, [r1])
    (res3a, x3a), = run(alone, [r3])
    alone.close()
    assert res3a["cache_hit"] == 1
    assert np.abs(x3 - x3a).max() / np.abs(x3a).max() <= 1e-12


def test_server_fails_requests_that_never_fit(gpu):
    """A request above the whole budget on every route completes with
    OOCDAG_EBUDGET instead of blocking the queue; later requests still run
    and the server shuts down."""
    ooc = gpu
    srv = ooc.Server(2 << 20, slots=1)   # 2 MiB
    t_big = srv.submit(precision="fp64", tol=TOL, matrix_id=1, n=1000, nrhs=8, mu=0.1, seed=1,
                       rhs_seed=2)
    t_ok = srv.submit(precision="fp64", tol=TOL, matrix_id=2, n=60, nrhs=1, mu=0.1, seed=3,
                      rhs_seed=4)
    (rb, _), (rk, _) = srv.wait(t_big), srv.wait(t_ok)
    srv.close()
    assert rb["status"] == ooc.EBUDGET
    assert rk["status"] == 0 and rk["backward_error"] <= TOL


def test_server_results_carry_solve_resid(gpu):
    ooc = gpu
    reqs = [dict(matrix_id=0, n=300, nrhs=3, mu=0.1, seed=5, rhs_seed=6),
            dict(matrix_id=1, n=2500, nrhs=2, mu=1e-4, seed=7, rhs_seed=8)]
    srv = ooc.Server(16 * GB, slots=2)
    out = run(srv, reqs)
    srv.close()
    for (res, x), r in zip(out, reqs):
        assert res["status"] == 0 and 0 < res["solve_resid"] <= 10.0, res


def test_server_stalled_hit_refactors_without_touching_the_shared_factor(gpu):
    """Cache hits whose refinement on the cached BF16 factor stalls (a target
    no factor reaches) refactor in FP64 into a factor of their own: two such
    hits run at once on two slots, each solves with its own FP64 factor, the
    shared cached factor is only replaced when one request is its sole user,
    and the memory reserved for the extra factor is returned."""
    ooc = gpu
    r1 = dict(matrix_id=7, n=8192, nrhs=2, mu=1e-2, seed=31, rhs_seed=32)
    srv = ooc.Server(64 * GB, slots=2)
    (res1, x1), = run(srv, [r1], prec="bf16")
    assert res1["status"] == 0 and res1["precision_used"] == 2 and res1["cache_hit"] == 0
    used_cached = srv.stats()["used_bytes"]
    hits = [dict(r1, rhs_seed=40 + k) for k in range(2)]
    tickets = [srv.submit(precision="bf16", tol=1e-300, **h) for h in hits]
    out = [srv.wait(t, want_x=True) for t in tickets]
    for res, x in out:
        assert res["status"] == 0 and res["cache_hit"] == 1, res
        assert res["fell_back"] == 1 and res["precision_used"] == 0, res
        assert res["backward_error"] <= TOL, res   # FP64 reaches 8 eps, not 1e-300
    assert srv.stats()["used_bytes"] == used_cached
    # the entry still serves: a later hit at the ordinary target
    (res3, x3), = run(srv, [dict(r1, rhs_seed=32)], prec="bf16")
    srv.close()
    assert res3["status"] == 0 and res3["cache_hit"] == 1 and res3["backward_error"] <= TOL
    assert np.abs(x3 - x1).max() / np.abs(x1).max() <= 1e-10
