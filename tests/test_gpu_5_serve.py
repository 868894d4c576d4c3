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
