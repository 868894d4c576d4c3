"""The benchmark configurations at their stated sizes, on the GPU.

  out of core : n = 196608, NB = 1024, PDMATGEN('S','D',100) in pageable
                host packed tiles (155.4 GB), 64 GiB device budget
  mixed prec. : n = 65536, NB = 512, A = G G^T / n + mu I, 16 right-hand
                sides, mu in {1e-2, 1e-6}, FP32 and BF16 trailing updates
  batched     : 10000 systems, n log-uniform in [32, 1024], 1-8 right-hand
                sides, FP64 and FP32 + refinement
  served      : Poisson arrivals, n log-uniform in [64, 32768], 1-64
                right-hand sides, 30 % matrix reuse, mu in {1e-1, 1e-4}

Size-independent properties are checked where the oracle cannot run at
these sizes: the LIN scaled residual of the out-of-core factor (formed on
the GPU against the regenerated matrix) and agreement with the
HBM-resident factor; backward errors recomputed on the host from the
returned solutions; each served result the same alone and in the mix.
"""
import threading
import time

import numpy as np
import pytest

from test_gpu_2_cholesky import expected_traffic

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53
TOL = 8 * EPS
GB = 1 << 30


@pytest.mark.timeout(900)
def test_out_of_core_n196608_budget64g(gpu):
    ooc = gpu
    n, nb, mem = 196608, 1024, 64 << 30
    plan = ooc.panel_plan(n, nb, mem)
    assert len(plan) >= 3
    elems = ooc.tiles_elems(n, nb)
    tile = nb * nb
    ntile = elems // tile
    # the HBM-resident factor of the same matrix: L(N-1, N-1) and sampled tiles
    buf = ooc.DeviceBuffer(elems * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    assert ooc.potrf_tiles(buf, n, nb) == 0
    picks = [0, 1, 191, 192, 5000, 12345, ntile - 2, ntile - 1]
    hbm = {}
    for k in picks:
        out = np.empty(tile)
        ooc.lib().oocdag_memcpy_d2h(out.ctypes.data, buf.ptr + k * tile * 8, tile * 8)
        hbm[k] = out
    t = ooc.HostTiles(n, nb, pinned=False)
    try:
        t.generate()
        secs, info = ooc.cholesky_host_tiles(t, mem)
        assert info == 0 and secs > 0
        st = ooc.last_stats()
        h2d, d2h = expected_traffic(n, nb, plan)
        assert (st["h2d_bytes"], st["d2h_bytes"]) == (h2d, d2h)
        arr = t.array()
        print(f"n={n} out of core {secs:.2f} s ({n ** 3 / 3 / secs / 1e12:.2f} TFLOP/s), "
              f"panels {plan}, L(N-1,N-1) {arr[-1]!r} vs HBM {hbm[ntile - 1][-1]!r}")
        assert arr[-1] == pytest.approx(hbm[ntile - 1][-1], rel=1e-12)
        for k in picks:
            got = arr[k * tile:(k + 1) * tile]
            assert np.max(np.abs(got - hbm[k])) <= 1e-12 * max(1.0, np.max(np.abs(hbm[k]))), k
        # LIN scaled residual of the out-of-core factor, on the GPU
        assert ooc.lib().oocdag_memcpy_h2d(buf.ptr, t.ptr, elems * 8) == 0
    finally:
        t.free()
    r = ooc.llt_resid_tiles(buf, n, nb)
    buf.free()
    print(f"scaled residual of the out-of-core factor {r:.3e}")
    assert r <= 10.0


def dense_symmetric(ooc, buf, n, nb):
    """A (lower tiles, full diagonal tiles) as a dense symmetric host array,
    mirrored block by block in place."""
    A = np.zeros((n, n), order="F")
    ooc.tiles_to_host(buf, A, nb)
    step = 4096
    for j0 in range(0, n, step):
        j1 = min(n, j0 + step)
        A[:j0, j0:j1] = A[j0:j1, :j0].T
        blk = A[j0:j1, j0:j1]
        A[j0:j1, j0:j1] = np.tril(blk) + np.tril(blk, -1).T
    return A


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mu", [1e-2, 1e-6])
def test_mixed_precision_n65536(gpu, mu):
    ooc = gpu
    n, nb, nrhs = 65536, 512, 16
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, mu, seed=1)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=1, stream_id=2)
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    sols, stats = {}, {}
    for prec in ("fp64", "fp32", "bf16"):
        st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=TOL, max_iter=30)
        stats[prec] = st
        sols[prec] = dX.to_array((n, nrhs))
        print(f"mu={mu} {prec}: {st['total_seconds']:.3f} s factor {st['factor_seconds']:.3f} s "
              f"it {st['iterations']} eta {st['backward_error']:.2e} lres {st['solve_resid']:.2e} "
              f"used {st['precision_used']} fell_back {st['fell_back']}")
    B = dB.to_array((n, nrhs))
    A = dense_symmetric(ooc, buf, n, nb)
    buf.free()
    anorm = np.abs(A).sum(1).max()
    kappa = (4.0 + mu) / mu   # eigenvalues of G G^T / n lie in (0, ~4)
    for prec, X in sols.items():
        st = stats[prec]
        assert st["backward_error"] <= TOL, (prec, st)
        R = B - A @ X
        xn = np.abs(X).max(0)
        rn = np.abs(R).max(0)
        eta = np.max(rn / (anorm * xn + np.abs(B).max(0)))
        lres = np.max(rn / (anorm * xn * EPS * n))
        assert eta <= 2 * TOL, (prec, eta)
        assert lres <= 10.0 and st["solve_resid"] <= 10.0, (prec, lres, st["solve_resid"])
        ferr = np.abs(X - sols["fp64"]).max() / np.abs(sols["fp64"]).max()
        assert ferr <= 1e3 * kappa * EPS, (prec, ferr)
    if mu == 1e-2:
        assert stats["fp32"]["fell_back"] == 0 and stats["bf16"]["fell_back"] == 0
    else:
        # kappa ~ 4e6: kappa * 2^-8 >> 1, a BF16 factor cannot refine to FP64;
        # it stops at its first bad pivot and the FP32 attempt follows, so
        # asking for BF16 is still faster than the plain FP64 path
        assert stats["fp32"]["fell_back"] == 0
        assert stats["bf16"]["fell_back"] == 1 and stats["bf16"]["precision_used"] == 1
        assert stats["bf16"]["total_seconds"] < stats["fp64"]["total_seconds"]


@pytest.mark.timeout(600)
def test_batched_10000_systems(gpu):
    ooc = gpu
    rng = np.random.default_rng(2024)
    n = np.exp(rng.uniform(np.log(32), np.log(1024), size=10000)).astype(np.int32)
    r = rng.integers(1, 9, size=10000).astype(np.int32)
    bt = ooc.Batch(n, r)
    bt.generate(mu=0.1, seed=2024)
    picks = np.random.default_rng(1).choice(10000, size=50, replace=False)
    for prec in ("fp64", "fp32"):
        st, berr, iters, status, lres = bt.solve(prec, tol=TOL, max_iter=10, reps=1,
                                                 with_resid=True)
        print(prec, {k: st[k] for k in ("seconds", "systems_per_second", "max_backward_error",
                                        "max_solve_resid", "variant_T")})
        assert st["failed"] == 0 and (status == 0).all()
        assert berr.max() <= TOL and lres.max() <= 10.0
        for s in picks:
            A, B, X = bt.host_system(int(s))
            Xref = np.linalg.solve(A, B)
            assert np.abs(X - Xref).max() / np.abs(Xref).max() < 1e-12, s
            R = B - A @ X
            eta = np.max(np.abs(R).max(0) / (np.abs(A).sum(1).max() * np.abs(X).max(0) +
                                             np.abs(B).max(0)))
            assert eta <= 2 * TOL, s


def poisson_workload(count, seed=7):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        if reqs and rng.random() < 0.3:
            src = int(rng.integers(0, len(reqs)))
            q = dict(reqs[src])
            q.update(nrhs=int(rng.integers(1, 65)), rhs_seed=10_000 + i)
        else:
            nn = int(np.exp(rng.uniform(np.log(64), np.log(32768))))
            q = dict(matrix_id=i, n=nn, nrhs=int(rng.integers(1, 65)),
                     mu=float(rng.choice([1e-1, 1e-4])), seed=1000 + i, rhs_seed=10_000 + i)
        reqs.append(q)
    return reqs


@pytest.mark.timeout(600)
def test_served_poisson_stream(gpu):
    """40 requests arriving as a Poisson stream at 4 requests/s into a
    3-slot server with a 64 GiB budget: every result at FP64 backward error,
    and sampled results (a fresh matrix, a cache hit, batched ones) equal to
    the same request served alone."""
    ooc = gpu
    reqs = poisson_workload(40)
    arrivals = np.cumsum(np.random.default_rng(11).exponential(1.0 / 4.0, size=len(reqs)))
    srv = ooc.Server(64 * GB, slots=3)
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def client():
        for i, q in enumerate(reqs):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(precision="fp32", tol=TOL, **q)

    th = threading.Thread(target=client)
    th.start()
    out = []
    for i in range(len(reqs)):
        while tickets[i] is None:
            time.sleep(0.001)
        out.append(srv.wait(tickets[i], want_x=True))
    th.join()
    srv.close()
    lat = np.array([o[0]["latency_seconds"] for o in out])
    print(f"p50 {np.percentile(lat, 50) * 1e3:.1f} ms p99 {np.percentile(lat, 99) * 1e3:.1f} ms, "
          f"routes {[o[0]['route'] for o in out].count(0)} batched, "
          f"hits {sum(o[0]['cache_hit'] for o in out)}")
    for (res, x), q in zip(out, reqs):
        assert res["status"] == 0, (q, res)
        assert res["backward_error"] <= TOL and res["solve_resid"] <= 10.0, (q, res)
    assert any(o[0]["cache_hit"] for o in out)
    # alone: a fresh matrix, a reuse (after its source), two small ones
    hit = next(i for i, o in enumerate(out) if o[0]["cache_hit"] and reqs[i]["n"] <= 12000)
    src = next(i for i in range(hit) if reqs[i]["matrix_id"] == reqs[hit]["matrix_id"])
    fresh = next(i for i, q in enumerate(reqs) if 2000 < q["n"] <= 12000 and i not in (hit, src))
    small = [i for i, q in enumerate(reqs) if q["n"] <= 1024][:2]
    for i in [fresh, hit] + small:
        one = ooc.Server(64 * GB, slots=1)
        seq = [reqs[src], reqs[i]] if i == hit else [reqs[i]]
        ts = [one.submit(precision="fp32", tol=TOL, **q) for q in seq]
        res1, x1 = [one.wait(t, want_x=True) for t in ts][-1]
        one.close()
        x = out[i][1]
        assert np.abs(x - x1).max() / np.abs(x1).max() <= 1e-12, (i, reqs[i])
