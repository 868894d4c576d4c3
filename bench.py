#!/usr/bin/env python3
"""Benchmark: FP64 out-of-core Cholesky TFLOP/s (n^3/3 flops) on one MI355X.

Default workload (the reference's own path, Cholesky() oocdag.cpp:18-112, at
n = 196608, NB = 1024): the PDMATGEN('S','D', ISEED=100) matrix
(oocdag.cpp:298-302) held in host memory as packed lower tiles (155.4 GB),
the device limited to a 64 GiB budget -- below the packed triangle, so the
factorization streams column panels over PCIe with double-buffered copies.
One step = one out-of-core factorization, timed PCIe-inclusive from the call
to its return (the reference times only Cholesky(), oocdag.cpp:51,90); the
matrix is regenerated into the host tiles between steps, outside the clock.

--steps and --warmup are upper bounds under a wall-clock budget (--budget-s,
from process start): a step starts only when the time it is expected to take
still fits; the JSON line reports the steps and warmups actually run.

Prints one JSON line (metric, value, unit, n_gpus, steps, timing, config ...) plus
  roofline     -- the panel-update GEMM (dominant kernel), timed by HIP events
                  on the stream each launch runs on, algorithmic flops / kernel
                  seconds vs the 78.6 TFLOP/s FP64 matrix peak; traffic from
                  the committed rocprofv3 PMC summary of the same command;
  cpu_baseline -- the CPU restatement of the reference driver (oracle/oracle.c)
                  on the host's cores at n = 4096, NB = 256, a 28 MiB budget
                  (the reference's 4-panel schedule [2, 3, 5, 6]), median of
                  repeated runs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

_T_START = time.monotonic()

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6  # TFLOP/s, MI355X FP64 matrix (dense)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def elapsed() -> float:
    return time.monotonic() - _T_START


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def verify_small_ooc(ooc) -> float:
    """Scaled residual of a small factorization through the timed entry point
    (host tiles, a budget of a few tile columns: several panels, streamed),
    formed on the GPU against the regenerated matrix.  A cheap guard that the
    timed code computes a correct factor; the full-size factor is checked by
    --verify-full and by the GPU tests."""
    n, nb = 12288 - 300, 1024
    t = ooc.HostTiles(n, nb, pinned=False)
    try:
        t.generate()
        _, info = ooc.cholesky_host_tiles(t, 40 * nb * nb * 8)
        if info != 0:
            raise RuntimeError(f"small out-of-core check: info={info}")
        buf = ooc.DeviceBuffer(t.elems * 8)
        ooc.lib().oocdag_memcpy_h2d(buf.ptr, t.ptr, t.elems * 8)
        r = ooc.llt_resid_tiles(buf, n, nb)
        buf.free()
    finally:
        t.free()
    return float(r)


def cpu_baseline(n: int = 4096, nb: int = 256, memsize: int = 28 << 20,
                 min_runs: int = 5, target_s: float = 10.0) -> dict:
    """The reference driver restated on the CPU (oracle/oracle.c: the
    reference's panel planner find_Yn, A2Y/A2X staging into a host buffer of
    memBlock tiles and the same tile-task order), one task per tile on all
    host cores, each tile kernel single-threaded MKL (the reference linked
    mkl_sequential for its host dpotrf, Makefile:17,45) when the MKL runtime
    is present, else the oracle's own loops.  n = 4096, NB = 256 with a
    28 MiB budget: memBlock = 56 tiles, panels [2, 3, 5, 6].  Median of at
    least `min_runs` runs (until about `target_s` seconds), each timed from
    after the buffer allocation to the final join (oocdag.cpp:51,90)."""
    import numpy as np
    import oracle
    oracle.build()
    backend = oracle.set_backend(1)
    threads = oracle.lib().ora_get_threads()
    A0 = oracle.pdmatgen(n, nb)
    plan = oracle.panel_plan(n, nb, memsize)
    times = []
    t_begin = time.perf_counter()
    while len(times) < min_runs or (time.perf_counter() - t_begin < target_s and len(times) < 1001):
        A = A0.copy(order="F")
        t, info = oracle.cholesky(A, nb, memsize)
        if info != 0:
            raise RuntimeError(f"CPU baseline factorization failed, info={info}")
        times.append(t)
    resid = float(oracle.llt_resid(A))
    oracle.set_backend(0)
    med = float(np.median(times))
    blas = "MKL (libmkl_rt, sequential per tile task)" if backend == 1 else \
        "the oracle's own FP64 tile loops (no MKL runtime found)"
    return {"value": round(n ** 3 / 3 / med / 1e12, 5), "unit": "TFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"CPU restatement of the reference driver (oracle/oracle.c) at n={n} "
                      f"nb={nb} memsize={memsize >> 20} MiB -> panels {plan}, median of "
                      f"{len(times)} runs = {med * 1e3:.1f} ms (min {min(times) * 1e3:.1f}, "
                      f"max {max(times) * 1e3:.1f}), scaled residual {resid:.2e}; "
                      f"{threads} threads on {os.cpu_count()} logical CPUs ({cpu_model()}); "
                      f"BLAS: {blas}",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "blas": blas,
            "median_s": med, "runs": len(times), "panels": plan}


def update_algorithmic_bytes(n: int, nb: int) -> int:
    """Average algorithmic HBM bytes per update-GEMM launch of the
    HBM-resident look-ahead factorization (--mode hbm): each launch reads its
    lower trapezoid C once, writes it once and reads the K = nb wide panel
    column once (B is rows of that same column); 16 B per C element + 8 B per
    panel element."""
    bb = -(-n // nb)
    tot = cnt = 0
    for j in range(bb - 1):
        m = (bb - j - 1) * nb
        tot += 16 * (m * nb - nb * (nb - 1) // 2) + 8 * m * nb
        cnt += 1
        if j + 2 < bb:
            m = (bb - j - 2) * nb
            tot += 16 * (m * (m + 1) // 2) + 8 * m * nb
            cnt += 1
    return tot // max(cnt, 1)


def measured_traffic(name: str, key: dict):
    """HBM bytes per update-GEMM launch from a committed PMC summary
    (profiles/<name>: FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    the reading of a copy kernel of known size).  The summary's
    `traffic_bytes_per_launch:` line must match every field of `key`; None
    when no summary covers the workload."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    fields = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
                    if all(fields.get(k) == str(v) for k, v in key.items()):
                        return int(fields["bytes"]), "profiles/" + name
    except (OSError, KeyError, ValueError):
        pass
    return None, None


class Budget:
    """Wall-clock budget of the whole run, counted from process start."""

    def __init__(self, seconds: float, reserve: float):
        self.deadline = seconds
        self.reserve = reserve   # kept for the work after the timed steps

    def fits(self, need: float) -> bool:
        return elapsed() + need + self.reserve <= self.deadline


def bench_ooc(ooc, n: int, nb: int, budget_gib: float, steps: int, warmup: int,
              budget: Budget, verify_full: bool = False) -> dict:
    """The out-of-core FP64 factorization as the reference runs it: the
    matrix in host memory (packed lower tiles, pageable), the device limited
    to a budget below the packed triangle, column panels streamed over PCIe.
    Each step regenerates the matrix (outside the clock), then times one
    oocdag_cholesky_host_tiles call PCIe-inclusive; the update GEMM's HIP
    events give the roofline figures, the copy streams' busy time GB/s."""
    mem = int(budget_gib * (1 << 30))
    plan = ooc.panel_plan(n, nb, mem)
    t0 = time.perf_counter()
    t = ooc.HostTiles(n, nb, pinned=False)
    log(f"[ooc] n={n} nb={nb}: {t.elems * 8 / 1e9:.1f} GB pageable host tiles; budget "
        f"{budget_gib} GiB, panels {plan}; {elapsed():.1f} s since start")
    gen_s, fac_s = [], []
    fresh = False
    ooc.set_progress(32)

    def factor(timed: bool):
        nonlocal fresh
        if not fresh:
            g0 = time.perf_counter()
            t.generate()
            gen_s.append(time.perf_counter() - g0)
            log(f"[ooc] generated in {gen_s[-1]:.1f} s")
        ooc.profile_enable(timed)
        t1 = time.perf_counter()
        secs, info = ooc.cholesky_host_tiles(t, mem)
        t2 = time.perf_counter()
        fresh = False
        prof = ooc.profile_read()
        ooc.profile_enable(False)
        if info != 0:
            raise RuntimeError(f"out-of-core factorization info={info}")
        fac_s.append(t2 - t1)
        return secs, t2 - t1, prof, ooc.last_stats()

    def step_cost():
        return (max(gen_s) if gen_s else 30.0) + (max(fac_s) if fac_s else 60.0)

    warm_done = 0
    for w in range(warmup):
        # the first warmup always runs; later ones only while every timed
        # step still fits after them
        if w >= 1 and not budget.fits((steps + 1) * step_cost()):
            break
        secs, wall, _, _ = factor(False)
        warm_done += 1
        log(f"[ooc] warmup {warm_done}: {secs:.3f} s ({n ** 3 / 3 / secs / 1e12:.2f} TFLOP/s); "
            f"{elapsed():.1f} s since start")
    rows = []
    for k in range(steps):
        if k >= 1 and not budget.fits(step_cost()):
            log(f"[ooc] step {k + 1} would cross the {budget.deadline:.0f} s budget: stopping "
                f"after {k} steps")
            break
        secs, wall, prof, st = factor(True)
        rows.append((secs, wall, prof, st))
        log(f"[ooc] step {k + 1}: {secs:.3f} s, {n ** 3 / 3 / secs / 1e12:.2f} TFLOP/s "
            f"PCIe-inclusive; update GEMM {prof['launches']} launches {prof['seconds']:.3f} s "
            f"{prof['flops'] / max(prof['seconds'], 1e-12) / 1e12:.2f} TFLOP/s; H2D "
            f"{st['h2d_bytes'] / 1e9:.1f} GB D2H {st['d2h_bytes'] / 1e9:.1f} GB; "
            f"{elapsed():.1f} s since start")
    tiles = t.array()
    last = float(tiles[-1])   # L(N-1, N-1) when N is a multiple of nb
    resid = None
    if verify_full:
        t3 = time.perf_counter()
        buf = ooc.DeviceBuffer(t.elems * 8)
        ooc.lib().oocdag_memcpy_h2d(buf.ptr, t.ptr, t.elems * 8)
        resid = ooc.llt_resid_tiles(buf, n, nb)
        buf.free()
        log(f"[ooc] scaled residual of the out-of-core factor {resid:.4e} "
            f"({time.perf_counter() - t3:.1f} s)")
    t.free()
    secs = [r[0] for r in rows]
    h2d_b = sum(r[3]["h2d_bytes"] for r in rows)
    h2d_t = sum(r[3]["h2d_seconds"] for r in rows)
    d2h_b = sum(r[3]["d2h_bytes"] for r in rows)
    d2h_t = sum(r[3]["d2h_seconds"] for r in rows)
    return {"step_seconds": secs, "warmup_done": warm_done, "steps_done": len(rows),
            "generate_seconds": gen_s, "panels": plan, "budget_bytes": mem,
            "h2d_bytes_per_step": h2d_b / max(len(rows), 1),
            "d2h_bytes_per_step": d2h_b / max(len(rows), 1),
            "h2d_GBps": h2d_b / max(h2d_t, 1e-12) / 1e9, "d2h_GBps": d2h_b / max(d2h_t, 1e-12) / 1e9,
            "prof_launches": sum(r[2]["launches"] for r in rows),
            "prof_seconds": sum(r[2]["seconds"] for r in rows),
            "prof_flops": sum(r[2]["flops"] for r in rows),
            "prof_bytes": sum(r[2]["bytes"] for r in rows),
            "L_last_diag": last, "scaled_residual_full": resid}


def bench_mpir(ooc, n: int, nb: int, nrhs: int, mus, precs) -> dict:
    """Mixed-precision + IR workload (n=65536, nb=512, 16 right-hand sides,
    A = G G^T / n + mu I): time to FP64-level backward error per precision,
    speedup over plain FP64, backward error, forward error against the FP64
    solution."""
    EPS = 2.0 ** -53
    tol = 8 * EPS
    out = {}
    A = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    B = ooc.DeviceBuffer(n * nrhs * 8)
    X64 = ooc.DeviceBuffer(n * nrhs * 8)
    X = ooc.DeviceBuffer(n * nrhs * 8)
    import numpy as np
    for mu in mus:
        t0 = time.perf_counter()
        ooc.gram_tiles(A, n, nb, mu, seed=1)
        ooc.normal_fill(B, n, nrhs, n, seed=1, stream_id=2)
        log(f"[mpir] mu={mu}: generated in {time.perf_counter() - t0:.1f} s")
        row = {}
        x64 = None
        for prec in precs:
            dst = X64 if prec == "fp64" else X
            st = ooc.mpir_solve_tiles(A, n, nb, B, nrhs, dst, prec, tol=tol, max_iter=30)
            x = dst.to_array((n, nrhs))
            if prec == "fp64":
                x64 = x
            ferr = float(np.abs(x - x64).max() / np.abs(x64).max()) if x64 is not None else None
            st["forward_error_vs_fp64"] = ferr
            row[prec] = st
            log(f"[mpir] mu={mu} {prec}: {st['total_seconds']:.3f} s factor {st['factor_seconds']:.3f} "
                f"it {st['iterations']} eta {st['backward_error']:.2e} fell_back {st['fell_back']} "
                f"ferr {ferr}")
        for prec in precs:
            if prec != "fp64" and "fp64" in row:
                row[prec]["speedup_vs_fp64"] = row["fp64"]["total_seconds"] / row[prec]["total_seconds"]
        out[f"mu={mu}"] = row
    return out


def batched_workload(nsys: int):
    """Sizes n (log-uniform in [32, 1024]) and right-hand-side counts (1..8)
    of the batched workload."""
    import numpy as np
    rng = np.random.default_rng(2024)
    n = np.exp(rng.uniform(np.log(32), np.log(1024), size=nsys)).astype(np.int32)
    r = rng.integers(1, 9, size=nsys).astype(np.int32)
    return n, r


def bench_batched(ooc, n, r, reps: int, cpu_rec=None, split: int = -1) -> dict:
    """10000 independent SPD systems, n log-uniform in [32, 1024], A = G G^T/n
    + 0.1 I, 1..8 N(0,1) right-hand sides; FP64 and FP32+IR; systems/s of the
    graph-replayed persistent kernel and per-system backward error.  split:
    the block-column threshold (-1: autotuned, 0: none, k: systems of n >= k)."""
    import numpy as np
    nsys = len(n)
    for prec in ("fp64", "fp32"):
        ooc.Batch.set_split(prec, split)
    bt = ooc.Batch(n, r)
    t0 = time.perf_counter()
    bt.generate(mu=0.1, seed=2024)
    log(f"[batched] {nsys} systems generated in {time.perf_counter() - t0:.1f} s "
        f"(sum n^3/3 = {float((n.astype(float) ** 3).sum()) / 3:.3e} flops)")
    out = {"systems": nsys, "flops_factor": float((n.astype(float) ** 3).sum() / 3)}
    for prec in ("fp64", "fp32"):
        st, berr, iters, status = bt.solve(prec, tol=8 * 2.0 ** -53, max_iter=10, reps=reps)
        st["median_backward_error"] = float(np.median(berr[status == 0]))
        st["mean_iterations"] = float(iters.mean())
        out[prec] = st
        log(f"[batched] {prec}: {st['systems_per_second']:.0f} systems/s ({st['seconds'] * 1e3:.1f} ms"
            f"/batch), variant T={st['variant_T']} RB={st['variant_RB']}, block-column launches "
            f"for {st['split_systems']} systems (n >= {st['split_min_n']}), max berr "
            f"{st['max_backward_error']:.2e}, failed {st['failed']}, mean iters {st['mean_iterations']:.2f}")
    if cpu_rec:
        out["cpu_baseline"] = cpu_rec
        out["speedup_vs_cpu"] = round(out["fp64"]["systems_per_second"] / cpu_rec["value"], 1)
    return out


_CPU_SYSTEMS = []


def _cpu_batched_share(args):
    """One worker's share of the CPU baseline: systems k, k + w, k + 2w ...,
    one BLAS thread; returns its own elapsed seconds."""
    k, w = args
    from scipy.linalg import cho_factor, cho_solve
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        mine = _CPU_SYSTEMS[k::w]
        t0 = time.perf_counter()
        for A, B in mine:
            cho_solve(cho_factor(A, lower=True, check_finite=False), B, check_finite=False)
        return time.perf_counter() - t0


def cpu_batched_baseline(n, r, sample: int = 2000, procs: int = 0, reps: int = 5) -> dict:
    """CPU reference timing of the batched workload: LAPACK dpotrf + dpotrs
    (scipy) over a sample of the batch's sizes and right-hand-side counts,
    one BLAS thread per call, the systems spread over forked worker processes
    (one per core; no exec, and run before the batched mode touches the
    GPU).  Throughput = sample / the slowest worker's time, median of reps
    passes.  The matrices are
    drawn on the host from the same distribution (A = G G^T/n + 0.1 I), not
    the device's values; factor + solve per system, generation outside the
    clock."""
    import multiprocessing as mp
    import numpy as np
    procs = procs or min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(7)
    idx = rng.choice(len(n), size=min(sample, len(n)), replace=False)
    _CPU_SYSTEMS.clear()
    for i in idx:
        k = int(n[i])
        G = rng.standard_normal((k, k))
        _CPU_SYSTEMS.append((np.asfortranarray(G @ G.T / k + 0.1 * np.eye(k)),
                             np.asfortranarray(rng.standard_normal((k, int(r[i]))))))
    with mp.get_context("fork").Pool(procs) as pool:
        pool.map(_cpu_batched_share, [(k, 8 * procs) for k in range(procs)])   # warm
        runs = sorted(max(pool.map(_cpu_batched_share, [(k, procs) for k in range(procs)]))
                      for _ in range(reps))
    dt = runs[len(runs) // 2]
    _CPU_SYSTEMS.clear()
    nsum = float(sum(int(n[i]) for i in idx))
    return {"value": round(len(idx) / dt, 1), "unit": "systems/s", "cores": procs,
            "kind": "lapack",
            "sample": f"{len(idx)} of the {len(n)} systems' (n, nrhs), host-drawn matrices, "
                      f"scipy cho_factor/cho_solve, 1 BLAS thread per process, "
                      f"slowest of {procs} processes {dt:.3f} s, median of {reps} passes "
                      f"(mean n {nsum / len(idx):.0f})"}


def serve_workload(count: int, seed: int = 7):
    """n log-uniform in [64, 32768], 1..64 right-hand sides, 30 % of requests
    reuse an earlier matrix with new right-hand sides, mu in {1e-1, 1e-4}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        if reqs and rng.random() < 0.3:
            src = reqs[int(rng.integers(0, len(reqs)))]
            r = dict(src)
            r.update(nrhs=int(rng.integers(1, 65)), rhs_seed=10_000 + i)
        else:
            n = int(np.exp(rng.uniform(np.log(64), np.log(32768))))
            r = dict(matrix_id=i, n=n, nrhs=int(rng.integers(1, 65)),
                     mu=float(rng.choice([1e-1, 1e-4])), seed=1000 + i, rhs_seed=10_000 + i)
        reqs.append(r)
    return reqs


def bench_serve(ooc, count: int, rate: float, slots: int, budget_gb: float) -> dict:
    """Poisson arrivals at `rate` requests/s into the continuous-batching
    server; p50/p99 latency and solves/s for mixed precision (FP32 factor +
    FP64 refinement) and plain FP64 on the same GPU."""
    import threading
    import numpy as np
    reqs = serve_workload(count)
    arrivals = np.cumsum(np.random.default_rng(11).exponential(1.0 / rate, size=count))
    out = {"requests": count, "offered_rate_per_s": rate, "slots": slots, "budget_gb": budget_gb}
    for prec in ("fp32", "fp64"):
        srv = ooc.Server(int(budget_gb * (1 << 30)), slots=slots)
        # warm the kernels (module load, MFMA probe, batched autotune) outside the clock
        w = [srv.submit(precision=prec, **dict(matrix_id=-1, n=2048, nrhs=2, mu=0.1, seed=1,
                                               rhs_seed=2)),
             srv.submit(precision=prec, **dict(matrix_id=-2, n=200, nrhs=2, mu=0.1, seed=3,
                                               rhs_seed=4))]
        for t in w:
            srv.wait(t)
        tickets = [None] * count
        t0 = time.perf_counter()

        def client():
            for i, r in enumerate(reqs):
                delay = t0 + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = srv.submit(precision=prec, tol=8 * 2.0 ** -53, **r)

        th = threading.Thread(target=client)
        th.start()
        results = []
        for i in range(count):
            while tickets[i] is None:
                time.sleep(0.001)
            res, _ = srv.wait(tickets[i])
            results.append(res)
            if (i + 1) % 40 == 0:
                log(f"[serve] {prec}: {i + 1}/{count} done, {time.perf_counter() - t0:.1f} s")
        th.join()
        t_all = time.perf_counter() - t0
        srv.close()
        lat = np.array([r["latency_seconds"] for r in results])
        # the same latencies without the on-device input materialization
        # (A = G G^T / n + mu I costs n^3 flops, twice the factorization)
        lat_solve = np.array([r["latency_seconds"] - r["input_seconds"] for r in results])
        ok = [r for r in results if r["status"] == 0]
        routes = [r["route"] for r in results]
        out[prec] = {
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "p50_latency_excl_input_s": float(np.percentile(lat_solve, 50)),
            "p99_latency_excl_input_s": float(np.percentile(lat_solve, 99)),
            "solves_per_s": count / t_all,
            "wall_s": t_all,
            "failed": count - len(ok),
            "max_backward_error": max(r["backward_error"] for r in ok) if ok else None,
            "cache_hits": sum(r["cache_hit"] for r in results),
            "fell_back": sum(r["fell_back"] for r in results),
            "routes": {"batched": routes.count(0), "hbm": routes.count(1), "ooc": routes.count(2)},
            "max_batch": max(r["batch_size"] for r in results),
        }
        log(f"[serve] {prec}: p50 {out[prec]['p50_latency_s'] * 1e3:.1f} ms p99 "
            f"{out[prec]['p99_latency_s'] * 1e3:.1f} ms (without input generation: p50 "
            f"{out[prec]['p50_latency_excl_input_s'] * 1e3:.1f} ms p99 "
            f"{out[prec]['p99_latency_excl_input_s'] * 1e3:.1f} ms), "
            f"{out[prec]['solves_per_s']:.2f} solves/s, "
            f"max berr {out[prec]['max_backward_error']:.2e}, hits {out[prec]['cache_hits']}, "
            f"routes {out[prec]['routes']}, max batch {out[prec]['max_batch']}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed steps (an upper bound under --budget-s)")
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed warmup steps (an upper bound under --budget-s)")
    ap.add_argument("--budget-s", type=float, default=420.0,
                    help="wall-clock budget of the whole run from process start; no step "
                         "starts that is expected to cross it")
    ap.add_argument("--n", type=int, default=196608)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--verify-full", action="store_true",
                    help="after the timed steps, form the LIN scaled residual of the full-size "
                         "factor on the GPU (another n^3/3 flops, outside the clock)")
    ap.add_argument("--mode", choices=["ooc", "hbm", "mpir", "batched", "serve"], default="ooc",
                    help="ooc: the headline out-of-core factorization (host tiles, device "
                         "budget, PCIe-inclusive); hbm: the same factorization with the packed "
                         "tiles resident in HBM; mpir: mixed precision + refinement (n=65536, "
                         "nb=512, 16 RHS); batched: 10000 small systems; serve: Poisson stream")
    ap.add_argument("--budget-gib", type=float, default=64.0,
                    help="ooc mode: device budget (GiB) for the out-of-core factorization")
    ap.add_argument("--mpir-precisions", default="fp64,fp32,bf16",
                    help="mpir mode: comma-separated subset of fp64,fp32,bf16")
    ap.add_argument("--mpir-mus", default="1e-2,1e-6",
                    help="mpir mode: comma-separated diagonal shifts mu")
    ap.add_argument("--mp-variant", type=int, default=0, metavar="BITS",
                    help="mpir mode: oocdag_mp_set_update_variant bits (4: the FP32 route's "
                         "super-column updates on FP32 MFMA instead of three BF16 pieces)")
    ap.add_argument("--batch-uniform", default="", metavar="COUNT,N",
                    help="batched mode: COUNT systems of order N with 4 right-hand sides "
                         "instead of the 10000-system workload")
    ap.add_argument("--batch-split", type=int, default=-1, metavar="MIN_N",
                    help="batched mode: factor the systems of n >= MIN_N by block-column "
                         "launches (0: none; -1: autotuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run device-to-device copies of BYTES (a plain copy kernel of "
                         "known size that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path "
              f"runs on one MI355X; multi-GPU sharding is a design note only).", file=sys.stderr)
        return 2

    cpu_batched = None
    if args.mode == "batched" and not args.no_cpu and not args.batch_uniform:
        # the CPU reference's worker processes fork before this process
        # touches the GPU
        cpu_batched = cpu_batched_baseline(*batched_workload(10000))
        log(f"[batched] CPU LAPACK: {cpu_batched['value']:.0f} systems/s "
            f"({cpu_batched['sample']})")

    import oocdag_amd as ooc
    if ooc.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    if args.copy_calibration:
        nbytes = args.copy_calibration
        src, dst = ooc.DeviceBuffer(nbytes), ooc.DeviceBuffer(nbytes)
        ooc.lib().oocdag_memset_d(src.ptr, 1, nbytes)
        for _ in range(3):
            ooc.lib().oocdag_memcpy_d2d(dst.ptr, src.ptr, nbytes)
        ooc.lib().oocdag_device_sync()
        log(f"[bench] copy calibration: 3 device copies of {nbytes} bytes")
        src.free()
        dst.free()

    if args.mode == "serve":
        res = bench_serve(ooc, 120, 4.0, 3, 200.0)
        print(json.dumps({"metric": "served SPD solves: p99 latency at offered load",
                          "value": round(res["fp32"]["p99_latency_s"] * 1e3, 2), "unit": "ms",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f64/f32",
                          "config": {"workload": "Poisson 4 req/s, n~logU[64,32768], 1-64 rhs, "
                                                 "30% matrix reuse", "results": res}}), flush=True)
        return 0

    if args.mode == "batched":
        workload = "10000 SPD systems n~logU[32,1024], 1-8 rhs"
        sizes = batched_workload(10000)
        if args.batch_uniform:
            import numpy as np
            count, order = (int(x) for x in args.batch_uniform.split(","))
            sizes = (np.full(count, order, dtype=np.int32), np.full(count, 4, dtype=np.int32))
            workload = f"{count} SPD systems of n={order}, 4 rhs"
        res = bench_batched(ooc, *sizes, max(1, args.steps) * 3, cpu_rec=cpu_batched,
                            split=args.batch_split)
        print(json.dumps({"metric": "batched small SPD factor+solve throughput", "value":
                          round(res["fp64"]["systems_per_second"], 1), "unit": "systems/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "config": {"workload": workload, "results": res}}), flush=True)
        return 0

    if args.mode == "mpir":
        n = 65536 if args.n == 196608 else args.n
        precs = [p for p in args.mpir_precisions.split(",") if p]
        mus = [float(m) for m in args.mpir_mus.split(",") if m]
        ooc.mp_set_update_variant(args.mp_variant)
        res = bench_mpir(ooc, n, 512, 16, mus, precs)
        best = max([v["speedup_vs_fp64"] for r in res.values() for k, v in r.items()
                    if "speedup_vs_fp64" in v and v["fell_back"] == 0] or [float("nan")])
        print(json.dumps({"metric": "mixed-precision+IR speedup over plain FP64 at FP64 backward "
                          "error", "value": round(best, 3), "unit": "x", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f64/f32/bf16",
                          "config": {"workload": f"Gram SPD n={n} nb=512 nrhs=16",
                                     "mp_variant": args.mp_variant, "results": res}}),
              flush=True)
        return 0

    n, nb = args.n, args.nb
    flops = n ** 3 / 3.0
    # kept back for the work after the timed steps: the small correctness
    # check, the CPU baseline and the full residual when asked for
    reserve = 15.0 + (0 if args.no_verify else 10.0) + (0 if args.no_cpu else 15.0) + \
        (60.0 if args.verify_full else 0.0)
    budget = Budget(args.budget_s, reserve)

    if args.mode == "ooc":
        res = bench_ooc(ooc, n, nb, args.budget_gib, max(1, args.steps), max(0, args.warmup),
                        budget, args.verify_full)
        times = res["step_seconds"]
        steps_done, warm_done = res["steps_done"], res["warmup_done"]
        traffic, traffic_src = measured_traffic(
            f"update_gemm_ooc_n{n}_budget{args.budget_gib:g}g_pmc_traffic.md",
            {"n": n, "nb": nb, "budget_gib": f"{args.budget_gib:g}"})
        workload = (f"FP64 out-of-core tiled Cholesky n={n} nb={nb}: PDMATGEN('S','D',100) in "
                    f"pageable host memory as packed lower tiles "
                    f"({ooc.tiles_elems(n, nb) * 8 / 1e9:.1f} GB), device budget "
                    f"{args.budget_gib:g} GiB (panels {res['panels']}), column panels streamed "
                    f"over PCIe with double-buffered copies; timed PCIe-inclusive")
        extra = {"panels": res["panels"], "budget_bytes": res["budget_bytes"],
                 "h2d_GB_per_step": round(res["h2d_bytes_per_step"] / 1e9, 2),
                 "d2h_GB_per_step": round(res["d2h_bytes_per_step"] / 1e9, 2),
                 "h2d_GBps": round(res["h2d_GBps"], 2), "d2h_GBps": round(res["d2h_GBps"], 2),
                 "generate_seconds": [round(x, 2) for x in res["generate_seconds"]],
                 "L_last_diag": res["L_last_diag"]}
        full_resid = res["scaled_residual_full"]
        data = "synthetic: PDMATGEN('S','D', ISEED=100) generated on the GPU into host tiles"
    else:
        log(f"[bench] n={n} nb={nb}: allocating {ooc.tiles_elems(n, nb) * 8 / 1e9:.1f} GB of "
            f"tiles in HBM")
        buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
        ooc.set_progress(32 if n >= 65536 else 0)
        step_cost = [0.0]
        warm_done = 0
        for w in range(max(0, args.warmup)):
            if w >= 1 and not budget.fits((args.steps + 1) * step_cost[0]):
                break
            t0 = time.perf_counter()
            ooc.pdmatgen_tiles(buf, n, nb)
            info = ooc.potrf_tiles(buf, n, nb)
            step_cost[0] = max(step_cost[0], time.perf_counter() - t0)
            warm_done += 1
            log(f"[bench] warmup {warm_done} done, info={info}")
        times = []
        prof = {"launches": 0, "seconds": 0.0, "flops": 0.0, "bytes": 0.0}
        for k in range(max(1, args.steps)):
            if k >= 1 and not budget.fits(step_cost[0]):
                break
            t0 = time.perf_counter()
            ooc.pdmatgen_tiles(buf, n, nb)   # synchronous; outside the timed region
            ooc.lib().oocdag_device_sync()
            ooc.profile_enable(True)
            t1 = time.perf_counter()
            info = ooc.potrf_tiles(buf, n, nb)  # returns after the device finished
            ooc.lib().oocdag_device_sync()
            t2 = time.perf_counter()
            p = ooc.profile_read()
            ooc.profile_enable(False)
            if info != 0:
                log(f"[bench] factorization reported info={info}")
                return 1
            times.append(t2 - t1)
            step_cost[0] = max(step_cost[0], t2 - t0)
            for key in prof:
                prof[key] += p[key]
            log(f"[bench] step {k + 1}: {t2 - t1:.3f} s, {flops / (t2 - t1) / 1e12:.2f} TFLOP/s; "
                f"update GEMM {p['launches']} launches {p['seconds']:.3f} s")
        full_resid = None
        if args.verify_full:
            full_resid = ooc.llt_resid_tiles(buf, n, nb)
            log(f"[bench] full-size scaled residual {full_resid:.4e}")
        buf.free()
        res = {"prof_launches": prof["launches"], "prof_seconds": prof["seconds"],
               "prof_flops": prof["flops"], "prof_bytes": prof["bytes"]}
        steps_done = len(times)
        traffic, traffic_src = measured_traffic(f"update_gemm_n{n}_pmc_traffic.md", {"nb": nb})
        workload = f"FP64 HBM-resident tiled Cholesky n={n} nb={nb}"
        extra = {}
        data = "synthetic: PDMATGEN('S','D', ISEED=100) generated on the GPU into HBM"

    t = sum(times) / len(times)
    value = flops / t / 1e12
    ach = res["prof_flops"] / max(res["prof_seconds"], 1e-12) / 1e12
    alg_bytes = res["prof_bytes"] / max(res["prof_launches"], 1)
    resid = None if args.no_verify else verify_small_ooc(ooc)
    if resid is not None:
        log(f"[bench] small out-of-core check: scaled residual {resid:.3e}")
    rec = {
        "metric": "Cholesky TFLOP/s (n^3/3) vs FP64 peak",
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": steps_done,
        "warmup": warm_done,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "ms_per_step": round(t * 1e3, 2),
        "step_ms": [round(x * 1e3, 1) for x in times],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": data,
        "config": {"workload": workload, "n": n, "nb": nb,
                   "fraction_of_fp64_peak": round(value / FP64_PEAK, 4),
                   "small_check_scaled_residual": resid,
                   "scaled_residual_full": full_resid, **extra},
        "roofline": {"bound": "mfma", "achieved": round(ach, 4), "peak": FP64_PEAK,
                     "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": round(alg_bytes),
                     "kernel": "k_dgemm_dtv<1,0,*> panel update (K inside one tile / across tiles)",
                     "launches": res["prof_launches"],
                     "kernel_seconds": round(res["prof_seconds"], 4),
                     "avg_launch_ms": round(res["prof_seconds"] / max(res["prof_launches"], 1) * 1e3,
                                            3)},
    }
    if not args.no_cpu:
        log("[bench] cpu baseline (n=4096, nb=256, 28 MiB budget) ...")
        rec["cpu_baseline"] = cpu_baseline()
    rec["wall_s"] = round(elapsed(), 1)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
