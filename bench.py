#!/usr/bin/env python3
"""Benchmark: FP64 Cholesky TFLOP/s (n^3/3 flops) on one MI355X.

One step = one factorization of the n x n PDMATGEN('S','D', ISEED=100)
matrix (the reference's input, oocdag.cpp:298-302) whose packed lower tiles
are already resident in HBM when the timed region starts; the generator runs
between steps, outside the timed region (the reference also times only
Cholesky(), oocdag.cpp:51,90).  Default workload: n = 196608, NB = 1024
(155.4 GB of packed lower tiles).

Prints one JSON line (metric, value, unit, n_gpus, steps, timing, config ...) plus
  roofline     -- the panel-update GEMM (dominant kernel), timed by HIP events
                  on its own stream, algorithmic flops / kernel seconds vs the
                  78.6 TFLOP/s FP64 matrix peak;
  cpu_baseline -- the CPU oracle (a port of the reference driver, run on the
                  host's cores) timed on a bounded sample (n = --cpu-n).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6  # TFLOP/s, MI355X FP64 matrix (dense)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def verify_small(ooc) -> float:
    """Scaled residual of a small factorization through the same path (the
    full-size L is checked by the GPU tests; here a cheap guard that the
    timed code computes a correct factor)."""
    import numpy as np
    import oracle
    n, nb = 4096, 1024
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    assert ooc.potrf_tiles(buf, n, nb) == 0
    A = oracle.pdmatgen(n, nb)
    ooc.tiles_to_host(buf, A, nb)
    buf.free()
    return float(oracle.llt_resid(A))


def cpu_baseline(n: int, nb: int) -> dict:
    """The reference driver restated on the CPU (oracle/oracle.c: same panel
    loop, staging and tile-task order), one task per tile on all host cores,
    each tile kernel single-threaded MKL when the MKL runtime is present (the
    reference linked MKL for its host dpotrf, Makefile:17,45), else the
    oracle's own loops.  The whole matrix is one panel (budget = lower
    triangle + one X column)."""
    import oracle
    oracle.build()
    backend = oracle.set_backend(1)
    threads = oracle.lib().ora_get_threads()
    A = oracle.pdmatgen(n, nb)
    bb = -(-n // nb)
    mem = (bb * (bb + 1) // 2 + bb) * nb * nb * 8
    t, info = oracle.cholesky(A, nb, mem)
    oracle.set_backend(0)
    if info != 0:
        raise RuntimeError(f"CPU baseline factorization failed, info={info}")
    kern = "MKL dgemm/dsyrk/dtrsm/dpotrf (sequential) per tile" if backend == 1 else \
        "the oracle's own FP64 tile loops"
    return {"value": round(n ** 3 / 3 / t / 1e12, 5), "unit": "TFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"CPU restatement of the reference driver (oracle/oracle.c), n={n} nb={nb}, "
                      f"{threads} threads, {kern}, {t:.2f} s"}


def update_algorithmic_bytes(n: int, nb: int) -> int:
    """Average algorithmic HBM bytes per update-GEMM launch of the look-ahead
    factorization: each launch reads its lower trapezoid C once, writes it
    once and reads the K = nb wide panel column once (B is rows of that same
    column); 16 B per C element + 8 B per panel element."""
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


def measured_traffic(n: int, nb: int):
    """HBM bytes per update-GEMM launch from the committed PMC summary of the
    same workload (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    reading of a copy kernel of known size); None when no summary covers this
    (n, nb)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"update_gemm_n{n}_pmc_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    fields = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
                    if int(fields["nb"]) == nb:
                        return int(fields["bytes"]), os.path.relpath(path, os.path.dirname(path) + "/..")
    except (OSError, KeyError, ValueError):
        pass
    return None, None


def bench_ooc(ooc, n: int, nb: int, budget_gib: float, verify_full: bool = False) -> dict:
    """Out-of-core FP64 factorization as the reference runs it: the matrix in
    host memory (packed lower tiles, pinned), the device limited to a budget
    below the packed triangle, panels streamed over PCIe.  PCIe-inclusive
    wall time of the factorization; copy-stream busy time gives GB/s."""
    mem = int(budget_gib * (1 << 30))
    plan = ooc.panel_plan(n, nb, mem)
    t0 = time.perf_counter()
    t = ooc.HostTiles(n, nb, pinned=None)
    t1 = time.perf_counter()
    t.generate()
    t2 = time.perf_counter()
    log(f"[ooc] n={n} nb={nb}: {t.elems * 8 / 1e9:.1f} GB {'pinned' if t.pinned else 'pageable'} "
        f"host tiles in {t1 - t0:.1f} s, "
        f"generated in {t2 - t1:.1f} s; budget {budget_gib} GiB, panels {plan}")
    ooc.set_progress(16)
    try:
        secs, info = ooc.cholesky_host_tiles(t, mem)
        st = ooc.last_stats()
        tiles = t.array()
        last = float(tiles[-1])   # L(N-1, N-1) when N is a multiple of nb
        resid = None
        if verify_full:
            # the factor that came back to the host, copied into HBM whole and
            # checked there (LIN residual against the regenerated matrix)
            t3 = time.perf_counter()
            buf = ooc.DeviceBuffer(t.elems * 8)
            ooc.lib().oocdag_memcpy_h2d(buf.ptr, t.ptr, t.elems * 8)
            resid = ooc.llt_resid_tiles(buf, n, nb)
            buf.free()
            log(f"[ooc] scaled residual of the out-of-core factor {resid:.4e} "
                f"({time.perf_counter() - t3:.1f} s)")
    finally:
        t.free()
    flops = n ** 3 / 3.0
    out = {"seconds": secs, "tflops": flops / secs / 1e12, "info": info, "panels": plan,
           "host_tiles_pinned": t.pinned,
           "budget_bytes": mem, "h2d_bytes": st["h2d_bytes"], "d2h_bytes": st["d2h_bytes"],
           "h2d_seconds": st["h2d_seconds"], "d2h_seconds": st["d2h_seconds"],
           "h2d_GBps": st["h2d_bytes"] / max(st["h2d_seconds"], 1e-12) / 1e9,
           "d2h_GBps": st["d2h_bytes"] / max(st["d2h_seconds"], 1e-12) / 1e9,
           "L_last_diag": last, "scaled_residual_full": resid}
    log(f"[ooc] {secs:.3f} s, {out['tflops']:.2f} TFLOP/s PCIe-inclusive; H2D "
        f"{out['h2d_bytes'] / 1e9:.1f} GB at {out['h2d_GBps']:.1f} GB/s, D2H "
        f"{out['d2h_bytes'] / 1e9:.1f} GB at {out['d2h_GBps']:.1f} GB/s, info {info}")
    return out


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


def bench_batched(ooc, n, r, reps: int, cpu_rec=None) -> dict:
    """10000 independent SPD systems, n log-uniform in [32, 1024], A = G G^T/n
    + 0.1 I, 1..8 N(0,1) right-hand sides; FP64 and FP32+IR; systems/s of the
    graph-replayed persistent kernel and per-system backward error."""
    import numpy as np
    nsys = len(n)
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
            f"/batch), variant T={st['variant_T']} RB={st['variant_RB']}, max berr "
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
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=196608)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--cpu-n", type=int, default=28672)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--verify-full", action="store_true",
                    help="after the timed steps, form the LIN scaled residual of the full-size "
                         "factor on the GPU (another n^3/3 flops, outside the clock)")
    ap.add_argument("--mode", choices=["fp64", "mpir", "batched", "serve", "ooc"], default="fp64",
                    help="fp64: the headline HBM-resident factorization; mpir: the mixed-"
                         "precision + refinement workload (n=65536, nb=512, 16 RHS)")
    ap.add_argument("--budget-gib", type=float, default=64.0,
                    help="ooc mode: device budget (GiB) for the out-of-core factorization")
    ap.add_argument("--mpir-precisions", default="fp64,fp32,bf16",
                    help="mpir mode: comma-separated subset of fp64,fp32,bf16")
    ap.add_argument("--mpir-mus", default="1e-2,1e-6",
                    help="mpir mode: comma-separated diagonal shifts mu")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run device-to-device copies of BYTES (a plain copy kernel of "
                         "known size that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path "
              f"runs on one MI355X; multi-GPU sharding is a design note only).", file=sys.stderr)
        return 2

    cpu_batched = None
    if args.mode == "batched" and not args.no_cpu:
        # the CPU reference's worker processes fork before this process
        # touches the GPU
        cpu_batched = cpu_batched_baseline(*batched_workload(10000))
        log(f"[batched] CPU LAPACK: {cpu_batched['value']:.0f} systems/s "
            f"({cpu_batched['sample']})")

    import oocdag_amd as ooc
    if ooc.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    if args.mode == "serve":
        res = bench_serve(ooc, 120, 4.0, 3, 200.0)
        print(json.dumps({"metric": "served SPD solves: p99 latency at offered load",
                          "value": round(res["fp32"]["p99_latency_s"] * 1e3, 2), "unit": "ms",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f64/f32",
                          "config": {"workload": "Poisson 4 req/s, n~logU[64,32768], 1-64 rhs, "
                                                 "30% matrix reuse", "results": res}}), flush=True)
        return 0

    if args.mode == "ooc":
        res = bench_ooc(ooc, args.n, args.nb, args.budget_gib, args.verify_full)
        print(json.dumps({"metric": "Cholesky TFLOP/s (n^3/3), out of core, PCIe-inclusive",
                          "value": round(res["tflops"], 4), "unit": "TFLOP/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic: PDMATGEN('S','D', ISEED=100) into pinned host tiles",
                          "config": {"workload": f"FP64 out-of-core tiled Cholesky n={args.n} "
                                                 f"nb={args.nb}, {args.budget_gib} GiB device budget",
                                     "results": res}}), flush=True)
        return 0

    if args.mode == "batched":
        res = bench_batched(ooc, *batched_workload(10000), max(1, args.steps) * 3,
                            cpu_rec=cpu_batched)
        print(json.dumps({"metric": "batched small SPD factor+solve throughput", "value":
                          round(res["fp64"]["systems_per_second"], 1), "unit": "systems/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "config": {"workload": "10000 SPD systems n~logU[32,1024], 1-8 rhs",
                                     "results": res}}), flush=True)
        return 0

    if args.mode == "mpir":
        n = 65536 if args.n == 196608 else args.n
        precs = [p for p in args.mpir_precisions.split(",") if p]
        mus = [float(m) for m in args.mpir_mus.split(",") if m]
        res = bench_mpir(ooc, n, 512, 16, mus, precs)
        best = max([v["speedup_vs_fp64"] for r in res.values() for k, v in r.items()
                    if "speedup_vs_fp64" in v and v["fell_back"] == 0] or [float("nan")])
        print(json.dumps({"metric": "mixed-precision+IR speedup over plain FP64 at FP64 backward "
                          "error", "value": round(best, 3), "unit": "x", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f64/f32/bf16",
                          "config": {"workload": f"Gram SPD n={n} nb=512 nrhs=16", "results": res}}),
              flush=True)
        return 0

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

    n, nb = args.n, args.nb
    flops = n ** 3 / 3.0
    log(f"[bench] n={n} nb={nb}: allocating {ooc.tiles_elems(n, nb) * 8 / 1e9:.1f} GB of tiles")
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.set_progress(16 if n >= 65536 else 0)

    for w in range(args.warmup):
        ooc.pdmatgen_tiles(buf, n, nb)
        info = ooc.potrf_tiles(buf, n, nb)
        log(f"[bench] warmup {w + 1}/{args.warmup} done, info={info}")

    times = []
    prof_t = prof_f = 0.0
    prof_n = 0
    for k in range(args.steps):
        t0 = time.perf_counter()
        ooc.pdmatgen_tiles(buf, n, nb)   # synchronous; outside the timed region
        log(f"[bench] step {k + 1}: generated in {time.perf_counter() - t0:.2f} s")
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
        prof_t += p["seconds"]
        prof_f += p["flops"]
        prof_n += p["launches"]
        log(f"[bench] step {k + 1}: {t2 - t1:.3f} s, {flops / (t2 - t1) / 1e12:.2f} TFLOP/s; "
            f"update GEMM {p['launches']} launches {p['seconds']:.3f} s "
            f"{p['flops'] / max(p['seconds'], 1e-12) / 1e12:.2f} TFLOP/s")
    full_resid = None
    if args.verify_full:
        # LIN scaled residual of the timed factor itself (L L^T rebuilt on the
        # device against the regenerated PDMATGEN matrix), outside the clock
        t0 = time.perf_counter()
        full_resid = ooc.llt_resid_tiles(buf, n, nb)
        log(f"[bench] full-size scaled residual {full_resid:.4e} ({time.perf_counter() - t0:.1f} s)")
    A_last = float("nan")
    buf.free()

    t = sum(times) / len(times)
    value = flops / t / 1e12
    ach = prof_f / max(prof_t, 1e-12) / 1e12
    traffic, traffic_note = measured_traffic(n, nb)
    resid = None if args.no_verify else verify_small(ooc)
    rec = {
        "metric": "Cholesky TFLOP/s (n^3/3) vs FP64 peak",
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: PDMATGEN('S','D', ISEED=100) generated on the GPU into HBM",
        "config": {"workload": f"FP64 HBM-resident tiled Cholesky n={n} nb={nb}",
                   "n": n, "nb": nb, "fraction_of_fp64_peak": round(value / FP64_PEAK, 4),
                   "small_check_scaled_residual_n4096": resid,
                   "scaled_residual_full": full_resid},
        "roofline": {"bound": "mfma", "achieved": round(ach, 4), "peak": FP64_PEAK,
                     "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK, 4), "traffic": traffic,
                     "traffic_source": traffic_note,
                     "algorithmic_bytes_per_launch": update_algorithmic_bytes(n, nb),
                     "kernel": "k_dgemm_dtv<1,0,true> panel update",
                     "launches": prof_n, "kernel_seconds": round(prof_t, 4)},
    }
    del A_last
    if not args.no_cpu:
        log(f"[bench] cpu baseline n={args.cpu_n} ...")
        rec["cpu_baseline"] = cpu_baseline(args.cpu_n, 512)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
