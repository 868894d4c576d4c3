"""GPU parity at the reference's own boundary: its panel schedule (find_Yn,
oocdag.cpp:478-492, one staged X tile column) and any tile edge NB, as
Cholesky() and main take them (oocdag.cpp:241-247; the reference's scripts
run NB = 2500 and NB = 3000, job.sh:3-4).

A tile whose edge is not a multiple of 64 sits on the device in a slot of
edge NB rounded up to 64, padded with an identity block; the reference's
tile grid, memBlock and panel schedule are kept.  Bars as in
test_gpu_2_cholesky: L within 1e-10 of the oracle (which runs the
reference's algorithm at the same NB), LIN scaled residual <= 10, strict
upper triangle untouched, exact copy volume.
"""
import os
import subprocess

import numpy as np
import pytest

from test_gpu_2_cholesky import L_TOL, expected_traffic, rel, slot

pytestmark = pytest.mark.gpu


def test_n4096_nb256_28mib_reference_schedule(gpu, ora):
    """n = 4096, NB = 256, memsize = 28 MiB (memBlock = 56 tiles) with one X
    column: the reference's four panels [2, 3, 5, 6], the reference's A2Y/A2X
    copy volume (oocquark.cpp:253-319) and the oracle's L."""
    ooc = gpu
    n, nb, mem = 4096, 256, 28 << 20
    plan = ooc.panel_plan(n, nb, mem, nx=1)
    assert plan == ora.panel_plan(n, nb, mem) == [2, 3, 5, 6]
    A = ooc.pdmatgen(n, nb)
    A0 = A.copy(order="F")
    Lref = A.copy(order="F")
    _, info_ref = ora.cholesky(Lref, nb, mem)
    t, info = ooc.cholesky(A, nb, mem, nx=1)
    assert info == info_ref == 0 and t > 0
    st = ooc.last_stats()
    h2d, d2h = expected_traffic(n, nb, plan)
    assert (st["h2d_bytes"], st["d2h_bytes"]) == (h2d, d2h)
    assert rel(A, Lref) < L_TOL
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))
    assert ora.llt_resid(A) <= 10.0
    assert A[n - 1, n - 1] == pytest.approx(Lref[n - 1, n - 1], rel=1e-12)


def test_schedules_agree_bit_for_bit(gpu):
    """One X column or two: every tile receives its left-looking updates in
    the same order through the same kernels, so the factors are identical
    whenever neither schedule splits its first panel."""
    ooc = gpu
    n, nb, mem = 3000, 128, 72 * 128 * 128 * 8
    p1, p2 = ooc.panel_plan(n, nb, mem, nx=1), ooc.panel_plan(n, nb, mem, nx=2)
    assert p1 != p2 and p1[0] < 4 and p2[0] < 4, (p1, p2)
    A1, A2 = ooc.pdmatgen(n, nb), ooc.pdmatgen(n, nb)
    assert ooc.cholesky(A1, nb, mem, nx=1)[1] == 0
    assert ooc.cholesky(A2, nb, mem, nx=2)[1] == 0
    assert np.array_equal(A1, A2)


@pytest.mark.parametrize("n,nb,budget_tiles,nx", [
    (1500, 500, 6, 1),      # NB = 500 -> 512-wide slots, reference schedule [1, 2]
    (1700, 500, 8, 0),      # ragged N, default schedule [1, 2, 1]
    (2600, 2500, 4, 1),     # the reference's NB = 2500 (job.sh:4), panels [1, 1]
    (3001, 300, 30, 1),     # ragged, NB = 300 -> 320, panels [1, 2, 3, 5]
    (777, 100, 40, 2),      # small odd NB (128-wide slots), double-buffered X [3, 5]
    (1000, 1000, 2, 1),     # N = NB: one tile, 1024-wide slot
    (450, 37, 40, 1),       # tiny NB: 64-wide slots, panels [2, 2, 4, 5]
    (450, 37, 100, 0),      # first panel split (default schedule [7, 6])
])
def test_cholesky_any_nb_matches_oracle(gpu, ora, n, nb, budget_tiles, nx):
    ooc = gpu
    mem = budget_tiles * nb * nb * 8
    plan = ooc.panel_plan(n, nb, mem, nx=nx)
    assert plan is not None and sum(plan) == -(-n // nb)
    if nx == 1:
        assert plan == ora.panel_plan(n, nb, mem)
    A = ooc.pdmatgen(n, nb)
    A0 = A.copy(order="F")
    Lref = A.copy(order="F")
    _, info_ref = ora.cholesky(Lref, nb, 1 << 34)
    _, info = ooc.cholesky(A, nb, mem, nx=nx)
    print(f"n={n} nb={nb} slot={slot(nb)} nx={nx} panels={plan}")
    assert info == info_ref == 0
    st = ooc.last_stats()
    h2d, d2h = expected_traffic(n, nb, plan)
    assert (st["h2d_bytes"], st["d2h_bytes"]) == (h2d, d2h)
    assert rel(A, Lref) < L_TOL
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))
    assert ora.llt_resid(A) <= 10.0


@pytest.mark.parametrize("n,nb,bad", [(1500, 500, 700), (1000, 100, 99), (777, 250, 776)])
def test_any_nb_reports_non_spd_column(gpu, n, nb, bad):
    """info is the LAPACK column of the matrix, not of the padded slots."""
    from scipy.linalg import lapack
    A = gpu.pdmatgen(n, nb)
    A[bad, bad] = -A[bad, bad]
    _, linfo = lapack.dpotrf(A.copy(order="F"), lower=1)
    _, info = gpu.cholesky(A, nb, 20 * nb * nb * 8, nx=1)
    assert info == linfo == bad + 1


@pytest.mark.parametrize("n,nb", [(1500, 500), (1000, 100), (2600, 2500)])
def test_tile_entries_any_nb(gpu, ora, n, nb):
    """The packed-tile entry points with padded slots: generator bit-exact
    against the oracle (and the host-tiles generator), HBM-resident
    factorization against the oracle, LIN residual on the device, the tile
    solve against numpy, the host-tiles out-of-core factorization bit for
    bit equal to the HBM-resident one."""
    ooc = gpu
    ts = ooc.tile_edge(nb)
    assert ts == slot(nb)
    elems = ooc.tiles_elems(n, nb)
    bb = -(-n // nb)
    assert elems == bb * (bb + 1) // 2 * ts * ts
    buf = ooc.DeviceBuffer(elems * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    A0 = ora.pdmatgen(n, nb)
    tiles = buf.to_array((elems,))
    assert np.array_equal(ooc.tiles_to_dense(tiles, n, nb), np.tril(A0))
    # slot padding: zero, unit diagonal on diagonal tiles
    t0 = tiles[:ts * ts].reshape(ts, ts, order="F")
    assert np.array_equal(t0[nb:, nb:], np.eye(ts - nb)) and not t0[nb:, :nb].any()
    ht = ooc.HostTiles(n, nb, pinned=False)
    try:
        ht.generate()
        assert np.array_equal(ht.array(), tiles)
        _, info = ooc.cholesky_host_tiles(ht, 3 * nb * nb * 8 * bb, nx=1)
        assert info == 0
        L_host = ht.array().copy()
    finally:
        ht.free()
    assert ooc.potrf_tiles(buf, n, nb) == 0
    L_hbm = buf.to_array((elems,))
    T = A0.copy(order="F")   # the strict upper triangle keeps A, as PDPOTRRV expects
    ooc.tiles_to_host(buf, T, nb)
    Lref = A0.copy(order="F")
    ora.cholesky(Lref, nb, 1 << 34)
    assert rel(T, Lref) < L_TOL
    assert np.max(np.abs(L_host - L_hbm)) <= 1e-13 * np.max(np.abs(L_hbm))
    r = ooc.llt_resid_tiles(buf, n, nb)
    assert r <= 10.0 and r == pytest.approx(ora.llt_resid(T), rel=0.5)
    rng = np.random.default_rng(n)
    B = np.asfortranarray(rng.standard_normal((n, 5)))
    dB = ooc.DeviceBuffer.from_array(B)
    ooc.potrs_tiles(buf, n, nb, dB, 5, n)
    X = dB.to_array((n, 5))
    Xref = np.linalg.solve(A0, B)
    assert np.max(np.abs(X - Xref)) / np.max(np.abs(Xref)) < 1e-12
    # tiles_from_host round trip keeps the slot padding
    buf2 = ooc.DeviceBuffer(elems * 8)
    ooc.tiles_from_host(buf2, A0, nb)
    assert np.array_equal(buf2.to_array((elems,)), tiles)


@pytest.mark.parametrize("n,lda", [(100, 100), (200, 201), (64, 65), (129, 130)])
def test_dpotrf_dtrsm_any_n(gpu, n, lda):
    """dpotrf_('L') and cublasDtrsm(R,L,T,N) replacements for sizes and
    leading dimensions the reference's tiles can have (a ragged edge tile)."""
    from scipy.linalg import cholesky, solve_triangular
    ooc = gpu
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + np.eye(n)
    Ap = np.zeros((lda, n), order="F")
    Ap[:n] = A
    Ap[:n][np.triu_indices(n, 1)] = 7.0   # strict upper must survive
    d = ooc.DeviceBuffer.from_array(Ap)
    assert ooc.dpotrf_l(d, lda, n) == 0
    out = d.to_array((lda, n))
    Lref = cholesky(A, lower=True)
    assert np.max(np.abs(np.tril(out[:n]) - Lref)) < 1e-12
    assert np.all(out[:n][np.triu_indices(n, 1)] == 7.0)
    m = 37
    B = np.asfortranarray(rng.standard_normal((m, n)))
    Bp = np.zeros((m + 1, n), order="F")
    Bp[:m] = B
    dB = ooc.DeviceBuffer.from_array(Bp)
    ooc.dtrsm_rltn(d, lda, dB, m + 1, m, n)
    got = dB.to_array((m + 1, n))
    ref = solve_triangular(Lref, B.T, lower=True).T
    assert np.max(np.abs(got[:m] - ref)) / np.max(np.abs(ref)) < 1e-12
    assert not got[m:].any()


def test_cli_reference_nb_2500(gpu):
    """The reference's GPU run shape, `oocdag N 2500 memsize_MB` (job.sh:4,
    run_gpu.sh:35), at a small N: NB = 2500 accepted, the reference's panel
    lines and summary line printed, residual <= 10, VERIFY identity holds."""
    exe = os.path.join(os.path.dirname(gpu.LIB_PATH), "oocdag")
    out = subprocess.run([exe, "6000", "2500", "400", "--resid", "--verify"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    panels = [l for l in lines if l.startswith("bb ")]
    # bb = 3, memBlock = 400 MiB / (8 * 2500^2) = 8 tiles: find_Yn gives [2, 1]
    assert panels == ["bb 3 jb 0 YM 6000 YN 5000 Ym 3 Yn 2",
                      "bb 3 jb 2 YM 1000 YN 1000 Ym 1 Yn 1"], panels
    i = lines.index("N NB memsize(MB) quark_pthreads time Gflops")
    f = lines[i + 1].split()
    assert f[0] == "6000" and f[1] == "2500" and float(f[2]) == 400.0
    resid = [l for l in lines if l.startswith("scaled residual")][0]
    assert float(resid.split()[-1]) <= 10
    maxerr = [l for l in lines if l.startswith("max error")][0]
    assert float(maxerr.split()[2]) < 1e-10
