"""GPU parity of the factorization paths, through the C ABI.

Bar: scaled residual ||A - L L^T||_inf / (n eps ||A||_inf) <= 10
(LIN/pdlafchk.f semantics, computed by the oracle) and L within 1e-10
(normwise relative, max-abs) of the oracle's factor, which follows the
reference's op order with host dpotrf tiles.  The strict upper triangle of A
must come back untouched.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L_TOL = 1e-10


def rel(a, b):
    return np.max(np.abs(np.tril(a) - np.tril(b))) / np.max(np.abs(np.tril(b)))


def slot(nb):
    """Edge of the slot holding one nb x nb tile: nb rounded up to 64."""
    return -(-nb // 64) * 64


def expected_traffic(n, nb, widths):
    """Bytes an out-of-core run must copy in and out for the panel widths:
    per panel of Yn tile columns at tile column jb (Ym = bb - jb tile rows),
    A2Y loads its packed tile columns, sum_jj (Ym - jj) tiles
    (oocquark.cpp:253-275), and A2X loads rows jb.. of every earlier tile
    column, jb * Ym tiles (oocquark.cpp:301-319); every tile of the packed
    lower triangle is copied out once (oocquark.cpp:389-419).  A tile moves
    as its slot (nb rounded up to 64)."""
    bb = -(-n // nb)
    h2d, jb = 0, 0
    for yn in widths:
        ym = bb - jb
        h2d += sum(ym - jj for jj in range(yn)) + jb * ym
        jb += yn
    assert jb == bb
    tile = slot(nb) ** 2 * 8
    return h2d * tile, bb * (bb + 1) // 2 * tile


@pytest.mark.parametrize("n,nb,budget_tiles", [
    (4096, 256, 56),     # 28 MiB, default schedule (two X columns): 6 panels
    (1000, 64, 96),      # ragged edge, many panels
    (2048, 256, 60),     # whole lower triangle fits next to the X columns: one panel
    (1536, 128, 40),
    (3000, 64, 1200),    # panels [40, 7]: the first panel split (8 columns first), ragged
])
def test_cholesky_host_ooc_matches_oracle(gpu, ora, n, nb, budget_tiles):
    ooc = gpu
    mem = budget_tiles * nb * nb * 8
    A = ooc.pdmatgen(n, nb)
    A0 = A.copy(order="F")
    Lref = A.copy(order="F")
    _, info_ref = ora.cholesky(Lref, nb, mem)
    t, info = ooc.cholesky(A, nb, mem)
    print(f"n={n} nb={nb} panels={ooc.panel_plan(n, nb, mem)} t={t:.4f}s")
    assert info == info_ref == 0
    st = ooc.last_stats()
    h2d, d2h = expected_traffic(n, nb, ooc.panel_plan(n, nb, mem))
    assert (st["h2d_bytes"], st["d2h_bytes"]) == (h2d, d2h)
    assert rel(A, Lref) < L_TOL
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))
    r = ora.llt_resid(A)
    assert r <= 10.0
    # the reference prints L(N-1,N-1) as its checksum (oocdag.cpp:92)
    assert A[n - 1, n - 1] == pytest.approx(Lref[n - 1, n - 1], rel=1e-12)


@pytest.mark.parametrize("n,nb", [(1000, 128), (1536, 256)])
def test_host_tiles_generator_matches_device(gpu, n, nb):
    """PDMATGEN into pinned host tiles == the device tile generator (bit-exact,
    padding included)."""
    ooc = gpu
    t = ooc.HostTiles(n, nb)
    try:
        t.generate()
        dev = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
        ooc.pdmatgen_tiles(dev, n, nb)
        assert np.array_equal(t.array(), dev.to_array((ooc.tiles_elems(n, nb),)))
    finally:
        t.free()


@pytest.mark.parametrize("pinned", [True, False])
@pytest.mark.parametrize("n,nb,budget_tiles", [
    (4096, 256, 56),     # 4 panels
    (1000, 64, 96),      # ragged edge, many panels
    (1536, 128, 40),
    (3000, 64, 1200),    # first panel split, next panel prefetched into its tiles
])
def test_cholesky_host_tiles_ooc(gpu, ora, n, nb, budget_tiles, pinned):
    """Out-of-core factorization of host packed tiles (page-locked: direct
    DMA; pageable: staged through pinned slots): the same device work as the
    column-major host path, so L is bit-identical to it, and within 1e-10 of
    the oracle."""
    ooc = gpu
    mem = budget_tiles * nb * nb * 8
    t = ooc.HostTiles(n, nb, pinned=pinned)
    assert t.pinned == pinned
    try:
        t.generate()
        secs, info = ooc.cholesky_host_tiles(t, mem)
        assert info == 0 and secs > 0
        st = ooc.last_stats()
        h2d, d2h = expected_traffic(n, nb, ooc.panel_plan(n, nb, mem))
        assert (st["h2d_bytes"], st["d2h_bytes"]) == (h2d, d2h)
        L = ooc.tiles_to_dense(t.array(), n, nb)
    finally:
        t.free()
    A = ooc.pdmatgen(n, nb)
    ooc.cholesky(A, nb, mem)
    assert np.array_equal(L, np.tril(A))
    Lref = ooc.pdmatgen(n, nb)
    ora.cholesky(Lref, nb, mem)
    assert rel(L, Lref) < L_TOL


def test_cholesky_matches_mkl_tiles(gpu, ora):
    """L against the driver run on MKL dpotrf/dtrsm/dsyrk/dgemm tiles, the
    library the reference linked for its host dpotrf (Makefile:17,45)."""
    if ora.set_backend(1) != 1:
        pytest.skip("MKL runtime not loadable")
    try:
        n, nb, mem = 2048, 256, 24 * 256 * 256 * 8
        A = gpu.pdmatgen(n, nb)
        Lref = A.copy(order="F")
        _, i0 = ora.cholesky(Lref, nb, mem)
        _, i1 = gpu.cholesky(A, nb, mem)
        assert i0 == i1 == 0
        assert rel(A, Lref) < L_TOL
    finally:
        ora.set_backend(0)


def test_cholesky_F_closed_form(gpu):
    from test_oracle import closed_form_F
    n, nb = 777, 64
    A = gpu.pdmatgen(n, nb, "F")
    _, info = gpu.cholesky(A, nb, 40 * nb * nb * 8)
    assert info == 0
    assert np.max(np.abs(np.tril(A) - closed_form_F(n))) < 1e-12


@pytest.mark.parametrize("bad", [5, 300, 1023])
def test_cholesky_reports_non_spd_column(gpu, ora, bad):
    from scipy.linalg import lapack
    n, nb = 1024, 128
    A = gpu.pdmatgen(n, nb)
    A[bad, bad] = -A[bad, bad]
    _, linfo = lapack.dpotrf(A.copy(order="F"), lower=1)
    _, info = gpu.cholesky(A, nb, 20 * nb * nb * 8)
    assert info == linfo == bad + 1


@pytest.mark.parametrize("n,nb,budget_tiles", [
    (1, 64, 2),          # 1 x 1: L = sqrt(a)
    (7, 64, 2),          # one tile, almost all of it padding
    (64, 64, 2),         # exactly one full tile
    (100, 256, 3),       # one ragged tile, NB above N
    (257, 256, 4),       # two panels, the second a one-row tile column
    (257, 256, 1 << 12),
])
def test_cholesky_small_sizes(gpu, ora, n, nb, budget_tiles):
    """N <= NB (a single tile, padded with the identity) and a one-row edge
    tile, through every entry of the path: the column-major out-of-core
    entry, device tiles (potrf_tiles), and both solves."""
    ooc = gpu
    A0 = ora.pdmatgen(n, nb)
    Lref = A0.copy(order="F")
    _, info_ref = ora.cholesky(Lref, nb, 1 << 30)
    assert info_ref == 0
    A = A0.copy(order="F")
    _, info = ooc.cholesky(A, nb, budget_tiles * nb * nb * 8)
    assert info == 0
    assert rel(A, Lref) < L_TOL
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))
    assert ora.llt_resid(A) <= 10.0
    if n == 1:
        assert A[0, 0] == pytest.approx(np.sqrt(A0[0, 0]), rel=1e-15)
    # the GPU residual: rounding-level, and a planted error seen as the oracle sees it
    assert ooc.llt_resid(A) <= 10.0
    P = A.copy(order="F")
    P[n - 1, 0] += 1e-9
    assert ooc.llt_resid(P) == pytest.approx(ora.llt_resid(P), rel=1e-3)

    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    assert ooc.potrf_tiles(buf, n, nb) == 0
    T = np.zeros((n, n), order="F")
    ooc.tiles_to_host(buf, T, nb)
    assert rel(T, Lref) < L_TOL
    assert ooc.llt_resid_tiles(buf, n, nb) <= 10.0

    rng = np.random.default_rng(n)
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    Xref = np.linalg.solve(A0, B)
    X = B.copy(order="F")
    ooc.potrs(A, X)
    assert np.max(np.abs(X - Xref)) / np.max(np.abs(Xref)) < 1e-12
    dB = ooc.DeviceBuffer.from_array(B)
    ooc.potrs_tiles(buf, n, nb, dB, 3, n)
    X = dB.to_array((n, 3))
    assert np.max(np.abs(X - Xref)) / np.max(np.abs(Xref)) < 1e-12


def test_cholesky_budget_error(gpu):
    A = np.eye(1024, order="F")
    with pytest.raises(gpu.OocdagError):
        gpu.cholesky(A, 256, 7 * 256 * 256 * 8)


@pytest.mark.parametrize("n,nb,aform", [(1024, 256, "S"), (1000, 128, "S"), (640, 64, "F"),
                                        (3000, 512, "S")])
def test_potrf_tiles_hbm_resident(gpu, ora, n, nb, aform):
    ooc = gpu
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.pdmatgen_tiles(buf, n, nb, aform, "D", 100)
    info = ooc.potrf_tiles(buf, n, nb)
    assert info == 0
    got = ora.pdmatgen(n, nb, aform, "D")
    A0 = got.copy(order="F")
    ooc.tiles_to_host(buf, got, nb)
    Lref = A0.copy(order="F")
    ora.cholesky(Lref, nb, 1 << 34)
    assert rel(got, Lref) < L_TOL
    assert ora.llt_resid(got, aform, "D") <= 10.0


def test_potrf_tiles_repeatable(gpu):
    """Same input twice -> bit-identical factor (no races between launches)."""
    ooc = gpu
    n, nb = 1536, 256
    outs = []
    for _ in range(2):
        buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
        ooc.pdmatgen_tiles(buf, n, nb)
        assert ooc.potrf_tiles(buf, n, nb) == 0
        outs.append(buf.to_array((ooc.tiles_elems(n, nb),)))
    assert np.array_equal(outs[0], outs[1])


def test_large_factor_out_of_core_matches_hbm(gpu):
    """n = 32768, NB = 1024 (528 tiles, 4.4 GB): the HBM-resident factor has a
    LIN scaled residual <= 10 (formed on the GPU), and the out-of-core factor
    of the same matrix from pageable host tiles under a 2 GiB budget (panels
    [6, 9, 17]: first panel split, later panels prefetched into finished
    columns' tiles) agrees with it to 1e-12 and has its own residual <= 10."""
    ooc = gpu
    n, nb, mem = 32768, 1024, 2 << 30
    plan = ooc.panel_plan(n, nb, mem)
    assert len(plan) >= 3 and plan[0] >= 4, plan
    elems = ooc.tiles_elems(n, nb)
    buf = ooc.DeviceBuffer(elems * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    assert ooc.potrf_tiles(buf, n, nb) == 0
    assert ooc.llt_resid_tiles(buf, n, nb) <= 10.0
    L_hbm = buf.to_array((elems,))
    t = ooc.HostTiles(n, nb, pinned=False)
    try:
        t.generate()
        secs, info = ooc.cholesky_host_tiles(t, mem)
        assert info == 0 and secs > 0
        st = ooc.last_stats()
        h2d, d2h = expected_traffic(n, nb, ooc.panel_plan(n, nb, mem))
        assert (st["h2d_bytes"], st["d2h_bytes"]) == (h2d, d2h)
        L_ooc = t.array().copy()
    finally:
        t.free()
    scale = np.max(np.abs(L_hbm))
    assert np.max(np.abs(L_ooc - L_hbm)) <= 1e-12 * scale
    ooc.lib().oocdag_memcpy_h2d(buf.ptr, L_ooc.ctypes.data, elems * 8)
    assert ooc.llt_resid_tiles(buf, n, nb) <= 10.0
    buf.free()


# ---------------------------------------------------------------- solve path

@pytest.mark.parametrize("n,nrhs", [(1000, 5), (4096, 16), (300, 1)])
def test_potrs_host_matches_numpy(gpu, ora, n, nrhs):
    A = gpu.pdmatgen(n, 128)
    A0 = A.copy(order="F")
    _, info = gpu.cholesky(A, 128, 1 << 31)
    assert info == 0
    rng = np.random.default_rng(n)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    X = B.copy(order="F")
    gpu.potrs(A, X)
    Xref = np.linalg.solve(A0, B)
    assert np.max(np.abs(X - Xref)) / np.max(np.abs(Xref)) < 1e-12
    assert ora.solve_resid(A0, X, B) <= 10.0


def test_potrs_verify_identity(gpu):
    """VERIFY block of oocdag.cpp:379-464: dpotrs(L, A) = I."""
    n = 768
    A = gpu.pdmatgen(n, 128)
    A2 = A.copy(order="F")
    gpu.cholesky(A, 128, 1 << 30)
    gpu.potrs(A, A2)
    assert np.max(np.abs(A2 - np.eye(n))) < 1e-12


def test_potrs_tiles_device(gpu, ora):
    ooc = gpu
    n, nb, nrhs = 2000, 256, 7
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    assert ooc.potrf_tiles(buf, n, nb) == 0
    rng = np.random.default_rng(11)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dB = ooc.DeviceBuffer.from_array(B)
    ooc.potrs_tiles(buf, n, nb, dB, nrhs, n)
    X = dB.to_array((n, nrhs))
    A0 = ora.pdmatgen(n, nb)
    assert np.max(np.abs(X - np.linalg.solve(A0, B))) < 1e-12


# ---------------------------------------------------------------- LIN residual on the GPU

def test_llt_resid_gpu_matches_oracle(gpu, ora):
    n, nb = 1500, 128
    A = gpu.pdmatgen(n, nb)
    gpu.cholesky(A, nb, 1 << 30)
    r_gpu, r_cpu = gpu.llt_resid(A), ora.llt_resid(A)
    assert r_gpu <= 10 and r_cpu <= 10
    assert r_gpu == pytest.approx(r_cpu, rel=0.5)  # both rounding-level
    # a planted error dominates both and must agree closely
    B = A.copy(order="F")
    B[n - 1, 3] += 1e-9
    assert gpu.llt_resid(B) == pytest.approx(ora.llt_resid(B), rel=1e-3)
    # the strict upper triangle counts
    C_ = A.copy(order="F")
    C_[2, n - 5] += 1e-6
    assert gpu.llt_resid(C_) == pytest.approx(ora.llt_resid(C_), rel=1e-3)


def test_llt_resid_tiles_hbm(gpu, ora):
    ooc = gpu
    n, nb = 3000, 512
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    assert ooc.potrf_tiles(buf, n, nb) == 0
    r = ooc.llt_resid_tiles(buf, n, nb)
    A = ora.pdmatgen(n, nb)
    ooc.tiles_to_host(buf, A, nb)
    assert r <= 10
    assert r == pytest.approx(ora.llt_resid(A), rel=0.5)


def test_cli_reference_output(gpu):
    """oocdag N NB memsize_MB prints the reference's summary line."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(gpu.LIB_PATH), "oocdag")
    out = subprocess.run([exe, "2048", "256", "10", "--resid", "--verify"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    i = lines.index("N NB memsize(MB) quark_pthreads time Gflops")
    f = lines[i + 1].split()
    assert f[0] == "2048" and f[1] == "256" and float(f[2]) == 10.0
    resid = [l for l in lines if l.startswith("scaled residual")][0]
    assert float(resid.split()[-1]) <= 10
    maxerr = [l for l in lines if l.startswith("max error")][0]
    assert float(maxerr.split()[2]) < 1e-10
