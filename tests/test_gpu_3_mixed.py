"""Mixed-precision factorization + FP64 iterative refinement (an addition with
no counterpart in the reference).  Checked against an FP64 oracle on the same
inputs: numpy/LAPACK solve of the same matrix copied back from the device.
Bars: normwise backward error <= tol (the FP64 level asked for), forward
error within kappa * eps * 1e2 of the FP64 solution, and a correct answer
whether or not the low-precision attempt fell back to FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def gram_host(ooc, buf, n, nb):
    A = np.zeros((n, n), order="F")
    ooc.tiles_to_host(buf, A, nb)
    return np.tril(A) + np.tril(A, -1).T


def test_gram_generator_properties(gpu):
    ooc = gpu
    n, nb, mu = 1024, 256, 1e-2
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, mu, seed=7)
    A = gram_host(ooc, buf, n, nb)
    # E[(G G^T / n)_ii] = 1, off-diagonal ~ N(0, 1/n)
    assert abs(np.mean(np.diag(A)) - (1 + mu)) < 0.02
    off = A[np.tril_indices(n, -1)]
    assert abs(np.std(off) * np.sqrt(n) - 1.0) < 0.05
    w = np.linalg.eigvalsh(A)
    assert w.min() > mu * 0.99 and w.max() < 4.5
    # diagonal tiles are stored full and symmetric
    tiles = buf.to_array((ooc.tiles_elems(n, nb),))
    d0 = tiles[: nb * nb].reshape(nb, nb, order="F")
    assert np.array_equal(d0, d0.T)
    # same seed -> same matrix
    buf2 = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf2, n, nb, mu, seed=7)
    assert np.array_equal(buf2.to_array((ooc.tiles_elems(n, nb),)), tiles)


# (2048, 256): one super-column (all updates FP64 inside it);
# (6144, 128): three super-columns of 16 tile columns -- narrow and bulk
# low-precision updates, both buffer parities; (100, 128) and (129, 256):
# a single padded tile (the second factored as two 128-column halves);
# (300, 128): one super-column narrower than 2048 columns, ragged;
# (4999, 128): 40 tile columns, the last one padded, in super-columns of
# 16, 16 and 8 -- narrow and bulk low-precision updates over a ragged edge.
@pytest.mark.parametrize("n,nb", [(2048, 256), (6144, 128), (100, 128), (129, 256), (300, 128),
                                  (4999, 128)])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("mu", [1e-2, 1e-6])
def test_mpir_matches_fp64_oracle(gpu, prec, mu, n, nb):
    ooc = gpu
    nrhs = 4
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, mu, seed=3)
    A = gram_host(ooc, buf, n, nb)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=3, stream_id=2)
    B = dB.to_array((n, nrhs))
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    tol = 8 * EPS
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=tol, max_iter=30)
    X = dX.to_array((n, nrhs))
    print(prec, mu, st)
    # backward error recomputed independently on the host
    R = B - A @ X
    eta = max(np.abs(R[:, k]).max() / (np.abs(A).sum(1).max() * np.abs(X[:, k]).max() +
                                        np.abs(B[:, k]).max()) for k in range(nrhs))
    assert st["backward_error"] <= tol
    assert eta <= 2 * tol
    Xref = np.linalg.solve(A, B)
    kappa = np.linalg.cond(A, 1)
    ferr = np.abs(X - Xref).max() / np.abs(Xref).max()
    assert ferr <= max(1e2 * kappa * EPS, 1e-13)
    if prec == "fp64":
        # the FP64 factor's solve alone meets tol or nearly so; a wrong
        # diagonal-tile inverse would show as many steps or a fallback
        assert st["fell_back"] == 0 and st["iterations"] <= 2, st
    if prec == "fp32" and mu == 1e-2:
        # kappa ~ 4e2: an FP32 factor contracts the error by ~kappa * 2^-24
        # per step, so a wrong update (which refinement alone would hide
        # behind many steps or the FP64 fallback) shows here
        assert st["fell_back"] == 0 and st["precision_used"] == 1
        assert st["iterations"] <= 3


def test_fp64_solve_tile_inverses_nb512_ragged(gpu):
    """The solves use Z_k = L_kk^-T of every diagonal tile, built blockwise
    from the 64 x 64 diagonal inverses (8 steps at NB = 512); a partial last
    tile (n = 2500) is padded with an identity diagonal."""
    ooc = gpu
    n, nb, nrhs = 2500, 512, 3
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 1e-2, seed=11)
    A = gram_host(ooc, buf, n, nb)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=11, stream_id=2)
    B = dB.to_array((n, nrhs))
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, "fp64", tol=8 * EPS, max_iter=30)
    X = dX.to_array((n, nrhs))
    assert st["fell_back"] == 0 and st["iterations"] <= 2, st
    Xref = np.linalg.solve(A, B)
    assert np.abs(X - Xref).max() / np.abs(Xref).max() <= 1e2 * np.linalg.cond(A, 1) * EPS


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_solve_large_grids(gpu, prec):
    """n = 20480, NB = 512: the early steps of each sweep and of the residual
    have 16384+ columns, which take the solve kernels' large-grid forms
    (unsplit K, 32 columns per wave); the later steps the split-K form.  The
    FP32 and BF16 factors' correction solves run those kernels in FP32.
    Backward error of the returned X recomputed on the host."""
    ooc = gpu
    n, nb, nrhs = 20480, 512, 2
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 1e-2, seed=13)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=13, stream_id=2)
    B = dB.to_array((n, nrhs))
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=8 * EPS, max_iter=30)
    X = dX.to_array((n, nrhs))
    # FP64: the factor's solve alone; FP32 (kappa ~ 4e2): ~kappa * 2^-24
    # contraction per step; BF16: ~kappa * 2^-8 per step
    assert st["fell_back"] == 0 and st["iterations"] <= {"fp64": 2, "fp32": 3, "bf16": 16}[prec], st
    A = gram_host(ooc, buf, n, nb)
    R = B - A @ X
    anorm = np.abs(A).sum(1).max()
    del A
    eta = max(np.abs(R[:, k]).max() / (anorm * np.abs(X[:, k]).max() + np.abs(B[:, k]).max())
              for k in range(nrhs))
    assert eta <= 2 * 8 * EPS


@pytest.mark.parametrize("n", [2048, 5000])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_mpir_nb1024(gpu, prec, n):
    """NB = 1024: super-columns of two tile columns, each tile column factored
    as halves of 512, 256 and 128 columns; n = 2048 is a single super-column
    (A copied whole before the factorization), n = 5000 has a ragged last
    tile and three super-columns."""
    ooc = gpu
    nb, nrhs = 1024, 3
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 1e-2, seed=19)
    A = gram_host(ooc, buf, n, nb)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=19, stream_id=2)
    B = dB.to_array((n, nrhs))
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=8 * EPS, max_iter=30)
    X = dX.to_array((n, nrhs))
    assert st["fell_back"] == 0 and st["iterations"] <= {"fp32": 3, "bf16": 16}[prec], st
    Xref = np.linalg.solve(A, B)
    assert np.abs(X - Xref).max() / np.abs(Xref).max() <= 1e2 * np.linalg.cond(A, 1) * EPS


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mpir_many_rhs(gpu, prec):
    """72 right-hand sides: the solves and the residual take the general GEMM
    kernels (not the few-right-hand-side ones), whose K runs across the 32
    tile columns (rows) of a super-step (n = 6144, NB = 128: two super-steps);
    an FP32 factor's corrections then solve in FP64."""
    ooc = gpu
    n, nb, nrhs = 6144, 128, 72
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 1e-2, seed=17)
    A = gram_host(ooc, buf, n, nb)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=17, stream_id=2)
    B = dB.to_array((n, nrhs))
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=8 * EPS, max_iter=30)
    X = dX.to_array((n, nrhs))
    assert st["fell_back"] == 0 and st["iterations"] <= 3, st
    R = B - A @ X
    eta = max(np.abs(R[:, k]).max() / (np.abs(A).sum(1).max() * np.abs(X[:, k]).max() +
                                        np.abs(B[:, k]).max()) for k in range(nrhs))
    assert eta <= 2 * 8 * EPS
    Xref = np.linalg.solve(A, B)
    assert np.abs(X - Xref).max() / np.abs(Xref).max() <= 1e2 * np.linalg.cond(A, 1) * EPS


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_mpir_update_variants_agree(gpu, prec):
    """Both FP32 update kernels accumulate every element over the same
    4-wide K slabs in the same order, and the bulk update touches other
    tiles than the next super-column's factorization whether or not the two
    overlap, so every variant gives a bit-identical factorization,
    refinement and solution."""
    ooc = gpu
    n, nb, nrhs = 6144, 128, 4
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=5, stream_id=2)
    out = []
    try:
        for variant in (3, 2, 1, 0):
            assert ooc.mp_set_update_variant(variant) == 8
            ooc.gram_tiles(buf, n, nb, 1e-2, seed=5)
            dX = ooc.DeviceBuffer(n * nrhs * 8)
            st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=8 * EPS, max_iter=30)
            if prec == "fp32":
                assert st["fell_back"] == 0 and st["iterations"] <= 3, st
            out.append((st["iterations"], dX.to_array((n, nrhs))))
    finally:
        ooc.mp_set_update_variant(0)
    for it, x in out[1:]:
        assert it == out[0][0]
        assert np.array_equal(x, out[0][1])


@pytest.mark.parametrize("mu", [1e-2, 1e-6])
def test_mpir_fp32_three_piece_updates_keep_fp32_accuracy(gpu, mu):
    """The FP32 route's super-column updates from three BF16 pieces per
    element (x = h + m + l, six BF16 products) against the same updates on
    FP32 MFMA (variant bit 4) and against the scaled-BF16 route.  With tol = 1
    the first solve is accepted, so its backward error measures the factor:
    the pieces' must be within a small factor of FP32's, and (mu = 1e-2; at
    1e-6 the BF16 factor is not positive definite) far below BF16's
    (u = 2^-8); the pieces' error without l would be ~2^-16, without m and l
    ~2^-8, and both fail.  Refined to 8 eps, the two FP32 factors take the
    same number of steps (+-1).  n = 8192, NB = 256: four super-columns,
    narrow and bulk updates."""
    ooc = gpu
    n, nb, nrhs = 8192, 256, 4
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, mu, seed=6)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=6, stream_id=2)
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    first, steps = {}, {}
    try:
        runs = [("pieces", "fp32", 0), ("fp32", "fp32", 4)]
        for name, prec, variant in runs + ([("bf16", "bf16", 0)] if mu == 1e-2 else []):
            ooc.mp_set_update_variant(variant)
            st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=1.0, max_iter=30)
            assert st["iterations"] == 0 and st["fell_back"] == 0, (name, st)
            first[name] = st["backward_error"]
            if prec == "fp32":
                st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec, tol=8 * EPS,
                                          max_iter=30)
                assert st["fell_back"] == 0 and st["backward_error"] <= 8 * EPS, (name, st)
                steps[name] = st["iterations"]
    finally:
        ooc.mp_set_update_variant(0)
    print(mu, first, steps)
    assert first["pieces"] <= 3 * first["fp32"], first
    if mu == 1e-2:
        assert first["bf16"] >= 30 * first["pieces"], first
    assert abs(steps["pieces"] - steps["fp32"]) <= 1, steps


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_mpir_non_spd_reports_info(gpu, prec):
    """An indefinite matrix: every attempt of the cascade (BF16 -> FP32 ->
    FP64) meets a bad pivot, the FP64 attempt reports LAPACK's info (the
    first column whose pivot is not positive, as numpy's Cholesky fails)."""
    ooc = gpu
    n, nb = 3072, 128   # BF16: two super-columns, narrow and bulk updates
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, -10.0, seed=1)   # shifted far below zero: indefinite
    dB = ooc.DeviceBuffer(n * 8)
    ooc.normal_fill(dB, n, 1, n)
    dX = ooc.DeviceBuffer(n * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, 1, dX, prec)
    assert st["info"] == 1 and st["fell_back"] == 1 and st["precision_used"] == 0


def test_mpir_bf16_cascades_through_fp32(gpu):
    """kappa ~ 4e5 (mu = 1e-5): kappa * 2^-8 >> 1, so the BF16 attempt cannot
    refine; the next attempt is FP32 (kappa * 2^-24 < 1), not FP64."""
    ooc = gpu
    n, nb, nrhs = 4096, 256, 2
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 1e-5, seed=4)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=4, stream_id=2)
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, "bf16", tol=8 * EPS, max_iter=30)
    assert st["fell_back"] == 1 and st["precision_used"] == 1, st
    assert st["backward_error"] <= 8 * EPS and st["info"] == 0, st
    A = gram_host(ooc, buf, n, nb)
    X, B = dX.to_array((n, nrhs)), dB.to_array((n, nrhs))
    Xref = np.linalg.solve(A, B)
    assert np.abs(X - Xref).max() / np.abs(Xref).max() <= 1e2 * np.linalg.cond(A, 1) * EPS


def test_mpir_bf16_per_tile_scales(gpu):
    """A = L' L'^T with L' = L C, C = diag(2^(t mod 3)) by tile column t (L the
    Cholesky factor of a Gram matrix): the tiles of one tile row of a
    super-column get different power-of-two scales, so the BF16 update's
    accumulator rescaling at each tile boundary along K is exercised (a wrong
    ratio or scale index breaks the factor and the refinement falls back).
    n = 5120, NB = 512: narrow and bulk updates, the bulk's scale rows offset."""
    ooc = gpu
    n, nb, nrhs = 5120, 512, 3
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 0.1, seed=5)
    L = np.linalg.cholesky(gram_host(ooc, buf, n, nb))
    L *= 2.0 ** (np.arange(n) // nb % 3)[None, :]
    A = L @ L.T
    A = np.asfortranarray((A + A.T) / 2)
    ooc.tiles_from_host(buf, A, nb)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=5, stream_id=2)
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, "bf16", tol=8 * EPS, max_iter=30)
    assert st["fell_back"] == 0 and st["precision_used"] == 2 and st["iterations"] <= 16, st
    X, B = dX.to_array((n, nrhs)), dB.to_array((n, nrhs))
    Xref = np.linalg.solve(A, B)
    assert np.abs(X - Xref).max() / np.abs(Xref).max() <= 1e2 * np.linalg.cond(A, 1) * EPS


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
def test_mpir_padded_leading_dimensions(gpu, prec):
    """B and X inside larger column-major arrays (odd ldb = N + 3, ldx = N + 1)
    through the C ABI: the same solution, bit for bit, as with ldb = ldx = N,
    and the padding rows of X left as they were."""
    import ctypes as C
    ooc = gpu
    n, nb, nrhs = 300, 128, 3
    ldb, ldx = n + 3, n + 1
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 1e-2, seed=4)
    rng = np.random.default_rng(4)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dB = ooc.DeviceBuffer.from_array(B)
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec)
    X = dX.to_array((n, nrhs))
    Bp = np.full((ldb, nrhs), 7.0, order="F")
    Bp[:n] = B
    Xp = np.full((ldx, nrhs), -3.0, order="F")
    dBp, dXp = ooc.DeviceBuffer.from_array(Bp), ooc.DeviceBuffer.from_array(Xp)
    stp = ooc.MpirStats()
    tol = 8 * EPS
    rc = ooc.lib().oocdag_mpir_solve_tiles(buf.ptr, n, nb, dBp.ptr, nrhs, ldb, dXp.ptr, ldx,
                                           ooc.PRECISIONS[prec], tol, 30, C.byref(stp))
    assert rc == 0
    Xp = dXp.to_array((ldx, nrhs))
    assert np.array_equal(Xp[:n], X)
    assert (Xp[n:] == -3.0).all()
    assert stp.as_dict()["iterations"] == st["iterations"]
