"""PDLASCHK solve residual on the GPU (LIN/pdlaschk.f:22-26,186,295):
max_k ||A x_k - b_k||_inf / (||A||_inf ||x_k||_inf eps N), eps = 2^-53.

Checked against the CPU oracle (ora_solve_resid, pinned against numpy in
test_oracle.py): a planted error in X dominates the residual and must be
reproduced closely; rounding-level residuals of a correct solve agree to
their summation-order noise.  The mixed-precision, batched and served
results carry the same residual for their returned solutions.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def np_resid(A, X, B):
    R = B - A @ X
    an = np.abs(A).sum(1).max()
    return max(np.abs(R[:, k]).max() / (np.abs(X[:, k]).max() * an * EPS * A.shape[0])
               for k in range(X.shape[1]))


@pytest.mark.parametrize("n,nrhs", [(1000, 3), (4500, 16), (77, 1)])
def test_solve_resid_host_matches_oracle(gpu, ora, n, nrhs):
    ooc = gpu
    A0 = ora.pdmatgen(n, 128)
    L = A0.copy(order="F")
    assert ooc.cholesky(L, 128, 1 << 30)[1] == 0
    rng = np.random.default_rng(n)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    X = B.copy(order="F")
    ooc.potrs(L, X)
    r_gpu, r_cpu = ooc.solve_resid(A0, X, B), ora.solve_resid(A0, X, B)
    assert r_gpu <= 10 and r_cpu <= 10
    # the lower triangle defines A: the factor in A's lower triangle would
    # not; A's strict upper triangle is not read
    Aup = np.tril(A0) + 5.0 * np.triu(np.ones((n, n)), 1)
    assert ooc.solve_resid(np.asfortranarray(Aup), X, B) == r_gpu
    # a planted error dominates and is reproduced
    Xp = X.copy(order="F")
    Xp[n // 3, nrhs - 1] += 1e-7
    ref = ora.solve_resid(A0, Xp, B)
    assert ref > 1e3
    assert ooc.solve_resid(A0, Xp, B) == pytest.approx(ref, rel=1e-6)
    assert ref == pytest.approx(np_resid(A0, Xp, B), rel=1e-6)


@pytest.mark.parametrize("n,nb", [(3000, 512), (1500, 500), (1000, 100)])
def test_solve_resid_tiles_matches_oracle(gpu, ora, n, nb):
    """A in device packed tiles (full diagonal tiles, NB = 500 and 100 in
    padded slots): the same values as the host entry point."""
    ooc = gpu
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.pdmatgen_tiles(buf, n, nb)
    A0 = ora.pdmatgen(n, nb)
    rng = np.random.default_rng(7)
    B = np.asfortranarray(rng.standard_normal((n, 4)))
    X = np.asfortranarray(np.linalg.solve(A0, B))
    X[n - 1, 2] += 1e-6
    dX, dB = ooc.DeviceBuffer.from_array(X), ooc.DeviceBuffer.from_array(B)
    r = ooc.solve_resid_tiles(buf, n, nb, dX, dB, 4, n, n)
    ref = ora.solve_resid(A0, X, B)
    assert r == pytest.approx(ref, rel=1e-6)
    assert r == pytest.approx(ooc.solve_resid(A0, X, B), rel=1e-9)


def test_mpir_reports_solve_resid(gpu):
    ooc = gpu
    n, nb, nrhs = 4096, 256, 3
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.gram_tiles(buf, n, nb, 1e-2, seed=5)
    dB = ooc.DeviceBuffer(n * nrhs * 8)
    ooc.normal_fill(dB, n, nrhs, n, seed=5, stream_id=2)
    dX = ooc.DeviceBuffer(n * nrhs * 8)
    for prec in ("fp64", "fp32", "bf16"):
        st = ooc.mpir_solve_tiles(buf, n, nb, dB, nrhs, dX, prec)
        assert st["backward_error"] <= 8 * EPS
        # the returned X gives the same residual through the standalone entry
        r = ooc.solve_resid_tiles(buf, n, nb, dX, dB, nrhs, n, n)
        assert st["solve_resid"] <= 1.0 and r <= 1.0, (prec, st["solve_resid"], r)


def test_batched_reports_solve_resid(gpu):
    ooc = gpu
    rng = np.random.default_rng(3)
    n = rng.integers(20, 300, size=40).astype(np.int32)
    r = rng.integers(1, 9, size=40).astype(np.int32)
    bt = ooc.Batch(n, r)
    bt.generate(mu=0.1, seed=3)
    for prec in ("fp64", "fp32"):
        st, berr, iters, status, lres = bt.solve(prec, reps=1, with_resid=True)
        assert (status == 0).all()
        assert np.all(lres <= 10.0) and st["max_solve_resid"] == lres.max()
        for s in (0, 17, 39):
            A, B, X = bt.host_system(s)
            Xp = X.copy(order="F")
            # the kernel's residual is rounding-level; numpy's of the same X too
            assert np_resid(A, Xp, B) <= 10.0
