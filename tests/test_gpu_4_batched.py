"""Batched small SPD factor + solve (north-star addition; no reference code).
Checked per system against an FP64 oracle (numpy/LAPACK solve of the same
matrix copied back from the device); a system's result must be bit-identical
alone and inside a batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def sizes(k, seed, lo=32, hi=1024):
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(lo), np.log(hi), size=k)).astype(np.int32)
    r = rng.integers(1, 9, size=k).astype(np.int32)
    return n, r


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batched_matches_fp64_oracle(gpu, prec):
    ooc = gpu
    n, r = sizes(120, 5)
    n[:3] = [1024, 33, 1]   # extremes: the largest order, an odd one, a scalar system
    r[3] = 37               # more right-hand sides than are resident at once (chunks of 8)
    r[4] = 64
    bt = ooc.Batch(n, r)
    bt.generate(mu=0.1, seed=5)
    tol = 8 * EPS
    st, berr, iters, status = bt.solve(prec, tol=tol)
    print(prec, st, "iters max", iters.max())
    assert st["failed"] == 0 and (status == 0).all()
    assert berr.max() <= tol
    for s in list(range(6)) + list(range(6, len(n), 17)):
        A, B, X = bt.host_system(s)
        Xref = np.linalg.solve(A, B)
        assert np.abs(X - Xref).max() / np.abs(Xref).max() < 1e-12
        R = B - A @ X
        eta = max(np.abs(R[:, k]).max() / (np.abs(A).sum(1).max() * np.abs(X[:, k]).max() +
                                            np.abs(B[:, k]).max()) for k in range(B.shape[1]))
        assert eta <= 2 * tol
    if prec == "fp32":
        assert iters.max() >= 1


def test_batched_alone_equals_in_mix(gpu):
    ooc = gpu
    n, r = sizes(40, 9)
    bt = ooc.Batch(n, r)
    bt.generate(mu=0.1, seed=9)
    bt.solve("fp32")
    for s in (0, 7, 23):
        A, B, X = bt.host_system(s)
        one = ooc.Batch(n[s:s + 1], r[s:s + 1])
        one.A.free(); one.B.free()
        one.A = ooc.DeviceBuffer.from_array(np.vstack([A, np.zeros((one.ld(0) - n[s], n[s]))]))
        one.B = ooc.DeviceBuffer.from_array(np.vstack([B, np.zeros((one.ld(0) - n[s], r[s]))]))
        one.solve("fp32")
        _, _, X1 = one.host_system(0)
        assert np.array_equal(X, X1)


def test_batched_non_spd_status(gpu):
    ooc = gpu
    n = np.array([64, 100, 200], dtype=np.int32)
    r = np.array([1, 2, 3], dtype=np.int32)
    bt = ooc.Batch(n, r)
    bt.generate(mu=-5.0, seed=1)  # shifted far below zero: indefinite
    st, berr, iters, status = bt.solve("fp64")
    assert (status > 0).all() and st["failed"] == 3
