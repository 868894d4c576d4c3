"""Batched small SPD factor + solve (an addition with no counterpart in the
reference).
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


@pytest.mark.parametrize("prec,variant", [("fp64", v) for v in (-1, 0, 1, 2)] +
                         [("fp32", v) for v in (-1, 0, 1, 2, 3, 4)])
def test_batched_matches_fp64_oracle(gpu, prec, variant):
    """Every kernel variant (and the autotuned choice, -1) against the oracle."""
    ooc = gpu
    ooc.Batch.set_variant(prec, variant)
    try:
        _check_batched(ooc, prec)
    finally:
        ooc.Batch.set_variant(prec, -1)


def _check_batched(ooc, prec):
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


# a fixed variant: the property is per kernel (separate autotuned calls may pick
# different kernels)
@pytest.mark.parametrize("variant", [0, 4])
def test_batched_alone_equals_in_mix(gpu, variant):
    ooc = gpu
    ooc.Batch.set_variant("fp32", variant)
    try:
        _alone_in_mix(ooc)
    finally:
        ooc.Batch.set_variant("fp32", -1)


def _alone_in_mix(ooc):
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


@pytest.mark.parametrize("variant", [-1, 2])
def test_batched_non_spd_status(gpu, variant):
    ooc = gpu
    ooc.Batch.set_variant("fp64", variant)
    try:
        _non_spd(ooc)
    finally:
        ooc.Batch.set_variant("fp64", -1)


def _non_spd(ooc):
    n = np.array([64, 100, 200], dtype=np.int32)
    r = np.array([1, 2, 3], dtype=np.int32)
    bt = ooc.Batch(n, r)
    bt.generate(mu=-5.0, seed=1)  # shifted far below zero: indefinite
    st, berr, iters, status = bt.solve("fp64")
    assert (status > 0).all() and st["failed"] == 3
