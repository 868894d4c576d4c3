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


LL_VARIANT = {"fp64": 2, "fp32": 4}   # the left-looking MFMA kernel


def _solve_with_split(ooc, bt, prec, min_n, **kw):
    ooc.Batch.set_variant(prec, LL_VARIANT[prec])
    ooc.Batch.set_split(prec, min_n)
    try:
        out = bt.solve(prec, **kw)
    finally:
        ooc.Batch.set_split(prec, -1)
        ooc.Batch.set_variant(prec, -1)
    total = int(bt.b_off[-1])
    return out, bt.X.to_array((total,))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batched_block_column_split_is_bit_identical(gpu, prec):
    """Systems factored by the block-column launches (one launch per 64-column
    step across all of them) give the same bits as inside one workgroup each:
    solutions, backward errors and refinement counts, for every threshold."""
    ooc = gpu
    n, r = sizes(60, 21, lo=32, hi=1024)
    n[:8] = [1024, 1023, 960, 769, 700, 513, 257, 65]   # full, ragged and one-step-past sizes
    r[:3] = [64, 9, 1]
    bt = ooc.Batch(n, r)
    bt.generate(mu=0.1, seed=21)
    (st0, berr0, it0, status0), x0 = _solve_with_split(ooc, bt, prec, 0)
    assert st0["split_systems"] == 0 and (status0 == 0).all()
    for min_n in (700, 257, 1):
        (st, berr, it, status), x = _solve_with_split(ooc, bt, prec, min_n)
        assert st["split_min_n"] == min_n
        assert st["split_systems"] == int((n >= min_n).sum())
        assert (status == 0).all()
        assert np.array_equal(x, x0), min_n
        assert np.array_equal(berr, berr0) and np.array_equal(it, it0)
    # and against the FP64 oracle on a few of the split systems
    for s in (0, 1, 3, 6):
        A, B, X = bt.host_system(s)
        Xref = np.linalg.solve(A, B)
        assert np.abs(X - Xref).max() / np.abs(Xref).max() < 1e-12


def test_batched_block_column_split_non_spd(gpu):
    """A bad pivot met by the block-column launches is reported at the same
    column as by the one-workgroup factorization (here in a later block column
    of two systems), the system's later steps are skipped and the others are
    solved."""
    ooc = gpu
    n = np.array([1024, 700, 300, 90], dtype=np.int32)
    r = np.array([2, 1, 3, 1], dtype=np.int32)
    bt = ooc.Batch(n, r)
    bt.generate(mu=0.1, seed=3)
    # plant a negative diagonal entry at row 700 of system 0 and row 130 of system 2
    total = int(bt.a_off[-1]) + bt.ld(3) * int(n[3])
    arena = bt.A.to_array((total,))
    for s, row in ((0, 700), (2, 130)):
        arena[int(bt.a_off[s]) + row + row * bt.ld(s)] = -1e3
    assert ooc.lib().oocdag_memcpy_h2d(bt.A.ptr, arena.ctypes.data, arena.nbytes) == 0
    (st0, _, _, status0), x0 = _solve_with_split(ooc, bt, "fp64", 0)
    (st1, _, _, status1), x1 = _solve_with_split(ooc, bt, "fp64", 64)
    assert st1["split_systems"] == 4
    assert list(status0) == [701, 0, 131, 0], status0
    assert np.array_equal(status0, status1)
    assert st0["failed"] == st1["failed"] == 2
    for s in (1, 3):   # the SPD systems: same bits
        lo, hi = int(bt.b_off[s]), int(bt.b_off[s]) + bt.ld(s) * int(r[s])
        assert np.array_equal(x0[lo:hi], x1[lo:hi])
