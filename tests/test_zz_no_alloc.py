"""The allocation-free variant of the dpotrf_('L') replacement,
oocdag_dpotrf_l_noalloc (scratch in a fixed device array of the library),
against oocdag_dpotrf_l (hipMalloc'ed scratch) through the C ABI: the same
kernels in the same order, so the factors and info agree bit for bit, on
consecutive calls that share the array (the second smaller than the first,
a non-SPD input between good ones so the info word must be reset)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spd(n, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return G @ G.T / n + np.eye(n)


def factor(ooc, fn, A, lda):
    n = A.shape[0]
    Ap = np.zeros((lda, n), order="F")
    Ap[:n] = A
    Ap[:n][np.triu_indices(n, 1)] = 3.0   # the strict upper triangle must survive
    d = ooc.DeviceBuffer.from_array(Ap)
    info = C.c_int(-7)
    rc = getattr(ooc.lib(), fn)(n, d.ptr, lda, C.byref(info), None)
    out = d.to_array((lda, n))
    d.free()
    return rc, info.value, out


def test_noalloc_matches_dpotrf_l_on_consecutive_calls(gpu):
    ooc = gpu
    from scipy.linalg import lapack
    cases = [(spd(1024, 1), 1026), (spd(512, 2), 512), (None, 256), (spd(256, 4), 258)]
    bad = spd(256, 3)
    bad[100, 100] = -5.0
    cases[2] = (bad, 256)
    for A, lda in cases:
        rc0, i0, ref = factor(ooc, "oocdag_dpotrf_l", A, lda)
        rc1, i1, got = factor(ooc, "oocdag_dpotrf_l_noalloc", A, lda)
        assert rc0 == rc1 == 0
        _, linfo = lapack.dpotrf(A.copy(), lower=1)
        assert i0 == i1 == linfo
        assert np.array_equal(got, ref, equal_nan=True)   # NaNs past a failed pivot too
        n = A.shape[0]
        assert np.all(got[:n][np.triu_indices(n, 1)] == 3.0)
        if linfo == 0:
            assert np.max(np.abs(np.tril(got[:n]) - np.linalg.cholesky(A))) < 1e-12


def test_noalloc_rejects_shapes_beyond_its_scratch(gpu):
    ooc = gpu
    d = ooc.DeviceBuffer(1 << 16)
    info = C.c_int(0)
    f = ooc.lib().oocdag_dpotrf_l_noalloc
    for n, lda in [(16384 + 64, 16384 + 64), (100, 100), (128, 129)]:
        assert f(n, d.ptr, lda, C.byref(info), None) == ooc.EARG
        assert "noalloc" in ooc.last_error()
    d.free()
