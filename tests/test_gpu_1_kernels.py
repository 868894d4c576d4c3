"""GPU parity of the tile kernels and generators, through the C ABI.

Each HIP kernel is compared with the CPU oracle / numpy on the same seeded
inputs: bit-exact for the generator (integer LCG), within FP64 rounding
bounds (stated per test) for the MFMA kernels.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _even(x):
    return x + (x & 1)


def _gemm_tol(k, A, B):
    # |error| of a k-term FP64 dot product <= k * eps * sum|a||b|
    return 4 * max(k, 1) * 2.0 ** -52 * (np.abs(A).max() * np.abs(B).max() * max(k, 1) + 1.0)


@pytest.mark.parametrize("m,n,k", [(128, 128, 16), (256, 128, 64), (200, 130, 48), (64, 50, 37),
                                   (1000, 300, 256), (1, 1, 1), (130, 257, 100)])
def test_dgemm_nt_matches_numpy(gpu, m, n, k):
    ooc = gpu
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    lda, ldb, ldc = _even(m), _even(n), _even(m)
    A = np.asfortranarray(rng.standard_normal((lda, k)))
    B = np.asfortranarray(rng.standard_normal((ldb, k)))
    C0 = np.asfortranarray(rng.standard_normal((ldc, n)))
    dA, dB, dC = (ooc.DeviceBuffer.from_array(x) for x in (A, B, C0))
    ooc.dgemm_nt(-1.0, dA, lda, dB, ldb, 0.5, dC, ldc, m, n, k)
    got = dC.to_array((ldc, n))
    ref = C0.copy()
    ref[:m, :n] = 0.5 * C0[:m, :n] - A[:m] @ B[:n].T
    assert np.max(np.abs(got - ref)) <= _gemm_tol(k, A, B)
    # rows/cols outside m x n untouched
    assert np.array_equal(got[m:], C0[m:])


def test_dgemm_nt_beta_zero_ignores_nan(gpu):
    ooc = gpu
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.standard_normal((128, 32)))
    C0 = np.full((128, 128), np.nan, order="F")
    dA, dC = ooc.DeviceBuffer.from_array(A), ooc.DeviceBuffer.from_array(C0)
    ooc.dgemm_nt(1.0, dA, 128, dA, 128, 0.0, dC, 128, 128, 128, 32)
    assert np.allclose(dC.to_array((128, 128)), A @ A.T, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,k", [(320, 96), (128, 1024), (77 * 2, 48)])
def test_dsyrk_ln_lower_only(gpu, n, k):
    ooc = gpu
    rng = np.random.default_rng(n + k)
    X = np.asfortranarray(rng.standard_normal((n, k)))
    S0 = np.asfortranarray(rng.standard_normal((n, n)))
    dX, dS = ooc.DeviceBuffer.from_array(X), ooc.DeviceBuffer.from_array(S0)
    ooc.dsyrk_ln(-1.0, dX, n, 1.0, dS, n, n, k)
    got = dS.to_array((n, n))
    assert np.max(np.abs(np.tril(got) - np.tril(S0 - X @ X.T))) <= _gemm_tol(k, X, X)
    assert np.array_equal(np.triu(got, 1), np.triu(S0, 1))


@pytest.mark.parametrize("n", [64, 256, 512])
def test_dpotrf_l_matches_oracle(gpu, ora, n):
    ooc = gpu
    A = ora.pdmatgen(n, 64)
    L = A.copy(order="F")
    assert ora.dpotrf(L) == 0
    dA = ooc.DeviceBuffer.from_array(A)
    assert ooc.dpotrf_l(dA, n, n) == 0
    got = dA.to_array((n, n))
    assert np.max(np.abs(np.tril(got) - np.tril(L))) / np.max(np.abs(L)) < 1e-13
    assert np.array_equal(np.triu(got, 1), np.triu(A, 1))


@pytest.mark.parametrize("bad", [0, 63, 64, 200])
def test_dpotrf_l_reports_first_bad_pivot(gpu, ora, bad):
    from scipy.linalg import lapack
    ooc = gpu
    A = ora.pdmatgen(256, 64)
    A[bad, bad] = -A[bad, bad]
    _, linfo = lapack.dpotrf(A.copy(order="F"), lower=1)
    dA = ooc.DeviceBuffer.from_array(A)
    assert ooc.dpotrf_l(dA, 256, 256) == linfo == bad + 1


@pytest.mark.parametrize("m,n", [(300, 128), (64, 64), (1024, 256)])
def test_dtrsm_rltn_matches_scipy(gpu, m, n):
    from scipy.linalg import solve_triangular
    ooc = gpu
    rng = np.random.default_rng(m + n)
    L = np.asfortranarray(np.tril(rng.standard_normal((n, n))) + n * np.eye(n))
    B = np.asfortranarray(rng.standard_normal((_even(m), n)))
    dL, dB = ooc.DeviceBuffer.from_array(L), ooc.DeviceBuffer.from_array(B)
    ooc.dtrsm_rltn(dL, n, dB, _even(m), m, n)
    got = dB.to_array(B.shape)
    ref = solve_triangular(L, B.T, lower=True).T
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-13


@pytest.mark.parametrize("n,nb,aform,diag,m", [
    (4, 2, "S", "D", None), (17, 5, "S", "D", None), (300, 64, "S", "N", None),
    (23, 4, "N", "D", 31), (23, 4, "T", "N", 31), (40, 16, "F", "D", None), (1000, 256, "H", "D", None),
])
def test_pdmatgen_bit_exact(gpu, ora, n, nb, aform, diag, m):
    assert np.array_equal(gpu.pdmatgen(n, nb, aform, diag, m=m),
                          ora.pdmatgen(n, nb, aform, diag, m=m))


def test_pdmatgen_matches_reference_fixture(gpu):
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "pdmatgen_golden.npz"))
    A = gpu.pdmatgen(4096, 256)
    idx = gold["S_D_4096_256_idx"]
    assert np.array_equal(A[idx[:, 0], idx[:, 1]], gold["S_D_4096_256_val"])
    assert np.array_equal(A.sum(axis=0), gold["S_D_4096_256_colsum"])


@pytest.mark.parametrize("n,nb,aform", [(1000, 256, "S"), (512, 128, "S"), (300, 64, "F"),
                                        (130, 64, "S")])
def test_pdmatgen_tiles_bit_exact(gpu, ora, n, nb, aform):
    ooc = gpu
    buf = ooc.DeviceBuffer(ooc.tiles_elems(n, nb) * 8)
    ooc.pdmatgen_tiles(buf, n, nb, aform, "D", 100)
    got = np.zeros((n, n), order="F")
    ooc.tiles_to_host(buf, got, nb)
    ref = ora.pdmatgen(n, nb, aform, "D")
    # tiles hold the lower triangle (plus the diagonal tiles' upper parts,
    # which tiles_to_host does not write back)
    assert np.array_equal(np.tril(got), np.tril(ref))
    # padding: rows/columns past n are 0 with 1 on the diagonal
    bb = -(-n // nb)
    if n % nb:
        tiles = buf.to_array((ooc.tiles_elems(n, nb),))
        last = tiles[-nb * nb:].reshape(nb, nb, order="F")
        v = n - (bb - 1) * nb
        pad = last.copy()
        pad[:v, :v] = 0
        assert np.array_equal(pad, np.diag((np.arange(nb) >= v).astype(float)))
