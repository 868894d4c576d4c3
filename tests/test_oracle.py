"""CPU tests that pin the oracle (oracle/oracle.c) before it is trusted as
the checker of the GPU path.

Each check is chosen so a plausible mistake fails it: generator against the
reference's own Fortran PDMATGEN (golden fixtures and, where built, the live
build), planner against a brute-force maximality check, factorization against
LAPACK via numpy/scipy and against the closed form of the AFORM='F' matrix,
residual against an independent numpy restatement and a planted error.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pdmatgen_golden.npz")
EPS = 2.0 ** -53


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


# ---------------------------------------------------------------- generator

@pytest.mark.parametrize("key,args", [
    ("S_D_4_2", dict(n=4, nb=2)),
    ("S_D_17_5", dict(n=17, nb=5)),
    ("S_D_64_16", dict(n=64, nb=16)),
    ("S_N_17_5", dict(n=17, nb=5, diag="N")),
    ("N_N_11x7_3", dict(n=7, nb=3, aform="N", diag="N", m=11)),
    ("T_N_11x7_3", dict(n=7, nb=3, aform="T", diag="N", m=11)),
    ("F_D_9_4", dict(n=9, nb=4, aform="F")),
    ("S_D_9_4_seed7", dict(n=9, nb=4, iseed=7)),
])
def test_pdmatgen_matches_reference_fixtures(ora, gold, key, args):
    got = ora.pdmatgen(**args)
    assert np.array_equal(got, gold[key])


def test_pdmatgen_4096_samples(ora, gold):
    A = ora.pdmatgen(4096, 256)
    idx = gold["S_D_4096_256_idx"]
    assert np.array_equal(A[idx[:, 0], idx[:, 1]], gold["S_D_4096_256_val"])
    assert np.array_equal(A.sum(axis=0), gold["S_D_4096_256_colsum"])


def test_pdmatgen_hand_derived_values(ora):
    # x1..x6 of x_{k+1} = (1103515245 x_k + 12345) mod 2^31 from ISEED = 100,
    # and A(0,0), A(1,0), A(1,1), A(3,3) of the 4x4 'S','D' matrix (worked by
    # hand from LIN/pdmatgen.f:210-303,516-543)
    xs = [ora.lcg_skip(100, k) for k in range(1, 7)]
    assert xs == [829870797, 1533044610, 1478614675, 1357823696, 413847241, 70351310]
    A = ora.pdmatgen(4, 2)
    assert A[0, 0] == 4.2271225927397609
    assert A[1, 0] == -0.42775905318558216
    assert A[1, 1] == 4.9344802368432283
    assert A[3, 3] == 4.2720769457519054


def test_lcg_full_period_and_composition(ora):
    # a = 1 mod 4 and c odd: full period 2^31 (Hull-Dobell)
    for x in (0, 100, 2**31 - 1):
        assert ora.lcg_skip(x, 2**31) == x
    assert ora.lcg_skip(ora.lcg_skip(100, 12345678), 987654321) == ora.lcg_skip(100, 999999999)


@pytest.mark.skipif(not __import__("oracle").ref_available(),
                    reason="reference generator build (oracle/_ref) not present")
@pytest.mark.parametrize("n,nb,aform,diag,m", [
    (33, 8, "S", "D", None), (100, 7, "S", "N", None), (50, 50, "H", "D", None),
    (23, 4, "N", "D", 31), (23, 4, "T", "N", 31), (40, 16, "F", "D", None),
])
def test_pdmatgen_matches_live_reference_build(ora, n, nb, aform, diag, m):
    assert np.array_equal(ora.pdmatgen(n, nb, aform, diag, m=m),
                          ora.ref_pdmatgen(n, nb, aform, diag, m=m))


def test_pdmatgen_argument_errors(ora):
    import ctypes as C
    A = np.zeros((4, 4), order="F")
    p = A.ctypes.data_as(C.POINTER(C.c_double))
    L = ora.lib()
    assert L.ora_pdmatgen(b"S", b"X", 4, 4, 2, 2, p, 4, 100) == 3
    assert L.ora_pdmatgen(b"S", b"D", 4, 3, 2, 2, p, 4, 100) == 5
    assert L.ora_pdmatgen(b"S", b"D", 4, 4, 2, 3, p, 4, 100) == 7


# ---------------------------------------------------------------- planner

# schedules of find_Yn (oocdag.cpp:478-492) for the configurations in use
PLAN_TABLE = [
    (4096, 256, 28 << 20, [2, 3, 5, 6]),
    (4096, 256, 32 << 20, [3, 4, 9]),
    (196608, 1024, 64 * 10**9, [43, 63, 86]),
    (196608, 1024, 160 << 30, [192]),
    (4096, 256, 1 << 30, [16]),
]


@pytest.mark.parametrize("n,nb,mem,widths", PLAN_TABLE)
def test_find_Yn_table(ora, n, nb, mem, widths):
    assert ora.panel_plan(n, nb, mem) == widths


@pytest.mark.parametrize("bb,memBlock", [(16, 32), (16, 56), (16, 64), (31, 80), (192, 7629),
                                         (7, 14), (40, 1000)])
def test_find_Yn_maximal_and_feasible(ora, bb, memBlock):
    T = lambda y, ym: y * ym - y * (y - 1) // 2
    jb, prev = 0, 0
    seq = []
    while jb < bb:
        seq.append(jb)
        ym = bb - jb
        # brute force: widest y (never below the previous width, the
        # function-static) whose trapezoid fits next to one X column
        fits = [y for y in range(1, ym + 1) if T(y, ym) <= memBlock - ym]
        best = ym if ym * (ym + 1) <= 2 * (memBlock - ym) else max([prev] + fits)
        got = ora.find_Yn_sequence(bb, memBlock, seq)[-1]
        assert got == best
        prev = got
        jb += got


def test_budget_too_small(ora):
    A = ora.pdmatgen(256, 64)
    t, info = ora.cholesky(A, 64, 64 * 64 * 8 * 7)  # memBlock 7 < 2*bb = 8
    assert info == -2


# ---------------------------------------------------------------- factor

def closed_form_F(n):
    # A = d I + c 11^T with c = (double)0.1f, d = 10 - c: L_kk = sqrt(d + c_k),
    # L_ik = c_k / L_kk (i > k), c_k = c d / (d + k c)
    c = float(np.float32(0.1))
    d = 10.0 - c
    L = np.zeros((n, n))
    for k in range(n):
        ck = c * d / (d + k * c)
        lkk = np.sqrt(d + ck)
        L[k, k] = lkk
        L[k + 1:, k] = ck / lkk
    return L


@pytest.mark.parametrize("n,nb,mem_tiles", [(200, 64, 10), (256, 64, 20), (97, 32, 40)])
def test_cholesky_F_closed_form(ora, n, nb, mem_tiles):
    A = ora.pdmatgen(n, nb, "F")
    A0 = A.copy(order="F")
    t, info = ora.cholesky(A, nb, mem_tiles * nb * nb * 8)
    assert info == 0
    L = np.tril(A)
    assert np.max(np.abs(L - closed_form_F(n))) < 1e-13
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))


@pytest.mark.parametrize("n,nb,mem_tiles", [(300, 64, 12), (512, 128, 8), (640, 64, 40)])
def test_cholesky_matches_lapack(ora, n, nb, mem_tiles):
    A = ora.pdmatgen(n, nb)
    A0 = A.copy(order="F")
    t, info = ora.cholesky(A, nb, mem_tiles * nb * nb * 8)
    assert info == 0
    Ln = np.linalg.cholesky(A0)
    assert np.max(np.abs(np.tril(A) - Ln)) / np.max(np.abs(Ln)) < 1e-13
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))


def test_cholesky_gram_matrix_lapack(ora):
    rng = np.random.default_rng(3)
    n = 256
    G = rng.standard_normal((n, n))
    A = np.asfortranarray(G @ G.T / n + 1e-2 * np.eye(n))
    A0 = A.copy(order="F")
    t, info = ora.cholesky(A, 64, 10 * 64 * 64 * 8)
    assert info == 0
    Ln = np.linalg.cholesky(A0)
    assert np.max(np.abs(np.tril(A) - Ln)) / np.max(np.abs(Ln)) < 1e-11


def test_cholesky_mkl_backend_agrees(ora):
    if ora.set_backend(1) != 1:
        pytest.skip("MKL runtime not loadable")
    try:
        A = ora.pdmatgen(512, 128)
        B = A.copy(order="F")
        _, i1 = ora.cholesky(A, 128, 8 * 128 * 128 * 8)
        ora.set_backend(0)
        _, i2 = ora.cholesky(B, 128, 8 * 128 * 128 * 8)
        assert i1 == i2 == 0
        assert np.max(np.abs(np.tril(A) - np.tril(B))) / np.max(np.abs(A)) < 1e-13
    finally:
        ora.set_backend(0)


@pytest.mark.parametrize("bad", [0, 70, 299])
def test_nonspd_info_matches_lapack(ora, bad):
    from scipy.linalg import lapack
    A = ora.pdmatgen(300, 64)
    A[bad, bad] = -A[bad, bad]
    _, linfo = lapack.dpotrf(A.copy(order="F"), lower=1)
    t, info = ora.cholesky(A, 64, 12 * 64 * 64 * 8)
    assert info == linfo == bad + 1


# ---------------------------------------------------------------- residuals

def numpy_llt_resid(Af, A0):
    n = A0.shape[0]
    L = np.tril(Af)
    R = np.tril(L @ L.T) + np.triu(Af, 1)
    return np.max(np.abs(R - A0).sum(axis=1)) / (n * EPS * np.max(np.abs(A0).sum(axis=1)))


def test_llt_resid_matches_numpy_and_detects_errors(ora):
    n, nb = 300, 64
    A = ora.pdmatgen(n, nb)
    A0 = A.copy(order="F")
    ora.cholesky(A, nb, 12 * nb * nb * 8)
    r = ora.llt_resid(A)
    # both are sums of rounding errors: equal up to summation order
    assert r == pytest.approx(numpy_llt_resid(A, A0), rel=0.05)
    assert r < 1.0
    # planted error: one entry of L off by 1e-11 relative must be visible
    B = A.copy(order="F")
    B[n - 1, 0] += 1e-11 * abs(B[n - 1, 0]) + 1e-11
    rb = ora.llt_resid(B)
    assert rb == pytest.approx(numpy_llt_resid(B, A0), rel=1e-2)
    assert rb > 2 * r
    # a change to the strict upper triangle counts too (PDPOTRRV keeps it)
    C_ = A.copy(order="F")
    C_[0, n - 1] += 1.0
    assert ora.llt_resid(C_) > 1e10


def test_solve_and_solve_resid(ora):
    n = 256
    A = ora.pdmatgen(n, 64)
    A0 = A.copy(order="F")
    ora.cholesky(A, 64, 12 * 64 * 64 * 8)
    rng = np.random.default_rng(0)
    B = np.asfortranarray(rng.standard_normal((n, 5)))
    X = B.copy(order="F")
    ora.dpotrs(A, X)
    assert np.max(np.abs(X - np.linalg.solve(A0, B))) < 1e-12
    r = ora.solve_resid(A0, X, B)
    assert r < 10

    def np_resid(Xc):
        Rn = B - A0 @ Xc
        return max(np.abs(Rn[:, j]).max() /
                   (np.abs(Xc[:, j]).max() * np.abs(A0).sum(1).max() * EPS * n) for j in range(5))

    # rounding-level residuals agree up to summation order ...
    assert r == pytest.approx(np_resid(X), rel=0.2)
    # ... and a planted error dominates and must be reproduced exactly
    Xp = X.copy(order="F")
    Xp[7, 3] += 1e-7
    assert ora.solve_resid(A0, Xp, B) == pytest.approx(np_resid(Xp), rel=1e-6)
    # VERIFY identity of oocdag.cpp:379-464: dpotrs(L, A) = I
    I = A0.copy(order="F")
    ora.dpotrs(A, I)
    assert np.max(np.abs(I - np.eye(n))) < 1e-12


def test_tile_kernels_match_numpy(ora):
    rng = np.random.default_rng(1)
    m, n, k = 70, 50, 30
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    C0 = np.asfortranarray(rng.standard_normal((m, n)))
    Cm = C0.copy(order="F")
    ora.dgemm_nt(-1.0, A, B, 1.0, Cm)
    assert np.allclose(Cm, C0 - A @ B.T, rtol=1e-13, atol=1e-13)
    X = np.asfortranarray(rng.standard_normal((n, k)))
    S0 = np.asfortranarray(rng.standard_normal((n, n)))
    S = S0.copy(order="F")
    ora.dsyrk_ln(-1.0, X, 1.0, S)
    assert np.allclose(np.tril(S), np.tril(S0 - X @ X.T), rtol=1e-13, atol=1e-13)
    assert np.array_equal(np.triu(S, 1), np.triu(S0, 1))
    from scipy.linalg import solve_triangular
    Lt = np.asfortranarray(np.tril(rng.standard_normal((n, n))) + 5 * np.eye(n))
    Bt = np.asfortranarray(rng.standard_normal((m, n)))
    Xt = Bt.copy(order="F")
    ora.dtrsm_rltn(Lt, Xt)
    assert np.allclose(Xt, solve_triangular(Lt, Bt.T, lower=True).T, rtol=1e-12, atol=1e-12)
