"""ctypes view of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests/, the smoke check and the benchmark's
CPU-baseline leg use this module as the checker.  The product library never
imports it.

Arrays are numpy float64 in column-major (Fortran) order, as the reference
stores A (oocdag.cpp:20, LDA = N at oocdag.cpp:251).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpdmatgen_ref.so")

_lib = None
_ref = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref/ where the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])


def build_ref() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_lcg_skip.restype = C.c_uint32
        L.ora_lcg_skip.argtypes = [C.c_uint32, C.c_uint64]
        L.ora_pdmatgen.restype = C.c_int
        L.ora_pdmatgen.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_int,
                                   _dp, C.c_int, C.c_int]
        L.ora_find_Yn.restype = C.c_int
        L.ora_find_Yn.argtypes = [C.c_int, C.c_int, C.c_int, _ip]
        L.ora_panel_plan.restype = C.c_int
        L.ora_panel_plan.argtypes = [C.c_int, C.c_int, C.c_size_t, _ip, C.c_int]
        L.ora_dpotrf_l.restype = C.c_int
        L.ora_dpotrf_l.argtypes = [C.c_int, _dp, C.c_int]
        L.ora_dtrsm_rltn.restype = None
        L.ora_dtrsm_rltn.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.ora_dsyrk_ln.restype = None
        L.ora_dsyrk_ln.argtypes = [C.c_int, C.c_int, C.c_double, _dp, C.c_int, C.c_double,
                                   _dp, C.c_int]
        L.ora_dgemm_nt.restype = None
        L.ora_dgemm_nt.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp,
                                   C.c_int, C.c_double, _dp, C.c_int]
        L.ora_set_backend.restype = C.c_int
        L.ora_set_backend.argtypes = [C.c_int]
        L.ora_get_backend.restype = C.c_int
        L.ora_set_threads.restype = None
        L.ora_set_threads.argtypes = [C.c_int]
        L.ora_get_threads.restype = C.c_int
        L.ora_cholesky.restype = C.c_double
        L.ora_cholesky.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_size_t, _ip]
        L.ora_llt_resid.restype = C.c_double
        L.ora_llt_resid.argtypes = [_dp, C.c_int, C.c_int, C.c_char, C.c_char, C.c_int, _dp]
        L.ora_llt_resid_explicit.restype = C.c_double
        L.ora_llt_resid_explicit.argtypes = [_dp, C.c_int, C.c_int, _dp, C.c_int]
        L.ora_dpotrs_l.restype = None
        L.ora_dpotrs_l.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.ora_solve_resid.restype = C.c_double
        L.ora_solve_resid.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.ora_norm_inf.restype = C.c_double
        L.ora_norm_inf.argtypes = [C.c_int, _dp, C.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(_dp)


def fmat(m: int, n: int) -> np.ndarray:
    return np.zeros((m, n), dtype=np.float64, order="F")


def lcg_skip(x: int, n: int) -> int:
    return int(lib().ora_lcg_skip(x, n))


def pdmatgen(n: int, nb: int = 1, aform: str = "S", diag: str = "D", iseed: int = 100,
             m: int | None = None) -> np.ndarray:
    m = n if m is None else m
    A = fmat(m, n)
    info = lib().ora_pdmatgen(aform.encode(), diag.encode(), m, n, nb, nb, _ptr(A), max(1, m), iseed)
    if info:
        raise ValueError(f"pdmatgen info={info}")
    return A


def find_Yn_sequence(bb: int, memBlock: int, jbs) -> list:
    st = C.c_int(0)
    return [lib().ora_find_Yn(bb, memBlock, jb, C.byref(st)) for jb in jbs]


def panel_plan(n: int, nb: int, memsize: int) -> list:
    w = (C.c_int * 4096)()
    np_ = lib().ora_panel_plan(n, nb, memsize, w, 4096)
    if np_ < 0:
        return None
    return [w[i] for i in range(np_)]


def set_backend(b: int) -> int:
    return lib().ora_set_backend(b)


def set_threads(t: int) -> None:
    lib().ora_set_threads(t)


def cholesky(A: np.ndarray, nb: int, memsize: int):
    """Run the restated driver in place on A (column-major).  Returns (seconds, info)."""
    info = C.c_int(0)
    n = A.shape[0]
    t = lib().ora_cholesky(_ptr(A), n, nb, A.shape[0], memsize, C.byref(info))
    return t, info.value


def dpotrf(A: np.ndarray) -> int:
    return lib().ora_dpotrf_l(A.shape[0], _ptr(A), A.shape[0])


def dtrsm_rltn(L: np.ndarray, B: np.ndarray) -> None:
    lib().ora_dtrsm_rltn(B.shape[0], B.shape[1], _ptr(L), L.shape[0], _ptr(B), B.shape[0])


def dsyrk_ln(alpha: float, X: np.ndarray, beta: float, Cm: np.ndarray) -> None:
    lib().ora_dsyrk_ln(Cm.shape[0], X.shape[1], alpha, _ptr(X), X.shape[0], beta, _ptr(Cm),
                       Cm.shape[0])


def dgemm_nt(alpha: float, A: np.ndarray, B: np.ndarray, beta: float, Cm: np.ndarray) -> None:
    lib().ora_dgemm_nt(Cm.shape[0], Cm.shape[1], A.shape[1], alpha, _ptr(A), A.shape[0],
                       _ptr(B), B.shape[0], beta, _ptr(Cm), Cm.shape[0])


def llt_resid(Af: np.ndarray, aform: str = "S", diag: str = "D", iseed: int = 100) -> float:
    an = C.c_double(0)
    return lib().ora_llt_resid(_ptr(Af), Af.shape[0], Af.shape[0], aform.encode(), diag.encode(),
                               iseed, C.byref(an))


def llt_resid_explicit(Af: np.ndarray, A0: np.ndarray) -> float:
    return lib().ora_llt_resid_explicit(_ptr(Af), Af.shape[0], Af.shape[0], _ptr(A0), A0.shape[0])


def dpotrs(L: np.ndarray, B: np.ndarray) -> None:
    lib().ora_dpotrs_l(L.shape[0], B.shape[1], _ptr(L), L.shape[0], _ptr(B), B.shape[0])


def solve_resid(A: np.ndarray, X: np.ndarray, B: np.ndarray) -> float:
    return lib().ora_solve_resid(A.shape[0], X.shape[1], _ptr(A), A.shape[0], _ptr(X), X.shape[0],
                                 _ptr(B), B.shape[0])


# ---- the reference's own Fortran generator (built only where the reference is) ----

def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_pdmatgen(n: int, nb: int, aform: str = "S", diag: str = "D", iseed: int = 100,
                 m: int | None = None) -> np.ndarray:
    """Call PDMATGEN (LIN/pdmatgen.f, BIND(C)) exactly as oocdag.cpp:298-302 does."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH, mode=C.RTLD_LOCAL)
    m = n if m is None else m
    A = fmat(m, n)
    ci = lambda v: C.byref(C.c_int(v))
    af = C.create_string_buffer(aform.encode())
    dg = C.create_string_buffer(diag.encode())
    _ref.pdmatgen(ci(0), af, dg, ci(m), ci(n), ci(nb), ci(nb), _ptr(A), ci(max(1, m)), ci(0), ci(0),
                  ci(iseed), ci(0), ci(m), ci(0), ci(n), ci(0), ci(0), ci(1), ci(1))
    return A
