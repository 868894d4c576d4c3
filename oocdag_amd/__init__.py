"""oocdag_amd -- MI355X out-of-core left-looking tiled Cholesky.

Python view of liboocdag.so (the C ABI declared in include/oocdag.h).  It
mirrors the reference entry point ``double Cholesky(Quark*, double *A, int N,
int NB, int LDA, size_t memsize)`` (oocdag.cpp:18) as :func:`cholesky`, the
generator ``pdmatgen`` (LIN/pdmatgen.f, oocdag.cpp:298-302) as
:func:`pdmatgen`, and exposes the HBM-resident tile path and the tile
kernels.

The library is HIP-only: every call runs on the GPU, and a missing or
unbuildable library raises instead of falling back to any CPU code.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboocdag.so")

# The out-of-core factorization drives four HIP streams (host->device,
# critical path, bulk update, device->host) beside HIP's null stream.  HIP
# maps streams onto a pool of GPU_MAX_HW_QUEUES hardware queues (4 by
# default); with fewer queues than busy streams two of them share one and the
# device->host copy of a finished column then holds back the next bulk
# update.  Unset or below 5, the pool is raised to 8 here.  HIP reads the
# variable once, when it initializes: a process that already initialized HIP
# keeps its pool.
_q = os.environ.get("GPU_MAX_HW_QUEUES", "")
if not _q.isdigit() or int(_q) < 5:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

OK, EARG, EBUDGET, ENOMEM, EHIP, ENODEV = 0, -1, -2, -3, -4, -5

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None

# name -> (restype, argtypes); every symbol declared in include/oocdag.h
SIGNATURES = {
    "oocdag_cholesky": (C.c_double, [_dp, C.c_int, C.c_int, C.c_int, C.c_size_t, _ip]),
    "oocdag_cholesky_host_tiles": (C.c_double, [C.c_void_p, C.c_int, C.c_int, C.c_size_t, _ip]),
    "oocdag_cholesky_nx": (C.c_double, [_dp, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_int, _ip]),
    "oocdag_cholesky_host_tiles_nx": (C.c_double, [C.c_void_p, C.c_int, C.c_int, C.c_size_t,
                                                   C.c_int, _ip]),
    "oocdag_tile_edge": (C.c_int, [C.c_int]),
    "oocdag_pdmatgen_host_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char, C.c_char,
                                             C.c_int]),
    "oocdag_host_alloc": (C.c_void_p, [C.c_size_t]),
    "oocdag_host_free": (C.c_int, [C.c_void_p]),
    "oocdag_panel_plan": (C.c_int, [C.c_int, C.c_int, C.c_size_t, _ip, C.c_int]),
    "oocdag_panel_plan_nx": (C.c_int, [C.c_int, C.c_int, C.c_size_t, C.c_int, _ip, C.c_int]),
    "oocdag_pdmatgen": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                  C.c_int, C.c_int]),
    "oocdag_potrf_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, _ip]),
    "oocdag_tiles_elems": (C.c_size_t, [C.c_int, C.c_int]),
    "oocdag_pdmatgen_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char, C.c_char,
                                        C.c_int, C.c_void_p]),
    "oocdag_tiles_from_host": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int]),
    "oocdag_tiles_to_host": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int]),
    "oocdag_dgemm_nt": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int,
                                  C.c_void_p]),
    "oocdag_dsyrk_ln": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_double,
                                  C.c_void_p, C.c_int, C.c_void_p]),
    "oocdag_dtrsm_rltn": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                    C.c_void_p]),
    "oocdag_dpotrf_l": (C.c_int, [C.c_int, C.c_void_p, C.c_int, _ip, C.c_void_p]),
    "oocdag_dpotrf_l_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_int, _ip, C.c_void_p]),
    "oocdag_potrs": (C.c_int, [_dp, C.c_int, C.c_int, _dp, C.c_int, C.c_int]),
    "oocdag_potrs_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                     C.c_void_p]),
    "oocdag_llt_resid": (C.c_int, [_dp, C.c_int, C.c_int, C.c_char, C.c_char, C.c_int, _dp]),
    "oocdag_llt_resid_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char, C.c_char,
                                         C.c_int, _dp]),
    "oocdag_solve_resid": (C.c_int, [_dp, C.c_int, C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_int,
                                     _dp]),
    "oocdag_solve_resid_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                           C.c_int, C.c_void_p, C.c_int, _dp]),
    "oocdag_dev_alloc": (C.c_void_p, [C.c_size_t]),
    "oocdag_dev_free": (C.c_int, [C.c_void_p]),
    "oocdag_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "oocdag_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "oocdag_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "oocdag_memset_d": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "oocdag_device_sync": (C.c_int, []),
    "oocdag_device_count": (C.c_int, []),
    "oocdag_last_error": (C.c_char_p, []),
    "oocdag_last_stats": (C.c_int, [_dp, _dp, _dp, _dp, _dp]),
    "oocdag_profile_enable": (C.c_int, [C.c_int]),
    "oocdag_profile_read": (C.c_int, [C.POINTER(C.c_longlong), _dp, _dp]),
    "oocdag_profile_bytes": (C.c_int, [_dp]),
    "oocdag_set_progress": (C.c_int, [C.c_int]),
    "oocdag_release_staging": (C.c_int, []),
}


class MpirStats(C.Structure):
    _fields_ = [("precision_used", C.c_int), ("fell_back", C.c_int), ("iterations", C.c_int),
                ("info", C.c_int), ("backward_error", C.c_double), ("factor_seconds", C.c_double),
                ("solve_seconds", C.c_double), ("total_seconds", C.c_double),
                ("solve_resid", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES.update({
    "oocdag_mpir_solve_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                          C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                          C.POINTER(MpirStats)]),
    "oocdag_gram_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_ulonglong,
                                    C.c_void_p]),
    "oocdag_normal_fill": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_ulonglong,
                                     C.c_ulonglong]),
})

class BatchStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("systems_per_second", C.c_double),
                ("max_backward_error", C.c_double), ("failed", C.c_int), ("variant_T", C.c_int),
                ("variant_RB", C.c_int), ("tune_seconds", C.c_double),
                ("max_solve_resid", C.c_double), ("split_min_n", C.c_int),
                ("split_systems", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_llp = C.POINTER(C.c_longlong)
SIGNATURES.update({
    "oocdag_batch_layout": (C.c_longlong, [C.c_int, _ip, _ip, _llp, _llp]),
    "oocdag_batch_generate": (C.c_int, [C.c_int, _ip, _ip, _llp, _llp, C.c_void_p, C.c_void_p,
                                        C.c_double, C.c_ulonglong]),
    "oocdag_batch_set_variant": (C.c_int, [C.c_int, C.c_int]),
    "oocdag_batch_set_split": (C.c_int, [C.c_int, C.c_int]),
    "oocdag_mp_set_update_variant": (C.c_int, [C.c_int]),
    "oocdag_batched_solve": (C.c_int, [C.c_int, _ip, _ip, _llp, _llp, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp,
                                       _ip, _ip, C.POINTER(BatchStats)]),
})

class Request(C.Structure):
    _fields_ = [("matrix_id", C.c_longlong), ("n", C.c_int), ("nrhs", C.c_int), ("mu", C.c_double),
                ("seed", C.c_ulonglong), ("rhs_seed", C.c_ulonglong), ("precision", C.c_int),
                ("tol", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("route", C.c_int), ("cache_hit", C.c_int),
                ("batch_size", C.c_int), ("precision_used", C.c_int), ("fell_back", C.c_int),
                ("iterations", C.c_int), ("backward_error", C.c_double),
                ("queue_seconds", C.c_double), ("service_seconds", C.c_double),
                ("latency_seconds", C.c_double), ("x_checksum", C.c_double),
                ("input_seconds", C.c_double), ("solve_resid", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES.update({
    "oocdag_server_create": (C.c_void_p, [C.c_size_t, C.c_int]),
    "oocdag_server_submit": (C.c_longlong, [C.c_void_p, C.POINTER(Request)]),
    "oocdag_server_wait": (C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(Result), _dp, C.c_size_t]),
    "oocdag_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), _ip, _ip, _ip]),
    "oocdag_server_destroy": (C.c_int, [C.c_void_p]),
    "oocdag_server_estimate": (C.c_int, [C.POINTER(Request), C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_size_t)]),
})

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}


class OocdagError(RuntimeError):
    pass


def lib():
    """Load liboocdag.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OocdagError(f"{LIB_PATH} is not built: run python -m oocdag_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().oocdag_last_error().decode()


def _check(rc: int, what: str) -> None:
    if rc != OK:
        raise OocdagError(f"{what} failed ({rc}): {last_error()}")


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(_dp)


def device_count() -> int:
    return lib().oocdag_device_count()


# ---------------------------------------------------------------------------
# reference entry points
# ---------------------------------------------------------------------------

def cholesky(A: np.ndarray, nb: int, memsize: int, nx: int | None = None):
    """Cholesky(A, N, NB, LDA, memsize) (oocdag.cpp:18): factors A in place.

    Returns (seconds, info) like the reference returns seconds; info follows
    LAPACK dpotrf (k > 0: leading minor k not positive definite).  nx fixes
    the staged X tile columns (1: the reference's find_Yn schedule)."""
    info = C.c_int(0)
    n = A.shape[1]
    if nx is None:
        t = lib().oocdag_cholesky(_ptr(A), n, nb, A.shape[0], memsize, C.byref(info))
    else:
        t = lib().oocdag_cholesky_nx(_ptr(A), n, nb, A.shape[0], memsize, nx, C.byref(info))
    if info.value < 0:
        raise OocdagError(f"oocdag_cholesky failed ({info.value}): {last_error()}")
    return t, info.value


class HostTiles:
    """Packed lower tiles of an n x n matrix in pinned host memory (the
    layout of the device tiles): the host side of
    :func:`cholesky_host_tiles`."""

    def __init__(self, n: int, nb: int, pinned: bool | None = True):
        """pinned=True: page-locked (error if refused); False: pageable;
        None: page-locked if the system grants it, else pageable."""
        self.n, self.nb = n, nb
        self.elems = tiles_elems(n, nb)
        self._np = None
        self.ptr = lib().oocdag_host_alloc(self.elems * 8) if pinned is not False else None
        self.pinned = bool(self.ptr)
        if not self.ptr:
            if pinned:
                raise OocdagError(f"pinned host allocation failed: {last_error()}")
            self._np = np.empty(self.elems, dtype=np.float64)
            self.ptr = self._np.ctypes.data

    def array(self) -> np.ndarray:
        """A numpy view of the tiles (no copy; valid until :meth:`free`)."""
        if self._np is not None:
            return self._np
        buf = (C.c_double * self.elems).from_address(self.ptr)
        return np.ctypeslib.as_array(buf)

    def generate(self, aform: str = "S", diag: str = "D", iseed: int = 100) -> None:
        _check(lib().oocdag_pdmatgen_host_tiles(self.ptr, self.n, self.nb, aform.encode(),
                                                diag.encode(), iseed), "pdmatgen_host_tiles")

    def free(self) -> None:
        if self.ptr and self._np is None:
            lib().oocdag_host_free(self.ptr)
        self.ptr = None
        self._np = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def cholesky_host_tiles(t: "HostTiles", memsize: int, nx: int = 0):
    """Out-of-core factorization of host packed tiles in place
    (Cholesky(), oocdag.cpp:18, on the tile layout).  Returns (seconds, info)."""
    info = C.c_int(0)
    secs = lib().oocdag_cholesky_host_tiles_nx(t.ptr, t.n, t.nb, memsize, nx, C.byref(info))
    if info.value < 0:
        raise OocdagError(f"oocdag_cholesky_host_tiles failed ({info.value}): {last_error()}")
    return secs, info.value


def tile_edge(nb: int) -> int:
    """Edge of the slot that holds one nb x nb tile (nb rounded up to 64)."""
    return int(lib().oocdag_tile_edge(nb))


def tiles_to_dense(tiles: np.ndarray, n: int, nb: int) -> np.ndarray:
    """Lower triangle of the matrix held in packed tiles, as a dense array."""
    bb = -(-n // nb)
    ts = -(-nb // 64) * 64
    out = np.zeros((bb * nb, bb * nb), order="F")
    k = 0
    for j in range(bb):
        for i in range(j, bb):
            out[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb] = \
                tiles[k * ts * ts:(k + 1) * ts * ts].reshape(ts, ts, order="F")[:nb, :nb]
            k += 1
    return np.tril(out[:n, :n])


def panel_plan(n: int, nb: int, memsize: int, nx: int | None = None):
    w = (C.c_int * 8192)()
    if nx is None:
        k = lib().oocdag_panel_plan(n, nb, memsize, w, 8192)
    else:
        k = lib().oocdag_panel_plan_nx(n, nb, memsize, nx, w, 8192)
    if k < 0:
        return None
    return [w[i] for i in range(k)]


def pdmatgen(n: int, nb: int = 1, aform: str = "S", diag: str = "D", iseed: int = 100,
             m: int | None = None) -> np.ndarray:
    """pdmatgen(...) of oocdag.cpp:298-302 (1x1 grid), computed on the GPU."""
    m = n if m is None else m
    A = np.zeros((m, n), dtype=np.float64, order="F")
    info = lib().oocdag_pdmatgen(aform.encode(), diag.encode(), m, n, nb, nb, _ptr(A),
                                 max(1, m), iseed)
    if info != 0:
        raise OocdagError(f"pdmatgen info={info}: {last_error()}")
    return A


def potrs(L: np.ndarray, B: np.ndarray) -> None:
    """dpotrs('L') with the factor left by :func:`cholesky` (oocdag.cpp:439);
    B (N x nrhs, Fortran order) is overwritten with the solution."""
    _check(lib().oocdag_potrs(_ptr(L), L.shape[1], L.shape[0], _ptr(B), B.shape[1], B.shape[0]),
           "potrs")


def llt_resid(A: np.ndarray, aform: str = "S", diag: str = "D", iseed: int = 100) -> float:
    """PDPOTRRV + PDLAFCHK scaled residual of the factor left in A."""
    r = C.c_double(0)
    _check(lib().oocdag_llt_resid(_ptr(A), A.shape[1], A.shape[0], aform.encode(), diag.encode(),
                                  iseed, C.byref(r)), "llt_resid")
    return r.value


def solve_resid(A: np.ndarray, X: np.ndarray, B: np.ndarray) -> float:
    """PDLASCHK residual max_k ||A x_k - b_k|| / (||A|| ||x_k|| eps N) on the
    GPU; the lower triangle of A defines the symmetric matrix."""
    r = C.c_double(0)
    _check(lib().oocdag_solve_resid(_ptr(A), A.shape[1], A.shape[0], _ptr(X), X.shape[1],
                                    X.shape[0], _ptr(B), B.shape[0], C.byref(r)),
           "solve_resid")
    return r.value


def solve_resid_tiles(buf: "DeviceBuffer", n: int, nb: int, X: "DeviceBuffer", B: "DeviceBuffer",
                      nrhs: int, ldx: int, ldb: int) -> float:
    """PDLASCHK residual with A in device packed tiles (full diagonal tiles)."""
    r = C.c_double(0)
    _check(lib().oocdag_solve_resid_tiles(buf.ptr, n, nb, X.ptr, nrhs, ldx, B.ptr, ldb,
                                          C.byref(r)), "solve_resid_tiles")
    return r.value


def last_stats() -> dict:
    v = [C.c_double(0) for _ in range(5)]
    lib().oocdag_last_stats(*[C.byref(x) for x in v])
    keys = ["seconds", "h2d_bytes", "d2h_bytes", "h2d_seconds", "d2h_seconds"]
    return {k: x.value for k, x in zip(keys, v)}


def profile_enable(on: bool = True) -> None:
    lib().oocdag_profile_enable(1 if on else 0)


def profile_read() -> dict:
    n, t, f = C.c_longlong(0), C.c_double(0), C.c_double(0)
    b = C.c_double(0)
    lib().oocdag_profile_read(C.byref(n), C.byref(t), C.byref(f))
    lib().oocdag_profile_bytes(C.byref(b))
    return {"launches": n.value, "seconds": t.value, "flops": f.value, "bytes": b.value}


def set_progress(every_columns: int) -> None:
    lib().oocdag_set_progress(int(every_columns))


# ---------------------------------------------------------------------------
# device buffers and the HBM-resident path
# ---------------------------------------------------------------------------

class DeviceBuffer:
    """A hipMalloc'ed block owned by Python."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = lib().oocdag_dev_alloc(self.nbytes)
        if not p:
            raise OocdagError(f"device allocation of {nbytes} bytes failed: {last_error()}")
        self.ptr = p

    def free(self):
        if self.ptr:
            lib().oocdag_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.asfortranarray(a, dtype=np.float64)
        b = cls(a.nbytes)
        _check(lib().oocdag_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes), "h2d")
        return b

    def to_array(self, shape, order="F") -> np.ndarray:
        out = np.empty(shape, dtype=np.float64, order=order)
        _check(lib().oocdag_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out


def tiles_elems(n: int, nb: int) -> int:
    return int(lib().oocdag_tiles_elems(n, nb))


def pdmatgen_tiles(buf: DeviceBuffer, n: int, nb: int, aform="S", diag="D", iseed=100):
    _check(lib().oocdag_pdmatgen_tiles(buf.ptr, n, nb, aform.encode(), diag.encode(), iseed, None),
           "pdmatgen_tiles")


def potrf_tiles(buf: DeviceBuffer, n: int, nb: int) -> int:
    info = C.c_int(0)
    _check(lib().oocdag_potrf_tiles(buf.ptr, n, nb, None, C.byref(info)), "potrf_tiles")
    return info.value


def potrs_tiles(buf: DeviceBuffer, n: int, nb: int, B: DeviceBuffer, nrhs: int, ldb: int):
    _check(lib().oocdag_potrs_tiles(buf.ptr, n, nb, B.ptr, nrhs, ldb, None), "potrs_tiles")


def llt_resid_tiles(buf: DeviceBuffer, n: int, nb: int, aform="S", diag="D", iseed=100) -> float:
    r = C.c_double(0)
    _check(lib().oocdag_llt_resid_tiles(buf.ptr, n, nb, aform.encode(), diag.encode(), iseed,
                                        C.byref(r)), "llt_resid_tiles")
    return r.value


def gram_tiles(buf: DeviceBuffer, n: int, nb: int, mu: float, seed: int = 1):
    _check(lib().oocdag_gram_tiles(buf.ptr, n, nb, mu, seed, None), "gram_tiles")


def normal_fill(buf: DeviceBuffer, rows: int, cols: int, ld: int, seed: int = 1, stream_id: int = 2):
    _check(lib().oocdag_normal_fill(buf.ptr, rows, cols, ld, seed, stream_id), "normal_fill")


def mpir_solve_tiles(A: DeviceBuffer, n: int, nb: int, B: DeviceBuffer, nrhs: int,
                     X: DeviceBuffer, precision="fp32", tol: float = None, max_iter: int = 30):
    """Mixed-precision factor + FP64 iterative refinement; returns the stats dict."""
    st = MpirStats()
    if tol is None:
        tol = 2.0 ** -53 * 8
    _check(lib().oocdag_mpir_solve_tiles(A.ptr, n, nb, B.ptr, nrhs, n, X.ptr, n,
                                         PRECISIONS.get(precision, precision), tol, max_iter,
                                         C.byref(st)), "mpir_solve_tiles")
    return st.as_dict()


def mp_set_update_variant(variant: int) -> int:
    """Choose how the mixed-precision factor runs its low-precision updates,
    as bits: 1 = LDS-staged FP32 kernel (else operands straight into MFMA
    registers), 2 = bulk update serialized after the next super-column (else
    overlapped on a second stream), 4 = the FP32 route's super-column updates
    on FP32 MFMA (else from three BF16 pieces per element); 0 is the
    default.  Returns the number of variants."""
    rc = lib().oocdag_mp_set_update_variant(variant)
    if rc < 0:
        raise OocdagError(f"mp_set_update_variant({variant}) failed")
    return rc


class Batch:
    """A batch of small SPD systems in device arenas (oocdag_batch_layout)."""

    def __init__(self, n, nrhs):
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
        k = len(self.n)
        self.a_off = np.zeros(k, dtype=np.int64)
        self.b_off = np.zeros(k + 1, dtype=np.int64)
        total_a = lib().oocdag_batch_layout(k, self._i(self.n), self._i(self.nrhs),
                                            self._l(self.a_off), self._l(self.b_off))
        self.A = DeviceBuffer(max(total_a, 1) * 8)
        self.B = DeviceBuffer(max(int(self.b_off[k]), 1) * 8)
        self.X = DeviceBuffer(max(int(self.b_off[k]), 1) * 8)

    @staticmethod
    def _i(a):
        return a.ctypes.data_as(_ip)

    @staticmethod
    def _l(a):
        return a.ctypes.data_as(_llp)

    def ld(self, s):
        return int(self.n[s] + (self.n[s] & 1))

    def generate(self, mu=0.1, seed=1):
        _check(lib().oocdag_batch_generate(len(self.n), self._i(self.n), self._i(self.nrhs),
                                           self._l(self.a_off), self._l(self.b_off), self.A.ptr,
                                           self.B.ptr, mu, seed), "batch_generate")

    @staticmethod
    def set_variant(precision: str, variant: int) -> int:
        """Fix the batched kernel variant (-1: autotune); returns the count."""
        k = {"fp64": 0, "fp32": 1}[precision]
        rc = lib().oocdag_batch_set_variant(k, variant)
        if rc < 0:
            raise OocdagError(f"batch_set_variant({precision}, {variant}) failed")
        return rc

    @staticmethod
    def set_split(precision: str, min_n: int) -> None:
        """Factor the systems with n >= min_n by block-column launches (0:
        none, -1: autotune the threshold)."""
        k = {"fp64": 0, "fp32": 1}[precision]
        if lib().oocdag_batch_set_split(k, min_n) != 0:
            raise OocdagError(f"batch_set_split({precision}, {min_n}) failed")

    def solve(self, precision="fp64", tol=None, max_iter=10, reps=1, with_resid=False):
        """Returns (stats, berr, iters, status), plus the per-system PDLASCHK
        residual as a fifth item when with_resid."""
        k = len(self.n)
        berr = np.zeros(k)
        lres = np.zeros(k)
        iters = np.zeros(k, dtype=np.int32)
        status = np.zeros(k, dtype=np.int32)
        st = BatchStats()
        if tol is None:
            tol = 8 * 2.0 ** -53
        _check(lib().oocdag_batched_solve(k, self._i(self.n), self._i(self.nrhs),
                                          self._l(self.a_off), self._l(self.b_off), self.A.ptr,
                                          self.B.ptr, self.X.ptr, PRECISIONS[precision], tol,
                                          max_iter, reps, berr.ctypes.data_as(_dp),
                                          lres.ctypes.data_as(_dp), self._i(iters),
                                          self._i(status), C.byref(st)),
               "batched_solve")
        if with_resid:
            return st.as_dict(), berr, iters, status, lres
        return st.as_dict(), berr, iters, status

    def host_system(self, s):
        """(A, B, X) of system s copied to the host."""
        n, r, ld = int(self.n[s]), int(self.nrhs[s]), self.ld(s)
        def get(buf, off, cols):
            out = np.empty(ld * cols)
            _check(lib().oocdag_memcpy_d2h(out.ctypes.data, buf.ptr + int(off) * 8, out.nbytes), "d2h")
            return out.reshape(ld, cols, order="F")[:n]
        return get(self.A, self.a_off[s], n), get(self.B, self.b_off[s], r), get(self.X, self.b_off[s], r)


class Server:
    """Continuous-batching SPD solve server (oocdag_server_*)."""

    def __init__(self, hbm_budget_bytes: int, slots: int = 3):
        self.h = lib().oocdag_server_create(int(hbm_budget_bytes), int(slots))
        if not self.h:
            raise OocdagError("oocdag_server_create failed (no GPU or bad arguments)")
        self.shapes = {}

    def submit(self, matrix_id, n, nrhs, mu, seed, rhs_seed, precision="fp32",
               tol=8 * 2.0 ** -53) -> int:
        rq = Request(matrix_id, n, nrhs, mu, seed, rhs_seed, PRECISIONS[precision], tol)
        t = lib().oocdag_server_submit(self.h, C.byref(rq))
        if t <= 0:
            raise OocdagError(f"submit failed ({t})")
        self.shapes[t] = (n, nrhs)
        return t

    def wait(self, ticket, want_x=False):
        res = Result()
        n, r = self.shapes.pop(ticket)
        x = np.zeros((n, r), order="F") if want_x else None
        _check(lib().oocdag_server_wait(self.h, ticket, C.byref(res),
                                        x.ctypes.data_as(_dp) if want_x else None,
                                        n * r if want_x else 0), "server_wait")
        return res.as_dict(), x

    @staticmethod
    def estimate(matrix_id, n, nrhs, mu=0.1, seed=1, rhs_seed=2, precision="fp32",
                 tol=8 * 2.0 ** -53) -> dict:
        """Admission's byte estimates for a request (hit, miss, cached, ooc_min)."""
        rq = Request(matrix_id, n, nrhs, mu, seed, rhs_seed, PRECISIONS[precision], tol)
        v = [C.c_size_t(0) for _ in range(4)]
        _check(lib().oocdag_server_estimate(C.byref(rq), *[C.byref(x) for x in v]),
               "server_estimate")
        return dict(zip(("hit", "miss", "cached", "ooc_min"), (x.value for x in v)))

    def stats(self):
        u, c, p, r = C.c_size_t(0), C.c_int(0), C.c_int(0), C.c_int(0)
        lib().oocdag_server_stats(self.h, C.byref(u), C.byref(c), C.byref(p), C.byref(r))
        return {"used_bytes": u.value, "cached": c.value, "pending": p.value, "running": r.value}

    def close(self):
        if self.h:
            lib().oocdag_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tiles_from_host(buf: DeviceBuffer, A: np.ndarray, nb: int):
    _check(lib().oocdag_tiles_from_host(buf.ptr, _ptr(A), A.shape[1], nb, A.shape[0]),
           "tiles_from_host")


def tiles_to_host(buf: DeviceBuffer, A: np.ndarray, nb: int):
    _check(lib().oocdag_tiles_to_host(buf.ptr, _ptr(A), A.shape[1], nb, A.shape[0]),
           "tiles_to_host")


# ---------------------------------------------------------------------------
# tile kernels (cuBLAS / MKL call sites of oocquark.cpp replaced)
# ---------------------------------------------------------------------------

def dgemm_nt(alpha, A: DeviceBuffer, lda, B: DeviceBuffer, ldb, beta, Cb: DeviceBuffer, ldc,
             m, n, k):
    _check(lib().oocdag_dgemm_nt(m, n, k, alpha, A.ptr, lda, B.ptr, ldb, beta, Cb.ptr, ldc, None),
           "dgemm_nt")
    _check(lib().oocdag_device_sync(), "sync")


def dsyrk_ln(alpha, A: DeviceBuffer, lda, beta, Cb: DeviceBuffer, ldc, n, k):
    _check(lib().oocdag_dsyrk_ln(n, k, alpha, A.ptr, lda, beta, Cb.ptr, ldc, None), "dsyrk_ln")
    _check(lib().oocdag_device_sync(), "sync")


def dtrsm_rltn(L: DeviceBuffer, ldl, B: DeviceBuffer, ldb, m, n):
    _check(lib().oocdag_dtrsm_rltn(m, n, L.ptr, ldl, B.ptr, ldb, None), "dtrsm_rltn")


def dpotrf_l(A: DeviceBuffer, lda, n) -> int:
    info = C.c_int(0)
    _check(lib().oocdag_dpotrf_l(n, A.ptr, lda, C.byref(info), None), "dpotrf_l")
    return info.value
