"""Test infrastructure only: ctypes access to the reference PROPACK oracle.

Loads oracle/_ref/libpropack_ref_{d,s,z,c}.so (built by ``make -C oracle
ref`` from the reference's own Fortran sources) and runs its drivers
DLANSVD (double/dlansvd.F:1) and DLANSVD_IRL (double/dlansvd_irl.F:1), the
S twins, or the complex ZLANSVD / ZLANSVD_IRL (complex16/zlansvd.F:1,
complex16/zlansvd_irl.F:1) and C twins, on a CSR matrix through the harness
APROD in ref_harness.c.
Only tests/, the smoke check and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_libs: dict = {}


def available(prec: str = "d") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libpropack_ref_{prec}.so"))


def lib(prec: str = "d"):
    if prec not in _libs:
        path = os.path.join(REF_DIR, f"libpropack_ref_{prec}.so")
        L = C.CDLL(path, mode=C.RTLD_LOCAL)
        getattr(L, f"ref_{prec}_init")()
        _libs[prec] = L
    return _libs[prec]


@dataclass
class RefResult:
    sigma: np.ndarray
    bnd: np.ndarray
    U: np.ndarray | None
    V: np.ndarray | None
    info: int
    k: int                      # number of converged triplets (K / NEIG out)
    counters: dict = field(default_factory=dict)
    dim: int = 0                # DIM out (IRL)
    anorm: float = 0.0          # doption(3) out


_CNT = ["nopx", "nreorth", "ndot", "nreorthu", "nreorthv", "nitref",
        "nrestart", "nbsvd", "nlandim", "nsing"]


class _Op:
    """Keeps the CSR arrays alive while the oracle holds raw pointers."""

    def __init__(self, L, prec, A, AT, dtype):
        self.arrs = [np.ascontiguousarray(A.rowptr, np.int32),
                     np.ascontiguousarray(A.col, np.int32),
                     np.ascontiguousarray(A.val, dtype),
                     np.ascontiguousarray(AT.rowptr, np.int32),
                     np.ascontiguousarray(AT.col, np.int32),
                     np.ascontiguousarray(AT.val, dtype)]
        ptrs = [a.ctypes.data_as(C.c_void_p) for a in self.arrs]
        getattr(L, f"ref_{prec}_set_operator")(C.c_int(A.m), C.c_int(A.n), *ptrs)


_DT = {"d": np.float64, "s": np.float32, "z": np.complex128, "c": np.complex64}
_RT = {"d": np.float64, "s": np.float32, "z": np.float64, "c": np.float32}


def _prep(prec, A, AT):
    L = lib(prec)
    dtype = _DT[prec]
    if AT is None:
        AT = A.transpose()
    return L, dtype, _Op(L, prec, A, AT, dtype)


def _counters(L, prec) -> dict:
    out = np.zeros(10, dtype=np.int32)
    getattr(L, f"ref_{prec}_get_counters")(out.ctypes.data_as(C.c_void_p))
    return dict(zip(_CNT, out.tolist()))


def set_threads(n: int, prec: str = "d") -> None:
    getattr(lib(prec), f"ref_{prec}_set_threads")(C.c_int(n))


def aprod(A, x, trans="n", prec="d", AT=None):
    L, dtype, op = _prep(prec, A, AT)
    x = np.ascontiguousarray(x, dtype)
    y = np.zeros(A.m if trans == "n" else A.n, dtype)
    getattr(L, f"ref_{prec}_aprod")(C.c_char(trans.encode()),
                                    x.ctypes.data_as(C.c_void_p),
                                    y.ctypes.data_as(C.c_void_p))
    return y


def lansvd(A, k, kmax, u0=None, tol=1e-12, jobu="y", jobv="y", prec="d",
           delta=None, eta=None, anorm=0.0, cgs=0, elr=1, AT=None) -> RefResult:
    """DLANSVD / SLANSVD / ZLANSVD / CLANSVD on CSR ``A``
    (double/dlansvd.F:1-291, complex16/zlansvd.F:1-301)."""
    L, dtype, op = _prep(prec, A, AT)
    rdt = _RT[prec]
    m, n = A.m, A.n
    eps = float(np.finfo(rdt).eps)
    U = np.zeros((kmax + 1, m), dtype)   # column-major U(m, kmax+1)
    V = np.zeros((kmax, n), dtype)
    if u0 is not None:
        U[0, :] = u0
    sigma = np.zeros(kmax, rdt)
    bnd = np.zeros(kmax, rdt)
    dopt = np.array([np.sqrt(eps) if delta is None else delta,
                     eps ** 0.75 if eta is None else eta, anorm, 0.0], rdt)
    iopt = np.array([cgs, elr], np.int32)
    kk = C.c_int(k)
    getattr(L, f"ref_{prec}_clear_counters")()
    fn = getattr(L, f"ref_{prec}_lansvd")
    R = C.c_double if rdt == np.float64 else C.c_float
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    info = fn(C.c_char(jobu.encode()), C.c_char(jobv.encode()), C.c_int(m),
              C.c_int(n), C.byref(kk), C.c_int(kmax), p(U), C.c_int(m),
              p(sigma), p(bnd), p(V), C.c_int(n), R(tol), p(dopt), p(iopt))
    kout = kk.value
    return RefResult(sigma[:kout].copy(), bnd[:kout].copy(),
                     U[:kout].T.copy() if jobu == "y" else None,
                     V[:kout].T.copy() if jobv == "y" else None,
                     info, kout, _counters(L, prec), anorm=float(dopt[2]))


def lansvd_irl(A, neig, dim, p=None, u0=None, tol=1e-10, which="L", jobu="y",
               jobv="y", maxiter=1000, prec="d", delta=None, eta=None,
               anorm=0.0, min_relgap=0.002, cgs=0, elr=1, AT=None) -> RefResult:
    """DLANSVD_IRL / SLANSVD_IRL / ZLANSVD_IRL / CLANSVD_IRL on CSR ``A``
    (double/dlansvd_irl.F:1-419, complex16/zlansvd_irl.F:1-421)."""
    L, dtype, op = _prep(prec, A, AT)
    rdt = _RT[prec]
    m, n = A.m, A.n
    eps = float(np.finfo(rdt).eps)
    if p is None:
        p = dim - neig
    U = np.zeros((dim + 1, m), dtype)
    V = np.zeros((dim, n), dtype)
    if u0 is not None:
        U[0, :] = u0
    sigma = np.zeros(dim, rdt)
    bnd = np.zeros(dim, rdt)
    dopt = np.array([np.sqrt(eps) if delta is None else delta,
                     eps ** 0.75 if eta is None else eta, anorm, min_relgap],
                    rdt)
    iopt = np.array([cgs, elr], np.int32)
    dd = C.c_int(dim)
    ne = C.c_int(neig)
    getattr(L, f"ref_{prec}_clear_counters")()
    fn = getattr(L, f"ref_{prec}_lansvd_irl")
    R = C.c_double if rdt == np.float64 else C.c_float
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    info = fn(C.c_char(which.encode()), C.c_char(jobu.encode()),
              C.c_char(jobv.encode()), C.c_int(m), C.c_int(n), C.byref(dd),
              C.c_int(p), C.byref(ne), C.c_int(maxiter), ptr(U), C.c_int(m),
              ptr(sigma), ptr(bnd), ptr(V), C.c_int(n), R(tol), ptr(dopt),
              ptr(iopt))
    k = ne.value
    return RefResult(sigma[:k].copy(), bnd[:k].copy(),
                     U[:k].T.copy() if jobu == "y" else None,
                     V[:k].T.copy() if jobv == "y" else None,
                     info, k, _counters(L, prec), dim=dd.value,
                     anorm=float(dopt[2]))
