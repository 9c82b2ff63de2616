"""Test infrastructure only: ctypes access to the plain-C restatement.

Loads oracle/_ref/libpk_port.so (``make -C oracle port``) and exposes the
restated drivers and scalar routines with the same call shapes as
``oracle.refpy``.  Used by tests/ as the checker and by bench.py's
cpu_baseline leg when the compiled reference is not present.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .refpy import RefResult, REF_DIR

_lib = None


def available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libpk_port.so"))


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(os.path.join(REF_DIR, "libpk_port.so"), mode=C.RTLD_LOCAL)
    return _lib


class _Stats(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("nopx", "nreorth", "ndot", "nbsvd", "nrestart", "nlandim", "nsing")]


def _pfx(prec):
    return "pkp_d_" if prec == "d" else "pkp_s_"


def _dt(prec):
    return np.float64 if prec == "d" else np.float32


def _R(prec):
    return C.c_double if prec == "d" else C.c_float


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class _Op:
    def __init__(self, prec, A, AT):
        dt = _dt(prec)
        if AT is None:
            AT = A.transpose()
        self.arrs = [np.ascontiguousarray(A.rowptr, np.int32),
                     np.ascontiguousarray(A.col, np.int32),
                     np.ascontiguousarray(A.val, dt),
                     np.ascontiguousarray(AT.rowptr, np.int32),
                     np.ascontiguousarray(AT.col, np.int32),
                     np.ascontiguousarray(AT.val, dt)]
        getattr(lib(), _pfx(prec) + "set_csr")(C.c_int(A.m), C.c_int(A.n),
                                               *[_ptr(a) for a in self.arrs])
        self.fn = C.cast(getattr(lib(), _pfx(prec) + "csr_aprod"), C.c_void_p)


def _stats(prec) -> dict:
    s = _Stats()
    getattr(lib(), _pfx(prec) + "get_stats")(C.byref(s))
    return {n: getattr(s, n) for n, _ in _Stats._fields_}


def lansvd(A, k, kmax, u0=None, tol=1e-12, jobu="y", jobv="y", prec="d",
           delta=None, eta=None, anorm=0.0, cgs=0, elr=1, AT=None) -> RefResult:
    """Restated DLANSVD (double/dlansvd.F:1-291) on CSR ``A``."""
    L, dt, R = lib(), _dt(prec), _R(prec)
    op = _Op(prec, A, AT)
    m, n = A.m, A.n
    eps = float(np.finfo(dt).eps) / 2 * 2
    U = np.zeros((kmax + 1, m), dt)
    V = np.zeros((kmax, n), dt)
    if u0 is not None:
        U[0] = u0
    sigma = np.zeros(kmax, dt)
    bnd = np.zeros(kmax, dt)
    dopt = np.array([np.sqrt(eps) if delta is None else delta,
                     eps ** 0.75 if eta is None else eta, anorm, 0.0], dt)
    iopt = np.array([cgs, elr], np.int32)
    kk = C.c_int(k)
    getattr(L, _pfx(prec) + "clear_stats")()
    info = getattr(L, _pfx(prec) + "lansvd")(
        C.c_char(jobu.encode()), C.c_char(jobv.encode()), C.c_int(m), C.c_int(n),
        C.byref(kk), C.c_int(kmax), op.fn, None, _ptr(U), C.c_int(m),
        _ptr(sigma), _ptr(bnd), _ptr(V), C.c_int(n), R(tol), _ptr(dopt), _ptr(iopt))
    k = kk.value
    return RefResult(sigma[:k].copy(), bnd[:k].copy(),
                     U[:k].T.copy() if jobu == "y" else None,
                     V[:k].T.copy() if jobv == "y" else None,
                     info, k, _stats(prec), anorm=float(dopt[2]))


def lansvd_irl(A, neig, dim, p=None, u0=None, tol=1e-10, which="L", jobu="y",
               jobv="y", maxiter=1000, prec="d", delta=None, eta=None,
               anorm=0.0, min_relgap=0.002, cgs=0, elr=1, AT=None) -> RefResult:
    """Restated DLANSVD_IRL (double/dlansvd_irl.F:1-419) on CSR ``A``."""
    L, dt, R = lib(), _dt(prec), _R(prec)
    op = _Op(prec, A, AT)
    m, n = A.m, A.n
    eps = float(np.finfo(dt).eps)
    if p is None:
        p = dim - neig
    U = np.zeros((dim + 1, m), dt)
    V = np.zeros((dim, n), dt)
    if u0 is not None:
        U[0] = u0
    sigma = np.zeros(dim, dt)
    bnd = np.zeros(dim, dt)
    dopt = np.array([np.sqrt(eps) if delta is None else delta,
                     eps ** 0.75 if eta is None else eta, anorm, min_relgap], dt)
    iopt = np.array([cgs, elr], np.int32)
    dd, ne = C.c_int(dim), C.c_int(neig)
    getattr(L, _pfx(prec) + "clear_stats")()
    info = getattr(L, _pfx(prec) + "lansvd_irl")(
        C.c_char(which.encode()), C.c_char(jobu.encode()), C.c_char(jobv.encode()),
        C.c_int(m), C.c_int(n), C.byref(dd), C.c_int(p), C.byref(ne),
        C.c_int(maxiter), op.fn, None, _ptr(U), C.c_int(m), _ptr(sigma),
        _ptr(bnd), _ptr(V), C.c_int(n), R(tol), _ptr(dopt), _ptr(iopt))
    k = ne.value
    return RefResult(sigma[:k].copy(), bnd[:k].copy(),
                     U[:k].T.copy() if jobu == "y" else None,
                     V[:k].T.copy() if jobv == "y" else None,
                     info, k, _stats(prec), dim=dd.value, anorm=float(dopt[2]))


def larnv_pm1(n: int, prec: str = "d") -> np.ndarray:
    x = np.zeros(n, _dt(prec))
    getattr(lib(), _pfx(prec) + "larnv_pm1")(C.c_int(n), _ptr(x))
    return x
