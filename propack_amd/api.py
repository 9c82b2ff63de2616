"""Python host-side access to libpropack_amd (the C ABI of include/propack_amd.h).

The product path is the HIP library; this module only marshals arguments.
If the library is missing the import of ``lib()`` raises -- there is no CPU
fallback.

* ``CsrOperator`` -- a device-resident CSR operator (A and its explicit A^T).
* ``lansvd`` / ``lansvd_irl`` -- call the Fortran drop-ins DLANSVD /
  DLANSVD_IRL (or SLANSVD / SLANSVD_IRL) with the reference calling
  sequence (double/dlansvd.F:1-3, double/dlansvd_irl.F:1-3), either with the
  device operator (``aprod=op``) or with a Python callable as APROD
  (host-callback path).
* ``Engine`` + ``lansvd_irl_dev`` / ``lansvd_dev`` -- device-resident drivers
  that leave the singular vectors on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libpropack_amd.so")

PK_F64, PK_F32, PK_C128, PK_C64 = 0, 1, 2, 3
PK_HOST, PK_DEVICE = 0, 1

EXPORTS = [
    "pk_last_error", "pk_csr_create", "pk_csr_destroy", "pk_csr_dims", "pk_csr_spmv",
    "pk_csr_to_iparm", "pk_dcsr_aprod_", "pk_scsr_aprod_", "pk_lbd_create",
    "pk_lbd_destroy", "pk_lbd_set_callback", "pk_lbd_put", "pk_lbd_get", "pk_lbd_basis",
    "pk_lbd_nrm2", "pk_lbd_getu0", "pk_lbd_lanbpro", "pk_lbd_update", "pk_lbd_stats",
    "pk_lbd_reset_stats", "pk_lbd_sync", "pk_lbd_shape", "pk_lansvd_run",
    "pk_lansvd_irl_run", "pk_lansvd_irl_mixed", "pk_lansvd_mixed", "pk_lbd_seed", "pk_lbd_set_mode", "pk_lbd_profile", "pk_lbd_profile_get", "pk_lbd_profile_reset",
    "pk_larnv_pm1", "pk_larnv_pm1_device", "pkh_update_mu", "pkh_update_nu", "pkh_compute_int", "pkh_set_mu",
    "pkh_lartg", "pkh_bdqr", "pkh_bsvdstep", "pkh_refinebounds", "dlansvd_", "dlansvd_irl_", "slansvd_", "slansvd_irl_",
    "clearstat_", "printstat_", "pk_lbd_set_rounding", "pk_lbd_eps", "pk_batch_create",
    "pk_batch_destroy", "pk_batch_lanbpro", "pk_lansvd_irl_batch", "pk_batch_set_bsvd",
    "pk_batch_get_bsvd", "pk_batch_bsvd", "pk_bsvd_batch", "pk_matrix_read",
    "pk_matrix_free", "pk_coo_to_csr", "pk_csr_create_coo", "pk_csr_read",
    "pk_csr_create_panels", "pk_csr_panels", "pk_csr_set_lanes", "pk_csr_lanes", "pk_csr_autotune",
    "pk_csr_set_sort", "pk_csr_sort", "pk_lbd_path_stats", "pk_batch_set_graph",
    "pk_csr_spmv_noalloc", "pk_lbd_residuals", "pk_lbd_promote",
    "zlansvd_", "zlansvd_irl_", "clansvd_", "clansvd_irl_", "pk_zcsr_aprod_", "pk_ccsr_aprod_",
    "pk_batch_update",
]

_lib = None


class PropackError(RuntimeError):
    pass


def lib():
    """Load libpropack_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PropackError(f"{LIB_PATH} is not built; run python -m propack_amd.build")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        L.pk_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().pk_last_error().decode(errors="replace")
        raise PropackError(f"{what} failed with code {rc}: {msg}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _dt(prec):
    """Element type of arrays of precision ``prec``."""
    return {PK_F64: np.float64, PK_F32: np.float32, PK_C128: np.complex128,
            PK_C64: np.complex64}[prec]


def _rt(prec):
    """Real type of precision ``prec`` (singular values, options)."""
    return np.float64 if prec in (PK_F64, PK_C128) else np.float32


def is_complex(prec) -> bool:
    return prec in (PK_C128, PK_C64)


def prec_of(dtype) -> int:
    """Precision code of a numpy dtype."""
    dt = np.dtype(dtype)
    if dt == np.complex128:
        return PK_C128
    if dt == np.complex64:
        return PK_C64
    return PK_F32 if dt == np.float32 else PK_F64


class CsrOperator:
    """Device CSR operator built from a host ``Csr`` (A^T built on device).
    Complex precisions (PK_C128 / PK_C64) apply A, A^T or A^H ('c') and
    have one layout (no panels, lanes or sorted layouts)."""

    def __init__(self, A, prec: int = PK_F64, panels=None, lanes=None, sort=None):
        """``panels``: (A, A^T) column-panel counts, ``lanes``: (A, A^T)
        lanes-per-row shifts, ``sort``: (A, A^T) block caps of the
        column-sorted layout (0: row-block layout); default the built-in
        choices (see autotune)."""
        L = lib()
        self.m, self.n, self.prec = A.m, A.n, prec
        rp = np.ascontiguousarray(A.rowptr, np.int32)
        ci = np.ascontiguousarray(A.col, np.int32)
        va = np.ascontiguousarray(A.val, _dt(prec))
        h = C.c_void_p()
        if panels is None:
            _check(L.pk_csr_create(C.c_int(A.m), C.c_int(A.n), C.c_longlong(int(rp[-1])),
                                   _p(rp), _p(ci), _p(va), C.c_int(prec), C.c_int(PK_HOST),
                                   C.byref(h)), "pk_csr_create")
        else:
            _check(L.pk_csr_create_panels(C.c_int(A.m), C.c_int(A.n), C.c_longlong(int(rp[-1])),
                                          _p(rp), _p(ci), _p(va), C.c_int(prec),
                                          C.c_int(PK_HOST), C.c_int(panels[0]),
                                          C.c_int(panels[1]), C.byref(h)),
                   "pk_csr_create_panels")
        self.handle = h
        if lanes is not None:
            self.set_lanes(*lanes)
        if sort is not None:
            self.set_sort(*sort)
        self.iparm = np.zeros(4, np.int32)
        _check(L.pk_csr_to_iparm(h, _p(self.iparm)), "pk_csr_to_iparm")

    @classmethod
    def from_file(cls, path, prec=PK_F64):
        """Device operator straight from a Matrix Market / Harwell-Boeing /
        coordinate file (pk_csr_read)."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        _check(lib().pk_csr_read(str(path).encode(), C.c_int(prec), C.byref(h)), "pk_csr_read")
        m, n, nnz = C.c_int(), C.c_int(), C.c_longlong()
        _check(lib().pk_csr_dims(h, C.byref(m), C.byref(n), C.byref(nnz)), "pk_csr_dims")
        self.handle, self.m, self.n, self.prec = h, m.value, n.value, prec
        self.nnz = nnz.value
        self.iparm = np.zeros(4, np.int32)
        _check(lib().pk_csr_to_iparm(h, _p(self.iparm)), "pk_csr_to_iparm")
        return self

    def panels(self):
        pa, pt = C.c_int(), C.c_int()
        _check(lib().pk_csr_panels(self.handle, C.byref(pa), C.byref(pt)), "pk_csr_panels")
        return pa.value, pt.value

    def set_lanes(self, shift_a: int, shift_t: int):
        """Lanes per row of (A, A^T): the built-in choice times 2^shift,
        pk_csr_set_lanes."""
        _check(lib().pk_csr_set_lanes(self.handle, C.c_int(shift_a), C.c_int(shift_t)),
               "pk_csr_set_lanes")

    def lanes(self):
        sa, st = C.c_int(), C.c_int()
        _check(lib().pk_csr_lanes(self.handle, C.byref(sa), C.byref(st)), "pk_csr_lanes")
        return sa.value, st.value

    def set_sort(self, cap_a: int, cap_t: int):
        """SpMV layout per direction (pk_csr_set_sort): cap > 0 the
        column-sorted layout with blocks of at most cap nonzeros, -cap the
        same with every column panel in one launch, 0 the row-block layout,
        -1 the row-wise layout."""
        _check(lib().pk_csr_set_sort(self.handle, C.c_int(cap_a), C.c_int(cap_t)),
               "pk_csr_set_sort")

    def sort(self):
        ca, ct = C.c_int(), C.c_int()
        _check(lib().pk_csr_sort(self.handle, C.byref(ca), C.byref(ct)), "pk_csr_sort")
        return ca.value, ct.value

    def spmv(self, x, trans="n"):
        x = np.ascontiguousarray(x, _dt(self.prec))
        y = np.zeros(self.m if trans == "n" else self.n, _dt(self.prec))
        _check(lib().pk_csr_spmv(self.handle, C.c_char(trans.encode()), _p(x), _p(y),
                                 C.c_int(self.prec), C.c_int(PK_HOST)), "pk_csr_spmv")
        return y

    def close(self):
        if self.handle:
            lib().pk_csr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class SvdResult:
    sigma: np.ndarray
    bnd: np.ndarray
    U: np.ndarray | None
    V: np.ndarray | None
    info: int
    k: int
    counters: dict = field(default_factory=dict)
    dim: int = 0
    anorm: float = 0.0


_TIMING = ["nopx", "nreorth", "ndot", "nreorthu", "nreorthv", "nitref", "nrestart", "nbsvd"]


class _Timing(C.Structure):
    # COMMON /timing/ layout (8 ints, 14 reals, nlandim, 3 reals, nsing)
    _fields_ = [(n, C.c_int) for n in _TIMING] + [("t1", C.c_float * 14),
                                                   ("nlandim", C.c_int),
                                                   ("t2", C.c_float * 3), ("nsing", C.c_int)]


def timing() -> dict:
    t = _Timing.in_dll(lib(), "timing_")
    d = {n: getattr(t, n) for n in _TIMING}
    d["nlandim"], d["nsing"] = t.nlandim, t.nsing
    return d


def clearstat() -> None:
    lib().clearstat_()


_APRODS = {PK_F64: "pk_dcsr_aprod_", PK_F32: "pk_scsr_aprod_", PK_C128: "pk_zcsr_aprod_",
           PK_C64: "pk_ccsr_aprod_"}


def _make_aprod(aprod, m, n, prec):
    """(function pointer, iparm, keepalive) for an APROD argument: a
    CsrOperator (its device path) or a Python callable aprod(t, x) with
    t = 'n' or 't' ('c' = A^H for complex precisions)."""
    if isinstance(aprod, CsrOperator):
        if aprod.prec != prec:
            raise PropackError("operator precision differs from the driver's")
        fn = C.cast(getattr(lib(), _APRODS[prec]), C.c_void_p)
        return fn, aprod.iparm, None
    dt = _dt(prec)
    ct = C.c_double if _rt(prec) == np.float64 else C.c_float
    w = 2 if is_complex(prec) else 1
    wide = np.complex128 if is_complex(prec) else np.float64

    def cb(transa, mp, np_, x, y, dparm, iparm, ln):
        t = transa[0:1].decode().lower()
        nin, nout = (n, m) if t == "n" else (m, n)
        xa = np.ctypeslib.as_array(C.cast(x, C.POINTER(ct)), shape=(w * nin,)).view(dt)
        ya = np.ctypeslib.as_array(C.cast(y, C.POINTER(ct)), shape=(w * nout,)).view(dt)
        ya[:] = np.asarray(aprod(t, xa.astype(wide)), dtype=dt)

    proto = C.CFUNCTYPE(None, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p,
                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t)
    f = proto(cb)
    return C.cast(f, C.c_void_p), np.zeros(4, np.int32), f


_DRV = {PK_F64: ("dlansvd_", "dlansvd_irl_"), PK_F32: ("slansvd_", "slansvd_irl_"),
        PK_C128: ("zlansvd_", "zlansvd_irl_"), PK_C64: ("clansvd_", "clansvd_irl_")}


def lansvd(aprod, m, n, k, kmax, u0=None, tol=1e-12, jobu="y", jobv="y", prec=PK_F64,
           delta=None, eta=None, anorm=0.0, cgs=1, elr=1) -> SvdResult:
    """DLANSVD / SLANSVD / ZLANSVD / CLANSVD through the library's Fortran
    drop-in."""
    L = lib()
    dt, rt = _dt(prec), _rt(prec)
    eps = float(np.finfo(rt).eps)
    U = np.zeros((kmax + 1, m), dt)
    V = np.zeros((kmax, n), dt)
    if u0 is not None:
        U[0] = u0
    sigma = np.zeros(kmax, rt)
    bnd = np.zeros(kmax, rt)
    dopt = np.array([np.sqrt(eps) if delta is None else delta,
                     eps ** 0.75 if eta is None else eta, anorm, 0.0], rt)
    iopt = np.array([cgs, elr], np.int32)
    lwork = m + n + 9 * kmax + 5 * kmax * kmax + 4 + max(3 * kmax * kmax + 4 * kmax + 4,
                                                         16 * max(m, n))
    work = np.zeros(1, rt)
    iwork = np.zeros(1, np.int32)
    fn, iparm, keep = _make_aprod(aprod, m, n, prec)
    dparm = np.zeros(1, dt)
    R = C.c_double if rt == np.float64 else C.c_float
    ii = lambda v: C.byref(C.c_int(v))
    kk, info = C.c_int(k), C.c_int(0)
    clearstat()
    name = _DRV[prec][0]
    # the complex drivers take ZWORK(LZWRK) after WORK(LWORK)
    zw = (_p(np.zeros(1, dt)), ii(1)) if is_complex(prec) else ()
    getattr(L, name)(C.c_char_p(jobu.encode()), C.c_char_p(jobv.encode()), ii(m), ii(n),
                     C.byref(kk), ii(kmax), fn, _p(U), ii(m), _p(sigma), _p(bnd), _p(V), ii(n),
                     C.byref(R(tol)), _p(work), ii(lwork), *zw, _p(iwork), ii(8 * kmax),
                     _p(dopt), _p(iopt), C.byref(info), _p(dparm), _p(iparm), C.c_size_t(1),
                     C.c_size_t(1))
    del keep
    if info.value <= -100:
        _check(info.value, name.rstrip("_"))
    kout = kk.value
    return SvdResult(sigma[:kout].copy(), bnd[:kout].copy(),
                     U[:kout].T.copy() if jobu == "y" else None,
                     V[:kout].T.copy() if jobv == "y" else None,
                     info.value, kout, timing(), anorm=float(dopt[2]))


def lansvd_irl(aprod, m, n, neig, dim, p=None, u0=None, tol=1e-10, which="L", jobu="y",
               jobv="y", maxiter=1000, prec=PK_F64, delta=None, eta=None, anorm=0.0,
               min_relgap=0.002, cgs=1, elr=1) -> SvdResult:
    """DLANSVD_IRL / SLANSVD_IRL / ZLANSVD_IRL / CLANSVD_IRL through the
    library's Fortran drop-in."""
    L = lib()
    dt, rt = _dt(prec), _rt(prec)
    eps = float(np.finfo(rt).eps)
    if p is None:
        p = dim - neig
    U = np.zeros((dim + 1, m), dt)
    V = np.zeros((dim, n), dt)
    if u0 is not None:
        U[0] = u0
    sigma = np.zeros(dim, rt)
    bnd = np.zeros(dim, rt)
    dopt = np.array([np.sqrt(eps) if delta is None else delta,
                     eps ** 0.75 if eta is None else eta, anorm, min_relgap], rt)
    iopt = np.array([cgs, elr], np.int32)
    work = np.zeros(1, rt)
    iwork = np.zeros(1, np.int32)
    fn, iparm, keep = _make_aprod(aprod, m, n, prec)
    dparm = np.zeros(1, dt)
    R = C.c_double if rt == np.float64 else C.c_float
    ii = lambda v: C.byref(C.c_int(v))
    dd, ne, info = C.c_int(dim), C.c_int(neig), C.c_int(0)
    clearstat()
    name = _DRV[prec][1]
    zw = (_p(np.zeros(1, dt)), ii(1)) if is_complex(prec) else ()
    getattr(L, name)(C.c_char_p(which.encode()), C.c_char_p(jobu.encode()),
                     C.c_char_p(jobv.encode()), ii(m), ii(n), C.byref(dd), ii(p), C.byref(ne),
                     ii(maxiter), fn, _p(U), ii(m), _p(sigma), _p(bnd), _p(V), ii(n),
                     C.byref(R(tol)), _p(work), ii(1), *zw, _p(iwork), ii(1), _p(dopt),
                     _p(iopt), C.byref(info), _p(dparm), _p(iparm), C.c_size_t(1),
                     C.c_size_t(1), C.c_size_t(1))
    del keep
    if info.value <= -100:
        _check(info.value, name.rstrip("_"))
    k = ne.value
    return SvdResult(sigma[:k].copy(), bnd[:k].copy(),
                     U[:k].T.copy() if jobu == "y" else None,
                     V[:k].T.copy() if jobv == "y" else None,
                     info.value, k, timing(), dim=dd.value, anorm=float(dopt[2]))


class Engine:
    """A Lanczos engine (device bases) bound to a device operator."""

    def __init__(self, op: CsrOperator, kmax: int, prec: int | None = None):
        self.op = op
        self.prec = op.prec if prec is None else prec
        self.m, self.n, self.kmax = op.m, op.n, kmax
        h = C.c_void_p()
        _check(lib().pk_lbd_create(C.c_int(self.prec), C.c_int(op.m), C.c_int(op.n),
                                   C.c_int(kmax), op.handle, C.byref(h)), "pk_lbd_create")
        self.handle = h

    def put_u1(self, u0):
        """U(:,1) <- u0: a host array, or a contiguous torch tensor on the GPU
        (float64, or complex128 for a complex engine; then a device-to-device
        copy)."""
        wide = PK_C128 if is_complex(self.prec) else PK_F64
        if hasattr(u0, "is_cuda") and u0.is_cuda:
            want = "torch.complex128" if wide == PK_C128 else "torch.float64"
            if str(u0.dtype) != want or not u0.is_contiguous() or u0.numel() != self.m:
                raise ValueError(f"put_u1: device start vector must be contiguous {want} of "
                                 "length m")
            ptr = C.c_void_p(u0.data_ptr())
            _check(lib().pk_lbd_put(self.handle, C.c_char(b"u"), C.c_int(1), ptr,
                                    C.c_int(wide)), "pk_lbd_put")
            return
        u0 = np.ascontiguousarray(u0, _dt(wide))
        _check(lib().pk_lbd_put(self.handle, C.c_char(b"u"), C.c_int(1), _p(u0),
                                C.c_int(wide)), "pk_lbd_put")

    def get(self, which: str, ncols: int, prec=None):
        """The first ncols columns of U or V as an (L x ncols) host array (in
        double or complex double unless prec says otherwise)."""
        if prec is None:
            prec = PK_C128 if is_complex(self.prec) else PK_F64
        L = self.m if which == "u" else self.n
        X = np.zeros((ncols, L), _dt(prec))
        if ncols:
            _check(lib().pk_lbd_get(self.handle, C.c_char(which.encode()), C.c_int(ncols), _p(X),
                                    C.c_int(L), C.c_int(prec)), "pk_lbd_get")
        return X.T

    def stats(self) -> dict:
        out = np.zeros(6, np.int64)
        _check(lib().pk_lbd_stats(self.handle, _p(out)), "pk_lbd_stats")
        return dict(zip(["nopx", "nreorth", "ndot", "ncols", "nsync", "nretry"], out.tolist()))

    def residuals(self, src: "Engine", sigma) -> np.ndarray:
        """Device residuals of the triplets in src's bases with this engine's
        operator: an array of (||A v - s u||, ||A^T u - s v||) per triplet
        (pk_lbd_residuals)."""
        sig = np.ascontiguousarray(sigma, np.float64)
        out = np.zeros(2 * sig.size)
        _check(lib().pk_lbd_residuals(self.handle, src.handle, C.c_int(sig.size), _p(sig),
                                      _p(out)), "pk_lbd_residuals")
        return out.reshape(-1, 2)

    def path_stats(self) -> dict:
        """How often the reference's rare DLANBPRO paths ran on this engine."""
        out = np.zeros(8, np.int64)
        _check(lib().pk_lbd_path_stats(self.handle, _p(out)), "pk_lbd_path_stats")
        return dict(zip(["v_invariant", "v_tiny", "v_zero", "u_invariant", "u_tiny", "u_zero",
                         "nwarn"], out[:7].tolist()))

    def set_mode(self, mode: str = "device", tries: int = 2):
        """'device': device-resident step chains; 'host': host executor."""
        _check(lib().pk_lbd_set_mode(self.handle, C.c_int(0 if mode == "device" else 1),
                                     C.c_int(tries)), "pk_lbd_set_mode")

    def reset_stats(self):
        lib().pk_lbd_reset_stats(self.handle)

    def set_rounding(self, model: int):
        """0: the reference's thresholds (max(m, n) * eps); 1: storage
        rounding only (FP32 storage with FP64 accumulation)."""
        _check(lib().pk_lbd_set_rounding(self.handle, C.c_int(model)), "pk_lbd_set_rounding")

    KCLASSES = ["spmv_n", "spmv_t", "gemvt", "gemv", "gemm", "vec"]

    PROF_SPMV, PROF_ALL = 0x03, 0x3F

    def profile(self, enable: bool = True, classes: int = None):
        """Live kernel timing of the classes in the bit mask (bit i: KCLASSES[i];
        all classes by default); enable=False turns it off."""
        mask = (self.PROF_ALL if classes is None else int(classes)) if enable else 0
        _check(lib().pk_lbd_profile(self.handle, C.c_int(mask)), "pk_lbd_profile")
        _check(lib().pk_lbd_profile_reset(self.handle), "pk_lbd_profile_reset")

    def profile_get(self) -> dict:
        out = np.zeros(4 * len(self.KCLASSES))
        _check(lib().pk_lbd_profile_get(self.handle, _p(out)), "pk_lbd_profile_get")
        return {c: dict(launches=int(out[4 * i]), ms=float(out[4 * i + 1]),
                        bytes=float(out[4 * i + 2]), flops=float(out[4 * i + 3]))
                for i, c in enumerate(self.KCLASSES)}

    def close(self):
        if self.handle:
            lib().pk_lbd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lansvd_irl_dev(eng: Engine, neig, dim, p=None, tol=1e-10, which="L", jobu="y", jobv="y",
                   maxiter=1000, delta=None, eta=None, anorm=0.0, min_relgap=0.002, cgs=1,
                   elr=1, fetch=True) -> SvdResult:
    """Device-resident DLANSVD_IRL (ZLANSVD_IRL for a complex engine) on
    ``eng`` (start vector = eng U(:,1))."""
    eps = 2.0 ** -53 if _rt(eng.prec) == np.float64 else 2.0 ** -24
    if p is None:
        p = dim - neig
    sigma = np.zeros(dim + 1)
    bnd = np.zeros(dim + 1)
    dopt = np.array([np.sqrt(2 * eps) if delta is None else delta,
                     (2 * eps) ** 0.75 if eta is None else eta, anorm, min_relgap])
    iopt = np.array([cgs, elr], np.int32)
    dd, ne = C.c_int(dim), C.c_int(neig)
    clearstat()
    info = lib().pk_lansvd_irl_run(eng.handle, C.c_char(which.encode()), C.c_char(jobu.encode()),
                                   C.c_char(jobv.encode()), C.byref(dd), C.c_int(p), C.byref(ne),
                                   C.c_int(maxiter), _p(sigma), _p(bnd), C.c_double(tol),
                                   _p(dopt), _p(iopt))
    if info <= -100:
        _check(info, "pk_lansvd_irl_run")
    k = ne.value
    U = eng.get("u", k) if (fetch and jobu == "y") else None
    V = eng.get("v", k) if (fetch and jobv == "y") else None
    return SvdResult(sigma[:k].copy(), bnd[:k].copy(), U, V, info, k, timing(), dim=dd.value,
                     anorm=float(dopt[2]))


@dataclass
class Tuning:
    """pk_csr_autotune's choice for (A, A^T): column panels, lane shifts,
    block caps of the column-sorted layout (0: row-block layout), and their
    microseconds per application."""
    panels: tuple
    lanes: tuple
    us: tuple
    sort: tuple = (0, 0)


def autotune(A, prec: int = PK_F64, lanes: bool = True, sort: bool = True) -> Tuning:
    """Fastest (A, A^T) column-panel counts (and, with ``lanes``, lanes-per-row
    shifts; with ``sort``, the column-sorted layout's block caps) for ``A``
    on this GPU (pk_csr_autotune)."""
    rp = np.ascontiguousarray(A.rowptr, np.int32)
    ci = np.ascontiguousarray(A.col, np.int32)
    va = np.ascontiguousarray(A.val, _dt(prec))
    pa, pt, ua, ut = C.c_int(), C.c_int(), C.c_double(), C.c_double()
    sa, st, ca, ct = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _check(lib().pk_csr_autotune(C.c_int(A.m), C.c_int(A.n), C.c_longlong(int(rp[-1])), _p(rp),
                                 _p(ci), _p(va), C.c_int(prec), C.c_int(PK_HOST), C.byref(pa),
                                 C.byref(pt), C.byref(ua), C.byref(ut),
                                 C.byref(sa) if lanes else None, C.byref(st) if lanes else None,
                                 C.byref(ca) if sort else None, C.byref(ct) if sort else None),
           "pk_csr_autotune")
    return Tuning((pa.value, pt.value), (sa.value, st.value), (ua.value, ut.value),
                  (ca.value, ct.value))


def read_matrix(path):
    """Coordinate form of a Matrix Market / Harwell-Boeing / coordinate file
    (host parse): (m, n, rows, cols, vals), 0-based, symmetric storage
    expanded, duplicates kept."""
    m, n, nnz = C.c_int(), C.c_int(), C.c_longlong()
    r, c, v = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    L = lib()
    _check(L.pk_matrix_read(str(path).encode(), C.byref(m), C.byref(n), C.byref(nnz), C.byref(r),
                            C.byref(c), C.byref(v)), "pk_matrix_read")
    k = nnz.value
    try:
        rows = np.ctypeslib.as_array(r, (max(k, 1),))[:k].copy()
        cols = np.ctypeslib.as_array(c, (max(k, 1),))[:k].copy()
        vals = np.ctypeslib.as_array(v, (max(k, 1),))[:k].copy()
    finally:
        for ptr in (r, c, v):
            L.pk_matrix_free(C.cast(ptr, C.c_void_p))
    return m.value, n.value, rows, cols, vals


def coo_to_csr(m, n, rows, cols, vals, base=0):
    """CSR (synth.Csr) of coordinate entries, built on the device: sorted by
    (row, column), duplicates summed in input order (complex values: real
    and imaginary parts each)."""
    vals = np.ascontiguousarray(vals)
    prec = prec_of(vals.dtype) if vals.dtype in (np.float32, np.complex64, np.complex128) \
        else PK_F64
    vals = vals.astype(_dt(prec))
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    nnz = rows.size
    rp = np.zeros(m + 1, np.int32)
    ci = np.zeros(max(nnz, 1), np.int32)
    va = np.zeros(max(nnz, 1), _dt(prec))
    nz = C.c_longlong()
    _check(lib().pk_coo_to_csr(C.c_int(m), C.c_int(n), C.c_longlong(nnz), _p(rows), _p(cols),
                               _p(vals), C.c_int(prec), C.c_int(PK_HOST), C.c_int(base), _p(rp),
                               _p(ci), _p(va), C.byref(nz)), "pk_coo_to_csr")
    from .synth import Csr
    return Csr(m, n, rp, ci[: nz.value].copy(), va[: nz.value].copy())


class Batch:
    """Engines whose Lanczos steps run together (pk_batch_create)."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.handle for e in self.engines])
        h = C.c_void_p()
        _check(lib().pk_batch_create(C.c_int(len(self.engines)), arr, C.byref(h)),
               "pk_batch_create")
        self.handle = h

    def set_graph(self, on: bool = True):
        """Replay each batched call's kernel chain as a captured HIP graph
        (default) or enqueue it directly (pk_batch_set_graph)."""
        _check(lib().pk_batch_set_graph(self.handle, C.c_int(int(on))), "pk_batch_set_graph")

    def set_bsvd(self, on: bool = True):
        """Run the restarted driver's bidiagonal SVDs on the device for the
        whole batch (pk_batch_set_bsvd) instead of the host DBDSQR."""
        _check(lib().pk_batch_set_bsvd(self.handle, C.c_int(int(on))), "pk_batch_set_bsvd")

    def close(self):
        if self.handle:
            lib().pk_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bsvd_batch(ds, es):
    """Singular values (decreasing) and the last components of the left
    singular vectors of upper bidiagonal matrices (diagonals ds[i],
    superdiagonals es[i]) on the device (pk_bsvd_batch)."""
    nb = len(ds)
    n = np.array([len(d) for d in ds], np.int32)
    ld = max(1, int(n.max()) if nb else 1)
    D = np.zeros((nb, ld))
    E = np.zeros((nb, ld))
    for i, (d, e) in enumerate(zip(ds, es)):
        D[i, :len(d)] = d
        E[i, :len(e)] = e
    S = np.zeros((nb, ld))
    X = np.zeros((nb, ld))
    _check(lib().pk_bsvd_batch(C.c_int(nb), _p(n), _p(D), _p(E), C.c_int(ld), _p(S), _p(X)),
           "pk_bsvd_batch")
    return [S[i, :n[i]].copy() for i in range(nb)], [X[i, :n[i]].copy() for i in range(nb)]


def lansvd_irl_batch(batch: Batch, neig, dim, p=None, tol=1e-10, which="L", jobu="n", jobv="n",
                     maxiter=1000, delta=None, eta=None, anorm=0.0, min_relgap=0.002, cgs=1,
                     elr=1, fetch=False):
    """DLANSVD_IRL on every engine of ``batch`` (start vectors = each
    engine's U(:,1)); their Lanczos steps advance together.  Returns one
    SvdResult per problem (vectors fetched when ``fetch``)."""
    engs = batch.engines
    nb = len(engs)
    eps = 2.0 ** -53 if engs[0].prec == PK_F64 else 2.0 ** -24
    if p is None:
        p = dim - neig
    ld = dim + 1
    sigma = np.zeros((nb, ld))
    bnd = np.zeros((nb, ld))
    nconv = np.zeros(nb, np.int32)
    info = np.zeros(nb, np.int32)
    dopt = np.array([np.sqrt(2 * eps) if delta is None else delta,
                     (2 * eps) ** 0.75 if eta is None else eta, anorm, min_relgap])
    iopt = np.array([cgs, elr], np.int32)
    arr = (C.c_void_p * nb)(*[e.handle for e in engs])
    clearstat()
    rc = lib().pk_lansvd_irl_batch(C.c_int(nb), arr, batch.handle, C.c_char(which.encode()),
                                   C.c_char(jobu.encode()), C.c_char(jobv.encode()),
                                   C.c_int(dim), C.c_int(p), C.c_int(neig), C.c_int(maxiter),
                                   C.c_double(tol), _p(sigma), _p(bnd), C.c_int(ld), _p(nconv),
                                   _p(info), _p(dopt), _p(iopt))
    _check(rc, "pk_lansvd_irl_batch")
    out = []
    for i, e in enumerate(engs):
        k = int(nconv[i])
        U = e.get("u", k) if (fetch and jobu == "y") else None
        V = e.get("v", k) if (fetch and jobv == "y") else None
        out.append(SvdResult(sigma[i, :k].copy(), bnd[i, :k].copy(), U, V, int(info[i]), k,
                             {"nopx": e.stats()["nopx"]}))
    return out


def lansvd_dev(eng: Engine, k, kmax, tol=1e-12, jobu="y", jobv="y", delta=None, eta=None,
               anorm=0.0, cgs=1, elr=1, fetch=True) -> SvdResult:
    """Device-resident DLANSVD (ZLANSVD for a complex engine) on ``eng``
    (start vector = eng U(:,1))."""
    eps = 2.0 ** -53 if _rt(eng.prec) == np.float64 else 2.0 ** -24
    sigma = np.zeros(kmax + 1)
    bnd = np.zeros(kmax + 1)
    dopt = np.array([np.sqrt(2 * eps) if delta is None else delta,
                     (2 * eps) ** 0.75 if eta is None else eta, anorm, 0.0])
    iopt = np.array([cgs, elr], np.int32)
    kk = C.c_int(k)
    clearstat()
    info = lib().pk_lansvd_run(eng.handle, C.c_char(jobu.encode()), C.c_char(jobv.encode()),
                               C.byref(kk), C.c_int(kmax), _p(sigma), _p(bnd), C.c_double(tol),
                               _p(dopt), _p(iopt))
    if info <= -100:
        _check(info, "pk_lansvd_run")
    k = kk.value
    U = eng.get("u", k) if (fetch and jobu == "y") else None
    V = eng.get("v", k) if (fetch and jobv == "y") else None
    return SvdResult(sigma[:k].copy(), bnd[:k].copy(), U, V, info, k, timing(), anorm=float(dopt[2]))


def lansvd_mixed(eng64: Engine, eng32: Engine, k, kmax, tol=1e-12, jobu="y", jobv="y",
                 cgs=1, elr=1, fetch=True):
    """Non-restarted FP32-basis solve (DLANSVD on ``eng32``, whose U(:,1) is
    the start vector) with the FP64 check / polish on ``eng64``.  Returns
    (polished SvdResult, dict with the FP32-phase sigma/bnd/info/k)."""
    sigma = np.zeros(kmax + 1)
    bnd = np.zeros(kmax + 1)
    s32 = np.zeros(kmax + 1)
    b32 = np.zeros(kmax + 1)
    eps = 2.0 ** -52
    dopt = np.array([np.sqrt(eps), eps ** 0.75, 0.0, 0.0])
    iopt = np.array([cgs, elr], np.int32)
    info32 = np.zeros(2, np.int32)
    kk = C.c_int(k)
    clearstat()
    info = lib().pk_lansvd_mixed(eng64.handle, eng32.handle, C.c_char(jobu.encode()),
                                 C.c_char(jobv.encode()), C.byref(kk), C.c_int(kmax), _p(sigma),
                                 _p(bnd), _p(s32), _p(b32), C.c_double(tol), _p(dopt), _p(iopt),
                                 _p(info32))
    if info <= -100:
        _check(info, "pk_lansvd_mixed")
    k = kk.value
    k32 = int(info32[1])
    U = eng64.get("u", k) if (fetch and jobu == "y") else None
    V = eng64.get("v", k) if (fetch and jobv == "y") else None
    res = SvdResult(sigma[:k].copy(), bnd[:k].copy(), U, V, info, k, timing(),
                    anorm=float(dopt[2]))
    return res, {"sigma": s32[:k32].copy(), "bnd": b32[:k32].copy(), "info": int(info32[0]),
                 "k": k32}


def lansvd_irl_mixed(eng64: Engine, eng32: Engine, neig, dim, p=None, tol=1e-10, which="L",
                     jobu="y", jobv="y", maxiter=1000, cgs=1, elr=1, min_relgap=0.002,
                     fetch=True):
    """FP32-basis solve on ``eng32`` then the seeded FP64 polish on ``eng64``
    (both engines start from their own U(:,1); eng32's is the start vector).
    Returns (polished SvdResult, dict with the FP32-phase sigma/bnd/info)."""
    if p is None:
        p = dim - neig
    sigma = np.zeros(dim + 1)
    bnd = np.zeros(dim + 1)
    s32 = np.zeros(dim + 1)
    b32 = np.zeros(dim + 1)
    eps = 2.0 ** -52
    dopt = np.array([np.sqrt(eps), eps ** 0.75, 0.0, min_relgap])
    iopt = np.array([cgs, elr], np.int32)
    info32 = np.zeros(2, np.int32)
    dd, ne = C.c_int(dim), C.c_int(neig)
    clearstat()
    info = lib().pk_lansvd_irl_mixed(eng64.handle, eng32.handle, C.c_char(which.encode()),
                                     C.c_char(jobu.encode()), C.c_char(jobv.encode()),
                                     C.byref(dd), C.c_int(p), C.byref(ne), C.c_int(maxiter),
                                     _p(sigma), _p(bnd), _p(s32), _p(b32), C.c_double(tol),
                                     _p(dopt), _p(iopt), _p(info32))
    if info <= -100:
        _check(info, "pk_lansvd_irl_mixed")
    k = ne.value
    k32 = int(info32[1])
    U = eng64.get("u", k) if (fetch and jobu == "y") else None
    V = eng64.get("v", k) if (fetch and jobv == "y") else None
    res = SvdResult(sigma[:k].copy(), bnd[:k].copy(), U, V, info, k, timing(), dim=dd.value)
    return res, {"sigma": s32[:k32].copy(), "bnd": b32[:k32].copy(), "info": int(info32[0]),
                 "k": k32}
