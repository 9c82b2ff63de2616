"""scipy-compatible front end: ``svdp`` with the signature and return
convention of scipy's PROPACK wrapper (scipy/sparse/linalg/_svdp.py,
``_svdp``), running on the MI355X library.

Dense arrays and scipy sparse matrices become a device CSR operator (A and
A^T resident in HBM; the drivers recognise the library's own APROD and keep
the whole solve on the GPU).  A ``scipy.sparse.linalg.LinearOperator`` is
called from the host as the APROD callback (matvec / rmatvec), the rest of
the Lanczos process still runs on the GPU.  float32 inputs use SLANSVD /
SLANSVD_IRL, complex128 ZLANSVD / ZLANSVD_IRL, complex64 CLANSVD /
CLANSVD_IRL, everything else DLANSVD / DLANSVD_IRL; for complex inputs the
returned ``vt`` is V^H.  ``svds`` is ``scipy.sparse.linalg.svds`` with
``solver='propack'`` over the same drivers.
"""
from __future__ import annotations

import numpy as np
from numpy.linalg import LinAlgError

from . import api, synth

_which = {"LM": "L", "SM": "S"}


def _as_operator(A, prec, tune=False):
    """(aprod, m, n, keepalive): a CsrOperator for explicit matrices (with
    ``tune``, in the SpMV layout pk_csr_autotune measures fastest), a Python
    callable for LinearOperators."""
    try:
        import scipy.sparse as sp
        from scipy.sparse.linalg import LinearOperator
    except Exception:  # pragma: no cover - scipy is optional for explicit matrices
        sp, LinearOperator = None, None
    dt = api._dt(prec)
    if LinearOperator is not None and isinstance(A, LinearOperator):
        m, n = A.shape

        def aprod(t, x):  # t = 't' (real) or 'c' (complex): rmatvec is A^H x
            return A.matvec(x) if t == "n" else A.rmatvec(x)
        return aprod, m, n, None
    if sp is not None and sp.issparse(A):
        C = sp.csr_matrix(A)
        C.sum_duplicates()
        C.sort_indices()
        csr = synth.Csr(C.shape[0], C.shape[1], C.indptr.astype(np.int32),
                        C.indices.astype(np.int32), np.asarray(C.data, dtype=dt))
    else:
        D = np.asarray(A)
        if D.ndim != 2:
            raise ValueError("A must be two-dimensional")
        r, c = np.nonzero(D)
        csr = synth.from_coo(D.shape[0], D.shape[1], r, c, D[r, c].astype(dt))
    if tune and not api.is_complex(prec):
        t = api.autotune(csr, prec=prec)
        op = api.CsrOperator(csr, prec=prec, panels=t.panels, lanes=t.lanes, sort=t.sort)
    else:
        op = api.CsrOperator(csr, prec=prec)
    return op, csr.m, csr.n, op


def svdp(A, k, which="LM", irl_mode=True, kmax=None, compute_u=True, compute_v=True,
         v0=None, full_output=False, tol=0, delta=None, eta=None, anorm=0, cgs=False,
         elr=True, min_relgap=0.002, shifts=None, maxiter=None, rng=None, *, tune=False):
    """Truncated SVD of A through the library's xLANSVD(_IRL) drop-ins.

    Returns ``(u, sigma, vt)`` or ``(u, sigma, vt, bnd)`` with
    ``full_output`` -- the layout of scipy's ``_svdp``.  ``tune`` (not in
    scipy's signature): time the SpMV layouts of an explicit matrix on the
    device first and solve with the fastest (about a second for 32M
    nonzeros; the singular values agree with the untuned solve to
    rounding).
    """
    which = which.upper()
    if which not in _which:
        raise ValueError("`which` must be either 'LM' or 'SM'")
    if not irl_mode and which == "SM":
        raise ValueError("`which`='SM' requires irl_mode=True")
    dtype = np.dtype(getattr(A, "dtype", np.float64))
    if dtype in (np.float32, np.complex64, np.complex128):
        prec = api.prec_of(dtype)
    elif np.issubdtype(dtype, np.complexfloating):
        prec = api.PK_C128
    else:
        prec = api.PK_F64
    cplx = api.is_complex(prec)
    aprod, m, n, keep = _as_operator(A, prec, tune)
    if k < 1 or k > min(m, n):
        raise ValueError("k must be positive and not greater than m or n")
    if kmax is None:
        kmax = 10 * k
    if maxiter is None:
        maxiter = 1000
    kmax = min(m + 1, n + 1, kmax)
    if kmax < k:
        raise ValueError(f"kmax must be greater than or equal to k, but kmax ({kmax}) < k ({k})")
    if v0 is None:
        if rng is None:
            rng = np.random.default_rng()
        u0 = rng.uniform(size=m)
        if cplx:
            u0 = u0 + 1j * rng.uniform(size=m)
    else:
        u0 = np.asarray(v0, dtype=np.complex128 if cplx else np.float64)
        if u0.shape != (m,):
            raise ValueError(f"v0 must be of length {m}")
    eps = float(np.finfo(api._rt(prec)).eps)
    delta = np.sqrt(eps) if delta is None else delta
    eta = eps ** 0.75 if eta is None else eta
    jobu = "y" if compute_u else "n"
    jobv = "y" if compute_v else "n"
    if irl_mode:
        if shifts is None:
            shifts = kmax - k
        if k > min(kmax - shifts, m, n):
            raise ValueError("shifts must satisfy k <= min(kmax-shifts, m, n)!")
        if shifts < 0:
            raise ValueError("shifts must be >= 0!")
        r = api.lansvd_irl(aprod, m, n, k, kmax, p=shifts, u0=u0, tol=tol, which=_which[which],
                           jobu=jobu, jobv=jobv, maxiter=maxiter, prec=prec, delta=delta,
                           eta=eta, anorm=anorm, min_relgap=min_relgap, cgs=int(bool(cgs)),
                           elr=int(bool(elr)))
    else:
        r = api.lansvd(aprod, m, n, k, kmax, u0=u0, tol=tol, jobu=jobu, jobv=jobv, prec=prec,
                       delta=delta, eta=eta, anorm=anorm, cgs=int(bool(cgs)),
                       elr=int(bool(elr)))
    del keep
    if r.info > 0:
        raise LinAlgError(f"An invariant subspace of dimension {r.info} was found.")
    if r.info < 0:
        raise LinAlgError(f"k={k} singular triplets did not converge within kmax={kmax} "
                          "iterations")
    u = r.U if compute_u else None
    vt = (r.V.conj().T.copy() if cplx else r.V.T.copy()) if compute_v else None
    if full_output:
        return u, r.sigma, vt, r.bnd
    return u, r.sigma, vt


def svds(A, k=6, ncv=None, tol=0, which="LM", v0=None, maxiter=None,
         return_singular_vectors=True, solver="propack", rng=None, options=None):
    """``scipy.sparse.linalg.svds`` with ``solver='propack'`` on the MI355X
    library: the arguments and results of scipy's propack branch
    (scipy/sparse/linalg/_eigen/_svds.py, ``svds`` and ``_iv``) -- ``tol``
    squared for the Lanczos solve, the restarted driver only for
    ``which='SM'``, ``maxiter`` as the Lanczos dimension, singular values in
    ascending order (u's columns and vh's rows with them).  Explicit
    matrices stay on the device (scipy wraps them in a LinearOperator); a
    LinearOperator is called from the host.  ``ncv`` and ``options`` are
    accepted and ignored, as scipy's propack branch does."""
    if str(solver).lower() != "propack":
        raise ValueError("solver must be 'propack'")
    shape = getattr(A, "shape", None)
    if shape is None:
        A = np.asarray(A)
        shape = A.shape
    if len(shape) != 2:
        raise ValueError("A must be two-dimensional")
    dtype = np.dtype(getattr(A, "dtype", np.float64))
    if not np.issubdtype(dtype, np.number):
        raise ValueError("`A` must be of numeric data type.")
    if shape[0] * shape[1] == 0:
        raise ValueError("`A` must not be empty.")
    if int(k) != k or not (0 < k <= min(shape)):
        raise ValueError("`k` must be an integer satisfying `0 < k < min(A.shape)`.")
    k = int(k)
    if tol < 0 or not np.isfinite(tol):
        raise ValueError("`tol` must be a non-negative floating point value.")
    which = str(which).upper()
    if which not in ("LM", "SM"):
        raise ValueError("`which` must be in {'LM', 'SM'}.")
    if v0 is not None:
        v0 = np.atleast_1d(v0)
        if not (np.issubdtype(v0.dtype, np.complexfloating)
                or np.issubdtype(v0.dtype, np.floating)):
            raise ValueError("`v0` must be of floating or complex floating data type.")
        if v0.shape != (shape[0],):
            raise ValueError(f"`v0` must have shape {(shape[0],)}.")
    if maxiter is not None and (int(maxiter) != maxiter or maxiter <= 0):
        raise ValueError("`maxiter` must be a positive integer.")
    if return_singular_vectors not in (True, False, "vh", "u"):
        raise ValueError("`return_singular_vectors` must be in {True, False, 'vh', 'u'}.")
    jobu = return_singular_vectors in (True, "u")
    jobv = return_singular_vectors in (True, "vh")
    if rng is not None and not isinstance(rng, np.random.Generator):
        rng = np.random.default_rng(rng)
    u, s, vh, _ = svdp(A, k=k, tol=float(tol) ** 2, which=which, maxiter=None,
                       compute_u=jobu, compute_v=jobv, irl_mode=(which == "SM"),
                       kmax=None if maxiter is None else int(maxiter), v0=v0,
                       full_output=True, rng=rng)
    s = s[::-1]
    u = u[:, ::-1] if jobu else None
    vh = vh[::-1] if jobv else None
    if return_singular_vectors:
        return u, s, vh
    return s
