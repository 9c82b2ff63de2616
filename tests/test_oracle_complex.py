"""Pins the complex oracle (ZLANSVD / ZLANSVD_IRL and the CLANSVD twins
compiled from complex16/ and complex8/ into oracle/_ref) on the CPU, against
what the mathematics fixes:

* singular values against the dense SVD of the complex matrix (a transposed
  instead of conjugated A^H, or a dropped imaginary part, moves them);
* the triplets satisfy A v = s u and A^H u = s v, with orthonormal U, V;
* a complex matrix with zero imaginary parts reproduces DLANSVD's singular
  values (the complex recurrences then carry real numbers);
* ZLARNV(IDIST=2) is the real xLARNV sequence read in (re, im) pairs: the
  reference's random start (u0 = 0) and the product's host generator give
  the same start through ZGETU0.

It also records a defect of the reference that the product does not copy:
ZDGEMM (complex16/zgemm_ovwr.F:41-61), the kernel of ZDGEMM_OVWR_LEFT,
ignores TRANSB and always forms A * B^T.  ZRITZVEC passes 'T'
(complex16/zritzvec.F:155), so ZLANSVD is right, but ZLANSVD_IRL's restarts
pass 'N' (complex16/zlansvd_irl.F:389-397): every restart applies Q^T where
Q is meant, and a restarted run reports converged values that are not
singular values.  Runs without a restart are unaffected.
"""
import numpy as np
import pytest

from oracle import refpy
from propack_amd import synth

pytestmark = pytest.mark.skipif(not (refpy.available("z") and refpy.available("c")),
                                reason="complex oracle not built (make -C oracle ref)")


@pytest.fixture(scope="module")
def zcase():
    refpy.set_threads(1, "z")
    refpy.set_threads(1, "c")
    A = synth.complex_uniform()
    D = A.todense()
    s = np.linalg.svd(D, compute_uv=False)
    u0 = synth.start_vector(A.m, 7, np.complex128)
    return A, D, s, u0


def _triplets_ok(r, D, s1, tol_res, tol_orth):
    res = np.linalg.norm(D @ r.V - r.U * r.sigma, axis=0).max()
    resh = np.linalg.norm(D.conj().T @ r.U - r.V * r.sigma, axis=0).max()
    k = r.k
    orth = max(np.abs(r.U.conj().T @ r.U - np.eye(k)).max(),
               np.abs(r.V.conj().T @ r.V - np.eye(k)).max())
    assert res <= tol_res * s1, res
    assert resh <= 1e3 * tol_res * s1, resh
    assert orth <= tol_orth, orth


def test_complex_values_are_complex(zcase):
    A, D, s, u0 = zcase
    assert np.abs(A.val.imag).max() > 1.0
    # the dense SVD of the complex matrix is not that of its real part
    sr = np.linalg.svd(D.real, compute_uv=False)
    assert abs(sr[0] - s[0]) > 1e-3 * s[0]


def test_zlansvd_against_dense_svd(zcase):
    A, D, s, u0 = zcase
    r = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, prec="z")
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-12)
    _triplets_ok(r, D, s[0], 1e-11, 1e-10)


def test_zlansvd_random_start(zcase):
    A, D, s, u0 = zcase
    r = refpy.lansvd(A, 8, 300, u0=None, tol=1e-12, prec="z")
    assert r.info == 0 and r.k == 8
    np.testing.assert_allclose(r.sigma, s[:8], rtol=1e-12)


def test_clansvd_against_dense_svd(zcase):
    A, D, s, u0 = zcase
    r = refpy.lansvd(A.astype(np.complex64), 10, 300, u0=u0.astype(np.complex64), tol=1e-5,
                     prec="c")
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-5)


def test_zlansvd_real_data_matches_dlansvd():
    A = synth.random_uniform()
    Az = synth.Csr(A.m, A.n, A.rowptr, A.col, A.val.astype(np.complex128))
    u0 = synth.start_vector(A.m, 7)
    rd = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, prec="d")
    rz = refpy.lansvd(Az, 10, 300, u0=u0.astype(np.complex128), tol=1e-12, prec="z")
    assert rz.info == 0 and rz.k == rd.k == 10
    np.testing.assert_allclose(rz.sigma, rd.sigma, rtol=1e-13)


def test_zlarnv_is_the_real_sequence_in_pairs():
    """ZGETU0's random vector (complex16/zgetu0.F:69) for a u0 = 0 start:
    the operator applied to the pairs of the real DLARNV sequence.  With
    A = I (square) and one Lanczos step, ZLANSVD's left vector is that
    vector scaled."""
    from propack_amd import api
    n = 300
    A = synth.Csr(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
                  np.ones(n, np.complex128))
    r = refpy.lansvd(A, 1, 1, u0=None, tol=1e-12, prec="z", jobu="y", jobv="n")
    x = np.zeros(2 * n)
    api.lib().pk_larnv_pm1(api.C.c_int(api.PK_F64), api.C.c_longlong(2 * n), api._p(x))
    z = x[0::2] + 1j * x[1::2]
    u = r.U[:, 0]
    cos = abs(np.vdot(u, z)) / (np.linalg.norm(u) * np.linalg.norm(z))
    assert cos > 1 - 1e-12, cos


def test_zlansvd_irl_without_restart_is_right(zcase):
    A, D, s, u0 = zcase
    r = refpy.lansvd_irl(A, 6, 120, p=20, u0=u0, tol=1e-10, prec="z")
    assert r.info == 0 and r.counters["nrestart"] == 0
    np.testing.assert_allclose(r.sigma, s[:6], rtol=1e-10)


def test_reference_zcgs_at_one_thread(zcase):
    """ZLANSVD with iterated classical Gram-Schmidt (ioption(1) = 1) at one
    OpenMP thread: right values (its 16-thread run races in ZCGS's interval
    loop, complex16/zreorth.F:140-201, and is not used as the oracle)."""
    A, D, s, u0 = zcase
    r = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=1, prec="z")
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-12)


def test_reference_zlansvd_irl_restarts_apply_q_transposed(zcase):
    """The documented defect: with restarts the reference's values are off
    by percents, not by rounding."""
    A, D, s, u0 = zcase
    r = refpy.lansvd_irl(A, 10, 30, p=10, u0=u0, tol=1e-10, prec="z")
    assert r.counters["nrestart"] > 0
    err = np.abs(r.sigma - s[:r.k]).max() / s[0]
    assert err > 1e-3, err
