"""GPU parity of the complex precisions: ZLANSVD / ZLANSVD_IRL and the
CLANSVD twins through the library's Fortran drop-ins, against the reference
compiled from complex16/ and complex8/ (oracle/_ref) and the dense SVD.

Tolerances: ZLANSVD singular values 1e-10 relative to the oracle's (and
1e-12 to the dense SVD's), residuals ||A v - s u|| <= 1e-11 s_1, singular
vectors within 1e-8 of the oracle's up to a unit complex phase (|u^H u_ref|),
the reference's counters exactly; the C twins 1e-5.  The reference's
restarted ZLANSVD_IRL applies Q^T where Q is meant (ZDGEMM ignores TRANSB,
complex16/zgemm_ovwr.F:41-61; pinned in test_oracle_complex.py), so
restarted runs are checked against the dense SVD, and on a complex matrix
with zero imaginary parts against the product's real path; runs without a
restart against the oracle."""
import numpy as np
import pytest

from oracle import refpy
from propack_amd import api, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zcase():
    # One reference thread: with 16 OpenMP threads the reference's complex
    # CGS (ZCGS, complex16/zreorth.F:140-201) races -- thread 0 refills the
    # shared projection WORK for the next INDEX interval while other threads
    # may still read it for the previous interval's update -- and loses
    # orthogonality (INFO = -1 here); at 1 or 4 threads it gives the counts
    # the device path gives.
    refpy.set_threads(1, "z")
    refpy.set_threads(1, "c")
    A = synth.complex_uniform()
    D = A.todense()
    s = np.linalg.svd(D, compute_uv=False)
    u0 = synth.start_vector(A.m, 7, np.complex128)
    return A, D, s, u0


def _pcos(a, b):
    """|a_j^H b_j| / (|a_j| |b_j|) per column: 1 for vectors equal up to a
    unit complex factor."""
    return np.abs(np.sum(a.conj() * b, axis=0)) / (np.linalg.norm(a, axis=0) *
                                                    np.linalg.norm(b, axis=0))


def _check(r, D, s1, tol_res, tol_orth):
    res = np.linalg.norm(D @ r.V - r.U * r.sigma, axis=0).max()
    assert res <= tol_res * s1, (res, tol_res * s1)
    k = r.k
    assert np.abs(r.U.conj().T @ r.U - np.eye(k)).max() < tol_orth
    assert np.abs(r.V.conj().T @ r.V - np.eye(k)).max() < tol_orth


def test_complex_spmv(zcase):
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    x = synth.start_vector(A.n, 3, np.complex128)
    y = synth.start_vector(A.m, 4, np.complex128)
    np.testing.assert_allclose(op.spmv(x, "n"), D @ x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(op.spmv(y, "t"), D.T @ y, rtol=0, atol=1e-12)
    np.testing.assert_allclose(op.spmv(y, "c"), D.conj().T @ y, rtol=0, atol=1e-12)
    op32 = api.CsrOperator(A, api.PK_C64)
    np.testing.assert_allclose(op32.spmv(x.astype(np.complex64), "n"), D @ x, rtol=0, atol=1e-4)
    np.testing.assert_allclose(op32.spmv(y.astype(np.complex64), "c"), D.conj().T @ y,
                               rtol=0, atol=1e-4)


@pytest.mark.parametrize("panels", [(2, 2), (4, 8), (8, 4)])
def test_complex_panels(zcase, panels):
    """Column panels of a complex operator (every panel in one launch, the
    panels' FP64 row sums added in panel order): products against the dense
    matrix and the one-panel operator, deterministic, and a ZLANSVD solve
    with the one-panel solve's values and operator count."""
    A, D, s, u0 = zcase
    op1 = api.CsrOperator(A, api.PK_C128)
    opp = api.CsrOperator(A, api.PK_C128, panels=panels)
    assert opp.panels() == panels
    x = synth.start_vector(A.n, 5, np.complex128)
    y = synth.start_vector(A.m, 6, np.complex128)
    for t, v, ref in (("n", x, D @ x), ("c", y, D.conj().T @ y), ("t", y, D.T @ y)):
        a = opp.spmv(v, t)
        np.testing.assert_allclose(a, ref, rtol=0, atol=1e-12)
        np.testing.assert_allclose(a, op1.spmv(v, t), rtol=0, atol=1e-13)
        assert np.array_equal(a, opp.spmv(v, t))
    r1 = api.lansvd(op1, A.m, A.n, 10, 300, u0=u0, tol=1e-12, prec=api.PK_C128)
    rp = api.lansvd(opp, A.m, A.n, 10, 300, u0=u0, tol=1e-12, prec=api.PK_C128)
    assert rp.info == 0 and rp.k == 10
    np.testing.assert_allclose(rp.sigma, r1.sigma, rtol=1e-12)
    assert rp.counters["nopx"] == r1.counters["nopx"]
    op32 = api.CsrOperator(A.astype(np.complex64), api.PK_C64, panels=panels)
    np.testing.assert_allclose(op32.spmv(y.astype(np.complex64), "c"), D.conj().T @ y,
                               rtol=0, atol=1e-4)


def test_complex_operator_refuses_layout_calls(zcase):
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    for call in (lambda: op.set_sort(8192, 8192), lambda: op.set_lanes(1, 1)):
        with pytest.raises(api.PropackError, match="column panels only"):
            call()
    # short rows (4.7 per row of A): the built-in choice keeps one panel
    assert op.panels() == (1, 1)


@pytest.mark.parametrize("cgs", [1, 0])
def test_zlansvd_device_operator(zcase, cgs):
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    r = api.lansvd(op, A.m, A.n, 10, 300, u0=u0, tol=1e-12, cgs=cgs, prec=api.PK_C128)
    ref = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=cgs, prec="z")
    print("gpu", r.info, r.k, r.counters)
    print("ref", ref.info, ref.k, ref.counters)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-12)
    _check(r, D, s[0], 1e-11, 1e-9)
    assert _pcos(r.U, ref.U).min() > 1 - 1e-8
    assert _pcos(r.V, ref.V).min() > 1 - 1e-8
    for key in ("nopx", "nreorth", "ndot", "nbsvd", "nlandim", "nsing"):
        assert r.counters[key] == ref.counters[key], key
    assert abs(r.anorm - ref.anorm) <= 1e-10 * ref.anorm


def test_zlansvd_random_start(zcase):
    """u0 = 0: ZGETU0's ZLARNV start, generated on the device."""
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    r = api.lansvd(op, A.m, A.n, 8, 300, u0=None, tol=1e-12, prec=api.PK_C128)
    ref = refpy.lansvd(A, 8, 300, u0=None, tol=1e-12, prec="z")
    assert r.info == 0 and r.k == 8
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    assert _pcos(r.U, ref.U).min() > 1 - 1e-8
    assert r.counters["nopx"] == ref.counters["nopx"]


def test_zlansvd_host_callback_matches_device_operator(zcase):
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    rd = api.lansvd(op, A.m, A.n, 6, 300, u0=u0, tol=1e-12, prec=api.PK_C128)

    def aprod(t, x):
        return D @ x if t == "n" else D.conj().T @ x

    rh = api.lansvd(aprod, A.m, A.n, 6, 300, u0=u0, tol=1e-12, prec=api.PK_C128)
    assert rh.info == 0 and rh.k == 6
    np.testing.assert_allclose(rh.sigma, rd.sigma, rtol=1e-11)
    assert rh.counters["nopx"] == rd.counters["nopx"]


def test_zlansvd_irl_without_restart_matches_reference(zcase):
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    r = api.lansvd_irl(op, A.m, A.n, 6, 120, p=20, u0=u0, tol=1e-10, prec=api.PK_C128)
    ref = refpy.lansvd_irl(A, 6, 120, p=20, u0=u0, tol=1e-10, prec="z")
    assert r.info == 0 and ref.counters["nrestart"] == 0
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    assert _pcos(r.U, ref.U).min() > 1 - 1e-8
    for key in ("nopx", "nreorth", "ndot", "nrestart"):
        assert r.counters[key] == ref.counters[key], key


@pytest.mark.parametrize("which", ["L", "S"])
def test_zlansvd_irl_restarted_against_dense(zcase, which):
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    r = api.lansvd_irl(op, A.m, A.n, 10, 30, p=10, u0=u0, tol=1e-10, which=which,
                       prec=api.PK_C128)
    print(r.info, r.k, r.counters)
    assert r.info == 0 and r.k == 10 and r.counters["nrestart"] > 0
    want = s[:10] if which == "L" else np.sort(s[-10:])[::-1]
    got = np.sort(r.sigma)[::-1]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * s[0])
    _check(r, D, s[0], 1e-9, 1e-8)


def test_zlansvd_irl_real_data_matches_dlansvd_irl():
    """On a complex matrix with zero imaginary parts and a real start the
    complex path's Lanczos vectors stay real (only DGETU0's norm estimate
    draws a complex vector): its restarted run finds the real path's
    values."""
    A = synth.random_uniform()
    Az = synth.Csr(A.m, A.n, A.rowptr, A.col, A.val.astype(np.complex128))
    u0 = synth.start_vector(A.m, 7)
    rd = api.lansvd_irl(api.CsrOperator(A), A.m, A.n, 10, 30, p=10, u0=u0, tol=1e-10)
    rz = api.lansvd_irl(api.CsrOperator(Az, api.PK_C128), A.m, A.n, 10, 30, p=10,
                        u0=u0.astype(np.complex128), tol=1e-10, prec=api.PK_C128)
    assert rz.info == 0 and rz.k == rd.k
    assert rz.counters["nrestart"] > 0 and rd.counters["nrestart"] > 0
    np.testing.assert_allclose(rz.sigma, rd.sigma, rtol=1e-10)


def test_clansvd_and_irl(zcase):
    A, D, s, u0 = zcase
    A32 = A.astype(np.complex64)
    op = api.CsrOperator(A32, api.PK_C64)
    u32 = u0.astype(np.complex64)
    r = api.lansvd(op, A.m, A.n, 10, 300, u0=u32, tol=1e-5, prec=api.PK_C64)
    ref = refpy.lansvd(A32, 10, 300, u0=u32, tol=1e-5, prec="c")
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-5)
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-5)
    ri = api.lansvd_irl(op, A.m, A.n, 10, 30, p=10, u0=u32, tol=1e-5, prec=api.PK_C64)
    assert ri.info == 0 and ri.k == 10
    np.testing.assert_allclose(ri.sigma, s[:10], rtol=1e-5)


def test_complex_engine_refuses_real_only_calls(zcase):
    A, D, s, u0 = zcase
    op = api.CsrOperator(A, api.PK_C128)
    eng = api.Engine(op, 20)
    with pytest.raises(api.PropackError):
        api.Batch([eng])
