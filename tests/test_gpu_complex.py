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


# ---- the reference's rare and failure paths, complex tree (one reference
# thread; INFO, K, nopx and nreorth exactly, sigma to 1e-10 of sigma_1) ----

def _zsame(r, ref, sig_rtol=1e-10, ndot_rtol=0.0):
    print("gpu", r.info, r.k, r.anorm, r.counters)
    print("ref", ref.info, ref.k, ref.anorm, ref.counters)
    assert r.info == ref.info and r.k == ref.k
    if ref.k:
        assert np.abs(r.sigma - ref.sigma).max() <= sig_rtol * ref.sigma.max()
    for key in ("nopx", "nreorth"):
        assert r.counters[key] == ref.counters[key], key
    assert abs(r.counters["ndot"] - ref.counters["ndot"]) <= ndot_rtol * ref.counters["ndot"]


def _dense_csr(D):
    rows, cols = np.nonzero(np.ones(D.shape))
    return synth.from_coo(D.shape[0], D.shape[1], rows, cols, D.ravel())


def test_zlansvd_kmax_exhausted(zcase):
    """kmax too small for k=10 at tol 1e-12: INFO = -1 with K converged
    (complex16/zlansvd.F:250-258)."""
    A, D, s, u0 = zcase
    r = api.lansvd(api.CsrOperator(A, api.PK_C128), A.m, A.n, 10, 20, u0=u0, tol=1e-12,
                   prec=api.PK_C128)
    ref = refpy.lansvd(A, 10, 20, u0=u0, tol=1e-12, cgs=1, prec="z")
    assert ref.info == -1
    _zsame(r, ref)


def test_zlansvd_wide_matrix(zcase):
    """m < n (A^H of the test matrix): DLANBPRO's norm estimate draws its
    random vector through A instead of A^H (complex16/zlanbpro.F:224-228)."""
    A, D, s, u0 = zcase
    T = A.transpose()
    AH = synth.Csr(T.m, T.n, T.rowptr, T.col, T.val.conj())
    v0 = synth.start_vector(AH.m, 9, np.complex128)
    r = api.lansvd(api.CsrOperator(AH, api.PK_C128), AH.m, AH.n, 8, 300, u0=v0, tol=1e-12,
                   prec=api.PK_C128)
    ref = refpy.lansvd(AH, 8, 300, u0=v0, tol=1e-12, cgs=1, prec="z")
    _zsame(r, ref)
    np.testing.assert_allclose(r.sigma, s[:8], rtol=1e-12)


def test_zlansvd_five_failed_cgs_passes():
    """40 x 20 complex with k = 20, kmax = 21: at step 21 the right Lanczos
    vectors span C^20 and after five CGS passes v_21 = 0
    (complex16/zreorth.F:85-99)."""
    B = synth.complex_uniform(40, 20, 400, seed=9)
    v0 = synth.start_vector(40, 2, np.complex128)
    s = np.linalg.svd(B.todense(), compute_uv=False)
    op = api.CsrOperator(B, api.PK_C128)
    r = api.lansvd(op, B.m, B.n, 20, 21, u0=v0, tol=1e-12, prec=api.PK_C128)
    ref = refpy.lansvd(B, 20, 21, u0=v0, tol=1e-12, cgs=1, prec="z")
    _zsame(r, ref)
    np.testing.assert_allclose(r.sigma, s, rtol=1e-12)
    eng = api.Engine(op, kmax=21)
    eng.put_u1(v0)
    api.lansvd_dev(eng, 20, 21, tol=1e-12)
    assert eng.path_stats()["v_zero"] >= 1


def test_zlansvd_invariant_subspace_new_direction():
    """Complex rank 15 exactly, k = 20: alpha falls below anorm*epsn and the
    step draws a new random direction (complex16/zlanbpro.F:390-427); the 15
    nonzero singular values match the reference, K within one (how many of
    the zero values count as converged is decided at the rounding level)."""
    g = np.random.Generator(np.random.Philox(4))
    m, n, rk = 200, 80, 15
    D = (g.standard_normal((m, rk)) + 1j * g.standard_normal((m, rk))) @ \
        (g.standard_normal((rk, n)) + 1j * g.standard_normal((rk, n)))
    A = _dense_csr(D)
    u0 = synth.start_vector(m, 6, np.complex128)
    op = api.CsrOperator(A, api.PK_C128)
    r = api.lansvd(op, m, n, 20, 40, u0=u0, tol=1e-12, prec=api.PK_C128)
    ref = refpy.lansvd(A, 20, 40, u0=u0, tol=1e-12, cgs=1, prec="z")
    print("gpu", r.info, r.k, r.counters)
    print("ref", ref.info, ref.k, ref.counters)
    assert r.info == ref.info and abs(r.k - ref.k) <= 1 and r.k >= rk
    np.testing.assert_allclose(r.sigma[:rk], ref.sigma[:rk], rtol=1e-10)
    eng = api.Engine(op, kmax=40)
    eng.put_u1(u0)
    api.lansvd_dev(eng, 20, 40, tol=1e-12)
    ps = eng.path_stats()
    assert ps["v_invariant"] + ps["u_invariant"] >= 1


def test_zlansvd_nan_operator(zcase):
    """A NaN value in the complex operator: INFO = 0 with K = 0 and a NaN
    ||A|| estimate, as the reference's control flow on NaN comparisons
    gives; the engine works afterwards."""
    A, D, s, u0 = zcase
    An = synth.Csr(A.m, A.n, A.rowptr, A.col, A.val.copy())
    An.val[100] = complex(np.nan, 0.0)
    ref = refpy.lansvd(An, 10, 300, u0=u0, tol=1e-12, cgs=1, prec="z")
    r = api.lansvd(api.CsrOperator(An, api.PK_C128), A.m, A.n, 10, 300, u0=u0, tol=1e-12,
                   prec=api.PK_C128)
    print("gpu", r.info, r.k, r.anorm, r.counters)
    print("ref", ref.info, ref.k, ref.anorm, ref.counters)
    assert r.info == ref.info and r.k == ref.k == 0
    assert np.isnan(r.anorm) and np.isnan(ref.anorm)
    assert r.counters["nopx"] == ref.counters["nopx"]
    r2 = api.lansvd(api.CsrOperator(A, api.PK_C128), A.m, A.n, 10, 300, u0=u0, tol=1e-12,
                    prec=api.PK_C128)
    np.testing.assert_allclose(r2.sigma, s[:10], rtol=1e-12)


def test_clansvd_tiny_alpha(zcase):
    """CLANSVD on the test matrix scaled by 1e-39 (complex64 values): every
    alpha and beta lies below FP32's safe minimum, so each scaling takes
    the stepwise path of CSAFESCAL (complex8/csafescal.F); the engine
    accumulates in FP64, and the values match the dense SVD of the stored
    matrix to FP32 storage precision."""
    A, D, s, u0 = zcase
    As = synth.Csr(A.m, A.n, A.rowptr, A.col, (A.val * 1e-39).astype(np.complex64))
    sd = np.linalg.svd(As.todense(), compute_uv=False)
    op = api.CsrOperator(As, api.PK_C64)
    eng = api.Engine(op, kmax=300)
    eng.put_u1(u0)
    r = api.lansvd_dev(eng, 10, 300, tol=1e-6)
    ps = eng.path_stats()
    print("paths", ps, "info", r.info, "k", r.k)
    assert ps["v_tiny"] >= 1 and ps["u_tiny"] >= 1
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, sd[:10], rtol=1e-5)
