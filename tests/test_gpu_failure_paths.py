"""GPU parity on the reference's rare and failure paths: the library's
drop-in drivers (and the device-resident drivers, whose engine reports
which rare DLANBPRO path ran) against the reference PROPACK compiled from
its own sources, on small inputs that force each path.

The reference's OpenMP build gives thread-count-dependent results on
vectors of a few dozen elements (8 threads: spurious INFO=-1 on the
40 x 20 case below), so these comparisons run it on one thread.

Tolerances: INFO, K and the counters nopx / nreorth / ndot exactly;
singular values 1e-10 relative; doption(3) (the ||A|| estimate written
back, double/dlanbpro.F:547) 1e-12 relative."""
import numpy as np
import pytest

from oracle import refpy
from propack_amd import api, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _one_thread_reference():
    refpy.set_threads(1)
    yield
    refpy.set_threads(8)


def _same(r, ref, sig_rtol=1e-10, ndot_rtol=0.0):
    """Singular values within sig_rtol * sigma_1 (the absolute accuracy
    Lanczos gives every singular value, noise-level ones included)."""
    print("gpu", r.info, r.k, r.anorm, r.counters)
    print("ref", ref.info, ref.k, ref.anorm, ref.counters)
    assert r.info == ref.info
    assert r.k == ref.k
    if ref.k:
        assert np.abs(r.sigma - ref.sigma).max() <= sig_rtol * ref.sigma.max()
    for key in ("nopx", "nreorth"):
        assert r.counters[key] == ref.counters[key], key
    assert abs(r.counters["ndot"] - ref.counters["ndot"]) <= ndot_rtol * ref.counters["ndot"]
    assert abs(r.anorm - ref.anorm) <= 1e-12 * abs(ref.anorm)


def _near_invariant():
    """300 x 120: rank 15 plus 1e-9 noise.  Once the Lanczos process has
    used up the rank-15 part, alpha falls to ~1e-9 ||A||: between
    anorm*epsn and anorm*eps/delta, the reference's near-invariance
    warning (IERR = j, double/dlanbpro.F:406-409)."""
    g = np.random.Generator(np.random.Philox(3))
    m, n, r = 300, 120, 15
    D = g.standard_normal((m, r)) @ g.standard_normal((r, n)) + 1e-9 * g.standard_normal((m, n))
    rows, cols = np.nonzero(np.ones((m, n)))
    return synth.from_coo(m, n, rows, cols, D.ravel()), D, synth.start_vector(m, 5)


def test_kmax_exhausted_info_minus_one(small_case):
    """kmax too small for k=10 at tol 1e-12: INFO = -1 and K = number
    converged (double/dlansvd.F:250-258)."""
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    r = api.lansvd(op, A.m, A.n, 10, 20, u0=u0, tol=1e-12)
    ref = refpy.lansvd(A, 10, 20, u0=u0, tol=1e-12, cgs=1)
    assert ref.info == -1
    _same(r, ref)


@pytest.mark.parametrize("delta,eta", [(0.0, None), (1e-13, 1e-12)])
def test_full_reorthogonalization(small_case, delta, eta):
    """delta = 0 or delta <= eta switch DLANBPRO to full reorthogonalization
    (double/dlanbpro.F:165-169)."""
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    r = api.lansvd(op, A.m, A.n, 10, 300, u0=u0, tol=1e-12, delta=delta, eta=eta)
    ref = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=1, delta=delta, eta=eta)
    assert ref.info == 0 and ref.counters["nreorth"] > ref.counters["nopx"] // 2 - 2
    _same(r, ref)
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-12)


def test_five_failed_cgs_passes():
    """40 x 20 with k = 20, kmax = 21: at step 21 the 20 right Lanczos
    vectors span R^20, every CGS pass leaves rounding noise in their span,
    and after five passes v_21 = 0 (double/dreorth.F:96-98)."""
    B = synth.random_uniform(40, 20, 400, seed=9)
    v0 = synth.start_vector(40, 2)
    s = np.linalg.svd(B.todense(), compute_uv=False)
    op = api.CsrOperator(B)
    r = api.lansvd(op, B.m, B.n, 20, 21, u0=v0, tol=1e-12)
    ref = refpy.lansvd(B, 20, 21, u0=v0, tol=1e-12, cgs=1)
    _same(r, ref)
    np.testing.assert_allclose(r.sigma, s, rtol=1e-12)
    eng = api.Engine(op, kmax=21)
    eng.put_u1(v0)
    rd = api.lansvd_dev(eng, 20, 21, tol=1e-12)
    ps = eng.path_stats()
    print("paths", ps)
    assert ps["v_zero"] >= 1
    assert rd.info == ref.info and rd.k == ref.k
    np.testing.assert_allclose(rd.sigma, ref.sigma, rtol=1e-10)


def test_near_invariance_warning():
    """Past the rank the Lanczos vectors are ~1e-9 ||A|| residues, and
    whether a CGS pass keeps kappa = 0.717 of the norm is decided at the
    rounding level: the reference projects each INDEX interval in turn, the
    engine all intervals in one block (equal in exact arithmetic), so the
    number of second passes, and ndot with it, may differ (10 %)."""
    A, D, u0 = _near_invariant()
    op = api.CsrOperator(A)
    r = api.lansvd(op, A.m, A.n, 20, 60, u0=u0, tol=1e-12)
    ref = refpy.lansvd(A, 20, 60, u0=u0, tol=1e-12, cgs=1)
    _same(r, ref, ndot_rtol=0.1)
    eng = api.Engine(op, kmax=60)
    eng.put_u1(u0)
    rd = api.lansvd_dev(eng, 20, 60, tol=1e-12)
    ps = eng.path_stats()
    print("paths", ps)
    assert ps["nwarn"] >= 1
    assert rd.info == ref.info and rd.k == ref.k


def test_invariant_subspace_new_direction():
    """Rank 15 exactly (no noise), k = 20: alpha drops below anorm*epsn, the
    step draws a new random direction (double/dlanbpro.F:372-409).  The 15
    nonzero singular values match the reference; past the rank, how many
    of the zero singular values count as converged depends on rounding
    (the reference projects each INDEX interval in turn, the engine all
    intervals in one block: equal in exact arithmetic), so K may differ
    by one."""
    g = np.random.Generator(np.random.Philox(4))
    m, n, rk = 200, 80, 15
    D = g.standard_normal((m, rk)) @ g.standard_normal((rk, n))
    rows, cols = np.nonzero(np.ones((m, n)))
    A = synth.from_coo(m, n, rows, cols, D.ravel())
    u0 = synth.start_vector(m, 6)
    op = api.CsrOperator(A)
    r = api.lansvd(op, m, n, 20, 40, u0=u0, tol=1e-12)
    ref = refpy.lansvd(A, 20, 40, u0=u0, tol=1e-12, cgs=1)
    print("gpu", r.info, r.k, r.counters)
    print("ref", ref.info, ref.k, ref.counters)
    assert r.info == ref.info and abs(r.k - ref.k) <= 1 and r.k >= rk
    np.testing.assert_allclose(r.sigma[:rk], ref.sigma[:rk], rtol=1e-10)
    eng = api.Engine(op, kmax=40)
    eng.put_u1(u0)
    api.lansvd_dev(eng, 20, 40, tol=1e-12)
    ps = eng.path_stats()
    print("paths", ps)
    assert ps["v_invariant"] + ps["u_invariant"] >= 1


@pytest.mark.parametrize("where", ["device", "host"])
def test_nan_operator(small_case, where):
    """A NaN in the operator (a device CSR value, or a host APROD that
    returns it): the reference's control flow on NaN comparisons gives
    INFO = 0 with K = 0 converged and a NaN ||A|| estimate; the drivers
    return the same, and every kernel exits normally."""
    A, u0, D, s = small_case
    An = synth.Csr(A.m, A.n, A.rowptr, A.col, A.val.copy())
    An.val[100] = np.nan
    ref = refpy.lansvd(An, 10, 300, u0=u0, tol=1e-12, cgs=1)
    if where == "device":
        r = api.lansvd(api.CsrOperator(An), A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    else:
        Dn = An.todense()

        def aprod(t, x):
            return Dn @ x if t == "n" else Dn.T @ x

        r = api.lansvd(aprod, A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    print("gpu", r.info, r.k, r.anorm, r.counters)
    print("ref", ref.info, ref.k, ref.anorm, ref.counters)
    assert r.info == ref.info and r.k == ref.k == 0
    assert np.isnan(r.anorm) and np.isnan(ref.anorm)
    assert r.counters["nopx"] == ref.counters["nopx"]
    # the engine still works afterwards
    r2 = api.lansvd(api.CsrOperator(A), A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    np.testing.assert_allclose(r2.sigma, s[:10], rtol=1e-12)


def test_tiny_alpha_single_precision(small_case):
    """SLANSVD on the 1850 x 712 matrix scaled by 1e-39 (FP32 values): every
    alpha and beta lies below FP32's safe minimum 2^-126, so each scaling
    takes DSAFESCAL's stepwise path (double/dsafescal.F:49-53).  The engine
    accumulates in FP64, so the norms stay exact and the singular values
    match the dense SVD of the stored matrix to FP32 storage precision.  (The
    reference's own S build returns NaN here, already at a 1e-36 scale: its
    FP32 arithmetic underflows; so the dense SVD, not the reference, is the
    oracle.  In FP64 the path cannot be reached: a sum of squares
    underflows long before alpha < 2^-1022.)"""
    A, u0, D, s = small_case
    c = 1e-39
    As = synth.Csr(A.m, A.n, A.rowptr, A.col, (A.val * c).astype(np.float32))
    sd = np.linalg.svd(As.todense(), compute_uv=False)
    op = api.CsrOperator(As, prec=api.PK_F32)
    eng = api.Engine(op, kmax=300)
    eng.put_u1(u0)
    r = api.lansvd_dev(eng, 10, 300, tol=1e-6)
    ps = eng.path_stats()
    print("paths", ps, "info", r.info, "k", r.k)
    assert ps["v_tiny"] >= 1 and ps["u_tiny"] >= 1
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, sd[:10], rtol=1e-5)


def test_irl_count_neig_above_kept(small_case):
    """neig > dim - p: the converged count of DLANSVD_IRL jumps to i = k+1
    after an unconverged value among the first k (double/dlansvd_irl.F:
    279-289).  Where the reference ends, the drivers agree with it; where
    an unconverged value lies past k the reference sets i = k+1 again and
    never leaves its loop (so scipy rejects k > kmax - shifts), and the
    drivers stop the count there instead."""
    # 150 x 70 diagonal, sigma_i = 10 * 0.85^i: the reference ends
    m, n = 150, 70
    sig = 10.0 * 0.85 ** np.arange(n)
    A = synth.from_coo(m, n, np.arange(n), np.arange(n), sig)
    u0 = synth.start_vector(m, 11)
    u0[0] = 1e-8
    r = api.lansvd_irl(api.CsrOperator(A), m, n, 12, 30, p=22, u0=u0, tol=1e-10, maxiter=2)
    ref = refpy.lansvd_irl(A, 12, 30, p=22, u0=u0, tol=1e-10, cgs=1, maxiter=2)
    print("gpu", r.info, r.k, r.counters)
    print("ref", ref.info, ref.k, ref.counters)
    assert r.info == ref.info and r.k == ref.k == 12
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    assert r.counters["nopx"] == ref.counters["nopx"]
    # the 1850 x 712 case, dim 30, p 22 (k = 8), neig 12: the reference loops
    # forever in its count; the driver returns, with true singular values
    A, u0, D, s = small_case
    r = api.lansvd_irl(api.CsrOperator(A), A.m, A.n, 12, 30, p=22, u0=u0, tol=1e-12,
                       maxiter=5)
    print("C1", r.info, r.k, r.sigma)
    assert r.k >= 1
    for v in r.sigma:
        assert np.min(np.abs(s - v)) <= 1e-10 * s[0]
