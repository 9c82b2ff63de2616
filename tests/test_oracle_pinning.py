"""Pins the oracle (compiled reference in oracle/_ref and the C restatement in
oracle/port) and the product's host-side scalar routines, on the CPU.

* DLARNV(IDIST=2) from the fixed seed: the reference, the restatement and
  the product are checked bit for bit against the closed form computed
  with exact integer arithmetic.
* The omega-recurrence routines (DUPDATE_MU/NU, DCOMPUTE_INT, DSET_MU) and
  the bidiagonal helpers (DLARTG, DBDQR, DBSVDSTEP, DREFINEBOUNDS): the
  restatement and the product's exported host routines equal the
  reference's output bit for bit on seeded inputs.
* End to end: the restatement reproduces the reference's counters exactly
  and its singular values to 1e-13 on the 1850 x 712 case (single thread,
  same BLAS), and both match the dense FP64 SVD.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import portpy, refpy
from propack_amd import api, synth

pytestmark = pytest.mark.skipif(not (refpy.available("d") and portpy.available()),
                                reason="oracle not built (make -C oracle)")

D = C.c_double
I = C.c_int


def P(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.fixture(scope="module")
def ref():
    refpy.set_threads(1)
    return refpy.lib("d")


@pytest.fixture(scope="module")
def port():
    return portpy.lib()


@pytest.fixture(scope="module")
def prod():
    return api.lib()


def closed_form_larnv(n):
    """Exact closed form of DLARNV(2, (1,3,5,7), n): the 48-bit MCG with
    multiplier 33952834046453; output g is 2*(s0*a^(g+1) mod 2^48)/2^48 - 1."""
    a, mod = 33952834046453, 1 << 48
    s = (1 << 36) + (3 << 24) + (5 << 12) + 7
    out = np.empty(n)
    for g in range(n):
        s = (s * a) % mod
        out[g] = 2.0 * (s / float(mod)) - 1.0  # s < 2^48: exact in double
    return out


def test_larnv_closed_form(ref, port, prod):
    n = 1000  # spans several 64-element DLARNV chunks
    exact = closed_form_larnv(n)
    x = np.zeros(n)
    ref.dlarnv_(C.byref(I(2)), P(np.array([1, 3, 5, 7], np.int32)), C.byref(I(n)), P(x))
    assert np.array_equal(x, exact)
    assert np.array_equal(portpy.larnv_pm1(n), exact)
    y = np.zeros(n)
    assert prod.pk_larnv_pm1(0, C.c_longlong(n), P(y)) == 0
    assert np.array_equal(y, exact)
    # the 128-entry multiplier table of dlaruv.f holds a^i mod 2^48
    a, mod = 33952834046453, 1 << 48
    assert pow(a, 1, mod) == 494 * 4096**3 + 322 * 4096**2 + 2508 * 4096 + 2549
    assert pow(a, 128, mod) == 545 * 4096**3 + 2366 * 4096**2 + 3801 * 4096 + 1537


def test_larnv_single(prod):
    refs = refpy.lib("s")
    n = 700
    x = np.zeros(n, np.float32)
    refs.slarnv_(C.byref(I(2)), P(np.array([1, 3, 5, 7], np.int32)), C.byref(I(n)), P(x))
    y = np.zeros(n, np.float32)
    assert prod.pk_larnv_pm1(1, C.c_longlong(n), P(y)) == 0
    assert np.array_equal(x, y)
    assert np.array_equal(portpy.larnv_pm1(n, "s"), x)
    # single-precision conversion rounds the 48-bit value: close to the exact one
    assert np.abs(x - closed_form_larnv(n)).max() < 1e-7


def _ab(rng, j):
    alpha = np.concatenate([[0.0], rng.uniform(0.5, 2.0, j + 1)])
    beta = np.concatenate([[0.0], rng.uniform(0.5, 2.0, j + 1)])
    return alpha, beta


@pytest.mark.parametrize("j", [1, 2, 3, 7, 40])
def test_update_mu_nu(ref, port, prod, j):
    rng = np.random.default_rng(j)
    alpha, beta = _ab(rng, j)
    mu0 = np.concatenate([[0.0], rng.standard_normal(j + 2) * 1e-9])
    nu0 = np.concatenate([[0.0], rng.standard_normal(j + 2) * 1e-9])
    anorm, eps1 = 3.7, 2.1e-14
    outs = []
    for who in ("ref", "port", "prod"):
        mu, nu, mx = mu0.copy(), nu0.copy(), D(0.0)
        if who == "ref":
            ref.dupdate_mu_(C.byref(mx), P(mu[1:]), P(nu[1:]), C.byref(I(j)), P(alpha[1:]),
                            P(beta[1:]), C.byref(D(anorm)), C.byref(D(eps1)))
        elif who == "port":
            port.pkp_d_update_mu(C.byref(mx), P(mu), P(nu), I(j), P(alpha), P(beta), D(anorm),
                                 D(eps1))
        else:
            prod.pkh_update_mu(C.byref(mx), P(mu), P(nu), I(j), P(alpha), P(beta), D(anorm),
                               D(eps1))
        mx2 = D(0.0)
        if who == "ref":
            ref.dupdate_nu_(C.byref(mx2), P(mu[1:]), P(nu[1:]), C.byref(I(j + 1)), P(alpha[1:]),
                            P(beta[1:]), C.byref(D(anorm)), C.byref(D(eps1)))
        elif who == "port":
            port.pkp_d_update_nu(C.byref(mx2), P(mu), P(nu), I(j + 1), P(alpha), P(beta),
                                 D(anorm), D(eps1))
        else:
            prod.pkh_update_nu(C.byref(mx2), P(mu), P(nu), I(j + 1), P(alpha), P(beta),
                               D(anorm), D(eps1))
        outs.append((mu, nu, mx.value, mx2.value))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert np.array_equal(o[1], outs[0][1])
        assert o[2] == outs[0][2] and o[3] == outs[0][3]


@pytest.mark.parametrize("seed", range(6))
def test_compute_int_and_set_mu(ref, port, prod, seed):
    rng = np.random.default_rng(100 + seed)
    j = int(rng.integers(1, 60))
    delta, eta = 1.5e-8, 1e-12
    mu = np.concatenate([[0.0], 10.0 ** rng.uniform(-16, -6, j + 1)
                         * rng.choice([-1, 1], j + 1)])
    res = []
    for who in ("ref", "port", "prod"):
        idx = np.zeros(2 * j + 4, np.int32)
        if who == "ref":
            ref.dcompute_int_(P(mu[1:]), C.byref(I(j)), C.byref(D(delta)), C.byref(D(eta)),
                              P(idx[1:]))
        elif who == "port":
            port.pkp_d_compute_int(P(mu), I(j), D(delta), D(eta), P(idx))
        else:
            prod.pkh_compute_int(P(mu), I(j), D(delta), D(eta), P(idx))
        m2 = mu.copy()
        if who == "ref":
            ref.dset_mu_(C.byref(I(j)), P(m2[1:]), P(idx[1:]), C.byref(D(7.0)))
        elif who == "port":
            port.pkp_d_set_mu(I(j), P(m2), P(idx), D(7.0))
        else:
            prod.pkh_set_mu(I(j), P(m2), P(idx), D(7.0))
        res.append((idx, m2))
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0])
        assert np.array_equal(r[1], res[0][1])
    # the intervals cover every |mu| > delta
    covered = res[0][1][1:j + 1] == 7.0
    assert np.all(covered[np.abs(mu[1:j + 1]) > delta])


def test_lartg(ref, prod):
    rng = np.random.default_rng(5)
    vals = list(rng.standard_normal(40)) + [0.0, 1e300, -1e300, 1e-300, -3e-310, 2.5]
    for f in vals:
        for g in vals[::3]:
            a = [D(), D(), D()]
            b = [D(), D(), D()]
            ref.dlartg_(C.byref(D(f)), C.byref(D(g)), *[C.byref(x) for x in a])
            prod.pkh_lartg(D(f), D(g), *[C.byref(x) for x in b])
            assert [x.value for x in a] == [x.value for x in b], (f, g)


@pytest.mark.parametrize("n,ignorelast,jobq", [(1, 0, b"Y"), (7, 0, b"Y"), (7, 1, b"Y"),
                                              (30, 0, b"N"), (30, 1, b"N")])
def test_bdqr(ref, port, prod, n, ignorelast, jobq):
    rng = np.random.default_rng(n)
    d0, e0 = rng.uniform(0.1, 3, n), rng.uniform(0.1, 3, n)
    out = []
    for who in ("ref", "port", "prod"):
        d, e = d0.copy(), e0.copy()
        c1, c2 = D(0.0), D(0.0)
        qt = np.zeros((n + 1) * (n + 1))
        if who == "ref":
            ref.dbdqr_(C.byref(I(ignorelast)), C.c_char_p(jobq), C.byref(I(n)), P(d), P(e),
                       C.byref(c1), C.byref(c2), P(qt), C.byref(I(n + 1)), C.c_size_t(1))
        elif who == "port":
            port.pkp_d_bdqr(I(ignorelast), C.c_char(jobq), I(n), P(d), P(e), C.byref(c1),
                            C.byref(c2), P(qt), I(n + 1))
        else:
            prod.pkh_bdqr(I(ignorelast), C.c_char(jobq), I(n), P(d), P(e), C.byref(c1),
                          C.byref(c2), P(qt), I(n + 1))
        out.append((d, e, c1.value, c2.value, qt))
    for o in out[1:]:
        for a, b in zip(o, out[0]):
            assert np.array_equal(a, b)
    if jobq == b"Y" and not ignorelast:
        # Q^T B = [R; 0]: check the factorisation itself
        B = np.zeros((n + 1, n))
        B[np.arange(n), np.arange(n)] = d0
        B[np.arange(1, n + 1), np.arange(n)] = e0
        Qt = out[0][4].reshape(n + 1, n + 1).T
        R = Qt @ B
        assert np.abs(R[n]).max() < 1e-12
        assert np.allclose(np.diag(R[:n]), out[0][0])


def test_bsvdstep(ref, port, prod):
    rng = np.random.default_rng(11)
    dim, k = 20, 12
    d0, e0 = rng.uniform(0.5, 2, dim), rng.uniform(0.1, 1, dim)
    out = []
    for who in ("ref", "port", "prod"):
        d, e = d0.copy(), e0.copy()
        Pm = np.eye(dim + 1).flatten(order="F")
        Qm = np.eye(dim).flatten(order="F")
        for i, sh in zip(range(dim, k, -1), rng.uniform(0, 0.3, dim - k)):
            if who == "ref":
                ref.dbsvdstep_(C.c_char_p(b"y"), C.c_char_p(b"y"), C.byref(I(dim + 1)),
                               C.byref(I(dim)), C.byref(I(i)), C.byref(D(sh)), P(d), P(e), P(Pm),
                               C.byref(I(dim + 1)), P(Qm), C.byref(I(dim)), C.c_size_t(1),
                               C.c_size_t(1))
            elif who == "port":
                port.pkp_d_bsvdstep(C.c_char(b"y"), C.c_char(b"y"), I(dim + 1), I(dim), I(i),
                                    D(sh), P(d), P(e), P(Pm), I(dim + 1), P(Qm), I(dim))
            else:
                prod.pkh_bsvdstep(I(dim + 1), I(dim), I(i), D(sh), P(d), P(e), P(Pm),
                                  I(dim + 1), P(Qm), I(dim))
        rng = np.random.default_rng(11)
        rng.uniform(0.5, 2, dim), rng.uniform(0.1, 1, dim)
        out.append((d, e, Pm, Qm))
    for o in out[1:]:
        # D, E bit for bit; the rotations applied to P, Q go through the
        # reference's BLAS DROT (vendor library, may fuse multiply-adds), so
        # the accumulated P, Q agree to rounding.
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])
        assert np.abs(o[2] - out[0][2]).max() < 1e-14
        assert np.abs(o[3] - out[0][3]).max() < 1e-14
    assert np.array_equal(out[1][2], out[2][2]) and np.array_equal(out[1][3], out[2][3])
    # orthogonality of the accumulated rotations
    Pm = out[0][2].reshape(dim + 1, dim + 1, order="F")
    assert np.abs(Pm.T @ Pm - np.eye(dim + 1)).max() < 1e-13


def test_refinebounds(ref, port, prod):
    rng = np.random.default_rng(3)
    k = 25
    theta = np.sort(rng.uniform(1, 10, k))[::-1].copy()
    theta[5] = theta[4] * (1 - 1e-14)  # a cluster
    b0 = 10.0 ** rng.uniform(-14, -2, k)
    out = []
    for who in ("ref", "port", "prod"):
        b = b0.copy()
        if who == "ref":
            ref.drefinebounds_(C.byref(I(100)), C.byref(I(k)), P(theta), P(b),
                               C.byref(D(1e-15)), C.byref(D(1e-12)))
        elif who == "port":
            port.pkp_d_refinebounds(I(100), I(k), P(theta), P(b), D(1e-15), D(1e-12))
        else:
            prod.pkh_refinebounds(I(100), I(k), P(theta), P(b), D(1e-15), D(1e-12))
        out.append(b)
    for o in out[1:]:
        assert np.array_equal(o, out[0])
    assert out[0][5] == 0.0 or out[0][4] == 0.0  # cluster merged


@pytest.mark.parametrize("cgs", [0, 1])
def test_port_matches_reference_end_to_end(small_case, cgs):
    A, u0, Dm, s = small_case
    r = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=cgs)
    q = portpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=cgs)
    assert r.info == q.info == 0 and r.k == q.k == 10
    for key in ("nopx", "nreorth", "ndot", "nbsvd"):
        assert r.counters[key] == q.counters[key], key
    np.testing.assert_allclose(q.sigma, r.sigma, rtol=1e-13)
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-13)
    # The reference's own triplets reach ||A v - s u|| ~ 7e-12 s_1 here with
    # MGS (tol = 1e-12 bounds the singular values, not the residual).
    res = np.linalg.norm(Dm @ r.V - r.U * r.sigma, axis=0)
    assert res.max() <= 1e-11 * s[0]


def test_port_matches_reference_irl(small_case):
    A, u0, Dm, s = small_case
    r = refpy.lansvd_irl(A, 10, 30, u0=u0, tol=1e-12, cgs=1)
    q = portpy.lansvd_irl(A, 10, 30, u0=u0, tol=1e-12, cgs=1)
    assert r.info == q.info == 0 and r.k == q.k == 10
    for key in ("nopx", "nreorth", "nrestart", "nbsvd"):
        assert r.counters[key] == q.counters[key], key
    np.testing.assert_allclose(q.sigma, r.sigma, rtol=1e-13)
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-13)


def test_random_start_matches(small_case):
    A, u0, Dm, s = small_case
    r = refpy.lansvd(A, 10, 300, u0=None, tol=1e-12, cgs=1)
    q = portpy.lansvd(A, 10, 300, u0=None, tol=1e-12, cgs=1)
    assert r.counters == {**r.counters, **{k: q.counters[k] for k in ("nopx", "nreorth")}}
    np.testing.assert_allclose(q.sigma, r.sigma, rtol=1e-13)


def test_single_precision_reference(small_case):
    A, u0, Dm, s = small_case
    r = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-6, prec="s", cgs=1)
    q = portpy.lansvd(A, 10, 300, u0=u0, tol=1e-6, prec="s", cgs=1)
    assert r.info == q.info == 0
    np.testing.assert_allclose(r.sigma, s[:10], rtol=2e-6)
    np.testing.assert_allclose(q.sigma, r.sigma, rtol=2e-6)
    assert r.counters["nopx"] == q.counters["nopx"]


@pytest.mark.parametrize("prec", ["d", "s"])
def test_port_invariant_subspace_lowers_dim(prec):
    """DIM > min(m, n): at step m the Lanczos process reaches an invariant
    subspace, DLANBPRO lowers its K argument, and the restarted driver goes
    on with the lowered DIM (double/dlanbpro.F:519, double/dlansvd_irl.F:213).
    Smallest triplets of a 10 x 20 matrix (scipy's test_svdp case)."""
    rng = np.random.RandomState(0)
    M = (-5 + 10 * rng.rand(10, 20)).astype(np.float32)
    M[M > 3] = 0
    r_, c_ = np.nonzero(M)
    A = synth.from_coo(10, 20, r_, c_, M[r_, c_].astype(np.float64))
    if prec == "s":
        A = A.astype(np.float32)
    u0 = np.random.default_rng(0).uniform(size=10)
    tol = 1e-4 if prec == "s" else 1e-8
    r = refpy.lansvd_irl(A, 3, 11, p=8, u0=u0, tol=tol, which="S", prec=prec)
    q = portpy.lansvd_irl(A, 3, 11, p=8, u0=u0, tol=tol, which="S", prec=prec)
    s = np.linalg.svd(M.astype(np.float64), compute_uv=False)
    assert r.info == q.info == 0 and r.k == q.k == 3
    assert r.dim == q.dim == 10 and r.counters["nlandim"] == q.counters["nlandim"] == 10
    np.testing.assert_allclose(np.sort(q.sigma), np.sort(s)[:3], rtol=1e-5 if prec == "s" else 1e-12)
    np.testing.assert_allclose(q.sigma, r.sigma, rtol=1e-5 if prec == "s" else 1e-12)
