"""GPU parity: the HIP engine behind the Fortran drop-ins against the
reference PROPACK (compiled from its own sources, oracle/_ref) and the dense
FP64 SVD, on the 1850 x 712 case and small restarted runs.

Tolerances (floating point): singular values 1e-10 relative to the oracle,
residual ||A v - s u|| <= tol * s_1, singular vectors within 1e-8 of the
oracle's up to sign (cosine), counters reported next to the oracle's."""
import numpy as np
import pytest

from oracle import refpy
from propack_amd import api, synth

pytestmark = pytest.mark.gpu


def _cos(a, b):
    return np.abs(np.sum(a * b, axis=0)) / (np.linalg.norm(a, axis=0) * np.linalg.norm(b, axis=0))


def _check_triplets(r, D, tol, s1):
    res = np.linalg.norm(D @ r.V - r.U * r.sigma, axis=0)
    assert res.max() <= tol * s1, (res.max(), tol * s1)
    # orthonormal singular vectors
    k = r.k
    assert np.abs(r.U.T @ r.U - np.eye(k)).max() < 1e-9
    assert np.abs(r.V.T @ r.V - np.eye(k)).max() < 1e-9


def test_spmv_against_host(small_case):
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    x = synth.start_vector(A.n, 3)
    y = synth.start_vector(A.m, 4)
    np.testing.assert_allclose(op.spmv(x, "n"), D @ x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(op.spmv(y, "t"), D.T @ y, rtol=0, atol=1e-12)


@pytest.mark.parametrize("cgs", [1, 0])
def test_lansvd_device_operator(small_case, cgs):
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    r = api.lansvd(op, A.m, A.n, 10, 300, u0=u0, tol=1e-12, cgs=cgs)
    ref = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=cgs)
    print("gpu", r.info, r.k, r.counters)
    print("ref", ref.info, ref.k, ref.counters)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-12)
    # residual bound: 10 x tol, the level the reference's own triplets reach
    _check_triplets(r, D, 1e-11, s[0])
    assert _cos(r.U, ref.U).min() > 1 - 1e-8
    assert _cos(r.V, ref.V).min() > 1 - 1e-8
    # the reference's counters exactly (ioption(1) = 0 runs block CGS here
    # against the reference's MGS; the counts agree on this case too)
    for key in ("nopx", "nreorth", "ndot", "nbsvd", "nlandim", "nsing"):
        assert r.counters[key] == ref.counters[key], key


def test_lansvd_irl_device_operator(small_case):
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    r = api.lansvd_irl(op, A.m, A.n, 10, 30, u0=u0, tol=1e-12)
    ref = refpy.lansvd_irl(A, 10, 30, u0=u0, tol=1e-12, cgs=1)
    print("gpu", r.info, r.k, r.counters)
    print("ref", ref.info, ref.k, ref.counters)
    assert r.info == 0 and r.k == 10 and r.dim == 30
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    _check_triplets(r, D, 1e-11, s[0])
    assert _cos(r.U, ref.U).min() > 1 - 1e-8
    for key in ("nopx", "nreorth", "ndot", "nrestart", "nbsvd", "nlandim"):
        assert r.counters[key] == ref.counters[key], key


def test_lansvd_host_callback(small_case):
    A, u0, D, s = small_case
    calls = []

    def aprod(t, x):
        calls.append(t)
        return D @ x if t == "n" else D.T @ x

    r = api.lansvd(aprod, A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-12)
    _check_triplets(r, D, 1e-11, s[0])
    assert len(calls) == r.counters["nopx"]
    assert r.counters["nsing"] == 10


def test_random_start_vector(small_case):
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    r = api.lansvd(op, A.m, A.n, 10, 300, u0=None, tol=1e-12)
    ref = refpy.lansvd(A, 10, 300, u0=None, tol=1e-12, cgs=1)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)


def test_single_precision_dropin(small_case):
    A, u0, D, s = small_case
    op = api.CsrOperator(A, prec=api.PK_F32)
    r = api.lansvd(op, A.m, A.n, 10, 300, u0=u0, tol=1e-6, prec=api.PK_F32)
    ref = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-6, prec="s", cgs=1)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, s[:10], rtol=2e-6)
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=5e-6)


def test_device_resident_irl(small_case):
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    eng = api.Engine(op, kmax=30)
    eng.put_u1(u0)
    r = api.lansvd_irl_dev(eng, 10, 30, tol=1e-12)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, s[:10], rtol=1e-12)
    _check_triplets(r, D, 1e-11, s[0])


@pytest.mark.parametrize("tries", [1, 2])
def test_device_chain_matches_host_executor(small_case, tries):
    """The device-resident steps (control kernels on the GPU) and the host
    executor run the same control code on the same deterministic
    reductions: identical singular values, bounds and counters."""
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    res = {}
    for mode in ("host", "device"):
        eng = api.Engine(op, kmax=300)
        eng.set_mode(mode, tries)
        eng.put_u1(u0)
        r = api.lansvd_dev(eng, 10, 300, tol=1e-12)
        res[mode] = (r, eng.stats(), eng.get("u", r.k))
    (rh, sh, Uh), (rd, sd, Ud) = res["host"], res["device"]
    assert rh.info == rd.info == 0 and rh.k == rd.k == 10
    assert np.array_equal(rh.sigma, rd.sigma)
    assert np.array_equal(rh.bnd, rd.bnd)
    assert np.array_equal(Uh, Ud)
    for key in ("nopx", "nreorth", "ndot"):
        assert sh[key] == sd[key], key
    assert sd["nsync"] < sh["nsync"] / 5


def test_device_chain_irl_large(small_case):
    """Restarted runs on a 200k x 100k matrix: device-resident and host
    executor agree bit for bit, and the triplets satisfy the residual."""
    A = synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=7)
    u0 = synth.start_vector(A.m, 8)
    op = api.CsrOperator(A)
    res = {}
    for mode in ("host", "device"):
        eng = api.Engine(op, kmax=40)
        eng.set_mode(mode, 2)
        eng.put_u1(u0)
        r = api.lansvd_irl_dev(eng, 20, 40, tol=1e-10)
        res[mode] = (r, eng.stats())
    (rh, sh), (rd, sd) = res["host"], res["device"]
    print("host", rh.counters, sh)
    print("device", rd.counters, sd)
    assert rh.info == rd.info == 0 and rh.k == rd.k == 20
    assert rd.counters["nrestart"] >= 1
    assert np.array_equal(rh.sigma, rd.sigma)
    for key in ("nopx", "nreorth", "ndot"):
        assert sh[key] == sd[key], key
    # residuals of the device run
    eng_r = rd
    worst = 0.0
    for i in range(eng_r.k):
        av = op.spmv(np.ascontiguousarray(eng_r.V[:, i]), "n")
        worst = max(worst, np.linalg.norm(av - eng_r.sigma[i] * eng_r.U[:, i]))
    assert worst <= 1e-9 * eng_r.sigma[0]


def test_fp32_basis_with_fp64_polish(small_case):
    """FP32 operator values and bases (FP64 reductions and scalars), then the
    seeded FP64 polish: polished sigma to 1e-10 of the FP64 oracle and the
    residual bound; the FP32 phase alone is accurate to ~1e-6."""
    A, u0, D, s = small_case
    op64 = api.CsrOperator(A)
    op32 = api.CsrOperator(A, prec=api.PK_F32)
    e64 = api.Engine(op64, kmax=30)
    e32 = api.Engine(op32, kmax=30)
    e32.put_u1(u0)
    r, f32 = api.lansvd_irl_mixed(e64, e32, 10, 30, tol=1e-12)
    ref = refpy.lansvd_irl(A, 10, 30, u0=u0, tol=1e-12, cgs=1)
    assert f32["info"] == 0 and f32["k"] == 10
    np.testing.assert_allclose(f32["sigma"], s[:10], rtol=1e-5)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    _check_triplets(r, D, 1e-11, s[0])


def test_device_residuals(small_case):
    """pk_lbd_residuals: ||A v - s u|| and ||A^T u - s v|| formed on the
    device against the same residuals formed on the host, for FP64
    triplets and for FP32-basis triplets checked with the FP64 operator."""
    A, u0, D, s = small_case
    op64 = api.CsrOperator(A)
    op32 = api.CsrOperator(A, prec=api.PK_F32)
    e64 = api.Engine(op64, kmax=30)
    e64.put_u1(u0)
    r = api.lansvd_irl_dev(e64, 10, 30, tol=1e-12)
    e32 = api.Engine(op32, kmax=30)
    e32.set_rounding(1)
    e32.put_u1(u0)
    r32 = api.lansvd_irl_dev(e32, 10, 30, tol=1e-6)
    for eng_src, rr in ((e64, r), (e32, r32)):
        got = e64.residuals(eng_src, rr.sigma)
        U, V = rr.U.astype(np.float64), rr.V.astype(np.float64)
        want = np.stack([np.linalg.norm(D @ V - U * rr.sigma, axis=0),
                         np.linalg.norm(D.T @ U - V * rr.sigma, axis=0)], axis=1)
        assert np.abs(got - want).max() <= 1e-13 * s[0] + 1e-6 * want.max()
    assert got[:, 0].max() > 1e-10 * s[0]  # the FP32 triplets: FP32-level residuals


@pytest.mark.parametrize("tol,polished", [(1e-6, False), (1e-10, True)])
def test_fp32_basis_skips_polish_when_triplets_meet_tol(small_case, tol, polished):
    """When tol is within the FP32 phase's reach (>= 16 eps32) its triplets
    are checked with the FP64 operator on the device and, when all meet
    tol, returned without the FP64 polish (the FP64 engine runs no Lanczos
    step); a tighter tol polishes.  Either way the result meets tol by the
    host residual."""
    A, u0, D, s = small_case
    e64 = api.Engine(api.CsrOperator(A), kmax=30)
    e32 = api.Engine(api.CsrOperator(A, prec=api.PK_F32), kmax=30)
    e32.put_u1(u0)
    e64.reset_stats()
    r, f32 = api.lansvd_irl_mixed(e64, e32, 10, 30, tol=tol)
    assert r.info == 0 and r.k == 10
    assert (e64.stats()["nopx"] > 0) == polished
    res = np.linalg.norm(D @ r.V - r.U * r.sigma, axis=0).max()
    assert res <= 10 * tol * s[0]
    np.testing.assert_allclose(r.sigma, s[:10], rtol=10 * tol)


def test_config1_fp32_basis_form(small_case):
    """The FP32-basis form of the non-restarted solve (DLANSVD, k=10,
    kmax=300, tol=1e-12 on the 1850 x 712 matrix): tol lies below 16 eps32,
    so the FP64 DLANSVD polish runs.  Per triplet, the FP32 phase's sigma
    against the FP64 reference result within 1e-5 relative (FP32 storage);
    the polished sigma within 1e-10 of the reference, residuals <= 1e-11 s_1
    by the host."""
    A, u0, D, s = small_case
    ref = refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=1)
    e64 = api.Engine(api.CsrOperator(A), kmax=300)
    e32 = api.Engine(api.CsrOperator(A, prec=api.PK_F32), kmax=300)
    e32.put_u1(u0)
    r, f32 = api.lansvd_mixed(e64, e32, 10, 300, tol=1e-12)
    print("fp32 phase", f32["info"], f32["k"], np.abs(f32["sigma"] - ref.sigma) / ref.sigma)
    assert f32["info"] == 0 and f32["k"] == 10
    np.testing.assert_allclose(f32["sigma"], ref.sigma, rtol=1e-5)
    assert r.info == 0 and r.k == 10
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    _check_triplets(r, D, 1e-11, s[0])


@pytest.mark.parametrize("tol,polished", [(1e-6, False), (1e-10, True)])
def test_config1_fp32_basis_skips_polish(small_case, tol, polished):
    """Non-restarted form: within the FP32 phase's reach its triplets are
    checked with the FP64 operator and returned without a polish."""
    A, u0, D, s = small_case
    e64 = api.Engine(api.CsrOperator(A), kmax=300)
    e32 = api.Engine(api.CsrOperator(A, prec=api.PK_F32), kmax=300)
    e32.put_u1(u0)
    e64.reset_stats()
    r, f32 = api.lansvd_mixed(e64, e32, 10, 300, tol=tol)
    assert r.info == 0 and r.k == 10
    assert (e64.stats()["nopx"] > 0) == polished
    res = np.linalg.norm(D @ r.V - r.U * r.sigma, axis=0).max()
    assert res <= 10 * tol * s[0]
    np.testing.assert_allclose(r.sigma, s[:10], rtol=10 * tol)


def test_single_precision_irl(small_case):
    A, u0, D, s = small_case
    op = api.CsrOperator(A, prec=api.PK_F32)
    r = api.lansvd_irl(op, A.m, A.n, 10, 30, u0=u0, tol=1e-6, prec=api.PK_F32)
    ref = refpy.lansvd_irl(A, 10, 30, u0=u0, tol=1e-6, prec="s", cgs=1)
    assert r.info == 0 and r.k == 10 and ref.k == 10
    np.testing.assert_allclose(r.sigma, s[:10], rtol=2e-6)
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-5)


def test_irl_smallest_triplets():
    """WHICH='S' (double/dlansvd_irl.F:245-344) on a small well-conditioned
    matrix: the k smallest singular values against the dense SVD and the
    reference run.  p <= dim/2: the reference documents WHICH='S' as unreliable
    for p > dim/2 (Changelog:56).  The reference runs single-threaded: its
    OpenMP build returns spurious values here with 8 or more threads (n = 120
    is split into tiny per-thread chunks), while 1-4 threads, the C
    restatement and the dense SVD agree."""
    A = synth.random_uniform(400, 120, 6000, seed=11)
    D = A.todense()
    s = np.linalg.svd(D, compute_uv=False)
    u0 = synth.start_vector(A.m, 12)
    op = api.CsrOperator(A)
    r = api.lansvd_irl(op, A.m, A.n, 5, 60, p=30, u0=u0, tol=1e-10, which="S", maxiter=2000)
    refpy.set_threads(1)
    ref = refpy.lansvd_irl(A, 5, 60, p=30, u0=u0, tol=1e-10, which="S", cgs=1, maxiter=2000)
    print("gpu", r.info, r.k, r.sigma, r.counters)
    print("ref", ref.info, ref.k, ref.sigma, ref.counters)
    assert r.k == ref.k and r.k > 0
    np.testing.assert_allclose(np.sort(r.sigma), np.sort(ref.sigma), rtol=1e-9)
    # counters: exact, but ndot within 1 % (a second CGS pass is decided at
    # the rounding level when the reference projects INDEX interval by
    # interval and the engine all intervals in one block)
    for key in ("nopx", "nreorth", "nrestart", "nbsvd"):
        assert r.counters[key] == ref.counters[key], key
    assert abs(r.counters["ndot"] - ref.counters["ndot"]) <= 0.01 * ref.counters["ndot"]
    smallest = np.sort(s)[: r.k]
    np.testing.assert_allclose(np.sort(r.sigma), smallest, rtol=1e-8)


@pytest.mark.parametrize("prec", ["d", "s"])
def test_invariant_subspace_lowers_dim(prec):
    """DIM = m + 1: the Lanczos process meets an invariant subspace at step
    m, DLANBPRO lowers DIM, and the restarted driver continues with it
    (double/dlanbpro.F:519, double/dlansvd_irl.F:213).  Smallest triplets of
    scipy's 10 x 20 test matrix; DIM is written back as 10."""
    rng = np.random.RandomState(0)
    M = (-5 + 10 * rng.rand(10, 20)).astype(np.float32)
    M[M > 3] = 0
    r_, c_ = np.nonzero(M)
    A = synth.from_coo(10, 20, r_, c_, M[r_, c_].astype(np.float64))
    pk = api.PK_F64
    if prec == "s":
        A, pk = A.astype(np.float32), api.PK_F32
    u0 = np.random.default_rng(0).uniform(size=10)
    tol = 1e-4 if prec == "s" else 1e-8
    op = api.CsrOperator(A, prec=pk)
    r = api.lansvd_irl(op, 10, 20, 3, 11, p=8, u0=u0, tol=tol, which="S", prec=pk)
    ref = refpy.lansvd_irl(A, 3, 11, p=8, u0=u0, tol=tol, which="S", prec=prec)
    s = np.linalg.svd(M.astype(np.float64), compute_uv=False)
    assert r.info == ref.info == 0 and r.k == ref.k == 3
    assert r.dim == ref.dim == 10 and r.counters["nlandim"] == 10
    rt = 1e-5 if prec == "s" else 1e-12
    np.testing.assert_allclose(np.sort(r.sigma), np.sort(s)[:3], rtol=rt)
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=rt)


def test_powerlaw_fp32_basis_polish():
    """The power-law bipartite shape (row degrees ~ d^-2.1, Zipf(1.8)
    columns: rows of A^T with tens of thousands of nonzeros) at 400k x 200k,
    k=16, dim=48, tol=1e-8, in the FP32-basis form with FP64 polish: the
    FP32 phase within 1e-5 sigma_1 of the reference's FP64 values, the polished
    values within 1e-10, residuals within the tolerance."""
    A = synth.powerlaw_bipartite(400_000, 200_000, seed=3, dtype=np.float64)
    u0 = synth.start_vector(A.m, 9)
    e64 = api.Engine(api.CsrOperator(A), kmax=48)
    op32 = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32)
    e32 = api.Engine(op32, kmax=48)
    e32.put_u1(u0)
    r, f32 = api.lansvd_irl_mixed(e64, e32, 16, 48, tol=1e-8)
    refpy.set_threads(8)
    ref = refpy.lansvd_irl(A, 16, 48, u0=u0, tol=1e-8, cgs=1, jobu="n", jobv="n")
    print("fp32", f32["k"], f32["info"], "polished", r.k, r.info, "ref", ref.k, ref.counters)
    assert ref.k == 16 and r.k == 16 and f32["k"] == 16
    # FP32 storage: errors of order eps32 * ||A|| (sigma_1 = 2.9e4 dominates here)
    np.testing.assert_allclose(f32["sigma"], ref.sigma, rtol=0, atol=1e-5 * ref.sigma[0])
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    op = api.CsrOperator(A)
    for i in range(r.k):
        res = np.linalg.norm(op.spmv(np.ascontiguousarray(r.V[:, i]), "n") - r.sigma[i] * r.U[:, i])
        assert res <= 1e-7 * r.sigma[0]
