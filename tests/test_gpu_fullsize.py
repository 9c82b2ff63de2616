"""Full-size properties.  The headline workload: DLANSVD_IRL k=50, dim=120,
tol=1e-10, FP64 on the 2,000,000 x 1,000,000 matrix with 32M nonzeros, a
size the oracle does not finish in a test.  Size-independent checks: every
returned triplet's residual ||A v - sigma u|| within tol * sigma_1 (A v
formed on the host with scipy, independent of the GPU SpMV) and ||A^T u -
sigma v|| (the Lanczos bound: within 10 tol sigma_1), U and V orthonormal to
the semi-orthogonality the partial reorthogonalization maintains
(5 sqrt(2 eps)), sigma descending and positive, the operator count of one
Lanczos pass, and the autotuned (column-sorted) layout agreeing with the
row-block layout to 1e-12."""
import numpy as np
import pytest

from propack_amd import api, synth

pytestmark = pytest.mark.gpu

M, N, LAM, DECAY, SEED = 2_000_000, 1_000_000, 16.0, 0.6, 2024
NEIG, DIM, TOL = 50, 120, 1e-10


def _solve(A, u0, **layout):
    op = api.CsrOperator(A, **layout)
    eng = api.Engine(op, kmax=DIM)
    eng.put_u1(u0)
    r = api.lansvd_irl_dev(eng, NEIG, DIM, tol=TOL)
    eng.close()
    op.close()
    return r


def test_headline_workload_properties():
    import scipy.sparse as sp
    A = synth.poisson_decay(M, N, LAM, DECAY, seed=SEED)
    u0 = synth.start_vector(M, SEED + 1)
    t = api.autotune(A)
    r = _solve(A, u0, panels=t.panels, lanes=t.lanes, sort=t.sort)
    assert r.info == 0 and r.k == NEIG
    assert r.counters["nopx"] <= 2 * DIM + 1
    s = r.sigma
    assert np.all(s > 0) and np.all(np.diff(s) <= 0)
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(M, N))
    R = S @ r.V - r.U * s
    assert np.linalg.norm(R, axis=0).max() <= TOL * s[0]
    R = S.T @ r.U - r.V * s
    assert np.linalg.norm(R, axis=0).max() <= 10 * TOL * s[0]
    del R
    semi = 5 * np.sqrt(2 * np.finfo(np.float64).eps)
    for Q in (r.U, r.V):
        G = Q.T @ Q
        assert np.abs(G - np.eye(NEIG)).max() <= semi
    r0 = _solve(A, u0)  # the row-block layout, built-in panels
    assert r0.info == 0 and r0.k == NEIG
    np.testing.assert_allclose(s, r0.sigma, rtol=1e-12)


def test_powerlaw_workload_properties():
    """The 8,000,000 x 4,000,000 power-law workload (37.8M nonzeros: each
    row's columns distinct, values 1.0), k=64, dim=160, tol=1e-8: the FP64 solve and the
    FP32-basis form with its FP64 polish.  The polished triplets' residuals
    within tol * sigma_1 (A v and A^T u formed on the host, the second within
    10 tol sigma_1), U and V orthonormal to 5 sqrt(2 eps), the polished sigma
    within 1e-12 of the FP64 solve's, the FP32 phase's within 1e-6."""
    import scipy.sparse as sp
    k, dim, tol = 64, 160, 1e-8
    A = synth.powerlaw_bipartite(8_000_000, 4_000_000, seed=8000, dtype=np.float64)
    u0 = synth.start_vector(A.m, 8001)
    op = api.CsrOperator(A)
    eng = api.Engine(op, kmax=dim)
    eng.put_u1(u0)
    r64 = api.lansvd_irl_dev(eng, k, dim, tol=tol, fetch=False)
    assert r64.info == 0 and r64.k == k
    op32 = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32)
    eng32 = api.Engine(op32, kmax=dim)
    eng32.put_u1(u0)
    rm, f32 = api.lansvd_irl_mixed(eng, eng32, k, dim, tol=tol)
    assert rm.info == 0 and rm.k == k and f32["k"] == k
    np.testing.assert_allclose(rm.sigma, r64.sigma, rtol=1e-12)
    np.testing.assert_allclose(f32["sigma"][:k], r64.sigma, rtol=1e-6)
    s = rm.sigma
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.m, A.n))
    R = S @ rm.V - rm.U * s
    assert np.linalg.norm(R, axis=0).max() <= tol * s[0]
    R = S.T @ rm.U - rm.V * s
    assert np.linalg.norm(R, axis=0).max() <= 10 * tol * s[0]
    del R
    semi = 5 * np.sqrt(2 * np.finfo(np.float64).eps)
    for Q in (rm.U, rm.V):
        assert np.abs(Q.T @ Q - np.eye(k)).max() <= semi
    for e in (eng, eng32):
        e.close()
