"""The scipy-compatible front end on the GPU, on the cases and tolerances of
scipy's own PROPACK tests (scipy/sparse/linalg/tests/test_propack.py:12-89):
random 10 x 20 matrices as dense / CSR / CSC, float32 and float64, IRL and
plain, largest and smallest triplets; singular values against the dense
SVD, singular vectors orthogonal to the dense ones up to sign."""
import numpy as np
import pytest
from numpy.testing import assert_allclose
from scipy.sparse import csc_array, csr_array
from scipy.sparse.linalg import aslinearoperator

from propack_amd.svdp import svdp

pytestmark = pytest.mark.gpu

TOLS = {np.float32: 1e-4, np.float64: 1e-8, np.complex64: 1e-4, np.complex128: 1e-8}


def generate_matrix(n, m, f, dtype, rseed=0):
    """Uniform entries in [-5, 5), a fraction 1 - f of them zeroed; complex
    types add an imaginary part drawn the same way."""
    rng = np.random.RandomState(rseed)
    M = -5 + 10 * rng.rand(n, m)
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        M = M + 1j * (-5 + 10 * rng.rand(n, m))
    M = M.astype(dtype)
    M[M.real > 10 * f - 5] = 0
    return M


def assert_orthogonal(u1, u2, rtol, atol):
    A = abs(np.dot(u1.conj().T, u2))
    assert_allclose(A, np.eye(u1.shape[1], u2.shape[1]), rtol=rtol, atol=atol)


@pytest.mark.parametrize("ctor", (np.array, csr_array, csc_array))
@pytest.mark.parametrize("dtype", (np.float32, np.float64, np.complex64, np.complex128))
@pytest.mark.parametrize("irl", (True, False))
@pytest.mark.parametrize("which", ("LM", "SM"))
def test_svdp(ctor, dtype, irl, which):
    n, m, k = 10, 20, 3
    tol = TOLS[dtype]
    M = generate_matrix(n, m, 0.8, dtype)
    if which == "SM" and not irl:
        with pytest.raises(ValueError, match="requires irl_mode=True"):
            svdp(ctor(M), k=k, which=which, irl_mode=irl, tol=tol)
        return
    wide = np.complex128 if np.iscomplexobj(M) else np.float64
    u1, s1, vt1 = np.linalg.svd(M.astype(wide), full_matrices=False)
    u2, s2, vt2 = svdp(ctor(M), k=k, which=which, irl_mode=irl, tol=tol,
                       rng=np.random.default_rng(0))
    if which == "SM":
        u1 = np.roll(u1, k, 1)
        vt1 = np.roll(vt1, k, 0)
        s1 = np.roll(s1, k)
    assert_allclose(s1[:k], s2, rtol=tol, atol=tol)
    assert_orthogonal(u1, u2.astype(wide), rtol=tol, atol=tol)
    # vt = V^H for complex inputs, as numpy's
    assert_orthogonal(vt1.conj().T, vt2.conj().T.astype(wide), rtol=tol, atol=tol)


def test_linear_operator_callback():
    """A LinearOperator runs as a host APROD callback."""
    M = generate_matrix(30, 50, 0.5, np.float64, rseed=3)
    u, s, vt, bnd = svdp(aslinearoperator(M), k=4, full_output=True, tol=1e-10,
                         rng=np.random.default_rng(1))
    assert_allclose(s, np.linalg.svd(M, compute_uv=False)[:4], rtol=1e-10)
    assert np.all(bnd >= 0)


def test_tuned_layout_same_triplets():
    """tune=True solves on the SpMV layout the autotuner picks (on this
    matrix a column-panel and column-sorted layout is among the candidates):
    the untuned solve's singular values to rounding, and triplets with small
    residuals."""
    from propack_amd import synth
    A = synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4)
    S = csr_array((A.val, A.col, A.rowptr), shape=(A.m, A.n))
    v0 = np.random.default_rng(3).standard_normal(A.m)
    u0, s0, vt0 = svdp(S, k=8, kmax=40, tol=1e-10, v0=v0)
    u1, s1, vt1 = svdp(S, k=8, kmax=40, tol=1e-10, v0=v0, tune=True)
    assert_allclose(s1, s0, rtol=1e-12)
    R = S @ vt1.T - u1 * s1
    assert np.linalg.norm(R, axis=0).max() <= 1e-8 * s1[0]


def test_argument_errors():
    M = generate_matrix(10, 20, 0.8, np.float64)
    with pytest.raises(ValueError):
        svdp(M, k=0)
    with pytest.raises(ValueError):
        svdp(M, k=3, which="XX")
    with pytest.raises(ValueError):
        svdp(M, k=3, v0=np.ones(5))


def test_complex_linear_operator_callback():
    """A complex LinearOperator: rmatvec supplies A^H x for ZLANSVD's 'c'."""
    M = generate_matrix(30, 50, 0.5, np.complex128, rseed=3)
    u, s, vt = svdp(aslinearoperator(M), k=4, tol=1e-10, rng=np.random.default_rng(1))
    assert_allclose(s, np.linalg.svd(M, compute_uv=False)[:4], rtol=1e-10)
    assert_allclose(M @ vt.conj().T, u * s, atol=1e-8 * s[0])


@pytest.mark.parametrize("shape", ((60, 90), (90, 60)))
@pytest.mark.parametrize("dtype", (np.float64, np.complex128))
def test_svds_propack_semantics(shape, dtype):
    """svds(solver='propack') as scipy's propack branch runs it
    (scipy/sparse/linalg/_eigen/_svds.py): singular values in ascending
    order with u's columns and vh's rows in the same order, the
    restarted driver for which='SM', the vectors each return mode asks for;
    the values agree with scipy's own svds(solver='propack') on the same
    start vector and with the dense SVD."""
    from scipy.sparse.linalg import svds as scipy_svds
    from propack_amd.svdp import svds
    M = generate_matrix(*shape, 0.5, dtype, rseed=5)
    S = csr_array(M)
    v0 = np.random.default_rng(4).standard_normal(shape[0])
    sd = np.linalg.svd(M, compute_uv=False)
    u, s, vh = svds(S, k=5, v0=v0, tol=1e-6)
    assert np.all(np.diff(s) >= 0)
    assert_allclose(s, sd[:5][::-1], rtol=1e-10)
    assert_allclose(M @ vh.conj().T, u * s, atol=1e-8 * s[-1])
    su, ss, svh = scipy_svds(S, k=5, v0=v0, tol=1e-6, solver="propack")
    assert_allclose(s, ss, rtol=1e-10)
    assert_orthogonal(u, su, rtol=1e-8, atol=1e-8)
    s_only = svds(S, k=5, v0=v0, tol=1e-6, return_singular_vectors=False)
    assert_allclose(s_only, s, rtol=1e-12)
    u_only = svds(S, k=5, v0=v0, tol=1e-6, return_singular_vectors="u")
    assert u_only[2] is None and u_only[0].shape == (shape[0], 5)
    vh_only = svds(S, k=5, v0=v0, tol=1e-6, return_singular_vectors="vh")
    assert vh_only[0] is None and vh_only[2].shape == (5, shape[1])
    # smallest triplets on the shape scipy's own svdp tests use for them (10 x
    # 20, k=3); on the 60 x 90 matrices scipy's PROPACK returns zeros for
    # which='SM' (the default shifts, kmax - k, exceed DIM/2: the range where
    # the reference's Changelog:56 lists WHICH='S' as incorrect)
    W = generate_matrix(10, 20, 0.8, dtype, rseed=0)
    w0 = np.random.default_rng(4).standard_normal(10)
    u, s, vh = svds(csr_array(W), k=3, v0=w0, which="SM", tol=1e-6)
    assert_allclose(s, np.linalg.svd(W, compute_uv=False)[-3:][::-1], rtol=1e-8)
    ss = scipy_svds(csr_array(W), k=3, v0=w0, which="SM", tol=1e-6, solver="propack",
                    return_singular_vectors=False)
    assert_allclose(s, ss, rtol=1e-8)
