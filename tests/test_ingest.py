"""Matrix ingest.  Host parsing (CPU): Matrix Market, Harwell-Boeing and
coordinate files written by scipy.io (and by hand) read back as the same
coordinate entries scipy's own readers give.  Device conversion (GPU):
coordinate -> CSR with duplicates summed, bit for bit scipy's CSR; a file
read straight into a device operator solves to the dense SVD."""
import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from propack_amd import api


def _dense(m, n, r, c, v):
    D = np.zeros((m, n))
    np.add.at(D, (r, c), v)
    return D


@pytest.fixture
def mat():
    A = sp.random(37, 23, density=0.2, random_state=3, format="csc")
    A.data = np.round(A.data * 1000) / 7.0
    return A


def test_matrix_market_general(tmp_path, mat):
    f = tmp_path / "a.mtx"
    sio.mmwrite(str(f), mat)
    m, n, r, c, v = api.read_matrix(f)
    assert (m, n) == mat.shape and r.size == mat.nnz
    np.testing.assert_array_equal(_dense(m, n, r, c, v), mat.toarray())


def test_matrix_market_symmetric_and_pattern(tmp_path):
    S = sp.random(15, 15, density=0.3, random_state=4)
    S = (S + S.T).tocoo()
    f = tmp_path / "s.mtx"
    sio.mmwrite(str(f), S, symmetry="symmetric")
    m, n, r, c, v = api.read_matrix(f)
    np.testing.assert_allclose(_dense(m, n, r, c, v), S.toarray(), rtol=1e-15)
    K = sp.random(9, 9, density=0.4, random_state=5)
    K = (K - K.T).tocoo()
    f = tmp_path / "k.mtx"
    sio.mmwrite(str(f), K, symmetry="skew-symmetric")
    m, n, r, c, v = api.read_matrix(f)
    np.testing.assert_allclose(_dense(m, n, r, c, v), K.toarray(), rtol=1e-15)
    f = tmp_path / "p.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n% c\n3 4 3\n1 1\n2 4\n3 2\n")
    m, n, r, c, v = api.read_matrix(f)
    assert (m, n) == (3, 4)
    np.testing.assert_array_equal(_dense(m, n, r, c, v),
                                  [[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0]])


def test_harwell_boeing(tmp_path, mat):
    f = tmp_path / "a.rua"
    sio.hb_write(str(f), mat)
    m, n, r, c, v = api.read_matrix(f)
    ref = sio.hb_read(str(f))
    assert (m, n) == ref.shape
    np.testing.assert_array_equal(_dense(m, n, r, c, v), ref.toarray())


def test_harwell_boeing_hand_written(tmp_path):
    """Fixed-width fields with no separating blanks, D exponents, a scale
    factor in the value format, symmetric storage (RSA)."""
    txt = ("Hand-written 3x3 symmetric" + " " * 46 + "HAND0001\n"
           "             4             1             1             2\n"
           "RSA                        3             3             4             0\n"
           "(4I2)           (4I2)           (1P,2D12.4)         \n"
           " 1 3 4 5\n"
           " 1 2 2 3\n"
           "  2.0000D+00 -1.0000D+00\n"
           "  3.5000D+00  4.2500D-01\n")
    f = tmp_path / "h.rsa"
    f.write_text(txt)
    m, n, r, c, v = api.read_matrix(f)
    want = np.array([[2.0, -1.0, 0], [-1.0, 3.5, 0], [0, 0, 0.425]])
    np.testing.assert_array_equal(_dense(m, n, r, c, v), want)


def test_coordinate_triples(tmp_path, mat):
    C = mat.tocoo()
    body = "".join(f"{i + 1} {j + 1} {float(x)!r}\n" for i, j, x in zip(C.row, C.col, C.data))
    f = tmp_path / "a.coord"
    f.write_text(f"{C.shape[0]} {C.shape[1]} {C.nnz}\n" + body)
    m, n, r, c, v = api.read_matrix(f)
    assert (m, n) == C.shape
    np.testing.assert_array_equal(_dense(m, n, r, c, v), mat.toarray())
    f2 = tmp_path / "b.coord"  # no size line: shape from the largest indices
    f2.write_text(body)
    m, n, r, c, v = api.read_matrix(f2)
    assert m == C.row.max() + 1 and n == C.col.max() + 1


def test_bad_files(tmp_path):
    f = tmp_path / "bad.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(api.PropackError):
        api.read_matrix(f)
    with pytest.raises(api.PropackError):
        api.read_matrix(tmp_path / "missing.mtx")
    # A value field strtod cannot read is an error, not a silent zero.
    f.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 np.float64(1.0)\n")
    with pytest.raises(api.PropackError, match="unreadable entry on line 3"):
        api.read_matrix(f)


@pytest.mark.gpu
@pytest.mark.parametrize("base", [0, 1])
def test_coo_to_csr_matches_scipy(base):
    rng = np.random.default_rng(7)
    m, n, nnz = 300, 200, 5000
    r = rng.integers(0, m, nnz)
    c = rng.integers(0, n, nnz)
    v = rng.integers(-50, 50, nnz).astype(np.float64)  # integer values: sums exact in any order
    A = api.coo_to_csr(m, n, r + base, c + base, v, base=base)
    B = sp.coo_matrix((v, (r, c)), shape=(m, n)).tocsr()
    B.sum_duplicates()
    B.sort_indices()
    np.testing.assert_array_equal(A.rowptr, B.indptr)
    np.testing.assert_array_equal(A.col, B.indices)
    np.testing.assert_array_equal(A.val, B.data)
    A32 = api.coo_to_csr(m, n, r, c, v.astype(np.float32))
    np.testing.assert_array_equal(A32.val, B.data.astype(np.float32))


@pytest.mark.gpu
def test_coo_to_csr_edge_cases():
    A = api.coo_to_csr(4, 3, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    np.testing.assert_array_equal(A.rowptr, np.zeros(5, np.int32))
    # one entry repeated many times, summed left to right
    v = np.full(1000, 0.1)
    A = api.coo_to_csr(2, 2, np.ones(1000, np.int32), np.zeros(1000, np.int32), v)
    s = 0.0
    for x in v:
        s += x
    assert A.val.tolist() == [s] and A.rowptr.tolist() == [0, 0, 1]
    with pytest.raises(api.PropackError):
        api.coo_to_csr(2, 2, np.array([2], np.int32), np.array([0], np.int32), np.ones(1))


@pytest.mark.gpu
def test_file_to_operator_solves(tmp_path):
    A = sp.random(400, 150, density=0.05, random_state=9, format="csc")
    f = tmp_path / "a.mtx"
    sio.mmwrite(str(f), A)
    op = api.CsrOperator.from_file(f)
    assert (op.m, op.n, op.nnz) == (400, 150, A.nnz)
    u0 = np.random.default_rng(1).standard_normal(400)
    r = api.lansvd(op, 400, 150, 5, 60, u0=u0, tol=1e-12)
    s = np.linalg.svd(A.toarray(), compute_uv=False)
    np.testing.assert_allclose(r.sigma, s[:5], rtol=1e-12)


@pytest.mark.gpu
def test_complex_coo_to_csr_matches_scipy():
    """Complex coordinate entries: real and imaginary parts each summed in
    input order (integer parts: exact in any order)."""
    rng = np.random.default_rng(8)
    m, n, nnz = 300, 200, 5000
    r = rng.integers(0, m, nnz)
    c = rng.integers(0, n, nnz)
    v = rng.integers(-50, 50, nnz) + 1j * rng.integers(-50, 50, nnz)
    A = api.coo_to_csr(m, n, r, c, v)
    B = sp.coo_matrix((v, (r, c)), shape=(m, n)).tocsr()
    B.sum_duplicates()
    B.sort_indices()
    np.testing.assert_array_equal(A.rowptr, B.indptr)
    np.testing.assert_array_equal(A.col, B.indices)
    np.testing.assert_array_equal(A.val, B.data)
    A64 = api.coo_to_csr(m, n, r, c, v.astype(np.complex64))
    np.testing.assert_array_equal(A64.val, B.data.astype(np.complex64))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["general", "hermitian", "skew-symmetric"])
def test_complex_matrix_market_to_operator_solves(tmp_path, kind):
    """A complex Matrix Market file (general; Hermitian or skew-symmetric
    with only the lower triangle stored, the upper one mirrored conjugated
    or negated) read straight into a complex operator; ZLANSVD's singular
    values against the dense SVD.  Real precisions and pk_matrix_read
    refuse the file."""
    rng = np.random.default_rng(12)
    D = (rng.standard_normal((150, 150)) + 1j * rng.standard_normal((150, 150)))
    D[rng.random((150, 150)) > 0.1] = 0
    if kind == "hermitian":
        D = np.tril(D, -1) + np.diag(rng.standard_normal(150))
        D = D + np.tril(D, -1).conj().T
        stored = sp.coo_matrix(np.tril(D))
    elif kind == "skew-symmetric":
        D = np.tril(D, -1)
        D = D - D.T
        stored = sp.coo_matrix(np.tril(D))
    else:
        stored = sp.coo_matrix(D)
    f = tmp_path / "z.mtx"
    with open(f, "w") as fh:
        fh.write(f"%%MatrixMarket matrix coordinate complex {kind}\n")
        fh.write(f"{D.shape[0]} {D.shape[1]} {stored.nnz}\n")
        for i, j, x in zip(stored.row, stored.col, stored.data):
            fh.write(f"{i + 1} {j + 1} {float(x.real)!r} {float(x.imag)!r}\n")
    op = api.CsrOperator.from_file(f, prec=api.PK_C128)
    assert (op.m, op.n) == D.shape
    x = rng.standard_normal(150) + 1j * rng.standard_normal(150)
    np.testing.assert_allclose(op.spmv(x, "n"), D @ x, rtol=0, atol=1e-12)
    u0 = rng.standard_normal(150) + 1j * rng.standard_normal(150)
    r = api.lansvd(op, 150, 150, 5, 100, u0=u0, tol=1e-12, prec=api.PK_C128)
    s = np.linalg.svd(D, compute_uv=False)
    if kind == "skew-symmetric":
        # a skew-symmetric matrix's singular values come in equal pairs; a
        # Lanczos process from one start vector meets each distinct value once
        s = s[np.r_[True, np.diff(s) < -1e-9 * s[0]]]
    np.testing.assert_allclose(r.sigma, s[:5], rtol=1e-12)
    with pytest.raises(api.PropackError, match="complex"):
        api.CsrOperator.from_file(f, prec=api.PK_F64)
    with pytest.raises(api.PropackError, match="complex"):
        api.read_matrix(f)
