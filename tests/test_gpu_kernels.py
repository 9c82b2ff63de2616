"""GPU kernel checks through the C ABI against float64 host references:
the basis-update GEMM (FP64 MFMA), the row-block SpMV on edge-case shapes
(empty rows, rows longer than one LDS block, single row / column), and the
column put/get round trip.  Tolerances: FP64 results within 1e-13 relative
to the norm of the host product."""
import ctypes as C

import numpy as np
import pytest

from propack_amd import api, synth

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _engine(m, n, kmax, prec=api.PK_F64):
    # a tiny operator just to own an engine of the right shape
    A = synth.random_uniform(m, n, 3 * m, seed=5)
    A = synth.Csr(A.m, A.n, A.rowptr, A.col,
                  A.val.astype(np.float64 if prec == api.PK_F64 else np.float32))
    op = api.CsrOperator(A, prec=prec)
    return op, api.Engine(op, kmax=kmax, prec=prec)


# (m, kmax, ncols, transb): every column-tile count of the LDS-staged
# kernel (ncols 3, 12, 20, 40, 64, 70 = 64 + 6), the restart and Ritz
# shapes' column counts (51, 50), row counts around its 256-row blocks and
# below one 16-row tile, k extents up to its LDS limit (kdim 201 and 313)
# and past it (kdim 321: the unstaged kernel)
@pytest.mark.parametrize("m,kmax,ncols,transb", [(1000, 37, 20, "n"), (2049, 80, 70, "n"),
                                                  (777, 30, 12, "t"), (64, 5, 3, "n"),
                                                  (10, 50, 33, "n"), (5000, 200, 64, "n"),
                                                  (3001, 312, 40, "t"), (2000, 320, 40, "n"),
                                                  (4097, 120, 51, "n"), (1500, 60, 50, "t")])
@pytest.mark.parametrize("prec", [api.PK_F64, api.PK_F32])
def test_basis_update_gemm(m, kmax, ncols, transb, prec):
    op, eng = _engine(m, 50, kmax, prec)
    rng = np.random.default_rng(m)
    dt = np.float64 if prec == api.PK_F64 else np.float32
    U = rng.standard_normal((m, kmax + 1)).astype(dt)
    for j in range(kmax + 1):
        assert api.lib().pk_lbd_put(eng.handle, C.c_char(b"u"), C.c_int(j + 1),
                                    _p(np.ascontiguousarray(U[:, j])), C.c_int(prec)) == 0
    kdim = kmax + 1
    ncols = min(ncols, kdim)
    B = rng.standard_normal((kdim, ncols))
    Bh = np.asfortranarray(B if transb == "n" else B.T)
    ldb = Bh.shape[0]
    rc = api.lib().pk_lbd_update(eng.handle, C.c_char(b"u"), C.c_char(transb.encode()),
                                 C.c_int(kdim), C.c_int(ncols), _p(Bh), C.c_int(ldb))
    assert rc == 0, api.lib().pk_last_error()
    got = eng.get("u", ncols, prec).astype(np.float64)
    want = U.astype(np.float64) @ B
    # FP64 accumulation; FP32 storage rounds each result once
    tol = 1e-13 * np.sqrt(kdim) if prec == api.PK_F64 else 2 ** -23
    assert np.abs(got - want).max() <= tol * np.abs(want).max()


def test_basis_update_chained():
    """Two updates in a row return with their GEMMs in flight: the second
    call's B staging must wait for the first call's copy, and its GEMM reads
    the first one's output."""
    m, kmax = 3001, 24
    op, eng = _engine(m, 50, kmax)
    rng = np.random.default_rng(7)
    U = rng.standard_normal((m, kmax + 1))
    for j in range(kmax + 1):
        assert api.lib().pk_lbd_put(eng.handle, C.c_char(b"u"), C.c_int(j + 1),
                                    _p(np.ascontiguousarray(U[:, j])), C.c_int(0)) == 0
    B1 = np.asfortranarray(rng.standard_normal((kmax + 1, 20)))
    B2 = np.asfortranarray(rng.standard_normal((20, 9)))
    for B in (B1, B2):
        rc = api.lib().pk_lbd_update(eng.handle, C.c_char(b"u"), C.c_char(b"n"),
                                     C.c_int(B.shape[0]), C.c_int(B.shape[1]), _p(B),
                                     C.c_int(B.shape[0]))
        assert rc == 0, api.lib().pk_last_error()
    want = U @ B1 @ B2
    got = eng.get("u", 9)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("prec", [api.PK_F64, api.PK_F32])
@pytest.mark.parametrize("n", [5, 1_000_003])
def test_larnv_device_matches_host(prec, n):
    """DGETU0's random vector generated on the device equals the host
    restatement (pinned against the reference's xLARNV) bit for bit; n beyond
    the generator's grid exercises the per-thread stepping."""
    import torch
    dt = np.float64 if prec == api.PK_F64 else np.float32
    host = np.zeros(n, dt)
    assert api.lib().pk_larnv_pm1(prec, C.c_longlong(n), _p(host)) == 0
    dev = torch.zeros(n, dtype=torch.float64 if prec == api.PK_F64 else torch.float32,
                      device="cuda")
    assert api.lib().pk_larnv_pm1_device(prec, C.c_longlong(n), C.c_void_p(dev.data_ptr())) == 0
    assert np.array_equal(dev.cpu().numpy().view(np.uint8), host.view(np.uint8))


def test_put_u1_from_device():
    import torch
    op, eng = _engine(4097, 300, 6)
    x = np.random.default_rng(3).standard_normal(4097)
    eng.put_u1(torch.from_numpy(x).to("cuda"))
    assert np.array_equal(eng.get("u", 1)[:, 0], x)


@pytest.mark.parametrize("prec", [api.PK_F64, api.PK_F32])
@pytest.mark.parametrize("m,j", [(100, 1), (2048, 4), (4097, 5), (6000, 11), (257, 9)])
def test_cgs_projection_edge_shapes(prec, m, j):
    """DGETU0 through the device Gram-Schmidt kernels: U(:, j+1) = op(A) x
    made orthogonal to U(:, 1:j), on row counts around the kernels'
    2048-row slices (a last slice of one row, fewer rows than a workgroup)
    and column counts around their 4-column tiles.  Reference: the
    projection (I - Q Q^T) A x in float64 on the host, with Q and A as the
    engine stores them.  Tolerances: 1e-13 (FP64 storage), 2e-5 (FP32
    storage) relative to ||A x||."""
    n = 300
    A = synth.random_uniform(m, n, 6 * m, seed=m + j)
    dt = np.float64 if prec == api.PK_F64 else np.float32
    A = A.astype(dt)
    op = api.CsrOperator(A, prec=prec)
    eng = api.Engine(op, kmax=j + 2, prec=prec)
    Q, _ = np.linalg.qr(np.random.default_rng(m * 7 + j).standard_normal((m, j)))
    Qs = Q.astype(dt).astype(np.float64)
    for c in range(j):
        assert api.lib().pk_lbd_put(eng.handle, C.c_char(b"u"), C.c_int(c + 1),
                                    _p(np.ascontiguousarray(Q[:, c])), C.c_int(api.PK_F64)) == 0
    u0norm, anormest, ierr = C.c_double(), C.c_double(), C.c_int()
    rc = api.lib().pk_lbd_getu0(eng.handle, C.c_char(b"n"), C.c_int(j), C.c_int(1),
                                C.byref(u0norm), C.c_int(1), C.byref(anormest), C.byref(ierr))
    assert rc == 0 and ierr.value == 0, api.lib().pk_last_error()
    x = np.zeros(n, dt)
    assert api.lib().pk_larnv_pm1(prec, C.c_longlong(n), _p(x)) == 0
    y = A.astype(np.float64).matvec(x.astype(np.float64))
    want = y - Qs @ (Qs.T @ y)
    got = eng.get("u", j + 1)[:, j]
    tol = 1e-13 if prec == api.PK_F64 else 2e-5
    ny = np.linalg.norm(y)
    assert np.linalg.norm(got - want) <= tol * ny
    assert np.abs(Qs.T @ got).max() <= tol * ny
    assert abs(u0norm.value - np.linalg.norm(got)) <= tol * ny


def test_put_get_roundtrip_f32():
    op, eng = _engine(300, 200, 10, prec=api.PK_F32)
    x = np.random.default_rng(1).standard_normal(300)
    eng.put_u1(x)
    got = eng.get("u", 1, prec=api.PK_F64)[:, 0]
    assert np.array_equal(got, x.astype(np.float32).astype(np.float64))


def _csr_from_dense(D):
    r, c = np.nonzero(D)
    return synth.from_coo(D.shape[0], D.shape[1], r, c, D[r, c])


def test_spmv_edge_shapes():
    rng = np.random.default_rng(9)
    # empty rows, a row of 9000 nonzeros (longer than one LDS row block),
    # a column hit 5000 times (long row of A^T)
    m, n = 300, 9500
    D = np.zeros((m, n))
    for i in range(0, m, 3):
        cols = rng.choice(n, rng.integers(1, 20), replace=False)
        D[i, cols] = rng.standard_normal(cols.size)
    D[7, :9000] = rng.standard_normal(9000)
    D[:, 42] = rng.standard_normal(m)
    A = _csr_from_dense(D)
    op = api.CsrOperator(A)
    x = rng.standard_normal(n)
    y = rng.standard_normal(m)
    np.testing.assert_allclose(op.spmv(x, "n"), D @ x, rtol=0, atol=1e-11)
    np.testing.assert_allclose(op.spmv(y, "t"), D.T @ y, rtol=0, atol=1e-11)
    AT = A.transpose()
    opt = api.CsrOperator(AT)  # 9500 x 300 with a 300-long row and a 9000-long column
    np.testing.assert_allclose(opt.spmv(y, "n"), D.T @ y, rtol=0, atol=1e-11)


@pytest.mark.parametrize("m,n", [(1, 1), (1, 50), (50, 1), (5000, 3)])
def test_spmv_tiny(m, n):
    rng = np.random.default_rng(m * 31 + n)
    D = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
    if not D.any():
        D[0, 0] = 1.0
    A = _csr_from_dense(D)
    op = api.CsrOperator(A)
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    np.testing.assert_allclose(op.spmv(x, "n"), D @ x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(op.spmv(y, "t"), D.T @ y, rtol=0, atol=1e-12)


def test_spmv_deterministic():
    A = synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4)
    op = api.CsrOperator(A)
    x = synth.start_vector(A.n, 2)
    y1, y2 = op.spmv(x, "n"), op.spmv(x, "n")
    assert np.array_equal(y1, y2)
    np.testing.assert_allclose(y1, A.matvec(x), rtol=0, atol=1e-12 * np.abs(y1).max())


def test_spmv_powerlaw_long_rows():
    """Zipf-distributed columns give rows of A^T with tens of thousands of
    nonzeros, split into chunks and finished by the fix-up kernel."""
    A = synth.powerlaw_bipartite(400_000, 200_000, seed=3)
    AT = A.transpose()
    assert np.diff(AT.rowptr).max() > 4 * 8192
    op = api.CsrOperator(A)
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(A.n), rng.standard_normal(A.m)
    want_n, want_t = A.matvec(x), AT.matvec(y)
    np.testing.assert_allclose(op.spmv(x, "n"), want_n, rtol=0, atol=1e-10 * np.abs(want_n).max())
    np.testing.assert_allclose(op.spmv(y, "t"), want_t, rtol=0, atol=1e-10 * np.abs(want_t).max())
    assert np.array_equal(op.spmv(y, "t"), op.spmv(y, "t"))


@pytest.fixture
def panels(monkeypatch):
    """Operators created inside the test use the given column panels."""
    def set_(spec):
        monkeypatch.setenv("PK_SPMV_PANELS", spec)
    yield set_


@pytest.mark.parametrize("spec", ["2", "4,8", "16"])
def test_spmv_column_panels(panels, spec):
    """Column-panel layout: one launch per panel, FP64 running row sums
    between launches; same products as the plain CSR on edge shapes (a
    long row split into chunks, empty rows, a dense column) and on the
    power-law matrix, and bit-identical run to run."""
    panels(spec)
    rng = np.random.default_rng(11)
    m, n = 300, 9500
    D = np.zeros((m, n))
    for i in range(0, m, 3):
        cols = rng.choice(n, rng.integers(1, 20), replace=False)
        D[i, cols] = rng.standard_normal(cols.size)
    D[7, :9000] = rng.standard_normal(9000)
    D[:, 42] = rng.standard_normal(m)
    op = api.CsrOperator(_csr_from_dense(D))
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    np.testing.assert_allclose(op.spmv(x, "n"), D @ x, rtol=0, atol=1e-11)
    np.testing.assert_allclose(op.spmv(y, "t"), D.T @ y, rtol=0, atol=1e-11)
    A = synth.powerlaw_bipartite(400_000, 200_000, seed=3)
    op = api.CsrOperator(A)
    x, y = rng.standard_normal(A.n), rng.standard_normal(A.m)
    want_n, want_t = A.matvec(x), A.transpose().matvec(y)
    got_t = op.spmv(y, "t")
    np.testing.assert_allclose(op.spmv(x, "n"), want_n, rtol=0, atol=1e-10 * np.abs(want_n).max())
    np.testing.assert_allclose(got_t, want_t, rtol=0, atol=1e-10 * np.abs(want_t).max())
    assert np.array_equal(got_t, op.spmv(y, "t"))
    # FP32 operator (FP64 row sums between panels, FP32 output)
    op32 = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32)
    got32 = op32.spmv(x.astype(np.float32), "n")
    np.testing.assert_allclose(got32, want_n, rtol=0, atol=1e-5 * np.abs(want_n).max())


def test_panels_same_triplets(panels, small_case):
    """DLANSVD through a panelled operator gives the plain operator's
    singular values (summation grouping differs: not bit-identical)."""
    A, u0, Dm, s = small_case
    r0 = api.lansvd(api.CsrOperator(A), A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    panels("4")
    r1 = api.lansvd(api.CsrOperator(A), A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    assert r0.k == r1.k == 10
    np.testing.assert_allclose(r1.sigma, r0.sigma, rtol=1e-12)
    np.testing.assert_allclose(r1.sigma, s[:10], rtol=1e-12)


def test_autotune_and_explicit_panels():
    """The autotuner returns panel counts and lanes-per-row shifts it
    measured; an operator built with them (and with any explicit layout)
    gives the plain layout's products up to rounding."""
    A = synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4)
    t = api.autotune(A)
    (pa, pt), (sa, st), (ua, ut), (ca, ct) = t.panels, t.lanes, t.us, t.sort
    assert pa in (1, 2, 4, 8, 16) and pt in (1, 2, 4, 8, 16) and ua > 0 and ut > 0
    assert sa in (-2, -1, 0, 1) and st in (-2, -1, 0, 1)
    # row-block 0, row-wise -1, sorted 8192 / 16384 panel by panel, or negated:
    # every panel in one launch
    caps = (0, -1, 8192, 16384, -8192, -16384)
    assert ca in caps and ct in caps
    t0 = api.autotune(A, sort=False)
    assert t0.sort == (0, 0) and t0.panels[0] in (1, 2, 4, 8) and t0.panels[1] in (1, 2, 4, 8)
    assert t0.lanes[0] in (-1, 0, 1) and t0.lanes[1] in (-1, 0, 1)
    x = synth.start_vector(A.n, 2)
    y = synth.start_vector(A.m, 3)
    want_n, want_t = A.matvec(x), A.transpose().matvec(y)
    for panels, lanes, sort in (((pa, pt), (sa, st), (ca, ct)), ((1, 1), None, None),
                                ((8, 2), (2, -2), (1024, 16384)), ((16, 16), (-1, 1), None)):
        op = api.CsrOperator(A, panels=panels, lanes=lanes, sort=sort)
        assert op.panels() == tuple(min(p, 16) for p in panels)
        assert op.lanes() == (lanes or (0, 0))
        # a one-panel direction runs a negated cap panel by panel and reports it plain
        want = tuple(-c if c < -1 and p == 1 else c
                     for c, p in zip(sort or (0, 0), op.panels()))
        assert op.sort() == want
        np.testing.assert_allclose(op.spmv(x, "n"), want_n, rtol=0,
                                   atol=1e-12 * np.abs(want_n).max())
        np.testing.assert_allclose(op.spmv(y, "t"), want_t, rtol=0,
                                   atol=1e-12 * np.abs(want_t).max())


def _edge_matrix(rng, m=300, n=9500):
    D = np.zeros((m, n))
    for i in range(0, m, 3):
        cols = rng.choice(n, rng.integers(1, 20), replace=False)
        D[i, cols] = rng.standard_normal(cols.size)
    D[7, :9000] = rng.standard_normal(9000)
    D[:, 42] = rng.standard_normal(m)
    return _csr_from_dense(D)


@pytest.mark.parametrize("spec", [None, "4"])
@pytest.mark.parametrize("cap", [1024, 8192, 16384])
def test_sorted_spmv_matches_rowblock(panels, spec, cap):
    """Column-sorted layout: each block's products are gathered in column
    order and summed per row from LDS by the same lanes, so every row of at
    most 1024 nonzeros is bit-identical to the row-block layout's; longer
    rows (chunk sums grouped differently) agree to rounding with the float64
    host product.  Edge shapes (empty rows, a 9000-nonzero row, a dense
    column), the power-law matrix (rows of A^T with tens of thousands of
    nonzeros, millions of empty ones), the uniform matrix, a matrix of
    600,000 columns (one panel of it is too wide to pack column and offset
    into one word at caps 8192 and 16384, so the separate offset array is
    used; every other case packs); FP32 operator."""
    if spec:
        panels(spec)
    rng = np.random.default_rng(cap + 7)
    for A in (_edge_matrix(rng), synth.powerlaw_bipartite(400_000, 200_000, seed=3),
              synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4),
              synth.random_uniform(2000, 600_000, 200_000, seed=9)):
        AT = A.transpose()
        op0 = api.CsrOperator(A)
        op1 = api.CsrOperator(A, sort=(cap, cap))
        assert op1.sort() == (cap, cap)
        x, y = rng.standard_normal(A.n), rng.standard_normal(A.m)
        for trans, v, B in (("n", x, A), ("t", y, AT)):
            short = np.diff(B.rowptr) <= 1024
            got, ref = op1.spmv(v, trans), op0.spmv(v, trans)
            assert np.array_equal(got[short], ref[short])
            want = B.matvec(v)
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
            assert np.array_equal(got, op1.spmv(v, trans))
        op32 = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32)
        op32s = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32, sort=(cap, cap))
        x32 = x.astype(np.float32)
        short = np.diff(A.rowptr) <= 1024
        assert np.array_equal(op32s.spmv(x32, "n")[short], op32.spmv(x32, "n")[short])
        # back to the row-block layout: the row-block results, bit for bit
        op1.set_sort(0, 0)
        assert op1.sort() == (0, 0)
        assert np.array_equal(op1.spmv(y, "t"), op0.spmv(y, "t"))


@pytest.mark.parametrize("spec", [None, "4"])
def test_rowwise_spmv_matches_rowblock(panels, spec):
    """Row-wise layout (cap -1: the row-block arrays, each row summed by its
    lanes straight from the index and value streams, no LDS staging): every
    row, long rows included, and the norm partials are bit for bit the
    row-block layout's for the same lanes; so a whole solve is too."""
    if spec:
        panels(spec)
    rng = np.random.default_rng(11)
    for A in (_edge_matrix(rng), synth.powerlaw_bipartite(400_000, 200_000, seed=3),
              synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4),
              synth.random_uniform(2000, 600_000, 200_000, seed=9)):
        x, y = rng.standard_normal(A.n), rng.standard_normal(A.m)
        for lanes in ((0, 0), (-2, 1)):
            op0 = api.CsrOperator(A, lanes=lanes)
            op1 = api.CsrOperator(A, lanes=lanes, sort=(-1, -1))
            assert op1.sort() == (-1, -1)
            for trans, v in (("n", x), ("t", y)):
                assert np.array_equal(op1.spmv(v, trans), op0.spmv(v, trans))
        op32 = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32)
        op32r = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32, sort=(-1, 0))
        x32 = x.astype(np.float32)
        assert np.array_equal(op32r.spmv(x32, "n"), op32.spmv(x32, "n"))
        want = A.matvec(x)
        np.testing.assert_allclose(op1.spmv(x, "n"), want, rtol=0,
                                   atol=1e-10 * np.abs(want).max())


def test_rowwise_same_solve(small_case):
    A, u0, Dm, s = small_case
    r0 = api.lansvd(api.CsrOperator(A), A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    r1 = api.lansvd(api.CsrOperator(A, sort=(-1, -1)), A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    assert r0.k == r1.k == 10 and np.array_equal(r0.sigma, r1.sigma)
    assert r0.counters == r1.counters
    for m, n in ((1, 1), (1, 50), (50, 1), (5000, 3), (3, 5000)):
        rng = np.random.default_rng(m + n)
        D = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
        D[0, 0] = 1.0
        B = _csr_from_dense(D)
        op0, op1 = api.CsrOperator(B), api.CsrOperator(B, sort=(-1, -1))
        for trans, v in (("n", rng.standard_normal(n)), ("t", rng.standard_normal(m))):
            assert np.array_equal(op1.spmv(v, trans), op0.spmv(v, trans))


@pytest.mark.parametrize("cap", [8192, 16384])
def test_concurrent_panels_match_sequential(cap):
    """Column-sorted layout with every column panel in one launch (the cap
    negated): each panel's row sums go to an FP64 array of their own and
    are added in panel order afterwards -- the sums the one-launch-per-panel
    path forms in its running sums -- so every output element is bit for
    bit that path's (only the norm partials are grouped differently).  2, 4,
    8 and 16 panels (16: two panels per XCD in turn); FP64 and FP32.  One
    panel, or long-row chunks in a panel (the power-law matrix), keep the
    panels sequential, and the operator reports the plain cap."""
    rng = np.random.default_rng(cap)
    A = synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4)
    x, y = rng.standard_normal(A.n), rng.standard_normal(A.m)
    for panels in ((2, 4), (4, 8), (16, 16)):
        op0 = api.CsrOperator(A, panels=panels, sort=(cap, cap))
        op1 = api.CsrOperator(A, panels=panels, sort=(-cap, -cap))
        assert op1.sort() == (-cap, -cap)
        for trans, v in (("n", x), ("t", y)):
            got = op1.spmv(v, trans)
            assert np.array_equal(got, op0.spmv(v, trans)), (panels, trans)
            assert np.array_equal(got, op1.spmv(v, trans))
    A32 = A.astype(np.float32)
    op0 = api.CsrOperator(A32, prec=api.PK_F32, panels=(4, 8), sort=(cap, cap))
    op1 = api.CsrOperator(A32, prec=api.PK_F32, panels=(4, 8), sort=(-cap, -cap))
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    assert np.array_equal(op1.spmv(x32, "n"), op0.spmv(x32, "n"))
    assert np.array_equal(op1.spmv(y32, "t"), op0.spmv(y32, "t"))
    assert api.CsrOperator(A, panels=(1, 1), sort=(-cap, -cap)).sort() == (cap, cap)
    P = synth.powerlaw_bipartite(400_000, 200_000, seed=3)
    opp = api.CsrOperator(P, panels=(4, 4), sort=(-cap, -cap))
    assert opp.sort()[1] == cap  # A^T has rows longer than 1024: sequential panels
    xp = rng.standard_normal(P.m)
    assert np.array_equal(opp.spmv(xp, "t"),
                          api.CsrOperator(P, panels=(4, 4), sort=(cap, cap)).spmv(xp, "t"))


def test_concurrent_panels_solve():
    """A restarted solve on concurrent panels: the singular values of the
    sequential-panel solve to 1e-12 and the same operator count.  An engine
    created before its operator switched to concurrent panels keeps the
    sequential path (its scratch holds one running-sum vector) and gives
    that path's singular values bit for bit."""
    B = synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4)
    u1 = synth.start_vector(B.m, 1)
    res = []
    for sort in ((8192, 8192), (-8192, -8192)):
        op = api.CsrOperator(B, panels=(4, 8), sort=sort)
        eng = api.Engine(op, kmax=60)
        eng.put_u1(u1)
        res.append(api.lansvd_irl_dev(eng, 20, 60, tol=1e-10))
        eng.close()
    a, b = res
    assert a.k == b.k == 20 and a.info == b.info == 0
    np.testing.assert_allclose(b.sigma, a.sigma, rtol=1e-12)
    assert a.counters["nopx"] == b.counters["nopx"]
    for i in range(5):
        r = B.matvec(b.V[:, i]) - b.sigma[i] * b.U[:, i]
        assert np.linalg.norm(r) / b.sigma[0] < 1e-11
    op = api.CsrOperator(B, panels=(4, 8), sort=(8192, 8192))
    eng = api.Engine(op, kmax=60)
    op.set_sort(-8192, -8192)
    eng.put_u1(u1)
    c = api.lansvd_irl_dev(eng, 20, 60, tol=1e-10)
    assert np.array_equal(c.sigma, a.sigma)


def test_sorted_layout_rejects_bad_caps():
    A = synth.random_uniform(3000, 1000, 20000, seed=5)
    op = api.CsrOperator(A)
    for caps in ((100, 0), (0, 16385), (-2, 1024), (-100, 0), (0, -16385)):
        with pytest.raises(api.PropackError):
            op.set_sort(*caps)
        assert op.sort() == (0, 0)


def test_sorted_same_triplets(small_case):
    """DLANSVD and DLANSVD_IRL through operators in the column-sorted layout
    give the row-block layout's singular values to rounding and the dense
    SVD's; the restarted solve on a 200k x 100k matrix keeps its operator
    application count."""
    A, u0, Dm, s = small_case
    r0 = api.lansvd(api.CsrOperator(A), A.m, A.n, 10, 300, u0=u0, tol=1e-12)
    r1 = api.lansvd(api.CsrOperator(A, sort=(16384, 8192)), A.m, A.n, 10, 300, u0=u0,
                    tol=1e-12)
    assert r0.k == r1.k == 10
    np.testing.assert_allclose(r1.sigma, r0.sigma, rtol=1e-12)
    np.testing.assert_allclose(r1.sigma, s[:10], rtol=1e-12)
    B = synth.poisson_decay(200_000, 100_000, 16.0, 0.6, seed=4)
    u1 = synth.start_vector(B.m, 1)
    res = []
    for sort in (None, (16384, 16384)):
        op = api.CsrOperator(B, panels=(4, 4), sort=sort)
        eng = api.Engine(op, kmax=60)
        eng.put_u1(u1)
        res.append(api.lansvd_irl_dev(eng, 20, 60, tol=1e-10))
        eng.close()
    a, b = res
    assert a.k == b.k == 20 and a.info == b.info == 0
    np.testing.assert_allclose(b.sigma, a.sigma, rtol=1e-12)
    assert abs(a.counters["nopx"] - b.counters["nopx"]) <= 2
    for i in range(5):
        r = B.matvec(b.V[:, i]) - b.sigma[i] * b.U[:, i]
        assert np.linalg.norm(r) / b.sigma[0] < 1e-11


@pytest.mark.parametrize("m,n", [(1, 1), (1, 50), (50, 1), (5000, 3), (3, 5000)])
@pytest.mark.parametrize("cap", [1024, 16384])
def test_sorted_spmv_tiny(m, n, cap):
    """Column-sorted layout on degenerate shapes (one row, one column, a
    block holding every row, rows of 5000 nonzeros split into chunks):
    the float64 host product, and bit for bit the row-block layout's on
    rows of at most 1024 nonzeros."""
    rng = np.random.default_rng(m * 31 + n + cap)
    D = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
    if not D.any():
        D[0, 0] = 1.0
    A = _csr_from_dense(D)
    op0, op1 = api.CsrOperator(A), api.CsrOperator(A, sort=(cap, cap))
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    for trans, v, want, lens in (("n", x, D @ x, np.count_nonzero(D, axis=1)),
                                 ("t", y, D.T @ y, np.count_nonzero(D, axis=0))):
        got = op1.spmv(v, trans)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))
        short = lens <= 1024
        assert np.array_equal(got[short], op0.spmv(v, trans)[short])


def test_sorted_spmv_empty_matrix():
    """An operator without nonzeros in the sorted layout returns zeros."""
    A = synth.from_coo(40, 30, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
    op = api.CsrOperator(A, sort=(8192, 8192))
    assert op.sort() == (8192, 8192)
    assert not op.spmv(np.ones(30), "n").any() and not op.spmv(np.ones(40), "t").any()


def test_sorted_slansvd_dropin(small_case):
    """SLANSVD through the Fortran drop-in on an FP32 operator in the
    column-sorted layout: the row-block operator's singular values to FP32
    rounding, and the dense FP64 SVD's to FP32-storage accuracy."""
    A, u0, Dm, s = small_case
    r0 = api.lansvd(api.CsrOperator(A, prec=api.PK_F32), A.m, A.n, 10, 300, u0=u0, tol=1e-6,
                    prec=api.PK_F32)
    r1 = api.lansvd(api.CsrOperator(A, prec=api.PK_F32, sort=(8192, 1024)), A.m, A.n, 10, 300,
                    u0=u0, tol=1e-6, prec=api.PK_F32)
    assert r0.k == r1.k == 10 and r0.info == r1.info == 0
    np.testing.assert_allclose(r1.sigma, r0.sigma, rtol=2e-6)
    np.testing.assert_allclose(r1.sigma, s[:10], rtol=2e-6)
