"""pk_csr_spmv_noalloc, the SpMV entry point that allocates nothing (its
scratch is a fixed device array per GPU), against pk_csr_spmv on the same
operators: bit-identical outputs (the same kernels in the same order; for
an operator whose panels pk_csr_spmv runs in one launch, the same row sums
panel by panel), on
two calls in a row on one GPU, the second on a smaller operator -- the
shared scratch still holds the first call's contents -- in both
directions and both precisions; and an operator whose scratch would not
fit is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from propack_amd import api, synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _long_row_case(seed):
    """40000 x 20000, Poisson(12) per row, plus a row of 6000 nonzeros (its
    chunk sums go through the scratch's long-row part)."""
    A = synth.poisson_decay(40000, 20000, 12.0, 0.6, seed=seed)
    g = np.random.Generator(np.random.Philox(seed))
    lr = g.choice(A.n, 6000, replace=False)
    rows = np.concatenate([np.repeat(np.arange(A.m), np.diff(A.rowptr)), np.full(lr.size, 7)])
    cols = np.concatenate([A.col, lr])
    vals = np.concatenate([A.val, g.standard_normal(lr.size)])
    return synth.from_coo(A.m, A.n, rows, cols, vals)


def _noalloc(op, trans, x, y):
    return api.lib().pk_csr_spmv_noalloc(op.handle, C.c_char(trans.encode()),
                                         C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                         C.c_int(op.prec), C.c_void_p(0))


def _check(op, A, trans, seed):
    tdt = torch.float64 if op.prec == api.PK_F64 else torch.float32
    ndt = np.float64 if op.prec == api.PK_F64 else np.float32
    nin, nout = (A.n, A.m) if trans == "n" else (A.m, A.n)
    x = synth.start_vector(nin, seed).astype(ndt)
    ref = op.spmv(x, trans)
    dx = torch.from_numpy(x).cuda()
    dy = torch.full((nout,), float("nan"), dtype=tdt, device="cuda")
    print("noalloc", A.m, A.n, A.nnz, trans, op.prec, flush=True)
    assert _noalloc(op, trans, dx, dy) == 0, api.lib().pk_last_error()
    torch.cuda.synchronize()
    assert np.array_equal(dy.cpu().numpy(), ref)


@pytest.mark.parametrize("prec", [api.PK_F64, api.PK_F32])
def test_noalloc_matches_spmv_two_calls(small_case, prec):
    dt = np.float64 if prec == api.PK_F64 else np.float32
    big = _long_row_case(41)
    big = synth.Csr(big.m, big.n, big.rowptr, big.col, big.val.astype(dt))
    A, u0, D, s = small_case
    small = synth.Csr(A.m, A.n, A.rowptr, A.col, A.val.astype(dt))
    op_big = api.CsrOperator(big, prec=prec, panels=(4, 4))
    op_small = api.CsrOperator(small, prec=prec)
    for trans in ("n", "t"):
        _check(op_big, big, trans, 1)      # four panels and a long row
        _check(op_small, small, trans, 2)  # then a smaller, single-panel operator
    # the sorted layout of the larger operator, after the smaller one ran
    op_big.set_sort(8192, 8192)
    _check(op_big, big, "t", 3)
    # every panel in one launch for pk_csr_spmv; the no-allocation call runs
    # the panels one launch each through its fixed scratch: the same bits
    conc = synth.poisson_decay(40000, 20000, 12.0, 0.6, seed=43)
    conc = synth.Csr(conc.m, conc.n, conc.rowptr, conc.col, conc.val.astype(dt))
    op_conc = api.CsrOperator(conc, prec=prec, panels=(4, 8), sort=(-8192, -8192))
    assert op_conc.sort() == (-8192, -8192)
    for trans in ("n", "t"):
        _check(op_conc, conc, trans, 4)


def test_noalloc_refuses_oversized_scratch():
    """7,000,000 rows in two column panels need 7M running row sums, more
    than the 6M doubles of the fixed scratch: refused, nothing launched."""
    m, n = 7_000_000, 16
    A = synth.Csr(m, n, np.arange(m + 1, dtype=np.int32),
                  (np.arange(m) % n).astype(np.int32), np.ones(m))
    op = api.CsrOperator(A, panels=(2, 1))
    assert op.panels()[0] == 2
    dx = torch.ones(n, dtype=torch.float64, device="cuda")
    dy = torch.zeros(m, dtype=torch.float64, device="cuda")
    rc = _noalloc(op, "n", dx, dy)
    assert rc == -105  # PK_ERR_NOTSUP
    assert b"scratch" in api.lib().pk_last_error()
    # the transposed direction (16 rows, one panel) fits
    dxt = torch.ones(m, dtype=torch.float64, device="cuda")
    dyt = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert _noalloc(op, "t", dxt, dyt) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dyt.cpu().numpy(), np.full(n, m / n))
