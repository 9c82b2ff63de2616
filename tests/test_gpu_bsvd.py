"""The batched device bidiagonal SVD (pk_bsvd_batch, pk_batch_set_bsvd):
the restarted driver's error-bound step (double/dlansvd_irl.F:226-229,
DBDSQR with one row of left vectors) for many problems in one launch.

Oracle: the dense LAPACK SVD (numpy) of each bidiagonal.  Tolerances:
singular values 1e-13 relative to the largest (both methods are backward
stable; one-sided Jacobi on a bidiagonal also keeps the small values
relatively accurate, checked at 1e-12 on values above 1e-8 of the
largest); the last components of the left singular vectors to 1e-9 in
magnitude for values separated from their neighbours by more than 1e-4
relative (the components of nearly equal values are not determined one by
one).  The batched driver with device SVDs against the reference
DLANSVD_IRL: singular values 1e-10 relative, the same converged count."""
import numpy as np
import pytest

from oracle import refpy
from propack_amd import api

from test_gpu_batch import NEIG, DIM, TOL, _problems, _batched

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(20)
    out = []
    for n in (1, 2, 3, 7, 47, 48, 64, 100, 128):
        out.append((rng.standard_normal(n), rng.standard_normal(n - 1)))
    # graded (Lanczos-like decaying), a zero on the diagonal, tiny and huge scales
    n = 48
    out.append((10.0 ** -np.linspace(0, 12, n) * (1 + rng.random(n)), rng.random(n - 1) * 1e-3))
    d = rng.standard_normal(n)
    d[20] = 0.0
    out.append((d, rng.standard_normal(n - 1)))
    out.append((1e-200 * rng.standard_normal(n), 1e-200 * rng.standard_normal(n - 1)))
    out.append((1e200 * rng.standard_normal(n), 1e200 * rng.standard_normal(n - 1)))
    return out


def test_bsvd_against_dense_svd():
    cases = _cases()
    S, X = api.bsvd_batch([c[0] for c in cases], [c[1] for c in cases])
    for (d, e), s, x in zip(cases, S, X):
        n = len(d)
        B = np.diag(d) + (np.diag(e, 1) if n > 1 else 0)
        U, sd, _ = np.linalg.svd(np.atleast_2d(B))
        assert np.all(np.diff(s) <= 0), "decreasing order"
        assert np.abs(s - sd).max() <= 1e-13 * sd[0], (n, np.abs(s - sd).max() / sd[0])
        big = sd > 1e-8 * sd[0]
        assert np.all(np.abs(s[big] - sd[big]) <= 1e-12 * sd[big])
        gap = np.ones(n, bool)
        if n > 1:
            rel = np.abs(np.diff(sd)) / sd[0]
            gap[:-1] &= rel > 1e-4
            gap[1:] &= rel > 1e-4
        assert np.abs(np.abs(x[gap]) - np.abs(U[-1, gap])).max(initial=0) <= 1e-9, n


def test_bsvd_rejects_large_order():
    with pytest.raises(api.PropackError):
        api.bsvd_batch([np.ones(129)], [np.ones(128)])


def test_batch_device_bsvd_matches_reference():
    """Four uneven problems with restarts: the batched driver with its
    bidiagonal SVDs on the device gives the reference's converged count and
    singular values (1e-10), and the host-DBDSQR batch's to 1e-12."""
    probs = _problems(4, 2_000, 40_000, seed=13)
    host = _batched(probs)
    for prm, A, op, eng, u0 in probs:
        eng.put_u1(u0)
        eng.reset_stats()
    b = api.Batch([p[3] for p in probs])
    b.set_bsvd(True)
    dev = api.lansvd_irl_batch(b, NEIG, DIM, tol=TOL)
    b.close()
    for rh, rd, (prm, A, op, eng, u0) in zip(host, dev, probs):
        print(prm["m"], rh.k, rd.k, rh.counters["nopx"], rd.counters["nopx"])
        assert rd.info == 0 and rd.k == rh.k == NEIG
        np.testing.assert_allclose(rd.sigma, rh.sigma, rtol=1e-12)
    for (prm, A, op, eng, u0), rd in list(zip(probs, dev))[:2]:
        ref = refpy.lansvd_irl(A, NEIG, DIM, u0=u0, tol=TOL, cgs=1, jobu="n", jobv="n")
        assert ref.k == rd.k
        np.testing.assert_allclose(rd.sigma, ref.sigma, rtol=1e-10)
