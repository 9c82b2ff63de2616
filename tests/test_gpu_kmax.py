"""Lanczos dimensions past 510 (the engine's control arrays are sized at run
time; steps past 500 run the device control on arrays in device memory
instead of the LDS window): the drop-in accepts every kmax the reference
does (double/dlansvd.F:146: lanmax = min(n+1, m+1, kmax)).

Tolerances: singular values 1e-10 relative to the reference; the device
control and the host executor bit-identical (same control code, same
deterministic reductions)."""
import numpy as np
import pytest

from oracle import refpy
from propack_amd import api, synth
from propack_amd.svdp import svdp

pytestmark = pytest.mark.gpu


def test_dlansvd_k200_kmax713(small_case):
    """k = 200 of the 1850 x 712 matrix with kmax = 713 (lanmax = n + 1):
    the Lanczos dimension reaches 524."""
    A, u0, D, s = small_case
    refpy.set_threads(1)
    ref = refpy.lansvd(A, 200, 713, u0=u0, tol=1e-12, cgs=1)
    refpy.set_threads(8)
    r = api.lansvd(api.CsrOperator(A), A.m, A.n, 200, 713, u0=u0, tol=1e-12)
    print("gpu", r.info, r.k, r.counters)
    print("ref", ref.info, ref.k, ref.counters)
    assert r.info == ref.info == 0 and r.k == ref.k == 200
    np.testing.assert_allclose(r.sigma, ref.sigma, rtol=1e-10)
    np.testing.assert_allclose(r.sigma, s[:200], rtol=1e-11)
    assert r.counters["nlandim"] > 500
    res = np.linalg.norm(D @ r.V - r.U * r.sigma, axis=0)
    assert res.max() <= 1e-11 * s[0]


def test_device_control_past_lds_window_matches_host(small_case):
    A, u0, D, s = small_case
    op = api.CsrOperator(A)
    out = {}
    for mode in ("host", "device"):
        eng = api.Engine(op, kmax=713)
        eng.set_mode(mode, 2)
        eng.put_u1(u0)
        r = api.lansvd_dev(eng, 200, 713, tol=1e-12)
        out[mode] = (r, eng.stats())
    (rh, sh), (rd, sd) = out["host"], out["device"]
    assert rh.info == rd.info == 0 and rh.k == rd.k == 200
    assert np.array_equal(rh.sigma, rd.sigma)
    assert np.array_equal(rh.bnd, rd.bnd)
    for key in ("nopx", "nreorth", "ndot"):
        assert sh[key] == sd[key], key


def test_svdp_default_kmax():
    """scipy's default kmax = 10 k: k = 60 gives a Lanczos dimension of 600
    on a 3000 x 1200 matrix."""
    A = synth.random_uniform(3000, 1200, 30000, seed=21)
    s = np.linalg.svd(A.todense(), compute_uv=False)
    import scipy.sparse as sp
    M = sp.csr_array((A.val, A.col, A.rowptr), shape=(A.m, A.n))
    u, sig, vt = svdp(M, 60)
    np.testing.assert_allclose(sig, s[:60], rtol=1e-8)
    D = A.todense()
    res = np.linalg.norm(D @ vt.T - u * sig, axis=0)
    assert res.max() <= 1e-8 * s[0]
