"""Compiled callers of the drop-in drivers (tests/fortran/dropin_example.f90,
tests/fortran/zdropin_example.f90, tests/c/dropin_example.c, built by
tests/build_programs.py): the
reference's calling sequences linked against libpropack_amd, with a host
APROD and with the device operator.  Their INFO and converged counts equal
the compiled reference's on the same matrix and start vector, and their
singular values the dense SVD of the matrix built here."""
import os
import subprocess

import numpy as np
import pytest

from oracle import refpy
from propack_amd import synth

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_bin")


def _run(name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (tests/build_programs.py)")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    out = {}
    for line in r.stdout.splitlines():
        w = line.split()
        if w:
            out[w[0]] = [float(x) for x in w[1:]]
    return out


def test_fortran_caller():
    m, n, per = 300, 200, 6
    D = np.zeros((m, n))
    for i in range(1, m + 1):
        for t in range(1, per + 1):
            D[i - 1, (i * 7 + t * 31 + t * t) % n] += np.sin(0.37 * i + 1.1 * t) + 0.1 * t
    s = np.linalg.svd(D, compute_uv=False)
    out = _run("dropin_example_f")
    r, c = np.nonzero(D)
    A = synth.from_coo(m, n, r, c, D[r, c])
    u0 = np.cos(0.11 * np.arange(1, m + 1))
    ref = refpy.lansvd(A, 8, 200, u0=u0, tol=1e-12, cgs=1)
    assert out["HOST_INFO_K"] == [ref.info, ref.k] == [0, 8]
    np.testing.assert_allclose(out["HOST_SIGMA"], s[:8], rtol=1e-11)
    assert out["DEV_INFO_K"][:2] == [0, 8]
    np.testing.assert_allclose(out["DEV_SIGMA"], s[:8], rtol=1e-11)
    assert out["DEV_RESID1"][0] < 1e-11


def test_c_caller():
    m, n, per = 250, 180, 5
    D = np.zeros((m, n))
    for i in range(m):
        for t in range(1, per + 1):
            D[i, (i * 11 + t * 29 + t * t) % n] += np.cos(0.23 * i + 0.7 * t) + 0.05 * t
    s = np.linalg.svd(D, compute_uv=False)
    out = _run("dropin_example_c")
    r, c = np.nonzero(D)
    A = synth.from_coo(m, n, r, c, D[r, c])
    u0 = np.sin(0.3 * np.arange(m) + 0.1)
    ref = refpy.lansvd(A, 6, 180, u0=u0, tol=1e-12, cgs=1)
    assert out["HOST_INFO_K"] == [ref.info, ref.k] == [0, 6]
    np.testing.assert_allclose(out["HOST_SIGMA"], s[:6], rtol=1e-11)
    assert out["DEV_INFO_K"][:2] == [0, 6]
    np.testing.assert_allclose(out["DEV_SIGMA"], s[:6], rtol=2e-5)


def test_fortran_complex_caller():
    """ZLANSVD with the caller's complex host APROD and ZLANSVD_IRL with the
    device operator (pk_zcsr_aprod), against the dense complex SVD and the
    compiled ZLANSVD."""
    m, n, per = 260, 190, 5
    D = np.zeros((m, n), np.complex128)
    for i in range(1, m + 1):
        for t in range(1, per + 1):
            D[i - 1, (i * 13 + t * 37 + t * t) % n] += complex(
                np.sin(0.29 * i + 0.9 * t) + 0.1 * t, np.cos(0.41 * i - 0.6 * t))
    s = np.linalg.svd(D, compute_uv=False)
    out = _run("zdropin_example_f")
    r, c = np.nonzero(D)
    A = synth.from_coo(m, n, r, c, D[r, c])
    u0 = np.cos(0.11 * np.arange(1, m + 1)) + 1j * np.sin(0.07 * np.arange(1, m + 1))
    refpy.set_threads(1, "z")
    ref = refpy.lansvd(A, 6, 190, u0=u0, tol=1e-12, cgs=0, prec="z")
    assert out["HOST_INFO_K"] == [ref.info, ref.k] == [0, 6]
    np.testing.assert_allclose(out["HOST_SIGMA"], s[:6], rtol=1e-11)
    np.testing.assert_allclose(out["HOST_SIGMA"], ref.sigma, rtol=1e-11)
    assert out["DEV_INFO_K"][:2] == [0, 6]
    np.testing.assert_allclose(out["DEV_SIGMA"], s[:6], rtol=1e-10)
    assert out["DEV_RESID1"][0] < 1e-10
