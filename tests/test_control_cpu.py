"""The engine's step control (propack_amd/csrc/pk_lbd.h: the phase machine
the GPU control kernels and the host executor share) driven on the CPU by
a test-only C++ executor (tests/lbd_cpu.cpp), against the oracle's
DLANBPRO restatement on the same inputs: B_k, the bases and the counters
must agree (1e-12 / identical counts) -- so control-logic errors show up
without a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import portpy

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("lbd") / "liblbd_cpu.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    os.path.join(HERE, "lbd_cpu.cpp"), "-o", out], check=True)
    return C.CDLL(out)


def P(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.skipif(not portpy.available(), reason="oracle not built")
@pytest.mark.parametrize("k,delta", [(50, None), (120, None), (40, 0.0)])
def test_phase_machine_matches_oracle_lanbpro(small_case, harness, k, delta):
    A, u0, D, s = small_case
    AT = A.transpose()
    eps = 2.0 ** -52
    dl = np.sqrt(eps) if delta is None else delta
    et = eps ** 0.75
    m, n = A.m, A.n
    arrs = [np.ascontiguousarray(x) for x in (A.rowptr, A.col, A.val, AT.rowptr, AT.col, AT.val)]
    # oracle restatement
    L = portpy.lib()
    L.pkp_d_set_csr(C.c_int(m), C.c_int(n), *[P(a) for a in (arrs[0], arrs[1], arrs[2],
                                                                arrs[3], arrs[4], arrs[5])])
    U1 = np.zeros((k + 1, m)); U1[0] = u0
    V1 = np.zeros((k, n))
    a1, b1 = np.zeros(k), np.zeros(k)
    rn1 = C.c_double(np.linalg.norm(u0))
    dopt = np.array([dl, et, 0.0, 0.0])
    iopt = np.array([1, 1], np.int32)
    kk = C.c_int(k)
    L.pkp_d_clear_stats()
    fn = C.cast(L.pkp_d_csr_aprod, C.c_void_p)
    ierr1 = L.pkp_d_lanbpro(C.c_int(m), C.c_int(n), C.c_int(0), C.byref(kk), fn, None, P(U1),
                            C.c_int(m), P(V1), C.c_int(n), P(a1), P(b1), C.byref(rn1), P(dopt),
                            P(iopt))
    st = portpy._stats("d")
    # product control on the CPU harness
    U2 = np.zeros((k + 1, m)); U2[0] = u0
    V2 = np.zeros((k, n))
    a2, b2 = np.zeros(k), np.zeros(k)
    rn2 = C.c_double(np.linalg.norm(u0))
    cnt = np.zeros(4, np.int64)
    rc = harness.lbd_cpu_lanbpro(C.c_int(m), C.c_int(n), *[P(a) for a in arrs], P(U2), P(V2),
                                 C.c_int(k), P(a2), P(b2), C.byref(rn2), C.c_double(dl),
                                 C.c_double(et), C.c_int(1), P(cnt))
    assert rc == 0
    assert kk.value == k
    np.testing.assert_allclose(a2, a1, rtol=1e-12)
    np.testing.assert_allclose(b2, b1, rtol=1e-12)
    assert abs(rn2.value - rn1.value) <= 1e-12 * abs(rn1.value)
    assert cnt[0] == st["nopx"] and cnt[1] == st["nreorth"] and cnt[2] == st["ndot"]
    assert int(cnt[3]) == ierr1
    assert np.abs(U2 - U1).max() < 1e-10 and np.abs(V2 - V1).max() < 1e-10
