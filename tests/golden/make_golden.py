"""Writes tests/golden/small_case.npz from the compiled reference PROPACK
(oracle/_ref, built by `make -C oracle ref` from the reference's own
sources) on the 1850 x 712 case: DLANSVD (k=10, kmax=300, tol=1e-12) with
MGS and with ICGS, DLANSVD_IRL (k=10, dim=30), SLANSVD (tol=1e-6), and the
random-start run.  Stored: singular values, bounds, INFO, the COMMON
/timing/ counters, and the first 8 entries of each singular vector.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")

from oracle import refpy  # noqa: E402
from propack_amd import synth  # noqa: E402


def main():
    refpy.set_threads(1)
    refpy.set_threads(1, "s")
    A = synth.random_uniform()
    u0 = synth.start_vector(A.m, 1)
    runs = {
        "lansvd_mgs": lambda: refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=0),
        "lansvd_cgs": lambda: refpy.lansvd(A, 10, 300, u0=u0, tol=1e-12, cgs=1),
        "lansvd_irl": lambda: refpy.lansvd_irl(A, 10, 30, u0=u0, tol=1e-12, cgs=1),
        "lansvd_random_start": lambda: refpy.lansvd(A, 10, 300, u0=None, tol=1e-12, cgs=1),
        "slansvd": lambda: refpy.lansvd(A, 10, 300, u0=u0, tol=1e-6, prec="s", cgs=1),
    }
    out = {"matrix_m": A.m, "matrix_n": A.n, "matrix_nnz": A.nnz,
           "matrix_colsum": float(A.val.astype(np.float64).sum())}
    for name, fn in runs.items():
        r = fn()
        out[f"{name}_sigma"] = r.sigma.astype(np.float64)
        out[f"{name}_bnd"] = r.bnd.astype(np.float64)
        out[f"{name}_info"] = r.info
        for key in ("nopx", "nreorth", "ndot", "nbsvd", "nrestart"):
            out[f"{name}_{key}"] = r.counters[key]
        out[f"{name}_U8"] = r.U[:8].astype(np.float64)
        out[f"{name}_V8"] = r.V[:8].astype(np.float64)
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_case.npz"), **out)
    print("wrote small_case.npz with", len(out), "arrays")


if __name__ == "__main__":
    main()
