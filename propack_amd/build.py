"""Build libpropack_amd.so in-tree.

  * HIP sources (propack_amd/csrc/*.hip) -> hipcc for gfx950 only.
  * Host drivers (propack_amd/fortran/*.f90) -> flang.
  * Linked into propack_amd/_lib/libpropack_amd.so against the HIP runtime,
    the flang runtime (static) and the image's MKL for host LAPACK
    (DBDSQR / DBDSDC on the small bidiagonal problems).

Run ``python -m propack_amd.build`` (or ``__graft_entry__.build()``).
Objects go to propack_amd/_build/ (git-ignored); the library is rebuilt on
every call.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libpropack_amd.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "lib", "llvm", "bin", "flang")
FLANG_RT = os.path.join(ROCM, "lib", "llvm", "lib", "clang", "22", "lib",
                        "x86_64-unknown-linux-gnu")
MKLDIR = "/opt/conda/lib"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions stay off for every kernel of this library.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-result",
]
F_FLAGS = ["-O2", "-fPIC"]

HIP_SOURCES = ["pk_kernels.hip", "pk_csr.hip", "pk_engine.hip", "pk_batch.hip", "pk_ingest.hip",
               "pk_cplx.hip",
               "pk_bsvd.hip"]
F_SOURCES = ["pk_api.f90", "pk_lansvd.f90", "pk_dropin.f90"]  # module order


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for src in HIP_SOURCES:
            obj = os.path.join(BUILD, src.replace(".hip", ".o"))
            objs.append(obj)
            jobs.append(ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c",
                                         os.path.join(CSRC, src), "-o", obj]))
        # Fortran modules depend on each other: compile in order.
        for src in F_SOURCES:
            obj = os.path.join(BUILD, src.replace(".f90", ".o"))
            objs.append(obj)
            _run([FLANG, *F_FLAGS, "-module-dir", BUILD, "-I" + BUILD, "-c",
                  os.path.join(FSRC, src), "-o", obj])
        for j in jobs:
            out = j.result()
            if verbose and out:
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "-o", tmp, *objs,
          "-L" + FLANG_RT, "-lflang_rt.runtime", "-lflang_rt.quadmath",
          "-L" + MKLDIR, "-lmkl_rt",
          # System and ROCm directories precede MKL's so that the older
          # C++ runtime shipped beside MKL never shadows the system one.
          "-Wl,-rpath," + ":".join([os.path.join(ROCM, "lib"), "/usr/lib/x86_64-linux-gnu",
                                    MKLDIR]),
          "-lm"])
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
