"""Build the drop-in caller programs of tests/fortran and tests/c against
propack_amd/_lib/libpropack_amd.so (outputs in tests/_bin/).  Run by
__graft_entry__.build() after the library; the GPU tests only run them."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(ROOT, "propack_amd", "_lib")
OUT = os.path.join(HERE, "_bin")
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def build():
    os.makedirs(OUT, exist_ok=True)
    link = ["-L" + LIB, "-lpropack_amd", "-Wl,-rpath," + LIB]
    cmds = [
        [FLANG, "-O2", os.path.join(HERE, "fortran", "dropin_example.f90"), "-o",
         os.path.join(OUT, "dropin_example_f"), "-module-dir", OUT, *link],
        [FLANG, "-O2", os.path.join(HERE, "fortran", "zdropin_example.f90"), "-o",
         os.path.join(OUT, "zdropin_example_f"), "-module-dir", OUT, *link],
        ["gcc", "-O2", "-std=c11", "-I" + os.path.join(ROOT, "include"),
         os.path.join(HERE, "c", "dropin_example.c"), "-o", os.path.join(OUT, "dropin_example_c"),
         *link, "-lm"],
    ]
    for c in cmds:
        r = subprocess.run(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("build of a drop-in test program failed:\n" + " ".join(c) + "\n" +
                               r.stdout)
    return OUT


if __name__ == "__main__":
    print(build())
