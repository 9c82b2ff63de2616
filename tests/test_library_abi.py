"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls)."""
import ctypes as C
import os
import re

import numpy as np

from propack_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("propack_amd.h", "propack.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(", txt):
            name = m.group(1)
            if name.startswith(("pk_", "pkh_")) or name in (
                    "dlansvd_", "dlansvd_irl_", "slansvd_", "slansvd_irl_", "zlansvd_",
                    "zlansvd_irl_", "clansvd_", "clansvd_irl_", "clearstat_", "printstat_"):
                names.add(name)
    # typedef'd function-pointer names are not exports
    return sorted(n for n in names if not n.endswith("_t"))


def test_library_exports_every_declared_symbol():
    L = api.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_export_list_is_declared():
    assert set(api.EXPORTS) <= set(declared_functions())


def test_timing_common_block_layout():
    api.clearstat()
    t = api.timing()
    assert all(v == 0 for v in t.values())
    assert C.sizeof(api._Timing) == 27 * 4


def test_host_rng_is_pure_cpu():
    x = np.zeros(5)
    assert api.lib().pk_larnv_pm1(0, C.c_longlong(5), x.ctypes.data_as(C.c_void_p)) == 0
    assert np.all(np.abs(x) < 1)


def test_errors_do_not_abort():
    L = api.lib()
    h = C.c_void_p()
    # invalid arguments return an error code instead of stopping the process
    assert L.pk_lbd_create(C.c_int(0), C.c_int(0), C.c_int(5), C.c_int(3), None,
                           C.byref(h)) == -101
    assert L.pk_lbd_destroy(C.c_void_p(0)) == -104
    assert L.pk_csr_destroy(C.c_void_p(0)) == -104
