"""The C-ABI library loads on a host without a GPU and exports every function
include/treesvd.h declares (no compute calls here)."""
import ctypes
import os
import re

from treesvd_amd import capi

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared():
    text = open(os.path.join(ROOT, "include", "treesvd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    names = set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text))
    return {n for n in names if n.startswith("tsvd_") or n == "frPCA"}


def test_header_and_binding_agree():
    assert declared() == set(capi.EXPORTS)


def test_library_exports_every_symbol():
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in capi.EXPORTS if not hasattr(L, s)]
    assert not missing, missing


def test_noalloc_frpca_rejects_sizes_beyond_its_workspace_before_any_device_call():
    """tsvd_frpca_csr_noalloc checks its size limits (m <= 4096, n <= 32768,
    k + 5 <= 136) before touching the device, so this runs without a GPU."""
    L = ctypes.CDLL(capi.LIB_PATH)
    L.tsvd_last_error.restype = ctypes.c_char_p
    f = L.tsvd_frpca_csr_noalloc
    f.argtypes = [ctypes.c_int32] * 3 + [ctypes.c_void_p] * 7 + [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64]
    for m, n, k, what in [(4097, 100, 8, "rows"), (100, 32769, 8, "columns"), (100, 100, 132, "rank"),
                          (0, 100, 8, "rows")]:
        rc = f(m, n, 0, None, None, None, None, None, None, None, k, 6, 1)
        assert rc == -1
        assert what in L.tsvd_last_error().decode()
