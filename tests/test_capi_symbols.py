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
