"""The C ABI library loads and exports every symbol include/leswht.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from leswht_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "leswht.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lw_[a-z_]+)\s*\(", txt)))


def test_header_declares_the_python_exports():
    assert set(solver.EXPORTS) <= set(declared())


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(solver.LIB_PATH)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_errors_are_reported_not_thrown():
    L = solver.load_library()
    h = ctypes.c_void_p()
    rc = L.lw_create(None, None, None, None, ctypes.byref(h))
    assert rc == -1
    assert b"null" in L.lw_last_error()
