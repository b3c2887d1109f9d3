"""The C ABI library loads and exports every symbol include/leswht.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

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


def test_fortran_host_reports_errors():
    """solver_exec run where there is no case ends with a non-zero status and
    the library's message (the reference would stop with status 0)."""
    import subprocess
    import tempfile
    from leswht_amd.build import EXE
    if not os.path.exists(EXE):
        pytest.skip("solver_exec not built (amdflang absent at build time)")
    with tempfile.TemporaryDirectory() as d:
        res = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
    assert res.returncode != 0
    assert "failed" in (res.stdout + res.stderr)
