"""The case builder's and the library's text writers against the
reference's own pre-processor writers (fixtures written by
tests/golden/make_preproc_golden.py, which calls lib_setgrid.setgrid and
lib_preprocessing's writers, lib_preprocessing.py:112-338): grid.bin from
the same grid.dat, and the ibmpre_*.bin tables, byte for byte."""
import json
import os
import tempfile

import numpy as np
import pytest

from leswht_amd import case as C

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "preproc")
NAMES = sorted(f[:-4] for f in os.listdir(FIX) if f.endswith(".npz"))
FILES = {"grid_bin": "output/grid/grid.bin", "fcpts": "output/ibmpre/ibmpre_fcpts.bin",
         "fcpts_t": "output/ibmpre/ibmpre_fcpts_t.bin", "nutzero": "output/ibmpre/ibmpre_nutzero.bin",
         "wallfdvm": "output/ibmpre/ibmpre_wallfdvm.bin"}


def _load(name):
    z = np.load(os.path.join(FIX, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    s = C.Settings(iles=1, idvmon=1, ihtrans=1)
    spec = C.CaseSpec(*meta["dims"], xlim=tuple(meta["xlim"]), ylim=tuple(meta["ylim"]), zlen=meta["zlen"],
                      dmin=meta["dmin"], body=meta["body"], body_args=tuple(meta["body_args"]), settings=s)
    return z, spec


@pytest.mark.parametrize("name", NAMES)
def test_case_builder_files_equal_reference_writers(name):
    z, spec = _load(name)
    case = C.build_case(spec)
    with tempfile.TemporaryDirectory() as root:
        C.write_case(case, root)
        for key, rel in FILES.items():
            got = open(os.path.join(root, rel), "rb").read()
            assert got == z[key].tobytes(), (name, key)
    # the grid lines the reference read: grid.dat holds the case's lines
    lines = z["grid_dat"].tobytes().decode().split("\n")
    xs = np.array([float(v) for v in lines[2].split()])
    assert np.array_equal(xs, case.ax.coord[1:case.ax.ncell + 2])


@pytest.mark.parametrize("name", NAMES)
def test_library_table_files_equal_reference_writers(name):
    from leswht_amd.solver import IBTablesNative
    z, spec = _load(name)
    case = C.build_case(spec)
    lines = tuple(a.coord[1:] for a in (case.ax, case.ay, case.az))
    nat = IBTablesNative(*lines, periodic=(0, 0, 1), body=C.body_params(spec.body, spec.body_args),
                         heat=True, les=True)
    try:
        with tempfile.TemporaryDirectory() as d:
            nat.write(d)
            for key in ("fcpts", "fcpts_t", "nutzero", "wallfdvm"):
                got = open(os.path.join(d, os.path.basename(FILES[key])), "rb").read()
                assert got == z[key].tobytes(), (name, key)
    finally:
        nat.close()
