"""The library's in-process immersed-boundary table builder (lw_ib_build)
against the reference's own table routines (the golden tables of
tests/golden/ibm, made by make_ib_golden.py from ibm_body.f90) and against
the case builder's tables and text files for other bodies and grids. Bar:
bit-exact (counts, indices, FP64 weights, masks, file bytes)."""
import os
import tempfile

import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import IBTablesNative
from tests.golden import make_ib_golden as G


def _native(lines, periodic, body=None, heat=True, les=True):
    return IBTablesNative(*lines, periodic=periodic, body=body, heat=heat, les=les)


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_native_tables_match_reference_routines(name):
    xs, ys, zs, zper = G.CASES[name]
    nat = _native((G._lines(*xs), G._lines(*ys), G._lines(*zs)), (0, 0, int(zper)))
    tab = nat.tables()
    z = np.load(os.path.join(G.OUT, name + ".npz"), allow_pickle=False)
    for c in G.COMPS:
        nbody, nintp, ninner = (int(v) for v in z[c + "_counts"])
        assert (tab[c]["nintp"], tab[c]["nbody"] - tab[c]["nintp"]) == (nintp, ninner), c
        assert np.array_equal(tab[c]["fcp"], z[c + "_fcp"]), c
        assert np.array_equal(tab[c]["indx"], z[c + "_indx"]), c
        got, want = tab[c]["geomfac_raw"], z[c + "_geomfac"]
        assert got.shape == want.shape and np.array_equal(got.view(np.int64), want.view(np.int64)), c
    assert tab["nzero"] == int(z["nzero"])
    assert np.array_equal(tab["iszero"], z["iszero"])


BODIES = [("cylinder", (0.5,)), ("square", (0.5, 0.0)), ("square", (0.45, 0.9)),
          ("ellipse", (0.5, 0.25, 0.0)), ("ellipse", (0.6, 0.2, 2.2))]


@pytest.mark.parametrize("body,args", BODIES)
@pytest.mark.parametrize("les", [True, False])
def test_native_case_files_equal_case_builder(body, args, les):
    """The ibmpre files the library writes are the case builder's, byte for
    byte, and the tables it hands to lw_create are the ones the solver would
    read back from them."""
    s = C.Settings(iles=1 if les else 0, ihtrans=1)
    spec = C.CaseSpec(40, 32, 6, xlim=(-3.0, 6.0), ylim=(-3.0, 3.0), zlen=1.0, dmin=0.05,
                      body=body, body_args=args, settings=s)
    case = C.build_case(spec)
    lines = tuple(case_axis.coord[1:] for case_axis in (case.ax, case.ay, case.az))
    nat = _native(lines, (0, 0, 1), body=C.body_params(body, args), les=les)
    with tempfile.TemporaryDirectory() as d:
        root = os.path.join(d, "py")
        C.write_case(case, root)
        nd = os.path.join(d, "native")
        os.makedirs(nd)
        nat.write(nd)
        names = ["ibmpre_fcpts.bin", "ibmpre_fcpts_t.bin", "ibmpre_nutzero.bin"] + \
            (["ibmpre_wallfdvm.bin"] if les else [])
        for f in names:
            a = open(os.path.join(root, "output", "ibmpre", f), "rb").read()
            b = open(os.path.join(nd, f), "rb").read()
            assert a == b, f
        assert not os.path.exists(os.path.join(nd, "ibmpre_wallfdvm.bin")) or les
    view = C.case_view(case)
    tab = nat.tables()
    for c in ("u", "v", "w", "t"):
        assert np.array_equal(tab[c]["geomfac"].view(np.int64),
                              np.asarray(view["tables"][c]["geomfac"]).reshape(-1, 3, 3, 3).view(np.int64)), c
        assert np.array_equal(tab[c]["fcp"], np.asarray(view["tables"][c]["fcp"]).reshape(-1, 3)), c


def test_native_builder_rejects_bad_input():
    from leswht_amd.solver import LwError
    with pytest.raises(LwError):
        _native((np.array([0.0, 1.0]), np.linspace(0, 1, 5), np.linspace(0, 1, 5)), (0, 0, 1))
