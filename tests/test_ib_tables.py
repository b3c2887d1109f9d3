"""The case builder's immersed-boundary tables against the reference's own
table routines (ibm_body.f90: find_inout, findbdy_intp, geomfac_preset,
geomfac_intp, find_zero_nu_sgs), run by tests/golden/make_ib_golden.py on
the same axes. Bar: bit-exact (counts, indices, FP64 weights, masks)."""
import os

import numpy as np
import pytest

from leswht_amd import case as C
from tests.golden import make_ib_golden as G

NAMES = sorted(G.CASES)


@pytest.mark.parametrize("name", NAMES)
def test_ib_tables_match_reference_routines(name):
    z = np.load(os.path.join(G.OUT, name + ".npz"), allow_pickle=False)
    ax, ay, az = G.axes_for(name)
    for key, a in (("x", ax.coord), ("xm", ax.center), ("y", ay.coord), ("ym", ay.center),
                   ("z", az.coord), ("zm", az.center)):
        assert np.array_equal(z[key], a), key  # the inputs the reference saw
    tab = C.build_ib_tables(C.body_cylinder, ax, ay, az, heat=True, les=True)
    for c in G.COMPS:
        nbody, nintp, ninner = (int(v) for v in z[c + "_counts"])
        assert (tab.nintp[c], tab.ninner[c]) == (nintp, ninner), c
        assert np.array_equal(tab.fcp[c], z[c + "_fcp"]), c
        assert np.array_equal(tab.indx[c], z[c + "_indx"]), c
        got = tab.geomfac_raw[c]
        want = z[c + "_geomfac"]
        assert got.shape == want.shape, c
        # bitwise, NaN-free: compare the IEEE bit patterns
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), \
            (c, int((got != want).sum()), float(np.abs(got - want).max()))
    assert tab.nzero == int(z["nzero"])
    assert np.array_equal(tab.iszero, z["iszero"])


def test_ib_fixtures_cover_exact_surface_hits():
    """The uniform 0.25 grid puts w/t points (cell centres) exactly on the
    cylinder surface, so the fixture set exercises the exact-zero exits of
    the surface search."""
    ax, ay, _ = G.axes_for("uniform_25x25x3")
    on = [(x, y) for x in ax.center for y in ay.center if C.body_cylinder(x, y, 0.0) == 0.0]
    assert len(on) >= 4
