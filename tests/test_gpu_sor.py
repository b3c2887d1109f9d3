"""Line-SOR wavenumbers (gsor2d, poiss_z_ft.f90:513-616; IMGSOR < n3m/2+1)
against the oracle at the 384x256x64 workload's kernel shapes (n2m = 256,
4 rows per lane, the 64-point z-DFT): 33 wavenumbers, 16 of them by line-SOR.
Bar: the FP64 path's, relative L2 <= 1e-10 with equal cycle and sweep
counts; the two-stream solve gives the same bits as one stream."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import Solver
from tests.helpers import rel, spec_for

pytestmark = pytest.mark.gpu


def _view():
    view = C.case_view(C.build_case(spec_for((96, 256, 64), "les", 2, 3)))
    view["settings"].imgsor = 17
    return view


def test_gpu_line_sor_matches_oracle():
    import oracle
    view = _view()
    print("sor 96x256x64 imgsor 17 vs oracle", flush=True)
    o = oracle.Oracle(view)
    do = o.step(2)
    g = Solver.from_case(view)
    dg = g.step(2)
    rows = g.residual_rows()
    f, fo = g.fields(), o.fields()
    g.close()
    for k in "uvwpt":
        assert rel(fo[k], f[k]) <= 1e-10, k
    assert dg["mg_cycles_sum"] == do.mg_cycles_sum and dg["mg_cycles_max"] == do.mg_cycles_max
    # the last step's rows: three mg2d rows per substep, then gsor2d's sweeps
    # of kv = 18..33, kv by kv
    kinds = [r[0] for r in rows]
    assert kinds.count(0) == 9 and kinds.count(1) > 0
    sor = [r for r in rows if r[0] == 1]
    assert all(18 <= r[1] <= 33 for r in sor)
    print("sor cycles", dg["mg_cycles_sum"], "rows", len(rows), flush=True)


def test_gpu_line_sor_two_streams_equal_one():
    view = _view()
    print("sor split vs one stream", flush=True)
    out = []
    for split in (0, 1):
        g = Solver.from_case(view)
        g.set_option("mg_split", split)
        st = g.step(2)
        out.append((st, g.fields(), g.residual_rows()))
        g.close()
    assert out[0][0]["mg_cycles_sum"] == out[1][0]["mg_cycles_sum"]
    for k in "uvwpt":
        assert np.array_equal(out[0][1][k], out[1][1][k]), k
    assert out[0][2] == out[1][2]
