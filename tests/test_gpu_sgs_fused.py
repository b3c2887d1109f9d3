"""The dynamic SGS models in one pass each (the test filter applied on the
fly) against the separate kernels and filter launches (LW_SGS_FUSED=0),
which the oracle tests pin: the filtered quantities are the same bits, the
Germano sums differ by the order of their additions, so the fields agree to
rounding over two steps, with the same multigrid cycle counts."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import Solver
from tests.helpers import rel, spec_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,nlev,filt", [((64, 64, 16), 2, 3), ((96, 64, 32), 3, 3), ((130, 40, 20), 1, 3),
                                            ((64, 32, 16), 2, 1), ((64, 32, 16), 2, 2)])
def test_gpu_fused_sgs_equals_separate_kernels(dims, nlev, filt, monkeypatch):
    sp = spec_for(dims, "les", 2, nlev)
    sp.settings.filter = filt
    view = C.case_view(C.build_case(sp))
    print("sgs fused", dims, filt, flush=True)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("LW_SGS_FUSED", flag)
        g = Solver.from_case(view)
        g.phase(0); g.phase(11); g.phase(1)
        nu = {k: g.array(k) for k in ("nusgs", "alsgs")}
        st = g.step(2)
        out.append((st["mg_cycles_sum"], nu, g.fields()))
        g.close()
    for k in ("nusgs", "alsgs"):
        assert rel(out[0][1][k], out[1][1][k]) <= 1e-13, k
    assert out[0][0] == out[1][0]
    for k in "uvwpt":
        assert rel(out[0][2][k], out[1][2][k]) <= 1e-12, k


@pytest.mark.parametrize("dims,filt", [((130, 32, 16), 3), ((62, 32, 8), 1)])
def test_gpu_fused_sgs_on_the_channel_equals_separate_kernels(dims, filt, monkeypatch):
    """The x-periodic channel: the x filter wraps at the end cells in both
    forms (the fused kernel's neighbour lanes take the cells across the
    period); n1m = 130 spans three 62-cell rows, n1m = 62 ends a row exactly."""
    sp = spec_for(dims, "chan_les", 2)
    sp.settings.filter = filt
    view = C.case_view(C.build_case(sp))
    print("sgs fused channel", dims, filt, flush=True)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("LW_SGS_FUSED", flag)
        g = Solver.from_case(view)
        g.phase(0); g.phase(11); g.phase(1)
        nu = g.array("nusgs")
        st = g.step(2)
        out.append((st["mg_cycles_sum"], nu, g.fields()))
        g.close()
    assert rel(out[0][1], out[1][1]) <= 1e-13
    for k in "uvwp":
        assert rel(out[0][2][k], out[1][2][k]) <= 1e-12, k
