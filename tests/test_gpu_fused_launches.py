"""The fused launches against the per-component ones they replace: the
momentum right-hand side of the three components in one pass (k_rhs_uvw
against k_rhs_u/v/w, LW_RHS_FUSED=0) and the immersed-boundary gather and
scatter of the three components in one launch each (k_ibm_gather3 and
k_ibm_scatter3 against one launch per component, LW_IBM_FUSED=0). Each
component's arithmetic is the same expression in the same order, so the
fields agree to the bit; both forms are also held to the oracle by the
parity tests (the fused one is the default). The x-periodic channel with a
body stores u at i = 1, the plane the fused kernel evaluates everywhere."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import Solver
from tests.helpers import spec_for

pytestmark = pytest.mark.gpu


def _run(view, nsteps, env, monkeypatch):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    g = Solver.from_case(view)
    st = g.step(nsteps)
    out = g.fields()
    g.close()
    for k in env:
        monkeypatch.delenv(k)
    return st, out


@pytest.mark.parametrize("dims,mode", [((96, 64, 16), "les"), ((64, 96, 8), "nosgs"), ((64, 32, 8), "chan_body")])
def test_gpu_fused_launches_bitwise_equal_per_component(dims, mode, monkeypatch):
    view = C.case_view(C.build_case(spec_for(dims, mode, 3, 3)))
    print("fused launches", dims, mode, flush=True)
    st1, f1 = _run(view, 3, {}, monkeypatch)
    st0, f0 = _run(view, 3, {"LW_RHS_FUSED": "0", "LW_IBM_FUSED": "0"}, monkeypatch)
    for k in "uvwpt":
        assert np.array_equal(f1[k], f0[k]), k
    assert st1["mg_cycles_sum"] == st0["mg_cycles_sum"]
