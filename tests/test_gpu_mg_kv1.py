"""kv = 1 of the two-stream Poisson solve as one launch (k_mg_kv1, a grid of
co-resident workgroups with grid-wide barriers between the stages of each
V-cycle) against the same cycles as one launch per level, and against the
oracle. The stages are the level kernels' own device code, so the fields are
expected bit-identical and the cycle counts equal; against the oracle the
bar is the FP64 path's, relative L2 <= 1e-10."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import Solver
from tests.helpers import c2_spec, rel, spec_for

pytestmark = pytest.mark.gpu


def _run(view, nsteps, kv1, mixed=False, split=None):
    g = Solver.from_case(view)
    if split is not None:
        g.set_option("mg_split", split)
    g.set_option("mg_kv1", kv1)
    if mixed:
        g.set_option("mg_fp32", 1)
    sts = [g.step(1) for _ in range(nsteps)]
    out = (sts, g.fields(), g.mg_residuals())
    g.close()
    return out


@pytest.mark.parametrize("mixed", [False, True])
def test_gpu_kv1_one_launch_matches_level_launches_c2(mixed):
    view = C.case_view(C.build_case(c2_spec(2)))
    print("kv1 c2 mixed=%s" % mixed, flush=True)
    a = _run(view, 2, 0, mixed)
    b = _run(view, 2, 1, mixed)
    for sa, sb in zip(a[0], b[0]):
        assert sa["mg_cycles_sum"] == sb["mg_cycles_sum"] and sa["mg_cycles_max"] == sb["mg_cycles_max"]
        assert sa["dt"] == sb["dt"]
    for k in "uvwpt":
        assert np.array_equal(a[1][k], b[1][k]), (k, rel(a[1][k], b[1][k]))
    assert a[2] == b[2]
    print("kv1 c2 cycles", [s["mg_cycles_sum"] for s in b[0]], flush=True)


def test_gpu_kv1_one_launch_matches_oracle():
    import oracle
    view = C.case_view(C.build_case(spec_for((96, 256, 64), "les", 2, 3)))
    print("kv1 96x256x64 vs oracle", flush=True)
    o = oracle.Oracle(view)
    do = o.step(2)
    g = Solver.from_case(view)
    g.set_option("mg_split", 1)
    g.set_option("mg_kv1", 1)
    dg = g.step(2)
    f, fo = g.fields(), o.fields()
    g.close()
    for k in "uvwpt":
        assert rel(fo[k], f[k]) <= 1e-10, k
    assert dg["mg_cycles_sum"] == do.mg_cycles_sum and dg["mg_cycles_max"] == do.mg_cycles_max
