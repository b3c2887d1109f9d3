"""The multigrid tile autotuner (lw_autotune): a tuned handle steps to the
same bits as an untuned one, and the chosen tiles are valid sets."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import Solver
from tests.helpers import spec_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,nlev,split", [((96, 64, 16), 3, "0"), ((64, 64, 16), 3, "1")])
def test_gpu_autotuned_tiles_give_the_same_bits(dims, nlev, split, monkeypatch):
    monkeypatch.setenv("LW_MG_SPLIT", split)
    view = C.case_view(C.build_case(spec_for(dims, "les", 3, nlev)))
    print("autotune", dims, "split", split, flush=True)
    plain = Solver.from_case(view)
    tuned = Solver.from_case(view)
    r = tuned.autotune(reps=1)
    assert r["candidates"] >= 7 * (2 if split == "1" else 1)
    for t in (r["tile"], r["tile1"]):
        assert all(v in (2, 4, 8) for v in t)
    assert 0.0 < r["ms_best"] <= r["ms_default"]
    sp, st = plain.step(3), tuned.step(3)
    assert (sp["mg_cycles_max"], sp["mg_cycles_sum"], sp["dt"]) == (st["mg_cycles_max"], st["mg_cycles_sum"], st["dt"])
    fp, ft = plain.fields(), tuned.fields()
    for k in "uvwpt":
        assert np.array_equal(fp[k], ft[k]), k
