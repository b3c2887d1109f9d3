"""lw_create_static, the entry point that allocates no device memory: a case
in the device's static area steps to the same bits as the same case created
by lw_create, on two cases in a row (the second smaller, in the area the
first one used); a second static handle and a case larger than the area are
refused with an error."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import LwError, Solver, load_library
from tests.helpers import spec_for

pytestmark = pytest.mark.gpu


def _view(dims, mode, nlev):
    return C.case_view(C.build_case(spec_for(dims, mode, 1, nlev)))


@pytest.mark.parametrize("dims,mode,nlev,nsteps", [((64, 64, 16), "les", 2, 3), ((32, 32, 8), "nosgs", 2, 4)])
def test_gpu_static_area_case_equals_allocated_case(dims, mode, nlev, nsteps):
    view = _view(dims, mode, nlev)
    print("static area", dims, mode, flush=True)
    a = Solver.from_case(view)
    sa = a.step(nsteps)
    fa = a.fields()
    a.close()
    b = Solver.from_case(view, static_area=True)
    assert 0 < b.device_bytes() <= load_library().lw_static_area_bytes()
    sb = b.step(nsteps)
    fb = b.fields()
    b.close()
    assert sa["mg_cycles_sum"] == sb["mg_cycles_sum"] and sa["time"] == sb["time"]
    for k in "uvwpt":
        assert np.array_equal(fa[k], fb[k]), k


def test_gpu_static_area_refuses_a_second_handle_and_oversized_cases():
    view = _view((32, 32, 8), "nosgs", 2)
    print("static area refusals", flush=True)
    b = Solver.from_case(view, static_area=True)
    try:
        with pytest.raises(LwError) as e:
            Solver.from_case(view, static_area=True)
        assert "another handle" in str(e.value)
    finally:
        b.close()
    big = C.case_view(C.build_case(C.CaseSpec(256, 192, 32, settings=C.Settings(iles=1, idvmon=1, nlev=3))))
    with pytest.raises(LwError) as e:
        Solver.from_case(big, static_area=True)
    assert "static device area" in str(e.value)
    # the refused case released the area
    c = Solver.from_case(view, static_area=True)
    c.step(1)
    c.close()
