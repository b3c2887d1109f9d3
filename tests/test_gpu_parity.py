"""Parity of the HIP time step (through the C ABI) with the reference's
golden outputs and with the oracle. Bar: relative L2 <= 1e-10 on u, v, w, p,
T (FP64 path), identical multigrid cycle counts."""
import os
import tempfile

import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import Solver
from tests.helpers import fld_regions, golden, golden_names, rel, spec_for

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _view(dims, mode, nend=1, nlev=None):
    return C.case_view(C.build_case(spec_for(dims, mode, nend, nlev)))


@pytest.mark.parametrize("name", golden_names())
def test_gpu_matches_reference_golden(name):
    z, meta = golden(name)
    view = _view(tuple(meta["dims"]), meta["mode"], meta["nend"])
    print("golden", name, flush=True)
    g = Solver.from_case(view)
    st = g.step(meta["nend"])
    f = g.fields()
    n = (g.n1, g.n2, g.n3)
    assert abs(st["time"] - float(z["time"])) <= 1e-14 * max(1.0, float(z["time"]))
    for k, sl in fld_regions(n).items():
        err = rel(z[k][sl], f[k][sl])
        assert err <= TOL, (name, k, err)


@pytest.mark.parametrize("mode", ["nosgs", "les"])
def test_gpu_phases_match_oracle(mode, oracle_lib):
    import oracle
    view = _view((32, 32, 8), mode)
    print("phases", mode, flush=True)
    o = oracle.Oracle(view)
    g = Solver.from_case(view)
    L, h = oracle_lib, o.h
    L.ora_stepinit(h); g.phase(0)
    for ms in (1, 2, 3):
        L.ora_substep_begin(h, ms); g.phase(10 + ms)
        L.ora_phase_sgs(h); g.phase(1)
        if ms == 1:
            for nm in ("nusgs", "nusgs1_1", "nusgs1_2", "nusgs1_3", "alsgs", "alsgs1_1", "alsgs1_2", "alsgs1_3"):
                assert rel(o.array(nm), g.array(nm)) <= 1e-12, nm
        L.ora_phase_rhsnlhs(h); g.phase(2)
        fo, fg = o.fields(), g.fields()
        for k in "uvw":
            assert rel(fo[k], fg[k]) <= 1e-12, ("rhsnlhs", ms, k)
        L.ora_phase_divgs(h); g.phase(3)
        assert rel(o.array("divgsum"), g.array("divgsum")) <= 1e-12
        q = g.array("qmass")[1:g.n1, 1:g.n2, 1:g.n3]
        assert rel(o.array("qmass"), q) <= 1e-12
        L.ora_phase_poisson(h); g.phase(4)
        assert rel(o.array("phi"), g.array("phi")) <= 1e-11
        L.ora_phase_ucalc_pcalc(h); g.phase(5)
        L.ora_phase_energy(h); g.phase(6)
        fo, fg = o.fields(), g.fields()
        for k in "uvwpt":
            assert rel(fo[k], fg[k]) <= 1e-11, ("substep", ms, k)


@pytest.mark.parametrize("dims,mode,nsteps,nlev", [((64, 32, 16), "les", 3, 2), ((48, 32, 16), "les_cfl", 4, 3),
                                                   ((96, 64, 32), "les", 2, 3)])
def test_gpu_steps_match_oracle(dims, mode, nsteps, nlev):
    import oracle
    view = _view(dims, mode, nsteps, nlev)
    print("steps", dims, mode, flush=True)
    o = oracle.Oracle(view)
    g = Solver.from_case(view)
    do = o.step(nsteps)
    dg = g.step(nsteps)
    assert dg["mg_cycles_max"] == do.mg_cycles_max
    assert dg["dt"] == pytest.approx(do.dt, rel=1e-13)
    fo, fg = o.fields(), g.fields()
    for k in "uvwpt":
        assert rel(fo[k], fg[k]) <= TOL, k


def test_gpu_full_c1_size_matches_oracle():
    """192x128x16 (the plain FP64 configuration's grid), one step."""
    import oracle
    s = C.Settings(re=100.0, iles=1, idvmon=0, csgsts=0.0, csgshf=0.0, resid1=1e-6, nlev=5, dt_size=0.005, nend=1)
    view = C.case_view(C.build_case(C.CaseSpec(192, 128, 16, dmin=0.01, noise=1e-3, settings=s)))
    print("c1 full size", flush=True)
    o = oracle.Oracle(view)
    g = Solver.from_case(view)
    do = o.step(1)
    dg = g.step(1)
    assert dg["mg_cycles_max"] == do.mg_cycles_max
    fo, fg = o.fields(), g.fields()
    for k in "uvwpt":
        assert rel(fo[k], fg[k]) <= TOL, k


def test_gpu_iles0_equals_zero_coefficient_les():
    """ILES=0 (no SGS) computes what the reference computes with ILES=1 and
    zero model constants (its ILES=0 path dereferences unallocated arrays)."""
    v0 = _view((32, 32, 8), "nosgs", 2)
    v1 = _view((32, 32, 8), "nosgs", 2)
    v0["settings"].iles = 0
    a = Solver.from_case(v0); b = Solver.from_case(v1)
    a.step(2); b.step(2)
    fa, fb = a.fields(), b.fields()
    for k in "uvwpt":
        assert np.array_equal(fa[k], fb[k]), k


def test_gpu_from_dir_is_the_same_case():
    """The C++ reader of the case directory (drop-in path) gives the state the
    Python host builds: identical results after one step."""
    spec = spec_for((32, 32, 8), "les", 1)
    case = C.build_case(spec)
    root = tempfile.mkdtemp()
    sol_dir = C.write_case(case, root)
    a = Solver.from_dir(sol_dir)
    b = Solver.from_case(C.case_view(case))
    a.step(1); b.step(1)
    fa, fb = a.fields(), b.fields()
    for k in "uvwpt":
        assert np.array_equal(fa[k], fb[k]), k
    out = os.path.join(root, "fld_out")
    a.write_field(out)
    r = C.read_fld(out)
    n = (a.n1, a.n2, a.n3)
    for k, sl in fld_regions(n).items():
        assert np.array_equal(r[k][sl], fa[k][sl]), k
    assert r["prdic"] == (0, 0, 1) and r["bcs"][:2] == (2, 3)


def test_gpu_rejects_out_of_scope_configuration():
    view = _view((32, 32, 8), "nosgs")
    view["settings"].ich = 1
    with pytest.raises(Exception) as e:
        Solver.from_case(view)
    assert "ICH" in str(e.value)
