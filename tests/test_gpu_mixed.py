"""The mixed-precision form of the Poisson multigrid (option "mg_fp32": the
level planes below the top level in FP32, the top level's relaxation,
residual, convergence test and solution in FP64) against the FP64 oracle:
the error of each field after a few steps, which the form trades for its
speed. The bar is the one the configuration states for the mixed form,
relative L2 <= 1e-6 against FP64; the errors are printed."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import LwError, Solver
from tests.helpers import c2_spec, rel, spec_for

pytestmark = pytest.mark.gpu
BAR = 1e-6


@pytest.mark.parametrize("dims,nlev,nsteps", [((96, 256, 64), 3, 2), ((192, 64, 16), 4, 5)])
def test_gpu_mixed_precision_against_fp64_oracle(dims, nlev, nsteps):
    import oracle
    view = C.case_view(C.build_case(spec_for(dims, "les", nsteps, nlev)))
    print("mixed", dims, flush=True)
    o = oracle.Oracle(view)
    g = Solver.from_case(view)
    g.set_option("mg_fp32", 1)
    do = o.step(nsteps)
    dg = g.step(nsteps)
    fo, fg = o.fields(), g.fields()
    errs = {k: rel(fo[k], fg[k]) for k in "uvwpt"}
    print("mixed", dims, "rel L2 vs FP64 oracle:", {k: "%.2e" % v for k, v in errs.items()},
          "cycles", dg["mg_cycles_sum"], "oracle", do.mg_cycles_sum, flush=True)
    for k, v in errs.items():
        assert v <= BAR, (k, v)
    assert abs(dg["mg_cycles_sum"] - do.mg_cycles_sum) <= 0.1 * do.mg_cycles_sum + 3


def test_gpu_mixed_precision_full_c2_step():
    """One step of the 384x256x64 LES in the mixed form against the FP64 GPU
    path (which matches the oracle within 1e-10 on this step)."""
    view = C.case_view(C.build_case(c2_spec(1)))
    print("mixed c2", flush=True)
    out = []
    for f32 in (0, 1):
        g = Solver.from_case(view)
        g.set_option("mg_fp32", f32)
        st = g.step(1)
        out.append((st, g.fields()))
        g.close()
    errs = {k: rel(out[0][1][k], out[1][1][k]) for k in "uvwpt"}
    print("mixed c2 rel L2 vs FP64:", {k: "%.2e" % v for k, v in errs.items()},
          "cycles", out[1][0]["mg_cycles_sum"], "fp64", out[0][0]["mg_cycles_sum"], flush=True)
    for k, v in errs.items():
        assert v <= BAR, (k, v)


def test_gpu_mixed_precision_rejected_where_unsupported():
    """The form needs the fused level kernels with the coarse levels in LDS
    (or no coarse levels) and the z-DFT solver: elsewhere the option is an
    error, and the FP64 path stays in place."""
    view = C.case_view(C.build_case(spec_for((64, 64, 16), "les", 1, 2)))
    g = Solver.from_case(view)
    with pytest.raises(LwError):
        g.set_option("mg_fp32", 1)
    g.close()
    view = C.case_view(C.build_case(spec_for((32, 32, 8), "les", 1, 2)))
    view["settings"].poisson = 1
    g = Solver.from_case(view)
    with pytest.raises(LwError):
        g.set_option("mg_fp32", 1)
    g.close()
