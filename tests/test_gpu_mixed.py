"""The mixed-precision form (option "mg_fp32": the multigrid level planes
below the top level in FP32, the top level's relaxation, residual,
convergence test and solution in FP64; option "sgs_fp32": the dynamic SGS
models' test filter in FP32, the Germano sums in FP64) against the FP64
oracle: the error of each field after a few steps, which the form trades for
its speed. The bar is the one the configuration states for the mixed form,
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


@pytest.mark.parametrize("opts", [("sgs_fp32",), ("mg_fp32", "sgs_fp32")])
def test_gpu_fp32_test_filter_against_fp64_oracle(opts):
    """The FP32 test filter alone and with the FP32 multigrid levels (the
    configuration's whole mixed form), 96x256x64 LES, 2 steps: every field
    within 1e-6 of the FP64 oracle. The model coefficients move by the
    filter's rounding (one that is not clipped to zero differs from the FP64
    path's, so the FP32 filter did run) but by less than 1e-4 relative."""
    import oracle
    view = C.case_view(C.build_case(spec_for((96, 256, 64), "les", 2, 3)))
    print("fp32 filter", opts, flush=True)
    o = oracle.Oracle(view)
    do = o.step(2)
    g64 = Solver.from_case(view)
    s64 = g64.step(2)
    g64.close()
    g = Solver.from_case(view)
    for name in opts:
        g.set_option(name, 1)
    dg = g.step(2)
    fo, fg = o.fields(), g.fields()
    errs = {k: rel(fo[k], fg[k]) for k in "uvwpt"}
    print("fp32 filter", opts, "rel L2 vs FP64 oracle:", {k: "%.2e" % v for k, v in errs.items()},
          "csgsts %.12e (fp64 %.12e) csgshf %.12e (fp64 %.12e)" % (dg["csgsts"], s64["csgsts"], dg["csgshf"],
                                                                 s64["csgshf"]), flush=True)
    for k, v in errs.items():
        assert v <= BAR, (k, v)
    moved = 0
    for c in ("csgsts", "csgshf"):
        # a coefficient the clip at zero holds (C = max(0, ...)) stays zero
        assert abs(dg[c] - s64[c]) <= 1e-4 * abs(s64[c]), (c, dg[c], s64[c])
        moved += dg[c] != s64[c]
    assert moved >= 1
    assert abs(dg["mg_cycles_sum"] - do.mg_cycles_sum) <= 0.1 * do.mg_cycles_sum + 3


def test_gpu_fp32_test_filter_needs_the_fused_models(monkeypatch):
    monkeypatch.setenv("LW_SGS_FUSED", "0")
    view = C.case_view(C.build_case(spec_for((32, 32, 8), "les", 1, 2)))
    g = Solver.from_case(view)
    with pytest.raises(LwError):
        g.set_option("sgs_fp32", 1)
    g.close()


def test_gpu_mixed_precision_full_c2_step():
    """One step of the 384x256x64 LES in the mixed form (FP32 multigrid
    levels and test filter) against the FP64 GPU path (which matches the
    oracle within 1e-10 on this step)."""
    view = C.case_view(C.build_case(c2_spec(1)))
    print("mixed c2", flush=True)
    out = []
    for f32 in (0, 1):
        g = Solver.from_case(view)
        g.set_option("mg_fp32", f32)
        g.set_option("sgs_fp32", f32)
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
