"""Batched stepping on the GPU: cases of different sizes, bodies, physics
and Poisson paths stepped together (every kernel launched once for all the
cases that issue it) give the same bits, dt and cycle counts as each case
stepped alone."""
import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd.solver import Batch, Solver
from tests.helpers import spec_for

pytestmark = pytest.mark.gpu


def _spec(dims, body, args, re=150.0, pr=0.71, iles=0, seed=5, nlev=2, **kw):
    s = C.Settings(re=re, pr=pr, iles=iles, idvmon=1 if iles else 0, idtopt=1, cflfac=0.5,
                   nlev=nlev, resid1=1e-8, **kw)
    return C.CaseSpec(*dims, xlim=(-3.0, 6.0), ylim=(-3.0, 3.0), zlen=1.0, dmin=0.05, noise=1e-2,
                      body=body, body_args=args, seed=seed, settings=s)


# n2m 64 (1 row per lane, full), 96 (2, partial), 128 (2, full), 160 (4,
# partial); n3m 8, 16, 32; NLEV 2 and 3; LES and laminar; three bodies;
# line-SOR above IMGSOR; a channel with poiss_ftft
MIX = [
    ("a", _spec((64, 64, 8), "cylinder", (0.5,), re=60.0)),
    ("b", _spec((96, 96, 16), "square", (0.5, 0.7), re=250.0, pr=7.0, nlev=3)),
    ("c", _spec((64, 160, 8), "ellipse", (0.5, 0.2, 2.1), re=120.0)),
    ("d", _spec((64, 64, 16), "square", (0.5, 0.3), re=3900.0, iles=1, seed=9)),
    ("e", _spec((32, 128, 32), "ellipse", (0.5, 0.3, 0.4), re=400.0, pr=7.0)),
    ("f", _spec((64, 64, 16), "cylinder", (0.5,), re=200.0, imgsor=5, wwsor=1.5)),
    ("g", spec_for((32, 32, 8), "chan_cfr")),
]


def _solver(spec):
    return Solver.from_case(C.case_view(C.build_case(spec)))


def _run_alone(nsteps):
    out = []
    for name, spec in MIX:
        print("alone", name, spec.n1m, spec.n2m, spec.n3m, flush=True)
        g = _solver(spec)
        st = g.step(nsteps)
        out.append((st, g.fields()))
        g.close()
    return out


def test_gpu_batched_cases_equal_alone():
    nsteps = 4
    alone = _run_alone(nsteps)
    solvers = [_solver(spec) for _, spec in MIX]
    b = Batch(solvers)
    print("batch of", len(MIX), flush=True)
    st = b.step(2)
    st = b.step(nsteps - 2)
    info = b.info()
    b.close()
    print("batch info", info, flush=True)
    # launches were shared: fewer batched launches than per-case launches
    # (the multigrid kernels of different line lengths are different
    # instances, so this mix shares mostly the other kernels)
    assert info["case_launches"] > 1.2 * info["launches"] > 0
    for (name, _), (sa, fa), sb, g in zip(MIX, alone, st, solvers):
        assert sa["time"] == sb["time"] and sa["dt"] == sb["dt"], name
        assert sa["mg_cycles_sum"] == sb["mg_cycles_sum"], name
        assert sa["mg_cycles_max"] == sb["mg_cycles_max"], name
        assert sa["dvmax"] == sb["dvmax"] and sa["cflmax"] == sb["cflmax"], name
        fb = g.fields()
        for k in fa:
            if fa[k] is None:
                continue
            assert np.array_equal(fa[k], fb[k]), (name, k)
    # the handles stay usable on their own after the batch: one more step
    # each, alone, equals one more step of a fresh alone run
    g = solvers[1]
    g.step(1)
    ref = _solver(MIX[1][1])
    ref.step(nsteps + 1)
    fr, fg = ref.fields(), g.fields()
    for k in ("u", "v", "w", "p", "t"):
        assert np.array_equal(fr[k], fg[k]), k
    ref.close()
    for g in solvers:
        g.close()


def test_gpu_batch_refuses_what_it_cannot_step():
    a = _solver(MIX[0][1])
    view = C.case_view(C.build_case(spec_for((32, 32, 8), "les")))
    view["settings"].poisson = 1
    m = Solver.from_case(view)
    with pytest.raises(Exception, match="Poisson"):
        Batch([a, m])
    with pytest.raises(Exception, match="twice"):
        Batch([a, a])
    b = Batch([a])
    with pytest.raises(Exception, match="belongs to a batch"):
        Batch([a])
    with pytest.raises(Exception, match="batch"):
        a.close()
    b.close()
    a.close()
    m.close()
