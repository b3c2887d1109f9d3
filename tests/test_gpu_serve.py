"""The request loop on the GPU: cases of different sizes, bodies and physics
give bit-identical results alone and in the mix; bodies other than the
cylinder match the oracle; admission keeps to the memory budget."""
import os
import tempfile

import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd import serve
from leswht_amd.solver import Solver, estimate_device_bytes
from tests.helpers import rel

pytestmark = pytest.mark.gpu


def _spec(dims, body, args, re=150.0, pr=0.71, iles=0, seed=5, nlev=2):
    s = C.Settings(re=re, pr=pr, iles=iles, idvmon=1 if iles else 0, idtopt=1, cflfac=0.5,
                   nlev=nlev, resid1=1e-8)
    return C.CaseSpec(*dims, xlim=(-3.0, 6.0), ylim=(-3.0, 3.0), zlen=1.0, dmin=0.05, noise=1e-2,
                      body=body, body_args=args, seed=seed, settings=s)


MIX = [
    serve.ServedCase("a", _spec((64, 64, 8), "cylinder", (0.5,), re=60.0), 12, 0.0),
    serve.ServedCase("b", _spec((96, 64, 16), "square", (0.5, 0.7), re=250.0, pr=7.0), 10, 0.0),
    serve.ServedCase("c", _spec((64, 96, 8), "ellipse", (0.5, 0.2, 2.1), re=120.0), 14, 0.0),
    serve.ServedCase("d", _spec((64, 64, 16), "square", (0.5, 0.3), re=3900.0, iles=1, seed=9), 8, 0.0),
    serve.ServedCase("e", _spec((32, 32, 8), "ellipse", (0.5, 0.3, 0.4), re=400.0, pr=7.0), 16, 0.0),
]


def _fields(solver_dir):
    return C.read_fld(serve.fld_path(solver_dir), heat=True)


def test_gpu_served_cases_equal_alone_and_in_the_mix():
    with tempfile.TemporaryDirectory() as root:
        alone_dirs = serve.prepare(MIX, os.path.join(root, "alone"))
        mix_dirs = serve.prepare(MIX, os.path.join(root, "mix"))
        alone = []
        for c, d in zip(MIX, alone_dirs):
            print("alone", c.name, c.spec.n1m, c.spec.n2m, c.spec.n3m, c.spec.body, flush=True)
            alone.append(serve.run_alone(c, d))
        print("mix of", len(MIX), flush=True)
        out = serve.run_stream(MIX, mix_dirs, budget_bytes=64 << 30, slots=len(MIX), honour_arrivals=False)
        assert not out["failed"], out["failed"]
        assert out["info"]["peak_running"] >= 2  # the cases did overlap
        for c, a, m, da, dm in zip(MIX, alone, out["results"], alone_dirs, mix_dirs):
            assert a["steps_done"] == m["steps_done"] == c.nsteps
            assert a["time"] == m["time"] and a["dt"] == m["dt"], c.name
            assert a["mg_cycles_sum"] == m["mg_cycles_sum"], c.name
            assert a["reserved_bytes"] >= a["device_bytes"] > 0
            fa, fm = _fields(da), _fields(dm)
            for k in ("u", "v", "w", "p", "t"):
                assert np.array_equal(fa[k], fm[k]), (c.name, k)
        # continuous batching: the cases of at most 70000 cells join the
        # server's batch as they arrive (staggered) and leave it as they
        # finish; the largest runs on its own streams. Same bits as alone.
        import dataclasses
        staggered = [dataclasses.replace(c, arrival=0.05 * q) for q, c in enumerate(MIX)]
        bat_dirs = serve.prepare(staggered, os.path.join(root, "batched"))
        print("batched mix of", len(MIX), flush=True)
        bout = serve.run_stream(staggered, bat_dirs, budget_bytes=64 << 30, slots=2, honour_arrivals=True,
                                batch_max_cells=70000)
        assert not bout["failed"], bout["failed"]
        for c, a, m, da, dm in zip(MIX, alone, bout["results"], alone_dirs, bat_dirs):
            assert a["steps_done"] == m["steps_done"] == c.nsteps, c.name
            assert a["time"] == m["time"] and a["dt"] == m["dt"], c.name
            assert a["mg_cycles_sum"] == m["mg_cycles_sum"], c.name
            fa, fm = _fields(da), _fields(dm)
            for k in ("u", "v", "w", "p", "t"):
                assert np.array_equal(fa[k], fm[k]), (c.name, k)
        # the served path is the single-handle path: same bits as a Solver
        # created from the same directory and stepped directly
        c, d = MIX[1], alone_dirs[1]
        g = Solver.from_dir(d, body=c.body)
        g.step(c.nsteps)
        f = g.fields()
        g.close()
        fa = _fields(d)
        n1, n2, n3 = g.n1, g.n2, g.n3
        assert np.array_equal(fa["u"][1:n1 + 1], f["u"][1:n1 + 1])
        assert np.array_equal(fa["p"][1:n1, 1:n2, 1:n3], f["p"][1:n1, 1:n2, 1:n3])
        assert np.array_equal(fa["t"][1:n1, 1:n2, 1:n3], f["t"][1:n1, 1:n2, 1:n3])


def test_gpu_cases_sharing_a_sweep_instance_equal_alone_and_in_the_mix():
    """A 256-row and a 192-row case run the same register-resident y-sweep
    instance (8 rows per lane), whose LDS limit is a property of the kernel:
    creating the smaller case while the larger one runs must not refuse the
    larger one's sweeps. Both give the bits they give alone."""
    pair = [serve.ServedCase("big", _spec((32, 256, 8), "cylinder", (0.5,), re=300.0), 6, 0.0),
            serve.ServedCase("small", _spec((32, 192, 8), "ellipse", (0.5, 0.3, 0.8), re=200.0), 40, 0.0)]
    with tempfile.TemporaryDirectory() as root:
        alone_dirs = serve.prepare(pair, os.path.join(root, "alone"))
        mix_dirs = serve.prepare(pair, os.path.join(root, "mix"))
        alone = []
        for c, d in zip(pair, alone_dirs):
            print("alone", c.name, flush=True)
            alone.append(serve.run_alone(c, d))
        print("mix", flush=True)
        out = serve.run_stream(pair, mix_dirs, budget_bytes=64 << 30, slots=2, honour_arrivals=False)
        assert not out["failed"], out["failed"]
        for c, a, m, da, dm in zip(pair, alone, out["results"], alone_dirs, mix_dirs):
            assert a["status"] == m["status"] == 0, c.name
            assert a["steps_done"] == m["steps_done"] == c.nsteps
            assert a["mg_cycles_sum"] == m["mg_cycles_sum"], c.name
            fa, fm = _fields(da), _fields(dm)
            for k in ("u", "v", "w", "p", "t"):
                assert np.array_equal(fa[k], fm[k]), (c.name, k)


@pytest.mark.parametrize("body,args,iles", [("square", (0.5, 0.6), 1), ("ellipse", (0.5, 0.25, 1.2), 0)])
def test_gpu_other_bodies_match_oracle(body, args, iles):
    import oracle
    spec = _spec((32, 32, 8), body, args, re=3900.0 if iles else 150.0, iles=iles)
    view = C.case_view(C.build_case(spec))
    print("body", body, args, flush=True)
    g = Solver.from_case(view)
    o = oracle.Oracle(view)
    sg, so = g.step(2), o.step(2)
    fg, fo = g.fields(), o.fields()
    for k in ("u", "v", "w", "p", "t"):
        assert rel(fo[k], fg[k]) <= 1e-10, (body, k)
    assert sg["mg_cycles_max"] == so.mg_cycles_max


def test_gpu_admission_keeps_to_the_budget():
    cases = [serve.ServedCase("m%d" % q, _spec((64, 64, 8), "cylinder", (0.5,), seed=20 + q), 6, 0.0)
             for q in range(5)]
    est = estimate_device_bytes(cases[0].spec.settings, 64, 64, 8)
    budget = 2 * est + est // 2  # two cases fit, three do not
    with tempfile.TemporaryDirectory() as root:
        dirs = serve.prepare(cases, root)
        print("budget", budget, flush=True)
        out = serve.run_stream(cases, dirs, budget_bytes=budget, slots=4, honour_arrivals=False,
                               write_fields=False)
    info = out["info"]
    assert info["peak_reserved_bytes"] <= budget
    assert info["peak_running"] <= 2
    for r in out["results"]:
        assert r["status"] == 0 and r["steps_done"] == 6
        assert r["reserved_bytes"] == est >= r["device_bytes"]
    # first-come first-served: no case is admitted before an earlier one
    adm = [r["t_admit"] for r in out["results"]]
    assert adm == sorted(adm)


def test_gpu_serve_exec_runs_a_manifest():
    """serve_exec runs a manifest of case directories through the request
    loop and writes each case's final field, the same bits as a Solver run of
    the same directory."""
    import subprocess
    from leswht_amd.build import SERVE_EXE
    if not os.path.exists(SERVE_EXE):
        pytest.skip("serve_exec not built")
    cases = MIX[1:3]
    with tempfile.TemporaryDirectory() as root:
        dirs = serve.prepare(cases, root)
        man = os.path.join(root, "manifest")
        with open(man, "w") as f:
            for c, d in zip(cases, dirs):
                kind, p = c.body
                f.write("%s %d %d %s\n" % (d, c.nsteps, kind, " ".join(repr(v) for v in p)))
        print("serve_exec", len(cases), flush=True)
        res = subprocess.run([SERVE_EXE, man, "--slots", "2"], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        assert len([ln for ln in res.stdout.splitlines() if not ln.startswith("#")]) == len(cases)
        for c, d in zip(cases, dirs):
            got = C.read_fld(os.path.join(d, "..", "output", "field", "fld%07d" % c.nsteps), heat=True)
            g = Solver.from_dir(d, body=c.body)
            g.step(c.nsteps)
            f = g.fields()
            g.close()
            n1, n2, n3 = g.n1, g.n2, g.n3
            assert np.array_equal(got["u"][1:n1 + 1], f["u"][1:n1 + 1]), c.name
            assert np.array_equal(got["p"][1:n1, 1:n2, 1:n3], f["p"][1:n1, 1:n2, 1:n3]), c.name
            assert np.array_equal(got["t"][1:n1, 1:n2, 1:n3], f["t"][1:n1, 1:n2, 1:n3]), c.name
