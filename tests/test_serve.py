"""Host side of the served form: body level sets of the library and the
oracle against the case builder's, the seeded case stream, the admission
estimate, and the request loop's error paths (no GPU needed)."""
import ctypes
import math
import os
import tempfile

import numpy as np
import pytest

from leswht_amd import case as C
from leswht_amd import serve
from leswht_amd import solver as S


def _points(n=4000, seed=7):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-2.0, 2.0, (n, 3))
    # points on and next to the axes and diagonals, where max/abs branches meet
    t = rng.uniform(-1.0, 1.0, 200)
    extra = np.concatenate([np.stack([t, 0 * t, t], 1), np.stack([t, t, t], 1), np.stack([t, -t, t], 1)])
    return np.concatenate([pts, extra])


BODY_CASES = [("cylinder", (0.5,)), ("cylinder", (0.37,)), ("square", (0.5, 0.0)), ("square", (0.5, 1.1)),
              ("ellipse", (0.5, 0.25, 0.0)), ("ellipse", (0.5, 0.2, 2.3)), ("ellipse", (0.3, 0.6, 0.7))]


@pytest.mark.parametrize("name,args", BODY_CASES)
def test_library_body_level_set_equals_case_builder(name, args):
    """The library evaluates the body on the host for its masks; its values
    must be the case builder's, bit for bit, or the masks and the IB tables
    would disagree at cells that sit on the surface."""
    L = S.load_library()
    b = S.body_struct(C.body_params(name, args))
    fn = C.BODIES[name]
    pts = _points()
    want = fn(pts[:, 0], pts[:, 1], pts[:, 2], *args)
    got = np.array([L.lw_body_eval(ctypes.byref(b), x, y, z) for x, y, z in pts])
    assert np.array_equal(want.view(np.int64), got.view(np.int64))


@pytest.mark.parametrize("name,args", BODY_CASES)
def test_oracle_body_level_set_equals_case_builder(name, args):
    import oracle
    from tests.helpers import spec_for
    spec = spec_for((16, 16, 4), "nosgs")
    spec = C.CaseSpec(16, 16, 4, xlim=spec.xlim, ylim=spec.ylim, zlen=1.0, dmin=0.05, body=name,
                      body_args=args, settings=spec.settings)
    o = oracle.Oracle(C.case_view(C.build_case(spec)))
    L = oracle.lib()
    pts = _points(1500)
    want = C.BODIES[name](pts[:, 0], pts[:, 1], pts[:, 2], *args)
    got = np.array([L.ora_funcbody(o.h, x, y, z) for x, y, z in pts])
    assert np.array_equal(want.view(np.int64), got.view(np.int64))


def test_body_shapes_match_closed_forms():
    """Closed forms: a square of half side h rotated by a quarter turn is the
    same square; an ellipse with equal semi-axes is a circle; the level sets
    are zero on the boundary points they name."""
    L = S.load_library()
    sq = S.body_struct(C.body_params("square", (0.5, 0.0)))
    el = S.body_struct(C.body_params("ellipse", (0.5, 0.2, 0.0)))
    ci = S.body_struct(C.body_params("ellipse", (0.4, 0.4, 0.3)))
    assert L.lw_body_eval(ctypes.byref(sq), 0.5, 0.2, 0.0) == 0.0
    assert L.lw_body_eval(ctypes.byref(sq), 0.7, 0.0, 0.0) == pytest.approx(0.2, abs=1e-15)
    assert L.lw_body_eval(ctypes.byref(el), 0.5, 0.0, 0.0) == 0.0
    assert L.lw_body_eval(ctypes.byref(el), 0.0, 0.2, 0.0) == 0.0
    assert L.lw_body_eval(ctypes.byref(el), 0.0, 0.0, 0.0) == pytest.approx(-0.2, abs=1e-15)
    for th in np.linspace(0, 2 * math.pi, 17):
        assert L.lw_body_eval(ctypes.byref(ci), 0.4 * math.cos(th), 0.4 * math.sin(th), 0.0) == \
            pytest.approx(0.0, abs=1e-15)


def test_bad_body_is_rejected():
    """lw_create refuses an unknown body kind or a non-positive size before
    it touches the device."""
    from tests.helpers import spec_for
    view = C.case_view(C.build_case(spec_for((16, 16, 4), "nosgs")))
    for body in ((7, (0.5,)), (0, (0.0,)), (2, (0.5, -1.0, 0.0))):
        with pytest.raises(S.LwError, match="body"):
            S.Solver.from_case(view, body=body)


def test_case_stream_is_seeded_and_in_range():
    a = serve.draw_cases(40, seed=11, rate=5.0)
    b = serve.draw_cases(40, seed=11, rate=5.0)
    assert [(c.spec, c.nsteps, c.arrival) for c in a] == [(c.spec, c.nsteps, c.arrival) for c in b]
    arr = [c.arrival for c in a]
    assert arr == sorted(arr) and arr[0] > 0
    for c in a:
        s = c.spec
        if c.large:
            assert (s.n1m, s.n2m, s.n3m) == (384, 256, 64) and s.settings.iles == 1
            continue
        assert s.n1m in range(64, 257, 32) and s.n2m in range(64, 193, 32) and s.n3m in (8, 16, 32)
        assert 40.0 <= s.settings.re <= 400.0 and s.settings.pr in (0.71, 7.0)
        assert s.n1m % (2 ** s.settings.nlev) == 0
        assert 200 <= c.nsteps <= 2000
        if s.body == "ellipse":
            assert 1.0 <= s.body_args[0] / s.body_args[1] <= 3.0


def test_probe_and_estimate():
    """lw_probe_dir reads the sizes of a case directory; the estimate grows
    with the grid and covers the field arrays it must hold."""
    from tests.helpers import spec_for
    L = S.load_library()
    with tempfile.TemporaryDirectory() as d:
        spec = spec_for((32, 24, 8), "les")
        sol = C.write_case(C.build_case(spec), d)
        st = S.lw_settings()
        nm = (ctypes.c_int64 * 3)()
        assert L.lw_probe_dir(sol.encode(), ctypes.byref(st), nm) == 0
        assert tuple(nm) == (32, 24, 8)
        assert st.iles == 1 and abs(st.re - 3900.0) < 1e-12
        assert L.lw_probe_dir(os.path.join(d, "nowhere").encode(), ctypes.byref(st), nm) == -3
    s = C.Settings(iles=1, ihtrans=1)
    e1 = S.estimate_device_bytes(s, 64, 64, 8)
    e2 = S.estimate_device_bytes(s, 128, 64, 8)
    assert e2 > e1 > 34 * 66 * 66 * 10 * 8
    assert S.estimate_device_bytes(C.Settings(iles=0), 64, 64, 8) < e1
    with pytest.raises(S.LwError):
        S.estimate_device_bytes(s, 1, 64, 8)


def test_server_error_paths_without_running_a_case():
    """Submission errors come back at once; a case that cannot be created is
    reported through its result, and the server still shuts down."""
    from tests.helpers import spec_for
    with tempfile.TemporaryDirectory() as d:
        sol = C.write_case(C.build_case(spec_for((16, 16, 4), "nosgs")), d)
        with pytest.raises(S.LwError):
            S.Server(0, 1)
        srv = S.Server(1 << 20, 2)  # 1 MiB: no case fits
        with pytest.raises(S.LwError, match="budget"):
            srv.submit(sol, 1)
        with pytest.raises(S.LwError):
            srv.submit(os.path.join(d, "missing"), 1)
        srv.close()
        # a case whose initial field is missing passes the probe (settings and
        # grid) and fails when the worker creates it, before any device work
        os.remove(os.path.join(d, "output", "field", "fld000000"))
        srv = S.Server(8 << 30, 2)
        t = srv.submit(sol, 0)
        r = srv.poll(t)
        assert r["status"] in (S.CASE_QUEUED, S.CASE_RUNNING, -3)
        with pytest.raises(S.LwError, match="fld|open"):
            srv.wait(t)
        info = srv.stats()
        assert info["submitted"] == 1 and info["slots"] == 2 and info["running"] == 0
        srv.close()


def test_serve_exec_reports_manifest_and_case_errors():
    """serve_exec (the manifest driver of the request loop): a malformed
    manifest line stops it with status 2; a case that cannot be read is
    reported and gives status 1 (no device work happens before that)."""
    import subprocess
    from leswht_amd.build import SERVE_EXE
    from tests.helpers import spec_for
    if not os.path.exists(SERVE_EXE):
        pytest.skip("serve_exec not built")
    with tempfile.TemporaryDirectory() as d:
        man = os.path.join(d, "manifest")
        with open(man, "w") as f:
            f.write("# comment\nsome/dir notanumber\n")
        res = subprocess.run([SERVE_EXE, man], capture_output=True, text=True, timeout=60)
        assert res.returncode == 2 and "expected SOLVER_DIR NSTEPS" in res.stderr
        sol = C.write_case(C.build_case(spec_for((16, 16, 4), "nosgs")), os.path.join(d, "case"))
        os.remove(os.path.join(d, "case", "output", "field", "fld000000"))
        with open(man, "w") as f:
            f.write("%s 2\n%s 1 1 0.5 0.3\n" % (sol, os.path.join(d, "missing")))
        res = subprocess.run([SERVE_EXE, man, "--slots", "2"], capture_output=True, text=True, timeout=60)
        assert res.returncode == 1
        assert "failed" in res.stderr or "cannot open" in res.stderr
