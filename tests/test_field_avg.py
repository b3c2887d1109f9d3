"""Time-averaged statistics (field_avg, misc_hist.f90:287-697).

CPU: the oracle's running sums against the fav files the reference solver
wrote (tests/golden/fav/cyl16_les_5.npz, made by make_fav_golden.py), and the
combined file as the sum of the interval files. GPU (marked): the library's
sums against the oracle, and the fav files it writes (intervals, manifest,
finalize) against the reference's."""
import dataclasses
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from leswht_amd import case as C
from tests.golden.make_fav_golden import fav_arrays, fav_spec
from tests.helpers import GOLDEN, rel

NAMES = ("uavg", "vavg", "wavg") + tuple("uiujavg_%d" % q for q in range(1, 7)) + ("pavg", "p2avg") + \
    tuple("voravg_%d" % q for q in range(1, 4)) + tuple("vor2avg_%d" % q for q in range(1, 7)) + \
    ("ssavg", "tavg", "t2avg")
# arrays per record of a fav file after the two header records
PER = (1, 1, 1, 6, 1, 1, 3, 6, 1, 1, 1)
TOL = 1e-10


def golden_fav():
    z = np.load(os.path.join(GOLDEN, "fav", "cyl16_les_5.npz"), allow_pickle=False)
    return z, json.loads(str(z["manifest"]))


def records_to_arrays(recs):
    """Splits the records of a fav file into the 23 named arrays."""
    out, l = {}, 0
    for r, per in zip(recs, PER):
        parts = np.split(np.asarray(r), per)
        for q in range(per):
            out[NAMES[l]] = parts[q]
            l += 1
    return out


def golden_arrays(z, tag):
    return records_to_arrays([z["%s_r%02d" % (tag, q)] for q in range(int(z[tag + "_nrec"]))])


def _oracle(npriavg, steps):
    import oracle
    spec = fav_spec()
    spec = dataclasses.replace(spec, settings=dataclasses.replace(spec.settings, npriavg=npriavg))
    o = oracle.Oracle(C.case_view(C.build_case(spec)))
    for _ in range(steps):
        o.step(1)
    return o


def _oracle_sums(o):
    out = {}
    for nm in NAMES:
        out[nm] = o.array(nm).ravel(order="F") if o.array(nm).ndim > 1 else o.array(nm)
    return out


def test_golden_manifest_and_headers():
    z, manifest = golden_fav()
    assert manifest == ["fav0000001-0000002", "fav0000002-0000004", "fav0000004-0000005"]
    for tag in ("first", "last", "combined"):
        assert list(z[tag + "_n"]) == [16, 16, 4] and float(z[tag + "_re"]) == 3900.0
        assert int(z[tag + "_nrec"]) == 11
    # combined: from the first interval's start to the last one's end
    assert z["combined_times"][0] == z["first_times"][0] == 0.0
    assert z["combined_times"][1] == z["last_times"][1]
    assert list(z["combined_times"][2:]) == [1.0, 5.0]


def test_oracle_first_interval_matches_reference():
    """Steps 1-2 accumulated by the oracle (no restart) equal the reference's
    first fav file."""
    z, _ = golden_fav()
    want = golden_arrays(z, "first")
    got = _oracle_sums(_oracle(1000, 2))
    for nm in NAMES:
        assert rel(want[nm], got[nm]) <= TOL, nm


def test_oracle_restarts_every_npriavg_steps():
    """With NPRIAVG = 2 the sums restart after steps 2 and 4, so after step 5
    they are the reference's last (partial) file."""
    z, _ = golden_fav()
    want = golden_arrays(z, "last")
    got = _oracle_sums(_oracle(2, 5))
    for nm in NAMES:
        assert rel(want[nm], got[nm]) <= TOL, nm


def test_combined_is_the_sum_of_the_intervals():
    """The reference's combined file equals the oracle's uninterrupted sum
    over the five steps (the interval sums only regroup the additions)."""
    z, _ = golden_fav()
    want = golden_arrays(z, "combined")
    got = _oracle_sums(_oracle(1000, 5))
    for nm in NAMES:
        assert rel(want[nm], got[nm]) <= TOL, nm
    # and a plausible mistake fails: the first interval alone is far off
    first = golden_arrays(z, "first")
    assert rel(want["uavg"], first["uavg"]) > 0.1


def _fav_case(root, npriavg=2):
    spec = fav_spec()
    spec = dataclasses.replace(spec, settings=dataclasses.replace(spec.settings, npriavg=npriavg))
    return C.write_case(C.build_case(spec), root)


def _compare_fav(path, z, tag):
    a = fav_arrays(path)
    assert list(a["n"]) == list(z[tag + "_n"]) and a["re"] == float(z[tag + "_re"])
    assert np.allclose(a["times"], z[tag + "_times"], rtol=1e-13, atol=1e-15)
    assert len(a["recs"]) == int(z[tag + "_nrec"])
    got, want = records_to_arrays(a["recs"]), golden_arrays(z, tag)
    for nm in NAMES:
        assert rel(want[nm], got[nm]) <= TOL, (tag, nm)


@pytest.mark.gpu
def test_gpu_sums_match_oracle():
    from leswht_amd.solver import Solver
    import oracle
    spec = fav_spec()
    spec = dataclasses.replace(spec, settings=dataclasses.replace(spec.settings, npriavg=1000))
    view = C.case_view(C.build_case(spec))
    print("field_avg sums", flush=True)
    g = Solver.from_case(view)
    o = oracle.Oracle(view)
    g.step(3)
    for _ in range(3):
        o.step(1)
    for nm in NAMES:
        assert rel(o.array(nm), g.avg_array(nm)) <= 1e-12, nm


@pytest.mark.gpu
def test_gpu_fav_files_match_reference():
    """The library run from a case directory writes the reference's fav
    files: the same names in the manifest, headers and records within 1e-10,
    and the combined file from lw_avg_finalize."""
    from leswht_amd.solver import Solver
    z, manifest = golden_fav()
    with tempfile.TemporaryDirectory() as root:
        sol = _fav_case(root)
        print("fav files", flush=True)
        g = Solver.from_dir(sol)
        g.step(5)
        g.avg_finalize()
        g.close()
        d = os.path.join(root, "output", "field_avg")
        got = [ln.strip() for ln in open(os.path.join(d, "fav_manifest.dat")) if ln.strip()]
        assert got == manifest
        _compare_fav(os.path.join(d, manifest[0]), z, "first")
        _compare_fav(os.path.join(d, manifest[-1]), z, "last")
        _compare_fav(os.path.join(d, "combined", "fav0000001-0000005"), z, "combined")


@pytest.mark.gpu
def test_gpu_fortran_host_writes_fav_files():
    from leswht_amd.build import EXE
    if not os.path.exists(EXE):
        pytest.skip("solver_exec not built")
    z, manifest = golden_fav()
    with tempfile.TemporaryDirectory() as root:
        sol = _fav_case(root)
        print("solver_exec with IAVG = 1", flush=True)
        res = subprocess.run([EXE], cwd=sol, capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
        d = os.path.join(root, "output", "field_avg")
        assert [ln.strip() for ln in open(os.path.join(d, "fav_manifest.dat")) if ln.strip()] == manifest
        _compare_fav(os.path.join(d, "combined", "fav0000001-0000005"), z, "combined")
