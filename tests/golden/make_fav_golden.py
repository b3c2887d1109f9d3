"""Writes the golden fixture of the time-averaged statistics (field_avg,
misc_hist.f90:287-697) by running the REFERENCE solver (oracle/_ref/solver_exec,
compiled from the reference's own sources by oracle/Makefile) on a tiny
heated-cylinder LES case with IAVG = 1: AVG_TST = 0, NPRIAVG = 2, NEND = 5.
The reference then writes fav0000001-0000002 and fav0000003-0000004 during
the run and, in field_avg_finalize, fav0000005-0000005 and the sum of the
three under combined/. The fixture keeps the first interval's file, the last
(partial) one and the combined one, record by record, and the manifest.

Run in the build container (the reference sources are not on the GPU box):
    make -C oracle ref && python tests/golden/make_fav_golden.py
"""
import json
import os
import shutil
import struct
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from leswht_amd import case as C  # noqa: E402
import oracle  # noqa: E402

DIMS = (16, 16, 4)


def fav_spec():
    s = C.Settings(nend=5, iles=1, idvmon=1, resid1=1e-8, nlev=2, re=3900.0, iavg=1, avg_tst=0.0, npriavg=2)
    return C.CaseSpec(*DIMS, xlim=(-3.0, 6.0), ylim=(-3.0, 3.0), zlen=1.0, dmin=0.05, noise=1e-2, settings=s)


def read_records(path):
    data = open(path, "rb").read()
    recs, pos = [], 0
    while pos < len(data):
        (ln,) = struct.unpack_from("<i", data, pos)
        recs.append(data[pos + 4:pos + 4 + ln])
        pos += 8 + ln
    return recs


def fav_arrays(path):
    """header ints, header reals and the arrays of a fav file, as numpy."""
    recs = read_records(path)
    n1m, n2m, n3m = struct.unpack("<3q", recs[0][:24])
    (re,) = struct.unpack("<d", recs[0][24:32])
    t = np.frombuffer(recs[1], "<f8").copy()
    body = [np.frombuffer(r, "<f8").copy() for r in recs[2:]]
    return dict(n=np.array([n1m, n2m, n3m]), re=re, times=t, recs=body)


def main():
    root = tempfile.mkdtemp(prefix="lw_fav_")
    try:
        C.write_case(C.build_case(fav_spec()), root)
        oracle.run_reference(root, threads=1)
        d = os.path.join(root, "output", "field_avg")
        manifest = [ln.strip() for ln in open(os.path.join(d, "fav_manifest.dat")) if ln.strip()]
        out = {"manifest": json.dumps(manifest)}
        for tag, path in (("first", os.path.join(d, manifest[0])), ("last", os.path.join(d, manifest[-1])),
                          ("combined", os.path.join(d, "combined", "fav0000001-0000005"))):
            a = fav_arrays(path)
            out[tag + "_n"] = a["n"]
            out[tag + "_re"] = a["re"]
            out[tag + "_times"] = a["times"]
            out[tag + "_nrec"] = len(a["recs"])
            for q, r in enumerate(a["recs"]):
                out["%s_r%02d" % (tag, q)] = r
        out["meta"] = json.dumps(dict(dims=DIMS, nend=5, npriavg=2, avg_tst=0.0, mode="les"))
        np.savez_compressed(os.path.join(HERE, "fav", "cyl16_les_5.npz"), **out)
        print("manifest", manifest)
    finally:
        shutil.rmtree(root)


if __name__ == "__main__":
    main()
