"""Writes the golden fixtures of the time step by running the REFERENCE
solver (oracle/_ref/solver_exec, compiled from the reference's own Fortran
sources by oracle/Makefile) on tiny cylinder cases built by
leswht_amd.case. Each fixture holds the case spec, a hash of the generated
input text, the fld fields the reference wrote after NEND steps and the rows
of its history files (fhist, fcdcl, fheatflux, ftemp).

Run in the build container (the reference sources are not on the GPU box):
    make -C oracle ref && python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from leswht_amd import case as C  # noqa: E402
import oracle  # noqa: E402

CASES = {
    # no SGS: ILES=1 with zero constant coefficients (the reference's ILES=0
    # path reads unallocated LES arrays and stops, see DESIGN)
    "cyl32_nosgs_1": dict(dims=(32, 32, 8), mode="nosgs", nend=1),
    "cyl32_nosgs_5": dict(dims=(32, 32, 8), mode="nosgs", nend=5),
    "cyl32_les_3": dict(dims=(32, 32, 8), mode="les", nend=3),
    "cyl48_les_cfl_4": dict(dims=(48, 32, 16), mode="les_cfl", nend=4),
}


def spec_for(dims, mode, nend):
    if mode == "nosgs":
        s = C.Settings(nend=nend, iles=1, idvmon=0, csgsts=0.0, csgshf=0.0, resid1=1e-8, nlev=2)
    elif mode == "les":
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=2, re=3900.0)
    else:
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=3, re=3900.0, idtopt=1, cflfac=0.5)
    return C.CaseSpec(*dims, xlim=(-3.0, 6.0), ylim=(-3.0, 3.0), zlen=1.0, dmin=0.05, noise=1e-2, settings=s)


def input_hash(root):
    h = hashlib.sha256()
    for rel in ("output/grid/grid.bin", "output/ibmpre/ibmpre_fcpts.bin", "output/ibmpre/ibmpre_fcpts_t.bin",
                "output/ibmpre/ibmpre_nutzero.bin", "02_solver/settings.input", "output/field/fld000000"):
        p = os.path.join(root, rel)
        if os.path.exists(p):
            h.update(open(p, "rb").read())
    return h.hexdigest()


def main():
    for name, c in CASES.items():
        spec = spec_for(c["dims"], c["mode"], c["nend"])
        root = tempfile.mkdtemp(prefix="lw_gold_")
        C.write_case(C.build_case(spec), root)
        oracle.run_reference(root, threads=1)
        fl = C.read_fld(os.path.join(root, "output/field/fld%07d" % c["nend"]))
        rows = {}
        for key, fn in (("hist", "fhist"), ("cdcl", "fcdcl"), ("hflux", "fheatflux"), ("temp", "ftemp")):
            path = os.path.join(root, "output/ftr/%s.dat" % fn)
            rows[key] = np.array([[float(x) for x in ln.split()] for ln in open(path) if ln.strip()],
                                 dtype=np.float64)
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            u=fl["u"], v=fl["v"], w=fl["w"], p=fl["p"], t=fl["t"],
                            time=fl["time"], dt=fl["dt"], ihist=fl["ihist"], **rows,
                            meta=json.dumps(dict(c, input_sha256=input_hash(root))))
        print(name, "time", fl["time"], "dvmax", rows["hist"][-1][3])
        shutil.rmtree(root)


if __name__ == "__main__":
    main()
