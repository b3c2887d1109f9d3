"""Writes byte fixtures of the pre-processor's text writers by calling the
REFERENCE's own Python pre-processor functions in this build container:
lib_setgrid.setgrid reads a grid.dat (in grid.py's %23.15f layout) and
lib_preprocessing.grid_preprocessing_data writes grid.bin from it;
ibm_preprocessing_data, ibm_preprocessing_data_htransfer and
les_preprocessing_data write the ibmpre_*.bin files from given tables.

The tables handed to the reference writers are the case builder's (their
values are pinned separately against the reference's Fortran ibm_body
routines, tests/golden/make_ib_golden.py); what these fixtures pin is the
writers' text: formats, order, separators and index bases. Each fixture
holds the case spec and the bytes of grid.dat, grid.bin and the ibmpre
files as uint8 arrays.

Run in the build container (the reference sources are not on the GPU box):
    python tests/golden/make_preproc_golden.py
"""
import dataclasses
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from leswht_amd import case as C  # noqa: E402

REF_PRE = "/root/reference/01_pre_processor"

# (name, sizes, domain, dmin, body, body args)
CASES = [
    ("cyl_24x24x4", (24, 24, 4), (-3.0, 6.0), (-3.0, 3.0), 1.0, 0.05, "cylinder", ()),
    ("square_40x32x8", (40, 32, 8), (-3.0, 6.0), (-3.0, 3.0), 1.0, 0.05, "square", (0.5, 0.7)),
    ("ellipse_64x48x8", (64, 48, 8), (-15.0, 30.0), (-15.0, 15.0), float(np.pi), 0.01, "ellipse",
     (0.5, 0.25, 1.1)),
]


def spec_of(entry):
    name, dims, xlim, ylim, zlen, dmin, body, args = entry
    s = C.Settings(iles=1, idvmon=1, ihtrans=1)
    return C.CaseSpec(*dims, xlim=xlim, ylim=ylim, zlen=zlen, dmin=dmin, body=body, body_args=args, settings=s)


def grid_dat_text(case):
    """grid.dat as grid.py:84-95 writes it, from the case's grid lines."""
    ax, ay, az = case.ax, case.ay, case.az
    out = ["%23d %22d %22d\n" % (ax.ncell + 1, ay.ncell + 1, az.ncell + 1),
           "%23.15f %22.15f %22.15f\n" % tuple(case.lengths)]
    for a, end in ((ax, "\n"), (ay, "\n"), (az, "")):
        out.append("".join("%23.15f" % v for v in a.coord[1:a.ncell + 2]) + end)
    return "".join(out)


def main():
    sys.path.insert(0, os.path.join(REF_PRE, "lib"))
    import lib_preprocessing as LP  # the reference's writers
    import lib_setgrid as LS        # the reference's grid metrics
    for entry in CASES:
        name = entry[0]
        spec = spec_of(entry)
        case = C.build_case(spec)
        tab = case.tables
        with tempfile.TemporaryDirectory() as root:
            pre = os.path.join(root, "01_pre_processor")
            for d in ("01_pre_processor", "output/grid", "output/ibmpre"):
                os.makedirs(os.path.join(root, d), exist_ok=True)
            gdat = os.path.join(root, "output/grid/grid.dat")
            with open(gdat, "w") as f:
                f.write(grid_dat_text(case))
            cwd = os.getcwd()
            os.chdir(pre)  # the writers open ../output/...
            try:
                grid = LS.setgrid("../output/grid/grid.dat", "OFF", "OFF", "ON")
                LP.grid_preprocessing_data(grid)
                comps = ("u", "v", "w")
                nintp = {c: tab.nintp[c] for c in comps + ("t",)}
                ninner = {c: tab.ninner[c] for c in comps + ("t",)}
                fcp = {c: tab.fcp[c] for c in comps + ("t",)}
                indx = {c: tab.indx[c] for c in comps + ("t",)}
                gf = {c: tab.geomfac_raw[c] for c in comps + ("t",)}
                LP.ibm_preprocessing_data(nintp, ninner, fcp, indx, gf)
                LP.ibm_preprocessing_data_htransfer(nintp, ninner, fcp, indx, gf)
                LP.les_preprocessing_data(tab.nzero, tab.iszero)
            finally:
                os.chdir(cwd)
            files = {"grid_dat": "output/grid/grid.dat", "grid_bin": "output/grid/grid.bin",
                     "fcpts": "output/ibmpre/ibmpre_fcpts.bin", "fcpts_t": "output/ibmpre/ibmpre_fcpts_t.bin",
                     "nutzero": "output/ibmpre/ibmpre_nutzero.bin", "wallfdvm": "output/ibmpre/ibmpre_wallfdvm.bin"}
            data = {k: np.frombuffer(open(os.path.join(root, p), "rb").read(), dtype=np.uint8) for k, p in files.items()}
        meta = dict(name=name, dims=list(entry[1]), xlim=list(entry[2]), ylim=list(entry[3]), zlen=entry[4],
                    dmin=entry[5], body=entry[6], body_args=list(entry[7]))
        os.makedirs(os.path.join(HERE, "preproc"), exist_ok=True)
        np.savez_compressed(os.path.join(HERE, "preproc", name + ".npz"), meta=json.dumps(meta), **data)
        print(name, {k: len(v) for k, v in data.items()})


if __name__ == "__main__":
    main()
