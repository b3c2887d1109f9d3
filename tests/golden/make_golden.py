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
from tests.helpers import spec_for  # noqa: E402

CASES = {
    # no SGS: ILES=1 with zero constant coefficients (the reference's ILES=0
    # path reads unallocated LES arrays and stops, see DESIGN)
    "cyl32_nosgs_1": dict(dims=(32, 32, 8), mode="nosgs", nend=1),
    "cyl32_nosgs_5": dict(dims=(32, 32, 8), mode="nosgs", nend=5),
    "cyl32_les_3": dict(dims=(32, 32, 8), mode="les", nend=3),
    "cyl48_les_cfl_4": dict(dims=(48, 32, 16), mode="les_cfl", nend=4),
    # constant-coefficient Vreman and eddy diffusivity (sgs_cvm, sgs_cedm)
    "cyl32_const_3": dict(dims=(32, 32, 8), mode="les_const", nend=3),
    # the configurations' domain and near-body spacing, five multigrid levels
    "cyl192_dom_nlev5_3": dict(dims=(192, 64, 16), mode="dom_les", nend=3, nlev=5),
    # the divergence floor after projection: RESID1 = 1e-11 (its fhist rows
    # hold the reference's own max |div u - qmass| per step)
    "cyl64_div_2": dict(dims=(64, 64, 16), mode="les_div", nend=2, nlev=3),
    # 100 steps of the dynamic LES
    "cyl32_les_100": dict(dims=(32, 32, 8), mode="les", nend=100),
    # makefld (IREAD = 0) and the tracer every 2 steps
    "cyl32_iread0_trace_4": dict(dims=(32, 32, 8), mode="les_trace_iread0", nend=4),
    # line-SOR (gsor2d) for the wavenumbers above IMGSOR = 2, with its RG rows
    "cyl32_sor_3": dict(dims=(32, 32, 8), mode="les_sor", nend=3),
    # plane channels, x periodic, with poiss_ftft.f90 (the reference Makefile's
    # default, oracle/_ref/solver_exec_ftft): constant pressure gradient and
    # constant flow rate, with their fcmfr.dat rows
    "chan32_cpg_3": dict(dims=(32, 32, 8), mode="chan_cpg", nend=3, poisson="ftft"),
    "chan32_cfr_4": dict(dims=(32, 32, 8), mode="chan_cfr", nend=4, poisson="ftft"),
    # the same channels with poiss_x_ft.f90 (oracle/_ref/solver_exec_xft): the
    # z multigrid per x wavenumber with its ' mg' rows, and line-SOR above
    # wavenumber 3 up to its iteration limit with its 'SOR' rows
    "xft_chan32_cfr_4": dict(dims=(32, 32, 8), mode="chan_cfr_xft", nend=4, poisson="x_ft"),
    "xft_chan32_sor_3": dict(dims=(32, 32, 8), mode="chan_cpg_xsor", nend=3, poisson="x_ft"),
    # IREAD = 0 channels: makefld's ICH profiles and addperturb, in the step
    # that reaches PTB_TST and before the first step (PTB_TST = 0)
    "chan32_ptb_4": dict(dims=(32, 32, 8), mode="chan_cfr_ptb", nend=4, poisson="ftft"),
    "chan32_ptb0_3": dict(dims=(32, 32, 8), mode="chan_cpg_ptb0", nend=3, poisson="ftft"),
    # the cylinder across a constant-flow-rate channel: IB forcing in meanpg
    "chan32_body_3": dict(dims=(32, 32, 8), mode="chan_body", nend=3, poisson="ftft"),
    # channel LES with the dynamic Vreman model (x- and z-periodic test filter)
    "chan32_les_3": dict(dims=(32, 32, 8), mode="chan_les", nend=3, poisson="ftft"),
    # the 3-D MG-preconditioned BiCGSTAB Poisson solver (poiss_mgrd.f90) linked
    # in place of poiss_z_ft.f90 (oracle/_ref/solver_exec_mgrd)
    "mgrd_cyl32_les_3": dict(dims=(32, 32, 8), mode="les", nend=3, poisson="mgrd"),
    "mgrd_cyl48_les_cfl_4": dict(dims=(48, 32, 16), mode="les_cfl", nend=4, nlev=3, poisson="mgrd"),
}


def input_hash(root):
    h = hashlib.sha256()
    for rel in ("output/grid/grid.bin", "output/ibmpre/ibmpre_fcpts.bin", "output/ibmpre/ibmpre_fcpts_t.bin",
                "output/ibmpre/ibmpre_nutzero.bin", "02_solver/settings.input", "output/field/fld000000"):
        p = os.path.join(root, rel)
        if os.path.exists(p):
            h.update(open(p, "rb").read())
    return h.hexdigest()


def text_rows(path):
    return np.array([[float(x) for x in ln.split()] for ln in open(path) if ln.strip()], dtype=np.float64)


def mg_rows(path):
    """mgftresiduemax.dat: 'KV=  k   KK=  n   RM= r' per line."""
    rows = []
    for ln in open(path):
        t = ln.replace("=", " ").split()
        if len(t) == 6:
            rows.append((float(t[1]), float(t[3]), float(t[5].replace("D", "E"))))
    return np.array(rows, dtype=np.float64).reshape(-1, 3)


def xft_rows(path):
    """poiss_x_ft's mgftresiduemax.dat: ' mg, time = t IV= v II= n RG= r'
    (mg2d, poiss_x_ft.f90:293) and 'SOR IV= v II= n RG= r' (gsor2d at its
    limit, :606), as (kind 0/1, time or 0, iv, ii, rg)."""
    rows = []
    for ln in open(path):
        t = ln.replace("=", " ").split()
        if len(t) == 9 and t[0] == "mg,":
            rows.append((0.0, float(t[2]), float(t[4]), float(t[6]), float(t[8].replace("D", "E"))))
        elif len(t) == 7 and t[0] == "SOR":
            rows.append((1.0, 0.0, float(t[2]), float(t[4]), float(t[6].replace("D", "E"))))
    return np.array(rows, dtype=np.float64).reshape(-1, 5)


def main(names=None):
    for name, c in CASES.items():
        if names and name not in names:
            continue
        spec = spec_for(c["dims"], c["mode"], c["nend"], c.get("nlev"))
        root = tempfile.mkdtemp(prefix="lw_gold_")
        C.write_case(C.build_case(spec), root)
        mgrd = c.get("poisson") == "mgrd"
        ftft = c.get("poisson") == "ftft"
        xft = c.get("poisson") == "x_ft"
        exe = (oracle.REF_EXEC_MGRD if mgrd else oracle.REF_EXEC_FTFT if ftft else
               oracle.REF_EXEC_XFT if xft else oracle.REF_EXEC)
        oracle.run_reference(root, threads=1, exe=exe)
        fl = C.read_fld(os.path.join(root, "output/field/fld%07d" % c["nend"]))
        rows = {}
        for key, fn in (("hist", "fhist"), ("cdcl", "fcdcl"), ("hflux", "fheatflux"), ("temp", "ftemp")):
            rows[key] = text_rows(os.path.join(root, "output/ftr/%s.dat" % fn))
        if mgrd:
            # poiss_itr.dat: iteration, residual*dtconst, i, j, k (poiss_mgrd.f90:185-186);
            # ftrpoittr.dat: time, dt, iterations, residual*dtconst per solve (:221-222)
            rows["bicg_itr"] = text_rows(os.path.join(root, "output/ftr/poiss_itr.dat"))
            rows["bicg_solve"] = text_rows(os.path.join(root, "output/ftr/ftrpoittr.dat"))
        elif xft:
            rows["xres"] = xft_rows(os.path.join(root, "output/ftr/mgftresiduemax.dat"))
        elif not ftft:
            rows["mgres"] = mg_rows(os.path.join(root, "output/ftr/mgftresiduemax.dat"))
        if spec.settings.ich != 0:
            # fcmfr.dat: time, qflux, ubulk, pmiavg, cmfravg(1:3) (misc_hist.f90:136)
            rows["cmfr"] = text_rows(os.path.join(root, "output/ftr/fcmfr.dat"))
        extra = {}
        if spec.settings.trace:
            for q in "uvwp":
                extra["trace_" + q] = text_rows(os.path.join(root, "output/ftr/f%strace.dat" % q))
        if spec.settings.iread == 0:
            f0 = C.read_fld(os.path.join(root, "output/field/fld0000000"))
            for q in ("u", "v", "w", "p", "t"):
                extra["init_" + q] = f0.get(q, np.zeros(0))
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            u=fl["u"], v=fl["v"], w=fl["w"], p=fl["p"], t=fl.get("t", np.zeros(0)),
                            time=fl["time"], dt=fl["dt"], ihist=fl["ihist"], **rows, **extra,
                            meta=json.dumps(dict(c, input_sha256=input_hash(root))))
        print(name, "time", fl["time"], "dvmax", rows["hist"][-1][3])
        shutil.rmtree(root)


if __name__ == "__main__":
    main(sys.argv[1:])
