"""Shared case definitions of the tests (tiny heated-cylinder cases)."""
import json
import os

import numpy as np

from leswht_amd import case as C

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


TRACE_CELLS = ((5, 6, 2), (20, 16, 4), (28, 25, 7))


def spec_for(dims, mode, nend=1, nlev=None):
    if mode == "dom_les":
        # the configurations' domain [-15,30] x [-15,15] x [0,pi], dmin 0.01
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=nlev or 5, re=3900.0, idtopt=1,
                       cflfac=0.5)
        return C.CaseSpec(*dims, dmin=0.01, noise=1e-2, settings=s)
    if mode == "chan_les":
        # the reference's own kind of case: a constant-flow-rate channel LES
        # with the dynamic Vreman model, test filter in x and z (both
        # periodic)
        s = C.Settings(nend=nend, iles=1, idvmon=1, filter=3, resid1=1e-8, nlev=nlev or 2, re=5000.0, ihtrans=0,
                       ich=1, udrv_i=1.0, bc_ybtm=0, bc_ytop=0)
        return C.CaseSpec(*dims, xlim=(1.0, 4.0), ylim=(-1.0, 1.0), zlen=1.0, noise=2e-1, settings=s,
                          layout="channel")
    if mode == "chan_body":
        # a constant-flow-rate channel with the cylinder (radius 0.5) across
        # its middle: the IB forcing enters meanpg's pmi(0) and its tables
        # wrap through the x-periodic boundary
        s = C.Settings(nend=nend, iles=1, idvmon=0, csgsts=0.0, csgshf=0.0, resid1=1e-8, nlev=nlev or 2, re=180.0,
                       ihtrans=0, ich=1, udrv_i=1.0, bc_ybtm=0, bc_ytop=0)
        return C.CaseSpec(*dims, xlim=(-1.5, 1.5), ylim=(-1.0, 1.0), zlen=1.0, noise=1e-2, settings=s,
                          layout="channel")
    if mode in ("chan_cpg", "chan_cfr", "chan_cfr_xft", "chan_cpg_xsor", "chan_cfr_ptb", "chan_cpg_ptb0"):
        # a plane channel: x and z periodic, no-slip walls at y = -1, 1, no
        # body in the domain (the cylinder of the case builder lies outside
        # x in [1, 4]); SGS with zero constants, no heat; ICH 2 (constant
        # pressure gradient -UDRV_I) or 1 (constant flow rate, PI control).
        # The _xft and _xsor forms take poiss_x_ft (x DFT, z multigrid), the
        # latter with line-SOR above x wavenumber 3 and an iteration limit
        # the SOR wavenumbers reach
        ich = 2 if mode.startswith("chan_cpg") else 1
        s = C.Settings(nend=nend, iles=1, idvmon=0, csgsts=0.0, csgshf=0.0, resid1=1e-8, nlev=nlev or 2, re=180.0,
                       ihtrans=0, ich=ich, udrv_i=1.0, bc_ybtm=0, bc_ytop=0)
        if mode == "chan_cfr_xft":
            s.poisson = 3
        elif mode == "chan_cpg_xsor":
            s.poisson, s.imgsor, s.wwsor, s.mgitr = 3, 3, 1.5, 25
        elif mode == "chan_cfr_ptb":
            # makefld's laminar profile, perturbed (addperturb) in the step
            # that reaches PTB_TST = 0.01
            s.iread, s.eps_ptr, s.ptb_tst = 0, 0.1, 0.01
        elif mode == "chan_cpg_ptb0":
            # makefld's bulk-12 profile, perturbed before the first step
            s.iread, s.eps_ptr, s.ptb_tst = 0, 0.05, 0.0
        return C.CaseSpec(*dims, xlim=(1.0, 4.0), ylim=(-1.0, 1.0), zlen=1.0, noise=1e-2, settings=s,
                          layout="channel")
    if mode == "nosgs":
        s = C.Settings(nend=nend, iles=1, idvmon=0, csgsts=0.0, csgshf=0.0, resid1=1e-8, nlev=nlev or 2)
    elif mode == "les":
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=nlev or 2, re=3900.0)
    elif mode == "les_div":
        # the divergence floor: the dynamic LES with the Poisson tolerance
        # RESID1 = 1e-11
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-11, nlev=nlev or 3, re=3900.0)
    elif mode == "les_const":
        s = C.Settings(nend=nend, iles=1, idvmon=0, csgsts=0.025, csgshf=0.003, resid1=1e-8, nlev=nlev or 2,
                       re=3900.0)
    elif mode == "les_trace_iread0":
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=nlev or 2, re=3900.0, iread=0,
                       trace=TRACE_CELLS, ntr=2)
    elif mode == "les_sor":
        # kv = 1, 2 by multigrid, kv = 3..n3m/2+1 by line-SOR (gsor2d)
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=nlev or 2, re=3900.0, imgsor=2, wwsor=1.8)
    elif mode == "les_cfl":
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=nlev or 3, re=3900.0, idtopt=1, cflfac=0.5)
    else:
        raise ValueError(mode)
    return C.CaseSpec(*dims, xlim=(-3.0, 6.0), ylim=(-3.0, 3.0), zlen=1.0, dmin=0.05, noise=1e-2, settings=s)


def c2_spec(nend=1):
    """The 384x256x64 heated-cylinder LES: Re 3900, dynamic Vreman and eddy
    diffusivity, box filter in x and z, CFL 0.5, NLEV 5, noise 1e-2; x and y
    lines clustered at the cylinder surface with spacing 0.01 there."""
    s = C.Settings(re=3900.0, pr=0.71, iles=1, insmdl=1, itemdl=0, idvmon=1, filter=3, idtopt=1,
                   cflfac=0.5, dt_size=0.005, resid1=1.0e-6, nlev=5, nbli=3, mgitr=500, imgsor=170,
                   nend=nend)
    return C.CaseSpec(384, 256, 64, dmin=0.01, noise=1.0e-2, settings=s, layout="tanh")


def golden_view(meta):
    """The case view of a golden fixture, with the Poisson solver it ran."""
    view = C.case_view(C.build_case(spec_for(tuple(meta["dims"]), meta["mode"], meta["nend"], meta.get("nlev"))))
    if meta.get("poisson") == "mgrd":
        view["settings"].poisson = 1
    elif meta.get("poisson") == "x_ft":
        view["settings"].poisson = 3
    return view


def golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return z, meta


def golden_names():
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(a)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a - b))


def fld_regions(n, heat=True):
    """The parts of the padded fields a fld file holds (misc_hist.f90:236-247);
    without heat transfer there is no T."""
    n1, n2, n3 = n
    out = {"u": (slice(1, n1 + 1), slice(0, n2 + 1), slice(0, n3 + 1)),
           "v": (slice(0, n1 + 1), slice(1, n2 + 1), slice(0, n3 + 1)),
           "w": (slice(0, n1 + 1), slice(0, n2 + 1), slice(1, n3 + 1)),
           "p": (slice(1, n1), slice(1, n2), slice(1, n3)),
           "t": (slice(1, n1), slice(1, n2), slice(1, n3))}
    if not heat:
        del out["t"]
    return out
