"""Shared case definitions of the tests (tiny heated-cylinder cases)."""
import json
import os

import numpy as np

from leswht_amd import case as C

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def spec_for(dims, mode, nend=1, nlev=None):
    if mode == "nosgs":
        s = C.Settings(nend=nend, iles=1, idvmon=0, csgsts=0.0, csgshf=0.0, resid1=1e-8, nlev=nlev or 2)
    elif mode == "les":
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=nlev or 2, re=3900.0)
    elif mode == "les_cfl":
        s = C.Settings(nend=nend, iles=1, idvmon=1, resid1=1e-8, nlev=nlev or 3, re=3900.0, idtopt=1, cflfac=0.5)
    else:
        raise ValueError(mode)
    return C.CaseSpec(*dims, xlim=(-3.0, 6.0), ylim=(-3.0, 3.0), zlen=1.0, dmin=0.05, noise=1e-2, settings=s)


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


def fld_regions(n):
    n1, n2, n3 = n
    return {"u": (slice(1, n1 + 1), slice(0, n2 + 1), slice(0, n3 + 1)),
            "v": (slice(0, n1 + 1), slice(1, n2 + 1), slice(0, n3 + 1)),
            "w": (slice(0, n1 + 1), slice(0, n2 + 1), slice(1, n3 + 1)),
            "p": (slice(1, n1), slice(1, n2), slice(1, n3)),
            "t": (slice(1, n1), slice(1, n2), slice(1, n3))}
