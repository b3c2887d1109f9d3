"""ctypes binding of the CPU restatement (oracle) of the LESwHT time step.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblw_oracle.so")
REF_EXEC = os.path.join(HERE, "_ref", "solver_exec")

_D = C.c_double
_I = C.c_int64
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int64)


class Settings(C.Structure):
    _fields_ = [(n, _D) for n in (
        "re", "pr", "gr", "tflu_i", "tsol_i", "eps_ptr", "udrv_i", "dt_size", "cflfac",
        "resid1", "wwsor", "csgsts", "csgshf", "tend", "ptb_tst", "avg_tst")] + \
        [(n, _I) for n in (
            "grdir", "ireset", "iread", "iavg", "ipzero", "nend", "nprint", "npriavg", "npin",
            "idtopt", "nlev", "nbli", "ioldv", "mgitr", "imgsor", "iles", "insmdl", "itemdl",
            "idvmon", "filter", "ibmon", "masson", "imovingon", "ihtrans", "idecomp",
            "bc_ybtm", "bc_ytop", "bc_zbtm", "bc_ztop", "bc_t_ybtm", "bc_t_ytop", "bc_t_zbtm",
            "bc_t_ztop", "ich", "iconjg", "xprdic", "yprdic", "zprdic", "iintp")] + \
        [(n, _D) for n in ("val_t_ybtm", "val_t_ytop", "val_t_zbtm", "val_t_ztop")]


_GRID_D = ("x", "y", "z")
_GRID_I = ("ipv", "jpv", "kpv", "imv", "jmv", "kmv", "fixil", "fixjl", "fixkl",
           "fixiu", "fixju", "fixku")
_GRID_M = ("c2cx", "c2cy", "c2cz", "c2cxi", "c2cyi", "c2czi", "f2fx", "f2fy", "f2fz",
           "f2fxi", "f2fyi", "f2fzi", "xmp", "ymp", "zmp")


class Grid(C.Structure):
    _fields_ = [(n, _I) for n in ("n1", "n2", "n3", "n1m", "n2m", "n3m")] + \
        [(n, _D) for n in ("xl", "yl", "zl")] + \
        [(n, _PD) for n in _GRID_D] + [(n, _PI) for n in _GRID_I] + \
        [(n, _PD) for n in _GRID_M]


class Tables(C.Structure):
    _fields_ = [("nintp", _I * 4), ("nbody", _I * 4), ("ifc", _PI * 4), ("jfc", _PI * 4),
                ("kfc", _PI * 4), ("intpindx", _PI * 4), ("geomfac", _PD * 4),
                ("nzero", _I), ("inz", _PI), ("jnz", _PI), ("knz", _PI), ("nwall_dvm", _PI)]


class Diag(C.Structure):
    _fields_ = [(n, _D) for n in ("dt", "time", "cflmax", "dvmax", "qmmax", "csgsts",
                                  "csgshf")] + [("cd", _D * 3), ("heatflux", _D * 3)] + \
        [(n, _I) for n in ("ntime", "mg_cycles_max", "mg_cycles_sum")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", HERE])
        L = C.CDLL(LIB_PATH)
        L.ora_create.restype = C.c_void_p
        L.ora_create.argtypes = [C.POINTER(Settings), C.POINTER(Grid), C.POINTER(Tables)]
        L.ora_destroy.argtypes = [C.c_void_p]
        L.ora_set_fields.argtypes = [C.c_void_p, _PD, _PD, _PD, _PD, _PD, _I, _I, _D, _D]
        L.ora_get_fields.argtypes = [C.c_void_p, _PD, _PD, _PD, _PD, _PD]
        L.ora_init.argtypes = [C.c_void_p]
        L.ora_set_body.argtypes = [C.c_void_p, _I, C.POINTER(_D)]
        L.ora_funcbody.argtypes = [C.c_void_p, _D, _D, _D]
        L.ora_funcbody.restype = _D
        L.ora_step.argtypes = [C.c_void_p, C.POINTER(Diag)]
        L.ora_get_array.argtypes = [C.c_void_p, C.c_char_p, _PD]
        L.ora_get_array.restype = _I
        for n in ("ora_phase_sgs", "ora_phase_rhsnlhs", "ora_phase_divgs", "ora_phase_poisson",
                  "ora_phase_ucalc_pcalc", "ora_phase_energy", "ora_stepinit"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.ora_substep_begin.argtypes = [C.c_void_p, C.c_int]
        L.ora_fft_forward.argtypes = [C.c_int, _PD, _PD]
        L.ora_fft_inverse.argtypes = [C.c_int, _PD, _PD]
        L.ora_thomas.argtypes = [_I, _PD, _PD, _PD, _PD]
        L.ora_thomas_periodic.argtypes = [_I, _PD, _PD, _PD, _PD]
        L.ora_set_poisson.argtypes = [C.c_void_p, _I]
        L.ora_set_poisson.restype = C.c_int
        L.ora_bicg_log.argtypes = [C.c_void_p, _PI, _PI, _PD, _PI, _I]
        L.ora_bicg_log.restype = _I
        L.ora_bicg_solves.argtypes = [C.c_void_p, _PI, _PD]
        L.ora_xft_rows.argtypes = [C.c_void_p, _PI, _PI, _PI, _PD, _I]
        L.ora_xft_rows.restype = _I
        L.ora_bicg_solves.restype = _I
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_PD)


def _pi(a):
    return a.ctypes.data_as(_PI)


def fortran_padded(a: np.ndarray) -> np.ndarray:
    """[i, j, k] numpy array -> contiguous buffer in Fortran (i fastest) order."""
    return np.asfortranarray(a, dtype=np.float64)


def settings_struct(s) -> Settings:
    st = Settings()
    for name, _ in Settings._fields_:
        src = {"imovingon": "imoving"}.get(name, name)
        val = getattr(s, src)
        setattr(st, name, val)
    return st


def fortran_1d(arr, lo):
    """Array read for Fortran indices lo..hi stored at its Fortran index."""
    out = np.zeros(len(arr) + lo + 1, dtype=arr.dtype)
    out[lo:lo + len(arr)] = arr
    return out


class Oracle:
    """One oracle instance built from :func:`leswht_amd.case.load_case_dir`."""

    def __init__(self, cd: dict, poisson: str = "z_ft"):
        L = lib()
        self._keep = []
        s = cd["settings"]
        g = cd["grid"]
        self.g = g
        self.heat = bool(s.ihtrans)
        self.st = settings_struct(s)
        gr = Grid()
        for n in ("n1", "n2", "n3", "n1m", "n2m", "n3m"):
            setattr(gr, n, g[n])
        gr.xl, gr.yl, gr.zl = g["xl"], g["yl"], g["zl"]
        for n in _GRID_D:
            a = fortran_1d(g[n].astype(np.float64), 1); self._keep.append(a); setattr(gr, n, _pd(a))
        for n in _GRID_I:
            a = fortran_1d(g[n].astype(np.int64), 1); self._keep.append(a); setattr(gr, n, _pi(a))
        for n in _GRID_M:
            a = fortran_1d(g[n].astype(np.float64), 0); self._keep.append(a); setattr(gr, n, _pd(a))
        tb = Tables()
        tabs = cd["tables"]
        for ci, c in enumerate(("u", "v", "w", "t")):
            if c not in tabs:
                continue
            t = tabs[c]
            tb.nintp[ci] = t["nintp"]
            tb.nbody[ci] = t["nbody"]
            fcp = np.ascontiguousarray(t["fcp"].reshape(-1, 3), dtype=np.int64)
            ifc = np.ascontiguousarray(fcp[:, 0]); jfc = np.ascontiguousarray(fcp[:, 1])
            kfc = np.ascontiguousarray(fcp[:, 2])
            ix = np.ascontiguousarray(t["indx"].reshape(-1, 3), dtype=np.int64)
            gf = np.ascontiguousarray(t["geomfac"].reshape(-1, 27), dtype=np.float64)
            self._keep += [ifc, jfc, kfc, ix, gf]
            tb.ifc[ci], tb.jfc[ci], tb.kfc[ci] = _pi(ifc), _pi(jfc), _pi(kfc)
            tb.intpindx[ci] = _pi(ix)
            tb.geomfac[ci] = _pd(gf)
        nz = cd["nutzero"]
        tb.nzero = nz["nzero"]
        for n in ("inz", "jnz", "knz"):
            a = np.ascontiguousarray(nz[n], dtype=np.int64); self._keep.append(a); setattr(tb, n, _pi(a))
        nw = np.asfortranarray(nz["nwall"], dtype=np.int64)
        self._keep.append(nw)
        tb.nwall_dvm = _pi(nw)
        self.h = L.ora_create(C.byref(self.st), C.byref(gr), C.byref(tb))
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        if cd.get("body") is not None:
            kind, bp = cd["body"]
            bp = tuple(bp) + (0.0,) * (4 - len(bp))
            L.ora_set_body(self.h, int(kind), (_D * 4)(*bp[:4]))
        f = cd["fields"]
        arrs = [fortran_padded(f[k]) for k in ("u", "v", "w", "p")]
        tt = fortran_padded(f["t"]) if f.get("t") is not None else None
        L.ora_set_fields(self.h, *[_pd(a) for a in arrs], _pd(tt) if tt is not None else None,
                         cd["ihist"], cd["m"], cd["time"], cd["dt"])
        L.ora_init(self.h)
        # the solver the settings name (poisson = 1 mgrd, 2 ftft, 3 x_ft); an
        # x-periodic case otherwise takes poiss_ftft, the reference Makefile's
        # default
        if poisson == "z_ft":
            poisson = {1: "mgrd", 2: "ftft", 3: "x_ft"}.get(int(getattr(cd["settings"], "poisson", 0)), "z_ft")
        if poisson == "z_ft" and cd["settings"].xprdic == 1:
            poisson = "ftft"
        if poisson != "z_ft":
            kind = {"mgrd": 1, "ftft": 2, "x_ft": 3}.get(poisson, -1)
            if kind < 0 or L.ora_set_poisson(self.h, kind) != 0:
                raise ValueError("oracle: Poisson path %r not available for this case" % poisson)

    def bicg_log(self):
        """Per BiCGSTAB iteration of the last step: (substep, iteration,
        residual * dtconst, i, j, k); per solve: (iteration, residual *
        dtconst)."""
        L = lib()
        cap = 3 * 4096
        sub, icg, loc = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(3 * cap, np.int64)
        rm = np.zeros(cap)
        n = L.ora_bicg_log(self.h, _pi(sub), _pi(icg), _pd(rm), _pi(loc), cap)
        it = [(int(sub[q]), int(icg[q]), float(rm[q]), int(loc[3 * q]), int(loc[3 * q + 1]), int(loc[3 * q + 2]))
              for q in range(n)]
        si, sr = np.zeros(3, np.int64), np.zeros(3)
        ns = L.ora_bicg_solves(self.h, _pi(si), _pd(sr))
        return it, [(int(si[q]), float(sr[q])) for q in range(ns)]

    def xft_rows(self):
        """poiss_x_ft's mgftresiduemax.dat rows of the last step: (kind,
        wavenumber, cycle counter, residual * dtconst * n1mh / n1m), kind 0
        for mg2d, 1 for gsor2d at its limit."""
        cap = 4096
        k, iv, ii = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(cap, np.int64)
        rg = np.zeros(cap)
        n = lib().ora_xft_rows(self.h, _pi(k), _pi(iv), _pi(ii), _pd(rg), cap)
        return [(int(k[q]), int(iv[q]), int(ii[q]), float(rg[q])) for q in range(min(n, cap))]

    def step(self, n: int = 1):
        d = Diag()
        for _ in range(n):
            lib().ora_step(self.h, C.byref(d))
        return d

    def shape(self):
        return (self.g["n1"] + 1, self.g["n2"] + 1, self.g["n3"] + 1)

    def fields(self) -> dict:
        sh = self.shape()
        out = {k: np.zeros(sh, order="F") for k in ("u", "v", "w", "p", "t")}
        lib().ora_get_fields(self.h, *[_pd(out[k]) for k in ("u", "v", "w", "p", "t")])
        if not self.heat:
            out["t"] = None
        return out

    def array(self, name: str) -> np.ndarray:
        n = lib().ora_get_array(self.h, name.encode(), None)
        out = np.zeros(n)
        lib().ora_get_array(self.h, name.encode(), _pd(out))
        sh = self.shape()
        if n == np.prod(sh):
            return out.reshape(sh, order="F")
        g = self.g
        if n == g["n1m"] * g["n2m"] * g["n3m"]:
            return out.reshape((g["n1m"], g["n2m"], g["n3m"]), order="F")
        return out  # scalars: "cmfr" (the fcmfr.dat row), "pmi"

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ora_destroy(self.h)
        except Exception:
            pass


REF_EXEC_MGRD = os.path.join(HERE, "_ref", "solver_exec_mgrd")
REF_EXEC_FTFT = os.path.join(HERE, "_ref", "solver_exec_ftft")
REF_EXEC_XFT = os.path.join(HERE, "_ref", "solver_exec_xft")


def run_reference(root: str, threads: int = 1, timeout: int = 600, exe: str = REF_EXEC) -> str:
    """Run the amdflang-built reference solver (oracle/_ref/solver_exec, or
    ``exe``, e.g. REF_EXEC_MGRD) in ``root``/02_solver. Returns its stdout."""
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    REF_EXEC_ = exe
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_DYNAMIC="FALSE",
               OMP_STACKSIZE="1G", OMP_PROC_BIND="close")
    res = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec " + REF_EXEC_],
                         cwd=os.path.join(root, "02_solver"), env=env, capture_output=True,
                         text=True, timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError("reference solver failed (%d): %s" % (res.returncode, res.stderr[-2000:]))
    return res.stdout
