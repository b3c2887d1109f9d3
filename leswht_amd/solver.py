"""Python host of the MI355X time step: a ctypes binding of the C ABI in
``include/leswht.h`` (library ``leswht_amd/_lib/libleswht.so``).

The reference's operator API is the sequence of Fortran calls in
``02_solver/f90/solver.f90``; :class:`Solver` mirrors it: construction does
readsettings .. poisinit (``solver.f90:28-61``), :meth:`Solver.step` runs
passes of the time loop (``solver.f90:85-181``) and :meth:`Solver.write_field`
is writefield (``misc_hist.f90:212-252``). There is no CPU fallback: without
the HIP library or a GPU every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libleswht.so")

_D, _I = C.c_double, C.c_int64
_PD, _PI = C.POINTER(C.c_double), C.POINTER(C.c_int64)


class LwError(RuntimeError):
    pass


class lw_settings(C.Structure):
    _fields_ = [(n, _D) for n in (
        "re", "pr", "gr", "tflu_i", "tsol_i", "eps_ptr", "udrv_i", "dt_size", "cflfac",
        "resid1", "wwsor", "csgsts", "csgshf", "tend", "ptb_tst", "avg_tst")] + \
        [(n, _I) for n in (
            "grdir", "ireset", "iread", "iavg", "ipzero", "nend", "nprint", "npriavg", "npin",
            "idtopt", "nlev", "nbli", "ioldv", "mgitr", "imgsor", "iles", "insmdl", "itemdl",
            "idvmon", "filter", "ibmon", "masson", "imovingon", "ihtrans", "idecomp",
            "bc_ybtm", "bc_ytop", "bc_zbtm", "bc_ztop", "bc_t_ybtm", "bc_t_ytop", "bc_t_zbtm",
            "bc_t_ztop", "ich", "iconjg", "xprdic", "yprdic", "zprdic", "iintp")] + \
        [(n, _D) for n in ("val_t_ybtm", "val_t_ytop", "val_t_zbtm", "val_t_ztop")] + \
        [("poisson", _I)]


GRID_D = ("x", "y", "z")
GRID_I = ("ipv", "jpv", "kpv", "imv", "jmv", "kmv", "fixil", "fixjl", "fixkl",
          "fixiu", "fixju", "fixku")
GRID_M = ("c2cx", "c2cy", "c2cz", "c2cxi", "c2cyi", "c2czi", "f2fx", "f2fy", "f2fz",
          "f2fxi", "f2fyi", "f2fzi", "xmp", "ymp", "zmp")


class lw_grid(C.Structure):
    _fields_ = [(n, _I) for n in ("n1", "n2", "n3", "n1m", "n2m", "n3m")] + \
        [(n, _D) for n in ("xl", "yl", "zl")] + \
        [(n, _PD) for n in GRID_D] + [(n, _PI) for n in GRID_I] + [(n, _PD) for n in GRID_M]


class lw_tables(C.Structure):
    _fields_ = [("nintp", _I * 4), ("nbody", _I * 4), ("ifc", _PI * 4), ("jfc", _PI * 4),
                ("kfc", _PI * 4), ("intpindx", _PI * 4), ("geomfac", _PD * 4),
                ("nzero", _I), ("inz", _PI), ("jnz", _PI), ("knz", _PI), ("nwall_dvm", _PI)]


class lw_step_stats(C.Structure):
    _fields_ = [(n, _D) for n in ("dt", "time", "cflmax", "dvmax", "qmmax", "csgsts", "csgshf")] + \
        [(n, _I) for n in ("ntime", "mg_cycles_max", "mg_cycles_sum")] + \
        [(n, _D) for n in ("ms_sgs", "ms_momentum", "ms_poisson", "ms_energy")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class lw_body(C.Structure):
    _fields_ = [("kind", _I), ("p", _D * 4)]


class lw_history(C.Structure):
    _fields_ = [(n, _I) for n in ("ntime", "ihist", "written")] + \
        [(n, _D) for n in ("time", "dt", "cflmax", "dvmax", "qmmax")] + \
        [("tbulk", _D * 2), ("tbulk_avg", _D * 2), ("cd", _D * 3), ("cd_avg", _D * 3),
         ("hflux", _D * 3), ("hflux_avg", _D * 3)] + \
        [(n, _D) for n in ("s_sgs", "s_rhsnlhs", "s_poisson", "s_energy")] + \
        [(n, _D) for n in ("qflux", "ubulk", "pmiavg")] + [("cmfravg", _D * 3)]

    def as_dict(self):
        out = {}
        for n, _ in self._fields_:
            v = getattr(self, n)
            out[n] = list(v) if hasattr(v, "__len__") else v
        return out


class lw_tune_result(C.Structure):
    _fields_ = [("tile", _I * 3), ("tile1", _I * 3), ("candidates", _I), ("ms_default", _D), ("ms_best", _D)]

    def as_dict(self):
        return dict(tile=list(self.tile), tile1=list(self.tile1), candidates=self.candidates,
                    ms_default=self.ms_default, ms_best=self.ms_best)


class lw_case_result(C.Structure):
    _fields_ = [(n, _I) for n in ("ticket", "status", "cells", "steps_done", "mg_cycles_sum",
                                  "reserved_bytes", "device_bytes")] + \
        [(n, _D) for n in ("t_submit", "t_admit", "t_done", "time", "dt")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class lw_server_info(C.Structure):
    _fields_ = [(n, _I) for n in ("submitted", "queued", "running", "slots", "reserved_bytes",
                                  "peak_reserved_bytes", "peak_running", "budget_bytes")] + \
        [("elapsed", _D)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


CASE_DONE, CASE_QUEUED, CASE_RUNNING = 0, 1, 2

EXPORTS = ("lw_create", "lw_create_from_dir", "lw_destroy", "lw_set_fields", "lw_get_fields",
           "lw_step", "lw_write_field", "lw_write_fld", "lw_get_dims", "lw_get_array", "lw_phase",
           "lw_device_bytes", "lw_profile_begin", "lw_profile_end", "lw_profile_count",
           "lw_profile_get", "lw_get_history", "lw_get_settings", "lw_last_error",
           "lw_estimate_device_bytes", "lw_probe_dir", "lw_body_eval", "lw_set_avg_dir", "lw_avg_finalize", "lw_set_option", "lw_autotune",
           "lw_ib_build", "lw_ib_tables", "lw_ib_geomfac_raw", "lw_ib_write", "lw_ib_free", "lw_server_create", "lw_server_submit",
           "lw_server_poll", "lw_server_wait", "lw_server_stats", "lw_server_destroy", "lw_server_create_batched",
           "lw_create_static", "lw_static_area_bytes", "lw_set_trace", "lw_trace_info", "lw_get_trace",
           "lw_get_mg_residuals", "lw_get_residual_rows", "lw_get_bicg_log", "lw_get_bicg_solves",
           "lw_batch_create", "lw_batch_step", "lw_batch_info", "lw_batch_destroy")

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads the HIP library; raises when it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LwError("libleswht.so not built: run `python -m leswht_amd.build` (%s)" % path)
    L = C.CDLL(path)
    H = C.c_void_p
    L.lw_create.argtypes = [C.POINTER(lw_settings), C.POINTER(lw_grid), C.POINTER(lw_tables),
                            C.POINTER(lw_body), C.POINTER(H)]
    L.lw_create_static.argtypes = L.lw_create.argtypes
    L.lw_set_trace.argtypes = [H, _I, _I, _PI]
    L.lw_trace_info.argtypes = [H, _PI, _PI, _PI]
    L.lw_get_trace.argtypes = [H, _PD, _I]
    L.lw_get_mg_residuals.argtypes = [H, _PI, _PI, _PD, _I]
    L.lw_get_mg_residuals.restype = _I
    L.lw_get_residual_rows.argtypes = [H, _PI, _PI, _PI, _PD, _I]
    L.lw_get_residual_rows.restype = _I
    L.lw_get_bicg_log.argtypes = [H, _PI, _PI, _PD, _PI, _I]
    L.lw_get_bicg_log.restype = _I
    L.lw_get_bicg_solves.argtypes = [H, _PI, _PD, _I]
    L.lw_get_bicg_solves.restype = _I
    L.lw_static_area_bytes.restype = _I
    L.lw_create_from_dir.argtypes = [C.c_char_p, C.POINTER(lw_body), C.POINTER(H)]
    L.lw_destroy.argtypes = [H]
    L.lw_set_fields.argtypes = [H, _PD, _PD, _PD, _PD, _PD, _I, _I, _D, _D]
    L.lw_get_fields.argtypes = [H, _PD, _PD, _PD, _PD, _PD]
    L.lw_step.argtypes = [H, _I, C.POINTER(lw_step_stats)]
    L.lw_write_field.argtypes = [H, C.c_char_p]
    L.lw_get_dims.argtypes = [H, _PI]
    L.lw_get_array.argtypes = [H, C.c_char_p, _PD, _I]
    L.lw_get_array.restype = _I
    L.lw_phase.argtypes = [H, C.c_int]
    L.lw_device_bytes.argtypes = [H]
    L.lw_device_bytes.restype = _I
    L.lw_get_history.argtypes = [H, C.POINTER(lw_history)]
    L.lw_get_settings.argtypes = [H, C.POINTER(lw_settings)]
    L.lw_profile_begin.argtypes = [H, C.c_char_p]
    L.lw_profile_end.argtypes = [H]
    L.lw_profile_count.argtypes = [H]
    L.lw_profile_count.restype = _I
    L.lw_profile_get.argtypes = [H, _I, C.c_char_p, _I, _PI, _PD, _PD]
    L.lw_last_error.restype = C.c_char_p
    L.lw_estimate_device_bytes.argtypes = [C.POINTER(lw_settings), _I, _I, _I]
    L.lw_estimate_device_bytes.restype = _I
    L.lw_probe_dir.argtypes = [C.c_char_p, C.POINTER(lw_settings), _PI]
    L.lw_ib_build.argtypes = [_PD, _I, _PD, _I, _PD, _I, _PI, C.POINTER(lw_body), _I, _I, C.POINTER(H)]
    L.lw_ib_tables.argtypes = [H, C.POINTER(lw_tables)]
    L.lw_ib_geomfac_raw.argtypes = [H, _I, _PD, _I]
    L.lw_ib_geomfac_raw.restype = _I
    L.lw_ib_write.argtypes = [H, C.c_char_p]
    L.lw_ib_free.argtypes = [H]
    L.lw_set_option.argtypes = [H, C.c_char_p, _I]
    L.lw_autotune.argtypes = [H, _I, C.POINTER(lw_tune_result)]
    L.lw_set_avg_dir.argtypes = [H, C.c_char_p]
    L.lw_avg_finalize.argtypes = [H]
    L.lw_body_eval.argtypes = [C.POINTER(lw_body), _D, _D, _D]
    L.lw_body_eval.restype = _D
    L.lw_server_create.argtypes = [_I, _I, C.POINTER(H)]
    L.lw_server_create_batched.argtypes = [_I, _I, _I, C.POINTER(H)]
    L.lw_server_submit.argtypes = [H, C.c_char_p, C.POINTER(lw_body), _I, C.c_char_p, _PI]
    L.lw_server_poll.argtypes = [H, _I, C.POINTER(lw_case_result)]
    L.lw_server_wait.argtypes = [H, _I, C.POINTER(lw_case_result)]
    L.lw_server_stats.argtypes = [H, C.POINTER(lw_server_info)]
    L.lw_server_destroy.argtypes = [H]
    L.lw_batch_create.argtypes = [C.POINTER(H), _I, C.POINTER(H)]
    L.lw_batch_step.argtypes = [H, _I, C.POINTER(lw_step_stats)]
    L.lw_batch_info.argtypes = [H, _PI]
    L.lw_batch_destroy.argtypes = [H]
    _lib = L
    return L


def last_error() -> str:
    return load_library().lw_last_error().decode(errors="replace")


def _check(rc: int):
    if rc != 0:
        msg = load_library().lw_last_error().decode(errors="replace")
        raise LwError("leswht error %d: %s" % (rc, msg))


def _pd(a):
    return a.ctypes.data_as(_PD)


def _pi(a):
    return a.ctypes.data_as(_PI)


def _at(arr, lo, dtype):
    out = np.zeros(len(arr) + lo + 1, dtype=dtype)
    out[lo:lo + len(arr)] = arr
    return out


def settings_struct(s) -> lw_settings:
    st = lw_settings()
    for name, _ in lw_settings._fields_:
        setattr(st, name, getattr(s, {"imovingon": "imoving"}.get(name, name)))
    return st


def body_struct(body) -> lw_body:
    """lw_body from (kind, params); None is the cylinder of radius 0.5."""
    if body is None:
        return lw_body(0, (_D * 4)(0.5, 0, 0, 0))
    kind, p = body
    p = tuple(p) + (0.0,) * (4 - len(p))
    return lw_body(int(kind), (_D * 4)(*p[:4]))


def estimate_device_bytes(settings, n1m: int, n2m: int, n3m: int) -> int:
    """Upper bound of a case's device footprint (lw_estimate_device_bytes)."""
    L = load_library()
    st = settings_struct(settings)
    b = L.lw_estimate_device_bytes(C.byref(st), n1m, n2m, n3m)
    if b < 0:
        _check(int(b))
    return int(b)


class IBTablesNative:
    """Immersed-boundary tables built by the library (lw_ib_build) from the
    grid lines of the three directions and a body (kind, params)."""

    COMPS = ("u", "v", "w", "t")

    def __init__(self, xlines, ylines, zlines, periodic=(0, 0, 1), body=None, heat=True, les=True):
        L = load_library()
        self._arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xlines, ylines, zlines)]
        per = (_I * 3)(*[int(v) for v in periodic])
        b = body_struct(body)
        self._h = C.c_void_p()
        x, y, z = self._arrs
        _check(L.lw_ib_build(_pd(x), len(x), _pd(y), len(y), _pd(z), len(z), per, C.byref(b),
                             int(heat), int(les), C.byref(self._h)))
        self.heat, self.les = heat, les
        self.shape = (len(x) - 1, len(y) - 1, len(z) - 1)  # cells

    def tables(self) -> dict:
        """Copies of the tables: per component nintp, nbody, fcp (nbody, 3),
        indx (nintp, 3), geomfac (nintp, 3, 3, 3) as the solver reads it,
        geomfac_raw; and nzero, iszero (cells, 0 = zero nu_t)."""
        L = load_library()
        t = lw_tables()
        _check(L.lw_ib_tables(self._h, C.byref(t)))
        out = {}
        for ci, c in enumerate(self.COMPS[:4 if self.heat else 3]):
            nb, ni = t.nbody[ci], t.nintp[ci]
            fcp = np.stack([np.ctypeslib.as_array(p, (nb,)) if nb else np.zeros(0, np.int64)
                            for p in (t.ifc[ci], t.jfc[ci], t.kfc[ci])], axis=1).astype(np.int64)
            ix = np.ctypeslib.as_array(t.intpindx[ci], (ni * 3,)).reshape(ni, 3).copy() if ni else \
                np.zeros((0, 3), np.int64)
            gf = np.ctypeslib.as_array(t.geomfac[ci], (ni * 27,)).reshape(ni, 3, 3, 3).copy() if ni else \
                np.zeros((0, 3, 3, 3))
            raw = np.zeros(ni * 27)
            if ni:
                _check(0 if L.lw_ib_geomfac_raw(self._h, ci, _pd(raw), ni * 27) == ni * 27 else -1)
            out[c] = dict(nintp=ni, nbody=nb, fcp=fcp, indx=ix, geomfac=gf, geomfac_raw=raw.reshape(ni, 3, 3, 3))
        out["nzero"] = t.nzero
        if self.les:
            cells = self.shape
            out["iszero"] = np.ctypeslib.as_array(t.nwall_dvm, (int(np.prod(cells)),)).reshape(
                cells, order="F").copy()
        return out

    def write(self, ibmpre_dir: str):
        _check(load_library().lw_ib_write(self._h, ibmpre_dir.encode()))

    def close(self):
        if self._h:
            load_library().lw_ib_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Solvers stepped together (lw_batch_*): every kernel of the step is
    launched once for all the cases that issue it. Each case gives the same
    results as stepped alone. The solvers stay usable for fields and
    histories between steps; destroy the batch (``close``) before the
    solvers."""

    def __init__(self, solvers):
        L = load_library()
        self.solvers = list(solvers)
        hs = (C.c_void_p * len(self.solvers))(*[sv._h.value for sv in self.solvers])
        self._b = C.c_void_p()
        _check(L.lw_batch_create(hs, len(self.solvers), C.byref(self._b)))

    def step(self, nsteps: int = 1) -> list:
        st = (lw_step_stats * len(self.solvers))()
        _check(load_library().lw_batch_step(self._b, int(nsteps), st))
        return [x.as_dict() for x in st]

    def info(self) -> dict:
        a = np.zeros(6, dtype=np.int64)
        _check(load_library().lw_batch_info(self._b, _pi(a)))
        return dict(rendezvous=int(a[0]), launches=int(a[1]), case_launches=int(a[2]), reads=int(a[3]),
                    prepare_us=int(a[4]), wait_us=int(a[5]))

    def close(self):
        if self._b:
            load_library().lw_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop over independent cases on one GPU (lw_server_*): case
    directories are submitted as they arrive, admitted first-come first-served
    by free slot and memory budget, run concurrently on their own streams,
    and written to ``fld_out`` when their steps are done."""

    def __init__(self, budget_bytes: int, slots: int, batch_max_cells: int = 0):
        """``batch_max_cells`` > 0: continuous batching of the cases of at
        most that many cells (lw_server_create_batched)."""
        L = load_library()
        self._s = C.c_void_p()
        _check(L.lw_server_create_batched(int(budget_bytes), int(slots), int(batch_max_cells), C.byref(self._s)))

    def submit(self, solver_dir: str, nsteps: int, fld_out: str = None, body=None) -> int:
        t = _I()
        b = body_struct(body)
        _check(load_library().lw_server_submit(self._s, solver_dir.encode(), C.byref(b), int(nsteps),
                                               fld_out.encode() if fld_out else None, C.byref(t)))
        return t.value

    def poll(self, ticket: int) -> dict:
        r = lw_case_result()
        _check(load_library().lw_server_poll(self._s, ticket, C.byref(r)))
        return r.as_dict()

    def wait(self, ticket: int, raise_on_error: bool = True) -> dict:
        """Blocks until the case is done. A case that failed raises, or with
        ``raise_on_error=False`` comes back with its negative status and the
        library's message under ``error``."""
        r = lw_case_result()
        _check(load_library().lw_server_wait(self._s, ticket, C.byref(r)))
        if r.status != CASE_DONE:
            if raise_on_error:
                _check(int(r.status))
            d = r.as_dict()
            d["error"] = last_error()
            return d
        return r.as_dict()

    def stats(self) -> dict:
        i = lw_server_info()
        _check(load_library().lw_server_stats(self._s, C.byref(i)))
        return i.as_dict()

    def close(self):
        """Runs every queued case to the end, then stops the workers."""
        if self._s:
            load_library().lw_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """One case resident in HBM."""

    def __init__(self, handle, heat: bool, dims, keep=None):
        self._h = handle
        self.heat = heat
        self.n1, self.n2, self.n3 = dims
        self._keep = keep

    # -- construction ------------------------------------------------------
    @classmethod
    def from_case(cls, cd: dict, body=None, static_area: bool = False) -> "Solver":
        """From :func:`leswht_amd.case.load_case_dir` output (arrays as the
        reference holds them after start-up). ``static_area``: create it with
        lw_create_static (no device allocation; the device's static area,
        one such case at a time)."""
        L = load_library()
        s, g = cd["settings"], cd["grid"]
        keep = []
        st = settings_struct(s)
        gr = lw_grid()
        for n in ("n1", "n2", "n3", "n1m", "n2m", "n3m"):
            setattr(gr, n, g[n])
        gr.xl, gr.yl, gr.zl = g["xl"], g["yl"], g["zl"]
        for n in GRID_D:
            a = _at(g[n], 1, np.float64); keep.append(a); setattr(gr, n, _pd(a))
        for n in GRID_I:
            a = _at(g[n], 1, np.int64); keep.append(a); setattr(gr, n, _pi(a))
        for n in GRID_M:
            a = np.ascontiguousarray(g[n], dtype=np.float64); keep.append(a); setattr(gr, n, _pd(a))
        tb = lw_tables()
        for ci, c in enumerate(("u", "v", "w", "t")):
            if c not in cd["tables"]:
                continue
            t = cd["tables"][c]
            tb.nintp[ci], tb.nbody[ci] = t["nintp"], t["nbody"]
            fcp = np.ascontiguousarray(t["fcp"].reshape(-1, 3), dtype=np.int64)
            cols = [np.ascontiguousarray(fcp[:, q]) for q in range(3)]
            ix = np.ascontiguousarray(t["indx"].reshape(-1, 3), dtype=np.int64)
            gf = np.ascontiguousarray(t["geomfac"].reshape(-1, 27), dtype=np.float64)
            keep += cols + [ix, gf]
            tb.ifc[ci], tb.jfc[ci], tb.kfc[ci] = (_pi(a) for a in cols)
            tb.intpindx[ci], tb.geomfac[ci] = _pi(ix), _pd(gf)
        nz = cd["nutzero"]
        tb.nzero = nz["nzero"]
        for n in ("inz", "jnz", "knz"):
            a = np.ascontiguousarray(nz[n], dtype=np.int64); keep.append(a); setattr(tb, n, _pi(a))
        nw = np.asfortranarray(nz["nwall"], dtype=np.int64); keep.append(nw)
        tb.nwall_dvm = _pi(nw)
        body = body_struct(body if body is not None else cd.get("body"))
        h = C.c_void_p()
        create = L.lw_create_static if static_area else L.lw_create
        _check(create(C.byref(st), C.byref(gr), C.byref(tb), C.byref(body), C.byref(h)))
        self = cls(h, bool(s.ihtrans), (g["n1"], g["n2"], g["n3"]))
        f = cd["fields"]
        self.set_fields(f["u"], f["v"], f["w"], f["p"], f.get("t"), cd["ihist"], cd["m"], cd["time"], cd["dt"])
        if getattr(s, "trace", ()):
            self.set_trace(s.trace, s.ntr)
        return self

    @classmethod
    def from_dir(cls, solver_dir: str, body=None) -> "Solver":
        """Reads a reference case directory (cwd of solver_exec). ``body`` is
        (kind, params) as :func:`leswht_amd.case.body_params` gives it; None
        is the reference's cylinder of radius 0.5."""
        L = load_library()
        body = body_struct(body)
        h = C.c_void_p()
        _check(L.lw_create_from_dir(solver_dir.encode(), C.byref(body), C.byref(h)))
        dims = (_I * 3)()
        _check(L.lw_get_dims(h, C.cast(dims, _PI)))
        from . import case as _case  # settings only, to know whether T is carried
        s = _case.parse_settings(open(os.path.join(solver_dir, "settings.input")).read(),
                                 open(os.path.join(solver_dir, "boundary.input")).read(),
                                 open(os.path.join(solver_dir, "../output/ibmpre/ibmpre_prdic.bin")).read())
        return cls(h, bool(s.ihtrans), tuple(dims))

    # -- state -------------------------------------------------------------
    def shape(self):
        return (self.n1 + 1, self.n2 + 1, self.n3 + 1)

    def set_fields(self, u, v, w, p, t=None, ihist=0, m=0, time=0.0, dt=0.0):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (u, v, w, p)]
        tt = np.asfortranarray(t, dtype=np.float64) if t is not None else None
        _check(load_library().lw_set_fields(self._h, *[_pd(a) for a in arrs],
                                            _pd(tt) if tt is not None else None, ihist, m, time, dt))

    def fields(self) -> dict:
        sh = self.shape()
        out = {k: np.zeros(sh, order="F") for k in ("u", "v", "w", "p", "t")}
        _check(load_library().lw_get_fields(self._h, *[_pd(out[k]) for k in ("u", "v", "w", "p", "t")]))
        if not self.heat:
            out["t"] = None
        return out

    def array(self, name: str) -> np.ndarray:
        L = load_library()
        n = L.lw_get_array(self._h, name.encode(), None, 0)
        if n < 0:
            _check(int(n))
        out = np.zeros(n)
        _check(0 if L.lw_get_array(self._h, name.encode(), _pd(out), n) == n else -1)
        if n == int(np.prod(self.shape())):
            return out.reshape(self.shape(), order="F")
        return out  # "pmi": the channel's pmi(0:4)

    # -- time-averaged statistics (field_avg) --------------------------------
    AVG_NAMES = ("uavg", "vavg", "wavg") + tuple("uiujavg_%d" % q for q in range(1, 7)) + \
        ("pavg", "p2avg") + tuple("voravg_%d" % q for q in range(1, 4)) + \
        tuple("vor2avg_%d" % q for q in range(1, 7)) + ("ssavg", "tavg", "t2avg")

    def avg_array(self, name: str) -> np.ndarray:
        """A running sum of field_avg as an (n1m, n2m, n3m) array."""
        L = load_library()
        n = L.lw_get_array(self._h, name.encode(), None, 0)
        if n < 0:
            _check(int(n))
        out = np.zeros(n)
        _check(0 if L.lw_get_array(self._h, name.encode(), _pd(out), n) == n else -1)
        return out.reshape((self.n1 - 1, self.n2 - 1, self.n3 - 1), order="F")

    def autotune(self, reps: int = 2) -> dict:
        """Picks the fused multigrid kernels' tile sizes by timing Poisson
        solves of the current field (lw_autotune); results stay bitwise."""
        r = lw_tune_result()
        _check(load_library().lw_autotune(self._h, int(reps), C.byref(r)))
        return r.as_dict()

    def set_option(self, name: str, value: int):
        _check(load_library().lw_set_option(self._h, name.encode(), int(value)))

    def set_avg_dir(self, path):
        _check(load_library().lw_set_avg_dir(self._h, path.encode() if path else None))

    def avg_finalize(self):
        _check(load_library().lw_avg_finalize(self._h))

    # -- time stepping -----------------------------------------------------
    def step(self, n: int = 1) -> dict:
        st = lw_step_stats()
        _check(load_library().lw_step(self._h, n, C.byref(st)))
        return st.as_dict()

    def phase(self, ph: int):
        _check(load_library().lw_phase(self._h, ph))

    def write_field(self, path: str):
        _check(load_library().lw_write_field(self._h, path.encode()))

    def history(self) -> dict:
        """History quantities of the last completed step: the values the
        reference writes to fhist/ftemp/fcdcl/fheatflux.dat (``written`` is 1
        on the steps it writes them) and the device seconds of the phases."""
        h = lw_history()
        _check(load_library().lw_get_history(self._h, C.byref(h)))
        return h.as_dict()

    def set_trace(self, cells, ntr: int = 1):
        """Tracer cells (i, j, k), traced every ``ntr`` steps (lw_set_trace)."""
        pts = np.ascontiguousarray(np.asarray(cells, dtype=np.int64).reshape(-1))
        _check(load_library().lw_set_trace(self._h, len(pts) // 3, int(ntr), _pi(pts) if len(pts) else None))

    def trace_info(self):
        """(ntrace, ntr, cells) of the handle."""
        n, r = _I(), _I()
        pts = np.zeros(192, dtype=np.int64)
        _check(load_library().lw_trace_info(self._h, C.byref(n), C.byref(r), _pi(pts)))
        return n.value, r.value, pts[:3 * n.value].reshape(-1, 3)

    def trace(self) -> dict:
        """What tracer writes for the current fields: u, v, w at the traced
        cell centres and p there (arrays of ntrace values)."""
        n = self.trace_info()[0]
        out = np.zeros(max(4 * n, 1))
        _check(load_library().lw_get_trace(self._h, _pd(out), len(out)))
        return {c: out[q * n:(q + 1) * n].copy() for q, c in enumerate("uvwp")}

    def residual_rows(self):
        """Every mgftresiduemax.dat row of the last step in file order: list of
        (kind, kv, kk, value), kind 0 for mg2d's RM rows and 1 for gsor2d's RG
        rows (one per sweep of its loop, kv above IMGSOR)."""
        L = load_library()
        n = int(L.lw_get_residual_rows(self._h, None, None, None, None, 0))
        if n < 0:
            _check(n)
        kind, kv, kk = (np.zeros(max(n, 1), np.int64) for _ in range(3))
        val = np.zeros(max(n, 1))
        L.lw_get_residual_rows(self._h, _pi(kind), _pi(kv), _pi(kk), _pd(val), n)
        return [(int(kind[q]), int(kv[q]), int(kk[q]), float(val[q])) for q in range(n)]

    def mg_residuals(self):
        """The mgftresiduemax.dat rows of the last step: list of (kv, kk, rm)."""
        L = load_library()
        kv, kk, rm = np.zeros(9, np.int64), np.zeros(9, np.int64), np.zeros(9)
        n = L.lw_get_mg_residuals(self._h, _pi(kv), _pi(kk), _pd(rm), 9)
        if n < 0:
            _check(int(n))
        return [(int(kv[q]), int(kk[q]), float(rm[q])) for q in range(min(n, 9))]

    def bicg_log(self):
        """With the 3-D Poisson solver: the rows of the last step, per
        BiCGSTAB iteration (substep, iteration, residual * dtconst, i, j, k)
        and per solve (iterations, residual * dtconst)."""
        L = load_library()
        cap = 3 * 4096
        sub, icg, loc = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(3 * cap, np.int64)
        rm = np.zeros(cap)
        n = L.lw_get_bicg_log(self._h, _pi(sub), _pi(icg), _pd(rm), _pi(loc), cap)
        if n < 0:
            _check(int(n))
        it = [(int(sub[q]), int(icg[q]), float(rm[q]), int(loc[3 * q]), int(loc[3 * q + 1]), int(loc[3 * q + 2]))
              for q in range(min(n, cap))]
        si, sr = np.zeros(3, np.int64), np.zeros(3)
        ns = L.lw_get_bicg_solves(self._h, _pi(si), _pd(sr), 3)
        return it, [(int(si[q]), float(sr[q])) for q in range(min(ns, 3))]

    def settings(self) -> dict:
        st = lw_settings()
        _check(load_library().lw_get_settings(self._h, C.byref(st)))
        return {n: getattr(st, n) for n, _ in lw_settings._fields_}

    def profile_begin(self, kernels=None):
        """Starts a fresh per-kernel event profile of the following steps;
        ``kernels`` None times every kernel, else an iterable of names."""
        arg = None if not kernels else ",".join(kernels).encode()
        _check(load_library().lw_profile_begin(self._h, arg))

    def profile_end(self):
        _check(load_library().lw_profile_end(self._h))

    def profile(self) -> list:
        """Per-kernel results of the current profile: dicts with name, count
        (launches), ms_total (summed event time) and bytes (summed
        algorithmic HBM bytes)."""
        L = load_library()
        out = []
        name = C.create_string_buffer(96)
        cnt, ms, by = _I(), _D(), _D()
        for q in range(L.lw_profile_count(self._h)):
            _check(L.lw_profile_get(self._h, q, name, 96, C.byref(cnt), C.byref(ms), C.byref(by)))
            out.append(dict(name=name.value.decode(), count=cnt.value, ms_total=ms.value, bytes=by.value))
        return out

    def device_bytes(self) -> int:
        return int(load_library().lw_device_bytes(self._h))

    def close(self):
        if self._h:
            _check(load_library().lw_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
