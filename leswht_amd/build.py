"""Builds the gfx950 library ``leswht_amd/_lib/libleswht.so`` with hipcc, and
the Fortran host ``leswht_amd/_lib/solver_exec`` (the drop-in for the
reference executable, calling the library through ISO_C_BINDING) with
amdflang when it is installed.

FP64 parity with the reference needs IEEE evaluation in the written order,
so every translation unit is compiled with -ffp-contract=off (no FMA
contraction) and without fast-math.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libleswht.so")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["lw_solver.hip", "lw_batch.hip", "lw_kern_flow.hip", "lw_kern_sgs.hip", "lw_kern_pois.hip", "lw_kern_stats.hip",
               "lw_static.hip", "lw_kern_mgrd.hip", "lw_kern_chan.hip",
               "lw_kern_xft.hip"]
CXX_SOURCES = ["lw_io.cpp", "lw_serve.cpp", "lw_ibm.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("LESWHT_FC", "/opt/rocm/bin/amdflang")
HOST = os.path.join(HERE, "host")
EXE = os.path.join(OUT_DIR, "solver_exec")
ARCH = os.environ.get("LESWHT_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError("compile failed: " + " ".join(cmd[-3:]))
    return res


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(os.path.join(OUT_DIR, "obj"), exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES + CXX_SOURCES:
        obj = os.path.join(OUT_DIR, "obj", src + ".o")
        objs.append(obj)
        cmd = [HIPCC] + COMMON
        if src.endswith(".hip"):
            cmd += DEVICE + ["-x", "hip"]
        cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    failed = []
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed.append((cmd, out + err))
        elif verbose and err.strip():
            sys.stderr.write(err)
    if failed:
        for cmd, msg in failed:
            sys.stderr.write(" ".join(cmd) + "\n" + msg + "\n")
        raise RuntimeError("hipcc failed on %d file(s)" % len(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_serve(verbose)
    build_host(verbose)
    return LIB


SERVE_EXE = os.path.join(OUT_DIR, "serve_exec")


def build_serve(verbose: bool = False) -> str:
    """serve_exec: the request loop over a manifest of case directories."""
    tmp = SERVE_EXE + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HOST, "serve_exec.cpp"), "-o", tmp, "-L" + OUT_DIR, "-lleswht", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, SERVE_EXE)
    return SERVE_EXE


def build_host(verbose: bool = False) -> str | None:
    """The Fortran solver_exec, linked against libleswht.so with an $ORIGIN
    rpath so the pair can move together. Skipped when amdflang is absent."""
    if not os.path.exists(FC):
        if verbose:
            print("amdflang not found: solver_exec not built")
        return None
    mod_dir = os.path.join(OUT_DIR, "obj", "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    tmp = EXE + ".tmp"
    cmd = [FC, "-O2", "-module-dir", mod_dir,
           os.path.join(HOST, "leswht_c.f90"), os.path.join(HOST, "solver_exec.f90"),
           "-o", tmp, "-L" + OUT_DIR, "-lleswht", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, EXE)
    return EXE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
