"""Writes the immersed-boundary table fixtures under tests/golden/ibm/ by
running the REFERENCE table routines (ibm_body.f90 with the cylinder body
function, compiled from the reference's sources by ``make -C oracle ibm``
into oracle/_ref/ibm_tables) on the axes that leswht_amd.case builds.

Each fixture holds the grid lines and periodic flags (the inputs) and, per
component u, v, w, t: nbody, nintp, ninner, the forcing-point indices, the
interpolation directions and the raw FP64 weights, plus nzero and the
zero-nu_t mask. tests/test_ib_tables.py compares the case builder with them
bit for bit.

Run in the build container (the reference sources are not on the GPU box):
    make -C oracle ibm && python tests/golden/make_ib_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from leswht_amd import case as C  # noqa: E402

OUT = os.path.join(HERE, "ibm")
EXE = os.path.join(ROOT, "oracle", "_ref", "ibm_tables")
COMPS = ("u", "v", "w", "t")


def _lines(kind, n, lo, hi, dmin=0.05):
    if kind == "sinh":
        return C._fmt15(C.gridlines_sinh(n, lo, hi, dmin))
    return C._fmt15(C.gridlines_uniform(n, lo, hi))


# (x lines, y lines, z lines, z periodic). The uniform 0.25 grid of 25 cells
# has cell centres at x = +-0.5, y = 0, so w and t points lie exactly on the
# surface and take the exact-zero exit of the surface search.
CASES = {
    "sinh_24x24x4": (("sinh", 24, -3.0, 6.0), ("sinh", 24, -3.0, 3.0), ("uniform", 4, 0.0, 1.0), True),
    "sinh_40x40x4": (("sinh", 40, -3.0, 6.0), ("sinh", 40, -3.0, 3.0), ("uniform", 4, 0.0, 1.0), True),
    "uniform_25x25x3": (("uniform", 25, -3.125, 3.125), ("uniform", 25, -3.125, 3.125), ("uniform", 3, 0.0, 1.0), True),
    "offset_37x29x3_wall": (("uniform", 37, -2.3, 4.1), ("sinh", 29, -2.05, 2.7), ("uniform", 3, 0.0, 0.7), False),
}


def axes_for(name):
    xs, ys, zs, zper = CASES[name]
    ax = C.make_axis(_lines(*xs), periodic=False)
    ay = C.make_axis(_lines(*ys), periodic=False)
    az = C.make_axis(_lines(*zs), periodic=zper)
    return ax, ay, az


def axes_bin(ax, ay, az, heat=1, les=1) -> bytes:
    n = [a.ncell + 1 for a in (ax, ay, az)]
    head = np.array(n + [heat, les] + [int(a.periodic) for a in (ax, ay, az)], dtype=np.int64)
    parts = [head.tobytes()]
    for a in (ax, ay, az):
        parts += [a.coord.astype(np.float64).tobytes(), a.center.astype(np.float64).tobytes()]
    for a in (ax, ay, az):
        parts += [a.upperfix.astype(np.int64).tobytes(), a.lowerfix.astype(np.int64).tobytes()]
    return b"".join(parts)


def parse_tables(raw: bytes, n, heat=1, les=1) -> dict:
    nx, ny, nz = n
    pos = 0

    def take(count, dt):
        nonlocal pos
        a = np.frombuffer(raw, dtype=dt, count=count, offset=pos)
        pos += a.nbytes
        return a

    out = {}
    for c in COMPS[:4 if heat else 3]:
        nbody, nintp, ninner = (int(v) for v in take(3, np.int64))
        nb = nintp + ninner
        out[c + "_counts"] = np.array([nbody, nintp, ninner], dtype=np.int64)
        out[c + "_fcp"] = take(nb * 3, np.int64).reshape((nb, 3), order="F").copy()
        out[c + "_indx"] = take(nintp * 3, np.int64).reshape((nintp, 3), order="F").copy()
        out[c + "_geomfac"] = take(nintp * 27, np.float64).reshape((nintp, 3, 3, 3), order="F").copy()
    if les:
        out["nzero"] = np.array(int(take(1, np.int64)[0]), dtype=np.int64)
        out["iszero"] = take((nx - 1) * (ny - 1) * (nz - 1), np.int64).reshape(
            (nx - 1, ny - 1, nz - 1), order="F").copy()
    assert pos == len(raw), (pos, len(raw))
    return out


def main():
    os.makedirs(OUT, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    for name in CASES:
        ax, ay, az = axes_for(name)
        with tempfile.TemporaryDirectory(prefix="lw_ib_") as tmp:
            fin, fout = os.path.join(tmp, "axes.bin"), os.path.join(tmp, "tables.bin")
            with open(fin, "wb") as f:
                f.write(axes_bin(ax, ay, az))
            subprocess.run([EXE, fin, fout], check=True, env=env)
            tabs = parse_tables(open(fout, "rb").read(), [a.ncell + 1 for a in (ax, ay, az)])
        np.savez_compressed(os.path.join(OUT, name + ".npz"),
                            x=ax.coord, xm=ax.center, y=ay.coord, ym=ay.center, z=az.coord, zm=az.center,
                            **tabs)
        print(name, {c: tabs[c + "_counts"].tolist() for c in COMPS}, "nzero", int(tabs["nzero"]))


if __name__ == "__main__":
    main()
