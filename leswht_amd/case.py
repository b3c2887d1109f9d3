"""Case builder for the LESwHT time step: grid, immersed-boundary tables,
settings files and Fortran-unformatted field files.

Everything written here follows the on-disk contract the reference solver
reads, byte for byte:

* ``grid.dat`` (``01_pre_processor/grid.py:84-95``, ``%23.15f`` coordinates),
* ``grid.bin`` metrics (``01_pre_processor/lib/lib_setgrid.py:24-110`` for the
  formulas, ``lib/lib_preprocessing.py:112-170`` for the ``%f`` text layout),
* ``ibmpre_*.bin`` tables (``lib_preprocessing.py:223-338``) built by the
  table algorithms of ``02_solver/f90/lib/ibm_body.f90:14-399``,
* ``settings.input`` / ``boundary.input`` (read by
  ``02_solver/f90/lib/misc_init.f90:19-158``),
* ``fld*`` restart files (``02_solver/f90/lib/misc_hist.f90:212-252``).

The grid-line distribution itself (a two-sided sinh clustering around the
body) is this module's own; the reference's ``grid.py`` only produces an input
fixture, so any distribution is a valid input.
"""
from __future__ import annotations

import dataclasses
import math
import os
import struct
from typing import Callable, Dict, Optional, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# bodies: level-set functions phi(x, y, z) > 0 in fluid
# (geometry/funcbody_example/funcbody_cylinder.f90:13)
# ---------------------------------------------------------------------------


def body_cylinder(x, y, z, radius=0.5):
    """sqrt(x**2 + y**2) - r, evaluated as sqrt(x*x + y*y) - r like the
    Fortran ``x**2`` integer power (funcbody_cylinder.f90:13)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    return np.sqrt(x * x + y * y) - radius


def body_square(x, y, z, half=0.5, angle=0.0):
    c, s = math.cos(angle), math.sin(angle)
    xr = c * np.asarray(x) + s * np.asarray(y)
    yr = -s * np.asarray(x) + c * np.asarray(y)
    return np.maximum(np.abs(xr), np.abs(yr)) - half


def body_ellipse(x, y, z, a=0.5, b=0.25, angle=0.0):
    c, s = math.cos(angle), math.sin(angle)
    xr = c * np.asarray(x) + s * np.asarray(y)
    yr = -s * np.asarray(x) + c * np.asarray(y)
    # scaled so that phi has unit slope along the minor axis
    return (np.sqrt((xr / a) ** 2 + (yr / b) ** 2) - 1.0) * min(a, b)


BODIES: Dict[str, Callable] = {
    "cylinder": body_cylinder,
    "square": body_square,
    "ellipse": body_ellipse,
}

# the solver library's lw_body kinds and the default parameters of each body
BODY_KINDS: Dict[str, Tuple[int, Tuple[float, ...]]] = {
    "cylinder": (0, (0.5,)),
    "square": (1, (0.5, 0.0)),
    "ellipse": (2, (0.5, 0.25, 0.0)),
}


def body_params(name: str, args: Tuple = ()) -> Tuple[int, Tuple[float, float, float, float]]:
    """(kind, p[0..3]) of the library's lw_body for a named body with the
    positional arguments its level-set function takes."""
    kind, defaults = BODY_KINDS[name]
    vals = list(args) + list(defaults[len(args):])
    if len(vals) > len(defaults):
        raise ValueError("too many arguments for body %r" % name)
    return kind, tuple(float(v) for v in vals) + (0.0,) * (4 - len(vals))

# ---------------------------------------------------------------------------
# grid lines
# ---------------------------------------------------------------------------


def gridlines_uniform(ncell: int, lo: float, hi: float) -> np.ndarray:
    return lo + (hi - lo) * np.arange(ncell + 1, dtype=np.float64) / ncell


def gridlines_sinh(ncell: int, lo: float, hi: float, dmin: float) -> np.ndarray:
    """Two-sided sinh clustering: x(s) = A sinh(beta (s - s0)), s in [0, 1],
    with x(0) = lo < 0 < hi = x(1) and smallest spacing about ``dmin`` at x=0."""
    if not (lo < 0.0 < hi):
        raise ValueError("sinh clustering needs lo < 0 < hi")
    target = dmin * ncell  # A * beta

    def mismatch(beta):
        a = target / beta
        return math.asinh(hi / a) - math.asinh(lo / a) - beta

    blo, bhi = 1e-6, 200.0
    if mismatch(blo) < 0:  # uniform spacing already finer than dmin
        return gridlines_uniform(ncell, lo, hi)
    for _ in range(200):
        mid = 0.5 * (blo + bhi)
        if mismatch(mid) > 0:
            blo = mid
        else:
            bhi = mid
    beta = 0.5 * (blo + bhi)
    a = target / beta
    s0 = -math.asinh(lo / a) / beta
    s = np.arange(ncell + 1, dtype=np.float64) / ncell
    xs = a * np.sinh(beta * (s - s0))
    xs[0], xs[-1] = lo, hi
    return xs


def _bisect(f, lo, hi, iters=200):
    """Root of a function decreasing on [lo, hi] (f(lo) > 0 > f(hi))."""
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def _tanh_expand(n: int, length: float, d1: float) -> np.ndarray:
    """n + 1 points from 0 to ``length`` with the hyperbolic-tangent
    expansion of the pre-processor's 'H' segments (lib_gridfunc.py:64-69 with
    Factor1 = 1): x(p) = L (1 - tanh(g (1 - p)) / tanh g), p = i / n, fine at
    0; g is chosen so that the first spacing is d1."""
    p = np.arange(n + 1, dtype=np.float64) / n
    if length / n <= d1:
        return length * p
    first = lambda g: length * (1.0 - math.tanh(g * (1.0 - 1.0 / n)) / math.tanh(g)) - d1
    g = _bisect(first, 1e-6, 30.0)
    x = length * (1.0 - np.tanh(g * (1.0 - p)) / math.tanh(g))
    x[0], x[-1] = 0.0, length
    return x


def _tanh_sym(n: int, length: float, d1: float) -> np.ndarray:
    """n + 1 points from 0 to ``length``, fine at both ends: the 'H' segment
    with Factor1 = 0.5 (lib_gridfunc.py:64-69), x(p) = L/2 (1 - tanh(g (1/2 -
    p)) / tanh(g/2)); g chosen so that the end spacings are d1."""
    p = np.arange(n + 1, dtype=np.float64) / n
    if length / n <= d1:
        return length * p
    first = lambda g: 0.5 * length * (1.0 - math.tanh(g * (0.5 - 1.0 / n)) / math.tanh(0.5 * g)) - d1
    g = _bisect(first, 1e-6, 60.0)
    x = 0.5 * length * (1.0 - np.tanh(g * (0.5 - p)) / math.tanh(0.5 * g))
    x[0], x[-1] = 0.0, length
    return x


def gridlines_tanh(ncell: int, lo: float, hi: float, dmin: float, half: float = 0.5,
                   ncore: int = 0) -> np.ndarray:
    """Lines clustered at the body surface x = -half and x = +half with
    spacing ``dmin`` there: hyperbolic-tangent segments as the pre-processor
    writes them (grid.input 'H'), an expansion from +half to hi, its mirror
    image from -half down to lo, and a symmetric segment across the body
    (coarsest at its centre). ``ncore`` cells across the body (default
    ncell / 6, even); the rest split between the two sides in proportion to
    the square root of their lengths."""
    if not (lo < -half < half < hi):
        raise ValueError("tanh clustering needs lo < -half < half < hi")
    nc = ncore or max(2, 2 * int(round(ncell / 12.0)))
    rest = ncell - nc
    wl, wr = math.sqrt(-half - lo), math.sqrt(hi - half)
    nl = int(round(rest * wl / (wl + wr)))
    nr = rest - nl
    if min(nl, nr, nc) < 2:
        raise ValueError("too few cells for the tanh layout")
    right = half + _tanh_expand(nr, hi - half, dmin)
    left = -half - _tanh_expand(nl, -half - lo, dmin)[::-1]
    core = -half + _tanh_sym(nc, 2.0 * half, dmin)
    x = np.concatenate([left[:-1], core[:-1], right])
    x[0], x[nl], x[nl + nc], x[-1] = lo, -half, half, hi
    return x


def _fmt15(values: np.ndarray) -> np.ndarray:
    """Round-trip through the ``%23.15f`` text that grid.dat holds."""
    return np.array([float("%23.15f" % v) for v in values], dtype=np.float64)


# ---------------------------------------------------------------------------
# metrics (lib_setgrid.py:24-110) and grid.bin (lib_preprocessing.py:112-170)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Axis:
    """One direction's metrics with the reference's index ranges: arrays of
    length N+1 are indexed 0..N (N = number of grid lines = cells + 1);
    per-cell arrays have length ncell and are indexed from 1 (stored at 0)."""

    ncell: int
    periodic: bool
    coord: np.ndarray  # 0..N
    f2fd: np.ndarray  # 0..N
    c2cd: np.ndarray  # 0..N
    f2fdi: np.ndarray  # 0..N
    c2cdi: np.ndarray  # 0..N
    center: np.ndarray  # 0..N
    idxplus: np.ndarray  # cells 1..ncell
    idxminus: np.ndarray
    lowerfix: np.ndarray
    upperfix: np.ndarray


def make_axis(lines: np.ndarray, periodic: bool) -> Axis:
    """Restates readcoord/setindex/setmesh/physcenter
    (lib_setgrid.py:1-79) for one direction."""
    lines = np.asarray(lines, dtype=np.float64)
    n = len(lines)  # number of grid lines N
    ncell = n - 1
    coord = np.empty(n + 1)
    coord[1:] = lines
    coord[0] = lines[0]
    idxplus = np.arange(2, ncell + 2, dtype=np.int64)
    idxminus = np.arange(0, ncell, dtype=np.int64)
    lowerfix = np.zeros(ncell, dtype=np.int64)
    upperfix = np.zeros(ncell, dtype=np.int64)
    lowerfix[0] = 1
    upperfix[-1] = 1
    if ncell == 1:
        lowerfix[0] = 1
        upperfix[0] = 1
    f2fd = np.zeros(n + 1)
    c2cd = np.zeros(n + 1)
    f2fdi = np.zeros(n + 1)
    c2cdi = np.zeros(n + 1)
    if ncell != 1:
        f2fd[1:ncell + 1] = coord[2:ncell + 2] - coord[1:ncell + 1]
        for i in range(2, ncell + 1):
            c2cd[i] = 0.5 * (f2fd[i] + f2fd[i - 1])
        c2cd[1] = 0.5 * f2fd[1]
        c2cd[ncell + 1] = 0.5 * f2fd[ncell]
        f2fdi[1:ncell + 1] = 1.0 / f2fd[1:ncell + 1]
        c2cdi[1:ncell + 2] = 1.0 / c2cd[1:ncell + 2]
    else:
        coord[2] = coord[1] + 1.0
        f2fd[1] = c2cd[1] = f2fdi[1] = c2cdi[1] = 1.0
        f2fd[2] = c2cd[2] = f2fdi[2] = c2cdi[2] = 1.0
    f2fd[0] = c2cd[0] = c2cdi[0] = f2fdi[0] = 0.0
    f2fd[ncell + 1] = 0.0
    f2fdi[ncell + 1] = 0.0
    if periodic:
        lowerfix[0] = 0
        idxminus[0] = ncell
        upperfix[-1] = 0
        idxplus[-1] = 1
        coord[0] = coord[1] - f2fd[ncell]
        f2fd[0] = f2fd[ncell]
        f2fdi[0] = f2fdi[ncell]
        c2cd[1] = 0.5 * (f2fd[1] + f2fd[ncell])
        f2fd[ncell + 1] = f2fd[1]
        f2fdi[ncell + 1] = f2fdi[1]
        c2cd[ncell + 1] = 0.5 * (f2fd[1] + f2fd[ncell])
        c2cdi[1:ncell + 2] = 1.0 / c2cd[1:ncell + 2]
    center = np.empty(n + 1)
    center[1:ncell + 1] = coord[1:ncell + 1] + 0.5 * f2fd[1:ncell + 1]
    center[ncell + 1] = coord[ncell + 1]
    center[0] = coord[0]
    return Axis(ncell, periodic, coord, f2fd, c2cd, f2fdi, c2cdi, center,
                idxplus, idxminus, lowerfix, upperfix)


def _fmt_f_line(values) -> str:
    return "".join(" %f " % v for v in values) + "\n"


def _fmt_d_line(values) -> str:
    return "".join(" %d " % int(v) for v in values) + "\n"


def grid_bin_text(ax: Axis, ay: Axis, az: Axis, lengths) -> str:
    """Text of grid.bin exactly as lib_preprocessing.py:112-170 writes it."""
    out = []
    out.append("%d %d %d\n" % (ax.ncell + 1, ay.ncell + 1, az.ncell + 1))
    out.append("%d %d %d\n" % (ax.ncell, ay.ncell, az.ncell))
    out.append("%f %f %f\n" % tuple(lengths))
    for a in (ax, ay, az):
        out.append(_fmt_f_line(a.coord[1:a.ncell + 2]))
    for name in ("idxplus", "idxminus", "lowerfix", "upperfix"):
        for a in (ax, ay, az):
            out.append(_fmt_d_line(getattr(a, name)))
    for name in ("c2cd", "c2cdi", "f2fd", "f2fdi"):
        for a in (ax, ay, az):
            out.append(_fmt_f_line(getattr(a, name)))
    out.append(_fmt_f_line(ax.center))
    out.append(_fmt_f_line(ay.center))
    out.append("".join(" %f " % v for v in az.center))
    return "".join(out)


# ---------------------------------------------------------------------------
# immersed-boundary tables (ibm_body.f90:14-399)
# ---------------------------------------------------------------------------


def find_inout(body, xc, yc, zc) -> Tuple[int, np.ndarray]:
    """inout(0:nx,0:ny,0:nz) = 0 where funcbody <= 1e-10
    (ibm_body.f90:14-46). Returned array is indexed [i, j, k]."""
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    val = body(X, Y, Z)
    inout = np.where(val <= 1.0e-10, 0, 1).astype(np.int64)
    return int((inout == 0).sum()), inout


def findbdy_intp(direction: int, inout: np.ndarray, ax: Axis, ay: Axis, az: Axis):
    """Classify body points into interpolation points (an in/out change along
    some non-boundary axis) and inner points, in k -> j -> i loop order,
    interpolation points first (ibm_body.f90:49-123)."""
    nx, ny, nz = (s - 1 for s in inout.shape)
    istart = 2 if direction == 1 else 1
    jstart = 2 if direction == 2 else 1
    kstart = 2 if direction == 3 else 1
    I = np.arange(istart, nx)
    J = np.arange(jstart, ny)
    K = np.arange(kstart, nz)
    sub = inout[np.ix_(I, J, K)]
    dx = inout[np.ix_(I + 1, J, K)] - inout[np.ix_(I - 1, J, K)]
    dy = inout[np.ix_(I, J + 1, K)] - inout[np.ix_(I, J - 1, K)]
    dz = inout[np.ix_(I, J, K + 1)] - inout[np.ix_(I, J, K - 1)]
    fx = ((1 - ax.upperfix[I - 1]) * (1 - ax.lowerfix[I - 1]))[:, None, None]
    fy = ((1 - ay.upperfix[J - 1]) * (1 - ay.lowerfix[J - 1]))[None, :, None]
    fz = ((1 - az.upperfix[K - 1]) * (1 - az.lowerfix[K - 1]))[None, None, :]
    ssum = np.abs(dx) * fx + np.abs(dy) * fy + np.abs(dz) * fz
    body = sub == 0
    intp = body & (ssum > 0)
    inner = body & (ssum == 0)

    def ordered(mask):
        # Fortran loop order k (outer) -> j -> i (inner)
        flat = np.flatnonzero(mask.ravel(order="F"))
        ii, jj, kk = np.unravel_index(flat, mask.shape, order="F")
        return ii, jj, kk

    ii, jj, kk = ordered(intp)
    fcp_intp = np.stack([I[ii], J[jj], K[kk]], axis=1)
    indx = np.stack([dx[ii, jj, kk], dy[ii, jj, kk], dz[ii, jj, kk]], axis=1)
    ii2, jj2, kk2 = ordered(inner)
    fcp_inner = np.stack([I[ii2], J[jj2], K[kk2]], axis=1)
    fcp = np.concatenate([fcp_intp, fcp_inner], axis=0).astype(np.int64)
    return len(fcp_intp), len(fcp_inner), fcp, indx.astype(np.int64)


def geomfac_preset(coord: np.ndarray, center: np.ndarray) -> np.ndarray:
    """i_adj(-1:ni+1, 3) (ibm_body.f90:126-157). The periodic branch compares
    the option string with lower-case 'on' while the pre-processor passes 'ON'
    (preprocessing.py:223-225), so it never runs; the ghost rows stay 0."""
    ni = len(coord) - 1
    adj = np.zeros((ni + 3, 3))
    adj[1:ni + 2, 0] = coord
    adj[1:ni + 2, 1] = center
    adj[1:ni + 2, 2] = center
    return adj


def geomfac_intp(body, xpre, ypre, zpre, fcp, intpindx) -> np.ndarray:
    """27 interpolation weights per interpolation point: 3-level 20-step
    surface search then 1-D linear weights (ibm_body.f90:160-322).
    xpre etc. are i_adj columns with Fortran index -1 stored at 0.
    Vectorised over points; per-point arithmetic follows the Fortran order."""
    npt = len(intpindx)
    g = np.zeros((npt, 3, 3, 3))
    if npt == 0:
        return g
    fi = fcp[:npt]
    x1 = xpre[fi[:, 0] + 1]
    y1 = ypre[fi[:, 1] + 1]
    z1 = zpre[fi[:, 2] + 1]
    x2 = xpre[fi[:, 0] + intpindx[:, 0] + 1]
    y2 = ypre[fi[:, 1] + intpindx[:, 1] + 1]
    z2 = zpre[fi[:, 2] + intpindx[:, 2] + 1]
    x0, y0, z0 = x1.copy(), y1.copy(), z1.copy()
    ffs = body(x1, y1, z1)
    ffe = body(x2, y2, z2)
    failsafe = ffs * ffe > 0.0
    active = ~failsafe  # still searching
    xt1, yt1, zt1 = x1.copy(), y1.copy(), z1.copy()
    xt2, yt2, zt2 = x2.copy(), y2.copy(), z2.copy()
    xx1 = np.zeros(npt); yy1 = np.zeros(npt); zz1 = np.zeros(npt)
    xx2 = np.zeros(npt); yy2 = np.zeros(npt); zz2 = np.zeros(npt)
    for m in (1, 2, 3):
        if m > 1:
            xt1, yt1, zt1 = xx1.copy(), yy1.copy(), zz1.copy()
            xt2, yt2, zt2 = xx2.copy(), yy2.copy(), zz2.copy()
        inloop = active.copy()  # points still scanning this level's n loop
        for n in range(20):
            if not inloop.any():
                break
            idx = np.flatnonzero(inloop)
            ddx = xt2[idx] - xt1[idx]
            ddy = yt2[idx] - yt1[idx]
            ddz = zt2[idx] - zt1[idx]
            a1 = xt1[idx] + ddx * float(n) / 20.0
            a2 = xt1[idx] + ddx * float(n + 1) / 20.0
            b1 = yt1[idx] + ddy * float(n) / 20.0
            b2 = yt1[idx] + ddy * float(n + 1) / 20.0
            c1 = zt1[idx] + ddz * float(n) / 20.0
            c2 = zt1[idx] + ddz * float(n + 1) / 20.0
            xx1[idx], xx2[idx], yy1[idx], yy2[idx], zz1[idx], zz2[idx] = a1, a2, b1, b2, c1, c2
            ff1 = body(a1, b1, c1)
            ff2 = body(a2, b2, c2)
            hit1 = ff1 == 0.0
            hit2 = (~hit1) & (ff2 == 0.0)
            cross = (~hit1) & (~hit2) & (ff1 * ff2 < 0.0)
            # exact zero found: stop the whole search (goto 33)
            for hit, xa, ya, za in ((hit1, a1, b1, c1), (hit2, a2, b2, c2)):
                if hit.any():
                    sel = idx[hit]
                    x0[sel], y0[sel], z0[sel] = xa[hit], ya[hit], za[hit]
                    active[sel] = False
                    inloop[sel] = False
            if cross.any():
                sel = idx[cross]
                x0[sel] = 0.5 * (a1[cross] + a2[cross])
                y0[sel] = 0.5 * (b1[cross] + b2[cross])
                z0[sel] = 0.5 * (c1[cross] + c2[cross])
                inloop[sel] = False
                if m == 3:
                    active[sel] = False
        # points that exhausted n without a crossing continue to the next
        # level from the last (xx1, xx2) pair, as the Fortran falls through
    x3 = xpre[fi[:, 0] + intpindx[:, 0] * 2 + 1]
    y3 = ypre[fi[:, 1] + intpindx[:, 1] * 2 + 1]
    z3 = zpre[fi[:, 2] + intpindx[:, 2] * 2 + 1]
    dx1, dx2, dx3 = np.abs(x1 - x0), np.abs(x2 - x0), np.abs(x3 - x0)
    dy1, dy2, dy3 = np.abs(y1 - y0), np.abs(y2 - y0), np.abs(y3 - y0)
    dz1, dz2, dz3 = np.abs(z1 - z0), np.abs(z2 - z0), np.abs(z3 - z0)

    def weights(ind, d1, d2, d3):
        w0 = np.empty(npt)
        w1 = np.empty(npt)
        zero = ind == 0
        near = (~zero) & (d2 >= d1)
        far = (~zero) & ~(d2 >= d1)
        w0[zero] = 1.0
        w1[zero] = 1.0
        with np.errstate(divide="ignore", invalid="ignore"):
            w0[near] = d2[near] / (d1[near] + d2[near])
            w1[near] = 1.0
            w0[far] = 0.5
            w1[far] = (d3[far] - d1[far]) / (d3[far] - d2[far])
        return w0, w1

    a0, a1 = weights(intpindx[:, 0], dx1, dx2, dx3)
    b0, b1 = weights(intpindx[:, 1], dy1, dy2, dy3)
    c0, c1 = weights(intpindx[:, 2], dz1, dz2, dz3)
    inv = 1.0 / (a0 * b0 * c0)
    # per-axis factors: offset 0 -> body point, 1 -> first, 2 -> second fluid point
    fa = [None, (1.0 - a0) * a1, (1.0 - a0) * (1.0 - a1)]
    g[:, 0, 0, 0] = inv
    # The Fortran multiplies left to right: -1./(a0*b0*c0) * A * B * C ...
    # in the exact factor order of ibm_body.f90:289-317; restated literally.
    A0, B0, C0 = a0, b0, c0
    nA0, nB0, nC0 = 1.0 - a0, 1.0 - b0, 1.0 - c0
    nA1, nB1, nC1 = 1.0 - a1, 1.0 - b1, 1.0 - c1
    t = lambda *fs: _neg_prod(inv, fs)  # noqa: E731
    g[:, 0, 0, 1] = t(A0, B0, nC0, c1)
    g[:, 0, 0, 2] = t(A0, B0, nC0, nC1)
    g[:, 0, 1, 0] = t(A0, nB0, C0, b1)
    g[:, 0, 1, 1] = t(A0, nB0, nC0, b1, c1)
    g[:, 0, 1, 2] = t(A0, nB0, nC0, b1, nC1)
    g[:, 0, 2, 0] = t(A0, nB0, C0, nB1)
    g[:, 0, 2, 1] = t(A0, nB0, nC0, nB1, c1)
    g[:, 0, 2, 2] = t(A0, nB0, nC0, nB1, nC1)
    g[:, 1, 0, 0] = t(nA0, B0, C0, a1)
    g[:, 1, 0, 1] = t(nA0, B0, nC0, a1, c1)
    g[:, 1, 0, 2] = t(nA0, B0, nC0, a1, nC1)
    g[:, 1, 1, 0] = t(nA0, nB0, C0, a1, b1)
    g[:, 1, 1, 1] = t(nA0, nB0, nC0, a1, b1, c1)
    g[:, 1, 1, 2] = t(nA0, nB0, nC0, a1, b1, nC1)
    g[:, 1, 2, 0] = t(nA0, nB0, C0, a1, nB1)
    g[:, 1, 2, 1] = t(nA0, nB0, nC0, a1, nB1, c1)
    g[:, 1, 2, 2] = t(nA0, nB0, nC0, a1, nB1, nC1)
    g[:, 2, 0, 0] = t(nA0, B0, C0, nA1)
    g[:, 2, 0, 1] = t(nA0, B0, nC0, nA1, c1)
    g[:, 2, 0, 2] = t(nA0, B0, nC0, nA1, nC1)
    g[:, 2, 1, 0] = t(nA0, nB0, C0, nA1, b1)
    g[:, 2, 1, 1] = t(nA0, nB0, nC0, nA1, b1, c1)
    g[:, 2, 1, 2] = t(nA0, nB0, nC0, nA1, b1, nC1)
    g[:, 2, 2, 0] = t(nA0, nB0, C0, nA1, nB1)
    g[:, 2, 2, 1] = t(nA0, nB0, nC0, nA1, nB1, c1)
    g[:, 2, 2, 2] = t(nA0, nB0, nC0, nA1, nB1, nC1)
    del fa
    g[failsafe] = 0.0
    g[failsafe, 0, 0, 0] = 1.0
    return g


def _neg_prod(inv, factors):
    acc = inv
    for f in factors:
        acc = acc * f
    return -acc


def find_zero_nu_sgs(body, xm, ym, zm) -> Tuple[int, np.ndarray]:
    """7-point test on cell centres (ibm_body.f90:357-399); returns
    iszero[i-1, j-1, k-1] for cells 1..n."""
    nx, ny, nz = len(xm) - 1, len(ym) - 1, len(zm) - 1
    X, Y, Z = np.meshgrid(xm, ym, zm, indexing="ij")
    solid = body(X, Y, Z) <= 1.0e-10
    c = (slice(1, nx), slice(1, ny), slice(1, nz))
    s = solid[c].copy()
    s |= solid[0:nx - 1, 1:ny, 1:nz]
    s |= solid[2:nx + 1, 1:ny, 1:nz]
    s |= solid[1:nx, 0:ny - 1, 1:nz]
    s |= solid[1:nx, 2:ny + 1, 1:nz]
    s |= solid[1:nx, 1:ny, 0:nz - 1]
    s |= solid[1:nx, 1:ny, 2:nz + 1]
    iszero = np.where(s, 0, 1).astype(np.int64)
    return int(s.sum()), iszero


@dataclasses.dataclass
class IBTables:
    """Forcing-point tables per component c in (u, v, w, t): ``fcp[c]`` is
    (nbody, 3) 1-based (i, j, k), interpolation points first; ``indx[c]`` is
    (nintp, 3); ``geomfac[c]`` is (nintp, 3, 3, 3) indexed [x-off, y-off,
    z-off] with values as the solver reads them back from ``%15.8f`` text."""

    nintp: Dict[str, int]
    ninner: Dict[str, int]
    fcp: Dict[str, np.ndarray]
    indx: Dict[str, np.ndarray]
    geomfac: Dict[str, np.ndarray]
    geomfac_raw: Dict[str, np.ndarray]
    nzero: int = 0
    iszero: Optional[np.ndarray] = None  # (n1m, n2m, n3m), 0 -> zero nu_t


def build_ib_tables(body, ax: Axis, ay: Axis, az: Axis, heat: bool,
                    les: bool) -> IBTables:
    """Restates preprocessing.py:117-276 (IBM-INTERP ON)."""
    X, XM = ax.coord, ax.center
    Y, YM = ay.coord, ay.center
    Z, ZM = az.coord, az.center
    grids = {"u": (X, YM, ZM), "v": (XM, Y, ZM), "w": (XM, YM, Z)}
    comps = ["u", "v", "w"] + (["t"] if heat else [])
    if heat:
        grids["t"] = (XM, YM, ZM)
    XX = geomfac_preset(X, XM)
    YY = geomfac_preset(Y, YM)
    ZZ = geomfac_preset(Z, ZM)
    pre = {"u": (XX[:, 0], YY[:, 2], ZZ[:, 1]),
           "v": (XX[:, 1], YY[:, 0], ZZ[:, 2]),
           "w": (XX[:, 2], YY[:, 1], ZZ[:, 0]),
           "t": (XX[:, 2], YY[:, 2], ZZ[:, 2])}
    nintp, ninner, fcp, indx, gf, gfr = {}, {}, {}, {}, {}, {}
    for d, c in enumerate(comps, start=1):
        _, inout = find_inout(body, *grids[c])
        ni, nn, f, ix = findbdy_intp(d, inout, ax, ay, az)
        raw = geomfac_intp(body, *pre[c], f, ix)
        nintp[c], ninner[c], fcp[c], indx[c], gfr[c] = ni, nn, f, ix, raw
        gf[c] = np.array([float(" %15.8f " % v) for v in raw.ravel()]).reshape(raw.shape) \
            if raw.size else raw.copy()
    tab = IBTables(nintp, ninner, fcp, indx, gf, gfr)
    if les:
        tab.nzero, tab.iszero = find_zero_nu_sgs(body, XM, YM, ZM)
    return tab


def fcpts_text(tab: IBTables, comps=("u", "v", "w")) -> str:
    """ibmpre_fcpts.bin (lib_preprocessing.py:223-277) or, for comps=('t',),
    ibmpre_fcpts_t.bin (279-303)."""
    out = []
    if comps == ("t",):
        out.append("%d\n" % tab.nintp["t"])
        out.append("%d\n" % tab.ninner["t"])
    else:
        out.append("%d %d %d\n" % tuple(tab.nintp[c] for c in comps))
        out.append("%d %d %d\n" % tuple(tab.ninner[c] for c in comps))
    for c in comps:
        for row in tab.fcp[c]:
            out.append("%d %d %d\n" % tuple(row))
    for c in comps:
        for row in tab.indx[c]:
            out.append("%d %d %d\n" % tuple(row))
    for c in comps:
        raw = tab.geomfac_raw[c]
        for n in range(len(raw)):
            out.append("".join(" %15.8f " % v for v in raw[n].ravel()) + "\n")
    return "".join(out)


def nutzero_text(tab: IBTables) -> str:
    """ibmpre_nutzero.bin: 0-based indices in i -> j -> k (C) order
    (lib_preprocessing.py:306-330)."""
    iszero = tab.iszero
    ii, jj, kk = np.nonzero(iszero == 0)  # C order: i outer, k inner
    out = ["%d \n" % tab.nzero]
    out.append("".join(" %d " % v for v in ii) + "\n")
    out.append("".join(" %d " % v for v in jj) + "\n")
    out.append("".join(" %d " % v for v in kk) + "\n")
    return "".join(out)


def wallfdvm_text(tab: IBTables) -> str:
    """ibmpre_wallfdvm.bin: mask with i fastest (lib_preprocessing.py:333-338)."""
    return "".join(" %d " % v for v in tab.iszero.ravel(order="F"))


# ---------------------------------------------------------------------------
# settings.input / boundary.input (misc_init.f90:19-158)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Settings:
    re: float = 100.0
    pr: float = 0.71
    gr: float = 0.0
    grdir: int = 0
    tflu_i: float = 1.0
    tsol_i: float = 0.0
    ireset: int = 1
    iread: int = 1
    iavg: int = 0
    ipzero: int = 0
    eps_ptr: float = 0.0
    udrv_i: float = 1.0
    nend: int = 1
    nprint: int = 100000
    npriavg: int = 100000
    npin: int = 1
    tend: float = 1.0e6
    ptb_tst: float = 1.0e6
    avg_tst: float = 1.0e6
    idtopt: int = 0
    dt_size: float = 0.005
    cflfac: float = 0.5
    resid1: float = 1.0e-6
    nlev: int = 3
    nbli: int = 3
    ioldv: int = 0
    mgitr: int = 500
    imgsor: int = 1000
    wwsor: float = 1.8
    iles: int = 0
    insmdl: int = 1
    itemdl: int = 0
    idvmon: int = 1
    csgsts: float = 0.025
    csgshf: float = 0.003
    filter: int = 3
    ibmon: int = 1
    masson: int = 1
    imoving: int = 0
    ihtrans: int = 1
    idecomp: int = 0
    gridfile: str = "../output/grid/grid.bin"
    prev_fld: str = "../output/field/fld000000"
    # TRACE block: cells (i, j, k) traced every ntr steps (at most 64)
    trace: tuple = ()
    ntr: int = 1
    # boundary.input
    bc_ybtm: int = 1
    bc_ytop: int = 1
    bc_zbtm: int = 1
    bc_ztop: int = 1
    bc_t_ybtm: int = 1
    bc_t_ytop: int = 1
    val_t_ybtm: float = 0.0
    val_t_ytop: float = 0.0
    bc_t_zbtm: int = 1
    bc_t_ztop: int = 1
    val_t_zbtm: float = 0.0
    val_t_ztop: float = 0.0
    ich: int = 0
    iconjg: int = 0
    # the Poisson solver the reference links (02_solver/Makefile:23-34), not
    # in any input file: 0 poiss_z_ft, 1 poiss_mgrd
    poisson: int = 0
    # ibmpre_prdic.bin
    xprdic: int = 0
    yprdic: int = 0
    zprdic: int = 1
    iintp: int = 1

    def settings_text(self) -> str:
        g = lambda v: repr(float(v))  # noqa: E731  full-precision reals
        lines = [
            "RE#     PR#     GR#     GRDIR   TFLU_I  TSOL_I",
            f"{g(self.re)} {g(self.pr)} {g(self.gr)} {self.grdir} {g(self.tflu_i)} {g(self.tsol_i)}",
            "IRESET  IREAD   IAVG    IPZERO  EPS_PTR UDRV_I",
            f"{self.ireset} {self.iread} {self.iavg} {self.ipzero} {g(self.eps_ptr)} {g(self.udrv_i)}",
            "NEND    NPRINT  NPRIAVG NPIN",
            f"{self.nend} {self.nprint} {self.npriavg} {self.npin}",
            "TEND    PTB_TST AVG_TST",
            f"{g(self.tend)} {g(self.ptb_tst)} {g(self.avg_tst)}",
            "IDTOPT  DT_SIZE CFLFAC",
            f"{self.idtopt} {g(self.dt_size)} {g(self.cflfac)}",
            "RESID1  NLEV    NBLI    IOLDV   MGITR   IMGSOR  WWSOR",
            f"{g(self.resid1)} {self.nlev} {self.nbli} {self.ioldv} {self.mgitr} {self.imgsor} {g(self.wwsor)}",
            "ILES    INSMDL  ITEMDL  IDVMON  CSGSTS  CSGSHF  FILTER",
            f"{self.iles} {self.insmdl} {self.itemdl} {self.idvmon} {g(self.csgsts)} {g(self.csgshf)} {self.filter}",
            "IBMON   MASSON  IMOVING IHTRANS IDECOMP",
            f"{self.ibmon} {self.masson} {self.imoving} {self.ihtrans} {self.idecomp}",
            "FILES",
            f'"{self.gridfile}"',
            f'"{self.prev_fld}"',
            "TRACE   NTR",
            f"{len(self.trace)} {self.ntr}",
        ] + [f"{int(i)} {int(j)} {int(k)}" for i, j, k in self.trace] + [
            "",
        ]
        return "\n".join(lines)

    def boundary_text(self) -> str:
        lines = [
            "BC_YBTM    BC_YTOP", f"{self.bc_ybtm} {self.bc_ytop}",
            "BC_ZBTM    BC_ZTOP", f"{self.bc_zbtm} {self.bc_ztop}",
            "BC_T_YBTM  BC_T_YTOP", f"{self.bc_t_ybtm} {self.bc_t_ytop}",
            "VAL_T_YBTM VAL_T_YTOP", f"{repr(float(self.val_t_ybtm))} {repr(float(self.val_t_ytop))}",
            "BC_T_ZBTM  BC_T_ZTOP", f"{self.bc_t_zbtm} {self.bc_t_ztop}",
            "VAL_T_ZBTM VAL_T_ZTOP", f"{repr(float(self.val_t_zbtm))} {repr(float(self.val_t_ztop))}",
            "ICH        ICONJG", f"{self.ich} {self.iconjg}",
            "",
        ]
        return "\n".join(lines)


# ---------------------------------------------------------------------------
# fld files: Fortran unformatted sequential, 4-byte record markers,
# int64 integers and float64 reals (misc_hist.f90:236-247)
# ---------------------------------------------------------------------------


def _record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def write_fld(path: str, n: Tuple[int, int, int], re: float, pr: float, gr: float,
              ihist: int, m: int, time: float, dt: float, prdic, bcs, ich: int,
              iconjg: int, u, v, w, p, t=None) -> None:
    """u, v, w, p, t are full padded arrays indexed [i, j, k] over 0..n1 etc."""
    n1, n2, n3 = n
    n1m, n2m, n3m = n1 - 1, n2 - 1, n3 - 1
    with open(path, "wb") as f:
        f.write(_record(struct.pack("<3q3d", n1, n2, n3, re, pr, gr)))
        f.write(_record(struct.pack("<2q2d", ihist, m, time, dt)))
        f.write(_record(struct.pack("<3q", *prdic)))
        f.write(_record(struct.pack("<6q", *bcs)))
        f.write(_record(struct.pack("<2q", ich, iconjg)))
        f.write(_record(np.ascontiguousarray(u[1:n1 + 1, 0:n2 + 1, 0:n3 + 1].transpose(2, 1, 0)).tobytes()))
        f.write(_record(np.ascontiguousarray(v[0:n1 + 1, 1:n2 + 1, 0:n3 + 1].transpose(2, 1, 0)).tobytes()))
        f.write(_record(np.ascontiguousarray(w[0:n1 + 1, 0:n2 + 1, 1:n3 + 1].transpose(2, 1, 0)).tobytes()))
        f.write(_record(np.ascontiguousarray(p[1:n1m + 1, 1:n2m + 1, 1:n3m + 1].transpose(2, 1, 0)).tobytes()))
        if t is not None:
            f.write(_record(np.ascontiguousarray(t[1:n1m + 1, 1:n2m + 1, 1:n3m + 1].transpose(2, 1, 0)).tobytes()))


def read_fld(path: str, heat: bool = True) -> dict:
    """Reads a fld file into padded [i, j, k] arrays (unread entries zero)."""
    with open(path, "rb") as f:
        data = f.read()
    recs = []
    pos = 0
    while pos < len(data):
        (ln,) = struct.unpack_from("<i", data, pos)
        recs.append(data[pos + 4:pos + 4 + ln])
        (ln2,) = struct.unpack_from("<i", data, pos + 4 + ln)
        if ln2 != ln:
            raise ValueError("corrupt record marker in %s" % path)
        pos += 8 + ln
    n1, n2, n3, re, pr, gr = struct.unpack("<3q3d", recs[0])
    ihist, m, time, dt = struct.unpack("<2q2d", recs[1])
    prdic = struct.unpack("<3q", recs[2])
    bcs = struct.unpack("<6q", recs[3])
    ich, iconjg = struct.unpack("<2q", recs[4])
    n1m, n2m, n3m = n1 - 1, n2 - 1, n3 - 1
    shape = (n1 + 1, n2 + 1, n3 + 1)

    def arr(rec, lo, hi):
        out = np.zeros(shape)
        ext = [h - l + 1 for l, h in zip(lo, hi)]
        blk = np.frombuffer(rec, dtype="<f8").reshape(ext[2], ext[1], ext[0]).transpose(2, 1, 0)
        out[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = blk
        return out

    res = dict(n=(n1, n2, n3), re=re, pr=pr, gr=gr, ihist=ihist, m=m, time=time, dt=dt,
               prdic=prdic, bcs=bcs, ich=ich, iconjg=iconjg)
    res["u"] = arr(recs[5], (1, 0, 0), (n1, n2, n3))
    res["v"] = arr(recs[6], (0, 1, 0), (n1, n2, n3))
    res["w"] = arr(recs[7], (0, 0, 1), (n1, n2, n3))
    res["p"] = arr(recs[8], (1, 1, 1), (n1m, n2m, n3m))
    if heat and len(recs) > 9:
        res["t"] = arr(recs[9], (1, 1, 1), (n1m, n2m, n3m))
    return res


# ---------------------------------------------------------------------------
# whole case
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class CaseSpec:
    """A heated-body case: sizes, domain, body, physics and the seed of the
    initial-field noise."""

    n1m: int
    n2m: int
    n3m: int
    xlim: Tuple[float, float] = (-15.0, 30.0)
    ylim: Tuple[float, float] = (-15.0, 15.0)
    zlen: float = math.pi
    dmin: float = 0.01
    body: str = "cylinder"
    body_args: Tuple = ()
    noise: float = 1.0e-3
    seed: int = 20261203
    settings: Settings = dataclasses.field(default_factory=Settings)
    # "body": sinh-clustered x and y lines around the body, inflow/outflow in
    # x; "tanh": x and y lines clustered at the body surface (|x|, |y| = 0.5)
    # with spacing dmin there, hyperbolic-tangent segments (gridlines_tanh);
    # "channel": uniform lines, x and z periodic, the y ends the settings'
    # boundaries (walls with BC_Y = 0), the body (if it reaches the domain)
    # immersed as usual
    layout: str = "body"


@dataclasses.dataclass
class Case:
    spec: CaseSpec
    ax: Axis
    ay: Axis
    az: Axis
    lengths: Tuple[float, float, float]
    tables: IBTables
    fields: dict  # padded u, v, w, p, t as the initial fld holds them


def build_case(spec: CaseSpec) -> Case:
    s = spec.settings
    channel = spec.layout == "channel"
    if channel:
        xl = gridlines_uniform(spec.n1m, spec.xlim[0], spec.xlim[1])
        yl = gridlines_uniform(spec.n2m, spec.ylim[0], spec.ylim[1])
    elif spec.layout == "tanh":
        xl = gridlines_tanh(spec.n1m, spec.xlim[0], spec.xlim[1], spec.dmin)
        yl = gridlines_tanh(spec.n2m, spec.ylim[0], spec.ylim[1], spec.dmin)
    else:
        xl = gridlines_sinh(spec.n1m, spec.xlim[0], spec.xlim[1], spec.dmin)
        yl = gridlines_sinh(spec.n2m, spec.ylim[0], spec.ylim[1], spec.dmin)
    zl = gridlines_uniform(spec.n3m, 0.0, spec.zlen)
    xl, yl, zl = _fmt15(xl), _fmt15(yl), _fmt15(zl)
    lengths = tuple(float("%23.15f" % v) for v in (
        spec.xlim[1] - spec.xlim[0], spec.ylim[1] - spec.ylim[0], spec.zlen))
    ax = make_axis(xl, periodic=channel)
    ay = make_axis(yl, periodic=False)
    az = make_axis(zl, periodic=True)
    fn = BODIES[spec.body]
    body = (lambda x, y, z: fn(x, y, z, *spec.body_args)) if spec.body_args else fn
    tables = build_ib_tables(body, ax, ay, az, heat=bool(s.ihtrans), les=bool(s.iles))
    fields = initial_fields(spec, ax, ay, az, body)
    if channel:
        # no-slip walls at the y ends: the wall-tangential ghosts and the
        # wall-normal face velocities are zero (the solver keeps them)
        n2 = ay.ncell + 1
        for q in ("u", "w"):
            if s.bc_ybtm == 0:
                fields[q][:, 0, :] = 0.0
            if s.bc_ytop == 0:
                fields[q][:, n2, :] = 0.0
        if s.bc_ybtm == 0:
            fields["v"][:, 1, :] = 0.0
        if s.bc_ytop == 0:
            fields["v"][:, n2, :] = 0.0
    return Case(spec, ax, ay, az, lengths, tables, fields)


def solver_metrics(case: Case) -> dict:
    """The metrics as the solver holds them after its list-directed read of
    the ``%f`` text of grid.bin (misc_init.f90:204-251)."""
    txt = grid_bin_text(case.ax, case.ay, case.az, case.lengths)
    return parse_grid_bin(txt)


def parse_grid_bin(txt: str) -> dict:
    lines = txt.split("\n")
    tok = [ln.split() for ln in lines]
    n1, n2, n3 = (int(v) for v in tok[0])
    n1m, n2m, n3m = (int(v) for v in tok[1])
    xl, yl, zl = (float(v) for v in tok[2])
    r = 3
    out = dict(n1=n1, n2=n2, n3=n3, n1m=n1m, n2m=n2m, n3m=n3m, xl=xl, yl=yl, zl=zl)
    for name in ("x", "y", "z"):
        out[name] = np.array([float(v) for v in tok[r]]); r += 1
    for name in ("ipv", "jpv", "kpv", "imv", "jmv", "kmv", "fixil", "fixjl", "fixkl",
                 "fixiu", "fixju", "fixku"):
        out[name] = np.array([int(v) for v in tok[r]], dtype=np.int64); r += 1
    for name in ("c2cx", "c2cy", "c2cz", "c2cxi", "c2cyi", "c2czi", "f2fx", "f2fy", "f2fz",
                 "f2fxi", "f2fyi", "f2fzi", "xmp", "ymp", "zmp"):
        out[name] = np.array([float(v) for v in tok[r]]); r += 1
    return out


def initial_fields(spec: CaseSpec, ax: Axis, ay: Axis, az: Axis, body) -> dict:
    """Seeded initial field in the theta = 1 - T form: u = udrv in fluid
    (funcbody >= 1e-10 at the u points) plus uniform noise of amplitude
    ``spec.noise`` on interior u, v, w; p = 0; theta = tflu in fluid, tsol in
    solid. The arrays are padded [i, j, k] over 0..n; entries a fld file does
    not carry stay zero."""
    s = spec.settings
    n1, n2, n3 = ax.ncell + 1, ay.ncell + 1, az.ncell + 1
    n1m, n2m, n3m = n1 - 1, n2 - 1, n3 - 1
    shape = (n1 + 1, n2 + 1, n3 + 1)
    xg = ax.coord  # 0..n1 (x(0) duplicates x(1))
    X, Y, Z = np.meshgrid(xg, ay.center, az.center, indexing="ij")
    u = np.where(body(X, Y, Z) >= 1.0e-10, s.udrv_i, 0.0)
    v = np.zeros(shape)
    w = np.zeros(shape)
    rng = np.random.default_rng(spec.seed)
    if spec.noise:
        amp = spec.noise
        u[2:n1m + 1, 1:n2m + 1, 1:n3m + 1] += rng.uniform(-amp, amp, (n1m - 1, n2m, n3m))
        v[1:n1m + 1, 2:n2m + 1, 1:n3m + 1] += rng.uniform(-amp, amp, (n1m, n2m - 1, n3m))
        w[1:n1m + 1, 1:n2m + 1, 1:n3m + 1] += rng.uniform(-amp, amp, (n1m, n2m, n3m))
        # zero the noise inside the body so the solid starts at rest
        Xc, Yc, Zc = np.meshgrid(ax.center, ay.coord, az.center, indexing="ij")
        v[body(Xc, Yc, Zc) < 1.0e-10] = 0.0
        Xw, Yw, Zw = np.meshgrid(ax.center, ay.center, az.coord, indexing="ij")
        w[body(Xw, Yw, Zw) < 1.0e-10] = 0.0
        u[body(X, Y, Z) < 1.0e-10] = 0.0
    u[0] = 0.0
    p = np.zeros(shape)
    Xt, Yt, Zt = np.meshgrid(ax.center, ay.center, az.center, indexing="ij")
    t = np.where(body(Xt, Yt, Zt) >= 1.0e-10, s.tflu_i, s.tsol_i)
    return dict(u=u, v=v, w=w, p=p, t=t)


def write_case(case: Case, root: str) -> str:
    """Writes the reference directory layout under ``root``:
    root/02_solver/{settings,boundary}.input and root/output/... .
    Returns the solver working directory (root/02_solver)."""
    s = case.spec.settings
    sol = os.path.join(root, "02_solver")
    for d in ("02_solver", "output/grid", "output/ibmpre", "output/field",
              "output/field_avg", "output/ftr"):
        os.makedirs(os.path.join(root, d), exist_ok=True)
    with open(os.path.join(sol, "settings.input"), "w") as f:
        f.write(s.settings_text())
    with open(os.path.join(sol, "boundary.input"), "w") as f:
        f.write(s.boundary_text())
    with open(os.path.join(root, "output/grid/grid.bin"), "w") as f:
        f.write(grid_bin_text(case.ax, case.ay, case.az, case.lengths))
    with open(os.path.join(root, "output/ibmpre/ibmpre_prdic.bin"), "w") as f:
        f.write("%d %d %d %d" % (case_prdic(case.spec) + (1,)))
    with open(os.path.join(root, "output/ibmpre/ibmpre_fcpts.bin"), "w") as f:
        f.write(fcpts_text(case.tables))
    if s.ihtrans:
        with open(os.path.join(root, "output/ibmpre/ibmpre_fcpts_t.bin"), "w") as f:
            f.write(fcpts_text(case.tables, comps=("t",)))
    with open(os.path.join(root, "output/ibmpre/ibmpre_nutzero.bin"), "w") as f:
        if s.iles:
            f.write(nutzero_text(case.tables))
        else:
            f.write("0\n\n\n\n")
    if s.iles:
        with open(os.path.join(root, "output/ibmpre/ibmpre_wallfdvm.bin"), "w") as f:
            f.write(wallfdvm_text(case.tables))
    n = (case.ax.ncell + 1, case.ay.ncell + 1, case.az.ncell + 1)
    fl = case.fields
    xp = case_prdic(case.spec)[0]
    write_fld(os.path.join(root, "output/field/fld000000"), n, s.re, s.pr, s.gr,
              0, 0, 0.0, s.dt_size, case_prdic(case.spec), (4 if xp else 2, 4 if xp else 3, s.bc_ybtm, s.bc_ytop, 4, 4),
              s.ich, s.iconjg, fl["u"], fl["v"], fl["w"], fl["p"],
              fl["t"] if s.ihtrans else None)
    return sol


# ---------------------------------------------------------------------------
# readers: the case directory as the solver sees it (misc_init.f90:19-786)
# ---------------------------------------------------------------------------


class ListDirected:
    """Fortran list-directed READ semantics for the inputs these files use:
    every READ starts on a new record (line), consumes blank/comma separated
    values across lines until satisfied, then drops the rest of the line;
    quoted strings are one value."""

    def __init__(self, text: str):
        self.lines = text.split("\n")
        self.pos = 0

    @staticmethod
    def _tokens(line: str):
        out, cur, q = [], "", None
        for ch in line:
            if q:
                if ch == q:
                    out.append(cur)
                    cur, q = "", None
                else:
                    cur += ch
            elif ch in "\"'":
                q = ch
            elif ch in " \t,":
                if cur:
                    out.append(cur)
                    cur = ""
            else:
                cur += ch
        if cur:
            out.append(cur)
        return out

    def read(self, n: int):
        vals = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise EOFError("list-directed read past end of file")
            vals.extend(self._tokens(self.lines[self.pos]))
            self.pos += 1
        return vals[:n]


def parse_settings(text: str, boundary: str, prdic: str) -> Settings:
    s = Settings()
    r = ListDirected(text)
    r.read(1)
    v = r.read(6); s.re, s.pr, s.gr, s.grdir, s.tflu_i, s.tsol_i = float(v[0]), float(v[1]), float(v[2]), int(v[3]), float(v[4]), float(v[5])
    r.read(1)
    v = r.read(6); s.ireset, s.iread, s.iavg, s.ipzero = (int(x) for x in v[:4]); s.eps_ptr, s.udrv_i = float(v[4]), float(v[5])
    r.read(1)
    v = r.read(4); s.nend, s.nprint, s.npriavg, s.npin = (int(x) for x in v)
    r.read(1)
    v = r.read(3); s.tend, s.ptb_tst, s.avg_tst = (float(x) for x in v)
    r.read(1)
    v = r.read(3); s.idtopt = int(v[0]); s.dt_size = float(v[1]); s.cflfac = float(v[2])
    r.read(1)
    v = r.read(7); s.resid1 = float(v[0]); s.nlev, s.nbli, s.ioldv, s.mgitr, s.imgsor = (int(x) for x in v[1:6]); s.wwsor = float(v[6])
    r.read(1)
    v = r.read(7); s.iles, s.insmdl, s.itemdl, s.idvmon = (int(x) for x in v[:4]); s.csgsts, s.csgshf = float(v[4]), float(v[5]); s.filter = int(v[6])
    r.read(1)
    v = r.read(5); s.ibmon, s.masson, s.imoving, s.ihtrans, s.idecomp = (int(x) for x in v)
    r.read(1)
    s.gridfile = r.read(1)[0][:25]
    s.prev_fld = r.read(1)[0][:25]
    r.read(1)
    v = r.read(2); ntrace, s.ntr = int(v[0]), int(v[1])
    s.trace = tuple(tuple(int(x) for x in r.read(3)) for _ in range(ntrace))
    b = ListDirected(boundary)
    b.read(1); s.bc_ybtm, s.bc_ytop = (int(x) for x in b.read(2))
    b.read(1); s.bc_zbtm, s.bc_ztop = (int(x) for x in b.read(2))
    b.read(1); s.bc_t_ybtm, s.bc_t_ytop = (int(x) for x in b.read(2))
    b.read(1); s.val_t_ybtm, s.val_t_ytop = (float(x) for x in b.read(2))
    b.read(1); s.bc_t_zbtm, s.bc_t_ztop = (int(x) for x in b.read(2))
    b.read(1); s.val_t_zbtm, s.val_t_ztop = (float(x) for x in b.read(2))
    b.read(1); s.ich, s.iconjg = (int(x) for x in b.read(2))
    p = ListDirected(prdic).read(4)
    s.xprdic, s.yprdic, s.zprdic, s.iintp = (int(x) for x in p)
    # the Poisson solver, as lw_create_from_dir picks it: an x-periodic grid
    # takes poiss_ftft (the reference Makefile's default) or, with
    # LW_POISSON = x_ft, poiss_x_ft; LW_POISSON = mgrd names poiss_mgrd
    env = os.environ.get("LW_POISSON", "")
    if s.xprdic == 1:
        s.poisson = 3 if env == "x_ft" else 2
    elif env == "mgrd":
        s.poisson = 1
    return s


def parse_fcpts(text: str, comps) -> dict:
    """ibmpre_fcpts[_t].bin as ibmpreread reads it (misc_init.f90:312-378)."""
    r = ListDirected(text)
    nc = len(comps)
    nintp = [int(x) for x in r.read(nc)]
    ninner = [int(x) for x in r.read(nc)]
    out = {}
    for c, ni, nn in zip(comps, nintp, ninner):
        out[c] = dict(nintp=ni, nbody=ni + nn)
    for c in comps:
        nb = out[c]["nbody"]
        out[c]["fcp"] = np.array([[int(x) for x in r.read(3)] for _ in range(nb)], dtype=np.int64).reshape(nb, 3)
    for c in comps:
        ni = out[c]["nintp"]
        out[c]["indx"] = np.array([[int(x) for x in r.read(3)] for _ in range(ni)], dtype=np.int64).reshape(ni, 3)
    for c in comps:
        ni = out[c]["nintp"]
        out[c]["geomfac"] = np.array([[float(x) for x in r.read(27)] for _ in range(ni)]).reshape(ni, 3, 3, 3)
    return out


def load_case_dir(root: str) -> dict:
    """Everything the solver reads at start-up from ``root`` (the parent of
    02_solver/ and output/), after prefld's ghost-cell rules
    (misc_init.f90:610-786). Field arrays are padded [i, j, k]."""
    sol = os.path.join(root, "02_solver")
    rd = lambda p: open(p).read()  # noqa: E731
    s = parse_settings(rd(os.path.join(sol, "settings.input")),
                       rd(os.path.join(sol, "boundary.input")),
                       rd(os.path.join(root, "output/ibmpre/ibmpre_prdic.bin")))
    g = parse_grid_bin(rd(os.path.normpath(os.path.join(sol, s.gridfile))))
    comps = ("u", "v", "w")
    tab = parse_fcpts(rd(os.path.join(root, "output/ibmpre/ibmpre_fcpts.bin")), comps)
    if s.ihtrans:
        tab.update(parse_fcpts(rd(os.path.join(root, "output/ibmpre/ibmpre_fcpts_t.bin")), ("t",)))
    nz = dict(nzero=0, inz=np.zeros(0, np.int64), jnz=np.zeros(0, np.int64), knz=np.zeros(0, np.int64),
              nwall=np.ones((g["n1m"], g["n2m"], g["n3m"]), np.int64))
    if s.iles:
        r = ListDirected(rd(os.path.join(root, "output/ibmpre/ibmpre_nutzero.bin")))
        n = int(r.read(1)[0])
        nz["nzero"] = n
        if n > 0:
            nz["inz"] = np.array([int(x) for x in r.read(n)], np.int64)
            nz["jnz"] = np.array([int(x) for x in r.read(n)], np.int64)
            nz["knz"] = np.array([int(x) for x in r.read(n)], np.int64)
        mask = np.array(open(os.path.join(root, "output/ibmpre/ibmpre_wallfdvm.bin")).read().split(), np.int64)
        nz["nwall"] = mask.reshape(g["n3m"], g["n2m"], g["n1m"]).transpose(2, 1, 0).copy()
    if s.iread == 0:
        # makefld with the reference's compiled-in body, the cylinder
        return dict(settings=s, grid=g, tables=tab, nutzero=nz, fields=makefld(s, g, body_cylinder),
                    ihist=0, m=0, time=0.0, dt=s.dt_size)
    fl = read_fld(os.path.normpath(os.path.join(sol, s.prev_fld)), heat=bool(s.ihtrans))
    if fl["n"] != (g["n1"], g["n2"], g["n3"]):
        raise ValueError("PREVIOUS FIELD DOES NOT MATCH WITH THE GRID SYSTEM.")
    fields = apply_prefld(fl, s, g)
    return dict(settings=s, grid=g, tables=tab, nutzero=nz, fields=fields,
                ihist=fl["ihist"], m=fl["m"], time=fl["time"], dt=fl["dt"])


def makefld(s: Settings, g: dict, body: Callable, body_args: Tuple = ()) -> dict:
    """The initial field of IREAD = 0 (makefld, misc_init.f90:443-607): u =
    the ICH's profile (UDRV_I; 1.5 UDRV_I (1 - y^2); 18 (1 - y^2)) where the
    level set at the u point is >= 1e-10, v = w =
    p = 0, u = 0 on the rows next to wall boundaries; T = TSOL_I, TFLU_I at
    fluid cell centres of rows 1..n2m, the inlet ghost column by the level
    set, y (and z) ghost rows by the thermal BCs. The u loop reads x(0), which
    the reference never allocated (x(1:n1), mod_common.f90:124); x(0) = 0 is
    taken here, and the u column i = 0 is read by no routine of the step."""
    n1, n2, n3, n2m, n3m = g["n1"], g["n2"], g["n3"], g["n2m"], g["n3m"]
    shape = (n1 + 1, n2 + 1, n3 + 1)
    x0 = np.concatenate([[0.0], g["x"]])  # Fortran x(0..n1)
    xu, yc, zc = np.meshgrid(x0, g["ymp"], g["zmp"], indexing="ij")
    phi = body(xu, yc, zc, *body_args)
    # the profile by ICH (misc_init.f90:471-481)
    if s.ich == 1:
        prof = 1.5 * s.udrv_i * (1.0 - yc * yc)
    elif s.ich == 2:
        prof = (1.5 * (1.0 - yc * yc)) * 12.0
    else:
        prof = np.full_like(yc, float(s.udrv_i))
    u = np.where(phi >= 1.e-10, prof, 0.0)
    if s.bc_ybtm == 0:
        u[1:, 0:2, :] = 0.0
    if s.bc_ytop == 0:
        u[1:, n2m, :] = 0.0
        u[1:, n2, :] = 0.0
    if s.bc_zbtm == 0:
        u[1:, :, 0:2] = 0.0
    if s.bc_ztop == 0:
        u[1:, :, n3m] = 0.0
        u[1:, :, n3] = 0.0
    out = dict(u=u, v=np.zeros(shape), w=np.zeros(shape), p=np.zeros(shape))
    if not s.ihtrans:
        return out
    t = np.full(shape, float(s.tsol_i))
    xc, yc, zc = np.meshgrid(g["xmp"], g["ymp"], g["zmp"], indexing="ij")
    fluid = body(xc, yc, zc, *body_args) >= 1.e-10
    inner = fluid[:, 1:n2m + 1, :]
    t[:, 1:n2m + 1, :] = np.where(inner, s.tflu_i, t[:, 1:n2m + 1, :])
    if s.xprdic == 0:
        t[0] = np.where(fluid[0], s.tflu_i, s.tsol_i)
    if s.yprdic == 0:
        if s.bc_t_ybtm == 0:
            t[:, 0, :] = s.val_t_ybtm
        elif s.bc_t_ybtm == 1:
            t[:, 0, :] = t[:, 1, :] - s.val_t_ybtm / g["c2cyi"][1]
        if s.bc_t_ytop == 0:
            t[:, n2, :] = s.val_t_ytop
        elif s.bc_t_ytop == 1:
            t[:, n2, :] = t[:, n2m, :] + s.val_t_ytop / g["c2cyi"][n2]
    if s.zprdic == 0:
        if s.bc_t_zbtm == 0:
            t[:, :, 0] = s.val_t_zbtm
        elif s.bc_t_zbtm == 1:
            t[:, :, 0] = t[:, :, 1] - s.val_t_zbtm / g["c2czi"][1]
        if s.bc_t_ztop == 0:
            t[:, :, n3] = s.val_t_ztop
        elif s.bc_t_ztop == 1:
            t[:, :, n3] = t[:, :, n3m] + s.val_t_ztop / g["c2czi"][n3]
    out["t"] = t
    return out


def apply_prefld(fl: dict, s: Settings, g: dict) -> dict:
    """Ghost-cell rules of prefld after the read (misc_init.f90:631-786):
    T starts as TSOL_I everywhere before its interior is read; the file's
    prdic flags and BC record override the inputs; periodic ghosts copied."""
    n1, n2, n3 = g["n1"], g["n2"], g["n3"]
    n1m, n2m, n3m = n1 - 1, n2 - 1, n3 - 1
    u, v, w, p = (fl[k].copy() for k in ("u", "v", "w", "p"))
    t = None
    if s.ihtrans:
        t = np.full_like(u, s.tsol_i)
        t[1:n1m + 1, 1:n2m + 1, 1:n3m + 1] = fl["t"][1:n1m + 1, 1:n2m + 1, 1:n3m + 1]
    if s.ipzero == 1:
        p[:] = 0.0
    xp, yp, zp = fl["prdic"]
    if xp == 1:
        u[0, :, :] = u[n1m, :, :]; u[n1, :, :] = u[1, :, :]
        v[0, 1:, :] = v[n1m, 1:, :]; v[n1, 1:, :] = v[1, 1:, :]
        w[0, :, 1:] = w[n1m, :, 1:]; w[n1, :, 1:] = w[1, :, 1:]
        if t is not None:
            t[0] = t[n1m]; t[n1] = t[1]
    if yp == 1:
        u[1:, 0, :] = u[1:, n2m, :]; u[1:, n2, :] = u[1:, 1, :]
        v[:, 0, :] = v[:, n2m, :]; v[:, n2, :] = v[:, 1, :]
        w[:, 0, 1:] = w[:, n2m, 1:]; w[:, n2, 1:] = w[:, 1, 1:]
        if t is not None:
            t[:, 0, :] = t[:, n2m, :]; t[:, n2, :] = t[:, 1, :]
    if zp == 1:
        u[1:, :, 0] = u[1:, :, n3m]; u[1:, :, n3] = u[1:, :, 1]
        v[:, 1:, 0] = v[:, 1:, n3m]; v[:, 1:, n3] = v[:, 1:, 1]
        w[:, :, 0] = w[:, :, n3m]; w[:, :, n3] = w[:, :, 1]
        if t is not None:
            t[:, :, 0] = t[:, :, n3m]; t[:, :, n3] = t[:, :, 1]
    return dict(u=u, v=v, w=w, p=p, t=t)


def case_prdic(spec: CaseSpec) -> Tuple[int, int, int]:
    """The periodicity flags of ibmpre_prdic.bin: z always, x for a channel."""
    return (1 if spec.layout == "channel" else 0, 0, 1)


def case_view(case: Case) -> dict:
    """The same dictionary :func:`load_case_dir` returns for the directory
    :func:`write_case` would write, built in memory (the values pass through
    the same text formatting, so they are bit-identical)."""
    s = case.spec.settings
    g = solver_metrics(case)
    tabs = {}
    comps = ("u", "v", "w") + (("t",) if s.ihtrans else ())
    for c in comps:
        t = case.tables
        tabs[c] = dict(nintp=t.nintp[c], nbody=t.nintp[c] + t.ninner[c], fcp=t.fcp[c],
                       indx=t.indx[c], geomfac=t.geomfac[c])
    nz = dict(nzero=0, inz=np.zeros(0, np.int64), jnz=np.zeros(0, np.int64), knz=np.zeros(0, np.int64),
              nwall=np.ones((g["n1m"], g["n2m"], g["n3m"]), np.int64))
    if s.iles:
        ii, jj, kk = np.nonzero(case.tables.iszero == 0)
        nz.update(nzero=case.tables.nzero, inz=ii.astype(np.int64), jnz=jj.astype(np.int64),
                  knz=kk.astype(np.int64), nwall=case.tables.iszero.copy())
    n1, n2, n3 = g["n1"], g["n2"], g["n3"]
    xp, yp, zp = case_prdic(case.spec)
    fl = dict(prdic=(xp, yp, zp))
    f = case.fields
    u = f["u"].copy(); u[0] = 0.0
    v = f["v"].copy(); v[:, 0, :] = 0.0
    w = f["w"].copy(); w[:, :, 0] = 0.0
    p = np.zeros_like(u)
    p[1:n1, 1:n2, 1:n3] = f["p"][1:n1, 1:n2, 1:n3]
    fl.update(u=u, v=v, w=w, p=p)
    if s.ihtrans:
        fl["t"] = f["t"]
    # an x-periodic grid takes poiss_ftft unless poiss_x_ft (3) is named
    st = dataclasses.replace(s, xprdic=xp, yprdic=yp, zprdic=zp, iintp=1,
                             poisson=(s.poisson if s.poisson == 3 else 2) if xp == 1 else s.poisson)
    if s.iread == 0:
        fields = makefld(st, g, BODIES[case.spec.body], case.spec.body_args)
    else:
        fields = apply_prefld(fl, st, g)
    return dict(settings=st, grid=g, tables=tabs, nutzero=nz, fields=fields,
                ihist=0, m=0, time=0.0, dt=s.dt_size,
                body=body_params(case.spec.body, case.spec.body_args))
