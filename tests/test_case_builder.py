"""Case builder: metric formulas, IB tables, file formats (inputs of both the
reference and the HIP path)."""
import os
import tempfile

import numpy as np
import pytest

from leswht_amd import case as C
from tests.helpers import spec_for


def test_axis_metrics_match_setgrid_rules():
    lines = C.gridlines_sinh(24, -3.0, 6.0, 0.05)
    ax = C.make_axis(lines, periodic=False)
    n = ax.ncell
    assert np.allclose(ax.f2fd[1:n + 1], np.diff(lines))
    assert ax.c2cd[1] == 0.5 * ax.f2fd[1] and ax.c2cd[n + 1] == 0.5 * ax.f2fd[n]
    assert ax.lowerfix[0] == 1 and ax.upperfix[-1] == 1
    assert ax.idxminus[0] == 0 and ax.idxplus[-1] == n + 1
    az = C.make_axis(C.gridlines_uniform(8, 0.0, 1.0), periodic=True)
    assert az.idxminus[0] == 8 and az.idxplus[-1] == 1
    assert az.lowerfix.sum() == 0 and az.upperfix.sum() == 0
    assert az.f2fd[0] == az.f2fd[8] and az.f2fd[9] == az.f2fd[1]


def test_interpolation_points_brute_force():
    """Interpolation points = body points with an in/out change along a
    non-boundary axis, in k->j->i order (ibm_body.f90:82-110)."""
    spec = spec_for((24, 24, 4), "les")
    case = C.build_case(spec)
    ax, ay, az = case.ax, case.ay, case.az
    _, inout = C.find_inout(C.body_cylinder, ax.coord, ay.center, az.center)
    want = []
    nx, ny, nz = (s - 1 for s in inout.shape)
    for k in range(1, nz):
        for j in range(1, ny):
            for i in range(2, nx):
                if inout[i, j, k] != 0:
                    continue
                s = abs(inout[i + 1, j, k] - inout[i - 1, j, k]) * (1 - ax.upperfix[i - 1]) * (1 - ax.lowerfix[i - 1]) \
                    + abs(inout[i, j + 1, k] - inout[i, j - 1, k]) * (1 - ay.upperfix[j - 1]) * (1 - ay.lowerfix[j - 1]) \
                    + abs(inout[i, j, k + 1] - inout[i, j, k - 1]) * (1 - az.upperfix[k - 1]) * (1 - az.lowerfix[k - 1])
                if s > 0:
                    want.append((i, j, k))
    t = case.tables
    got = [tuple(r) for r in t.fcp["u"][:t.nintp["u"]]]
    assert got == want


def test_geomfac_reproduces_linear_field_through_surface():
    """For a field linear along the interpolation line and zero at the body
    surface, the forcing-point target equals the field at the forcing point
    (ibm_body.f90:250-317): sum over fluid points of geomfac * f = f(x1).
    The surface point is found here independently by bisection; the builder's
    20-step three-level search resolves it to (x2-x1)/8000."""
    spec = spec_for((40, 40, 4), "nosgs")
    case = C.build_case(spec)
    t = case.tables
    xs = {"u": case.ax.coord, "v": case.ax.center}
    ys = {"u": case.ay.center, "v": case.ay.coord}
    checked = 0
    for comp in ("u", "v"):
        g = t.geomfac_raw[comp]
        for n in range(t.nintp[comp]):
            ix = t.indx[comp][n]
            if abs(ix).sum() != 1:
                continue
            i, j, k = t.fcp[comp][n]
            ax_ = 0 if ix[0] != 0 else 1
            s = ix[ax_]
            pts = xs[comp] if ax_ == 0 else ys[comp]
            idx = (i, j)[ax_]
            p1, p2, p3 = pts[idx], pts[idx + s], pts[idx + 2 * s]
            other = ys[comp][j] if ax_ == 0 else xs[comp][i]

            def phi(q):
                return C.body_cylinder(q, other, 0.0) if ax_ == 0 else C.body_cylinder(other, q, 0.0)

            lo, hi = p1, p2
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if (phi(lo) <= 0) == (phi(mid) <= 0):
                    lo = mid
                else:
                    hi = mid
            x0 = 0.5 * (lo + hi)
            w1 = g[n][1, 0, 0] if ax_ == 0 else g[n][0, 1, 0]
            w2 = g[n][2, 0, 0] if ax_ == 0 else g[n][0, 2, 0]
            target = w1 * (p2 - x0) + w2 * (p3 - x0)
            assert abs(target - (p1 - x0)) <= 2e-3 * abs(p2 - p1), (comp, n, target, p1 - x0)
            checked += 1
    assert checked > 10


def test_text_round_trip_is_bit_exact():
    spec = spec_for((32, 32, 8), "les")
    case = C.build_case(spec)
    root = tempfile.mkdtemp()
    C.write_case(case, root)
    a = C.load_case_dir(root)
    b = C.case_view(case)
    for k in ("u", "v", "w", "p", "t"):
        assert np.array_equal(a["fields"][k], b["fields"][k])
    for c in ("u", "v", "w", "t"):
        for key in ("fcp", "indx", "geomfac"):
            assert np.array_equal(a["tables"][c][key], b["tables"][c][key])
    for key in ("inz", "jnz", "knz", "nwall"):
        assert np.array_equal(a["nutzero"][key], b["nutzero"][key])
    for key in ("f2fx", "c2cyi", "xmp", "zmp"):
        assert np.array_equal(a["grid"][key], b["grid"][key])


def test_fld_writer_reader_round_trip():
    rng = np.random.default_rng(0)
    n = (9, 7, 5)
    sh = (n[0] + 1, n[1] + 1, n[2] + 1)
    arrs = {k: rng.standard_normal(sh) for k in "uvwpt"}
    path = os.path.join(tempfile.mkdtemp(), "fld0000001")
    C.write_fld(path, n, 100.0, 0.71, 0.0, 3, 4, 0.5, 0.01, (0, 0, 1), (2, 3, 1, 1, 4, 4), 0, 0,
                arrs["u"], arrs["v"], arrs["w"], arrs["p"], arrs["t"])
    r = C.read_fld(path)
    assert r["n"] == n and r["ihist"] == 3 and r["m"] == 4 and r["time"] == 0.5
    assert np.array_equal(r["u"][1:, :, :], arrs["u"][1:, :, :])
    assert np.array_equal(r["p"][1:8, 1:6, 1:4], arrs["p"][1:8, 1:6, 1:4])
    # record markers are 4-byte little endian lengths
    raw = open(path, "rb").read()
    assert int.from_bytes(raw[:4], "little") == 48


@pytest.mark.parametrize("n,lo,hi", [(384, -15.0, 30.0), (256, -15.0, 15.0), (192, -15.0, 30.0), (128, -15.0, 15.0)])
def test_tanh_lines_are_clustered_at_the_surface(n, lo, hi):
    """The configurations' stretching: spacing dmin on both sides of the
    body surface x = -0.5 and x = 0.5, lines exactly at the ends and at the
    surface, strictly increasing, neighbouring spacings within 40 % (a
    plausible slip, such as a segment's points in the wrong order or a
    spacing matched on the wrong side, breaks one of these)."""
    x = C.gridlines_tanh(n, lo, hi, 0.01)
    assert len(x) == n + 1 and x[0] == lo and x[-1] == hi
    d = np.diff(x)
    assert (d > 0).all()
    for a in (-0.5, 0.5):
        i = int(np.flatnonzero(x == a)[0])
        assert abs(d[i - 1] - 0.01) < 1e-12 and abs(d[i] - 0.01) < 1e-12
    assert (d[1:] / d[:-1]).max() < 1.4 and (d[:-1] / d[1:]).max() < 1.4
    # the expansion segment follows the pre-processor's formula: with its
    # own g it is reproduced point for point
    i5 = int(np.flatnonzero(x == 0.5)[0])
    seg = x[i5:] - 0.5
    m, L = len(seg) - 1, hi - 0.5
    g = C._bisect(lambda g: L * (1.0 - np.tanh(g * (1.0 - 1.0 / m)) / np.tanh(g)) - 0.01, 1e-6, 30.0)
    p = np.arange(m + 1) / m
    ref = L * (1.0 - np.tanh(g * (1.0 - p)) / np.tanh(g))
    assert np.abs(seg[1:-1] - ref[1:-1]).max() < 1e-12
