"""The oracle (CPU restatement) pinned against the reference's own outputs
(golden fixtures written by the amdflang-built reference solver) and against
closed forms and library routines for its building blocks."""
import numpy as np
import pytest

from leswht_amd import case as C
from tests.helpers import fld_regions, golden, golden_names, rel, spec_for


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_golden(name, oracle_lib):
    import oracle
    z, meta = golden(name)
    spec = spec_for(tuple(meta["dims"]), meta["mode"], meta["nend"], meta.get("nlev"))
    view = C.case_view(C.build_case(spec))
    o = oracle.Oracle(view, poisson=meta.get("poisson", "z_ft"))
    o.step(meta["nend"])
    f = o.fields()
    n = (view["grid"]["n1"], view["grid"]["n2"], view["grid"]["n3"])
    # the oracle restates the reference operation by operation; only libm
    # calls and the FFT's radix split differ, so a few steps stay within
    # 1e-11 (1e-10 over 100 steps, where the bar for 100 steps is 1e-6)
    # In the x-periodic channels p holds only the small fluctuations (the
    # mean gradient is pmi), so rounding of the two FFTs weighs more: 1e-10.
    bar = 1e-11 if meta["nend"] <= 10 and meta.get("poisson") not in ("ftft", "x_ft") else 1e-10
    for k, sl in fld_regions(n, z["t"].size > 0).items():
        err = rel(z[k][sl], f[k][sl])
        assert err <= bar, (name, k, err)
    if "cmfr" in z.files:
        # fcmfr.dat of the last step: time, qflux, ubulk, pmiavg, cmfravg(1:3)
        np.testing.assert_allclose(o.array("cmfr"), z["cmfr"][-1], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("name", [n for n in golden_names() if n.startswith("xft_")])
def test_oracle_x_ft_rows_match_reference(name, oracle_lib):
    """poiss_x_ft's mgftresiduemax.dat over every step: mg2d's rows for x
    wavenumbers 1 and 2 (cycle counts exact, the residual to 1e-3: it is the
    small difference of A phi and the right-hand side) and gsor2d's rows where
    a wavenumber reached the iteration limit (MGITR+1 sweeps)."""
    import oracle
    z, meta = golden(name)
    view = C.case_view(C.build_case(spec_for(tuple(meta["dims"]), meta["mode"], meta["nend"], meta.get("nlev"))))
    o = oracle.Oracle(view, poisson="x_ft")
    rows = []
    for _ in range(meta["nend"]):
        o.step(1)
        rows += o.xft_rows()
    ref = z["xres"]
    got = np.array([[r[0], r[1], r[2], r[3]] for r in rows])
    assert got.shape == (len(ref), 4), (got.shape, ref.shape)
    assert np.array_equal(got[:, :3], ref[:, [0, 2, 3]])
    np.testing.assert_allclose(got[:, 3], ref[:, 4], rtol=1e-3, atol=1e-14)
    # a mg2d row carries the time of the step it belongs to
    assert np.all(ref[ref[:, 0] == 0, 1] >= 0.0)


@pytest.mark.parametrize("name", [n for n in golden_names() if n.startswith("mgrd_")])
def test_oracle_bicgstab_logs_match_reference(name, oracle_lib):
    """poiss_mgrd's logs: per solve (ftrpoittr.dat: iterations and final
    residual * dtconst) over every step, per iteration of the last step
    (poiss_itr.dat: residual and the cell holding it). The oracle's sums run
    in the reference's order, so the rows agree to the digits printed."""
    import oracle
    z, meta = golden(name)
    view = C.case_view(C.build_case(spec_for(tuple(meta["dims"]), meta["mode"], meta["nend"], meta.get("nlev"))))
    o = oracle.Oracle(view, poisson="mgrd")
    solves = []
    for _ in range(meta["nend"]):
        o.step(1)
        it, sv = o.bicg_log()
        solves += sv
    ref = z["bicg_solve"]
    assert len(solves) == len(ref)
    assert [s[0] for s in solves] == [int(r[2]) for r in ref]
    np.testing.assert_allclose([s[1] for s in solves], ref[:, 3], rtol=1e-6)
    last = z["bicg_itr"][-len(it):]
    assert [r[1] for r in it] == [int(r[0]) for r in last]
    np.testing.assert_allclose([r[2] for r in it], last[:, 1], rtol=1e-6)
    assert [tuple(r[3:]) for r in it] == [tuple(int(v) for v in r[2:]) for r in last]


@pytest.mark.parametrize("name", ["cyl32_iread0_trace_4", "chan32_ptb_4", "chan32_ptb0_3"])
def test_makefld_matches_reference_initial_field(name):
    """IREAD = 0: the case builder's makefld gives the field the reference
    wrote as fld0000000 (misc_init.f90:443-607, solver.f90:50-53) bit for bit:
    the uniform inflow, and the channel profiles of ICH = 1 and 2 (written
    before addperturb changes the field)."""
    z, meta = golden(name)
    view = C.case_view(C.build_case(spec_for(tuple(meta["dims"]), meta["mode"], meta["nend"])))
    f = view["fields"]
    n = (view["grid"]["n1"], view["grid"]["n2"], view["grid"]["n3"])
    for k, sl in fld_regions(n, z["t"].size > 0).items():
        assert np.array_equal(z["init_" + k][sl], f[k][sl]), k


def test_oracle_addperturb_is_what_moves_the_channel(oracle_lib):
    """The PTB_TST golden pins addperturb: without the perturbation (EPS_PTR
    = 0) the oracle's w after 4 steps stays zero while the reference's is
    not, and with it the oracle matches the reference (the golden test)."""
    import dataclasses
    import oracle
    z, meta = golden("chan32_ptb_4")
    spec = spec_for(tuple(meta["dims"]), meta["mode"], meta["nend"])
    assert np.max(np.abs(z["w"])) > 1e-3
    spec0 = dataclasses.replace(spec, settings=dataclasses.replace(spec.settings, eps_ptr=0.0))
    o = oracle.Oracle(C.case_view(C.build_case(spec0)))
    o.step(meta["nend"])
    assert np.max(np.abs(o.fields()["w"])) < 1e-12


def test_case_inputs_unchanged_since_fixtures():
    """The case builder regenerates bit-identical inputs for the fixtures."""
    import hashlib
    import os
    import tempfile
    for name in golden_names():
        z, meta = golden(name)
        spec = spec_for(tuple(meta["dims"]), meta["mode"], meta["nend"], meta.get("nlev"))
        root = tempfile.mkdtemp()
        C.write_case(C.build_case(spec), root)
        h = hashlib.sha256()
        for rel_ in ("output/grid/grid.bin", "output/ibmpre/ibmpre_fcpts.bin", "output/ibmpre/ibmpre_fcpts_t.bin",
                     "output/ibmpre/ibmpre_nutzero.bin", "02_solver/settings.input", "output/field/fld000000"):
            p = os.path.join(root, rel_)
            if os.path.exists(p):
                h.update(open(p, "rb").read())
        assert h.hexdigest() == meta["input_sha256"], name


def _pd(a):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("n", [2, 8, 16, 64, 12])
def test_fft_conventions(n, oracle_lib):
    """FFTE conventions (zfft1d.f:41-83): forward unnormalised e^{-2 pi i jk/n},
    inverse = conj(FFT(conj(X))) / n; checked against numpy.fft."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    re, im = x.real.copy(), x.imag.copy()
    oracle_lib.ora_fft_forward(n, _pd(re), _pd(im))
    np.testing.assert_allclose(re + 1j * im, np.fft.fft(x), rtol=0, atol=1e-12 * n)
    oracle_lib.ora_fft_inverse(n, _pd(re), _pd(im))
    np.testing.assert_allclose(re + 1j * im, x, rtol=0, atol=1e-13 * n)


def test_thomas_against_banded_solve(oracle_lib):
    from scipy.linalg import solve_banded
    rng = np.random.default_rng(1)
    n = 37
    a = rng.uniform(-1, 0, n); c = rng.uniform(-1, 0, n); b = 2.5 + rng.uniform(0, 1, n)
    r = rng.standard_normal(n)
    ab = np.zeros((3, n)); ab[0, 1:] = c[:-1]; ab[1] = b; ab[2, :-1] = a[1:]
    want = solve_banded((1, 1), ab, r)
    got = r.copy()
    oracle_lib.ora_thomas(n, _pd(a), _pd(b), _pd(c), _pd(got))
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)


def test_periodic_thomas_against_dense_solve(oracle_lib):
    rng = np.random.default_rng(2)
    n = 16
    a = rng.uniform(-1, 0, n); c = rng.uniform(-1, 0, n); b = 2.5 + rng.uniform(0, 1, n)
    r = rng.standard_normal(n)
    A = np.diag(b)
    for j in range(n):
        A[j, (j - 1) % n] += a[j]
        A[j, (j + 1) % n] += c[j]
    want = np.linalg.solve(A, r)
    got = r.copy()
    oracle_lib.ora_thomas_periodic(n, _pd(a), _pd(b), _pd(c), _pd(got))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)


def test_oracle_rejects_out_of_scope(oracle_lib):
    import oracle
    spec = spec_for((32, 32, 8), "nosgs")
    view = C.case_view(C.build_case(spec))
    view["settings"].ich = 1
    with pytest.raises(ValueError):
        oracle.Oracle(view)
