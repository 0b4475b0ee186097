"""The basis stage on the CPU: the oracle's HEALPix pixelisation and cell
cubature pinned by identities, closed forms and the reference's own
adapt_integrate_v, and the basis program's command line and maskfile errors."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from opsec_amd import formats, lib, synth
from tests import oracles as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BASIS = os.path.join(ROOT, "opsec_amd", "bin", "basis")


def pix2ang(nside, p, nest):
    t, f = C.c_double(), C.c_double()
    (O.oracle().orc_pix2ang_nest if nest else O.oracle().orc_pix2ang_ring)(nside, int(p), C.byref(t), C.byref(f))
    return t.value, f.value


def ang2pix(nside, theta, phi, nest):
    fn = O.oracle().orc_ang2pix_nest if nest else O.oracle().orc_ang2pix_ring
    return fn(nside, float(theta), float(phi))


def random_dirs(n, seed):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, n)
    return np.arccos(z), rng.uniform(0, 2 * np.pi, n)


def test_healpix_base_pixels_known_centres():
    """nside 1: pixels 0-3 centred at z = 2/3, 4-7 on the equator, 8-11 at
    z = -2/3, at phi = pi/4 + k pi/2 in the caps and k pi/2 on the equator;
    RING and NESTED agree at nside 1."""
    for p in range(12):
        k = p % 4
        z, phi = ((2 / 3, np.pi / 4 + k * np.pi / 2) if p < 4 else
                  (0.0, k * np.pi / 2) if p < 8 else (-2 / 3, np.pi / 4 + k * np.pi / 2))
        for nest in (0, 1):
            t, f = pix2ang(1, p, nest)
            assert abs(np.cos(t) - z) < 1e-14 and abs(f - phi) < 1e-14, (p, nest)
            assert ang2pix(1, np.arccos(z), phi, nest) == p
    # nside 2: the first RING pixel (next to the north pole, phi = pi/4) is the
    # last pixel of NESTED face 0
    assert np.allclose(pix2ang(2, 0, 0), pix2ang(2, 3, 1))
    assert abs(np.cos(pix2ang(2, 0, 0)[0]) - (1 - 1 / 12)) < 1e-14


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16])
def test_healpix_pixel_centres_invert(nside):
    for nest in (0, 1):
        for p in range(12 * nside * nside):
            t, f = pix2ang(nside, p, nest)
            assert ang2pix(nside, t, f, nest) == p, (nest, p)


@pytest.mark.parametrize("nside", [2, 4, 8, 32])
def test_healpix_ring_and_nest_same_pixel(nside):
    th, ph = random_dirs(4000, nside)
    th = np.concatenate([th, [0.0, np.pi, np.arccos(2 / 3), np.arccos(-2 / 3), np.pi / 2]])
    ph = np.concatenate([ph, [0.0, 6.283185307179586, 1e-17, np.pi / 2, 2 * np.pi - 1e-12]])
    for t, f in zip(th, ph):
        a = pix2ang(nside, ang2pix(nside, t, f, 0), 0)
        b = pix2ang(nside, ang2pix(nside, t, f, 1), 1)
        assert abs(a[0] - b[0]) < 1e-12 and abs(np.cos(a[1] - b[1]) - 1) < 1e-12, (t, f)


def test_healpix_nested_hierarchy():
    th, ph = random_dirs(3000, 1)
    for nside in (2, 4, 16, 64, 1024):
        for t, f in zip(th, ph):
            assert ang2pix(nside, t, f, 1) >> 2 == ang2pix(nside // 2, t, f, 1)


def test_healpix_equal_areas_and_ring_order():
    nside, n = 4, 192000
    th, ph = random_dirs(n, 7)
    for nest in (0, 1):
        counts = np.bincount([ang2pix(nside, t, f, nest) for t, f in zip(th, ph)], minlength=192)
        exp = n / 192
        chi2 = ((counts - exp) ** 2 / exp).sum()
        assert counts.min() > 0 and chi2 < 191 + 5 * np.sqrt(2 * 191), chi2
    z = [np.cos(pix2ang(nside, p, 0)[0]) for p in range(192)]
    assert np.all(np.diff(z) <= 1e-15)


def _cell(lo, hi, sph=False):
    c = np.zeros(1, dtype=lib.CELL_DTYPE)
    c["min1"], c["max1"], c["min2"], c["max2"], c["min3"], c["max3"] = lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]
    return c


def test_cell_integral_uniform_is_nbar_volume():
    sel = O.Selection(0, nbar=3e-4)
    c = _cell((10.0, 20.0, 30.0), (22.5, 32.5, 42.5))
    ne = C.c_int64()
    N = O.oracle().orc_cell_integral(1, C.byref(sel.s), c.ctypes.data, 1e-5, 1e-10, 5000000, C.byref(ne))
    assert abs(N - 3e-4 * 12.5 ** 3) < 1e-15 * N + 1e-17 and ne.value == 33


def test_cell_integral_linear_radial_closed_form():
    """nbar(r) = a + b r on a shell cell: integral over (r^3, mu, phi) is
    dmu dphi [a (r2^3 - r1^3) + (3/4) b (r2^4 - r1^4)], to the 1e-5 absolute
    tolerance the reference's argument order gives."""
    a, b = 2e-4, -1e-7
    sel = O.Selection(1, rr=[0.0, 2000.0], rn=[a, a + b * 2000.0])
    for r1, r2, m1, m2, p1, p2 in ((300.0, 350.0, 0.5, 0.6, 0.0, 0.4), (100.0, 900.0, -0.2, 0.9, 1.0, 3.0)):
        c = _cell((r1, m1, p1), (r2, m2, p2))
        N = O.oracle().orc_cell_integral(2, C.byref(sel.s), c.ctypes.data, 1e-5, 1e-10, 5000000, None)
        exact = (m2 - m1) * (p2 - p1) * (a * (r2 ** 3 - r1 ** 3) + 0.75 * b * (r2 ** 4 - r1 ** 4))
        assert abs(N - exact) <= 1e-5 + 1e-12 * abs(exact), (N, exact)


def test_cell_integral_vs_reference():
    if not O.ref_available():
        pytest.skip("reference build not present")
    rng = np.random.default_rng(3)
    rr = np.linspace(150.0, 1100.0, 40)
    rn = 5e-4 * np.exp(-(rr / 600.0) ** 2) * (1 + 0.3 * rng.random(40))
    sel = O.Selection(1, rr=rr, rn=rn)
    cases = [(1, (0.0, 0.0, 0.0), (12.5, 12.5, 12.5)), (2, (200.0, 0.5, 0.0), (212.5, 0.515625, 0.19634954)),
             (2, (200.0, 0.5, 0.0), (1000.0, 1.0, 6.283185307179586)), (2, (140.0, 0.9, 1.0), (260.0, 0.95, 1.2))]
    for cs, lo, hi in cases:
        c = _cell(lo, hi)
        if cs == 1:
            s_o, kind = O.Selection(0, nbar=3e-4), 0
        else:
            s_o, kind = sel, 1
        N = O.oracle().orc_cell_integral(cs, C.byref(s_o.s), c.ctypes.data, 1e-5, 1e-10, 5000000, None)
        lo_u = np.array([lo[0] ** 3 if cs == 2 else lo[0], lo[1], lo[2]])
        hi_u = np.array([hi[0] ** 3 if cs == 2 else hi[0], hi[1], hi[2]])
        R = O.ref().ref_cell_integral(kind, 3e-4, len(rr), O.ptr(rr), O.ptr(rn), O.ptr(lo_u), O.ptr(hi_u), 1e-5, 1e-10,
                                      None)
        assert abs(N - R) <= 1e-14 * abs(R), (cs, lo, hi, N, R)


def test_oracle_basis_cells_and_drops():
    """Grid cells with Survey::CreateEmptyCell bounds; with a radial selection
    that is zero beyond r = 500 the outer shells are dropped and the rest
    renumbered; spherical Nbar = the r^3 integral / 3."""
    sv = lib.survey_sph(4, 2, 3, 200.0, 1000.0, 0.5, 1.0, 0.0, 2 * np.pi)
    sel = O.Selection(1, rr=[0.0, 500.0], rn=[1e-3, 1e-3])
    cells = O.basis(2, sv, sel)
    assert len(cells) == 2 * 2 * 3 and np.all(cells["a"] == np.arange(len(cells)))
    assert np.all(cells["d1"] < 2)
    re = np.array([200.0 + d * 800.0 / 4 for d in range(5)])
    np.testing.assert_array_equal(cells["min1"], re[cells["d1"]])
    V = (cells["max1"] ** 3 - cells["min1"] ** 3) / 3 * (cells["max2"] - cells["min2"]) * (cells["max3"] - cells["min3"])
    np.testing.assert_allclose(cells["Veff"], V, rtol=1e-15)
    # r < 400 for d1 = 0 (exact integral of the constant 1e-3) and 400..600 cut at 500 for d1 = 1
    inner = cells[cells["d1"] == 0]
    np.testing.assert_allclose(inner["Nbar"], 1e-3 * inner["Veff"], rtol=1e-12)
    box = O.basis(1, lib.survey_box(3, 0.0, 150.0), O.Selection(0, nbar=3e-4))
    np.testing.assert_allclose(box["Nbar"], 3e-4 * 50.0 ** 3, rtol=1e-14)


def run(*args, cwd=None):
    return subprocess.run([BASIS, *args], capture_output=True, text=True, cwd=cwd, timeout=120)


def test_basis_help_and_missing_keys(tmp_path):
    r = run("-h")
    assert r.returncode == 0 and "Usage:" in r.stdout
    r = run("coordsys=cartesian")
    assert r.returncode == 1 and "missing config option 'cellfile'" in r.stderr
    r = run(f"cellfile={tmp_path}/c.abn", "coordsys=polar")
    assert r.returncode == 1 and "invalid or missing config option: coordsys" in r.stderr


def test_basis_rejects_bad_maskfile(tmp_path):
    (tmp_path / "radial.dat").write_text("0 1e-3\n2000 1e-3\n")
    (tmp_path / "mask.fits").write_bytes(b"SIMPLE  =                    F" + b" " * 2849)
    args = [f"cellfile={tmp_path}/c.abn", "coordsys=spherical", "survey.type=SphericalSurvey", "survey.RMin=200",
            "survey.RMax=1000", "survey.MuMin=0.5", "survey.MuMax=1", "survey.PhiMin=0", "survey.PhiMax=6.283185307179586",
            "survey.Nr=2", "survey.Nmu=2", "survey.Nphi=2", f"survey.radial={tmp_path}/radial.dat"]
    r = run(*args, f"survey.maskfile={tmp_path}/mask.fits")
    assert r.returncode == 1 and "could not read HEALPix file" in r.stderr and "not a FITS file" in r.stderr
    # a FITS map whose ORDERING is neither RING nor NESTED
    synth.write_healpix_fits(tmp_path / "m2.fits", np.ones(12, np.float32))
    raw = (tmp_path / "m2.fits").read_bytes().replace(b"'RING    '", b"'SPIRAL  '")
    (tmp_path / "m2.fits").write_bytes(raw)
    r = run(*args, f"survey.maskfile={tmp_path}/m2.fits")
    assert r.returncode == 1 and "unrecognized HEALPix ordering: SPIRAL" in r.stderr
    # a map whose table is cut short, and one claiming an absurd NSIDE
    synth.write_healpix_fits(tmp_path / "m3.fits", np.ones(12 * 64 * 64, np.float32), per_row=1024)
    raw = (tmp_path / "m3.fits").read_bytes()
    (tmp_path / "m3.fits").write_bytes(raw[:2 * 2880 + 5000])
    r = run(*args, f"survey.maskfile={tmp_path}/m3.fits")
    assert r.returncode == 1 and "truncated table" in r.stderr
    raw = (tmp_path / "m2.fits").read_bytes().replace(b"NSIDE   =                    1", b"NSIDE   =           1073741824")
    (tmp_path / "m4.fits").write_bytes(raw)
    r = run(*args, f"survey.maskfile={tmp_path}/m4.fits")
    assert r.returncode == 1 and "bad NSIDE" in r.stderr
    r = run(*args[:-1], "survey.radial=compute", f"survey.maskfile={tmp_path}/m2.fits")
    assert r.returncode == 1 and "'radial = compute' not implemented" in r.stderr
    r = run(*args[:1], "coordsys=cartesian", *args[2:], f"survey.maskfile={tmp_path}/m2.fits")
    assert r.returncode == 1 and "does not match survey.type" in r.stderr
