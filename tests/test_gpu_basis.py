"""The basis stage on the GPU (opsec_basis, opsec_healpix_pixels and the basis
program) against the oracle."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from opsec_amd import formats, lib, synth
from tests import oracles as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "opsec_amd", "bin")


def cap_mask(nside, nest, mumin, seed=0):
    """Completeness 1 for pixel centres with mu >= mumin, 0 elsewhere, and a
    few random partial values inside, in RING or NESTED order."""
    rng = np.random.default_rng(seed)
    npix = 12 * nside * nside
    t, f = C.c_double(), C.c_double()
    fn = O.oracle().orc_pix2ang_nest if nest else O.oracle().orc_pix2ang_ring
    mu = np.empty(npix)
    for p in range(npix):
        fn(nside, p, C.byref(t), C.byref(f))
        mu[p] = np.cos(t.value)
    m = (mu >= mumin).astype(np.float32)
    inside = np.nonzero(m)[0]
    m[rng.choice(inside, size=min(len(inside), 20), replace=False)] = rng.uniform(0.3, 0.9, min(len(inside), 20))
    return m


def assert_nbar_close(cells, ref, nevals):
    """Nbar of converged cells to 1e-12; cells whose cubature stops at the
    5e6-evaluation cap to 1e-5: their error estimates stay above the
    tolerance (a mask edge or the kinks of the radial table inside the cell),
    regions of nearly equal error sit along the kinks, and one-ulp differences
    between the device's and the host's cbrt and acos can pick different
    regions to split before the cap."""
    capped = nevals >= 5000000
    print("cells at the evaluation cap:", int(capped.sum()), "of", len(ref))
    np.testing.assert_allclose(cells["Nbar"][~capped], ref["Nbar"][~capped], rtol=1e-12)
    np.testing.assert_allclose(cells["Nbar"][capped], ref["Nbar"][capped], rtol=1e-5)


def test_healpix_pixels_vs_oracle(ctx):
    rng = np.random.default_rng(2)
    n = 20000
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(-1, 7.5, n)   # includes phi < 0 and > 2 pi
    edge_t = np.arccos(np.array([1.0, -1.0, 2 / 3, -2 / 3, 0.0, np.nextafter(2 / 3, 1), np.nextafter(-2 / 3, -1)]))
    th = np.concatenate([th, edge_t, edge_t])
    ph = np.concatenate([ph, np.zeros(7), np.full(7, 2 * np.pi)])
    for nside in (1, 2, 16, 1024, 1 << 20):
        for nest in (0, 1):
            got = ctx.healpix_pixels(nside, nest, th, ph)
            fn = O.oracle().orc_ang2pix_nest if nest else O.oracle().orc_ang2pix_ring
            ref = np.array([fn(nside, float(a), float(b)) for a, b in zip(th, ph)])
            np.testing.assert_array_equal(got, ref, err_msg=f"nside {nside} nest {nest}")
            assert got.min() >= 0 and got.max() < 12 * nside * nside


def test_basis_box_uniform(ctx):
    sv = lib.survey_box(6, 0.0, 150.0)
    cells, nev = ctx.basis(sv, lib.OPSEC_SEL_UNIFORM, nbar=3e-4)
    ref = O.basis(1, sv, O.Selection(0, nbar=3e-4))
    assert len(cells) == 216 and nev == 216 * 33
    assert cells.tobytes() == ref.tobytes()
    np.testing.assert_allclose(cells["Nbar"], 3e-4 * 25.0 ** 3, rtol=1e-14)
    np.testing.assert_array_equal(cells[["a", "d1", "d2", "d3"]], synth.box_cells(6, 0.0, 150.0)[["a", "d1", "d2", "d3"]])


def test_basis_spherical_smooth_converges(ctx):
    """A linear radial profile without a mask: every cell converges (a few
    hundred evaluations) and Nbar agrees with the oracle to 1e-12."""
    sv = lib.survey_sph(4, 3, 5, 200.0, 1000.0, 0.4, 1.0, 0.0, 2 * np.pi)
    cells, nev = ctx.basis(sv, lib.OPSEC_SEL_RADIAL, rad_r=[100.0, 1200.0], rad_n=[6e-4, 1e-4])
    sel = O.Selection(1, rr=[100.0, 1200.0], rn=[6e-4, 1e-4])
    ref = O.basis(2, sv, sel)
    assert len(cells) == 60 and nev < 60 * 1000
    np.testing.assert_array_equal(cells["Veff"], ref["Veff"])
    np.testing.assert_allclose(cells["Nbar"], ref["Nbar"], rtol=1e-12)


@pytest.mark.parametrize("nest", [0, 1])
def test_basis_spherical_radial_and_mask(ctx, nest):
    """Radial spline x HEALPix cap mask with partial pixels: cells crossing the
    mask edge need many adaptive regions; cells outside the cap or beyond the
    radial table are dropped."""
    rr = np.linspace(180.0, 900.0, 25)
    rn = 5e-4 * np.exp(-(rr / 600.0) ** 2)
    mask = cap_mask(8, nest, 0.55, seed=nest)
    sv = lib.survey_sph(5, 4, 6, 200.0, 1000.0, 0.4, 1.0, 0.0, 2 * np.pi)
    cells, nev = ctx.basis(sv, lib.OPSEC_SEL_RADIAL, rad_r=rr, rad_n=rn, mask=mask, nest=nest)
    sel = O.Selection(1, rr=rr, rn=rn, mask=mask, nside=8, nest=nest)
    ref, ne = O.basis(2, sv, sel, with_evals=True)
    print("spherical basis: kept", len(cells), "of", 5 * 4 * 6, "integrand evaluations", nev)
    assert len(cells) == len(ref) and 0 < len(cells) < 120
    for k in ("a", "d1", "d2", "d3", "min1", "max1", "min2", "max2", "min3", "max3", "Veff"):
        np.testing.assert_array_equal(cells[k], ref[k], err_msg=k)
    assert_nbar_close(cells, ref, ne)
    assert nev > len(cells) * 33   # the mask edge forces subdivisions


def test_basis_program_end_to_end(tmp_path):
    """basis writes the cellfile from a SphericalSurvey configuration with a
    FITS maskfile and a radial table; the cells equal the oracle's and klt
    accepts the file."""
    rr = np.linspace(150.0, 1100.0, 30)
    rn = 5e-4 * np.exp(-(rr / 600.0) ** 2)
    np.savetxt(tmp_path / "radial.dat", np.c_[rr, rn], fmt="%.17g", header="r nbar")
    mask = cap_mask(4, 1, 0.5, seed=3)
    synth.write_healpix_fits(tmp_path / "mask.fits", mask, nest=True, per_row=48)
    nr, nmu, nphi = 3, 3, 4
    args = [f"cellfile={tmp_path}/cells.abn", "coordsys=spherical", "survey.type=SphericalSurvey",
            "survey.RMin=200", "survey.RMax=1000", "survey.MuMin=0.4", "survey.MuMax=1",
            "survey.PhiMin=0", "survey.PhiMax=6.283185307179586", f"survey.Nr={nr}", f"survey.Nmu={nmu}",
            f"survey.Nphi={nphi}", f"survey.radial={tmp_path}/radial.dat", f"survey.maskfile={tmp_path}/mask.fits",
            "survey.galfile=gals.abn"]
    r = subprocess.run([os.path.join(BIN, "basis"), *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "non-empty cells out of 36 possible" in r.stdout
    opts, cells = formats.read_cells(tmp_path / "cells.abn")
    sv = lib.survey_sph(nr, nmu, nphi, 200.0, 1000.0, 0.4, 1.0, 0.0, 6.283185307179586)
    sel = O.Selection(1, rr=rr, rn=rn, mask=mask, nside=4, nest=1)
    ref, ne = O.basis(2, sv, sel, with_evals=True)
    assert opts["Ncells"] == str(len(ref)) and opts["coordsys"] == "spherical" and opts["Nphi"] == str(nphi)
    assert opts["maskfile"] == f"{tmp_path}/mask.fits" and opts["galfile"] == "gals.abn"
    np.testing.assert_array_equal(cells[["a", "d1", "d2", "d3"]], ref[["a", "d1", "d2", "d3"]])
    assert_nbar_close(cells, ref, ne)
    k, P = synth.pk_table()
    formats.write_pkfile(tmp_path / "pk.dat", k, P)
    r = subprocess.run([os.path.join(BIN, "klt"), *args[2:], f"cellfile={tmp_path}/cells.abn", "Nmodes=5",
                        "model=RealModel", f"RealModel.pkfile={tmp_path}/pk.dat", "RealModel.Nbands=2",
                        "RealModel.kmin=0.01", "RealModel.kmax=0.1", "solver=parpack", f"modefile={tmp_path}/m.abn",
                        f"evalfile={tmp_path}/e.abn"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
