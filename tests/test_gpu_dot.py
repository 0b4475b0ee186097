"""GPU parity of the dot stage (weighted galaxy counts and pixel values)
against the oracle, through the C ABI, and the dot program end to end."""
import os
import subprocess

import numpy as np
import pytest

from opsec_amd import formats, lib, synth
from tests import oracles as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "opsec_amd", "bin")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_counts_bit_exact_box(ctx):
    n, lo, hi = 12, 0.0, 400.0
    cells = synth.drop_cells(synth.box_cells(n, lo, hi), 0.15, seed=11)
    sv = lib.survey_box(n, lo, hi)
    ctx.cells_upload(cells, sv)
    g = synth.catalogue(300000, lo, hi, seed=12, weights="wide")
    got = ctx.count_galaxies(g)
    ref = O.count_galaxies(cells, sv, g)
    assert np.array_equal(_bits(got), _bits(ref))
    assert ref[-1] > 0 and np.count_nonzero(ref[:-1]) == len(cells)


def test_counts_accumulate_in_pieces(ctx):
    n, lo, hi = 8, -100.0, 300.0
    cells = synth.box_cells(n, lo, hi)
    sv = lib.survey_box(n, lo, hi)
    ctx.cells_upload(cells, sv)
    g = synth.catalogue(100000, lo, hi, seed=13, weights="wide")
    whole = ctx.count_galaxies(g)
    c = None
    for a, b in ((0, 1), (1, 33333), (33333, 100000)):
        c = ctx.count_galaxies(g[a:b], c)
    assert np.array_equal(_bits(c), _bits(whole))
    assert np.array_equal(_bits(whole), _bits(O.count_galaxies(cells, sv, g)))


def test_counts_edge_cases(ctx):
    cells = synth.box_cells(2, 0.0, 2.0)
    cells = cells[np.arange(8) != 5].copy()
    cells["a"] = np.arange(7)
    sv = lib.survey_box(2, 0.0, 2.0)
    ctx.cells_upload(cells, sv)
    empty = np.zeros(0, dtype=lib.GALAXY_DTYPE)
    assert np.array_equal(ctx.count_galaxies(empty), np.zeros(8))
    big = float(2 ** 60)
    g = np.array([(0.5, 0.5, 0.5, 1.0), (1.5, 0.5, 1.5, 7.0), (2.0, 0.5, 0.5, 2.0), (-0.25, 0.5, 0.5, 3.0),
                  (0.5, 0.5, 1.5, big), (0.5, 0.5, 1.25, 1.0), (0.75, 0.5, 1.5, -big),
                  (0.5, 1.5, 0.5, big), (0.5, 1.5, 0.5, -big), (0.5, 1.5, 0.5, 1.0),
                  (1.5, 1.5, 1.5, 0.5), (1.0, 1.0, 1.0, 0.25), (np.nan, 0.5, 0.5, 4.0), (0.5, np.inf, 0.5, 8.0)],
                 dtype=lib.GALAXY_DTYPE)
    got = ctx.count_galaxies(g)
    assert list(got) == [1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.75, 24.0]
    assert np.array_equal(_bits(got), _bits(O.count_galaxies(cells, sv, g)))
    # two cells at one grid position: the later cell gets the galaxies
    dup = np.concatenate([cells, cells[:1]])
    dup["a"] = np.arange(len(dup))
    ctx.cells_upload(dup, sv)
    got = ctx.count_galaxies(g)
    assert got[0] == 0.0 and got[7] == 1.0
    assert np.array_equal(_bits(got), _bits(O.count_galaxies(dup, sv, g)))


def test_counts_spherical(ctx):
    nr, nmu, nphi = 8, 4, 16
    rmin, rmax = 200.0, 1000.0
    cells = synth.shell_cells(nr, nmu, nphi, rmin=rmin, rmax=rmax)
    sv = lib.survey_sph(nr, nmu, nphi, rmin, rmax)
    ctx.cells_upload(cells, sv)
    g = synth.catalogue(200000, (rmin, 0.5, 0.0), (rmax, 1.0, 2 * np.pi), seed=14, weights="wide")
    got = ctx.count_galaxies(g)
    ref = O.count_galaxies(cells, sv, g)
    assert np.array_equal(_bits(got), _bits(ref))
    assert ref[-1] > 0


@pytest.mark.parametrize("n,M,ngal", [(6, 40, 30000), (32, 2000, 2000000)])
def test_pixel_values_vs_oracle(ctx, n, M, ngal):
    """x bit-exact (IEEE subtract, sqrt and divide); y = B^T x within 1e-13 of
    sum_a |B_ia x_a| (the two sums differ only in association order).  The
    second case is the 32^3 grid with 2000 modes."""
    lo, hi = 0.0, 25.0 * n
    cells = synth.box_cells(n, lo, hi, nbar=ngal / (hi - lo) ** 3)
    sv = lib.survey_box(n, lo, hi)
    ctx.cells_upload(cells, sv)
    B = np.random.default_rng(n).normal(size=(M, len(cells)))
    ctx.modes_upload(B)
    g = synth.catalogue(ngal, lo, hi, seed=15, margin=0.0)
    counts = ctx.count_galaxies(g)
    assert np.array_equal(_bits(counts), _bits(O.count_galaxies(cells, sv, g)))
    x, y = ctx.pixel_values(counts)
    xr, yr = O.pixel_values(cells, B, counts)
    assert np.array_equal(_bits(x), _bits(xr))
    scale = np.abs(B) @ np.abs(xr)
    assert np.all(np.abs(y - yr) <= 1e-13 * scale)


def test_dot_program_end_to_end(tmp_path):
    n, M, lo, hi = 8, 50, 0.0, 200.0
    cells = synth.drop_cells(synth.box_cells(n, lo, hi, nbar=0.05), 0.1, seed=16)
    N = len(cells)
    formats.write_cells(tmp_path / "cells.abn", cells, {"coordsys": "cartesian"})
    B = np.random.default_rng(17).normal(size=(M, N))
    formats.write(tmp_path / "modes.abn", B.ravel(), "d", {"Nmodes": M, "Ncells": N})
    g = synth.catalogue(200000, lo, hi, seed=18, weights="wide")
    formats.write(tmp_path / "gals.abn", g, "4f")
    cfg = {"cellfile": tmp_path / "cells.abn", "modefile": tmp_path / "modes.abn", "pixfile": tmp_path / "pix.abn",
           "Nmodes": M, "countfile": tmp_path / "counts.txt", "survey.type": "BoxSurvey",
           "survey.XMin": lo, "survey.XMax": hi, "survey.YMin": lo, "survey.YMax": hi, "survey.ZMin": lo,
           "survey.ZMax": hi, "survey.Nx": n, "survey.Ny": n, "survey.Nz": n, "survey.nbar": 0.05,
           "survey.L": hi - lo, "survey.galfile": tmp_path / "gals.abn"}
    formats.write_config(tmp_path / "dot.cfg", cfg)
    r = subprocess.run([os.path.join(BIN, "dot"), "-c", str(tmp_path / "dot.cfg")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    raw = open(tmp_path / "pix.abn", "rb").read()
    assert raw.startswith(b"# Pixel values\n")
    opts, (y,) = formats.read_doubles(tmp_path / "pix.abn")
    assert opts == {} and len(y) == M
    counts = O.count_galaxies(cells, lib.survey_box(n, lo, hi), g)
    xr, yr = O.pixel_values(cells, B, counts)
    scale = np.abs(B) @ np.abs(xr)
    assert np.all(np.abs(y - yr) <= 1e-13 * scale)
    tab = np.loadtxt(tmp_path / "counts.txt")
    assert np.array_equal(tab[:, 1], counts[:-1]) and np.array_equal(tab[:, 2], cells["Nbar"])
    assert f"# {counts[-1]:g} (weighted) galaxies outside any cell" in open(tmp_path / "counts.txt").read()
