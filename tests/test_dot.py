"""The dot stage on the CPU: the oracle's grid index against the reference's
own Survey::GetGridIndex, the galfile format against the reference's reader,
known answers for the weighted counts, and the pixel-value arithmetic against
numpy.  The device path is checked against this oracle in test_gpu_dot.py."""
import numpy as np
import pytest

from opsec_amd import formats, lib, synth
from tests import oracles as O


def _grid_points(lo, hi, n, rng, count=1500):
    """Interior, boundary, cell-edge and outside points (float32 triples)."""
    h = (hi - lo) / n
    edges = (lo + h * np.arange(n + 1)).astype(np.float32)
    pts = [rng.uniform(lo - 0.1 * (hi - lo), hi + 0.1 * (hi - lo), (count, 3))]
    e = rng.choice(edges, (400, 3))
    pts.append(e)
    pts.append(np.nextafter(e, np.float32(-np.inf)))
    pts.append(np.nextafter(e, np.float32(np.inf)))
    pts.append(np.array([[lo, lo, lo], [hi, hi, hi], [lo, hi, lo], [hi - 1e-3, lo, lo], [hi - 1e-3, hi - 1e-3, hi - 1e-3]]))
    return np.ascontiguousarray(np.concatenate(pts).astype(np.float32))


@pytest.mark.parametrize("n,lo,hi", [(8, 0.0, 400.0), (7, -55.5, 333.25), (32, 0.0, 2000.0)])
def test_grid_index_vs_reference(n, lo, hi):
    if not O.ref_available():
        pytest.skip("reference build not present")
    pts = _grid_points(lo, hi, n, np.random.default_rng(n))
    ref = np.zeros(len(pts), dtype=np.int32)
    O.ref().ref_grid_index(O.survey_text(n, lo, hi), len(pts), O.ptr(pts), O.ptr(ref))
    sv = lib.survey_box(n, lo, hi)
    geom, nn = O.survey_geom(sv)
    got = np.array([O.oracle().orc_grid_index(O.ptr(geom), O.ptr(nn), float(p[0]), float(p[1]), float(p[2]))
                    for p in pts])
    assert np.array_equal(got, ref)
    assert (ref == 0).any() and (ref == -1).any() and (ref == n ** 3 - 1).any()


def test_grid_index_nan_and_huge_are_outside():
    sv = lib.survey_box(4, 0.0, 100.0)
    geom, nn = O.survey_geom(sv)
    o = O.oracle()
    for p in ((np.nan, 1.0, 1.0), (1.0, np.inf, 1.0), (1.0, 1.0, -np.inf), (3e38, 1.0, 1.0), (-3e38, 1.0, 1.0)):
        assert o.orc_grid_index(O.ptr(geom), O.ptr(nn), *p) == -1


def test_galfile_format_vs_reference_reader(tmp_path):
    if not O.ref_available():
        pytest.skip("reference build not present")
    g = synth.catalogue(1000, 0.0, 400.0, seed=3, weights="wide")
    path = tmp_path / "gals.abn"
    formats.write(path, g, "4f")
    txt = O.survey_text(8) + f"survey.galfile = {path}\n".encode()
    out = np.zeros(len(g), dtype=lib.GALAXY_DTYPE)
    assert O.ref().ref_read_galaxies(txt, O.ptr(out), len(out)) == len(g)
    assert out.tobytes() == g.tobytes()


def _tiny_survey():
    """2x2x2 box [0,2)^3 whose grid position e = 5 has no cell."""
    cells = synth.box_cells(2, 0.0, 2.0)
    cells = cells[np.arange(8) != 5].copy()
    cells["a"] = np.arange(7)
    return cells, lib.survey_box(2, 0.0, 2.0)


def _gals(rows):
    g = np.zeros(len(rows), dtype=lib.GALAXY_DTYPE)
    for i, r in enumerate(rows):
        g[i] = r
    return g


def test_count_known_answers():
    cells, sv = _tiny_survey()
    big = float(2 ** 60)
    g = _gals([
        (0.5, 0.5, 0.5, 1.0),     # grid position 0 -> cell 0 (the reference drops it: e > 0)
        (1.5, 0.5, 1.5, 7.0),     # e = 5 has no cell -> outside
        (2.0, 0.5, 0.5, 2.0),     # x = XMax -> outside
        (-0.25, 0.5, 0.5, 3.0),   # below XMin -> outside
        (0.5, 0.5, 1.5, big),     # cell 1: 2^60 + 1 rounds to 2^60, then -2^60 -> 0
        (0.5, 0.5, 1.25, 1.0),
        (0.75, 0.5, 1.5, -big),
        (0.5, 1.5, 0.5, big),     # cell 2: 2^60 - 2^60 + 1 -> 1
        (0.5, 1.5, 0.5, -big),
        (0.5, 1.5, 0.5, 1.0),
        (1.5, 1.5, 1.5, 0.5),     # e = 7 -> cell 6
        (1.0, 1.0, 1.0, 0.25),    # exactly on the inner edges -> e = 7
    ])
    c = O.count_galaxies(cells, sv, g)
    assert list(c) == [1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.75, 12.0]
    # accumulation continues the fold
    c2 = O.count_galaxies(cells, sv, g[:5])
    c2 = O.count_galaxies(cells, sv, g[5:], c2)
    assert np.array_equal(c2, c)


def test_count_integer_weights_vs_histogram():
    n, lo, hi = 8, 0.0, 400.0
    cells = synth.drop_cells(synth.box_cells(n, lo, hi), 0.2, seed=1)
    sv = lib.survey_box(n, lo, hi)
    g = synth.catalogue(50000, lo, hi, seed=2, weights="int")
    c = O.count_galaxies(cells, sv, g)
    # independent histogram: numpy binning, cell lookup by dictionary, np.add.at
    pos = np.stack([g["x1"], g["x2"], g["x3"]], 1).astype(np.float64)
    d = np.floor(n * (pos - lo) / (hi - lo))
    inside = np.all((d >= 0) & (d < n), axis=1)
    lookup = {(int(c_["d1"]), int(c_["d2"]), int(c_["d3"])): int(c_["a"]) for c_ in cells}
    idx = np.full(len(g), len(cells))
    for i in np.nonzero(inside)[0]:
        idx[i] = lookup.get(tuple(int(v) for v in d[i]), len(cells))
    ref = np.zeros(len(cells) + 1)
    np.add.at(ref, idx, g["w"].astype(np.float64))
    assert np.array_equal(c, ref)
    assert c.sum() == g["w"].astype(np.float64).sum()
    assert 0 < c[-1] < c.sum()


def test_count_rounding_depends_on_order():
    """With weights spanning six decades the per-cell sums are order-dependent,
    so bit-exact agreement of another implementation checks the order."""
    n, lo, hi = 6, 0.0, 400.0
    cells = synth.box_cells(n, lo, hi)
    sv = lib.survey_box(n, lo, hi)
    g = synth.catalogue(200000, lo, hi, seed=4, weights="wide")
    c = O.count_galaxies(cells, sv, g)
    p = np.random.default_rng(5).permutation(len(g))
    cp = O.count_galaxies(cells, sv, g[p])
    assert not np.array_equal(c, cp)
    np.testing.assert_allclose(c, cp, rtol=1e-13)


def test_pixel_values_vs_numpy():
    n = 6
    cells = synth.box_cells(n, 0.0, 150.0, nbar=2e-3)
    sv = lib.survey_box(n, 0.0, 150.0)
    g = synth.catalogue(30000, 0.0, 150.0, seed=6)
    counts = O.count_galaxies(cells, sv, g)
    B = np.random.default_rng(7).normal(size=(40, len(cells)))
    x, y = O.pixel_values(cells, B, counts)
    xr = (counts[:-1] - cells["Nbar"]) / np.sqrt(cells["Nbar"])
    assert np.array_equal(x, xr)
    scale = np.abs(B) @ np.abs(x)
    assert np.all(np.abs(y - B @ x) <= 1e-13 * scale)
