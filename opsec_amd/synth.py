"""Synthetic survey configurations (box surveys of the benchmark configs).

Cells are laid out the way the reference's `basis` program lays them out for a
box survey (src/basis.cpp:170-190, src/Survey.cpp:71-91): grid order
(d1, d2, d3) with d3 fastest, bounds Min + d (Max - Min)/N, Veff = volume and
Nbar = nbar * volume (constant selection, src/BoxSurvey.cpp:61-75).
"""
from __future__ import annotations

import numpy as np

from .lib import CELL_DTYPE


def shell_cells(nr, nmu, nphi, rmin=200.0, rmax=1000.0, mumin=0.5, mumax=1.0, phimin=0.0, phimax=2 * np.pi,
                nbar=lambda r: 5e-4 * np.exp(-(r / 600.0) ** 2)):
    """Spherical-shell cells (d1, d2, d3) = (d_r, d_mu, d_phi) in the
    reference's order (r slowest, phi fastest), radial edges uniform in r;
    Nbar = nbar(r at the cell's volume-weighted mid radius) x V."""
    re = np.linspace(rmin, rmax, nr + 1)
    me = np.linspace(mumin, mumax, nmu + 1)
    pe = np.linspace(phimin, phimax, nphi + 1)
    cells = np.zeros(nr * nmu * nphi, dtype=CELL_DTYPE)
    a = 0
    for i in range(nr):
        for j in range(nmu):
            for k in range(nphi):
                c = cells[a]
                c["a"], c["d1"], c["d2"], c["d3"] = a, i, j, k
                c["min1"], c["max1"] = re[i], re[i + 1]
                c["min2"], c["max2"] = me[j], me[j + 1]
                c["min3"], c["max3"] = pe[k], pe[k + 1]
                V = (re[i + 1] ** 3 - re[i] ** 3) / 3 * (me[j + 1] - me[j]) * (pe[k + 1] - pe[k])
                rm = ((re[i] ** 3 + re[i + 1] ** 3) / 2) ** (1 / 3)
                c["Veff"] = V
                c["Nbar"] = nbar(rm) * V
                a += 1
    return cells


def box_cells(n, lo=0.0, hi=400.0, nbar=3e-4):
    d = np.arange(n)
    g1, g2, g3 = np.meshgrid(d, d, d, indexing="ij")
    cells = np.zeros(n ** 3, dtype=CELL_DTYPE)
    cells["a"] = np.arange(n ** 3)
    cells["d1"], cells["d2"], cells["d3"] = g1.ravel(), g2.ravel(), g3.ravel()
    for k, g in (("1", g1), ("2", g2), ("3", g3)):
        cells["min" + k] = lo + g.ravel() * (hi - lo) / n
        cells["max" + k] = lo + (g.ravel() + 1) * (hi - lo) / n
    V = (cells["max1"] - cells["min1"]) * (cells["max2"] - cells["min2"]) * (cells["max3"] - cells["min3"])
    cells["Veff"] = V
    cells["Nbar"] = nbar * V
    return cells


def fiducial_pk(k):
    """P(k) = 2e4 (k/0.02) / (1 + (k/0.02)^2)^1.5  (Mpc/h)^3."""
    x = np.asarray(k, dtype=np.float64) / 0.02
    return 2e4 * x / (1 + x * x) ** 1.5


def pk_table(npts=2000, kmax=10.5):
    """pkfile samples: k = 0 then log-spaced over [1e-4, kmax] (spans [0,10])."""
    k = np.concatenate([[0.0], np.logspace(-4, np.log10(kmax), npts)])
    return k, fiducial_pk(k)


def linear_bands(nb, kmin, kmax):
    """RealModel regular bands (src/RealModel.cpp:60-71)."""
    lo = np.array([kmin + n * (kmax - kmin) / nb for n in range(nb)])
    hi = np.array([kmin + (n + 1) * (kmax - kmin) / nb for n in range(nb)])
    return lo, hi


# The benchmark configurations (box part).
CONFIGS = {
    "C1": dict(n=12, nbands=20, kmin=0.01, kmax=0.20, nmodes=500, nodes=8, nbar=3e-4),
    "C2": dict(n=32, nbands=40, kmin=0.005, kmax=0.25, nmodes=4000, nodes=6, nbar=3e-4),
}


def drop_cells(cells, frac, seed=0):
    """A cell list with a random fraction of grid positions removed and the
    rest renumbered a = 0..n-1 (an irregular survey footprint)."""
    rng = np.random.default_rng(seed)
    keep = rng.random(len(cells)) >= frac
    out = cells[keep].copy()
    out["a"] = np.arange(len(out))
    return out


def catalogue(n, lo, hi, seed=0, margin=0.05, weights="unit"):
    """n synthetic galaxies (GALAXY_DTYPE) uniform in the box [lo, hi) per
    axis widened by `margin` of its extent on each side, so some fall outside
    the survey.  lo/hi: scalars or 3-sequences.  weights: "unit", "int"
    (1..5), or "wide" (log-uniform over 1e-3..1e3, so that the rounding of a
    per-cell sum depends on the order of its terms)."""
    from .lib import GALAXY_DTYPE
    rng = np.random.default_rng(seed)
    lo = np.broadcast_to(np.asarray(lo, dtype=np.float64), (3,))
    hi = np.broadcast_to(np.asarray(hi, dtype=np.float64), (3,))
    ext = hi - lo
    g = np.zeros(n, dtype=GALAXY_DTYPE)
    for k, name in enumerate(("x1", "x2", "x3")):
        g[name] = rng.uniform(lo[k] - margin * ext[k], hi[k] + margin * ext[k], n)
    if weights == "unit":
        g["w"] = 1.0
    elif weights == "int":
        g["w"] = rng.integers(1, 6, n)
    elif weights == "wide":
        g["w"] = 10.0 ** rng.uniform(-3, 3, n)
    else:
        raise ValueError(weights)
    return g


def write_healpix_fits(path, values, nest=False, coordsys="C", per_row=None):
    """A HEALPix map as a FITS binary table (the maskfile of a SphericalSurvey):
    an empty primary HDU, then a BINTABLE whose column 1 holds the 12 nside^2
    float32 values, `per_row` to a row (default: all in rows of up to 1024),
    with NSIDE, ORDERING and COORDSYS in its header."""
    v = np.asarray(values, dtype=">f4").ravel()
    npix = len(v)
    nside = int(round(np.sqrt(npix / 12)))
    assert 12 * nside * nside == npix, "length must be 12 nside^2"
    if per_row is None:
        per_row = min(npix, 1024)
    assert npix % per_row == 0
    nrows = npix // per_row

    def card(key, val):
        if isinstance(val, bool):
            s = f"{key:<8}= {'T' if val else 'F':>20}"
        elif isinstance(val, (int, np.integer)):
            s = f"{key:<8}= {int(val):>20}"
        else:
            s = f"{key:<8}= '{str(val):<8}'"
        return s.ljust(80).encode("ascii")

    def header(cards):
        b = b"".join(cards) + b"END".ljust(80)
        return b + b" " * (-len(b) % 2880)

    prim = header([card("SIMPLE", True), card("BITPIX", 8), card("NAXIS", 0), card("EXTEND", True)])
    ext = header([card("XTENSION", "BINTABLE"), card("BITPIX", 8), card("NAXIS", 2), card("NAXIS1", 4 * per_row),
                  card("NAXIS2", nrows), card("PCOUNT", 0), card("GCOUNT", 1), card("TFIELDS", 1),
                  card("TTYPE1", "SIGNAL"), card("TFORM1", f"{per_row}E"), card("PIXTYPE", "HEALPIX"),
                  card("ORDERING", "NESTED" if nest else "RING"), card("NSIDE", nside), card("COORDSYS", coordsys),
                  card("FIRSTPIX", 0), card("LASTPIX", npix - 1)])
    data = v.tobytes()
    data += b"\0" * (-len(data) % 2880)
    with open(path, "wb") as f:
        f.write(prim + ext + data)
