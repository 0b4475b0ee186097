"""ctypes bindings of the two checkers: oracle/liboracle.so (the FP64 C
restatement) and oracle/_ref/libopsec_ref.so (the reference's own serial
slice, built from its sources by oracle/Makefile when /root/reference is
present).  Used only by the tests, smoke() and bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libopsec_ref.so")

_P, _I, _D = C.c_void_p, C.c_int, C.c_double


class OrcXi(C.Structure):
    _fields_ = [("kind", C.c_int), ("r768", _P), ("y768", _P), ("a", _D), ("b", _D)]


class OrcSelection(C.Structure):
    _fields_ = [("kind", C.c_int), ("nest", C.c_int), ("nside", C.c_int64), ("nbar", _D), ("rr", _P), ("rn", _P),
                ("nrad", C.c_int), ("mask", _P)]


_ORC = {
    "orc_fmm_spline": ([_I, _P, _P, _P, _P, _P], None),
    "orc_cubic_eval": ([_I, _P, _P, _P, _P, _P, _I, _P, _P], None),
    "orc_xilm": ([_I, _I, _P, _P, _I, _P, _P, _I, _D, _D, _D], None),
    "orc_qspline_grid": ([_P], None),
    "orc_qspline_eval": ([_P, _P, _D], _D),
    "orc_band_xi": ([_D, _D, _D], _D),
    "orc_band_param": ([_I, _P, _P, _D, _D], _D),
    "orc_xi_eval": ([C.POINTER(OrcXi), _D], _D),
    "orc_box_r": ([_P, _P, _D], _D),
    "orc_sph_r": ([_P, _P], _D),
    "orc_sig_block_sph": ([C.POINTER(OrcXi), _P, _I, _I, _I, _I, _P, _I, _P, _I, _P, _I, _D, _D, C.c_uint], _I),
    "orc_brute6_sph": ([C.POINTER(OrcXi), _P, _P, _I], _D),
    "orc_gm_integrate": ([C.POINTER(OrcXi), _I, _D, _P, _P, _D, _D, C.c_uint, C.POINTER(_D),
                          C.POINTER(C.c_int64)], _D),
    "orc_tent_gauss": ([C.POINTER(OrcXi), _D, _P, _P, _I], _D),
    "orc_tent_composite": ([C.POINTER(OrcXi), _D, _P, _P, _I, _I], _D),
    "orc_brute6": ([C.POINTER(OrcXi), _D, _P, _P, _I], _D),
    "orc_gauss_legendre": ([_I, _P, _P], None),
    "orc_gauss_jacobi10": ([_I, _P, _P], None),
    "orc_sig_block": ([C.POINTER(OrcXi), _P, _I, _I, _I, _I, _D, _P, _I, _P, _I, _P, _I, _I, _D, _D, C.c_uint, _I,
                       _I], _I),
    "orc_project": ([_I, _P, _I, _P, _P], None),
    "orc_gemm_rows": ([_I, _I, _P, _I, _P, _P], None),
    "orc_estimate": ([_I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "orc_inverse": ([_I, _P, _P], _I),
    "orc_sqrtm": ([_I, _P, _P], None),
    "orc_sym_eigen": ([_I, _P, _P, _P], None),
    "orc_mlp_nweights": ([], _I),
    "orc_mlp_forward": ([_P, _I, _P, _P], None),
    "orc_mlp_loss_grad": ([_P, _I, _P, _P, _P], _D),
    "orc_grid_index": ([_P, _P, _D, _D, _D], C.c_int64),
    "orc_count_galaxies": ([_P, _I, _P, _P, _P, C.c_int64, _P], None),
    "orc_pixel_values": ([_P, _I, _I, _P, _P, _P, _P], None),
    "orc_selection_eval": ([C.POINTER(OrcSelection), _I, _P], _D),
    "orc_cell_integral": ([_I, C.POINTER(OrcSelection), _P, _D, _D, C.c_uint, C.POINTER(C.c_int64)], _D),
    "orc_basis": ([_I, _P, _P, C.POINTER(OrcSelection), _D, _D, _P, _P], _I),
    "orc_ang2pix_ring": ([C.c_int64, _D, _D], C.c_int64),
    "orc_ang2pix_nest": ([C.c_int64, _D, _D], C.c_int64),
    "orc_pix2ang_ring": ([C.c_int64, C.c_int64, C.POINTER(_D), C.POINTER(_D)], None),
    "orc_pix2ang_nest": ([C.c_int64, C.c_int64, C.POINTER(_D), C.POINTER(_D)], None),
}

_REF = {
    "ref_model_new": ([C.c_char_p], _P),
    "ref_model_free": ([_P], None),
    "ref_model_nparams": ([_P], _I),
    "ref_model_param": ([_P, _I], _D),
    "ref_model_xi": ([_P, _I, _I, _P, _P], None),
    "ref_xilm": ([_I, _I, _I, _P, _P, _I, _P, _P, _I, _D, _D, _D], None),
    "ref_cubic_spline": ([_I, _P, _P, _I, _P, _P], None),
    "ref_box_sep": ([C.c_char_p, _I, _P, _P, _P], None),
    "ref_signal_c": ([_P, _I, C.c_char_p, _P, _P, _D, _D, C.POINTER(_D)], _D),
    "ref_sig_block": ([_P, _I, C.c_char_p, _P, _I, _P, _I, _P, _I, _P, _I, _D, _D], _I),
    "ref_read_cells": ([C.c_char_p, _P, _I], _I),
    "ref_sizeof_cell": ([], _I),
    "ref_read_modes": ([C.c_char_p, C.POINTER(_I), _P, C.c_long], _I),
    "ref_read_abn_doubles": ([C.c_char_p, _P, C.c_long], C.c_long),
    "ref_grid_index": ([C.c_char_p, _I, _P, _P], None),
    "ref_read_galaxies": ([C.c_char_p, _P, _I], _I),
    "ref_cell_integral": ([_I, _D, _I, _P, _P, _P, _P, _D, _D, C.POINTER(C.c_int64)], _D),
}


def _load(path, sigs):
    lib = C.CDLL(path)
    for name, (a, r) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = a
        fn.restype = r
    return lib


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run __graft_entry__.build()")
        _orc = _load(ORACLE_SO, _ORC)
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = _load(REF_SO, _REF)
    return _ref


def ptr(a):
    """Pointer to an array's data that keeps the array alive for the call
    (numpy's data_as holds a reference), so temporaries are safe."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- convenience wrappers -----------------------------------------------------

class XiHandle:
    """An oracle xi function; keeps the table arrays alive."""

    def __init__(self, kind, r768=None, y768=None, a=0.0, b=0.0):
        self.r = None if r768 is None else np.ascontiguousarray(r768, dtype=np.float64)
        self.y = None if y768 is None else np.ascontiguousarray(y768, dtype=np.float64)
        self.s = OrcXi(kind, ptr(self.r), ptr(self.y), a, b)

    def __call__(self, r):
        o = oracle()
        return np.array([o.orc_xi_eval(C.byref(self.s), float(x)) for x in np.atleast_1d(r)])


def oracle_tables(pk_k, pk_P, band_lo, band_hi):
    """Reference tables (prior + bands) computed by the oracle."""
    o = oracle()
    r = np.zeros(768)
    o.orc_qspline_grid(ptr(r))
    k, P = np.ascontiguousarray(pk_k, float), np.ascontiguousarray(pk_P, float)
    tabs = []
    y = np.zeros(768)
    o.orc_xilm(0, len(k), ptr(k), ptr(P), 768, ptr(r), ptr(y), 32768, 0.0, 10.0, 1.0)
    tabs.append(y.copy())
    for lo, hi in zip(band_lo, band_hi):
        o.orc_xilm(1, 0, None, None, 768, ptr(r), ptr(y), 8192, float(lo), float(hi), 1.0)
        tabs.append(y.copy())
    return r, tabs


def survey_text(n, lo=0.0, hi=400.0, L=None, nbar=3e-4):
    lo, hi = float(lo), float(hi)
    L = (hi - lo) if L is None else float(L)
    nbar = float(nbar)
    s = "".join(f"survey.{a}Min = {lo!r}\nsurvey.{a}Max = {hi!r}\n" for a in "XYZ")
    s += f"survey.Nx = {n}\nsurvey.Ny = {n}\nsurvey.Nz = {n}\nsurvey.nbar = {nbar!r}\nsurvey.L = {L!r}\n"
    return s.encode()


def write_pkfile(path, k, P):
    np.savetxt(path, np.c_[k, P], fmt="%.17g")


def ref_real_model(pkfile, band_lo, band_hi):
    """Reference RealModel with explicit linear bands (Nbands, kmin, kmax)."""
    nb = len(band_lo)
    txt = (f"model = RealModel\nRealModel.pkfile = {pkfile}\nRealModel.Nbands = {nb}\n"
           f"RealModel.kmin = {float(band_lo[0])!r}\nRealModel.kmax = {float(band_hi[-1])!r}\n")
    return ref().ref_model_new(txt.encode())


def survey_geom(sv):
    """(geom, n) arrays of an opsec_survey for the oracle's grid functions."""
    geom = np.array([sv.min1, sv.max1, sv.min2, sv.max2, sv.min3, sv.max3])
    return geom, np.array([sv.n1, sv.n2, sv.n3], dtype=np.int32)


def count_galaxies(cells, sv, gals, counts=None):
    """Oracle weighted counts (Ncells + 1), accumulated into `counts` if given."""
    geom, n = survey_geom(sv)
    out = np.zeros(len(cells) + 1) if counts is None else np.array(counts, dtype=np.float64)
    g = np.ascontiguousarray(gals)
    oracle().orc_count_galaxies(ptr(cells), len(cells), ptr(geom), ptr(n), ptr(g), len(g), ptr(out))
    return out


def pixel_values(cells, B, counts):
    """Oracle (x, y) for mode-major B (nmodes x ncells)."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    c = np.ascontiguousarray(counts, dtype=np.float64)
    x = np.zeros(len(cells))
    y = np.zeros(B.shape[0])
    oracle().orc_pixel_values(ptr(cells), len(cells), B.shape[0], ptr(B), ptr(c), ptr(x), ptr(y))
    return x, y


class Selection:
    """An oracle selection function; keeps its arrays alive."""

    def __init__(self, kind=0, nbar=0.0, rr=None, rn=None, mask=None, nside=0, nest=0):
        self.rr = None if rr is None else np.ascontiguousarray(rr, dtype=np.float64)
        self.rn = None if rn is None else np.ascontiguousarray(rn, dtype=np.float64)
        self.mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        self.s = OrcSelection(kind, nest, nside, nbar, ptr(self.rr), ptr(self.rn),
                              0 if self.rr is None else len(self.rr), ptr(self.mask))


def basis(coordsys, sv, sel, epsrel=1e-5, epsabs=1e-10, with_evals=False):
    """Oracle basis cells of survey `sv` (an opsec_survey) with selection
    `sel`; with_evals: also the integrand evaluations of each cell."""
    from opsec_amd.lib import CELL_DTYPE
    g = np.array([sv.min1, sv.max1, sv.min2, sv.max2, sv.min3, sv.max3])
    n = np.array([sv.n1, sv.n2, sv.n3], dtype=np.int32)
    out = np.zeros(int(n.prod()), dtype=CELL_DTYPE)
    ne = np.zeros(int(n.prod()), dtype=np.int64)
    k = oracle().orc_basis(coordsys, ptr(g), ptr(n), C.byref(sel.s), epsrel, epsabs, out.ctypes.data, ptr(ne))
    return (out[:k].copy(), ne[:k].copy()) if with_evals else out[:k].copy()
