"""ctypes binding of libopsec_hip.so (the C ABI declared in include/opsec_hip.h).

This is the host-side mirror used by the Python tests and bench.py.  Every
call checks the return code and raises OpsecError with the library's message.
Loading fails loudly when the shared library has not been built: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libopsec_hip.so")

OPSEC_OK = 0
OPSEC_COORD_CARTESIAN = 1
OPSEC_COORD_SPHERICAL = 2
OPSEC_XI_TABLE = 0
OPSEC_XI_CLOSED = 1
OPSEC_QUAD_GM = 0
OPSEC_QUAD_GAUSS = 1
OPSEC_QUAD_SURROGATE = 2
OPSEC_PREC_FP64 = 0
OPSEC_PREC_FP32 = 1
OPSEC_PREC_BF16X3 = 2
OPSEC_PROJ_DENSE = 0
OPSEC_PROJ_TOEPLITZ = 1
OPSEC_MIX_INVERSE = 0
OPSEC_MIX_INVERSE_SQRT = 1
OPSEC_MIX_IDENTITY = 2
OPSEC_SEL_UNIFORM = 0
OPSEC_SEL_RADIAL = 1

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ECONV", -5: "ESTATE", -6: "ENODEV"}

# 80-byte cell record, src/Cell.h:74-96 ("4i8d")
CELL_DTYPE = np.dtype([("a", "<i4"), ("d1", "<i4"), ("d2", "<i4"), ("d3", "<i4"),
                       ("min1", "<f8"), ("max1", "<f8"), ("min2", "<f8"), ("max2", "<f8"),
                       ("min3", "<f8"), ("max3", "<f8"), ("Veff", "<f8"), ("Nbar", "<f8")])
assert CELL_DTYPE.itemsize == 80

# galfile record, src/Survey.h:13-19 ("4f")
GALAXY_DTYPE = np.dtype([("x1", "<f4"), ("x2", "<f4"), ("x3", "<f4"), ("w", "<f4")])


class Survey(C.Structure):
    _fields_ = [("coordsys", C.c_int32), ("n1", C.c_int32), ("n2", C.c_int32), ("n3", C.c_int32),
                ("min1", C.c_double), ("max1", C.c_double), ("min2", C.c_double), ("max2", C.c_double),
                ("min3", C.c_double), ("max3", C.c_double), ("L", C.c_double)]


class Quad(C.Structure):
    _fields_ = [("method", C.c_int32), ("nodes", C.c_int32), ("near", C.c_int32), ("maxeval", C.c_int32),
                ("epsrel", C.c_double), ("epsabs", C.c_double), ("spot_frac", C.c_double),
                ("spot_tol", C.c_double), ("spot_seed", C.c_uint64)]


class Selection(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nest", C.c_int32), ("nside", C.c_int64), ("nbar", C.c_double),
                ("rad_r", C.c_void_p), ("rad_n", C.c_void_p), ("nrad", C.c_int32), ("reserved", C.c_int32),
                ("mask", C.c_void_p)]


class Estimate(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("phat", "cov", "F", "W", "M", "q", "f")]


class MlpTrainOpts(C.Structure):
    _fields_ = [("nsamples", C.c_int32), ("epochs", C.c_int32), ("batch", C.c_int32), ("seed", C.c_int32),
                ("lr", C.c_double), ("near", C.c_double)]


class OpsecError(RuntimeError):
    pass


_lib = None

_P = C.c_void_p
_I = C.c_int
_D = C.c_double

_SIGS = {
    "opsec_abi_version": ([], _I),
    "opsec_ctx_create": ([_I, C.POINTER(_P)], _I),
    "opsec_ctx_destroy": ([_P], _I),
    "opsec_last_error": ([_P], C.c_char_p),
    "opsec_ctx_sync": ([_P], _I),
    "opsec_ctx_stream": ([_P], _P),
    "opsec_device_info": ([_P, C.POINTER(_I), C.c_char_p, _I], _I),
    "opsec_model_real": ([_P, _P, _P, _I, _I, _P, _P, _I], _I),
    "opsec_model_test": ([_P, _I, _D, _D], _I),
    "opsec_model_tables": ([_P, _I, _P, _P], _I),
    "opsec_model_nparams": ([_P, C.POINTER(_I)], _I),
    "opsec_model_params": ([_P, _P], _I),
    "opsec_model_xi": ([_P, _I, _I, _P, _P], _I),
    "opsec_model_table": ([_P, _I, _P, _P], _I),
    "opsec_quad_defaults": ([_I, C.POINTER(Quad)], _I),
    "opsec_sig_block": ([_P, _P, _I, _P, _I, _P, _I, _I, C.POINTER(Survey), C.POINTER(Quad), _P, _I], _I),
    "opsec_pair_integral": ([_P, _P, _P, _I, C.POINTER(Survey), C.POINTER(Quad), C.POINTER(_D), C.POINTER(_D)], _I),
    "opsec_cells_upload": ([_P, _P, _I, C.POINTER(Survey)], _I),
    "opsec_pair_tables": ([_P, _I, _P, C.POINTER(Quad)], _I),
    "opsec_pair_nkeys": ([_P, C.POINTER(_I)], _I),
    "opsec_pair_table_get": ([_P, _I, _P], _I),
    "opsec_pair_stats": ([_P] + [C.POINTER(C.c_int64)] * 5, _I),
    "opsec_assemble_device": ([_P, _I, _P, _I], _I),
    "opsec_dev_alloc": ([_P, C.c_size_t, C.POINTER(_P)], _I),
    "opsec_dev_free": ([_P, _P], _I),
    "opsec_h2d": ([_P, _P, _P, C.c_size_t], _I),
    "opsec_d2h": ([_P, _P, _P, C.c_size_t], _I),
    "opsec_klt": ([_P, _I, _D, _I, _I, _P, _P, C.POINTER(_I)], _I),
    "opsec_lanczos_dense": ([_P, _I, _P, _I, _D, _I, _I, _P, _P, C.POINTER(_I)], _I),
    "opsec_modes_upload": ([_P, _P, _I, _I], _I),
    "opsec_project": ([_P, _I, _P, _I, _P], _I),
    "opsec_set_projection": ([_P, _I], _I),
    "opsec_project_dense": ([_P, _I, _P, _I, _P, _I, _P], _I),
    "opsec_count_galaxies": ([_P, _P, C.c_int64, _I, _P], _I),
    "opsec_pixel_values": ([_P, _P, _P, _P], _I),
    "opsec_basis": ([_P, C.POINTER(Survey), C.POINTER(Selection), C.c_double, C.c_double, _P, C.POINTER(_I),
                     C.POINTER(C.c_int64)], _I),
    "opsec_healpix_pixels": ([_P, C.c_int64, _I, C.c_int64, _P, _P, _P], _I),
    "opsec_estimate_run": ([_P, _I, _I, _P, _P, _P, _P, _I, _I, C.POINTER(Estimate)], _I),
    "opsec_estimate_device": ([_P, _P, _P, _P, _I, _I, C.POINTER(Estimate)], _I),
    "opsec_fisher_gram_noalloc": ([_P, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P], _I),
    "opsec_mlp_defaults": ([C.POINTER(MlpTrainOpts)], _I),
    "opsec_mlp_train": ([_P, C.POINTER(MlpTrainOpts), C.POINTER(Quad), C.POINTER(_D)], _I),
    "opsec_mlp_nweights": ([_P, C.POINTER(_I)], _I),
    "opsec_mlp_get_weights": ([_P, _P], _I),
    "opsec_mlp_set_weights": ([_P, _P], _I),
    "opsec_mlp_forward": ([_P, _I, _P, _P], _I),
    "opsec_mlp_loss_grad": ([_P, _I, _P, _P, C.POINTER(_D), _P], _I),
}

# include/opsec_hip_diag.h: timings and tuning knobs (benchmark and tests only)
_F = C.POINTER(C.c_float)
_DIAG_SIGS = {
    "opsec_diag_copy": ([_P, _P, _P, C.c_size_t, _F], _I),
    "opsec_diag_sincos_rate": ([_P, C.POINTER(_D)], _I),
    "opsec_diag_tridiag_top": ([_I, _P, _P, _I, _P, _P], _I),
    "opsec_diag_klt_times": ([_P, _I], _I),
    "opsec_diag_project_times": ([_P, _F, _F], _I),
    "opsec_diag_project_kernel_ms": ([_P, _F, _F, _F, C.POINTER(_I)], _I),
    "opsec_diag_project_launches": ([_P, C.POINTER(_I), C.POINTER(_I)], _I),
    "opsec_diag_set_nt_tile": ([_I], _I),
    "opsec_diag_set_nt_gen": ([_I], _I),
    "opsec_diag_set_stage2_split": ([_I], _I),
    "opsec_diag_set_toeplitz_stage2_split": ([_I], _I),
    "opsec_diag_set_stage2_group": ([_I], _I),
    "opsec_diag_set_nt_order": ([_I, _I], _I),
    "opsec_diag_set_g3_dl": ([_I], _I),
    "opsec_diag_set_gm_overlap": ([_I], _I),
    "opsec_diag_set_gm_bands": ([_I], _I),
    "opsec_diag_set_gm_deep_slots": ([_I], _I),
}


def exported_symbols():
    """Names of every entry point include/opsec_hip.h declares."""
    return sorted(_SIGS)


def diag_symbols():
    """Names of the diagnostics include/opsec_hip_diag.h declares."""
    return sorted(_DIAG_SIGS)


def load(path: str | None = None):
    """Load the C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OpsecError(f"{p} is missing: build it with __graft_entry__.build() or make -C opsec_amd")
    lib = C.CDLL(p)
    for name, (argt, rest) in list(_SIGS.items()) + list(_DIAG_SIGS.items()):
        fn = getattr(lib, name, None)
        if fn is None:   # reported by tests/test_abi.py
            continue
        fn.argtypes = argt
        fn.restype = rest
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    # data_as keeps a reference to the array for as long as the pointer lives
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One context = one HIP stream on one device (not thread safe)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.opsec_ctx_create(device, C.byref(h))
        if rc != OPSEC_OK:
            raise OpsecError(f"opsec_ctx_create({device}) failed: {ERRORS.get(rc, rc)} "
                             "(needs a gfx950 device)")
        self.h = h

    def close(self):
        if self.h:
            self.lib.opsec_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc, what=""):
        if rc != OPSEC_OK:
            msg = self.lib.opsec_last_error(self.h)
            raise OpsecError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    # ---- device -----------------------------------------------------------
    def device_info(self):
        n = C.c_int()
        buf = C.create_string_buffer(64)
        self.check(self.lib.opsec_device_info(self.h, C.byref(n), buf, 64), "device_info")
        return n.value, buf.value.decode()

    def sync(self):
        self.check(self.lib.opsec_ctx_sync(self.h), "sync")

    # ---- model ------------------------------------------------------------
    def model_real(self, pk_k, pk_P, band_lo, band_hi, xi_mode=OPSEC_XI_TABLE):
        k, P = f64(pk_k), f64(pk_P)
        lo, hi = f64(band_lo), f64(band_hi)
        self.check(self.lib.opsec_model_real(self.h, _ptr(k), _ptr(P), len(k), len(lo), _ptr(lo), _ptr(hi),
                                             xi_mode), "model_real")

    def model_tables(self, xi768, params=None):
        """Upload a model given by its QSpline node values (nfuncs x 768)."""
        t = np.ascontiguousarray(xi768, dtype=np.float64)
        assert t.ndim == 2 and t.shape[1] == 768
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        self.check(self.lib.opsec_model_tables(self.h, t.shape[0], _ptr(t), _ptr(p)), "model_tables")

    def model_test(self, form, A, b):
        self.check(self.lib.opsec_model_test(self.h, form, A, b), "model_test")

    def nparams(self):
        n = C.c_int()
        self.check(self.lib.opsec_model_nparams(self.h, C.byref(n)), "nparams")
        return n.value

    def params(self):
        out = np.zeros(self.nparams())
        self.check(self.lib.opsec_model_params(self.h, _ptr(out)), "params")
        return out

    def xi(self, f, r):
        r = f64(r)
        out = np.zeros_like(r)
        self.check(self.lib.opsec_model_xi(self.h, f, len(r), _ptr(r), _ptr(out)), "model_xi")
        return out

    def table(self, f):
        r = np.zeros(768)
        y = np.zeros(768)
        self.check(self.lib.opsec_model_table(self.h, f, _ptr(r), _ptr(y)), "model_table")
        return r, y

    # ---- quadrature -------------------------------------------------------
    @staticmethod
    def quad(method=OPSEC_QUAD_GM, **kw):
        q = Quad()
        load().opsec_quad_defaults(method, C.byref(q))
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    def sig_block(self, cells, rows, cols, f, survey, quad, lld=0):
        cells = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        ld = lld if lld > 0 else len(cols)
        s = np.zeros((len(rows), ld))
        self.check(self.lib.opsec_sig_block(self.h, _ptr(cells), len(cells), _ptr(rows), len(rows), _ptr(cols),
                                            len(cols), f, C.byref(survey), C.byref(quad), _ptr(s), ld),
                   "sig_block")
        return s

    def pair_integral(self, c1, c2, f, survey, quad):
        c1 = np.ascontiguousarray(c1, dtype=CELL_DTYPE).reshape(1)
        c2 = np.ascontiguousarray(c2, dtype=CELL_DTYPE).reshape(1)
        q, e = C.c_double(), C.c_double()
        self.check(self.lib.opsec_pair_integral(self.h, _ptr(c1), _ptr(c2), f, C.byref(survey), C.byref(quad),
                                                C.byref(q), C.byref(e)), "pair_integral")
        return q.value, e.value

    def cells_upload(self, cells, survey):
        cells = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        self._cells = cells
        self.check(self.lib.opsec_cells_upload(self.h, _ptr(cells), len(cells), C.byref(survey)), "cells_upload")

    def pair_tables(self, funcs, quad):
        fs = np.ascontiguousarray(funcs, dtype=np.int32)
        self.check(self.lib.opsec_pair_tables(self.h, len(fs), _ptr(fs), C.byref(quad)), "pair_tables")

    def nkeys(self):
        n = C.c_int()
        self.check(self.lib.opsec_pair_nkeys(self.h, C.byref(n)), "pair_nkeys")
        return n.value

    def pair_table(self, f):
        out = np.zeros(self.nkeys())
        self.check(self.lib.opsec_pair_table_get(self.h, f, _ptr(out)), "pair_table_get")
        return out

    def pair_stats(self):
        v = [C.c_int64() for _ in range(5)]
        self.check(self.lib.opsec_pair_stats(self.h, *[C.byref(x) for x in v]), "pair_stats")
        return dict(zip(("integrals", "nodes", "surrogate", "spot", "spot_fail"), [x.value for x in v]))

    def dev_alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.opsec_dev_alloc(self.h, nbytes, C.byref(p)), "dev_alloc")
        return p.value

    def dev_free(self, p):
        self.check(self.lib.opsec_dev_free(self.h, p), "dev_free")

    def d2h(self, dst, src):
        self.check(self.lib.opsec_d2h(self.h, _ptr(dst), src, dst.nbytes), "d2h")

    def h2d(self, dst, src):
        self.check(self.lib.opsec_h2d(self.h, dst, _ptr(src), src.nbytes), "h2d")

    def assemble(self, f, ncells):
        """Full S_f (device assembly, copied back)."""
        d = self.dev_alloc(ncells * ncells * 8)
        try:
            self.check(self.lib.opsec_assemble_device(self.h, f, d, ncells), "assemble_device")
            out = np.zeros((ncells, ncells))
            self.d2h(out, d)
        finally:
            self.dev_free(d)
        return out

    # ---- eigenmodes -------------------------------------------------------
    def klt(self, nev, tol=1e-5, block=0, maxdim=0, ncells=None):
        n = ncells if ncells is not None else len(self._cells)
        ev = np.zeros(nev)
        vec = np.zeros((nev, n))
        nconv = C.c_int()
        self.check(self.lib.opsec_klt(self.h, nev, tol, block, maxdim, _ptr(ev), _ptr(vec), C.byref(nconv)), "klt")
        return ev, vec, nconv.value

    def lanczos_dense(self, A, nev, tol=1e-10, block=0, maxdim=0):
        A = f64(A)
        n = A.shape[0]
        ev = np.zeros(nev)
        vec = np.zeros((nev, n))
        nconv = C.c_int()
        self.check(self.lib.opsec_lanczos_dense(self.h, n, _ptr(A), nev, tol, block, maxdim, _ptr(ev), _ptr(vec),
                                                C.byref(nconv)), "lanczos_dense")
        return ev, vec, nconv.value

    # ---- projection -------------------------------------------------------
    def modes_upload(self, B):
        """B: (nmodes, ncells) mode-major array (modefile order)."""
        B = f64(B)
        self.check(self.lib.opsec_modes_upload(self.h, _ptr(B), B.shape[1], B.shape[0]), "modes_upload")
        self._nmodes = B.shape[0]

    def count_galaxies(self, gals, counts=None):
        """Weighted counts per cell of a GALAXY_DTYPE array; Ncells + 1 values,
        the last for galaxies in no cell.  With `counts` given the sums continue
        from it (the catalogue passed in pieces, in order)."""
        g = np.ascontiguousarray(gals, dtype=GALAXY_DTYPE)
        acc = counts is not None
        out = np.array(counts, dtype=np.float64) if acc else np.zeros(len(self._cells) + 1)
        assert out.shape == (len(self._cells) + 1,)
        self.check(self.lib.opsec_count_galaxies(self.h, _ptr(g), len(g), int(acc), _ptr(out)), "count_galaxies")
        return out

    def pixel_values(self, counts):
        """(x, y): x_a = (N_a - Nbar_a)/sqrt(Nbar_a) and y = B^T x."""
        c = f64(counts)
        assert len(c) >= len(self._cells)
        x = np.zeros(len(self._cells))
        y = np.zeros(self._nmodes)
        self.check(self.lib.opsec_pixel_values(self.h, _ptr(c), _ptr(x), _ptr(y)), "pixel_values")
        return x, y

    def basis(self, survey, kind=OPSEC_SEL_UNIFORM, nbar=0.0, rad_r=None, rad_n=None, mask=None, nest=0,
              epsrel=1e-5, epsabs=1e-10):
        """Basis cells of `survey` with their Nbar (opsec_basis); returns
        (cells, integrand evaluations)."""
        rr = None if rad_r is None else f64(rad_r)
        rn = None if rad_n is None else f64(rad_n)
        mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        nside = 0 if mk is None else int(round(np.sqrt(len(mk) / 12)))
        sel = Selection(kind, nest, nside, nbar, _ptr(rr), _ptr(rn), 0 if rr is None else len(rr), 0, _ptr(mk))
        out = np.zeros(survey.n1 * survey.n2 * survey.n3, dtype=CELL_DTYPE)
        n, ev = C.c_int(), C.c_int64()
        self.check(self.lib.opsec_basis(self.h, C.byref(survey), C.byref(sel), epsrel, epsabs, _ptr(out), C.byref(n),
                                        C.byref(ev)), "basis")
        return out[:n.value].copy(), ev.value

    def healpix_pixels(self, nside, nest, theta, phi):
        t, p = f64(theta), f64(phi)
        out = np.zeros(len(t), dtype=np.int64)
        self.check(self.lib.opsec_healpix_pixels(self.h, nside, int(nest), len(t), _ptr(t), _ptr(p), _ptr(out)),
                   "healpix_pixels")
        return out

    def set_projection(self, method):
        """OPSEC_PROJ_DENSE or OPSEC_PROJ_TOEPLITZ (DFT along x of a box or phi of a shell; every precision)."""
        self.check(self.lib.opsec_set_projection(self.h, method), "set_projection")

    def project(self, funcs, prec=OPSEC_PREC_FP64, keep_on_device=False):
        fs = np.ascontiguousarray(funcs, dtype=np.int32)
        m = self._nmodes
        out = None if keep_on_device else np.zeros((len(fs), m * (m + 1) // 2))
        self.check(self.lib.opsec_project(self.h, len(fs), _ptr(fs), prec, _ptr(out)), "project")
        return out

    def project_dense(self, S, B, prec=OPSEC_PREC_FP64):
        S, B = f64(S), f64(B)
        n, m = S.shape[0], B.shape[0]
        out = np.zeros(m * (m + 1) // 2)
        self.check(self.lib.opsec_project_dense(self.h, n, _ptr(S), m, _ptr(B), prec, _ptr(out)), "project_dense")
        return out

    # ---- estimate ---------------------------------------------------------
    def _est_out(self, np_):
        o = {k: np.zeros(np_) for k in ("phat", "q", "f")}
        o.update({k: np.zeros((np_, np_)) for k in ("cov", "F", "W", "M")})
        e = Estimate(*[_ptr(o[k]) for k in ("phat", "cov", "F", "W", "M", "q", "f")])
        return o, e

    def estimate(self, lam, D, y, prior, mixing=OPSEC_MIX_INVERSE, prec=OPSEC_PREC_FP64):
        lam, D, y, prior = f64(lam), f64(D), f64(y), f64(prior)
        o, e = self._est_out(len(prior))
        self.check(self.lib.opsec_estimate_run(self.h, len(lam), len(prior), _ptr(lam), _ptr(D), _ptr(y),
                                               _ptr(prior), mixing, prec, C.byref(e)), "estimate_run")
        return o

    def estimate_device(self, lam, y, prior, mixing=OPSEC_MIX_INVERSE, prec=OPSEC_PREC_FP64):
        lam, y, prior = f64(lam), f64(y), f64(prior)
        o, e = self._est_out(len(prior))
        self.check(self.lib.opsec_estimate_device(self.h, _ptr(lam), _ptr(y), _ptr(prior), mixing, prec,
                                                  C.byref(e)), "estimate_device")
        return o

    # ---- surrogate --------------------------------------------------------
    @staticmethod
    def mlp_opts(**kw):
        o = MlpTrainOpts()
        load().opsec_mlp_defaults(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def mlp_train(self, opts, quad):
        loss = C.c_double()
        self.check(self.lib.opsec_mlp_train(self.h, C.byref(opts), C.byref(quad), C.byref(loss)), "mlp_train")
        return loss.value

    def mlp_nweights(self):
        n = C.c_int()
        self.check(self.lib.opsec_mlp_nweights(self.h, C.byref(n)), "mlp_nweights")
        return n.value

    def mlp_weights(self):
        w = np.zeros(self.mlp_nweights(), dtype=np.float32)
        self.check(self.lib.opsec_mlp_get_weights(self.h, _ptr(w)), "mlp_get_weights")
        return w

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self.check(self.lib.opsec_mlp_set_weights(self.h, _ptr(w)), "mlp_set_weights")

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros(len(x), dtype=np.float32)
        self.check(self.lib.opsec_mlp_forward(self.h, len(x), _ptr(x), _ptr(out)), "mlp_forward")
        return out

    def mlp_loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.mlp_nweights(), dtype=np.float32)
        loss = C.c_double()
        self.check(self.lib.opsec_mlp_loss_grad(self.h, len(x), _ptr(x), _ptr(t), C.byref(loss), _ptr(g)),
                   "mlp_loss_grad")
        return loss.value, g


def survey_box(n, lo=0.0, hi=400.0, L=None, ny=None, nz=None):
    """Box survey descriptor: n^3 cells over [lo,hi]^3, period L (default hi-lo)."""
    s = Survey()
    s.coordsys = OPSEC_COORD_CARTESIAN
    s.n1, s.n2, s.n3 = n, ny or n, nz or n
    s.min1 = s.min2 = s.min3 = lo
    s.max1 = s.max2 = s.max3 = hi
    s.L = (hi - lo) if L is None else L
    return s


def survey_sph(nr, nmu, nphi, rmin=200.0, rmax=1000.0, mumin=0.5, mumax=1.0, phimin=0.0, phimax=2 * 3.141592653589793):
    """Spherical survey descriptor (SphericalSurvey keys RMin..PhiMax, Nr, Nmu, Nphi)."""
    s = Survey()
    s.coordsys = OPSEC_COORD_SPHERICAL
    s.n1, s.n2, s.n3 = nr, nmu, nphi
    s.min1, s.max1 = rmin, rmax
    s.min2, s.max2 = mumin, mumax
    s.min3, s.max3 = phimin, phimax
    s.L = 0.0
    return s
