"""Pin the FP64 oracle: reference golden vectors, the reference's own known
answers, closed forms, scipy, and brute force.  CPU only."""
import ctypes as C
import math
import os

import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as sp

from opsec_amd import synth
from tests import oracles as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def mk_cell(d, lo, hi, N=1.0, a=0):
    from opsec_amd.lib import CELL_DTYPE
    c = np.zeros(1, dtype=CELL_DTYPE)
    c["a"] = a
    c["d1"], c["d2"], c["d3"] = d
    c["min1"], c["min2"], c["min3"] = lo
    c["max1"], c["max2"], c["max3"] = hi
    V = np.prod(np.array(hi) - np.array(lo))
    c["Veff"], c["Nbar"] = V, N
    return c


@pytest.fixture(scope="module")
def orc():
    return O.oracle()


@pytest.fixture(scope="module")
def tables():
    return O.oracle_tables(GOLD["pk_k"], GOLD["pk_P"], GOLD["c1_band_lo"], GOLD["c1_band_hi"])


# ---- Gauss rules ---------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 16])
def test_gauss_rules_match_scipy(orc, n):
    t, w = np.zeros(n), np.zeros(n)
    orc.orc_gauss_jacobi10(n, O.ptr(t), O.ptr(w))
    x, wx = sp.roots_jacobi(n, 1, 0)
    np.testing.assert_allclose(t, (x + 1) / 2, atol=1e-14)
    np.testing.assert_allclose(w, wx / 4, rtol=1e-12)
    orc.orc_gauss_legendre(n, O.ptr(t), O.ptr(w))
    x, wx = sp.roots_legendre(n)
    np.testing.assert_allclose(t, x, atol=1e-14)
    np.testing.assert_allclose(w, wx, rtol=1e-12)


# ---- splines and xi ---------------------------------------------------------------

def test_cubic_spline_matches_reference(orc):
    k, P = GOLD["pk_k"], GOLD["pk_P"]
    n = len(k)
    b, c, d = np.zeros(n), np.zeros(n), np.zeros(n)
    orc.orc_fmm_spline(n, O.ptr(k), O.ptr(P), O.ptr(b), O.ptr(c), O.ptr(d))
    ke = GOLD["spline_k"]
    ye = np.zeros_like(ke)
    orc.orc_cubic_eval(n, O.ptr(k), O.ptr(P), O.ptr(b), O.ptr(c), O.ptr(d), len(ke), O.ptr(ke), O.ptr(ye))
    np.testing.assert_allclose(ye, GOLD["spline_P"], rtol=1e-13, atol=1e-12)


def test_cubic_spline_reproduces_cubics(orc):
    # the FMM end conditions make the spline exact for any cubic
    x = np.sort(np.random.default_rng(1).uniform(0, 3, 12))
    y = 1.5 - 2 * x + 0.7 * x ** 2 - 0.3 * x ** 3
    n = len(x)
    b, c, d = np.zeros(n), np.zeros(n), np.zeros(n)
    orc.orc_fmm_spline(n, O.ptr(x), O.ptr(y), O.ptr(b), O.ptr(c), O.ptr(d))
    xe = np.linspace(x[0], x[-1] - 1e-9, 50)
    ye = np.zeros_like(xe)
    orc.orc_cubic_eval(n, O.ptr(x), O.ptr(y), O.ptr(b), O.ptr(c), O.ptr(d), 50, O.ptr(xe), O.ptr(ye))
    np.testing.assert_allclose(ye, 1.5 - 2 * xe + 0.7 * xe ** 2 - 0.3 * xe ** 3, atol=1e-11)


def test_xi_tables_match_reference(tables):
    r768, tabs = tables
    rq = GOLD["xi_r"]
    for f in range(21):
        xi = O.XiHandle(0, r768, tabs[f])(rq)
        ref = GOLD["c1_xi"][f]
        scale = np.abs(ref).max()
        np.testing.assert_allclose(xi, ref, rtol=1e-12, atol=1e-12 * scale, err_msg=f"function {f}")


def test_band_params_match_reference(orc):
    k, P = GOLD["pk_k"], GOLD["pk_P"]
    got = [orc.orc_band_param(len(k), O.ptr(k), O.ptr(P), lo, hi)
           for lo, hi in zip(GOLD["c1_band_lo"], GOLD["c1_band_hi"])]
    np.testing.assert_allclose(got, GOLD["c1_params"], rtol=1e-13)


@pytest.mark.parametrize("klo,khi", [(0.005, 0.01125), (0.01, 0.0195), (0.2438, 0.25)])
def test_band_closed_form(orc, klo, khi):
    # against direct numerical integration of (1/2pi^2) k^2 j0(kr) over the band
    for r in [0.0, 1e-6, 0.3, 2.0, 12.5, 70.0, 399.0, 1500.0]:
        num = si.quad(lambda k: k * k * (np.sinc(k * r / np.pi)), klo, khi, epsabs=1e-300, epsrel=1e-13, limit=200)[0]
        num /= 2 * np.pi ** 2
        got = orc.orc_band_xi(klo, khi, r)
        assert abs(got - num) <= 1e-13 * (khi ** 3 / (6 * np.pi ** 2)), (r, got, num)
    assert orc.orc_band_xi(klo, khi, 0.0) == pytest.approx((khi ** 3 - klo ** 3) / (6 * np.pi ** 2), rel=1e-15)


def test_box_separation_known_answers(orc):
    # tests/testBoxSurvey.cpp:15-29
    p1 = np.array([10.0, 20.0, 30.0])
    p2 = np.array([70.0, 80.0, 90.0])
    assert orc.orc_box_r(O.ptr(p1), O.ptr(p2), 100.0) == pytest.approx(math.sqrt(3 * 40 * 40), abs=1e-10)
    p2[2] = 50
    assert orc.orc_box_r(O.ptr(p1), O.ptr(p2), 100.0) == pytest.approx(math.sqrt(2 * 40 * 40 + 20 * 20), abs=1e-10)
    got = [orc.orc_box_r(O.ptr(a), O.ptr(b), 100.0) for a, b in zip(GOLD["sep_p1"], GOLD["sep_p2"])]
    np.testing.assert_allclose(got, GOLD["sep_r"], rtol=1e-15)


# ---- pair integrals ---------------------------------------------------------------

def _c1_cells():
    return synth.box_cells(12)


def test_gm_matches_reference_pairs(orc, tables):
    r768, tabs = tables
    cells = _c1_cells()
    for t, f in enumerate(GOLD["pair_funcs"]):
        xi = O.XiHandle(0, r768, tabs[f])
        for i, (a, b) in enumerate(GOLD["pairs"]):
            err = C.c_double()
            q = orc.orc_gm_integrate(C.byref(xi.s), 1, 400.0, cells[a:a + 1].ctypes.data, cells[b:b + 1].ctypes.data,
                                     1e-5, 1e-10, 50000000, C.byref(err), None)
            ref = GOLD["pair_Q"][t, i]
            assert q == pytest.approx(ref, rel=1e-12, abs=1e-17), (f, a, b)
            assert err.value == pytest.approx(GOLD["pair_err"][t, i], rel=1e-6, abs=1e-20)


def test_gm_constant_xi_known_answer(orc):
    # tests/testsig.cpp:19-39: xi = 1 -> Q = 1 (S = 0.125 with Nbar 1, 0.125)
    xi = O.XiHandle(3, a=1.0, b=0.0)
    c1 = mk_cell((0, 0, 0), (0, 0, 0), (1, 1, 1), 1.0)
    c2 = mk_cell((1, 1, 1), (1, 1, 1), (1.5, 1.5, 1.5), 0.125)
    q = orc.orc_gm_integrate(C.byref(xi.s), 1, 1000.0, c1.ctypes.data, c2.ctypes.data, 1e-5, 1e-10, 50000000, None,
                             None)
    assert 1.0 * 0.125 * q == pytest.approx(0.125, abs=1e-12)


def test_gaussian_known_answer_tight(orc):
    # tests/testsig.cpp:41-66 with tolerances that are not swapped into the
    # absolute slot: periodic Gaussian over the whole torus integrates exactly.
    L, sig = 10.0, 1 / math.sqrt(2)
    A, b = (2 * math.pi * sig * sig) ** -1.5, 1 / (2 * sig * sig)
    xi = O.XiHandle(3, a=A, b=b)
    c = mk_cell((0, 0, 0), (0, 0, 0), (L, L, L), L ** 3)
    expect = (L * math.erf(L / (2 * math.sqrt(2) * sig))) ** 3 / L ** 6
    # the tent rule: exact reduction, converges spectrally for this smooth xi
    q_tent = orc.orc_tent_composite(C.byref(xi.s), L, c.ctypes.data, c.ctypes.data, 8, 8)
    assert q_tent == pytest.approx(expect, rel=1e-10)


def test_gm_tight_against_tent_rule(orc):
    # two independent discretisations of the same 6-D mean must agree when the
    # adaptive rule is run to a tight absolute tolerance (smooth, aperiodic)
    xi = O.XiHandle(3, a=1.0, b=0.3)
    for d in [(0, 0, 0), (1, 0, 0), (1, 2, 1)]:
        c1 = mk_cell((0, 0, 0), (0, 0, 0), (1.0, 1.0, 1.0))
        c2 = mk_cell(d, tuple(float(x) for x in d), tuple(x + 1.0 for x in d))
        q_gm = orc.orc_gm_integrate(C.byref(xi.s), 1, 1e9, c1.ctypes.data, c2.ctypes.data, 1e-12, 1e-14, 50000000,
                                    None, None)
        q_t = orc.orc_tent_composite(C.byref(xi.s), 1e9, c1.ctypes.data, c2.ctypes.data, 8, 8)
        assert q_gm == pytest.approx(q_t, rel=1e-10), d


def test_gaussian_reference_tolerance_quirk(orc):
    # With the reference's own call (epsrel=1e-5 lands in the absolute slot) the
    # Gaussian test misses its 5e-5 tolerance: the oracle reproduces that.
    L, sig = 10.0, 1 / math.sqrt(2)
    A, b = (2 * math.pi * sig * sig) ** -1.5, 1 / (2 * sig * sig)
    xi = O.XiHandle(3, a=A, b=b)
    c = mk_cell((0, 0, 0), (0, 0, 0), (L, L, L), L ** 3)
    expect = (L * math.erf(L / (2 * math.sqrt(2) * sig))) ** 3 / L ** 6
    q = orc.orc_gm_integrate(C.byref(xi.s), 1, L, c.ctypes.data, c.ctypes.data, 1e-5, 1e-12, 50000000, None, None)
    assert abs(q * 1e6 - 1000.6299173715882) < 1e-9   # measured from the reference build
    assert abs(q / expect - 1) > 5e-5


def test_tent_gauss_against_brute_force(orc):
    # the exact tent reduction equals the 6-D mean: compare with a brute-force
    # 6-D Gauss-Legendre product rule on a smooth xi (tiny cells)
    xi = O.XiHandle(2, a=1.3, b=0.05)   # exponential, smooth away from r=0
    for d in [(2, 0, 0), (1, 3, 2), (4, 4, 4)]:
        lo1, hi1 = (0.0, 0.0, 0.0), (5.0, 5.0, 5.0)
        lo2 = tuple(5.0 * x for x in d)
        hi2 = tuple(5.0 * x + 5.0 for x in d)
        c1, c2 = mk_cell((0, 0, 0), lo1, hi1), mk_cell(d, lo2, hi2)
        brute = orc.orc_brute6(C.byref(xi.s), 1e9, c1.ctypes.data, c2.ctypes.data, 10)
        tent = orc.orc_tent_gauss(C.byref(xi.s), 1e9, c1.ctypes.data, c2.ctypes.data, 16)
        assert tent == pytest.approx(brute, rel=1e-9), d


def test_tent_gauss_converges_for_bands(orc):
    # a wrong weight or node sign would not converge to the composite value
    xi = O.XiHandle(1, a=0.2438, b=0.25)
    h = 12.5
    for d in [(1, 0, 0), (2, 3, 1), (7, 0, 5)]:
        c1 = mk_cell((0, 0, 0), (0, 0, 0), (h, h, h))
        c2 = mk_cell(d, tuple(h * x for x in d), tuple(h * x + h for x in d))
        ref = orc.orc_tent_composite(C.byref(xi.s), 400.0, c1.ctypes.data, c2.ctypes.data, 8, 8)
        g6 = orc.orc_tent_gauss(C.byref(xi.s), 400.0, c1.ctypes.data, c2.ctypes.data, 6)
        g16 = orc.orc_tent_gauss(C.byref(xi.s), 400.0, c1.ctypes.data, c2.ctypes.data, 16)
        scale = (0.25 ** 3 - 0.2438 ** 3) / (6 * np.pi ** 2)
        assert abs(g16 - ref) < 1e-12 * scale
        assert abs(g6 - ref) < 2e-3 * scale


def test_sig_block_matches_reference(orc, tables):
    r768, tabs = tables
    cells = _c1_cells()
    rows = GOLD["blk_rows"].astype(np.int32)
    cols = np.arange(1728, dtype=np.int32)
    for t, f in enumerate(GOLD["blk_funcs"]):
        xi = O.XiHandle(0, r768, tabs[f])
        s = np.zeros((len(rows), 1728))
        rc = orc.orc_sig_block(C.byref(xi.s), cells.ctypes.data, 1728, 12, 12, 12, 400.0, O.ptr(rows), len(rows),
                               O.ptr(cols), 1728, O.ptr(s), 1728, 0, 1e-5, 1e-10, 50000000, 0, 0)
        assert rc == 0
        ref = GOLD["blk_S"][t]
        np.testing.assert_allclose(s, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max(), err_msg=f"f={f}")


def test_sig_block_subblock_stride(orc, tables):
    r768, tabs = tables
    cells = synth.box_cells(6, 0.0, 120.0)
    rows, cols = GOLD["b6_rows"].astype(np.int32), GOLD["b6_cols"].astype(np.int32)
    for t, f in enumerate((0, 5)):
        xi = O.XiHandle(0, r768, tabs[f])
        s = np.full((len(rows), 9), -7.0)
        orc.orc_sig_block(C.byref(xi.s), cells.ctypes.data, 216, 6, 6, 6, 1e9, O.ptr(rows), len(rows), O.ptr(cols),
                          len(cols), O.ptr(s), 9, 0, 1e-5, 1e-10, 50000000, 0, 0)
        ref = GOLD["b6_S"][t]
        np.testing.assert_allclose(s, ref, rtol=1e-12, atol=1e-12 * np.abs(ref[:, :7]).max())
        assert np.all(s[:, 7:] == -7.0)


def test_testmodel_gm_matches_reference(orc):
    cells = _c1_cells()
    forms = [(3, 2.5, 0.01), (3, 1.0, 0.01), (4, -2.5 * 0.01, 0.01)]   # GetXi, GetXiDeriv(0), GetXiDeriv(1)
    for f, (kind, a, b) in enumerate(forms):
        xi = O.XiHandle(kind, a=a, b=b)
        for i, (p, q) in enumerate(GOLD["pairs"]):
            v = orc.orc_gm_integrate(C.byref(xi.s), 1, 400.0, cells[p:p + 1].ctypes.data, cells[q:q + 1].ctypes.data,
                                     1e-5, 1e-10, 50000000, None, None)
            assert v == pytest.approx(GOLD["test_gauss_Q"][f, i], rel=1e-12, abs=1e-300)


# ---- dense linear algebra -----------------------------------------------------------

def test_inverse_and_sqrtm_known_answers(orc):
    # tests/testMyMatrix.cpp:67-104
    X = np.array([[5, -4, 1, 0, 0], [-4, 6, -4, 1, 0], [1, -4, 6, -4, 1], [0, 1, -4, 6, -4], [0, 0, 1, -4, 5]], float)
    Y = np.zeros((5, 5))
    orc.orc_sqrtm(5, O.ptr(X), O.ptr(Y))
    T = 2 * np.eye(5) - np.eye(5, k=1) - np.eye(5, k=-1)
    np.testing.assert_allclose(Y, T, atol=1e-10)
    A = np.array([[1, 1, 2], [2, 4, -3], [3, 6, -5]], float)
    B = np.zeros((3, 3))
    assert orc.orc_inverse(3, O.ptr(A), O.ptr(B)) == 0
    np.testing.assert_allclose(B, [[2, -17, 11], [-1, 11, -7], [0, 3, -2]], atol=1e-10)


def test_sym_eigen_jx_known_answer(orc):
    # tests/testeig.cpp:78-121: J_x for j=5 has eigenvalues -5..5
    n, j = 11, 5.0
    A = np.zeros((n, n))
    for a in range(n):
        for b in range(n):
            m = j - b
            if a == b - 1:
                A[a, b] = 0.5 * math.sqrt(j * (j + 1) - m * (m + 1))
            elif a == b + 1:
                A[a, b] = 0.5 * math.sqrt(j * (j + 1) - m * (m - 1))
    A = 0.5 * (A + A.T)
    ev, V = np.zeros(n), np.zeros((n, n))
    orc.orc_sym_eigen(n, O.ptr(A), O.ptr(ev), O.ptr(V))
    np.testing.assert_allclose(ev, np.arange(5, -6, -1), atol=1e-12)
    np.testing.assert_allclose(V @ A @ V.T, np.diag(ev), atol=1e-12)


def test_project_brute_force(orc):
    rng = np.random.default_rng(3)
    n, m = 37, 9
    S = rng.normal(size=(n, n))
    S = S + S.T
    B = rng.normal(size=(m, n))
    P = np.zeros(m * (m + 1) // 2)
    orc.orc_project(n, O.ptr(S), m, O.ptr(B), O.ptr(P))
    full = B @ S @ B.T
    k = 0
    for j in range(m):
        for i in range(j + 1):
            assert P[k] == pytest.approx(full[i, j], rel=1e-12, abs=1e-12)
            k += 1


def _packed(Mfull):
    m = Mfull.shape[0]
    return np.array([Mfull[i, j] for j in range(m) for i in range(j + 1)])


def test_estimate_diagonal_case(orc):
    # D_p diagonal: F_pq = 1/2 sum_i D_p,ii D_q,ii / C_i^2; identity mixing -> M = diag(1/rowsum(F))
    rng = np.random.default_rng(4)
    M, npar = 6, 3
    lam = rng.uniform(0.1, 2.0, M)
    Dd = rng.uniform(0.1, 1.0, (npar, M))
    D = np.stack([_packed(np.diag(Dd[p])) for p in range(npar)])
    y = rng.normal(size=M)
    prior = rng.uniform(0.5, 2, npar)
    out = {k: np.zeros(npar) for k in ("phat", "q", "f")}
    out.update({k: np.zeros((npar, npar)) for k in ("cov", "F", "W", "M")})
    args = [O.ptr(out[k]) for k in ("phat", "cov", "F", "W", "M", "q", "f")]
    assert orc.orc_estimate(M, npar, O.ptr(lam), O.ptr(D), O.ptr(y), O.ptr(prior), 2, *args) == 0
    C_ = 1 + lam
    Fx = 0.5 * np.einsum("pi,qi->pq", Dd / C_, Dd / C_)
    np.testing.assert_allclose(out["F"], Fx, rtol=1e-13)
    qx = 0.5 * np.einsum("pi,i->p", Dd / C_ ** 2, y ** 2)
    np.testing.assert_allclose(out["q"], qx, rtol=1e-13)
    R = C_ - prior @ Dd
    fx = 0.5 * np.einsum("pi,i->p", Dd / C_ ** 2, R)
    np.testing.assert_allclose(out["f"], fx, rtol=1e-12)
    Mx = np.diag(1 / Fx.sum(1))
    np.testing.assert_allclose(out["M"], Mx, rtol=1e-13)
    np.testing.assert_allclose(out["W"].sum(1), 1, rtol=1e-13)
    np.testing.assert_allclose(out["phat"], Mx @ (qx - fx), rtol=1e-12)
    np.testing.assert_allclose(out["cov"], Mx @ Fx @ Mx.T, rtol=1e-12)


def test_estimate_inverse_mixing_gives_identity_window(orc):
    rng = np.random.default_rng(6)
    M, npar = 8, 3
    lam = rng.uniform(0.1, 2.0, M)
    D = []
    for p in range(npar):
        A = rng.normal(size=(M, M))
        D.append(_packed(A + A.T))
    D = np.stack(D)
    y, prior = rng.normal(size=M), rng.uniform(0.5, 2, npar)
    for mixing in (0, 1):
        out = {k: np.zeros(npar) for k in ("phat", "q", "f")}
        out.update({k: np.zeros((npar, npar)) for k in ("cov", "F", "W", "M")})
        args = [O.ptr(out[k]) for k in ("phat", "cov", "F", "W", "M", "q", "f")]
        assert orc.orc_estimate(M, npar, O.ptr(lam), O.ptr(D), O.ptr(y), O.ptr(prior), mixing, *args) == 0
        np.testing.assert_allclose(out["W"].sum(1), 1, rtol=1e-12)
        if mixing == 0:
            np.testing.assert_allclose(out["W"], np.eye(npar), atol=1e-12)
        # F from full matrices: 1/2 tr(Ci Dp Ci Dq)
        Ci = np.diag(1 / (1 + lam))
        full = []
        for p in range(npar):
            Fm = np.zeros((M, M))
            k = 0
            for j in range(M):
                for i in range(j + 1):
                    Fm[i, j] = Fm[j, i] = D[p][k]
                    k += 1
            full.append(Fm)
        Fx = np.array([[0.5 * np.trace(Ci @ full[p] @ Ci @ full[q]) for q in range(npar)] for p in range(npar)])
        np.testing.assert_allclose(out["F"], Fx, rtol=1e-12)


# ---- MLP oracle -----------------------------------------------------------------------

def test_mlp_gradient_finite_differences(orc):
    nw = orc.orc_mlp_nweights()
    assert nw == 8 * 128 + 128 + 3 * (128 * 128 + 128) + 128 + 1
    rng = np.random.default_rng(8)
    w = rng.normal(scale=0.1, size=nw)
    x = rng.normal(size=(5, 8))
    t = rng.normal(size=5)
    g = np.zeros(nw)
    orc.orc_mlp_loss_grad(O.ptr(w), 5, O.ptr(x), O.ptr(t), O.ptr(g))
    for idx in [0, 17, 1100, 1152 + 5000, nw - 129, nw - 2, nw - 1]:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        scratch = np.zeros(nw)
        lp = orc.orc_mlp_loss_grad(O.ptr(wp), 5, O.ptr(x), O.ptr(t), O.ptr(scratch))
        lm = orc.orc_mlp_loss_grad(O.ptr(wm), 5, O.ptr(x), O.ptr(t), O.ptr(scratch))
        assert g[idx] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-10)


# ---- spherical surveys (src/SphericalSurvey.cpp:19-23, src/sig.cpp:125-213,436-482) ----

def _sph_cell(r, mu, phi, d=(0, 0, 0), nbar=1.0):
    c = np.zeros(1, dtype=synth.CELL_DTYPE)
    c["d1"], c["d2"], c["d3"] = d
    c["min1"], c["max1"] = r
    c["min2"], c["max2"] = mu
    c["min3"], c["max3"] = phi
    c["Nbar"] = nbar
    return c


def test_spherical_separation_known_answers(orc):
    def sep(p1, p2):
        a, b = np.array(p1, float), np.array(p2, float)
        return orc.orc_sph_r(O.ptr(a), O.ptr(b))
    assert sep((300.0, 0.3, 1.0), (500.0, 0.3, 1.0)) == pytest.approx(200.0, abs=1e-9)    # same ray
    assert sep((400.0, 0.2, 0.5), (400.0, -0.2, 0.5 + np.pi)) == pytest.approx(800.0, rel=1e-12)   # antipodes
    assert sep((300.0, 1.0, 0.0), (400.0, 0.0, 2.0)) == pytest.approx(500.0, rel=1e-12)   # pole vs equator
    assert sep((250.0, 0.7, 3.0), (250.0, 0.7, 3.0)) == 0.0


def test_spherical_pair_xi_one(orc):
    xi = O.XiHandle(3, a=1.0, b=0.0)
    c1 = _sph_cell((200.0, 300.0), (0.5, 0.75), (0.0, 0.4))
    c2 = _sph_cell((300.0, 420.0), (0.75, 1.0), (0.4, 1.1))
    q = orc.orc_gm_integrate(C.byref(xi.s), 2, 0.0, c1.ctypes.data, c2.ctypes.data, 1e-5, 1e-10, 50000000, None, None)
    assert q == pytest.approx(1.0, abs=1e-13)


def test_spherical_pair_vs_brute_force(orc):
    """Adaptive Genz-Malik on a spherical cell pair vs tensor Gauss-Legendre
    of the same volume-uniform integrand (smooth Gaussian xi)."""
    xi = O.XiHandle(3, a=2.0, b=1.0 / (2 * 150.0 ** 2))
    c1 = _sph_cell((300.0, 360.0), (0.6, 0.7), (0.0, 0.3))
    c2 = _sph_cell((360.0, 420.0), (0.7, 0.8), (0.3, 0.6))
    q = orc.orc_gm_integrate(C.byref(xi.s), 2, 0.0, c1.ctypes.data, c2.ctypes.data, 1e-13, 1e-13, 50000000, None, None)
    b = orc.orc_brute6_sph(C.byref(xi.s), c1.ctypes.data, c2.ctypes.data, 6)
    assert q == pytest.approx(b, rel=1e-10)


def test_spherical_block_semantics(orc):
    """The arc cache (one integral per arc pair and |d_phi|) against every
    element integrated on its own: phi rotations leave the integrand
    unchanged, so the two agree to rounding; xi = 1 gives sqrt(Nbar Nbar)."""
    cells = synth.shell_cells(2, 2, 5, rmin=300.0, rmax=500.0)
    N = len(cells)
    rows = np.array([0, 3, 7, 12, 19], dtype=np.int32)
    cols = np.arange(N, dtype=np.int32)
    band = O.XiHandle(1, a=0.02, b=0.05)
    S = np.zeros((len(rows), N))
    assert orc.orc_sig_block_sph(C.byref(band.s), cells.ctypes.data, N, 2, 2, 5, O.ptr(rows), len(rows), O.ptr(cols),
                                 N, O.ptr(S), N, 1e-5, 1e-10, 50000000) == 0
    for i, a in enumerate(rows):
        for b in range(N):
            lo, hi = (a, b) if a <= b else (b, a)
            q = orc.orc_gm_integrate(C.byref(band.s), 2, 0.0, cells[lo:lo + 1].ctypes.data,
                                     cells[hi:hi + 1].ctypes.data, 1e-5, 1e-10, 50000000, None, None)
            exp = np.sqrt(cells["Nbar"][a] * cells["Nbar"][b]) * q
            assert S[i, b] == pytest.approx(exp, rel=1e-9, abs=1e-12 * abs(S).max()), (a, b)
    one = O.XiHandle(3, a=1.0, b=0.0)
    assert orc.orc_sig_block_sph(C.byref(one.s), cells.ctypes.data, N, 2, 2, 5, O.ptr(rows), len(rows), O.ptr(cols),
                                 N, O.ptr(S), N, 1e-5, 1e-10, 50000000) == 0
    np.testing.assert_allclose(S, np.sqrt(np.outer(cells["Nbar"][rows], cells["Nbar"])), rtol=1e-13)
