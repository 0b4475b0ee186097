"""GPU parity of the xi model, pair quadrature and assembly
against the reference golden vectors and the FP64 oracle."""
import ctypes as C
import os

import numpy as np
import pytest

from opsec_amd import lib, synth
from tests import oracles as O

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
LO, HI = GOLD["c1_band_lo"], GOLD["c1_band_hi"]


@pytest.fixture
def c1(ctx):
    """The C1 RealModel (reference xi tables), set up afresh for every test."""
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], LO, HI, lib.OPSEC_XI_TABLE)
    return ctx


def test_device_is_gfx950(ctx):
    ncu, arch = ctx.device_info()
    print("device", arch, ncu, "CUs", flush=True)
    assert arch.startswith("gfx950") and ncu > 0


def test_xi_tables_vs_reference(c1):
    rq = GOLD["xi_r"]
    for f in range(21):
        got = c1.xi(f, rq)
        ref = GOLD["c1_xi"][f]
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max(), err_msg=f"f={f}")


def test_band_params_vs_reference(c1):
    np.testing.assert_allclose(c1.params(), GOLD["c1_params"], rtol=1e-13)


def test_closed_form_bands_vs_oracle(ctx):
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], LO, HI, lib.OPSEC_XI_CLOSED)
    orc = O.oracle()
    r = np.concatenate([[0.0, 1e-9, 1e-3], np.linspace(0.01, 800, 301)])
    for p in (0, 7, 19):
        got = ctx.xi(1 + p, r)
        exp = np.array([orc.orc_band_xi(LO[p], HI[p], x) for x in r])
        scale = HI[p] ** 3 / (6 * np.pi ** 2)
        assert np.abs(got - exp).max() < 1e-14 * scale


def test_gm_pairs_vs_reference(c1):
    cells = synth.box_cells(12)
    sv = lib.survey_box(12)
    q = c1.quad(lib.OPSEC_QUAD_GM)
    for t, f in enumerate(GOLD["pair_funcs"]):
        for i, (a, b) in enumerate(GOLD["pairs"]):
            Q, err = c1.pair_integral(cells[a], cells[b], int(f), sv, q)
            assert Q == pytest.approx(GOLD["pair_Q"][t, i], rel=1e-12, abs=1e-17), (f, a, b)
            assert err == pytest.approx(GOLD["pair_err"][t, i], rel=1e-6, abs=1e-20)


def test_sig_block_gm_vs_reference(c1):
    cells = synth.box_cells(12)
    sv = lib.survey_box(12)
    q = c1.quad(lib.OPSEC_QUAD_GM)
    rows = GOLD["blk_rows"]
    cols = np.arange(1728)
    for t, f in enumerate(GOLD["blk_funcs"]):
        s = c1.sig_block(cells, rows, cols, int(f), sv, q)
        ref = GOLD["blk_S"][t]
        np.testing.assert_allclose(s, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max(), err_msg=f"f={f}")


def test_uploaded_xifunc_tables_bind_sig_block(c1):
    """A caller's XiFunc given as its 768 QSpline node values
    (opsec_model_tables): the RealModel's own tables uploaded that way give the
    reference's xi and ComputeSignalMatrixBlock blocks, and the same bits as
    the RealModel path."""
    tabs = np.array([c1.table(f)[1] for f in range(21)])
    params = c1.params()
    cells = synth.box_cells(12)
    sv = lib.survey_box(12)
    q = c1.quad(lib.OPSEC_QUAD_GM)
    rows, cols = GOLD["blk_rows"], np.arange(1728)
    native = [c1.sig_block(cells, rows, cols, int(f), sv, q) for f in GOLD["blk_funcs"]]
    c1.model_tables(tabs, params)
    assert c1.nparams() == 20
    np.testing.assert_array_equal(c1.params(), params)
    for f in range(21):
        ref = GOLD["c1_xi"][f]
        np.testing.assert_allclose(c1.xi(f, GOLD["xi_r"]), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    for t, f in enumerate(GOLD["blk_funcs"]):
        s = c1.sig_block(cells, rows, cols, int(f), sv, q)
        np.testing.assert_array_equal(s, native[t])
        ref = GOLD["blk_S"][t]
        np.testing.assert_allclose(s, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max(), err_msg=f"f={f}")
    with pytest.raises(lib.OpsecError, match="not finite"):
        bad = tabs.copy()
        bad[3, 7] = np.nan
        c1.model_tables(bad)


def test_sig_block_stride_and_subset(c1):
    cells = synth.box_cells(6, 0.0, 120.0)
    sv = lib.survey_box(6, 0.0, 120.0, L=1e9)
    q = c1.quad(lib.OPSEC_QUAD_GM)
    rows, cols = GOLD["b6_rows"], GOLD["b6_cols"]
    for t, f in enumerate((0, 5)):
        s = c1.sig_block(cells, rows, cols, f, sv, q, lld=9)
        ref = GOLD["b6_S"][t][:, :7]
        np.testing.assert_allclose(s[:, :7], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_sig_block_rejects_bad_stride(c1):
    cells = synth.box_cells(4)
    sv = lib.survey_box(4)
    q = c1.quad(lib.OPSEC_QUAD_GM)
    out = np.zeros((4, 8))
    rows = np.arange(4, dtype=np.int32)
    cols = np.arange(8, dtype=np.int32)
    rc = c1.lib.opsec_sig_block(c1.h, cells.ctypes.data, len(cells), rows.ctypes.data, 4, cols.ctypes.data, 8, 0,
                                C.byref(sv), C.byref(q), out.ctypes.data, 5)
    assert rc == -1
    assert b"invalid array stride" in c1.lib.opsec_last_error(c1.h)


def test_testmodel_vs_reference(ctx):
    ctx.model_test(1, 2.5, 0.01)
    cells = synth.box_cells(12)
    sv = lib.survey_box(12)
    q = ctx.quad(lib.OPSEC_QUAD_GM)
    for f in range(3):
        for i, (a, b) in enumerate(GOLD["pairs"]):
            Q, _ = ctx.pair_integral(cells[a], cells[b], f, sv, q)
            assert Q == pytest.approx(GOLD["test_gauss_Q"][f, i], rel=1e-12, abs=1e-300), (f, a, b)


def test_gauss_sig_block_vs_oracle(c1):
    """Tent-kernel Gauss: GPU vs oracle, 8 nodes per axis, composite self pair."""
    orc = O.oracle()
    r768, tabs = O.oracle_tables(GOLD["pk_k"], GOLD["pk_P"], LO, HI)
    cells = synth.box_cells(12)
    sv = lib.survey_box(12)
    q = c1.quad(lib.OPSEC_QUAD_GAUSS, nodes=8, near=1)
    rows = np.array([0, 1, 57, 1727], dtype=np.int32)
    cols = np.arange(1728, dtype=np.int32)
    for f in (0, 1, 20):
        s = c1.sig_block(cells, rows, cols, f, sv, q)
        xi = O.XiHandle(0, r768, tabs[f])
        ref = np.zeros((len(rows), 1728))
        orc.orc_sig_block(C.byref(xi.s), cells.ctypes.data, 1728, 12, 12, 12, 400.0, O.ptr(rows), len(rows),
                          O.ptr(cols), 1728, O.ptr(ref), 1728, 1, 0, 0, 0, 8, 1)
        np.testing.assert_allclose(s, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max(), err_msg=f"f={f}")


def test_pair_tables_and_assembly_vs_reference(c1):
    """Device-resident pair tables and assembly for the full 12^3 matrix vs the reference rows."""
    cells = synth.box_cells(12)
    sv = lib.survey_box(12)
    c1.cells_upload(cells, sv)
    c1.pair_tables([0, 1, 20], c1.quad(lib.OPSEC_QUAD_GM))
    for t, f in enumerate(GOLD["blk_funcs"]):
        S = c1.assemble(int(f), 1728)
        assert np.array_equal(S, S.T)
        ref = GOLD["blk_S"][t]
        np.testing.assert_allclose(S[GOLD["blk_rows"]], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    st = c1.pair_stats()
    assert st["integrals"] == 3 * 1728 and st["nodes"] >= 149 * 3 * 1728


def test_gauss_pair_tables_c2_shape(c1):
    """32^3 grid: all 32768 keys x 3 functions; translation invariance and the
    periodic mirror symmetry Q(k) = Q(N-k) hold to rounding."""
    n = 32
    cells = synth.box_cells(n)
    sv = lib.survey_box(n)
    c1.cells_upload(cells, sv)
    c1.pair_tables([0, 1, 20], c1.quad(lib.OPSEC_QUAD_GAUSS, nodes=6, near=1))
    for f in (0, 1, 20):
        Q = c1.pair_table(f).reshape(n, n, n)
        assert np.all(np.isfinite(Q))
        for k in (1, 5, 13):
            m = n - k
            np.testing.assert_allclose(Q[k], Q[m], rtol=1e-9, atol=1e-12 * np.abs(Q).max())
            np.testing.assert_allclose(Q[:, k], Q[:, m], rtol=1e-9, atol=1e-12 * np.abs(Q).max())
        # permutation symmetry of the cubic grid
        np.testing.assert_allclose(Q, Q.transpose(1, 0, 2), rtol=1e-12, atol=1e-15 * np.abs(Q).max())
        np.testing.assert_allclose(Q, Q.transpose(2, 1, 0), rtol=1e-12, atol=1e-15 * np.abs(Q).max())


def test_spherical_sig_block_vs_oracle(ctx):
    """Spherical shell: the device arc-cache block (GM pairs) against the
    oracle's restatement of ComputeSignalMatrixBlockS, on row/column subsets
    with a padded stride, for the closed-form band and the test Gaussian."""
    nr, nmu, nphi = 3, 2, 6
    cells = synth.shell_cells(nr, nmu, nphi, rmin=250.0, rmax=550.0)
    N = len(cells)
    sv = lib.survey_sph(nr, nmu, nphi, 250.0, 550.0)
    rows = np.array([0, 5, 6, 13, 20, 35], dtype=np.int32)
    cols = np.array([1, 2, 3, 7, 11, 12, 18, 24, 29, 30, 33, 35], dtype=np.int32)
    lld = len(cols) + 3
    orc = O.oracle()
    for kind, setup in (("band", lambda: ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], [0.02], [0.06], lib.OPSEC_XI_CLOSED)),
                        ("gauss", lambda: ctx.model_test(1, 2.5, 1.0 / (2 * 120.0 ** 2)))):
        setup()
        f = 1 if kind == "band" else 0
        xi = O.XiHandle(1, a=0.02, b=0.06) if kind == "band" else O.XiHandle(3, a=2.5, b=1.0 / (2 * 120.0 ** 2))
        q = ctx.quad(lib.OPSEC_QUAD_GM)
        got = ctx.sig_block(cells, rows, cols, f, sv, q, lld=lld)
        ref = np.zeros((len(rows), lld))
        assert orc.orc_sig_block_sph(C.byref(xi.s), cells.ctypes.data, N, nr, nmu, nphi, O.ptr(rows), len(rows),
                                     O.ptr(cols), len(cols), O.ptr(ref), lld, q.epsrel, q.epsabs, q.maxeval) == 0
        np.testing.assert_allclose(got[:, :len(cols)], ref[:, :len(cols)], rtol=1e-12,
                                   atol=1e-13 * np.abs(ref).max(), err_msg=kind)
