"""GPU parity of the projection and the Fisher / estimate step
against the FP64 oracle."""
import ctypes as C
import os

import numpy as np
import pytest

from opsec_amd import lib, synth
from tests import oracles as O

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def orc_project(S, B):
    n, m = S.shape[0], B.shape[0]
    out = np.zeros(m * (m + 1) // 2)
    S = np.ascontiguousarray(S)
    B = np.ascontiguousarray(B)
    O.oracle().orc_project(n, O.ptr(S), m, O.ptr(B), O.ptr(out))
    return out


def rand_sym(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n))
    return A + A.T


@pytest.mark.parametrize("n,m", [(300, 70), (128, 128), (257, 129)])
def test_project_dense_fp64(ctx, n, m):
    S = rand_sym(n, n)
    B = np.random.default_rng(m).normal(size=(m, n))
    got = ctx.project_dense(S, B, lib.OPSEC_PREC_FP64)
    ref = orc_project(S, B)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("prec,tol", [(lib.OPSEC_PREC_FP32, 2e-5), (lib.OPSEC_PREC_BF16X3, 5e-5)])
def test_project_dense_mixed(ctx, prec, tol):
    n, m = 400, 96
    S = rand_sym(n, 1)
    B = np.random.default_rng(2).normal(size=(m, n))
    got = ctx.project_dense(S, B, prec)
    ref = orc_project(S, B)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("mixed precision", prec, "max rel err", err)
    assert err < tol


def _box_setup(ctx, n=8, nb=4, lo=0.0, hi=200.0):
    cells = synth.box_cells(n, lo, hi)
    sv = lib.survey_box(n, lo, hi)
    k, P = GOLD["pk_k"], GOLD["pk_P"]
    blo, bhi = synth.linear_bands(nb, 0.02, 0.2)
    ctx.model_real(k, P, blo, bhi, lib.OPSEC_XI_TABLE)
    ctx.cells_upload(cells, sv)
    return cells, (k, P, blo, bhi)


def test_project_pair_tables_fp64_vs_oracle(ctx):
    """Fused S_p generation + projection vs oracle assembly + oracle projection."""
    n, nb = 8, 4
    cells, (k, P, blo, bhi) = _box_setup(ctx, n, nb)
    N = len(cells)
    q = ctx.quad(lib.OPSEC_QUAD_GM)
    funcs = list(range(nb + 1))
    ctx.pair_tables(funcs, q)
    rng = np.random.default_rng(9)
    m = 100
    B = np.linalg.qr(rng.normal(size=(N, m)))[0].T.copy()
    ctx.modes_upload(B)
    got = ctx.project(funcs, lib.OPSEC_PREC_FP64)
    r768, tabs = O.oracle_tables(k, P, blo, bhi)
    rows = np.arange(N, dtype=np.int32)
    for t, f in enumerate(funcs):
        xi = O.XiHandle(0, r768, tabs[f])
        S = np.zeros((N, N))
        O.oracle().orc_sig_block(C.byref(xi.s), cells.ctypes.data, N, n, n, n, 200.0, O.ptr(rows), N, O.ptr(rows), N,
                                 O.ptr(S), N, 0, 1e-5, 1e-10, 50000000, 0, 0)
        ref = orc_project(S, B)
        np.testing.assert_allclose(got[t], ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max(), err_msg=f"f={f}")
    # mixed precisions against the FP64 result
    for prec in (lib.OPSEC_PREC_FP32, lib.OPSEC_PREC_BF16X3):
        mp = ctx.project(funcs, prec)
        err = np.abs(mp - got).max(axis=1) / np.abs(got).max(axis=1)
        print("pair-table projection, precision", prec, "max rel err per function", err)
        assert err.max() < 5e-5


def _packed(Mfull):
    m = Mfull.shape[0]
    return np.array([Mfull[i, j] for j in range(m) for i in range(j + 1)])


def _orc_estimate(lam, D, y, prior, mixing):
    npar, M = len(prior), len(lam)
    out = {k: np.zeros(npar) for k in ("phat", "q", "f")}
    out.update({k: np.zeros((npar, npar)) for k in ("cov", "F", "W", "M")})
    lam, D, y, prior = [np.ascontiguousarray(a, dtype=float) for a in (lam, D, y, prior)]
    rc = O.oracle().orc_estimate(M, npar, O.ptr(lam), O.ptr(D), O.ptr(y), O.ptr(prior), mixing,
                                 *[O.ptr(out[k]) for k in ("phat", "cov", "F", "W", "M", "q", "f")])
    assert rc == 0
    return out


@pytest.mark.parametrize("mixing", [0, 1, 2])
def test_estimate_fp64_vs_oracle(ctx, mixing):
    rng = np.random.default_rng(11 + mixing)
    M, npar = 150, 7
    lam = np.sort(rng.uniform(0.05, 5, M))[::-1].copy()
    D = []
    for p in range(npar):
        A = rng.normal(size=(M, M)) * 0.1
        A = A + A.T + np.diag(rng.uniform(0.5, 1.5, M))
        D.append(_packed(A))
    D = np.array(D)
    y = np.sqrt(1 + lam) * rng.normal(size=M)
    prior = rng.uniform(0.5, 2.0, npar)
    got = ctx.estimate(lam, D, y, prior, mixing, lib.OPSEC_PREC_FP64)
    ref = _orc_estimate(lam, D, y, prior, mixing)
    for k in ("F", "q", "f", "M", "W", "phat", "cov"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-10, atol=1e-10 * np.abs(ref[k]).max(), err_msg=k)
    np.testing.assert_allclose(got["W"].sum(1), 1.0, rtol=1e-12)


@pytest.mark.parametrize("npar,prec,tol", [(46, lib.OPSEC_PREC_FP64, 1e-10), (47, lib.OPSEC_PREC_FP64, 1e-10),
                                            (50, lib.OPSEC_PREC_FP64, 1e-10), (62, lib.OPSEC_PREC_FP32, 1e-5)])
def test_estimate_many_bands(ctx, npar, prec, tol):
    """Up to 46 parameters the Gram product is 48 x 48, above it 64 x 64 (the
    50 bands of the reference's RealModel.bands); 63 is rejected."""
    rng = np.random.default_rng(npar)
    M = 90
    lam = np.sort(rng.uniform(0.05, 5, M))[::-1].copy()
    D = np.array([_packed((lambda A: A + A.T + np.diag(rng.uniform(0.5, 1.5, M)))(rng.normal(size=(M, M)) * 0.1))
                  for _ in range(npar)])
    y = np.sqrt(1 + lam) * rng.normal(size=M)
    prior = rng.uniform(0.5, 2.0, npar)
    got = ctx.estimate(lam, D, y, prior, lib.OPSEC_MIX_IDENTITY, prec)
    ref = _orc_estimate(lam, D, y, prior, lib.OPSEC_MIX_IDENTITY)
    for k in ("F", "q", "f"):
        np.testing.assert_allclose(got[k], ref[k], rtol=tol, atol=tol * np.abs(ref[k]).max(), err_msg=k)
    if npar == 62:
        D2 = np.concatenate([D, D[:1]])
        with pytest.raises(lib.OpsecError, match="Nparams must be 1..62"):
            ctx.estimate(lam, D2, y, np.append(prior, 1.0), lib.OPSEC_MIX_IDENTITY, prec)


def test_estimate_fp32_traces(ctx):
    rng = np.random.default_rng(5)
    M, npar = 300, 5
    lam = rng.uniform(0.05, 5, M)
    D = np.array([_packed((lambda A: A + A.T)(rng.normal(size=(M, M)))) for _ in range(npar)])
    y = np.sqrt(1 + lam) * rng.normal(size=M)
    prior = rng.uniform(0.5, 2, npar)
    got = ctx.estimate(lam, D, y, prior, 0, lib.OPSEC_PREC_FP32)
    ref = _orc_estimate(lam, D, y, prior, 0)
    errF = np.abs(got["F"] - ref["F"]).max() / np.abs(ref["F"]).max()
    print("FP32 trace F rel err", errF)
    assert errF < 1e-5


def test_estimate_device_after_projection(ctx):
    n, nb = 6, 3
    cells, _ = _box_setup(ctx, n, nb, 0.0, 150.0)
    N = len(cells)
    ctx.pair_tables(list(range(nb + 1)), ctx.quad(lib.OPSEC_QUAD_GM))
    m = 50
    B = np.linalg.qr(np.random.default_rng(3).normal(size=(N, m)))[0].T.copy()
    ctx.modes_upload(B)
    D = ctx.project(list(range(1, nb + 1)), lib.OPSEC_PREC_FP64)
    rng = np.random.default_rng(4)
    lam = rng.uniform(0.1, 3, m)
    y = rng.normal(size=m)
    prior = ctx.params()
    a = ctx.estimate_device(lam, y, prior)
    b = ctx.estimate(lam, D, y, prior)
    for k in ("F", "q", "f", "phat"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-14 * np.abs(b[k]).max())


@pytest.mark.parametrize("n,shuffle", [(6, True), (20, False)])
def test_project_bf16x3_generated_layouts(ctx, n, shuffle):
    """Split-BF16x3 projection with S_p generated from the pair table inside the
    GEMM, against the FP64 projection of the same tables: padded tiles
    (n^3 not a multiple of 256), a shuffled cell order (general key path) and
    z-lines that straddle 16-column chunks (mixed fast and general paths)."""
    cells = synth.box_cells(n, 0.0, 25.0 * n)
    if shuffle:
        cells = cells[np.random.default_rng(3).permutation(len(cells))].copy()
    sv = lib.survey_box(n, 0.0, 25.0 * n)
    blo, bhi = synth.linear_bands(3, 0.02, 0.2)
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], blo, bhi, lib.OPSEC_XI_CLOSED)
    ctx.cells_upload(cells, sv)
    funcs = [0, 1, 2, 3]
    ctx.pair_tables(funcs, ctx.quad(lib.OPSEC_QUAD_GAUSS, nodes=4, near=1))
    N = len(cells)
    m = 72
    B = np.linalg.qr(np.random.default_rng(n).normal(size=(N, m)))[0].T.copy()
    ctx.modes_upload(B)
    ref = ctx.project(funcs, lib.OPSEC_PREC_FP64)
    got = ctx.project(funcs, lib.OPSEC_PREC_BF16X3)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("generated-S_p BF16x3 projection max rel err per function", err)
    assert err.max() < 5e-5


@pytest.mark.parametrize("tile,gen,dl", [(1, 0, 0), (2, 0, 0), (2, 1, 0), (2, 0, 1), (2, 1, 1), (2, 0, 2)])
def test_project_bf16x3_gemm_variants(ctx, tile, gen, dl):
    """The split-BF16x3 GEMM variants (256 x 128; 256 x 256 double-buffered,
    with S_p planes or S_p generated from the pair table, tiles staged through
    VGPRs or by DMA loads into LDS), with split-K in stage 2, against the FP64
    projection of the same pair tables; the Toeplitz path too with B in
    registers."""
    n = 10
    cells = synth.box_cells(n, 0.0, 250.0)
    sv = lib.survey_box(n, 0.0, 250.0)
    blo, bhi = synth.linear_bands(2, 0.02, 0.2)
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], blo, bhi, lib.OPSEC_XI_CLOSED)
    ctx.cells_upload(cells, sv)
    funcs = [0, 1, 2]
    ctx.pair_tables(funcs, ctx.quad(lib.OPSEC_QUAD_GAUSS, nodes=4, near=1))
    N = len(cells)
    m = 300
    B = np.linalg.qr(np.random.default_rng(11).normal(size=(N, m)))[0].T.copy()
    ctx.modes_upload(B)
    ref = ctx.project(funcs, lib.OPSEC_PREC_FP64)
    try:
        assert ctx.lib.opsec_diag_set_nt_tile(tile) == 0
        assert ctx.lib.opsec_diag_set_nt_gen(gen) == 0
        assert ctx.lib.opsec_diag_set_stage2_split(3) == 0
        assert ctx.lib.opsec_diag_set_toeplitz_stage2_split(3) == 0
        assert ctx.lib.opsec_diag_set_g3_dl(dl) == 0
        got = ctx.project(funcs, lib.OPSEC_PREC_BF16X3)
        ctx.set_projection(lib.OPSEC_PROJ_TOEPLITZ)
        got_t = ctx.project(funcs, lib.OPSEC_PREC_BF16X3)
    finally:
        ctx.set_projection(lib.OPSEC_PROJ_DENSE)
        ctx.lib.opsec_diag_set_nt_tile(2)
        ctx.lib.opsec_diag_set_nt_gen(0)
        ctx.lib.opsec_diag_set_stage2_split(8)
        ctx.lib.opsec_diag_set_toeplitz_stage2_split(4)
        ctx.lib.opsec_diag_set_g3_dl(2)
    for name, g in (("dense", got), ("toeplitz", got_t)):
        err = np.abs(g - ref).max(axis=1) / np.abs(ref).max(axis=1)
        print(name, "tile", tile, "gen", gen, "dl", dl, "max rel err per function", err)
        assert err.max() < 5e-5


@pytest.mark.parametrize("dims,drop,shuffle", [((8, 8, 8), 0.0, False), ((6, 5, 4), 0.2, True), ((3, 7, 9), 0.0, True),
                                               ((1, 8, 6), 0.0, False), ((2, 6, 6), 0.1, True)])
def test_project_toeplitz_vs_fp64(ctx, dims, drop, shuffle):
    """Toeplitz projection (DFT along x, 2 (Nx + 1) slab GEMMs in split-BF16x3)
    against the dense FP64 projection of the same pair tables: a full cubic
    grid, an irregular subset of a non-cubic grid in shuffled cell order, and a
    grid whose slab (63 positions) is not a multiple of the 256-row tile, and
    the shortest x extents (1 and 2 planes)."""
    n1, n2, n3 = dims
    h = 25.0
    cells = synth.box_cells(max(dims), 0.0, h * max(dims))
    keep = (cells["d1"] < n1) & (cells["d2"] < n2) & (cells["d3"] < n3)
    cells = cells[keep].copy()
    rng = np.random.default_rng(sum(dims))
    if drop:
        cells = cells[rng.random(len(cells)) >= drop].copy()
    if shuffle:
        cells = cells[rng.permutation(len(cells))].copy()
    cells["a"] = np.arange(len(cells))
    sv = lib.survey_box(n1, 0.0, h * n1, ny=n2, nz=n3)
    sv.max2, sv.max3 = h * n2, h * n3
    blo, bhi = synth.linear_bands(3, 0.02, 0.2)
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], blo, bhi, lib.OPSEC_XI_CLOSED)
    ctx.cells_upload(cells, sv)
    funcs = [0, 1, 2, 3]
    ctx.pair_tables(funcs, ctx.quad(lib.OPSEC_QUAD_GAUSS, nodes=4, near=1))
    N = len(cells)
    m = min(N - 1, 90)
    B = np.linalg.qr(rng.normal(size=(N, m)))[0].T.copy()
    ctx.modes_upload(B)
    ref = ctx.project(funcs, lib.OPSEC_PREC_FP64)
    try:
        ctx.set_projection(lib.OPSEC_PROJ_TOEPLITZ)
        got = ctx.project(funcs, lib.OPSEC_PREC_BF16X3)
        again = ctx.project(funcs, lib.OPSEC_PREC_BF16X3)
        got64 = ctx.project(funcs, lib.OPSEC_PREC_FP64)
        got32 = ctx.project(funcs, lib.OPSEC_PREC_FP32)
    finally:
        ctx.set_projection(lib.OPSEC_PROJ_DENSE)
    assert np.array_equal(got, again)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("Toeplitz BF16x3 projection", dims, "max rel err per function", err)
    assert err.max() < 5e-5
    # the same identity in FP64 (FP64 DFTs and MFMA) and in FP32 MFMA
    err64 = np.abs(got64 - ref).max(axis=1) / np.abs(ref).max(axis=1)
    err32 = np.abs(got32 - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("Toeplitz FP64 / FP32 projection", dims, "max rel err", err64.max(), err32.max())
    assert err64.max() < 1e-12
    assert err32.max() < 2e-5


def test_project_toeplitz_rejects_shared_grid_position(ctx):
    cells = synth.box_cells(4, 0.0, 100.0)
    dup = np.concatenate([cells, cells[:1]])
    dup["a"] = np.arange(len(dup))
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], [0.02], [0.1], lib.OPSEC_XI_CLOSED)
    ctx.cells_upload(dup, lib.survey_box(4, 0.0, 100.0))
    ctx.pair_tables([1], ctx.quad(lib.OPSEC_QUAD_GAUSS, nodes=4, near=1))
    ctx.modes_upload(np.linalg.qr(np.random.default_rng(0).normal(size=(len(dup), 8)))[0].T.copy())
    try:
        ctx.set_projection(lib.OPSEC_PROJ_TOEPLITZ)
        for prec in (lib.OPSEC_PREC_BF16X3, lib.OPSEC_PREC_FP64):
            with pytest.raises(lib.OpsecError, match="share a grid position"):
                ctx.project([1], prec)
    finally:
        ctx.set_projection(lib.OPSEC_PROJ_DENSE)


def test_new_cells_invalidate_modes(ctx):
    """Modes are weighted by the cells' Nbar: a new cell upload drops them until the next
    mode upload, so a projection cannot run on modes of another cell set."""
    cells = synth.box_cells(4, 0.0, 100.0)
    sv = lib.survey_box(4, 0.0, 100.0)
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], [0.02], [0.1], lib.OPSEC_XI_CLOSED)
    ctx.cells_upload(cells, sv)
    ctx.pair_tables([1], ctx.quad(lib.OPSEC_QUAD_GAUSS, nodes=4, near=1))
    B = np.linalg.qr(np.random.default_rng(1).normal(size=(len(cells), 8)))[0].T.copy()
    ctx.modes_upload(B)
    first = ctx.project([1], lib.OPSEC_PREC_FP64)
    ctx.cells_upload(cells, sv)
    ctx.pair_tables([1], ctx.quad(lib.OPSEC_QUAD_GAUSS, nodes=4, near=1))
    with pytest.raises(lib.OpsecError, match="no modes uploaded"):
        ctx.project([1], lib.OPSEC_PREC_FP64)
    ctx.modes_upload(B)
    assert np.array_equal(ctx.project([1], lib.OPSEC_PREC_FP64), first)


@pytest.mark.parametrize("prec", [lib.OPSEC_PREC_FP64, lib.OPSEC_PREC_FP32])
def test_estimate_bit_reproducible(ctx, prec):
    """The Gram product's workgroup partials are added in a fixed order: the
    same inputs give the same F, q and f bits on every call (M = 3000: 4.5e6
    packed entries over hundreds of workgroups)."""
    rng = np.random.default_rng(21)
    M, npar = 3000, 12
    lam = np.sort(rng.uniform(0.05, 5, M))[::-1].copy()
    D = rng.normal(size=(npar, M * (M + 1) // 2))
    y = np.sqrt(1 + lam) * rng.normal(size=M)
    prior = rng.uniform(0.5, 2.0, npar)
    a = ctx.estimate(lam, D, y, prior, lib.OPSEC_MIX_IDENTITY, prec)
    for _ in range(3):
        b = ctx.estimate(lam, D, y, prior, lib.OPSEC_MIX_IDENTITY, prec)
        for k in ("F", "q", "f"):
            assert np.array_equal(a[k], b[k]), k
