"""Spherical-shell surveys on the device: the arc-pair tables of every ordered
arc pair, the assembled S, the KL modes, the projections and the programs end
to end, against the oracle's restatement of ComputeSignalMatrixBlockS
(src/sig.cpp:125-213) and the brute-force projection."""
import ctypes as C
import os

import numpy as np
import pytest

from opsec_amd import formats, lib, synth
from tests import oracles as O

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
RMIN, RMAX = 250.0, 650.0


def _shell(ctx, dims=(3, 3, 8), drop=0.0, nb=3, shuffle=False):
    nr, nmu, nphi = dims
    cells = synth.shell_cells(nr, nmu, nphi, rmin=RMIN, rmax=RMAX)
    if drop:
        cells = synth.drop_cells(cells, drop, seed=3)
    if shuffle:
        cells = cells[np.random.default_rng(5).permutation(len(cells))].copy()
        cells["a"] = np.arange(len(cells))
    sv = lib.survey_sph(nr, nmu, nphi, RMIN, RMAX)
    k, P = GOLD["pk_k"], GOLD["pk_P"]
    blo, bhi = synth.linear_bands(nb, 0.01, 0.1)
    ctx.model_real(k, P, blo, bhi, lib.OPSEC_XI_TABLE)
    ctx.cells_upload(cells, sv)
    return cells, (k, P, blo, bhi)


def _oracle_mats(cells, dims, model, q):
    """S_f for the prior and every band from the oracle's arc-cache block (all
    rows and columns, so the cache holds the same first pairs)."""
    k, P, blo, bhi = model
    r768, tabs = O.oracle_tables(k, P, blo, bhi)
    N = len(cells)
    rows = np.arange(N, dtype=np.int32)
    mats = []
    for t in tabs:
        xi = O.XiHandle(0, r768, t)
        S = np.zeros((N, N))
        assert O.oracle().orc_sig_block_sph(C.byref(xi.s), cells.ctypes.data, N, *dims, O.ptr(rows), N, O.ptr(rows),
                                            N, O.ptr(S), N, q.epsrel, q.epsabs, q.maxeval) == 0
        mats.append(S)
    return mats


def _orc_project(S, B):
    m = B.shape[0]
    out = np.zeros(m * (m + 1) // 2)
    O.oracle().orc_project(S.shape[0], O.ptr(np.ascontiguousarray(S)), m, O.ptr(np.ascontiguousarray(B)), O.ptr(out))
    return out


@pytest.mark.parametrize("dims,drop,shuffle", [((3, 3, 8), 0.0, False), ((3, 2, 7), 0.25, True)])
def test_arc_tables_and_assembly_vs_oracle(ctx, dims, drop, shuffle):
    """S_f assembled on the device from the arc-pair tables equals the oracle's
    block of all rows and columns, on a full shell and on an irregular,
    shuffled subset (arcs with missing cells)."""
    cells, model = _shell(ctx, dims, drop, nb=3, shuffle=shuffle)
    q = ctx.quad(lib.OPSEC_QUAD_GM)
    ctx.pair_tables(range(4), q)
    nr, nmu, nphi = dims
    assert ctx.nkeys() == (nr * nmu) ** 2 * nphi
    mats = _oracle_mats(cells, dims, model, q)
    for f, ref in enumerate(mats):
        got = ctx.assemble(f, len(cells))
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max(), err_msg=f"f={f}")
    # xi = 1 known answer: S = sqrt(Nbar_a Nbar_b) exactly up to the rule's rounding
    ctx.model_test(0, 1.0, 0.0)
    ctx.pair_tables([0], q)
    got = ctx.assemble(0, len(cells))
    nb_ = np.sqrt(np.outer(cells["Nbar"], cells["Nbar"]))
    np.testing.assert_allclose(got, nb_, rtol=1e-13)


def test_spherical_projection_all_precisions(ctx):
    dims = (3, 3, 8)
    cells, model = _shell(ctx, dims, 0.2, nb=3, shuffle=True)
    q = ctx.quad(lib.OPSEC_QUAD_GM)
    funcs = list(range(1, 4))
    ctx.pair_tables([0] + funcs, q)
    mats = _oracle_mats(cells, dims, model, q)
    N, m = len(cells), 24
    B = np.linalg.qr(np.random.default_rng(4).normal(size=(N, m)))[0].T.copy()
    ctx.modes_upload(B)
    ref = np.array([_orc_project(mats[f], B) for f in funcs])
    got = ctx.project(funcs, lib.OPSEC_PREC_FP64)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    for prec, tol in ((lib.OPSEC_PREC_FP32, 2e-5), (lib.OPSEC_PREC_BF16X3, 5e-5)):
        got = ctx.project(funcs, prec)
        err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
        print("spherical projection", prec, "max rel err per band", err)
        assert err.max() < tol
    # the Toeplitz projection (DFT along phi) against the same brute-force products
    ctx.set_projection(lib.OPSEC_PROJ_TOEPLITZ)
    try:
        got = ctx.project(funcs, lib.OPSEC_PREC_FP64)
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
        for prec, tol in ((lib.OPSEC_PREC_FP32, 2e-5), (lib.OPSEC_PREC_BF16X3, 5e-5)):
            got = ctx.project(funcs, prec)
            err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
            print("spherical Toeplitz projection", prec, "max rel err per band", err)
            assert err.max() < tol
    finally:
        ctx.set_projection(lib.OPSEC_PROJ_DENSE)


@pytest.mark.parametrize("dims,drop,shuffle", [((17, 16, 4), 0.0, False), ((3, 5, 7), 0.3, True), ((4, 4, 1), 0.0, True),
                                               ((2, 3, 2), 0.2, False), ((2, 3, 40), 0.1, True)])
def test_spherical_toeplitz_vs_dense(ctx, dims, drop, shuffle):
    """Toeplitz projection of a shell (length-2 Nphi DFT along phi, dense arc x arc
    That_k) against the dense FP64 projection of the same arc-pair tables: 272
    arcs (two column tiles of the That kernel, a padded 512-row GEMM tile), an
    odd Nphi with a third of the cells missing in shuffled order (table keys that
    no pair carries), a single phi plane, two planes, and 40 planes (the kernels
    for 32 < Nphi <= 64)."""
    cells, _ = _shell(ctx, dims, drop, nb=2, shuffle=shuffle)
    funcs = [0, 1, 2]
    ctx.pair_tables(funcs, ctx.quad(lib.OPSEC_QUAD_GM))
    N = len(cells)
    m = min(N - 1, 70)
    B = np.linalg.qr(np.random.default_rng(sum(dims)).normal(size=(N, m)))[0].T.copy()
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
    rel = lambda x: (np.abs(x - ref).max(axis=1) / np.abs(ref).max(axis=1)).max()
    print("spherical Toeplitz", dims, "BF16x3", rel(got), "FP64", rel(got64), "FP32", rel(got32))
    assert rel(got64) < 1e-12
    assert rel(got32) < 2e-5
    assert rel(got) < 5e-5


def test_overlapped_sweeps_give_the_same_tables(ctx):
    """The prior's sweep on a second stream beside the bands' sweep (the default)
    and after it give the same bits in every table."""
    cells, _ = _shell(ctx, (4, 3, 8), 0.15, nb=3, shuffle=True)
    q = ctx.quad(lib.OPSEC_QUAD_GM)
    funcs = [0, 1, 2, 3]
    ctx.pair_tables(funcs, q)
    overlapped = [ctx.pair_table(f) for f in funcs]
    try:
        assert ctx.lib.opsec_diag_set_gm_overlap(0) == 0
        ctx.pair_tables(funcs, q)
        serial = [ctx.pair_table(f) for f in funcs]
    finally:
        ctx.lib.opsec_diag_set_gm_overlap(1)
    for f, a, b in zip(funcs, overlapped, serial):
        assert np.array_equal(a, b), f
    assert ctx.lib.opsec_diag_set_gm_overlap(2) == -1   # OPSEC_EINVAL


def test_band_first_regions_vs_per_job_kernel_and_oracle(ctx):
    """Closed-form band tables of a shell: the first-region kernel that forms a
    pair's geometry once and shares the band-edge terms between adjacent bands
    gives the bits of the one-thread-per-job kernel, and agrees with the
    oracle's arc-cache block to 1e-12 (1e-11 of the largest entry for the
    entries that cancel; 12 log bands: two chunks of the kernel's band loop;
    near and self pairs take band_xi's series branch)."""
    nr, nmu, nphi = 3, 3, 6
    cells = synth.drop_cells(synth.shell_cells(nr, nmu, nphi, rmin=RMIN, rmax=RMAX), 0.1, seed=11)
    sv = lib.survey_sph(nr, nmu, nphi, RMIN, RMAX)
    edges = np.logspace(np.log10(0.003), np.log10(0.3), 13)
    blo, bhi = edges[:-1], edges[1:]
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], blo, bhi, lib.OPSEC_XI_CLOSED)
    ctx.cells_upload(cells, sv)
    q = ctx.quad(lib.OPSEC_QUAD_GM)
    funcs = list(range(13))
    ctx.pair_tables(funcs, q)
    shared = [ctx.pair_table(f) for f in funcs]
    N = len(cells)
    S_gpu = {p: ctx.assemble(1 + p, N) for p in (0, 6, 11)}
    try:
        assert ctx.lib.opsec_diag_set_gm_bands(0) == 0
        ctx.pair_tables(funcs, q)
        per_job = [ctx.pair_table(f) for f in funcs]
    finally:
        ctx.lib.opsec_diag_set_gm_bands(1)
    for f in funcs:
        assert np.array_equal(shared[f], per_job[f]), f
    assert ctx.lib.opsec_diag_set_gm_bands(2) == -1   # OPSEC_EINVAL
    rows = np.arange(N, dtype=np.int32)
    for p, got in S_gpu.items():
        xi = O.XiHandle(1, a=blo[p], b=bhi[p])
        S = np.zeros((N, N))
        assert O.oracle().orc_sig_block_sph(C.byref(xi.s), cells.ctypes.data, N, nr, nmu, nphi, O.ptr(rows), N,
                                            O.ptr(rows), N, O.ptr(S), N, q.epsrel, q.epsabs, q.maxeval) == 0
        # band_xi forms the closed form over one division by r^3, the oracle over two
        # (k^3 g(kr) per edge): near a node of an oscillating band the two differ by a
        # few ulps of the edge terms, 2e-12 of the largest entry at most
        np.testing.assert_allclose(got, S, rtol=1e-12, atol=1e-11 * np.abs(S).max(), err_msg=f"band {p}")


def test_spherical_rejects_tent_rule(ctx):
    _shell(ctx, (2, 2, 4))
    with pytest.raises(lib.OpsecError, match="adaptive"):
        ctx.pair_tables([0], ctx.quad(lib.OPSEC_QUAD_GAUSS))


def test_spherical_klt_vs_oracle(ctx):
    dims = (4, 3, 8)
    cells, model = _shell(ctx, dims, 0.1, nb=1)
    q = ctx.quad(lib.OPSEC_QUAD_GM)
    ctx.pair_tables([0], q)
    S = _oracle_mats(cells, dims, model, q)[0]
    nev = 30
    lam, V, nconv = ctx.klt(nev, tol=1e-10, ncells=len(cells))
    assert nconv == nev
    Ssym = 0.5 * (S + S.T)
    print("spherical S asymmetry", np.abs(S - S.T).max() / np.abs(S).max())
    w = np.linalg.eigvalsh(Ssym)[::-1][:nev]
    np.testing.assert_allclose(lam, w, atol=1e-9 * w[0])
    np.testing.assert_allclose(V @ Ssym @ V.T, np.diag(lam), atol=1e-9 * w[0])


def test_spherical_programs_end_to_end(tmp_path):
    """klt -> comma (FP64 and split-BF16x3, dense and Toeplitz projections) ->
    estimate on a spherical shell given by a SphericalSurvey configuration."""
    nr, nmu, nphi, nb, M = 3, 3, 8, 3, 30
    cells = synth.drop_cells(synth.shell_cells(nr, nmu, nphi, rmin=RMIN, rmax=RMAX), 0.1, seed=8)
    N = len(cells)
    cf = tmp_path / "cells.abn"
    formats.write_cells(cf, cells, {"coordsys": "spherical"})
    k, P = synth.pk_table()
    pk = tmp_path / "pk.dat"
    formats.write_pkfile(pk, k, P)
    cfg = {"cellfile": cf, "model": "RealModel", "RealModel.pkfile": pk, "RealModel.Nbands": nb,
           "RealModel.kmin": 0.01, "RealModel.kmax": 0.1, "survey.type": "SphericalSurvey",
           "survey.RMin": RMIN, "survey.RMax": RMAX, "survey.MuMin": 0.5, "survey.MuMax": 1.0,
           "survey.PhiMin": 0.0, "survey.PhiMax": 2 * np.pi, "survey.Nr": nr, "survey.Nmu": nmu,
           "survey.Nphi": nphi, "survey.galfile": "none", "survey.maskfile": "none", "survey.radial": "none",
           "solver": "parpack"}
    cfile = tmp_path / "opsec.cfg"
    formats.write_config(cfile, cfg)
    run = lambda prog, *a: __import__("subprocess").run(
        [os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opsec_amd", "bin", prog), *a],
        capture_output=True, text=True, timeout=600)
    r = run("klt", "-c", str(cfile), f"Nmodes={M}", f"modefile={tmp_path}/modes.abn",
            f"evalfile={tmp_path}/evals.abn", "hip.tol=1e-10")
    assert r.returncode == 0, r.stderr
    _, (lam,) = formats.read_doubles(tmp_path / "evals.abn")
    _, (modes,) = formats.read_doubles(tmp_path / "modes.abn")
    B = modes.reshape(M, N)
    kP = np.loadtxt(pk)
    lo, hi = synth.linear_bands(nb, 0.01, 0.1)
    q = lib.Context.quad(lib.OPSEC_QUAD_GM)
    mats = _oracle_mats(cells, (nr, nmu, nphi), (kP[:, 0], kP[:, 1], lo, hi), q)
    w = np.linalg.eigvalsh(0.5 * (mats[0] + mats[0].T))[::-1][:M]
    np.testing.assert_allclose(lam, w, atol=1e-9 * w[0])
    K = M * (M + 1) // 2
    for extra, tol in (((), 1e-10), (("hip.precision=bf16x3",), 5e-5), (("hip.projection=toeplitz",), 1e-10),
                       (("hip.precision=bf16x3", "hip.projection=toeplitz"), 5e-5)):
        r = run("comma", "-c", str(cfile), f"Nmodes={M}", f"modefile={tmp_path}/modes.abn",
                f"covfile={tmp_path}/cov.abn", *extra)
        assert r.returncode == 0, r.stderr
        _, (D,) = formats.read_doubles(tmp_path / "cov.abn")
        D = D.reshape(nb, K)
        for p in range(nb):
            ref = _orc_project(mats[p + 1], B)
            err = np.abs(D[p] - ref).max() / np.abs(ref).max()
            assert err < tol, (extra, p, err)
    # the FP64 covfile for the estimate
    r = run("comma", "-c", str(cfile), f"Nmodes={M}", f"modefile={tmp_path}/modes.abn", f"covfile={tmp_path}/cov.abn")
    assert r.returncode == 0, r.stderr
    _, (D,) = formats.read_doubles(tmp_path / "cov.abn")
    D = D.reshape(nb, K)
    y = np.sqrt(1 + lam) * np.random.default_rng(42).normal(size=M)
    formats.write(tmp_path / "pix.abn", y, "d", comments=("Pixel values",))
    r = run("estimate", "-c", str(cfile), f"Nmodes={M}", f"evalfile={tmp_path}/evals.abn",
            f"covfile={tmp_path}/cov.abn", f"pixfile={tmp_path}/pix.abn", f"estfile={tmp_path}/est.abn")
    assert r.returncode == 0, r.stderr
    _, blocks = formats.read_doubles(tmp_path / "est.abn")
    F = blocks[2]
    prior = np.array([O.oracle().orc_band_param(len(kP), O.ptr(np.ascontiguousarray(kP[:, 0])),
                                                O.ptr(np.ascontiguousarray(kP[:, 1])), a, b) for a, b in zip(lo, hi)])
    out = {k2: np.zeros(nb) for k2 in ("phat", "q", "f")}
    out.update({k2: np.zeros((nb, nb)) for k2 in ("cov", "F", "W", "M")})
    O.oracle().orc_estimate(M, nb, O.ptr(np.ascontiguousarray(lam)), O.ptr(np.ascontiguousarray(D)),
                            O.ptr(np.ascontiguousarray(y)), O.ptr(prior), 0,
                            *[O.ptr(out[k2]) for k2 in ("phat", "cov", "F", "W", "M", "q", "f")])
    np.testing.assert_allclose(F, out["F"].ravel(), rtol=1e-9, atol=1e-10 * np.abs(out["F"]).max())
