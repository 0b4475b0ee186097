"""The klt / comma / estimate drop-ins: command line, config and file-format
behaviour (CPU), and the end-to-end pipeline against the oracle (GPU)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from opsec_amd import formats, lib, synth
from tests import oracles as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "opsec_amd", "bin")


def run(prog, *args, cwd=None):
    return subprocess.run([os.path.join(BIN, prog), *args], capture_output=True, text=True, cwd=cwd, timeout=600)


@pytest.mark.parametrize("prog", ["klt", "comma", "dot", "estimate"])
def test_help_exits_zero(prog):
    r = run(prog, "-h")
    assert r.returncode == 0 and "Usage:" in r.stdout


@pytest.mark.parametrize("prog,keys", [("klt", "cellfile"), ("comma", "covfile"), ("dot", "pixfile"),
                                       ("estimate", "pixfile")])
def test_missing_keys_exit_one(prog, keys):
    r = run(prog, "Nmodes=3")
    assert r.returncode == 1
    assert "missing configuration options" in r.stderr


def test_dot_does_not_require_countfile(tmp_path):
    """The reference requires countfile and never uses it (src/dot.cpp:54);
    here it is optional: without it dot gets past the key check."""
    r = run("dot", f"cellfile={tmp_path}/nope.abn", "Nmodes=1", f"modefile={tmp_path}/m", f"pixfile={tmp_path}/p")
    assert "missing configuration options" not in r.stderr
    assert "ReadCells: could not open" in r.stderr


def test_dot_rejects_inconsistent_modefile(tmp_path):
    cells = synth.box_cells(3)
    formats.write_cells(tmp_path / "cells.abn", cells, {"coordsys": "cartesian"})
    formats.write(tmp_path / "modes.abn", np.zeros(2 * 27), "d", {"Nmodes": 2, "Ncells": 27})
    r = run("dot", f"cellfile={tmp_path}/cells.abn", "Nmodes=3", f"modefile={tmp_path}/modes.abn",
            f"pixfile={tmp_path}/p")
    assert r.returncode != 0
    assert "modefile has Nmodes = 2, expected 3" in r.stderr


def test_unreadable_cellfile_aborts(tmp_path):
    r = run("klt", f"cellfile={tmp_path}/nope.abn", "Nmodes=1", f"modefile={tmp_path}/m", f"evalfile={tmp_path}/e")
    assert r.returncode < 0 or r.returncode == 134
    assert "ReadCells: could not open" in r.stderr


def test_klt_rejects_too_many_modes(tmp_path):
    cells = synth.box_cells(3)
    cf = tmp_path / "cells.abn"
    formats.write_cells(cf, cells, {"coordsys": "cartesian"})
    r = run("klt", f"cellfile={cf}", "Nmodes=30", f"modefile={tmp_path}/m", f"evalfile={tmp_path}/e")
    assert r.returncode == 1
    assert "requesting 30 modes from a 27-dimensional space" in r.stderr


def test_cellfile_roundtrip_with_reference_reader(tmp_path):
    if not O.ref_available():
        pytest.skip("reference build not present")
    cells = synth.box_cells(4, 0.0, 100.0)
    cf = str(tmp_path / "cells.abn")
    formats.write_cells(cf, cells, {"coordsys": "cartesian", "Nx": 4})
    out = np.zeros(len(cells), dtype=lib.CELL_DTYPE)
    n = O.ref().ref_read_cells(cf.encode(), out.ctypes.data, len(out))
    assert n == 64 and out.tobytes() == cells.tobytes()


def _klt_headers(path_modes, path_evals, cellfile, Ncells, Nmodes, nconv, evals, modes):
    """The modefile and evalfile exactly as the klt program writes them
    (opsec_amd/host/klt.cpp): comment lines, opts Ncells, Nmodes, cellfile,
    nconv, all Nmodes vectors."""
    opts = {"Ncells": Ncells, "Nmodes": Nmodes, "cellfile": cellfile, "nconv": nconv}
    formats.write(path_evals, evals, "d", opts, comments=("Signal matrix eigenvalues",))
    formats.write(path_modes, modes.ravel(), "d", opts,
                  comments=("Signal matrix eigenmodes", "Each mode is stored consecutively, as below:",
                            "  v_1,1   v_1,2   ... v_1,n", "  ...", "  v_m,1   v_m,2   ... v_m,n"))


def test_modefile_read_by_reference_reader(tmp_path):
    """The reference's ReadModes (src/sig.cpp:621-697) and abn_read read the
    klt output layout; with fewer converged modes the reference's own header
    (n = nconv Ncells, Nmodes from the config, src/klt.cpp:210-225) is
    rejected by its own reader, which is why klt writes all Nmodes vectors."""
    R = O.ref()
    N, M = 50, 7
    rng = np.random.default_rng(3)
    modes, evals = rng.normal(size=(M, N)), np.sort(rng.uniform(0, 3, M))[::-1].copy()
    mf, ef = tmp_path / "modes.abn", tmp_path / "evals.abn"
    _klt_headers(mf, ef, "cells.abn", N, M, M - 2, evals, modes)
    out = np.zeros(M * N)
    nc = C.c_int()
    assert R.ref_read_modes(str(mf).encode(), C.byref(nc), O.ptr(out), out.size) == M
    assert nc.value == N and np.array_equal(out.reshape(M, N), modes)
    ev = np.zeros(M)
    assert R.ref_read_abn_doubles(str(ef).encode(), O.ptr(ev), M) == M and np.array_equal(ev, evals)
    # the reference's inconsistent header when nconv < Nmodes
    bad = tmp_path / "bad.abn"
    formats.write(bad, modes[:M - 2].ravel(), "d", {"Ncells": N, "Nmodes": M, "nconv": M - 2})
    assert R.ref_read_modes(str(bad).encode(), C.byref(nc), None, 0) == -1


def _write_inputs(tmp_path, n=6, nb=3, lo=0.0, hi=150.0):
    cells = synth.box_cells(n, lo, hi)
    cf = tmp_path / "cells.abn"
    formats.write_cells(cf, cells, {"coordsys": "cartesian"})
    k, P = synth.pk_table()
    pk = tmp_path / "pk.dat"
    formats.write_pkfile(pk, k, P)
    cfg = {
        "cellfile": cf, "model": "RealModel", "RealModel.pkfile": pk, "RealModel.Nbands": nb,
        "RealModel.kmin": 0.02, "RealModel.kmax": 0.14, "survey.type": "BoxSurvey",
        "survey.XMin": lo, "survey.XMax": hi, "survey.YMin": lo, "survey.YMax": hi, "survey.ZMin": lo,
        "survey.ZMax": hi, "survey.Nx": n, "survey.Ny": n, "survey.Nz": n, "survey.nbar": 3e-4,
        "survey.L": hi - lo, "survey.galfile": "none", "solver": "parpack",
    }
    cfile = tmp_path / "opsec.cfg"
    formats.write_config(cfile, cfg)
    return cells, cfile, np.loadtxt(pk)


def _split(path, piece):
    """Cut a file into pieces as `split -d -a 2 -b piece` does; returns the
    first piece's name (path + ".00")."""
    data = open(path, "rb").read()
    for i in range(0, len(data), piece):
        with open(f"{path}.{i // piece:02d}", "wb") as f:
            f.write(data[i:i + piece])
    return f"{path}.00"


def test_split_pieces_named_like_split(tmp_path):
    p = tmp_path / "x.abn"
    p.write_bytes(bytes(range(250)))
    first = _split(p, 100)
    assert first.endswith("x.abn.00")
    assert [len(open(f"{p}.{i:02d}", "rb").read()) for i in range(3)] == [100, 100, 50]


def _estimate_inputs(tmp_path, nb=3, M=10):
    """Model config, evalfile and pixfile for estimate, and a packed covfile
    of nb bands (estimate reads all three before it opens the device)."""
    _, cfile, _ = _write_inputs(tmp_path, 3, nb)
    formats.write(tmp_path / "evals.abn", np.linspace(2.0, 1.0, M), "d", {"Nmodes": M})
    formats.write(tmp_path / "pix.abn", np.ones(M), "d", comments=("Pixel values",))
    K = M * (M + 1) // 2
    formats.write(tmp_path / "cov.abn", np.arange(nb * K, dtype=np.float64), "d",
                  {"Nmodes": M, "Nparams": nb, "packed": 1})
    return cfile


def test_estimate_reads_covfile_pieces(tmp_path):
    """A covfile cut into pieces reads through (SplitFile); with a piece
    missing the read fails with the reference's message."""
    cfile = _estimate_inputs(tmp_path)
    first = _split(tmp_path / "cov.abn", 500)
    args = ["-c", str(cfile), "Nmodes=10", f"evalfile={tmp_path}/evals.abn", f"pixfile={tmp_path}/pix.abn",
            f"estfile={tmp_path}/est.abn", f"covfile={first}"]
    r = run("estimate", *args)
    assert "ReadCovMatrix" not in r.stderr and "Reading pixel values" in r.stdout + r.stderr
    os.remove(f"{tmp_path}/cov.abn.02")
    r = run("estimate", *args)
    assert r.returncode != 0 and "ReadCovMatrix: read error" in r.stderr


@pytest.mark.gpu
def test_pipeline_end_to_end_vs_oracle(tmp_path):
    """klt -> comma -> estimate.  comma reads the modefile and estimate the
    covfile cut into pieces (SplitFile); with 6 bands comma projects and
    writes in two groups."""
    n, nb, M = 6, 6, 40
    cells, cfile, kP = _write_inputs(tmp_path, n, nb)
    N = len(cells)
    r = run("klt", "-c", str(cfile), f"Nmodes={M}", f"modefile={tmp_path}/modes.abn", f"evalfile={tmp_path}/evals.abn",
            "hip.tol=1e-10")
    assert r.returncode == 0, r.stderr
    eopts, (lam,) = formats.read_doubles(tmp_path / "evals.abn")
    assert eopts["Nmodes"] == str(M) and eopts["Ncells"] == str(N) and int(eopts["nconv"]) == M
    _, (modes,) = formats.read_doubles(tmp_path / "modes.abn")
    B = modes.reshape(M, N)
    # the reference's own readers accept klt's files (src/sig.cpp:621-697, src/abn.c)
    got = np.zeros(M * N)
    nc = C.c_int()
    assert O.ref().ref_read_modes(str(tmp_path / "modes.abn").encode(), C.byref(nc), O.ptr(got), got.size) == M
    assert nc.value == N and np.array_equal(got, modes)
    gl = np.zeros(M)
    assert O.ref().ref_read_abn_doubles(str(tmp_path / "evals.abn").encode(), O.ptr(gl), M) == M
    assert np.array_equal(gl, lam)
    # oracle S and its spectrum
    lo, hi = synth.linear_bands(nb, 0.02, 0.14)
    r768, tabs = O.oracle_tables(kP[:, 0], kP[:, 1], lo, hi)
    rows = np.arange(N, dtype=np.int32)
    mats = []
    for f in range(nb + 1):
        xi = O.XiHandle(0, r768, tabs[f])
        S = np.zeros((N, N))
        O.oracle().orc_sig_block(C.byref(xi.s), cells.ctypes.data, N, n, n, n, 150.0, O.ptr(rows), N, O.ptr(rows), N,
                                 O.ptr(S), N, 0, 1e-5, 1e-10, 50000000, 0, 0)
        mats.append(S)
    w = np.linalg.eigvalsh(mats[0])[::-1][:M]
    np.testing.assert_allclose(lam, w, atol=1e-9 * w[0])
    np.testing.assert_allclose(B @ mats[0] @ B.T, np.diag(lam), atol=1e-9 * w[0])

    modes_split = _split(tmp_path / "modes.abn", 20000)
    r = run("comma", "-c", str(cfile), f"Nmodes={M}", f"modefile={modes_split}", f"covfile={tmp_path}/cov.abn")
    assert r.returncode == 0, r.stderr
    copts, (D,) = formats.read_doubles(tmp_path / "cov.abn")
    assert copts == {"Nmodes": str(M), "Nparams": str(nb), "packed": "1"}
    K = M * (M + 1) // 2
    D = D.reshape(nb, K)
    for p in range(nb):
        ref = np.zeros(K)
        O.oracle().orc_project(N, O.ptr(mats[p + 1]), M, O.ptr(np.ascontiguousarray(B)), O.ptr(ref))
        np.testing.assert_allclose(D[p], ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max())

    # the FP64 Toeplitz projection (DFT along x, FP64 MFMA) against the oracle
    r = run("comma", "-c", str(cfile), f"Nmodes={M}", f"modefile={modes_split}", f"covfile={tmp_path}/cov_t64.abn",
            "hip.precision=fp64", "hip.projection=toeplitz")
    assert r.returncode == 0, r.stderr
    _, (Dt64,) = formats.read_doubles(tmp_path / "cov_t64.abn")
    Dt64 = Dt64.reshape(nb, K)
    for p in range(nb):
        ref = np.zeros(K)
        O.oracle().orc_project(N, O.ptr(mats[p + 1]), M, O.ptr(np.ascontiguousarray(B)), O.ptr(ref))
        np.testing.assert_allclose(Dt64[p], ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max())

    # the mixed-precision Toeplitz projection of the same bands
    r = run("comma", "-c", str(cfile), f"Nmodes={M}", f"modefile={modes_split}", f"covfile={tmp_path}/cov_t.abn",
            "hip.precision=bf16x3", "hip.projection=toeplitz")
    assert r.returncode == 0, r.stderr
    _, (Dt,) = formats.read_doubles(tmp_path / "cov_t.abn")
    Dt = Dt.reshape(nb, K)
    assert (np.abs(Dt - D).max(axis=1) / np.abs(D).max(axis=1)).max() < 5e-5

    y = np.sqrt(1 + lam) * np.random.default_rng(42).normal(size=M)
    formats.write(tmp_path / "pix.abn", y, "d", comments=("Pixel values",))
    r = run("estimate", "-c", str(cfile), f"Nmodes={M}", f"evalfile={tmp_path}/evals.abn",
            f"covfile={_split(tmp_path / 'cov.abn', 30000)}", f"pixfile={tmp_path}/pix.abn",
            f"estfile={tmp_path}/est.abn")
    assert r.returncode == 0, r.stderr
    assert "p_n          error" in r.stdout
    eo, blocks = formats.read_doubles(tmp_path / "est.abn")
    assert eo == {"Nparams": str(nb)} and len(blocks) == 7
    phat, cov, F, W, Mm, q, f = blocks
    prior = np.array([O.oracle().orc_band_param(len(kP), O.ptr(np.ascontiguousarray(kP[:, 0])),
                                                O.ptr(np.ascontiguousarray(kP[:, 1])), a, b) for a, b in zip(lo, hi)])
    out = {k2: np.zeros(nb) for k2 in ("phat", "q", "f")}
    out.update({k2: np.zeros((nb, nb)) for k2 in ("cov", "F", "W", "M")})
    lam_c, D_c, y_c = np.ascontiguousarray(lam), np.ascontiguousarray(D), np.ascontiguousarray(y)
    O.oracle().orc_estimate(M, nb, O.ptr(lam_c), O.ptr(D_c), O.ptr(y_c), O.ptr(prior), 0,
                            *[O.ptr(out[k2]) for k2 in ("phat", "cov", "F", "W", "M", "q", "f")])
    # phat = F^-1 (...) inherits the condition number of F (6 bands on 40 modes)
    cond = np.linalg.cond(out["F"])
    for name, got in (("F", F), ("q", q), ("f", f), ("phat", phat)):
        ref = out[name].ravel()
        rtol = max(1e-9, 1e-14 * cond) if name == "phat" else 1e-9
        np.testing.assert_allclose(got, ref, rtol=rtol, atol=1e-10 * np.abs(ref).max(), err_msg=name)
