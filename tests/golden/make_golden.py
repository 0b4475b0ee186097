"""Generate the golden fixtures from the reference's own compiled serial slice.

Run in a container where /root/reference exists and oracle/_ref has been
built (`make -C oracle ref`):

    python tests/golden/make_golden.py

Writes tests/golden/*.npz -- input/output vectors only.  Every number comes
from calling the reference's functions (RealModel::GetXi/GetXiDeriv/GetParam,
CubicSpline, BoxSeparationFunc, ComputeSignalC, ComputeSignalMatrixBlock)
through oracle/ref_shim.cpp.
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from opsec_amd import synth  # noqa: E402
from tests import oracles as O  # noqa: E402


def main():
    R = O.ref()
    out = {}
    tmp = tempfile.mkdtemp()
    pkfile = os.path.join(tmp, "pk.dat")
    k, P = synth.pk_table()
    O.write_pkfile(pkfile, k, P)
    # the pkfile is read back from its text form by the reference; keep what it read
    kP = np.loadtxt(pkfile)
    out["pk_k"], out["pk_P"] = kP[:, 0], kP[:, 1]

    # cubic spline of the pkfile at assorted k
    ke = np.concatenate([np.linspace(0, 10.4, 97), [1e-5, 0.0141, 0.02, 0.25, 9.999]])
    ye = np.zeros_like(ke)
    kk, PP = np.ascontiguousarray(kP[:, 0]), np.ascontiguousarray(kP[:, 1])
    R.ref_cubic_spline(len(kP), O.ptr(kk), O.ptr(PP), len(ke), O.ptr(ke), O.ptr(ye))
    out["spline_k"], out["spline_P"] = ke, ye

    # C1 model: 20 bands over [0.01, 0.20]
    lo, hi = synth.linear_bands(20, 0.01, 0.20)
    m = O.ref_real_model(pkfile, lo, hi)
    out["c1_band_lo"], out["c1_band_hi"] = lo, hi
    out["c1_params"] = np.array([R.ref_model_param(m, i) for i in range(20)])
    r = np.zeros(768)
    O.oracle().orc_qspline_grid(O.ptr(r))
    # tables are read back at the nodes themselves (QSpline reproduces node values)
    rq = np.concatenate([r[:-1], [0.0, 1e-3, 13.7, 199.99, 200.0, 333.3, 19999.0, 25000.0]])
    xis = np.zeros((21, len(rq)))
    for f in range(21):
        R.ref_model_xi(m, f, len(rq), O.ptr(rq), O.ptr(xis[f]))
    out["xi_r"], out["c1_xi"] = rq, xis

    # separations: box known answers plus random pairs (L = 100)
    rng = np.random.default_rng(5)
    p1 = rng.uniform(0, 100, (64, 3))
    p2 = rng.uniform(0, 100, (64, 3))
    rr = np.zeros(64)
    R.ref_box_sep(O.survey_text(1, 0.0, 100.0, 100.0), 64, O.ptr(p1), O.ptr(p2), O.ptr(rr))
    out["sep_p1"], out["sep_p2"], out["sep_r"] = p1, p2, rr

    # single pair integrals (GM, reference tolerances) on the C1 grid
    cells = synth.box_cells(12)
    sv = O.survey_text(12)
    pairs = [(0, 0), (0, 1), (0, 13), (0, 157), (0, 1727), (5, 900), (100, 101), (0, 6 * 144 + 6 * 12 + 6)]
    Q = np.zeros((3, len(pairs)))
    E = np.zeros((3, len(pairs)))
    for t, f in enumerate((0, 1, 20)):
        for i, (a, b) in enumerate(pairs):
            e = C.c_double()
            Q[t, i] = R.ref_signal_c(m, f, sv, cells[a:a + 1].ctypes.data, cells[b:b + 1].ctypes.data, 1e-5, 1e-10,
                                     C.byref(e))
            E[t, i] = e.value
    out["pairs"], out["pair_funcs"], out["pair_Q"], out["pair_err"] = np.array(pairs), np.array([0, 1, 20]), Q, E

    # signal-matrix blocks: 12^3 box, rows 0..7 and a scattered set, all columns
    rows = np.array(list(range(8)) + [100, 863, 1727], dtype=np.int32)
    cols = np.arange(1728, dtype=np.int32)
    blocks = np.zeros((3, len(rows), len(cols)))
    for t, f in enumerate((0, 1, 20)):
        R.ref_sig_block(m, f, sv, cells.ctypes.data, 1728, O.ptr(rows), len(rows), O.ptr(cols), 1728,
                        O.ptr(blocks[t]), 1728, 1e-5, 1e-10)
    out["blk_rows"], out["blk_funcs"], out["blk_S"] = rows, np.array([0, 1, 20]), blocks

    # a non-trivial sub-block with ld > ncols, aperiodic 6^3 box (L = 1e9)
    cells6 = synth.box_cells(6, 0.0, 120.0)
    sv6 = O.survey_text(6, 0.0, 120.0, 1e9)
    r6 = np.array([3, 7, 50, 215], dtype=np.int32)
    c6 = np.array([0, 1, 2, 36, 100, 150, 215], dtype=np.int32)
    b6 = np.full((2, len(r6), 9), -7.0)
    for t, f in enumerate((0, 5)):
        R.ref_sig_block(m, f, sv6, cells6.ctypes.data, 216, O.ptr(r6), len(r6), O.ptr(c6), len(c6),
                        O.ptr(b6[t]), 9, 1e-5, 1e-10)
    out["b6_rows"], out["b6_cols"], out["b6_S"] = r6, c6, b6
    R.ref_model_free(m)

    # TestModel (exponential and gaussian) pair integrals
    tm = R.ref_model_new(b"model = TestModel\nTestModel.form = gaussian\nTestModel.A = 2.5\nTestModel.b = 0.01\n")
    tq = np.zeros((3, len(pairs)))
    for f in range(3):
        for i, (a, b) in enumerate(pairs):
            tq[f, i] = R.ref_signal_c(tm, f, sv, cells[a:a + 1].ctypes.data, cells[b:b + 1].ctypes.data, 1e-5,
                                      1e-10, None)
    out["test_gauss_Q"] = tq
    R.ref_model_free(tm)

    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "reference_vectors.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
