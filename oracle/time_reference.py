"""Times the reference's own serial pair-quadrature slice against the oracle's
restatement of it, on the CPU.  TEST INFRASTRUCTURE ONLY (a measurement).

Workload (C1): the 12^3 periodic box, 400 Mpc/h, nbar 3e-4, RealModel with 20
linear bands over [0.01, 0.20] h/Mpc; the full 1728 x 1728 block of the prior
S and of every C_{,p} (21 matrices) by ComputeSignalMatrixBlock
(src/sig.cpp:50-123): adaptive Genz-Malik pairs with the reference's
tolerances (sig.epsrel 1e-5, sig.epsabs 1e-10) and its translation cache.

  reference    oracle/_ref/libopsec_ref.so: the reference's sig.cpp, cubature.c,
               RealModel.cpp, ... compiled from its sources (OpenMP inside the
               integrand, as the reference runs it)
  restatement  oracle/liboracle.so: orc_sig_block in Genz-Malik mode (OpenMP
               over the distinct keys)

Usage: python oracle/time_reference.py [--runs 5] [--threads 8]
Prints one JSON line: per-run seconds, medians, and the ratio.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    # OpenMP reads the thread count when the libraries load
    os.environ["OMP_NUM_THREADS"] = str(args.threads)
    sys.path.insert(0, ROOT)
    import ctypes as C

    import numpy as np

    from opsec_amd import synth
    from tests import oracles as O

    n, nb = 12, 20
    N = n ** 3
    cells = synth.box_cells(n)
    k, P = synth.pk_table()
    lo, hi = synth.linear_bands(nb, 0.01, 0.20)
    rows = np.arange(N, dtype=np.int32)
    S = np.zeros((N, N))

    R = O.ref()
    with tempfile.TemporaryDirectory() as d:
        pk = os.path.join(d, "pk.dat")
        O.write_pkfile(pk, k, P)
        m = O.ref_real_model(pk, lo, hi)
        sv = O.survey_text(n)

        def ref_pass():
            t0 = time.perf_counter()
            for f in range(nb + 1):
                R.ref_sig_block(m, f, sv, cells.ctypes.data, N, O.ptr(rows), N, O.ptr(rows), N, O.ptr(S), N, 1e-5,
                                1e-10)
            return time.perf_counter() - t0

        t_ref = [ref_pass() for _ in range(args.runs)]
        S_ref = S.copy()
        R.ref_model_free(m)

    orc = O.oracle()
    r768, tabs = O.oracle_tables(k, P, lo, hi)
    xis = [O.XiHandle(0, r768, t) for t in tabs]

    def orc_pass():
        t0 = time.perf_counter()
        for xi in xis:
            orc.orc_sig_block(C.byref(xi.s), cells.ctypes.data, N, n, n, n, 400.0, O.ptr(rows), N, O.ptr(rows), N,
                              O.ptr(S), N, 0, 1e-5, 1e-10, 50000000, 0, 0)
        return time.perf_counter() - t0

    t_orc = [orc_pass() for _ in range(args.runs)]
    # the last function of both passes: the same block
    agree = float(np.abs(S - S_ref).max() / np.abs(S_ref).max())
    out = {
        "workload": "C1: 12^3 box, prior + 20 bands, full 1728 x 1728 blocks, Genz-Malik (reference tolerances)",
        "threads": args.threads, "runs": args.runs,
        "reference_s": t_ref, "restatement_s": t_orc,
        "reference_median_s": float(np.median(t_ref)), "restatement_median_s": float(np.median(t_orc)),
        "ratio_reference_over_restatement": float(np.median(t_ref) / np.median(t_orc)),
        "cell_pairs_per_s_reference": (nb + 1) * N * N / float(np.median(t_ref)),
        "last_block_max_rel_diff": agree,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
