"""Benchmark of the hot path: pair integrals -> S and all C_{,p} -> projection
onto the KL modes -> Fisher matrix, on one MI355X.

Workload (default): the 32^3 periodic box (400 Mpc/h side, nbar = 3e-4),
40 linear bands over [0.005, 0.25] h/Mpc, 4000 KL modes, 6^3-node tent-kernel
Gauss pair quadrature with closed-form band correlation functions.

One step = pair tables for the prior and the 40 bands, the batched projection
P_p = B^T C_{,p} B of all bands (S_p generated on the fly inside the GEMM), and
the Fisher matrix / bias / quadratic estimates with the 40 x 40 post-processing.
Two paths are timed on the same inputs:
  fp64       plain method: Gauss quadrature, FP64 MFMA projection, FP64 traces
  surrogate  learned pair integrals for pairs >= 2 cells apart (quadrature
             below, 1% spot checks), split-BF16x3 MFMA projection, FP32 traces.
value = cell-pair integrals delivered per second by the surrogate path,
i.e. Ncells^2 x (1 + Nbands) / (S->Fisher wall time).

The KL modes are the eigenvectors of the prior S.  On the periodic box with a
constant selection S is a 3-level circulant, so its eigenvectors are real
Fourier modes and its eigenvalues the DFT of the first row; the bench builds
B that way from the GPU pair table (setup, untimed).  The klt program and
opsec_klt compute the same modes by block Lanczos.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-pair integrals/s + S→Fisher wall time, surrogate vs FP64 at ΔF/F≤1e-3"
PEAK_TF = {"fp64": 78.6, "fp32": 157.3, "bf16": 2516.6}
# HBM bytes per launch of the dominant projection kernel at 32^3 cells x 4000 modes:
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 KiB from two rocprofv3 --pmc passes, the factors being what
# the same counters read for copy16_kernel on a 2 GiB copy (profiles/projection_c2_hbm_counters.txt).
# Keys: (config, projection, kernel).
TRAFFIC_BYTES = {("C2", "dense", "stage1"): (18874466 * 2.0 + 883503 * 1.0) * 1024,
                 ("C2", "toeplitz", "stage1"): (1321387 * 2.0 + 1059234 * 1.0) * 1024,
                 ("C2", "toeplitz", "stage2"): (12218166 * 2.0 + 1150502 * 1.0) * 1024}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fourier_kl_modes(g, nmodes):
    """Top eigenpairs of the real symmetric 3-level circulant with generator g
    (n x n x n): eigenvalues = real DFT of g, eigenvectors = cos/sin modes."""
    n = g.shape[0]
    lam = np.fft.fftn(g).real
    idx = np.indices((n, n, n)).reshape(3, -1).T
    neg = (-idx) % n
    lin = idx[:, 0] * n * n + idx[:, 1] * n + idx[:, 2]
    nlin = neg[:, 0] * n * n + neg[:, 1] * n + neg[:, 2]
    rep = lin <= nlin                      # one representative per +-m pair
    order = np.argsort(-lam.ravel()[rep], kind="stable")
    reps = np.nonzero(rep)[0][order]
    x = np.indices((n, n, n)).reshape(3, -1).astype(np.float64)
    modes, evals = [], []
    for r in reps:
        m = idx[r]
        ph = 2 * np.pi * (m @ x) / n
        self_conj = lin[r] == nlin[r]
        c = np.cos(ph)
        modes.append(c / np.linalg.norm(c))
        evals.append(lam.ravel()[r])
        if not self_conj and len(modes) < nmodes:
            s = np.sin(ph)
            modes.append(s / np.linalg.norm(s))
            evals.append(lam.ravel()[r])
        if len(modes) >= nmodes:
            break
    return np.array(evals[:nmodes]), np.ascontiguousarray(np.array(modes[:nmodes]))


def run_path(ctx, L, funcs, quad, prec, trace_prec, lam, y, prior, steps, warmup, tag):
    """Time `steps` S->Fisher passes; returns (ms per step, est, kernel ms)."""
    import ctypes as C
    s1, s2 = C.c_float(), C.c_float()
    kg, k1, k2, kn = C.c_float(), C.c_float(), C.c_float(), C.c_int()
    l1, l2 = C.c_int(), C.c_int()
    times, est, kms, qms = [], None, [], []
    for it in range(warmup + steps):
        ctx.sync()
        t0 = time.perf_counter()
        ctx.pair_tables(funcs, quad)
        ctx.sync()
        tq = time.perf_counter() - t0
        ctx.project(funcs[1:], prec, keep_on_device=True)
        est = ctx.estimate_device(lam, y, prior, L.OPSEC_MIX_INVERSE, trace_prec)
        ctx.sync()
        dt = time.perf_counter() - t0
        ctx.lib.opsec_diag_project_times(ctx.h, C.byref(s1), C.byref(s2))
        ctx.lib.opsec_diag_project_kernel_ms(ctx.h, C.byref(kg), C.byref(k1), C.byref(k2), C.byref(kn))
        ctx.lib.opsec_diag_project_launches(ctx.h, C.byref(l1), C.byref(l2))
        if it >= warmup:
            times.append(dt)
            qms.append(tq)
            kms.append((s1.value, s2.value, kg.value, k1.value, k2.value, kn.value, 1e3 * tq, l1.value, l2.value))
        log(f"[{tag}] step {it - warmup if it >= warmup else 'warmup'}: {dt * 1e3:.1f} ms "
            f"(projection stage 1 {s1.value:.1f} ms, stage 2 {s2.value:.1f} ms)")
    return 1e3 * float(np.mean(times)), est, np.mean(np.array(kms), axis=0)


def stage1_kernel_name(tile, gen):
    if gen:
        return "gemm3_kernel<ASrcTable,EpiPlanes> (stage 1, S_p from the pair table, one launch per band)"
    if tile == 2:
        return "gemm3_kernel<ASrcPlanes,EpiPlanes> (stage 1, S_p planes, one launch per band)"
    return "nt_bf16x3_kernel<EpiPlanes> (stage 1, S_p planes, one launch per band)"


def compare_estimates(e64, e, info_floor=1e-8):
    """Error of an estimate against the FP64 one.

    F: entry-wise in correlation units |dF_pq| / sqrt(F_pp F_qq), and on its
    diagonal.  Band powers: with 4000 KL modes the highest bands are almost
    unconstrained (F is numerically singular), so p = F^-1 (q - f) is compared
    on the constrained subspace -- the eigen-directions of the FP64 F holding
    at least info_floor of the largest eigenvalue -- through the pseudo-inverse
    F+ there: dp = F+ d(q - f) in units of sigma_p = sqrt(F+_pp), and as the
    whitened shift w^-1/2 V^T d(q - f) (one unit = one sigma along an
    independent band combination)."""
    F64, F = e64["F"], e["F"]
    d = np.sqrt(np.abs(np.diag(F64)))
    w, V = np.linalg.eigh(F64)
    keep = w >= info_floor * w.max()
    Vc, wc = V[:, keep], w[keep]
    r = (e["q"] - e["f"]) - (e64["q"] - e64["f"])
    z = (Vc.T @ r) / np.sqrt(wc)
    Fp = (Vc / wc) @ Vc.T
    dp = Fp @ r
    sig = np.sqrt(np.diag(Fp))
    return {
        "fisher_max_rel_err_diag": float(np.max(np.abs(np.diag(F) - np.diag(F64)) / np.abs(np.diag(F64)))),
        "fisher_max_err_corr_units": float(np.max(np.abs(F - F64) / np.outer(d, d))),
        "fisher_max_abs_err_over_max": float(np.max(np.abs(F - F64)) / np.max(np.abs(F64))),
        "fisher_condition_number": float(w.max() / max(abs(w.min()), 1e-300)),
        "constrained_directions": int(keep.sum()),
        "band_power_max_shift_sigma": float(np.max(np.abs(dp) / sig)),
        "band_power_max_shift_sigma_whitened": float(np.max(np.abs(z))),
    }


def diagnose(ctx, L, funcs, q64, qs, lam, y, prior, e64):
    """Split the error: quadrature tables with the mixed-precision projection
    and traces, and surrogate tables with the FP64 projection and traces."""
    out = {}
    for name, quad, prec, tprec in (("mixed_precision_only", q64, L.OPSEC_PREC_BF16X3, L.OPSEC_PREC_FP32),
                                    ("bf16x3_projection_only", q64, L.OPSEC_PREC_BF16X3, L.OPSEC_PREC_FP64),
                                    ("fp32_traces_only", q64, L.OPSEC_PREC_FP64, L.OPSEC_PREC_FP32),
                                    ("surrogate_only", qs, L.OPSEC_PREC_FP64, L.OPSEC_PREC_FP64)):
        ctx.pair_tables(funcs, quad)
        ctx.project(funcs[1:], prec, keep_on_device=True)
        e = ctx.estimate_device(lam, y, prior, L.OPSEC_MIX_INVERSE, tprec)
        c = compare_estimates(e64, e)
        c["fisher_rel_err_diag_per_band"] = [float(x) for x in
                                             np.abs(np.diag(e["F"]) - np.diag(e64["F"])) / np.abs(np.diag(e64["F"]))]
        out[name] = c
    return out


def library_ceiling(shapes):
    """Plain BF16 GEMM rate of the vendor library (torch.matmul / torch.bmm on
    ROCm: hipBLASLt, FP32 accumulation) on the projection GEMM shapes: the MFMA
    rate this device sustains in practice, against which the split-BF16x3
    kernels' raw rate (3 MFMA products per useful flop) is read; shapes named
    *_fp64 run in FP64 (rocBLAS / hipBLASLt dgemm), the ceiling of the FP64
    path.  shapes: name -> (batch, m, n, k).  Runs in a child process: torch
    brings its own HIP runtime, which does not share a process with the
    library's."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--library-ceiling-child", json.dumps(shapes)],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr.strip().splitlines()[-1] if r.stderr.strip() else f"exit {r.returncode}"}
    return json.loads(r.stdout.strip().splitlines()[-1])


def _library_ceiling_child(shapes, reps=10):
    import torch
    dev = torch.device("cuda:0")
    out = {}
    for name, (bt, m, n, k) in shapes.items():
        dt = torch.float64 if name.endswith("_fp64") else torch.bfloat16
        a = torch.randn((bt, m, k) if bt > 1 else (m, k), device=dev, dtype=dt)
        b = torch.randn((bt, k, n) if bt > 1 else (k, n), device=dev, dtype=dt)
        f = (lambda: torch.bmm(a, b)) if bt > 1 else (lambda: a @ b)
        for _ in range(3):
            c = f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            c = f()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"shape": [bt, m, n, k], "ms": ms, "tflops": 2.0 * bt * m * n * k / ms / 1e9}
        del a, b, c
        torch.cuda.empty_cache()
    return out


def _blas_rate(N, B, R=1024):
    """FP64 GEMM rate of the host BLAS (numpy's OpenBLAS) on an R x N by N x M
    block of the projection's first product, S_p B (flop/s)."""
    S = np.random.default_rng(0).normal(size=(R, N))
    Bt = np.ascontiguousarray(B.T)            # N x M, the reference's column-major B
    S[:8] @ Bt                                 # warm the thread pool
    t0 = time.perf_counter()
    S @ Bt
    return 2.0 * R * N * B.shape[0] / (time.perf_counter() - t0)


def cpu_baseline(cfg, cells, k, P, lo, hi, B):
    """The reference's own CPU path on a bounded sample of the same workload,
    extrapolated to one S->Fisher pass.

    Pair integrals and the S_p fill: the reference's ComputeSignalMatrixBlock
    (src/sig.cpp:50-123, compiled from its sources into oracle/_ref) with its own
    method, adaptive Genz-Malik pairs at sig.epsrel/epsabs and the translation
    cache.  A call on rows [0, R) x all columns costs a_f (the N distinct
    integrals of function f, all met in row 0) plus R b (the cache fill of a
    row): a_f is timed for the prior and the lowest and highest band at R = 1,
    b from the highest band at a larger R.  The projection: the reference's
    comma does two pdgemm per band (S_p B, then B^T C; src/comma-scalapack.cpp:
    248-252), 2 N^2 M + 2 N M^2 flops, timed here at the rate of the host BLAS
    pdgemm sits on (numpy's OpenBLAS) for an R x N by N x M block.  The
    estimate's packed Gram products (src/estimate-mpi.cpp:501-546) at the
    same BLAS on a sample of the packed entries."""
    import ctypes as Cc
    import tempfile
    from tests import oracles as O
    n = cfg["n"]
    N = n ** 3
    M = cfg["nmodes"]
    nb = cfg["nbands"]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    R = O.ref()
    with tempfile.TemporaryDirectory() as d:
        pk = os.path.join(d, "pk.dat")
        O.write_pkfile(pk, k, P)
        m = O.ref_real_model(pk, lo, hi)
        sv = O.survey_text(n, 0.0, 400.0, nbar=cfg["nbar"])
        cols = np.arange(N, dtype=np.int32)

        def block(f, nrows):
            rows = np.arange(nrows, dtype=np.int32)
            S = np.zeros((nrows, N))
            t0 = time.perf_counter()
            R.ref_sig_block(m, f, sv, cells.ctypes.data, N, O.ptr(rows), nrows, O.ptr(cols), N, O.ptr(S), N, 1e-5,
                            1e-10)
            return time.perf_counter() - t0

        a_prior, a_lo, a_hi = block(0, 1), block(1, 1), block(nb, 1)
        R2 = 2048
        b_row = max(0.0, (block(nb, R2) - a_hi) / (R2 - 1))
        R.ref_model_free(m)
    t_tables = a_prior + nb * 0.5 * (a_lo + a_hi) + (nb + 1) * N * b_row
    rate = _blas_rate(N, B)
    t_proj = nb * (2.0 * N * N * M + 2.0 * N * M * M) / rate
    # estimate: X (nb + 2 rows) times D^T over the packed entries
    Ks = 1 << 20
    Dk = np.random.default_rng(1).normal(size=(nb, Ks))
    Xk = np.random.default_rng(2).normal(size=(nb + 2, Ks))
    t0 = time.perf_counter()
    Xk @ Dk.T
    t_est = (time.perf_counter() - t0) * (M * (M + 1) / 2) / Ks
    t_full = t_tables + t_proj + t_est
    pairs = float(N) * N * (nb + 1)
    return {
        "value": pairs / t_full, "unit": "cell-pair integrals/s", "cores": cores, "kind": "reference",
        "sample": (f"the reference's ComputeSignalMatrixBlock (oracle/_ref, its Genz-Malik pairs and translation "
                   f"cache, OpenMP in the integrand) on row 0 of the prior and of the lowest and highest band "
                   f"({a_prior:.2f} / {a_lo:.2f} / {a_hi:.2f} s: the {N} distinct integrals of each) and on "
                   f"{R2} rows of the highest band (fill {b_row * 1e3:.3f} ms per row); the two pdgemm products "
                   f"per band and the estimate's Gram products at the host BLAS rate (numpy OpenBLAS, "
                   f"{rate / 1e9:.0f} GFLOP/s on a 1024 x {N} by {N} x {M} block); {cores} threads; extrapolated "
                   f"to {nb + 1} matrices: tables {t_tables:.0f} s + projection {t_proj:.0f} s + estimate "
                   f"{t_est:.1f} s = {t_full:.0f} s per S->Fisher pass"),
        "s_per_step_extrapolated": t_full,
        "parts_s": {"tables": t_tables, "projection": t_proj, "estimate": t_est},
    }


def cpu_baseline_sph(cells, dims, k, P, lo, hi, B, budget_s=25.0):
    """Oracle on a bounded sample of the spherical workload, extrapolated: the
    arc-cache block (Genz-Malik pairs, src/sig.cpp:125-213) of the rows of a few
    arcs against all columns, for the prior and the lowest and highest bands,
    one arc per host thread; then S_p B for a block of modes."""
    import ctypes as Cc
    from concurrent.futures import ThreadPoolExecutor
    from tests import oracles as O
    orc = O.oracle()
    nr, nmu, nphi = dims
    narcs = nr * nmu
    N, M, nb = len(cells), B.shape[0], len(lo)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    r768, tabs = O.oracle_tables(k, P, [lo[0], lo[-1]], [hi[0], hi[-1]])
    xis = {"prior": O.XiHandle(0, r768, tabs[0]), "band_lo": O.XiHandle(1, a=float(lo[0]), b=float(hi[0])),
           "band_hi": O.XiHandle(1, a=float(lo[-1]), b=float(hi[-1]))}
    arc = cells["d1"] * nmu + cells["d2"]
    cols = np.arange(N, dtype=np.int32)
    q = (1e-5, 1e-10, 50000000)

    def block(args):
        name, h = args
        rows = np.nonzero(arc == h)[0].astype(np.int32)
        S = np.zeros((len(rows), N))
        t0 = time.perf_counter()
        orc.orc_sig_block_sph(Cc.byref(xis[name].s), cells.ctypes.data, N, nr, nmu, nphi, O.ptr(rows), len(rows),
                              O.ptr(cols), N, O.ptr(S), N, *q)
        return name, time.perf_counter() - t0

    nsamp = max(1, min(cores, narcs))
    arcs = [int(h) for h in np.linspace(0, narcs - 1, nsamp)]
    times = {n: [] for n in xis}
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        for name, dt in ex.map(block, [(n, h) for n in xis for h in arcs]):
            times[name].append(dt)
    t_sample = time.perf_counter() - t0
    per_arc = {n: float(np.mean(v)) for n, v in times.items()}
    t_tables = narcs * (per_arc["prior"] + nb * 0.5 * (per_arc["band_lo"] + per_arc["band_hi"])) / cores
    # the two pdgemm products per band at the host BLAS rate (numpy's OpenBLAS)
    rate = _blas_rate(N, B)
    t_full = t_tables + nb * (2.0 * N * N * M + 2.0 * N * M * M) / rate
    pairs = float(N) * N * (nb + 1)
    return {"value": pairs / t_full, "unit": "cell-pair integrals/s", "cores": cores, "kind": "port",
            "sample": (f"oracle arc-cache blocks (Genz-Malik) of the rows of {nsamp} of {narcs} arcs against all "
                       f"{N} columns for the prior, the lowest and the highest band ({t_sample:.1f} s on {cores} "
                       f"threads; per arc {per_arc['prior']:.2f} / {per_arc['band_lo']:.2f} / "
                       f"{per_arc['band_hi']:.2f} s), bands taken as the mean of the lowest and highest (the reference's "
                       f"spherical survey needs cfitsio and is not built, so this is the oracle's restatement of "
                       f"ComputeSignalMatrixBlockS), and the two pdgemm products per band at the host BLAS rate "
                       f"(numpy OpenBLAS, {rate / 1e9:.0f} GFLOP/s); extrapolated: {t_full:.0f} s per S->Fisher pass"),
            "s_per_step_extrapolated": t_full}


def spherical_main(args):
    """The spherical-shell workload: r in [200, 1000] Mpc/h over the 60-degree cap
    mu >= 0.5, Nr x Nmu x Nphi = (64, 32, 32) / scale cells, nbar(r) = 5e-4 exp(-(r/600)^2)
    as a radial table, 50 log-spaced bands over [0.003, 0.3] h/Mpc, 6000 / scale^3 KL modes.
    Setup (untimed): basis cells on the GPU, the prior's arc-pair tables and klt.  One step:
    arc-pair tables of the prior and all bands (adaptive Genz-Malik pairs; spherical pairs
    have no surrogate), the projection and the estimate -- FP64 with the dense projection
    (the reference's algorithm), and split-BF16x3 + FP32 traces with the --projection method
    (default: Toeplitz, a length-2 Nphi DFT along phi); with Toeplitz also the FP64 path
    with the same DFT."""
    import ctypes as Cc
    from opsec_amd import lib as L
    from opsec_amd import synth
    s = args.scale
    nr, nmu, nphi = 64 // s, 32 // s, 32 // s
    nb, M = 50, (args.modes if args.modes > 0 else max(16, 6000 // s ** 3))
    ctx = L.Context(0)
    ncu, arch = ctx.device_info()
    sv = L.survey_sph(nr, nmu, nphi, 200.0, 1000.0, 0.5, 1.0, 0.0, 2 * np.pi)
    rr = np.linspace(150.0, 1100.0, 1901)
    rn = 5e-4 * np.exp(-(rr / 600.0) ** 2)
    t0 = time.perf_counter()
    cells, nev = ctx.basis(sv, L.OPSEC_SEL_RADIAL, rad_r=rr, rad_n=rn)
    t_basis = time.perf_counter() - t0
    N = len(cells)
    log(f"device {arch} with {ncu} CUs; spherical shell {nr} x {nmu} x {nphi} = {N} cells, {nb} bands, {M} modes; "
        f"basis {t_basis:.2f} s ({nev:.3g} integrand evaluations)")
    edges = np.logspace(np.log10(0.003), np.log10(0.3), nb + 1)
    lo, hi = edges[:-1], edges[1:]
    k, P = synth.pk_table()
    ctx.model_real(k, P, lo, hi, L.OPSEC_XI_CLOSED)
    ctx.cells_upload(cells, sv)
    q = ctx.quad(L.OPSEC_QUAD_GM)
    funcs = list(range(nb + 1))
    t0 = time.perf_counter()
    ctx.pair_tables([0], q)
    t_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    os.environ["OPSEC_KLT_PROGRESS"] = "1"    # one line per Rayleigh-Ritz step (long solves stay visible)
    # the shell's spectrum converges more slowly than the box's: a Krylov space of 6 M
    lam, B, nconv = ctx.klt(M, tol=1e-8, maxdim=min(N, 6 * M + 512))
    t_klt = time.perf_counter() - t0
    ph = (Cc.c_double * 9)()
    ctx.lib.opsec_diag_klt_times(ph, 9)
    klt_phases = {nm: float(ph[i]) for i, nm in enumerate(
        ("block_matvecs", "orthogonalisation", "ritz_tridiagonal_reduction_gpu", "ritz_tridiagonal_eigen_host",
         "ritz_back_transformation", "ritz_vectors"))}
    log(f"setup: prior arc-pair table {t_s:.2f} s, klt {M} modes {t_klt:.1f} s (nconv {nconv}); phases "
        f"{klt_phases}, Krylov dimension {int(ph[6])}, {int(ph[7])} Rayleigh-Ritz steps")
    ctx.modes_upload(B)
    prior = ctx.params()
    y = np.sqrt(1 + lam) * np.random.Generator(np.random.Philox(42)).normal(size=M)
    r64 = run_path(ctx, L, funcs, q, L.OPSEC_PREC_FP64, L.OPSEC_PREC_FP64, lam, y, prior, args.steps, args.warmup,
                   "spherical fp64")
    st64 = ctx.pair_stats()
    toep = args.projection == "toeplitz"
    if toep:
        ctx.set_projection(L.OPSEC_PROJ_TOEPLITZ)
        r64t = run_path(ctx, L, funcs, q, L.OPSEC_PREC_FP64, L.OPSEC_PREC_FP64, lam, y, prior, args.steps,
                        args.warmup, "spherical fp64, Toeplitz projection")
    rmx = run_path(ctx, L, funcs, q, L.OPSEC_PREC_BF16X3, L.OPSEC_PREC_FP32, lam, y, prior, args.steps, args.warmup,
                   "spherical mixed")
    pairs = float(N) * N * (nb + 1)
    qm = float(rmx[2][6])
    out = {
        "metric": METRIC, "value": pairs / (rmx[0] / 1e3), "unit": "cell-pair integrals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": rmx[0], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3+fp32", "data": "synthetic",
        "config": {"workload": (f"C5/{s}: spherical shell r 200-1000 Mpc/h, mu >= 0.5 (60 deg cap), {nr} x {nmu} x "
                                f"{nphi} = {N} cells, nbar(r) = 5e-4 exp(-(r/600)^2), {nb} log bands "
                                f"[0.003, 0.3] h/Mpc, {M} KL modes, adaptive Genz-Malik pairs"),
                   "path": "mixed", "projection": args.projection},
        "fp64_path": {"ms_per_step": r64[0], "value": pairs / (r64[0] / 1e3), "projection": "dense",
                      "projection_ms": [float(r64[2][0]), float(r64[2][1])]},
        "speedup_vs_fp64": r64[0] / rmx[0],
        "projection_ms": [float(rmx[2][0]), float(rmx[2][1])],
        "quadrature": {"ms_per_step": qm, "integrals": int(st64["integrals"]), "node_evaluations": int(st64["nodes"]),
                       "integrals_per_s": st64["integrals"] / (qm / 1e3),
                       "node_evaluations_per_s": st64["nodes"] / (qm / 1e3)},
        "basis": {"s": t_basis, "cells": N, "integrand_evaluations": int(nev), "evaluations_per_s": nev / t_basis},
        "klt": {"s": t_klt, "nconv": int(nconv), "prior_table_s": t_s, "tol": 1e-8, "phases_ms": klt_phases,
                "krylov_dim": int(ph[6]), "rayleigh_ritz_steps": int(ph[7]), "block_steps": int(ph[8])},
    }
    out.update(compare_estimates(r64[1], rmx[1]))
    if toep:
        out["fp64_toeplitz_path"] = {"ms_per_step": r64t[0], "value": pairs / (r64t[0] / 1e3),
                                     "projection_ms": [float(r64t[2][0]), float(r64t[2][1])]}
        out["fp64_toeplitz_path"].update(compare_estimates(r64[1], r64t[1]))
        out["speedup_vs_fp64_same_algorithm"] = r64t[0] / rmx[0]
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline_sph(cells, (nr, nmu, nphi), k, P, lo, hi, B)
        log(f"cpu baseline sample took {time.perf_counter() - t0:.1f} s")
    kg_ms, k1_ms, k2_ms = (float(x) for x in rmx[2][2:5])
    l1, l2 = max(1, int(round(float(rmx[2][7])))), max(1, int(round(float(rmx[2][8]))))
    if toep:
        # stage 1 = 2 Nphi real GEMMs of G x G by G x M per band (G = Nr Nmu arcs), one launch
        # per band; stage 2 = N M^2 per band, one launch per group of bands
        G = nr * nmu
        s1_flops, s2_flops = 2.0 * (2 * nphi) * G * G * M * nb, float(N) * M * M * nb
        dom = "stage1" if k1_ms >= k2_ms else "stage2"
        out["roofline"] = {
            "bound": "mfma", "peak": PEAK_TF["bf16"] / 3, "unit": "TFLOP/s", "traffic": None,
            "achieved": (s1_flops / k1_ms if dom == "stage1" else s2_flops / k2_ms) * 1e3 / 1e12,
            "launch_ms": k1_ms / l1 if dom == "stage1" else k2_ms / l2,
            "kernel": ("gemm3_kernel<ASrcPlanes,EpiYhat> (Toeplitz stage 1, 2 Nphi arc GEMMs per launch, one "
                       "launch per band)" if dom == "stage1" else
                       f"gemm3_kernel<ASrcPlanes,EpiPackedParts> (stage 2, split-K, {nb // l2} bands per launch)"),
            "note": ("useful flops per band: stage 1 2 (2 Nphi) G^2 M, stage 2 N M^2; "
                     "peak = dense bf16 2516.6 TF / 3 products"),
            "per_step_ms": {"that_inverse_dft_and_split_k_sum": kg_ms, "stage1_gemm": k1_ms, "stage2_gemm": k2_ms},
            "stage1_tflops": s1_flops / k1_ms * 1e3 / 1e12, "stage2_tflops": s2_flops / k2_ms * 1e3 / 1e12}
    else:
        s1 = k1_ms / l1
        out["roofline"] = {"bound": "mfma", "achieved": 2.0 * N * N * M / (s1 / 1e3) / 1e12,
                           "peak": PEAK_TF["bf16"] / 3, "unit": "TFLOP/s", "traffic": None, "launch_ms": s1,
                           "kernel": "gemm3_kernel<ASrcPlanes,EpiPlanes> (stage 1, S_p planes, one launch per band)"}
    out["roofline"]["frac"] = out["roofline"]["achieved"] / out["roofline"]["peak"]
    ctx.close()
    print(json.dumps(out))


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--library-ceiling-child":
        print(json.dumps(_library_ceiling_child({k: tuple(v) for k, v in json.loads(sys.argv[2]).items()})))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C2", choices=["C1", "C2", "C5"])
    ap.add_argument("--scale", type=int, default=2, choices=[1, 2, 4],
                    help="C5 only: grid divisor per dimension (1 = the full 65536-cell shell)")
    ap.add_argument("--modes", type=int, default=0,
                    help="C5 only: KL modes (default 6000 / scale^3)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--path", default="both", choices=["both", "fp64", "surrogate"])
    ap.add_argument("--mlp-samples", type=int, default=200000, help="surrogate training pairs (setup)")
    ap.add_argument("--mlp-epochs", type=int, default=40, help="surrogate training epochs (setup)")
    ap.add_argument("--diagnose", action="store_true",
                    help="also attribute the Fisher error to the surrogate and to the mixed precision (untimed)")
    ap.add_argument("--nt-tile", type=int, default=2, choices=[0, 1, 2],
                    help="split-BF16x3 GEMM tile: 0 = 128 x 128, 1 = 256 x 128, 2 = 256 x 256 double-buffered")
    ap.add_argument("--kl", default="lanczos", choices=["fourier", "lanczos"],
                    help="KL modes (setup, untimed): the klt solver (block Lanczos on the device-resident S, "
                         "checked against the circulant's spectrum) or the box's analytic Fourier modes")
    ap.add_argument("--aux-steps", type=int, default=3,
                    help="timed steps of the comparison paths (FP64 dense and Toeplitz, quadrature-only mixed, "
                         "dense mixed); the headline path runs --steps")
    ap.add_argument("--nt-gen", type=int, default=0, choices=[0, 1],
                    help="stage 1 generates S_p from the pair table (1) or reads S_p planes (0)")
    ap.add_argument("--projection", default="toeplitz", choices=["dense", "toeplitz"],
                    help="split-BF16x3 projection: dense S_p (2 N^2 M stage 1) or Toeplitz (DFT along x of the "
                         "box, 2 Nx slab GEMMs; along phi of the shell, 2 Nphi arc GEMMs); the FP64 path is "
                         "dense (the reference's algorithm), timed with Toeplitz too as a comparison")
    ap.add_argument("--s2-split", type=int, default=8, help="split-K factor of the split-BF16x3 stage 2")
    ap.add_argument("--s2-split-toep", type=int, default=4,
                    help="split-K factor of the split-BF16x3 Toeplitz stage 2 (eight bands per launch)")
    ap.add_argument("--nt-group", type=int, default=8, help="split-BF16x3 GEMM tile-group edge (stage 1)")
    ap.add_argument("--s2-group", type=int, default=2, help="split-BF16x3 stage-2 tile-group edge")
    ap.add_argument("--nt-remap", type=int, default=0, choices=[0, 1], help="deal tile groups to XCDs")
    ap.add_argument("--g3-dl", type=int, default=2, choices=[0, 1, 2],
                    help="256 x 256 split-BF16x3 GEMM: operand tiles through VGPRs (0) or by DMA loads into LDS "
                         "(1; 2: issued in two halves per step)")
    ap.add_argument("--library-ceiling", action="store_true",
                    help="also time the vendor library's plain BF16 GEMM on the projection shapes (imports torch)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="also time a plain device copy of this many GiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (sharding the bands over "
              "several GPUs is designed but not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.config == "C5":
        spherical_main(args)
        return

    from opsec_amd import lib as L
    from opsec_amd import synth

    cfg = dict(synth.CONFIGS[args.config])
    n, nb, M = cfg["n"], cfg["nbands"], cfg["nmodes"]
    N = n ** 3
    cells = synth.box_cells(n, 0.0, 400.0, cfg["nbar"])
    sv = L.survey_box(n, 0.0, 400.0)
    k, P = synth.pk_table()
    lo, hi = synth.linear_bands(nb, cfg["kmin"], cfg["kmax"])
    funcs = list(range(nb + 1))

    ctx = L.Context(0)
    ctx.lib.opsec_diag_set_nt_tile(args.nt_tile)
    ctx.lib.opsec_diag_set_nt_order(args.nt_group, args.nt_remap)
    ctx.lib.opsec_diag_set_nt_gen(args.nt_gen)
    ctx.lib.opsec_diag_set_stage2_split(args.s2_split)
    ctx.lib.opsec_diag_set_toeplitz_stage2_split(args.s2_split_toep)
    ctx.lib.opsec_diag_set_stage2_group(args.s2_group)
    ctx.lib.opsec_diag_set_g3_dl(args.g3_dl)
    ctx.set_projection(L.OPSEC_PROJ_TOEPLITZ if args.projection == "toeplitz" else L.OPSEC_PROJ_DENSE)
    ncu, arch = ctx.device_info()
    log(f"device {arch} with {ncu} CUs; workload {args.config}: {N} cells, {nb} bands, {M} modes")
    ctx.model_real(k, P, lo, hi, L.OPSEC_XI_CLOSED)
    ctx.cells_upload(cells, sv)
    q64 = ctx.quad(L.OPSEC_QUAD_GAUSS, nodes=cfg["nodes"], near=1)

    # --- setup (untimed): KL modes of the prior S, pixel values --------------------
    t0 = time.perf_counter()
    ctx.pair_tables([0], q64)
    g = ctx.pair_table(0).reshape(n, n, n) * cells["Nbar"][0]
    lam, B = fourier_kl_modes(g, M)
    klt_info = None
    if args.kl == "lanczos":
        # the klt path: block Lanczos on the device-resident S (8.6 GB at 32^3),
        # checked against the circulant's spectrum (real DFT of the generator)
        import ctypes as C
        t1 = time.perf_counter()
        ev, V, nconv = ctx.klt(M, tol=1e-8)
        t_klt = time.perf_counter() - t1
        err = float(np.max(np.abs(ev - lam)) / lam[0])
        ph = (C.c_double * 9)()
        ctx.lib.opsec_diag_klt_times(ph, 9)
        names = ("block_matvecs", "orthogonalisation", "ritz_tridiagonal_reduction_gpu", "ritz_tridiagonal_eigen_host",
                 "ritz_back_transformation", "ritz_vectors")
        klt_info = {"s": t_klt, "nconv": int(nconv), "eigenvalue_max_err_over_lambda_max_vs_dft": err,
                    "phases_ms": {nm: float(ph[i]) for i, nm in enumerate(names)},
                    "krylov_dim": int(ph[6]), "rayleigh_ritz_steps": int(ph[7]), "block_steps": int(ph[8]),
                    "tol": 1e-8}
        log(f"setup: klt (block Lanczos) {M} modes in {t_klt:.1f} s, nconv {nconv}, "
            f"max |lambda - DFT| / lambda_max {err:.2e}; phases {klt_info['phases_ms']}, K {int(ph[6])}")
        lam, B = ev, V
    ctx.modes_upload(B)
    prior = ctx.params()
    # Assembly of the dense S_0 (what klt's matvecs read; the projection never
    # stores S_p): 8 N^2 bytes written, timed once in the setup
    dS = ctx.dev_alloc(N * N * 8)
    ctx.sync()
    ta = []
    for _ in range(3):
        t1 = time.perf_counter()
        ctx.check(ctx.lib.opsec_assemble_device(ctx.h, 0, dS, N), "assemble")
        ctx.sync()
        ta.append(time.perf_counter() - t1)
    ctx.dev_free(dS)
    assembly = {"bytes": 8 * N * N, "ms": 1e3 * min(ta), "GB_per_s": 8 * N * N / min(ta) / 1e9,
                "kernel": "assemble_kernel (S_0, FP64, row-major)", "timer": "host clock around the call, best of 3"}
    z = np.random.Generator(np.random.Philox(42)).normal(size=M)
    y = np.sqrt(1 + lam) * z
    log(f"setup: KL modes (lambda {lam[0]:.3g} .. {lam[-1]:.3g}) in {time.perf_counter() - t0:.1f} s")

    # surrogate training (setup, untimed): exact pair integrals of seeded pairs
    t0 = time.perf_counter()
    qs = ctx.quad(L.OPSEC_QUAD_SURROGATE, nodes=cfg["nodes"])
    qs.near = 2
    opts = ctx.mlp_opts(nsamples=args.mlp_samples, epochs=args.mlp_epochs, batch=4096, seed=7, lr=2e-3, near=2.0)
    loss = ctx.mlp_train(opts, q64)
    t_train = time.perf_counter() - t0
    log(f"surrogate trained on {opts.nsamples} exact integrals in {t_train:.1f} s, final loss {loss:.3e}")

    results = {}
    aux_steps, aux_warm = min(args.steps, args.aux_steps), min(args.warmup, 1)
    if args.path in ("both", "fp64"):
        # the reference's algorithm in FP64: dense projection (comma's two pdgemm per band)
        ctx.set_projection(L.OPSEC_PROJ_DENSE)
        ms, est64, k64 = run_path(ctx, L, funcs, q64, L.OPSEC_PREC_FP64, L.OPSEC_PREC_FP64, lam, y, prior,
                                  aux_steps, aux_warm, "fp64")
        st = ctx.pair_stats()
        results["fp64"] = dict(ms=ms, est=est64, kms=k64, stats=st)
        tables64 = [ctx.pair_table(f) for f in funcs]   # the quadrature tables of S and every C_{,p}
        # the same Toeplitz algorithm as the headline path, in FP64
        ctx.set_projection(L.OPSEC_PROJ_TOEPLITZ)
        ms_t, est_t, k_t = run_path(ctx, L, funcs, q64, L.OPSEC_PREC_FP64, L.OPSEC_PREC_FP64, lam, y, prior,
                                    aux_steps, aux_warm, "fp64 toeplitz")
        results["fp64_toeplitz"] = dict(ms=ms_t, est=est_t, kms=k_t)
        ctx.set_projection(L.OPSEC_PROJ_TOEPLITZ if args.projection == "toeplitz" else L.OPSEC_PROJ_DENSE)
        if args.path == "both":
            # the mixed path on the quadrature tables: what the surrogate alone saves
            ms_q, est_q, k_q = run_path(ctx, L, funcs, q64, L.OPSEC_PREC_BF16X3, L.OPSEC_PREC_FP32, lam, y, prior,
                                        aux_steps, aux_warm, "mixed, quadrature tables")
            results["mixed_quadrature"] = dict(ms=ms_q, est=est_q, kms=k_q)
    if args.path in ("both", "surrogate"):
        ms, ests, ks = run_path(ctx, L, funcs, qs, L.OPSEC_PREC_BF16X3, L.OPSEC_PREC_FP32, lam, y, prior,
                                args.steps, args.warmup, "surrogate")
        st = ctx.pair_stats()
        results["surrogate"] = dict(ms=ms, est=ests, kms=ks, stats=st)
        if "fp64" in results:
            # error of the surrogate path's S and C_{,p} (their distinct pair integrals) against quadrature:
            # per function max |dQ| / max |Q|, and the largest per-entry error relative to the entry
            errs, rel = [], []
            for f in funcs:
                q_s, q_64 = ctx.pair_table(f), tables64[f]
                errs.append(float(np.max(np.abs(q_s - q_64)) / np.max(np.abs(q_64))))
                big = np.abs(q_64) > 1e-3 * np.max(np.abs(q_64))
                rel.append(float(np.max(np.abs(q_s - q_64)[big] / np.abs(q_64[big]))))
            results["s_error"] = {"max_abs_over_function_max": max(errs), "prior_max_abs_over_max": errs[0],
                                  "max_rel_on_entries_above_1e-3_of_max": max(rel),
                                  "note": "S and all C_{,p}: the surrogate path's pair tables against the 6^3 "
                                          "Gauss quadrature tables of the FP64 path, per function"}
        if args.projection == "toeplitz":
            # the same mixed path with the dense projection, for comparison
            ctx.set_projection(L.OPSEC_PROJ_DENSE)
            ms_d, est_d, ks_d = run_path(ctx, L, funcs, qs, L.OPSEC_PREC_BF16X3, L.OPSEC_PREC_FP32, lam, y, prior,
                                         aux_steps, aux_warm, "surrogate, dense projection")
            ctx.set_projection(L.OPSEC_PROJ_TOEPLITZ)
            results["surrogate_dense"] = dict(ms=ms_d, est=est_d, kms=ks_d)

    pairs = float(N) * N * (nb + 1)
    main_path = "surrogate" if "surrogate" in results else "fp64"
    R = results[main_path]
    out = {
        "metric": METRIC,
        "value": pairs / (R["ms"] / 1e3),
        "unit": "cell-pair integrals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+fp32" if main_path == "surrogate" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {n}^3 periodic box 400 Mpc/h, nbar 3e-4, {nb} linear bands "
                               f"[{cfg['kmin']},{cfg['kmax']}] h/Mpc, {M} KL modes, {cfg['nodes']}^3 Gauss pair "
                               f"quadrature", "path": main_path,
                   "projection": args.projection if main_path == "surrogate" else "dense"},
    }
    # distinct (pair, function) integrals behind one step: the value counts the
    # N^2 (1 + nb) matrix entries the step delivers, the Toeplitz path never
    # forms an entry of S, and a step integrates each distinct key once
    nkeys = ctx.nkeys()
    out["distinct_integrals_per_step"] = nkeys * (nb + 1)
    out["distinct_integrals_per_s"] = nkeys * (nb + 1) / (R["ms"] / 1e3)
    if "fp64" in results:
        k64 = results["fp64"]["kms"]
        out["fp64_path"] = {"ms_per_step": results["fp64"]["ms"], "steps": aux_steps,
                            "value": pairs / (results["fp64"]["ms"] / 1e3),
                            "projection": "dense (the reference's two GEMMs per band)",
                            "projection_ms": [float(k64[0]), float(k64[1])],
                            "stage1_tflops": 2.0 * N * N * M * nb / (float(k64[0]) / 1e3) / 1e12}
        kt = results["fp64_toeplitz"]["kms"]
        out["fp64_toeplitz_path"] = {"ms_per_step": results["fp64_toeplitz"]["ms"], "steps": aux_steps,
                                     "value": pairs / (results["fp64_toeplitz"]["ms"] / 1e3),
                                     "projection_ms": [float(kt[0]), float(kt[1])],
                                     "stage1_tflops": 2.0 * (2 * n) * n ** 4 * M * nb / (float(kt[3]) / 1e3) / 1e12,
                                     "stage2_tflops": float(N) * M * M * nb / (float(kt[4]) / 1e3) / 1e12}
        out["fp64_toeplitz_path"].update(compare_estimates(results["fp64"]["est"], results["fp64_toeplitz"]["est"]))
    # pair quadrature: distinct integrals and node evaluations per second
    # against the device's measured FP64 sincos rate (band nodes need two pairs)
    import ctypes as C
    sc = C.c_double()
    ctx.check(ctx.lib.opsec_diag_sincos_rate(ctx.h, C.byref(sc)), "sincos probe")
    out["quadrature"] = {"sincos_pairs_per_s": sc.value}
    out["assembly"] = assembly
    for name in ("fp64", "surrogate"):
        if name in results:
            qm = float(results[name]["kms"][6])
            st = results[name]["stats"]
            out["quadrature"][name + "_path"] = {
                "ms_per_step": qm, "integrals": int(st["integrals"]), "node_evaluations": int(st["nodes"]),
                "surrogate_served": int(st["surrogate"]),
                "integrals_per_s": st["integrals"] / (qm / 1e3),
                "node_evaluations_per_s": st["nodes"] / (qm / 1e3),
                "node_rate_over_half_sincos_rate": st["nodes"] / (qm / 1e3) / (sc.value / 2)}
    if "surrogate_dense" in results:
        kd = results["surrogate_dense"]["kms"]
        out["dense_projection_path"] = {
            "ms_per_step": results["surrogate_dense"]["ms"],
            "stage1_gemm_ms_per_band": float(kd[3]) / float(kd[5]),
            "stage1_useful_tflops": 2.0 * N * N * M / (float(kd[3]) / float(kd[5]) / 1e3) / 1e12,
            "per_step_ms": {"s_planes": float(kd[2]), "stage1_gemm": float(kd[3]), "stage2_gemm": float(kd[4])}}
        if "fp64" in results:
            out["dense_projection_path"]["speedup_vs_fp64"] = results["fp64"]["ms"] / results["surrogate_dense"]["ms"]
            out["dense_projection_path"].update(compare_estimates(results["fp64"]["est"],
                                                                  results["surrogate_dense"]["est"]))
    if "fp64" in results and "surrogate" in results:
        out["speedup_vs_fp64"] = results["fp64"]["ms"] / results["surrogate"]["ms"]
        out["speedup_vs_fp64_same_algorithm"] = results["fp64_toeplitz"]["ms"] / results["surrogate"]["ms"]
        if "mixed_quadrature" in results:
            mq = results["mixed_quadrature"]
            out["mixed_quadrature_path"] = {"ms_per_step": mq["ms"], "steps": aux_steps,
                                            "pair_tables_ms": float(mq["kms"][6])}
            out["mixed_quadrature_path"].update(compare_estimates(results["fp64"]["est"], mq["est"]))
            out["surrogate_share"] = {
                "step_ms_saved": mq["ms"] - results["surrogate"]["ms"],
                "fraction_of_step": (mq["ms"] - results["surrogate"]["ms"]) / mq["ms"],
                "pair_tables_ms_quadrature": float(mq["kms"][6]),
                "pair_tables_ms_surrogate": float(results["surrogate"]["kms"][6])}
        out.update(compare_estimates(results["fp64"]["est"], results["surrogate"]["est"]))
        st = results["surrogate"]["stats"]
        out["surrogate_stats"] = {k2: int(v) for k2, v in st.items()}
        if "s_error" in results:
            out["s_error"] = results["s_error"]
        out["surrogate_train_s"] = t_train
        out["surrogate_train_loss"] = loss
        if args.diagnose:
            out["error_sources"] = diagnose(ctx, L, funcs, q64, qs, lam, y, prior, results["fp64"]["est"])
    # roofline of the dominant kernel: the stage-1 projection GEMM C_p = S_p B'
    # (useful flops 2 N^2 M per band), average launch time from HIP events on
    # the context's stream
    if main_path == "surrogate" and args.projection == "toeplitz":
        # stage 1 = 2 Nx real GEMMs of G x G by G x M per band (G = Ny Nz), one launch per band;
        # stage 2 = N M^2 per band, one launch per group of bands
        kg_ms, k1_ms, k2_ms, launches = [float(x) for x in R["kms"][2:6]]
        l1, l2 = int(round(float(R["kms"][7]))), int(round(float(R["kms"][8])))
        G = n * n
        s1_flops, s2_flops = 2.0 * (2 * n) * G * G * M * nb, float(N) * M * M * nb   # per step
        dom = "stage1" if k1_ms >= k2_ms else "stage2"
        per_launch = k1_ms / l1 if dom == "stage1" else k2_ms / l2
        achieved = (s1_flops / k1_ms if dom == "stage1" else s2_flops / k2_ms) * 1e3 / 1e12
        peak = PEAK_TF["bf16"] / 3
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                           "frac": achieved / peak,
                           "traffic": TRAFFIC_BYTES.get((args.config, "toeplitz", dom)),
                           "kernel": ("gemm3_kernel<ASrcPlanes,EpiYhat> (Toeplitz stage 1, 2 Nx slab GEMMs "
                                      "per launch, one launch per band)" if dom == "stage1" else
                                      f"gemm3_kernel<ASrcPlanes,EpiPackedParts> (stage 2, split-K, "
                                      f"{nb // l2} bands per launch)"),
                           "launch_ms": per_launch,
                           "note": ("useful flops per band: stage 1 2 (2 Nx) G^2 M, stage 2 N M^2; "
                                    "peak = dense bf16 2516.6 TF / 3 products"),
                           "per_step_ms": {"that_inverse_dft_and_split_k_sum": kg_ms, "stage1_gemm": k1_ms,
                                           "stage2_gemm": k2_ms},
                           "stage1_tflops": s1_flops / k1_ms * 1e3 / 1e12,
                           "stage2_tflops": s2_flops / k2_ms * 1e3 / 1e12}
    elif main_path == "surrogate":
        kg_ms, k1_ms, k2_ms, launches = [float(x) for x in R["kms"][2:6]]
        per_launch = k1_ms / launches
        achieved = 2.0 * N * N * M / (per_launch / 1e3) / 1e12
        peak = PEAK_TF["bf16"] / 3
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                           "frac": achieved / peak,
                           "traffic": (TRAFFIC_BYTES.get((args.config, "dense", "stage1"))
                                       if (args.nt_tile, args.nt_gen) == (2, 0) else None),
                           "kernel": stage1_kernel_name(args.nt_tile, args.nt_gen),
                           "launch_ms": per_launch,
                           "note": "useful flops 2 N^2 M per launch; peak = dense bf16 2516.6 TF / 3 products",
                           "mfma_utilisation": 3 * achieved / PEAK_TF["bf16"],
                           "per_step_ms": {"s_planes": kg_ms, "stage1_gemm": k1_ms, "stage2_gemm": k2_ms}}
    else:
        s1_ms = float(R["kms"][0])
        achieved = 2.0 * N * N * M * nb / (s1_ms / 1e3) / 1e12
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TF["fp64"], "unit": "TFLOP/s",
                           "frac": achieved / PEAK_TF["fp64"], "traffic": None,
                           "kernel": "gemm_kernel<double> (projection stage 1, all bands in one launch)",
                           "launch_ms": s1_ms}
    if args.copy_gib > 0:
        import ctypes as C
        nbytes = int(args.copy_gib * (1 << 30)) // 16 * 16
        src, dst = ctx.dev_alloc(nbytes), ctx.dev_alloc(nbytes)
        ms = C.c_float()
        best = None
        for _ in range(3):
            ctx.check(ctx.lib.opsec_diag_copy(ctx.h, C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes),
                                                  C.byref(ms)), "copy")
            best = ms.value if best is None else min(best, ms.value)
        out["copy_calibration"] = {"bytes": nbytes, "ms": best, "GB_per_s": 2 * nbytes / (best / 1e3) / 1e9,
                                   "kernel": "copy16_kernel"}
        ctx.dev_free(src)
        ctx.dev_free(dst)
    if klt_info is not None:
        out["klt"] = klt_info
    if args.library_ceiling and main_path == "surrogate" and args.projection == "toeplitz":
        Mp = (M + 255) // 256 * 256
        lc = library_ceiling({"toeplitz_stage1": (2 * n, n * n, Mp, n * n), "toeplitz_stage2": (1, Mp, Mp, N),
                              "dense_stage1": (1, N, Mp, N), "toeplitz_stage1_fp64": (2 * n, n * n, M, n * n),
                              "stage2_fp64": (1, M, M, N), "dense_stage1_fp64": (1, N // 4, M, N)})
        rf = out["roofline"]
        if "error" not in lc:
            lc["kernel_raw_mfma_tflops"] = {"toeplitz_stage1": 3 * rf["stage1_tflops"],
                                            "toeplitz_stage2": 3 * rf["stage2_tflops"]}
            lc["kernel_over_library"] = {k2: lc["kernel_raw_mfma_tflops"][k2] / lc[k2]["tflops"]
                                         for k2 in ("toeplitz_stage1", "toeplitz_stage2")}
        out["library_ceiling"] = lc
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(cfg, cells, k, P, lo, hi, B)
        log(f"cpu baseline sample took {time.perf_counter() - t0:.1f} s")
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
