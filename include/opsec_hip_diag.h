/* opsec_hip_diag.h -- diagnostics and tuning knobs of libopsec_hip.
 *
 * Not part of the drop-in boundary (opsec_hip.h): none of these has a
 * counterpart in the reference.  The benchmark uses them to time kernels and
 * to calibrate HBM counters, and the tests use them to run every GEMM
 * variant and the host tridiagonal eigen-solver.  The tuning setters change
 * process-wide defaults; leave them at their defaults in production.
 */
#ifndef OPSEC_HIP_DIAG_H
#define OPSEC_HIP_DIAG_H

#include <stddef.h>

#include "opsec_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device-to-device copy of `bytes` (a multiple of 16) by a plain 16-byte copy
 * kernel on the context stream; *ms = its duration.  The calibration kernel
 * for the HBM byte counters. */
int opsec_diag_copy(opsec_ctx* ctx, void* dst, const void* src, size_t bytes, float* ms);
/* Throughput of a kernel doing nothing but FP64 sin/cos pairs (pairs/s): the
 * ceiling against which the quadrature's node rate is read. */
int opsec_diag_sincos_rate(opsec_ctx* ctx, double* pairs_per_s);
/* Top-nev eigenpairs of the symmetric tridiagonal (d[n], e[n-1]) by the
 * solver's host QL + inverse iteration; Z is n x nev column-major. */
int opsec_diag_tridiag_top(int n, const double* d, const double* e, int nev, double* evals, double* Z);

/* Phases of the last Lanczos solve (opsec_klt / opsec_lanczos_dense; the
 * last pass when retries ran): out[0..5] wall-clock ms of the block matvecs,
 * the orthogonalisation, the Rayleigh-Ritz tridiagonal reduction (GPU), the
 * host tridiagonal eigenpairs, the back-transformation and the Ritz vectors;
 * out[6] the final Krylov dimension, out[7] the Rayleigh-Ritz steps, out[8]
 * the block steps.  Returns the number of entries (9) or OPSEC_EINVAL. */
int opsec_diag_klt_times(double* out, int n);

/* Timings of the last opsec_project call (ms): stage 1 and stage 2 totals;
 * for the split-BF16x3 paths also plane generation / DFTs, the stage-1 and
 * stage-2 GEMMs and the launch counts. */
int opsec_diag_project_times(opsec_ctx* ctx, float* stage1, float* stage2);
int opsec_diag_project_kernel_ms(opsec_ctx* ctx, float* gen, float* gemm1, float* gemm2, int* launches);
int opsec_diag_project_launches(opsec_ctx* ctx, int* stage1, int* stage2);

/* Split-BF16x3 GEMM variants: tile (0: 128 x 128, 1: 256 x 128, 2: 256 x 256
 * double-buffered, the default), stage-1 A operand generated from the pair
 * table (1) or read from S_p planes (0, default), stage-2 split-K factor
 * (default 8; the Toeplitz path's, eight bands per launch, default 4) and
 * tile-group edge (default 2), stage-1 tile order (group edge, XCD remap). */
int opsec_diag_set_nt_tile(int v);
int opsec_diag_set_nt_gen(int v);
int opsec_diag_set_stage2_split(int v);
int opsec_diag_set_toeplitz_stage2_split(int v);
int opsec_diag_set_stage2_group(int v);
int opsec_diag_set_nt_order(int group, int remap);
/* Operand tiles of the 256 x 256 split-BF16x3 GEMM (A from plane pairs):
 * through VGPRs and LDS stores (0) or by global-to-LDS DMA loads, all of a
 * step's loads issued at its start (1) or in two halves, at its start and
 * mid-step (2, the default). */
int opsec_diag_set_g3_dl(int v);
/* Pair tables of a spherical survey: the sweep of the functions without band
 * edges (the prior, whose near-diagonal integrals run to the evaluation cap)
 * on a second stream from a second host thread, beside the band functions'
 * sweep (1, the default), or after it (0).  The tables are the same bits. */
int opsec_diag_set_gm_overlap(int v);
/* Pair tables of a spherical survey, the sweep of closed-form bands: first
 * regions of every band of an arc pair in one block, the pair's geometry and
 * the band-edge terms sin kr - kr cos kr formed once and shared by adjacent
 * bands (1, the default), or one thread per (pair, band) job (0).  The same
 * bits either way. */
int opsec_diag_set_gm_bands(int v);
/* Wavefront slots of the adaptive cubature's deep rerun (pools for the whole
 * evaluation cap): 0 (the default) takes as many as half the free device memory
 * holds, up to eight per CU (three per CU for a shell's prior sweep beside its
 * band sweep); v > 0 caps them at v.  The tables are the same bits. */
int opsec_diag_set_gm_deep_slots(int v);

#ifdef __cplusplus
}
#endif
#endif /* OPSEC_HIP_DIAG_H */
