/* opsec_hip.h -- C ABI of the MI355X-native OPSEC hot path.
 *
 * Plain C: fixed-size records, pointers and sizes, no C++ or torch types.
 * Every call returns int: 0 on success, a negative OPSEC_E* code otherwise;
 * opsec_last_error(ctx) returns the message of the last failure on that
 * context.  Host buffers belong to the caller; device buffers made by a call
 * belong to the context and live until it is destroyed (or the matching
 * *_free call).  No call keeps a caller pointer after it returns.
 *
 * A context owns one HIP stream on one device.  A context is not thread
 * safe: use one context per host thread.
 *
 * Reference interfaces replaced (paths relative to code/opsec/):
 *   ComputeSignalMatrixBlock      src/sig.h:20-25, src/sig.cpp:50-123
 *   ComputeSignalC / cubature     src/sig.cpp:388-434, src/cubature.c:885-1042
 *   RealModel / TestModel xi      src/RealModel.cpp:98-127, src/Model.cpp:84-123,
 *                                 src/TestModel.cpp:13-85
 *   Solver::Solve (PARPACK)       src/Solver.h:11-25, src/ParpackSolver.cpp:68-227
 *   comma projection loop         src/comma-scalapack.cpp:228-266, src/slp.cpp:107-120
 *   estimate compute + post       src/estimate-mpi.cpp:480-654, src/MyMatrix.hpp:478-569
 */
#ifndef OPSEC_HIP_H
#define OPSEC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPSEC_HIP_ABI_VERSION 1

/* ---- error codes ------------------------------------------------------- */
enum {
    OPSEC_OK = 0,
    OPSEC_EINVAL = -1,      /* invalid argument (sizes, strides, modes)      */
    OPSEC_ENOMEM = -2,      /* device or host allocation failed              */
    OPSEC_EHIP = -3,        /* HIP runtime error                             */
    OPSEC_ECONV = -4,       /* iterative method did not converge             */
    OPSEC_ESTATE = -5,      /* call out of order (e.g. no model uploaded)    */
    OPSEC_ENODEV = -6       /* no usable gfx950 device                       */
};

/* ---- records ----------------------------------------------------------- */

/* The 80-byte cell record of a cellfile ("4i8d"), src/Cell.h:74-101. */
typedef struct opsec_cell {
    int32_t a;                       /* cell number, 0 <= a < Ncells         */
    int32_t d1, d2, d3;              /* grid index (dx,dy,dz) or (dr,dmu,dphi) */
    double min1, max1, min2, max2, min3, max3;   /* cell bounds             */
    double Veff;                     /* effective volume                     */
    double Nbar;                     /* expected galaxy count                */
} opsec_cell;

enum { OPSEC_COORD_CARTESIAN = 1, OPSEC_COORD_SPHERICAL = 2 };   /* src/Cell.h:52-56 */

/* Survey geometry, the "survey." keys, src/Survey.cpp:18-53, src/BoxSurvey.cpp:85-91.
 * For a box survey L is the period of the minimum-image separation
 * (src/BoxSurvey.cpp:26-37); L >= 2 x extent makes the box aperiodic. */
typedef struct opsec_survey {
    int32_t coordsys;                /* OPSEC_COORD_*                        */
    int32_t n1, n2, n3;              /* Nx,Ny,Nz or Nr,Nmu,Nphi              */
    double min1, max1, min2, max2, min3, max3;
    double L;                        /* box period (Cartesian only)          */
} opsec_survey;

/* ---- context ----------------------------------------------------------- */
typedef struct opsec_ctx opsec_ctx;

int opsec_ctx_create(int device, opsec_ctx** out);
int opsec_ctx_destroy(opsec_ctx* ctx);
const char* opsec_last_error(const opsec_ctx* ctx);
int opsec_ctx_sync(opsec_ctx* ctx);                 /* wait for the context stream */
int opsec_abi_version(void);
/* Device properties: number of CUs and gfx arch name (e.g. "gfx950"). */
int opsec_device_info(opsec_ctx* ctx, int* num_cus, char* arch, int arch_len);
/* The context stream as an opaque handle (a hipStream_t). */
void* opsec_ctx_stream(opsec_ctx* ctx);

/* ---- model: the xi functions ------------------------------------------- */
/* Function index f: 0 = prior xi (Model::GetXi), 1+p = dxi/dp_p
 * (Model::GetXiDeriv(p)), 0 <= p < nparams. */

enum {
    OPSEC_XI_TABLE = 0,     /* reference tables: Simpson xi_0^2 on a QSpline
                               (src/Model.cpp:84-123, src/QSpline.cpp:6-70)   */
    OPSEC_XI_CLOSED = 1     /* band derivatives in closed form,
                               xi_p(r) = [sin kr - kr cos kr]_{klo}^{khi}/(2 pi^2 r^3),
                               prior still from its table                     */
};

/* RealModel (src/RealModel.cpp:33-127).  pk_k/pk_P: the pkfile samples (must
 * span [0,10]; the prior integral runs over [0,10] with Nk=32768 and
 * exp(-k^2) damping).  band_lo/band_hi: the nbands band edges.  The tables
 * are built on the device. */
int opsec_model_real(opsec_ctx* ctx, const double* pk_k, const double* pk_P, int npk,
                     int nbands, const double* band_lo, const double* band_hi, int xi_mode);

/* TestModel (src/TestModel.cpp:13-85): form 0 = A exp(-b r), 1 = A exp(-b r^2).
 * Two parameters; derivative functions follow the reference's definitions. */
int opsec_model_test(opsec_ctx* ctx, int form, double A, double b);

/* A model given by its tables: nfuncs functions (f = 0 the prior xi, f = 1+p
 * the derivative for parameter p), each as its values at the 768 QSpline
 * nodes r_i (i < 512: r = 200 i / 512; then 256 nodes log-spaced over
 * [200, 20000], QSpline::MakeArray, src/QSpline.cpp:6-17; the nodes are what
 * opsec_model_table returns in r768).  xi768 holds nfuncs x 768 values;
 * params the nfuncs - 1 prior parameter values p_n (NULL: zeros).  A RealModel
 * XiFunc (src/RealModel.cpp:19-30, src/XiFunc.h:22-44) is exactly such a
 * QSpline, so a caller holding an XiFunc -- ComputeSignalMatrixBlock(...,
 * const XiFunc& xi, ...), src/sig.h:20-25 -- uploads xi at these nodes and
 * passes its index as f: the device interpolates as QSpline::operator()
 * does (src/QSpline.cpp:61-70) and gives the same values. */
int opsec_model_tables(opsec_ctx* ctx, int nfuncs, const double* xi768, const double* params);

int opsec_model_nparams(opsec_ctx* ctx, int* nparams);
/* Prior parameter values p_n (RealModel::GetParam, src/RealModel.cpp:85-96). */
int opsec_model_params(opsec_ctx* ctx, double* params);
/* Evaluate function f at separations r[0..n-1] on the device. */
int opsec_model_xi(opsec_ctx* ctx, int f, int n, const double* r, double* xi);
/* The 768-node QSpline table of function f (for table-mode functions of a
 * RealModel or an uploaded table model) and its node separations. */
int opsec_model_table(opsec_ctx* ctx, int f, double* r768, double* xi768);

/* ---- pair quadrature and assembly ------------------------------------------ */
enum {
    OPSEC_QUAD_GM = 0,        /* reference: adaptive Genz-Malik 7/5 on [0,1]^6
                                 with the reference's argument order
                                 (epsrel -> absolute, epsabs -> relative)     */
    OPSEC_QUAD_GAUSS = 1,     /* tent-kernel product Gauss, `nodes` per axis,
                                 converged composite rule for near pairs      */
    OPSEC_QUAD_SURROGATE = 2  /* trained MLP for far pairs, GAUSS otherwise   */
};

typedef struct opsec_quad {
    int32_t method;           /* OPSEC_QUAD_*                                 */
    int32_t nodes;            /* GAUSS: nodes per axis (even, 2..32)          */
    int32_t near;             /* GAUSS: pairs with max|dd| < near use the near rule */
    int32_t maxeval;          /* GM: evaluation cap (reference 50000000)      */
    double epsrel;            /* GM: reference sig.epsrel (1e-5)              */
    double epsabs;            /* GM: reference sig.epsabs (1e-10)             */
    double spot_frac;         /* SURROGATE: fraction of far pairs spot-checked */
    double spot_tol;          /* SURROGATE: max |Q_mlp - Q_quad| / scale      */
    uint64_t spot_seed;       /* SURROGATE: spot-check selection seed         */
} opsec_quad;

/* Fill a quad record with the reference defaults for `method`. */
int opsec_quad_defaults(int method, opsec_quad* q);

/* Drop-in for ComputeSignalMatrixBlock (src/sig.cpp:50-70):
 *   s[i*lld + j] = sqrt(Nbar_a Nbar_b) Q_ab,  a = rows[i], b = cols[j],
 * for function f of the current model.  lld <= 0 means lld = ncols; 0 < lld <
 * ncols is an error.  cells/rows/cols/s are host buffers. */
int opsec_sig_block(opsec_ctx* ctx, const opsec_cell* cells, int ncells,
                    const int* rows, int nrows, const int* cols, int ncols, int f,
                    const opsec_survey* sv, const opsec_quad* q, double* s, int lld);

/* Single cell-pair average of xi over c1 x c2 (ComputeSignalC, src/sig.cpp:416-434). */
int opsec_pair_integral(opsec_ctx* ctx, const opsec_cell* c1, const opsec_cell* c2, int f,
                        const opsec_survey* sv, const opsec_quad* q, double* Q, double* err);

/* Device-resident survey: upload cells + geometry once. */
int opsec_cells_upload(opsec_ctx* ctx, const opsec_cell* cells, int ncells,
                       const opsec_survey* sv);
/* Pair-integral tables Q_f[key] for functions funcs[0..nf-1] on the device
 * survey.  Box: key = (|dx| Ny + |dy|) Nz + |dz| (src/sig.cpp:109-112), every
 * method.  Spherical: the arc cache of src/sig.cpp:150-190 for the block of all
 * rows and columns, key = (h_a narcs + h_b) Nphi + |dphi_a - dphi_b| with arcs
 * h = d_r Nmu + d_mu, narcs = Nr Nmu; adaptive (GM) rule only. */
int opsec_pair_tables(opsec_ctx* ctx, int nf, const int* funcs, const opsec_quad* q);
/* Copy table of function f to the host (nkeys doubles); nkeys via opsec_pair_nkeys. */
int opsec_pair_nkeys(opsec_ctx* ctx, int* nkeys);
int opsec_pair_table_get(opsec_ctx* ctx, int f, double* Q);
/* Statistics of the last opsec_pair_tables call: evaluated integrals, node
 * evaluations, surrogate-served pairs, spot checks, spot failures. */
int opsec_pair_stats(opsec_ctx* ctx, int64_t* integrals, int64_t* nodes,
                     int64_t* surrogate, int64_t* spot, int64_t* spot_fail);
/* Assemble S_f (Ncells x Ncells, row-major, leading dimension ld) into a
 * DEVICE buffer dS. */
int opsec_assemble_device(opsec_ctx* ctx, int f, double* dS, int ld);

/* ---- device buffers ------------------------------------------------------ */
int opsec_dev_alloc(opsec_ctx* ctx, size_t bytes, void** dptr);
int opsec_dev_free(opsec_ctx* ctx, void* dptr);
int opsec_h2d(opsec_ctx* ctx, void* dst, const void* src, size_t bytes);
int opsec_d2h(opsec_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- KL eigenmodes --------------------------------------------------------- */
/* Largest-eigenvalue modes of S_0 (the prior signal matrix on the device
 * survey) by block Lanczos over the device-resident S with full FP64
 * re-orthogonalisation.  evals: nev values in descending order; evecs: nev x
 * Ncells, mode-major (the modefile layout, src/klt.cpp:219-248).  *nconv =
 * number of modes whose residual |S v - lambda v| <= tol |lambda_max|.
 * block <= 0 picks the default block size; maxdim caps the Krylov dimension
 * (maxdim <= 0: 3 nev + 4 block, doubled up to twice, to at most Ncells,
 * while some of the nev modes have not converged; a retry's result replaces
 * the earlier one only if it converged more modes).  A Krylov space that
 * would span all but one block of the space is replaced by the dense
 * symmetric eigen-solver on S itself (same answer, all modes converged).
 * With the environment variable OPSEC_KLT_PROGRESS set, each Rayleigh-Ritz
 * step prints one progress line to stderr. */
int opsec_klt(opsec_ctx* ctx, int nev, double tol, int block, int maxdim,
              double* evals, double* evecs, int* nconv);
/* Same solver (and the same automatic Krylov dimension) on an arbitrary
 * symmetric host matrix A (n x n row-major); used for the reference's eigen
 * known answers (tests/testeig.cpp:78-121). */
int opsec_lanczos_dense(opsec_ctx* ctx, int n, const double* A, int nev, double tol,
                        int block, int maxdim, double* evals, double* evecs, int* nconv);

/* ---- projection ------------------------------------------------------------ */
enum {
    OPSEC_PREC_FP64 = 0,      /* FP64 MFMA, FP64 accumulation                  */
    OPSEC_PREC_FP32 = 1,      /* FP32 operands and accumulation                */
    OPSEC_PREC_BF16X3 = 2     /* split-BF16x3 MFMA, FP32 accumulation          */
};

/* Upload the mode matrix: B is Ncells x Nmodes column-major = nmodes
 * consecutive mode vectors (modefile order). */
int opsec_modes_upload(opsec_ctx* ctx, const double* B, int ncells, int nmodes);
/* P_f = B^T S_f B for each f in funcs (S_f generated on the fly from the pair
 * table, never stored).  Output: nf packed upper triangles, column-major
 * (column j holds rows 0..j), M(M+1)/2 doubles each -- the covfile layout of
 * src/comma-scalapack.cpp:257-265.  P_packed may be NULL to keep the results
 * on the device only (for opsec_fisher_device). */
int opsec_project(opsec_ctx* ctx, int nf, const int* funcs, int prec, double* P_packed);
/* Projection method (no counterpart in the reference, whose comma always
 * forms the dense product, src/comma-scalapack.cpp:228-266).
 * DENSE: stage 1 multiplies S_p (N x N, generated from the pair table) by B'.
 * TOEPLITZ (one cell per grid position, at most 64 grid planes along the
 * transformed axis): on a box S_p[a][b] depends on |dx|, |dy|, |dz| only, so a
 * length-2Nx DFT along x turns stage 1 into 2 Nx products of Ny Nz x Ny Nz
 * slab matrices with the transformed modes; on a spherical shell S_p[a][b]
 * depends on the two cells' (r, mu) arcs and |dphi| only (the arc cache of
 * src/sig.cpp:151-199), so a length-2Nphi DFT along phi gives 2 Nphi products
 * of (Nr Nmu) x (Nr Nmu) arc matrices.  The same result up to rounding, Nx/2
 * (Nphi/2) times fewer flops.  Applies to every precision: FP64 (FP64 DFTs
 * and MFMA), FP32 (FP32 MFMA, FP64 DFTs) and split-BF16x3. */
enum { OPSEC_PROJ_DENSE = 0, OPSEC_PROJ_TOEPLITZ = 1 };
int opsec_set_projection(opsec_ctx* ctx, int method);
/* Projection of an explicit host matrix S (N x N row-major) -- used to pin the
 * projection kernels against the oracle. */
int opsec_project_dense(opsec_ctx* ctx, int n, const double* S, int nmodes,
                        const double* B, int prec, double* P_packed);

/* ---- pixel values of a galaxy catalogue (the dot stage) ------------------ */
/* Galaxy record of a galfile: "4f" abn records of (x1, x2, x3, w) -- (x, y, z)
 * or (r, mu, phi) and a weight (src/Survey.h:13-19, src/BoxSurvey.cpp:105-133). */
typedef struct opsec_galaxy {
    float x1, x2, x3, w;
} opsec_galaxy;

/* Weighted galaxy counts per cell of the device survey (src/dot.cpp:105-128):
 * counts[a] (a < Ncells) = sum of the weights of the galaxies in cell a, each
 * added in catalogue order as the reference loop adds them, so the sums are
 * bit-identical to it; counts[Ncells] = the weight of galaxies in no cell.
 * A galaxy's grid position is Survey::GetGridIndex (src/Survey.cpp:93-101).
 * Unlike the reference, grid position 0 counts as inside (src/dot.cpp:127
 * tests e > 0) and a grid position that has no cell counts as outside (the
 * reference indexes counts[-1]).  counts holds Ncells + 1 doubles; with
 * accumulate != 0 the sums continue from its contents, so a catalogue passed
 * in pieces, in order, gives the same bits as one call. */
int opsec_count_galaxies(opsec_ctx* ctx, const opsec_galaxy* gals, int64_t ngals, int accumulate,
                         double* counts);
/* Pixel values y_i = sum_a B_ia x_a, x_a = (counts[a] - Nbar_a)/sqrt(Nbar_a),
 * with the uploaded modes (src/dot.cpp:130-153).  counts: Ncells doubles
 * (an extra outside bin is ignored); x (optional): Ncells; y: Nmodes. */
int opsec_pixel_values(opsec_ctx* ctx, const double* counts, double* x, double* y);

/* ---- basis cells and their expected counts (the basis stage) ------------ */
enum { OPSEC_SEL_UNIFORM = 0, OPSEC_SEL_RADIAL = 1 };

/* Selection function nbar of a survey.  UNIFORM: the constant nbar of a
 * BoxSurvey (src/BoxSurvey.cpp:59-75).  RADIAL: SphericalSurvey's radial
 * linear spline (rad_r ascending; zero outside [rad_r[0], rad_r[nrad-1]))
 * times the HEALPix completeness mask[pixel(acos(mu), phi)] of 12 nside^2
 * values in RING (nest = 0) or NESTED (nest = 1) order, nside = 0 meaning
 * completeness 1 (src/SphericalSurvey.cpp:51-82, :104-140). */
typedef struct opsec_selection {
    int32_t kind;          /* OPSEC_SEL_*                                   */
    int32_t nest;
    int64_t nside;
    double nbar;           /* UNIFORM                                       */
    const double* rad_r;   /* RADIAL: nrad nodes                            */
    const double* rad_n;
    int32_t nrad;
    int32_t reserved;
    const float* mask;     /* RADIAL: 12 nside^2 values, or NULL (nside 0)  */
} opsec_selection;

/* The cells of the survey grid with their volume Veff and expected count
 * Nbar = integral of nbar over the cell -- basis's FinalizeCell and cell loop
 * (src/basis.cpp:89-128, :170-190).  Cells come in grid order with the bounds
 * and volume of Survey::CreateEmptyCell (src/Survey.cpp:71-91).  Each Nbar is
 * the reference's adaptive Genz-Malik cubature (maxeval 5e6, epsrel and epsabs
 * in the reference's swapped argument order, so the absolute tolerance is
 * epsrel), one GPU wavefront per cell.  Spherical cells are integrated over
 * r^3 as in the reference and the result divided by 3 (the reference keeps
 * the factor: its spherical Nbar is 3x the expected count).  Cells with
 * Nbar <= 0 are dropped and the rest numbered a = 0, 1, ...  out holds
 * n1 n2 n3 records; *ncells receives the number kept; *nevals (optional) the
 * integrand evaluations. */
int opsec_basis(opsec_ctx* ctx, const opsec_survey* sv, const opsec_selection* sel, double epsrel,
                double epsabs, opsec_cell* out, int* ncells, int64_t* nevals);

/* HEALPix pixel indices of n directions (theta, phi) on the device (nest: 1
 * NESTED, 0 RING), as the completeness lookup of opsec_basis computes them. */
int opsec_healpix_pixels(opsec_ctx* ctx, int64_t nside, int nest, int64_t n, const double* theta,
                         const double* phi, int64_t* pix);

/* ---- Fisher matrix and band powers ----------------------------------------- */
enum { OPSEC_MIX_INVERSE = 0, OPSEC_MIX_INVERSE_SQRT = 1, OPSEC_MIX_IDENTITY = 2 };

typedef struct opsec_estimate {
    double* phat;    /* nparams                      */
    double* cov;     /* nparams x nparams row-major  */
    double* F;       /* nparams x nparams            */
    double* W;       /* nparams x nparams            */
    double* M;       /* nparams x nparams            */
    double* q;       /* nparams                      */
    double* f;       /* nparams                      */
} opsec_estimate;

/* estimate (src/estimate-mpi.cpp:480-654) from host inputs: lambda (M S/N
 * eigenvalues), D (nparams packed matrices, covfile layout), y (M pixel
 * values), prior (nparams p_n, 1 <= nparams <= 62).  prec selects FP64 or
 * FP32-accumulated traces.  Outputs are host buffers; any pointer in *out may
 * be NULL. */
int opsec_estimate_run(opsec_ctx* ctx, int nmodes, int nparams, const double* lambda,
                       const double* D, const double* y, const double* prior, int mixing,
                       int prec, opsec_estimate* out);
/* Same, with D = the device results of the last opsec_project call. */
int opsec_estimate_device(opsec_ctx* ctx, const double* lambda, const double* y,
                          const double* prior, int mixing, int prec, opsec_estimate* out);

/* The Gram step of opsec_estimate_run -- F (nparams x nparams row-major), f
 * and q (nparams each) -- from DEVICE inputs (d_lambda, d_y: nmodes; d_D:
 * nparams packed matrices in the covfile layout; d_prior: nparams) into
 * DEVICE outputs, allocating nothing: no hipMalloc, hipFree or memset, only
 * kernel launches on the context stream (asynchronous: opsec_ctx_sync waits).
 * Same kernels, same workgroup partition and the same fixed-order sums as
 * opsec_estimate_run, so the same bits.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops).  Its partial sums use a fixed scratch array of 16 MiB per
 * device, held until the process ends and shared by every call on that
 * device: calls that overlap on one device race.  1 <= nparams <= 62, any
 * nmodes >= 1. */
int opsec_fisher_gram_noalloc(opsec_ctx* ctx, int nmodes, int nparams, const double* d_lambda,
                              const double* d_D, const double* d_y, const double* d_prior, int prec,
                              double* d_F, double* d_f, double* d_q);

/* ---- learned pair-integral surrogate --------------------------------------- */
typedef struct opsec_mlp_train_opts {
    int32_t nsamples;         /* exact training integrals (2e5)               */
    int32_t epochs;
    int32_t batch;
    int32_t seed;
    double lr;                /* Adam step size                               */
    double near;              /* min separation in cell widths (2)            */
} opsec_mlp_train_opts;

int opsec_mlp_defaults(opsec_mlp_train_opts* o);
/* Generate exact training integrals on the device survey with quadrature q,
 * then train the 8 -> 128^4 -> 1 SiLU network on the GPU, starting from the
 * initial weights of o->seed (every reduction runs in a fixed order, so the
 * same options on the same survey give the same network bit for bit).
 * final_loss gets the last epoch's mean squared error (normalised units). */
int opsec_mlp_train(opsec_ctx* ctx, const opsec_mlp_train_opts* o, const opsec_quad* q,
                    double* final_loss);
/* Raw network access for gradient checks: nparams_total weights as FP32. */
int opsec_mlp_nweights(opsec_ctx* ctx, int* n);
int opsec_mlp_get_weights(opsec_ctx* ctx, float* w);
int opsec_mlp_set_weights(opsec_ctx* ctx, const float* w);
/* Forward on n feature rows (n x 8, FP32) -> out[n]. */
int opsec_mlp_forward(opsec_ctx* ctx, int n, const float* x, float* out);
/* Loss 0.5 mean (f(x)-t)^2 and its gradient w.r.t. all weights (FP32). */
int opsec_mlp_loss_grad(opsec_ctx* ctx, int n, const float* x, const float* t, double* loss,
                        float* grad);

#ifdef __cplusplus
}
#endif
#endif /* OPSEC_HIP_H */
