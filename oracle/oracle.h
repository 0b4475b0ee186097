/* oracle.h -- FP64 CPU restatement of OPSEC's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the CPU baseline; the product never links it.
 *
 * Parity pinning of each function (what checks it besides the GPU path):
 *   orc_fmm_spline / orc_cubic_eval    oracle/_ref (reference CubicSpline)
 *   orc_xilm                           oracle/_ref (reference ComputeXiLM)
 *   orc_qspline_*                      oracle/_ref (reference QSpline via RealModel)
 *   orc_band_xi                        numerical k-integral (scipy) + r->0 limit
 *   orc_box_r                          reference known answers (tests/testBoxSurvey.cpp:22-26)
 *   orc_gm_integrate                   oracle/_ref ComputeSignalC; xi=1 known answer
 *   orc_tent_gauss                     brute-force 6-D Gauss on tiny cells; Gaussian
 *                                      closed form (tests/testsig.cpp:41-66)
 *   orc_gauss_jacobi10                 scipy.special.roots_jacobi
 *   orc_sig_block                      oracle/_ref ComputeSignalMatrixBlock
 *   orc_sph_r                          law-of-cosines known answers (same ray, same radius,
 *                                      antipodes); the reference's SphericalSurvey needs
 *                                      cfitsio and is not built
 *   orc_gm_integrate (spherical)       xi=1 known answer; orc_brute6_sph on smooth xi
 *   orc_sig_block_sph                  xi=1 (S = sqrt(Nbar Nbar)); every element against its
 *                                      own pair integral (phi-rotation invariance)
 *   orc_project                        brute force triple loop + B^T S B = diag(lambda)
 *   orc_estimate                       hand-derived diagonal cases, W row sums,
 *                                      inverse/sqrtm known answers (tests/testMyMatrix.cpp:67-104)
 *   orc_sym_eigen                      J_x spectrum (tests/testeig.cpp:78-121), numpy eigh
 *   orc_mlp_*                          finite differences of orc_mlp_forward
 *   orc_grid_index                     oracle/_ref (reference Survey::GetGridIndex) on
 *                                      interior, boundary and outside points
 *   orc_count_galaxies                 known answers: order-dependent rounding
 *                                      (2^60 + 1 - 2^60), grid cell 0, missing cells,
 *                                      outside points; integer weights against an
 *                                      independent numpy histogram
 *   orc_pixel_values                   x against numpy IEEE arithmetic; y against
 *                                      numpy's BLAS product B @ x
 *   orc_cell_integral                  oracle/_ref (the reference's adapt_integrate_v in
 *                                      3-D with its LinearSpline, src/basis.cpp:89-128);
 *                                      uniform nbar = nbar V; polynomial radial profiles
 *                                      in closed form
 *   orc_ang2pix_ring / _nest,          pix2ang(ang2pix) identities, NESTED hierarchy
 *   orc_pix2ang_ring / _nest           (pixel >> 2 at nside / 2), RING and NESTED give the
 *                                      same pixel centre, equal pixel areas, the known
 *                                      base-pixel centres
 *   orc_basis                          cell bounds and volumes against Survey::CreateEmptyCell
 *                                      formulas; box Nbar = nbar V
 */
#ifndef OPSEC_ORACLE_H
#define OPSEC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 80-byte cell record (src/Cell.h:74-96) */
typedef struct orc_cell {
    int32_t a, d1, d2, d3;
    double min1, max1, min2, max2, min3, max3, Veff, Nbar;
} orc_cell;

/* ---- splines and xi ---- */
void orc_fmm_spline(int n, const double* x, const double* y, double* b, double* c, double* d);
void orc_cubic_eval(int n, const double* x, const double* y, const double* b, const double* c,
                    const double* d, int m, const double* xe, double* ye);
/* Simpson xi_0^2 (src/Model.cpp:84-123).  kind 0: P = cubic spline (n, x, y);
 * kind 1: P = 1 on [kmin, kmax). */
void orc_xilm(int kind, int n, const double* px, const double* py, int nr, const double* r,
              double* xi, int Nk, double kmin, double kmax, double kdamp);
/* QSpline grid (src/QSpline.cpp:6-17): 768 r nodes */
void orc_qspline_grid(double* r768);
double orc_qspline_eval(const double* r768, const double* y768, double r);
double orc_band_xi(double klo, double khi, double r);
/* Prior band power (src/RealModel.cpp:85-96) */
double orc_band_param(int n, const double* px, const double* py, double kmin, double kmax);

/* xi function handle for quadrature */
typedef struct orc_xi {
    int kind;            /* 0 table (r768,y768), 1 band closed, 2 exp, 3 gauss, 4 dgauss */
    const double* r768;
    const double* y768;
    double a, b;
} orc_xi;
double orc_xi_eval(const orc_xi* f, double r);

/* ---- geometry ---- */
double orc_box_r(const double* p1, const double* p2, double L);
/* spherical separation, points (r, mu, phi) (src/SphericalSurvey.cpp:19-23) */
double orc_sph_r(const double* p1, const double* p2);

/* ---- pair integrals ---- */
/* Genz-Malik 7/5 adaptive cubature of the cell-pair mean with the reference's
 * argument order (src/sig.cpp:416-434, src/cubature.c:885-1042).  coordsys 1
 * box (L period), 2 spherical.  Returns Q; *err, *neval. */
double orc_gm_integrate(const orc_xi* f, int coordsys, double L, const orc_cell* c1, const orc_cell* c2,
                        double epsrel, double epsabs, unsigned maxeval, double* err, int64_t* neval);
/* tent-kernel Gauss: n nodes per axis (n/2 Gauss nodes for weight (1-t) per half) */
double orc_tent_gauss(const orc_xi* f, double L, const orc_cell* c1, const orc_cell* c2, int n);
/* composite near rule: panels x q Gauss-Legendre nodes per half axis */
double orc_tent_composite(const orc_xi* f, double L, const orc_cell* c1, const orc_cell* c2, int panels, int q);
/* brute-force 6-D product Gauss-Legendre (n per dimension) of the pair mean */
double orc_brute6(const orc_xi* f, double L, const orc_cell* c1, const orc_cell* c2, int n);
void orc_gauss_legendre(int n, double* x, double* w);
void orc_gauss_jacobi10(int n, double* t, double* w);

/* Signal-matrix block with the translation cache (src/sig.cpp:82-123).
 * method 0: GM (epsrel, epsabs, maxeval), 1: GAUSS (nodes, near). */
/* spherical block with the arc cache (src/sig.cpp:125-213); Genz-Malik pairs */
int orc_sig_block_sph(const orc_xi* f, const orc_cell* cells, int ncells, int Nr, int Nmu, int Nphi,
                      const int* rows, int nrows, const int* cols, int ncols, double* s, int lld, double epsrel,
                      double epsabs, unsigned maxeval);
/* tensor Gauss-Legendre (n per dimension) of a spherical cell-pair mean */
double orc_brute6_sph(const orc_xi* f, const orc_cell* c1, const orc_cell* c2, int n);
int orc_sig_block(const orc_xi* f, const orc_cell* cells, int ncells, int Nx, int Ny, int Nz, double L,
                  const int* rows, int nrows, const int* cols, int ncols, double* s, int lld, int method,
                  double epsrel, double epsabs, unsigned maxeval, int nodes, int nearc);

/* ---- projection and estimate ---- */
/* P = B^T S B, S n x n row-major, B n x m column-major; P packed upper
 * column-major (covfile layout). */
void orc_project(int n, const double* S, int m, const double* B, double* Ppacked);
/* C = S[rows,:] B for a row block (R x n row-major S), C column-major R x m */
void orc_gemm_rows(int R, int n, const double* S, int m, const double* B, double* C);
/* estimate-mpi math (src/estimate-mpi.cpp:480-654).  D: nparams packed
 * matrices. Outputs row-major nparams^2 (cov, F, W, M), vectors phat, q, f. */
int orc_estimate(int M, int np, const double* lambda, const double* D, const double* y, const double* prior,
                 int mixing, double* phat, double* cov, double* F, double* W, double* Mm, double* q, double* f);
int orc_inverse(int n, const double* A, double* B);
void orc_sqrtm(int n, const double* A, double* B);
/* dense symmetric eigen (cyclic Jacobi): eigenvalues descending, vectors as
 * rows of V (n x n, row k = eigenvector k) */
void orc_sym_eigen(int n, const double* A, double* evals, double* V);

/* ---- MLP 8 -> 128^4 -> 1, SiLU (FP64) ---- */
int orc_mlp_nweights(void);
void orc_mlp_forward(const double* w, int n, const double* x, double* out);
double orc_mlp_loss_grad(const double* w, int n, const double* x, const double* t, double* grad);

/* ---- basis cells (basis) ---- */
/* Selection function: kind 0 uniform nbar; kind 1 radial linear spline (rr, rn;
 * zero outside [rr[0], rr[nrad-1])) times a HEALPix completeness mask
 * (nside 0: none) at (acos(mu), phi) (src/SphericalSurvey.cpp:65-82). */
typedef struct orc_selection {
    int kind, nest;
    int64_t nside;
    double nbar;
    const double* rr;
    const double* rn;
    int nrad;
    const float* mask;
} orc_selection;
/* value at u = (x, y, z) (coordsys 1) or (r^3, mu, phi) (coordsys 2) */
double orc_selection_eval(const orc_selection* s, int coordsys, const double* u);
/* FinalizeCell (src/basis.cpp:89-128): 3-D Genz-Malik adaptive cubature of the
 * selection over the cell (r^3 for spherical cells), the reference's swapped
 * tolerances; returns the reference's integral (spherical: 3x the count). */
double orc_cell_integral(int coordsys, const orc_selection* s, const orc_cell* c, double epsrel, double epsabs,
                         unsigned maxeval, int64_t* neval);
/* All grid cells (src/basis.cpp:170-190) with Veff and Nbar (spherical: the
 * integral / 3); cells with Nbar <= 0 dropped.  geom = {min1, max1, min2,
 * max2, min3, max3}, n = {N1, N2, N3}; out holds N1 N2 N3 records, nevals
 * (optional) the integrand evaluations of each kept cell; returns the number
 * kept. */
int orc_basis(int coordsys, const double* geom, const int* n, const orc_selection* s, double epsrel, double epsabs,
              orc_cell* out, int64_t* nevals);
/* HEALPix (Gorski et al. 2005): pixel of a direction and centre of a pixel */
int64_t orc_ang2pix_ring(int64_t nside, double theta, double phi);
int64_t orc_ang2pix_nest(int64_t nside, double theta, double phi);
void orc_pix2ang_ring(int64_t nside, int64_t pix, double* theta, double* phi);
void orc_pix2ang_nest(int64_t nside, int64_t pix, double* theta, double* phi);

/* ---- pixel values (dot) ---- */
/* geom = {min1, max1, min2, max2, min3, max3}, n = {N1, N2, N3}; -1 outside */
int64_t orc_grid_index(const double* geom, const int* n, double x1, double x2, double x3);
/* gals: ngals x (x1, x2, x3, w) floats; counts (ncells + 1) is accumulated into */
void orc_count_galaxies(const orc_cell* cells, int ncells, const double* geom, const int* n,
                        const float* gals, int64_t ngals, double* counts);
void orc_pixel_values(const orc_cell* cells, int ncells, int nmodes, const double* B,
                      const double* counts, double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
