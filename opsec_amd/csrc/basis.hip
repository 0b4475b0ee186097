// Basis cells and their expected galaxy counts -- the basis stage
// (src/basis.cpp:89-190).
//
// Every cell of the survey grid gets Nbar = integral of the selection function
// over the cell.  The reference integrates each cell with the adaptive
// Genz-Malik 7/5 rule in 3 dimensions (33 nodes per region, max-error heap,
// src/cubature.c:885-998); here one 64-lane wavefront runs one cell at a time
// from a work counter: the lanes evaluate a region's 33 nodes, lane 0 forms the
// rule in the reference's node order and keeps the region heap.  Regions live
// in a per-wavefront pool in HBM; cells that outgrow it are run again on a few
// wavefronts with pools that hold every region the evaluation cap allows.
//
// Spherical cells are integrated over (r^3, mu, phi) with r = cbrt(u) as the
// reference does (src/basis.cpp:67-79, 104-108) and the result divided by 3,
// since d(r^3) = 3 r^2 dr; the reference keeps the factor.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "healpix.h"
#include "xi_dev.h"

namespace opsec {
namespace {

constexpr int B3_NPTS = 33;   // 1 + 2*3 + 2*3 + 2*3*2 + 2^3
constexpr double B3_L2 = 0.35856858280031809199;   // sqrt(9/70)
constexpr double B3_L4 = 0.94868329805051379960;   // sqrt(9/10)
constexpr double B3_L5 = 0.68824720161168529772;   // sqrt(9/19)

struct SelDev {
    int kind, nest;
    long long nside;
    double nbar;
    const double* rr;
    const double* rn;
    int nrad;
    const float* mask;
};

struct CellJob {
    double lo[3], hi[3];   // integration bounds: (x, y, z) or (r^3, mu, phi)
};

struct B3Params {
    double absTol, relTol;
    unsigned maxEval;
    int rmax;
    int coordsys;
};

struct B3Slot {
    double* c;       // 3 x R centres
    double* hw;      // 3 x R half-widths
    double* val;
    double* err;
    int* split;
    int* heap;
    int* list;
};

__device__ __forceinline__ B3Slot b3_slot(double* dpool, int* ipool, int slot, int R) {
    B3Slot s;
    double* d = dpool + (size_t)slot * 8 * R;
    int* q = ipool + (size_t)slot * 3 * R;
    s.c = d;
    s.hw = d + 3 * R;
    s.val = d + 6 * R;
    s.err = d + 7 * R;
    s.split = q;
    s.heap = q + R;
    s.list = q + 2 * R;
    return s;
}

// Selection function at an integration point (src/BoxSurvey.cpp:59-75,
// src/SphericalSurvey.cpp:65-82 with the volume-uniform radius of
// src/basis.cpp:67-79).
__device__ __forceinline__ double sel_eval(const SelDev& s, int coordsys, const double* u) {
    if (s.kind == OPSEC_SEL_UNIFORM) return s.nbar;
    const double r = cbrt(u[0]), mu = u[1], phi = u[2];
    (void)coordsys;
    const double rad = linear_spline(r, s.nrad, s.rr, s.rn);
    double ang = 1.0;
    if (s.nside > 0) {
        const double theta = acos(mu);
        const int64_t ip = s.nest ? hp_ang2pix_nest(s.nside, theta, phi) : hp_ang2pix_ring(s.nside, theta, phi);
        ang = (double)s.mask[ip];
    }
    return ang * rad;
}

// Node p of the 3-D rule in the reference's order (src/cubature.c:347-443):
// centre; per axis -l2, +l2, -l4, +l4; per axis pair (i, j) the four
// (-l4, -l4), (+, -), (+, +), (-, +); the 8 corners at l5 in Gray-code order
// starting from all '+'.
__device__ __forceinline__ void b3_node(int p, const double* c, const double* hw, double* u) {
    for (int d = 0; d < 3; d++) u[d] = c[d];
    if (p == 0) return;
    if (p <= 12) {
        const int q = p - 1, i = q >> 2, s = q & 3;
        const double r = hw[i] * (s < 2 ? B3_L2 : B3_L4);
        u[i] = (s & 1) ? c[i] + r : c[i] - r;
        return;
    }
    if (p <= 24) {
        const int q = p - 13, pair = q >> 2, s = q & 3;
        const int i = pair == 2 ? 1 : 0, j = pair == 0 ? 1 : 2;
        const double ri = hw[i] * B3_L4, rj = hw[j] * B3_L4;
        const bool ip = (s == 1 || s == 2), jp = (s >= 2);
        u[i] = ip ? c[i] + ri : c[i] - ri;
        u[j] = jp ? c[j] + rj : c[j] - rj;
        return;
    }
    const int g = p - 25, signs = g ^ (g >> 1);
    for (int d = 0; d < 3; d++) {
        const double r = hw[d] * B3_L5;
        u[d] = ((signs >> d) & 1) ? c[d] - r : c[d] + r;
    }
}

__device__ void b3_eval_region(const SelDev& sel, const B3Params& P, B3Slot& S, int R, int ri, double* vals) {
    const int lane = threadIdx.x;
    double c[3], hw[3];
    for (int d = 0; d < 3; d++) {
        c[d] = S.c[d * R + ri];
        hw[d] = S.hw[d * R + ri];
    }
    if (lane < B3_NPTS) {
        double u[3];
        b3_node(lane, c, hw, u);
        vals[lane] = sel_eval(sel, P.coordsys, u);
    }
    __syncthreads();
    if (lane == 0) {
        const int n = 3;
        const double ratio = (B3_L2 * B3_L2) / (B3_L4 * B3_L4);
        const double w1 = (double)(12824 - 9120 * n + 400 * n * n) / 19683.0;
        const double w2 = 980. / 6561.;
        const double w3 = (double)(1820 - 400 * n) / 19683.0;
        const double w4 = 200. / 19683.;
        const double w5 = 6859.0 / 19683.0 / 8.0;
        const double e1 = (double)(729 - 950 * n + 50 * n * n) / 729.0;
        const double e2 = 245. / 486.;
        const double e3 = (double)(265 - 100 * n) / 1458.0;
        const double e4 = 25. / 729.;
        double val0 = vals[0], sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0, maxdiff = 0;
        int split = 0;
        for (int k = 0; k < 3; k++) {
            const double v0 = vals[1 + 4 * k], v1 = vals[2 + 4 * k], v2 = vals[3 + 4 * k], v3 = vals[4 + 4 * k];
            sum2 += v0 + v1;
            sum3 += v2 + v3;
            const double diff = 0.0 + fabs(v0 + v1 - 2 * val0 - ratio * (v2 + v3 - 2 * val0));
            if (diff > maxdiff) {
                maxdiff = diff;
                split = k;
            }
        }
        for (int k = 0; k < 12; k++) sum4 += vals[13 + k];
        for (int k = 0; k < 8; k++) sum5 += vals[25 + k];
        double vol = 1;
        for (int d = 0; d < 3; d++) vol *= 2 * hw[d];
        const double result = vol * (w1 * val0 + w2 * sum2 + w3 * sum3 + w4 * sum4 + w5 * sum5);
        const double res5th = vol * (e1 * val0 + e2 * sum2 + e3 * sum3 + e4 * sum4);
        S.val[ri] = result;
        S.err[ri] = fabs(res5th - result);
        S.split[ri] = split;
    }
    __syncthreads();
}

// Binary max-heap on region errors, lane 0 only.
__device__ void b3_push(B3Slot& S, int& n, int ri, double& tv, double& te) {
    tv += S.val[ri];
    te += S.err[ri];
    int at = n++;
    const double key = S.err[ri];
    while (at) {
        const int parent = (at - 1) / 2;
        if (key <= S.err[S.heap[parent]]) break;
        S.heap[at] = S.heap[parent];
        at = parent;
    }
    S.heap[at] = ri;
}

__device__ int b3_pop(B3Slot& S, int& n, double& tv, double& te) {
    const int top = S.heap[0];
    int i = 0, child;
    S.heap[0] = S.heap[--n];
    while ((child = i * 2 + 1) < n) {
        int largest = (S.err[S.heap[child]] <= S.err[S.heap[i]]) ? i : child;
        if (++child < n && S.err[S.heap[largest]] < S.err[S.heap[child]]) largest = child;
        if (largest == i) break;
        const int t = S.heap[i];
        S.heap[i] = S.heap[largest];
        S.heap[largest] = t;
        i = largest;
    }
    tv -= S.val[top];
    te -= S.err[top];
    return top;
}

__device__ __forceinline__ bool b3_converged(double val, double err, const B3Params& P) {
    const double rel = (val == 0.0) ? HUGE_VAL : fabs(err / val);
    return err <= P.absTol || rel <= P.relTol;
}

// One wavefront per cell; jobs taken from a counter.  out[jb] = integral,
// fail[jb] = 1 when the region pool ran out (the host reruns those cells).
__global__ __launch_bounds__(64) void basis_gm_kernel(SelDev sel, const CellJob* jobs, int njobs, int* counter,
                                                      double* dpool, int* ipool, B3Params P, double* out, int* fail,
                                                      unsigned long long* nevals) {
    __shared__ double vals[B3_NPTS];
    __shared__ int sh_job, sh_nR, sh_done;
    const int lane = threadIdx.x;
    const int R = P.rmax;
    B3Slot S = b3_slot(dpool, ipool, blockIdx.x, R);
    for (;;) {
        if (lane == 0) sh_job = atomicAdd(counter, 1);
        __syncthreads();
        const int jb = sh_job;
        __syncthreads();
        if (jb >= njobs) break;
        const CellJob J = jobs[jb];
        if (lane < 3) {
            S.c[lane * R] = 0.5 * (J.lo[lane] + J.hi[lane]);   // src/cubature.c:165-177
            S.hw[lane * R] = 0.5 * (J.hi[lane] - J.lo[lane]);
        }
        __syncthreads();
        b3_eval_region(sel, P, S, R, 0, vals);
        int nheap = 0, nregions = 1;
        double tv = 0, te = 0;
        unsigned numEval = B3_NPTS;
        if (lane == 0) b3_push(S, nheap, 0, tv, te);
        int status = 0;
        for (;;) {
            if (lane == 0) {
                int done = 0, nR = 0;
                if (!(numEval < P.maxEval || !P.maxEval)) done = 1;
                else if (b3_converged(tv, te, P)) done = 1;
                else {
                    const double eev = tv;
                    double eee = te;
                    do {
                        if (nregions + 1 > R) {
                            done = 2;
                            break;
                        }
                        const int ri = b3_pop(S, nheap, tv, te);
                        eee -= S.err[ri];
                        const int d = S.split[ri];
                        const int r2 = nregions++;
                        for (int k = 0; k < 3; k++) {
                            S.c[k * R + r2] = S.c[k * R + ri];
                            S.hw[k * R + r2] = S.hw[k * R + ri];
                        }
                        const double h = S.hw[d * R + ri] * 0.5;
                        S.hw[d * R + ri] = h;
                        S.hw[d * R + r2] = h;
                        S.c[d * R + ri] -= h;
                        S.c[d * R + r2] += h;
                        numEval += B3_NPTS * 2;
                        S.list[nR++] = ri;
                        S.list[nR++] = r2;
                        if (b3_converged(eev, eee, P)) break;
                    } while (nheap > 0 && (numEval < P.maxEval || !P.maxEval));
                }
                sh_nR = nR;
                sh_done = done;
            }
            __syncthreads();
            const int nR = sh_nR, done = sh_done;
            __syncthreads();
            if (done) {
                status = done;
                break;
            }
            __threadfence_block();
            for (int t = 0; t < nR; t++) b3_eval_region(sel, P, S, R, S.list[t], vals);
            if (lane == 0)
                for (int t = 0; t < nR; t++) b3_push(S, nheap, S.list[t], tv, te);
        }
        if (lane == 0) {
            if (status == 2) {
                fail[jb] = 1;
            } else {
                double v = 0;
                for (int i = 0; i < nheap; i++) v += S.val[S.heap[i]];
                out[jb] = v;
                fail[jb] = 0;
                atomicAdd(nevals, (unsigned long long)numEval);
            }
        }
        __syncthreads();
    }
}

__global__ void healpix_kernel(long long nside, int nest, long long n, const double* theta, const double* phi,
                               long long* pix) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pix[i] = nest ? hp_ang2pix_nest(nside, theta[i], phi[i]) : hp_ang2pix_ring(nside, theta[i], phi[i]);
}

struct GrowBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~GrowBuf() { dev_free(&p); }
};

inline double cube(double x) { return x * x * x; }

}  // namespace
}  // namespace opsec

using namespace opsec;

extern "C" {

int opsec_basis(opsec_ctx* ctx, const opsec_survey* sv, const opsec_selection* sel, double epsrel, double epsabs,
                opsec_cell* out, int* ncells, int64_t* nevals) {
    if (!ctx) return OPSEC_EINVAL;
    SurveyDev s;
    int rc = survey_from_abi(ctx, sv, &s);
    if (rc) return rc;
    OPSEC_CHECK_ARG(ctx, sel && out && ncells, "NULL selection or output");
    OPSEC_CHECK_ARG(ctx, sel->kind == OPSEC_SEL_UNIFORM || sel->kind == OPSEC_SEL_RADIAL, "unknown selection kind %d",
                    sel->kind);
    const bool sph = s.coordsys == OPSEC_COORD_SPHERICAL;
    if (sel->kind == OPSEC_SEL_RADIAL) {
        OPSEC_CHECK_ARG(ctx, sph, "a radial selection needs a spherical survey");
        OPSEC_CHECK_ARG(ctx, sel->nrad >= 2 && sel->rad_r && sel->rad_n, "radial selection needs >= 2 nodes");
        for (int i = 0; i + 1 < sel->nrad; i++)
            OPSEC_CHECK_ARG(ctx, sel->rad_r[i] < sel->rad_r[i + 1], "radial nodes must increase");
        OPSEC_CHECK_ARG(ctx, sel->nside >= 0 && sel->nside <= (1 << 29), "HEALPix nside %lld out of range",
                        (long long)sel->nside);
        if (sel->nside > 0) {
            OPSEC_CHECK_ARG(ctx, sel->mask, "HEALPix mask is NULL");
            if (sel->nest)
                OPSEC_CHECK_ARG(ctx, (sel->nside & (sel->nside - 1)) == 0, "NESTED ordering needs nside = 2^k");
        }
    }
    if (sph)
        OPSEC_CHECK_ARG(ctx, s.min1 >= 0 && s.min2 >= -1 && s.max2 <= 1, "spherical bounds need r >= 0, -1 <= mu <= 1");
    const long long ngrid = (long long)s.n1 * s.n2 * s.n3;
    OPSEC_CHECK_ARG(ctx, ngrid <= INT32_MAX, "survey grid too large");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));

    // the grid cells (src/Survey.cpp:71-91) and their integration bounds
    std::vector<opsec_cell> cells((size_t)ngrid);
    std::vector<CellJob> jobs((size_t)ngrid);
    size_t a = 0;
    for (int d1 = 0; d1 < s.n1; d1++)
        for (int d2 = 0; d2 < s.n2; d2++)
            for (int d3 = 0; d3 < s.n3; d3++, a++) {
                opsec_cell& c = cells[a];
                c.a = -1;
                c.d1 = d1;
                c.d2 = d2;
                c.d3 = d3;
                c.min1 = s.min1 + d1 * (s.max1 - s.min1) / s.n1;
                c.max1 = s.min1 + (d1 + 1) * (s.max1 - s.min1) / s.n1;
                c.min2 = s.min2 + d2 * (s.max2 - s.min2) / s.n2;
                c.max2 = s.min2 + (d2 + 1) * (s.max2 - s.min2) / s.n2;
                c.min3 = s.min3 + d3 * (s.max3 - s.min3) / s.n3;
                c.max3 = s.min3 + (d3 + 1) * (s.max3 - s.min3) / s.n3;
                // the volume FinalizeCell stores (src/basis.cpp:111-119)
                c.Veff = sph ? (cube(c.max1) - cube(c.min1)) / 3 * (c.max2 - c.min2) * (c.max3 - c.min3)
                             : (c.max1 - c.min1) * (c.max2 - c.min2) * (c.max3 - c.min3);
                c.Nbar = 0;
                CellJob& J = jobs[a];
                J.lo[0] = sph ? cube(c.min1) : c.min1;
                J.hi[0] = sph ? cube(c.max1) : c.max1;
                J.lo[1] = c.min2;
                J.hi[1] = c.max2;
                J.lo[2] = c.min3;
                J.hi[2] = c.max3;
            }

    // selection tables on the device
    GrowBuf srad, smask, sjob, sout, sfail, sdp, sip, sctr;
    SelDev sd{};
    sd.kind = sel->kind;
    sd.nest = sel->nest;
    sd.nside = sel->kind == OPSEC_SEL_RADIAL ? sel->nside : 0;
    sd.nbar = sel->nbar;
    hipStream_t st = ctx->stream;
    if (sel->kind == OPSEC_SEL_RADIAL) {
        if ((rc = dev_realloc(ctx, &srad.p, &srad.cap, (size_t)sel->nrad * 16))) return rc;
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(srad.p, sel->rad_r, (size_t)sel->nrad * 8, hipMemcpyHostToDevice, st));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync((double*)srad.p + sel->nrad, sel->rad_n, (size_t)sel->nrad * 8,
                                          hipMemcpyHostToDevice, st));
        sd.rr = (const double*)srad.p;
        sd.rn = (const double*)srad.p + sel->nrad;
        sd.nrad = sel->nrad;
        if (sd.nside > 0) {
            const size_t npix = (size_t)12 * sd.nside * sd.nside;
            if ((rc = dev_realloc(ctx, &smask.p, &smask.cap, npix * 4))) return rc;
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(smask.p, sel->mask, npix * 4, hipMemcpyHostToDevice, st));
            sd.mask = (const float*)smask.p;
        }
    }
    std::vector<double> N((size_t)ngrid, 0.0);
    std::vector<int> idx((size_t)ngrid);
    for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
    B3Params P;
    P.absTol = epsrel;   // the reference's argument order, src/basis.cpp:97-107
    P.relTol = epsabs;
    P.maxEval = 5000000;   // src/basis.cpp:95
    P.coordsys = s.coordsys;
    unsigned long long ev_total = 0;
    if ((rc = dev_realloc(ctx, &sctr.p, &sctr.cap, 64))) return rc;
    // pass 0: 1024-region pools on many wavefronts; pass 1: the cells that ran
    // out, with pools for every region the evaluation cap allows
    for (int pass = 0; pass < 2 && !idx.empty(); pass++) {
        const int nj = (int)idx.size();
        std::vector<CellJob> pj(nj);
        for (int i = 0; i < nj; i++) pj[i] = jobs[idx[i]];
        P.rmax = pass == 0 ? 1024 : (int)(P.maxEval / B3_NPTS + 2);
        int nslots = pass == 0 ? std::min(nj, std::max(1, ctx->num_cus) * 8) : std::min(nj, 64);
        if ((rc = dev_realloc(ctx, &sjob.p, &sjob.cap, sizeof(CellJob) * nj))) return rc;
        if ((rc = dev_realloc(ctx, &sout.p, &sout.cap, 8 * (size_t)nj))) return rc;
        if ((rc = dev_realloc(ctx, &sfail.p, &sfail.cap, 4 * (size_t)nj))) return rc;
        if ((rc = dev_realloc(ctx, &sdp.p, &sdp.cap, 8 * 8 * (size_t)P.rmax * nslots))) return rc;
        if ((rc = dev_realloc(ctx, &sip.p, &sip.cap, 4 * 3 * (size_t)P.rmax * nslots))) return rc;
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(sjob.p, pj.data(), sizeof(CellJob) * nj, hipMemcpyHostToDevice, st));
        OPSEC_HIP_TRY(ctx, hipMemsetAsync(sctr.p, 0, 64, st));
        hipLaunchKernelGGL(basis_gm_kernel, dim3(nslots), dim3(64), 0, st, sd, (const CellJob*)sjob.p, nj,
                           (int*)sctr.p, (double*)sdp.p, (int*)sip.p, P, (double*)sout.p, (int*)sfail.p,
                           (unsigned long long*)((char*)sctr.p + 8));
        OPSEC_HIP_TRY(ctx, hipGetLastError());
        std::vector<double> o(nj);
        std::vector<int> f(nj);
        unsigned long long ev = 0;
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(o.data(), sout.p, 8 * (size_t)nj, hipMemcpyDeviceToHost, st));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(f.data(), sfail.p, 4 * (size_t)nj, hipMemcpyDeviceToHost, st));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&ev, (char*)sctr.p + 8, 8, hipMemcpyDeviceToHost, st));
        OPSEC_HIP_TRY(ctx, hipStreamSynchronize(st));
        ev_total += ev;
        std::vector<int> again;
        for (int i = 0; i < nj; i++) {
            if (f[i]) again.push_back(idx[i]);
            else N[idx[i]] = o[i];
        }
        if (!again.empty() && pass == 1)
            return set_error(ctx, OPSEC_ECONV, "adaptive cubature exceeded %d regions", P.rmax);
        idx.swap(again);
    }
    // keep cells with a positive count (src/basis.cpp:121-127), numbered in grid order
    int kept = 0;
    for (size_t i = 0; i < cells.size(); i++) {
        const double Nb = sph ? N[i] / 3 : N[i];
        if (Nb > 0) {
            opsec_cell c = cells[i];
            c.a = kept;
            c.Nbar = Nb;
            out[kept++] = c;
        }
    }
    *ncells = kept;
    if (nevals) *nevals = (int64_t)ev_total;
    return OPSEC_OK;
}

int opsec_healpix_pixels(opsec_ctx* ctx, int64_t nside, int nest, int64_t n, const double* theta, const double* phi,
                         int64_t* pix) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, nside > 0 && nside <= (1 << 29), "nside out of range");
    OPSEC_CHECK_ARG(ctx, !nest || (nside & (nside - 1)) == 0, "NESTED ordering needs nside = 2^k");
    OPSEC_CHECK_ARG(ctx, n >= 0 && (n == 0 || (theta && phi && pix)), "NULL arrays");
    if (n == 0) return OPSEC_OK;
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    GrowBuf sb;
    int rc;
    if ((rc = dev_realloc(ctx, &sb.p, &sb.cap, (size_t)n * 24))) return rc;
    double* dt = (double*)sb.p;
    double* dp = dt + n;
    long long* dq = (long long*)(dp + n);
    hipStream_t st = ctx->stream;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dt, theta, (size_t)n * 8, hipMemcpyHostToDevice, st));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dp, phi, (size_t)n * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(healpix_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (long long)nside, nest,
                       (long long)n, dt, dp, dq);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(pix, dq, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(st));
    return OPSEC_OK;
}

}  // extern "C"
