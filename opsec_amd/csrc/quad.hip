// Pair quadrature and signal-matrix assembly.
//
// Pair integrals Q_ab = mean of xi(|x - y|) over x in cell a, y in cell b are
// cached by the reference's translation key (src/sig.cpp:94-123):
//   box:   key = (|dx_a-dx_b| Ny + |dy_a-dy_b|) Nz + |dz_a-dz_b|
// and each distinct key is integrated once, on the first (row, col) pair that
// carries it in row-major order, with the lower cell index first
// (src/sig.cpp:114-118).  The first pairs are found on the device with one
// atomicMin per (row, col).
//
// Methods:
//  GM     one wavefront per integral: the Genz-Malik degree-7/5 embedded rule
//         (Genz & Malik 1980, 149 nodes in 6-D) with h-adaptive bisection of
//         the worst regions, max-error binary heap and Gladwell's batch
//         selection, run with the reference's arguments (epsrel lands in the
//         absolute slot, epsabs in the relative slot: src/sig.cpp:427-430 vs
//         src/cubature.h:70-73).  The 149 values are summed by one lane in the
//         reference's node order so that equal xi values give equal sums.
//  GAUSS  exact tent-kernel reduction of the 6-D cube-pair mean to 3-D:
//           Q = int xi(|mi(D + s)|) prod_i T(s_i) ds,  T(s) = (1-|s|/h)/h,
//         integrated by an n/2-point Gauss rule for weight (1-t) on each half
//         of every axis (n^3 nodes), one thread per integral; pairs with
//         max|dd| < near use a composite rule (4 panels x 8 Gauss-Legendre
//         nodes per half axis, 64^3 nodes) reduced over a workgroup.
//
// This file is compiled with -ffp-contract=off: the reference's arithmetic is
// plain IEEE multiply/add (x86-64 baseline code has no FMA).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

#include "common.h"
#include "xi_dev.h"

namespace opsec {

// ---- Genz-Malik constants (Genz & Malik, SIAM J. Numer. Anal. 20, 1983) -----
// lambda2^2 = 9/70, lambda4^2 = 9/10, lambda5^2 = 9/19 (correctly rounded).
#define GM_L2 0.35856858280031809199
#define GM_L4 0.94868329805051379960
#define GM_L5 0.68824720161168529772
constexpr int GM_NPTS = 149;   // 1 + 2*6 + 2*6 + 2*6*5 + 2^6

struct GMJob {
    int f;
    int key;
    double lo[6], hi[6];   // integrand bounds in unit-cube coordinate order
};

// The jobs of a quadrature call: job j is template j % npairs with function
// funcs[j / npairs] (funcs set: formed where it is read, so the nf x npairs
// list is never written), or p[j] (funcs null: a gathered list).
struct GMJobs {
    const GMJob* p;
    const int* funcs;
    long long npairs;
    __device__ __forceinline__ GMJob get(int j) const {
        if (!funcs) return p[j];
        const long long t = j / npairs;
        GMJob J = p[j - t * npairs];
        J.f = funcs[t];
        return J;
    }
};

struct GaussJob {
    int f;
    int key;
    double D[3];           // centre offset, cell 2 minus cell 1
    double h[3];           // cell widths (equal for both cells)
};

struct GMParams {
    double absTol, relTol;
    unsigned maxEval;
    int rmax;              // region pool size per wave
    int coordsys;
    double L;
};

struct TentRule {
    int n;                 // nodes per axis
    double s[64];          // unit offsets in [-1,1] (times h)
    double w[64];          // weights, sum 1
};

// ---- GM: one wavefront per integral ----------------------------------------

struct GMSlot {
    double* c;       // 6 x R centres
    double* hw;      // 6 x R half-widths
    double* val;     // R
    double* err;     // R
    double* hkey;    // R: err of the region at each heap position
    int* split;      // R
    int* heap;       // R (region indices)
    int* list;       // R
};

__device__ __forceinline__ GMSlot gm_slot(double* dpool, int* ipool, int slot, int R) {
    GMSlot s;
    double* d = dpool + (size_t)slot * 15 * R;
    int* q = ipool + (size_t)slot * 3 * R;
    s.c = d;
    s.hw = d + 6 * R;
    s.val = d + 12 * R;
    s.err = d + 13 * R;
    s.hkey = d + 14 * R;
    s.split = q;
    s.heap = q + R;
    s.list = q + 2 * R;
    return s;
}

__device__ __forceinline__ double gm_integrand(const XiModelDev& m, const XiDesc& xd, const GMJob& J,
                                               const GMParams& P, const double* u) {
    if (P.coordsys == OPSEC_COORD_CARTESIAN) {
        // src/sig.cpp:388-414
        double x1 = (1 - u[0]) * J.lo[0] + u[0] * J.hi[0];
        double y1 = (1 - u[1]) * J.lo[1] + u[1] * J.hi[1];
        double z1 = (1 - u[2]) * J.lo[2] + u[2] * J.hi[2];
        double x2 = (1 - u[3]) * J.lo[3] + u[3] * J.hi[3];
        double y2 = (1 - u[4]) * J.lo[4] + u[4] * J.hi[4];
        double z2 = (1 - u[5]) * J.lo[5] + u[5] * J.hi[5];
        return xi_eval(m, xd, box_r(x1, y1, z1, x2, y2, z2, P.L));
    }
    // src/sig.cpp:436-462: lo/hi = {r1^3, r2^3, mu1, mu2, phi1, phi2}
    double r1 = cbrt((1 - u[0]) * J.lo[0] + u[0] * J.hi[0]);
    double r2 = cbrt((1 - u[1]) * J.lo[1] + u[1] * J.hi[1]);
    double mu1 = (1 - u[2]) * J.lo[2] + u[2] * J.hi[2];
    double mu2 = (1 - u[3]) * J.lo[3] + u[3] * J.hi[3];
    double ph1 = (1 - u[4]) * J.lo[4] + u[4] * J.hi[4];
    double ph2 = (1 - u[5]) * J.lo[5] + u[5] * J.hi[5];
    return xi_eval(m, xd, sph_r(r1, mu1, ph1, r2, mu2, ph2));
}

// Node p of the rule in the reference's node order (src/cubature.c:347-443).
__device__ __forceinline__ void gm_node(int p, const double* c, const double* hw, double* u) {
    for (int d = 0; d < 6; d++) u[d] = c[d];
    if (p == 0) return;
    if (p <= 24) {
        int q = p - 1, i = q >> 2, s = q & 3;
        double r = hw[i] * (s < 2 ? GM_L2 : GM_L4);
        u[i] = (s & 1) ? c[i] + r : c[i] - r;
        return;
    }
    if (p <= 84) {
        int q = p - 25, pair = q >> 2, s = q & 3;
        int i = 0, j = 1, k = pair;
        for (i = 0; i < 5; i++) {
            int cnt = 5 - i;
            if (k < cnt) { j = i + 1 + k; break; }
            k -= cnt;
        }
        double ri = hw[i] * GM_L4, rj = hw[j] * GM_L4;
        // order: (i-, j-), (i+, j-), (i+, j+), (i-, j+)
        bool ip = (s == 1 || s == 2), jp = (s >= 2);
        u[i] = ip ? c[i] + ri : c[i] - ri;
        u[j] = jp ? c[j] + rj : c[j] - rj;
        return;
    }
    int g = p - 85;
    int signs = g ^ (g >> 1);   // Gray-code walk starting from all '+'
    for (int d = 0; d < 6; d++) {
        double r = hw[d] * GM_L5;
        u[d] = ((signs >> d) & 1) ? c[d] - r : c[d] + r;
    }
}

// ---- spherical integrand with per-region tables ------------------------------
// Along each axis the 149 nodes of a region take 7 coordinates: value index 0
// the centre c, 1/2 c -/+ hw L2, 3/4 c -/+ hw L4, 5/6 c -/+ hw L5 (the same
// c +- hw lambda sums gm_node forms).  The transcendentals of the spherical
// integrand depend on one axis (cbrt of r^3) or on an axis pair (the mu and phi
// pairs) and take at most 17 value-index pairs there, so they are computed once
// per region: 14 cbrt, 17 sqrt and 17 cos instead of 2, 2 and 1 per node.  Each
// table entry is the expression the per-node integrand forms, evaluated the same
// way, so the node values are bit-identical.

// Value index of node p along each axis (the node order of gm_node).
__device__ __forceinline__ void gm_node_vi(int p, int* vi) {
    for (int d = 0; d < 6; d++) vi[d] = 0;
    if (p == 0) return;
    if (p <= 24) {
        const int q = p - 1;
        vi[q >> 2] = 1 + (q & 3);
        return;
    }
    if (p <= 84) {
        int q = p - 25, pair = q >> 2, s = q & 3;
        int i = 0, j = 1, k = pair;
        for (i = 0; i < 5; i++) {
            int cnt = 5 - i;
            if (k < cnt) { j = i + 1 + k; break; }
            k -= cnt;
        }
        vi[i] = (s == 1 || s == 2) ? 4 : 3;
        vi[j] = (s >= 2) ? 4 : 3;
        return;
    }
    const int g = p - 85, signs = g ^ (g >> 1);
    for (int d = 0; d < 6; d++) vi[d] = ((signs >> d) & 1) ? 5 : 6;
}

// The 17 value-index pairs of an axis pair: (0,0), (1..4,0), (0,1..4),
// (3|4,3|4), (5|6,5|6).
constexpr int SPH_NCMB = 17;
__device__ __forceinline__ int sph_cmb(int a, int b) {
    if (b == 0) return a;
    if (a == 0) return 4 + b;
    if (a <= 4) return 9 + 2 * (a - 3) + (b - 3);
    return 13 + 2 * (a - 5) + (b - 5);
}
__device__ __forceinline__ void sph_cmb_pair(int k, int& a, int& b) {
    if (k <= 4) { a = k; b = 0; return; }
    if (k <= 8) { a = 0; b = k - 4; return; }
    if (k <= 12) { a = 3 + ((k - 9) >> 1); b = 3 + ((k - 9) & 1); return; }
    a = 5 + ((k - 13) >> 1);
    b = 5 + ((k - 13) & 1);
}

// Coordinate of value index v on an axis with centre c and half-width hw.
__device__ __forceinline__ double gm_axis_value(double c, double hw, int v) {
    if (v == 0) return c;
    const double r = hw * (v <= 2 ? GM_L2 : v <= 4 ? GM_L4 : GM_L5);
    return (v & 1) ? c - r : c + r;
}

// Per-region tables, one layout for registers (one thread per region) and LDS
// (a wavefront's regions): 7 values of r1, r2, mu1, mu2, then the 17
// sqrt((1 - mu1^2)(1 - mu2^2)) and cos(phi1 - phi2) of the pairs.
constexpr int SPH_TAB = 4 * 7 + 2 * SPH_NCMB;
struct SphIdx {
    enum { R1 = 0, R2 = 7, M1 = 14, M2 = 21, SQ = 28, CS = 28 + SPH_NCMB };
};

// Table entries of region (c, hw) of job J (src/sig.cpp:436-462 mapping:
// lo/hi = {r1^3, r2^3, mu1, mu2, phi1, phi2}); entry e of SPH_TAB, with the
// phi values it needs formed on the spot.
__device__ __forceinline__ double sph_entry(const GMJob& J, const double* c, const double* hw, int e) {
    auto w = [&](int d, int v) {
        const double u = gm_axis_value(c[d], hw[d], v);
        return (1 - u) * J.lo[d] + u * J.hi[d];
    };
    if (e < SphIdx::SQ) {
        const int q = e / 7, v = e - 7 * q;
        return q < 2 ? cbrt(w(q, v)) : w(q, v);
    }
    int a, b;
    if (e < SphIdx::CS) {
        sph_cmb_pair(e - SphIdx::SQ, a, b);
        const double m1 = w(2, a), m2 = w(3, b);
        return sqrt((1 - m1 * m1) * (1 - m2 * m2));
    }
    sph_cmb_pair(e - SphIdx::CS, a, b);
    return cos(w(4, a) - w(5, b));
}

// xi at node p from a region's table T (the expression of sph_r).
template <typename Tab>
__device__ __forceinline__ double sph_node_value(const XiModelDev& m, const XiDesc& xd, const Tab& T, int p) {
    int vi[6];
    gm_node_vi(p, vi);
    const double r1 = T[SphIdx::R1 + vi[0]], r2 = T[SphIdx::R2 + vi[1]];
    const double cosg = T[SphIdx::SQ + sph_cmb(vi[2], vi[3])] * T[SphIdx::CS + sph_cmb(vi[4], vi[5])] +
                        T[SphIdx::M1 + vi[2]] * T[SphIdx::M2 + vi[3]];
    return xi_eval(m, xd, sqrt(fmax(r1 * r1 + r2 * r2 - 2 * r1 * r2 * cosg, 0.0)));
}

// The degree-7 value, the degree-5 error and the split axis of one region from
// its 149 node values, summed in node order (src/cubature.c:445-613).
__device__ __forceinline__ void gm_region_sum(const double* vals, const double* hw, double& result, double& err,
                                              int& split) {
    const double ratio = (GM_L2 * GM_L2) / (GM_L4 * GM_L4);
    const double w1 = (double)(12824 - 9120 * 6 + 400 * 36) / 19683.0;
    const double w2 = 980. / 6561.;
    const double w3 = (double)(1820 - 400 * 6) / 19683.0;
    const double w4 = 200. / 19683.;
    const double w5 = 6859.0 / 19683.0 / 64.0;
    const double e1 = (double)(729 - 950 * 6 + 50 * 36) / 729.0;
    const double e2 = 245. / 486.;
    const double e3 = (double)(265 - 100 * 6) / 1458.0;
    const double e4 = 25. / 729.;
    double val0 = vals[0], sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
    double maxdiff = 0;
    split = 0;
    for (int k = 0; k < 6; k++) {
        double v0 = vals[1 + 4 * k], v1 = vals[2 + 4 * k], v2 = vals[3 + 4 * k], v3 = vals[4 + 4 * k];
        sum2 += v0 + v1;
        sum3 += v2 + v3;
        double diff = 0.0 + fabs(v0 + v1 - 2 * val0 - ratio * (v2 + v3 - 2 * val0));
        if (diff > maxdiff) {
            maxdiff = diff;
            split = k;
        }
    }
    for (int k = 0; k < 60; k++) sum4 += vals[25 + k];
    for (int k = 0; k < 64; k++) sum5 += vals[85 + k];
    double vol = 1;
    for (int d = 0; d < 6; d++) vol *= 2 * hw[d];
    result = vol * (w1 * val0 + w2 * sum2 + w3 * sum3 + w4 * sum4 + w5 * sum5);
    const double res5th = vol * (e1 * val0 + e2 * sum2 + e3 * sum3 + e4 * sum4);
    err = fabs(res5th - result);
}

constexpr int GM_SUB = 16;   // regions evaluated together by one wavefront

// Evaluate the regions list[0..n): GM_SUB at a time, the lanes share out their
// GM_SUB x 149 nodes, then lane t forms the rule of region t.
__device__ void gm_eval_regions(const XiModelDev& m, const XiDesc& xd, const GMJob& J, const GMParams& P,
                                GMSlot& S, int R, const int* list, int n, double (*vals)[GM_NPTS],
                                double (*tab)[SPH_TAB]) {
    const int lane = threadIdx.x;
    const bool sph = P.coordsys != OPSEC_COORD_CARTESIAN;
    for (int t0 = 0; t0 < n; t0 += GM_SUB) {
        const int nt = min(GM_SUB, n - t0);
        if (sph) {   // the regions' tables, then the nodes from them
            for (int i = lane; i < nt * SPH_TAB; i += blockDim.x) {
                const int t = i / SPH_TAB, e = i - t * SPH_TAB;
                const int ri = list[t0 + t];
                double c[6], hw[6];
                for (int d = 0; d < 6; d++) {
                    c[d] = S.c[d * R + ri];
                    hw[d] = S.hw[d * R + ri];
                }
                tab[t][e] = sph_entry(J, c, hw, e);
            }
            __syncthreads();
            for (int i = lane; i < nt * GM_NPTS; i += blockDim.x) {
                const int t = i / GM_NPTS, p = i - t * GM_NPTS;
                vals[t][p] = sph_node_value(m, xd, tab[t], p);
            }
        } else {
            for (int i = lane; i < nt * GM_NPTS; i += blockDim.x) {
                const int t = i / GM_NPTS, p = i - t * GM_NPTS;
                const int ri = list[t0 + t];
                double c[6], hw[6], u[6];
                for (int d = 0; d < 6; d++) {
                    c[d] = S.c[d * R + ri];
                    hw[d] = S.hw[d * R + ri];
                }
                gm_node(p, c, hw, u);
                vals[t][p] = gm_integrand(m, xd, J, P, u);
            }
        }
        __syncthreads();
        if (lane < nt) {
            const int ri = list[t0 + lane];
            double hw[6];
            for (int d = 0; d < 6; d++) hw[d] = S.hw[d * R + ri];
            double result, err;
            int split;
            gm_region_sum(vals[lane], hw, result, err, split);
            S.val[ri] = result;
            S.err[ri] = err;
            S.split[ri] = split;
        }
        __syncthreads();
    }
}

// Binary max-heap on region errors (Cormen-Leiserson-Rivest), lane 0 only.
// Storage of the region heap: position p holds region idx(p) with key(p) =
// its error, keys beside the indices so a sift level costs independent loads.
// Positions below nl sit in LDS (the top levels, which every pop walks), the
// rest in the slot's arrays.
struct HeapStore {
    double* gk;
    int* gi;
    double* lk;
    int* li;
    int nl;
    __device__ __forceinline__ double key(int p) const { return p < nl ? lk[p] : gk[p]; }
    __device__ __forceinline__ int idx(int p) const { return p < nl ? li[p] : gi[p]; }
    __device__ __forceinline__ void set(int p, double k, int i) const {
        if (p < nl) {
            lk[p] = k;
            li[p] = i;
        } else {
            gk[p] = k;
            gi[p] = i;
        }
    }
};

// Binary max-heap on region errors (Cormen-Leiserson-Rivest), one lane.
__device__ void heap_push(const GMSlot& S, const HeapStore& H, int& n, int ri, double& tv, double& te) {
    tv += S.val[ri];
    te += S.err[ri];
    int insert = n++;
    const double key = S.err[ri];
    while (insert) {
        const int parent = (insert - 1) / 2;
        const double pk = H.key(parent);
        if (key <= pk) break;
        H.set(insert, pk, H.idx(parent));
        insert = parent;
    }
    H.set(insert, key, ri);
}

__device__ int heap_pop(const GMSlot& S, const HeapStore& H, int& n, double& tv, double& te) {
    const int ret = H.idx(0);
    --n;
    const double ki = H.key(n);   // the last element, sifted down from the root
    const int ii = H.idx(n);
    H.set(0, ki, ii);
    int i = 0, child;
    while ((child = i * 2 + 1) < n) {
        const double kc = H.key(child);
        const double kc2 = child + 1 < n ? H.key(child + 1) : 0.0;
        int largest = (kc <= ki) ? i : child;
        double kl = (kc <= ki) ? ki : kc;
        if (++child < n && kl < kc2) {
            largest = child;
            kl = kc2;
        }
        if (largest == i) break;
        H.set(i, kl, H.idx(largest));
        H.set(largest, ki, ii);
        i = largest;
    }
    tv -= S.val[ret];
    te -= S.err[ret];
    return ret;
}

__device__ __forceinline__ bool gm_converged(double val, double err, const GMParams& P) {
    double rel = (val == 0.0) ? HUGE_VAL : fabs(err / val);
    return err <= P.absTol || rel <= P.relTol;
}

// One wavefront per integral; jobs idx[0..njobs) (all jobs when idx is null)
// are taken from a counter.
// NT threads per integral: 64 in the first pass; in the rerun of the deep
// integrals, whose split batches are short and follow one another, up to 512
// when there are few of them, so the nodes of a batch are spread over more
// lanes while all of them stay resident.
// NL > 0: the first NL heap positions in LDS.  A register limit of 256 per
// wave (two waves per SIMD by registers; the LDS of the 64-thread variant holds
// five per CU): 11 % faster on the shell's prior than 262 registers.
template <int NT, int NL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void gm_kernel(XiModelDev m, GMJobs jobs, const int* idx, int njobs,
                                                int* counter, double* dpool, int* ipool, GMParams P, double* Q,
                                                size_t qstride, double* Qerr, int* jobfail,
                                                unsigned long long* nevals) {
    __shared__ double vals[GM_SUB][GM_NPTS];
    __shared__ double sphtab[GM_SUB][SPH_TAB];
    __shared__ double lkey[NL > 0 ? NL : 1];
    __shared__ int lidx[NL > 0 ? NL : 1];
    __shared__ int sh_job, sh_nR, sh_done;
    const int lane = threadIdx.x;
    const int R = P.rmax;
    GMSlot S = gm_slot(dpool, ipool, blockIdx.x, R);
    const HeapStore H{S.hkey, S.heap, lkey, lidx, NL < R ? NL : R};
    for (;;) {
        if (lane == 0) sh_job = atomicAdd(counter, 1);
        __syncthreads();
        const int jl = sh_job;
        __syncthreads();
        if (jl >= njobs) break;
        const int jb = idx ? idx[jl] : jl;
        const GMJob J = jobs.get(jb);
        const XiDesc xd = m.desc[J.f];
        // region 0 = [0,1]^6
        if (lane < 6) {
            S.c[lane * R] = 0.5;
            S.hw[lane * R] = 0.5;
        }
        if (lane == 0) S.list[0] = 0;
        __syncthreads();
        gm_eval_regions(m, xd, J, P, S, R, S.list, 1, vals, sphtab);
        // lane-0 state (heap size, totals, evaluation count)
        int nheap = 0, nregions = 1;
        double tv = 0, te = 0;
        unsigned numEval = GM_NPTS;
        if (lane == 0) heap_push(S, H, nheap, 0, tv, te);
        int status = 0;
        for (;;) {
            if (lane == 0) {
                int done = 0, nR = 0;
                if (!(numEval < P.maxEval || !P.maxEval)) done = 1;
                else if (gm_converged(tv, te, P)) done = 1;
                else {
                    double eev = tv, eee = te;
                    do {
                        if (nregions + 1 > R) {   // pool exhausted
                            done = 2;
                            break;
                        }
                        int ri = heap_pop(S, H, nheap, tv, te);
                        eee -= S.err[ri];
                        // cut along the split dimension (src/cubature.c:210-224)
                        int d = S.split[ri];
                        int r2 = nregions++;
                        // the region's geometry through registers: all 12 loads
                        // first, so they overlap, then the stores
                        double cc[6], hh[6];
#pragma unroll
                        for (int k = 0; k < 6; k++) {
                            cc[k] = S.c[k * R + ri];
                            hh[k] = S.hw[k * R + ri];
                        }
                        const double h = hh[d] * 0.5;
#pragma unroll
                        for (int k = 0; k < 6; k++) {
                            S.c[k * R + r2] = k == d ? cc[k] + h : cc[k];
                            S.hw[k * R + r2] = k == d ? h : hh[k];
                        }
                        S.hw[d * R + ri] = h;
                        S.c[d * R + ri] = cc[d] - h;
                        numEval += GM_NPTS * 2;
                        S.list[nR++] = ri;
                        S.list[nR++] = r2;
                        if (gm_converged(eev, eee, P)) break;
                    } while (nheap > 0 && (numEval < P.maxEval || !P.maxEval));
                }
                sh_nR = nR;
                sh_done = done;
            }
            __syncthreads();
            int nR = sh_nR, done = sh_done;
            __syncthreads();
            if (done) {   // 2: region pool exhausted, the host reruns this job with a larger pool
                status = done;
                break;
            }
            __threadfence_block();
            gm_eval_regions(m, xd, J, P, S, R, S.list, nR, vals, sphtab);
            if (lane == 0)
                for (int t = 0; t < nR; t++) heap_push(S, H, nheap, S.list[t], tv, te);
        }
        if (lane == 0) {
            if (status == 2) {
                jobfail[jb] = 1;
            } else {
                double val = 0, err = 0;
                for (int i = 0; i < nheap; i++) {
                    val += S.val[H.idx(i)];
                    err += S.err[H.idx(i)];
                }
                Q[(size_t)J.f * qstride + J.key] = val;
                if (Qerr) Qerr[(size_t)J.f * qstride + J.key] = err;
                jobfail[jb] = 0;
                atomicAdd(nevals, (unsigned long long)numEval);
            }
        }
        __syncthreads();
    }
}

// The rerun's jobs: job j (template j % npairs, function funcs[j / npairs]) of
// the sorted list failed[0..n).
__global__ void gather_gm_jobs_kernel(const GMJob* __restrict__ tmpl, long long npairs, const int* __restrict__ funcs,
                                      const int* __restrict__ failed, int n, GMJob* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = failed[i];
    GMJob J = tmpl[j % npairs];
    J.f = funcs[j / npairs];
    out[i] = J;
}

// Indices of the jobs (idx[0..n) or 0..n) whose region pool ran out.
__global__ void list_failed_kernel(const int* __restrict__ fail, const int* __restrict__ idx, int n,
                                   int* __restrict__ out, int* __restrict__ nout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int j = idx ? idx[t] : t;
    if (fail[j]) out[atomicAdd(nout, 1)] = j;
}

// First region of every job, one thread per job: the rule on [0,1]^6 with the
// node values summed in the order of gm_region_sum.  A job whose first region
// already meets the tolerance (the driver's test before any split) ends here
// with the value gm_kernel would give; the others are listed in `pending` for
// gm_kernel.  Most band integrals end here: one region, no heap, no atomics
// per job.
__global__ __launch_bounds__(256) void gm_first_kernel(XiModelDev m, GMJobs jobs, int njobs, GMParams P,
                                                       double* Q, size_t qstride, double* Qerr, int* jobfail,
                                                       int* pending, int* npending, unsigned long long* nevals) {
    __shared__ unsigned long long blk;
    if (threadIdx.x == 0) blk = 0;
    __syncthreads();
    const int jb = blockIdx.x * blockDim.x + threadIdx.x;
    if (jb < njobs) {
        const GMJob J = jobs.get(jb);
        const XiDesc xd = m.desc[J.f];
        double c[6], hw[6], u[6];
        for (int d = 0; d < 6; d++) {
            c[d] = 0.5;
            hw[d] = 0.5;
        }
        const bool sph = P.coordsys != OPSEC_COORD_CARTESIAN;
        double T[SPH_TAB];
        if (sph) {
#pragma unroll
            for (int e = 0; e < SPH_TAB; e++) T[e] = sph_entry(J, c, hw, e);
        }
        auto node = [&](int p) {
            if (sph) return sph_node_value(m, xd, T, p);
            gm_node(p, c, hw, u);
            return gm_integrand(m, xd, J, P, u);
        };
        const double ratio = (GM_L2 * GM_L2) / (GM_L4 * GM_L4);
        const double w1 = (double)(12824 - 9120 * 6 + 400 * 36) / 19683.0;
        const double w2 = 980. / 6561.;
        const double w3 = (double)(1820 - 400 * 6) / 19683.0;
        const double w4 = 200. / 19683.;
        const double w5 = 6859.0 / 19683.0 / 64.0;
        const double e1 = (double)(729 - 950 * 6 + 50 * 36) / 729.0;
        const double e2 = 245. / 486.;
        const double e3 = (double)(265 - 100 * 6) / 1458.0;
        const double e4 = 25. / 729.;
        const double val0 = node(0);
        double sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
        for (int k = 0; k < 6; k++) {
            double v[4];
            for (int e = 0; e < 4; e++) v[e] = node(1 + 4 * k + e);
            sum2 += v[0] + v[1];
            sum3 += v[2] + v[3];
        }
        for (int p = 25; p < 85; p++) sum4 += node(p);
        for (int p = 85; p < GM_NPTS; p++) sum5 += node(p);
        double vol = 1;
        for (int d = 0; d < 6; d++) vol *= 2 * hw[d];
        const double result = vol * (w1 * val0 + w2 * sum2 + w3 * sum3 + w4 * sum4 + w5 * sum5);
        const double res5th = vol * (e1 * val0 + e2 * sum2 + e3 * sum3 + e4 * sum4);
        const double err = fabs(res5th - result);
        (void)ratio;
        // the heap totals after the first push are 0 + val and 0 + err
        const double tv = 0.0 + result, te = 0.0 + err;
        if (!((unsigned)GM_NPTS < P.maxEval || !P.maxEval) || gm_converged(tv, te, P)) {
            Q[(size_t)J.f * qstride + J.key] = tv;
            if (Qerr) Qerr[(size_t)J.f * qstride + J.key] = te;
            jobfail[jb] = 0;
            atomicAdd(&blk, (unsigned long long)GM_NPTS);
        } else {
            pending[atomicAdd(npending, 1)] = jb;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && blk) atomicAdd(nevals, blk);
}

// First regions of a shell's band sweep, GMB_PAIRS arc pairs per block for
// every band of the sweep at once.  A pair's region table, its 149 node
// separations r and r^3 are formed once instead of once per band, and the
// band-edge term F_k(r) = sin kr - kr cos kr is shared by the two bands that
// meet at edge k: one sincos per band and node instead of two when the sweep's
// bands follow one another.  A node value is band_xi's, bit for bit (the same
// operations; its series branch, klo r < 0.5, calls band_xi), and each band's
// rule is summed by one thread in gm_first_kernel's node order, so the results,
// the convergence test and the pending jobs (job t npairs + i for band t of pair
// i) are those of gm_first_kernel.  Thread q GM_NPTS + n forms node n of pair
// slot q; the bands go through LDS GMB_CHUNK at a time.
constexpr int GMB_PAIRS = 3, GMB_CHUNK = 8;
constexpr int GMB_THREADS = 448;   // >= GMB_PAIRS * GM_NPTS, whole wavefronts
static_assert(GMB_THREADS >= GMB_PAIRS * GM_NPTS && GMB_THREADS % 64 == 0 &&
                  GMB_PAIRS * GMB_CHUNK <= GMB_THREADS,
              "a node thread per pair node, a summing thread per chunk band");
__global__ __launch_bounds__(GMB_THREADS) void gm_first_bands_kernel(
    XiModelDev m, const GMJob* __restrict__ tmpl, int npairs, const int* __restrict__ funcs, int nf,
    const double* __restrict__ edge, const int* __restrict__ fedge, GMParams P, double* Q, size_t qstride,
    double* Qerr, int* jobfail, int* pending, int* npending, unsigned long long* nevals) {
    __shared__ double tab[GMB_PAIRS][SPH_TAB];
    __shared__ double vals[GMB_PAIRS][GMB_CHUNK][GM_NPTS];
    __shared__ unsigned long long blk;
    const int t = threadIdx.x;
    const int q = t / GM_NPTS, n = t - q * GM_NPTS;
    const int i0 = blockIdx.x * GMB_PAIRS;
    const bool live = q < GMB_PAIRS && i0 + q < npairs;
    if (t == 0) blk = 0;
    for (int k = t; k < GMB_PAIRS * SPH_TAB; k += GMB_THREADS) {
        const int qq = k / SPH_TAB, e = k - qq * SPH_TAB;
        if (i0 + qq < npairs) {
            const GMJob J = tmpl[i0 + qq];
            double c[6], hw[6];
            for (int d = 0; d < 6; d++) {
                c[d] = 0.5;
                hw[d] = 0.5;
            }
            tab[qq][e] = sph_entry(J, c, hw, e);
        }
    }
    __syncthreads();
    double r = 0.0, r3 = 0.0;
    if (live) {   // the expression of sph_node_value
        int vi[6];
        gm_node_vi(n, vi);
        const double* T = tab[q];
        const double r1 = T[SphIdx::R1 + vi[0]], r2 = T[SphIdx::R2 + vi[1]];
        const double cosg = T[SphIdx::SQ + sph_cmb(vi[2], vi[3])] * T[SphIdx::CS + sph_cmb(vi[4], vi[5])] +
                            T[SphIdx::M1 + vi[2]] * T[SphIdx::M2 + vi[3]];
        r = sqrt(fmax(r1 * r1 + r2 * r2 - 2 * r1 * r2 * cosg, 0.0));
        r3 = r * r * r;
    }
    int ce = -1;        // edge of the last closed-form term of this node
    double cF = 0.0;
    for (int f0 = 0; f0 < nf; f0 += GMB_CHUNK) {
        const int nc = min(GMB_CHUNK, nf - f0);
        if (live)
            for (int u = 0; u < nc; u++) {
                const int ilo = fedge[2 * (f0 + u)], ihi = fedge[2 * (f0 + u) + 1];
                const double klo = edge[ilo], khi = edge[ihi];
                const double xl = klo * r;
                double v;
                if (xl >= 0.5) {   // band_xi's closed form
                    double Fl;
                    if (ce == ilo) {
                        Fl = cF;
                    } else {
                        double sl, cl;
                        sincos(xl, &sl, &cl);
                        Fl = sl - xl * cl;
                    }
                    const double xh = khi * r;
                    double sh, ch;
                    sincos(xh, &sh, &ch);
                    const double Fh = sh - xh * ch;
                    v = 0.050660591821168885 * (Fh - Fl) / r3;
                    ce = ihi;
                    cF = Fh;
                } else {
                    v = band_xi(klo, khi, r);
                    ce = -1;
                }
                vals[q][u][n] = v;
            }
        __syncthreads();
        if (t < GMB_PAIRS * GMB_CHUNK) {
            const int qq = t / GMB_CHUNK, u = t - qq * GMB_CHUNK, ii = i0 + qq;
            if (u < nc && ii < npairs) {
                const double* V = vals[qq][u];
                const double w1 = (double)(12824 - 9120 * 6 + 400 * 36) / 19683.0;
                const double w2 = 980. / 6561.;
                const double w3 = (double)(1820 - 400 * 6) / 19683.0;
                const double w4 = 200. / 19683.;
                const double w5 = 6859.0 / 19683.0 / 64.0;
                const double e1 = (double)(729 - 950 * 6 + 50 * 36) / 729.0;
                const double e2 = 245. / 486.;
                const double e3 = (double)(265 - 100 * 6) / 1458.0;
                const double e4 = 25. / 729.;
                const double val0 = V[0];
                double sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
                for (int k = 0; k < 6; k++) {
                    sum2 += V[1 + 4 * k] + V[2 + 4 * k];
                    sum3 += V[3 + 4 * k] + V[4 + 4 * k];
                }
                for (int p = 25; p < 85; p++) sum4 += V[p];
                for (int p = 85; p < GM_NPTS; p++) sum5 += V[p];
                double vol = 1;
                for (int d = 0; d < 6; d++) vol *= 2 * 0.5;
                const double result = vol * (w1 * val0 + w2 * sum2 + w3 * sum3 + w4 * sum4 + w5 * sum5);
                const double res5th = vol * (e1 * val0 + e2 * sum2 + e3 * sum3 + e4 * sum4);
                const double err = fabs(res5th - result);
                const double tv = 0.0 + result, te = 0.0 + err;
                const int jb = (f0 + u) * npairs + ii;
                if (!((unsigned)GM_NPTS < P.maxEval || !P.maxEval) || gm_converged(tv, te, P)) {
                    const size_t o = (size_t)funcs[f0 + u] * qstride + tmpl[ii].key;
                    Q[o] = tv;
                    if (Qerr) Qerr[o] = te;
                    jobfail[jb] = 0;
                    atomicAdd(&blk, (unsigned long long)GM_NPTS);
                } else {
                    pending[atomicAdd(npending, 1)] = jb;
                }
            }
        }
        __syncthreads();
    }
    if (t == 0 && blk) atomicAdd(nevals, blk);
}

// The rule on region ri of S evaluated by one thread: node values summed on the
// fly in the order of gm_region_sum.
__device__ void gm_thread_region(const XiModelDev& m, const XiDesc& xd, const GMJob& J, const GMParams& P, GMSlot& S,
                                 int R, int ri) {
    double c[6], hw[6], u[6];
    for (int d = 0; d < 6; d++) {
        c[d] = S.c[d * R + ri];
        hw[d] = S.hw[d * R + ri];
    }
    const bool sph = P.coordsys != OPSEC_COORD_CARTESIAN;
    double T[SPH_TAB];
    if (sph) {
#pragma unroll
        for (int e = 0; e < SPH_TAB; e++) T[e] = sph_entry(J, c, hw, e);
    }
    auto node = [&](int p) {
        if (sph) return sph_node_value(m, xd, T, p);
        gm_node(p, c, hw, u);
        return gm_integrand(m, xd, J, P, u);
    };
    const double ratio = (GM_L2 * GM_L2) / (GM_L4 * GM_L4);
    const double w1 = (double)(12824 - 9120 * 6 + 400 * 36) / 19683.0;
    const double w2 = 980. / 6561.;
    const double w3 = (double)(1820 - 400 * 6) / 19683.0;
    const double w4 = 200. / 19683.;
    const double w5 = 6859.0 / 19683.0 / 64.0;
    const double e1 = (double)(729 - 950 * 6 + 50 * 36) / 729.0;
    const double e2 = 245. / 486.;
    const double e3 = (double)(265 - 100 * 6) / 1458.0;
    const double e4 = 25. / 729.;
    const double val0 = node(0);
    double sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0, maxdiff = 0;
    int split = 0;
    for (int k = 0; k < 6; k++) {
        double v[4];
        for (int e = 0; e < 4; e++) v[e] = node(1 + 4 * k + e);
        sum2 += v[0] + v[1];
        sum3 += v[2] + v[3];
        const double diff = 0.0 + fabs(v[0] + v[1] - 2 * val0 - ratio * (v[2] + v[3] - 2 * val0));
        if (diff > maxdiff) {
            maxdiff = diff;
            split = k;
        }
    }
    for (int p = 25; p < 85; p++) sum4 += node(p);
    for (int p = 85; p < GM_NPTS; p++) sum5 += node(p);
    double vol = 1;
    for (int d = 0; d < 6; d++) vol *= 2 * hw[d];
    const double result = vol * (w1 * val0 + w2 * sum2 + w3 * sum3 + w4 * sum4 + w5 * sum5);
    const double res5th = vol * (e1 * val0 + e2 * sum2 + e3 * sum3 + e4 * sum4);
    S.val[ri] = result;
    S.err[ri] = fabs(res5th - result);
    S.split[ri] = split;
}

constexpr int GM_TR = 16;   // regions a thread keeps in private memory

// The adaptive driver of gm_kernel's lane 0, one thread per integral, regions
// in private memory: integrals that converge within GM_TR regions end here
// with the same value; the others are listed in `deeper` for gm_kernel.
__global__ __launch_bounds__(256) void gm_thread_kernel(XiModelDev m, GMJobs jobs, const int* idx, int n,
                                                        GMParams P, double* Q, size_t qstride, double* Qerr,
                                                        int* jobfail, int* deeper, int* ndeeper,
                                                        unsigned long long* nevals) {
    __shared__ unsigned long long blk;
    if (threadIdx.x == 0) blk = 0;
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        const int jb = idx[t];
        const GMJob J = jobs.get(jb);
        const XiDesc xd = m.desc[J.f];
        double dreg[15 * GM_TR];
        int ireg[3 * GM_TR];
        GMSlot S;
        S.c = dreg;
        S.hw = dreg + 6 * GM_TR;
        S.val = dreg + 12 * GM_TR;
        S.err = dreg + 13 * GM_TR;
        S.hkey = dreg + 14 * GM_TR;
        S.split = ireg;
        S.heap = ireg + GM_TR;
        S.list = ireg + 2 * GM_TR;
        const int R = GM_TR;
        const HeapStore H{S.hkey, S.heap, nullptr, nullptr, 0};
        for (int d = 0; d < 6; d++) {
            S.c[d * R] = 0.5;
            S.hw[d * R] = 0.5;
        }
        gm_thread_region(m, xd, J, P, S, R, 0);
        int nheap = 0, nregions = 1;
        double tv = 0, te = 0;
        unsigned numEval = GM_NPTS;
        heap_push(S, H, nheap, 0, tv, te);
        bool overflow = false;
        for (;;) {
            if (!(numEval < P.maxEval || !P.maxEval)) break;
            if (gm_converged(tv, te, P)) break;
            const double eev = tv;
            double eee = te;
            int nR = 0;
            do {
                if (nregions + 1 > R) {
                    overflow = true;
                    break;
                }
                const int ri = heap_pop(S, H, nheap, tv, te);
                eee -= S.err[ri];
                const int d = S.split[ri];
                const int r2 = nregions++;
                double cc[6], hh[6];
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    cc[k] = S.c[k * R + ri];
                    hh[k] = S.hw[k * R + ri];
                }
                const double h = hh[d] * 0.5;
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    S.c[k * R + r2] = k == d ? cc[k] + h : cc[k];
                    S.hw[k * R + r2] = k == d ? h : hh[k];
                }
                S.hw[d * R + ri] = h;
                S.c[d * R + ri] = cc[d] - h;
                numEval += GM_NPTS * 2;
                S.list[nR++] = ri;
                S.list[nR++] = r2;
                if (gm_converged(eev, eee, P)) break;
            } while (nheap > 0 && (numEval < P.maxEval || !P.maxEval));
            if (overflow) break;
            for (int q = 0; q < nR; q++) gm_thread_region(m, xd, J, P, S, R, S.list[q]);
            for (int q = 0; q < nR; q++) heap_push(S, H, nheap, S.list[q], tv, te);
        }
        if (overflow) {
            deeper[atomicAdd(ndeeper, 1)] = jb;
        } else {
            double val = 0, err = 0;
            for (int i = 0; i < nheap; i++) {
                val += S.val[H.idx(i)];
                err += S.err[H.idx(i)];
            }
            Q[(size_t)J.f * qstride + J.key] = val;
            if (Qerr) Qerr[(size_t)J.f * qstride + J.key] = err;
            jobfail[jb] = 0;
            atomicAdd(&blk, (unsigned long long)numEval);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && blk) atomicAdd(nevals, blk);
}

// ---- GAUSS: the tensor rule, a node slab per thread -------------------------

// One thread per (integral, a): the b and c sums of x-node a; the a == 0
// thread of each integral then adds the n slab sums in node order, so the
// result is the plain a-outer triple sum.  256 / n integrals per workgroup.
__global__ __launch_bounds__(256) void tent_kernel(XiModelDev m, const GaussJob* jobs, int njobs, TentRule rule,
                                                   double L, double* Q, size_t qstride) {
    __shared__ double s_s[64], s_w[64];
    __shared__ double slab[256];
    const int n = rule.n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s_s[i] = rule.s[i];
        s_w[i] = rule.w[i];
    }
    __syncthreads();
    const int per = blockDim.x / n;   // integrals per workgroup
    const int local = threadIdx.x / n, a = threadIdx.x - local * n;
    const long long jb = (long long)blockIdx.x * per + local;
    const bool live = local < per && jb < njobs;
    if (live) {
        const GaussJob J = jobs[jb];
        const XiDesc xd = m.desc[J.f];
        const double rx = min_image(J.D[0] + s_s[a] * J.h[0], L);
        double accy = 0;
        for (int b = 0; b < n; b++) {
            double ry = min_image(J.D[1] + s_s[b] * J.h[1], L);
            double rxy2 = rx * rx + ry * ry;
            double accz = 0;
            for (int c = 0; c < n; c++) {
                double rz = min_image(J.D[2] + s_s[c] * J.h[2], L);
                accz += s_w[c] * xi_eval(m, xd, sqrt(rxy2 + rz * rz));
            }
            accy += s_w[b] * accz;
        }
        slab[threadIdx.x] = accy;
    }
    __syncthreads();
    if (live && a == 0) {
        double acc = 0;
        for (int i = 0; i < n; i++) acc += s_w[i] * slab[threadIdx.x + i];
        const GaussJob& J = jobs[jb];
        Q[(size_t)J.f * qstride + J.key] = acc;
    }
}

// Near pairs: composite rule.  Each integral is spread over gridDim.y
// workgroups of 256 threads, one (a, b) node column per thread; a workgroup's
// tree-reduced partial goes to part[job * gridDim.y + y].
__global__ __launch_bounds__(256) void tent_near_kernel(XiModelDev m, const GaussJob* jobs, TentRule rule, double L,
                                                        double* part) {
    __shared__ double s_s[64], s_w[64];
    __shared__ double red[256];
    const int n = rule.n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s_s[i] = rule.s[i];
        s_w[i] = rule.w[i];
    }
    __syncthreads();
    const GaussJob J = jobs[blockIdx.x];
    const XiDesc xd = m.desc[J.f];
    const int ab = blockIdx.y * blockDim.x + threadIdx.x;
    double acc = 0;
    if (ab < n * n) {
        const int a = ab / n, b = ab % n;
        double rx = min_image(J.D[0] + s_s[a] * J.h[0], L);
        double ry = min_image(J.D[1] + s_s[b] * J.h[1], L);
        double rxy2 = rx * rx + ry * ry;
        double accz = 0;
        for (int c = 0; c < n; c++) {
            double rz = min_image(J.D[2] + s_s[c] * J.h[2], L);
            accz += s_w[c] * xi_eval(m, xd, sqrt(rxy2 + rz * rz));
        }
        acc = s_w[a] * s_w[b] * accz;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * gridDim.y + blockIdx.y] = red[0];
}

// The near integrals: the nsplit partials of each, added in order.
__global__ void tent_near_sum_kernel(const GaussJob* jobs, int njobs, int nsplit, const double* part, double* Q,
                                     size_t qstride) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    double acc = 0;
    for (int i = 0; i < nsplit; i++) acc += part[(size_t)j * nsplit + i];
    Q[(size_t)jobs[j].f * qstride + jobs[j].key] = acc;
}

// ---- first-pair search and job construction --------------------------------

struct CellGeo {
    int d1, d2, d3;
    double Nbar;
};

__device__ __forceinline__ int box_key(int ax, int ay, int az, int bx, int by, int bz, int Ny, int Nz) {
    return (abs(ax - bx) * Ny + abs(ay - by)) * Nz + abs(az - bz);
}

__global__ void first_pair_kernel(const int* cellidx, const int* rows, int nrows, const int* cols, int ncols,
                                  int Ny, int Nz, int row0, unsigned long long* first) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= ncols || i >= nrows) return;
    int a = rows[i], b = cols[j];
    int key = box_key(cellidx[3 * a], cellidx[3 * a + 1], cellidx[3 * a + 2], cellidx[3 * b], cellidx[3 * b + 1],
                      cellidx[3 * b + 2], Ny, Nz);
    atomicMin(&first[key], (unsigned long long)(i + row0) * (unsigned long long)ncols + (unsigned long long)j);
}

__global__ void fill_u64_kernel(unsigned long long* p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Assembly: s[i*ld + j] = sqrt(Nbar_a Nbar_b) Q[key(a,b)] (src/sig.cpp:120, 203).
// Tile of 8 rows x 512 columns per 256-thread block; column geometry in LDS.
__global__ __launch_bounds__(256) void assemble_kernel(const int* cellidx, const double* nbar, const int* rows,
                                                       int nrows, const int* cols, int ncols, PairKey K,
                                                       const double* Qf, double* s, size_t ld) {
    __shared__ int cx[512], cy[512], cz[512];
    __shared__ double cn[512];
    const int j0 = blockIdx.x * 512, i0 = blockIdx.y * 8;
    for (int t = threadIdx.x; t < 512; t += 256) {
        int j = j0 + t;
        if (j < ncols) {
            int b = cols ? cols[j] : j;
            cx[t] = cellidx[3 * b];
            cy[t] = cellidx[3 * b + 1];
            cz[t] = cellidx[3 * b + 2];
            cn[t] = nbar[b];
        }
    }
    __syncthreads();
    for (int r = 0; r < 8; r++) {
        int i = i0 + r;
        if (i >= nrows) break;
        int a = rows ? rows[i] : i;
        int ax = cellidx[3 * a], ay = cellidx[3 * a + 1], az = cellidx[3 * a + 2];
        double na = nbar[a];
        for (int t = threadIdx.x; t < 512; t += 256) {
            int j = j0 + t;
            if (j >= ncols) break;
            const int key = K(ax, ay, az, cx[t], cy[t], cz[t]);
            s[(size_t)i * ld + j] = sqrt(na * cn[t]) * Qf[key];
        }
    }
}

// ---- host side ------------------------------------------------------------


struct Scratch {
    void* p = nullptr;
    size_t cap = 0;
    ~Scratch() { dev_free(&p); }
};

// Jobs of nf functions from one per-pair template: job t * npairs + i is
// template i with function funcs[t].
__global__ void expand_jobs_kernel(const GaussJob* __restrict__ tmpl, int npairs, const int* __restrict__ funcs,
                                   int nf, GaussJob* __restrict__ out) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)npairs * nf) return;
    const int t = (int)(j / npairs), i = (int)(j - (size_t)t * npairs);
    GaussJob J = tmpl[i];
    J.f = funcs[t];
    out[j] = J;
}

static int make_tent_rule(opsec_ctx* ctx, int nodes, TentRule* r) {
    OPSEC_CHECK_ARG(ctx, nodes >= 2 && nodes <= 32 && nodes % 2 == 0, "gauss nodes per axis must be even, 2..32");
    int m = nodes / 2;
    double t[16], w[16];
    gauss_jacobi10(m, t, w);
    r->n = nodes;
    // negative half (mirror, reversed) then positive half
    for (int k = 0; k < m; k++) {
        r->s[k] = -t[m - 1 - k];
        r->w[k] = w[m - 1 - k];
        r->s[m + k] = t[k];
        r->w[m + k] = w[k];
    }
    return OPSEC_OK;
}

static void make_near_rule(TentRule* r) {
    const int panels = 4, q = 8;
    double x[8], w[8];
    gauss_legendre(q, x, w);
    int m = panels * q;   // nodes per half
    double t[32], wt[32];
    for (int p = 0; p < panels; p++)
        for (int k = 0; k < q; k++) {
            double a = (double)p / panels, b = (double)(p + 1) / panels;
            double tt = 0.5 * (a + b) + 0.5 * (b - a) * x[k];
            t[p * q + k] = tt;
            wt[p * q + k] = 0.5 * (b - a) * w[k] * (1.0 - tt);
        }
    r->n = 2 * m;
    for (int k = 0; k < m; k++) {
        r->s[k] = -t[m - 1 - k];
        r->w[k] = wt[m - 1 - k];
        r->s[m + k] = t[k];
        r->w[m + k] = wt[k];
    }
}

// Geometry of cell d in the survey grid, the way Survey::CreateEmptyCell does
// it (src/Survey.cpp:71-91).
static void grid_cell(const SurveyDev& sv, int d1, int d2, int d3, double* lo, double* hi) {
    lo[0] = sv.min1 + d1 * (sv.max1 - sv.min1) / sv.n1;
    hi[0] = sv.min1 + (d1 + 1) * (sv.max1 - sv.min1) / sv.n1;
    lo[1] = sv.min2 + d2 * (sv.max2 - sv.min2) / sv.n2;
    hi[1] = sv.min2 + (d2 + 1) * (sv.max2 - sv.min2) / sv.n2;
    lo[2] = sv.min3 + d3 * (sv.max3 - sv.min3) / sv.n3;
    hi[2] = sv.min3 + (d3 + 1) * (sv.max3 - sv.min3) / sv.n3;
}

PairGeo grid_pair_geo(const SurveyDev& sv, int kx, int ky, int kz) {
    PairGeo g;
    grid_cell(sv, 0, 0, 0, g.lo1, g.hi1);
    grid_cell(sv, kx, ky, kz, g.lo2, g.hi2);
    return g;
}

// Run method q->method on the given key pairs for functions funcs; results to
// dQ[f * nkeys + key].  keys[i] with geometry geo[i].
// Genz-Malik integrals of the functions funcs[0..nf) over the npairs device
// templates d_tmpl (f unset): the nf x npairs jobs, function-major (the prior,
// with the most adaptive work, first), are formed from the templates where the
// kernels read them (GMJobs).
// Device buffers of run_gm_device, kept across the calls of one pair-table
// sweep (its chunks) so the multi-GB pools and lists are allocated once.
struct GMScratch {
    Scratch sj, sd, si, sc, sf, sp, st, sx, sa, se;
};
static std::mutex g_gm_stats_mutex;
// Shell tables: the deep sweep beside the shallow one on a second stream (1, the
// default) or after it (0).
static int g_gm_overlap = 1;
// Shell band sweeps: first regions by gm_first_bands_kernel (1, the default) or
// by gm_first_kernel (0); the same results.
static int g_gm_bands = 1;
// Wavefront slots of the rerun with whole-cap pools (0: as many as memory allows).
static int g_gm_deep_slots = 0;

static int run_gm_device(opsec_ctx* ctx, int nf, const int* funcs, const GMJob* d_tmpl, size_t npairs,
                         const opsec_quad* q, double* dQ, size_t nkeys, double* dQerr, GMScratch* gs = nullptr,
                         hipStream_t stm = nullptr, int deep_cap = 0) {
    if (!stm) stm = ctx->stream;
    XiModelDev mv = model_view(ctx);
    if (npairs == 0 || nf == 0) return OPSEC_OK;
    GMScratch local;
    GMScratch& S = gs ? *gs : local;
    Scratch& sa = S.sa;
    Scratch& sj = S.sj;
    Scratch& sd = S.sd;
    Scratch& si = S.si;
    Scratch& sc = S.sc;
    Scratch& sf = S.sf;
    Scratch& sp = S.sp;
    Scratch& st = S.st;
    Scratch& sx = S.sx;
    Scratch& se = S.se;
    GMParams P;
    P.absTol = q->epsrel;   // reference argument order, src/sig.cpp:427-430
    P.relTol = q->epsabs;
    P.maxEval = (unsigned)q->maxeval;
    P.coordsys = ctx->sv.coordsys;
    P.L = ctx->sv.L;
    // Pass 0: every job's first region (one thread per job), then the jobs
    // that need splits with a 4096-region pool per wavefront slot.  Pass 1:
    // jobs that exhausted it, with pools that hold every region the
    // evaluation cap allows (maxeval / 149 + 2), on fewer slots.
    const size_t njobs_all = npairs * (size_t)nf;
    if (njobs_all > (size_t)INT32_MAX) return set_error(ctx, OPSEC_EINVAL, "too many pair integrals in one call");
    // A shell's sweep of closed-form bands takes its first regions from
    // gm_first_bands_kernel: the distinct band edges, and each band's two edges
    // as indices into them.
    bool bands = g_gm_bands && ctx->sv.coordsys == OPSEC_COORD_SPHERICAL;
    for (int t = 0; t < nf && bands; t++) bands = ctx->funcs[funcs[t]].kind == XI_BAND;
    std::vector<double> edges;
    std::vector<int> fedge;
    if (bands) {
        for (int t = 0; t < nf; t++) {
            edges.push_back(ctx->funcs[funcs[t]].a);
            edges.push_back(ctx->funcs[funcs[t]].b);
        }
        std::sort(edges.begin(), edges.end());
        edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
        for (int t = 0; t < nf; t++)
            for (double k : {ctx->funcs[funcs[t]].a, ctx->funcs[funcs[t]].b})
                fedge.push_back((int)(std::lower_bound(edges.begin(), edges.end(), k) - edges.begin()));
        int rc;
        if ((rc = dev_realloc(ctx, &se.p, &se.cap, edges.size() * 8 + fedge.size() * 4))) return rc;
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(se.p, edges.data(), edges.size() * 8, hipMemcpyHostToDevice, stm));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync((double*)se.p + edges.size(), fedge.data(), fedge.size() * 4,
                                          hipMemcpyHostToDevice, stm));
    }
    unsigned long long nev_total = 0;
    int nagain = 0;             // pass-1 jobs, gathered on the device into sa
    for (int pass = 0; pass < 2; pass++) {
        const int njobs = pass == 0 ? (int)njobs_all : nagain;
        if (njobs == 0) break;
        long long cap_regions = P.maxEval ? (long long)P.maxEval / GM_NPTS + 2 : (1LL << 20);
        P.rmax = pass == 0 ? 4096 : (int)std::min<long long>(cap_regions, 1LL << 22);
        int nslots = pass == 0 ? std::min(njobs, ctx->num_cus * 8) : std::min(njobs, 64);
        if (pass == 1) {
            // as many wavefronts as half the free device memory holds pools for
            size_t fre = 0, tot = 0;
            if (hipMemGetInfo(&fre, &tot) != hipSuccess) fre = (size_t)16 << 30;
            const size_t per_slot = (size_t)P.rmax * (15 * 8 + 3 * 4);
            const size_t budget = std::max<size_t>((size_t)1 << 30, fre / 2);
            nslots = std::min(njobs, ctx->num_cus * 8);
            nslots = (int)std::max<size_t>(1, std::min<size_t>(nslots, budget / per_slot));
            const int cap = g_gm_deep_slots > 0 ? g_gm_deep_slots : deep_cap;
            if (cap > 0) nslots = std::min(nslots, cap);
        }
        int rc;
        if (pass == 1 && (rc = dev_realloc(ctx, &sj.p, &sj.cap, sizeof(GMJob) * njobs))) return rc;
        if ((rc = dev_realloc(ctx, &sd.p, &sd.cap, sizeof(double) * 15 * (size_t)P.rmax * nslots))) return rc;
        if ((rc = dev_realloc(ctx, &si.p, &si.cap, sizeof(int) * 3 * (size_t)P.rmax * nslots))) return rc;
        if ((rc = dev_realloc(ctx, &sc.p, &sc.cap, 64))) return rc;
        if ((rc = dev_realloc(ctx, &sf.p, &sf.cap, sizeof(int) * njobs))) return rc;
        if ((rc = dev_realloc(ctx, &sp.p, &sp.cap, sizeof(int) * njobs))) return rc;
        if ((rc = dev_realloc(ctx, &sx.p, &sx.cap, sizeof(int) * njobs))) return rc;
        OPSEC_HIP_TRY(ctx, hipMemsetAsync(sc.p, 0, 64, stm));
        int* d_counter = (int*)sc.p;
        unsigned long long* d_nev = (unsigned long long*)((char*)sc.p + 8);
        int* d_npend = (int*)((char*)sc.p + 16);
        int* d_nfail = (int*)((char*)sc.p + 20);
        int* d_ndeep = (int*)((char*)sc.p + 24);
        const int* d_idx = nullptr;
        int nrun = njobs;
        GMJobs jobs{(const GMJob*)sj.p, nullptr, 0};   // the rerun's gathered jobs
        if (pass == 0) {
            if ((rc = dev_realloc(ctx, &st.p, &st.cap, sizeof(int) * nf))) return rc;
            int* dfun = (int*)st.p;
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dfun, funcs, sizeof(int) * nf, hipMemcpyHostToDevice, stm));
            jobs = GMJobs{d_tmpl, dfun, (long long)npairs};
            {
                // first regions, then the shallow integrals one thread each
                if (bands)
                    hipLaunchKernelGGL(gm_first_bands_kernel,
                                       dim3((unsigned)((npairs + GMB_PAIRS - 1) / GMB_PAIRS)), dim3(GMB_THREADS), 0,
                                       stm, mv, d_tmpl, (int)npairs, (const int*)st.p, nf, (const double*)se.p,
                                       (const int*)((const double*)se.p + edges.size()), P, dQ, nkeys, dQerr,
                                       (int*)sf.p, (int*)sx.p, d_npend, d_nev);
                else
                    hipLaunchKernelGGL(gm_first_kernel, dim3((unsigned)((njobs + 255) / 256)), dim3(256), 0,
                                       stm, mv, jobs, njobs, P, dQ, nkeys, dQerr, (int*)sf.p, (int*)sx.p, d_npend,
                                       d_nev);
                OPSEC_HIP_TRY(ctx, hipGetLastError());
                int npend = 0;
                OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&npend, d_npend, 4, hipMemcpyDeviceToHost, stm));
                OPSEC_HIP_TRY(ctx, hipStreamSynchronize(stm));
                nrun = 0;
                if (npend > 0) {
                    hipLaunchKernelGGL(gm_thread_kernel, dim3((unsigned)((npend + 255) / 256)), dim3(256), 0,
                                       stm, mv, jobs, (const int*)sx.p, npend, P, dQ, nkeys,
                                       dQerr, (int*)sf.p, (int*)sp.p, d_ndeep, d_nev);
                    OPSEC_HIP_TRY(ctx, hipGetLastError());
                    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&nrun, d_ndeep, 4, hipMemcpyDeviceToHost, stm));
                    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(stm));
                }
                d_idx = (const int*)sp.p;
            }
        } else {
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(sj.p, sa.p, sizeof(GMJob) * njobs, hipMemcpyDeviceToDevice, stm));
        }
        if (nrun > 0) {
            if (pass == 0)
                hipLaunchKernelGGL((gm_kernel<64, 0>), dim3(std::min(nslots, nrun)), dim3(64), 0, stm, mv,
                                   jobs, d_idx, nrun, d_counter, (double*)sd.p, (int*)si.p, P, dQ,
                                   nkeys, dQerr, (int*)sf.p, d_nev);
            else if (nrun <= 2 * ctx->num_cus)   // few deep integrals: 8 waves each
                hipLaunchKernelGGL((gm_kernel<512, 8192>), dim3(std::min(nslots, nrun)), dim3(512), 0, stm, mv,
                                   jobs, d_idx, nrun, d_counter, (double*)sd.p, (int*)si.p, P, dQ,
                                   nkeys, dQerr, (int*)sf.p, d_nev);
            else if (nrun <= 4 * ctx->num_cus)
                hipLaunchKernelGGL((gm_kernel<256, 1024>), dim3(std::min(nslots, nrun)), dim3(256), 0, stm, mv,
                                   jobs, d_idx, nrun, d_counter, (double*)sd.p, (int*)si.p, P, dQ,
                                   nkeys, dQerr, (int*)sf.p, d_nev);
            else   // many: one wave each
                hipLaunchKernelGGL((gm_kernel<64, 0>), dim3(std::min(nslots, nrun)), dim3(64), 0, stm, mv,
                                   jobs, d_idx, nrun, d_counter, (double*)sd.p, (int*)si.p, P, dQ,
                                   nkeys, dQerr, (int*)sf.p, d_nev);
            OPSEC_HIP_TRY(ctx, hipGetLastError());
        }
        // jobs whose region pool ran out, listed on the device
        int nfail = 0;
        if (nrun > 0) {
            hipLaunchKernelGGL(list_failed_kernel, dim3((unsigned)((nrun + 255) / 256)), dim3(256), 0, stm,
                               (const int*)sf.p, d_idx, nrun, (int*)sx.p, d_nfail);
            OPSEC_HIP_TRY(ctx, hipGetLastError());
        }
        unsigned long long nev = 0;
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&nev, d_nev, 8, hipMemcpyDeviceToHost, stm));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&nfail, d_nfail, 4, hipMemcpyDeviceToHost, stm));
        OPSEC_HIP_TRY(ctx, hipStreamSynchronize(stm));
        nev_total += nev;
        if (nfail > 0 && pass == 1)
            return set_error(ctx, OPSEC_ECONV, "adaptive cubature exceeded %d regions", P.rmax);
        std::vector<int> failed(nfail);
        if (nfail > 0) {
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(failed.data(), sx.p, sizeof(int) * nfail, hipMemcpyDeviceToHost,
                                              stm));
            OPSEC_HIP_TRY(ctx, hipStreamSynchronize(stm));
        }
        std::sort(failed.begin(), failed.end());
        nagain = nfail;
        if (nfail > 0) {
            // the rerun's jobs: template j % npairs with function funcs[j / npairs]
            if ((rc = dev_realloc(ctx, &sa.p, &sa.cap, sizeof(GMJob) * nfail + sizeof(int) * (nfail + nf)))) return rc;
            int* d_failed = (int*)((GMJob*)sa.p + nfail);
            int* d_fun = d_failed + nfail;
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_failed, failed.data(), sizeof(int) * nfail, hipMemcpyHostToDevice,
                                              stm));
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_fun, funcs, sizeof(int) * nf, hipMemcpyHostToDevice, stm));
            hipLaunchKernelGGL(gather_gm_jobs_kernel, dim3((unsigned)((nfail + 255) / 256)), dim3(256), 0, stm,
                               d_tmpl, (long long)npairs, d_fun, d_failed, nfail, (GMJob*)sa.p);
            OPSEC_HIP_TRY(ctx, hipGetLastError());
            OPSEC_HIP_TRY(ctx, hipStreamSynchronize(stm));
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_gm_stats_mutex);   // sweeps may run on two host threads
        ctx->st_integrals += (int64_t)njobs_all;
        ctx->st_nodes += (int64_t)nev_total;
    }
    return OPSEC_OK;
}

static int run_quadrature(opsec_ctx* ctx, int nf, const int* funcs, const std::vector<int>& keys,
                          const std::vector<PairGeo>& geo, const std::vector<int>& nearflag, const opsec_quad* q,
                          double* dQ, size_t nkeys, double* dQerr = nullptr) {
    XiModelDev mv = model_view(ctx);
    const size_t npairs = keys.size();
    if (npairs == 0 || nf == 0) return OPSEC_OK;
    int method = q->method;
    if (method == OPSEC_QUAD_GM) {
        // one template per pair (f unset); the nf x npairs jobs, function-major
        // (the prior, with the most adaptive work, first), are formed from them on
        // the device
        std::vector<GMJob> tmpl(npairs);
        for (size_t i = 0; i < npairs; i++) {
            GMJob& J = tmpl[i];
            J.f = -1;
            J.key = keys[i];
            const PairGeo& g = geo[i];
            if (ctx->sv.coordsys == OPSEC_COORD_CARTESIAN) {
                for (int d = 0; d < 3; d++) {
                    J.lo[d] = g.lo1[d];
                    J.hi[d] = g.hi1[d];
                    J.lo[3 + d] = g.lo2[d];
                    J.hi[3 + d] = g.hi2[d];
                }
            } else {
                J.lo[0] = g.lo1[0] * g.lo1[0] * g.lo1[0];
                J.hi[0] = g.hi1[0] * g.hi1[0] * g.hi1[0];
                J.lo[1] = g.lo2[0] * g.lo2[0] * g.lo2[0];
                J.hi[1] = g.hi2[0] * g.hi2[0] * g.hi2[0];
                J.lo[2] = g.lo1[1];
                J.hi[2] = g.hi1[1];
                J.lo[3] = g.lo2[1];
                J.hi[3] = g.hi2[1];
                J.lo[4] = g.lo1[2];
                J.hi[4] = g.hi1[2];
                J.lo[5] = g.lo2[2];
                J.hi[5] = g.hi2[2];
            }
        }
        Scratch st;
        int rc;
        if ((rc = dev_realloc(ctx, &st.p, &st.cap, sizeof(GMJob) * npairs))) return rc;
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(st.p, tmpl.data(), sizeof(GMJob) * npairs, hipMemcpyHostToDevice,
                                          ctx->stream));
        return run_gm_device(ctx, nf, funcs, (const GMJob*)st.p, npairs, q, dQ, nkeys, dQerr);
    }
    if (method == OPSEC_QUAD_GAUSS) {
        OPSEC_CHECK_ARG(ctx, ctx->sv.coordsys == OPSEC_COORD_CARTESIAN, "GAUSS quadrature needs a box survey");
        TentRule rule, nrule;
        int rc = make_tent_rule(ctx, q->nodes, &rule);
        if (rc) return rc;
        make_near_rule(&nrule);
        // one template per pair on the host; the nf x pairs jobs are expanded on
        // the device into the context's job buffer (far jobs, then near jobs)
        std::vector<GaussJob> far, nearj;
        for (size_t i = 0; i < npairs; i++) {
            const PairGeo& g = geo[i];
            GaussJob J;
            J.f = -1;
            J.key = keys[i];
            for (int d = 0; d < 3; d++) {
                double h1 = g.hi1[d] - g.lo1[d], h2 = g.hi2[d] - g.lo2[d];
                if (std::fabs(h1 - h2) > 1e-9 * std::fabs(h1))
                    return set_error(ctx, OPSEC_EINVAL, "GAUSS quadrature needs equal cell widths");
                J.h[d] = h1;
                J.D[d] = 0.5 * (g.lo2[d] + g.hi2[d]) - 0.5 * (g.lo1[d] + g.hi1[d]);
            }
            (nearflag[i] ? nearj : far).push_back(J);
        }
        const size_t nfar = far.size() * nf, nnear = nearj.size() * nf, nj = nfar + nnear;
        const size_t ntmpl = far.size() + nearj.size();
        // near partials after the jobs, templates and function list
        const int nsplit = (nrule.n * nrule.n + 255) / 256;
        const size_t part_off = (sizeof(GaussJob) * (nj + ntmpl) + sizeof(int) * nf + 255) & ~(size_t)255;
        if ((rc = dev_realloc(ctx, &ctx->d_jobs, &ctx->cap_jobs, part_off + sizeof(double) * nnear * nsplit)))
            return rc;
        GaussJob* dj = (GaussJob*)ctx->d_jobs;
        GaussJob* dt = dj + nj;
        int* dfun = (int*)(dt + ntmpl);
        double* dpart = (double*)((char*)ctx->d_jobs + part_off);
        if (!far.empty())
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dt, far.data(), sizeof(GaussJob) * far.size(), hipMemcpyHostToDevice,
                                              ctx->stream));
        if (!nearj.empty())
            OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dt + far.size(), nearj.data(), sizeof(GaussJob) * nearj.size(),
                                              hipMemcpyHostToDevice, ctx->stream));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dfun, funcs, sizeof(int) * nf, hipMemcpyHostToDevice, ctx->stream));
        if (!far.empty())
            hipLaunchKernelGGL(expand_jobs_kernel, dim3((unsigned)((nfar + 255) / 256)), dim3(256), 0, ctx->stream, dt,
                               (int)far.size(), dfun, nf, dj);
        if (!nearj.empty())
            hipLaunchKernelGGL(expand_jobs_kernel, dim3((unsigned)((nnear + 255) / 256)), dim3(256), 0, ctx->stream,
                               dt + far.size(), (int)nearj.size(), dfun, nf, dj + nfar);
        if (nfar) {
            const size_t per = 256 / rule.n;
            hipLaunchKernelGGL(tent_kernel, dim3((unsigned)((nfar + per - 1) / per)), dim3(256), 0, ctx->stream, mv,
                               dj, (int)nfar, rule, ctx->sv.L, dQ, nkeys);
            OPSEC_HIP_TRY(ctx, hipGetLastError());
        }
        if (nnear) {
            hipLaunchKernelGGL(tent_near_kernel, dim3((unsigned)nnear, (unsigned)nsplit), dim3(256), 0, ctx->stream, mv,
                               dj + nfar, nrule, ctx->sv.L, dpart);
            hipLaunchKernelGGL(tent_near_sum_kernel, dim3((unsigned)((nnear + 255) / 256)), dim3(256), 0, ctx->stream,
                               dj + nfar, (int)nnear, nsplit, dpart, dQ, nkeys);
            OPSEC_HIP_TRY(ctx, hipGetLastError());
        }
        OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        ctx->st_integrals += (int64_t)nj;
        ctx->st_nodes += (int64_t)nfar * rule.n * rule.n * rule.n + (int64_t)nnear * nrule.n * nrule.n * nrule.n;
        return OPSEC_OK;
    }
    return set_error(ctx, OPSEC_EINVAL, "unknown quadrature method %d", method);
}

static int check_model(opsec_ctx* ctx, int f) {
    if (ctx->model_kind < 0) return set_error(ctx, OPSEC_ESTATE, "no model loaded");
    OPSEC_CHECK_ARG(ctx, f >= 0 && f < (int)ctx->funcs.size(), "function index %d out of range", f);
    return OPSEC_OK;
}

static int check_quad(opsec_ctx* ctx, const opsec_quad* q) {
    OPSEC_CHECK_ARG(ctx, q, "quadrature options are NULL");
    OPSEC_CHECK_ARG(ctx, q->method >= 0 && q->method <= 2, "unknown quadrature method %d", q->method);
    if (q->method == OPSEC_QUAD_GM) OPSEC_CHECK_ARG(ctx, q->maxeval >= 0, "maxeval must be >= 0");
    return OPSEC_OK;
}

int surrogate_tables(opsec_ctx* ctx, int nf, const int* funcs, const opsec_quad* q, const std::vector<int>& keys,
                     const std::vector<PairGeo>& geo, const std::vector<int>& nearflag);

// Pair geometry for a key of the box grid from an explicit cell pair.
static PairGeo pair_geo(const opsec_cell& c1, const opsec_cell& c2) {
    PairGeo g;
    g.lo1[0] = c1.min1; g.hi1[0] = c1.max1;
    g.lo1[1] = c1.min2; g.hi1[1] = c1.max2;
    g.lo1[2] = c1.min3; g.hi1[2] = c1.max3;
    g.lo2[0] = c2.min1; g.hi2[0] = c2.max1;
    g.lo2[1] = c2.min2; g.hi2[1] = c2.max2;
    g.lo2[2] = c2.min3; g.hi2[2] = c2.max3;
    return g;
}

static int near_of(const opsec_cell& a, const opsec_cell& b, int near) {
    int m = std::max(std::abs(a.d1 - b.d1), std::max(std::abs(a.d2 - b.d2), std::abs(a.d3 - b.d3)));
    return m < near ? 1 : 0;
}

// First cell pairs of the arc pairs (ha, hb), ha in [ha0, ha0 + nha), hb in
// [0, narcs): for each k = |dphi_a - dphi_b| the first pair met in cell order
// (row cell of arc ha, then column cell of arc hb), as a job template of the
// pair with the lower cell index first.  Arc members are listed in cell order
// (arc_start / arc_mem), cb holds each cell's {min, max} of r, mu and phi.
// The templates are appended in no particular order (each integral is
// independent; count receives their number).  W 32-bit words hold the k seen.
template <int W>
__global__ __launch_bounds__(256) void sph_templates_kernel(const int* __restrict__ arc_start,
                                                            const int* __restrict__ arc_mem,
                                                            const int* __restrict__ cd3, const double* __restrict__ cb,
                                                            int narcs, int Nphi, int ha0, int nha,
                                                            GMJob* __restrict__ out,
                                                            unsigned long long* __restrict__ count) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nha * narcs) return;
    const int ha = ha0 + (int)(t / narcs), hb = (int)(t % narcs);
    const int a0 = arc_start[ha], a1 = arc_start[ha + 1], b0 = arc_start[hb], b1 = arc_start[hb + 1];
    unsigned seen[W];
#pragma unroll
    for (int w = 0; w < W; w++) seen[w] = 0u;
    int left = Nphi;
    for (int ia = a0; ia < a1 && left > 0; ia++) {
        const int a = arc_mem[ia], da = cd3[a];
        for (int ib = b0; ib < b1; ib++) {
            const int b = arc_mem[ib];
            const int k = abs(da - cd3[b]);
            if ((seen[k >> 5] >> (k & 31)) & 1u) continue;
            seen[k >> 5] |= 1u << (k & 31);
            const double* c1 = cb + 6 * (size_t)min(a, b);
            const double* c2 = cb + 6 * (size_t)max(a, b);
            GMJob J;
            J.f = -1;
            J.key = (ha * narcs + hb) * Nphi + k;
            J.lo[0] = c1[0] * c1[0] * c1[0];
            J.hi[0] = c1[1] * c1[1] * c1[1];
            J.lo[1] = c2[0] * c2[0] * c2[0];
            J.hi[1] = c2[1] * c2[1] * c2[1];
            J.lo[2] = c1[2];
            J.hi[2] = c1[3];
            J.lo[3] = c2[2];
            J.hi[3] = c2[3];
            J.lo[4] = c1[4];
            J.hi[4] = c1[5];
            J.lo[5] = c2[4];
            J.hi[5] = c2[5];
            out[atomicAdd(count, 1ull)] = J;
            if (--left == 0) break;
        }
    }
}

// Arc-pair tables of a spherical survey (src/sig.cpp:125-213) for the block of
// all rows and all columns: for every ordered arc pair (h_a, h_b) and each
// k = |dphi_a - dphi_b| the integral of the first cell pair met in cell order
// (row cell of arc h_a, then column cell of arc h_b), taken with the lower
// cell index first.  Keys that no cell pair carries stay zero.  The integrals
// run in chunks of row arcs, so the host job list stays bounded.
static int build_pair_tables_sph(opsec_ctx* ctx, int nf, const int* funcs, const opsec_quad* q) {
    OPSEC_CHECK_ARG(ctx, q->method == OPSEC_QUAD_GM,
                    "spherical surveys use the adaptive (GM) rule; the tent reduction and the surrogate need box cells");
    const SurveyDev& sv = ctx->sv;
    const int Nmu = sv.n2, Nphi = sv.n3, narcs = sv.n1 * sv.n2, N = ctx->ncells;
    const std::vector<opsec_cell>& cells = ctx->h_cells;
    std::vector<std::vector<int>> arcs(narcs);
    for (int a = 0; a < N; a++) arcs[cells[a].d1 * Nmu + cells[a].d2].push_back(a);
    for (int t = 0; t < nf; t++)
        OPSEC_HIP_TRY(ctx, hipMemsetAsync(ctx->d_Q + (size_t)funcs[t] * ctx->nkeys, 0, (size_t)ctx->nkeys * 8,
                                          ctx->stream));
    // Functions without band edges (the prior, whose near-diagonal integrals
    // run to the evaluation cap) and band functions are integrated in separate
    // sweeps.  A quadrature call takes up to 2^28 jobs (30 GB of job records);
    // its duration is set by its deepest integrals, so grouping the deep ones
    // into few calls pays that wait few times.
    std::vector<int> deep, shallow;
    for (int t = 0; t < nf; t++) {
        const XiDesc& d = ctx->funcs[funcs[t]];
        (d.b > d.a && d.b > 0 ? shallow : deep).push_back(funcs[t]);
    }
    if (Nphi <= 1024) {
        // templates generated on the device, a range of row arcs at a time
        std::vector<int> arc_start(narcs + 1, 0), arc_mem, cd3(N);
        std::vector<double> cb(6 * (size_t)N);
        for (int h = 0; h < narcs; h++) {
            arc_start[h + 1] = arc_start[h] + (int)arcs[h].size();
            arc_mem.insert(arc_mem.end(), arcs[h].begin(), arcs[h].end());
        }
        for (int a = 0; a < N; a++) {
            const opsec_cell& c = cells[a];
            cd3[a] = c.d3;
            const double v[6] = {c.min1, c.max1, c.min2, c.max2, c.min3, c.max3};
            std::copy(v, v + 6, cb.begin() + 6 * (size_t)a);
        }
        Scratch sg, st;
        int rc;
        const size_t geo_bytes = 4 * (size_t)(narcs + 1) + 8 * (size_t)N + 48 * (size_t)N + 64;
        if ((rc = dev_realloc(ctx, &sg.p, &sg.cap, geo_bytes))) return rc;
        int* d_start = (int*)sg.p;
        int* d_mem = d_start + (narcs + 1);
        int* d_d3 = d_mem + N;
        double* d_cb = (double*)(((uintptr_t)(d_d3 + N) + 15) & ~(uintptr_t)15);
        unsigned long long* d_count = (unsigned long long*)(d_cb + 6 * (size_t)N);   // two counters
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_start, arc_start.data(), 4 * (size_t)(narcs + 1), hipMemcpyHostToDevice,
                                          ctx->stream));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_mem, arc_mem.data(), 4 * (size_t)N, hipMemcpyHostToDevice, ctx->stream));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_d3, cd3.data(), 4 * (size_t)N, hipMemcpyHostToDevice, ctx->stream));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_cb, cb.data(), 48 * (size_t)N, hipMemcpyHostToDevice, ctx->stream));
        const size_t per_row = (size_t)narcs * Nphi;   // keys of one row arc at most
        // One sweep: a group's functions over every arc pair, a range of row arcs
        // at a time, on stream s with its own template buffer tb and counter cnt.
        auto sweep = [&](const std::vector<int>& group, hipStream_t s, Scratch& tb, unsigned long long* cnt,
                         int deep_cap) -> int {
            GMScratch gs;   // the sweep's job buffers, kept across its chunks
            const int ng = (int)group.size();
            const size_t chunk = std::max<size_t>(4096, ((size_t)1 << 28) / (size_t)ng);
            const int nha = (int)std::max<size_t>(1, std::min<size_t>(narcs, chunk / per_row));
            int rc2;
            if ((rc2 = dev_realloc(ctx, &tb.p, &tb.cap, sizeof(GMJob) * (size_t)nha * per_row))) return rc2;
            for (int ha0 = 0; ha0 < narcs; ha0 += nha) {
                const int nh = std::min(nha, narcs - ha0);
                OPSEC_HIP_TRY(ctx, hipMemsetAsync(cnt, 0, 8, s));
                const long long nt = (long long)nh * narcs;
                if (Nphi <= 64)
                    hipLaunchKernelGGL(sph_templates_kernel<2>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s,
                                       d_start, d_mem, d_d3, d_cb, narcs, Nphi, ha0, nh, (GMJob*)tb.p, cnt);
                else
                    hipLaunchKernelGGL(sph_templates_kernel<32>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s,
                                       d_start, d_mem, d_d3, d_cb, narcs, Nphi, ha0, nh, (GMJob*)tb.p, cnt);
                OPSEC_HIP_TRY(ctx, hipGetLastError());
                unsigned long long npairs = 0;
                OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&npairs, cnt, 8, hipMemcpyDeviceToHost, s));
                OPSEC_HIP_TRY(ctx, hipStreamSynchronize(s));
                if ((rc2 = run_gm_device(ctx, ng, group.data(), (const GMJob*)tb.p, (size_t)npairs, q, ctx->d_Q,
                                         (size_t)ctx->nkeys, nullptr, &gs, s, deep_cap)))
                    return rc2;
            }
            return OPSEC_OK;
        };
        if (deep.empty() || shallow.empty() || !g_gm_overlap) {
            for (const std::vector<int>* group : {&deep, &shallow})
                if (!group->empty() && (rc = sweep(*group, ctx->stream, st, d_count, 0))) return rc;
            return OPSEC_OK;
        }
        // The deep sweep (few integrals, most of them run to the evaluation cap
        // by latency-bound wavefronts that leave the CUs mostly idle) runs on a
        // second stream from a second host thread, beside the shallow sweep's
        // full-device passes.  The sweeps write different functions' tables and
        // every integral is computed alone, so the tables are the same bits.  Its
        // rerun takes at most three wavefronts per CU, so that the registers of
        // its waiting waves leave room for the shallow sweep's (at the 65536-cell
        // shell 8.3-8.5 s for both sweeps against 9.0 s with eight per CU).
        hipStream_t s2 = nullptr;
        hipEvent_t ready = nullptr;
        if (hipStreamCreateWithFlags(&s2, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(ready, ctx->stream) != hipSuccess || hipStreamWaitEvent(s2, ready, 0) != hipSuccess) {
            const hipError_t e = hipGetLastError();
            if (ready) hipEventDestroy(ready);
            if (s2) hipStreamDestroy(s2);
            return set_error(ctx, OPSEC_EHIP, "second quadrature stream: %s", hipGetErrorString(e));
        }
        Scratch st2;
        int rc_deep = OPSEC_OK;
        std::thread worker;
        try {
            worker = std::thread([&] {
                if (hipSetDevice(ctx->device) != hipSuccess)
                    rc_deep = set_error(ctx, OPSEC_EHIP, "quadrature thread: hipSetDevice(%d) failed", ctx->device);
                else
                    rc_deep = sweep(deep, s2, st2, d_count + 1, 3 * ctx->num_cus);
            });
        } catch (...) {   // no thread: the deep sweep runs after the shallow one, on s2
        }
        rc = sweep(shallow, ctx->stream, st, d_count, 0);
        if (worker.joinable())
            worker.join();
        else if (!rc)
            rc_deep = sweep(deep, s2, st2, d_count + 1, 0);
        const hipError_t e2 = hipStreamSynchronize(s2);
        hipEventDestroy(ready);
        hipStreamDestroy(s2);
        if (rc) return rc;
        if (rc_deep) return rc_deep;
        if (e2 != hipSuccess) return set_error(ctx, OPSEC_EHIP, "quadrature stream: %s", hipGetErrorString(e2));
        return OPSEC_OK;
    }
    std::vector<int> keys, nearflag, tab(Nphi);
    std::vector<PairGeo> geo;
    for (const std::vector<int>* group : {&deep, &shallow}) {
        if (group->empty()) continue;
        const int ng = (int)group->size();
        const size_t chunk = std::max<size_t>(4096, ((size_t)1 << 28) / (size_t)ng);
        keys.clear();
        geo.clear();
        for (int ha = 0; ha < narcs; ha++) {
            if (!arcs[ha].empty())
                for (int hb = 0; hb < narcs; hb++) {
                    if (arcs[hb].empty()) continue;
                    std::fill(tab.begin(), tab.end(), 0);
                    int left = Nphi;
                    for (int a : arcs[ha]) {
                        for (int b : arcs[hb]) {
                            const int k = std::abs(cells[a].d3 - cells[b].d3);
                            if (tab[k]) continue;
                            tab[k] = 1;
                            keys.push_back((ha * narcs + hb) * Nphi + k);
                            geo.push_back(pair_geo(cells[std::min(a, b)], cells[std::max(a, b)]));
                            if (--left == 0) break;
                        }
                        if (left == 0) break;
                    }
                }
            if (keys.size() >= chunk || (ha == narcs - 1 && !keys.empty())) {
                nearflag.assign(keys.size(), 0);
                int rc = run_quadrature(ctx, ng, group->data(), keys, geo, nearflag, q, ctx->d_Q, (size_t)ctx->nkeys);
                if (rc) return rc;
                keys.clear();
                geo.clear();
            }
        }
    }
    return OPSEC_OK;
}

int build_pair_tables(opsec_ctx* ctx, int nf, const int* funcs, const opsec_quad* q) {
    if (ctx->sv.coordsys == OPSEC_COORD_SPHERICAL) return build_pair_tables_sph(ctx, nf, funcs, q);
    // Keys of the full survey matrix: first occurrences along row 0, then the
    // grid-canonical pair for keys that row 0 does not carry.
    const SurveyDev& sv = ctx->sv;
    const int nk = ctx->nkeys;
    const int N = ctx->ncells;
    std::vector<int> owner(nk, -1);
    const opsec_cell& c0 = ctx->h_cells[0];
    for (int b = 0; b < N; b++) {
        const opsec_cell& cb = ctx->h_cells[b];
        int key = (std::abs(c0.d1 - cb.d1) * sv.n2 + std::abs(c0.d2 - cb.d2)) * sv.n3 + std::abs(c0.d3 - cb.d3);
        if (owner[key] < 0) owner[key] = b;
    }
    std::vector<int> keys;
    std::vector<PairGeo> geo;
    std::vector<int> nearflag;
    keys.reserve(nk);
    for (int key = 0; key < nk; key++) {
        int kz = key % sv.n3, ky = (key / sv.n3) % sv.n2, kx = key / (sv.n3 * sv.n2);
        PairGeo g;
        int nflag;
        if (owner[key] >= 0) {
            const opsec_cell& cb = ctx->h_cells[owner[key]];
            g = pair_geo(c0, cb);   // row 0 is the lower index
            nflag = near_of(c0, cb, q->near);
        } else {
            grid_cell(sv, 0, 0, 0, g.lo1, g.hi1);
            grid_cell(sv, kx, ky, kz, g.lo2, g.hi2);
            nflag = std::max(kx, std::max(ky, kz)) < q->near;
        }
        keys.push_back(key);
        geo.push_back(g);
        nearflag.push_back(nflag);
    }
    if (q->method == OPSEC_QUAD_SURROGATE) return surrogate_tables(ctx, nf, funcs, q, keys, geo, nearflag);
    return run_quadrature(ctx, nf, funcs, keys, geo, nearflag, q, ctx->d_Q, nk);
}

int run_quadrature_public(opsec_ctx* ctx, int nf, const int* funcs, const std::vector<int>& keys,
                          const std::vector<PairGeo>& geo, const std::vector<int>& nearflag, const opsec_quad* q,
                          double* dQ, size_t nkeys) {
    return run_quadrature(ctx, nf, funcs, keys, geo, nearflag, q, dQ, nkeys);
}

}  // namespace opsec

using namespace opsec;

extern "C" {

int opsec_quad_defaults(int method, opsec_quad* q) {
    if (!q) return OPSEC_EINVAL;
    std::memset(q, 0, sizeof *q);
    q->method = method;
    q->nodes = 6;
    q->near = 1;
    q->maxeval = 50000000;   // SIGNAL_MAXEVAL, src/sig.cpp:35-37
    q->epsrel = 1e-5;        // src/sig.h:24
    q->epsabs = 1e-10;
    q->spot_frac = 0.01;
    q->spot_tol = 1e-3;
    q->spot_seed = 11;
    if (method == OPSEC_QUAD_SURROGATE) q->near = 2;
    return (method >= 0 && method <= 2) ? OPSEC_OK : OPSEC_EINVAL;
}

int opsec_cells_upload(opsec_ctx* ctx, const opsec_cell* cells, int ncells, const opsec_survey* sv) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, cells && ncells > 0, "no cells");
    SurveyDev s;
    int rc = survey_from_abi(ctx, sv, &s);
    if (rc) return rc;
    for (int a = 0; a < ncells; a++) {
        const opsec_cell& c = cells[a];
        OPSEC_CHECK_ARG(ctx, c.d1 >= 0 && c.d1 < s.n1 && c.d2 >= 0 && c.d2 < s.n2 && c.d3 >= 0 && c.d3 < s.n3,
                        "cell %d has grid index outside the survey grid", a);
    }
    OPSEC_CHECK_ARG(ctx, pair_nkeys_of(s) <= INT32_MAX, "survey grid too large for a pair table (%lld keys)",
                    pair_nkeys_of(s));
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    invalidate_modes(ctx);   // modes are weighted by the old cells' Nbar
    ctx->sv = s;
    ctx->ncells = ncells;
    ctx->nkeys = (int)pair_nkeys_of(s);
    ctx->h_cells.assign(cells, cells + ncells);
    std::vector<int> idx(3 * (size_t)ncells);
    std::vector<double> nb(ncells);
    for (int a = 0; a < ncells; a++) {
        idx[3 * a] = cells[a].d1;
        idx[3 * a + 1] = cells[a].d2;
        idx[3 * a + 2] = cells[a].d3;
        nb[a] = cells[a].Nbar;
    }
    size_t cap = 0;
    dev_free((void**)&ctx->d_cellidx);
    dev_free((void**)&ctx->d_nbar);
    dev_free((void**)&ctx->d_Q);
    if ((rc = dev_realloc(ctx, (void**)&ctx->d_cellidx, &cap, idx.size() * 4))) return rc;
    cap = 0;
    if ((rc = dev_realloc(ctx, (void**)&ctx->d_nbar, &cap, nb.size() * 8))) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_cellidx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_nbar, nb.data(), nb.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->Q_valid.assign(ctx->funcs.size(), 0);
    return OPSEC_OK;
}

int opsec_pair_nkeys(opsec_ctx* ctx, int* nkeys) {
    if (!ctx || !nkeys) return OPSEC_EINVAL;
    if (ctx->ncells == 0) return set_error(ctx, OPSEC_ESTATE, "no survey uploaded");
    *nkeys = ctx->nkeys;
    return OPSEC_OK;
}

int opsec_pair_tables(opsec_ctx* ctx, int nf, const int* funcs, const opsec_quad* q) {
    if (!ctx) return OPSEC_EINVAL;
    if (ctx->ncells == 0) return set_error(ctx, OPSEC_ESTATE, "no survey uploaded");
    if (ctx->model_kind < 0) return set_error(ctx, OPSEC_ESTATE, "no model loaded");
    int rc = check_quad(ctx, q);
    if (rc) return rc;
    OPSEC_CHECK_ARG(ctx, nf >= 1 && funcs, "no functions");
    for (int t = 0; t < nf; t++)
        if ((rc = check_model(ctx, funcs[t]))) return rc;
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    size_t need = (size_t)ctx->funcs.size() * ctx->nkeys * 8;
    if (!ctx->d_Q) {
        size_t cap = 0;
        if ((rc = dev_realloc(ctx, (void**)&ctx->d_Q, &cap, need))) return rc;
    }
    ctx->st_integrals = ctx->st_nodes = ctx->st_surrogate = ctx->st_spot = ctx->st_spot_fail = 0;
    rc = build_pair_tables(ctx, nf, funcs, q);
    if (rc) return rc;
    for (int t = 0; t < nf; t++) ctx->Q_valid[funcs[t]] = 1;
    return OPSEC_OK;
}

int opsec_pair_table_get(opsec_ctx* ctx, int f, double* Q) {
    if (!ctx || !Q) return OPSEC_EINVAL;
    int rc = check_model(ctx, f);
    if (rc) return rc;
    if (!ctx->d_Q || !ctx->Q_valid[f]) return set_error(ctx, OPSEC_ESTATE, "pair table %d not built", f);
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(Q, ctx->d_Q + (size_t)f * ctx->nkeys, (size_t)ctx->nkeys * 8,
                                      hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

int opsec_pair_stats(opsec_ctx* ctx, int64_t* integrals, int64_t* nodes, int64_t* surrogate, int64_t* spot,
                     int64_t* spot_fail) {
    if (!ctx) return OPSEC_EINVAL;
    if (integrals) *integrals = ctx->st_integrals;
    if (nodes) *nodes = ctx->st_nodes;
    if (surrogate) *surrogate = ctx->st_surrogate;
    if (spot) *spot = ctx->st_spot;
    if (spot_fail) *spot_fail = ctx->st_spot_fail;
    return OPSEC_OK;
}

int opsec_assemble_device(opsec_ctx* ctx, int f, double* dS, int ld) {
    if (!ctx || !dS) return OPSEC_EINVAL;
    int rc = check_model(ctx, f);
    if (rc) return rc;
    if (!ctx->d_Q || !ctx->Q_valid[f]) return set_error(ctx, OPSEC_ESTATE, "pair table %d not built", f);
    int N = ctx->ncells;
    if (ld <= 0) ld = N;
    OPSEC_CHECK_ARG(ctx, ld >= N, "leading dimension %d < Ncells %d", ld, N);
    hipLaunchKernelGGL(assemble_kernel, dim3((N + 511) / 512, (N + 7) / 8), dim3(256), 0, ctx->stream, ctx->d_cellidx,
                       ctx->d_nbar, (const int*)nullptr, N, (const int*)nullptr, N, pair_key_of(ctx->sv),
                       ctx->d_Q + (size_t)f * ctx->nkeys, dS, (size_t)ld);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    return OPSEC_OK;
}

// Drop-in for ComputeSignalMatrixBlock (box surveys).
// Spherical signal block (src/sig.cpp:125-213): cells are grouped into arcs
// h = d_r * Nmu + d_mu (cell order within an arc).  For each (row arc, column
// arc) the first (row, column) pair met with a given k = |d_phi_a - d_phi_b|
// is integrated by the adaptive rule (lower cell index first) and reused for
// every pair of the two arcs with that k.  The distinct integrals run as one
// batch of wavefront jobs; the assembly is a gather on the device.
__global__ void sph_assemble_kernel(const int* __restrict__ rarc, const int* __restrict__ rphi,
                                    const double* __restrict__ rnb, int nrows, const int* __restrict__ carc,
                                    const int* __restrict__ cphi, const double* __restrict__ cnb, int ncols,
                                    const int* __restrict__ table, int ncarcs, int Nphi, const double* __restrict__ Q,
                                    double* __restrict__ out, size_t lld, int row0) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = row0 + blockIdx.y;
    if (j >= ncols || i >= nrows) return;
    const int k = abs(rphi[i] - cphi[j]);
    const int id = table[((size_t)rarc[i] * ncarcs + carc[j]) * Nphi + k];
    out[(size_t)i * lld + j] = sqrt(rnb[i] * cnb[j]) * Q[id];
}

static int sig_block_sph(opsec_ctx* ctx, const opsec_cell* cells, int ncells, const int* rows, int nrows,
                         const int* cols, int ncols, int f, const SurveyDev& svd, const opsec_quad* q, double* s,
                         int lld) {
    OPSEC_CHECK_ARG(ctx, q->method == OPSEC_QUAD_GM,
                    "spherical surveys use the adaptive (GM) rule; the tent reduction needs box cells");
    const int Nmu = svd.n2, Nphi = svd.n3, narcs = svd.n1 * svd.n2;
    std::vector<int> rg(ncells, -1), cg(ncells, -1);
    for (int i = 0; i < nrows; i++) rg[rows[i]] = i;
    for (int j = 0; j < ncols; j++) cg[cols[j]] = j;
    std::vector<std::vector<int>> arcs(narcs);
    for (int a = 0; a < ncells; a++) arcs[cells[a].d1 * Nmu + cells[a].d2].push_back(a);
    // arcs holding rows / columns of the block
    std::vector<int> rai(narcs, -1), cai(narcs, -1), rlist, clist;
    for (int h = 0; h < narcs; h++) {
        bool hr = false, hc = false;
        for (int a : arcs[h]) {
            hr = hr || rg[a] >= 0;
            hc = hc || cg[a] >= 0;
        }
        if (hr) {
            rai[h] = (int)rlist.size();
            rlist.push_back(h);
        }
        if (hc) {
            cai[h] = (int)clist.size();
            clist.push_back(h);
        }
    }
    const size_t tsize = rlist.size() * clist.size() * (size_t)Nphi;
    std::vector<int> table(tsize, -1);
    std::vector<PairGeo> geo;
    std::vector<int> keys;
    for (size_t x = 0; x < rlist.size(); x++)
        for (size_t y = 0; y < clist.size(); y++) {
            int* tab = &table[(x * clist.size() + y) * Nphi];
            int left = Nphi;
            for (int a : arcs[rlist[x]]) {
                if (rg[a] < 0) continue;
                for (int b : arcs[clist[y]]) {
                    if (cg[b] < 0) continue;
                    const int k = std::abs(cells[a].d3 - cells[b].d3);
                    if (tab[k] >= 0) continue;
                    tab[k] = (int)keys.size();
                    keys.push_back((int)keys.size());
                    geo.push_back(pair_geo(cells[std::min(a, b)], cells[std::max(a, b)]));
                    if (--left == 0) break;
                }
                if (left == 0) break;
            }
        }
    const int nd = (int)keys.size();
    std::vector<int> rarc(nrows), rphi(nrows), carc(ncols), cphi(ncols);
    std::vector<double> rnb(nrows), cnb(ncols);
    for (int i = 0; i < nrows; i++) {
        const opsec_cell& c = cells[rows[i]];
        rarc[i] = rai[c.d1 * Nmu + c.d2];
        rphi[i] = c.d3;
        rnb[i] = c.Nbar;
    }
    for (int j = 0; j < ncols; j++) {
        const opsec_cell& c = cells[cols[j]];
        carc[j] = cai[c.d1 * Nmu + c.d2];
        cphi[j] = c.d3;
        cnb[j] = c.Nbar;
    }
    SurveyDev saved = ctx->sv;
    ctx->sv = svd;
    Scratch sQ, sT, sR, sC, sRn, sCn, sOut;
    int rc;
    const size_t nfun = ctx->funcs.size();
    hipStream_t st = ctx->stream;
    std::vector<int> nearflag(nd, 0);
    int fs[1] = {f};
    if ((rc = dev_realloc(ctx, &sQ.p, &sQ.cap, std::max<size_t>(1, nfun * nd) * 8))) goto out;
    if ((rc = dev_realloc(ctx, &sT.p, &sT.cap, std::max<size_t>(1, tsize) * 4))) goto out;
    if ((rc = dev_realloc(ctx, &sR.p, &sR.cap, (size_t)nrows * 8))) goto out;
    if ((rc = dev_realloc(ctx, &sC.p, &sC.cap, (size_t)ncols * 8))) goto out;
    if ((rc = dev_realloc(ctx, &sRn.p, &sRn.cap, (size_t)nrows * 8))) goto out;
    if ((rc = dev_realloc(ctx, &sCn.p, &sCn.cap, (size_t)ncols * 8))) goto out;
    if ((rc = dev_realloc(ctx, &sOut.p, &sOut.cap, (size_t)nrows * lld * 8))) goto out;
    rc = run_quadrature(ctx, 1, fs, keys, geo, nearflag, q, (double*)sQ.p, (size_t)nd);
    if (rc) goto out;
    {
        hipError_t e = hipMemcpyAsync(sT.p, table.data(), tsize * 4, hipMemcpyHostToDevice, st);
        int* dR = (int*)sR.p;
        int* dC = (int*)sC.p;
        if (e == hipSuccess) e = hipMemcpyAsync(dR, rarc.data(), (size_t)nrows * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dR + nrows, rphi.data(), (size_t)nrows * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dC, carc.data(), (size_t)ncols * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dC + ncols, cphi.data(), (size_t)ncols * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(sRn.p, rnb.data(), (size_t)nrows * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(sCn.p, cnb.data(), (size_t)ncols * 8, hipMemcpyHostToDevice, st);
        for (int r0 = 0; e == hipSuccess && r0 < nrows; r0 += 32768) {
            const int nr = std::min(32768, nrows - r0);
            hipLaunchKernelGGL(sph_assemble_kernel, dim3((ncols + 255) / 256, nr), dim3(256), 0, st, dR, dR + nrows,
                               (const double*)sRn.p, nrows, dC, dC + ncols, (const double*)sCn.p, ncols,
                               (const int*)sT.p, (int)clist.size(), Nphi, (const double*)sQ.p + (size_t)f * nd,
                               (double*)sOut.p, (size_t)lld, r0);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(s, (size_t)lld * 8, sOut.p, (size_t)lld * 8, (size_t)ncols * 8, nrows,
                                 hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "spherical block: %s", hipGetErrorString(e));
    }
out:
    ctx->sv = saved;
    return rc;
}

int opsec_sig_block(opsec_ctx* ctx, const opsec_cell* cells, int ncells, const int* rows, int nrows, const int* cols,
                    int ncols, int f, const opsec_survey* sv, const opsec_quad* q, double* s, int lld) {
    if (!ctx) return OPSEC_EINVAL;
    int rc = check_model(ctx, f);
    if (rc) return rc;
    if ((rc = check_quad(ctx, q))) return rc;
    OPSEC_CHECK_ARG(ctx, cells && ncells > 0, "no cells");
    OPSEC_CHECK_ARG(ctx, nrows >= 0 && ncols >= 0, "negative block size");
    if (lld <= 0) lld = ncols;
    OPSEC_CHECK_ARG(ctx, lld >= ncols, "ComputeSignalMatrixBlock: invalid array stride (%d,%d,%d)", lld, nrows, ncols);
    if (nrows == 0 || ncols == 0) return OPSEC_OK;
    OPSEC_CHECK_ARG(ctx, rows && cols && s, "NULL rows/cols/output");
    SurveyDev svd;
    if ((rc = survey_from_abi(ctx, sv, &svd))) return rc;
    for (int i = 0; i < nrows; i++) OPSEC_CHECK_ARG(ctx, rows[i] >= 0 && rows[i] < ncells, "row index out of range");
    for (int j = 0; j < ncols; j++) OPSEC_CHECK_ARG(ctx, cols[j] >= 0 && cols[j] < ncells, "column index out of range");
    for (int a = 0; a < ncells; a++) {
        const opsec_cell& c = cells[a];
        OPSEC_CHECK_ARG(ctx, c.d1 >= 0 && c.d1 < svd.n1 && c.d2 >= 0 && c.d2 < svd.n2 && c.d3 >= 0 && c.d3 < svd.n3,
                        "cell %d has grid index outside the survey grid", a);
    }
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (svd.coordsys == OPSEC_COORD_SPHERICAL)
        return sig_block_sph(ctx, cells, ncells, rows, nrows, cols, ncols, f, svd, q, s, lld);
    SurveyDev saved = ctx->sv;
    ctx->sv = svd;
    const int nk = svd.n1 * svd.n2 * svd.n3;

    std::vector<int> idx(3 * (size_t)ncells);
    std::vector<double> nb(ncells);
    for (int a = 0; a < ncells; a++) {
        idx[3 * a] = cells[a].d1;
        idx[3 * a + 1] = cells[a].d2;
        idx[3 * a + 2] = cells[a].d3;
        nb[a] = cells[a].Nbar;
    }
    Scratch s_idx, s_nb, s_rows, s_cols, s_first, s_Q, s_out;
    if ((rc = dev_realloc(ctx, &s_idx.p, &s_idx.cap, idx.size() * 4))) goto done;
    if ((rc = dev_realloc(ctx, &s_nb.p, &s_nb.cap, nb.size() * 8))) goto done;
    if ((rc = dev_realloc(ctx, &s_rows.p, &s_rows.cap, (size_t)nrows * 4))) goto done;
    if ((rc = dev_realloc(ctx, &s_cols.p, &s_cols.cap, (size_t)ncols * 4))) goto done;
    if ((rc = dev_realloc(ctx, &s_first.p, &s_first.cap, (size_t)nk * 8))) goto done;
    if ((rc = dev_realloc(ctx, &s_Q.p, &s_Q.cap, (size_t)nk * ctx->funcs.size() * 8))) goto done;
    if ((rc = dev_realloc(ctx, &s_out.p, &s_out.cap, (size_t)nrows * lld * 8))) goto done;
    {
        hipStream_t st = ctx->stream;
#define TRY(call)                                                                                       \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            rc = set_error(ctx, OPSEC_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));             \
            goto done;                                                                                  \
        }                                                                                               \
    } while (0)
        TRY(hipMemcpyAsync(s_idx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(s_nb.p, nb.data(), nb.size() * 8, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(s_rows.p, rows, (size_t)nrows * 4, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(s_cols.p, cols, (size_t)ncols * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(fill_u64_kernel, dim3((nk + 255) / 256), dim3(256), 0, st, (unsigned long long*)s_first.p,
                           (size_t)nk, ~0ull);
        for (int r0 = 0; r0 < nrows; r0 += 32768) {
            int nr = std::min(32768, nrows - r0);
            hipLaunchKernelGGL(first_pair_kernel, dim3((ncols + 255) / 256, nr), dim3(256), 0, st, (const int*)s_idx.p,
                               (const int*)s_rows.p + r0, nr, (const int*)s_cols.p, ncols, svd.n2, svd.n3, r0,
                               (unsigned long long*)s_first.p);
        }
        TRY(hipGetLastError());
        std::vector<unsigned long long> first(nk);
        TRY(hipMemcpyAsync(first.data(), s_first.p, (size_t)nk * 8, hipMemcpyDeviceToHost, st));
        TRY(hipStreamSynchronize(st));
        std::vector<int> keys;
        std::vector<PairGeo> geo;
        std::vector<int> nearflag;
        for (int key = 0; key < nk; key++) {
            if (first[key] == ~0ull) continue;
            int i = (int)(first[key] / (unsigned long long)ncols), j = (int)(first[key] % (unsigned long long)ncols);
            int a = rows[i], b = cols[j];
            const opsec_cell& ca = cells[std::min(a, b)];
            const opsec_cell& cb = cells[std::max(a, b)];
            keys.push_back(key);
            geo.push_back(pair_geo(ca, cb));
            nearflag.push_back(near_of(ca, cb, q->near));
        }
        int fs[1] = {f};
        if (q->method == OPSEC_QUAD_SURROGATE) {
            rc = set_error(ctx, OPSEC_EINVAL, "opsec_sig_block: use the device-resident path for the surrogate");
            goto done;
        }
        rc = run_quadrature(ctx, 1, fs, keys, geo, nearflag, q, (double*)s_Q.p, (size_t)nk);
        if (rc) goto done;
        hipLaunchKernelGGL(assemble_kernel, dim3((ncols + 511) / 512, (nrows + 7) / 8), dim3(256), 0, st,
                           (const int*)s_idx.p, (const double*)s_nb.p, (const int*)s_rows.p, nrows,
                           (const int*)s_cols.p, ncols, pair_key_of(svd), (const double*)s_Q.p + (size_t)f * nk,
                           (double*)s_out.p, (size_t)lld);
        TRY(hipGetLastError());
        // copy rows back, only the ncols entries of each row (the rest of s is untouched)
        TRY(hipMemcpy2DAsync(s, (size_t)lld * 8, s_out.p, (size_t)lld * 8, (size_t)ncols * 8, nrows,
                             hipMemcpyDeviceToHost, st));
        TRY(hipStreamSynchronize(st));
#undef TRY
    }
done:
    ctx->sv = saved;
    return rc;
}

int opsec_pair_integral(opsec_ctx* ctx, const opsec_cell* c1, const opsec_cell* c2, int f, const opsec_survey* sv,
                        const opsec_quad* q, double* Q, double* err) {
    if (!ctx || !c1 || !c2 || !Q) return OPSEC_EINVAL;
    int rc = check_model(ctx, f);
    if (rc) return rc;
    if ((rc = check_quad(ctx, q))) return rc;
    OPSEC_CHECK_ARG(ctx, q->method != OPSEC_QUAD_SURROGATE, "single integrals use GM or GAUSS");
    SurveyDev svd;
    if ((rc = survey_from_abi(ctx, sv, &svd))) return rc;
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    SurveyDev saved = ctx->sv;
    ctx->sv = svd;
    std::vector<int> keys{0};
    std::vector<PairGeo> geo{pair_geo(*c1, *c2)};
    std::vector<int> nearflag{near_of(*c1, *c2, q->near)};
    Scratch sQ;
    size_t nfun = ctx->funcs.size();
    if ((rc = dev_realloc(ctx, &sQ.p, &sQ.cap, 2 * nfun * 8))) {
        ctx->sv = saved;
        return rc;
    }
    OPSEC_HIP_TRY(ctx, hipMemsetAsync(sQ.p, 0, 2 * nfun * 8, ctx->stream));
    int fs[1] = {f};
    rc = run_quadrature(ctx, 1, fs, keys, geo, nearflag, q, (double*)sQ.p, 1, (double*)sQ.p + nfun);
    ctx->sv = saved;
    if (rc) return rc;
    double qe[2];
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&qe[0], (double*)sQ.p + f, 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(&qe[1], (double*)sQ.p + nfun + f, 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    *Q = qe[0];
    if (err) *err = qe[1];
    return OPSEC_OK;
}

// Shell pair tables: the deep sweep beside the shallow one (1) or after it (0).
int opsec_diag_set_gm_overlap(int v) {
    if (v != 0 && v != 1) return OPSEC_EINVAL;
    g_gm_overlap = v;
    return OPSEC_OK;
}

// Wavefront slots of the deep rerun (0: as many as memory allows, up to eight
// per CU).
int opsec_diag_set_gm_deep_slots(int v) {
    if (v < 0) return OPSEC_EINVAL;
    g_gm_deep_slots = v;
    return OPSEC_OK;
}

// Shell band sweeps: first regions with shared band-edge terms (1) or one
// thread per (pair, band) job (0).
int opsec_diag_set_gm_bands(int v) {
    if (v != 0 && v != 1) return OPSEC_EINVAL;
    g_gm_bands = v;
    return OPSEC_OK;
}

}  // extern "C"
