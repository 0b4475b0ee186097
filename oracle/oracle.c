/* oracle.c -- FP64 CPU restatement of OPSEC's hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; the product never links it.
 * Each function names the reference lines (relative to code/opsec/) whose
 * behaviour it restates; see oracle.h for how each one is pinned.
 * Compiled by gcc without -march flags, i.e. without FMA contraction, as the
 * reference is.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* splines                                                                   */

/* Interval of x in increasing X (src/Spline.cpp:13-37): proportional guess,
 * then bisection until X[i] <= x <= X[i+1]. */
static int find_interval(double x, int N, const double* X) {
    double I = (N - 1) * (x - X[0]) / (X[N - 1] - X[0]);
    int i = (int)I, lo = 0, hi = N - 2;
    while (lo < hi) {
        if (x < X[i]) {
            hi = i - 1;
            i = (lo + hi) / 2;
        } else if (x > X[i + 1]) {
            lo = i + 1;
            i = (lo + hi) / 2;
        } else
            break;
    }
    return i;
}

/* Forsythe-Malcolm-Moler cubic spline coefficients (src/Spline.cpp:258-333). */
void orc_fmm_spline(int n, const double* x, const double* y, double* b, double* c, double* d) {
    int i;
    if (n < 2) return;
    if (n == 2) {
        b[0] = b[1] = (y[1] - y[0]) / (x[1] - x[0]);
        c[0] = c[1] = d[0] = d[1] = 0;
        return;
    }
    d[0] = x[1] - x[0];
    c[1] = (y[1] - y[0]) / d[0];
    for (i = 1; i < n - 1; i++) {
        d[i] = x[i + 1] - x[i];
        b[i] = 2 * (d[i - 1] + d[i]);
        c[i + 1] = (y[i + 1] - y[i]) / d[i];
        c[i] = c[i + 1] - c[i];
    }
    b[0] = -d[0];
    b[n - 1] = -d[n - 2];
    c[0] = c[n - 1] = 0;
    if (n > 3) {
        double c0 = c[2] / (x[3] - x[1]) - c[1] / (x[2] - x[0]);
        double cn = c[n - 2] / (x[n - 1] - x[n - 3]) - c[n - 3] / (x[n - 2] - x[n - 4]);
        c[0] = c0 * d[0] * d[0] / (x[3] - x[0]);
        c[n - 1] = -cn * d[n - 2] * d[n - 2] / (x[n - 1] - x[n - 4]);
    }
    for (i = 1; i < n; i++) {
        double t = d[i - 1] / b[i - 1];
        b[i] = b[i] - t * d[i - 1];
        c[i] = c[i] - t * c[i - 1];
    }
    c[n - 1] = c[n - 1] / b[n - 1];
    for (i = n - 2; i >= 0; i--) c[i] = (c[i] - d[i] * c[i + 1]) / b[i];
    b[n - 1] = (y[n - 1] - y[n - 2]) / d[n - 2] + d[n - 2] * (c[n - 2] + 2 * c[n - 1]);
    for (i = 0; i < n - 1; i++) {
        b[i] = (y[i + 1] - y[i]) / d[i] - d[i] * (c[i + 1] + 2 * c[i]);
        d[i] = (c[i + 1] - c[i]) / d[i];
        c[i] = 3 * c[i];
    }
    c[n - 1] = 3 * c[n - 1];
    d[n - 1] = d[n - 2];
}

static double cubic_at(int n, const double* x, const double* y, const double* b, const double* c, const double* d,
                       double xe) {
    if (xe < x[0] || xe >= x[n - 1]) return 0.0;   /* Spline::Evaluate, src/Spline.cpp:416-421 */
    int i = find_interval(xe, n, x);
    double u = xe - x[i];
    return y[i] + b[i] * u + c[i] * u * u + d[i] * u * u * u;
}

void orc_cubic_eval(int n, const double* x, const double* y, const double* b, const double* c, const double* d,
                    int m, const double* xe, double* ye) {
    for (int i = 0; i < m; i++) ye[i] = cubic_at(n, x, y, b, c, d, xe[i]);
}

typedef struct {
    int kind;
    int n;
    const double *x, *y;
    double *b, *c, *d;
    double kmin, kmax;
} pfun;

static double pfun_at(const pfun* P, double k) {
    if (P->kind == 1) return (k < P->kmin || k >= P->kmax) ? 0.0 : 1.0;
    return cubic_at(P->n, P->x, P->y, P->b, P->c, P->d, k);
}

static double sj0(double x) {
    double x2 = x * x;
    if (x2 < 1e-4) return 1 - x2 * (1 / 6. + x2 * (1 / 120. - x2 * (1 / 5040.)));
    return sin(x) / x;
}

void orc_xilm(int kind, int n, const double* px, const double* py, int nr, const double* r, double* xi, int Nk,
              double kmin, double kmax, double kdamp) {
    pfun P = {kind, n, px, py, NULL, NULL, NULL, kmin, kmax};
    if (kind == 0) {
        P.b = malloc(sizeof(double) * n);
        P.c = malloc(sizeof(double) * n);
        P.d = malloc(sizeof(double) * n);
        orc_fmm_spline(n, px, py, P.b, P.c, P.d);
    }
    double dk = (kmax - kmin) / Nk;
    double* k = malloc(sizeof(double) * (Nk + 1));
    double* w = malloc(sizeof(double) * (Nk + 1));
    for (int j = 0; j <= Nk; j++) {
        k[j] = kmin + j * dk;
        double sw = 2 + 2 * (j % 2) - (j == 0) - (j == Nk);   /* Simpson 1,4,2,...,4,1 */
        sw *= pfun_at(&P, k[j]) * (k[j] * k[j]) * exp(-k[j] * k[j] / (kdamp * kdamp));
        sw *= (dk / 3) / (2 * M_PI * M_PI);
        w[j] = sw;
    }
#pragma omp parallel for
    for (int i = 0; i < nr; i++) {
        double s = 0;
        for (int j = 0; j <= Nk; j++) s += w[j] * sj0(k[j] * r[i]);
        xi[i] = s;
    }
    free(k);
    free(w);
    if (kind == 0) {
        free(P.b);
        free(P.c);
        free(P.d);
    }
}

double orc_band_param(int n, const double* px, const double* py, double kmin, double kmax) {
    double *b = malloc(8 * n), *c = malloc(8 * n), *d = malloc(8 * n);
    orc_fmm_spline(n, px, py, b, c, d);
    double sum = 0;
    for (int i = 0; i < 256; i++) sum += cubic_at(n, px, py, b, c, d, kmin + (i + 0.5) * (kmax - kmin) / 256);
    free(b);
    free(c);
    free(d);
    return sum / 256;
}

void orc_qspline_grid(double* r) {
    for (int i = 0; i < 512; i++) r[i] = i * 200.0 / 512;
    for (int j = 0; j < 256; j++) r[512 + j] = 200.0 * exp(j * log(20000.0 / 200.0) / 255);
}

static double linear_at(double x, int N, const double* X, const double* Y) {
    if (x < X[0] || x >= X[N - 1]) return 0.0;
    int i = find_interval(x, N, X);
    return Y[i] + (Y[i + 1] - Y[i]) * (x - X[i]) / (X[i + 1] - X[i]);
}

double orc_qspline_eval(const double* r768, const double* y768, double r) {
    /* src/QSpline.cpp:61-70 */
    double u[256];
    if (r < 0) return 0;
    if (r < r768[512]) return linear_at(r, 513, r768, y768);
    if (r < r768[767]) {
        for (int j = 0; j < 256; j++) u[j] = log(r768[512 + j]);
        return linear_at(log(r), 256, u, y768 + 512);
    }
    return y768[767];
}

/* (1/2pi^2) int_{klo}^{khi} k^2 j0(kr) dk = [sin kr - kr cos kr]/(2 pi^2 r^3) */
static double g3(double x) {   /* (sin x - x cos x)/x^3 */
    if (x < 0.5) {
        double x2 = x * x, t = 1.0 / 3.0, term = 1.0;
        /* series sum_n (-1)^{n+1} 2n x^{2n-2} / (2n+1)! */
        double s = 0;
        double fact = 6.0;   /* (2n+1)! for n=1 */
        for (int n = 1; n <= 9; n++) {
            if (n > 1) fact *= (2.0 * n) * (2.0 * n + 1);
            double c = ((n % 2) ? 1.0 : -1.0) * 2.0 * n / fact;
            s += c * term;
            term *= x2;
        }
        (void)t;
        return s;
    }
    return (sin(x) - x * cos(x)) / (x * x * x);
}

double orc_band_xi(double klo, double khi, double r) {
    return (khi * khi * khi * g3(khi * r) - klo * klo * klo * g3(klo * r)) / (2 * M_PI * M_PI);
}

double orc_xi_eval(const orc_xi* f, double r) {
    switch (f->kind) {
        case 0: return orc_qspline_eval(f->r768, f->y768, r);
        case 1: return orc_band_xi(f->a, f->b, r);
        case 2: return f->a * exp(-f->b * r);
        case 3: return f->a * exp(-f->b * r * r);
        case 4: return f->a * r * exp(-f->b * r * r);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* geometry (src/BoxSurvey.cpp:26-37, src/SphericalSurvey.cpp:20-23)          */

static double wrap(double dx, double L) {
    if (dx < -L / 2) return dx + L;
    if (dx >= L / 2) return dx - L;
    return dx;
}

double orc_box_r(const double* p1, const double* p2, double L) {
    double x = wrap(p2[0] - p1[0], L), y = wrap(p2[1] - p1[1], L), z = wrap(p2[2] - p1[2], L);
    return sqrt(x * x + y * y + z * z);
}

static double sph_sep(double r1, double mu1, double ph1, double r2, double mu2, double ph2) {
    double cg = sqrt((1 - mu1 * mu1) * (1 - mu2 * mu2)) * cos(ph1 - ph2) + mu1 * mu2;
    return sqrt(fmax(r1 * r1 + r2 * r2 - 2 * r1 * r2 * cg, 0.0));
}

/* src/SphericalSurvey.cpp:19-23 */
double orc_sph_r(const double* p1, const double* p2) { return sph_sep(p1[0], p1[1], p1[2], p2[0], p2[1], p2[2]); }

/* ------------------------------------------------------------------------ */
/* Genz-Malik 7/5 rule and the adaptive driver                               */

typedef struct {
    int ndim;                  /* 6: a cell pair's xi; 3: a selection function over one cell */
    const orc_xi* f;
    const orc_selection* sel;
    int coordsys;
    double L;
    double lo[6], hi[6];
} integrand_t;

static double integrand(const integrand_t* I, const double* u) {
    if (I->ndim == 3) return orc_selection_eval(I->sel, I->coordsys, u);
    if (I->coordsys == 1) {
        double p1[3], p2[3];
        for (int d = 0; d < 3; d++) {
            p1[d] = (1 - u[d]) * I->lo[d] + u[d] * I->hi[d];
            p2[d] = (1 - u[3 + d]) * I->lo[3 + d] + u[3 + d] * I->hi[3 + d];
        }
        return orc_xi_eval(I->f, orc_box_r(p1, p2, I->L));
    }
    double v[6];
    for (int d = 0; d < 6; d++) v[d] = (1 - u[d]) * I->lo[d] + u[d] * I->hi[d];
    return orc_xi_eval(I->f, sph_sep(cbrt(v[0]), v[2], v[4], cbrt(v[1]), v[3], v[5]));
}

typedef struct {
    double c[6], h[6];
    double val, err;
    int split;
} region_t;

/* One application of the degree-7 rule with degree-5 error estimate
 * (Genz & Malik 1980) on region R; nodes summed in the order centre,
 * (+-l2, +-l4 per axis), (+-l4,+-l4 pairs), (+-l5)^6 corners. */
static void gm_rule(const integrand_t* I, region_t* R) {
    const double l2 = sqrt(9.0 / 70.0), l4 = sqrt(9.0 / 10.0), l5 = sqrt(9.0 / 19.0);
    const int n = I->ndim;
    double u[6], v0;
    int i, j, g;
    memcpy(u, R->c, sizeof u);
    v0 = integrand(I, u);
    double s2 = 0, s3 = 0, s4 = 0, s5 = 0, best = 0;
    int split = 0;
    for (i = 0; i < n; i++) {
        double a[4];
        double r2 = R->h[i] * l2, r4 = R->h[i] * l4;
        u[i] = R->c[i] - r2; a[0] = integrand(I, u);
        u[i] = R->c[i] + r2; a[1] = integrand(I, u);
        u[i] = R->c[i] - r4; a[2] = integrand(I, u);
        u[i] = R->c[i] + r4; a[3] = integrand(I, u);
        u[i] = R->c[i];
        s2 += a[0] + a[1];
        s3 += a[2] + a[3];
        double dd = 0.0 + fabs(a[0] + a[1] - 2 * v0 - (l2 * l2) / (l4 * l4) * (a[2] + a[3] - 2 * v0));
        if (dd > best) {
            best = dd;
            split = i;
        }
    }
    for (i = 0; i < n - 1; i++)
        for (j = i + 1; j < n; j++) {
            double ri = R->h[i] * l4, rj = R->h[j] * l4;
            u[i] = R->c[i] - ri; u[j] = R->c[j] - rj; s4 += integrand(I, u);
            u[i] = R->c[i] + ri; s4 += integrand(I, u);
            u[j] = R->c[j] + rj; s4 += integrand(I, u);
            u[i] = R->c[i] - ri; s4 += integrand(I, u);
            u[i] = R->c[i]; u[j] = R->c[j];
        }
    for (g = 0; g < (1 << n); g++) {
        int signs = g ^ (g >> 1);
        for (i = 0; i < n; i++) {
            double r5 = R->h[i] * l5;
            u[i] = ((signs >> i) & 1) ? R->c[i] - r5 : R->c[i] + r5;
        }
        s5 += integrand(I, u);
    }
    double vol = 1;
    for (i = 0; i < n; i++) vol *= 2 * R->h[i];
    double w1 = (12824.0 - 9120 * n + 400 * n * n) / 19683, w2 = 980.0 / 6561, w3 = (1820.0 - 400 * n) / 19683;
    double w4 = 200.0 / 19683, w5 = 6859.0 / 19683 / (1 << n);
    double e1 = (729.0 - 950 * n + 50 * n * n) / 729, e2 = 245.0 / 486, e3 = (265.0 - 100 * n) / 1458, e4 = 25.0 / 729;
    double r7 = vol * (w1 * v0 + w2 * s2 + w3 * s3 + w4 * s4 + w5 * s5);
    double r5 = vol * (e1 * v0 + e2 * s2 + e3 * s3 + e4 * s4);
    R->val = r7;
    R->err = fabs(r5 - r7);
    R->split = split;
}

/* max-heap of regions keyed on err */
typedef struct {
    region_t* a;
    int n, cap;
    double tv, te;
} rheap;

static void rh_push(rheap* H, region_t r) {
    H->tv += r.val;
    H->te += r.err;
    if (H->n == H->cap) {
        H->cap = H->cap ? 2 * H->cap : 64;
        H->a = realloc(H->a, sizeof(region_t) * H->cap);
    }
    int k = H->n++;
    while (k > 0) {
        int p = (k - 1) / 2;
        if (r.err <= H->a[p].err) break;
        H->a[k] = H->a[p];
        k = p;
    }
    H->a[k] = r;
}

static region_t rh_pop(rheap* H) {
    region_t top = H->a[0];
    region_t last = H->a[--H->n];
    int k = 0;
    H->a[0] = last;
    for (;;) {
        int ch = 2 * k + 1, big;
        if (ch >= H->n) break;
        big = (H->a[ch].err <= H->a[k].err) ? k : ch;
        if (ch + 1 < H->n && H->a[big].err < H->a[ch + 1].err) big = ch + 1;
        if (big == k) break;
        region_t t = H->a[k];
        H->a[k] = H->a[big];
        H->a[big] = t;
        k = big;
    }
    H->tv -= top.val;
    H->te -= top.err;
    return top;
}

static int small_enough(double val, double err, double absTol, double relTol) {
    return err <= absTol || (val == 0 ? HUGE_VAL : fabs(err / val)) <= relTol;
}

/* The adaptive driver (src/cubature.c:885-998) from the first region R0. */
static double gm_adapt(const integrand_t* I, region_t R0, double absTol, double relTol, unsigned maxeval,
                       double* err, int64_t* neval);

double orc_gm_integrate(const orc_xi* f, int coordsys, double L, const orc_cell* c1, const orc_cell* c2,
                        double epsrel, double epsabs, unsigned maxeval, double* err, int64_t* neval) {
    integrand_t I;
    I.ndim = 6;
    I.sel = NULL;
    I.f = f;
    I.coordsys = coordsys;
    I.L = L;
    if (coordsys == 1) {
        double l1[3] = {c1->min1, c1->min2, c1->min3}, h1[3] = {c1->max1, c1->max2, c1->max3};
        double l2[3] = {c2->min1, c2->min2, c2->min3}, h2[3] = {c2->max1, c2->max2, c2->max3};
        for (int d = 0; d < 3; d++) {
            I.lo[d] = l1[d]; I.hi[d] = h1[d];
            I.lo[3 + d] = l2[d]; I.hi[3 + d] = h2[d];
        }
    } else {
        I.lo[0] = c1->min1 * c1->min1 * c1->min1; I.hi[0] = c1->max1 * c1->max1 * c1->max1;
        I.lo[1] = c2->min1 * c2->min1 * c2->min1; I.hi[1] = c2->max1 * c2->max1 * c2->max1;
        I.lo[2] = c1->min2; I.hi[2] = c1->max2;
        I.lo[3] = c2->min2; I.hi[3] = c2->max2;
        I.lo[4] = c1->min3; I.hi[4] = c1->max3;
        I.lo[5] = c2->min3; I.hi[5] = c2->max3;
    }
    region_t R0;
    for (int d = 0; d < 6; d++) {
        R0.c[d] = 0.5;
        R0.h[d] = 0.5;
    }
    /* the reference passes (epsrel, epsabs) into (absolute, relative) */
    return gm_adapt(&I, R0, epsrel, epsabs, maxeval, err, neval);
}

static double gm_adapt(const integrand_t* I, region_t R0, double absTol, double relTol, unsigned maxeval,
                       double* err, int64_t* neval) {
    const int n = I->ndim;
    const unsigned npts = 1 + 4 * n + 2 * n * (n - 1) + (1u << n);   /* 149 in 6-D, 33 in 3-D */
    rheap H = {0};
    gm_rule(I, &R0);
    rh_push(&H, R0);
    uint64_t ne = npts;
    region_t* batch = NULL;
    int bcap = 0;
    while (ne < maxeval || !maxeval) {
        if (small_enough(H.tv, H.te, absTol, relTol)) break;
        int nb = 0;
        double rem = H.te, tot = H.tv;
        do {
            if (nb + 2 > bcap) {
                bcap = 2 * (nb + 2);
                batch = realloc(batch, sizeof(region_t) * bcap);
            }
            region_t r = rh_pop(&H);
            rem -= r.err;
            int d = r.split;
            region_t a = r, b = r;
            a.h[d] = b.h[d] = r.h[d] * 0.5;
            a.c[d] = r.c[d] - a.h[d];
            b.c[d] = r.c[d] + a.h[d];
            batch[nb++] = a;
            batch[nb++] = b;
            ne += 2 * npts;
            if (small_enough(tot, rem, absTol, relTol)) break;
        } while (H.n > 0 && (ne < maxeval || !maxeval));
        for (int t = 0; t < nb; t++) gm_rule(I, &batch[t]);
        for (int t = 0; t < nb; t++) rh_push(&H, batch[t]);
    }
    double v = 0, e = 0;
    for (int i = 0; i < H.n; i++) {
        v += H.a[i].val;
        e += H.a[i].err;
    }
    free(H.a);
    free(batch);
    if (err) *err = e;
    if (neval) *neval = (int64_t)ne;
    return v;
}

/* ------------------------------------------------------------------------ */
/* Gauss rules                                                               */

/* Legendre P_n and derivative by the three-term recurrence; Newton from
 * Chebyshev-like guesses. */
void orc_gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < n; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) {
                p1 = z;
                p0 = 1;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
        }
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2 / ((1 - z * z) * dp * dp);
    }
}

/* Gauss rule for weight (1-t) on [0,1]: nodes are the zeros of the Jacobi
 * polynomial P_n^{(1,0)}(2t-1); Newton on the three-term recurrence, weights
 * from the Christoffel numbers. */
static void jacobi10(int n, double x, double* p, double* dp) {
    /* P_k^{(1,0)}: (2k(k+1)(2k-1)) P_k = (2k)(2k-1)... use general recurrence */
    const double a = 1, b = 0;
    double p0 = 1, p1 = 0.5 * (a - b + (a + b + 2) * x);
    double d0 = 0, d1 = 0.5 * (a + b + 2);
    if (n == 0) {
        *p = 1;
        *dp = 0;
        return;
    }
    for (int k = 2; k <= n; k++) {
        double c = 2 * k + a + b;
        double a1 = 2 * k * (k + a + b) * (c - 2);
        double a2 = (c - 1) * (a * a - b * b);
        double a3 = (c - 2) * (c - 1) * c;
        double a4 = 2 * (k + a - 1) * (k + b - 1) * c;
        double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
        double d2 = ((a2 + a3 * x) * d1 + a3 * p1 - a4 * d0) / a1;
        p0 = p1;
        p1 = p2;
        d0 = d1;
        d1 = d2;
    }
    *p = p1;
    *dp = d1;
}

void orc_gauss_jacobi10(int n, double* t, double* w) {
    double* xs = malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) {
        double z = -cos(M_PI * (i + 0.75) / (n + 0.5));
        /* deflated Newton (avoid converging to a previous root) */
        for (int it = 0; it < 200; it++) {
            double p, dp;
            jacobi10(n, z, &p, &dp);
            double s = 0;
            for (int j = 0; j < i; j++) s += 1 / (z - xs[j]);
            double dz = p / (dp - s * p);
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        xs[i] = z;
    }
    /* sort ascending */
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (xs[j] < xs[i]) {
                double tt = xs[i];
                xs[i] = xs[j];
                xs[j] = tt;
            }
    /* weights for (1-x) on [-1,1]: w_i = 2^{a+b+1} Gamma(n+a+1)Gamma(n+b+1) /
     * (Gamma(n+a+b+1) n!) / ((1-x^2) P'_n(x)^2) with a=1, b=0 -> 4/((1-x^2)P'^2) */
    for (int i = 0; i < n; i++) {
        double p, dp;
        jacobi10(n, xs[i], &p, &dp);
        double wx = 4.0 / ((1 - xs[i] * xs[i]) * dp * dp);
        t[i] = 0.5 * (xs[i] + 1);
        w[i] = 0.25 * wx;
    }
    free(xs);
}

static void tent_nodes(int n, double* s, double* w) {
    int m = n / 2;
    double t[32], wt[32];
    orc_gauss_jacobi10(m, t, wt);
    for (int k = 0; k < m; k++) {
        s[k] = -t[m - 1 - k];
        w[k] = wt[m - 1 - k];
        s[m + k] = t[k];
        w[m + k] = wt[k];
    }
}

static double tent_sum(const orc_xi* f, double L, const orc_cell* c1, const orc_cell* c2, int n, const double* s,
                       const double* w) {
    double D[3], h[3];
    D[0] = 0.5 * (c2->min1 + c2->max1) - 0.5 * (c1->min1 + c1->max1);
    D[1] = 0.5 * (c2->min2 + c2->max2) - 0.5 * (c1->min2 + c1->max2);
    D[2] = 0.5 * (c2->min3 + c2->max3) - 0.5 * (c1->min3 + c1->max3);
    h[0] = c1->max1 - c1->min1;
    h[1] = c1->max2 - c1->min2;
    h[2] = c1->max3 - c1->min3;
    double acc = 0;
    for (int a = 0; a < n; a++) {
        double rx = wrap(D[0] + s[a] * h[0], L), ay = 0;
        for (int b = 0; b < n; b++) {
            double ry = wrap(D[1] + s[b] * h[1], L), az = 0;
            for (int c = 0; c < n; c++) {
                double rz = wrap(D[2] + s[c] * h[2], L);
                az += w[c] * orc_xi_eval(f, sqrt(rx * rx + ry * ry + rz * rz));
            }
            ay += w[b] * az;
        }
        acc += w[a] * ay;
    }
    return acc;
}

double orc_tent_gauss(const orc_xi* f, double L, const orc_cell* c1, const orc_cell* c2, int n) {
    double s[64], w[64];
    tent_nodes(n, s, w);
    return tent_sum(f, L, c1, c2, n, s, w);
}

double orc_tent_composite(const orc_xi* f, double L, const orc_cell* c1, const orc_cell* c2, int panels, int q) {
    double x[32], wl[32], s[256], w[256];
    orc_gauss_legendre(q, x, wl);
    int m = panels * q;
    double t[128], wt[128];
    for (int p = 0; p < panels; p++)
        for (int k = 0; k < q; k++) {
            double a = (double)p / panels, b = (double)(p + 1) / panels;
            double tt = 0.5 * (a + b) + 0.5 * (b - a) * x[k];
            t[p * q + k] = tt;
            wt[p * q + k] = 0.5 * (b - a) * wl[k] * (1 - tt);
        }
    for (int k = 0; k < m; k++) {
        s[k] = -t[m - 1 - k];
        w[k] = wt[m - 1 - k];
        s[m + k] = t[k];
        w[m + k] = wt[k];
    }
    return tent_sum(f, L, c1, c2, 2 * m, s, w);
}

double orc_brute6(const orc_xi* f, double L, const orc_cell* c1, const orc_cell* c2, int n) {
    double x[64], w[64];
    orc_gauss_legendre(n, x, w);
    double acc = 0;
    double lo1[3] = {c1->min1, c1->min2, c1->min3}, hi1[3] = {c1->max1, c1->max2, c1->max3};
    double lo2[3] = {c2->min1, c2->min2, c2->min3}, hi2[3] = {c2->max1, c2->max2, c2->max3};
    int idx[6];
    long total = 1;
    for (int d = 0; d < 6; d++) total *= n;
    for (long t = 0; t < total; t++) {
        long q = t;
        double wt = 1, p1[3], p2[3];
        for (int d = 0; d < 6; d++) {
            idx[d] = (int)(q % n);
            q /= n;
            wt *= 0.5 * w[idx[d]];
        }
        for (int d = 0; d < 3; d++) {
            p1[d] = lo1[d] + 0.5 * (x[idx[d]] + 1) * (hi1[d] - lo1[d]);
            p2[d] = lo2[d] + 0.5 * (x[idx[3 + d]] + 1) * (hi2[d] - lo2[d]);
        }
        acc += wt * orc_xi_eval(f, orc_box_r(p1, p2, L));
    }
    return acc;
}

/* ------------------------------------------------------------------------ */
/* translation-cache assembly (src/sig.cpp:82-123)                            */

int orc_sig_block(const orc_xi* f, const orc_cell* cells, int ncells, int Nx, int Ny, int Nz, double L,
                  const int* rows, int nrows, const int* cols, int ncols, double* s, int lld, int method,
                  double epsrel, double epsabs, unsigned maxeval, int nodes, int nearc) {
    (void)ncells;
    if (lld <= 0) lld = ncols;
    if (lld < ncols) return -1;
    const double unset = -1e100;
    double* cache = malloc(sizeof(double) * (size_t)Nx * Ny * Nz);
    for (long k = 0; k < (long)Nx * Ny * Nz; k++) cache[k] = unset;
    for (int i = 0; i < nrows; i++) {
        const orc_cell* ca = &cells[rows[i]];
        for (int j = 0; j < ncols; j++) {
            const orc_cell* cb = &cells[cols[j]];
            int kx = abs(ca->d1 - cb->d1), ky = abs(ca->d2 - cb->d2), kz = abs(ca->d3 - cb->d3);
            long k = ((long)kx * Ny + ky) * Nz + kz;
            if (cache[k] == unset) {
                const orc_cell* lo = rows[i] <= cols[j] ? ca : cb;
                const orc_cell* hi = rows[i] <= cols[j] ? cb : ca;
                if (method == 0)
                    cache[k] = orc_gm_integrate(f, 1, L, lo, hi, epsrel, epsabs, maxeval, NULL, NULL);
                else {
                    int m = kx > ky ? kx : ky;
                    m = m > kz ? m : kz;
                    cache[k] = (m < nearc) ? orc_tent_composite(f, L, lo, hi, 4, 8) : orc_tent_gauss(f, L, lo, hi, nodes);
                }
            }
            s[(size_t)i * lld + j] = sqrt(ca->Nbar * cb->Nbar) * cache[k];
        }
    }
    free(cache);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* projection (src/comma-scalapack.cpp:245-265)                              */

void orc_project(int n, const double* S, int m, const double* B, double* P) {
    double* C = malloc(sizeof(double) * (size_t)n * m);   /* C = S B, column-major n x m */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < m; j++) {
        const double* bj = B + (size_t)j * n;
        double* cj = C + (size_t)j * n;
        for (int a = 0; a < n; a++) {
            const double* srow = S + (size_t)a * n;
            double acc = 0;
            for (int b = 0; b < n; b++) acc += srow[b] * bj[b];
            cj[a] = acc;
        }
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < m; j++) {
        const double* cj = C + (size_t)j * n;
        for (int i = 0; i <= j; i++) {
            const double* bi = B + (size_t)i * n;
            double acc = 0;
            for (int a = 0; a < n; a++) acc += bi[a] * cj[a];
            P[(size_t)j * (j + 1) / 2 + i] = acc;
        }
    }
    free(C);
}

/* ------------------------------------------------------------------------ */
/* dense helpers                                                             */

/* Gauss-Jordan inverse with partial pivoting; returns 0 or -1 if singular. */
int orc_inverse(int n, const double* A, double* B) {
    double* M = malloc(sizeof(double) * n * 2 * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 2 * n; j++) M[i * 2 * n + j] = j < n ? A[i * n + j] : (j - n == i);
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (fabs(M[r * 2 * n + c]) > fabs(M[p * 2 * n + c])) p = r;
        if (M[p * 2 * n + c] == 0) {
            free(M);
            return -1;
        }
        if (p != c)
            for (int j = 0; j < 2 * n; j++) {
                double t = M[c * 2 * n + j];
                M[c * 2 * n + j] = M[p * 2 * n + j];
                M[p * 2 * n + j] = t;
            }
        double piv = M[c * 2 * n + c];
        for (int j = 0; j < 2 * n; j++) M[c * 2 * n + j] /= piv;
        for (int r = 0; r < n; r++)
            if (r != c) {
                double fct = M[r * 2 * n + c];
                if (fct != 0)
                    for (int j = 0; j < 2 * n; j++) M[r * 2 * n + j] -= fct * M[c * 2 * n + j];
            }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) B[i * n + j] = M[i * 2 * n + n + j];
    free(M);
    return 0;
}

void orc_sym_eigen(int n, const double* A0, double* evals, double* Vout) {
    double* A = malloc(sizeof(double) * n * n);
    double* V = malloc(sizeof(double) * n * n);
    memcpy(A, A0, sizeof(double) * n * n);
    for (int i = 0; i < n * n; i++) V[i] = 0;
    for (int i = 0; i < n; i++) V[i * n + i] = 1;
    for (int sweep = 0; sweep < 200; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                double t = (th >= 0 ? 1 : -1) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort descending; output eigenvectors as rows */
    int* idx = malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (A[idx[j] * n + idx[j]] > A[idx[i] * n + idx[i]]) {
                int t = idx[i];
                idx[i] = idx[j];
                idx[j] = t;
            }
    for (int k = 0; k < n; k++) {
        evals[k] = A[idx[k] * n + idx[k]];
        if (Vout)
            for (int i = 0; i < n; i++) Vout[k * n + i] = V[i * n + idx[k]];
    }
    free(idx);
    free(A);
    free(V);
}

/* principal square root of a symmetric positive (semi)definite matrix */
void orc_sqrtm(int n, const double* A, double* B) {
    double* ev = malloc(sizeof(double) * n);
    double* V = malloc(sizeof(double) * n * n);
    orc_sym_eigen(n, A, ev, V);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += V[k * n + i] * sqrt(ev[k]) * V[k * n + j];
            B[i * n + j] = s;
        }
    free(ev);
    free(V);
}

/* ------------------------------------------------------------------------ */
/* estimate (src/estimate-mpi.cpp:480-654)                                    */

int orc_estimate(int M, int np, const double* lambda, const double* D, const double* y, const double* prior,
                 int mixing, double* phat, double* cov, double* F, double* W, double* Mm, double* q, double* f) {
    long K = (long)M * (M + 1) / 2;
    double* Ci = malloc(sizeof(double) * M);
    double* R = malloc(sizeof(double) * K);
    for (int i = 0; i < M; i++) Ci[i] = 1.0 / (1.0 + lambda[i]);
    /* packed index k = j(j+1)/2 + i, i <= j */
    for (int j = 0; j < M; j++)
        for (int i = 0; i <= j; i++) {
            long k = (long)j * (j + 1) / 2 + i;
            double r = (i == j) ? 1.0 + lambda[i] : 0.0;
            for (int n = 0; n < np; n++) r -= prior[n] * D[(size_t)n * K + k];
            R[k] = r;
        }
    for (int m = 0; m < np; m++)
        for (int n = 0; n <= m; n++) {
            const double *Dm = D + (size_t)m * K, *Dn = D + (size_t)n * K;
            double s = 0;
            for (int j = 0; j < M; j++)
                for (int i = 0; i <= j; i++) {
                    long k = (long)j * (j + 1) / 2 + i;
                    s += ((i == j) ? 0.5 : 1.0) * Ci[i] * Dm[k] * Ci[j] * Dn[k];
                }
            F[m * np + n] = F[n * np + m] = s;
        }
    for (int n = 0; n < np; n++) {
        const double* Dn = D + (size_t)n * K;
        double sf = 0, sq = 0;
        for (int j = 0; j < M; j++)
            for (int i = 0; i <= j; i++) {
                long k = (long)j * (j + 1) / 2 + i;
                double mult = (i == j) ? 0.5 : 1.0;
                sf += mult * Ci[i] * Dn[k] * Ci[j] * R[k];
                sq += mult * y[i] * Ci[i] * Dn[k] * Ci[j] * y[j];
            }
        f[n] = sf;
        q[n] = sq;
    }
    free(Ci);
    free(R);
    /* mixing matrix */
    if (mixing == 0) {
        if (orc_inverse(np, F, Mm)) return -1;
    } else if (mixing == 1) {
        double* T = malloc(sizeof(double) * np * np);
        if (orc_inverse(np, F, T)) {
            free(T);
            return -1;
        }
        orc_sqrtm(np, T, Mm);
        free(T);
    } else {
        for (int i = 0; i < np * np; i++) Mm[i] = 0;
        for (int i = 0; i < np; i++) Mm[i * np + i] = 1;
    }
    /* W = M F, rows of M scaled so rows of W sum to one, W recomputed */
    for (int pass = 0; pass < 2; pass++) {
        for (int i = 0; i < np; i++)
            for (int j = 0; j < np; j++) {
                double s = 0;
                for (int k = 0; k < np; k++) s += Mm[i * np + k] * F[k * np + j];
                W[i * np + j] = s;
            }
        if (pass == 0)
            for (int i = 0; i < np; i++) {
                double rs = 0;
                for (int j = 0; j < np; j++) rs += W[i * np + j];
                for (int j = 0; j < np; j++) Mm[i * np + j] /= rs;
            }
    }
    for (int i = 0; i < np; i++) {
        double s = 0;
        for (int k = 0; k < np; k++) s += Mm[i * np + k] * (q[k] - f[k]);
        phat[i] = s;
    }
    /* cov = M F M^T */
    for (int i = 0; i < np; i++)
        for (int j = 0; j < np; j++) {
            double s = 0;
            for (int k = 0; k < np; k++) s += W[i * np + k] * Mm[j * np + k];
            cov[i * np + j] = s;
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* MLP 8 -> 128 -> 128 -> 128 -> 128 -> 1, SiLU hidden activations.
 * Weight layout: for each layer l: W_l (out x in, row-major) then b_l.      */

static const int mlp_dims[6] = {8, 128, 128, 128, 128, 1};

int orc_mlp_nweights(void) {
    int n = 0;
    for (int l = 0; l < 5; l++) n += mlp_dims[l + 1] * mlp_dims[l] + mlp_dims[l + 1];
    return n;
}

static double silu(double z) { return z / (1 + exp(-z)); }
static double dsilu(double z) {
    double s = 1 / (1 + exp(-z));
    return s * (1 + z * (1 - s));
}

void orc_mlp_forward(const double* w, int n, const double* x, double* out) {
    double a[128], z[128];
    for (int s = 0; s < n; s++) {
        const double* p = w;
        double in[128];
        for (int i = 0; i < 8; i++) in[i] = x[s * 8 + i];
        for (int l = 0; l < 5; l++) {
            int di = mlp_dims[l], dout = mlp_dims[l + 1];
            const double* W = p;
            const double* b = p + di * dout;
            for (int o = 0; o < dout; o++) {
                double acc = b[o];
                for (int i = 0; i < di; i++) acc += W[o * di + i] * in[i];
                z[o] = acc;
            }
            for (int o = 0; o < dout; o++) a[o] = (l < 4) ? silu(z[o]) : z[o];
            for (int o = 0; o < dout; o++) in[o] = a[o];
            p += di * dout + dout;
        }
        out[s] = in[0];
    }
}

double orc_mlp_loss_grad(const double* w, int n, const double* x, const double* t, double* grad) {
    int nw = orc_mlp_nweights();
    for (int i = 0; i < nw; i++) grad[i] = 0;
    double loss = 0;
    static double acts[6][128], pre[6][128];
    for (int s = 0; s < n; s++) {
        const double* p = w;
        for (int i = 0; i < 8; i++) acts[0][i] = x[s * 8 + i];
        int off[5];
        for (int l = 0; l < 5; l++) {
            int di = mlp_dims[l], dout = mlp_dims[l + 1];
            off[l] = (int)(p - w);
            for (int o = 0; o < dout; o++) {
                double acc = p[di * dout + o];
                for (int i = 0; i < di; i++) acc += p[o * di + i] * acts[l][i];
                pre[l + 1][o] = acc;
                acts[l + 1][o] = (l < 4) ? silu(acc) : acc;
            }
            p += di * dout + dout;
        }
        double e = acts[5][0] - t[s];
        loss += 0.5 * e * e;
        double delta[128], nd[128];
        delta[0] = e / n;
        for (int l = 4; l >= 0; l--) {
            int di = mlp_dims[l], dout = mlp_dims[l + 1];
            double* gW = grad + off[l];
            double* gb = gW + di * dout;
            const double* W = w + off[l];
            for (int o = 0; o < dout; o++) {
                gb[o] += delta[o];
                for (int i = 0; i < di; i++) gW[o * di + i] += delta[o] * acts[l][i];
            }
            if (l > 0) {
                for (int i = 0; i < di; i++) {
                    double acc = 0;
                    for (int o = 0; o < dout; o++) acc += W[o * di + i] * delta[o];
                    nd[i] = acc * dsilu(pre[l][i]);
                }
                for (int i = 0; i < di; i++) delta[i] = nd[i];
            }
        }
    }
    return loss / n;
}

/* ------------------------------------------------------------------------ */
/* Row block of the first projection product, C = S[rows, :] B (CPU baseline
 * sample of src/comma-scalapack.cpp:249).  S: R x n row-major, B: n x m
 * column-major, C: R x m column-major (ld R). */
void orc_gemm_rows(int R, int n, const double* S, int m, const double* B, double* C) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < m; j++) {
        const double* bj = B + (size_t)j * n;
        for (int a = 0; a < R; a++) {
            const double* srow = S + (size_t)a * n;
            double acc = 0;
            for (int b = 0; b < n; b++) acc += srow[b] * bj[b];
            C[(size_t)j * R + a] = acc;
        }
    }
}


/* ------------------------------------------------------------------------ */
/* spherical signal block (src/sig.cpp:125-213)                              */
/* Cells are grouped into arcs h = d_r * Nmu + d_mu (cell order within an arc);
 * for each (row arc, column arc) the first (row, column) pair met with a given
 * |d_phi_a - d_phi_b| is integrated (lower cell index first) and reused for
 * every pair of the two arcs with that |d_phi| difference. */
int orc_sig_block_sph(const orc_xi* f, const orc_cell* cells, int ncells, int Nr, int Nmu, int Nphi,
                      const int* rows, int nrows, const int* cols, int ncols, double* s, int lld, double epsrel,
                      double epsabs, unsigned maxeval) {
    if (lld <= 0) lld = ncols;
    if (lld < ncols) return -1;
    const double unset = -1e100;
    int* rg = malloc(sizeof(int) * ncells);
    int* cg = malloc(sizeof(int) * ncells);
    for (int a = 0; a < ncells; a++) rg[a] = cg[a] = -1;
    for (int i = 0; i < nrows; i++) rg[rows[i]] = i;
    for (int j = 0; j < ncols; j++) cg[cols[j]] = j;
    const int narcs = Nr * Nmu;
    int* count = calloc(narcs, sizeof(int));
    for (int a = 0; a < ncells; a++) count[cells[a].d1 * Nmu + cells[a].d2]++;
    int* start = malloc(sizeof(int) * (narcs + 1));
    start[0] = 0;
    for (int h = 0; h < narcs; h++) start[h + 1] = start[h] + count[h];
    int* members = malloc(sizeof(int) * ncells);
    int* fill = calloc(narcs, sizeof(int));
    for (int a = 0; a < ncells; a++) {
        int h = cells[a].d1 * Nmu + cells[a].d2;
        members[start[h] + fill[h]++] = a;
    }
    double* cache = malloc(sizeof(double) * Nphi);
    for (int ha = 0; ha < narcs; ha++)
        for (int hb = 0; hb < narcs; hb++) {
            for (int k = 0; k < Nphi; k++) cache[k] = unset;
            for (int ia = start[ha]; ia < start[ha + 1]; ia++) {
                int a = members[ia], i = rg[a];
                if (i < 0) continue;
                for (int ib = start[hb]; ib < start[hb + 1]; ib++) {
                    int b = members[ib], j = cg[b];
                    if (j < 0) continue;
                    int k = abs(cells[a].d3 - cells[b].d3);
                    if (cache[k] == unset)
                        cache[k] = (a <= b) ? orc_gm_integrate(f, 2, 0.0, &cells[a], &cells[b], epsrel, epsabs, maxeval, NULL, NULL)
                                            : orc_gm_integrate(f, 2, 0.0, &cells[b], &cells[a], epsrel, epsabs, maxeval, NULL, NULL);
                    s[(size_t)i * lld + j] = sqrt(cells[a].Nbar * cells[b].Nbar) * cache[k];
                }
            }
        }
    free(rg);
    free(cg);
    free(count);
    free(start);
    free(members);
    free(fill);
    free(cache);
    return 0;
}

/* Brute-force tensor Gauss-Legendre (n points per dimension) of the
 * spherical cell-pair mean, in the volume-uniform variables of
 * src/sig.cpp:436-462 (an independent check of the adaptive rule). */
double orc_brute6_sph(const orc_xi* f, const orc_cell* c1, const orc_cell* c2, int n) {
    double x[64], w[64];
    orc_gauss_legendre(n, x, w);
    double lo[6] = {c1->min1 * c1->min1 * c1->min1, c2->min1 * c2->min1 * c2->min1, c1->min2, c2->min2, c1->min3, c2->min3};
    double hi[6] = {c1->max1 * c1->max1 * c1->max1, c2->max1 * c2->max1 * c2->max1, c1->max2, c2->max2, c1->max3, c2->max3};
    double acc = 0;
    long tot = 1;
    for (int d = 0; d < 6; d++) tot *= n;
    for (long t = 0; t < tot; t++) {
        long q = t;
        double wt = 1, v[6];
        for (int d = 0; d < 6; d++) {
            int id = (int)(q % n);
            q /= n;
            double u = 0.5 * (x[id] + 1);
            wt *= 0.5 * w[id];
            v[d] = (1 - u) * lo[d] + u * hi[d];
        }
        acc += wt * orc_xi_eval(f, sph_sep(cbrt(v[0]), v[2], v[4], cbrt(v[1]), v[3], v[5]));
    }
    return acc;
}

/* ---- pixel values (the dot stage) ---- */

/* Survey::GetGridIndex (src/Survey.cpp:93-101).  The reference casts
 * floor(.) to int and then range-tests; testing the floored double gives the
 * same answer wherever that cast is defined, and sends NaN and values beyond
 * the int range outside, as the x86 conversion does. */
static int orc_grid_coord(double x, double lo, double hi, int n) {
    double t = floor(n * (x - lo) / (hi - lo));
    if (!(t >= 0 && t < n)) return -1;
    return (int)t;
}

int64_t orc_grid_index(const double* geom, const int* n, double x1, double x2, double x3) {
    int d1 = orc_grid_coord(x1, geom[0], geom[1], n[0]);
    int d2 = orc_grid_coord(x2, geom[2], geom[3], n[1]);
    int d3 = orc_grid_coord(x3, geom[4], geom[5], n[2]);
    if (d1 < 0 || d2 < 0 || d3 < 0) return -1;
    return ((int64_t)d1 * n[1] + d2) * n[2] + d3;
}

/* Weighted counts (src/dot.cpp:105-128): cellmap filled in cell order (the
 * last cell at a grid position wins), then every galaxy's weight added to its
 * cell in catalogue order.  Grid position 0 is inside (the reference's e > 0
 * is a defect) and a grid position with no cell goes to counts[ncells] (the
 * reference indexes counts[-1]).  counts: ncells + 1, accumulated into. */
void orc_count_galaxies(const orc_cell* cells, int ncells, const double* geom, const int* n,
                        const float* gals, int64_t ngals, double* counts) {
    int64_t ngrid = (int64_t)n[0] * n[1] * n[2];
    int* cellmap = (int*)malloc(sizeof(int) * (size_t)ngrid);
    for (int64_t e = 0; e < ngrid; e++) cellmap[e] = -1;
    for (int a = 0; a < ncells; a++) {
        int64_t e = ((int64_t)cells[a].d1 * n[1] + cells[a].d2) * n[2] + cells[a].d3;
        cellmap[e] = a;
    }
    for (int64_t g = 0; g < ngals; g++) {
        const float* p = gals + 4 * g;
        int64_t e = orc_grid_index(geom, n, p[0], p[1], p[2]);
        int a = (e >= 0 && cellmap[e] >= 0) ? cellmap[e] : ncells;
        counts[a] += p[3];
    }
    free(cellmap);
}

/* x_a = (N_a - Nbar_a)/sqrt(Nbar_a), y_i = sum_a B_ia x_a in index order
 * (src/dot.cpp:130-153); B is nmodes x ncells, mode-major. */
void orc_pixel_values(const orc_cell* cells, int ncells, int nmodes, const double* B,
                      const double* counts, double* x, double* y) {
    for (int a = 0; a < ncells; a++) x[a] = (counts[a] - cells[a].Nbar) / sqrt(cells[a].Nbar);
    for (int i = 0; i < nmodes; i++) {
        const double* b = B + (size_t)i * ncells;
        double s = 0;
        for (int a = 0; a < ncells; a++) s += b[a] * x[a];
        y[i] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* Basis cells: selection function and FinalizeCell (src/basis.cpp:89-190)   */

/* HEALPix (Gorski et al. 2005, ApJ 622, 759): z = cos(theta), t = phi/(pi/2)
 * in [0, 4).  In |z| <= 2/3 the pixel edges are the lines
 * n (1/2 + t) -+ (3/4) n z; in the caps the same lines on sqrt(3 (1 - |z|)). */
static double hp_t(double phi) {
    double p = fmod(phi, 2 * M_PI);
    if (p < 0) p += 2 * M_PI;
    double t = p / (M_PI / 2);
    return t >= 4 ? 0 : t;
}

int64_t orc_ang2pix_ring(int64_t n, double theta, double phi) {
    double z = cos(theta), za = fabs(z), t = hp_t(phi);
    if (za <= 2.0 / 3.0) {
        double up = n * (0.5 + t) - n * z * 0.75, dn = n * (0.5 + t) + n * z * 0.75;
        int64_t jp = (int64_t)up, jm = (int64_t)dn;
        int64_t ring = n + 1 + jp - jm;               /* 1 .. 2n+1 below z = 2/3 */
        int64_t odd = ring & 1;
        int64_t j = ((jp + jm - n + (1 - odd) + 1 + 8 * n) / 2) % (4 * n);
        return 2 * n * (n - 1) + (ring - 1) * 4 * n + j;
    }
    double frac = t - floor(t), w = n * sqrt(3 * (1 - za));
    int64_t jp = (int64_t)(frac * w), jm = (int64_t)((1 - frac) * w);
    int64_t ring = jp + jm + 1;
    int64_t j = ((int64_t)(t * ring)) % (4 * ring);
    if (z > 0) return 2 * ring * (ring - 1) + j;
    return 12 * n * n - 2 * ring * (ring + 1) + j;
}

static uint64_t interleave0(uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 32; b++) r |= ((v >> b) & 1) << (2 * b);
    return r;
}

static uint64_t deinterleave0(uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 32; b++) r |= ((v >> (2 * b)) & 1) << b;
    return r;
}

int64_t orc_ang2pix_nest(int64_t n, double theta, double phi) {
    double z = cos(theta), za = fabs(z), t = hp_t(phi);
    int64_t face, x, y;
    if (za <= 2.0 / 3.0) {
        int64_t jp = (int64_t)(n * (0.5 + t) - n * z * 0.75), jm = (int64_t)(n * (0.5 + t) + n * z * 0.75);
        int64_t fp = jp / n, fm = jm / n;
        if (fp == fm) face = fp | 4;
        else if (fp < fm) face = fp;
        else face = fm + 8;
        x = jm % n;
        y = n - (jp % n) - 1;
    } else {
        int64_t q = (int64_t)t;
        if (q > 3) q = 3;
        double frac = t - q, w = n * sqrt(3 * (1 - za));
        int64_t jp = (int64_t)(frac * w), jm = (int64_t)((1 - frac) * w);
        if (jp > n - 1) jp = n - 1;
        if (jm > n - 1) jm = n - 1;
        if (z >= 0) {
            face = q;
            x = n - jm - 1;
            y = n - jp - 1;
        } else {
            face = q + 8;
            x = jp;
            y = jm;
        }
    }
    return face * n * n + (int64_t)(interleave0((uint64_t)x) | (interleave0((uint64_t)y) << 1));
}

static int64_t isqrt64(int64_t v) {
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r > v) r--;
    while ((r + 1) * (r + 1) <= v) r++;
    return r;
}

void orc_pix2ang_ring(int64_t n, int64_t p, double* theta, double* phi) {
    int64_t npix = 12 * n * n, ncap = 2 * n * (n - 1);
    double z, ph;
    if (p < ncap) {
        int64_t i = (1 + isqrt64(1 + 2 * p)) / 2, j = p + 1 - 2 * i * (i - 1);
        z = 1 - (double)(i * i) / (3.0 * n * n);
        ph = (j - 0.5) * M_PI / (2.0 * i);
    } else if (p < npix - ncap) {
        int64_t q = p - ncap, i = q / (4 * n) + n, j = q % (4 * n) + 1;
        double shift = ((i + n) & 1) ? 1.0 : 0.5;
        z = (2 * n - i) * 2.0 / (3.0 * n);
        ph = (j - shift) * M_PI / (2.0 * n);
    } else {
        int64_t q = npix - p, i = (1 + isqrt64(2 * q - 1)) / 2, j = 4 * i + 1 - (q - 2 * i * (i - 1));
        z = -1 + (double)(i * i) / (3.0 * n * n);
        ph = (j - 0.5) * M_PI / (2.0 * i);
    }
    *theta = acos(z);
    *phi = ph;
}

void orc_pix2ang_nest(int64_t n, int64_t p, double* theta, double* phi) {
    static const int ring_of_face[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
    static const int col_of_face[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
    int64_t face = p / (n * n), q = p % (n * n);
    int64_t x = (int64_t)deinterleave0((uint64_t)q), y = (int64_t)deinterleave0((uint64_t)q >> 1);
    int64_t jr = ring_of_face[face] * n - x - y - 1, nr, shift;
    double z;
    if (jr < n) {
        nr = jr;
        z = 1 - (double)(nr * nr) / (3.0 * n * n);
        shift = 0;
    } else if (jr > 3 * n) {
        nr = 4 * n - jr;
        z = (double)(nr * nr) / (3.0 * n * n) - 1;
        shift = 0;
    } else {
        nr = n;
        z = (2 * n - jr) * 2.0 / (3.0 * n);
        shift = (jr - n) & 1;
    }
    int64_t jp = (col_of_face[face] * nr + x - y + 1 + shift) / 2;
    if (jp > 4 * n) jp -= 4 * n;
    if (jp < 1) jp += 4 * n;
    *theta = acos(z);
    *phi = (jp - (shift + 1) * 0.5) * (M_PI / 2) / nr;
}

/* src/BoxSurvey.cpp:59-75 (constant) and src/SphericalSurvey.cpp:65-82 with
 * r = cbrt(u) (src/basis.cpp:67-79) */
double orc_selection_eval(const orc_selection* s, int coordsys, const double* u) {
    (void)coordsys;
    if (s->kind == 0) return s->nbar;
    double r = cbrt(u[0]), mu = u[1], ph = u[2];
    double rad = linear_at(r, s->nrad, s->rr, s->rn);
    double ang = 1.0;
    if (s->nside > 0) {
        double th = acos(mu);
        int64_t ip = s->nest ? orc_ang2pix_nest(s->nside, th, ph) : orc_ang2pix_ring(s->nside, th, ph);
        ang = s->mask[ip];
    }
    return ang * rad;
}

static double cube3(double x) { return x * x * x; }

double orc_cell_integral(int coordsys, const orc_selection* s, const orc_cell* c, double epsrel, double epsabs,
                         unsigned maxeval, int64_t* neval) {
    integrand_t I;
    memset(&I, 0, sizeof I);
    I.ndim = 3;
    I.sel = s;
    I.coordsys = coordsys;
    double lo[3] = {c->min1, c->min2, c->min3}, hi[3] = {c->max1, c->max2, c->max3};
    if (coordsys == 2) {
        lo[0] = cube3(c->min1);
        hi[0] = cube3(c->max1);
    }
    region_t R0;
    memset(&R0, 0, sizeof R0);
    for (int d = 0; d < 3; d++) {
        R0.c[d] = 0.5 * (lo[d] + hi[d]);   /* src/cubature.c:165-177 */
        R0.h[d] = 0.5 * (hi[d] - lo[d]);
    }
    return gm_adapt(&I, R0, epsrel, epsabs, maxeval, NULL, neval);
}

int orc_basis(int coordsys, const double* g, const int* n, const orc_selection* s, double epsrel, double epsabs,
              orc_cell* out, int64_t* nevals) {
    const int64_t ng = (int64_t)n[0] * n[1] * n[2];
    orc_cell* all = malloc(sizeof(orc_cell) * (size_t)ng);
    int64_t* ne = malloc(sizeof(int64_t) * (size_t)ng);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < ng; i++) {
        int d1 = (int)(i / ((int64_t)n[1] * n[2])), d2 = (int)((i / n[2]) % n[1]), d3 = (int)(i % n[2]);
        orc_cell c;
        c.a = -1;
        c.d1 = d1;
        c.d2 = d2;
        c.d3 = d3;
        c.min1 = g[0] + d1 * (g[1] - g[0]) / n[0];
        c.max1 = g[0] + (d1 + 1) * (g[1] - g[0]) / n[0];
        c.min2 = g[2] + d2 * (g[3] - g[2]) / n[1];
        c.max2 = g[2] + (d2 + 1) * (g[3] - g[2]) / n[1];
        c.min3 = g[4] + d3 * (g[5] - g[4]) / n[2];
        c.max3 = g[4] + (d3 + 1) * (g[5] - g[4]) / n[2];
        c.Veff = coordsys == 2 ? (cube3(c.max1) - cube3(c.min1)) / 3 * (c.max2 - c.min2) * (c.max3 - c.min3)
                               : (c.max1 - c.min1) * (c.max2 - c.min2) * (c.max3 - c.min3);
        double N = orc_cell_integral(coordsys, s, &c, epsrel, epsabs, 5000000u, &ne[i]);
        c.Nbar = coordsys == 2 ? N / 3 : N;
        all[i] = c;
    }
    int kept = 0;
    for (int64_t i = 0; i < ng; i++)
        if (all[i].Nbar > 0) {
            out[kept] = all[i];
            out[kept].a = kept;
            if (nevals) nevals[kept] = ne[i];
            kept++;
        }
    free(all);
    free(ne);
    return kept;
}
