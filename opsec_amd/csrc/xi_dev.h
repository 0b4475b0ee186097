// Device-side evaluation of the xi functions and the separation functions.
#pragma once

#include "common.h"

namespace opsec {

// Interval search of the reference's linear splines (src/Spline.cpp:13-37):
// proportional first guess, then a bisection that stops as soon as
// X[i] <= x <= X[i+1].  Returns i in [0, N-2].
__device__ __forceinline__ int lookup_index(double x, int N, const double* X) {
    double I = (N - 1) * (x - X[0]) / (X[N - 1] - X[0]);
    int i = (int)I;
    int ilo = 0, ihi = N - 2;
    while (ilo < ihi) {
        if (x < X[i]) {
            ihi = i - 1;
            i = (ilo + ihi) / 2;
        } else if (x > X[i + 1]) {
            ilo = i + 1;
            i = (ilo + ihi) / 2;
        } else {
            break;
        }
    }
    return i;
}

// Linear spline value on [X[0], X[N-1]); zero outside (Spline::Evaluate,
// src/Spline.cpp:416-421 and LinearSplineImpl::y :145-150).
__device__ __forceinline__ double linear_spline(double x, int N, const double* X, const double* Y) {
    if (x < X[0] || x >= X[N - 1]) return 0.0;
    int i = lookup_index(x, N, X);
    double x1 = X[i], x2 = X[i + 1], y1 = Y[i], y2 = Y[i + 1];
    return y1 + __ddiv_rn(__dmul_rn(y2 - y1, x - x1), x2 - x1);
}

// QSpline (src/QSpline.cpp:61-70): 0 for r<0, linear in r below 200, linear in
// ln r up to 20000, constant beyond.
__device__ __forceinline__ double qspline(const XiModelDev& m, const double* y, double r) {
    if (r < 0) return 0.0;
    if (r < QS_X1) return linear_spline(r, QS_N1 + 1, m.xgrid, y);
    if (r < m.x2) return linear_spline(log(r), QS_N2, m.ugrid, y + QS_N1);
    return y[QS_N - 1];
}

// (sin x - x cos x)/x^3, stable for all x >= 0 (series below 0.5).
__device__ __forceinline__ double tophat_kernel(double x) {
    if (x < 0.5) {
        double x2 = x * x;
        // sum_{n>=1} (-1)^{n+1} 2n x^{2n-2}/(2n+1)!
        return 1.0 / 3.0 +
               x2 * (-1.0 / 30.0 +
                     x2 * (1.0 / 840.0 +
                           x2 * (-1.0 / 45360.0 +
                                 x2 * (1.0 / 3991680.0 +
                                       x2 * (-1.0 / 518918400.0 + x2 * (1.0 / 93405312000.0))))));
    }
    double s, c;
    sincos(x, &s, &c);
    return (s - x * c) / (x * x * x);
}

// Closed-form top-hat band: (1/2pi^2) int_{klo}^{khi} k^2 j0(kr) dk
// = [sin kr - kr cos kr]_{klo}^{khi} / (2 pi^2 r^3).  With klo r >= 0.5 both
// terms take the closed form over one division by r^3; below, k^3 times
// tophat_kernel (its series near 0).
__device__ __forceinline__ double band_xi(double klo, double khi, double r) {
    const double inv2pi2 = 0.050660591821168885;   // 1/(2 pi^2)
    const double xh = khi * r, xl = klo * r;
    if (xl >= 0.5) {
        double sh, ch, sl, cl;
        sincos(xh, &sh, &ch);
        sincos(xl, &sl, &cl);
        return inv2pi2 * ((sh - xh * ch) - (sl - xl * cl)) / (r * r * r);
    }
    return inv2pi2 * (khi * khi * khi * tophat_kernel(xh) - klo * klo * klo * tophat_kernel(xl));
}

__device__ __forceinline__ double xi_eval(const XiModelDev& m, const XiDesc& d, double r) {
    switch (d.kind) {
        case XI_TABLE:
            return qspline(m, m.y + (size_t)d.table * QS_N, r);
        case XI_BAND:
            return band_xi(d.a, d.b, r);
        case XI_TEST_EXP:
            return d.a * exp(-d.b * r);
        case XI_TEST_GAUSS:
            return d.a * exp(-d.b * r * r);
        case XI_TEST_DGAUSS:
            return d.a * r * exp(-d.b * r * r);
    }
    return 0.0;
}

// Periodic minimum-image component, translated into [-L/2, L/2)
// (src/BoxSurvey.cpp:26-30).
__device__ __forceinline__ double min_image(double dx, double L) {
    return dx + L * (double)((int)(dx < -L / 2) - (int)(dx >= L / 2));
}

__device__ __forceinline__ double box_r(double x1, double y1, double z1, double x2, double y2,
                                        double z2, double L) {
    double rx = min_image(x2 - x1, L);
    double ry = min_image(y2 - y1, L);
    double rz = min_image(z2 - z1, L);
    return sqrt(rx * rx + ry * ry + rz * rz);
}

// Spherical separation by the law of cosines (src/SphericalSurvey.cpp:19-46).
__device__ __forceinline__ double sph_r(double r1, double mu1, double phi1, double r2, double mu2,
                                        double phi2) {
    double cosg = sqrt((1 - mu1 * mu1) * (1 - mu2 * mu2)) * cos(phi1 - phi2) + mu1 * mu2;
    return sqrt(fmax(r1 * r1 + r2 * r2 - 2 * r1 * r2 * cosg, 0.0));
}

}  // namespace opsec
