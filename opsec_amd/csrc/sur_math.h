// Closed-form pieces of the surrogate's analytic baseline (host and device).
//
// The cube-pair tent kernel's Fourier transform prod_i sinc^2(k_i a / 2) is
// cubic-invariant, so its expansion in spherical harmonics of the direction
// keeps only the cubic harmonics K_0 = 1, K_4, K_6, K_8, ... (one per degree
// up to 10).  On the unit sphere every cubic-invariant polynomial of degree
// <= 8 is a combination of 1, e2, e3, e2^2 with
//   e2 = x^2 y^2 + y^2 z^2 + z^2 x^2,   e3 = x^2 y^2 z^2,
// and K_l is the Gram-Schmidt orthogonalisation of that sequence.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace opsec {

constexpr int SUR_NL = 4;   // harmonics l = 0, 4, 6, 8

// Coefficients of K_0, K_4, K_6, K_8 over (1, e2, e3, e2^2).
struct SurHarm {
    double c[SUR_NL][4];
};

__host__ __device__ inline void sur_invariants(double x, double y, double z, double m[4]) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double e2 = x2 * y2 + y2 * z2 + z2 * x2;
    m[0] = 1.0;
    m[1] = e2;
    m[2] = x2 * y2 * z2;
    m[3] = e2 * e2;
}

// Spherical Bessel function j_l(x), x >= 0, by its power series
//   j_l(x) = x^l / (2l+1)!! sum_s (-x^2/2)^s / (s! (2l+3)(2l+5)...(2l+2s+1)),
// used for x <= 8 where the upward recurrence loses accuracy for l > x.
__host__ __device__ inline double sph_bessel_series(int l, double x) {
    double t = 1.0;
    for (int i = 1; i <= l; i++) t *= x / (2 * i + 1);
    double sum = t;
    const double h = -0.5 * x * x;
    for (int s = 1; s < 60; s++) {
        t *= h / (s * (2.0 * l + 2 * s + 1));
        sum += t;
        if (fabs(t) <= 1e-17 * fabs(sum)) break;
    }
    return sum;
}

// j_0, j_4, j_6, j_8 at x >= 0.
__host__ __device__ inline void sph_bessel_0468(double x, double j[SUR_NL]) {
    if (x > 8.0) {
        const double s = sin(x), c = cos(x);
        double jm = s / x, jl = s / (x * x) - c / x;   // j_0, j_1
        j[0] = jm;
        for (int l = 1; l < 8; l++) {                  // upward: stable for x > l
            const double jn = (2 * l + 1) / x * jl - jm;
            jm = jl;
            jl = jn;
            if (l + 1 == 4) j[1] = jl;
            if (l + 1 == 6) j[2] = jl;
        }
        j[3] = jl;
    } else {
        j[0] = (x < 1e-4) ? 1.0 - x * x / 6.0 : sin(x) / x;
        j[1] = sph_bessel_series(4, x);
        j[2] = sph_bessel_series(6, x);
        j[3] = sph_bessel_series(8, x);
    }
}

}  // namespace opsec
