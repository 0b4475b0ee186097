// HEALPix pixel index of a direction (theta, phi), RING and NESTED schemes,
// for the angular completeness maps of spherical surveys
// (src/SphericalSurvey.cpp:65-82 looks the map up by pixel).
//
// Written from the published construction (Gorski et al. 2005, ApJ 622, 759,
// section 4): with z = cos(theta) and t = phi / (pi/2) in [0, 4), the
// equatorial zone |z| <= 2/3 is cut by the two families of lines
// n (1/2 + t) -/+ (3/4) n z, whose integer parts jp, jm index the pixel edges;
// the polar caps use the same construction on sqrt(3 (1 - |z|)).
// Host and device code share these functions.
#pragma once

#include <cmath>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace opsec {

// Interleave the low 32 bits of v with zeros (bit b -> bit 2b).
__host__ __device__ __forceinline__ uint64_t hp_spread(uint64_t v) {
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

// t = phi / (pi/2) reduced to [0, 4).
__host__ __device__ __forceinline__ double hp_quarter_turns(double phi) {
    const double two_pi = 6.283185307179586476925286766559;
    double p = fmod(phi, two_pi);
    if (p < 0) p += two_pi;
    double t = p * (2.0 / 3.141592653589793238462643383280);
    return t >= 4.0 ? 0.0 : t;
}

// RING index: rings of constant z numbered from the north pole, pixels in a
// ring numbered eastwards from phi = 0.
__host__ __device__ __forceinline__ int64_t hp_ang2pix_ring(int64_t nside, double theta, double phi) {
    const double z = cos(theta), za = fabs(z);
    const double t = hp_quarter_turns(phi);
    if (za <= 2.0 / 3.0) {
        const double a = nside * (0.5 + t), b = nside * z * 0.75;
        const int64_t jp = (int64_t)(a - b), jm = (int64_t)(a + b);
        const int64_t ir = nside + 1 + jp - jm;   // ring below z = 2/3, 1 .. 2 nside + 1
        const int64_t shift = 1 - (ir & 1);
        const int64_t ip = ((jp + jm - nside + shift + 1 + 8 * nside) / 2) % (4 * nside);
        return 2 * nside * (nside - 1) + (ir - 1) * 4 * nside + ip;
    }
    const double tp = t - floor(t), w = nside * sqrt(3.0 * (1.0 - za));
    const int64_t jp = (int64_t)(tp * w), jm = (int64_t)((1.0 - tp) * w);
    const int64_t ir = jp + jm + 1;   // ring counted from the nearer pole
    const int64_t ip = ((int64_t)(t * ir)) % (4 * ir);
    return z > 0 ? 2 * ir * (ir - 1) + ip : 12 * nside * nside - 2 * ir * (ir + 1) + ip;
}

// NESTED index (nside a power of two): base face f in 0..11, then the
// bit-interleaved (x, y) position inside the face.
__host__ __device__ __forceinline__ int64_t hp_ang2pix_nest(int64_t nside, double theta, double phi) {
    const double z = cos(theta), za = fabs(z);
    const double t = hp_quarter_turns(phi);
    int64_t face, ix, iy;
    if (za <= 2.0 / 3.0) {
        const double a = nside * (0.5 + t), b = nside * z * 0.75;
        const int64_t jp = (int64_t)(a - b), jm = (int64_t)(a + b);
        const int64_t fp = jp / nside, fm = jm / nside;
        face = fp == fm ? (fp | 4) : (fp < fm ? fp : fm + 8);
        ix = jm & (nside - 1);
        iy = nside - (jp & (nside - 1)) - 1;
    } else {
        int64_t nt = (int64_t)t;
        if (nt > 3) nt = 3;
        const double tp = t - nt, w = nside * sqrt(3.0 * (1.0 - za));
        int64_t jp = (int64_t)(tp * w), jm = (int64_t)((1.0 - tp) * w);
        if (jp > nside - 1) jp = nside - 1;
        if (jm > nside - 1) jm = nside - 1;
        if (z >= 0) {
            face = nt;
            ix = nside - jm - 1;
            iy = nside - jp - 1;
        } else {
            face = nt + 8;
            ix = jp;
            iy = jm;
        }
    }
    return face * nside * nside + (int64_t)(hp_spread((uint64_t)ix) | (hp_spread((uint64_t)iy) << 1));
}

}  // namespace opsec
