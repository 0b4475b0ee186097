// MFMA GEMM building blocks for gfx950 (wave64).
//
//   C = A * B,  A: M x K, B: K x N, with the A operand either loaded from a
//   dense matrix or generated on the fly from the pair table of a box survey
//   (A[a][b] = Q[key(a,b)], the signal matrix without its sqrt(Nbar) factors),
//   so that C = T B' never reads an N x N matrix from HBM.
//
// Workgroup tile 128 x 128 x BK, 256 threads = 4 waves, each wave owning a
// 64 x 64 sub-tile as 4 x 4 MFMA fragments of 16 x 16.  Operand tiles are
// staged through LDS k-major (As[k][m], Bs[k][n]) so that the 16 lanes of a
// fragment row read consecutive words; the next tile's global loads are issued
// before the current tile's MFMAs (register double buffering).
//
// Fragment layouts (measured with an in-bounds probe on gfx950):
//   f64 16x16x4:    A lane l -> A[l%16][l/16],  B lane l -> B[l/16][l%16],
//                   D lane l, item r -> D[l/16 + 4r][l%16]
//   f32 16x16x4:    A, B as f64;                D lane l, item r -> D[4(l/16)+r][l%16]
//   bf16 16x16x32:  A lane l -> A[l%16][8(l/16)+e], B lane l -> B[8(l/16)+e][l%16],
//                   D as f32.
#pragma once

#include <cstdint>
#include <type_traits>
#include <utility>

#include "common.h"

namespace opsec {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

constexpr int GT_M = 128, GT_N = 128, GT_THREADS = 256;

// Row (within a 16 x 16 output fragment) of accumulator item r of `lane`.
template <typename T>
__device__ __forceinline__ int acc_row(int lane, int r) {
    return 4 * (lane >> 4) + r;
}
template <>
__device__ __forceinline__ int acc_row<double>(int lane, int r) {
    return (lane >> 4) + 4 * r;
}

// ---- operand sources -------------------------------------------------------
// Each source exposes  T get(m_or_n, k)  with bounds handled by the caller,
// and a flag telling which index is contiguous in memory.

// Eight consecutive doubles at q by four 16-byte loads; false (nothing loaded)
// when q is not 16-byte aligned.
__device__ __forceinline__ bool load8_f64(const double* q, double (&r)[8]) {
    if ((reinterpret_cast<uintptr_t>(q) & 15) != 0) return false;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const double2 v = reinterpret_cast<const double2*>(q)[e];
        r[2 * e] = v.x;
        r[2 * e + 1] = v.y;
    }
    return true;
}

struct SrcDenseRowMajor {      // X[i][k] = p[i*ld + k]   (k contiguous)
    const double* p;
    size_t ld;
    static constexpr bool k_contig = true;
    __device__ __forceinline__ double get(int i, int k) const { return p[(size_t)i * ld + k]; }
    __device__ __forceinline__ bool load8(int i, int k, double (&r)[8]) const {
        return load8_f64(p + (size_t)i * ld + k, r);
    }
};

struct SrcDenseColMajor {      // X[i][k] = p[k*ld + i]   (i contiguous)
    const double* p;
    size_t ld;
    static constexpr bool k_contig = false;
    __device__ __forceinline__ double get(int i, int k) const { return p[(size_t)k * ld + i]; }
};

struct SrcKColMajor {          // X[i][k] = p[i*ld + k], i.e. a column-major K x N operand read as [n][k]
    const double* p;
    size_t ld;
    static constexpr bool k_contig = true;
    __device__ __forceinline__ double get(int i, int k) const { return p[(size_t)i * ld + k]; }
    __device__ __forceinline__ bool load8(int i, int k, double (&r)[8]) const {
        return load8_f64(p + (size_t)i * ld + k, r);
    }
};

// Whether a k-contiguous source can load eight consecutive k at once (load8).
template <typename S, typename = void>
struct HasLoad8 : std::false_type {};
template <typename S>
struct HasLoad8<S, std::void_t<decltype(std::declval<const S&>().load8(0, 0, std::declval<double (&)[8]>()))>>
    : std::true_type {};

// Signal matrix generated from the pair table: box key (SPH false) or the
// arc-pair key of a spherical survey (SPH true, narcs = Nr Nmu).
template <bool SPH>
struct SrcPairTableT {
    const double* Q;           // pair table of one function
    const int* cellidx;        // ncells x 3 grid indices
    int Ny, Nz, narcs;
    static constexpr bool k_contig = true;
    __device__ __forceinline__ double get(int a, int b) const {
        const int* ca = cellidx + 3 * a;
        const int* cb = cellidx + 3 * b;
        int key;
        if constexpr (SPH)
            key = ((ca[0] * Ny + ca[1]) * narcs + cb[0] * Ny + cb[1]) * Nz + abs(ca[2] - cb[2]);
        else
            key = (abs(ca[0] - cb[0]) * Ny + abs(ca[1] - cb[1])) * Nz + abs(ca[2] - cb[2]);
        return Q[key];
    }
};
typedef SrcPairTableT<false> SrcPairTable;

}  // namespace opsec
