// Projection of the band matrices onto the KL modes, P_p = B^T C_{,p} B.
//
// Replaces the comma band loop (src/comma-scalapack.cpp:228-266): per band
// p, fill S_p, C = S_p B (pdgemm), P = B^T C (pdgemm), write the packed upper
// triangle.  Here every band runs in one batched launch per stage (grid.z =
// band) and S_p is never stored:
//   stage 1  C_p = T_p B'      T_p[a][b] = Q_p[key(a,b)] generated in the A
//                              loader, B' = diag(sqrt(Nbar)) B (once per upload)
//   stage 2  P_p = B'^T C_p    upper-triangle tiles only, written packed
//                              (column j holds rows 0..j: the covfile layout).
// Precisions: FP64 (v_mfma_f64_16x16x4f64), FP32 (v_mfma_f32_16x16x4f32) and
// split-BF16x3 (x = hi + lo in bf16, hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation).
//
// The split-BF16x3 path runs per band on pre-split operand planes (bf16 hi
// and lo, k contiguous, zero padded to multiples of 128):
//   gen      S_p planes from the pair table (HBM-write bound, ~1 ms per band at
//            32^3 cells)
//   stage 1  C_p = S_p B'     NT GEMM, epilogue writes C_p already split
//   stage 2  P_p = B'^T C_p   NT GEMM on the same kernel, packed FP64 output
// B' is split once per mode upload.  Splitting outside the GEMM keeps its
// inner loop to 16-byte loads, LDS staging and MFMAs.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "common.h"
#include "gemm.h"

namespace opsec {

struct ProjState {
    double* d_Bw = nullptr;      // B' = diag(sqrt(Nbar)) B, ncells x nmodes column-major, pitch ldw
    size_t ldw = 0;              // ncells + 32 (a power-of-two pitch aliases L2 sets and HBM channels)
    size_t cap_Bw = 0;
    void* d_C = nullptr;         // stage-1 products, nb x ncells x nmodes (double or float)
    size_t cap_C = 0;
    void* d_P = nullptr;         // packed results, nb x M(M+1)/2 (double)
    size_t cap_P = 0;
    float* d_Ppart = nullptr;    // split-K partials of one band's stage 2 (FP32)
    size_t cap_Ppart = 0;
    int nb = 0;                  // bands in d_P
    int prec = 0;
    std::vector<int> funcs;
    float last_ms_stage1 = 0, last_ms_stage2 = 0;
    // split-BF16x3 planes: B' (Mp x Kp), one band's S_p (Kp x Kp), its C_p (Mp x Kp)
    __bf16* d_Bp = nullptr;      // hi plane then lo plane
    size_t cap_Bp = 0;
    bool bp_valid = false;
    __bf16* d_Sp = nullptr;
    size_t cap_Sp = 0;
    __bf16* d_Cp = nullptr;
    size_t cap_Cp = 0;
    unsigned* d_Qhl = nullptr;   // one band's pair table as (hi | lo << 16) words
    size_t cap_Qhl = 0;
    unsigned* d_cpk = nullptr;   // packed cell coordinates, padded to Kp
    size_t cap_cpk = 0;
    // per-kernel sums over the bands of the last split-BF16x3 projection (ms)
    float last_ms_gen = 0, last_ms_gemm1 = 0, last_ms_gemm2 = 0;
    int last_launches = 0;       // stage-1 launches (one per band)
    int last_launches2 = 0;      // stage-2 launches (one per band, or per group of bands)
    // Toeplitz projection (grid order, DFT along x of a box or phi of a shell)
    bool toep_valid = false;     // grid planes of B' and its DFT match the uploaded modes
    int* d_gmap = nullptr;       // grid index -> cell number, -1 where the grid has no cell
    size_t cap_gmap = 0;
    __bf16* d_Bg = nullptr;      // B' in grid order: Mp x Pg, hi plane then lo plane
    size_t cap_Bg = 0;
    __bf16* d_Bhat = nullptr;    // DFT of B' along x: nbat planes Mp x Gp (hi), then nbat (lo)
    size_t cap_Bhat = 0;
    __bf16* d_That = nullptr;    // one band's That_k, k = 0..n1: Gr x Gp each (hi), then (lo)
    size_t cap_That = 0;
    double* d_Qhat = nullptr;    // one band's Qhat_k(|dy|, |dz|) of a box: (n1 + 1) x G
    size_t cap_Qhat = 0;
    float* d_Yhat = nullptr;     // stage-1 products in frequency: M x nbat x G
    size_t cap_Yhat = 0;
    __bf16* d_Cg = nullptr;      // Y = S_p B' in grid order, one per band of a group: Mp x Pg, hi then lo
    size_t cap_Cg = 0;
    double* d_trig = nullptr;    // cos(2 pi j / L), then sin(2 pi j / L), j < L
    size_t cap_trig = 0;
    float* d_W = nullptr;        // inverse-DFT weights, transposed: nbat x n1p (n1p = n1 rounded up to 4)
    size_t cap_W = 0;
    // Toeplitz projection in FP64 / FP32 (same DFT along x, no operand split)
    bool toep64_valid = false;   // Bg64 and Bhat64 match the uploaded modes
    int* d_gmap64 = nullptr;     // grid index -> cell number (as d_gmap)
    size_t cap_gmap64 = 0;
    int* d_slab = nullptr;       // slab position r -> (y << 16) | z
    size_t cap_slab = 0;
    double* d_trig64 = nullptr;  // cos, then sin of 2 pi j / L
    size_t cap_trig64 = 0;
    double* d_W64 = nullptr;     // inverse-DFT weights W[x][b], n1 x nbat
    size_t cap_W64 = 0;
    double* d_Bg64 = nullptr;    // B' in grid order: M x Ng (pitch ldg), zero where the grid has no cell
    size_t cap_Bg64 = 0;
    double* d_Bhat64 = nullptr;  // DFT of B' along x: nbat x M x G ([b][m][r])
    size_t cap_Bhat64 = 0;
    double* d_Yhat64 = nullptr;  // one band's stage-1 products: nbat x M x G
    size_t cap_Yhat64 = 0;
    double* d_Y64 = nullptr;     // Y = S_p B' in grid order, up to four bands: M x Ng each (pitch ldg)
    size_t cap_Y64 = 0;
    double* d_That64 = nullptr;  // one band's That_k of a shell: (Nphi + 1) x G x G
    size_t cap_That64 = 0;
};

// ---- FP64 / FP32 MFMA GEMM -----------------------------------------------------

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef d4_t acc_t;
    __device__ __forceinline__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<float> {
    typedef f4_t acc_t;
    __device__ __forceinline__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// Tile loads: 128 (m or n) x 16 (k) elements, 8 per thread.
template <typename Src>
__device__ __forceinline__ void tile_fetch(const Src& s, int i0, int k0, int nI, int nK, double (&r)[8]) {
    const int t = threadIdx.x;
    if (Src::k_contig) {
        int i = i0 + (t >> 1), kb = k0 + (t & 1) * 8;
        // whole, aligned runs of 8 by 16-byte loads
        if constexpr (HasLoad8<Src>::value)
            if (i < nI && kb + 8 <= nK && s.load8(i, kb, r)) return;
#pragma unroll
        for (int e = 0; e < 8; e++) r[e] = (i < nI && kb + e < nK) ? s.get(i, kb + e) : 0.0;
    } else {
        int k = k0 + (t >> 4), ib = i0 + (t & 15) * 8;
#pragma unroll
        for (int e = 0; e < 8; e++) r[e] = (k < nK && ib + e < nI) ? s.get(ib + e, k) : 0.0;
    }
}

template <typename Src, typename T>
__device__ __forceinline__ void tile_store(T (*lds)[GT_M + 4], const double (&r)[8]) {
    const int t = threadIdx.x;
    if (Src::k_contig) {
        int i = t >> 1, kb = (t & 1) * 8;
#pragma unroll
        for (int e = 0; e < 8; e++) lds[kb + e][i] = (T)r[e];
    } else {
        int k = t >> 4, ib = (t & 15) * 8;
#pragma unroll
        for (int e = 0; e < 8; e++) lds[k][ib + e] = (T)r[e];
    }
}

// Epilogues
struct OutColMajor {             // C[n*ld + m]
    void* p;
    size_t ld;
    size_t batch_stride;
};

struct OutPackedUpper {          // P[n(n+1)/2 + m] for m <= n
    double* p;
    size_t batch_stride;
};

template <typename T, typename SrcA, typename SrcB, typename Out, bool UPPER>
__global__ __launch_bounds__(GT_THREADS) void gemm_kernel(SrcA A, size_t a_batch, SrcB B, size_t b_batch, Out O,
                                                          int M, int N, int K) {
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[16][GT_M + 4];
    __shared__ T Bs[16][GT_N + 4];
    const int m0 = blockIdx.x * GT_M, n0 = blockIdx.y * GT_N;
    if (UPPER && m0 > n0 + GT_N - 1) return;   // tile entirely below the diagonal
    // batch offsets
    A.p_shift(blockIdx.z, a_batch);
    B.p_shift(blockIdx.z, b_batch);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = acc_t{0, 0, 0, 0};

    double ra[8], rb[8];
    tile_fetch(A, m0, 0, M, K, ra);
    tile_fetch(B, n0, 0, N, K, rb);
    for (int k0 = 0; k0 < K; k0 += 16) {
        __syncthreads();
        tile_store<SrcA, T>(As, ra);
        tile_store<SrcB, T>(Bs, rb);
        __syncthreads();
        if (k0 + 16 < K) {
            tile_fetch(A, m0, k0 + 16, M, K, ra);
            tile_fetch(B, n0, k0 + 16, N, K, rb);
        }
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            T af[4], bf[4];
            const int kr = kk * 4 + (lane >> 4);
#pragma unroll
            for (int i = 0; i < 4; i++) af[i] = As[kr][wm + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; j++) bf[j] = Bs[kr][wn + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = Mfma<T>::mma(af[i], bf[j], acc[i][j]);
        }
    }
    // epilogue
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + wn + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + acc_row<T>(lane, r);
                O.put(blockIdx.z, m, n, M, N, (double)acc[i][j][r]);
            }
        }
}

// Sources with batch shifting (z index = band)
template <bool SPH>
struct GA_PairTableT : SrcPairTableT<SPH> {
    size_t qstride;
    const int* funcs;
    __device__ __forceinline__ void p_shift(int z, size_t) { this->Q += (size_t)funcs[z] * qstride; }
};
typedef GA_PairTableT<false> GA_PairTable;
typedef GA_PairTableT<true> GA_PairTableSph;
struct GA_RowMajor : SrcDenseRowMajor {
    __device__ __forceinline__ void p_shift(int z, size_t s) { p += (size_t)z * s; }
};
struct GB_ColMajor : SrcKColMajor {   // B[k][n] = p[n*ld + k]
    __device__ __forceinline__ void p_shift(int z, size_t s) { p += (size_t)z * s; }
};
struct GA_ColMajorT : SrcKColMajor {  // A = X^T with X column-major: A[m][k] = X[k][m] = p[m*ld + k]
    __device__ __forceinline__ void p_shift(int z, size_t s) { p += (size_t)z * s; }
};

template <typename T>
struct PutColMajor {
    T* p;
    size_t ld, batch;
    __device__ __forceinline__ void put(int z, int m, int n, int M, int N, double v) const {
        if (m < M && n < N) p[(size_t)z * batch + (size_t)n * ld + m] = (T)v;
    }
};

struct PutPacked {
    double* p;
    size_t batch;
    __device__ __forceinline__ void put(int z, int m, int n, int M, int N, double v) const {
        if (m < M && n < N && m <= n) p[(size_t)z * batch + (size_t)n * (n + 1) / 2 + m] = v;
    }
};

// Stage-2 source reading the stage-1 product stored as T (double or float).
template <typename T>
struct GB_ColMajorT {
    const T* p;
    size_t ld;
    static constexpr bool k_contig = true;
    __device__ __forceinline__ double get(int n, int k) const { return (double)p[(size_t)n * ld + k]; }
    __device__ __forceinline__ bool load8(int n, int k, double (&r)[8]) const {
        const T* q = p + (size_t)n * ld + k;
        if constexpr (std::is_same<T, double>::value) {
            return load8_f64(q, r);
        } else {
            if ((reinterpret_cast<uintptr_t>(q) & 15) != 0) return false;
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const float4 v = reinterpret_cast<const float4*>(q)[e];
                r[4 * e] = v.x;
                r[4 * e + 1] = v.y;
                r[4 * e + 2] = v.z;
                r[4 * e + 3] = v.w;
            }
            return true;
        }
    }
    __device__ __forceinline__ void p_shift(int z, size_t s) { p += (size_t)z * s; }
};

// ---- split-BF16x3 on pre-split planes --------------------------------------------
//
// A plane pair (hi, lo) holds X = hi + lo for a rows x ld row-major matrix in
// bf16 with k contiguous; ld and the row count are multiples of 256 and the
// padding is zero, so the GEMM needs no bounds checks.

constexpr int PT = 128;      // tile edge
constexpr int PK = 32;       // k per step: one 64-byte LDS row per tile row

// LDS slot of 16-byte chunk c (0..3) of tile row r: chunks are XOR-swizzled by
// ((r >> 1) ^ (r >> 2)) & 3, which makes both the staging writes (4 lanes per
// row) and the fragment reads (16 rows, one chunk) free of bank conflicts for
// 32 or 64 LDS banks.
__device__ __forceinline__ int lds_slot(int r, int c) { return r * PK + ((c ^ (((r >> 1) ^ (r >> 2)) & 3)) << 3); }

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Row pitch of a plane (bf16): the K extent plus 256 bytes.  A power-of-two
// pitch (64 KB at 32^3) maps every row of a tile onto the same L2 sets and
// HBM channels; the offset spreads them (measured: stage 1 at 32^3 x 4000
// modes, 26 ms -> 18.5 ms per band).  The padding columns are never read.
inline size_t plane_pitch(size_t k) { return k + 128; }

__device__ __forceinline__ void split2(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// planes[r][c] = split(X[r * ldx + c]) for r < rows, c < cols; zero elsewhere
// (r < rows_p, c < ld).  One thread per 8 consecutive columns.
__global__ __launch_bounds__(256) void split_plane_kernel(const double* __restrict__ X, size_t ldx, int rows, int cols,
                                                          __bf16* __restrict__ hi, __bf16* __restrict__ lo, size_t ld,
                                                          int rows_p) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per_row = ld / 8;
    const size_t r = t / per_row;
    if (r >= (size_t)rows_p) return;
    const int c0 = (int)(t % per_row) * 8;
    bf8_t h, l;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int c = c0 + e;
        const float v = (r < (size_t)rows && c < cols) ? (float)X[r * ldx + c] : 0.0f;
        __bf16 a, b;
        split2(v, a, b);
        h[e] = a;
        l[e] = b;
    }
    *(bf8_t*)(hi + r * ld + c0) = h;
    *(bf8_t*)(lo + r * ld + c0) = l;
}

// S_p planes from the pair table: S[a][b] = Q[key(a, b)] for a, b < N (the
// sqrt(Nbar) factors live in B').  One thread per 8 consecutive columns.
__global__ __launch_bounds__(256) void gen_split_kernel(const double* __restrict__ Q, const int* __restrict__ cellidx,
                                                        PairKey K, int N, __bf16* __restrict__ hi,
                                                        __bf16* __restrict__ lo, size_t ld, int rows_p) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per_row = ld / 8;
    const size_t a = t / per_row;
    if (a >= (size_t)rows_p) return;
    const int b0 = (int)(t % per_row) * 8;
    bf8_t h, l;
    int xa = 0, ya = 0, za = 0;
    if (a < (size_t)N) {
        xa = cellidx[3 * a];
        ya = cellidx[3 * a + 1];
        za = cellidx[3 * a + 2];
    }
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int b = b0 + e;
        float v = 0.0f;
        if (a < (size_t)N && b < N) {
            const int* cb = cellidx + 3 * b;
            v = (float)Q[K(xa, ya, za, cb[0], cb[1], cb[2])];
        }
        __bf16 x, y;
        split2(v, x, y);
        h[e] = x;
        l[e] = y;
    }
    *(bf8_t*)(hi + a * ld + b0) = h;
    *(bf8_t*)(lo + a * ld + b0) = l;
}

// Tile order: tiles are grouped 8 m-blocks x 8 n-blocks so that the blocks
// resident at one time share A and B slices in L2, and consecutive groups are
// dealt to the eight XCDs (workgroups go to XCDs round-robin) so that each XCD
// works through its own range of groups.
__device__ __forceinline__ void tile_of(int id, int nmb, int nnb, int G, int remap, int& mb, int& nbk) {
    const int total = nmb * nnb;
    if (remap && total % 8 == 0) id = (id % 8) * (total / 8) + id / 8;
    const int group = id / (G * nnb);
    const int first = group * G;
    const int gs = min(nmb - first, G);
    const int in = id % (G * nnb);
    mb = first + in % gs;
    nbk = in / gs;
}

// D[m][n] -> planes [n][m] (the stage-1 product C_p, cells contiguous), split.
struct EpiPlanes {
    __bf16* hi;
    __bf16* lo;
    size_t ld;
    __device__ __forceinline__ EpiPlanes part(int) const { return *this; }
    __device__ __forceinline__ EpiPlanes batch(int) const { return *this; }
    __device__ __forceinline__ void put4(int m, int n, const f4_t& v) const {
        typedef __bf16 bf4_t __attribute__((ext_vector_type(4)));
        bf4_t h, l;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            __bf16 x, y;
            split2(v[r], x, y);
            h[r] = x;
            l[r] = y;
        }
        *(bf4_t*)(hi + (size_t)n * ld + m) = h;
        *(bf4_t*)(lo + (size_t)n * ld + m) = l;
    }
    static constexpr bool upper = false;
};

// D[m][n] -> packed upper triangle, FP64 (m <= n < M).  With split-K, part s
// of the K range writes to p + s * part_stride.
struct EpiPacked {
    double* p;
    int M;
    size_t part_stride = 0;
    __device__ __forceinline__ EpiPacked part(int s) const {
        EpiPacked e = *this;
        e.p += (size_t)s * part_stride;
        return e;
    }
    __device__ __forceinline__ EpiPacked batch(int) const { return *this; }
    __device__ __forceinline__ void put4(int m, int n, const f4_t& v) const {
        if (n >= M) return;
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (m + r <= n) p[(size_t)n * (n + 1) / 2 + m + r] = (double)v[r];
    }
    static constexpr bool upper = true;
};

// Split-K partials of the packed upper triangle, FP32 (the accumulators'
// own precision, so nothing is lost); part s at p + s * part_stride, summed in
// FP64 by sum_parts_kernel.
struct EpiPackedParts {
    float* p;
    int M;
    size_t part_stride;
    size_t batch_stride = 0;   // batch b (one band) writes its parts at p + b * batch_stride
    __device__ __forceinline__ EpiPackedParts part(int s) const {
        EpiPackedParts e = *this;
        e.p += (size_t)s * part_stride;
        return e;
    }
    __device__ __forceinline__ EpiPackedParts batch(int b) const {
        EpiPackedParts e = *this;
        e.p += (size_t)b * batch_stride;
        return e;
    }
    __device__ __forceinline__ void put4(int m, int n, const f4_t& v) const {
        if (n >= M) return;
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (m + r <= n) p[(size_t)n * (n + 1) / 2 + m + r] = v[r];
    }
    static constexpr bool upper = true;
};

// D = A B^T on plane pairs: A (nmb*BM rows) and B (nnb*BN rows), both with
// K (multiple of 32) contiguous.  4 waves in a 2 x 2 arrangement, each owning
// FM x FN fragments of 16 x 16 (BM = 32 FM, BN = 32 FN), on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Operand tiles are staged
// through LDS (64-byte rows, swizzled chunks) with the next tile's global
// loads issued before the current tile's MFMAs.
template <typename Epi, int FM, int FN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void nt_bf16x3_kernel(const __bf16* __restrict__ Ah, const __bf16* __restrict__ Al,
                                                        size_t lda, const __bf16* __restrict__ Bh,
                                                        const __bf16* __restrict__ Bl, size_t ldb, int K, int nmb,
                                                        int nnb, int G, int remap, Epi E) {
    constexpr int BM = 32 * FM, BN = 32 * FN;
    constexpr int CA = BM * 4 / 256, CB = BN * 4 / 256;   // 16-byte chunks per thread and plane
    __shared__ __attribute__((aligned(16))) __bf16 sAh[BM * PK];
    __shared__ __attribute__((aligned(16))) __bf16 sAl[BM * PK];
    __shared__ __attribute__((aligned(16))) __bf16 sBh[BN * PK];
    __shared__ __attribute__((aligned(16))) __bf16 sBl[BN * PK];
    int mb, nbk;
    tile_of(blockIdx.x, nmb, nnb, G, remap, mb, nbk);
    const int m0 = mb * BM, n0 = nbk * BN;
    if (Epi::upper && m0 > n0 + BN - 1) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * (16 * FM), wn = (wave & 1) * (16 * FN);
    // chunk c of a plane tile: row c / 4, k offset (c % 4) * 8
    const int kc = (tid & 3) * 8, r0 = tid >> 2;
    const __bf16* pA = Ah + (size_t)(m0 + r0) * lda + kc;
    const __bf16* pAl = Al + (size_t)(m0 + r0) * lda + kc;
    const __bf16* pB = Bh + (size_t)(n0 + r0) * ldb + kc;
    const __bf16* pBl = Bl + (size_t)(n0 + r0) * ldb + kc;
    const size_t sa = (size_t)64 * lda, sb = (size_t)64 * ldb;   // 64 rows between a thread's chunks
    f4_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = f4_t{0, 0, 0, 0};
    bf8_t ga[2 * CA], gb[2 * CB];
#pragma unroll
    for (int c = 0; c < CA; c++) {
        ga[c] = *(const bf8_t*)(pA + c * sa);
        ga[CA + c] = *(const bf8_t*)(pAl + c * sa);
    }
#pragma unroll
    for (int c = 0; c < CB; c++) {
        gb[c] = *(const bf8_t*)(pB + c * sb);
        gb[CB + c] = *(const bf8_t*)(pBl + c * sb);
    }
    const int kq = lane >> 4, fr = lane & 15, kcw = tid & 3;
    for (int k0 = 0; k0 < K; k0 += PK) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < CA; c++) {
            const int o = lds_slot(r0 + 64 * c, kcw);
            *(bf8_t*)&sAh[o] = ga[c];
            *(bf8_t*)&sAl[o] = ga[CA + c];
        }
#pragma unroll
        for (int c = 0; c < CB; c++) {
            const int o = lds_slot(r0 + 64 * c, kcw);
            *(bf8_t*)&sBh[o] = gb[c];
            *(bf8_t*)&sBl[o] = gb[CB + c];
        }
        __syncthreads();
        if (k0 + PK < K) {
            const int o = k0 + PK;
#pragma unroll
            for (int c = 0; c < CA; c++) {
                ga[c] = *(const bf8_t*)(pA + c * sa + o);
                ga[CA + c] = *(const bf8_t*)(pAl + c * sa + o);
            }
#pragma unroll
            for (int c = 0; c < CB; c++) {
                gb[c] = *(const bf8_t*)(pB + c * sb + o);
                gb[CB + c] = *(const bf8_t*)(pBl + c * sb + o);
            }
        }
        bf8_t bh[FN], bl[FN];
#pragma unroll
        for (int j = 0; j < FN; j++) {
            const int o = lds_slot(wn + j * 16 + fr, kq);
            bh[j] = *(const bf8_t*)&sBh[o];
            bl[j] = *(const bf8_t*)&sBl[o];
        }
#pragma unroll
        for (int i = 0; i < FM; i++) {
            const int o = lds_slot(wm + i * 16 + fr, kq);
            const bf8_t ah = *(const bf8_t*)&sAh[o];
            const bf8_t al = *(const bf8_t*)&sAl[o];
            // three sweeps over independent accumulators (lo*hi, hi*lo, hi*hi):
            // consecutive MFMAs never wait on each other's result
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
            E.put4(m0 + wm + i * 16 + 4 * (lane >> 4), n0 + wn + j * 16 + fr, acc[i][j]);
}

// ---- stage 1 with S_p generated from the pair table ------------------------------
//
// The A operand of C_p = S_p B' is never stored: each of the 256 threads owns
// one row a of the 256-row tile and builds its 32 entries per k-step from the
// band's pair table, pre-split into 32-bit words (hi bf16 | lo bf16 << 16),
// Q_p[key(a, b)] with key = (|dx| Ny + |dy|) Nz + |dz|.  The table (128 KB at
// 32^3) stays in L1/L2; a 256 x 32 tile touches about 256 distinct keys.  Cell
// coordinates are packed z | y << 10 | x << 20 (padding columns 0xffffffff).
// When the 16 columns of a half-chunk are consecutive-z cells of one (x, y)
// line -- always so for a box in the reference's cell order -- the test is
// wave-uniform and each key costs one v_sad_u32: key = base + |za - zb|.

constexpr unsigned CELL_NONE = 0xffffffffu;

__global__ void pack_table_kernel(const double* __restrict__ Q, unsigned* __restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = (float)Q[i];
    const __bf16 h = (__bf16)v;
    const __bf16 l = (__bf16)(v - (float)h);
    out[i] = (unsigned)__builtin_bit_cast(unsigned short, h) | ((unsigned)__builtin_bit_cast(unsigned short, l) << 16);
}

__global__ void pack_cells_kernel(const int* __restrict__ cellidx, int N, int Kp, unsigned* __restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= Kp) return;
    out[a] = (a < N) ? ((unsigned)cellidx[3 * a + 2] | ((unsigned)cellidx[3 * a + 1] << 10) |
                        ((unsigned)cellidx[3 * a] << 20))
                     : CELL_NONE;
}

// hi halves / lo halves of two table words as one dword of two bf16
__device__ __forceinline__ unsigned pair_hi(unsigned w0, unsigned w1) { return (w0 & 0xffffu) | (w1 << 16); }
__device__ __forceinline__ unsigned pair_lo(unsigned w0, unsigned w1) { return (w0 >> 16) | (w1 & 0xffff0000u); }

// 16 table words for row (xa, ya, za) and the columns of c[0..15].
__device__ __forceinline__ void gather16(const unsigned* __restrict__ Qhl, const unsigned (&c)[16], unsigned xa,
                                         unsigned ya, unsigned za, bool rowv, int Ny, int Nz, unsigned* w) {
    // consecutive z within one (x, y) line, no carry out of the z field
    bool line = c[0] != CELL_NONE && (c[0] & 1023u) <= 1008u;
#pragma unroll
    for (int e = 1; e < 16; e++) line = line && (c[e] == c[0] + e);
    if (line) {
        const unsigned xb = c[0] >> 20, yb = (c[0] >> 10) & 1023u, zb = c[0] & 1023u;
        const unsigned base = (__sad((int)xa, (int)xb, 0u) * Ny + __sad((int)ya, (int)yb, 0u)) * Nz;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const unsigned key = rowv ? __sad((int)za, (int)(zb + e), base) : 0u;   // padding rows read key 0
            const unsigned v = Qhl[key];
            w[e] = rowv ? v : 0u;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const unsigned cb = c[e];
            const bool ok = rowv && cb != CELL_NONE;
            const unsigned key = ok ? (__sad((int)xa, (int)(cb >> 20), 0u) * Ny + __sad((int)ya, (int)((cb >> 10) & 1023u), 0u)) * Nz +
                                          __sad((int)za, (int)(cb & 1023u), 0u)
                                    : 0u;
            const unsigned v = Qhl[key];
            w[e] = ok ? v : 0u;
        }
    }
}

// S_p planes from the packed pair table.  Block (x, y) writes row y, columns
// 2048 x .. 2048 x + 2047 of the pitch; each lane takes column pairs 512 apart,
// so the coordinate loads, the table gathers (neighbouring columns have
// neighbouring keys) and the 4-byte stores of both planes coalesce across the
// wave.  Columns at or beyond Kp (the pitch padding) are written as zero.
__global__ __launch_bounds__(256) void gen_planes_kernel(const unsigned* __restrict__ Qhl,
                                                         const unsigned* __restrict__ cpk, int Ny, int Nz, int Kp,
                                                         __bf16* __restrict__ hi, __bf16* __restrict__ lo, size_t ld,
                                                         int row0) {
    const size_t a = (size_t)row0 + blockIdx.y;
    const unsigned ca = cpk[a];
    const bool rowv = ca != CELL_NONE;
    const int xa = (int)(ca >> 20), ya = (int)((ca >> 10) & 1023u), za = (int)(ca & 1023u);
    unsigned* ph = (unsigned*)(hi + a * ld);
    unsigned* pl = (unsigned*)(lo + a * ld);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int b = 2048 * blockIdx.x + 512 * e + 2 * threadIdx.x;
        if (b >= (int)ld) break;
        unsigned w0 = 0u, w1 = 0u;
        if (b < Kp) {
            const uint2 cb = *(const uint2*)(cpk + b);
            if (rowv && cb.x != CELL_NONE) {
                const unsigned key = (__sad(xa, (int)(cb.x >> 20), 0u) * Ny + __sad(ya, (int)((cb.x >> 10) & 1023u), 0u)) * Nz +
                                     __sad(za, (int)(cb.x & 1023u), 0u);
                w0 = Qhl[key];
            }
            if (rowv && cb.y != CELL_NONE) {
                const unsigned key = (__sad(xa, (int)(cb.y >> 20), 0u) * Ny + __sad(ya, (int)((cb.y >> 10) & 1023u), 0u)) * Nz +
                                     __sad(za, (int)(cb.y & 1023u), 0u);
                w1 = Qhl[key];
            }
        }
        ph[b >> 1] = pair_hi(w0, w1);
        pl[b >> 1] = pair_lo(w0, w1);
    }
}

// ---- 256 x 256 split-BF16x3 GEMM, double-buffered LDS ------------------------------
//
// D = A B^T with B from plane pairs and A either from plane pairs (ASrcPlanes)
// or generated from the pair table (ASrcTable).  512 threads = 8 waves in a
// 2 (m) x 4 (n) arrangement, each owning 8 x 4 fragments of 16 x 16; K in
// steps of 32 through two LDS stages of 64 KB, one barrier per step: the next
// step's global loads (or table gathers) are issued before the current step's
// MFMAs and stored to the other stage before its last two row fragments.

constexpr int G3 = 256;          // tile edge
constexpr int G3_STAGE = G3 * PK; // bf16 per plane and stage

struct ASrcPlanes {
    const __bf16* hi;
    const __bf16* lo;
    size_t ld;
    struct Regs {
        bf8_t v[4];
    };
    __device__ __forceinline__ ASrcPlanes at(size_t off) const { return ASrcPlanes{hi + off, lo + off, ld}; }
    __device__ __forceinline__ void init(int, Regs&) const {}
    // thread t: rows (t >> 2) and (t >> 2) + 128, chunk t & 3, both planes
    __device__ __forceinline__ void fetch(int m0, int k, Regs& R) const {
        const int t = threadIdx.x, r = t >> 2, kc = (t & 3) * 8;
        const size_t o0 = (size_t)(m0 + r) * ld + k + kc, o1 = o0 + (size_t)128 * ld;
        R.v[0] = *(const bf8_t*)(hi + o0);
        R.v[1] = *(const bf8_t*)(hi + o1);
        R.v[2] = *(const bf8_t*)(lo + o0);
        R.v[3] = *(const bf8_t*)(lo + o1);
    }
    __device__ __forceinline__ void store(__bf16* sh, __bf16* sl, const Regs& R) const {
        const int t = threadIdx.x, r = t >> 2, c = t & 3;
        *(bf8_t*)&sh[lds_slot(r, c)] = R.v[0];
        *(bf8_t*)&sh[lds_slot(r + 128, c)] = R.v[1];
        *(bf8_t*)&sl[lds_slot(r, c)] = R.v[2];
        *(bf8_t*)&sl[lds_slot(r + 128, c)] = R.v[3];
    }
};

struct ASrcTable {
    const unsigned* Qhl;
    const unsigned* cpk;
    int Ny, Nz;
    struct Regs {
        unsigned w[16];
        unsigned xa, ya, za;
        bool rowv;
    };
    __device__ __forceinline__ ASrcTable at(size_t) const { return *this; }
    // thread t: row t >> 1 of the tile, columns 16 (t & 1) .. + 15 of the step
    __device__ __forceinline__ void init(int m0, Regs& R) const {
        const unsigned ca = cpk[m0 + (threadIdx.x >> 1)];
        R.rowv = ca != CELL_NONE;
        R.xa = ca >> 20;
        R.ya = (ca >> 10) & 1023u;
        R.za = ca & 1023u;
    }
    __device__ __forceinline__ void fetch(int, int k, Regs& R) const {
        const int h = threadIdx.x & 1;
        unsigned c[16];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 v = *(const uint4*)(cpk + k + 16 * h + 4 * q);
            c[4 * q] = v.x;
            c[4 * q + 1] = v.y;
            c[4 * q + 2] = v.z;
            c[4 * q + 3] = v.w;
        }
        gather16(Qhl, c, R.xa, R.ya, R.za, R.rowv, Ny, Nz, R.w);
    }
    __device__ __forceinline__ void store(__bf16* sh, __bf16* sl, const Regs& R) const {
        const int r = threadIdx.x >> 1, h = threadIdx.x & 1;
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
            const unsigned* w = R.w + 8 * cc;
            uint4 a, b;
            a.x = pair_hi(w[0], w[1]);
            a.y = pair_hi(w[2], w[3]);
            a.z = pair_hi(w[4], w[5]);
            a.w = pair_hi(w[6], w[7]);
            b.x = pair_lo(w[0], w[1]);
            b.y = pair_lo(w[2], w[3]);
            b.z = pair_lo(w[4], w[5]);
            b.w = pair_lo(w[6], w[7]);
            const int o = lds_slot(r, 2 * h + cc);
            *(uint4*)&sh[o] = a;
            *(uint4*)&sl[o] = b;
        }
    }
};

// Batched form: workgroup id = (batch * nmb * nnb + tile) * ksplit + split;
// batch b multiplies A at offset ((b + a_add) >> a_shift) * a_bstride by B at
// offset b * b_bstride into the epilogue's batch b.
// DL (A from plane pairs only): the operand tiles go from global memory to
// LDS by the DMA loads of gfx950 (global_load_lds_dwordx4), with no VGPR
// staging and no LDS store instructions.  A wave-instruction fills 1 KB of LDS,
// 16 B per lane in lane order: 16 tile rows of one plane; lane L takes row
// 16 rb + L/4 and the chunk that the swizzle puts in its slot L % 4.  The 64
// instructions of a step are spread over the 8 waves.  DL = 2 issues half of a
// wave's DMA loads at the start of the step and the other half after its
// fourth row fragment, which spreads the LDS writes over the step.
template <typename ASrc, typename Epi, int DL = 0>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void gemm3_kernel(
    ASrc A0, const __bf16* __restrict__ Bh0, const __bf16* __restrict__ Bl0, size_t ldb, int K, int nmb, int nnb, int G,
    int remap, int ksplit, Epi E0, size_t a_bstride, size_t b_bstride, int a_add, int a_shift) {
    extern __shared__ __attribute__((aligned(16))) __bf16 g3_lds[];   // 2 stages x (Ah, Al, Bh, Bl)
    // split-K: workgroup b takes part b % ksplit of the K steps of tile b / ksplit
    const int split = blockIdx.x % ksplit;
    const int tiles = nmb * nnb, id = blockIdx.x / ksplit, bat = id / tiles;
    const Epi E = E0.part(split).batch(bat);
    const ASrc A = A0.at((size_t)((bat + a_add) >> a_shift) * a_bstride);
    const __bf16* __restrict__ Bh = Bh0 + (size_t)bat * b_bstride;
    const __bf16* __restrict__ Bl = Bl0 + (size_t)bat * b_bstride;
    int mb, nbk;
    tile_of(id - bat * tiles, nmb, nnb, G, remap, mb, nbk);
    const int m0 = mb * G3, n0 = nbk * G3;
    if (Epi::upper && m0 > n0 + G3 - 1) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 2) * 128, wn = (wave & 3) * 64;
    typename ASrc::Regs ra;
    A.init(m0, ra);
    // B chunks: rows (t >> 2), + 128, chunk t & 3, both planes
    const int br = tid >> 2, bkc = (tid & 3) * 8;
    const __bf16* pBh = Bh + (size_t)(n0 + br) * ldb + bkc;
    const __bf16* pBl = Bl + (size_t)(n0 + br) * ldb + bkc;
    const size_t b128 = (size_t)128 * ldb;
    bf8_t rb[4];
    auto fetchB = [&](int k) {
        rb[0] = *(const bf8_t*)(pBh + k);
        rb[1] = *(const bf8_t*)(pBh + b128 + k);
        rb[2] = *(const bf8_t*)(pBl + k);
        rb[3] = *(const bf8_t*)(pBl + b128 + k);
    };
    auto storeB = [&](__bf16* sh, __bf16* sl) {
        const int c = tid & 3;
        *(bf8_t*)&sh[lds_slot(br, c)] = rb[0];
        *(bf8_t*)&sh[lds_slot(br + 128, c)] = rb[1];
        *(bf8_t*)&sl[lds_slot(br, c)] = rb[2];
        *(bf8_t*)&sl[lds_slot(br + 128, c)] = rb[3];
    };
    auto stage = [&](int s, int plane) { return g3_lds + (size_t)(4 * s + plane) * G3_STAGE; };
    const int kq = lane >> 4, fr = lane & 15;
    // DL: the 8 wave-instructions of this wave (u = 8 wave + q): operand u / 32
    // (A, B), plane (u / 16) % 2, rows 16 (u % 16) .. + 15; this lane's source
    const __bf16* dl_src[8];
    int dl_off[8];
    if constexpr (DL) {
        const int lr = lane >> 2, slot = lane & 3;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int u = wave * 8 + q, opnd = u >> 5, plane = (u >> 4) & 1, rb = u & 15;
            const int r = rb * 16 + lr;
            const int c = slot ^ (((r >> 1) ^ (r >> 2)) & 3);
            const __bf16* base = opnd == 0 ? (plane == 0 ? A.hi : A.lo) : (plane == 0 ? Bh : Bl);
            const size_t ldx = opnd == 0 ? A.ld : ldb;
            dl_src[q] = base + (size_t)((opnd == 0 ? m0 : n0) + r) * ldx + c * 8;
            dl_off[q] = (2 * opnd + plane) * G3_STAGE + rb * 16 * PK;
        }
    }
    auto dl_issue = [&](int s, int k, int q0, int q1) {
#pragma unroll
        for (int q = q0; q < q1; q++)
            __builtin_amdgcn_global_load_lds(dl_src[q] + k,
                                             (__attribute__((address_space(3))) void*)(g3_lds + (size_t)s * 4 * G3_STAGE +
                                                                                      dl_off[q]),
                                             16, 0, 0);
    };
    f4_t acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4_t{0, 0, 0, 0};
    // balanced K ranges: every part has at least one step (the host keeps
    // ksplit <= K / PK)
    const int nkall = K / PK;
    const int kt0 = (int)((long long)split * nkall / ksplit), nk = (int)((long long)(split + 1) * nkall / ksplit);
    if constexpr (DL) {
        dl_issue(0, kt0 * PK, 0, 8);
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
        A.fetch(m0, kt0 * PK, ra);
        fetchB(kt0 * PK);
        A.store(stage(0, 0), stage(0, 1), ra);
        storeB(stage(0, 2), stage(0, 3));
    }
    __syncthreads();
    for (int kt = kt0; kt < nk; kt++) {
        const int cur = (kt - kt0) & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            if constexpr (DL) {
                dl_issue(cur ^ 1, (kt + 1) * PK, 0, DL == 2 ? 4 : 8);
            } else {
                A.fetch(m0, (kt + 1) * PK, ra);
                fetchB((kt + 1) * PK);
            }
        }
        const __bf16* sAh = stage(cur, 0);
        const __bf16* sAl = stage(cur, 1);
        const __bf16* sBh = stage(cur, 2);
        const __bf16* sBl = stage(cur, 3);
        bf8_t bh[4], bl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = lds_slot(wn + j * 16 + fr, kq);
            bh[j] = *(const bf8_t*)&sBh[o];
            bl[j] = *(const bf8_t*)&sBl[o];
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            // the next step's LDS stores go out while the last two row
            // fragments' MFMAs run (measured 2-3 % faster than after them)
            if (!DL && i == 6 && more) {
                A.store(stage(cur ^ 1, 0), stage(cur ^ 1, 1), ra);
                storeB(stage(cur ^ 1, 2), stage(cur ^ 1, 3));
            }
            if (DL == 2 && i == 4 && more) dl_issue(cur ^ 1, (kt + 1) * PK, 4, 8);
            const int o = lds_slot(wm + i * 16 + fr, kq);
            const bf8_t ah = *(const bf8_t*)&sAh[o];
            const bf8_t al = *(const bf8_t*)&sAl[o];
            // three sweeps over independent accumulators (lo*hi, hi*lo, hi*hi):
            // consecutive MFMAs never wait on each other's result
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
        }
        // DL: this step's DMA loads into the other stage have landed
        if constexpr (DL != 0) __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) E.put4(m0 + wm + i * 16 + 4 * (lane >> 4), n0 + wn + j * 16 + fr, acc[i][j]);
}

constexpr size_t G3_LDS_BYTES = (size_t)8 * G3_STAGE * sizeof(__bf16);   // 128 KB

// out[i] = sum_s part[s][i] in FP64, in a fixed order (split-K reduction of stage 2).
// Four entries per thread; with up to 8 parts and n a multiple of 4 all part
// loads are 16-byte and issued before the sums.
constexpr int SUM_PARTS_MAX = 8;
__global__ void sum_parts_kernel(const float* __restrict__ part, int nparts, size_t n, double* __restrict__ out) {
    const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= n) return;
    if ((n & 3) == 0 && nparts <= SUM_PARTS_MAX) {
        f4_t v[SUM_PARTS_MAX];
#pragma unroll
        for (int s = 0; s < SUM_PARTS_MAX; s++)
            if (s < nparts) v[s] = *(const f4_t*)&part[(size_t)s * n + i0];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            double acc = 0.0;
#pragma unroll
            for (int s = 0; s < SUM_PARTS_MAX; s++)
                if (s < nparts) acc += (double)v[s][q];
            out[i0 + q] = acc;
        }
        return;
    }
    for (size_t i = i0; i < i0 + 4 && i < n; i++) {
        double acc = 0.0;
        for (int s = 0; s < nparts; s++) acc += (double)part[(size_t)s * n + i];
        out[i] = acc;
    }
}

// B' = diag(w) B with w = sqrt(Nbar)
__global__ void scale_rows_kernel(const double* B, const double* nbar, int N, int M, double* Bw, size_t ldw) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t tot = (size_t)N * M;
    if (i >= tot) return;
    const size_t a = i % N, n = i / N;
    Bw[n * ldw + a] = B[i] * sqrt(nbar[a]);
}

// ---- Toeplitz projection of box surveys ---------------------------------------------
//
// On a box grid S_p[a][b] = Q_p(|dx|, |dy|, |dz|) (the cache key of
// src/sig.cpp:109-112).  Along x, S_p is a symmetric block-Toeplitz matrix of
// blocks T_d (d = |dx|), each G x G over the (y, z) slab (G = n2 n3).  Embedded
// in a circulant of length L = 2 n1 (T_d at d and L - d, zero at n1), the product
// Y = S_p B' becomes, after a length-L DFT along x,
//     Yhat_k = That_k Bhat_k,  That_k = sum_d w_d cos(2 pi d k / L) T_d  (real),
// for k = 0..n1 (the others follow by Hermitian symmetry), and Y is the inverse
// DFT of Yhat at x < n1.  That_k is a 2-level Toeplitz matrix generated from
// Qhat_k(|dy|, |dz|).  Stage 1 becomes nbat = 2 n1 GEMMs of G x G by G x M:
// batch 0 is R_0, batch 2k - 1 the real part R_k (k = 1..n1) and batch 2k the
// imaginary part I_k (k = 1..n1-1) of Bhat_k (I_0 = I_n1 = 0), so batch b uses
// That_((b+1)/2); about n1 / 2 times fewer flops than N x N by N x M.  Grid positions without a cell have B' = 0, so any cell subset of the
// grid is exact.

// B' in grid order, split: planes [m][g] of Mp x Pg, zero where no cell / padding.
__global__ void grid_planes_kernel(const double* __restrict__ Bw, size_t ldw, const int* __restrict__ gmap, int Ng,
                                   int M, __bf16* __restrict__ hi, __bf16* __restrict__ lo, size_t Pg) {
    const size_t g0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    const int m = blockIdx.y;
    if (g0 >= Pg) return;
    bf8_t h, l;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const size_t g = g0 + q;
        double v = 0.0;
        if (m < M && g < (size_t)Ng) {
            const int a = gmap[g];
            if (a >= 0) v = Bw[(size_t)m * ldw + a];
        }
        __bf16 x, y;
        split2((float)v, x, y);
        h[q] = x;
        l[q] = y;
    }
    *(bf8_t*)(hi + (size_t)m * Pg + g0) = h;
    *(bf8_t*)(lo + (size_t)m * Pg + g0) = l;
}

// Bhat planes: batch b holds R_k = sum_x cos(2 pi x k / L) B'_x (b = 0 or odd) or
// I_k = -sum_x sin(2 pi x k / L) B'_x (b even, > 0), k = (b + 1) / 2 (FP64 sums,
// then split).  Block: one mode, 256 slab positions; the n1 grid values of each
// position sit in LDS.
__global__ __launch_bounds__(256) void bhat_kernel(const double* __restrict__ Bw, size_t ldw,
                                                    const int* __restrict__ gmap, int n1, int G, int M,
                                                    const double* __restrict__ trig, __bf16* __restrict__ hi,
                                                    __bf16* __restrict__ lo, size_t Gp, size_t plane, int nbat) {
    extern __shared__ double bh_v[];   // n1 x 256
    const int m = blockIdx.y, t = threadIdx.x;
    const int r = blockIdx.x * 256 + t;
    const int L = 2 * n1;
    for (int x = 0; x < n1; x++) {
        double v = 0.0;
        if (m < M && r < G) {
            const int a = gmap[(size_t)x * G + r];
            if (a >= 0) v = Bw[(size_t)m * ldw + a];
        }
        bh_v[x * 256 + t] = v;
    }
    if (r >= (int)Gp) return;
    for (int b = 0; b < nbat; b++) {
        const int k = (b + 1) >> 1;
        double acc = 0.0;
        if (b > 0 && !(b & 1)) {
            for (int x = 0; x < n1; x++) acc -= trig[L + (x * k) % L] * bh_v[x * 256 + t];
        } else {
            for (int x = 0; x < n1; x++) acc += trig[(x * k) % L] * bh_v[x * 256 + t];
        }
        __bf16 h, l;
        split2((float)acc, h, l);
        const size_t o = (size_t)b * plane + (size_t)m * Gp + r;
        hi[o] = h;
        lo[o] = l;
    }
}

// Qhat_k(e) = Q(0, e) + 2 sum_{d=1}^{n1-1} cos(2 pi d k / L) Q(d, e), e = |dy| n3 + |dz|.
__global__ void qhat_kernel(const double* __restrict__ Q, int n1, int G, const double* __restrict__ trig,
                            double* __restrict__ qhat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n1 + 1) * G) return;
    const int k = i / G, e = i - k * G, L = 2 * n1;
    double acc = Q[e];
    for (int d = 1; d < n1; d++) acc += 2.0 * trig[(d * k) % L] * Q[(size_t)d * G + e];
    qhat[i] = acc;
}

// That_k planes: row r, column c (slab positions) -> Qhat_k(|y_r - y_c| n3 + |z_r - z_c|),
// zero outside G x G; k = blockIdx.z, 8 columns per thread.
__global__ void that_kernel(const double* __restrict__ qhat, int G, int n3, __bf16* __restrict__ hi,
                            __bf16* __restrict__ lo, size_t Gr, size_t Gp) {
    const size_t c0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    const int r = blockIdx.y, k = blockIdx.z;
    if (c0 >= Gp) return;
    const double* qk = qhat + (size_t)k * G;
    const int yr = r / n3, zr = r - yr * n3;
    // (y, z) of column c0, then stepped along the 8 columns
    int yc = (int)(c0 / n3), zc = (int)(c0 - (size_t)yc * n3);
    bf8_t h, l;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int c = (int)(c0 + q);
        float v = 0.f;
        if (r < G && c < G) v = (float)qk[abs(yr - yc) * n3 + abs(zr - zc)];
        if (++zc == n3) {
            zc = 0;
            yc++;
        }
        __bf16 x, y;
        split2(v, x, y);
        h[q] = x;
        l[q] = y;
    }
    const size_t o = ((size_t)k * Gr + r) * Gp + c0;
    *(bf8_t*)(hi + o) = h;
    *(bf8_t*)(lo + o) = l;
}

// Stage-1 epilogue of the Toeplitz path: D[r][m] of batch b -> Yhat[m][b][r] (FP32).
struct EpiYhat {
    float* y;
    int M, G, nbat;
    __device__ __forceinline__ EpiYhat part(int) const { return *this; }
    __device__ __forceinline__ EpiYhat batch(int b) const {
        EpiYhat e = *this;
        e.y += (size_t)b * G;
        return e;
    }
    __device__ __forceinline__ void put4(int m, int n, const f4_t& v) const {
        if (n >= M) return;
        float* d = y + (size_t)n * nbat * G + m;
        if (m + 3 < G && (G & 3) == 0) {
            *(f4_t*)d = v;
            return;
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (m + r < G) d[r] = v[r];
    }
    static constexpr bool upper = false;
};

// Inverse DFT along x: Y[m][x G + r] = sum_b W[x][b] Yhat[m][b][r] for x < n1,
// written as split planes (the stage-2 operand).  Block: one mode, 128 slab
// positions; the Yhat slice and W (stored as [b][x]) sit in LDS; each thread
// forms a 4 (x) x 4 (r) block, so two 16-byte LDS reads feed 16 FMAs, and
// writes 4 consecutive positions per plane with one 8-byte store.
__global__ __launch_bounds__(256) void idft_kernel(const float* __restrict__ yhat, int nbat, int G, int n1,
                                                    const float* __restrict__ W, __bf16* __restrict__ hi,
                                                    __bf16* __restrict__ lo, size_t Pg) {
    extern __shared__ float id_s[];   // nbat x 128 Yhat, then nbat x n1p W^T
    const int n1p = (n1 + 3) & ~3;
    float* sY = id_s;
    float* sW = id_s + nbat * 128;
    const int m = blockIdx.y, r0 = blockIdx.x * 128, t = threadIdx.x;
    const float* ym = yhat + (size_t)m * nbat * G;
    if ((G & 3) == 0) {
        // every load of the thread is issued before the first LDS store
        // (nbat = 2 Nx <= 128: at most 16 per thread)
        constexpr int MAXL = 16;
        const int nl = (nbat * 32 + 255) / 256;
        f4_t v[MAXL];
#pragma unroll
        for (int u = 0; u < MAXL; u++) {
            const int i = t + u * 256;
            const int b = i >> 5, r = (i & 31) * 4;
            v[u] = (u < nl && i < nbat * 32 && r0 + r < G) ? *(const f4_t*)&ym[(size_t)b * G + r0 + r]
                                                           : f4_t{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < MAXL; u++) {
            const int i = t + u * 256;
            if (u < nl && i < nbat * 32) *(f4_t*)&sY[(i >> 5) * 128 + (i & 31) * 4] = v[u];
        }
    } else {
        for (int i = t; i < nbat * 128; i += 256) {
            const int b = i >> 7, r = i & 127;
            sY[i] = (r0 + r < G) ? ym[(size_t)b * G + r0 + r] : 0.f;
        }
    }
    for (int i = t; i < nbat * n1p / 4; i += 256) *(f4_t*)&sW[4 * i] = *(const f4_t*)&W[4 * i];
    __syncthreads();
    typedef __bf16 bf4_t __attribute__((ext_vector_type(4)));
    const int rq = (t & 31) * 4;
    for (int xq = (t >> 5) * 4; xq < n1; xq += 32) {
        f4_t acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        for (int b = 0; b < nbat; b++) {
            const f4_t y = *(const f4_t*)&sY[b * 128 + rq];
            const f4_t w = *(const f4_t*)&sW[b * n1p + xq];
#pragma unroll
            for (int u = 0; u < 4; u++) acc[u] += w[u] * y;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int x = xq + u;
            if (x >= n1) break;
            const size_t o = (size_t)m * Pg + (size_t)x * G + r0 + rq;
            if ((G & 3) == 0 && r0 + rq + 3 < G) {
                bf4_t h, l;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    __bf16 a, c;
                    split2(acc[u][q], a, c);
                    h[q] = a;
                    l[q] = c;
                }
                *(bf4_t*)(hi + o) = h;
                *(bf4_t*)(lo + o) = l;
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (r0 + rq + q >= G) break;
                    __bf16 a, c;
                    split2(acc[u][q], a, c);
                    hi[o + q] = a;
                    lo[o + q] = c;
                }
            }
        }
    }
}

// ---- Toeplitz projection in FP64 / FP32 -------------------------------------------
//
// The same identity without splitting the operands: B' in grid order and its
// DFT along x in FP64, stage 1 as nbat = 2 n1 batched GEMMs (batch b:
// That_((b+1)/2) times the b-th transformed block, That generated inside the
// GEMM's A loader from Qhat), the inverse DFT in FP64, stage 2 as the upper
// triangle of Bg'^T Y over the grid.  MFMA in FP64 (v_mfma_f64_16x16x4f64) or
// FP32 (v_mfma_f32_16x16x4f32, operands rounded to FP32 at the LDS store).

// B' in grid order: Bg[m][g] = B'[m][gmap[g]] (0 where the grid has no cell).
__global__ void grid_b64_kernel(const double* __restrict__ Bw, size_t ldw, const int* __restrict__ gmap, int Ng, int M,
                                double* __restrict__ Bg, size_t ldg) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * Ng) return;
    const size_t m = i / Ng, g = i - m * Ng;
    const int a = gmap[g];
    Bg[m * ldg + g] = a >= 0 ? Bw[m * ldw + a] : 0.0;
}

// Bhat[b][m][r] = sum_x cos(2 pi x k / L) Bg[m][x G + r] (b = 0 or odd) or
// -sum_x sin(2 pi x k / L) Bg[m][x G + r] (b even, > 0), k = (b + 1) / 2.
// Block (r tile, m); each thread keeps the n1 <= 64 grid values of its (m, r)
// in registers.
constexpr int TOEP_MAXX = 64;
__global__ __launch_bounds__(256) void bhat64_kernel(const double* __restrict__ Bg, size_t ldg, int n1, int G, int M,
                                                     const double* __restrict__ trig, double* __restrict__ Bhat,
                                                     int nbat) {
    const int m = blockIdx.y, r = blockIdx.x * 256 + threadIdx.x;
    if (r >= G) return;
    const int L = 2 * n1;
    double v[TOEP_MAXX];
#pragma unroll
    for (int x = 0; x < TOEP_MAXX; x++) v[x] = x < n1 ? Bg[(size_t)m * ldg + (size_t)x * G + r] : 0.0;
    for (int b = 0; b < nbat; b++) {
        const int k = (b + 1) >> 1;
        const bool imag = b > 0 && !(b & 1);
        const double* t = trig + (imag ? L : 0);
        double acc = 0.0;
#pragma unroll
        for (int x = 0; x < TOEP_MAXX; x++)
            if (x < n1) acc += t[(x * k) % L] * v[x];
        Bhat[((size_t)b * M + m) * G + r] = imag ? -acc : acc;
    }
}

// A operand of the FP64 Toeplitz stage 1: That_k[r][c] = Qhat_k(|y_r - y_c| n3
// + |z_r - z_c|) with k = (batch + 1) / 2; slab[r] = (y << 16) | z.
struct GA_That {
    const double* qhat;
    const int* slab;
    int G, n3;
    static constexpr bool k_contig = true;
    __device__ __forceinline__ void p_shift(int z, size_t) { qhat += (size_t)((z + 1) >> 1) * G; }
    __device__ __forceinline__ double get(int r, int c) const {
        const int a = slab[r], b = slab[c];
        const int dy = abs((a >> 16) - (b >> 16)), dz = abs((a & 0xffff) - (b & 0xffff));
        return qhat[dy * n3 + dz];
    }
};

// Y[m][x G + r] = sum_b W[b][x] Yhat[b][m][r] for x < n1 (FP64), summed in b
// order.  Block (r tile, m); W (nbat x n1) in LDS, the n1 <= NMAX sums of a
// thread in registers; the Yhat values of eight batches are loaded together.
template <int NMAX>
__global__ __launch_bounds__(256) void idft64_kernel(const double* __restrict__ Yhat, int nbat, int G, int n1, int M,
                                                     const double* __restrict__ W, double* __restrict__ Y, size_t ldg) {
    extern __shared__ double id64_w[];   // nbat x n1
    const int m = blockIdx.y, r = blockIdx.x * 256 + threadIdx.x;
    for (int i = threadIdx.x; i < nbat * n1; i += 256) id64_w[i] = W[i];
    __syncthreads();
    if (r >= G) return;
    double acc[NMAX];
#pragma unroll
    for (int x = 0; x < NMAX; x++) acc[x] = 0.0;
    for (int b0 = 0; b0 < nbat; b0 += 8) {
        double yh[8];
#pragma unroll
        for (int u = 0; u < 8; u++) yh[u] = b0 + u < nbat ? Yhat[((size_t)(b0 + u) * M + m) * G + r] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (b0 + u >= nbat) break;
            const double* w = id64_w + (b0 + u) * n1;
#pragma unroll
            for (int x = 0; x < NMAX; x++)
                if (x < n1) acc[x] += w[x] * yh[u];
        }
    }
#pragma unroll
    for (int x = 0; x < NMAX; x++)
        if (x < n1) Y[(size_t)m * ldg + (size_t)x * G + r] = acc[x];
}

// ---- Toeplitz projection of spherical shells (DFT along phi) -----------------------
//
// The arc cache of src/sig.cpp:151-199 gives S_p[a][b] = sqrt(Nbar_a Nbar_b) Q_p(h_a, h_b,
// |dphi_a - dphi_b|), h = dr Nmu + dmu the arc of a cell: along phi, S_p is a symmetric
// block-Toeplitz matrix of blocks T_d[h][h'] = Q_p(h, h', d) over the G = Nr Nmu arcs.  In
// the grid order g = dphi G + h the box identity above holds with x -> phi, and B' in grid
// order, its DFT, the batched slab GEMMs, the inverse DFT and stage 2 are the same kernels.
// Only That_k differs: a shell's arcs are not translation invariant, so That_k is a dense
// G x G matrix, That_k[h][h'] = Q_p(h, h', 0) + 2 sum_{d=1}^{Nphi-1} cos(2 pi d k / L)
// Q_p(h, h', d) (L = 2 Nphi, k = 0..Nphi), formed from the Nphi contiguous table entries
// of each arc pair.  Stage 1 costs 2 Nphi G^2 M flops instead of 2 N^2 M (Nphi / 2 times
// fewer).  Table keys that no cell pair carries hold zero, so every That_k is finite; their
// terms meet B' = 0 (a missing column cell) or are dropped with the rows of missing cells.

// One block: arc row h (blockIdx.y), CB arc columns (one per thread).  The row's table
// slice (CB x Nphi contiguous doubles) is staged through LDS by coalesced reads, next to
// the DFT weights
// C[k][d] (1 at d = 0, else 2 cos(2 pi d k / L)); each thread then forms the Nphi + 1
// transformed values of its column, That_k = sum_d C[k][d] q[d] in d order, from
// wave-uniform LDS rows.  NMAX >= Nphi bounds the registers; CB = 128 for NMAX = 64 keeps
// the block's LDS under 100 KB.  SPLIT: hi / lo BF16 planes
// [k][Gr][Gp], zero outside G x G (the grid covers Gr x Gp); else FP64 T[k][G][G] (the
// grid covers G x G).
template <bool SPLIT, int NMAX>
__global__ __launch_bounds__(256) void that_sph_kernel(const double* __restrict__ Q, int G, int nph,
                                                       const double* __restrict__ trig, __bf16* __restrict__ hi,
                                                       __bf16* __restrict__ lo, size_t Gr, size_t Gp,
                                                       double* __restrict__ T) {
    constexpr int CB = NMAX <= 32 ? 256 : 128;
    extern __shared__ double ts_q[];   // CB x (nph + 1) table entries, then C: (nph + 1) x NMAX
    const int L = 2 * nph, ldq = nph + 1;
    double* sC = ts_q + CB * ldq;
    const int h = blockIdx.y, c0 = blockIdx.x * CB, t = threadIdx.x;
    const int ncol = (h < G && c0 < G) ? min(CB, G - c0) : 0;
    for (int i = t; i < (nph + 1) * NMAX; i += CB) {
        const int k = i / NMAX, d = i - k * NMAX;
        sC[i] = d == 0 ? 1.0 : (d < nph ? 2.0 * trig[(d * k) % L] : 0.0);
    }
    if (ncol > 0) {
        // all of a thread's loads (element t + CB u, u < nph) in flight before the LDS stores
        const double* src = Q + ((size_t)h * G + c0) * nph;
        const int n = ncol * nph;
        double v[NMAX];
#pragma unroll
        for (int u = 0; u < NMAX; u++) {
            const int i = t + u * CB;
            v[u] = (u < nph && i < n) ? src[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NMAX; u++) {
            const int i = t + u * CB;
            if (u < nph && i < n) {
                const int c = i / nph;
                ts_q[c * ldq + (i - c * nph)] = v[u];
            }
        }
    }
    __syncthreads();
    const bool live = t < ncol;
    double q[NMAX];
#pragma unroll
    for (int d = 0; d < NMAX; d++) q[d] = (live && d < nph) ? ts_q[t * ldq + d] : 0.0;
    const int c = c0 + t;
    for (int k = 0; k <= nph; k++) {
        const double* Ck = sC + k * NMAX;
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < NMAX; d++)
            if (d < nph) acc += Ck[d] * q[d];
        if constexpr (SPLIT) {
            if ((size_t)c < Gp) {
                __bf16 x, y;
                split2((float)acc, x, y);
                const size_t o = ((size_t)k * Gr + h) * Gp + c;
                hi[o] = x;
                lo[o] = y;
            }
        } else if (live) {
            T[((size_t)k * G + h) * G + c] = acc;
        }
    }
}

// That_k of one band: the instantiation for Nphi <= 32 holds half the registers.
static int that_sph_cb(int nph) { return nph <= 32 ? 256 : 128; }
static size_t that_sph_lds(int nph) {
    return ((size_t)that_sph_cb(nph) * (nph + 1) + (size_t)(nph + 1) * (nph <= 32 ? 32 : TOEP_MAXX)) * 8;
}
template <bool SPLIT>
static int that_sph_attr(opsec_ctx* ctx, int nph) {
    const void* fn = nph <= 32 ? (const void*)that_sph_kernel<SPLIT, 32> : (const void*)that_sph_kernel<SPLIT, TOEP_MAXX>;
    OPSEC_HIP_TRY(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)that_sph_lds(nph)));
    return OPSEC_OK;
}
template <bool SPLIT>
static void launch_that_sph(opsec_ctx* ctx, const double* Q, int G, int nph, const double* trig, __bf16* hi,
                            __bf16* lo, size_t Gr, size_t Gp, double* T) {
    const size_t cb = that_sph_cb(nph), cols = SPLIT ? Gp : (size_t)G;
    const dim3 grid((unsigned)((cols + cb - 1) / cb), SPLIT ? (unsigned)Gr : (unsigned)G);
    if (nph <= 32)
        hipLaunchKernelGGL((that_sph_kernel<SPLIT, 32>), grid, dim3((unsigned)cb), that_sph_lds(nph), ctx->stream, Q, G,
                           nph, trig, hi, lo, Gr, Gp, T);
    else
        hipLaunchKernelGGL((that_sph_kernel<SPLIT, TOEP_MAXX>), grid, dim3((unsigned)cb), that_sph_lds(nph), ctx->stream,
                           Q, G, nph, trig, hi, lo, Gr, Gp, T);
}

// A operand of the FP64 / FP32 Toeplitz stage 1 of a shell: batch b reads the dense
// That_((b + 1) / 2), row-major G x G.
struct GA_ThatDense : SrcDenseRowMajor {
    __device__ __forceinline__ void p_shift(int z, size_t s) { p += (size_t)((z + 1) >> 1) * s; }
};

// ---- host --------------------------------------------------------------------------

static dim3 grid_for(int M, int N, int nb) {
    return dim3((M + GT_M - 1) / GT_M, (N + GT_N - 1) / GT_N, nb);
}

static dim3 blocks_for(size_t threads) { return dim3((unsigned)((threads + 255) / 256)); }

// Split a row-major double matrix (rows x cols, ldx) into a plane pair of
// rows_p x ld.
static int split_planes(opsec_ctx* ctx, const double* X, size_t ldx, int rows, int cols, __bf16* planes, int rows_p,
                        size_t ld) {
    __bf16* hi = planes;
    __bf16* lo = planes + (size_t)rows_p * ld;
    hipLaunchKernelGGL(split_plane_kernel, blocks_for((size_t)rows_p * ld / 8), dim3(256), 0, ctx->stream, X, ldx,
                       rows, cols, hi, lo, ld, rows_p);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    return OPSEC_OK;
}

// Tile shape of the split-BF16x3 GEMM (0 = 128 x 128, 1 = 256 x 128), tile
// group edge and XCD remapping of the tile order.
static int g_nt_tile = 2, g_nt_group = 8, g_nt_remap = 0;
// Stage 1 generates S_p from the pair table (1) or reads S_p planes written by
// gen_split_kernel (0).
static int g_nt_gen = 0;
// Split-K factor of stage 2 on the 256 x 256 kernel, and its tile-group edge
// (2 measured best for stage 2; 8 for the stage-1 launches).
static int g_s2_split = 8;
// split-K factor of the Toeplitz path's stage 2, which runs eight bands per launch
static int g_s2_split_toep = 4;
static int g_s2_group = 2;

// Operand tiles of the 256 x 256 GEMM (A from plane pairs) into LDS by DMA
// loads (1, measured 4-6 % faster at 32^3 x 4000 modes than through VGPRs, 0),
// issued in two halves per step (2, another 1-3 %).
static int g_g3_dl = 2;

template <typename ASrc, typename Epi, int DL>
static void set_g3_attr() {
    (void)hipFuncSetAttribute((const void*)gemm3_kernel<ASrc, Epi, DL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)G3_LDS_BYTES);
}

template <typename ASrc, typename Epi>
static void launch_g3(opsec_ctx* ctx, const ASrc& A, int arows, const __bf16* B, int brows, size_t ld, int K,
                      const Epi& E, int ksplit = 1, int nbat = 1, size_t a_bstride = 0, size_t b_bstride = 0,
                      int a_add = 0, int a_shift = 0, size_t b_lo_offset = 0, int remap = -1, int group = -1) {
    static bool attr = false;
    if (!attr) {   // 128 KB of dynamic LDS per workgroup
        set_g3_attr<ASrcPlanes, EpiPlanes, 1>();
        set_g3_attr<ASrcPlanes, EpiPacked, 1>();
        set_g3_attr<ASrcPlanes, EpiYhat, 1>();
        set_g3_attr<ASrcPlanes, EpiPackedParts, 1>();
        set_g3_attr<ASrcPlanes, EpiPlanes, 2>();
        set_g3_attr<ASrcPlanes, EpiPacked, 2>();
        set_g3_attr<ASrcPlanes, EpiYhat, 2>();
        set_g3_attr<ASrcPlanes, EpiPackedParts, 2>();
        (void)hipFuncSetAttribute((const void*)gemm3_kernel<ASrcPlanes, EpiPlanes>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)G3_LDS_BYTES);
        (void)hipFuncSetAttribute((const void*)gemm3_kernel<ASrcPlanes, EpiPacked>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)G3_LDS_BYTES);
        (void)hipFuncSetAttribute((const void*)gemm3_kernel<ASrcTable, EpiPlanes>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)G3_LDS_BYTES);
        (void)hipFuncSetAttribute((const void*)gemm3_kernel<ASrcPlanes, EpiYhat>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)G3_LDS_BYTES);
        (void)hipFuncSetAttribute((const void*)gemm3_kernel<ASrcPlanes, EpiPackedParts>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)G3_LDS_BYTES);
        attr = true;
    }
    const int nmb = arows / G3, nnb = brows / G3;
    ksplit = std::max(1, std::min(ksplit, K / PK));
    // the lo plane follows the hi plane (brows x ld) unless an offset is given
    const size_t blo = b_lo_offset ? b_lo_offset : (size_t)brows * ld;
    if constexpr (std::is_same<ASrc, ASrcPlanes>::value) {
        if (g_g3_dl == 1) {
            hipLaunchKernelGGL((gemm3_kernel<ASrcPlanes, Epi, 1>), dim3(nmb * nnb * ksplit * nbat), dim3(512),
                               G3_LDS_BYTES, ctx->stream, A, B, B + blo, ld, K, nmb, nnb,
                               group < 0 ? g_nt_group : group, remap < 0 ? g_nt_remap : remap, ksplit, E, a_bstride,
                               b_bstride, a_add, a_shift);
            return;
        }
        if (g_g3_dl == 2) {
            hipLaunchKernelGGL((gemm3_kernel<ASrcPlanes, Epi, 2>), dim3(nmb * nnb * ksplit * nbat), dim3(512),
                               G3_LDS_BYTES, ctx->stream, A, B, B + blo, ld, K, nmb, nnb,
                               group < 0 ? g_nt_group : group, remap < 0 ? g_nt_remap : remap, ksplit, E, a_bstride,
                               b_bstride, a_add, a_shift);
            return;
        }
    }
        hipLaunchKernelGGL((gemm3_kernel<ASrc, Epi, 0>), dim3(nmb * nnb * ksplit * nbat), dim3(512), G3_LDS_BYTES,
                           ctx->stream, A, B, B + blo, ld, K, nmb, nnb, group < 0 ? g_nt_group : group,
                           remap < 0 ? g_nt_remap : remap, ksplit, E, a_bstride, b_bstride, a_add, a_shift);
}

// C = A B^T (stage 1, epilogue to split planes) or P = A B^T (stage 2, packed).
// Plane row counts are multiples of 256.
template <typename Epi>
static void launch_nt(opsec_ctx* ctx, const __bf16* A, int arows, const __bf16* B, int brows, size_t ld, int K,
                      const Epi& E) {
    if (g_nt_tile == 2) {
        launch_g3(ctx, ASrcPlanes{A, A + (size_t)arows * ld, ld}, arows, B, brows, ld, K, E);
    } else if (g_nt_tile == 0) {
        const int nmb = arows / 128, nnb = brows / 128;
        hipLaunchKernelGGL((nt_bf16x3_kernel<Epi, 4, 4>), dim3(nmb * nnb), dim3(256), 0, ctx->stream, A,
                           A + (size_t)arows * ld, ld, B, B + (size_t)brows * ld, ld, K, nmb, nnb, g_nt_group,
                           g_nt_remap, E);
    } else {
        const int nmb = arows / 256, nnb = brows / 128;
        hipLaunchKernelGGL((nt_bf16x3_kernel<Epi, 8, 4>), dim3(nmb * nnb), dim3(256), 0, ctx->stream, A,
                           A + (size_t)arows * ld, ld, B, B + (size_t)brows * ld, ld, K, nmb, nnb, g_nt_group,
                           g_nt_remap, E);
    }
}

// Split-BF16x3 projection of the bands funcs[0..nb) into d_P, band by band.
static int run_project_bf16x3(opsec_ctx* ctx, int nb, const int* funcs) {
    ProjState* ps = ctx->proj;
    const int N = ctx->ncells, M = ctx->nmodes;
    const size_t Kp = round_up(N, 256), Mp = round_up(M, 256), Pp = plane_pitch(Kp);
    int rc;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Bp, &ps->cap_Bp, 2 * Mp * Pp * sizeof(__bf16)))) return rc;
    // packed coordinates (box grids with indices below 1024) let the S_p planes be
    // generated from 32-bit cell words; other grids use the generic generator
    const bool packable = ctx->sv.coordsys == OPSEC_COORD_CARTESIAN && ctx->sv.n1 < 1024 && ctx->sv.n2 < 1024 &&
                          ctx->sv.n3 < 1024;
    const bool gen = g_nt_gen && packable;
    if (!gen)
        if ((rc = dev_realloc(ctx, (void**)&ps->d_Sp, &ps->cap_Sp, 2 * Kp * Pp * sizeof(__bf16)))) return rc;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Cp, &ps->cap_Cp, 2 * Mp * Pp * sizeof(__bf16)))) return rc;
    if ((rc = dev_realloc(ctx, &ps->d_P, &ps->cap_P, (size_t)nb * packed_size(M) * 8))) return rc;
    if (g_s2_split > 1 &&
        (rc = dev_realloc(ctx, (void**)&ps->d_Ppart, &ps->cap_Ppart, (size_t)g_s2_split * packed_size(M) * 4)))
        return rc;
    if (!ps->bp_valid) {
        // B' is column-major N x M: row n of the plane is mode n, cells contiguous
        if ((rc = split_planes(ctx, ps->d_Bw, ps->ldw, M, N, ps->d_Bp, (int)Mp, Pp))) return rc;
        ps->bp_valid = true;
    }
    if (packable) {
        if ((rc = dev_realloc(ctx, (void**)&ps->d_Qhl, &ps->cap_Qhl, (size_t)ctx->nkeys * 4))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_cpk, &ps->cap_cpk, Kp * 4))) return rc;
        hipLaunchKernelGGL(pack_cells_kernel, dim3((unsigned)((Kp + 255) / 256)), dim3(256), 0, ctx->stream,
                           ctx->d_cellidx, N, (int)Kp, ps->d_cpk);
    }
    std::vector<hipEvent_t> ev(4 * (size_t)nb);
    for (auto& e : ev) OPSEC_HIP_TRY(ctx, hipEventCreate(&e));
    __bf16* Sh = ps->d_Sp;
    __bf16* Sl = ps->d_Sp + Kp * Pp;
    EpiPlanes toC{ps->d_Cp, ps->d_Cp + Mp * Pp, Pp};
    for (int z = 0; z < nb; z++) {
        const double* Qf = ctx->d_Q + (size_t)funcs[z] * ctx->nkeys;
        hipEventRecord(ev[4 * z], ctx->stream);
        if (packable)
            hipLaunchKernelGGL(pack_table_kernel, dim3((ctx->nkeys + 255) / 256), dim3(256), 0, ctx->stream, Qf,
                               ps->d_Qhl, ctx->nkeys);
        if (!gen && packable)
            for (size_t r0 = 0; r0 < Kp; r0 += 32768)
                hipLaunchKernelGGL(gen_planes_kernel, dim3((unsigned)((Pp + 2047) / 2048), (unsigned)std::min<size_t>(32768, Kp - r0)),
                                   dim3(256), 0, ctx->stream, ps->d_Qhl, ps->d_cpk, ctx->sv.n2, ctx->sv.n3, (int)Kp, Sh, Sl,
                                   Pp, (int)r0);
        else if (!gen)
            hipLaunchKernelGGL(gen_split_kernel, blocks_for(Kp * Pp / 8), dim3(256), 0, ctx->stream, Qf,
                               ctx->d_cellidx, pair_key_of(ctx->sv), N, Sh, Sl, Pp, (int)Kp);
        hipEventRecord(ev[4 * z + 1], ctx->stream);
        if (gen) {
            launch_g3(ctx, ASrcTable{ps->d_Qhl, ps->d_cpk, ctx->sv.n2, ctx->sv.n3}, (int)Kp, ps->d_Bp, (int)Mp, Pp,
                      (int)Kp, toC);
        } else {
            launch_nt(ctx, ps->d_Sp, (int)Kp, ps->d_Bp, (int)Mp, Pp, (int)Kp, toC);
        }
        hipEventRecord(ev[4 * z + 2], ctx->stream);
        double* Pz = (double*)ps->d_P + (size_t)z * packed_size(M);
        const int s2 = std::max(1, std::min(g_s2_split, (int)(Kp / PK)));
        if (g_nt_tile >= 2 && s2 > 1) {
            const size_t np = (size_t)packed_size(M);
            EpiPackedParts toP{ps->d_Ppart, M, np};
            launch_g3(ctx, ASrcPlanes{ps->d_Bp, ps->d_Bp + Mp * Pp, Pp}, (int)Mp, ps->d_Cp, (int)Mp, Pp, (int)Kp, toP,
                      s2, 1, 0, 0, 0, 0, 0, -1, g_s2_group);
            hipLaunchKernelGGL(sum_parts_kernel, blocks_for((np + 3) / 4), dim3(256), 0, ctx->stream, ps->d_Ppart, s2, np,
                               Pz);
        } else {
            EpiPacked toP{Pz, M};
            launch_nt(ctx, ps->d_Bp, (int)Mp, ps->d_Cp, (int)Mp, Pp, (int)Kp, toP);
        }
        hipEventRecord(ev[4 * z + 3], ctx->stream);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection launch: %s", hipGetErrorString(e));
    hipError_t es = hipEventSynchronize(ev[4 * nb - 1]);
    if (!rc && es != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection: %s", hipGetErrorString(es));
    float tg = 0, t1 = 0, t2 = 0;
    if (!rc)
        for (int z = 0; z < nb; z++) {
            float a, b, c;
            hipEventElapsedTime(&a, ev[4 * z], ev[4 * z + 1]);
            hipEventElapsedTime(&b, ev[4 * z + 1], ev[4 * z + 2]);
            hipEventElapsedTime(&c, ev[4 * z + 2], ev[4 * z + 3]);
            tg += a;
            t1 += b;
            t2 += c;
        }
    for (auto& x : ev) hipEventDestroy(x);
    if (rc) return rc;
    ps->last_ms_gen = tg;
    ps->last_ms_gemm1 = t1;
    ps->last_ms_gemm2 = t2;
    ps->last_launches = nb;
    ps->last_launches2 = nb;
    ps->last_ms_stage1 = tg + t1;
    ps->last_ms_stage2 = t2;
    ps->nb = nb;
    ps->prec = OPSEC_PREC_BF16X3;
    return OPSEC_OK;
}

constexpr int TOEP_GROUP = 8;     // bands per stage-2 launch of the split-BF16x3 Toeplitz path
constexpr int TOEP64_GROUP = 4;   // bands per stage-2 launch of the FP64 / FP32 Toeplitz path

// cos / sin of 2 pi j / L, exact at the quarter points
static void trig_table(int L, std::vector<double>& t) {
    t.assign(2 * (size_t)L, 0.0);
    for (int j = 0; j < L; j++) {
        double c, sn;
        if (j == 0) { c = 1; sn = 0; }
        else if (2 * j == L) { c = -1; sn = 0; }
        else if (4 * j == L) { c = 0; sn = 1; }
        else if (4 * j == 3 * L) { c = 0; sn = -1; }
        else { c = std::cos(2.0 * M_PI * j / L); sn = std::sin(2.0 * M_PI * j / L); }
        t[j] = c;
        t[L + j] = sn;
    }
}

// Geometry of the Toeplitz projection: the transformed axis (x of a box, phi of a
// shell) of n1 grid planes, the slab of G positions ((y, z) of a box, the Nr Nmu arcs
// of a shell) and the grid index g = x G + slab position of a cell.
struct ToepGrid {
    bool sph;
    int n1, G;
};
static ToepGrid toep_grid(const SurveyDev& sv) {
    if (sv.coordsys == OPSEC_COORD_SPHERICAL) return {true, sv.n3, sv.n1 * sv.n2};
    return {false, sv.n1, sv.n2 * sv.n3};
}
static size_t toep_index(const SurveyDev& sv, const opsec_cell& c) {
    if (sv.coordsys == OPSEC_COORD_SPHERICAL)
        return (size_t)c.d3 * sv.n1 * sv.n2 + (size_t)c.d1 * sv.n2 + c.d2;
    return ((size_t)c.d1 * sv.n2 + c.d2) * sv.n3 + c.d3;
}
// grid index -> cell number, -1 where the grid has no cell (cells must occupy
// distinct grid positions)
static int toep_gmap(opsec_ctx* ctx, size_t Ng, std::vector<int>& gmap) {
    gmap.assign(Ng, -1);
    for (int a = 0; a < ctx->ncells; a++) {
        const size_t g = toep_index(ctx->sv, ctx->h_cells[a]);
        if (gmap[g] >= 0) return set_error(ctx, OPSEC_EINVAL, "cells %d and %d share a grid position", gmap[g], a);
        gmap[g] = a;
    }
    return OPSEC_OK;
}
static int toep_check(opsec_ctx* ctx, const ToepGrid& tg) {
    if (tg.n1 > TOEP_MAXX)
        return set_error(ctx, OPSEC_EINVAL, "Toeplitz projection supports at most %d grid planes along %s (%d)",
                         TOEP_MAXX, tg.sph ? "phi" : "x", tg.n1);
    return OPSEC_OK;
}

static int run_project_toeplitz(opsec_ctx* ctx, int nb, const int* funcs) {
    ProjState* ps = ctx->proj;
    const SurveyDev& sv = ctx->sv;
    const int M = ctx->nmodes;
    const ToepGrid tgrid = toep_grid(sv);
    const int n1 = tgrid.n1, G = tgrid.G, Ng = n1 * G, L = 2 * n1, nbat = 2 * n1;
    int rc;
    if ((rc = toep_check(ctx, tgrid))) return rc;
    const size_t Mp = round_up(M, 256), Gr = round_up(G, 256), Gk = round_up(G, 32), Gp = plane_pitch(Gk);
    const size_t Kg = round_up(Ng, 256), Pg = plane_pitch(Kg);
    const size_t plane_b = Mp * Gp, plane_t = Gr * Gp;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Bg, &ps->cap_Bg, 2 * Mp * Pg * sizeof(__bf16)))) return rc;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Bhat, &ps->cap_Bhat, 2 * (size_t)nbat * plane_b * sizeof(__bf16))))
        return rc;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_That, &ps->cap_That, 2 * (size_t)(n1 + 1) * plane_t * sizeof(__bf16))))
        return rc;
    if (!tgrid.sph && (rc = dev_realloc(ctx, (void**)&ps->d_Qhat, &ps->cap_Qhat, (size_t)(n1 + 1) * G * 8)))
        return rc;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Yhat, &ps->cap_Yhat, (size_t)M * nbat * G * 4))) return rc;
    if ((rc = dev_realloc(ctx, &ps->d_P, &ps->cap_P, (size_t)nb * packed_size(M) * 8))) return rc;
    const int s2 = std::max(1, std::min(g_s2_split_toep, (int)(Kg / PK)));
    // stage 2 runs once per group of up to TOEP_GROUP bands: the 136 upper tiles x 4
    // split-K parts of one band fill 2.1 rounds of the CUs, eight bands 17 (against
    // four bands x 8 parts: the same tiles, half the partial sums, 1.1 % faster steps)
    const int grp = std::min(nb, TOEP_GROUP);
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Ppart, &ps->cap_Ppart, (size_t)grp * s2 * packed_size(M) * 4)))
        return rc;
    const size_t cg_pair = 2 * Mp * Pg;   // one band's Y plane pair (bf16 elements)
    if (!ps->toep_valid) {
        std::vector<int> gmap;
        if ((rc = toep_gmap(ctx, Ng, gmap))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_gmap, &ps->cap_gmap, (size_t)Ng * 4))) return rc;
        std::vector<double> trig;
        trig_table(L, trig);
        if ((rc = dev_realloc(ctx, (void**)&ps->d_trig, &ps->cap_trig, trig.size() * 8))) return rc;
        // inverse weights W[x][b]: (c_k / L) cos and -(c_k / L) sin, c_0 = c_n1 = 1, else 2
        // stored transposed and padded, [b][x] with x < n1p = n1 rounded up to 4
        const int n1p = (n1 + 3) & ~3;
        std::vector<float> W((size_t)nbat * n1p, 0.f);
        for (int x = 0; x < n1; x++)
            for (int b = 0; b < nbat; b++) {
                const int k = (b + 1) >> 1, j = (x * k) % L;
                const double ck = (k == 0 || k == n1) ? 1.0 : 2.0;
                const bool imag = b > 0 && !(b & 1);
                W[(size_t)b * n1p + x] = (float)(imag ? -ck * trig[L + j] / L : ck * trig[j] / L);
            }
        if ((rc = dev_realloc(ctx, (void**)&ps->d_W, &ps->cap_W, W.size() * 4))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_Cg, &ps->cap_Cg, TOEP_GROUP * cg_pair * sizeof(__bf16)))) return rc;
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, ps->d_gmap, gmap.data(), gmap.size() * 4, hipMemcpyHostToDevice));
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, ps->d_trig, trig.data(), trig.size() * 8, hipMemcpyHostToDevice));
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, ps->d_W, W.data(), W.size() * 4, hipMemcpyHostToDevice));
        // the inverse DFT writes rows m < M and columns g < Ng of Cg; the rest stays zero
        OPSEC_HIP_TRY(ctx, hipMemsetAsync(ps->d_Cg, 0, TOEP_GROUP * cg_pair * sizeof(__bf16), ctx->stream));
        hipLaunchKernelGGL(grid_planes_kernel, dim3((unsigned)((Pg / 8 + 255) / 256), (unsigned)Mp), dim3(256), 0,
                           ctx->stream, ps->d_Bw, ps->ldw, ps->d_gmap, Ng, M, ps->d_Bg, ps->d_Bg + Mp * Pg, Pg);
        const size_t bh_lds = (size_t)n1 * 256 * 8;
        OPSEC_HIP_TRY(ctx, hipFuncSetAttribute((const void*)bhat_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)bh_lds));
        hipLaunchKernelGGL(bhat_kernel, dim3((unsigned)((Gp + 255) / 256), (unsigned)Mp), dim3(256),
                           bh_lds, ctx->stream, ps->d_Bw, ps->ldw, ps->d_gmap, n1, G, M, ps->d_trig,
                           ps->d_Bhat, ps->d_Bhat + (size_t)nbat * plane_b, Gp, plane_b, nbat);
        OPSEC_HIP_TRY(ctx, hipGetLastError());
        ps->toep_valid = true;
    }
    const size_t id_lds = ((size_t)nbat * 128 + (size_t)((n1 + 3) & ~3) * nbat) * 4;
    OPSEC_HIP_TRY(ctx, hipFuncSetAttribute((const void*)idft_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)id_lds));
    if (tgrid.sph && (rc = that_sph_attr<true>(ctx, n1))) return rc;
    __bf16* Th = ps->d_That;
    __bf16* Tl = ps->d_That + (size_t)(n1 + 1) * plane_t;
    EpiYhat toY{ps->d_Yhat, M, G, nbat};
    const size_t np = (size_t)packed_size(M);
    // events per band: start, That generated, stage-1 GEMMs, inverse DFT; per
    // group: stage-2 GEMM, split-K reductions
    const int ngroups = (nb + grp - 1) / grp;
    std::vector<hipEvent_t> ev(4 * (size_t)nb + 2 * (size_t)ngroups);
    for (auto& e : ev) OPSEC_HIP_TRY(ctx, hipEventCreate(&e));
    hipEvent_t* evb = ev.data();
    hipEvent_t* evg = ev.data() + 4 * (size_t)nb;
    float tg = 0, t1 = 0, t2 = 0;
    for (int z0 = 0, gi = 0; z0 < nb; z0 += grp, gi++) {
        const int g = std::min(grp, nb - z0);
        for (int q = 0; q < g; q++) {
            const int z = z0 + q;
            __bf16* Cz = ps->d_Cg + (size_t)q * cg_pair;
            const double* Qf = ctx->d_Q + (size_t)funcs[z] * ctx->nkeys;
            hipEventRecord(evb[4 * z], ctx->stream);
            if (tgrid.sph) {
                launch_that_sph<true>(ctx, Qf, G, n1, ps->d_trig, Th, Tl, Gr, Gp, nullptr);
            } else {
                hipLaunchKernelGGL(qhat_kernel, dim3((unsigned)(((n1 + 1) * G + 255) / 256)), dim3(256), 0,
                                   ctx->stream, Qf, n1, G, ps->d_trig, ps->d_Qhat);
                hipLaunchKernelGGL(that_kernel,
                                   dim3((unsigned)((Gp / 8 + 255) / 256), (unsigned)Gr, (unsigned)(n1 + 1)), dim3(256),
                                   0, ctx->stream, ps->d_Qhat, G, sv.n3, Th, Tl, Gr, Gp);
            }
            hipEventRecord(evb[4 * z + 1], ctx->stream);
            // XCD-aware tile order: the four row tiles that share a transformed-mode
            // block run on one XCD and meet in its L2 (measured 7 % faster here;
            // slower for the long-K stage 2)
            launch_g3(ctx, ASrcPlanes{Th, Tl, Gp}, (int)Gr, ps->d_Bhat, (int)Mp, Gp, (int)Gk, toY, 1, nbat, plane_t,
                      plane_b, 1, 1, (size_t)nbat * plane_b, 1);
            hipEventRecord(evb[4 * z + 2], ctx->stream);
            hipLaunchKernelGGL(idft_kernel, dim3((unsigned)((G + 127) / 128), (unsigned)M), dim3(256), id_lds,
                               ctx->stream, ps->d_Yhat, nbat, G, n1, ps->d_W, Cz, Cz + Mp * Pg, Pg);
            hipEventRecord(evb[4 * z + 3], ctx->stream);
        }
        // stage 2 of the group's bands in one launch: batch q = band z0 + q
        EpiPackedParts toP{ps->d_Ppart, M, np, (size_t)s2 * np};
        launch_g3(ctx, ASrcPlanes{ps->d_Bg, ps->d_Bg + Mp * Pg, Pg}, (int)Mp, ps->d_Cg, (int)Mp, Pg, (int)Kg, toP, s2, g,
                  0, cg_pair, 0, 0, Mp * Pg, -1, g_s2_group);
        hipEventRecord(evg[2 * gi], ctx->stream);
        for (int q = 0; q < g; q++)
            hipLaunchKernelGGL(sum_parts_kernel, blocks_for((np + 3) / 4), dim3(256), 0, ctx->stream,
                               ps->d_Ppart + (size_t)q * s2 * np, s2, np, (double*)ps->d_P + (size_t)(z0 + q) * np);
        hipEventRecord(evg[2 * gi + 1], ctx->stream);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection launch: %s", hipGetErrorString(e));
    hipError_t es = hipEventSynchronize(evg[2 * ngroups - 1]);
    if (!rc && es != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection: %s", hipGetErrorString(es));
    if (!rc) {
        for (int z = 0; z < nb; z++) {
            float a, b, c;
            hipEventElapsedTime(&a, evb[4 * z], evb[4 * z + 1]);
            hipEventElapsedTime(&b, evb[4 * z + 1], evb[4 * z + 2]);
            hipEventElapsedTime(&c, evb[4 * z + 2], evb[4 * z + 3]);
            tg += a + c;   // That generation, inverse DFT
            t1 += b;
        }
        for (int gi = 0; gi < ngroups; gi++) {
            const int zl = std::min(nb, (gi + 1) * grp) - 1;   // last band of the group
            float d, e2;
            hipEventElapsedTime(&d, evb[4 * zl + 3], evg[2 * gi]);
            hipEventElapsedTime(&e2, evg[2 * gi], evg[2 * gi + 1]);
            t2 += d;
            tg += e2;      // split-K reductions
        }
    }
    for (auto& x : ev) hipEventDestroy(x);
    if (rc) return rc;
    ps->last_ms_gen = tg;
    ps->last_ms_gemm1 = t1;
    ps->last_ms_gemm2 = t2;
    ps->last_launches = nb;
    ps->last_launches2 = ngroups;
    ps->last_ms_stage1 = tg + t1;
    ps->last_ms_stage2 = t2;
    ps->nb = nb;
    ps->prec = OPSEC_PREC_BF16X3;
    return OPSEC_OK;
}

// Toeplitz projection without operand splitting: FP64 (prec FP64) or FP32
// (prec FP32) MFMA, FP64 DFTs.  Per band: Qhat (FP64 DFT of the table along
// x), the batched stage-1 GEMMs into Yhat, the inverse DFT into Y (grid
// order), the upper-triangle stage-2 GEMM into the packed result.
static int run_project_toeplitz64(opsec_ctx* ctx, int nb, const int* funcs, int prec) {
    ProjState* ps = ctx->proj;
    const SurveyDev& sv = ctx->sv;
    const int M = ctx->nmodes;
    const ToepGrid tgrid = toep_grid(sv);
    const int n1 = tgrid.n1, G = tgrid.G, Ng = n1 * G, L = 2 * n1, nbat = 2 * n1;
    int rc;
    if ((rc = toep_check(ctx, tgrid))) return rc;
    if (!tgrid.sph && (sv.n2 >= 65536 || sv.n3 >= 65536))
        return set_error(ctx, OPSEC_EINVAL, "Toeplitz projection: Ny, Nz < 65536");
    const size_t ldg = (size_t)Ng + 32;   // pitch of the grid-order operands (not a power of two)
    // stage 2 runs once per group of up to TOEP64_GROUP bands (grid.z = band): the upper
    // tiles of one band alone fill the CUs unevenly (528 at 4000 modes: 2.06 rounds of 256)
    const int grp = std::min(nb, TOEP64_GROUP);
    const size_t ystride = (size_t)M * ldg;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Yhat64, &ps->cap_Yhat64, (size_t)nbat * M * G * 8))) return rc;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Y64, &ps->cap_Y64, (size_t)grp * ystride * 8))) return rc;
    if (tgrid.sph) {
        if ((rc = dev_realloc(ctx, (void**)&ps->d_That64, &ps->cap_That64, (size_t)(n1 + 1) * G * G * 8))) return rc;
    } else if ((rc = dev_realloc(ctx, (void**)&ps->d_Qhat, &ps->cap_Qhat, (size_t)(n1 + 1) * G * 8))) {
        return rc;
    }
    if ((rc = dev_realloc(ctx, &ps->d_P, &ps->cap_P, (size_t)nb * packed_size(M) * 8))) return rc;
    if (!ps->toep64_valid) {
        std::vector<int> gmap;
        if ((rc = toep_gmap(ctx, Ng, gmap))) return rc;
        // slab position -> (y << 16) | z (box only)
        std::vector<int> slab((size_t)G, 0);
        if (!tgrid.sph)
            for (int r = 0; r < G; r++) slab[r] = ((r / sv.n3) << 16) | (r % sv.n3);
        std::vector<double> trig;
        trig_table(L, trig);
        // inverse weights W[b][x]: (c_k / L) cos and -(c_k / L) sin, c_0 = c_n1 = 1, else 2
        std::vector<double> W((size_t)nbat * n1);
        for (int b = 0; b < nbat; b++)
            for (int x = 0; x < n1; x++) {
                const int k = (b + 1) >> 1, j = (x * k) % L;
                const double ck = (k == 0 || k == n1) ? 1.0 : 2.0;
                const bool imag = b > 0 && !(b & 1);
                W[(size_t)b * n1 + x] = imag ? -ck * trig[L + j] / L : ck * trig[j] / L;
            }
        if ((rc = dev_realloc(ctx, (void**)&ps->d_gmap64, &ps->cap_gmap64, (size_t)Ng * 4))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_slab, &ps->cap_slab, (size_t)G * 4))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_trig64, &ps->cap_trig64, trig.size() * 8))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_W64, &ps->cap_W64, W.size() * 8))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_Bg64, &ps->cap_Bg64, (size_t)M * ldg * 8))) return rc;
        if ((rc = dev_realloc(ctx, (void**)&ps->d_Bhat64, &ps->cap_Bhat64, (size_t)nbat * M * G * 8))) return rc;
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, ps->d_gmap64, gmap.data(), gmap.size() * 4, hipMemcpyHostToDevice));
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, ps->d_slab, slab.data(), slab.size() * 4, hipMemcpyHostToDevice));
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, ps->d_trig64, trig.data(), trig.size() * 8, hipMemcpyHostToDevice));
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, ps->d_W64, W.data(), W.size() * 8, hipMemcpyHostToDevice));
        const size_t tot = (size_t)M * Ng;
        hipLaunchKernelGGL(grid_b64_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, ps->d_Bw,
                           ps->ldw, ps->d_gmap64, Ng, M, ps->d_Bg64, ldg);
        hipLaunchKernelGGL(bhat64_kernel, dim3((unsigned)((G + 255) / 256), (unsigned)M), dim3(256), 0, ctx->stream,
                           ps->d_Bg64, ldg, n1, G, M, ps->d_trig64, ps->d_Bhat64, nbat);
        OPSEC_HIP_TRY(ctx, hipGetLastError());
        ps->toep64_valid = true;
    }
    const size_t id_lds = (size_t)nbat * n1 * 8;
    OPSEC_HIP_TRY(ctx, hipFuncSetAttribute(n1 <= 32 ? (const void*)idft64_kernel<32> : (const void*)idft64_kernel<TOEP_MAXX>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)id_lds));
    if (tgrid.sph && (rc = that_sph_attr<false>(ctx, n1))) return rc;
    const int ngroups = (nb + grp - 1) / grp;
    // events per band: start, That / Qhat generated, stage 1 and inverse DFT; per group: stage 2
    std::vector<hipEvent_t> ev(3 * (size_t)nb + ngroups);
    for (auto& e : ev) OPSEC_HIP_TRY(ctx, hipEventCreate(&e));
    hipEvent_t* evb = ev.data();
    hipEvent_t* evg = ev.data() + 3 * (size_t)nb;
    const size_t np = (size_t)packed_size(M);
    GA_That A;
    A.qhat = ps->d_Qhat;
    A.slab = ps->d_slab;
    A.G = G;
    A.n3 = sv.n3;
    GA_ThatDense As;
    As.p = ps->d_That64;
    As.ld = G;
    GB_ColMajor Bop;   // batch b: Bhat[b] (B[k = r][n = m] = p[m G + r])
    Bop.p = ps->d_Bhat64;
    Bop.ld = G;
    PutColMajor<double> toYhat{ps->d_Yhat64, (size_t)G, (size_t)M * G};
    GA_ColMajorT A2;   // Bg'^T: A2[m][g] = Bg[m ldg + g]
    A2.p = ps->d_Bg64;
    A2.ld = ldg;
    GB_ColMajorT<double> B2{ps->d_Y64, ldg};
    const dim3 g1 = grid_for(G, M, nbat);
    for (int z0 = 0, gi = 0; z0 < nb; z0 += grp, gi++) {
        const int g = std::min(grp, nb - z0);
        for (int q = 0; q < g; q++) {
            const int z = z0 + q;
            const double* Qf = ctx->d_Q + (size_t)funcs[z] * ctx->nkeys;
            hipEventRecord(evb[3 * z], ctx->stream);
            if (tgrid.sph)
                launch_that_sph<false>(ctx, Qf, G, n1, ps->d_trig64, nullptr, nullptr, 0, 0, ps->d_That64);
            else
                hipLaunchKernelGGL(qhat_kernel, dim3((unsigned)(((n1 + 1) * G + 255) / 256)), dim3(256), 0,
                                   ctx->stream, Qf, n1, G, ps->d_trig64, ps->d_Qhat);
            hipEventRecord(evb[3 * z + 1], ctx->stream);
            const size_t tplane = (size_t)G * G;
            if (tgrid.sph && prec == OPSEC_PREC_FP64)
                hipLaunchKernelGGL((gemm_kernel<double, GA_ThatDense, GB_ColMajor, PutColMajor<double>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, As, tplane, Bop, (size_t)M * G, toYhat, G, M, G);
            else if (tgrid.sph)
                hipLaunchKernelGGL((gemm_kernel<float, GA_ThatDense, GB_ColMajor, PutColMajor<double>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, As, tplane, Bop, (size_t)M * G, toYhat, G, M, G);
            else if (prec == OPSEC_PREC_FP64)
                hipLaunchKernelGGL((gemm_kernel<double, GA_That, GB_ColMajor, PutColMajor<double>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, (size_t)M * G, toYhat, G, M, G);
            else
                hipLaunchKernelGGL((gemm_kernel<float, GA_That, GB_ColMajor, PutColMajor<double>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, (size_t)M * G, toYhat, G, M, G);
            if (n1 <= 32)
                hipLaunchKernelGGL(idft64_kernel<32>, dim3((unsigned)((G + 255) / 256), (unsigned)M), dim3(256), id_lds,
                                   ctx->stream, ps->d_Yhat64, nbat, G, n1, M, ps->d_W64,
                                   ps->d_Y64 + (size_t)q * ystride, ldg);
            else
                hipLaunchKernelGGL(idft64_kernel<TOEP_MAXX>, dim3((unsigned)((G + 255) / 256), (unsigned)M), dim3(256),
                                   id_lds, ctx->stream, ps->d_Yhat64, nbat, G, n1, M, ps->d_W64,
                                   ps->d_Y64 + (size_t)q * ystride, ldg);
            hipEventRecord(evb[3 * z + 2], ctx->stream);
        }
        // stage 2 of the group's bands in one launch: batch q = band z0 + q
        PutPacked outP{(double*)ps->d_P + (size_t)z0 * np, np};
        const dim3 g2 = grid_for(M, M, g);
        if (prec == OPSEC_PREC_FP64)
            hipLaunchKernelGGL((gemm_kernel<double, GA_ColMajorT, GB_ColMajorT<double>, PutPacked, true>), g2,
                               dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, ystride, outP, M, M, Ng);
        else
            hipLaunchKernelGGL((gemm_kernel<float, GA_ColMajorT, GB_ColMajorT<double>, PutPacked, true>), g2,
                               dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, ystride, outP, M, M, Ng);
        hipEventRecord(evg[gi], ctx->stream);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection launch: %s", hipGetErrorString(e));
    hipError_t es = hipEventSynchronize(evg[ngroups - 1]);
    if (!rc && es != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection: %s", hipGetErrorString(es));
    float tg = 0, t1 = 0, t2 = 0;
    if (!rc) {
        for (int z = 0; z < nb; z++) {
            float a, b;
            hipEventElapsedTime(&a, evb[3 * z], evb[3 * z + 1]);
            hipEventElapsedTime(&b, evb[3 * z + 1], evb[3 * z + 2]);
            tg += a;
            t1 += b;
        }
        for (int gi = 0; gi < ngroups; gi++) {
            const int zl = std::min(nb, (gi + 1) * grp) - 1;   // last band of the group
            float c;
            hipEventElapsedTime(&c, evb[3 * zl + 2], evg[gi]);
            t2 += c;
        }
    }
    for (auto& x : ev) hipEventDestroy(x);
    if (rc) return rc;
    ps->last_ms_gen = tg;
    ps->last_ms_gemm1 = t1;     // stage-1 GEMMs and the inverse DFT
    ps->last_ms_gemm2 = t2;
    ps->last_launches = nb;
    ps->last_launches2 = ngroups;
    ps->last_ms_stage1 = tg + t1;
    ps->last_ms_stage2 = t2;
    ps->nb = nb;
    ps->prec = prec;
    return OPSEC_OK;
}

// Stage 1 + stage 2 for nb functions (device pair tables), results to d_P.
// Dense FP64 / FP32 projection of the bands funcs[0..nb) in groups: stage 1
// (all bands of a group in one launch, grid.z = band) keeps the group's
// products C_p = T_p B' (N x M each), so a group holds at most
// DENSE_GROUP_BYTES of them (all 40 bands of the 32^3 box fit; the 50 bands
// of the 65536-cell shell at 6000 modes take seven groups).
constexpr size_t DENSE_GROUP_BYTES = (size_t)48 << 30;

static int run_project(opsec_ctx* ctx, int nb, const int* d_funcs, const int* funcs, int prec) {
    if (prec == OPSEC_PREC_BF16X3 && ctx->projection == OPSEC_PROJ_TOEPLITZ) return run_project_toeplitz(ctx, nb, funcs);
    if (ctx->projection == OPSEC_PROJ_TOEPLITZ) return run_project_toeplitz64(ctx, nb, funcs, prec);
    if (prec == OPSEC_PREC_BF16X3) return run_project_bf16x3(ctx, nb, funcs);
    ProjState* ps = ctx->proj;
    const int N = ctx->ncells, M = ctx->nmodes;
    const size_t esize = (prec == OPSEC_PREC_FP64) ? 8 : 4;
    const size_t ldc = ps->ldw;   // stage-1 products use the same padded pitch as B'
    const size_t per_band = ldc * M * esize;
    const int grp = (int)std::max<size_t>(1, std::min<size_t>(nb, DENSE_GROUP_BYTES / per_band));
    int rc;
    if ((rc = dev_realloc(ctx, &ps->d_C, &ps->cap_C, (size_t)grp * per_band))) return rc;
    if ((rc = dev_realloc(ctx, &ps->d_P, &ps->cap_P, (size_t)nb * packed_size(M) * 8))) return rc;
    const bool sph = ctx->sv.coordsys == OPSEC_COORD_SPHERICAL;
    GA_PairTable A;
    A.Q = ctx->d_Q;
    A.cellidx = ctx->d_cellidx;
    A.Ny = ctx->sv.n2;
    A.Nz = ctx->sv.n3;
    A.narcs = ctx->sv.n1 * ctx->sv.n2;
    A.qstride = ctx->nkeys;
    GA_PairTableSph As;
    As.Q = A.Q;
    As.cellidx = A.cellidx;
    As.Ny = A.Ny;
    As.Nz = A.Nz;
    As.narcs = A.narcs;
    As.qstride = A.qstride;
    GB_ColMajor Bop;
    Bop.p = ps->d_Bw;
    Bop.ld = ps->ldw;
    GA_ColMajorT A2;
    A2.p = ps->d_Bw;
    A2.ld = ps->ldw;
    const size_t np = (size_t)packed_size(M);
    const int ngroups = (nb + grp - 1) / grp;
    std::vector<hipEvent_t> ev(2 * (size_t)ngroups + 1);
    for (auto& e : ev) OPSEC_HIP_TRY(ctx, hipEventCreate(&e));
    hipEventRecord(ev[0], ctx->stream);
    for (int gi = 0; gi < ngroups; gi++) {
        const int z0 = gi * grp, g = std::min(grp, nb - z0);
        A.funcs = d_funcs + z0;
        As.funcs = d_funcs + z0;
        dim3 g1 = grid_for(N, M, g), g2 = grid_for(M, M, g);
        PutPacked outP{(double*)ps->d_P + (size_t)z0 * np, np};
        if (prec == OPSEC_PREC_FP64) {
            PutColMajor<double> outC{(double*)ps->d_C, ldc, ldc * M};
            if (sph)
                hipLaunchKernelGGL((gemm_kernel<double, GA_PairTableSph, GB_ColMajor, PutColMajor<double>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, As, 0, Bop, 0, outC, N, M, N);
            else
                hipLaunchKernelGGL((gemm_kernel<double, GA_PairTable, GB_ColMajor, PutColMajor<double>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
            hipEventRecord(ev[2 * gi + 1], ctx->stream);
            GB_ColMajorT<double> B2{(const double*)ps->d_C, ldc};
            hipLaunchKernelGGL((gemm_kernel<double, GA_ColMajorT, GB_ColMajorT<double>, PutPacked, true>), g2,
                               dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, ldc * M, outP, M, M, N);
        } else {
            PutColMajor<float> outC{(float*)ps->d_C, ldc, ldc * M};
            if (sph)
                hipLaunchKernelGGL((gemm_kernel<float, GA_PairTableSph, GB_ColMajor, PutColMajor<float>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, As, 0, Bop, 0, outC, N, M, N);
            else
                hipLaunchKernelGGL((gemm_kernel<float, GA_PairTable, GB_ColMajor, PutColMajor<float>, false>), g1,
                                   dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
            hipEventRecord(ev[2 * gi + 1], ctx->stream);
            GB_ColMajorT<float> B2{(const float*)ps->d_C, ldc};
            hipLaunchKernelGGL((gemm_kernel<float, GA_ColMajorT, GB_ColMajorT<float>, PutPacked, true>), g2,
                               dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, ldc * M, outP, M, M, N);
        }
        hipEventRecord(ev[2 * gi + 2], ctx->stream);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection launch: %s", hipGetErrorString(e));
    hipError_t es = hipEventSynchronize(ev[2 * ngroups]);
    if (!rc && es != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection: %s", hipGetErrorString(es));
    float t1 = 0, t2 = 0;
    if (!rc)
        for (int gi = 0; gi < ngroups; gi++) {
            float a, b;
            hipEventElapsedTime(&a, ev[2 * gi], ev[2 * gi + 1]);
            hipEventElapsedTime(&b, ev[2 * gi + 1], ev[2 * gi + 2]);
            t1 += a;
            t2 += b;
        }
    for (auto& x : ev) hipEventDestroy(x);
    if (rc) return rc;
    ps->last_ms_stage1 = t1;
    ps->last_ms_stage2 = t2;
    ps->nb = nb;
    ps->prec = prec;
    return OPSEC_OK;
}

// A new cell set invalidates the uploaded modes and everything derived from them.
void invalidate_modes(opsec_ctx* ctx) {
    ctx->nmodes = 0;
    if (!ctx->proj) return;
    ctx->proj->bp_valid = false;
    ctx->proj->toep_valid = false;
    ctx->proj->toep64_valid = false;
    ctx->proj->nb = 0;
}

void free_proj(opsec_ctx* ctx) {
    if (!ctx->proj) return;
    dev_free((void**)&ctx->proj->d_Bw);
    dev_free(&ctx->proj->d_C);
    dev_free(&ctx->proj->d_P);
    dev_free((void**)&ctx->proj->d_Bp);
    dev_free((void**)&ctx->proj->d_Sp);
    dev_free((void**)&ctx->proj->d_Cp);
    dev_free((void**)&ctx->proj->d_Qhl);
    dev_free((void**)&ctx->proj->d_Ppart);
    dev_free((void**)&ctx->proj->d_cpk);
    dev_free((void**)&ctx->proj->d_gmap);
    dev_free((void**)&ctx->proj->d_Bg);
    dev_free((void**)&ctx->proj->d_Bhat);
    dev_free((void**)&ctx->proj->d_That);
    dev_free((void**)&ctx->proj->d_Qhat);
    dev_free((void**)&ctx->proj->d_Yhat);
    dev_free((void**)&ctx->proj->d_Cg);
    dev_free((void**)&ctx->proj->d_trig);
    dev_free((void**)&ctx->proj->d_W);
    dev_free((void**)&ctx->proj->d_gmap64);
    dev_free((void**)&ctx->proj->d_slab);
    dev_free((void**)&ctx->proj->d_trig64);
    dev_free((void**)&ctx->proj->d_W64);
    dev_free((void**)&ctx->proj->d_Bg64);
    dev_free((void**)&ctx->proj->d_Bhat64);
    dev_free((void**)&ctx->proj->d_Yhat64);
    dev_free((void**)&ctx->proj->d_Y64);
    dev_free((void**)&ctx->proj->d_That64);
    delete ctx->proj;
    ctx->proj = nullptr;
}

}  // namespace opsec

using namespace opsec;

extern "C" {


int opsec_modes_upload(opsec_ctx* ctx, const double* B, int ncells, int nmodes) {
    if (!ctx) return OPSEC_EINVAL;
    if (ctx->ncells == 0) return set_error(ctx, OPSEC_ESTATE, "no survey uploaded");
    OPSEC_CHECK_ARG(ctx, B && ncells == ctx->ncells, "mode vectors have %d cells, survey has %d", ncells, ctx->ncells);
    OPSEC_CHECK_ARG(ctx, nmodes >= 1 && nmodes <= ncells, "Nmodes %d out of range", nmodes);
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!ctx->proj) ctx->proj = new ProjState;
    ProjState* ps = ctx->proj;
    size_t bytes = (size_t)ncells * nmodes * 8;
    int rc;
    dev_free((void**)&ctx->d_B);
    size_t capB = 0;
    if ((rc = dev_realloc(ctx, (void**)&ctx->d_B, &capB, bytes))) return rc;
    ps->ldw = (size_t)ncells + 32;
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Bw, &ps->cap_Bw, ps->ldw * nmodes * 8))) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_B, B, bytes, hipMemcpyHostToDevice, ctx->stream));
    size_t tot = (size_t)ncells * nmodes;
    hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, ctx->d_B,
                       ctx->d_nbar, ncells, nmodes, ps->d_Bw, ps->ldw);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->nmodes = nmodes;
    ps->bp_valid = false;
    ps->toep_valid = false;
    ps->toep64_valid = false;
    return OPSEC_OK;
}

int opsec_project(opsec_ctx* ctx, int nf, const int* funcs, int prec, double* P_packed) {
    if (!ctx) return OPSEC_EINVAL;
    if (!ctx->proj || ctx->nmodes == 0) return set_error(ctx, OPSEC_ESTATE, "no modes uploaded");
    OPSEC_CHECK_ARG(ctx, nf >= 1 && funcs, "no functions");
    OPSEC_CHECK_ARG(ctx, prec >= 0 && prec <= 2, "unknown precision %d", prec);
    for (int t = 0; t < nf; t++) {
        OPSEC_CHECK_ARG(ctx, funcs[t] >= 0 && funcs[t] < (int)ctx->funcs.size(), "function %d out of range", funcs[t]);
        if (!ctx->Q_valid[funcs[t]]) return set_error(ctx, OPSEC_ESTATE, "pair table %d not built", funcs[t]);
    }
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    int* d_funcs = nullptr;
    DevScratch tmp;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&d_funcs, nf * sizeof(int)));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_funcs, funcs, nf * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    int rc = run_project(ctx, nf, d_funcs, funcs, prec);
    hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    ctx->proj->funcs.assign(funcs, funcs + nf);
    if (P_packed) {
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(P_packed, ctx->proj->d_P, (size_t)nf * packed_size(ctx->nmodes) * 8,
                                          hipMemcpyDeviceToHost, ctx->stream));
        OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    return OPSEC_OK;
}

// Timing of the last projection's two stages (ms), for the benchmark.
int opsec_diag_project_times(opsec_ctx* ctx, float* s1, float* s2) {
    if (!ctx || !ctx->proj) return OPSEC_EINVAL;
    *s1 = ctx->proj->last_ms_stage1;
    *s2 = ctx->proj->last_ms_stage2;
    return OPSEC_OK;
}

// Tile shape of the split-BF16x3 GEMM (0: 128 x 128, 1: 256 x 128, 2: 256 x 256
// with double-buffered LDS), for tuning.
int opsec_diag_set_nt_tile(int v) {
    if (v < 0 || v > 2) return OPSEC_EINVAL;
    g_nt_tile = v;
    return OPSEC_OK;
}

// Split-K factor of the split-BF16x3 stage 2 (its upper-triangle tiles alone
// fill about half of the CUs), for tuning.
int opsec_diag_set_stage2_split(int v) {
    if (v < 1 || v > 16) return OPSEC_EINVAL;
    g_s2_split = v;
    return OPSEC_OK;
}

// Split-K factor of the Toeplitz path's split-BF16x3 stage 2 (eight bands per
// launch), for tuning.
int opsec_diag_set_toeplitz_stage2_split(int v) {
    if (v < 1 || v > 16) return OPSEC_EINVAL;
    g_s2_split_toep = v;
    return OPSEC_OK;
}

// Tile-group edge of the split-BF16x3 stage 2, for tuning.
int opsec_diag_set_stage2_group(int v) {
    if (v < 1 || v > 128) return OPSEC_EINVAL;
    g_s2_group = v;
    return OPSEC_OK;
}

int opsec_set_projection(opsec_ctx* ctx, int method) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, method == OPSEC_PROJ_DENSE || method == OPSEC_PROJ_TOEPLITZ, "unknown projection method %d",
                    method);
    ctx->projection = method;
    return OPSEC_OK;
}

// Stage-1 A operand of the split-BF16x3 projection, for tuning: generated
// from the pair table (1) or read from S_p planes (0).
int opsec_diag_set_nt_gen(int v) {
    if (v != 0 && v != 1) return OPSEC_EINVAL;
    g_nt_gen = v;
    return OPSEC_OK;
}

// Operand tiles of the 256 x 256 split-BF16x3 GEMM: through VGPRs (0) or by
// global-to-LDS DMA loads, issued at a step's start (1) or in two halves (2).
int opsec_diag_set_g3_dl(int v) {
    if (v < 0 || v > 2) return OPSEC_EINVAL;
    g_g3_dl = v;
    return OPSEC_OK;
}

// Tile-order knobs of the split-BF16x3 GEMM, for tuning: group edge (>= 1)
// and XCD remapping (0 or 1).
int opsec_diag_set_nt_order(int group, int remap) {
    if (group < 1 || (remap != 0 && remap != 1)) return OPSEC_EINVAL;
    g_nt_group = group;
    g_nt_remap = remap;
    return OPSEC_OK;
}

// Per-kernel times of the last split-BF16x3 projection, summed over its bands
// (ms): S_p plane generation, stage-1 GEMM, stage-2 GEMM; *launches = bands.
int opsec_diag_project_kernel_ms(opsec_ctx* ctx, float* gen, float* gemm1, float* gemm2, int* launches) {
    if (!ctx || !ctx->proj) return OPSEC_EINVAL;
    *gen = ctx->proj->last_ms_gen;
    *gemm1 = ctx->proj->last_ms_gemm1;
    *gemm2 = ctx->proj->last_ms_gemm2;
    *launches = ctx->proj->last_launches;
    return OPSEC_OK;
}

// Launch counts of the last split-BF16x3 projection: stage 1, stage 2.
int opsec_diag_project_launches(opsec_ctx* ctx, int* stage1, int* stage2) {
    if (!ctx || !ctx->proj) return OPSEC_EINVAL;
    *stage1 = ctx->proj->last_launches;
    *stage2 = ctx->proj->last_launches2;
    return OPSEC_OK;
}

int opsec_project_dense(opsec_ctx* ctx, int n, const double* S, int nmodes, const double* B, int prec,
                        double* P_packed) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, n >= 1 && nmodes >= 1 && nmodes <= n && S && B && P_packed, "bad projection arguments");
    OPSEC_CHECK_ARG(ctx, prec >= 0 && prec <= 2, "unknown precision %d", prec);
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int N = n, M = nmodes;
    double *dS = nullptr, *dB = nullptr, *dP = nullptr;
    void* dC = nullptr;
    size_t esize = prec == OPSEC_PREC_FP64 ? 8 : 4;
    DevScratch tmp;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dS, (size_t)N * N * 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dB, (size_t)N * M * 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dC, (size_t)N * M * esize));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dP, (size_t)packed_size(M) * 8));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dS, S, (size_t)N * N * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dB, B, (size_t)N * M * 8, hipMemcpyHostToDevice, ctx->stream));
    GA_RowMajor A;
    A.p = dS;
    A.ld = N;
    GB_ColMajor Bop;
    Bop.p = dB;
    Bop.ld = N;
    GA_ColMajorT A2;
    A2.p = dB;
    A2.ld = N;
    PutPacked outP{dP, (size_t)packed_size(M)};
    dim3 g1 = grid_for(N, M, 1), g2 = grid_for(M, M, 1);
    if (prec == OPSEC_PREC_FP64) {
        PutColMajor<double> outC{(double*)dC, (size_t)N, 0};
        hipLaunchKernelGGL((gemm_kernel<double, GA_RowMajor, GB_ColMajor, PutColMajor<double>, false>), g1,
                           dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
        GB_ColMajorT<double> B2{(const double*)dC, (size_t)N};
        hipLaunchKernelGGL((gemm_kernel<double, GA_ColMajorT, GB_ColMajorT<double>, PutPacked, true>), g2,
                           dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, 0, outP, M, M, N);
    } else if (prec == OPSEC_PREC_FP32) {
        PutColMajor<float> outC{(float*)dC, (size_t)N, 0};
        hipLaunchKernelGGL((gemm_kernel<float, GA_RowMajor, GB_ColMajor, PutColMajor<float>, false>), g1,
                           dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
        GB_ColMajorT<float> B2{(const float*)dC, (size_t)N};
        hipLaunchKernelGGL((gemm_kernel<float, GA_ColMajorT, GB_ColMajorT<float>, PutPacked, true>), g2,
                           dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, 0, outP, M, M, N);
    } else {
        const size_t Kp = round_up(N, 256), Mp = round_up(M, 256), Pp = plane_pitch(Kp);
        __bf16 *Sp = nullptr, *Bp = nullptr, *Cp = nullptr;
        DevScratch planes;
        OPSEC_HIP_TRY(ctx, planes.alloc(&Sp, 2 * Kp * Pp * sizeof(__bf16)));
        OPSEC_HIP_TRY(ctx, planes.alloc(&Bp, 2 * Mp * Pp * sizeof(__bf16)));
        OPSEC_HIP_TRY(ctx, planes.alloc(&Cp, 2 * Mp * Pp * sizeof(__bf16)));
        int rc2 = split_planes(ctx, dS, (size_t)N, N, N, Sp, (int)Kp, Pp);
        // B is row-major nmodes x n here: row = mode, cells contiguous
        if (!rc2) rc2 = split_planes(ctx, dB, (size_t)N, M, N, Bp, (int)Mp, Pp);
        if (!rc2) {
            launch_nt(ctx, Sp, (int)Kp, Bp, (int)Mp, Pp, (int)Kp, EpiPlanes{Cp, Cp + Mp * Pp, Pp});
            launch_nt(ctx, Bp, (int)Mp, Cp, (int)Mp, Pp, (int)Kp, EpiPacked{dP, M});
        }
        hipStreamSynchronize(ctx->stream);
        if (rc2) {
            return rc2;
        }
    }
    hipError_t e = hipGetLastError();
    int rc = OPSEC_OK;
    if (e != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "projection launch: %s", hipGetErrorString(e));
    if (!rc) {
        hipError_t e2 = hipMemcpyAsync(P_packed, dP, (size_t)packed_size(M) * 8, hipMemcpyDeviceToHost, ctx->stream);
        hipError_t e3 = hipStreamSynchronize(ctx->stream);
        if (e2 != hipSuccess || e3 != hipSuccess)
            rc = set_error(ctx, OPSEC_EHIP, "projection copy-back: %s", hipGetErrorString(e2 ? e2 : e3));
    }
    return rc;
}

}  // extern "C"

namespace opsec {
// Device-resident results of the last opsec_project call (nb packed matrices).
void* proj_device_results(opsec_ctx* ctx, int* nb, int* M) {
    if (!ctx->proj || !ctx->proj->d_P || ctx->proj->nb == 0) return nullptr;
    *nb = ctx->proj->nb;
    *M = ctx->nmodes;
    return ctx->proj->d_P;
}
}  // namespace opsec
