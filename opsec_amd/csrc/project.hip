// K3: projection of the band matrices onto the KL modes, P_p = B^T C_{,p} B.
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
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"
#include "gemm.h"

namespace opsec {

struct ProjState {
    double* d_Bw = nullptr;      // B' = diag(sqrt(Nbar)) B, ncells x nmodes column-major
    size_t cap_Bw = 0;
    void* d_C = nullptr;         // stage-1 products, nb x ncells x nmodes (double or float)
    size_t cap_C = 0;
    void* d_P = nullptr;         // packed results, nb x M(M+1)/2 (double)
    size_t cap_P = 0;
    int nb = 0;                  // bands in d_P
    int prec = 0;
    std::vector<int> funcs;
    float last_ms_stage1 = 0, last_ms_stage2 = 0;
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
struct GA_PairTable : SrcPairTable {
    size_t qstride;
    const int* funcs;
    __device__ __forceinline__ void p_shift(int z, size_t) { Q += (size_t)funcs[z] * qstride; }
};
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
    __device__ __forceinline__ void p_shift(int z, size_t s) { p += (size_t)z * s; }
};

// ---- split-BF16x3 GEMM -----------------------------------------------------------
//
// Tile 128 x 128 x 32.  LDS holds hi and lo bf16 tiles in [m][k] order (k
// contiguous, 8 bf16 per lane per fragment).  Accumulation in FP32.

constexpr int BKB = 32;
constexpr int BPAD = 8;   // bf16 padding per LDS row

__device__ __forceinline__ void split_bf16(double x, __bf16& hi, __bf16& lo) {
    float f = (float)x;
    hi = (__bf16)f;
    lo = (__bf16)(f - (float)hi);
}

template <typename Src>
__device__ __forceinline__ void tile_fetch32(const Src& s, int i0, int k0, int nI, int nK, double (&r)[16]) {
    const int t = threadIdx.x;
    if (Src::k_contig) {
        int i = i0 + (t >> 1), kb = k0 + (t & 1) * 16;
#pragma unroll
        for (int e = 0; e < 16; e++) r[e] = (i < nI && kb + e < nK) ? s.get(i, kb + e) : 0.0;
    } else {
        int k = k0 + (t >> 3), ib = i0 + (t & 7) * 16;
#pragma unroll
        for (int e = 0; e < 16; e++) r[e] = (k < nK && ib + e < nI) ? s.get(ib + e, k) : 0.0;
    }
}

template <typename Src>
__device__ __forceinline__ void tile_store32(__bf16 (*hi)[BKB + BPAD], __bf16 (*lo)[BKB + BPAD], const double (&r)[16]) {
    const int t = threadIdx.x;
    if (Src::k_contig) {
        int i = t >> 1, kb = (t & 1) * 16;
#pragma unroll
        for (int e = 0; e < 16; e++) split_bf16(r[e], hi[i][kb + e], lo[i][kb + e]);
    } else {
        int k = t >> 3, ib = (t & 7) * 16;
#pragma unroll
        for (int e = 0; e < 16; e++) split_bf16(r[e], hi[ib + e][k], lo[ib + e][k]);
    }
}

template <typename SrcA, typename SrcB, typename Out, bool UPPER>
__global__ __launch_bounds__(GT_THREADS) void gemm_bf16x3_kernel(SrcA A, size_t a_batch, SrcB B, size_t b_batch, Out O,
                                                                 int M, int N, int K) {
    __shared__ __bf16 Ah[GT_M][BKB + BPAD], Al[GT_M][BKB + BPAD];
    __shared__ __bf16 Bh[GT_N][BKB + BPAD], Bl[GT_N][BKB + BPAD];
    const int m0 = blockIdx.x * GT_M, n0 = blockIdx.y * GT_N;
    if (UPPER && m0 > n0 + GT_N - 1) return;
    A.p_shift(blockIdx.z, a_batch);
    B.p_shift(blockIdx.z, b_batch);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4_t{0, 0, 0, 0};
    double ra[16], rb[16];
    tile_fetch32(A, m0, 0, M, K, ra);
    tile_fetch32(B, n0, 0, N, K, rb);
    for (int k0 = 0; k0 < K; k0 += BKB) {
        __syncthreads();
        tile_store32<SrcA>(Ah, Al, ra);
        tile_store32<SrcB>(Bh, Bl, rb);
        __syncthreads();
        if (k0 + BKB < K) {
            tile_fetch32(A, m0, k0 + BKB, M, K, ra);
            tile_fetch32(B, n0, k0 + BKB, N, K, rb);
        }
        const int kq = 8 * (lane >> 4);
        bf8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ah[i] = *(const bf8_t*)&Ah[wm + i * 16 + (lane & 15)][kq];
            al[i] = *(const bf8_t*)&Al[wm + i * 16 + (lane & 15)][kq];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            bh[j] = *(const bf8_t*)&Bh[wn + j * 16 + (lane & 15)][kq];
            bl[j] = *(const bf8_t*)&Bl[wn + j * 16 + (lane & 15)][kq];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + wn + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                O.put(blockIdx.z, m, n, M, N, (double)acc[i][j][r]);
            }
        }
}

// B' = diag(w) B with w = sqrt(Nbar)
__global__ void scale_rows_kernel(const double* B, const double* nbar, int N, int M, double* Bw) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t tot = (size_t)N * M;
    if (i >= tot) return;
    int a = (int)(i % N);
    Bw[i] = B[i] * sqrt(nbar[a]);
}

// ---- host --------------------------------------------------------------------------

static dim3 grid_for(int M, int N, int nb) {
    return dim3((M + GT_M - 1) / GT_M, (N + GT_N - 1) / GT_N, nb);
}

// Stage 1 + stage 2 for nb functions (device pair tables), results to d_P.
static int run_project(opsec_ctx* ctx, int nb, const int* d_funcs, int prec) {
    ProjState* ps = ctx->proj;
    const int N = ctx->ncells, M = ctx->nmodes;
    const size_t esize = (prec == OPSEC_PREC_FP64) ? 8 : 4;
    int rc;
    if ((rc = dev_realloc(ctx, &ps->d_C, &ps->cap_C, (size_t)nb * N * M * esize))) return rc;
    if ((rc = dev_realloc(ctx, &ps->d_P, &ps->cap_P, (size_t)nb * packed_size(M) * 8))) return rc;
    GA_PairTable A;
    A.Q = ctx->d_Q;
    A.cellidx = ctx->d_cellidx;
    A.Ny = ctx->sv.n2;
    A.Nz = ctx->sv.n3;
    A.qstride = ctx->nkeys;
    A.funcs = d_funcs;
    GB_ColMajor Bop;
    Bop.p = ps->d_Bw;
    Bop.ld = N;
    hipEvent_t e0, e1, e2;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventCreate(&e2);
    hipEventRecord(e0, ctx->stream);
    dim3 g1 = grid_for(N, M, nb), g2 = grid_for(M, M, nb);
    GA_ColMajorT A2;
    A2.p = ps->d_Bw;
    A2.ld = N;
    PutPacked outP{(double*)ps->d_P, (size_t)packed_size(M)};
    if (prec == OPSEC_PREC_FP64) {
        PutColMajor<double> outC{(double*)ps->d_C, (size_t)N, (size_t)N * M};
        hipLaunchKernelGGL((gemm_kernel<double, GA_PairTable, GB_ColMajor, PutColMajor<double>, false>), g1,
                           dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
        hipEventRecord(e1, ctx->stream);
        GB_ColMajorT<double> B2{(const double*)ps->d_C, (size_t)N};
        hipLaunchKernelGGL((gemm_kernel<double, GA_ColMajorT, GB_ColMajorT<double>, PutPacked, true>), g2,
                           dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, (size_t)N * M, outP, M, M, N);
    } else if (prec == OPSEC_PREC_FP32) {
        PutColMajor<float> outC{(float*)ps->d_C, (size_t)N, (size_t)N * M};
        hipLaunchKernelGGL((gemm_kernel<float, GA_PairTable, GB_ColMajor, PutColMajor<float>, false>), g1,
                           dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
        hipEventRecord(e1, ctx->stream);
        GB_ColMajorT<float> B2{(const float*)ps->d_C, (size_t)N};
        hipLaunchKernelGGL((gemm_kernel<float, GA_ColMajorT, GB_ColMajorT<float>, PutPacked, true>), g2,
                           dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, (size_t)N * M, outP, M, M, N);
    } else {
        PutColMajor<float> outC{(float*)ps->d_C, (size_t)N, (size_t)N * M};
        hipLaunchKernelGGL((gemm_bf16x3_kernel<GA_PairTable, GB_ColMajor, PutColMajor<float>, false>), g1,
                           dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
        hipEventRecord(e1, ctx->stream);
        GB_ColMajorT<float> B2{(const float*)ps->d_C, (size_t)N};
        hipLaunchKernelGGL((gemm_bf16x3_kernel<GA_ColMajorT, GB_ColMajorT<float>, PutPacked, true>), g2,
                           dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, (size_t)N * M, outP, M, M, N);
    }
    hipEventRecord(e2, ctx->stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ctx, OPSEC_EHIP, "projection launch: %s", hipGetErrorString(e));
    OPSEC_HIP_TRY(ctx, hipEventSynchronize(e2));
    hipEventElapsedTime(&ps->last_ms_stage1, e0, e1);
    hipEventElapsedTime(&ps->last_ms_stage2, e1, e2);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    ps->nb = nb;
    ps->prec = prec;
    return OPSEC_OK;
}

}  // namespace opsec

using namespace opsec;

extern "C" {

void opsec_internal_free_proj(opsec_ctx* ctx) {
    if (!ctx->proj) return;
    dev_free((void**)&ctx->proj->d_Bw);
    dev_free(&ctx->proj->d_C);
    dev_free(&ctx->proj->d_P);
    delete ctx->proj;
    ctx->proj = nullptr;
}

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
    if ((rc = dev_realloc(ctx, (void**)&ps->d_Bw, &ps->cap_Bw, bytes))) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_B, B, bytes, hipMemcpyHostToDevice, ctx->stream));
    size_t tot = (size_t)ncells * nmodes;
    hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, ctx->d_B,
                       ctx->d_nbar, ncells, nmodes, ps->d_Bw);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->nmodes = nmodes;
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
    OPSEC_HIP_TRY(ctx, hipMalloc(&d_funcs, nf * sizeof(int)));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_funcs, funcs, nf * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    int rc = run_project(ctx, nf, d_funcs, prec);
    hipStreamSynchronize(ctx->stream);
    hipFree(d_funcs);
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
int opsec_internal_project_times(opsec_ctx* ctx, float* s1, float* s2) {
    if (!ctx || !ctx->proj) return OPSEC_EINVAL;
    *s1 = ctx->proj->last_ms_stage1;
    *s2 = ctx->proj->last_ms_stage2;
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
    OPSEC_HIP_TRY(ctx, hipMalloc(&dS, (size_t)N * N * 8));
    OPSEC_HIP_TRY(ctx, hipMalloc(&dB, (size_t)N * M * 8));
    OPSEC_HIP_TRY(ctx, hipMalloc(&dC, (size_t)N * M * esize));
    OPSEC_HIP_TRY(ctx, hipMalloc(&dP, (size_t)packed_size(M) * 8));
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
        PutColMajor<float> outC{(float*)dC, (size_t)N, 0};
        hipLaunchKernelGGL((gemm_bf16x3_kernel<GA_RowMajor, GB_ColMajor, PutColMajor<float>, false>), g1,
                           dim3(GT_THREADS), 0, ctx->stream, A, 0, Bop, 0, outC, N, M, N);
        GB_ColMajorT<float> B2{(const float*)dC, (size_t)N};
        hipLaunchKernelGGL((gemm_bf16x3_kernel<GA_ColMajorT, GB_ColMajorT<float>, PutPacked, true>), g2,
                           dim3(GT_THREADS), 0, ctx->stream, A2, 0, B2, 0, outP, M, M, N);
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
    hipFree(dS);
    hipFree(dB);
    hipFree(dC);
    hipFree(dP);
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
