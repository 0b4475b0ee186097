// Fisher matrix, bias and quadratic estimates from the packed projected
// matrices, then the estimate post-processing (src/estimate-mpi.cpp:480-654).
//
// Over the packed upper triangle k = j(j+1)/2 + i (i <= j), with
// C = diag(1 + lambda) and w_k = (i == j ? 1/2 : 1) / (C_i C_j):
//   F_mn = sum_k w_k D_m[k] D_n[k]
//   f_n  = sum_k w_k D_n[k] R[k],      R[k] = delta_ij C_i - sum_n p_n D_n[k]
//   q_n  = sum_k w_k D_n[k] y_i y_j
// i.e. one Gram product G = X D^T with X = [w D_0 .. w D_{nb-1}, w R, w Y]
// (nb + 2 rows, padded to 48, or to 64 for nb > 46) and D^T (padded likewise),
// run on MFMA (FP64 16x16x4, or FP32 16x16x4 for the FP32 traces).  Each
// workgroup writes its partial Gram matrix; a second kernel adds the partials
// in workgroup order, so F, f and q are the same bits on every run.  The
// 48 x 48 (64 x 64) result is tiny; the inverse / matrix square root / window
// normalisation on nb x nb runs on the host in FP64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "gemm.h"

namespace opsec {

// NF 16-row fragments per side: the padded Gram size is 16 NF (48 or 64).
// Packed entries per workgroup iteration (one per thread while loading): 256
// for FP32 traces at 48, 128 for FP64 at 48, 48 for 64 (LDS stays within
// 100 KB).
template <typename T, int NF>
struct FShape {
    static constexpr int FG = 16 * NF;
    static constexpr int KT = NF <= 3 ? (sizeof(T) == 4 ? 256 : 128) : 48;
};
constexpr int FG_MAX = 64;
// Partial Gram matrices: one per workgroup, at most FISHER_MAX_PARTS
// workgroups (with one 100 KB-LDS workgroup resident per CU, two per CU keep
// every CU busy through the grid-stride loop).
constexpr int FISHER_MAX_PARTS = 512;
inline int fisher_parts(const opsec_ctx* ctx, long long tiles) {
    return (int)std::min<long long>(tiles, std::min(FISHER_MAX_PARTS, 2 * std::max(1, ctx->num_cus)));
}

template <typename T>
struct FMfma;
template <>
struct FMfma<double> {
    typedef d4_t acc_t;
    __device__ static acc_t mma(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct FMfma<float> {
    typedef f4_t acc_t;
    __device__ static acc_t mma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

// column j of packed index k (k = j(j+1)/2 + i, 0 <= i <= j)
__device__ __forceinline__ int packed_col(long long k) {
    long long j = (long long)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while (j * (j + 1) / 2 > k) j--;
    while ((j + 1) * (j + 2) / 2 <= k) j++;
    return (int)j;
}

template <typename T, int NF>
__global__ __launch_bounds__(256) void fisher_gram_kernel(const double* D, long long K, int nb, const double* lambda,
                                                          const double* y, const double* prior, double* Gpart) {
    typedef typename FMfma<T>::acc_t acc_t;
    constexpr int FG = FShape<T, NF>::FG, FK_TILE = FShape<T, NF>::KT;
    // FP32 MFMA partial sums are closed at the end of every tile and carried
    // in FP64 registers, so each FP32 accumulation chain is FK_TILE / 4 terms
    constexpr bool fp32 = sizeof(T) == 4;
    __shared__ T X[FG][FK_TILE + 4];     // scaled rows
    __shared__ T Dt[FG][FK_TILE + 4];    // plain rows
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    acc_t acc[NF][NF];
    double acc64[NF][NF][4];
    for (int i = 0; i < NF; i++)
        for (int j = 0; j < NF; j++) {
            acc[i][j] = acc_t{0, 0, 0, 0};
            for (int r = 0; r < 4; r++) acc64[i][j][r] = 0.0;
        }
    __shared__ double sp[FG];   // prior band powers
    for (int n = threadIdx.x; n < FG; n += blockDim.x) sp[n] = n < nb ? prior[n] : 0.0;
    // padding rows never change: zero them once
    for (int e = threadIdx.x; e < FG * FK_TILE; e += blockDim.x) {
        const int row = e / FK_TILE, t = e % FK_TILE;
        if (row >= nb) Dt[row][t] = (T)0;
        if (row >= nb + 2) X[row][t] = (T)0;
    }
    for (long long k0 = (long long)blockIdx.x * FK_TILE; k0 < K; k0 += (long long)gridDim.x * FK_TILE) {
        __syncthreads();   // the previous tile's MFMAs are done with X and Dt
        // one thread per packed entry: its weight, the rows of D (coalesced
        // along k, loads independent of each other), the residual
        // R = delta_ij C_i - sum_n p_n D_n in FP64 from the FP64 inputs, and
        // the scaled rows
        if (threadIdx.x < FK_TILE) {
            const int t = threadIdx.x;
            const long long k = k0 + t;
            // all FG - 2 row loads are issued first (rows at and beyond nb
            // load row nb - 1 and count as 0), so one tile costs one memory
            // latency
            const long long kc = k < K ? k : K - 1;
            double v[FG - 2];
#pragma unroll
            for (int n = 0; n < FG - 2; n++) v[n] = D[(size_t)min(n, nb - 1) * K + kc];
            double w = 0, r = 0, yy = 0;
            if (k < K) {
                int j = packed_col(k);
                int i = (int)(k - (long long)j * (j + 1) / 2);
                double ci = 1.0 + lambda[i], cj = 1.0 + lambda[j];
                w = (i == j ? 0.5 : 1.0) / (ci * cj);
                r = (i == j) ? ci : 0.0;
                yy = y[i] * y[j];
            }
#pragma unroll
            for (int n = 0; n < FG - 2; n++) {
                const double vn = (n < nb && k < K) ? v[n] : 0.0;
                const T vt = (T)vn;
                Dt[n][t] = vt;
                X[n][t] = (T)(w * (double)vt);
                r -= sp[n] * vn;
            }
            X[nb][t] = (T)(w * r);
            X[nb + 1][t] = (T)(w * yy);
        }
        __syncthreads();
        // each wave takes FK_TILE / 4 entries = FK_TILE / 16 k-steps of 4
#pragma unroll
        for (int s = 0; s < FK_TILE / 16; s++) {
            const int kk = wave * (FK_TILE / 4) + s * 4 + (lane >> 4);
            T a[NF], b[NF];
            for (int i = 0; i < NF; i++) a[i] = X[i * 16 + (lane & 15)][kk];
            for (int j = 0; j < NF; j++) b[j] = Dt[j * 16 + (lane & 15)][kk];
            for (int i = 0; i < NF; i++)
                for (int j = 0; j < NF; j++) acc[i][j] = FMfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        if (fp32)
            for (int i = 0; i < NF; i++)
                for (int j = 0; j < NF; j++) {
                    for (int r = 0; r < 4; r++) acc64[i][j][r] += (double)acc[i][j][r];
                    acc[i][j] = acc_t{0, 0, 0, 0};
                }
    }
    // every wave holds the full FG x FG product of its quarter of the tile's
    // entries: the four waves' sums go through LDS (waves in order), then the
    // workgroup's partial is written with plain stores
    __shared__ double red[FG * FG];
    double* out = Gpart + (size_t)blockIdx.x * FG * FG;
    for (int w = 0; w < 4; w++) {
        if (wave == w)
            for (int i = 0; i < NF; i++)
                for (int j = 0; j < NF; j++)
                    for (int r = 0; r < 4; r++) {
                        const int row = i * 16 + acc_row<T>(lane, r), col = j * 16 + (lane & 15);
                        const double v = fp32 ? acc64[i][j][r] : (double)acc[i][j][r];
                        red[row * FG + col] = w == 0 ? v : red[row * FG + col] + v;
                    }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < FG * FG; e += blockDim.x) out[e] = red[e];
}

// Scratch of the allocation-free Gram entry point: FISHER_MAX_PARTS partial
// Gram matrices of FG_MAX x FG_MAX doubles, then the summed FG_MAX x FG_MAX
// matrix -- 16.03 MiB per device, held until the process ends.  Host code
// never uses this name as an address: it uses the device address that
// hipGetSymbolAddress returns for the current device (fisher_scratch_setup).
constexpr size_t FISHER_SCRATCH_DOUBLES = (size_t)(FISHER_MAX_PARTS + 1) * FG_MAX * FG_MAX;
static __device__ __align__(256) double scratch[FISHER_SCRATCH_DOUBLES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[SCRATCH_MAX_DEVICES];
static bool g_scratch_done[SCRATCH_MAX_DEVICES];

// Once per device, with that device current: the device address of `scratch`.
int fisher_scratch_setup(opsec_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= SCRATCH_MAX_DEVICES)
        return set_error(ctx, OPSEC_EINVAL, "device %d beyond %d", ctx->device, SCRATCH_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[ctx->device]) return OPSEC_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(ctx, OPSEC_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = (double*)p;
    g_scratch_done[ctx->device] = true;
    return OPSEC_OK;
}

static double* fisher_scratch_ptr(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_done[device] ? g_scratch_ptr[device] : nullptr;
}

// G[e] = sum over the partials in workgroup order (FP64, fixed order).
__global__ void fisher_sum_parts_kernel(const double* __restrict__ Gpart, int nparts, int FG, double* __restrict__ G) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= FG * FG) return;
    double acc = 0.0;
    for (int s = 0; s < nparts; s++) acc += Gpart[(size_t)s * FG * FG + e];
    G[e] = acc;
}

// The Gram outputs on the device: F (nb x nb row-major, symmetrised as
// 0.5 (G_mn + G_nm), the same operation the host post-processing does),
// f = G row nb, q = G row nb + 1.
__global__ void fisher_outputs_kernel(const double* __restrict__ G, int FG, int nb, double* __restrict__ F,
                                      double* __restrict__ f, double* __restrict__ q) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nb * nb) {
        const int m = e / nb, n = e - m * nb;
        F[e] = 0.5 * (G[m * FG + n] + G[n * FG + m]);
    }
    if (e < nb) {
        f[e] = G[nb * FG + e];
        q[e] = G[(nb + 1) * FG + e];
    }
}


// ---- host FP64 dense helpers for nb x nb ----------------------------------

// Inverse by LU with partial pivoting (the reference uses LAPACK GETRF/GETRI,
// src/MyMatrix.hpp:531-569).  Returns false if singular.
static bool lu_inverse(int n, const std::vector<double>& A, std::vector<double>& Ainv) {
    std::vector<double> LU(A);
    std::vector<int> piv(n);
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(LU[r * n + c]) > std::fabs(LU[p * n + c])) p = r;
        piv[c] = p;
        if (LU[p * n + c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j < n; j++) std::swap(LU[c * n + j], LU[p * n + j]);
        for (int r = c + 1; r < n; r++) {
            LU[r * n + c] /= LU[c * n + c];
            double l = LU[r * n + c];
            for (int j = c + 1; j < n; j++) LU[r * n + j] -= l * LU[c * n + j];
        }
    }
    Ainv.assign((size_t)n * n, 0.0);
    for (int col = 0; col < n; col++) {
        std::vector<double> b(n, 0.0);
        b[col] = 1.0;
        for (int c = 0; c < n; c++) std::swap(b[c], b[piv[c]]);
        for (int r = 0; r < n; r++)
            for (int j = 0; j < r; j++) b[r] -= LU[r * n + j] * b[j];
        for (int r = n - 1; r >= 0; r--) {
            for (int j = r + 1; j < n; j++) b[r] -= LU[r * n + j] * b[j];
            b[r] /= LU[r * n + r];
        }
        for (int r = 0; r < n; r++) Ainv[r * n + col] = b[r];
    }
    return true;
}

// Principal square root of a symmetric matrix through its eigenbasis
// (src/MyMatrix.hpp:478-527), eigenvectors by cyclic Jacobi.
static void sym_sqrtm(int n, const std::vector<double>& A0, std::vector<double>& out) {
    std::vector<double> A(A0), V((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                double t = (th >= 0 ? 1 : -1) / (std::fabs(th) + std::sqrt(th * th + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
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
    out.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += V[i * n + k] * std::sqrt(A[k * n + k]) * V[j * n + k];
            out[i * n + j] = s;
        }
}

static int post_process(opsec_ctx* ctx, int nb, const std::vector<double>& G, int FG, int mixing,
                        opsec_estimate* out) {
    std::vector<double> F((size_t)nb * nb), f(nb), q(nb);
    for (int m = 0; m < nb; m++)
        for (int n = 0; n < nb; n++) F[m * nb + n] = 0.5 * (G[m * FG + n] + G[n * FG + m]);
    for (int n = 0; n < nb; n++) {
        f[n] = G[nb * FG + n];
        q[n] = G[(nb + 1) * FG + n];
    }
    std::vector<double> M((size_t)nb * nb, 0.0), W((size_t)nb * nb);
    if (mixing == OPSEC_MIX_INVERSE || mixing == OPSEC_MIX_INVERSE_SQRT) {
        if (!lu_inverse(nb, F, M)) return set_error(ctx, OPSEC_ECONV, "Fisher matrix is singular");
        if (mixing == OPSEC_MIX_INVERSE_SQRT) {
            std::vector<double> Ms;
            sym_sqrtm(nb, M, Ms);
            M.swap(Ms);
        }
    } else {
        for (int i = 0; i < nb; i++) M[i * nb + i] = 1.0;
    }
    auto mul = [&](const std::vector<double>& A, const std::vector<double>& B, std::vector<double>& Cm) {
        Cm.assign((size_t)nb * nb, 0.0);
        for (int i = 0; i < nb; i++)
            for (int k = 0; k < nb; k++)
                for (int j = 0; j < nb; j++) Cm[i * nb + j] += A[i * nb + k] * B[k * nb + j];
    };
    mul(M, F, W);
    for (int m = 0; m < nb; m++) {
        double rs = 0;
        for (int n = 0; n < nb; n++) rs += W[m * nb + n];
        for (int n = 0; n < nb; n++) M[m * nb + n] /= rs;
    }
    mul(M, F, W);
    std::vector<double> phat(nb, 0.0), cov((size_t)nb * nb, 0.0);
    for (int i = 0; i < nb; i++)
        for (int k = 0; k < nb; k++) phat[i] += M[i * nb + k] * (q[k] - f[k]);
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++) {
            double s = 0;
            for (int k = 0; k < nb; k++) s += W[i * nb + k] * M[j * nb + k];
            cov[i * nb + j] = s;
        }
    auto cp = [](double* dst, const std::vector<double>& v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * 8);
    };
    cp(out->phat, phat);
    cp(out->cov, cov);
    cp(out->F, F);
    cp(out->W, W);
    cp(out->M, M);
    cp(out->q, q);
    cp(out->f, f);
    return OPSEC_OK;
}

// Padded Gram size and packed entries per tile for nb parameters.
static void gram_shape(int nb, int prec, int* FG, int* KT) {
    const bool big = nb + 2 > FShape<double, 3>::FG;
    const bool f64 = prec == OPSEC_PREC_FP64;
    *FG = big ? FShape<double, 4>::FG : FShape<double, 3>::FG;
    *KT = f64 ? (big ? FShape<double, 4>::KT : FShape<double, 3>::KT)
              : (big ? FShape<float, 4>::KT : FShape<float, 3>::KT);
}

// Launch the Gram product of the nb packed matrices dD (M modes) into the
// padded FG x FG matrix d_G: `nparts` workgroup partials in d_parts (nparts x
// FG x FG doubles, every entry written), then their fixed-order sum.  Device
// inputs; nothing is allocated.
static void launch_gram(opsec_ctx* ctx, const double* dD, int M, int nb, const double* d_lambda, const double* d_y,
                        const double* d_prior, int prec, int nparts, double* d_parts, double* d_G) {
    const long long K = (long long)M * (M + 1) / 2;
    int FG, KT;
    gram_shape(nb, prec, &FG, &KT);
    const bool big = FG > FShape<double, 3>::FG;
    if (prec == OPSEC_PREC_FP64 && !big)
        hipLaunchKernelGGL((fisher_gram_kernel<double, 3>), dim3(nparts), dim3(256), 0, ctx->stream, dD, K, nb,
                           d_lambda, d_y, d_prior, d_parts);
    else if (prec == OPSEC_PREC_FP64)
        hipLaunchKernelGGL((fisher_gram_kernel<double, 4>), dim3(nparts), dim3(256), 0, ctx->stream, dD, K, nb,
                           d_lambda, d_y, d_prior, d_parts);
    else if (!big)
        hipLaunchKernelGGL((fisher_gram_kernel<float, 3>), dim3(nparts), dim3(256), 0, ctx->stream, dD, K, nb,
                           d_lambda, d_y, d_prior, d_parts);
    else
        hipLaunchKernelGGL((fisher_gram_kernel<float, 4>), dim3(nparts), dim3(256), 0, ctx->stream, dD, K, nb,
                           d_lambda, d_y, d_prior, d_parts);
    hipLaunchKernelGGL(fisher_sum_parts_kernel, dim3((FG * FG + 255) / 256), dim3(256), 0, ctx->stream, d_parts,
                       nparts, FG, d_G);
}

static int check_gram_args(opsec_ctx* ctx, int M, int nb, int prec) {
    OPSEC_CHECK_ARG(ctx, nb >= 1 && nb + 2 <= FG_MAX, "Nparams must be 1..%d", FG_MAX - 2);
    OPSEC_CHECK_ARG(ctx, M >= 1, "Nmodes must be positive");
    OPSEC_CHECK_ARG(ctx, prec == OPSEC_PREC_FP64 || prec == OPSEC_PREC_FP32 || prec == OPSEC_PREC_BF16X3,
                    "unknown precision %d", prec);
    return OPSEC_OK;
}

static int run_fisher(opsec_ctx* ctx, const double* dD, int M, int nb, const double* lambda, const double* y,
                      const double* prior, int mixing, int prec, opsec_estimate* out) {
    int rc = check_gram_args(ctx, M, nb, prec);
    if (rc) return rc;
    OPSEC_CHECK_ARG(ctx, mixing >= 0 && mixing <= 2, "unknown mixing mode %d", mixing);
    OPSEC_CHECK_ARG(ctx, lambda && y && prior && out, "NULL input");
    const long long K = (long long)M * (M + 1) / 2;
    int FG, KT;
    gram_shape(nb, prec, &FG, &KT);
    const int nparts = fisher_parts(ctx, (K + KT - 1) / KT);
    double *d_aux = nullptr, *d_G = nullptr, *d_parts = nullptr;
    DevScratch tmp;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&d_aux, ((size_t)2 * M + nb) * 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&d_G, (size_t)FG * FG * 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&d_parts, (size_t)nparts * FG * FG * 8));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_aux, lambda, M * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_aux + M, y, M * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_aux + 2 * M, prior, nb * 8, hipMemcpyHostToDevice, ctx->stream));
    launch_gram(ctx, dD, M, nb, d_aux, d_aux + M, d_aux + 2 * M, prec, nparts, d_parts, d_G);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    std::vector<double> G((size_t)FG * FG);
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(G.data(), d_G, (size_t)FG * FG * 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return post_process(ctx, nb, G, FG, mixing, out);
}

void* proj_device_results(opsec_ctx* ctx, int* nb, int* M);

}  // namespace opsec

using namespace opsec;

extern "C" {

int opsec_estimate_run(opsec_ctx* ctx, int nmodes, int nparams, const double* lambda, const double* D, const double* y,
                       const double* prior, int mixing, int prec, opsec_estimate* out) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, D, "NULL covariance derivatives");
    OPSEC_CHECK_ARG(ctx, nmodes >= 1 && nparams >= 1, "bad sizes");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    size_t bytes = (size_t)nparams * packed_size(nmodes) * 8;
    double* dD = nullptr;
    DevScratch tmp;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dD, bytes));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dD, D, bytes, hipMemcpyHostToDevice, ctx->stream));
    int rc = run_fisher(ctx, dD, nmodes, nparams, lambda, y, prior, mixing, prec, out);
    hipStreamSynchronize(ctx->stream);
    return rc;
}

int opsec_fisher_gram_noalloc(opsec_ctx* ctx, int nmodes, int nparams, const double* d_lambda, const double* d_D,
                              const double* d_y, const double* d_prior, int prec, double* d_F, double* d_f,
                              double* d_q) {
    if (!ctx) return OPSEC_EINVAL;
    int rc = check_gram_args(ctx, nmodes, nparams, prec);
    if (rc) return rc;
    OPSEC_CHECK_ARG(ctx, d_lambda && d_D && d_y && d_prior && d_F && d_f && d_q, "NULL device buffer");
    double* base = fisher_scratch_ptr(ctx->device);
    if (!base) return set_error(ctx, OPSEC_ESTATE, "Gram scratch of device %d not set up", ctx->device);
    const long long K = (long long)nmodes * (nmodes + 1) / 2;
    int FG, KT;
    gram_shape(nparams, prec, &FG, &KT);
    const int nparts = fisher_parts(ctx, (K + KT - 1) / KT);
    OPSEC_CHECK_ARG(ctx, nparts <= FISHER_MAX_PARTS && FG <= FG_MAX, "Gram scratch too small");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    // every partial the sum reads is written by its workgroup, and G is
    // written whole by the sum: nothing needs clearing
    double* d_G = base + (size_t)FISHER_MAX_PARTS * FG_MAX * FG_MAX;
    launch_gram(ctx, d_D, nmodes, nparams, d_lambda, d_y, d_prior, prec, nparts, base, d_G);
    hipLaunchKernelGGL(fisher_outputs_kernel, dim3((nparams * nparams + 255) / 256), dim3(256), 0, ctx->stream, d_G,
                       FG, nparams, d_F, d_f, d_q);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    return OPSEC_OK;
}

int opsec_estimate_device(opsec_ctx* ctx, const double* lambda, const double* y, const double* prior, int mixing,
                          int prec, opsec_estimate* out) {
    if (!ctx) return OPSEC_EINVAL;
    int nb = 0, M = 0;
    void* dP = proj_device_results(ctx, &nb, &M);
    if (!dP) return set_error(ctx, OPSEC_ESTATE, "no projection results on the device");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return run_fisher(ctx, (const double*)dP, M, nb, lambda, y, prior, mixing, prec, out);
}

}  // extern "C"
