// KL eigenmodes by block Lanczos (replaces ParpackSolver,
// src/ParpackSolver.cpp:68-227, and the klt solve, src/klt.cpp:160-248).
//
// The reference runs PARPACK's implicitly restarted Lanczos (which = LA,
// tol = 1e-5, ncv = 2 nev) over a pdgemm matvec.  On one MI355X the whole
// Krylov basis stays in HBM, so this solver never restarts:
//   * block Lanczos, block size b, operator = the device-resident signal
//     matrix (dense FP64, streamed from HBM by the MFMA GEMM);
//   * full re-orthogonalisation against every previous block, twice (CGS2),
//     as FP64 MFMA GEMMs;
//   * block orthonormalisation by Cholesky QR, or, for ill-conditioned
//     blocks, by SVQB (eigen-decomposition of the b x b Gram matrix), which
//     replaces numerically dependent directions by fresh random ones -- the
//     periodic box's signal matrix has cubic-symmetry multiplets that
//     single-vector Lanczos cannot resolve;
//   * Rayleigh-Ritz on the block-tridiagonal T: blocked GPU Householder
//     tridiagonalisation (panels of deferred two-sided updates), host
//     bisection for the wanted eigenvalues and inverse iteration (with
//     re-orthogonalisation inside clusters) for their eigenvectors on host
//     threads, compact-WY GPU back-transformation;
//   * convergence: |B_j z_last| <= tol |theta_max| for each wanted Ritz pair.
// Output: eigenvalues descending, eigenvectors mode-major (modefile layout).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

#include "common.h"
#include "gemm.h"

namespace opsec {

// ---- FP64 GEMM variants used by the solver ------------------------------------

struct LA_RowMajor : SrcDenseRowMajor {
    __device__ __forceinline__ void p_shift(int, size_t) {}
};
struct LA_ColMajor : SrcDenseColMajor {
    __device__ __forceinline__ void p_shift(int, size_t) {}
};
struct LA_KCol : SrcKColMajor {   // operand X with X[i][k] = p[i*ld + k]
    __device__ __forceinline__ void p_shift(int, size_t) {}
};

template <int MODE>   // 0 store, 1 subtract
struct LPut {
    double* p;
    size_t ld;
    __device__ __forceinline__ void put(int, int m, int n, int M, int N, double v) const {
        if (m < M && n < N) {
            if (MODE == 0) p[(size_t)n * ld + m] = v;
            else p[(size_t)n * ld + m] -= v;
        }
    }
};

// With split-K (gridDim.z > 1) part z of the K range is written to
// parts[(z N + n) M + m] and summed in order by lsum_parts_kernel.
template <typename SrcA, typename SrcB, typename Out>
__global__ __launch_bounds__(GT_THREADS) void lgemm_kernel(SrcA A, SrcB B, Out O, int M, int N, int K, double* parts) {
    __shared__ double As[16][GT_M + 4];
    __shared__ double Bs[16][GT_N + 4];
    const int m0 = blockIdx.x * GT_M, n0 = blockIdx.y * GT_N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    // split-K: blockIdx.z owns k range [kz0, kz1)
    const int kchunk = ((K + gridDim.z - 1) / gridDim.z + 15) / 16 * 16;
    const int kz0 = blockIdx.z * kchunk, kz1 = min(K, kz0 + kchunk);
    d4_t acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = d4_t{0, 0, 0, 0};
    const int t = threadIdx.x;
    for (int k0 = kz0; k0 < kz1; k0 += 16) {
        __syncthreads();
        if (SrcA::k_contig) {
            int i = t >> 1, kb = (t & 1) * 8;
            for (int e = 0; e < 8; e++) {
                int m = m0 + i, k = k0 + kb + e;
                As[kb + e][i] = (m < M && k < kz1) ? A.get(m, k) : 0.0;
            }
        } else {
            int k = t >> 4, ib = (t & 15) * 8;
            for (int e = 0; e < 8; e++) {
                int m = m0 + ib + e, kk = k0 + k;
                As[k][ib + e] = (m < M && kk < kz1) ? A.get(m, kk) : 0.0;
            }
        }
        {
            int i = t >> 1, kb = (t & 1) * 8;
            for (int e = 0; e < 8; e++) {
                int n = n0 + i, k = k0 + kb + e;
                Bs[kb + e][i] = (n < N && k < kz1) ? B.get(n, k) : 0.0;
            }
        }
        __syncthreads();
        for (int kk = 0; kk < 4; kk++) {
            double af[4], bf[4];
            const int kr = kk * 4 + (lane >> 4);
            for (int i = 0; i < 4; i++) af[i] = As[kr][wm + i * 16 + (lane & 15)];
            for (int j = 0; j < 4; j++) bf[j] = Bs[kr][wn + j * 16 + (lane & 15)];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const int n = n0 + wn + j * 16 + (lane & 15);
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + acc_row<double>(lane, r);
                if (gridDim.z == 1) O.put(0, m, n, M, N, acc[i][j][r]);
                else if (m < M && n < N) parts[((size_t)blockIdx.z * N + n) * M + m] = acc[i][j][r];
            }
        }
}

// H[n ld + m] = sum_z parts[(z N + n) M + m], z in order.
__global__ void lsum_parts_kernel(const double* __restrict__ parts, int nparts, int M, int N, double* H, size_t ld) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    double acc = 0.0;
    for (int z = 0; z < nparts; z++) acc += parts[(size_t)z * M * N + i];
    H[(i / M) * ld + i % M] = acc;
}

__global__ void fill_kernel(double* p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Wall-clock phases of the last solve (ms): block matvecs, orthogonalisation,
// Rayleigh-Ritz tridiagonal reduction on the GPU, host tridiagonal eigenpairs,
// back-transformation, Ritz vectors; then the Krylov dimension, the number of
// Rayleigh-Ritz steps and of block steps.
enum { KT_MATVEC, KT_ORTH, KT_TRIRED, KT_TRIEIG, KT_BACK, KT_RITZVEC, KT_KDIM, KT_NRR, KT_NSTEPS, KT_N };
static double g_klt_times[KT_N];

struct PhaseTimer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double lap() {
        auto t1 = std::chrono::steady_clock::now();
        double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        t0 = t1;
        return ms;
    }
};

struct Dev {
    opsec_ctx* ctx;
    hipStream_t s;
    int err = 0;
    // split-K partials of gemm_tn: at most 2 x CUs workgroups of 128 x 128
    double* parts = nullptr;
    DevScratch mem;
    Dev(opsec_ctx* c, hipStream_t st) : ctx(c), s(st) {}
    hipError_t init() {
        return mem.alloc(&parts, (size_t)2 * std::max(1, ctx->num_cus) * GT_M * GT_N * 8);
    }
    void fill(double* p, size_t n, double v) {
        if (n == 0) return;
        hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    }
    // H (p x q, ld p) = X^T Y, X: N x p col-major (ld ldx), Y: N x q col-major
    void gemm_tn(const double* X, size_t ldx, const double* Y, size_t ldy, int N, int p, int q, double* H) {
        LA_KCol A;
        A.p = X;
        A.ld = ldx;
        LA_KCol B;
        B.p = Y;
        B.ld = ldy;
        int gx = (p + GT_M - 1) / GT_M, gy = (q + GT_N - 1) / GT_N;
        // split-K fills the CUs when the output has few tiles; the partials
        // (splits x p x q <= 2 CUs x 128 x 128 doubles) are summed in order
        int splits = std::max(1, std::min(64, (2 * ctx->num_cus) / std::max(1, gx * gy)));
        splits = std::min(splits, std::max(1, N / 512));
        hipLaunchKernelGGL((lgemm_kernel<LA_KCol, LA_KCol, LPut<0>>), dim3(gx, gy, splits), dim3(GT_THREADS), 0, s, A,
                           B, LPut<0>{H, (size_t)p}, p, q, N, parts);
        if (splits > 1)
            hipLaunchKernelGGL(lsum_parts_kernel, dim3((unsigned)(((size_t)p * q + 255) / 256)), dim3(256), 0, s, parts,
                               splits, p, q, H, (size_t)p);
    }
    // Y (N x q) op= X (N x p) H (p x q); op: 0 store, 1 subtract
    void gemm_nn(int mode, double* Y, size_t ldy, const double* X, size_t ldx, const double* H, size_t ldh, int N,
                 int p, int q) {
        LA_ColMajor A;
        A.p = X;
        A.ld = ldx;
        LA_KCol B;
        B.p = H;
        B.ld = ldh;
        dim3 g((N + GT_M - 1) / GT_M, (q + GT_N - 1) / GT_N, 1);
        if (mode == 0)
            hipLaunchKernelGGL((lgemm_kernel<LA_ColMajor, LA_KCol, LPut<0>>), g, dim3(GT_THREADS), 0, s, A, B,
                               LPut<0>{Y, ldy}, N, q, p, nullptr);
        else
            hipLaunchKernelGGL((lgemm_kernel<LA_ColMajor, LA_KCol, LPut<1>>), g, dim3(GT_THREADS), 0, s, A, B,
                               LPut<1>{Y, ldy}, N, q, p, nullptr);
    }
    // Y (N x q) op= X (N x p) Xt^T, Xt: q x p column-major (ld ldt); op: 1 subtract
    void gemm_nt_sub(double* Y, size_t ldy, const double* X, size_t ldx, const double* Xt, size_t ldt, int N, int p,
                     int q) {
        LA_ColMajor A;
        A.p = X;
        A.ld = ldx;
        LA_ColMajor B;   // B[k][n] = Xt(n, k) = Xt[k ldt + n]
        B.p = Xt;
        B.ld = ldt;
        dim3 g((N + GT_M - 1) / GT_M, (q + GT_N - 1) / GT_N, 1);
        hipLaunchKernelGGL((lgemm_kernel<LA_ColMajor, LA_ColMajor, LPut<1>>), g, dim3(GT_THREADS), 0, s, A, B,
                           LPut<1>{Y, ldy}, N, q, p, nullptr);
    }
    // Y (N x q) = S X, S dense row-major N x N (symmetric)
    void apply(const double* S, int N, const double* X, int q, double* Y) {
        LA_RowMajor A;
        A.p = S;
        A.ld = N;
        LA_KCol B;
        B.p = X;
        B.ld = N;
        dim3 g((N + GT_M - 1) / GT_M, (q + GT_N - 1) / GT_N, 1);
        hipLaunchKernelGGL((lgemm_kernel<LA_RowMajor, LA_KCol, LPut<0>>), g, dim3(GT_THREADS), 0, s, A, B,
                           LPut<0>{Y, (size_t)N}, N, q, N, nullptr);
    }
    int sync() {
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return set_error(ctx, OPSEC_EHIP, "eigen-solver: %s", hipGetErrorString(e));
        return OPSEC_OK;
    }
};

// ---- Householder tridiagonalisation of a dense symmetric K x K (col-major) --

// One workgroup: Householder vector of column k (rows k+1..K-1), in place.
__global__ __launch_bounds__(1024) void hh_vector_kernel(double* A, int K, int k, double* tau, double* e) {
    __shared__ double red[1024];
    __shared__ double sh_scale, sh_tau, sh_beta;
    const int m = K - k - 1;
    double* x = A + (size_t)k * K + (k + 1);
    double s = 0;
    for (int i = 1 + threadIdx.x; i < m; i += blockDim.x) s += x[i] * x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sigma = red[0], alpha = x[0];
        if (sigma == 0.0) {
            sh_tau = 0.0;
            sh_beta = alpha;
            sh_scale = 0.0;
        } else {
            double nrm = sqrt(alpha * alpha + sigma);
            double beta = alpha <= 0 ? nrm : -nrm;
            sh_tau = (beta - alpha) / beta;
            sh_scale = 1.0 / (alpha - beta);
            sh_beta = beta;
        }
        tau[k] = sh_tau;
        e[k] = sh_beta;
    }
    __syncthreads();
    double sc = sh_scale;
    for (int i = 1 + threadIdx.x; i < m; i += blockDim.x) x[i] *= sc;
    if (threadIdx.x == 0) x[0] = 1.0;
}

// part[c][i] = sum_{j in chunk c} A22[i][j] v[j]   (A22 = A[k+1:, k+1:], v = A[k+1:, k]);
// chunk c = blockIdx.y covers columns 256 c .. 256 c + 255
__global__ __launch_bounds__(256) void hh_symv_kernel(const double* A, int K, int k, double* part) {
    __shared__ double vs[256];
    const int m = K - k - 1;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j0 = blockIdx.y * 256;
    const double* v = A + (size_t)k * K + (k + 1);
    vs[threadIdx.x] = (j0 + threadIdx.x < m) ? v[j0 + threadIdx.x] : 0.0;
    __syncthreads();
    if (i >= m) return;
    double acc = 0;
    const int jn = min(256, m - j0);
    const double* col = A + (size_t)(k + 1 + j0) * K + (k + 1) + i;
    for (int jj = 0; jj < jn; jj++) acc += col[(size_t)jj * K] * vs[jj];
    part[(size_t)blockIdx.y * m + i] = acc;
}

// p[i] = sum_c part[c][i], chunks in order
__global__ __launch_bounds__(256) void hh_symv_sum_kernel(const double* __restrict__ part, int nc, int m,
                                                          double* __restrict__ p) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double acc = 0.0;
    for (int c = 0; c < nc; c++) acc += part[(size_t)c * m + i];
    p[i] = acc;
}

// w = tau p - (tau/2)(tau p . v) v  (one workgroup)
__global__ __launch_bounds__(1024) void hh_w_kernel(const double* A, int K, int k, const double* tau, double* p) {
    __shared__ double red[1024];
    const int m = K - k - 1;
    const double* v = A + (size_t)k * K + (k + 1);
    const double t = tau[k];
    double s = 0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) s += t * p[i] * v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double kap = 0.5 * t * red[0];
    for (int i = threadIdx.x; i < m; i += blockDim.x) p[i] = t * p[i] - kap * v[i];
}


// ---- blocked reduction (panels of columns, deferred two-sided updates) -------
// Within a panel the rank-2 updates of its reflectors are kept as the columns
// of V (reflectors) and W (update vectors), K x nb column-major, zero above
// their first row; the trailing matrix is updated once per panel.

// Column k (rows k .. K-1) gets the panel's deferred updates from its first i
// reflectors: A(r, k) -= sum_j V(r, j) W(k, j) + W(r, j) V(k, j).
__global__ __launch_bounds__(256) void hh_colupd_kernel(double* A, int K, int k, int i, const double* __restrict__ V,
                                                        const double* __restrict__ W) {
    const int r = k + blockIdx.x * 256 + threadIdx.x;
    if (r >= K) return;
    double acc = 0.0;
    for (int j = 0; j < i; j++)
        acc += V[(size_t)j * K + r] * W[(size_t)j * K + k] + W[(size_t)j * K + r] * V[(size_t)j * K + k];
    A[(size_t)k * K + r] -= acc;
}

// ab[j] = W(:, j) . v and ab[i + j] = V(:, j) . v over rows k+1 .. K-1 (v in
// column k of A), j < i; one workgroup per dot, fixed-order reduction.
__global__ __launch_bounds__(256) void hh_dots_kernel(const double* A, int K, int k, int i, const double* __restrict__ V,
                                                      const double* __restrict__ W, double* __restrict__ ab) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    const double* x = (j < i ? W + (size_t)j * K : V + (size_t)(j - i) * K);
    const double* v = A + (size_t)k * K;
    double acc = 0.0;
    for (int r = k + 1 + threadIdx.x; r < K; r += 256) acc += x[r] * v[r];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) ab[j] = red[0];
}

// p -= V' (W'^T v) + W' (V'^T v): the deferred updates' share of A22 v, p
// indexed from row k+1.
__global__ __launch_bounds__(256) void hh_corr_kernel(double* p, int K, int k, int i, const double* __restrict__ V,
                                                      const double* __restrict__ W, const double* __restrict__ ab) {
    const int r = k + 1 + blockIdx.x * 256 + threadIdx.x;
    if (r >= K) return;
    double acc = 0.0;
    for (int j = 0; j < i; j++) acc += V[(size_t)j * K + r] * ab[j] + W[(size_t)j * K + r] * ab[i + j];
    p[r - k - 1] -= acc;
}

// Column i of the panel: V(:, i) = v_k (rows k+1 .. K-1, from A), W(:, i) = w
// (from p); zero above.
__global__ __launch_bounds__(256) void hh_store_kernel(const double* A, int K, int k, int i, const double* __restrict__ w,
                                                       double* __restrict__ V, double* __restrict__ W) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= K) return;
    const bool in = r >= k + 1;
    V[(size_t)i * K + r] = in ? A[(size_t)k * K + r] : 0.0;
    W[(size_t)i * K + r] = in ? w[r - k - 1] : 0.0;
}

// Panel of reflectors k0 .. k0 + nb - 1 as a dense K x nb column-major
// matrix: column j is v_(k0+j), nonzero in rows k0 + j + 1 .. K - 1 (its first
// entry, 1, is stored in A by hh_vector_kernel), zero above.
__global__ void hh_panel_kernel(const double* __restrict__ A, int K, int k0, int nb, double* __restrict__ Vp) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= K || j >= nb) return;
    const int k = k0 + j;
    Vp[(size_t)j * K + i] = i >= k + 1 ? A[(size_t)k * K + i] : 0.0;
}

__global__ void diag_kernel(const double* A, int K, double* d, double* elast) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K) d[i] = A[(size_t)i * K + i];
    if (i == 0 && K >= 2) *elast = A[(size_t)(K - 2) * K + (K - 1)];
}

// ---- host: symmetric tridiagonal eigenpairs --------------------------------------

// f(i) for i in [0, n) on up to 16 host threads, items taken in chunks from a
// shared counter (the results must not depend on which thread runs an item).
template <typename F>
static void parallel_for(int n, int chunk, F f) {
    const int nt = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (n <= chunk || nt == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next{0};
    auto work = [&]() {
        for (;;) {
            const int i0 = next.fetch_add(chunk);
            if (i0 >= n) break;
            for (int i = i0; i < std::min(n, i0 + chunk); i++) f(i);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < std::min(nt, (n + chunk - 1) / chunk); t++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// Number of eigenvalues of the symmetric tridiagonal (d, e2 = e^2) below x,
// from the signs of the pivots of the LDL^T factorisation of T - x I (Sturm
// count); a zero pivot is replaced by -pivmin.
static int sturm_below(int n, const double* d, const double* e2, double x, double pivmin) {
    int cnt = 0;
    double q = d[0] - x;
    if (std::fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0;
    for (int i = 1; i < n; i++) {
        q = d[i] - x - e2[i - 1] / q;
        if (std::fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0;
    }
    return cnt;
}

// Eigenvalues of indices j[0..NB) (ascending, 0-based) by bisection on
// [lo, hi] (Gershgorin bounds) to full precision; the NB Sturm sequences of a
// pass run interleaved (independent dependency chains).
template <int NB>
static void bisect_batch(int n, const double* d, const double* e2, const int* j, double lo0, double hi0,
                         double pivmin, double* out) {
    const double eps = 2.2204460492503131e-16;
    double lo[NB], hi[NB];
    for (int b = 0; b < NB; b++) lo[b] = lo0, hi[b] = hi0;
    for (int it = 0; it < 200; it++) {
        double x[NB], q[NB];
        int cnt[NB];
        bool any = false;
        for (int b = 0; b < NB; b++) {
            x[b] = 0.5 * (lo[b] + hi[b]);
            const bool done = hi[b] - lo[b] <= 2 * eps * std::max(std::fabs(lo[b]), std::fabs(hi[b])) + pivmin ||
                              x[b] <= lo[b] || x[b] >= hi[b];
            any |= !done;
            q[b] = d[0] - x[b];
            if (std::fabs(q[b]) < pivmin) q[b] = -pivmin;
            cnt[b] = q[b] < 0;
        }
        if (!any) break;
        for (int i = 1; i < n; i++) {
            const double di = d[i], ei = e2[i - 1];
#pragma unroll
            for (int b = 0; b < NB; b++) {
                double v = di - x[b] - ei / q[b];
                if (std::fabs(v) < pivmin) v = -pivmin;
                q[b] = v;
                cnt[b] += v < 0;
            }
        }
        for (int b = 0; b < NB; b++) {
            if (hi[b] - lo[b] <= 2 * eps * std::max(std::fabs(lo[b]), std::fabs(hi[b])) + pivmin ||
                x[b] <= lo[b] || x[b] >= hi[b])
                continue;
            if (cnt[b] > j[b]) hi[b] = x[b];
            else lo[b] = x[b];
        }
    }
    for (int b = 0; b < NB; b++) out[b] = 0.5 * (lo[b] + hi[b]);
}

// The eigenvalue of index j (ascending, 0-based) of the unreduced tridiagonal
// by bisection on [lo, hi] (Gershgorin bounds) to full precision.
static double bisect_eigenvalue(int n, const double* d, const double* e2, int j, double lo, double hi,
                                double pivmin) {
    const double eps = 2.2204460492503131e-16;
    for (int it = 0; it < 200; it++) {
        const double mid = 0.5 * (lo + hi);
        if (hi - lo <= 2 * eps * std::max(std::fabs(lo), std::fabs(hi)) + pivmin || mid <= lo || mid >= hi) break;
        if (sturm_below(n, d, e2, mid, pivmin) > j) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// LU factors (partial pivoting, as LAPACK's dgttrf) of T - lam I for the
// symmetric tridiagonal (d, e); zero pivots are replaced by `tiny`.
struct TriLU {
    std::vector<double> dl, dd, du, du2;
    std::vector<char> swp;
    void factor(int n, const double* d, const double* e, double lam, double tiny) {
        dl.assign(e, e + std::max(0, n - 1));
        du.assign(e, e + std::max(0, n - 1));
        dd.resize(n);
        du2.assign(std::max(0, n - 2), 0.0);
        swp.assign(n, 0);
        for (int i = 0; i < n; i++) dd[i] = d[i] - lam;
        for (int i = 0; i < n - 1; i++) {
            if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
                if (dd[i] == 0.0) dd[i] = tiny;
                double fact = dl[i] / dd[i];
                dl[i] = fact;
                dd[i + 1] -= fact * du[i];
            } else {
                double fact = dd[i] / dl[i];
                dd[i] = dl[i];
                dl[i] = fact;
                double temp = dd[i + 1];
                dd[i + 1] = du[i] - fact * temp;
                if (i < n - 2) {
                    du2[i] = du[i + 1];
                    du[i + 1] = -fact * du2[i];
                }
                du[i] = temp;
                swp[i] = 1;
            }
        }
        if (dd[n - 1] == 0.0) dd[n - 1] = tiny;
    }
    // x <- (T - lam I)^-1 x
    void solve(int n, double* x) const {
        if (n == 1) {
            x[0] /= dd[0];
            return;
        }
        for (int i = 0; i < n - 1; i++) {
            if (swp[i]) std::swap(x[i], x[i + 1]);
            x[i + 1] -= dl[i] * x[i];
        }
        x[n - 1] /= dd[n - 1];
        x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dd[n - 2];
        for (int i = n - 3; i >= 0; i--) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dd[i];
    }
};

struct TriBlock {
    int off, n;
};

// Top-nev eigenpairs of the symmetric tridiagonal (d, e), vectors as columns of
// Z (n x nev, column-major), eigenvalues descending.
//   * unreduced blocks split at negligible off-diagonals;
//   * eigenvalues by bisection (Sturm counts), only the wanted ones: the
//     nev-th largest of the whole matrix first, then every eigenvalue of each
//     block above it, in parallel over eigenvalues;
//   * eigenvectors by inverse iteration (LAPACK dstein's scheme): clusters of
//     eigenvalues closer than 1e-7 |T| are orthogonalised against each other,
//     coincident shifts separated by 10 eps |T|; the clusters run in parallel,
//     each from its own seed, so the result does not depend on the threads.
static void tridiag_top(int n, const std::vector<double>& d, const std::vector<double>& e, int nev,
                        std::vector<double>& evals, std::vector<double>& Z) {
    const double eps = 2.2204460492503131e-16;
    double tnorm = 0, glo = 0, ghi = 0;
    for (int i = 0; i < n; i++) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0) + (i < n - 1 ? std::fabs(e[i]) : 0);
        tnorm = std::max(tnorm, std::fabs(d[i]) + r);
        glo = i ? std::min(glo, d[i] - r) : d[i] - r;
        ghi = i ? std::max(ghi, d[i] + r) : d[i] + r;
    }
    if (tnorm == 0) tnorm = 1;
    glo -= 2 * eps * tnorm + 1e-300;
    ghi += 2 * eps * tnorm + 1e-300;
    std::vector<double> e2(std::max(1, n - 1), 0.0);
    double emax2 = 0;
    for (int i = 0; i + 1 < n; i++) {
        e2[i] = e[i] * e[i];
        emax2 = std::max(emax2, e2[i]);
    }
    const double pivmin = std::max(1e-300, 1e-300 * emax2);
    // split into unreduced blocks
    std::vector<TriBlock> blocks;
    int start = 0;
    for (int i = 0; i < n - 1; i++)
        if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1])) + 1e-300) {
            blocks.push_back({start, i + 1 - start});
            start = i + 1;
        }
    blocks.push_back({start, n - start});
    nev = std::min(nev, n);
    // the nev-th largest eigenvalue of the whole matrix (split off-diagonals as zero)
    std::vector<double> dsplit(d.begin(), d.end()), e2split(e2);
    for (const TriBlock& B : blocks)
        if (B.off > 0) e2split[B.off - 1] = 0.0;
    const double tau = bisect_eigenvalue(n, dsplit.data(), e2split.data(), n - nev, glo, ghi, pivmin);
    // every eigenvalue of each block at or above tau (with a margin for ties)
    struct EV {
        double v;
        int blk, idx;
    };
    std::vector<EV> cand;
    // floating-point Sturm counts are not exactly monotone, so inside a tight
    // cluster the blocks' counts at tcut can leave fewer than nev candidates:
    // widen the margin until there are enough
    for (double margin = 4 * eps * tnorm; cand.size() < (size_t)nev && margin < 2 * tnorm; margin *= 16) {
        cand.clear();
        const double tcut = tau - margin;
        for (size_t bi = 0; bi < blocks.size(); bi++) {
            const TriBlock& B = blocks[bi];
            const int below = sturm_below(B.n, &d[B.off], &e2[B.off], tcut, pivmin);
            for (int j = below; j < B.n; j++) cand.push_back({0.0, (int)bi, j});
        }
    }
    // batches of 8 candidates of one block (each bisection converges alone)
    constexpr int NB = 8;
    std::vector<int> bstart;
    for (size_t c = 0; c < cand.size();) {
        bstart.push_back((int)c);
        size_t e0 = c;
        while (e0 < cand.size() && e0 - c < (size_t)NB && cand[e0].blk == cand[c].blk) e0++;
        c = e0;
    }
    parallel_for((int)bstart.size(), 1, [&](int bi) {
        const size_t c0 = bstart[bi];
        const TriBlock& B = blocks[cand[c0].blk];
        int js[NB];
        double out[NB];
        int nb = 0;
        while (nb < NB && c0 + nb < cand.size() && cand[c0 + nb].blk == cand[c0].blk) {
            js[nb] = cand[c0 + nb].idx;
            nb++;
        }
        for (int t = nb; t < NB; t++) js[t] = js[nb - 1];
        bisect_batch<NB>(B.n, &d[B.off], &e2[B.off], js, glo, ghi, pivmin, out);
        for (int t = 0; t < nb; t++) cand[c0 + t].v = out[t];
    });
    std::sort(cand.begin(), cand.end(), [](const EV& a, const EV& b) { return a.v > b.v; });
    cand.resize(std::min((size_t)nev, cand.size()));
    nev = (int)cand.size();
    evals.resize(nev);
    Z.assign((size_t)n * nev, 0.0);
    // clusters: per block, the wanted eigenvalues in descending order, split at gaps > ortol
    const double ortol = 1e-7 * tnorm;
    std::vector<std::vector<int>> per_block(blocks.size());
    for (int t = 0; t < nev; t++) {
        evals[t] = cand[t].v;
        per_block[cand[t].blk].push_back(t);
    }
    struct Cluster {
        int blk, first, count;
    };
    std::vector<Cluster> clusters;
    for (size_t bi = 0; bi < blocks.size(); bi++) {
        std::vector<int>& idx = per_block[bi];
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return evals[a] > evals[b]; });
        for (size_t t = 0; t < idx.size(); t++)
            if (t == 0 || std::fabs(evals[idx[t - 1]] - evals[idx[t]]) > ortol)
                clusters.push_back({(int)bi, (int)t, 1});
            else
                clusters.back().count++;
    }
    parallel_for((int)clusters.size(), 1, [&](int ci) {
        const Cluster& cl = clusters[ci];
        const TriBlock& B = blocks[cl.blk];
        const std::vector<int>& idx = per_block[cl.blk];
        const double* db = &d[B.off];
        const double* eb = B.n > 1 ? &e[B.off] : nullptr;
        std::vector<double> x(B.n);
        TriLU lu;
        std::mt19937_64 rng(1234567 + 1000003ULL * cl.blk + (uint64_t)cl.first);
        std::normal_distribution<double> nd;
        double prev = 0;
        for (int t = cl.first; t < cl.first + cl.count; t++) {
            const int col = idx[t];
            double lam = evals[col];
            // separate coincident shifts slightly (as LAPACK's dstein does)
            if (t > cl.first && std::fabs(prev - lam) < 10 * eps * tnorm) lam = prev - 10 * eps * tnorm;
            prev = lam;
            double* z = &Z[(size_t)col * n + B.off];
            if (B.n == 1) {
                z[0] = 1.0;
                continue;
            }
            lu.factor(B.n, db, eb, lam, eps * tnorm);
            for (int i = 0; i < B.n; i++) x[i] = nd(rng);
            for (int it = 0; it < 5; it++) {
                lu.solve(B.n, x.data());
                // orthogonalise against the earlier vectors of the cluster
                for (int s2 = cl.first; s2 < t; s2++) {
                    const double* zz = &Z[(size_t)idx[s2] * n + B.off];
                    double dot = 0;
                    for (int i = 0; i < B.n; i++) dot += zz[i] * x[i];
                    for (int i = 0; i < B.n; i++) x[i] -= dot * zz[i];
                }
                double nrm = 0;
                for (int i = 0; i < B.n; i++) nrm += x[i] * x[i];
                nrm = std::sqrt(nrm);
                if (nrm == 0) {
                    for (int i = 0; i < B.n; i++) x[i] = nd(rng);
                    continue;
                }
                for (int i = 0; i < B.n; i++) x[i] /= nrm;
            }
            for (int i = 0; i < B.n; i++) z[i] = x[i];
        }
    });
}

// Top-nev eigenpairs of the dense symmetric K x K device matrix dA (destroyed).
// evals (host, descending), dZ (device K x nev col-major).
static int dense_eigen_top(Dev& D, double* dA, int K, int nev, std::vector<double>& evals, double* dZ) {
    opsec_ctx* ctx = D.ctx;
    nev = std::min(nev, K);
    std::vector<double> d(K), e(std::max(1, K - 1), 0.0);
    double *dtau = nullptr, *de = nullptr, *dp = nullptr, *dd = nullptr, *dpart = nullptr;
    DevScratch tmp;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dpart, (size_t)((K + 255) / 256) * K * 8 + 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dtau, (size_t)K * 8 + 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&de, (size_t)K * 8 + 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dp, (size_t)std::max(K, nev) * 8 + 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dd, (size_t)K * 8 + 8));
    PhaseTimer pt;
    // Householder tridiagonalisation in panels of TR_NB columns: per column
    // the deferred updates of the panel's earlier reflectors are applied to
    // the column, its reflector formed, and w = tau p - (tau^2/2)(p.v) v from
    // p = A22 v corrected for the deferred updates; after the panel, the
    // trailing matrix A22 -= V W^T + W V^T as two MFMA GEMMs.  The symmetric
    // matrix-vector product is the only per-column pass over A22.
    constexpr int TR_NB = 32;
    double *dVr = nullptr, *dWr = nullptr, *dab = nullptr;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dVr, (size_t)K * TR_NB * 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dWr, (size_t)K * TR_NB * 8));
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dab, (size_t)2 * TR_NB * 8));
    for (int k0 = 0; k0 + 2 < K; k0 += TR_NB) {
        const int nbp = std::min(TR_NB, K - 2 - k0);
        for (int i = 0; i < nbp; i++) {
            const int k = k0 + i, m = K - k - 1;
            if (i > 0)
                hipLaunchKernelGGL(hh_colupd_kernel, dim3((K - k + 255) / 256), dim3(256), 0, D.s, dA, K, k, i, dVr,
                                   dWr);
            hipLaunchKernelGGL(hh_vector_kernel, dim3(1), dim3(1024), 0, D.s, dA, K, k, dtau, de);
            const int nc = (m + 255) / 256;
            hipLaunchKernelGGL(hh_symv_kernel, dim3(nc, nc), dim3(256), 0, D.s, dA, K, k, dpart);
            hipLaunchKernelGGL(hh_symv_sum_kernel, dim3(nc), dim3(256), 0, D.s, dpart, nc, m, dp);
            if (i > 0) {
                hipLaunchKernelGGL(hh_dots_kernel, dim3(2 * i), dim3(256), 0, D.s, dA, K, k, i, dVr, dWr, dab);
                hipLaunchKernelGGL(hh_corr_kernel, dim3(nc), dim3(256), 0, D.s, dp, K, k, i, dVr, dWr, dab);
            }
            hipLaunchKernelGGL(hh_w_kernel, dim3(1), dim3(1024), 0, D.s, dA, K, k, dtau, dp);
            hipLaunchKernelGGL(hh_store_kernel, dim3((K + 255) / 256), dim3(256), 0, D.s, dA, K, k, i, dp, dVr, dWr);
        }
        const int r0 = k0 + nbp, m2 = K - r0;
        if (m2 > 0) {
            double* A22 = dA + (size_t)r0 * K + r0;
            D.gemm_nt_sub(A22, K, dVr + r0, K, dWr + r0, K, m2, nbp, m2);
            D.gemm_nt_sub(A22, K, dWr + r0, K, dVr + r0, K, m2, nbp, m2);
        }
    }
    hipLaunchKernelGGL(diag_kernel, dim3((K + 255) / 256), dim3(256), 0, D.s, dA, K, dd, de + (K >= 2 ? K - 2 : 0));
    int rc = D.sync();
    if (rc) return rc;
    OPSEC_HIP_TRY(ctx, stream_copy(ctx, d.data(), dd, (size_t)K * 8, hipMemcpyDeviceToHost));
    if (K >= 2) OPSEC_HIP_TRY(ctx, stream_copy(ctx, e.data(), de, (size_t)(K - 1) * 8, hipMemcpyDeviceToHost));
    g_klt_times[KT_TRIRED] += pt.lap();
    std::vector<double> Z;
    tridiag_top(K, d, e, nev, evals, Z);
    if ((int)evals.size() < nev || Z.size() < (size_t)K * nev)
        return set_error(ctx, OPSEC_ECONV, "tridiagonal eigen-solver: %d of %d eigenpairs", (int)evals.size(), nev);
    OPSEC_HIP_TRY(ctx, stream_copy(ctx, dZ, Z.data(), (size_t)K * nev * 8, hipMemcpyHostToDevice));
    g_klt_times[KT_TRIEIG] += pt.lap();
    // Back-transformation Z <- H_0 H_1 ... H_(K-3) Z in panels of HH_NB
    // reflectors, last panel first: H_k0 ... H_(k0+nb-1) = I - V T V^T
    // (compact WY, T upper triangular: T_jj = tau_j, T[0:j, j] = -tau_j
    // T[0:j, 0:j] V[:, 0:j]^T v_j), applied as three GEMMs on rows k0+1..K-1.
    const int nrefl = std::max(0, K - 2);
    if (nrefl > 0) {
        constexpr int HH_NB = 64;
        std::vector<double> tau(nrefl);
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, tau.data(), dtau, (size_t)nrefl * 8, hipMemcpyDeviceToHost));
        double *dVp = nullptr, *dS = nullptr, *dT = nullptr, *dW = nullptr, *dW2 = nullptr;
        OPSEC_HIP_TRY(ctx, tmp.alloc(&dVp, (size_t)K * HH_NB * 8));
        OPSEC_HIP_TRY(ctx, tmp.alloc(&dS, (size_t)HH_NB * HH_NB * 8));
        OPSEC_HIP_TRY(ctx, tmp.alloc(&dT, (size_t)HH_NB * HH_NB * 8));
        OPSEC_HIP_TRY(ctx, tmp.alloc(&dW, (size_t)HH_NB * nev * 8));
        OPSEC_HIP_TRY(ctx, tmp.alloc(&dW2, (size_t)HH_NB * nev * 8));
        std::vector<double> S((size_t)HH_NB * HH_NB), T((size_t)HH_NB * HH_NB);
        for (int k0 = ((nrefl - 1) / HH_NB) * HH_NB; k0 >= 0; k0 -= HH_NB) {
            const int nb = std::min(HH_NB, nrefl - k0), mp = K - k0 - 1;
            hipLaunchKernelGGL(hh_panel_kernel, dim3((K + 255) / 256, nb), dim3(256), 0, D.s, dA, K, k0, nb, dVp);
            D.gemm_tn(dVp + k0 + 1, K, dVp + k0 + 1, K, mp, nb, nb, dS);   // S = V^T V
            if ((rc = D.sync())) return rc;
            OPSEC_HIP_TRY(ctx, stream_copy(ctx, S.data(), dS, (size_t)nb * nb * 8, hipMemcpyDeviceToHost));
            std::fill(T.begin(), T.end(), 0.0);
            for (int j = 0; j < nb; j++) {
                const double tj = tau[k0 + j];
                T[(size_t)j * nb + j] = tj;
                for (int i = 0; i < j; i++) {
                    double acc = 0.0;   // (T[0:j, 0:j] V[:, 0:j]^T v_j)_i, T upper triangular
                    for (int l = i; l < j; l++) acc += T[(size_t)l * nb + i] * S[(size_t)j * nb + l];
                    T[(size_t)j * nb + i] = -tj * acc;
                }
            }
            OPSEC_HIP_TRY(ctx, stream_copy(ctx, dT, T.data(), (size_t)nb * nb * 8, hipMemcpyHostToDevice));
            D.gemm_tn(dVp + k0 + 1, K, dZ + k0 + 1, K, mp, nb, nev, dW);          // W = V^T Z
            D.gemm_nn(0, dW2, nb, dT, nb, dW, nb, nb, nb, nev);                  // W2 = T W
            D.gemm_nn(1, dZ + k0 + 1, K, dVp + k0 + 1, K, dW2, nb, mp, nb, nev);  // Z -= V W2
        }
    }
    int rc2 = D.sync();
    g_klt_times[KT_BACK] += pt.lap();
    g_klt_times[KT_NRR] += 1;
    return rc2;
}

// ---- block orthonormalisation (SVQB) -------------------------------------------

static void host_sym_eigen(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[i * n + i] = 1;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-32 * tot || tot == 0) break;
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
    w.resize(n);
    for (int i = 0; i < n; i++) w[i] = A[i * n + i];
}

// Cholesky QR of a block with Gram matrix G (b x b, symmetric): G = R^T R,
// X R^-1 orthonormal.  Mh = R^-1 (column-major), Bnew = R (row-major).
// Returns false -- and the caller uses the eigen-decomposition instead -- when
// the block is too ill-conditioned for it (smallest pivot^2 below 1e-10 of
// the largest diagonal entry, i.e. cond(X) above about 1e5).
static bool host_cholesky_qr(int b, const std::vector<double>& G, std::vector<double>& Mh, std::vector<double>& Bnew) {
    std::vector<double> R((size_t)b * b, 0.0);   // row-major upper triangle
    double gmax = 0;
    for (int i = 0; i < b; i++) gmax = std::max(gmax, G[(size_t)i * b + i]);
    if (!(gmax > 0)) return false;
    for (int j = 0; j < b; j++) {
        double djj = G[(size_t)j * b + j];
        for (int l = 0; l < j; l++) djj -= R[(size_t)l * b + j] * R[(size_t)l * b + j];
        if (!(djj > 1e-10 * gmax)) return false;
        const double rjj = std::sqrt(djj);
        R[(size_t)j * b + j] = rjj;
        for (int i = j + 1; i < b; i++) {
            double v = G[(size_t)i * b + j];   // G symmetric: G(j, i)
            for (int l = 0; l < j; l++) v -= R[(size_t)l * b + j] * R[(size_t)l * b + i];
            R[(size_t)j * b + i] = v / rjj;
        }
    }
    // R^-1 (upper triangular), column by column
    std::vector<double> Ri((size_t)b * b, 0.0);   // row-major
    for (int c = 0; c < b; c++) {
        Ri[(size_t)c * b + c] = 1.0 / R[(size_t)c * b + c];
        for (int r = c - 1; r >= 0; r--) {
            double v = 0.0;
            for (int l = r + 1; l <= c; l++) v += R[(size_t)r * b + l] * Ri[(size_t)l * b + c];
            Ri[(size_t)r * b + c] = -v / R[(size_t)r * b + r];
        }
    }
    Mh.assign((size_t)b * b, 0.0);
    for (int i = 0; i < b; i++)
        for (int k = 0; k < b; k++) Mh[(size_t)i * b + k] = Ri[(size_t)k * b + i];   // (k, i) column-major
    Bnew = R;
    return true;
}

struct LanczosWork {
    double* V = nullptr;      // N x (Kmax + b)
    double* W = nullptr;      // N x b
    double* H = nullptr;      // (Kmax + b) x b
    double* small = nullptr;  // b x b
    DevScratch mem;           // owns the buffers above: freed on every return path
};

// Orthonormalise the N x b block X against V[:, 0:kprev] (twice) and itself
// (Cholesky QR or SVQB, twice).  On return X = X_orth * Bmat (b x b, row-major host) for the
// part of X inside the new directions; dependent directions are replaced by
// random orthonormal ones with zero coefficient.
static int orthonormalise_block(Dev& D, LanczosWork& L, int N, int b, int kprev, double* X, std::vector<double>& Bmat,
                                std::mt19937_64& rng) {
    opsec_ctx* ctx = D.ctx;
    std::normal_distribution<double> nd;
    Bmat.assign((size_t)b * b, 0.0);
    for (int i = 0; i < b; i++) Bmat[i * b + i] = 1.0;
    for (int round = 0; round < 3; round++) {
        // project out previous basis, twice
        for (int pass = 0; pass < 2 && kprev > 0; pass++) {
            D.gemm_tn(L.V, N, X, N, N, kprev, b, L.H);
            D.gemm_nn(1, X, N, L.V, N, L.H, kprev, N, kprev, b);
        }
        int rc = D.sync();
        if (rc) return rc;
        // SVQB twice
        bool replaced = false;
        for (int pass = 0; pass < 2; pass++) {
            D.gemm_tn(X, N, X, N, N, b, b, L.small);
            if ((rc = D.sync())) return rc;
            std::vector<double> G((size_t)b * b), w, U;
            OPSEC_HIP_TRY(ctx, stream_copy(ctx, G.data(), L.small, (size_t)b * b * 8, hipMemcpyDeviceToHost));
            for (int i = 0; i < b; i++)
                for (int j = 0; j < i; j++) G[i * b + j] = G[j * b + i] = 0.5 * (G[i * b + j] + G[j * b + i]);
            std::vector<double> Mh, Bnew;
            std::vector<int> bad;
            // well-conditioned blocks (the usual case): Cholesky QR; otherwise
            // SVQB through the eigen-decomposition, which finds and replaces
            // the dependent directions
            if (!host_cholesky_qr(b, G, Mh, Bnew)) {
                host_sym_eigen(b, G, w, U);   // G = U diag(w) U^T, U column i = eigvec i (U[k*b+i])
                double wmax = *std::max_element(w.begin(), w.end());
                const double thr = std::max(wmax, 1e-300) * 1e-20;
                // X <- X U diag(1/sqrt(w)); coefficients Bnew = diag(sqrt(w)) U^T
                Mh.assign((size_t)b * b, 0.0);
                Bnew.assign((size_t)b * b, 0.0);
                for (int i = 0; i < b; i++) {
                    bool ok = w[i] > thr;
                    if (!ok) bad.push_back(i);
                    double sw = ok ? std::sqrt(w[i]) : 0.0;
                    for (int k = 0; k < b; k++) {
                        Mh[(size_t)i * b + k] = ok ? U[k * b + i] / sw : 0.0;   // column-major b x b: (k, i)
                        Bnew[i * b + k] = sw * U[k * b + i];
                    }
                }
            }
            OPSEC_HIP_TRY(ctx, stream_copy(ctx, L.small, Mh.data(), (size_t)b * b * 8, hipMemcpyHostToDevice));
            OPSEC_HIP_TRY(ctx, stream_copy(ctx, L.W, X, (size_t)N * b * 8, hipMemcpyDeviceToDevice));
            D.gemm_nn(0, X, N, L.W, N, L.small, b, N, b, b);
            if ((rc = D.sync())) return rc;
            // Bmat <- Bnew * Bmat
            std::vector<double> T((size_t)b * b, 0.0);
            for (int i = 0; i < b; i++)
                for (int k = 0; k < b; k++)
                    for (int j = 0; j < b; j++) T[i * b + j] += Bnew[i * b + k] * Bmat[k * b + j];
            Bmat.swap(T);
            if (!bad.empty()) {
                // replace dependent directions with random vectors (zero coefficient rows)
                std::vector<double> r((size_t)N);
                for (int c : bad) {
                    for (int i = 0; i < N; i++) r[i] = nd(rng);
                    OPSEC_HIP_TRY(ctx, stream_copy(ctx, X + (size_t)c * N, r.data(), (size_t)N * 8, hipMemcpyHostToDevice));
                    for (int j = 0; j < b; j++) Bmat[c * b + j] = 0.0;
                }
                replaced = true;
                break;   // redo projection against the basis
            }
        }
        if (!replaced) return OPSEC_OK;
    }
    return OPSEC_OK;
}

// Block Lanczos on the dense symmetric device matrix S (N x N row-major).
static int block_lanczos(opsec_ctx* ctx, const double* S, int N, int nev, double tol, int b, int maxdim, double* evals,
                         double* evecs, int* nconv) {
    Dev D(ctx, ctx->stream);
    OPSEC_HIP_TRY(ctx, D.init());
    for (double& t : g_klt_times) t = 0.0;
    const auto t_solve = std::chrono::steady_clock::now();
    nev = std::min(nev, N);
    if (b <= 0) b = std::min(128, std::max(1, N / 4));
    b = std::max(1, std::min(b, N));
    if (maxdim <= 0) maxdim = std::max(3 * nev + 4 * b, 16 * b);
    if (maxdim + b > N) {
        // the Krylov space would span (nearly) the whole space: the dense
        // symmetric eigen-solver on S itself gives the same answer directly
        DevScratch mem;
        double *dA = nullptr, *dZ = nullptr;
        OPSEC_HIP_TRY(ctx, mem.alloc(&dA, (size_t)N * N * 8));
        OPSEC_HIP_TRY(ctx, mem.alloc(&dZ, (size_t)N * nev * 8));
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dA, S, (size_t)N * N * 8, hipMemcpyDeviceToDevice, ctx->stream));
        std::vector<double> ev;
        int rc = dense_eigen_top(D, dA, N, nev, ev, dZ);
        if (rc) return rc;
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, evecs, dZ, (size_t)N * nev * 8, hipMemcpyDeviceToHost));
        for (int c = 0; c < nev; c++) evals[c] = ev[c];
        *nconv = nev;
        return OPSEC_OK;
    }
    int Kmax = std::min(N, ((maxdim + b - 1) / b) * b);
    if (Kmax < nev) Kmax = std::min(N, ((nev + b - 1) / b) * b);
    Kmax = std::max(b, (Kmax / b) * b);
    LanczosWork L;
    size_t ncols = (size_t)Kmax + b;
    if (L.mem.alloc(&L.V, (size_t)N * ncols * 8) != hipSuccess || L.mem.alloc(&L.W, (size_t)N * b * 8) != hipSuccess ||
        L.mem.alloc(&L.H, ncols * b * 8) != hipSuccess || L.mem.alloc(&L.small, (size_t)b * b * 8 + 8) != hipSuccess)
        return set_error(ctx, OPSEC_ENOMEM, "Lanczos workspace (%d x %zu) does not fit", N, ncols);
    std::mt19937_64 rng(42);
    std::normal_distribution<double> nd;
    // starting block
    {
        std::vector<double> r((size_t)N * b);
        for (auto& x : r) x = nd(rng);
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, L.V, r.data(), r.size() * 8, hipMemcpyHostToDevice));
        std::vector<double> Bm;
        int rc = orthonormalise_block(D, L, N, b, 0, L.V, Bm, rng);
        if (rc) return rc;
    }
    std::vector<double> T((size_t)Kmax * Kmax, 0.0);   // column-major block tridiagonal
    int check_at = std::min(Kmax, std::max(((nev * 3 / 2 + b - 1) / b) * b, 2 * b));
    std::vector<double> ritz;
    // Ritz workspace: T of the current step and its wanted eigenvectors,
    // owned here so that every return path frees them
    struct RitzBufs {
        double* T = nullptr;
        double* Z = nullptr;
        ~RitzBufs() {
            if (T) (void)hipFree(T);
            if (Z) (void)hipFree(Z);
        }
    } rb;
    double*& dZ = rb.Z;
    int converged = 0, Kr = 0;
    int rc = OPSEC_OK;
    for (int j = 0;; j++) {
        double* Vj = L.V + (size_t)j * b * N;
        const int K = (j + 1) * b;              // basis size including block j
        double* Xn = L.V + (size_t)K * N;       // next block slot (ncols = Kmax + b)
        PhaseTimer pt;
        D.apply(S, N, Vj, b, Xn);               // Xn = S V_j
        if ((rc = D.sync())) break;
        g_klt_times[KT_MATVEC] += pt.lap();
        g_klt_times[KT_NSTEPS] += 1;
        D.gemm_tn(Vj, N, Xn, N, N, b, b, L.small);   // A_j = V_j^T S V_j
        if ((rc = D.sync())) break;
        std::vector<double> Aj((size_t)b * b);
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, Aj.data(), L.small, (size_t)b * b * 8, hipMemcpyDeviceToHost));
        for (int c = 0; c < b; c++)
            for (int r = 0; r < b; r++)
                T[(size_t)(j * b + c) * Kmax + j * b + r] = 0.5 * (Aj[(size_t)c * b + r] + Aj[(size_t)r * b + c]);
        const bool last = K >= Kmax;
        std::vector<double> Bm((size_t)b * b, 0.0);   // coupling: S V_j = ... + V_{j+1} Bm
        if (K < N) {
            if ((rc = orthonormalise_block(D, L, N, b, K, Xn, Bm, rng))) break;
            g_klt_times[KT_ORTH] += pt.lap();
            if (!last)
                for (int r = 0; r < b; r++)
                    for (int c = 0; c < b; c++) {
                        double v = Bm[(size_t)r * b + c];
                        T[(size_t)(j * b + c) * Kmax + (j + 1) * b + r] = v;
                        T[(size_t)((j + 1) * b + r) * Kmax + j * b + c] = v;
                    }
        }
        if (!last && K < check_at) continue;
        // Rayleigh-Ritz on T[0:K, 0:K]; residual of Ritz pair z: |Bm z[K-b:K]|
        if (rb.T) (void)hipFree(rb.T);
        rb.T = nullptr;
        if (hipMalloc(&rb.T, (size_t)K * K * 8) != hipSuccess) {
            rb.T = nullptr;
            rc = set_error(ctx, OPSEC_ENOMEM, "Ritz workspace");
            break;
        }
        double* dT = rb.T;
        std::vector<double> Tk((size_t)K * K);
        for (int c = 0; c < K; c++)
            for (int r = 0; r < K; r++) Tk[(size_t)c * K + r] = T[(size_t)c * Kmax + r];
        OPSEC_HIP_TRY(ctx, stream_copy(ctx, dT, Tk.data(), Tk.size() * 8, hipMemcpyHostToDevice));
        hipFree(dZ);
        dZ = nullptr;
        const int want = std::min(nev, K);
        if (hipMalloc(&dZ, (size_t)K * want * 8) != hipSuccess) {
            dZ = nullptr;
            rc = set_error(ctx, OPSEC_ENOMEM, "Ritz vectors");
            break;
        }
        rc = dense_eigen_top(D, dT, K, want, ritz, dZ);
        (void)hipFree(rb.T);
        rb.T = nullptr;
        if (rc) break;
        Kr = K;
        std::vector<double> Zl((size_t)b * want);
        for (int c = 0; c < want; c++)
            OPSEC_HIP_TRY(ctx, stream_copy(ctx, &Zl[(size_t)c * b], dZ + (size_t)c * K + (K - b), (size_t)b * 8,
                                         hipMemcpyDeviceToHost));
        double tmax = 0;
        for (double t : ritz) tmax = std::max(tmax, std::fabs(t));
        converged = 0;
        for (int c = 0; c < want; c++) {
            double s2 = 0;
            for (int r = 0; r < b; r++) {
                double acc = 0;
                for (int q2 = 0; q2 < b; q2++) acc += Bm[(size_t)r * b + q2] * Zl[(size_t)c * b + q2];
                s2 += acc * acc;
            }
            if (std::sqrt(s2) <= tol * tmax) converged++;
            else break;   // the leading run of converged pairs
        }
        if (std::getenv("OPSEC_KLT_PROGRESS"))
            std::fprintf(stderr, "klt: Krylov dimension %d of %d, %d of %d Ritz pairs converged (%.1f s)\n", K, Kmax,
                         converged, nev, std::chrono::duration<double>(
                                             std::chrono::steady_clock::now() - t_solve).count());
        if (converged >= nev || last) break;
        check_at = std::min(Kmax, ((K * 5 / 4 + b - 1) / b) * b);
    }
    if (!rc) {
        if (!dZ) {
            rc = set_error(ctx, OPSEC_ECONV, "Lanczos stopped before a Rayleigh-Ritz step");
        } else {
            int want = std::min(nev, Kr);
            // eigenvectors Y = V[:, 0:Kr] Z  (N x want, column-major = mode-major)
            double* dY = nullptr;
            DevScratch ymem;
            if (ymem.alloc(&dY, (size_t)N * want * 8) != hipSuccess) {
                rc = set_error(ctx, OPSEC_ENOMEM, "Ritz vector output");
            } else {
                PhaseTimer pt;
                D.gemm_nn(0, dY, N, L.V, N, dZ, Kr, N, Kr, want);
                rc = D.sync();
                g_klt_times[KT_RITZVEC] += pt.lap();
                g_klt_times[KT_KDIM] = Kr;
                if (!rc) {
                    hipError_t e = stream_copy(ctx, evecs, dY, (size_t)N * want * 8, hipMemcpyDeviceToHost);
                    if (e != hipSuccess) rc = set_error(ctx, OPSEC_EHIP, "copy eigenvectors: %s", hipGetErrorString(e));
                }
            }
            for (int c = 0; c < want; c++) evals[c] = ritz[c];
            for (int c = want; c < nev; c++) evals[c] = 0.0;
            *nconv = std::min(converged, nev);
        }
    }
    return rc;
}

// block_lanczos with the automatic Krylov dimension (maxdim <= 0): when some
// of the nev modes have not converged (slowly decaying spectra, e.g. spherical
// shells), run again with twice the space, up to N, at most twice.  A retry
// solves into its own buffers and is adopted only when it succeeds with more
// converged modes; a failed retry (e.g. a larger space that does not fit)
// keeps the earlier result.
static int lanczos_auto(opsec_ctx* ctx, const double* dS, int N, int nev, double tol, int block, int maxdim,
                        double* evals, double* evecs, int* nconv) {
    int rc = block_lanczos(ctx, dS, N, nev, tol, block, maxdim, evals, evecs, nconv);
    if (rc || maxdim > 0) return rc;
    const int b = block > 0 ? std::min(block, N) : std::min(128, std::max(1, N / 4));
    int md = std::max(3 * std::min(nev, N) + 4 * b, 16 * b);
    std::vector<double> ev2, evec2;
    for (int retry = 0; retry < 2 && *nconv < nev && md < N; retry++) {
        md = std::min(N, 2 * md);
        int nconv2 = 0;
        ev2.resize((size_t)nev);
        evec2.resize((size_t)nev * N);
        if (block_lanczos(ctx, dS, N, nev, tol, block, md, ev2.data(), evec2.data(), &nconv2)) break;
        if (nconv2 > *nconv) {
            std::copy(ev2.begin(), ev2.end(), evals);
            std::copy(evec2.begin(), evec2.end(), evecs);
            *nconv = nconv2;
        }
    }
    return OPSEC_OK;
}

}  // namespace opsec

using namespace opsec;

extern "C" {

int opsec_lanczos_dense(opsec_ctx* ctx, int n, const double* A, int nev, double tol, int block, int maxdim,
                        double* evals, double* evecs, int* nconv) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, n >= 1 && A && evals && evecs && nconv, "bad arguments");
    OPSEC_CHECK_ARG(ctx, nev >= 1 && nev <= n, "requesting %d modes from a %d-dimensional space", nev, n);
    OPSEC_CHECK_ARG(ctx, tol > 0, "tolerance must be positive");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* dS = nullptr;
    DevScratch tmp;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dS, (size_t)n * n * 8));
    OPSEC_HIP_TRY(ctx, stream_copy(ctx, dS, A, (size_t)n * n * 8, hipMemcpyHostToDevice));
    return lanczos_auto(ctx, dS, n, nev, tol, block, maxdim, evals, evecs, nconv);
}

int opsec_klt(opsec_ctx* ctx, int nev, double tol, int block, int maxdim, double* evals, double* evecs, int* nconv) {
    if (!ctx) return OPSEC_EINVAL;
    if (ctx->ncells == 0) return set_error(ctx, OPSEC_ESTATE, "no survey uploaded");
    if (ctx->Q_valid.empty() || !ctx->Q_valid[0]) return set_error(ctx, OPSEC_ESTATE, "prior pair table not built");
    const int N = ctx->ncells;
    OPSEC_CHECK_ARG(ctx, nev >= 1 && nev <= N, "klt: requesting %d modes from a %d-dimensional space", nev, N);
    OPSEC_CHECK_ARG(ctx, evals && evecs && nconv, "NULL outputs");
    OPSEC_CHECK_ARG(ctx, tol > 0, "tolerance must be positive");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* dS = nullptr;
    DevScratch tmp;
    OPSEC_HIP_TRY(ctx, tmp.alloc(&dS, (size_t)N * N * 8));
    int rc = opsec_assemble_device(ctx, 0, dS, N);
    if (!rc) rc = lanczos_auto(ctx, dS, N, nev, tol, block, maxdim, evals, evecs, nconv);
    return rc;
}

int opsec_diag_klt_times(double* out, int n) {
    if (!out || n < 1) return OPSEC_EINVAL;
    for (int i = 0; i < n && i < KT_N; i++) out[i] = g_klt_times[i];
    return KT_N;
}

// Host tridiagonal eigen-solver, exposed for CPU unit tests (not in the public header).
int opsec_diag_tridiag_top(int n, const double* d, const double* e, int nev, double* evals, double* Z) {
    if (n < 1 || nev < 1 || nev > n) return OPSEC_EINVAL;
    std::vector<double> dd(d, d + n), ee(e, e + std::max(0, n - 1)), ev, Zv;
    if (ee.empty()) ee.push_back(0.0);
    tridiag_top(n, dd, ee, nev, ev, Zv);
    if ((int)ev.size() < nev) return OPSEC_ECONV;
    std::copy(ev.begin(), ev.end(), evals);
    std::copy(Zv.begin(), Zv.end(), Z);
    return OPSEC_OK;
}

}  // extern "C"
