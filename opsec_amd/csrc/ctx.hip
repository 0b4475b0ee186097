// Context, device buffers and the xi model (the xi tables on the device).
//
// The RealModel tables follow src/RealModel.cpp:98-127 and src/Model.cpp:84-123:
//   xi_0^2(r) = (1/2pi^2) int_{kmin}^{kmax} k^2 P(k) j0(kr) e^{-(k/kdamp)^2} dk
// by composite Simpson on Nk intervals, evaluated at the 768 QSpline nodes.
// The prior uses the pkfile cubic spline, Nk = 32768 over [0,10]; each band
// uses P = 1 on [klo, khi) (a constant spline that is zero at k >= khi),
// Nk = 8192 over [klo, khi].  Both run on the GPU: one thread per (function,
// r node) accumulating the Simpson sum in the reference's order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.h"
#include "xi_dev.h"

namespace opsec {

int set_error(opsec_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        static std::mutex m;   // a shell's pair-table sweeps run on two host threads
        std::lock_guard<std::mutex> lk(m);
        ctx->err = buf;
    }
    return code;
}

int dev_realloc(opsec_ctx* ctx, void** p, size_t* cap, size_t bytes) {
    if (*p && *cap >= bytes) return OPSEC_OK;
    if (*p) {
        hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    if (bytes == 0) return OPSEC_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(ctx, OPSEC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    *cap = bytes;
    return OPSEC_OK;
}

void dev_free(void** p) {
    if (*p) hipFree(*p);
    *p = nullptr;
}

XiModelDev model_view(const opsec_ctx* ctx) {
    XiModelDev m;
    m.desc = ctx->d_desc;
    m.y = ctx->d_tables;
    m.xgrid = ctx->d_xgrid;
    m.ugrid = ctx->d_ugrid;
    m.x2 = ctx->h_xgrid.empty() ? QS_X2 : ctx->h_xgrid[QS_N - 1];
    m.uhi = ctx->h_ugrid.empty() ? std::log(QS_X2) : ctx->h_ugrid[QS_N2 - 1];
    return m;
}

// ---- small dense symmetric eigen-solver (cyclic Jacobi), host -------------
static void jacobi_eigen(int n, std::vector<double>& A, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; i++) {
            diag += A[(size_t)i * n + i] * A[(size_t)i * n + i];
            for (int j = i + 1; j < n; j++) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        }
        if (off <= 1e-34 * diag) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
}

// Golub-Welsch: nodes/weights from the Jacobi matrix of the orthogonal
// polynomials (diag a[0..n-1], off-diag b[1..n-1]) and the weight's mass mu0.
static void golub_welsch(int n, const double* a, const double* b, double mu0, double* x, double* w) {
    std::vector<double> A((size_t)n * n, 0.0), V;
    for (int i = 0; i < n; i++) A[(size_t)i * n + i] = a[i];
    for (int i = 1; i < n; i++) A[(size_t)i * n + i - 1] = A[(size_t)(i - 1) * n + i] = b[i];
    jacobi_eigen(n, A, V);
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int i, int j) { return A[(size_t)i * n + i] < A[(size_t)j * n + j]; });
    for (int k = 0; k < n; k++) {
        int e = idx[k];
        x[k] = A[(size_t)e * n + e];
        w[k] = mu0 * V[e] * V[e];   // first component of eigenvector e
    }
}

void gauss_legendre(int n, double* x, double* w) {
    std::vector<double> a(n, 0.0), b(n, 0.0);
    for (int k = 1; k < n; k++) b[k] = k / std::sqrt(4.0 * k * k - 1.0);
    golub_welsch(n, a.data(), b.data(), 2.0, x, w);
}

void gauss_jacobi10(int n, double* t, double* w) {
    // Weight (1-x) on [-1,1]: Jacobi alpha=1, beta=0.
    std::vector<double> a(n), b(n, 0.0), x(n), wx(n);
    for (int k = 0; k < n; k++) a[k] = -1.0 / ((2.0 * k + 1) * (2.0 * k + 3));
    for (int k = 1; k < n; k++) b[k] = std::sqrt((double)k * (k + 1)) / (2.0 * k + 1);
    golub_welsch(n, a.data(), b.data(), 2.0, x.data(), wx.data());
    // map to [0,1] with weight (1-t): t = (x+1)/2, w_t = w_x / 4
    for (int k = 0; k < n; k++) {
        t[k] = 0.5 * (x[k] + 1.0);
        w[k] = 0.25 * wx[k];
    }
}

int survey_from_abi(opsec_ctx* ctx, const opsec_survey* in, SurveyDev* out) {
    OPSEC_CHECK_ARG(ctx, in, "survey is NULL");
    OPSEC_CHECK_ARG(ctx, in->coordsys == OPSEC_COORD_CARTESIAN || in->coordsys == OPSEC_COORD_SPHERICAL,
                    "unknown coordinate system %d", in->coordsys);
    OPSEC_CHECK_ARG(ctx, in->n1 > 0 && in->n2 > 0 && in->n3 > 0, "grid sizes must be positive");
    OPSEC_CHECK_ARG(ctx, in->min1 < in->max1 && in->min2 < in->max2 && in->min3 < in->max3,
                    "survey bounds must satisfy Min < Max");
    if (in->coordsys == OPSEC_COORD_CARTESIAN)
        OPSEC_CHECK_ARG(ctx, in->L > 0, "box period L must be positive");
    out->coordsys = in->coordsys;
    out->n1 = in->n1;
    out->n2 = in->n2;
    out->n3 = in->n3;
    out->min1 = in->min1;
    out->max1 = in->max1;
    out->min2 = in->min2;
    out->max2 = in->max2;
    out->min3 = in->min3;
    out->max3 = in->max3;
    out->L = in->L;
    return OPSEC_OK;
}

// ---- cubic spline of the pkfile (Forsythe-Malcolm-Moler end conditions) ----
// Builds b, c, d with y(x) = Y_i + b_i u + c_i u^2 + d_i u^3, u = x - X_i:
// not-a-knot-like ends whose third derivatives match cubic divided
// differences of the first/last four points (src/Spline.cpp:258-333).
static void fmm_spline(int n, const double* X, const double* Y, double* b, double* c, double* d) {
    if (n == 2) {
        b[0] = b[1] = (Y[1] - Y[0]) / (X[1] - X[0]);
        c[0] = c[1] = d[0] = d[1] = 0;
        return;
    }
    int nm1 = n - 1;
    // tridiagonal system: b = diagonal, d = off-diagonal, c = right-hand side
    d[0] = X[1] - X[0];
    c[1] = (Y[1] - Y[0]) / d[0];
    for (int i = 1; i < nm1; i++) {
        d[i] = X[i + 1] - X[i];
        b[i] = 2 * (d[i - 1] + d[i]);
        c[i + 1] = (Y[i + 1] - Y[i]) / d[i];
        c[i] = c[i + 1] - c[i];
    }
    b[0] = -d[0];
    b[nm1] = -d[n - 2];
    c[0] = 0;
    c[nm1] = 0;
    if (n > 3) {
        c[0] = c[2] / (X[3] - X[1]) - c[1] / (X[2] - X[0]);
        c[nm1] = c[n - 2] / (X[nm1] - X[n - 3]) - c[n - 3] / (X[n - 2] - X[n - 4]);
        c[0] = c[0] * d[0] * d[0] / (X[3] - X[0]);
        c[nm1] = -c[nm1] * d[n - 2] * d[n - 2] / (X[nm1] - X[n - 4]);
    }
    for (int i = 1; i < n; i++) {   // elimination
        double t = d[i - 1] / b[i - 1];
        b[i] -= t * d[i - 1];
        c[i] -= t * c[i - 1];
    }
    c[nm1] /= b[nm1];               // back substitution
    for (int i = n - 2; i >= 0; i--) c[i] = (c[i] - d[i] * c[i + 1]) / b[i];
    b[nm1] = (Y[nm1] - Y[n - 2]) / d[n - 2] + d[n - 2] * (c[n - 2] + 2 * c[nm1]);
    for (int i = 0; i < nm1; i++) {
        b[i] = (Y[i + 1] - Y[i]) / d[i] - d[i] * (c[i + 1] + 2 * c[i]);
        d[i] = (c[i + 1] - c[i]) / d[i];
        c[i] = 3 * c[i];
    }
    c[nm1] = 3 * c[nm1];
    d[nm1] = d[n - 2];
}

// ---- device kernels --------------------------------------------------------

struct SplineDev {
    int kind;          // 0 cubic (pk), 1 constant
    int n;
    const double* X;
    const double* Y;
    const double* b;
    const double* c;
    const double* d;
    double value, xmin, xmax;   // constant spline
};

__device__ double spline_eval(const SplineDev& s, double x) {
    if (s.kind == 1) return (x < s.xmin || x >= s.xmax) ? 0.0 : s.value;
    if (x < s.X[0] || x >= s.X[s.n - 1]) return 0.0;
    int i = lookup_index(x, s.n, s.X);
    double u = x - s.X[i];
    // Y + b u + c u^2 + d u^3, evaluated left to right without contraction
    double r = __dadd_rn(s.Y[i], __dmul_rn(s.b[i], u));
    r = __dadd_rn(r, __dmul_rn(__dmul_rn(s.c[i], u), u));
    r = __dadd_rn(r, __dmul_rn(__dmul_rn(__dmul_rn(s.d[i], u), u), u));
    return r;
}

struct SimpsonJob {
    SplineDev P;
    int Nk;
    double kmin, kmax, kdamp;
    double* mult;     // Nk+1 workspace
    double* kgrid;    // Nk+1 workspace
    double* out;      // QS_N outputs
};

// Simpson multipliers (src/Model.cpp:104-114), l = 0, m = 2.
__global__ void simpson_mult_kernel(const SimpsonJob* jobs) {
    const SimpsonJob& J = jobs[blockIdx.y];
    const double dk = (J.kmax - J.kmin) / J.Nk;
    const double norm = (dk / 3) / (2 * M_PI * M_PI);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j <= J.Nk; j += gridDim.x * blockDim.x) {
        double k = __dadd_rn(J.kmin, __dmul_rn((double)j, dk));
        double m = (double)(2 + 2 * (j % 2) - (j == 0) - (j == J.Nk));
        double damp = exp(__ddiv_rn(-k * k, J.kdamp * J.kdamp));
        m = __dmul_rn(m, __dmul_rn(__dmul_rn(spline_eval(J.P, k), __dmul_rn(k, k)), damp));
        m = __dmul_rn(m, norm);
        J.kgrid[j] = k;
        J.mult[j] = m;
    }
}

__device__ __forceinline__ double sj0_ref(double x) {
    double x2 = x * x;
    if (x2 < 1e-4)
        return 1 - __dmul_rn(x2, __dadd_rn(1 / 6., __dmul_rn(x2, __dsub_rn(1 / 120., __dmul_rn(x2, 1 / 5040.)))));
    return sin(x) / x;
}

// One thread per (job, r node): xi[i] = sum_j mult[j] j0(k_j r_i) in order.
__global__ void simpson_sum_kernel(const SimpsonJob* jobs, const double* r) {
    const SimpsonJob& J = jobs[blockIdx.y];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= QS_N) return;
    double ri = r[i], s = 0.0;
    for (int j = 0; j <= J.Nk; j++) s = __dadd_rn(s, __dmul_rn(J.mult[j], sj0_ref(J.kgrid[j] * ri)));
    J.out[i] = s;
}

// Band averages of P at 256 midpoints (src/RealModel.cpp:85-96).
__global__ void band_param_kernel(SplineDev P, int nb, const double* lo, const double* hi, double* out) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nb) return;
    const int Nk = 256;
    double kmin = lo[n], kmax = hi[n], sum = 0;
    for (int i = 0; i < Nk; i++) {
        double k = __dadd_rn(kmin, __ddiv_rn(__dmul_rn(i + 0.5, kmax - kmin), (double)Nk));
        sum = __dadd_rn(sum, spline_eval(P, k));
    }
    out[n] = sum / Nk;
}

__global__ void xi_eval_kernel(XiModelDev m, int f, int n, const double* r, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = xi_eval(m, m.desc[f], r[i]);
}

static void free_model(opsec_ctx* ctx) {
    dev_free((void**)&ctx->d_desc);
    dev_free((void**)&ctx->d_tables);
    dev_free((void**)&ctx->d_xgrid);
    dev_free((void**)&ctx->d_ugrid);
    dev_free((void**)&ctx->d_pk);
    ctx->funcs.clear();
    ctx->params.clear();
    ctx->nparams = 0;
    ctx->ntables = 0;
    ctx->model_kind = -1;
    ctx->Q_valid.clear();
}

static int upload_desc(opsec_ctx* ctx) {
    size_t cap = 0;
    int rc = dev_realloc(ctx, (void**)&ctx->d_desc, &cap, sizeof(XiDesc) * ctx->funcs.size());
    if (rc) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_desc, ctx->funcs.data(), sizeof(XiDesc) * ctx->funcs.size(),
                                      hipMemcpyHostToDevice, ctx->stream));
    return OPSEC_OK;
}

}  // namespace opsec

using namespace opsec;

namespace opsec {
__global__ __launch_bounds__(256) void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// FP64 sin + cos throughput: each thread evaluates `iters` (sin x, cos x) pairs
// at arguments spread over [0.1, 0.1 + 0.37 iters] rad (the range of k r in the
// band correlation functions); the sum is stored so nothing is optimised away.
__global__ __launch_bounds__(256) void sincos_probe_kernel(double* out, int iters) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double x = 0.1 + 1e-7 * (double)(i & 1023), acc = 0.0;
    for (int k = 0; k < iters; k++) {
        double sn, cs;
        sincos(x, &sn, &cs);
        acc += sn * cs;
        x += 0.37;
    }
    out[i] = acc;
}
}  // namespace opsec

extern "C" {

int opsec_abi_version(void) { return OPSEC_HIP_ABI_VERSION; }

int opsec_ctx_create(int device, opsec_ctx** out) {
    if (!out) return OPSEC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return OPSEC_ENODEV;
    if (device < 0 || device >= ndev) return OPSEC_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return OPSEC_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return OPSEC_ENODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return OPSEC_ENODEV;
    opsec_ctx* ctx = new opsec_ctx;
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return OPSEC_EHIP;
    }
    // device address of the allocation-free Gram entry point's scratch
    if (fisher_scratch_setup(ctx) != OPSEC_OK) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return OPSEC_EHIP;
    }
    *out = ctx;
    return OPSEC_OK;
}

int opsec_ctx_destroy(opsec_ctx* ctx) {
    if (!ctx) return OPSEC_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    free_model(ctx);
    dev_free((void**)&ctx->d_cellidx);
    dev_free((void**)&ctx->d_nbar);
    dev_free((void**)&ctx->d_Q);
    dev_free((void**)&ctx->d_B);
    dev_free(&ctx->d_jobs);
    free_proj(ctx);
    free_mlp(ctx);
    free_dot(ctx);
    for (void* p : ctx->user_bufs) hipFree(p);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return OPSEC_OK;
}

const char* opsec_last_error(const opsec_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int opsec_ctx_sync(opsec_ctx* ctx) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

void* opsec_ctx_stream(opsec_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int opsec_device_info(opsec_ctx* ctx, int* num_cus, char* arch, int arch_len) {
    if (!ctx) return OPSEC_EINVAL;
    if (num_cus) *num_cus = ctx->num_cus;
    if (arch && arch_len > 0) {
        std::strncpy(arch, ctx->arch.c_str(), arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return OPSEC_OK;
}

int opsec_dev_alloc(opsec_ctx* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return OPSEC_EINVAL;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) return set_error(ctx, OPSEC_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    ctx->user_bufs.push_back(*dptr);
    return OPSEC_OK;
}

int opsec_dev_free(opsec_ctx* ctx, void* dptr) {
    if (!ctx) return OPSEC_EINVAL;
    auto it = std::find(ctx->user_bufs.begin(), ctx->user_bufs.end(), dptr);
    if (it == ctx->user_bufs.end()) return set_error(ctx, OPSEC_EINVAL, "pointer not owned by context");
    ctx->user_bufs.erase(it);
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    OPSEC_HIP_TRY(ctx, hipFree(dptr));
    return OPSEC_OK;
}

// Plain 16-byte-per-lane copy of `bytes` (a multiple of 16) between device
// buffers on the context's stream, timed with HIP events: the reference
// kernel whose FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic
// figures of the other kernels.
int opsec_diag_copy(opsec_ctx* ctx, void* dst, const void* src, size_t bytes, float* ms) {
    if (!ctx || !dst || !src || bytes % 16) return OPSEC_EINVAL;
    hipEvent_t e0, e1;
    OPSEC_HIP_TRY(ctx, hipEventCreate(&e0));
    OPSEC_HIP_TRY(ctx, hipEventCreate(&e1));
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 32);
    hipEventRecord(e0, ctx->stream);
    hipLaunchKernelGGL(opsec::copy16_kernel, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4*)src, (uint4*)dst, n);
    hipEventRecord(e1, ctx->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess && ms) e = hipEventElapsedTime(ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e != hipSuccess) return set_error(ctx, OPSEC_EHIP, "copy: %s", hipGetErrorString(e));
    return OPSEC_OK;
}

// Measured ceiling of FP64 (sin, cos) pairs per second on the whole device.
int opsec_diag_sincos_rate(opsec_ctx* ctx, double* pairs_per_s) {
    if (!ctx || !pairs_per_s) return OPSEC_EINVAL;
    const unsigned blocks = (unsigned)ctx->num_cus * 16;
    const int iters = 2000;
    double* d_out = nullptr;
    OPSEC_HIP_TRY(ctx, hipMalloc(&d_out, (size_t)blocks * 256 * 8));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipLaunchKernelGGL(opsec::sincos_probe_kernel, dim3(blocks), dim3(256), 0, ctx->stream, d_out, 8);   // warm-up
    hipEventRecord(e0, ctx->stream);
    hipLaunchKernelGGL(opsec::sincos_probe_kernel, dim3(blocks), dim3(256), 0, ctx->stream, d_out, iters);
    hipEventRecord(e1, ctx->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(d_out);
    if (e != hipSuccess) return set_error(ctx, OPSEC_EHIP, "sincos probe: %s", hipGetErrorString(e));
    *pairs_per_s = (double)blocks * 256 * iters / (ms * 1e-3);
    return OPSEC_OK;
}

int opsec_h2d(opsec_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

int opsec_d2h(opsec_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

int opsec_model_real(opsec_ctx* ctx, const double* pk_k, const double* pk_P, int npk, int nbands,
                     const double* band_lo, const double* band_hi, int xi_mode) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, pk_k && pk_P && npk >= 2, "pkfile needs at least two samples");
    OPSEC_CHECK_ARG(ctx, nbands >= 1 && band_lo && band_hi, "need at least one band");
    OPSEC_CHECK_ARG(ctx, xi_mode == OPSEC_XI_TABLE || xi_mode == OPSEC_XI_CLOSED, "bad xi mode %d", xi_mode);
    for (int i = 1; i < npk; i++)
        OPSEC_CHECK_ARG(ctx, pk_k[i] > pk_k[i - 1], "pkfile k values must increase (row %d)", i);
    OPSEC_CHECK_ARG(ctx, pk_k[0] <= 0.0 && pk_k[npk - 1] > 10.0,
                    "pkfile must span [0,10] (the prior integral runs over k in [0,10])");
    for (int n = 0; n < nbands; n++)
        OPSEC_CHECK_ARG(ctx, band_hi[n] > band_lo[n] && band_lo[n] >= 0, "band %d has bad edges", n);
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    free_model(ctx);

    ctx->model_kind = 0;
    ctx->nparams = nbands;
    ctx->band_lo.assign(band_lo, band_lo + nbands);
    ctx->band_hi.assign(band_hi, band_hi + nbands);
    ctx->funcs.resize(nbands + 1);
    ctx->funcs[0] = XiDesc{XI_TABLE, 0, 0.0, 0.0};
    int ntab = 1;
    for (int n = 0; n < nbands; n++) {
        if (xi_mode == OPSEC_XI_TABLE)
            ctx->funcs[n + 1] = XiDesc{XI_TABLE, ntab++, band_lo[n], band_hi[n]};
        else
            ctx->funcs[n + 1] = XiDesc{XI_BAND, -1, band_lo[n], band_hi[n]};
    }
    ctx->ntables = ntab;

    // QSpline nodes (QSpline::MakeArray, src/QSpline.cpp:6-17) and log nodes.
    ctx->h_xgrid.resize(QS_N);
    ctx->h_ugrid.resize(QS_N2);
    for (int i = 0; i < QS_N1; i++) ctx->h_xgrid[i] = i * QS_X1 / QS_N1;
    for (int j = 0; j < QS_N2; j++) ctx->h_xgrid[QS_N1 + j] = QS_X1 * std::exp(j * std::log(QS_X2 / QS_X1) / (QS_N2 - 1));
    for (int j = 0; j < QS_N2; j++) ctx->h_ugrid[j] = std::log(ctx->h_xgrid[QS_N1 + j]);

    // pk spline coefficients
    std::vector<double> pk(5 * (size_t)npk);
    std::copy(pk_k, pk_k + npk, pk.begin());
    std::copy(pk_P, pk_P + npk, pk.begin() + npk);
    fmm_spline(npk, pk_k, pk_P, &pk[2 * npk], &pk[3 * npk], &pk[4 * npk]);
    ctx->npk = npk;

    size_t cap;
    int rc;
    cap = 0; if ((rc = dev_realloc(ctx, (void**)&ctx->d_pk, &cap, pk.size() * 8))) return rc;
    cap = 0; if ((rc = dev_realloc(ctx, (void**)&ctx->d_tables, &cap, (size_t)ntab * QS_N * 8))) return rc;
    cap = 0; if ((rc = dev_realloc(ctx, (void**)&ctx->d_xgrid, &cap, QS_N * 8))) return rc;
    cap = 0; if ((rc = dev_realloc(ctx, (void**)&ctx->d_ugrid, &cap, QS_N2 * 8))) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_pk, pk.data(), pk.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_xgrid, ctx->h_xgrid.data(), QS_N * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_ugrid, ctx->h_ugrid.data(), QS_N2 * 8, hipMemcpyHostToDevice, ctx->stream));

    SplineDev pkdev{};
    pkdev.kind = 0;
    pkdev.n = npk;
    pkdev.X = ctx->d_pk;
    pkdev.Y = ctx->d_pk + npk;
    pkdev.b = ctx->d_pk + 2 * npk;
    pkdev.c = ctx->d_pk + 3 * npk;
    pkdev.d = ctx->d_pk + 4 * npk;

    // Simpson jobs: prior + tabulated bands
    std::vector<SimpsonJob> jobs;
    size_t work = 32769;
    for (int n = 0; n < nbands; n++)
        if (xi_mode == OPSEC_XI_TABLE) work += 8193;
    double* d_work = nullptr;
    cap = 0;
    if ((rc = dev_realloc(ctx, (void**)&d_work, &cap, 2 * work * 8))) return rc;
    size_t off = 0;
    {
        SimpsonJob J{};
        J.P = pkdev;
        J.Nk = 32768;
        J.kmin = 0.0;
        J.kmax = 10.0;
        J.kdamp = 1.0;
        J.mult = d_work + off;
        J.kgrid = d_work + work + off;
        J.out = ctx->d_tables;
        off += J.Nk + 1;
        jobs.push_back(J);
    }
    if (xi_mode == OPSEC_XI_TABLE)
        for (int n = 0; n < nbands; n++) {
            SimpsonJob J{};
            J.P.kind = 1;
            J.P.value = 1.0;
            J.P.xmin = band_lo[n];
            J.P.xmax = band_hi[n];
            J.Nk = 8192;
            J.kmin = band_lo[n];
            J.kmax = band_hi[n];
            J.kdamp = 1.0;
            J.mult = d_work + off;
            J.kgrid = d_work + work + off;
            J.out = ctx->d_tables + (size_t)(n + 1) * QS_N;
            off += J.Nk + 1;
            jobs.push_back(J);
        }
    SimpsonJob* d_jobs = nullptr;
    cap = 0;
    if ((rc = dev_realloc(ctx, (void**)&d_jobs, &cap, jobs.size() * sizeof(SimpsonJob)))) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(SimpsonJob), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(simpson_mult_kernel, dim3(64, (unsigned)jobs.size()), dim3(256), 0, ctx->stream, d_jobs);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    hipLaunchKernelGGL(simpson_sum_kernel, dim3((QS_N + 63) / 64, (unsigned)jobs.size()), dim3(64), 0, ctx->stream,
                       d_jobs, ctx->d_xgrid);
    OPSEC_HIP_TRY(ctx, hipGetLastError());

    // prior band powers
    double *d_lo = nullptr, *d_par = nullptr;
    cap = 0;
    if ((rc = dev_realloc(ctx, (void**)&d_lo, &cap, 3 * (size_t)nbands * 8))) return rc;
    d_par = d_lo + 2 * nbands;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_lo, band_lo, nbands * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d_lo + nbands, band_hi, nbands * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(band_param_kernel, dim3((nbands + 63) / 64), dim3(64), 0, ctx->stream, pkdev, nbands, d_lo,
                       d_lo + nbands, d_par);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    ctx->params.resize(nbands);
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->params.data(), d_par, nbands * 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(d_lo);
    hipFree(d_jobs);
    hipFree(d_work);
    rc = upload_desc(ctx);
    if (rc) return rc;
    ctx->Q_valid.assign(ctx->funcs.size(), 0);
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

int opsec_model_test(opsec_ctx* ctx, int form, double A, double b) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, form == 0 || form == 1, "TestModel form must be 0 (exponential) or 1 (gaussian)");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    free_model(ctx);
    ctx->model_kind = 1;
    ctx->nparams = 2;
    ctx->params = {A, b};
    ctx->funcs.resize(3);
    // src/TestModel.cpp:64-85
    if (form == 0) {
        ctx->funcs[0] = XiDesc{XI_TEST_EXP, -1, A, b};
        ctx->funcs[1] = XiDesc{XI_TEST_EXP, -1, 1.0, b};
        ctx->funcs[2] = XiDesc{XI_TEST_EXP, -1, -A * b, b};
    } else {
        ctx->funcs[0] = XiDesc{XI_TEST_GAUSS, -1, A, b};
        ctx->funcs[1] = XiDesc{XI_TEST_GAUSS, -1, 1.0, b};
        ctx->funcs[2] = XiDesc{XI_TEST_DGAUSS, -1, -A * b, b};
    }
    ctx->h_xgrid.clear();
    ctx->h_ugrid.clear();
    int rc = upload_desc(ctx);
    if (rc) return rc;
    ctx->Q_valid.assign(ctx->funcs.size(), 0);
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

int opsec_model_tables(opsec_ctx* ctx, int nfuncs, const double* xi768, const double* params) {
    if (!ctx) return OPSEC_EINVAL;
    OPSEC_CHECK_ARG(ctx, nfuncs >= 1 && xi768, "need at least one table");
    for (size_t i = 0; i < (size_t)nfuncs * QS_N; i++)
        OPSEC_CHECK_ARG(ctx, std::isfinite(xi768[i]), "table %d node %d is not finite", (int)(i / QS_N),
                        (int)(i % QS_N));
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    free_model(ctx);
    ctx->model_kind = 2;
    ctx->nparams = nfuncs - 1;
    ctx->params.assign(params ? params : xi768, params ? params + (nfuncs - 1) : xi768);
    if (!params) ctx->params.assign(nfuncs - 1, 0.0);
    ctx->funcs.resize(nfuncs);
    for (int f = 0; f < nfuncs; f++) ctx->funcs[f] = XiDesc{XI_TABLE, f, 0.0, 0.0};
    ctx->ntables = nfuncs;
    // the QSpline nodes of a RealModel table (QSpline::MakeArray, src/QSpline.cpp:6-17)
    ctx->h_xgrid.resize(QS_N);
    ctx->h_ugrid.resize(QS_N2);
    for (int i = 0; i < QS_N1; i++) ctx->h_xgrid[i] = i * QS_X1 / QS_N1;
    for (int j = 0; j < QS_N2; j++) ctx->h_xgrid[QS_N1 + j] = QS_X1 * std::exp(j * std::log(QS_X2 / QS_X1) / (QS_N2 - 1));
    for (int j = 0; j < QS_N2; j++) ctx->h_ugrid[j] = std::log(ctx->h_xgrid[QS_N1 + j]);
    size_t cap;
    int rc;
    cap = 0; if ((rc = dev_realloc(ctx, (void**)&ctx->d_tables, &cap, (size_t)nfuncs * QS_N * 8))) return rc;
    cap = 0; if ((rc = dev_realloc(ctx, (void**)&ctx->d_xgrid, &cap, QS_N * 8))) return rc;
    cap = 0; if ((rc = dev_realloc(ctx, (void**)&ctx->d_ugrid, &cap, QS_N2 * 8))) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_tables, xi768, (size_t)nfuncs * QS_N * 8, hipMemcpyHostToDevice,
                                      ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_xgrid, ctx->h_xgrid.data(), QS_N * 8, hipMemcpyHostToDevice, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_ugrid, ctx->h_ugrid.data(), QS_N2 * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = upload_desc(ctx))) return rc;
    ctx->Q_valid.assign(ctx->funcs.size(), 0);
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

int opsec_model_nparams(opsec_ctx* ctx, int* nparams) {
    if (!ctx || !nparams) return OPSEC_EINVAL;
    if (ctx->model_kind < 0) return set_error(ctx, OPSEC_ESTATE, "no model loaded");
    *nparams = ctx->nparams;
    return OPSEC_OK;
}

int opsec_model_params(opsec_ctx* ctx, double* params) {
    if (!ctx || !params) return OPSEC_EINVAL;
    if (ctx->model_kind < 0) return set_error(ctx, OPSEC_ESTATE, "no model loaded");
    std::copy(ctx->params.begin(), ctx->params.end(), params);
    return OPSEC_OK;
}

int opsec_model_xi(opsec_ctx* ctx, int f, int n, const double* r, double* xi) {
    if (!ctx) return OPSEC_EINVAL;
    if (ctx->model_kind < 0) return set_error(ctx, OPSEC_ESTATE, "no model loaded");
    OPSEC_CHECK_ARG(ctx, f >= 0 && f < (int)ctx->funcs.size(), "function index %d out of range", f);
    OPSEC_CHECK_ARG(ctx, n >= 0 && (n == 0 || (r && xi)), "bad arrays");
    if (n == 0) return OPSEC_OK;
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* d = nullptr;
    size_t cap = 0;
    int rc = dev_realloc(ctx, (void**)&d, &cap, 2 * (size_t)n * 8);
    if (rc) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d, r, n * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(xi_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, model_view(ctx), f, n, d, d + n);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(xi, d + n, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(d);
    return OPSEC_OK;
}

int opsec_model_table(opsec_ctx* ctx, int f, double* r768, double* xi768) {
    if (!ctx) return OPSEC_EINVAL;
    if (ctx->model_kind != 0 && ctx->model_kind != 2) return set_error(ctx, OPSEC_ESTATE, "no tabulated model loaded");
    OPSEC_CHECK_ARG(ctx, f >= 0 && f < (int)ctx->funcs.size(), "function index %d out of range", f);
    const XiDesc& d = ctx->funcs[f];
    if (d.kind != XI_TABLE) return set_error(ctx, OPSEC_EINVAL, "function %d is not tabulated", f);
    if (r768) std::copy(ctx->h_xgrid.begin(), ctx->h_xgrid.end(), r768);
    if (xi768) {
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(xi768, ctx->d_tables + (size_t)d.table * QS_N, QS_N * 8,
                                          hipMemcpyDeviceToHost, ctx->stream));
        OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    return OPSEC_OK;
}

}  // extern "C"
