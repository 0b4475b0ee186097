// Pixel values of a galaxy catalogue: weighted galaxy counts per cell and
// y = B^T x with x_a = (N_a - Nbar_a)/sqrt(Nbar_a) -- the dot stage
// (src/dot.cpp:105-157).
//
// Counting.  The reference adds each galaxy's weight into counts[a] in
// catalogue order (src/dot.cpp:124-128), so the per-cell sums are left folds
// whose rounding depends on that order.  Here every galaxy gets its bin key
// (cell number, or Ncells for "outside every cell"), a stable radix sort
// groups the keys while keeping catalogue order inside each group, and one
// thread per bin folds its group in that order: the counts are bit-identical
// to the sequential loop.  Catalogues are taken in chunks of CHUNK galaxies;
// a chunk's fold starts from the previous chunk's count, so chunking does not
// change a bit either.
//
// Two defects of the reference loop are not reproduced: a galaxy in grid cell
// e = 0 is counted as outside (`e > 0`, src/dot.cpp:127), and a galaxy in a
// grid position with no cell indexes counts[-1].  Here e >= 0 is inside, and
// grid positions without a cell count as outside.
#include <rocprim/device/device_radix_sort.hpp>

#include "common.h"

namespace opsec {

struct DotState {
    int* d_map = nullptr;          // grid index -> cell number, -1 for none
    size_t cap_map = 0;
    float4* d_gal = nullptr;       // one chunk of galaxies
    size_t cap_gal = 0;
    uint32_t* d_key[2] = {nullptr, nullptr};
    size_t cap_key[2] = {0, 0};
    float* d_w[2] = {nullptr, nullptr};
    size_t cap_w[2] = {0, 0};
    void* d_tmp = nullptr;         // radix-sort temporary storage
    size_t cap_tmp = 0;
    double* d_counts = nullptr;    // Ncells + 1
    size_t cap_counts = 0;
    double* d_x = nullptr;         // Ncells
    size_t cap_x = 0;
    double* d_y = nullptr;         // Nmodes
    size_t cap_y = 0;
};

namespace {

constexpr int64_t CHUNK = int64_t(1) << 26;

struct GridGeom {
    double lo1, hi1, lo2, hi2, lo3, hi3;
    int n1, n2, n3;
};

// floor(n (x - lo)/(hi - lo)) in the operation order of Survey::GetGridIndex
// (src/Survey.cpp:93-101); -1 outside [0, n).  The comparison is made on the
// floored double, which also sends NaN and out-of-int-range values outside.
__device__ __forceinline__ int grid_coord(float x, double lo, double hi, int n) {
    double t = floor((double)n * ((double)x - lo) / (hi - lo));
    return (t >= 0.0 && t < (double)n) ? (int)t : -1;
}

__global__ void fill_int_kernel(int* p, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// cellmap[e] = a for every cell; when two cells share a grid position the
// larger cell number wins, as the reference's in-order loop leaves it
// (src/dot.cpp:106-115).
__global__ void cellmap_kernel(const int* __restrict__ cellidx, int ncells, int n2, int n3, int* map) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ncells) return;
    int64_t e = ((int64_t)cellidx[3 * a] * n2 + cellidx[3 * a + 1]) * n3 + cellidx[3 * a + 2];
    atomicMax(map + e, a);
}

__global__ void galaxy_keys_kernel(const float4* __restrict__ gal, int64_t n, GridGeom g,
                                   const int* __restrict__ map, uint32_t nbins_out,
                                   uint32_t* __restrict__ key, float* __restrict__ w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = gal[i];
    int d1 = grid_coord(p.x, g.lo1, g.hi1, g.n1);
    int d2 = grid_coord(p.y, g.lo2, g.hi2, g.n2);
    int d3 = grid_coord(p.z, g.lo3, g.hi3, g.n3);
    uint32_t k = nbins_out;
    if (d1 >= 0 && d2 >= 0 && d3 >= 0) {
        int a = map[((int64_t)d1 * g.n2 + d2) * g.n3 + d3];
        if (a >= 0) k = (uint32_t)a;
    }
    key[i] = k;
    w[i] = p.w;
}

__device__ __forceinline__ int64_t lower_bound_u32(const uint32_t* __restrict__ k, int64_t n, uint32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (k[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// counts[b] += w over bin b's galaxies, one by one in catalogue order.
__global__ void fold_bins_kernel(const uint32_t* __restrict__ key, const float* __restrict__ w, int64_t n,
                                 int nbins, double* __restrict__ counts) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    int64_t lo = lower_bound_u32(key, n, (uint32_t)b);
    int64_t hi = lower_bound_u32(key, n, (uint32_t)b + 1);
    if (lo == hi) return;
    double s = counts[b];
    int64_t i = lo;
    for (; i + 8 <= hi; i += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = w[i + j];
#pragma unroll
        for (int j = 0; j < 8; j++) s += (double)v[j];
    }
    for (; i < hi; i++) s += (double)w[i];
    counts[b] = s;
}

// x_a = (N_a - Nbar_a)/sqrt(Nbar_a)  (src/dot.cpp:130-134)
__global__ void density_kernel(const double* __restrict__ counts, const double* __restrict__ nbar, int n,
                               double* __restrict__ x) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) x[a] = (counts[a] - nbar[a]) / sqrt(nbar[a]);
}

// y_i = sum_a B_ia x_a: one 256-thread workgroup per mode.  Each thread folds
// a stride-256 subsequence, then a fixed LDS tree combines the 256 partials,
// so y does not depend on timing.
constexpr int GEMV_THREADS = 256;
__global__ __launch_bounds__(GEMV_THREADS) void modes_gemv_kernel(const double* __restrict__ B, int64_t ncells,
                                                                  const double* __restrict__ x,
                                                                  double* __restrict__ y) {
    __shared__ double part[GEMV_THREADS];
    const double* b = B + (int64_t)blockIdx.x * ncells;
    int t = threadIdx.x;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int64_t a = t;
    for (; a + 3 * GEMV_THREADS < ncells; a += 4 * GEMV_THREADS) {
        s0 += b[a] * x[a];
        s1 += b[a + GEMV_THREADS] * x[a + GEMV_THREADS];
        s2 += b[a + 2 * GEMV_THREADS] * x[a + 2 * GEMV_THREADS];
        s3 += b[a + 3 * GEMV_THREADS] * x[a + 3 * GEMV_THREADS];
    }
    for (; a < ncells; a += GEMV_THREADS) s0 += b[a] * x[a];
    part[t] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int h = GEMV_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) part[t] += part[t + h];
        __syncthreads();
    }
    if (t == 0) y[blockIdx.x] = part[0];
}

unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

int ensure_dot(opsec_ctx* ctx) {
    if (!ctx->dot) ctx->dot = new DotState;
    return OPSEC_OK;
}

}  // namespace

void free_dot(opsec_ctx* ctx) {
    DotState* d = ctx->dot;
    if (!d) return;
    dev_free((void**)&d->d_map);
    dev_free((void**)&d->d_gal);
    for (int i = 0; i < 2; i++) {
        dev_free((void**)&d->d_key[i]);
        dev_free((void**)&d->d_w[i]);
    }
    dev_free(&d->d_tmp);
    dev_free((void**)&d->d_counts);
    dev_free((void**)&d->d_x);
    dev_free((void**)&d->d_y);
    delete d;
    ctx->dot = nullptr;
}

}  // namespace opsec

using namespace opsec;

extern "C" {


int opsec_count_galaxies(opsec_ctx* ctx, const opsec_galaxy* gals, int64_t ngals, int accumulate, double* counts) {
    if (!ctx) return OPSEC_EINVAL;
    if (ctx->ncells == 0) return set_error(ctx, OPSEC_ESTATE, "no survey uploaded");
    OPSEC_CHECK_ARG(ctx, counts, "no counts buffer");
    OPSEC_CHECK_ARG(ctx, ngals >= 0 && (gals || ngals == 0), "bad galaxy array");
    static_assert(sizeof(opsec_galaxy) == sizeof(float4), "galaxy record is four floats");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_dot(ctx);
    if (rc) return rc;
    DotState* d = ctx->dot;
    const SurveyDev& sv = ctx->sv;
    const int ncells = ctx->ncells, nbins = ncells + 1;
    const int64_t ngrid = (int64_t)sv.n1 * sv.n2 * sv.n3;

    if ((rc = dev_realloc(ctx, (void**)&d->d_map, &d->cap_map, (size_t)ngrid * 4))) return rc;
    if ((rc = dev_realloc(ctx, (void**)&d->d_counts, &d->cap_counts, (size_t)nbins * 8))) return rc;
    hipLaunchKernelGGL(fill_int_kernel, dim3(blocks_for(ngrid, 256)), dim3(256), 0, ctx->stream, d->d_map, ngrid, -1);
    hipLaunchKernelGGL(cellmap_kernel, dim3(blocks_for(ncells, 256)), dim3(256), 0, ctx->stream, ctx->d_cellidx,
                       ncells, sv.n2, sv.n3, d->d_map);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    if (accumulate)
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d->d_counts, counts, (size_t)nbins * 8, hipMemcpyHostToDevice, ctx->stream));
    else
        OPSEC_HIP_TRY(ctx, hipMemsetAsync(d->d_counts, 0, (size_t)nbins * 8, ctx->stream));

    GridGeom g{sv.min1, sv.max1, sv.min2, sv.max2, sv.min3, sv.max3, sv.n1, sv.n2, sv.n3};
    unsigned end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= (uint64_t)ncells) end_bit++;

    const int64_t chunk_max = ngals < CHUNK ? ngals : CHUNK;
    if (chunk_max > 0) {
        if ((rc = dev_realloc(ctx, (void**)&d->d_gal, &d->cap_gal, (size_t)chunk_max * 16))) return rc;
        for (int i = 0; i < 2; i++) {
            if ((rc = dev_realloc(ctx, (void**)&d->d_key[i], &d->cap_key[i], (size_t)chunk_max * 4))) return rc;
            if ((rc = dev_realloc(ctx, (void**)&d->d_w[i], &d->cap_w[i], (size_t)chunk_max * 4))) return rc;
        }
        size_t tmp = 0;
        OPSEC_HIP_TRY(ctx, rocprim::radix_sort_pairs(nullptr, tmp, d->d_key[0], d->d_key[1], d->d_w[0], d->d_w[1],
                                                     (size_t)chunk_max, 0u, end_bit, ctx->stream));
        if ((rc = dev_realloc(ctx, &d->d_tmp, &d->cap_tmp, tmp))) return rc;
    }
    for (int64_t off = 0; off < ngals; off += CHUNK) {
        const int64_t n = (ngals - off) < CHUNK ? (ngals - off) : CHUNK;
        OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d->d_gal, gals + off, (size_t)n * 16, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(galaxy_keys_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, d->d_gal, n, g,
                           d->d_map, (uint32_t)ncells, d->d_key[0], d->d_w[0]);
        OPSEC_HIP_TRY(ctx, hipGetLastError());
        size_t tmp = d->cap_tmp;
        OPSEC_HIP_TRY(ctx, rocprim::radix_sort_pairs(d->d_tmp, tmp, d->d_key[0], d->d_key[1], d->d_w[0], d->d_w[1],
                                                     (size_t)n, 0u, end_bit, ctx->stream));
        hipLaunchKernelGGL(fold_bins_kernel, dim3(blocks_for(nbins, 256)), dim3(256), 0, ctx->stream, d->d_key[1],
                           d->d_w[1], n, nbins, d->d_counts);
        OPSEC_HIP_TRY(ctx, hipGetLastError());
    }
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(counts, d->d_counts, (size_t)nbins * 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

int opsec_pixel_values(opsec_ctx* ctx, const double* counts, double* x, double* y) {
    if (!ctx) return OPSEC_EINVAL;
    if (ctx->ncells == 0) return set_error(ctx, OPSEC_ESTATE, "no survey uploaded");
    if (ctx->nmodes == 0 || !ctx->d_B) return set_error(ctx, OPSEC_ESTATE, "no modes uploaded");
    OPSEC_CHECK_ARG(ctx, counts && y, "null counts or pixel buffer");
    OPSEC_HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_dot(ctx);
    if (rc) return rc;
    DotState* d = ctx->dot;
    const int N = ctx->ncells, M = ctx->nmodes;
    if ((rc = dev_realloc(ctx, (void**)&d->d_counts, &d->cap_counts, (size_t)(N + 1) * 8))) return rc;
    if ((rc = dev_realloc(ctx, (void**)&d->d_x, &d->cap_x, (size_t)N * 8))) return rc;
    if ((rc = dev_realloc(ctx, (void**)&d->d_y, &d->cap_y, (size_t)M * 8))) return rc;
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(d->d_counts, counts, (size_t)N * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(density_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, d->d_counts, ctx->d_nbar,
                       N, d->d_x);
    hipLaunchKernelGGL(modes_gemv_kernel, dim3(M), dim3(GEMV_THREADS), 0, ctx->stream, ctx->d_B, (int64_t)N, d->d_x,
                       d->d_y);
    OPSEC_HIP_TRY(ctx, hipGetLastError());
    if (x) OPSEC_HIP_TRY(ctx, hipMemcpyAsync(x, d->d_x, (size_t)N * 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipMemcpyAsync(y, d->d_y, (size_t)M * 8, hipMemcpyDeviceToHost, ctx->stream));
    OPSEC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return OPSEC_OK;
}

}  // extern "C"
