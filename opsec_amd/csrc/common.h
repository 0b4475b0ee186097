// Internal declarations shared by the HIP translation units of libopsec_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "opsec_hip.h"
#include "opsec_hip_diag.h"

namespace opsec {

// ---- xi function descriptors ---------------------------------------------
enum XiKind : int {
    XI_TABLE = 0,        // QSpline table (reference tables)
    XI_BAND = 1,         // closed-form top-hat band
    XI_TEST_EXP = 2,     // A exp(-b r)
    XI_TEST_GAUSS = 3,   // A exp(-b r^2)
    XI_TEST_DGAUSS = 4   // A r exp(-b r^2)
};

struct XiDesc {
    int kind;
    int table;           // row of the table array (XI_TABLE)
    double a, b;         // band [klo, khi] or test (A, b)
};

// QSpline geometry of the reference tables (src/RealModel.cpp:98-127):
// 512 linear nodes over [0,200) then 256 log nodes over [200,20000].
constexpr int QS_N1 = 512;
constexpr int QS_N2 = 256;
constexpr int QS_N = QS_N1 + QS_N2;
constexpr double QS_X1 = 200.0;
constexpr double QS_X2 = 20000.0;

// Device view of the model: descriptors + tables.
struct XiModelDev {
    const XiDesc* desc;     // nf descriptors
    const double* y;        // ntables x QS_N values
    const double* xgrid;    // QS_N r nodes (first QS_N1+1 are the linear part)
    const double* ugrid;    // QS_N2 log-r nodes
    double x2, uhi;         // xgrid[QS_N-1], ugrid[QS_N2-1]
};

// Survey geometry for pair integrals on the device.
struct SurveyDev {
    int coordsys;
    int n1, n2, n3;
    double min1, max1, min2, max2, min3, max3;
    double L;
};

// Pair-table key of two cells from their grid indices.  Box surveys:
// (|dx| Ny + |dy|) Nz + |dz| (src/sig.cpp:109-112).  Spherical surveys: the
// arc-pair cache of src/sig.cpp:150-190, (h_a narcs + h_b) Nphi + |dphi_a - dphi_b|
// with arcs h = d_r Nmu + d_mu.  narcs = 0 selects the box key.
struct PairKey {
    int n2, n3, narcs;
    __host__ __device__ __forceinline__ int operator()(int a1, int a2, int a3, int b1, int b2, int b3) const {
        const int dz = a3 > b3 ? a3 - b3 : b3 - a3;
        if (narcs) return ((a1 * n2 + a2) * narcs + b1 * n2 + b2) * n3 + dz;
        return ((a1 > b1 ? a1 - b1 : b1 - a1) * n2 + (a2 > b2 ? a2 - b2 : b2 - a2)) * n3 + dz;
    }
};
inline PairKey pair_key_of(const SurveyDev& s) {
    return PairKey{s.n2, s.n3, s.coordsys == OPSEC_COORD_SPHERICAL ? s.n1 * s.n2 : 0};
}
// Number of pair-table keys: n1 n2 n3 (box), (n1 n2)^2 n3 (spherical).
inline long long pair_nkeys_of(const SurveyDev& s) {
    const long long g = (long long)s.n1 * s.n2;
    return s.coordsys == OPSEC_COORD_SPHERICAL ? g * g * s.n3 : g * s.n3;
}

// ---- context -------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct MlpState;  // surrogate network (mlp.hip)
struct ProjState; // projection results (project.hip)
struct DotState;  // galaxy counts and pixel values (dot.hip)

}  // namespace opsec

struct opsec_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cus = 0;
    std::string arch;

    // model
    int model_kind = -1;     // 0 real, 1 test, 2 uploaded tables
    int nparams = 0;
    std::vector<opsec::XiDesc> funcs;     // nparams + 1
    std::vector<double> params;           // prior p_n
    opsec::XiDesc* d_desc = nullptr;
    double* d_tables = nullptr;           // ntables x QS_N
    double* d_xgrid = nullptr;
    double* d_ugrid = nullptr;
    std::vector<double> h_xgrid, h_ugrid;
    int ntables = 0;
    double pk_kmin = 0, pk_kmax = 0;      // pkfile range (for band params)
    double* d_pk = nullptr;               // pk spline: X,Y,b,c,d (5 x npk)
    int npk = 0;
    std::vector<double> band_lo, band_hi;

    // survey (device-resident)
    opsec::SurveyDev sv{};
    int ncells = 0;
    int nkeys = 0;
    int* d_cellidx = nullptr;             // ncells x 3 grid indices
    double* d_nbar = nullptr;             // ncells Nbar
    std::vector<opsec_cell> h_cells;
    double* d_Q = nullptr;                // (nparams+1) x nkeys pair tables
    std::vector<char> Q_valid;
    int64_t st_integrals = 0, st_nodes = 0, st_surrogate = 0, st_spot = 0, st_spot_fail = 0;

    // modes
    int nmodes = 0;
    int projection = OPSEC_PROJ_DENSE;    // method of the split-BF16x3 projection
    double* d_B = nullptr;                // ncells x nmodes, column-major
    opsec::ProjState* proj = nullptr;
    opsec::MlpState* mlp = nullptr;
    opsec::DotState* dot = nullptr;

    std::vector<void*> user_bufs;
    void* d_jobs = nullptr;               // quadrature job buffer (reused across calls)
    size_t cap_jobs = 0;
};

namespace opsec {

int set_error(opsec_ctx* ctx, int code, const char* fmt, ...);

#define OPSEC_HIP_TRY(ctx, call)                                                        \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::opsec::set_error((ctx), OPSEC_EHIP, "%s failed: %s (%s:%d)", #call, \
                                      hipGetErrorString(e_), __FILE__, __LINE__);       \
    } while (0)

#define OPSEC_CHECK_ARG(ctx, cond, ...)                                         \
    do {                                                                        \
        if (!(cond)) return ::opsec::set_error((ctx), OPSEC_EINVAL, __VA_ARGS__); \
    } while (0)

// Allocate (or re-allocate to at least `bytes`) a device buffer owned by ctx.
int dev_realloc(opsec_ctx* ctx, void** p, size_t* cap, size_t bytes);
void dev_free(void** p);

// Device buffers owned by one call: freed when the scope ends, error returns
// included (hipFree waits for the work that uses them).
struct DevScratch {
    std::vector<void*> bufs;
    DevScratch() = default;
    DevScratch(const DevScratch&) = delete;
    DevScratch& operator=(const DevScratch&) = delete;
    ~DevScratch() {
        for (void* q : bufs) (void)hipFree(q);
    }
    template <typename T>
    hipError_t alloc(T** out, size_t bytes) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        *out = e == hipSuccess ? (T*)q : nullptr;
        if (e == hipSuccess) bufs.push_back(q);
        return e;
    }
};

// Model view for kernels.
XiModelDev model_view(const opsec_ctx* ctx);

// Gauss rules (host): n-point Gauss-Legendre on [-1,1]; n-point Gauss rule for
// weight (1-t) on [0,1].
void gauss_legendre(int n, double* x, double* w);
void gauss_jacobi10(int n, double* t, double* w);

// Survey checks / conversions.
int survey_from_abi(opsec_ctx* ctx, const opsec_survey* in, SurveyDev* out);

// Pair-table builder shared by sig_block and pair_tables: fills d_Q rows for
// the functions in funcs.
int build_pair_tables(opsec_ctx* ctx, int nf, const int* funcs, const opsec_quad* q);

// Bounds of the two cells of a pair integral (cell 1 then cell 2).
struct PairGeo {
    double lo1[3], hi1[3], lo2[3], hi2[3];
};

// Quadrature of the given keys x functions into dQ[f * nkeys + key].
int run_quadrature_public(opsec_ctx* ctx, int nf, const int* funcs, const std::vector<int>& keys,
                          const std::vector<PairGeo>& geo, const std::vector<int>& nearflag, const opsec_quad* q,
                          double* dQ, size_t nkeys);
// Geometry of key (kx, ky, kz) of the device survey's grid: the grid cell at
// the origin and the cell at offset k, as Survey::CreateEmptyCell lays them out.
PairGeo grid_pair_geo(const SurveyDev& sv, int kx, int ky, int kz);

// Per-unit state teardown (project.hip, mlp.hip, dot.hip) and the mode
// invalidation a new cell upload triggers (project.hip).
void free_proj(opsec_ctx* ctx);
void free_mlp(opsec_ctx* ctx);
void free_dot(opsec_ctx* ctx);
void invalidate_modes(opsec_ctx* ctx);
// Device address of the Gram scratch array on ctx's device (fisher.hip), once
// per device, with that device current.
int fisher_scratch_setup(opsec_ctx* ctx);

// Rows of the packed upper triangle: column j starts at j(j+1)/2.
inline int64_t packed_size(int64_t m) { return m * (m + 1) / 2; }

// Blocking copy ordered on the context's stream.  (hipMemcpy runs on the null
// stream, which does not order against the context's non-blocking stream, so
// a kernel queued on the context's stream could overtake it.)
inline hipError_t stream_copy(opsec_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e;
}

}  // namespace opsec
