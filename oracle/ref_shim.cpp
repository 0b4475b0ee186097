/* ref_shim.cpp -- C entry points over the reference's own serial sources.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
 * reference's pair-quadrature slice (src/sig.cpp, src/cubature.c,
 * src/Model.cpp, src/RealModel.cpp, src/TestModel.cpp, src/QSpline.cpp,
 * src/Spline.cpp, src/BoxSurvey.cpp, src/Survey.cpp, src/SeparationFunc.cpp,
 * src/SelectionFunc.cpp, src/XiFunc.cpp, src/Cell.cpp, src/cfg.c, src/abn.c,
 * src/opsec.c) by oracle/Makefile into oracle/_ref/libopsec_ref.so.  Only the
 * tests and the CPU-baseline leg of bench.py load it; nothing in the product
 * links it.
 *
 * The reference code is called exactly as its own programs call it:
 *   - models are built the way InitializeModel does (src/Model.cpp:26-51),
 *     from a "<ModelName>." sub-config;
 *   - the box survey is built from a "survey." sub-config (src/Survey.cpp:140-165,
 *     src/BoxSurvey.cpp:85-91);
 *   - matrix blocks come from ComputeSignalMatrixBlock (src/sig.cpp:50-70).
 */

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "BoxSurvey.h"
#include "Cell.h"
#include "Model.h"
#include "RealModel.h"
#include "SeparationFunc.h"
#include "Spline.h"
#include "Survey.h"
#include "TestModel.h"
#include "XiFunc.h"
#include "abn.h"
#include "cfg.h"
#include "cubature.h"
#include "sig.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

struct RefModel {
    Model* model;
    std::vector<XiFunc> xis;   // index 0 = prior xi, 1+p = derivative p
};

Config config_from_text(const char* text) {
    Config cfg = cfg_new();
    std::string s(text ? text : "");
    size_t pos = 0;
    while (pos <= s.size()) {
        size_t nl = s.find('\n', pos);
        if (nl == std::string::npos) nl = s.size();
        std::string line = s.substr(pos, nl - pos);
        if (!line.empty()) cfg_read_line(cfg, line.c_str());
        pos = nl + 1;
    }
    return cfg;
}

}  // namespace

/* Build a model from "key = value" lines, as InitializeModel does. */
REF_API void* ref_model_new(const char* cfgtext) {
    Config cfg = config_from_text(cfgtext);
    const char* name = cfg_get(cfg, "model");
    Model* m = NULL;
    if (strcmp(name, "RealModel") == 0) {
        Config sub = cfg_new_sub(cfg, "RealModel.");
        m = new RealModel(sub);
        cfg_destroy(sub);
    } else if (strcmp(name, "TestModel") == 0) {
        Config sub = cfg_new_sub(cfg, "TestModel.");
        m = new TestModel(sub);
        cfg_destroy(sub);
    }
    cfg_destroy(cfg);
    if (!m) return NULL;
    RefModel* rm = new RefModel;
    rm->model = m;
    rm->xis.push_back(m->GetXi());
    for (int p = 0; p < m->NumParams(); p++) rm->xis.push_back(m->GetXiDeriv(p));
    return rm;
}

REF_API void ref_model_free(void* h) {
    RefModel* rm = (RefModel*)h;
    if (!rm) return;
    rm->xis.clear();
    delete rm->model;
    delete rm;
}

REF_API int ref_model_nparams(void* h) { return ((RefModel*)h)->model->NumParams(); }

REF_API double ref_model_param(void* h, int n) { return ((RefModel*)h)->model->GetParam(n); }

/* xi(r) of function f (0 = prior, 1+p = derivative p) at separations r[i]. */
REF_API void ref_model_xi(void* h, int f, int n, const double* r, double* out) {
    RefModel* rm = (RefModel*)h;
    BoxSurvey box("none", 1.0, 1e30);
    SeparationFunc sep = box.GetSeparationFunction();
    for (int i = 0; i < n; i++) {
        Point p1 = {{0.0}, {0.0}, {0.0}};
        Point p2 = {{r[i]}, {0.0}, {0.0}};
        out[i] = rm->xis[f](p1, p2, sep);
    }
}

/* The reference's ComputeXiLM on a cubic spline through (k, P) samples. */
REF_API void ref_xilm(int l, int m, int nk_tab, const double* ktab, const double* ptab,
                      int nr, const double* r, double* xi, int Nk, double kmin,
                      double kmax, double kdamp) {
    Spline P = CubicSpline(nk_tab, ktab, ptab);
    ComputeXiLM(l, m, P, nr, r, xi, Nk, kmin, kmax, kdamp);
}

/* Cubic spline (netlib fmm) evaluation, src/Spline.cpp:258-359. */
REF_API void ref_cubic_spline(int n, const double* x, const double* y, int m,
                              const double* xe, double* ye) {
    Spline S = CubicSpline(n, x, y);
    for (int i = 0; i < m; i++) ye[i] = S(xe[i]);
}

static Survey* make_box(const char* survey_cfgtext) {
    Config cfg = config_from_text(survey_cfgtext);
    Config sub = cfg_new_sub(cfg, "survey.");
    Survey* s = new BoxSurvey(sub);
    cfg_destroy(sub);
    cfg_destroy(cfg);
    return s;
}

/* Minimum-image distance of the box survey (src/BoxSurvey.cpp:26-37). */
REF_API void ref_box_sep(const char* survey_cfgtext, int n, const double* p1,
                         const double* p2, double* r) {
    Survey* s = make_box(survey_cfgtext);
    SeparationFunc sep = s->GetSeparationFunction();
    for (int i = 0; i < n; i++) {
        Point a = {{p1[3 * i]}, {p1[3 * i + 1]}, {p1[3 * i + 2]}};
        Point b = {{p2[3 * i]}, {p2[3 * i + 1]}, {p2[3 * i + 2]}};
        r[i] = sep.r(a, b);
    }
    delete s;
}

/* One cell-pair integral, src/sig.cpp:416-434. cells are 80-byte records. */
REF_API double ref_signal_c(void* h, int f, const char* survey_cfgtext, const void* c1,
                            const void* c2, double epsrel, double epsabs, double* err) {
    RefModel* rm = (RefModel*)h;
    Survey* s = make_box(survey_cfgtext);
    SeparationFunc sep = s->GetSeparationFunction();
    double e = 0;
    double q = ComputeSignalC(*(const Cell*)c1, *(const Cell*)c2, rm->xis[f], sep, epsrel,
                              epsabs, &e);
    if (err) *err = e;
    delete s;
    return q;
}

/* A block of S (f = 0) or C_{,p} (f = 1+p), src/sig.cpp:50-123. */
REF_API int ref_sig_block(void* h, int f, const char* survey_cfgtext, const void* cells,
                          int ncells, const int* rows, int nrows, const int* cols,
                          int ncols, double* s, int lld, double epsrel, double epsabs) {
    RefModel* rm = (RefModel*)h;
    Survey* sv = make_box(survey_cfgtext);
    std::vector<int> R(rows, rows + nrows), C(cols, cols + ncols);
    ComputeSignalMatrixBlock(ncells, R, C, s, lld, rm->xis[f], sv, (const Cell*)cells,
                             epsrel, epsabs);
    delete sv;
    return 0;
}

/* Read a cellfile with the reference's reader (src/Cell.cpp:16-80). */
REF_API int ref_read_cells(const char* path, void* out, int maxcells) {
    int n = 0;
    Cell* c = ReadCells(path, &n);
    if (!c) return -1;
    if (n <= maxcells) memcpy(out, c, sizeof(Cell) * n);
    free(c);
    return n;
}

REF_API int ref_sizeof_cell(void) { return (int)sizeof(Cell); }

/* Read a modefile with the reference's ReadModes (src/sig.cpp:646-697):
 * returns Nmodes, or -1 when ReadModesHeader rejects the file; *ncells gets
 * Ncells; up to maxvals doubles are copied to out. */
REF_API int ref_read_modes(const char* path, int* ncells, double* out, long maxvals) {
    int Nmodes = -1, Ncells = -1;
    real* m = ReadModes(path, &Nmodes, &Ncells);
    if (ncells) *ncells = Ncells;
    if (!m) return -1;
    const long n = (long)Nmodes * Ncells;
    if (out) memcpy(out, m, sizeof(double) * (size_t)(n < maxvals ? n : maxvals));
    free(m);
    return Nmodes;
}

/* The first data block of an .abn file read by the reference's abn_read
 * (src/abn.c): returns the element count n, or -1 on a read error or an
 * element size other than 8; up to maxvals doubles are copied to out. */
REF_API long ref_read_abn_doubles(const char* path, double* out, long maxvals) {
    FILE* fp = fopen(path, "rb");
    if (!fp) return -1;
    void* data = NULL;
    size_t n = 0, size = 0;
    int err = abn_read(fp, &data, &n, &size, NULL, NULL, NULL);
    fclose(fp);
    if (err || size != sizeof(double)) {
        free(data);
        return -1;
    }
    if (out) memcpy(out, data, sizeof(double) * (n < (size_t)maxvals ? n : (size_t)maxvals));
    free(data);
    return (long)n;
}

/* Survey::GetGridIndex of the box survey (src/Survey.cpp:93-101) at n points
 * given as float triples (galfile coordinates). */
REF_API void ref_grid_index(const char* survey_cfgtext, int n, const float* xyz, int* out) {
    Survey* s = make_box(survey_cfgtext);
    for (int i = 0; i < n; i++) out[i] = s->GetGridIndex(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    delete s;
}

/* Galaxies of survey.galfile read by BoxSurvey::GetGalaxies
 * (src/BoxSurvey.cpp:105-133): returns the count, copies up to maxgals
 * (x1, x2, x3, w) records. */
REF_API int ref_read_galaxies(const char* survey_cfgtext, float* out, int maxgals) {
    Survey* s = make_box(survey_cfgtext);
    std::vector<Galaxy> gals;
    s->GetGalaxies(gals);
    delete s;
    int n = (int)gals.size();
    if (n <= maxgals && n > 0) memcpy(out, &gals[0], sizeof(Galaxy) * n);
    return n;
}

/* FinalizeCell's integral (src/basis.cpp:89-110) through the reference's own
 * adapt_integrate_v, called with basis's arguments: maxeval 5e6 and
 * (epsrel, epsabs) in its argument order.  The integrand restates basis's
 * (src/basis.cpp:54-80): a constant nbar (kind 0) or, for spherical cells, the
 * reference's LinearSpline of (rr, rn) at r = cbrt(u0) (kind 1, no angular
 * mask: the HEALPix reader needs cfitsio).  lo/hi: the integration box
 * ((x, y, z), or (r^3, mu, phi)).  Returns the reference's integral. */
namespace {
struct ShimSel {
    int kind;
    double nbar;
    Spline radial;
};
void shim_sel_integrand(unsigned, unsigned npt, const double* uu, void* data, unsigned, double* fval) {
    ShimSel* s = (ShimSel*)data;
    for (unsigned i = 0; i < npt; i++) {
        const double* u = &uu[3 * i];
        fval[i] = s->kind == 0 ? s->nbar : 1.0 * s->radial(cbrt(u[0]));
    }
}
}  // namespace

REF_API double ref_cell_integral(int kind, double nbar, int nrad, const double* rr, const double* rn,
                                 const double* lo, const double* hi, double epsrel, double epsabs, int64_t* neval) {
    ShimSel s;
    s.kind = kind;
    s.nbar = nbar;
    if (kind == 1) s.radial = LinearSpline(nrad, rr, rn);
    double N = 0, err = 0;
    adapt_integrate_v(1, shim_sel_integrand, &s, 3, lo, hi, 5000000, epsrel, epsabs, &N, &err);
    if (neval) *neval = 0;
    return N;
}
