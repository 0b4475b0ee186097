// Program start-up helpers (see setup.h).
#include "setup.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace opsec_host {

void prog_exit(int status) {
    std::printf("OPSEC is now exiting.\n\n");
    std::fflush(stdout);
    std::fflush(stderr);
    std::exit(status);
}

void prog_abort(int status) {
    (void)status;
    std::fprintf(stderr, "OPSEC is aborting!!!.\n\n");
    std::fflush(stdout);
    std::fflush(stderr);
    std::abort();
}

std::vector<opsec_cell> read_cells(const std::string& path, Config* cellopts) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        error("ReadCells: could not open '%s'\n", path.c_str());
        prog_abort(1);
    }
    AbnHeader h;
    if (!abn_read_header(f, &h)) {
        error("ReadCells: error reading cells from '%s'\n", path.c_str());
        prog_abort(1);
    }
    int ncells = 0;
    if (!h.opts.get_int("Ncells", &ncells) || (size_t)ncells != h.n) {
        error("ReadCells: consistency check fail for Ncells: %d != %zu\n", ncells, h.n);
        prog_abort(1);
    }
    if (h.size != sizeof(opsec_cell)) {
        error("ReadCells: consistency check fail for sizeof(Cell): %zu != %zu\n", sizeof(opsec_cell), h.size);
        prog_abort(1);
    }
    std::vector<opsec_cell> cells(h.n);
    if (std::fread(cells.data(), sizeof(opsec_cell), h.n, f) != h.n) {
        error("ReadCells: error reading cells from '%s'\n", path.c_str());
        prog_abort(1);
    }
    if (h.endian != abn_endian()) abn_byte_swap((char*)cells.data(), h.n, h.fmt);
    std::fclose(f);
    if (cellopts) *cellopts = h.opts;
    return cells;
}

std::vector<double> read_modes(const std::string& path, int Nmodes, int Ncells, bool exact, const char* prog) {
    SplitReader fm;
    if (!fm.open(path)) {
        error("%s: cannot open modes file %s\n", prog, path.c_str());
        prog_abort(1);
    }
    AbnHeader mh;
    if (!abn_read_header(fm.stream(), &mh)) prog_abort(1);
    int hv = 0;
    if ((exact || mh.opts.has("Nmodes")) && (!mh.opts.get_int("Nmodes", &hv) || hv != Nmodes)) {
        error("%s: modefile has Nmodes = %s, expected %d\n", prog, mh.opts.get("Nmodes").c_str(), Nmodes);
        prog_abort(1);
    }
    if ((exact || mh.opts.has("Ncells")) && (!mh.opts.get_int("Ncells", &hv) || hv != Ncells)) {
        error("%s: modefile has Ncells = %s, expected %d\n", prog, mh.opts.get("Ncells").c_str(), Ncells);
        prog_abort(1);
    }
    if (exact && (mh.n != (size_t)Nmodes * Ncells || mh.size != sizeof(double))) {
        error("%s: modefile holds %zu items of %zu bytes, expected %zu doubles\n", prog, mh.n, mh.size,
              (size_t)Nmodes * Ncells);
        prog_abort(1);
    }
    std::vector<double> B((size_t)Nmodes * Ncells);
    for (int j = 0; j < Nmodes; j++) {
        size_t nread = fm.read(&B[(size_t)j * Ncells], sizeof(double) * Ncells);
        if (nread != sizeof(double) * Ncells) {
            error("%s: error reading modes (%s,%d,%zu,%d)\n", prog, fm.filename().c_str(), j, nread, Ncells);
            prog_abort(1);
        }
    }
    fm.close();
    if (mh.endian != abn_endian()) abn_byte_swap((char*)B.data(), B.size(), "d");
    return B;
}

std::vector<opsec_galaxy> read_galaxies(const std::string& path, const char* who) {
    std::vector<opsec_galaxy> gals;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        error("%s: could not open file '%s'\n", who, path.c_str());
        return gals;
    }
    AbnHeader h;
    if (!abn_read_header(f, &h) || h.size != sizeof(opsec_galaxy) || h.fmt != "4f") {
        error("%s: error reading galaxies from '%s'\n", who, path.c_str());
        std::fclose(f);
        return gals;
    }
    gals.resize(h.n);
    size_t nread = std::fread(gals.data(), sizeof(opsec_galaxy), h.n, f);
    if (nread != h.n) {
        error("%s: expecting %zu galaxies from '%s', got %zu\n", who, h.n, path.c_str(), nread);
        gals.resize(nread);
    }
    std::fclose(f);
    if (h.endian != abn_endian()) abn_byte_swap((char*)gals.data(), gals.size(), h.fmt);
    return gals;
}

// Two-column text table (src/Spline.cpp:49-88): '#' lines skipped.
bool load_columns(const std::string& path, std::vector<double>& k, std::vector<double>& P) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) {
        error("LoadData: could not read file '%s'\n", path.c_str());
        return false;
    }
    char buf[4096];
    while (std::fgets(buf, sizeof buf, f)) {
        if (buf[0] == '\0' || buf[0] == '#') continue;
        double a, b;
        if (std::sscanf(buf, "%lf %lf", &a, &b) == 2) {
            k.push_back(a);
            P.push_back(b);
        }
    }
    std::fclose(f);
    return true;
}

bool load_model(const Config& cfg, ModelSpec* m) {
    m->name = cfg.get("model");
    if (m->name == "RealModel") {
        Config sub = cfg.sub("RealModel.");
        if (!sub.has("pkfile")) {
            error("RealModel: must set config option 'pkfile'\n");
            return false;
        }
        if (!load_columns(sub.get("pkfile"), m->pk_k, m->pk_P) || m->pk_k.size() < 2) return false;
        if (sub.has("bands")) {
            // RealModel.bands: explicit edges (src/RealModel.cpp:44-59).  The
            // reference always sizes the edge list to 100 (50 bands); here the
            // list is read as given (pairs of edges).
            std::vector<double> kv = sub.get_doubles("bands", 1000);
            if (kv.size() % 2) {
                error("RealModel: odd number of kvals\n");
                kv.pop_back();
            }
            for (size_t n = 0; n + 1 < kv.size(); n += 2) {
                m->band_lo.push_back(kv[n]);
                m->band_hi.push_back(kv[n + 1]);
            }
        } else if (sub.has_all("Nbands,kmin,kmax")) {
            int nb = 0;
            double kmin = 0, kmax = 0;
            sub.get_int("Nbands", &nb);
            sub.get_double("kmin", &kmin);
            sub.get_double("kmax", &kmax);
            for (int n = 0; n < nb; n++) {
                m->band_lo.push_back(kmin + n * (kmax - kmin) / nb);
                m->band_hi.push_back(kmin + (n + 1) * (kmax - kmin) / nb);
            }
        } else {
            error("RealModel: k-bands not specified in configuration\n");
            return false;
        }
        return !m->band_lo.empty();
    }
    if (m->name == "TestModel") {
        Config sub = cfg.sub("TestModel.");
        m->test_form = sub.get_enum("form", {{"exponential", 0}, {"gaussian", 1}}, -1);
        if (m->test_form == -1) {
            error("TestModel: unrecognized functional form: '%s'\n           (defaulting to 'exponential')\n",
                  sub.get("form").c_str());
            m->test_form = 0;
        }
        if (sub.has("A")) sub.get_double("A", &m->test_A);
        if (sub.has("b")) sub.get_double("b", &m->test_b);
        return true;
    }
    error("InitializeModel: unrecognized model '%s'\n", m->name.c_str());
    return false;
}

int upload_model(opsec_ctx* ctx, const ModelSpec& m, int xi_mode) {
    if (m.name == "TestModel") return opsec_model_test(ctx, m.test_form, m.test_A, m.test_b);
    return opsec_model_real(ctx, m.pk_k.data(), m.pk_P.data(), (int)m.pk_k.size(), (int)m.band_lo.size(),
                            m.band_lo.data(), m.band_hi.data(), xi_mode);
}

bool load_survey(const Config& cfg, opsec_survey* sv) {
    if (!cfg.has("survey.type")) {
        error("InitializeSurvey: missing survey.type configuration option\n");
        return false;
    }
    std::string type = cfg.get("survey.type");
    Config s = cfg.sub("survey.");
    std::memset(sv, 0, sizeof *sv);
    if (type == "BoxSurvey") {
        std::string miss = s.missing("XMin,XMax,YMin,YMax,ZMin,ZMax,Nx,Ny,Nz");
        if (!miss.empty()) {
            error("Survey: missing configuration options %s\n", miss.c_str());
            prog_abort(1);
        }
        sv->coordsys = OPSEC_COORD_CARTESIAN;
        s.get_double("XMin", &sv->min1);
        s.get_double("XMax", &sv->max1);
        s.get_double("YMin", &sv->min2);
        s.get_double("YMax", &sv->max2);
        s.get_double("ZMin", &sv->min3);
        s.get_double("ZMax", &sv->max3);
        s.get_int("Nx", &sv->n1);
        s.get_int("Ny", &sv->n2);
        s.get_int("Nz", &sv->n3);
        if (!s.get_double("L", &sv->L)) sv->L = 0;
        return true;
    }
    if (type == "SphericalSurvey") {
        std::string miss = s.missing("RMin,RMax,MuMin,MuMax,PhiMin,PhiMax,Nr,Nmu,Nphi");
        if (!miss.empty()) {
            error("Survey: missing configuration options %s\n", miss.c_str());
            prog_abort(1);
        }
        // MuMin comes from "MuMin"; the reference reads the misspelled key
        // "MuMuin" (src/Survey.cpp:39) and so always gets 0.
        sv->coordsys = OPSEC_COORD_SPHERICAL;
        s.get_double("RMin", &sv->min1);
        s.get_double("RMax", &sv->max1);
        s.get_double("MuMin", &sv->min2);
        s.get_double("MuMax", &sv->max2);
        s.get_double("PhiMin", &sv->min3);
        s.get_double("PhiMax", &sv->max3);
        s.get_int("Nr", &sv->n1);
        s.get_int("Nmu", &sv->n2);
        s.get_int("Nphi", &sv->n3);
        sv->L = 0;
        return true;
    }
    error("InitializeSurvey: unrecognized survey type '%s'\n", type.c_str());
    return false;
}

opsec_quad quad_from_config(const Config& cfg, double epsrel, double epsabs) {
    int method = cfg.get_enum("hip.quad", {{"gm", OPSEC_QUAD_GM}, {"gauss", OPSEC_QUAD_GAUSS},
                                            {"surrogate", OPSEC_QUAD_SURROGATE}}, OPSEC_QUAD_GM);
    opsec_quad q;
    opsec_quad_defaults(method, &q);
    q.epsrel = epsrel;
    q.epsabs = epsabs;
    int v;
    if (cfg.get_int("hip.nodes", &v)) q.nodes = v;
    if (cfg.get_int("hip.near", &v)) q.near = v;
    return q;
}

int xi_mode_from_config(const Config& cfg) {
    return cfg.get_enum("hip.xi", {{"table", OPSEC_XI_TABLE}, {"closed", OPSEC_XI_CLOSED}}, OPSEC_XI_TABLE);
}

int projection_from_config(const Config& cfg) {
    return cfg.get_enum("hip.projection", {{"dense", OPSEC_PROJ_DENSE}, {"toeplitz", OPSEC_PROJ_TOEPLITZ}},
                        OPSEC_PROJ_DENSE);
}

int precision_from_config(const Config& cfg) {
    return cfg.get_enum("hip.precision", {{"fp64", OPSEC_PREC_FP64}, {"fp32", OPSEC_PREC_FP32},
                                          {"bf16x3", OPSEC_PREC_BF16X3}}, OPSEC_PREC_FP64);
}

opsec_ctx* open_context(const Config& cfg) {
    int dev = 0;
    if (cfg.has("hip.device")) cfg.get_int("hip.device", &dev);
    opsec_ctx* ctx = nullptr;
    int rc = opsec_ctx_create(dev, &ctx);
    if (rc) {
        error("could not open gfx950 device %d (error %d)\n", dev, rc);
        prog_exit(1);
    }
    return ctx;
}

void check(opsec_ctx* ctx, int rc, const char* what) {
    if (rc) {
        error("%s failed (%d): %s\n", what, rc, opsec_last_error(ctx));
        prog_abort(1);
    }
}

}  // namespace opsec_host
