// estimate -- Fisher matrix and quadratic band-power estimates.
//
// Drop-in for the reference's estimate-mpi (src/estimate-mpi.cpp:332-661):
// same keys, stdout table ("p_n          error") and estfile (nine comment
// lines, then p_hat, Cov, F, W, M, q, f as consecutive .abn blocks).  The
// traces over the packed matrices run on the GPU; the Nparams x Nparams
// mixing-matrix algebra runs on the host in FP64, as on the reference's root.
#include <cstdio>
#include <cstring>
#include <ctime>
#include <cmath>
#include <vector>

#include "formats.h"
#include "setup.h"

using namespace opsec_host;

static const char* usage =
    "Usage: estimate [SWITCHES] [OPTIONS]\n"
    "Switches:\n"
    "  -h             Display this help\n"
    "  -c FILE        Read additional configuration options from FILE\n"
    "Configuration options:\n"
    "  estfile=FILE   Write parameter estimates to FILE (required)\n"
    "  evalfile=FILE  Read S/N eigenvalues from FILE (required)\n"
    "  covfile=FILE   Read covariance matrix derivatives from FILE (required)\n"
    "  pixfile=FILE   Read pixel values from FILE (required)\n"
    "  mixing=MODE    Mixing matrix choice: inverse (default), inverse-sqrt, or identity\n"
    "  hip.precision=fp64|fp32  Trace accumulation precision\n";

static std::vector<double> read_vector(const char* what, const std::string& path, AbnHeader* h) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        error("%s: could not read from file '%s'\n", what, path.c_str());
        prog_abort(1);
    }
    if (!abn_read_header(f, h)) {
        error("%s: error reading header (%s)\n", what, path.c_str());
        prog_abort(1);
    }
    if (h->size != sizeof(double) || h->fmt != "d") {
        error("%s: inconsistent data types (%s,%zu,%zu,%s,%s)\n", what, path.c_str(), h->size, sizeof(double),
              h->fmt.c_str(), "d");
        prog_abort(1);
    }
    std::vector<double> v(h->n);
    if (std::fread(v.data(), sizeof(double), h->n, f) != h->n) {
        error("%s: error reading data (%s)\n", what, path.c_str());
        prog_abort(1);
    }
    std::fclose(f);
    if (h->endian != abn_endian()) abn_byte_swap((char*)v.data(), v.size(), "d");
    return v;
}

int main(int argc, char** argv) {
    Config cfg;
    int code = 0;
    if (!parse_command_line(argc, argv, usage, cfg, &code)) prog_exit(code);
    std::string miss = cfg.missing("estfile,evalfile,covfile,pixfile,Nmodes");
    if (!miss.empty()) {
        error("estimate: missing configuration options %s\n", miss.c_str());
        std::fputs(usage, stderr);
        prog_exit(1);
    }
    const std::string estfile = cfg.get("estfile"), evalfile = cfg.get("evalfile"), covfile = cfg.get("covfile"),
                      pixfile = cfg.get("pixfile");
    int Nmodes = 0;
    cfg.get_int("Nmodes", &Nmodes);
    FILE* fest = std::fopen(estfile.c_str(), "wb");
    if (!fest) {
        error("estimate: cannot write to '%s'\n", estfile.c_str());
        prog_exit(1);
    }
    const int mixing = cfg.get_enum("mixing", {{"inverse", OPSEC_MIX_INVERSE}, {"inverse-sqrt", OPSEC_MIX_INVERSE_SQRT},
                                               {"identity", OPSEC_MIX_IDENTITY}}, OPSEC_MIX_INVERSE);
    ModelSpec model;
    if (!load_model(cfg, &model)) prog_exit(1);
    const int Nparams = model.nparams();
    const size_t Ncoeff = (size_t)Nmodes * (Nmodes + 1) / 2;

    // eigenvalues (src/estimate-mpi.cpp:53-100)
    AbnHeader eh;
    std::vector<double> lambda = read_vector("ReadEvals", evalfile, &eh);
    if ((int)lambda.size() < Nmodes) {
        error("ReadEvals: inconsistent evals length (%s,%zu,%d)\n", evalfile.c_str(), lambda.size(), Nmodes);
        prog_abort(1);
    }
    lambda.resize(Nmodes);

    // covariance derivatives (src/estimate-mpi.cpp:152-330)
    info("Reading covariance matrix derivatives...\n");
    SplitReader fc;
    if (!fc.open(covfile)) {
        error("estimate: cannot open covfile (%s)\n", covfile.c_str());
        prog_abort(1);
    }
    AbnHeader ch;
    if (!abn_read_header(fc.stream(), &ch)) {
        error("ReadCovHeader: error reading covfile header (%s)\n", covfile.c_str());
        prog_abort(1);
    }
    if (ch.endian != abn_endian()) {
        error("ReadCovHeader: endianness mismatch (%s,%c,%c)\n", covfile.c_str(), abn_endian(), ch.endian);
        prog_abort(1);
    }
    if (ch.size != sizeof(double) || ch.fmt != "d") {
        error("ReadCovHeader: data type inconsistent (%s)\n", covfile.c_str());
        prog_abort(1);
    }
    int np_check = -1, nm_check = -1, packed = 0;
    ch.opts.get_int("Nparams", &np_check);
    ch.opts.get_int("Nmodes", &nm_check);
    if (ch.opts.has("packed")) ch.opts.get_int("packed", &packed);
    if (np_check != Nparams || nm_check != Nmodes) {
        error("estimate: inconsistent covfile (%d,%d,%d,%d)\n", Nparams, np_check, Nmodes, nm_check);
        prog_abort(1);
    }
    std::vector<double> D((size_t)Nparams * Ncoeff);
    for (int m = 0; m < Nparams; m++) {
        double* dst = &D[(size_t)m * Ncoeff];
        if (packed) {
            if (fc.read(dst, sizeof(double) * Ncoeff) != sizeof(double) * Ncoeff) {
                error("ReadCovMatrix: read error (%s,%d)\n", fc.filename().c_str(), m);
                prog_abort(1);
            }
        } else {
            // full storage: keep rows 0..j of each column j
            std::vector<double> col(Nmodes);
            for (int j = 0; j < Nmodes; j++) {
                if (fc.read(col.data(), sizeof(double) * Nmodes) != sizeof(double) * Nmodes) {
                    error("ReadCovMatrix: read error (%s,%d,%d)\n", fc.filename().c_str(), m, j);
                    prog_abort(1);
                }
                std::memcpy(dst + (size_t)j * (j + 1) / 2, col.data(), sizeof(double) * (j + 1));
            }
        }
    }
    fc.close();

    info("Reading pixel values...\n");
    AbnHeader ph;
    std::vector<double> y = read_vector("ReadPixels", pixfile, &ph);
    if ((int)y.size() != Nmodes) {
        error("estimate: error reading pixel values\n");
        prog_exit(1);
    }

    opsec_ctx* ctx = open_context(cfg);
    check(ctx, upload_model(ctx, model, OPSEC_XI_TABLE), "model");
    std::vector<double> prior(Nparams);
    check(ctx, opsec_model_params(ctx, prior.data()), "model parameters");
    int prec = precision_from_config(cfg);
    if (prec == OPSEC_PREC_BF16X3) prec = OPSEC_PREC_FP32;
    std::vector<double> phat(Nparams), q(Nparams), f(Nparams), cov((size_t)Nparams * Nparams),
        F((size_t)Nparams * Nparams), W((size_t)Nparams * Nparams), M((size_t)Nparams * Nparams);
    opsec_estimate out{phat.data(), cov.data(), F.data(), W.data(), M.data(), q.data(), f.data()};
    info("Computing Fisher matrix...\n");
    check(ctx, opsec_estimate_run(ctx, Nmodes, Nparams, lambda.data(), D.data(), y.data(), prior.data(), mixing, prec,
                                  &out),
          "estimate");
    opsec_ctx_destroy(ctx);

    std::printf("p_n          error\n");
    for (int n = 0; n < Nparams; n++) std::printf("%e %e\n", phat[n], std::sqrt(cov[(size_t)n * Nparams + n]));
    info("Writing estimates to '%s'\n", estfile.c_str());
    std::fprintf(fest, "# Parameter estimates and associated matrices.\n");
    std::fprintf(fest, "# This file consists of several consecutive real-valued binary arrays:\n");
    std::fprintf(fest, "# - parameter estimates $\\hat{p}_n$ (vector of length Nparams)\n");
    std::fprintf(fest, "# - covariance matrix $Cov[\\hat{p}_m,\\hat{p}_n]$ (Nparams-by-Nparams matrix)\n");
    std::fprintf(fest, "# - Fisher matrix $F_{mn}$ (Nparams-by-Nparams matrix)\n");
    std::fprintf(fest, "# - window matrix $W_{mn}$ (Nparams-by-Nparams matrix)\n");
    std::fprintf(fest, "# - mixing matrix $M_{mn}$ (Nparams-by-Nparams matrix)\n");
    std::fprintf(fest, "# - quadratic combinations $q_n$ (vector of length Nparams)\n");
    std::fprintf(fest, "# - bias corrections $f_n$ (vector of length Nparams)\n");
    time_t now = time(nullptr);
    std::fprintf(fest, "# Generated %s\n", ctime(&now));
    Config opts;
    opts.set_int("Nparams", Nparams);
    abn_write(fest, phat.data(), Nparams, "d", &opts);
    abn_write(fest, cov.data(), (size_t)Nparams * Nparams, "d", nullptr);
    abn_write(fest, F.data(), (size_t)Nparams * Nparams, "d", nullptr);
    abn_write(fest, W.data(), (size_t)Nparams * Nparams, "d", nullptr);
    abn_write(fest, M.data(), (size_t)Nparams * Nparams, "d", nullptr);
    abn_write(fest, q.data(), Nparams, "d", nullptr);
    abn_write(fest, f.data(), Nparams, "d", nullptr);
    std::fclose(fest);
    return 0;
}
