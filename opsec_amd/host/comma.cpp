// comma -- derivatives of the covariance matrix projected onto the KL modes.
//
// Drop-in for the reference's comma-scalapack (src/comma-scalapack.cpp:100-289):
// same keys and covfile (Nparams consecutive packed upper triangles, column j
// holding rows 0..j).  The band pair tables, the fused S_p generation and both
// projection GEMMs for all bands run on the GPU (libopsec_hip).  As in the
// reference, the sig.* keys are ignored and the default cubature tolerances
// are used (src/comma-scalapack.cpp:237-243).
#include <algorithm>
#include <cstdio>
#include <future>
#include <vector>

#include "formats.h"
#include "setup.h"

using namespace opsec_host;

static const char* usage =
    "Usage: comma [SWITCHES] [OPTIONS]\n"
    "Switches:\n"
    "  -h             Display this help\n"
    "  -c FILE        Read additional configuration options from FILE\n"
    "Configuration options:\n"
    "  cellfile=FILE  Read list of cells from FILE (required)\n"
    "  modefile=FILE  Read KL modes from FILE (required)\n"
    "  covfile=FILE   Write covariance matrices to FILE (required)\n"
    "  Nmodes=NUM     Number of KL modes (required)\n"
    "  model=NAME     Use specified model (required)\n"
    "  hip.quad=gm|gauss|surrogate, hip.nodes=N, hip.xi=table|closed, hip.precision=fp64|fp32|bf16x3,\n"
    "  hip.projection=dense|toeplitz\n";

int main(int argc, char** argv) {
    Config cfg;
    int code = 0;
    if (!parse_command_line(argc, argv, usage, cfg, &code)) prog_exit(code);
    std::string miss = cfg.missing("cellfile,modefile,covfile,Nmodes");
    if (!miss.empty()) {
        error("comma: missing configuration options %s\n", miss.c_str());
        std::fputs(usage, stderr);
        prog_exit(1);
    }
    const std::string cellfile = cfg.get("cellfile"), modefile = cfg.get("modefile"), covfile = cfg.get("covfile");
    int Nmodes = 0;
    cfg.get_int("Nmodes", &Nmodes);
    std::vector<opsec_cell> cells = read_cells(cellfile);
    const int Ncells = (int)cells.size();
    if (Ncells < Nmodes) {
        error("comma: Nmodes %d exceeds Ncells %d\n", Nmodes, Ncells);
        prog_abort(1);
    }
    ModelSpec model;
    if (!load_model(cfg, &model)) prog_exit(1);
    const int Nparams = model.nparams();
    opsec_survey sv;
    if (!load_survey(cfg, &sv)) prog_exit(1);

    // modes (src/comma-scalapack.cpp:52-98)
    std::vector<double> B = read_modes(modefile, Nmodes, Ncells, false, "comma");

    FILE* fout = std::fopen(covfile.c_str(), "wb");
    if (!fout) {
        error("comma: could not open '%s' for writing\n", covfile.c_str());
        prog_exit(1);
    }
    std::fprintf(fout, "# Derivatives of the covariance matrix in KL basis, C_{ij,m}.\n");
    std::fprintf(fout, "# Binary data consists of %d consecutive symmetric %dx%d matrices\n", Nparams, Nmodes, Nmodes);
    std::fprintf(fout, "# in packed storage.  Each matrix consists of %d elements.\n", Nmodes * (Nmodes + 1) / 2);
    Config opts;
    opts.set_int("Nmodes", Nmodes);
    opts.set_int("Nparams", Nparams);
    opts.set_int("packed", 1);
    const size_t K = (size_t)Nmodes * (Nmodes + 1) / 2;
    abn_write_header(fout, (size_t)Nparams * K, "d", &opts);

    opsec_quad q = quad_from_config(cfg, 1e-5, 1e-10);
    const int prec = precision_from_config(cfg);
    opsec_ctx* ctx = open_context(cfg);
    check(ctx, opsec_set_projection(ctx, projection_from_config(cfg)), "projection method");
    check(ctx, upload_model(ctx, model, xi_mode_from_config(cfg)), "model");
    check(ctx, opsec_cells_upload(ctx, cells.data(), Ncells, &sv), "cells");
    check(ctx, opsec_modes_upload(ctx, B.data(), Ncells, Nmodes), "modes");
    std::vector<int> funcs(Nparams);
    for (int p = 0; p < Nparams; p++) funcs[p] = p + 1;
    info("Computing pair integrals for %d parameters\n", Nparams);
    check(ctx, opsec_pair_tables(ctx, Nparams, funcs.data(), &q), "pair integrals");
    // Project in groups of bands and write each group while the next one is
    // computed: two host buffers, the file writes on a second thread.  A group
    // is at most 4 bands (64 MB of packed output per band at 4000 modes) and
    // its stage-1 products (Ncells x Nmodes per band) stay well inside HBM.
    const size_t per_band = (size_t)Ncells * Nmodes * (prec == OPSEC_PREC_FP64 ? 8 : 4);
    int group = (int)std::max<size_t>(1, std::min<size_t>({(size_t)Nparams, (size_t)4,
                                                           ((size_t)96 << 30) / std::max<size_t>(per_band, 1)}));
    std::vector<double> P[2] = {std::vector<double>((size_t)group * K), std::vector<double>((size_t)group * K)};
    std::future<size_t> pending;
    size_t expect = 0;
    int buf = 0;
    for (int p0 = 0; p0 < Nparams; p0 += group) {
        int ng = std::min(group, Nparams - p0);
        info("Projecting parameters %d-%d onto %d modes\n", p0 + 1, p0 + ng, Nmodes);
        check(ctx, opsec_project(ctx, ng, funcs.data() + p0, prec, P[buf].data()), "projection");
        if (pending.valid() && pending.get() != expect) {
            error("comma: error writing '%s'\n", covfile.c_str());
            prog_exit(1);
        }
        expect = (size_t)ng * K;
        const double* src = P[buf].data();
        pending = std::async(std::launch::async, [=] { return std::fwrite(src, sizeof(double), expect, fout); });
        buf ^= 1;
    }
    if (pending.valid() && pending.get() != expect) {
        error("comma: error writing '%s'\n", covfile.c_str());
        prog_exit(1);
    }
    std::fclose(fout);
    opsec_ctx_destroy(ctx);
    return 0;
}
