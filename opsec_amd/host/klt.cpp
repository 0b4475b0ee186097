// klt -- signal matrix in the cell basis and its KL eigenmodes.
//
// Drop-in for the reference's klt (src/klt.cpp:100-267): same switches, keys
// and output files (evalfile: eigenvalues descending; modefile: one mode per
// Ncells consecutive doubles).  The pair quadrature, assembly and block
// Lanczos run on the GPU through libopsec_hip.
#include <cstdio>
#include <cstring>
#include <vector>

#include "formats.h"
#include "setup.h"

using namespace opsec_host;

static const char* usage =
    "Usage: klt [SWITCHES] [OPTIONS]\n"
    "Switches:\n"
    "  -h             Display this help\n"
    "  -c FILE        Read additional configuration options from FILE\n"
    "Configuration options:\n"
    "  cellfile=FILE  Read list of cells from FILE (required)\n"
    "  Nmodes=NUM     Compute NUM mode functions of greatest S/N (required)\n"
    "  modefile=FILE  Write mode function coefficients to FILE (required)\n"
    "  evalfile=FILE  Write S/N eigenvalues to FILE (required)\n"
    "  model=NAME     Use specified model (required)\n"
    "  solver=NAME    Eigensolver: parpack, anasazi or lanczos (all run the GPU block Lanczos)\n"
    "  hip.quad=gm|gauss, hip.nodes=N, hip.xi=table|closed, hip.tol=T, hip.block=B\n";

int main(int argc, char** argv) {
    Config cfg;
    int code = 0;
    if (!parse_command_line(argc, argv, usage, cfg, &code)) prog_exit(code);
    if (!cfg.has_all("cellfile,Nmodes,modefile,evalfile")) {
        error("klt: missing configuration options\n");
        std::fputs(usage, stderr);
        prog_exit(1);
    }
    const std::string cellfile = cfg.get("cellfile"), modefile = cfg.get("modefile"), evalfile = cfg.get("evalfile");
    int Nmodes = 0;
    cfg.get_int("Nmodes", &Nmodes);

    std::vector<opsec_cell> cells = read_cells(cellfile);
    const int Ncells = (int)cells.size();
    if (Nmodes > Ncells) {
        error("klt: requesting %d modes from a %d-dimensional space\n", Nmodes, Ncells);
        prog_exit(1);
    }
    ModelSpec model;
    if (!load_model(cfg, &model)) prog_exit(1);
    opsec_survey sv;
    if (!load_survey(cfg, &sv)) prog_exit(1);

    std::string solver = cfg.get("solver");
    if (solver != "parpack" && solver != "anasazi" && solver != "lanczos") {
        error("klt: unrecognized solver: %s\n", solver.c_str());
        prog_exit(1);
    }

    FILE* fmode = std::fopen(modefile.c_str(), "wb");
    if (!fmode) {
        error("klt: could not open '%s' for writing\n", modefile.c_str());
        prog_abort(1);
    }
    FILE* feval = std::fopen(evalfile.c_str(), "wb");
    if (!feval) {
        error("klt: could not open '%s' for writing\n", evalfile.c_str());
        prog_abort(1);
    }

    double epsrel = 1e-5, epsabs = 1e-10;   // sig.epsrel / sig.epsabs (src/klt.cpp:163-165)
    if (cfg.has("sig.epsrel")) cfg.get_double("sig.epsrel", &epsrel);
    if (cfg.has("sig.epsabs")) cfg.get_double("sig.epsabs", &epsabs);
    opsec_quad q = quad_from_config(cfg, epsrel, epsabs);
    double tol = 1e-8;
    int block = 0, maxdim = 0;
    if (cfg.has("hip.tol")) cfg.get_double("hip.tol", &tol);
    if (cfg.has("hip.block")) cfg.get_int("hip.block", &block);
    if (cfg.has("hip.maxdim")) cfg.get_int("hip.maxdim", &maxdim);

    opsec_ctx* ctx = open_context(cfg);
    check(ctx, upload_model(ctx, model, xi_mode_from_config(cfg)), "model");
    check(ctx, opsec_cells_upload(ctx, cells.data(), Ncells, &sv), "cells");
    int f0 = 0;
    info("Computing signal matrix (%d cells)...\n", Ncells);
    check(ctx, opsec_pair_tables(ctx, 1, &f0, &q), "pair integrals");
    info("Solving for %d eigenmodes...\n", Nmodes);
    std::vector<double> evals(Nmodes), evecs((size_t)Nmodes * Ncells);
    int nconv = 0;
    check(ctx, opsec_klt(ctx, Nmodes, tol, block, maxdim, evals.data(), evecs.data(), &nconv), "eigen-solver");
    if (nconv < Nmodes) error("klt: not all eigenvalues converged (%d, %d)\n", Nmodes, nconv);

    info("Writing eigenmodes and eigenvalues to file...\n");
    Config opts;
    opts.set("cellfile", cellfile);
    opts.set_int("Ncells", Ncells);
    opts.set_int("Nmodes", Nmodes);
    opts.set_int("nconv", nconv);
    // All Nmodes Ritz pairs are written (unconverged ones are the best
    // available), so the header stays consistent with Nmodes.
    std::fprintf(feval, "# Signal matrix eigenvalues\n");
    abn_write_header(feval, Nmodes, "d", &opts);
    std::fwrite(evals.data(), sizeof(double), Nmodes, feval);
    std::fprintf(fmode, "# Signal matrix eigenmodes\n");
    std::fprintf(fmode, "# Each mode is stored consecutively, as below:\n");
    std::fprintf(fmode, "#   v_1,1   v_1,2   ... v_1,n\n");
    std::fprintf(fmode, "#   ...\n");
    std::fprintf(fmode, "#   v_m,1   v_m,2   ... v_m,n\n");
    abn_write_header(fmode, (size_t)Nmodes * Ncells, "d", &opts);
    std::fwrite(evecs.data(), sizeof(double), evecs.size(), fmode);
    std::fclose(feval);
    std::fclose(fmode);
    opsec_ctx_destroy(ctx);
    return 0;
}
