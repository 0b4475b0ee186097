// dot -- pixel values y_i of a galaxy catalogue in the KL basis.
//
// Drop-in for the reference's dot (src/dot.cpp:50-190): same keys, same
// pixfile ("# Pixel values" and Nmodes doubles).  The galaxies of
// survey.galfile are binned into the cells, x_a = (N_a - Nbar_a)/sqrt(Nbar_a),
// and y = B^T x; binning and the product run on the GPU (libopsec_hip).
//
// Differences from the reference, all deliberate:
//   - countfile is optional (the reference requires it and never opens it);
//     when given, the weighted count of every cell is written to it as text;
//   - a galaxy in grid cell 0 is counted in its cell (src/dot.cpp:127 tests
//     e > 0), and a galaxy in a grid position with no cell counts as outside
//     (the reference writes counts[-1]);
//   - SphericalSurvey reads MuMin from "MuMin" (src/Survey.cpp:39 reads "MuMuin").
#include <cmath>
#include <cstdio>
#include <vector>

#include "formats.h"
#include "setup.h"

using namespace opsec_host;

static const char* usage =
    "Usage: dot [SWITCHES] [OPTIONS]\n"
    "Switches:\n"
    "  -h            Display this help\n"
    "  -c FILE       Read additional configuration options from FILE\n"
    "Configuration options:\n"
    "  pixfile=FILE  Write pixel values to FILE (required)\n"
    "  cellfile=FILE Read list of cells from FILE (required)\n"
    "  modefile=FILE Read KL mode coefficients from FILE (required)\n"
    "  Nmodes=NUM    Number of KL modes (required)\n"
    "  countfile=FILE Write weighted galaxy counts per cell to FILE (optional)\n"
    "  survey.type=BoxSurvey|SphericalSurvey, survey.galfile=FILE and the survey geometry\n";

int main(int argc, char** argv) {
    Config cfg;
    int code = 0;
    if (!parse_command_line(argc, argv, usage, cfg, &code)) prog_exit(code);
    std::string miss = cfg.missing("cellfile,modefile,pixfile,Nmodes");
    if (!miss.empty()) {
        error("dot: missing configuration options %s\n", miss.c_str());
        std::fputs(usage, stderr);
        prog_exit(1);
    }
    const std::string cellfile = cfg.get("cellfile"), modefile = cfg.get("modefile"), pixfile = cfg.get("pixfile");
    int Nmodes = 0;
    cfg.get_int("Nmodes", &Nmodes);

    FILE* fpix = std::fopen(pixfile.c_str(), "wb");
    if (!fpix) {
        error("dot: could not open '%s' for writing\n", pixfile.c_str());
        prog_exit(1);
    }
    std::vector<opsec_cell> cells = read_cells(cellfile);
    const int Ncells = (int)cells.size();
    for (int a = 0; a < Ncells; a++)
        if (cells[a].a != a) {
            error("dot: cell %d of '%s' is numbered %d\n", a, cellfile.c_str(), cells[a].a);
            prog_abort(1);
        }
    if (Nmodes < 1 || Nmodes > Ncells) {
        error("dot: inconsistent modes file\n");
        prog_exit(1);
    }
    std::vector<double> B = read_modes(modefile, Nmodes, Ncells, true, "dot");
    info("Read %d modes from '%s'\n", Nmodes, modefile.c_str());

    opsec_survey sv;
    if (!load_survey(cfg, &sv)) prog_exit(1);
    const char* who = sv.coordsys == OPSEC_COORD_SPHERICAL ? "SphericalSurvey" : "BoxSurvey";
    std::vector<opsec_galaxy> gals = read_galaxies(cfg.get("survey.galfile"), who);
    if (gals.empty()) {
        error("dot: error reading galaxies\n");
        prog_exit(1);
    }
    info("Read %zu galaxies\n", gals.size());

    opsec_ctx* ctx = open_context(cfg);
    check(ctx, opsec_cells_upload(ctx, cells.data(), Ncells, &sv), "cells");
    check(ctx, opsec_modes_upload(ctx, B.data(), Ncells, Nmodes), "modes");
    std::vector<double> counts(Ncells + 1, 0.0), y(Nmodes);
    check(ctx, opsec_count_galaxies(ctx, gals.data(), (int64_t)gals.size(), 0, counts.data()), "galaxy counts");
    check(ctx, opsec_pixel_values(ctx, counts.data(), nullptr, y.data()), "pixel values");
    opsec_ctx_destroy(ctx);

    std::fprintf(fpix, "# Pixel values\n");
    abn_write(fpix, y.data(), (size_t)Nmodes, "d", nullptr);
    std::fclose(fpix);

    if (cfg.has("countfile")) {
        const std::string countfile = cfg.get("countfile");
        FILE* fc = std::fopen(countfile.c_str(), "w");
        if (!fc) {
            error("dot: could not open '%s' for writing\n", countfile.c_str());
            prog_exit(1);
        }
        std::fprintf(fc, "# Cell counts\n");
        std::fprintf(fc, "# %g (weighted) galaxies outside any cell\n", counts[Ncells]);
        std::fprintf(fc, "# Columns are:  cell number -- observed count -- expected count\n");
        for (int a = 0; a < Ncells; a++) std::fprintf(fc, "%6d %.17g %.17g\n", a, counts[a], cells[a].Nbar);
        std::fclose(fc);
    }
    return 0;
}
