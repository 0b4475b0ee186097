// basis -- the survey's basis cells with their expected galaxy counts.
//
// Drop-in for the reference's basis (src/basis.cpp): same keys, same cellfile
// ("4i8d" records, the survey's options plus Ncells in the header).  Every
// cell of the survey grid gets Nbar = integral of the selection function over
// the cell, computed on the GPU with the reference's adaptive cubature
// (libopsec_hip, opsec_basis); cells with Nbar <= 0 are dropped.
//
// Differences from the reference, all deliberate:
//   - spherical cells get the expected count itself: the reference integrates
//     over r^3 without the 1/3 of d(r^3) = 3 r^2 dr (src/basis.cpp:67-79,
//     104-108), so its spherical Nbar is 3x too large;
//   - SphericalSurvey reads MuMin from "MuMin" (src/Survey.cpp:39 reads "MuMuin");
//   - survey.maskfile = none selects completeness 1 everywhere (a survey
//     bounded by its own Mu / Phi limits, e.g. a polar cap), where the
//     reference needs a HEALPix file;
//   - coordsys must agree with survey.type.
#include <cstdio>
#include <string>
#include <vector>

#include "formats.h"
#include "healpix_fits.h"
#include "setup.h"

using namespace opsec_host;

static const char* usage =
    "Usage: basis [SWITCHES] [OPTIONS]\n"
    "Switches:\n"
    "  -h                Print this usage information\n"
    "  -c FILE           Read configuration options from FILE\n"
    "Configuration options:\n"
    "  coordsys=TYPE Coordinate system; either spherical or cartesian (required)\n"
    "  cellfile=FILE Output file for basis cells (required)\n"
    "  survey.type=BoxSurvey|SphericalSurvey and the survey keys (required):\n"
    "    BoxSurvey:       XMin,XMax,YMin,YMax,ZMin,ZMax,Nx,Ny,Nz, nbar, L, galfile\n"
    "    SphericalSurvey: RMin,RMax,MuMin,MuMax,PhiMin,PhiMax,Nr,Nmu,Nphi,\n"
    "                     radial (two-column r, nbar table), maskfile (HEALPix FITS, or none), galfile\n";

int main(int argc, char** argv) {
    Config cfg;
    int code = 0;
    if (!parse_command_line(argc, argv, usage, cfg, &code)) prog_exit(code);
    if (!cfg.has("cellfile")) {
        error("basis: missing config option 'cellfile'\n");
        prog_exit(1);
    }
    const std::string cellfile = cfg.get("cellfile");
    FILE* fcells = std::fopen(cellfile.c_str(), "w");
    if (!fcells) {
        error("basis: cannot write to '%s'\n", cellfile.c_str());
        prog_exit(1);
    }
    const int coordsys = cfg.get_enum("coordsys", {{"cartesian", OPSEC_COORD_CARTESIAN},
                                                   {"spherical", OPSEC_COORD_SPHERICAL}}, -1);
    if (coordsys == -1 || !cfg.has("coordsys")) {
        error("basis: invalid or missing config option: coordsys = %s\n", cfg.get("coordsys").c_str());
        prog_exit(1);
    }
    opsec_survey sv;
    if (!load_survey(cfg, &sv)) prog_exit(1);
    if (sv.coordsys != coordsys) {
        error("basis: coordsys = %s does not match survey.type = %s\n", cfg.get("coordsys").c_str(),
              cfg.get("survey.type").c_str());
        prog_exit(1);
    }
    Config s = cfg.sub("survey.");

    // selection function (src/BoxSurvey.cpp:59-75, src/SphericalSurvey.cpp:88-124)
    opsec_selection sel = {};
    std::vector<double> rr, rn;
    HealpixMap mask;
    Config opts;
    if (coordsys == OPSEC_COORD_CARTESIAN) {
        double nbar = 0;
        if (!s.get_double("nbar", &nbar) || !(nbar > 0)) {
            error("BoxSurvey: survey.nbar must be positive\n");
            prog_exit(1);
        }
        sel.kind = OPSEC_SEL_UNIFORM;
        sel.nbar = nbar;
        opts.set("coordsys", "cartesian");
        opts.set_double("XMin", sv.min1);
        opts.set_double("XMax", sv.max1);
        opts.set_double("YMin", sv.min2);
        opts.set_double("YMax", sv.max2);
        opts.set_double("ZMin", sv.min3);
        opts.set_double("ZMax", sv.max3);
        opts.set_int("Nx", sv.n1);
        opts.set_int("Ny", sv.n2);
        opts.set_int("Nz", sv.n3);
        opts.set("galfile", s.get("galfile"));
        opts.set_double("nbar", nbar);
        opts.set_double("L", sv.L);
    } else {
        const std::string radial = s.get("radial"), maskfile = s.get("maskfile");
        if (radial == "compute") {
            error("SphericalSurvey: 'radial = compute' not implemented yet (FIXME)\n");
            prog_exit(1);
        }
        if (!load_columns(radial, rr, rn) || rr.size() < 2) {
            error("SphericalSurvey: could not read the radial selection function '%s'\n", radial.c_str());
            prog_exit(1);
        }
        sel.kind = OPSEC_SEL_RADIAL;
        sel.rad_r = rr.data();
        sel.rad_n = rn.data();
        sel.nrad = (int)rr.size();
        if (maskfile != "none") {
            std::string why;
            if (!read_healpix_map(maskfile, &mask, &why)) {
                error("SphericalSurvey: could not read HEALPix file '%s' (%s)\n", maskfile.c_str(), why.c_str());
                prog_exit(1);
            }
            if (mask.ordering.compare(0, 6, "NESTED") == 0) {
                sel.nest = 1;
            } else if (mask.ordering.compare(0, 4, "RING") == 0) {
                sel.nest = 0;
            } else {
                error("SphericalSurvey: unrecognized HEALPix ordering: %s\n", mask.ordering.c_str());
                prog_exit(1);
            }
            sel.nside = mask.nside;
            sel.mask = mask.map.data();
        }
        opts.set("coordsys", "spherical");
        opts.set_double("RMin", sv.min1);
        opts.set_double("RMax", sv.max1);
        opts.set_double("MuMin", sv.min2);
        opts.set_double("MuMax", sv.max2);
        opts.set_double("PhiMin", sv.min3);
        opts.set_double("PhiMax", sv.max3);
        opts.set_int("Nr", sv.n1);
        opts.set_int("Nmu", sv.n2);
        opts.set_int("Nphi", sv.n3);
        opts.set("galfile", s.get("galfile"));
        opts.set("maskfile", maskfile);
        opts.set("radial", radial);
    }

    opsec_ctx* ctx = open_context(cfg);
    const long long ngrid = (long long)sv.n1 * sv.n2 * sv.n3;
    std::vector<opsec_cell> cells((size_t)ngrid);
    int Ncells = 0;
    int64_t nevals = 0;
    check(ctx, opsec_basis(ctx, &sv, &sel, 1e-5, 1e-10, cells.data(), &Ncells, &nevals), "basis cells");
    info("Got %d non-empty cells out of %lld possible.\n", Ncells, ngrid);
    opts.set_int("Ncells", Ncells);
    info("Writing cells to '%s'\n", cellfile.c_str());
    abn_write(fcells, cells.data(), (size_t)Ncells, "4i8d", &opts);
    std::fclose(fcells);
    opsec_ctx_destroy(ctx);
    return 0;
}
