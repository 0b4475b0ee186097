// Shared start-up of the klt / comma / estimate programs: cellfile, model and
// survey from the reference's configuration keys, HIP context and options.
#pragma once

#include <string>
#include <vector>

#include "formats.h"
#include "opsec_hip.h"

namespace opsec_host {

// opsec_exit / opsec_abort equivalents (src/opsec.c:58-116)
[[noreturn]] void prog_exit(int status);
[[noreturn]] void prog_abort(int status);

// ReadCells (src/Cell.cpp:16-80): abn file of 80-byte "4i8d" records whose
// header carries Ncells.  Aborts on unreadable or inconsistent files.
std::vector<opsec_cell> read_cells(const std::string& path, Config* cellopts = nullptr);

// Mode vectors of a modefile (klt output, src/klt.cpp:210-248): Nmodes
// consecutive vectors of Ncells doubles.  With exact, the header must carry
// Nmodes and Ncells equal to the arguments and exactly Nmodes x Ncells
// doubles, as ReadModesHeader requires (src/sig.cpp:621-645); otherwise header
// values, where present, must match and the first Nmodes vectors are read.
// Aborts on any mismatch.
std::vector<double> read_modes(const std::string& path, int Nmodes, int Ncells, bool exact, const char* prog);

// Galaxies of a galfile: "4f" abn records (x1, x2, x3, w), byte-swapped if
// written on a machine of the other endianness (src/BoxSurvey.cpp:105-133,
// src/SphericalSurvey.cpp:142-170).  Prints why and returns what was read
// (nothing if the file or its header is unusable).
std::vector<opsec_galaxy> read_galaxies(const std::string& path, const char* who);

// First two columns of a text table, '#' lines skipped (LoadData,
// src/Spline.cpp:49-88).  Returns false (after printing why) if unreadable.
bool load_columns(const std::string& path, std::vector<double>& x, std::vector<double>& y);

// Model from the "model" key and its "<Model>." sub-config
// (src/Model.cpp:26-51, src/RealModel.cpp:33-76, src/TestModel.cpp:34-45).
// Returns false (after printing why) if the model cannot be built.
struct ModelSpec {
    std::string name;                 // RealModel or TestModel
    std::vector<double> pk_k, pk_P;   // RealModel pkfile
    std::vector<double> band_lo, band_hi;
    int test_form = 0;
    double test_A = 1.0, test_b = 0.01;
    int nparams() const { return name == "TestModel" ? 2 : (int)band_lo.size(); }
};
bool load_model(const Config& cfg, ModelSpec* m);
int upload_model(opsec_ctx* ctx, const ModelSpec& m, int xi_mode);

// Survey from "survey.type" and the "survey." sub-config (src/Survey.cpp:140-165):
// BoxSurvey (src/BoxSurvey.cpp:85-91) or SphericalSurvey (src/Survey.cpp:31-44).
bool load_survey(const Config& cfg, opsec_survey* sv);

// New keys of this implementation (prefix "hip."):
//   hip.quad      = gm | gauss | surrogate   (default gm: the reference's cubature)
//   hip.nodes     = Gauss nodes per axis     (default 6)
//   hip.near      = near-pair threshold in cells
//   hip.xi        = table | closed           (default table: the reference's tables)
//   hip.precision = fp64 | fp32 | bf16x3     (default fp64)
//   hip.projection = dense | toeplitz        (projection method, every precision: dense S_p, or the DFT along
//                                             x of a box / phi of a shell; default dense)
//   hip.device    = device index             (default 0)
//   hip.tol, hip.block, hip.maxdim           (klt eigen-solver)
opsec_quad quad_from_config(const Config& cfg, double epsrel, double epsabs);
int xi_mode_from_config(const Config& cfg);
int precision_from_config(const Config& cfg);
int projection_from_config(const Config& cfg);

opsec_ctx* open_context(const Config& cfg);
void check(opsec_ctx* ctx, int rc, const char* what);   // aborts with the library message

}  // namespace opsec_host
