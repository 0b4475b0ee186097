// Entry points of the surrogate and eigen-solver units until they land.
#include "common.h"
namespace opsec {
struct PairGeo;
int surrogate_tables(opsec_ctx* ctx, int, const int*, const opsec_quad*, const std::vector<int>&,
                     const std::vector<struct PairGeo>&, const std::vector<int>&) {
    return set_error(ctx, OPSEC_ESTATE, "surrogate not available");
}
}
extern "C" void opsec_internal_free_mlp(opsec_ctx*) {}
