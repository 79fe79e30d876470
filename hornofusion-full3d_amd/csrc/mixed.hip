// mixed.hip -- placeholder, replaced by the mixed-precision solver.
#include "batch.hpp"
namespace eafp {
struct MixedWork {};
void free_mixed(MixedWork* w) { delete w; }
int solve_mixed(Batch&, const eafp_mixed_opts&, eafp_stats*) { return EAFP_ERR_UNSUPPORTED; }
}  // namespace eafp
