// ox_fp64.hip -- FP64 plain-form specialisations.  Built with
// -ffp-contract=off: every multiply and add rounds separately, in the
// evaluation order of the reference expressions.
#include "ox_stage_impl.hpp"

namespace ox {
StageFn find_stage_fp64(int order, int scheme, int visc, int vorder) { OX_STAGE_TABLE(double, double) }
}  // namespace ox
