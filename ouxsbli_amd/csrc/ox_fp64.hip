// ox_fp64.hip -- FP64 plain-form specialisations (single case).  Built with
// -ffp-contract=off: every multiply and add rounds separately, in the
// evaluation order of the reference expressions.
#include "ox_stage_impl.hpp"

namespace ox {
StageFn find_stage_fp64(int order, int scheme, int visc, int vorder) {
  using R_ = double;
  using S_ = double;
  switch (stage_key(order, scheme, visc, vorder)) {
    OX_FOR_EACH_STAGE(OX_STAGE_CASE)
    OX_FOR_EACH_LES_STAGE(OX_STAGE_CASE)
    default: return nullptr;
  }
}
}  // namespace ox
