// ox_fp64_batch.hip -- FP64 plain-form specialisations of the batched
// (multi-case, persistent) stage launches.  Same kernels bodies as the
// single-case launches; built with -ffp-contract=off like them.
#include "ox_stage_impl.hpp"

namespace ox {
bool find_batch_fp64(int order, int scheme, int visc, int vorder, BatchOps* ops) {
  using R_ = double;
  using S_ = double;
  switch (stage_key(order, scheme, visc, vorder)) {
    OX_FOR_EACH_STAGE(OX_BATCH_CASE)
    OX_FOR_EACH_LES_STAGE(OX_BATCH_CASE)
    default: return false;
  }
}
}  // namespace ox
