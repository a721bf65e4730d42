// ox_fp32_batch.hip -- FP32-bulk specialisations of the batched (multi-case,
// persistent) stage launches.
#include "ox_stage_impl.hpp"

namespace ox {
bool find_batch_fp32(int order, int scheme, int visc, int vorder, BatchOps* ops) {
  using R_ = float;
  using S_ = float;
  switch (stage_key(order, scheme, visc, vorder)) {
    OX_FOR_EACH_STAGE(OX_BATCH_CASE)
    OX_FOR_EACH_LES_STAGE(OX_BATCH_CASE)
    default: return false;
  }
}
}  // namespace ox
