// ox_fp32.hip -- FP32-bulk specialisations (single case): primitive fields and
// the residual accumulator stored in FP32, face-flux and stencil arithmetic in
// FP32; primitive recovery and the RK state stay FP64 (each face flux is
// widened to FP64 before it is differenced, and the RK combination runs in
// FP64).
#include "ox_stage_impl.hpp"

namespace ox {
StageFn find_stage_fp32(int order, int scheme, int visc, int vorder) {
  using R_ = float;
  using S_ = float;
  switch (stage_key(order, scheme, visc, vorder)) {
    OX_FOR_EACH_STAGE(OX_STAGE_CASE)
    OX_FOR_EACH_LES_STAGE(OX_STAGE_CASE)
    default: return nullptr;
  }
}
}  // namespace ox
