// ox_fp32.hip -- FP32-bulk specialisations: primitive fields stored in FP32,
// face-flux and stencil arithmetic in FP32; primitive recovery, the residual
// accumulation and the RK state stay FP64 (the kernels widen each face flux to
// FP64 before it is differenced and accumulated).
#include "ox_stage_impl.hpp"

namespace ox {
StageFn find_stage_fp32(int order, int scheme, int visc, int vorder) { OX_STAGE_TABLE(float, float) }
}  // namespace ox
