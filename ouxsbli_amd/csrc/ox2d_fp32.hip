// ox2d_fp32.hip -- FP32-bulk 2-D stage kernels (flux arithmetic in FP32;
// primitive recovery, residual and RK combination in FP64).
#include "ox2d_impl.hpp"

namespace ox2 {
bool find_launch2_fp32(int order, int scheme, Launch2 out[3]) {
  using R_ = float;
  switch (key2(order, scheme)) {
    OX2_FOR_EACH(OX2_CASE)
    default: return false;
  }
}
}  // namespace ox2
