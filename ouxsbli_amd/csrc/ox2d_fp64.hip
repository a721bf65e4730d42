// ox2d_fp64.hip -- FP64 plain-form 2-D stage kernels.  Built with
// -ffp-contract=off: every multiply and add rounds separately, in the
// evaluation order of the reference expressions.
#include "ox2d_impl.hpp"

namespace ox2 {
bool find_launch2_fp64(int order, int scheme, Launch2 out[3]) {
  using R_ = double;
  switch (key2(order, scheme)) {
    OX2_FOR_EACH(OX2_CASE)
    default: return false;
  }
}
}  // namespace ox2
