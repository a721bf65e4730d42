// ox2d_impl.hpp -- instantiation of the 2-D stage kernels and their launcher
// table for one arithmetic type (included once by ox2d_fp64.hip and once by
// ox2d_fp32.hip, which differ in their contraction flags).
#pragma once

#include "ox2d_core.hpp"

namespace ox2 {

template <class R, int ORDER, int SCHEME, int EP>
void launch2(const void* args, int ncase, int ntiles, int grid, hipStream_t stream) {
  k2d_stage<R, ORDER, SCHEME, EP><<<grid, 256, 0, stream>>>((const Args2<R>*)args, ncase, ntiles);
}

template <class R, int ORDER, int SCHEME, int EP> int max_blocks2() {
  static int cap = 0;
  if (cap == 0) {
    int occ = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k2d_stage<R, ORDER, SCHEME, EP>, 256, 0) != hipSuccess)
      occ = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    cap = (occ < 1 ? 1 : occ) * (cus < 1 ? 1 : cus);
  }
  return cap;
}

template <class R, int ORDER, int SCHEME> void table2(Launch2 out[3]) {
  out[0] = {&launch2<R, ORDER, SCHEME, EP2_S1>, &max_blocks2<R, ORDER, SCHEME, EP2_S1>};
  out[1] = {&launch2<R, ORDER, SCHEME, EP2_S2>, &max_blocks2<R, ORDER, SCHEME, EP2_S2>};
  out[2] = {&launch2<R, ORDER, SCHEME, EP2_S3>, &max_blocks2<R, ORDER, SCHEME, EP2_S3>};
}

// Supported (ORDER, SCHEME) pairs of the 2-D path.
#define OX2_FOR_EACH(X) \
  X(2, OX_SCHEME_SLAU)  \
  X(4, OX_SCHEME_SLAU)  \
  X(6, OX_SCHEME_SLAU)  \
  X(2, OX_SCHEME_ROE)   \
  X(4, OX_SCHEME_ROE)   \
  X(6, OX_SCHEME_ROE)

#define OX2_CASE(O, SC)      \
  case O * 10 + SC:          \
    table2<R_, O, SC>(out);  \
    return true;

}  // namespace ox2
