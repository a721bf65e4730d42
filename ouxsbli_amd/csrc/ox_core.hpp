// ox_core.hpp -- the per-handle device context shared by the C ABI
// (ox_capi.hip) and the stage launchers (ox_fp64.hip, ox_fp32.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "ox_physics.hpp"
#include "ox_types.hpp"

namespace ox {

struct Core {
  ox_config cfg;
  Geom g;
  HostConsts hc;
  int vorder;        // viscous order actually used (0 for Euler)
  int tvd, variant;  // runtime switches (uniform across the launch)
  // Two sets of primitive arrays (element type per precision mode):
  // rho,u,v,w,p,T,mu,sensor,mut,qc2.  A stage reads set `cur`; its RK
  // epilogue writes the primitives of the stage's new state into the other
  // set, which becomes current.  `prim_for` is the state buffer the current
  // set was computed from (null: stale, k_prim runs first).
  void* prim[2][10];
  int cur;
  const double* prim_for;
  // residual accumulator between the sweeps, state layout, element type S
  // (FP64 in the plain form; FP32 in the FP32-bulk form, where it carries
  // FP32 flux differences and the RK combination itself stays FP64)
  void* Racc;
  hipStream_t stream;
  // profiling: when `prof` is set, mark() records an event after a kernel
  bool prof;
  void (*mark)(Core* C, const char* name);
  void* prof_state;
};

inline void prof_mark(Core& C, const char* name) {
  if (C.prof) C.mark(&C, name);
}

// Launches one right-hand-side evaluation on state Qsrc followed by the
// stage epilogue `sa`.  The ghost layers of sa.Qout are not filled (the
// primitives of the ghost cells are written directly); launch_ghost_fill
// does that where the state itself is read.
// cx, cy, cz are the calc_R coefficients (coef * dt * face area).
using StageFn = void (*)(Core& C, const double* Qsrc, const StageArgs& sa, double cx, double cy, double cz);

StageFn find_stage_fp64(int order, int scheme, int visc, int vorder);
StageFn find_stage_fp32(int order, int scheme, int visc, int vorder);

void launch_ghost_fill(Core& C, double* Q);

}  // namespace ox
