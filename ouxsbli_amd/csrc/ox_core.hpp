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
  // rho,u,v,w,p,sensor,mut,qc2 (mut, qc2 for LES only).  A stage reads set `cur`; its RK
  // epilogue writes the primitives of the stage's new state into the other
  // set, which becomes current.  `prim_for` is the state buffer the current
  // set was computed from (null: stale, k_prim runs first).
  void* prim[2][8];
  int cur;
  const double* prim_for;
  // set s was last written by a z-sweep epilogue, which stores at every ghost
  // cell its periodic image's primitives (k_prim instead takes the ghost
  // layers of the state as they are)
  bool prim_images[2];
  // residual accumulator between the sweeps, state layout, element type S
  // (FP64 in the plain form; FP32 in the FP32-bulk form, where it carries
  // FP32 flux differences and the RK combination itself stays FP64)
  void* Racc;
  int lseg[2];       // segment lengths of the y and z sweeps (tuned per case)
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

// Key of a kernel specialisation (ORDER, SCHEME, VISC, VISC_ORDER).
inline int stage_key(int order, int scheme, int visc, int vorder) {
  return order * 1000 + scheme * 100 + visc * 10 + (visc ? vorder : 0);
}

StageFn find_stage_fp64(int order, int scheme, int visc, int vorder);
StageFn find_stage_fp32(int order, int scheme, int visc, int vorder);

// After a stage whose epilogue wrote the primitives of sa.Qout into the
// other primitive set, that set becomes current.
inline void finish_stage(Core& C, const StageArgs& sa) {
  if (sa.ep != EP_RHS) {
    C.cur = 1 - C.cur;
    C.prim_for = sa.Qout;
    C.prim_images[C.cur] = true;
  }
}

// ---------------------------------------------------------------------------
// Batched stages.  The cases of one group share a kernel specialisation; one
// stage of the whole group is three persistent launches (x, y, z sweeps)
// over the concatenated tiles of its cases.
// ---------------------------------------------------------------------------
struct TileMap;

struct BatchGroupLaunch {
  hipStream_t stream;
  int ncase;
  const void* args[3];     // device arrays of the per-case sweep descriptors (x, y, z)
  const TileMap* map[3];   // device tile maps (x, y, z)
  int ntiles[3];
  int grid[3];             // persistent grid sizes
  int ep;                  // RK epilogue of the stage
  // LES groups: the SGS viscosity and its two ghost-copy passes
  const TileMap* map_les[3];
  int ntiles_les[3];
  int grid_les[3];
};

struct BatchOps {
  size_t args_bytes;  // size of one sweep descriptor
  int xt;             // interior cells per x tile
  // k_prim when the primitives are stale; the LES viscosity of the stage
  void (*prepare)(Core& C, const double* Qsrc);
  // the three sweep descriptors (x, y, z) of one case for one stage
  void (*make_args)(Core& C, const StageArgs& sa, double cx, double cy, double cz, void* out3);
  void (*launch)(const BatchGroupLaunch& L);
  // after a stage of one case with non-periodic directions: the ghost
  // primitives of the set its epilogue wrote (prims_new, 10 arrays), then the
  // Ducros sensors of the inner ghost layers
  void (*bc_after)(const Geom& g, void* const* prims_new, hipStream_t st);
};

bool find_batch_fp64(int order, int scheme, int visc, int vorder, BatchOps* ops);
bool find_batch_fp32(int order, int scheme, int visc, int vorder, BatchOps* ops);

void launch_ghost_fill(Core& C, double* Q);

}  // namespace ox
