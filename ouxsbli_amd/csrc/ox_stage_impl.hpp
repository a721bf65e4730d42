// ox_stage_impl.hpp -- stage launcher templates and the specialisation tables.
//
// The reference specialises each case at build time with fypp
// (src/mod_constant.f90.fypp:37-100: VISC, ORDER, SCHEME encoded as kinds);
// here (ORDER, SCHEME, VISC, VISC_ORDER) are template parameters and every
// supported combination is instantiated once.  TVD kind and SLAU variant are
// uniform runtime switches.
#pragma once

#include "ox_bc.hpp"
#include "ox_core.hpp"
#include "ox_kernels.hpp"

namespace ox {

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

template <class S> Prim<S> prim_set(const Core& C, int set) {
  Prim<S> P;
  S** a[kPrimArrays] = {&P.rho, &P.u, &P.v, &P.w, &P.p, &P.s, &P.mut, &P.qc2};
  for (int m = 0; m < kPrimArrays; ++m) *a[m] = (S*)C.prim[set][m];
  return P;
}

// Launch grids of the three sweeps of one case (tiles along the three
// launch dimensions).
template <int ORDER, int VISC, int VO> dim3 grid_x(const Geom& g) {
  return dim3(cdiv(g.ix, xtile<ORDER, VISC, VO>()), cdiv(g.iy, XROWS), g.iz);
}
// k_sweep_xrow: 256 / ix rows per workgroup, bands of XROW_BAND rows per plane
inline dim3 grid_xrow(const Geom& g) {
  return dim3((unsigned)(cdiv(g.iy, XROW_BAND) * g.iz * (XROW_BAND / (256 / g.ix))));
}
inline dim3 grid_y(const Geom& g, int lseg) {
  return dim3(cdiv(g.ix, LINE_X), cdiv(g.iz, LINE_Y), cdiv(g.iy, lseg));
}
inline dim3 grid_z(const Geom& g, int lseg) {
  return dim3(cdiv(g.ix, LINE_X), cdiv(g.iy, LINE_Y), cdiv(g.iz, lseg));
}
inline dim3 line_block() { return dim3(LINE_X, LINE_Y, 1); }

// What precedes the sweeps of a stage on one case: primitives when they do not
// describe Qsrc, and for LES the SGS viscosity of the stage.
template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
void prepare_stage(Core& C, const double* Qsrc) {
  const Geom& g = C.g;
  const Prim<S> P = prim_set<S>(C, C.cur);
  const dim3 blk(64, 4, 1);
  if (C.prim_for != Qsrc) {
    k_prim<S><<<dim3(cdiv(g.nx, 64), cdiv(g.ny, 4), g.nz), blk, 0, C.stream>>>(
        Qsrc, P, g, make_physc<double>(C.hc));
    C.prim_for = Qsrc;
    C.prim_images[C.cur] = false;
    if (any_bc(g)) launch_bc_sensors<R, S, ORDER, SCHEME>(g, P, C.stream);
    prof_mark(C, "prim");
  }
  if constexpr (VISC == 2) {
    const PhysC<R> cr = make_physc<R>(C.hc);
    k_les_mut<S, R><<<dim3(cdiv(g.nx - 2, 64), cdiv(g.ny - 2, 4), g.nz - 2), blk, 0, C.stream>>>(P, g, cr);
    k_les_bc<S><<<(unsigned)((les_bc_items(g, 0) + 255) / 256), 256, 0, C.stream>>>(P, g, 0);
    k_les_bc<S><<<(unsigned)((les_bc_items(g, 1) + 255) / 256), 256, 0, C.stream>>>(P, g, 1);
    prof_mark(C, "les_mut");
  }
}

// The x, y and z sweep descriptors of one stage of one case.
template <class R, class S>
void make_sweep_args(const Core& C, const StageArgs& sa, double cx, double cy, double cz, SweepArgs<R, S>* a3) {
  SweepArgs<R, S> a{};
  a.P = prim_set<S>(C, C.cur);
  a.Pn = prim_set<S>(C, 1 - C.cur);
  a.Racc = (S*)C.Racc;
  a.g = C.g;
  a.c = make_physc<R>(C.hc);
  a.c64 = make_physc<double>(C.hc);
  a.sa = sa;
  a.tvd = C.tvd;
  a.variant = C.variant;
  a3[0] = a;
  a3[0].cd = cx;
  a3[0].lseg = 0;
  a3[1] = a;
  a3[1].cd = cy;
  a3[1].lseg = C.lseg[0];
  a3[2] = a;
  a3[2].cd = cz;
  a3[2].lseg = C.lseg[1];
}

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
void run_stage(Core& C, const double* Qsrc, const StageArgs& sa, double cx, double cy, double cz) {
  const Geom& g = C.g;
  const dim3 blk(64, 4, 1);
  prepare_stage<R, S, ORDER, SCHEME, VISC, VO>(C, Qsrc);
  SweepArgs<R, S> a[3];
  make_sweep_args<R, S>(C, sa, cx, cy, cz, a);
  bool rows = false;
  if constexpr (kXRowSweep<ORDER, SCHEME, VISC>) {
    rows = xrow_width(g.ix) && !any_bc(g) && C.prim_images[C.cur];
    if (rows) k_sweep_xrow<R, S, ORDER, SCHEME, VISC, VO><<<grid_xrow(g), 256, 0, C.stream>>>(a[0]);
  }
  if (!rows) k_sweep_x<R, S, ORDER, SCHEME, VISC, VO><<<grid_x<ORDER, VISC, VO>(g), blk, 0, C.stream>>>(a[0]);
  prof_mark(C, "sweep_x");
  k_sweep_line<R, S, ORDER, SCHEME, VISC, VO, 1, EP_RHS><<<grid_y(g, a[1].lseg), line_block(), 0, C.stream>>>(a[1]);
  prof_mark(C, "sweep_y");
  const dim3 gz = grid_z(g, a[2].lseg);
#define OX_Z(E) k_sweep_line<R, S, ORDER, SCHEME, VISC, VO, 2, E><<<gz, line_block(), 0, C.stream>>>(a[2])
  switch (sa.ep) {
    case EP_RK3_S1: OX_Z(EP_RK3_S1); break;
    case EP_RK3_S23: OX_Z(EP_RK3_S23); break;
    case EP_RK4_S: OX_Z(EP_RK4_S); break;
    case EP_RK4_S4: OX_Z(EP_RK4_S4); break;
    default: OX_Z(EP_RHS); break;
  }
#undef OX_Z
  prof_mark(C, "sweep_z_rk");
  if (sa.ep != EP_RHS && any_bc(g)) {
    // the case's set_bc on the new state, as ghost primitives (and sensors)
    const Prim<S> Pn = prim_set<S>(C, 1 - C.cur);
    launch_bc_prims<R, S, ORDER, SCHEME, VISC>(g, Pn, C.stream);
    launch_bc_sensors<R, S, ORDER, SCHEME>(g, Pn, C.stream);
    prof_mark(C, "bc");
  }
  finish_stage(C, sa);
}

// One stage of a group of cases: three persistent launches (LES: the SGS
// viscosity and its two ghost-copy passes first, as in prepare_stage).
template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
void run_stage_batch(const BatchGroupLaunch& L) {
  const dim3 blk(64, 4, 1);
  using A = SweepArgs<R, S>;
  if constexpr (VISC == 2) {
    k_les_mut_batch<R, S><<<L.grid_les[0], blk, 0, L.stream>>>((const A*)L.args[0], L.map_les[0], L.ncase,
                                                                L.ntiles_les[0]);
    for (int pass = 0; pass < 2; ++pass)
      k_les_bc_batch<R, S><<<L.grid_les[1 + pass], 256, 0, L.stream>>>((const A*)L.args[0], L.map_les[1 + pass],
                                                                        L.ncase, L.ntiles_les[1 + pass], pass);
  }
  k_sweep_x_batch<R, S, ORDER, SCHEME, VISC, VO>
      <<<L.grid[0], blk, 0, L.stream>>>((const A*)L.args[0], L.map[0], L.ncase, L.ntiles[0]);
  k_sweep_line_batch<R, S, ORDER, SCHEME, VISC, VO, 1, EP_RHS>
      <<<L.grid[1], line_block(), 0, L.stream>>>((const A*)L.args[1], L.map[1], L.ncase, L.ntiles[1]);
#define OX_ZB(E)                                         \
  k_sweep_line_batch<R, S, ORDER, SCHEME, VISC, VO, 2, E> \
      <<<L.grid[2], line_block(), 0, L.stream>>>((const A*)L.args[2], L.map[2], L.ncase, L.ntiles[2])
  switch (L.ep) {
    case EP_RK3_S1: OX_ZB(EP_RK3_S1); break;
    case EP_RK3_S23: OX_ZB(EP_RK3_S23); break;
    case EP_RK4_S: OX_ZB(EP_RK4_S); break;
    case EP_RK4_S4: OX_ZB(EP_RK4_S4); break;
    default: OX_ZB(EP_RHS); break;
  }
#undef OX_ZB
}

template <class R, class S>
void make_sweep_args_erased(Core& C, const StageArgs& sa, double cx, double cy, double cz, void* out3) {
  make_sweep_args<R, S>(C, sa, cx, cy, cz, (SweepArgs<R, S>*)out3);
}

template <class R, class S, int ORDER, int SCHEME, int VISC>
void bc_after_erased(const Geom& g, void* const* pn, hipStream_t st) {
  Prim<S> P;
  S** a[kPrimArrays] = {&P.rho, &P.u, &P.v, &P.w, &P.p, &P.s, &P.mut, &P.qc2};
  for (int m = 0; m < kPrimArrays; ++m) *a[m] = (S*)pn[m];
  launch_bc_prims<R, S, ORDER, SCHEME, VISC>(g, P, st);
  launch_bc_sensors<R, S, ORDER, SCHEME>(g, P, st);
}

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO> BatchOps batch_ops() {
  BatchOps o;
  o.args_bytes = sizeof(SweepArgs<R, S>);
  o.xt = xtile<ORDER, VISC, VO>();
  o.prepare = &prepare_stage<R, S, ORDER, SCHEME, VISC, VO>;
  o.make_args = &make_sweep_args_erased<R, S>;
  o.launch = &run_stage_batch<R, S, ORDER, SCHEME, VISC, VO>;
  o.bc_after = &bc_after_erased<R, S, ORDER, SCHEME, VISC>;
  return o;
}

// Supported (ORDER, SCHEME, VISC, VO) combinations.  VO is ignored for
// Euler (pass 0); VO=6 needs ORDER=6 (three ghost layers).
#define OX_FOR_EACH_STAGE(X) \
  X(2, KEEP, 0, 0)           \
  X(4, KEEP, 0, 0)           \
  X(6, KEEP, 0, 0)           \
  X(4, KEEP, 1, 4)           \
  X(6, KEEP, 1, 4)           \
  X(6, KEEP, 1, 6)           \
  X(2, SLAU, 0, 0)           \
  X(4, SLAU, 0, 0)           \
  X(6, SLAU, 0, 0)           \
  X(4, SLAU, 1, 4)           \
  X(6, SLAU, 1, 4)           \
  X(6, SLAU, 1, 6)           \
  X(2, ROE, 0, 0)            \
  X(4, ROE, 0, 0)            \
  X(6, ROE, 0, 0)            \
  X(4, ROE, 1, 4)            \
  X(6, ROE, 1, 6)            \
  X(2, HYBRID, 0, 0)         \
  X(4, HYBRID, 0, 0)         \
  X(6, HYBRID, 0, 0)         \
  X(4, HYBRID, 1, 4)         \
  X(6, HYBRID, 1, 4)         \
  X(6, HYBRID, 1, 6)         \
  X(6, KEEP, 1, 5)           \
  X(6, SLAU, 1, 5)           \
  X(6, HYBRID, 1, 5)
// ... and with the LES model.  VO 5 is ME4Base.
#define OX_FOR_EACH_LES_STAGE(X) \
  X(4, KEEP, 2, 4)               \
  X(6, KEEP, 2, 6)               \
  X(6, SLAU, 2, 6)               \
  X(4, HYBRID, 2, 4)             \
  X(6, HYBRID, 2, 6)             \
  X(6, SLAU, 2, 5)               \
  X(6, HYBRID, 2, 5)

// Table bodies: R_ and S_ name the arithmetic and storage types.
#define OX_STAGE_CASE(O, SC, V, VO) \
  case O * 1000 + SC * 100 + V * 10 + VO: return &run_stage<R_, S_, O, SC, V, VO>;
#define OX_BATCH_CASE(O, SC, V, VO)           \
  case O * 1000 + SC * 100 + V * 10 + VO:     \
    *ops = batch_ops<R_, S_, O, SC, V, VO>(); \
    return true;

}  // namespace ox
