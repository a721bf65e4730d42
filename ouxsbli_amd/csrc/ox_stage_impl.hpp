// ox_stage_impl.hpp -- stage launcher templates and the specialisation table.
//
// The reference specialises each case at build time with fypp
// (src/mod_constant.f90.fypp:37-100: VISC, ORDER, SCHEME encoded as kinds);
// here (ORDER, SCHEME, VISC, VISC_ORDER) are template parameters and every
// supported combination is instantiated once.  TVD kind and SLAU variant are
// uniform runtime switches.
#pragma once

#include "ox_core.hpp"
#include "ox_kernels.hpp"

namespace ox {

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

template <class S> Prim<S> prim_set(Core& C, int set) {
  Prim<S> P;
  S** a[10] = {&P.rho, &P.u, &P.v, &P.w, &P.p, &P.T, &P.mu, &P.s, &P.mut, &P.qc2};
  for (int m = 0; m < 10; ++m) *a[m] = (S*)C.prim[set][m];
  return P;
}

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
void run_stage(Core& C, const double* Qsrc, const StageArgs& sa, double cx, double cy, double cz) {
  const Geom& g = C.g;
  const Prim<S> P = prim_set<S>(C, C.cur);
  const Prim<S> Pn = prim_set<S>(C, 1 - C.cur);
  const PhysC<double> c64 = make_physc<double>(C.hc);
  const PhysC<R> cr = make_physc<R>(C.hc);
  const dim3 blk(64, 4, 1);
  if (C.prim_for != Qsrc) {
    k_prim<S, (VISC != 0)><<<dim3(cdiv(g.nx, 64), cdiv(g.ny, 4), g.nz), blk, 0, C.stream>>>(Qsrc, P, g, c64);
    C.prim_for = Qsrc;
    prof_mark(C, "prim");
  }
  if constexpr (VISC == 2) {
    k_les_mut<S, R><<<dim3(cdiv(g.nx - 2, 64), cdiv(g.ny - 2, 4), g.nz - 2), blk, 0, C.stream>>>(P, g, cr);
    const long long n0 = (long long)(g.ny - 2) * (g.nz - 2) + (long long)(g.nx - 2) * (g.nz - 2) + (g.nz - 2);
    k_les_bc<S><<<(unsigned)((n0 + 255) / 256), 256, 0, C.stream>>>(P, g, 0);
    k_les_bc<S><<<(unsigned)(((long long)g.nx * g.ny + 255) / 256), 256, 0, C.stream>>>(P, g, 1);
    prof_mark(C, "les_mut");
  }
  k_sweep_x<R, S, ORDER, SCHEME, VISC, VO><<<dim3(cdiv(g.ix, xtile<ORDER, VISC, VO>()), cdiv(g.iy, XROWS), g.iz),
                                             blk, 0, C.stream>>>(P, (S*)C.Racc, g, cr, C.tvd, C.variant, cx);
  prof_mark(C, "sweep_x");
  k_sweep_line<R, S, ORDER, SCHEME, VISC, VO, 1, EP_RHS>
      <<<dim3(cdiv(g.ix, 64), cdiv(g.iz, 4), cdiv(g.iy, LSEG)), blk, 0, C.stream>>>(P, (S*)C.Racc, g, cr, C.tvd,
                                                                                    C.variant, cy, sa, Pn, c64);
  prof_mark(C, "sweep_y");
  const dim3 gz(cdiv(g.ix, 64), cdiv(g.iy, 4), cdiv(g.iz, LSEG));
#define OX_Z(E)                                                                                                 \
  k_sweep_line<R, S, ORDER, SCHEME, VISC, VO, 2, E><<<gz, blk, 0, C.stream>>>(P, (S*)C.Racc, g, cr, C.tvd, C.variant, \
                                                                              cz, sa, Pn, c64)
  switch (sa.ep) {
    case EP_RK3_S1: OX_Z(EP_RK3_S1); break;
    case EP_RK3_S23: OX_Z(EP_RK3_S23); break;
    case EP_RK4_S: OX_Z(EP_RK4_S); break;
    case EP_RK4_S4: OX_Z(EP_RK4_S4); break;
    default: OX_Z(EP_RHS); break;
  }
#undef OX_Z
  prof_mark(C, "sweep_z_rk");
  if (sa.ep != EP_RHS) {
    // the epilogue wrote the primitives of sa.Qout into the other set
    C.cur = 1 - C.cur;
    C.prim_for = sa.Qout;
  }
}

// Supported (ORDER, SCHEME, VISC, VO) combinations.  VO is ignored for
// Euler (pass 0); VO=6 needs ORDER=6 (three ghost layers).
#define OX_STAGE_TABLE(R, S)                                                  \
  switch (order * 1000 + scheme * 100 + visc * 10 + (visc ? vorder : 0)) {    \
    case 2000: return &run_stage<R, S, 2, KEEP, 0, 0>;                        \
    case 4000: return &run_stage<R, S, 4, KEEP, 0, 0>;                        \
    case 6000: return &run_stage<R, S, 6, KEEP, 0, 0>;                        \
    case 4014: return &run_stage<R, S, 4, KEEP, 1, 4>;                        \
    case 6014: return &run_stage<R, S, 6, KEEP, 1, 4>;                        \
    case 6016: return &run_stage<R, S, 6, KEEP, 1, 6>;                        \
    case 2100: return &run_stage<R, S, 2, SLAU, 0, 0>;                        \
    case 4100: return &run_stage<R, S, 4, SLAU, 0, 0>;                        \
    case 6100: return &run_stage<R, S, 6, SLAU, 0, 0>;                        \
    case 4114: return &run_stage<R, S, 4, SLAU, 1, 4>;                        \
    case 6114: return &run_stage<R, S, 6, SLAU, 1, 4>;                        \
    case 6116: return &run_stage<R, S, 6, SLAU, 1, 6>;                        \
    case 2200: return &run_stage<R, S, 2, ROE, 0, 0>;                         \
    case 4200: return &run_stage<R, S, 4, ROE, 0, 0>;                         \
    case 6200: return &run_stage<R, S, 6, ROE, 0, 0>;                         \
    case 4214: return &run_stage<R, S, 4, ROE, 1, 4>;                         \
    case 6216: return &run_stage<R, S, 6, ROE, 1, 6>;                         \
    case 2300: return &run_stage<R, S, 2, HYBRID, 0, 0>;                      \
    case 4300: return &run_stage<R, S, 4, HYBRID, 0, 0>;                      \
    case 6300: return &run_stage<R, S, 6, HYBRID, 0, 0>;                      \
    case 4314: return &run_stage<R, S, 4, HYBRID, 1, 4>;                      \
    case 6314: return &run_stage<R, S, 6, HYBRID, 1, 4>;                      \
    case 6316: return &run_stage<R, S, 6, HYBRID, 1, 6>;                      \
    case 4024: return &run_stage<R, S, 4, KEEP, 2, 4>;                        \
    case 6026: return &run_stage<R, S, 6, KEEP, 2, 6>;                        \
    case 6126: return &run_stage<R, S, 6, SLAU, 2, 6>;                        \
    case 4324: return &run_stage<R, S, 4, HYBRID, 2, 4>;                      \
    case 6326: return &run_stage<R, S, 6, HYBRID, 2, 6>;                      \
    default: return nullptr;                                                  \
  }

}  // namespace ox
