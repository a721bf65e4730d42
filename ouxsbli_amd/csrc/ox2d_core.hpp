// ox2d_core.hpp -- host side of the 2-D path shared by the C ABI
// (ox2d_capi.hip), the batch engine (ox_batch.hip) and the kernel
// instantiation units (ox2d_fp64.hip, ox2d_fp32.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "ox2d_kernels.hpp"
#include "ox_types.hpp"

namespace ox2 {

// A stage launcher of one kernel specialisation and the number of its
// workgroups the device holds at once (the persistent grid's cap).
struct Launch2 {
  void (*launch)(const void* args, int ncase, int ntiles, int grid, hipStream_t stream);
  int (*max_blocks)();
};

// Specialisation key of a 2-D case: (ORDER, SCHEME).
inline int key2(int order, int scheme) { return order * 10 + scheme; }

bool find_launch2_fp64(int order, int scheme, Launch2 out[3]);
bool find_launch2_fp32(int order, int scheme, Launch2 out[3]);

// Grid metrics and boundary values of one case, computed once on the host
// in the reference's operation order: set_grid (OS/set.f90:8-26),
// set_Jacobian_xy2 (src/set_coordinate.f90:78-96), pre_calc
// (2D_solver/src/preprocess.f90.fypp:93-102), the boundary states of
// OS/set.f90:55-65 divided by the Jacobian as 1/J products.
struct Geom2 {
  double J, dtdx, dtdy, sx, sy, g1;
  double qin[4], qpost[4];
  ox::HostConsts hc;
  int nbx, nby;

  void init(const ox2d_config& c, int tile_rows) {
    ox_config c3{};
    c3.nx = c3.ny = c3.nz = 8;
    c3.order = 2;
    c3.Lx = c3.Ly = c3.Lz = 1.0;
    c3.gamma = c.gamma;
    c3.R = c.R;
    c3.Pr = 0.72;
    c3.Prt = 0.9;
    c3.threshold = c.threshold;
    hc.init(c3);
    const double dx0 = c.Lx / (double)(c.nx - 1), dy0 = c.Ly / (double)(c.ny - 1);
    J = 4.0 / ((dx0 + dx0) * (dy0 + dy0));
    const double xix = 1.0 / dx0, etay = 1.0 / dy0;
    sx = 0.25 * (xix + xix);
    sy = 0.25 * (etay + etay);
    dtdx = 0.5 * c.dt * (dx0 + dx0);
    dtdy = 0.5 * c.dt * (dy0 + dy0);
    g1 = hc.gamma_1;
    const double Jt = 1.0 / J;
    bc_state(c.inflow, Jt, qin);
    bc_state(c.post, Jt, qpost);
    nbx = (c.nx - 2 + TX - 1) / TX;
    nby = (c.ny - 2 + tile_rows - 1) / tile_rows;
  }
  void bc_state(const double* s, double Jt, double* q) const {
    q[0] = s[0] * Jt;
    q[1] = s[0] * s[1] * Jt;
    q[2] = s[0] * s[2] * Jt;
    q[3] = (s[3] * hc.over_gamma_1 + 0.5 * s[0] * (s[1] * s[1] + s[2] * s[2])) * Jt;
  }
};

inline int tile_rows_of(int precision) { return precision == OX_PREC_FP64 ? tile_rows<double>() : tile_rows<float>(); }

// RK3 stage st (0-based) on the buffers A (= Q^n), B, C: source, Q^n,
// output and the coefficients of calc_time_dev.f90.fypp:79-93.
struct Stage2 {
  int src, cen, out;  // buffer indices (cen < 0: unused)
  double coef, c1, c2, c4;
};
inline Stage2 rk3_stage(int st, double one_third) {
  switch (st) {
    case 0: return {0, -1, 1, 1.0, 0.0, 0.0, 1.0};
    case 1: return {1, 0, 2, 0.25, 0.75, 0.25, 1.0};
    default: return {2, 0, 0, 2.0, 2.0, 1.0, one_third};
  }
}

// Descriptor of one case for RK3 stage st.
template <class R>
Args2<R> make_args2(const ox2d_config& c, const Geom2& g, double* const buf[3], int st, int tile0) {
  const Stage2 s = rk3_stage(st, g.hc.one_third);
  Args2<R> a{};
  a.Qs = buf[s.src];
  a.Qc = s.cen >= 0 ? buf[s.cen] : nullptr;
  a.Qo = buf[s.out];
  a.nx = c.nx;
  a.ny = c.ny;
  a.nbx = g.nbx;
  a.tile0 = tile0;
  a.tvd = c.tvd;
  a.variant = c.slau_variant;
  a.split = c.top_split;
  a.J = g.J;
  a.g1 = g.g1;
  a.cx = s.coef * g.dtdy;
  a.cy = s.coef * g.dtdx;
  a.c1 = s.c1;
  a.c2 = s.c2;
  a.c4 = s.c4;
  for (int l = 0; l < 4; ++l) {
    a.qin[l] = g.qin[l];
    a.qpost[l] = g.qpost[l];
  }
  a.sx = (R)g.sx;
  a.sy = (R)g.sy;
  a.c = ox::make_physc<R>(g.hc);
  return a;
}

inline size_t args2_bytes(int precision) {
  return precision == OX_PREC_FP64 ? sizeof(Args2<double>) : sizeof(Args2<float>);
}

// Writes the descriptor of stage st of one case in the case's precision.
inline void write_args2(const ox2d_config& c, const Geom2& g, double* const buf[3], int st, int tile0, void* out) {
  if (c.precision == OX_PREC_FP64) *(Args2<double>*)out = make_args2<double>(c, g, buf, st, tile0);
  else *(Args2<float>*)out = make_args2<float>(c, g, buf, st, tile0);
}

// Persistent grid: the tiles, rounded up to a multiple of 8 (whole XCD
// rounds), capped at what the device holds at once.
inline int grid2(int ntiles, int cap) {
  int g = ntiles > 8 ? (ntiles + 7) / 8 * 8 : ntiles;
  if (g > cap) g = cap / 8 * 8 > 0 ? cap / 8 * 8 : cap;
  return g < 1 ? 1 : g;
}

}  // namespace ox2

// The handle behind ox2d_solver.
struct ox2d_solver {
  ox2d_config cfg{};
  ox2::Geom2 g{};
  double* buf[3] = {nullptr, nullptr, nullptr};  // A = Q^n, B, C
  char* d_args = nullptr;                          // the three stage descriptors
  size_t args_bytes = 0;
  ox2::Launch2 launch[3]{};
  int ntiles = 0, grid = 0;
  hipStream_t stream = nullptr, own_stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
};

size_t solver2d_footprint(const ox2d_config& c);
