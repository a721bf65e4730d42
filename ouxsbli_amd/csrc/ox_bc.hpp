// ox_bc.hpp -- ghost fills of non-periodic directions (OX_BC_EXTRAPOLATE,
// OX_BC_SLIP) for the state, the primitives and the Ducros sensor.
//
// The reference fills the ghost layers of the state after every RK stage
// with the case's set_bc (3D_solver/STZ/set.f90:77-138: slip walls mirror
// the state about the boundary face and negate the normal momentum,
// zero-gradient ends copy the adjacent interior cell), then recomputes the
// primitives of every cell.  Here the stage epilogue writes the primitives of
// interior cells, and the fills below write those of the ghost cells: a
// copied or mirrored state gives the same primitives bit for bit as the
// interior cell it came from (the normal velocity negated, an exact
// operation), so the primitives are filled directly.  Directions are filled
// in the order x, y, z over the whole extent of the other two, as set_bc
// does; the fills of different directions commute.
#pragma once

#include "ox_kernels.hpp"

namespace ox {

// Source index along one direction of extent h + m + h (h ghost layers, m
// interior cells) for index p, and whether the normal component changes sign.
__device__ __forceinline__ int bc_source(int kind, int p, int h, int m, bool& flip) {
  flip = false;
  if (p >= h && p < h + m) return p;
  if (kind == OX_BC_PERIODIC) return p < h ? p + m : p - m;
  if (kind == OX_BC_EXTRAPOLATE) return p < h ? h : h + m - 1;
  flip = true;  // OX_BC_SLIP: ghost h-1-n mirrors interior h+n
  return p < h ? 2 * h - 1 - p : 2 * (h + m) - 1 - p;
}

// Cell (i, j, k) of the 2h ghost layers of direction d for linear index t
// over (2h) x (extent of the other two directions); false past the end.
__device__ __forceinline__ bool bc_cell(const Geom& g, int d, long long t, int& i, int& j, int& k) {
  const int h = g.h;
  if (d == 0) {
    if (t >= 2LL * h * g.ny * g.nz) return false;
    const int s = (int)(t % (2 * h));
    const long long r = t / (2 * h);
    i = s < h ? s : g.nx - 2 * h + s;
    j = (int)(r % g.ny);
    k = (int)(r / g.ny);
  } else if (d == 1) {
    if (t >= 2LL * h * g.nx * g.nz) return false;
    i = (int)(t % g.nx);
    const long long r = t / g.nx;
    const int s = (int)(r % (2 * h));
    j = s < h ? s : g.ny - 2 * h + s;
    k = (int)(r / (2 * h));
  } else {
    if (t >= 2LL * h * g.nx * g.ny) return false;
    i = (int)(t % g.nx);
    const long long r = t / g.nx;
    j = (int)(r % g.ny);
    const int s = (int)(r / g.ny);
    k = s < h ? s : g.nz - 2 * h + s;
  }
  return true;
}

inline long long bc_cells(const Geom& g, int d) {
  return 2LL * g.h * (d == 0 ? (long long)g.ny * g.nz : d == 1 ? (long long)g.nx * g.nz : (long long)g.nx * g.ny);
}

// The primitives' ghost layers of direction d.
template <class S>
__global__ __launch_bounds__(256) void k_bc_prim(Prim<S> P, Geom g, int d) {
  int i, j, k;
  if (!bc_cell(g, d, (long long)blockIdx.x * 256 + threadIdx.x, i, j, k)) return;
  bool flip;
  int si = i, sj = j, sk = k;
  if (d == 0) si = bc_source(g.bc[0], i, g.h, g.ix, flip);
  else if (d == 1) sj = bc_source(g.bc[1], j, g.h, g.iy, flip);
  else sk = bc_source(g.bc[2], k, g.h, g.iz, flip);
  const int m = cidx(g, i, j, k), sm = cidx(g, si, sj, sk);
  G(P.rho)[m] = G(P.rho)[sm];
  const S u = G(P.u)[sm], v = G(P.v)[sm], w = G(P.w)[sm];
  G(P.u)[m] = (flip && d == 0) ? -u : u;
  G(P.v)[m] = (flip && d == 1) ? -v : v;
  G(P.w)[m] = (flip && d == 2) ? -w : w;
  G(P.p)[m] = G(P.p)[sm];
}

// The Ducros sensor of the ghost layer next to the interior along a
// non-periodic direction d (1: y, 2: z) over the interior of the other two:
// the cells whose sensor the faces of interior cells read (calc_Ducros,
// 3D_solver/src/calc_hybrid.f90:11-72, computes it at every cell 2..n-1 from
// the filled primitives).  The x sweep forms the sensors of its own ghost
// cells from its row tile.
template <class S, class R>
__global__ __launch_bounds__(256) void k_bc_sensor(Prim<S> P, Geom g, int d) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nt = 2LL * g.ix * (d == 1 ? g.iz : g.iy);
  if (t >= nt) return;
  const int i = g.h + (int)(t % g.ix);
  const long long r = t / g.ix;
  const int side = (int)(r & 1);
  const int o = (int)(r >> 1);  // k for d 1, j for d 2
  int j, k;
  if (d == 1) {
    j = side ? g.h + g.iy : g.h - 1;
    k = g.h + o;
  } else {
    j = g.h + o;
    k = side ? g.h + g.iz : g.h - 1;
  }
  const int q = cidx(g, i, j, k), sy = g.nx, sz = g.nx * g.ny;
  const R fd = ducros<R>(g, R(G(P.u)[q - 1]), R(G(P.u)[q + 1]), R(G(P.v)[q - 1]), R(G(P.v)[q + 1]),
                         R(G(P.w)[q - 1]), R(G(P.w)[q + 1]), R(G(P.u)[q - sy]), R(G(P.u)[q + sy]),
                         R(G(P.v)[q - sy]), R(G(P.v)[q + sy]), R(G(P.w)[q - sy]), R(G(P.w)[q + sy]),
                         R(G(P.u)[q - sz]), R(G(P.u)[q + sz]), R(G(P.v)[q - sz]), R(G(P.v)[q + sz]),
                         R(G(P.w)[q - sz]), R(G(P.w)[q + sz]));
  G(P.s)[q] = (S)fd;
}

inline bool any_bc(const Geom& g) { return g.bc[0] || g.bc[1] || g.bc[2]; }

// After a stage whose epilogue wrote the interior primitives of its new state
// into Pn: the ghost primitives of the non-periodic directions, then the
// sensors of their inner ghost layers.
template <class R, class S, int ORDER, int SCHEME, int VISC>
void launch_bc_prims(const Geom& g, const Prim<S>& Pn, hipStream_t st) {
  for (int d = 0; d < 3; ++d)
    if (g.bc[d]) k_bc_prim<S><<<(unsigned)((bc_cells(g, d) + 255) / 256), 256, 0, st>>>(Pn, g, d);
}
template <class R, class S, int ORDER, int SCHEME>
void launch_bc_sensors(const Geom& g, const Prim<S>& P, hipStream_t st) {
  if constexpr (kUsesSensor<SCHEME, ORDER> && ORDER > 2) {
    for (int d = 1; d < 3; ++d)
      if (g.bc[d]) {
        const long long n = 2LL * g.ix * (d == 1 ? g.iz : g.iy);
        k_bc_sensor<S, R><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(P, g, d);
      }
  }
}

}  // namespace ox
