// ox2d_kernels.hpp -- the gfx950 kernel of the two-dimensional path: one
// fused launch per RK stage over 2-D tiles of one or many cases.
//
// The reference evaluates a 2-D stage in five launches
// (2D_solver/src/calc_time_dev.f90.fypp:77-94): primitives
// (calc_quantities_2D, src/calc_physical_quantities.f90:7-29), the Ducros
// sensor (calc_hybrid.f90:11-52), the x and y face fluxes with the boundary
// order fallback (calc_slau_kernel.f90.fypp:17-200), the RK update
// (calc_steps.f90:31-54, :89-117) and set_bc (OS/set.f90:48-86), with the
// face fluxes E, F and the primitives round-tripping through HBM.  Here one
// workgroup takes a tile of TX x TY interior cells and does all of it:
//
//   1. loads the state of the tile plus a halo of HL points into LDS as
//      primitives (rho, u, v, p), recovered in FP64 and stored in the
//      arithmetic type R;
//   2. the Ducros sensor on the cross-shaped set of points the tile's faces
//      read (boundary points take their inner neighbour's value);
//   3. the x faces and y faces of the tile into LDS, each with the order its
//      distance to the boundary allows (6-, 4- or 2-point);
//   4. per cell: R = cx dE + cy dF in FP64 (the x part as soon as the x
//      faces are done, kept in registers while the y faces reuse their
//      LDS), the RK combination in FP64, the new state and -- for cells next
//      to the boundary -- the boundary points of the oblique-shock
//      conditions, from the new values.
//
// Stages read one state buffer and write another, so tiles never race on
// halo points (three buffers per case for RK3).  The FP64 form follows the
// reference's operation order and matches the FP64 oracle bit for bit; the
// FP32-bulk form computes primitives and fluxes' arithmetic in FP32 after an
// FP64 primitive recovery and keeps the residual and the state in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "ox_physics.hpp"

namespace ox2 {

using ox::PhysC;

// Stage epilogues of RK3 (calc_steps.f90):
enum Ep2 : int {
  EP2_S1 = 0,  // calc_step1:    Qo = Qs - R
  EP2_S2 = 1,  // calc_step2_3:  Qo = (c1*Qc + c2*Qs - R) * c4   (Qc = Q^n)
  EP2_S3 = 2,  // calc_step2_3:  Qo = (c1*Qs + c2*Qc - R) * c4   (Qo = Qc = Q^n, in place)
};

// One case for one stage.
template <class R> struct Args2 {
  const double* Qs;  // state the right-hand side is evaluated on (read with halos)
  const double* Qc;  // Q^n of the RK combination (stages 2, 3)
  double* Qo;        // state written (interior and boundary points)
  int nx, ny;        // points incl. the boundary layer
  int nbx;           // tiles along x
  int tile0;         // first tile of this case in the launch
  int tvd, variant;  // MUSCL kind, SLAU variant
  int split;         // top boundary: points i < split (1-based) take the inflow state
  int pad;
  double J;          // Jacobian (uniform grid)
  double g1;         // gamma - 1 (FP64 pressure recovery in both forms)
  double cx, cy;     // coef * dt * dy (x faces), coef * dt * dx (y faces)
  double c1, c2, c4;
  double qin[4], qpost[4];  // boundary values of the inflow and post-shock states (QJ units)
  R sx, sy;          // Ducros metric 0.25 * (1/dx + 1/dx), 0.25 * (1/dy + 1/dy)
  PhysC<R> c;
};

constexpr int TX = 64;
// Rows per tile.  11 FP32 rows: 715 x faces, 768 y faces and 704 cells
// keep the 256 threads busy in three passes each, and 34 KiB of LDS (107
// VGPRs) lets four workgroups share a CU.  4 FP64 rows: 35 KiB of LDS and
// 124 VGPRs, also four per CU.
#ifndef OX2_ROWS_FP64
#define OX2_ROWS_FP64 4
#endif
#ifndef OX2_ROWS_FP32
#define OX2_ROWS_FP32 11
#endif
template <class R> constexpr int tile_rows() { return sizeof(R) == 8 ? OX2_ROWS_FP64 : OX2_ROWS_FP32; }
// halo of the LDS tile: the stencil half-width, at least 2 where the sensor
// (2nd-order central differences at the points next to the faces) is needed
template <int ORDER> constexpr int halo() { return ORDER == 6 ? 3 : (ORDER == 4 ? 2 : 1); }

template <class R, int ORDER> struct Tile2 {
  static constexpr int TY = tile_rows<R>();
  static constexpr int HL = halo<ORDER>();
  static constexpr int W = TX + 2 * HL, H = TY + 2 * HL;
  R prim[4][H][W];               // rho, u, v, p
  R sens[TY + 2][TX + 2];        // Ducros sensor, points x0-1..x0+TX, y0-1..y0+TY
  // face fluxes: first the x faces x0-1..x0+TX-1 of the tile rows, then --
  // once each cell holds its x differences in registers -- the y faces
  // y0-1..y0+TY-1 of the tile columns in the same storage
  union {
    R E[4][TY][TX + 1];
    R F[4][TY + 1][TX];
  } fl;
};

// The Ducros sensor of one interior point from the velocity neighbours
// (calc_hybrid.f90:24-37); pu, pv point at (i, j) in the LDS planes.
template <class R, int W>
__device__ __forceinline__ R ducros2(const R* pu, const R* pv, R sx, R sy) {
  const R dudx = (-pu[-1] + pu[1]) * sx;
  const R dvdx = (-pv[-1] + pv[1]) * sx;
  const R dudy = (-pu[-W] + pu[W]) * sy;
  const R dvdy = (-pv[-W] + pv[W]) * sy;
  const R div = dudx + dvdy;
  const R rot = dvdx - dudy;
  const R fd = (div * div) / (div * div + rot * rot + R(1.0e-12));
  return ox::fmin2(R(1), fd);
}

// The face flux of the 2-D scheme from reconstructed states (rho, u, v, p).
// SLAU: calc_slau_2d.f90 = the 3-D routine with w = 0.  Roe: the 3-D routine
// (calc_roe_3d.f90:1-88) with w = 0 in the local frame (normal, t1, t2) with
// t1, t2 cyclic after the normal: (u, v, w) for x, (v, w, u) for y.
template <class R, int SCHEME, int DIR>
__device__ __forceinline__ void flux2(const PhysC<R>& c, int variant, const R* L, const R* Rr, R hr, R* F) {
  if constexpr (SCHEME == OX_SCHEME_SLAU) {
    R F5[5];
    ox::slau<R, DIR>(c, variant, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], R(0), R(0), L[3], Rr[3], hr, F5);
    F[0] = F5[0];
    F[1] = F5[1];
    F[2] = F5[2];
    F[3] = F5[4];
  } else {
    // roe_local2: the 3-D routine's values with the zero component's terms
    // removed (normal u, tangential v for x faces; normal v, tangential u for y)
    R o[4];
    if constexpr (DIR == 0) {
      ox::roe_local2<R>(c, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], L[3], Rr[3], o);
      F[0] = o[0];
      F[1] = o[1];
      F[2] = o[2];
      F[3] = o[3];
    } else {
      ox::roe_local2<R>(c, L[0], Rr[0], L[2], Rr[2], L[1], Rr[1], L[3], Rr[3], o);
      F[0] = o[0];
      F[1] = o[2];
      F[2] = o[1];
      F[3] = o[3];
    }
  }
}

// MUSCL states of the four primitives in one form (delta4 / delta6 of
// src/calc_muscl.f90:142-169 with the limiter switch taken once per face).
template <class R, int N, bool LIM>
__device__ __forceinline__ void recon4(const PhysC<R>& c, const R (&st)[4][6], R* L, R* Rr) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    if constexpr (N == 6) ox::delta6_form<R, LIM>(c, &st[v][0], L[v], Rr[v]);
    else ox::delta4_form<R, LIM>(c, &st[v][1], L[v], Rr[v]);
  }
}

// One face: f = Fortran face index along the direction (1..n-1), between the
// points f and f+1; st[v][m] = primitive v at point f-2+m; fs = face sensor.
// The order falls back next to the boundaries as in
// calc_slau_kernel.f90.fypp:45-95: 6-point for 3 <= f <= n-3, 4-point for
// 2 <= f <= n-2, first order otherwise (HR = 1 there).
template <class R, int ORDER, int SCHEME, int DIR>
__device__ __forceinline__ void face2(const Args2<R>& a, int f, int n, const R (&st)[4][6], R fs, R* F) {
  R L[4], Rr[4], hr;
  const bool six = ORDER == 6 && 3 <= f && f <= n - 3;
  const bool four = !six && ORDER >= 4 && 2 <= f && f <= n - 2;
  if (ORDER == 6 && six) {
    if (ox::muscl_limited(a.c, a.tvd, fs)) recon4<R, 6, true>(a.c, st, L, Rr);
    else recon4<R, 6, false>(a.c, st, L, Rr);
    hr = ox::wiggle(&st[3][1]);
  } else if (ORDER >= 4 && four) {
    if (ox::muscl_limited(a.c, a.tvd, fs)) recon4<R, 4, true>(a.c, st, L, Rr);
    else recon4<R, 4, false>(a.c, st, L, Rr);
    hr = ox::wiggle(&st[3][1]);
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      L[v] = st[v][2];
      Rr[v] = st[v][3];
    }
    hr = R(1);
  }
  flux2<R, SCHEME, DIR>(a.c, a.variant, L, Rr, hr, F);
}

// Conserved state of one point of a 2-D case (Fortran order (nx,4,ny)).
// The state pointers come from a descriptor in memory, so the compiler cannot
// tell that they point to global memory and would use flat loads and stores,
// which also count on the LDS wait counter: every LDS wait of the tile would
// then wait for the outstanding state stores too.  They are cast to the global
// address space here.
using gdouble = __attribute__((address_space(1))) double;
__device__ __forceinline__ void load_q(const double* __restrict__ Q, int nx, int ny, int i, int j, double* q) {
  const gdouble* G = (const gdouble*)Q;
  const size_t b = (size_t)i + (size_t)nx * 4 * (size_t)j;
#pragma unroll
  for (int l = 0; l < 4; ++l) q[l] = G[b + (size_t)l * nx];
}
__device__ __forceinline__ void store_q(double* __restrict__ Q, int nx, int i, int j, const double* q) {
  gdouble* G = (gdouble*)Q;
  const size_t b = (size_t)i + (size_t)nx * 4 * (size_t)j;
#pragma unroll
  for (int l = 0; l < 4; ++l) G[b + (size_t)l * nx] = q[l];
}

// One tile: interior cells x0..x0+TX-1, y0..y0+TY-1 (0-based point indices;
// interior points are 1..nx-2, 1..ny-2).
template <class R, int ORDER, int SCHEME, int EP>
__device__ __forceinline__ void stage_tile(const Args2<R>& a, int tx, int ty, Tile2<R, ORDER>& T) {
  using TL = Tile2<R, ORDER>;
  constexpr int TY = TL::TY, HL = TL::HL, W = TL::W, H = TL::H;
  constexpr int M0 = 3 - ORDER / 2, M1 = 2 + ORDER / 2;  // stencil points f-2+m read, m in [M0, M1]
  const int tid = threadIdx.x;
  const int nx = a.nx, ny = a.ny;
  const int x0 = 1 + tx * TX, y0 = 1 + ty * TY;
  const int xlast = min(x0 + TX - 1, nx - 2), ylast = min(y0 + TY - 1, ny - 2);

  // 1. primitives of the tile and its halo (calc_quantities_2D), recovered in
  //    FP64; points outside the case are zero and never read by a face
  for (int k = tid; k < W * H; k += 256) {
    const int rx = k % W, ry = k / W;
    const int i = x0 - HL + rx, j = y0 - HL + ry;
    R pr[4] = {R(0), R(0), R(0), R(0)};
    if (i >= 0 && j >= 0 && i < nx && j < ny) {
      double q[4];
      load_q(a.Qs, nx, ny, i, j, q);
      const double over_Q1 = 1.0 / q[0];
      const double rho = a.J * q[0];
      const double u = q[1] * over_Q1;
      const double v = q[2] * over_Q1;
      const double p = a.g1 * (a.J * q[3] - 0.5 * rho * (u * u + v * v));
      pr[0] = (R)rho;
      pr[1] = (R)u;
      pr[2] = (R)v;
      pr[3] = (R)p;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) T.prim[v][ry][rx] = pr[v];
  }
  __syncthreads();

  // 2. Ducros sensor at the points x0-1..x0+TX of the tile rows and
  //    y0-1..y0+TY of the tile columns; a boundary point takes the value of
  //    its inner neighbour (calc_hybrid.f90:39-50)
  const bool sensed = ORDER >= 4 && a.tvd == OX_TVD_HYBRID;
  if (sensed) {
    for (int k = tid; k < (TX + 2) * (TY + 2); k += 256) {
      const int sx = k % (TX + 2), sy = k / (TX + 2);
      if ((sx == 0 || sx == TX + 1) && (sy == 0 || sy == TY + 1)) continue;
      const int i = x0 - 1 + sx, j = y0 - 1 + sy;
      if (i > nx - 1 || j > ny - 1) continue;
      const int ci = min(max(i, 1), nx - 2), cj = min(max(j, 1), ny - 2);
      const int lx = ci - (x0 - HL), ly = cj - (y0 - HL);
      T.sens[sy][sx] = ducros2<R, W>(&T.prim[1][ly][lx], &T.prim[2][ly][lx], a.sx, a.sy);
    }
    __syncthreads();
  }

  // 3a. x faces f = x0-1..xlast (0-based: between points f and f+1) of rows y0..ylast
  for (int k = tid; k < (TX + 1) * TY; k += 256) {
    const int fx = k % (TX + 1), fy = k / (TX + 1);
    const int f = x0 - 1 + fx, j = y0 + fy;
    if (f > xlast || j > ylast) continue;
    R st[4][6] = {};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      if (m < M0 || m > M1) continue;
      const int lx = f - 2 + m - (x0 - HL);
#pragma unroll
      for (int v = 0; v < 4; ++v) st[v][m] = T.prim[v][fy + HL][lx];
    }
    const R fs = sensed ? R(0.5) * (T.sens[fy + 1][fx] + T.sens[fy + 1][fx + 1]) : R(0);
    R Fl[4];
    face2<R, ORDER, SCHEME, 0>(a, f + 1, nx, st, fs, Fl);
#pragma unroll
    for (int l = 0; l < 4; ++l) T.fl.E[l][fy][fx] = Fl[l];
  }
  __syncthreads();

  // 4a. x part of the residual, R = cx * (-E_i + E_i+1) (calc_R,
  //     calc_steps.f90:17-21), in registers: cell k = tid + 256 m
  constexpr int NC = (TX * TY + 255) / 256;
  double Rc[NC][4];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int k = tid + 256 * m;
    const int rx = k % TX, ry = k / TX;
    if (k < TX * TY && x0 + rx <= xlast && y0 + ry <= ylast) {
#pragma unroll
      for (int l = 0; l < 4; ++l) Rc[m][l] = a.cx * (-(double)T.fl.E[l][ry][rx] + (double)T.fl.E[l][ry][rx + 1]);
    }
  }
  __syncthreads();

  // 3b. y faces g = y0-1..ylast of columns x0..xlast
  for (int k = tid; k < TX * (TY + 1); k += 256) {
    const int gx = k % TX, gy = k / TX;
    const int i = x0 + gx, g = y0 - 1 + gy;
    if (i > xlast || g > ylast) continue;
    R st[4][6] = {};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      if (m < M0 || m > M1) continue;
      const int ly = g - 2 + m - (y0 - HL);
#pragma unroll
      for (int v = 0; v < 4; ++v) st[v][m] = T.prim[v][ly][gx + HL];
    }
    const R fs = sensed ? R(0.5) * (T.sens[gy][gx + 1] + T.sens[gy + 1][gx + 1]) : R(0);
    R Fl[4];
    face2<R, ORDER, SCHEME, 1>(a, g + 1, ny, st, fs, Fl);
#pragma unroll
    for (int l = 0; l < 4; ++l) T.fl.F[l][gy][gx] = Fl[l];
  }
  __syncthreads();

  // 4b. y part of the residual (calc_steps.f90:23-27), RK combination,
  //     boundary points
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int k = tid + 256 * m;
    const int rx = k % TX, ry = k / TX;
    const int i = x0 + rx, j = y0 + ry;
    if (k >= TX * TY || i > xlast || j > ylast) continue;
    double Rr[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) Rr[l] = Rc[m][l] + a.cy * (-(double)T.fl.F[l][ry][rx] + (double)T.fl.F[l][ry + 1][rx]);
    double qs[4], qn[4];
    load_q(a.Qs, nx, ny, i, j, qs);
    if constexpr (EP == EP2_S1) {
#pragma unroll
      for (int l = 0; l < 4; ++l) qn[l] = qs[l] - Rr[l];
    } else {
      double qc[4];
      load_q(a.Qc, nx, ny, i, j, qc);
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        if constexpr (EP == EP2_S2) qn[l] = (a.c1 * qc[l] + a.c2 * qs[l] - Rr[l]) * a.c4;
        else qn[l] = (a.c1 * qs[l] + a.c2 * qc[l] - Rr[l]) * a.c4;
      }
    }
    store_q(a.Qo, nx, i, j, qn);
    // oblique-shock boundary points next to this cell (OS/set.f90:48-86)
    const double slip[4] = {qn[0], qn[1], -qn[2], qn[3]};
    if (j == 1) {
      store_q(a.Qo, nx, i, 0, slip);                     // slip wall
      if (i == nx - 2) store_q(a.Qo, nx, nx - 1, 0, slip);
    }
    if (i == nx - 2) store_q(a.Qo, nx, nx - 1, j, qn);  // extrapolation
    if (i == 1) {                                         // inflow
      store_q(a.Qo, nx, 0, j, a.qin);
      if (j == 1) store_q(a.Qo, nx, 0, 0, a.qin);
      if (j == ny - 2) store_q(a.Qo, nx, 0, ny - 1, a.qin);
    }
    if (j == ny - 2) {                                    // top: inflow / post-shock
      store_q(a.Qo, nx, i, ny - 1, (i + 1 < a.split) ? a.qin : a.qpost);
      if (i == nx - 2) store_q(a.Qo, nx, nx - 1, ny - 1, (nx - 1 < a.split) ? a.qin : a.qpost);
    }
  }
}

// Tile t of the launch -> case (binary search over the first tiles).
template <class R>
__device__ __forceinline__ int find_case2(const Args2<R>* __restrict__ args, int ncase, int t) {
  int lo = 0, hi = ncase - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (args[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One RK stage over the tiles of `ncase` cases; persistent: each workgroup
// strides over the tiles.  Workgroups are dealt to the eight XCDs round
// robin; the logical index groups consecutive tiles (which share halo rows)
// on one XCD's L2 when the grid is a multiple of 8.
template <class R, int ORDER, int SCHEME, int EP>
__global__ __launch_bounds__(256) void k2d_stage(const Args2<R>* __restrict__ args, int ncase, int ntiles) {
  __shared__ Tile2<R, ORDER> T;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int L = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  for (int t = L; t < ntiles; t += G) {
    const int ci = find_case2(args, ncase, t);
    const Args2<R>& a = args[ci];
    const int local = t - a.tile0;
    stage_tile<R, ORDER, SCHEME, EP>(a, local % a.nbx, local / a.nbx, T);
    __syncthreads();  // the next tile overwrites the LDS
  }
}

}  // namespace ox2
