// ox_kernels.hpp -- gfx950 kernels of one RK stage (templated; instantiated
// by ox_fp64.hip / ox_fp32.hip).
//
// Data layout in HBM (all on one GPU, ghosts included):
//   state  QJ(nx,5,ny,nz) FP64, the reference layout (x fastest, then the
//          variable, then y, z): drop-in for set_state/get_state/checkpoints.
//   prims  rho,u,v,w,p,sensor as separate (nx,ny,nz) arrays of type S
//          (FP64 plain form, FP32 in the FP32-bulk form); T and mu are
//          formed from rho and p where a sweep reads a cell.
//   Racc   residual accumulator in the state layout, type S.
//
// One stage = k_sweep_x (LDS row tiles; Ducros sensor; writes R = c_x*dE; on
// periodic rows of 64, 128 or 256 cells k_sweep_xrow, whole rows in LDS) -> k_sweep_line<DIR=1>
// (register sliding window along y; R += c_y*dF) -> k_sweep_line<DIR=2>
// (sliding window along z; R += c_z*dG, then the RK combination, and the
// primitives of the new state for the next stage).  k_prim (QJ -> primitives)
// runs only after the state is uploaded, k_ghost_fill at the end of a call.  Face fluxes (convective minus viscous) live only in
// registers: E, F and G are never written to memory.  The summation order
// x, then y, then z is the order of calc_R (3D_solver/src/calc_steps.f90:8-38).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "ox_physics.hpp"
#include "ox_types.hpp"

namespace ox {

template <class S> struct Prim {
  S *rho, *u, *v, *w, *p, *s, *mut, *qc2;
};
constexpr int kPrimArrays = 8;

// The pointer seen as one into global memory.  Pointers a kernel receives as
// arguments are known global; those the batched kernels read from their
// descriptor arrays are not, and accesses through them would compile to flat
// loads and stores (addresses in VGPRs only, and counted on the LDS wait
// counter as well).  Every access to primitives, residual and state goes
// through G.
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T* G(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

constexpr int KEEP = OX_SCHEME_KEEP;
constexpr int SLAU = OX_SCHEME_SLAU;
constexpr int ROE = OX_SCHEME_ROE;
constexpr int HYBRID = OX_SCHEME_HYBRID;
// schemes that read the Ducros sensor (calc_flux_base.f90.fypp:56-58)
// (MUSCL reads it at ORDER 4/6; the hybrid switch at every order)
template <int SCHEME, int ORDER> constexpr bool kUsesSensor = SCHEME == HYBRID || (SCHEME != KEEP && ORDER > 2);
// schemes whose face flux reads T (KEEP's internal-energy term)
template <int SCHEME> constexpr bool kUsesT = SCHEME == KEEP || SCHEME == HYBRID;

// Viscous stencil codes VO: 4 and 6 are the central stencils of that order
// (calc_visc_cent.f90.fypp); 5 is ME4Base (calc_visc_me4_base.f90): a 6-cell
// face window with 4th-order cross derivatives.
template <int VO> constexpr int vwin() { return VO == 5 ? 6 : VO; }    // face window (cells)
template <int VO> constexpr int vcross() { return VO == 5 ? 4 : VO; }  // cross-derivative order

// Stencil width of a face: the convective stencil (ORDER cells) and the
// viscous one (vwin cells) are both centred on the face.
template <int ORDER, int VISC, int VO> struct Sten {
  static constexpr int SW = (VISC && vwin<VO>() > ORDER) ? vwin<VO>() : ORDER;
  static constexpr int HW = SW / 2;
  static constexpr int CO = HW - ORDER / 2;  // first convective cell in the window
  static constexpr int VOFF = VISC ? HW - vwin<VO>() / 2 : 0;
  static constexpr int VH = vcross<VO>() / 2;  // half width of the cross-derivative stencil
};

// state index of (i,l,j,k), 0-based: 64-bit, since the state holds five
// values per cell and a grid of up to 2^31 cells fits one GPU's HBM (the
// bracket, 5 * (j + ny * k) + l, stays below 5 * ny * nz)
__device__ __forceinline__ long long qidx(const Geom& g, int i, int l, int j, int k) {
  return i + (long long)g.nx * (l + 5 * (j + g.ny * k));
}
__device__ __forceinline__ int cidx(const Geom& g, int i, int j, int k) { return i + g.nx * (j + g.ny * k); }

// ---------------------------------------------------------------------------
// Primitive recovery.  calc_quantities_T_3D / calc_quantities_3D
// (src/calc_physical_quantities.f90:60-116), always in FP64.
// ---------------------------------------------------------------------------
// Primitives of one cell from its conserved values (J * Q = physical).
struct PrimVals {
  double rho, u, v, w, p;
};
__device__ __forceinline__ PrimVals prim_of(double q1, double q2, double q3, double q4, double q5, double J,
                                            const PhysC<double>& c) {
  PrimVals r;
  const double over_Q1 = 1.0 / q1;
  r.rho = J * q1;
  r.u = q2 * over_Q1;
  r.v = q3 * over_Q1;
  r.w = q4 * over_Q1;
  r.p = c.gamma_1 * (J * q5 - 0.5 * r.rho * (r.u * r.u + r.v * r.v + r.w * r.w));
  return r;
}
// T and mu are not stored: the sweeps form them from the stored rho and p
// where they read a cell (finish_cell), with the reference's expressions
// (calc_quantities_T_3D, src/calc_physical_quantities.f90:60-116), the same
// values bit for bit as forming them here in FP64.  That takes two arrays off
// the z sweep's writes and off every sweep's reads, for one division, one
// square root and one more division per window cell.
template <class S>
__device__ __forceinline__ void store_prim(const Prim<S>& P, int m, const PrimVals& r) {
  G(P.rho)[m] = (S)r.rho;
  G(P.u)[m] = (S)r.u;
  G(P.v)[m] = (S)r.v;
  G(P.w)[m] = (S)r.w;
  G(P.p)[m] = (S)r.p;
}

template <class S>
__global__ __launch_bounds__(256) void k_prim(const double* __restrict__ QJ, Prim<S> P, Geom g, PhysC<double> c) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const long long q = qidx(g, i, 0, j, k);
  const int st = g.nx;
  const PrimVals r = prim_of(QJ[q], QJ[q + st], QJ[q + 2 * st], QJ[q + 3 * st], QJ[q + 4 * st], g.jac, c);
  store_prim<S>(P, cidx(g, i, j, k), r);
}

// ---------------------------------------------------------------------------
// Selective mixed-scale SGS viscosity (calc_mut, 3D_solver/src/calc_les.f90:
// 12-125) over the cells the reference computes (Fortran 2..n-1).  Only the
// seven test-filtered values the central differences and qc2 read are formed
// (the centre and its six face neighbours); exponents 1/4 as sqrt(sqrt(.)).
// ---------------------------------------------------------------------------
template <class S, class R>
__device__ __forceinline__ R tfilt(const S* __restrict__ f, int m, int sy, int sz) {
  const R c0 = R(G(f)[m]);
  return c0 + R(0.25) * (R(G(f)[m - 1]) - R(2) * c0 + R(G(f)[m + 1]) + R(G(f)[m - sy]) - R(2) * c0 +
                         R(G(f)[m + sy]) + R(G(f)[m - sz]) - R(2) * c0 + R(G(f)[m + sz]));
}

template <class R>
__device__ __forceinline__ void grad9(R um, R up, R vm, R vp, R wm, R wp, R dx, R dy, R dz, const R* uy, const R* vy,
                                      const R* wy, const R* uz, const R* vz, const R* wz, R* g) {
  g[0] = R(0.5) * (-um + up) * dx;        // dudx
  g[1] = R(0.5) * (-uy[0] + uy[1]) * dy;  // dudy
  g[2] = R(0.5) * (-uz[0] + uz[1]) * dz;  // dudz
  g[3] = R(0.5) * (-vm + vp) * dx;
  g[4] = R(0.5) * (-vy[0] + vy[1]) * dy;
  g[5] = R(0.5) * (-vz[0] + vz[1]) * dz;
  g[6] = R(0.5) * (-wm + wp) * dx;
  g[7] = R(0.5) * (-wy[0] + wy[1]) * dy;
  g[8] = R(0.5) * (-wz[0] + wz[1]) * dz;
}

// nu_t and qc2 of cell (i, j, k), 1 <= i <= nx-2, 1 <= j <= ny-2, 1 <= k <= nz-2.
template <class S, class R>
__device__ __forceinline__ void les_mut_cell(const Prim<S>& P, const Geom& g, const PhysC<R>& c, int i, int j,
                                             int k) {
  const int sy = g.nx, sz = g.nx * g.ny;
  const int m = cidx(g, i, j, k);
  const bool filt = (i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3 && k >= 2 && k <= g.nz - 3);
  const S* U[3] = {P.u, P.v, P.w};
  R raw[3][7], fil[3][7];  // centre, -x, +x, -y, +y, -z, +z
  const int off[7] = {0, -1, 1, -sy, sy, -sz, sz};
  // The filter is formed at every stencil point whatever `filt` says, around
  // the centre itself where the cell is not filtered (its neighbours always
  // exist there; the value is not used), so that all loads sit in one block
  // the compiler can issue ahead of their use instead of one branch per point.
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      raw[f][t] = R(G(U[f])[m + off[t]]);
      const R ft = tfilt<S, R>(U[f], filt ? m + off[t] : m, sy, sz);
      fil[f][t] = filt ? ft : raw[f][t];
    }
  const R dx = R(g.inv_dx), dy = R(g.inv_dy), dz = R(g.inv_dz);
  R gr[9], gh[9];
  {
    const R uy[2] = {raw[0][3], raw[0][4]}, vy[2] = {raw[1][3], raw[1][4]}, wy[2] = {raw[2][3], raw[2][4]};
    const R uz[2] = {raw[0][5], raw[0][6]}, vz[2] = {raw[1][5], raw[1][6]}, wz[2] = {raw[2][5], raw[2][6]};
    grad9<R>(raw[0][1], raw[0][2], raw[1][1], raw[1][2], raw[2][1], raw[2][2], dx, dy, dz, uy, vy, wy, uz, vz, wz,
             gr);
  }
  {
    const R uy[2] = {fil[0][3], fil[0][4]}, vy[2] = {fil[1][3], fil[1][4]}, wy[2] = {fil[2][3], fil[2][4]};
    const R uz[2] = {fil[0][5], fil[0][6]}, vz[2] = {fil[1][5], fil[1][6]}, wz[2] = {fil[2][5], fil[2][6]};
    grad9<R>(fil[0][1], fil[0][2], fil[1][1], fil[1][2], fil[2][1], fil[2][2], dx, dy, dz, uy, vy, wy, uz, vz, wz,
             gh);
  }
  const R du = raw[0][0] - fil[0][0], dv = raw[1][0] - fil[1][0], dw = raw[2][0] - fil[2][0];
  const R q2 = R(0.5) * (du * du + dv * dv + dw * dw);
  // SMS (calc_les.f90:30-86)
  const R vor1 = gr[7] - gr[5], vor2_ = gr[2] - gr[6], vor3 = gr[3] - gr[1];
  const R vh1 = gh[7] - gh[5], vh2 = gh[2] - gh[6], vh3 = gh[3] - gh[1];
  const R d1 = vor1 - vh1, d2 = vor2_ - vh2, d3 = vor3 - vh3;
  const R vor2 = vor1 * vor1 + vor2_ * vor2_ + vor3 * vor3;
  const R vorh2 = vh1 * vh1 + vh2 * vh2 + vh3 * vh3;
  const R vdiff2 = d1 * d1 + d2 * d2 + d3 * d3;
  const R svor2 = R(2) * fsqrt(vor2) * fsqrt(vorh2) + vorh2 + vor2 - vdiff2;
  R tan2;
  if (fabs_(svor2) < R(1.0e-30)) tan2 = R(0);
  else tan2 = (R(2) * fsqrt(vorh2) * fsqrt(vor2) - vorh2 - vor2 + vdiff2) / svor2;
  const R rtheta = tan2 / c.tan10deg2;
  const R ftheta = fmin2(R(1), rtheta * rtheta);
  const R S2 = R(2) * (gr[0] * gr[0] + gr[4] * gr[4] + gr[8] * gr[8]) +
               ((gr[3] + gr[1]) * (gr[3] + gr[1]) + (gr[7] + gr[5]) * (gr[7] + gr[5]) +
                (gr[2] + gr[6]) * (gr[2] + gr[6]));
  const R nut = ftheta * R(0.06) * fsqrt(fsqrt(S2)) * fsqrt(fsqrt(q2)) * c.les_delta15;
  G(P.qc2)[m] = (S)q2;
  G(P.mut)[m] = (S)(R(G(P.rho)[m]) * nut);
}

template <class S, class R>
__global__ __launch_bounds__(256) void k_les_mut(Prim<S> P, Geom g, PhysC<R> c) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.nx - 2 || j > g.ny - 2) return;
  les_mut_cell<S, R>(P, g, c, i, j, k);
}

// set_bc_mut_common (3D_solver/src/set_bc_common.f90:296-335): pass 0 = x
// faces, y faces and xy corners over k = Fortran 2..nz-1; pass 1 = z planes.
// Item t of a pass (items past the pass's count do nothing); pass 1 reads
// what pass 0 wrote, so the two run as separate launches.
template <class S>
__device__ __forceinline__ void les_bc_item(const Prim<S>& P, const Geom& g, int pass, long long t) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  __attribute__((address_space(1))) S* A[2] = {G(P.mut), G(P.qc2)};
  if (pass == 0) {
    const int nk = nz - 2;
    const long long nxf = (long long)(ny - 2) * nk, nyf = (long long)(nx - 2) * nk;
    if (t < nxf) {
      const int j = 1 + (int)(t % (ny - 2)), k = 1 + (int)(t / (ny - 2));
      for (int a = 0; a < 2; ++a) {
        A[a][cidx(g, 0, j, k)] = A[a][cidx(g, nx - 2, j, k)];
        A[a][cidx(g, nx - 1, j, k)] = A[a][cidx(g, 1, j, k)];
      }
    } else if (t < nxf + nyf) {
      const long long u = t - nxf;
      const int i = 1 + (int)(u % (nx - 2)), k = 1 + (int)(u / (nx - 2));
      for (int a = 0; a < 2; ++a) {
        A[a][cidx(g, i, 0, k)] = A[a][cidx(g, i, ny - 2, k)];
        A[a][cidx(g, i, ny - 1, k)] = A[a][cidx(g, i, 1, k)];
      }
    } else if (t < nxf + nyf + nk) {
      const int k = 1 + (int)(t - nxf - nyf);
      for (int a = 0; a < 2; ++a) {
        A[a][cidx(g, 0, 0, k)] = A[a][cidx(g, nx - 2, ny - 2, k)];
        A[a][cidx(g, nx - 1, 0, k)] = A[a][cidx(g, 1, ny - 2, k)];
        A[a][cidx(g, 0, ny - 1, k)] = A[a][cidx(g, nx - 2, 1, k)];
        A[a][cidx(g, nx - 1, ny - 1, k)] = A[a][cidx(g, 1, 1, k)];
      }
    }
  } else {
    if (t < (long long)nx * ny) {
      const int i = (int)(t % nx), j = (int)(t / nx);
      for (int a = 0; a < 2; ++a) {
        A[a][cidx(g, i, j, 0)] = A[a][cidx(g, i, j, nz - 2)];
        A[a][cidx(g, i, j, nz - 1)] = A[a][cidx(g, i, j, 1)];
      }
    }
  }
}

template <class S>
__global__ __launch_bounds__(256) void k_les_bc(Prim<S> P, Geom g, int pass) {
  les_bc_item<S>(P, g, pass, (long long)blockIdx.x * 256 + threadIdx.x);
}

// Items of the two passes of les_bc_item over a grid.
inline long long les_bc_items(const Geom& g, int pass) {
  return pass == 0 ? (long long)(g.ny - 2) * (g.nz - 2) + (long long)(g.nx - 2) * (g.nz - 2) + (g.nz - 2)
                   : (long long)g.nx * g.ny;
}

// ---------------------------------------------------------------------------
// One cell's worth of window data.
// ---------------------------------------------------------------------------
template <class R, int DIR, int ORDER, int SCHEME, int VISC, int VO> struct CellData {
  R rho, u, v, w, p, T, mu, s, d1, d2, d3, d4, mt, H;  // mt, H: LES only
};

// The values one window cell is built from, as they sit in memory: its
// primitives and, for viscous runs, the transverse stencil values of the four
// cross derivatives a face in direction DIR needs (centre excluded).  Kept
// separate from CellData so that a sweep can issue the loads of the next cell
// one face ahead of their use.
template <class S, int VISC, int VO> struct RawCell {
  static constexpr int NX = VISC ? vcross<VO>() : 1;
  S rho, u, v, w, p, s, mt, qc2;
  S x[4][NX];
};

// The four cross derivatives by face direction:
//   DIR 0: uy, vy, uz, wz   DIR 1: ux, vx, vz, wz   DIR 2: ux, wx, vy, wy
template <class S, int DIR>
__device__ __forceinline__ void cross_fields(const Prim<S>& P, const Geom& g, const S* (&f)[4], int (&st)[4]) {
  const int sx = 1, sy = g.nx, sz = g.nx * g.ny;
  if constexpr (DIR == 0) {
    f[0] = P.u; st[0] = sy; f[1] = P.v; st[1] = sy; f[2] = P.u; st[2] = sz; f[3] = P.w; st[3] = sz;
  } else if constexpr (DIR == 1) {
    f[0] = P.u; st[0] = sx; f[1] = P.v; st[1] = sx; f[2] = P.v; st[2] = sz; f[3] = P.w; st[3] = sz;
  } else {
    f[0] = P.u; st[0] = sx; f[1] = P.w; st[1] = sx; f[2] = P.v; st[2] = sy; f[3] = P.w; st[3] = sy;
  }
}

// Issues the loads of cell m (sensor read at ms, its periodic image).
template <class S, int DIR, int ORDER, int SCHEME, int VISC, int VO, bool LOAD_S = true>
__device__ __forceinline__ void fetch_cell(const Prim<S>& P, const Geom& g, int m, int ms, RawCell<S, VISC, VO>& r) {
  r.rho = G(P.rho)[m];
  r.u = G(P.u)[m];
  r.v = G(P.v)[m];
  r.w = G(P.w)[m];
  r.p = G(P.p)[m];
  if constexpr (LOAD_S && kUsesSensor<SCHEME, ORDER>) r.s = G(P.s)[ms];
  if constexpr (VISC) {
    if constexpr (VISC == 2) {
      r.mt = G(P.mut)[m];
      r.qc2 = G(P.qc2)[m];
    }
    const S* f[4];
    int st[4];
    cross_fields<S, DIR>(P, g, f, st);
    constexpr int VC = vcross<VO>();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int t = 0; t < VC; ++t) {
        const int off = t < VC / 2 ? t - VC / 2 : t - VC / 2 + 1;
        r.x[q][t] = G(f[q])[m + off * st[q]];
      }
    }
  }
}

// cross_d over the raw stencil values (centre entry zero, never read)
template <class R, class S, int VO>
__device__ __forceinline__ R cross_raw(const PhysC<R>& c, const S* x, R inv) {
  constexpr int VC = vcross<VO>();
  R a[VC + 1];
#pragma unroll
  for (int t = 0; t <= VC; ++t) a[t] = (t == VC / 2) ? R(0) : R(x[t < VC / 2 ? t : t - 1]);
  return cross_d<R, VC>(c, inv, a);
}

// Turns the loaded values into the window cell (conversions, cross
// derivatives, LES total enthalpy).
template <class R, class S, int DIR, int ORDER, int SCHEME, int VISC, int VO, bool LOAD_S = true>
__device__ __forceinline__ void finish_cell(const Geom& g, const PhysC<R>& c, const RawCell<S, VISC, VO>& r,
                                            CellData<R, DIR, ORDER, SCHEME, VISC, VO>& d) {
  d.rho = R(r.rho);
  d.u = R(r.u);
  d.v = R(r.v);
  d.w = R(r.w);
  d.p = R(r.p);
  // T = p / (R rho) and Sutherland's law (calc_quantities_T_3D,
  // src/calc_physical_quantities.f90:60-116)
  if constexpr (kUsesT<SCHEME> || VISC) d.T = d.p / (c.Rgas * d.rho);
  if constexpr (VISC) d.mu = c.mu0c / (d.T + R(111.0)) * (d.T * fsqrt(d.T));
  if constexpr (LOAD_S && kUsesSensor<SCHEME, ORDER>) d.s = R(r.s);
  if constexpr (VISC) {
    if constexpr (VISC == 2) {
      d.mt = R(r.mt);
      d.H = c.Cp * d.T + R(0.5) * (d.u * d.u + d.v * d.v + d.w * d.w) + R(r.qc2);
    }
    const R ix = R(g.inv_dx), iy = R(g.inv_dy), iz = R(g.inv_dz);
    const R i0 = DIR == 0 ? iy : ix, i1 = DIR == 0 ? iy : ix;
    const R i2 = DIR == 2 ? iy : iz, i3 = DIR == 2 ? iy : iz;
    d.d1 = cross_raw<R, S, VO>(c, r.x[0], i0);
    d.d2 = cross_raw<R, S, VO>(c, r.x[1], i1);
    d.d3 = cross_raw<R, S, VO>(c, r.x[2], i2);
    d.d4 = cross_raw<R, S, VO>(c, r.x[3], i3);
  }
}

// Loads the primitive values (and, for viscous runs, the cross derivatives a
// face in direction DIR needs) of cell (i,j,k) = linear m.
template <class R, class S, int DIR, int ORDER, int SCHEME, int VISC, int VO, bool LOAD_S = true>
__device__ __forceinline__ void load_cell(const Prim<S>& P, const Geom& g, const PhysC<R>& c, int m, int ms,
                                          CellData<R, DIR, ORDER, SCHEME, VISC, VO>& d) {
  RawCell<S, VISC, VO> r;
  fetch_cell<S, DIR, ORDER, SCHEME, VISC, VO, LOAD_S>(P, g, m, ms, r);
  finish_cell<R, S, DIR, ORDER, SCHEME, VISC, VO, LOAD_S>(g, c, r, d);
}

// Convective face flux from SW window cells w[0..SW-1]; the face lies
// between w[HW-1] and w[HW].
template <class R, int DIR, int ORDER, int SCHEME, int VISC, int VO>
__device__ __forceinline__ void face_conv(const PhysC<R>& c, int tvd, int variant,
                                          const CellData<R, DIR, ORDER, SCHEME, VISC, VO>* w, R* F) {
  using St = Sten<ORDER, VISC, VO>;
  constexpr int N = ORDER;
  R rho[N], u[N], v[N], wv[N], p[N], T[N];
#pragma unroll
  for (int m = 0; m < N; ++m) {
    rho[m] = w[St::CO + m].rho;
    u[m] = w[St::CO + m].u;
    v[m] = w[St::CO + m].v;
    wv[m] = w[St::CO + m].w;
    p[m] = w[St::CO + m].p;
    if constexpr (kUsesT<SCHEME>) T[m] = w[St::CO + m].T;
  }
  R fs = R(0);
  if constexpr (kUsesSensor<SCHEME, ORDER>) fs = R(0.5) * (w[St::HW - 1].s + w[St::HW].s);
  if constexpr (SCHEME == KEEP) {
    keep_face<R, DIR, N>(c, rho, u, v, wv, p, T, F);
  } else if constexpr (SCHEME == SLAU) {
    slau_face<R, DIR, N>(c, tvd, variant, rho, u, v, wv, p, fs, F);
  } else if constexpr (SCHEME == ROE) {
    roe_face<R, DIR, N>(c, tvd, rho, u, v, wv, p, fs, F);
  } else {
    // calc_hybrid_{x,y,z}_in (calc_hybrid_kernel_internal.f90.fypp:96-121)
    if (fs <= c.threshold) {
      keep_face<R, DIR, N>(c, rho, u, v, wv, p, T, F);
    } else {
      R L[5], Rr[5];
      reconstruct<R, N>(c, tvd, fs, rho, u, v, wv, p, L, Rr);
      slau<R, DIR>(c, variant, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], L[3], Rr[3], L[4], Rr[4], R(1), F);
    }
  }
}

// Subtracts the viscous face flux from F (convective in, total out).
template <class R, int DIR, int ORDER, int SCHEME, int VISC, int VO>
__device__ __forceinline__ void face_visc(const PhysC<R>& c, R inv_d,
                                          const CellData<R, DIR, ORDER, SCHEME, VISC, VO>* w, R* F) {
  using St = Sten<ORDER, VISC, VO>;
  if constexpr (VISC) {
    constexpr int M = vwin<VO>();
    R mu_[M], T_[M], u_[M], v_[M], w_[M], a1[M], a2[M], a3[M], a4[M], mt_[M], H_[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const auto& x = w[St::VOFF + m];
      mu_[m] = x.mu; T_[m] = x.T; u_[m] = x.u; v_[m] = x.v; w_[m] = x.w;
      a1[m] = x.d1; a2[m] = x.d2; a3[m] = x.d3; a4[m] = x.d4;
      if constexpr (VISC == 2) { mt_[m] = x.mt; H_[m] = x.H; }
    }
    R tau[3], en;
    if constexpr (VO == 5) visc_face_me4<R, DIR, (VISC == 2)>(c, inv_d, mu_, T_, u_, v_, w_, a1, a2, a3, a4, tau, en, mt_, H_);
    else visc_face<R, DIR, M, (VISC == 2)>(c, inv_d, mu_, T_, u_, v_, w_, a1, a2, a3, a4, tau, en, mt_, H_);
    F[1] = F[1] - tau[0];
    F[2] = F[2] - tau[1];
    F[3] = F[3] - tau[2];
    F[4] = F[4] - en;
  }
}

// Face flux (convective minus viscous) from SW window cells w[0..SW-1].
template <class R, int DIR, int ORDER, int SCHEME, int VISC, int VO>
__device__ __forceinline__ void face_flux(const PhysC<R>& c, int tvd, int variant, R inv_d,
                                          const CellData<R, DIR, ORDER, SCHEME, VISC, VO>* w, R* F) {
  face_conv<R, DIR, ORDER, SCHEME, VISC, VO>(c, tvd, variant, w, F);
  face_visc<R, DIR, ORDER, SCHEME, VISC, VO>(c, inv_d, w, F);
}

// ---------------------------------------------------------------------------
// Launch descriptors.  SweepArgs carries everything one sweep of one case
// needs; the single-case kernels take it by value, the batched kernels read
// an array of them (one per case) and map their blocks onto the cases' tiles
// through a TileMap (first global tile index and tile counts per case).
// ---------------------------------------------------------------------------
template <class R, class S> struct SweepArgs {
  Prim<S> P;        // primitives read by the stage
  Prim<S> Pn;       // primitives of the new state (written by the z epilogue)
  S* Racc;          // residual accumulator
  Geom g;
  PhysC<R> c;
  PhysC<double> c64;
  StageArgs sa;
  double cd;        // c_x, c_y or c_z of calc_R (stage coefficient * dt * face area)
  int tvd, variant;
  int lseg;         // cells per thread segment in the y / z sweeps
};

struct TileMap {
  int start;     // first global tile of the case
  int nbx, nby;  // tiles along the three launch dimensions
  int nbz;
};

// Workgroups are dispatched round-robin over the eight XCDs, and each XCD has
// its own L2.  For n workgroups with n a multiple of 8, tile
// (b % 8) * (n / 8) + b / 8 gives XCD b % 8 one contiguous run of n / 8
// tiles, so the tiles whose stencils overlap (neighbours along the second and
// third launch dimensions) share an L2 instead of fetching the same rows into
// two.  The map is a permutation of [0, n).
#ifndef OX_XCD_SWIZZLE
#define OX_XCD_SWIZZLE 1
#endif
#ifndef OX_XSWEEP_ZMAJOR
#define OX_XSWEEP_ZMAJOR 1
#endif
__device__ __forceinline__ int xcd_tile(int b, int n) {
  if (!OX_XCD_SWIZZLE || (n & 7)) return b;
  return (b & 7) * (n >> 3) + (b >> 3);
}

// This workgroup's tile (bx, by, bz) of a 3-D launch under xcd_tile.
__device__ __forceinline__ void launch_tile(int& bx, int& by, int& bz) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
  bx = t % gx;
  by = (t / gx) % gy;
  bz = t / (gx * gy);
}

// The x sweep's tiles (bx, row band by, plane bz) in x, z, y order under
// xcd_tile: one XCD's run covers a band of rows through many planes, so both
// the y and the z neighbours a row tile's cross-derivative stencils read
// (rows by +- 1, planes bz +- 3) are fetched by tiles resident at about the
// same time in one L2.
__device__ __forceinline__ void launch_tile_xzy(int& bx, int& by, int& bz) {
  const int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const int t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gz);
#if OX_XSWEEP_ZMAJOR
  bx = t % gx;
  bz = (t / gx) % gz;
  by = t / (gx * gz);
#else
  bx = t % gx;
  by = (t / gx) % gy;
  bz = t / (gx * gy);
#endif
}

// Case owning global tile t (map[].start ascending, map[0].start == 0).
__device__ __forceinline__ int find_case(const TileMap* __restrict__ map, int ncase, int t) {
  int lo = 0, hi = ncase - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (map[mid].start <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// x sweep.  One wavefront per grid row: the tile of a row is exactly 64 cells
// wide (XT = 64 - SW interior cells plus the stencil halo), so every lane
// loads one cell, and lanes 0..XT compute the XT+1 faces the interior cells
// need; the right-face flux of each cell comes from the neighbouring lane by
// a cross-lane shuffle.
// ---------------------------------------------------------------------------
#ifndef OX_SPLIT_FENCE
#define OX_SPLIT_FENCE() asm volatile("" ::: "memory")
#endif
template <int ORDER, int VISC, int VO> constexpr int xtile() { return 64 - Sten<ORDER, VISC, VO>::SW; }
constexpr int XROWS = 4;

// Ducros sensor of one cell (calc_Ducros, 3D_solver/src/calc_hybrid.f90:11-72)
// from its six face neighbours' velocities.
template <class R>
__device__ __forceinline__ R ducros(const Geom& g, R uxm, R uxp, R vxm, R vxp, R wxm, R wxp, R uym, R uyp, R vym,
                                    R vyp, R wym, R wyp, R uzm, R uzp, R vzm, R vzp, R wzm, R wzp) {
  const R dxt = R(0.25) * (R(g.inv_dx) + R(g.inv_dx));
  const R dyt = R(0.25) * (R(g.inv_dy) + R(g.inv_dy));
  const R dzt = R(0.25) * (R(g.inv_dz) + R(g.inv_dz));
  const R dudx = (-uxm + uxp) * dxt;
  const R dvdx = (-vxm + vxp) * dxt;
  const R dwdx = (-wxm + wxp) * dxt;
  const R dudy = (-uym + uyp) * dyt;
  const R dvdy = (-vym + vyp) * dyt;
  const R dwdy = (-wym + wyp) * dyt;
  const R dudz = (-uzm + uzp) * dzt;
  const R dvdz = (-vzm + vzp) * dzt;
  const R dwdz = (-wzm + wzp) * dzt;
  const R div = dudx + dvdy + dwdz;
  const R r1 = dwdy - dvdz;
  const R r2 = dudz - dwdx;
  const R r3 = dvdx - dudy;
  R fd = (div * div) / (div * div + (r1 * r1 + r2 * r2 + r3 * r3) + R(1.0e-12));
  return fmin2(R(1), fd);
}

// The same sensor in two parts: the y/z-derivative terms (from loads the
// cell's own thread issues) and the completion with the x derivatives (from
// the LDS row).  Same operations in the same order as ducros().
template <class R> struct DucrosPart {
  R dudy, dvdy, dwdz, dudz, r1;
};
template <class R>
__device__ __forceinline__ DucrosPart<R> ducros_yz(const Geom& g, R uym, R uyp, R vym, R vyp, R wym, R wyp, R uzm,
                                                   R uzp, R vzm, R vzp, R wzm, R wzp) {
  const R dyt = R(0.25) * (R(g.inv_dy) + R(g.inv_dy));
  const R dzt = R(0.25) * (R(g.inv_dz) + R(g.inv_dz));
  DucrosPart<R> d;
  d.dudy = (-uym + uyp) * dyt;
  d.dvdy = (-vym + vyp) * dyt;
  const R dwdy = (-wym + wyp) * dyt;
  d.dudz = (-uzm + uzp) * dzt;
  const R dvdz = (-vzm + vzp) * dzt;
  d.dwdz = (-wzm + wzp) * dzt;
  d.r1 = dwdy - dvdz;
  return d;
}
template <class R>
__device__ __forceinline__ R ducros_x(const Geom& g, const DucrosPart<R>& d, R uxm, R uxp, R vxm, R vxp, R wxm,
                                      R wxp) {
  const R dxt = R(0.25) * (R(g.inv_dx) + R(g.inv_dx));
  const R dudx = (-uxm + uxp) * dxt;
  const R dvdx = (-vxm + vxp) * dxt;
  const R dwdx = (-wxm + wxp) * dxt;
  const R div = dudx + d.dvdy + d.dwdz;
  const R r2 = d.dudz - dwdx;
  const R r3 = dvdx - d.dudy;
  R fd = (div * div) / (div * div + (d.r1 * d.r1 + r2 * r2 + r3 * r3) + R(1.0e-12));
  return fmin2(R(1), fd);
}

// Writes Racc = c_x * (E_right - E_left) (the first term of calc_R).  For
// the schemes that read the Ducros sensor the sweep also computes it for its
// row and stores it for the y and z sweeps.  The LDS tile is
// structure-of-arrays so that lane-consecutive reads are bank-conflict free.
template <int VISC> constexpr int xsweep_nf() { return VISC == 2 ? 14 : 12; }  // rho,u,v,w,p,T,mu,s,d1..d4[,mt,H]

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
__device__ __forceinline__ void sweep_x_tile(const SweepArgs<R, S>& a, int bx, int by, int bz,
                                             R (&tile)[xsweep_nf<VISC>()][XROWS][64]) {
  const Prim<S>& P = a.P;
  S* __restrict__ Racc = a.Racc;
  const Geom& g = a.g;
  const PhysC<R>& c = a.c;
  const int tvd = a.tvd, variant = a.variant;
  const double cx = a.cd;
  using St = Sten<ORDER, VISC, VO>;
  using CD = CellData<R, 0, ORDER, SCHEME, VISC, VO>;
  using RC = RawCell<S, VISC, VO>;
  constexpr bool SENS = kUsesSensor<SCHEME, ORDER>;
  constexpr int XT = xtile<ORDER, VISC, VO>();
  constexpr int W = XT + St::SW;
  static_assert(W == 64, "one tile cell per lane");
  // With one ghost layer (ORDER 2) a ghost cell's sensor is the boundary copy
  // of its interior neighbour (calc_hybrid.f90:54-71), taken from a clamped
  // index; with more layers every sensor the faces read belongs to a cell in
  // 1..nx-2 whose x neighbours are in the row tile.
  constexpr bool SPLIT_SENS = SENS && ORDER > 2;
  constexpr int NF = xsweep_nf<VISC>();
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = g.h + bx * XT;
  const int j = g.h + by * XROWS + ty;
  const int k = g.h + bz;
  const bool rowok = j < g.h + g.iy;
  const int m = tx;                // this lane's tile position
  const int i = i0 - St::HW + m;   // and its cell
  const int sy = g.nx, sz = g.nx * g.ny;
  // sensor positions: the cells either side of the XT+1 faces
  const bool spos = m >= St::HW - 1 && m <= St::HW + XT && i <= g.nx - 2;
  DucrosPart<R> dp{};
  if (rowok && i < g.nx) {
    const int q = cidx(g, i, j, k);
    RC r;
    fetch_cell<S, 0, ORDER, SCHEME, VISC, VO, false>(P, g, q, 0, r);
    if constexpr (SPLIT_SENS) {
      if (spos) {
        constexpr int H = vcross<VO>() / 2;  // r.x[.][H-1], r.x[.][H]: the -1 and +1 neighbours
        R uym, uyp, vym, vyp, uzm, uzp, wzm, wzp;
        if constexpr (VISC) {  // DIR 0 cross stencils: u,v along y; u,w along z
          uym = R(r.x[0][H - 1]); uyp = R(r.x[0][H]);
          vym = R(r.x[1][H - 1]); vyp = R(r.x[1][H]);
          uzm = R(r.x[2][H - 1]); uzp = R(r.x[2][H]);
          wzm = R(r.x[3][H - 1]); wzp = R(r.x[3][H]);
        } else {
          uym = R(G(P.u)[q - sy]); uyp = R(G(P.u)[q + sy]);
          vym = R(G(P.v)[q - sy]); vyp = R(G(P.v)[q + sy]);
          uzm = R(G(P.u)[q - sz]); uzp = R(G(P.u)[q + sz]);
          wzm = R(G(P.w)[q - sz]); wzp = R(G(P.w)[q + sz]);
        }
        dp = ducros_yz<R>(g, uym, uyp, vym, vyp, R(G(P.w)[q - sy]), R(G(P.w)[q + sy]), uzm, uzp, R(G(P.v)[q - sz]),
                          R(G(P.v)[q + sz]), wzm, wzp);
      }
    }
    CD d;
    finish_cell<R, S, 0, ORDER, SCHEME, VISC, VO, false>(g, c, r, d);
    tile[0][ty][m] = d.rho; tile[1][ty][m] = d.u; tile[2][ty][m] = d.v; tile[3][ty][m] = d.w;
    tile[4][ty][m] = d.p;
    if constexpr (kUsesT<SCHEME> || VISC) tile[5][ty][m] = d.T;
    if constexpr (VISC) {
      tile[6][ty][m] = d.mu;
      tile[8][ty][m] = d.d1; tile[9][ty][m] = d.d2; tile[10][ty][m] = d.d3; tile[11][ty][m] = d.d4;
      if constexpr (VISC == 2) { tile[NF - 2][ty][m] = d.mt; tile[NF - 1][ty][m] = d.H; }
    }
  }
  __syncthreads();
  if constexpr (SENS) {
    if constexpr (SPLIT_SENS) {
      if (rowok && spos) {
        const R fd = ducros_x<R>(g, dp, tile[1][ty][m - 1], tile[1][ty][m + 1], tile[2][ty][m - 1],
                                 tile[2][ty][m + 1], tile[3][ty][m - 1], tile[3][ty][m + 1]);
        tile[7][ty][m] = fd;
        if (i >= i0 && i < i0 + XT && i < g.h + g.ix) G(P.s)[cidx(g, i, j, k)] = (S)fd;  // one tile stores a cell
      }
    } else {
      if (rowok && m >= St::HW - 1 && m <= St::HW + XT && i <= g.nx - 1) {
        const int ic = min(max(i, 1), g.nx - 2);
        const int q = cidx(g, ic, j, k);
        const R fd = ducros<R>(g, R(G(P.u)[q - 1]), R(G(P.u)[q + 1]), R(G(P.v)[q - 1]), R(G(P.v)[q + 1]), R(G(P.w)[q - 1]),
                               R(G(P.w)[q + 1]), R(G(P.u)[q - sy]), R(G(P.u)[q + sy]), R(G(P.v)[q - sy]), R(G(P.v)[q + sy]),
                               R(G(P.w)[q - sy]), R(G(P.w)[q + sy]), R(G(P.u)[q - sz]), R(G(P.u)[q + sz]), R(G(P.v)[q - sz]),
                               R(G(P.v)[q + sz]), R(G(P.w)[q - sz]), R(G(P.w)[q + sz]));
        tile[7][ty][m] = fd;
        if (i >= i0 && i < i0 + XT && i < g.h + g.ix) G(P.s)[q] = (S)fd;
      }
    }
    __syncthreads();
  }
  if (!rowok) return;
  // face f = i0 - 1 + tx: window cells f-HW+1 .. f+HW = tile positions
  // tx .. tx+SW-1 (< W for the XT+1 faces the tile's cells need)
  R F[5] = {R(0), R(0), R(0), R(0), R(0)};
  const int f = i0 - 1 + tx;
  const bool face_ok = tx <= XT && f <= g.h + g.ix - 1;
  if (face_ok) {
    // The convective and the viscous halves each read their window from the
    // LDS row just before use, which keeps the two halves' inputs out of
    // registers at the same time.
    {
      CD win[St::SW];
#pragma unroll
      for (int n = 0; n < St::SW; ++n) {
        win[n].rho = tile[0][ty][tx + n]; win[n].u = tile[1][ty][tx + n]; win[n].v = tile[2][ty][tx + n];
        win[n].w = tile[3][ty][tx + n]; win[n].p = tile[4][ty][tx + n];
        if constexpr (kUsesT<SCHEME>) win[n].T = tile[5][ty][tx + n];
        if constexpr (SENS) {
          if (n == St::HW - 1 || n == St::HW) win[n].s = tile[7][ty][tx + n];
        }
      }
      face_conv<R, 0, ORDER, SCHEME, VISC, VO>(c, tvd, variant, win, F);
    }
    if constexpr (VISC) {
      OX_SPLIT_FENCE();
      CD win[St::SW];
#pragma unroll
      for (int n = 0; n < St::SW; ++n) {
        win[n].u = tile[1][ty][tx + n]; win[n].v = tile[2][ty][tx + n]; win[n].w = tile[3][ty][tx + n];
        win[n].T = tile[5][ty][tx + n];
        win[n].mu = tile[6][ty][tx + n];
        win[n].d1 = tile[8][ty][tx + n]; win[n].d2 = tile[9][ty][tx + n];
        win[n].d3 = tile[10][ty][tx + n]; win[n].d4 = tile[11][ty][tx + n];
        if constexpr (VISC == 2) { win[n].mt = tile[NF - 2][ty][tx + n]; win[n].H = tile[NF - 1][ty][tx + n]; }
      }
      face_visc<R, 0, ORDER, SCHEME, VISC, VO>(c, R(g.inv_dx), win, F);
    }
  }
  double Fr[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) Fr[l] = __shfl_down((double)F[l], 1, 64);
  const int ic = i0 + tx;
  if (tx < XT && ic < g.h + g.ix) {
#pragma unroll
    for (int l = 0; l < 5; ++l) G(Racc)[qidx(g, ic, l, j, k)] = (S)(cx * (-(double)F[l] + Fr[l]));
  }
}

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
__global__ __launch_bounds__(256) void k_sweep_x(SweepArgs<R, S> a) {
  __shared__ R tile[xsweep_nf<VISC>()][XROWS][64];
  int bx, by, bz;
  launch_tile_xzy(bx, by, bz);
  sweep_x_tile<R, S, ORDER, SCHEME, VISC, VO>(a, bx, by, bz, tile);
}

// ---------------------------------------------------------------------------
// x sweep over whole rows of a periodic grid whose interior width ix is 64,
// 128 or 256.  A row tile above spends SW - 1 of its 64 lanes on halo cells,
// so a row of 256 cells at SW = 6 takes five waves, the fifth with 24 cells.
// Here a workgroup of 256 lanes holds 256 / ix whole rows, one lane per cell
// and one face per lane (its cell's left face), and the halo positions of the
// LDS row hold the periodic images of the row's other end.  The face left of
// the first cell is then the face right of the last: its window cells are the
// periodic images of that face's window cells.  The row tile computes the
// same face from the ghost cells, which hold exact copies of those images
// once a z-sweep epilogue has written the primitive set, so both tiles give
// the same bits.  Used for such primitive sets only (k_prim takes the ghost
// layers of an uploaded state as they are), on fully periodic grids, without
// LES (its ghost SGS fields come from k_les_bc), and not for ORDER 2 with the
// sensor, whose ghost cells copy the adjacent interior sensor
// (calc_hybrid.f90:54-71) instead of the image's.
// ---------------------------------------------------------------------------
template <int ORDER, int SCHEME, int VISC>
constexpr bool kXRowSweep = VISC != 2 && !(kUsesSensor<SCHEME, ORDER> && ORDER == 2);
__host__ __device__ constexpr bool xrow_width(int ix) { return ix == 64 || ix == 128 || ix == 256; }
// Consecutive workgroups take the rows of a band of XROW_BAND rows in one
// plane, then the next plane: the y neighbours that the cross-derivative and
// sensor stencils read are shared by workgroups resident together in one L2,
// and the z neighbours by the bands of the next planes.  (Whole rows one after
// another through the planes re-read the y neighbours of every row from HBM.)
constexpr int XROW_BAND = 16;

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
__global__ __launch_bounds__(256) void k_sweep_xrow(SweepArgs<R, S> a) {
  using St = Sten<ORDER, VISC, VO>;
  using CD = CellData<R, 0, ORDER, SCHEME, VISC, VO>;
  using RC = RawCell<S, VISC, VO>;
  constexpr bool SENS = kUsesSensor<SCHEME, ORDER>;
  constexpr int NF = xsweep_nf<VISC>();
  constexpr int HW = St::HW, SW = St::SW;
  static_assert(kXRowSweep<ORDER, SCHEME, VISC>, "row sweep for image-ghost schemes only");
  // rows of L + SW - 1 positions: 4 * (63 + SW) covers L = 64 (four rows) to 256 (one)
  __shared__ R tile[NF][4 * (63 + SW)];
  __shared__ double edge[4][5];  // each wave's lane-0 face flux
  const Prim<S>& P = a.P;
  const Geom& g = a.g;
  const PhysC<R>& c = a.c;
  const int L = g.ix, rw = 256 / L;  // row length, rows per workgroup
  const int tid = threadIdx.x, lane = tid & 63;
  const int r = tid / L, p = tid - r * L;  // row in the workgroup, cell in the row
  const int nbb = XROW_BAND / rw;  // workgroups per band
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = g.h + (t / (nbb * g.iz)) * XROW_BAND + (t % nbb) * rw + r, k = g.h + (t / nbb) % g.iz;
  const bool rowok = j < g.h + g.iy;
  const int i = g.h + p;
  const int rb = r * (L + SW - 1);  // the row's first position, cell h - HW
  const int m = p + HW;             // this lane's position
  const int q = cidx(g, i, j, k);
  DucrosPart<R> dp{};
  if (rowok) {
    RC rc;
    fetch_cell<S, 0, ORDER, SCHEME, VISC, VO, false>(P, g, q, 0, rc);
    if constexpr (SENS) {
      const int sy = g.nx, sz = g.nx * g.ny;
      constexpr int H = vcross<VO>() / 2;
      R uym, uyp, vym, vyp, uzm, uzp, wzm, wzp;
      if constexpr (VISC) {
        uym = R(rc.x[0][H - 1]); uyp = R(rc.x[0][H]);
        vym = R(rc.x[1][H - 1]); vyp = R(rc.x[1][H]);
        uzm = R(rc.x[2][H - 1]); uzp = R(rc.x[2][H]);
        wzm = R(rc.x[3][H - 1]); wzp = R(rc.x[3][H]);
      } else {
        uym = R(G(P.u)[q - sy]); uyp = R(G(P.u)[q + sy]);
        vym = R(G(P.v)[q - sy]); vyp = R(G(P.v)[q + sy]);
        uzm = R(G(P.u)[q - sz]); uzp = R(G(P.u)[q + sz]);
        wzm = R(G(P.w)[q - sz]); wzp = R(G(P.w)[q + sz]);
      }
      dp = ducros_yz<R>(g, uym, uyp, vym, vyp, R(G(P.w)[q - sy]), R(G(P.w)[q + sy]), uzm, uzp, R(G(P.v)[q - sz]),
                        R(G(P.v)[q + sz]), wzm, wzp);
    }
    CD d;
    finish_cell<R, S, 0, ORDER, SCHEME, VISC, VO, false>(g, c, rc, d);
    auto put = [&](int at) {
      at += rb;
      tile[0][at] = d.rho; tile[1][at] = d.u; tile[2][at] = d.v; tile[3][at] = d.w; tile[4][at] = d.p;
      if constexpr (kUsesT<SCHEME> || VISC) tile[5][at] = d.T;
      if constexpr (VISC) {
        tile[6][at] = d.mu;
        tile[8][at] = d.d1; tile[9][at] = d.d2; tile[10][at] = d.d3; tile[11][at] = d.d4;
      }
    };
    put(m);
    if (p >= L - HW) put(m - L);      // left halo: images of the row's last HW cells
    if (p < SW - HW - 1) put(m + L);  // right halo: images of its first SW - HW - 1
  }
  __syncthreads();
  if constexpr (SENS) {
    if (rowok) {
      const int at = rb + m;
      const R fd = ducros_x<R>(g, dp, tile[1][at - 1], tile[1][at + 1], tile[2][at - 1], tile[2][at + 1],
                               tile[3][at - 1], tile[3][at + 1]);
      tile[7][at] = fd;
      if (p == L - 1) tile[7][rb + HW - 1] = fd;  // the image left of the row
      G(P.s)[q] = (S)fd;
    }
    __syncthreads();
  }
  R F[5] = {R(0), R(0), R(0), R(0), R(0)};
  if (rowok) {
    const int w0 = rb + p;  // window cells h + p - HW .. h + p + HW - 1
    {
      CD win[SW];
#pragma unroll
      for (int n = 0; n < SW; ++n) {
        win[n].rho = tile[0][w0 + n]; win[n].u = tile[1][w0 + n]; win[n].v = tile[2][w0 + n];
        win[n].w = tile[3][w0 + n]; win[n].p = tile[4][w0 + n];
        if constexpr (kUsesT<SCHEME>) win[n].T = tile[5][w0 + n];
        if constexpr (SENS) {
          if (n == HW - 1 || n == HW) win[n].s = tile[7][w0 + n];
        }
      }
      face_conv<R, 0, ORDER, SCHEME, VISC, VO>(c, a.tvd, a.variant, win, F);
    }
    if constexpr (VISC) {
      OX_SPLIT_FENCE();
      CD win[SW];
#pragma unroll
      for (int n = 0; n < SW; ++n) {
        win[n].u = tile[1][w0 + n]; win[n].v = tile[2][w0 + n]; win[n].w = tile[3][w0 + n];
        win[n].T = tile[5][w0 + n];
        win[n].mu = tile[6][w0 + n];
        win[n].d1 = tile[8][w0 + n]; win[n].d2 = tile[9][w0 + n];
        win[n].d3 = tile[10][w0 + n]; win[n].d4 = tile[11][w0 + n];
      }
      face_visc<R, 0, ORDER, SCHEME, VISC, VO>(c, R(g.inv_dx), win, F);
    }
  }
  // the right face is the next cell's left face: the next lane's, across a
  // wave boundary the next wave's lane 0, and after the row's last cell the
  // row's first (with L = 64 the row is the wave, and the rotation covers it)
  double Fr[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) Fr[l] = __shfl((double)F[l], (lane + 1) & 63, 64);
  if (L > 64) {
    if (lane == 0) {
#pragma unroll
      for (int l = 0; l < 5; ++l) edge[tid >> 6][l] = (double)F[l];
    }
    __syncthreads();
    if (lane == 63) {
      const int wn = (r * L + ((p + 1) & (L - 1))) >> 6;
#pragma unroll
      for (int l = 0; l < 5; ++l) Fr[l] = edge[wn][l];
    }
  }
  if (rowok) {
    const double cx = a.cd;
#pragma unroll
    for (int l = 0; l < 5; ++l) G(a.Racc)[qidx(g, i, l, j, k)] = (S)(cx * (-(double)F[l] + Fr[l]));
  }
}

// Persistent batched x sweep: block b handles global tiles b, b + gridDim.x, ...
template <class R, class S, int ORDER, int SCHEME, int VISC, int VO>
__global__ __launch_bounds__(256) void k_sweep_x_batch(const SweepArgs<R, S>* __restrict__ cases,
                                                       const TileMap* __restrict__ map, int ncase, int ntiles) {
  __shared__ R tile[xsweep_nf<VISC>()][XROWS][64];
  for (int t = xcd_tile(blockIdx.x, gridDim.x); t < ntiles; t += gridDim.x) {
    const int ci = find_case(map, ncase, t);
    const int local = t - map[ci].start, nbx = map[ci].nbx;
#if OX_XSWEEP_ZMAJOR
    const int nbz = map[ci].nbz;
    sweep_x_tile<R, S, ORDER, SCHEME, VISC, VO>(cases[ci], local % nbx, local / (nbx * nbz), (local / nbx) % nbz,
                                                tile);
#else
    const int nby = map[ci].nby;
    sweep_x_tile<R, S, ORDER, SCHEME, VISC, VO>(cases[ci], local % nbx, (local / nbx) % nby, local / (nbx * nby),
                                                tile);
#endif
    __syncthreads();  // the next tile overwrites the LDS rows
  }
}

// The residual of one cell and, for the z sweep, the state values its RK
// epilogue reads (A = Qin, B = Qout, C = Rs), per conserved variable.
template <int DIR, int EP> struct CellState {
  double R[5], A[5], B[5], C[5];
};
template <int DIR, int EP, class S>
__device__ __forceinline__ void fetch_state(const S* __restrict__ Racc, const StageArgs& sa, int q, int st,
                                            CellState<DIR, EP>& cs) {
#pragma unroll
  for (int l = 0; l < 5; ++l) cs.R[l] = (double)G(Racc)[q + l * st];
  if constexpr (DIR == 2) {
    if constexpr (EP == EP_RK3_S1 || EP == EP_RK3_S23 || EP == EP_RK4_S) {
#pragma unroll
      for (int l = 0; l < 5; ++l) cs.A[l] = G(sa.Qin)[q + l * st];
    }
    if constexpr (EP == EP_RK3_S23 || EP == EP_RK4_S4) {
#pragma unroll
      for (int l = 0; l < 5; ++l) cs.B[l] = G(sa.Qout)[q + l * st];
    }
    if constexpr (EP == EP_RK4_S || EP == EP_RK4_S4) {
#pragma unroll
      for (int l = 0; l < 5; ++l) cs.C[l] = G(sa.Rs)[q + l * st];
    }
  }
}

// ---------------------------------------------------------------------------
// y / z sweeps: each thread owns one (x, transverse) column and marches a
// segment of LSEG cells along DIR with a register window of SW cells; loads
// are coalesced along x across the wavefront.  DIR=1 accumulates
// R += c_y*(F_right - F_left); DIR=2 adds c_z*(G_right - G_left) and applies
// the stage epilogue.
// ---------------------------------------------------------------------------
constexpr int LSEG_DEFAULT = 64;
// Line-sweep workgroup: LINE_X consecutive x columns by LINE_Y rows (z rows
// for the y sweep, y rows for the z sweep), 256 threads.
#ifndef OX_LINE_X
#define OX_LINE_X 64
#endif
constexpr int LINE_X = OX_LINE_X;
constexpr int LINE_Y = 256 / LINE_X;
static_assert(LINE_X % 64 == 0 && LINE_X * LINE_Y == 256, "line-sweep workgroup of whole wavefronts");

// Longest segment of the y (dir 1) or z (dir 2) sweep whose state offsets
// from a thread's column pointer (the cell before the segment) stay below
// 2^31: a thread reaches lseg + 1 cells along dir past it, plus the five
// variables of a cell, so (lseg + 2) cells of 5 nx (y) or 5 nx ny (z) values.
// At least 1 for any grid ox_create accepts (15 nx ny < 2^31).
inline int lseg_limit(const Geom& g, int dir) {
  const long long step = dir == 1 ? 5LL * g.nx : 5LL * g.nx * g.ny;
  const long long n = 2147483647LL / step - 2;
  return (int)std::max(1LL, std::min(n, 4096LL));
}
inline int clamp_lseg(const Geom& g, int dir, int lseg) { return std::min(lseg, lseg_limit(g, dir)); }

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO, int DIR, int EP>
__device__ __forceinline__ void sweep_line_tile(const SweepArgs<R, S>& a, int bx, int by, int bz) {
  const Prim<S>& P = a.P;
  const Prim<S>& Pn = a.Pn;
  const Geom& g = a.g;
  const PhysC<R>& c = a.c;
  const PhysC<double>& c64 = a.c64;
  const int tvd = a.tvd, variant = a.variant;
  const double cd = a.cd;
  using St = Sten<ORDER, VISC, VO>;
  using CD = CellData<R, DIR, ORDER, SCHEME, VISC, VO>;
  const int i = g.h + bx * LINE_X + threadIdx.x;
  const int t = g.h + by * LINE_Y + threadIdx.y;  // k for DIR 1, j for DIR 2
  const int s0 = g.h + bz * a.lseg;
  const int nt = DIR == 1 ? g.iz : g.iy;
  const int nd = DIR == 1 ? g.iy : g.iz;
  if (i >= g.h + g.ix || t >= g.h + nt) return;
  const int s_end = min(s0 + a.lseg, g.h + nd);
  const int stride = DIR == 1 ? g.nx : g.nx * g.ny;
  const int mbase = DIR == 1 ? cidx(g, i, 0, t) : cidx(g, i, t, 0);
  // State, residual and RK operands: each thread's pointers to its column
  // at the cell before the segment (64-bit: the state of a large grid spans
  // more than 2^31 values), and offsets along DIR that are the same for the
  // whole workgroup (scalar registers, 32-bit: below 2^31 by the
  // segment-length limit, lseg_limit).
  const long long qcol = DIR == 1 ? qidx(g, i, 0, s0 - 1, t) : qidx(g, i, 0, t, s0 - 1);
  S* __restrict__ Racc = a.Racc + qcol;
  StageArgs sa = a.sa;
  if constexpr (DIR == 2) {
    // only the operands this epilogue touches (the others may be null)
    if constexpr (EP == EP_RK3_S1 || EP == EP_RK3_S23 || EP == EP_RK4_S) sa.Qin += qcol;
    sa.Qout += qcol;
    if constexpr (EP == EP_RK4_S || EP == EP_RK4_S4) sa.Rs += qcol;
  }
  const int qstride = DIR == 1 ? 5 * g.nx : 5 * g.nx * g.ny;
  const R inv_d = DIR == 1 ? R(g.inv_dy) : R(g.inv_dz);
  // the sensor is stored for interior cells only; a ghost cell's sensor is
  // bitwise that of its periodic image (its inputs are exact copies), except
  // with one ghost layer (ORDER 2), where the reference copies the adjacent
  // interior value into the ghost cell (calc_hybrid.f90:54-71), and along a
  // non-periodic direction, where the inner ghost layer holds its own sensor
  // (the shift is a select, not a branch: a branch in the loop would make the
  // compiler wait for every load in flight at its merge point)
  const int period = g.bc[DIR] == OX_BC_PERIODIC ? nd : 0;
  auto wrap = [&](int sp) {
    if constexpr (ORDER == 2) return sp < g.h ? g.h : (sp >= g.h + nd ? g.h + nd - 1 : sp);
    return sp < g.h ? sp + period : (sp >= g.h + nd ? sp - period : sp);
  };
  using RC = RawCell<S, VISC, VO>;
  CD win[St::SW];
#pragma unroll
  for (int m = 0; m < St::SW - 1; ++m) {
    const int sp = s0 - St::HW + m;
    load_cell<R, S, DIR, ORDER, SCHEME, VISC, VO>(P, g, c, mbase + sp * stride, mbase + wrap(sp) * stride, win[m]);
  }
  // Software pipeline: the values of the cell that enters the window at face
  // f+1, and the residual (and state) of the cell that face f+1 completes,
  // are loaded while face f is computed, so that one face's arithmetic
  // covers the memory latency of the next.
  RC nxt;
  {
    const int sp = s0 - 1 + St::HW;
    fetch_cell<S, DIR, ORDER, SCHEME, VISC, VO>(P, g, mbase + sp * stride, mbase + wrap(sp) * stride, nxt);
  }
  R Fp[5];
  CellState<DIR, EP> cs;
  // The prefetches below are issued on every iteration, the last one included
  // (its values are never used): a prefetch behind a branch makes the
  // compiler wait for every outstanding memory operation, the residual stores
  // of the iteration included, at the loop's merge points.  On the last
  // iteration the cell entering the window is clamped to the array's last
  // layer along DIR, and cell f+1 = s_end is at most that layer as well.
  const int n_along = DIR == 1 ? g.ny : g.nz;
  for (int f = s0 - 1; f < s_end; ++f) {
    finish_cell<R, S, DIR, ORDER, SCHEME, VISC, VO>(g, c, nxt, win[St::SW - 1]);
    {
      const int sp = min(f + 1 + St::HW, n_along - 1);
      fetch_cell<S, DIR, ORDER, SCHEME, VISC, VO>(P, g, mbase + sp * stride, mbase + wrap(sp) * stride, nxt);
    }
    // residual (and state) of cell f+1, consumed one iteration later
    CellState<DIR, EP> ncs;
    const int q = (f - s0 + 1) * qstride;  // cell f, from the column pointers
    fetch_state<DIR, EP>(Racc, sa, q + qstride, g.nx, ncs);
    const double *Ra = cs.R, *Qa = cs.A, *Qb = cs.B, *Qc = cs.C;
    R F[5];
    face_flux<R, DIR, ORDER, SCHEME, VISC, VO>(c, tvd, variant, inv_d, win, F);
    if (f >= s0) {
      double Rl[5];
#pragma unroll
      for (int l = 0; l < 5; ++l) Rl[l] = Ra[l] + cd * (-(double)Fp[l] + (double)F[l]);
      if constexpr (DIR == 1) {
#pragma unroll
        for (int l = 0; l < 5; ++l) G(Racc)[q + l * g.nx] = (S)Rl[l];
      } else {
        // stage epilogue (3D_solver/src/calc_steps.f90:60-69, :98-106, :137-145, :173-183)
        double Qn[5];
        if constexpr (EP == EP_RK3_S1) {
#pragma unroll
          for (int l = 0; l < 5; ++l) Qn[l] = Qa[l] - Rl[l];
        } else if constexpr (EP == EP_RK3_S23) {
#pragma unroll
          for (int l = 0; l < 5; ++l) Qn[l] = (sa.c1 * Qa[l] + sa.c2 * Qb[l] - Rl[l]) * sa.c4inv;
        } else if constexpr (EP == EP_RK4_S) {
#pragma unroll
          for (int l = 0; l < 5; ++l) {
            Qn[l] = Qa[l] - sa.c1 * Rl[l];
            G(sa.Rs)[q + l * g.nx] = Qc[l] + sa.c2 * Rl[l];
          }
        } else if constexpr (EP == EP_RK4_S4) {
#pragma unroll
          for (int l = 0; l < 5; ++l) {
            const double r = Qc[l] + Rl[l];
            Qn[l] = Qb[l] - r * sa.c1;  // c1 = one_sixth
            G(sa.Rs)[q + l * g.nx] = 0.0;
          }
        } else {
#pragma unroll
          for (int l = 0; l < 5; ++l) Qn[l] = Rl[l];
        }
#pragma unroll
        for (int l = 0; l < 5; ++l) G(sa.Qout)[q + l * g.nx] = Qn[l];
        if constexpr (EP != EP_RHS) {
          // primitives of the new state for the next stage (what k_prim
          // would compute from it), into the other primitive set, at the
          // cell and at every ghost cell that is its periodic image (the
          // ghost fill of set_bc_cyclic applied to the primitives)
          const PrimVals pv = prim_of(Qn[0], Qn[1], Qn[2], Qn[3], Qn[4], g.jac, c64);
          const int jj = t, kk = f;
          store_prim<S>(Pn, cidx(g, i, jj, kk), pv);
          const int ai = i < 2 * g.h ? i + g.ix : (i >= g.ix ? i - g.ix : -1);
          const int aj = jj < 2 * g.h ? jj + g.iy : (jj >= g.iy ? jj - g.iy : -1);
          const int ak = kk < 2 * g.h ? kk + g.iz : (kk >= g.iz ? kk - g.iz : -1);
          if (ai >= 0 || aj >= 0 || ak >= 0) {
            const int xs[2] = {i, ai}, ys[2] = {jj, aj}, zs[2] = {kk, ak};
            for (int c3 = 1; c3 < 8; ++c3) {
              const int x = xs[c3 & 1], y = ys[(c3 >> 1) & 1], z = zs[(c3 >> 2) & 1];
              if (x >= 0 && y >= 0 && z >= 0) store_prim<S>(Pn, cidx(g, x, y, z), pv);
            }
          }
        }
      }
    }
#pragma unroll
    for (int l = 0; l < 5; ++l) Fp[l] = F[l];
#pragma unroll
    for (int m = 0; m < St::SW - 1; ++m) win[m] = win[m + 1];
    cs = ncs;
  }
}

template <class R, class S, int ORDER, int SCHEME, int VISC, int VO, int DIR, int EP>
__global__ __launch_bounds__(256) void k_sweep_line(SweepArgs<R, S> a) {
  int bx, by, bz;
  launch_tile(bx, by, bz);
  sweep_line_tile<R, S, ORDER, SCHEME, VISC, VO, DIR, EP>(a, bx, by, bz);
}

// Persistent batched y / z sweep over the tiles of several cases.
template <class R, class S, int ORDER, int SCHEME, int VISC, int VO, int DIR, int EP>
__global__ __launch_bounds__(256) void k_sweep_line_batch(const SweepArgs<R, S>* __restrict__ cases,
                                                          const TileMap* __restrict__ map, int ncase, int ntiles) {
  for (int t = xcd_tile(blockIdx.x, gridDim.x); t < ntiles; t += gridDim.x) {
    const int ci = find_case(map, ncase, t);
    const int local = t - map[ci].start, nbx = map[ci].nbx, nby = map[ci].nby;
    sweep_line_tile<R, S, ORDER, SCHEME, VISC, VO, DIR, EP>(cases[ci], local % nbx, (local / nbx) % nby,
                                                            local / (nbx * nby));
  }
}

// Persistent batched SGS viscosity of LES cases over the tiles of several
// cases (tile = 64 x 4 cells of one plane, the single-case launch's block),
// reading each case's x-sweep descriptor: the primitive set the stage reads,
// the geometry and the constants.  Blocks are 64 x 4.
template <class R, class S>
__global__ __launch_bounds__(256) void k_les_mut_batch(const SweepArgs<R, S>* __restrict__ cases,
                                                       const TileMap* __restrict__ map, int ncase, int ntiles) {
  for (int t = xcd_tile(blockIdx.x, gridDim.x); t < ntiles; t += gridDim.x) {
    const int ci = find_case(map, ncase, t);
    const int local = t - map[ci].start, nbx = map[ci].nbx, nby = map[ci].nby;
    const SweepArgs<R, S>& a = cases[ci];
    const int i = 1 + (local % nbx) * 64 + (int)threadIdx.x;
    const int j = 1 + ((local / nbx) % nby) * 4 + (int)threadIdx.y;
    const int k = 1 + local / (nbx * nby);
    if (i <= a.g.nx - 2 && j <= a.g.ny - 2) les_mut_cell<S, R>(a.P, a.g, a.c, i, j, k);
  }
}

// Persistent batched ghost copy of nu_t and qc2, one pass (tile = 256 items).
// Blocks are 256 x 1.
template <class R, class S>
__global__ __launch_bounds__(256) void k_les_bc_batch(const SweepArgs<R, S>* __restrict__ cases,
                                                      const TileMap* __restrict__ map, int ncase, int ntiles,
                                                      int pass) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int ci = find_case(map, ncase, t);
    les_bc_item<S>(cases[ci].P, cases[ci].g, pass, (long long)(t - map[ci].start) * 256 + threadIdx.x);
  }
}

}  // namespace ox
