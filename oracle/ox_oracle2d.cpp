// ox_oracle2d.cpp -- FP64 CPU transcription of the reference's two-dimensional
// solver with boundaries, for its inviscid oblique-shock case (the ORACLE of
// the 2-D path).
//
// TEST INFRASTRUCTURE ONLY, like ox_oracle.hpp: only tests/, the smoke check
// and the benchmark's CPU baseline may build, link or call it.
//
// Restated routines (2D_solver/src unless noted), Fortran 1-based indices and
// the reference's loop ranges kept:
//   set_grid                 OS/set.f90:8-26 (uniform grid)
//   set_Jacobian_xy2         src/set_coordinate.f90:78-96
//   pre_calc                 preprocess.f90.fypp:59-111 (QJ = Q/J, dtdx, dtdy, 1/dx)
//   calc_quantities_2D       src/calc_physical_quantities.f90:7-29
//   calc_Ducros (2-D)        calc_hybrid.f90:11-52
//   calc_slau_{x,y}          calc_slau_kernel.f90.fypp:17-200 (boundary kernels:
//                            order fallback next to the boundaries)
//   SLAU1 / HRSLAU2          calc_slau_2d.f90 (= the 3-D routines with w = 0;
//                            the oracle calls those with w = 0, which adds
//                            exact zeros only)
//   calc_R, calc_step1,      calc_steps.f90:8-28, :31-54, :89-117
//   calc_step2_3
//   set_bc (oblique shock)   OS/set.f90:48-86
//   RungeKutta (RK3)         calc_time_dev.f90.fypp:77-94
//
// Roe.  The 2-D solver has no Roe scheme (2D_solver/src holds KEEP, SLAU and
// Hybrid only), and the 3-D Roe boundary kernel compares io with 4 and 6 and
// is therefore first order everywhere (3D_solver/src/calc_roe_kernel.f90.fypp:43-92).
// The 2-D Roe path here is defined by this library: the 3-D Roe flux
// (3D_solver/src/calc_roe_3d.f90:1-88, local frame of
// calc_roe_kernel_internal.f90.fypp:161-163 for y) with w = 0, behind the
// same order fallback as the SLAU boundary kernels, without the wiggle flag.
//
// Parity pins (tests/test_oracle2d.py): free stream with consistent
// boundaries is steady to rounding; the face flux reduces to the Euler flux
// for equal states and upwinds supersonic flow; mirror symmetry in y of the
// interior update; the oblique-shock run reproduces the Rankine-Hugoniot
// states of the reference's own test (ouxsbli/tests/test_os.py, 3 %) and of
// the exact regular reflection.
#include <omp.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ox_oracle.hpp"

namespace oxo {

struct Oracle2D {
  ox2d_config cfg{};
  Consts c{};
  int nx = 0, ny = 0, io = 0;
  std::vector<double> dx, dy, xix, etay, Jac, dtdx, dtdy;
  std::vector<double> QJ, QJ2, Qp, T, sensor, E, F;
  double qin[4] = {0, 0, 0, 0}, qpost[4] = {0, 0, 0, 0};  // inflow / post-shock primitives
  std::string err;

  size_t iq(int i, int l, int j) const { return (size_t)(i - 1) + (size_t)nx * ((size_t)(l - 1) + 4 * (size_t)(j - 1)); }
  size_t i2(int i, int j) const { return (size_t)(i - 1) + (size_t)nx * (size_t)(j - 1); }
  size_t iE(int l, int i, int j) const { return (size_t)(l - 1) + 4 * ((size_t)(i - 1) + (size_t)(nx - 1) * (size_t)(j - 1)); }
  size_t iF(int l, int i, int j) const { return (size_t)(l - 1) + 4 * ((size_t)(i - 1) + (size_t)(nx - 2) * (size_t)(j - 1)); }
  size_t nstate() const { return (size_t)nx * 4 * ny; }

  bool init(const ox2d_config& cf) {
    cfg = cf;
    if (cf.nx < 8 || cf.ny < 8) return fail("nx, ny must be >= 8");
    if (cf.order != 2 && cf.order != 4 && cf.order != 6) return fail("order must be 2, 4 or 6");
    if (cf.scheme != OX_SCHEME_SLAU && cf.scheme != OX_SCHEME_ROE) return fail("2-D scheme must be SLAU or Roe");
    if (cf.rk != 3) return fail("2-D cases advance with RK3");
    if (cf.bc != OX2D_BC_OBLIQUE_SHOCK) return fail("unknown 2-D boundary kind");
    nx = cf.nx;
    ny = cf.ny;
    io = cf.order / 2 - 1;
    ox_config c3{};
    c3.gamma = cf.gamma;
    c3.R = cf.R;
    c3.Pr = 0.72;
    c3.Prt = 0.9;
    c3.threshold = cf.threshold;
    c.init(c3);
    for (int l = 0; l < 4; ++l) qin[l] = cf.inflow[l], qpost[l] = cf.post[l];
    // set_grid (OS/set.f90:8-26)
    const double dx0 = cf.Lx / (double)(nx - 1), dy0 = cf.Ly / (double)(ny - 1);
    dx.assign(nx - 1, dx0);
    dy.assign(ny - 1, dy0);
    // set_Jacobian_xy2 (src/set_coordinate.f90:78-96)
    Jac.assign((size_t)nx * ny, 0.0);
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) Jac[i2(i, j)] = 4.0 / ((dx[i - 2] + dx[i - 1]) * (dy[j - 2] + dy[j - 1]));
    for (int i = 1; i <= nx; ++i) {
      Jac[i2(i, 1)] = Jac[i2(i, 2)];
      Jac[i2(i, ny)] = Jac[i2(i, ny - 1)];
    }
    for (int j = 1; j <= ny; ++j) {
      Jac[i2(1, j)] = Jac[i2(2, j)];
      Jac[i2(nx, j)] = Jac[i2(nx - 1, j)];
    }
    // pre_calc (preprocess.f90.fypp:93-102)
    xix.resize(nx - 1);
    etay.resize(ny - 1);
    for (int i = 0; i < nx - 1; ++i) xix[i] = 1.0 / dx[i];
    for (int j = 0; j < ny - 1; ++j) etay[j] = 1.0 / dy[j];
    dtdx.resize(nx - 2);
    dtdy.resize(ny - 2);
    for (int i = 1; i <= nx - 2; ++i) dtdx[i - 1] = 0.5 * cf.dt * (dx[i - 1] + dx[i]);
    for (int j = 1; j <= ny - 2; ++j) dtdy[j - 1] = 0.5 * cf.dt * (dy[j - 1] + dy[j]);
    QJ.assign(nstate(), 0.0);
    QJ2.assign(nstate(), 0.0);
    Qp.assign(nstate(), 0.0);
    T.assign((size_t)nx * ny, 0.0);
    sensor.assign((size_t)nx * ny, 0.0);
    E.assign((size_t)4 * (nx - 1) * (ny - 2), 0.0);
    F.assign((size_t)4 * (nx - 2) * (ny - 1), 0.0);
    return true;
  }
  bool fail(const char* m) {
    err = m;
    return false;
  }

  // calc_quantities_2D (src/calc_physical_quantities.f90:7-29)
  void primitives(const std::vector<double>& q) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double over_Q1 = 1.0 / q[iq(i, 1, j)];
        const double rho = Jac[i2(i, j)] * q[iq(i, 1, j)];
        const double u = q[iq(i, 2, j)] * over_Q1;
        const double v = q[iq(i, 3, j)] * over_Q1;
        const double p = c.gamma_1 * (Jac[i2(i, j)] * q[iq(i, 4, j)] - 0.5 * rho * (u * u + v * v));
        Qp[iq(i, 1, j)] = rho;
        Qp[iq(i, 2, j)] = u;
        Qp[iq(i, 3, j)] = v;
        Qp[iq(i, 4, j)] = p;
        T[i2(i, j)] = p / (c.Rg * rho);
      }
  }

  // calc_Ducros, 2-D (calc_hybrid.f90:11-52); dx, dy there are 1/dx, 1/dy
  void ducros() {
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        const double dx_tmp = 0.25 * (xix[i - 2] + xix[i - 1]);
        const double dy_tmp = 0.25 * (etay[j - 2] + etay[j - 1]);
        const double dudx = (-Qp[iq(i - 1, 2, j)] + Qp[iq(i + 1, 2, j)]) * dx_tmp;
        const double dvdx = (-Qp[iq(i - 1, 3, j)] + Qp[iq(i + 1, 3, j)]) * dx_tmp;
        const double dudy = (-Qp[iq(i, 2, j - 1)] + Qp[iq(i, 2, j + 1)]) * dy_tmp;
        const double dvdy = (-Qp[iq(i, 3, j - 1)] + Qp[iq(i, 3, j + 1)]) * dy_tmp;
        const double div = dudx + dvdy;
        const double rot = dvdx - dudy;
        double fd = (div * div) / (div * div + rot * rot + 1.0e-12);
        fd = fmin2(1.0, fd);
        sensor[i2(i, j)] = fd;
      }
    // boundary copies (:39-50), after the interior values they read
    for (int j = 2; j <= ny - 1; ++j) {
      sensor[i2(1, j)] = sensor[i2(2, j)];
      sensor[i2(nx, j)] = sensor[i2(nx - 1, j)];
    }
    for (int i = 2; i <= nx - 1; ++i) {
      sensor[i2(i, 1)] = sensor[i2(i, 2)];
      sensor[i2(i, ny)] = sensor[i2(i, ny - 1)];
    }
  }

  // One face of the boundary kernels.  dir 0: x face between points f and
  // f+1 of a row; dir 1: y face.  st[v][m] holds primitive v (rho, u, v, p) at
  // the point f-2+m; n = points along the direction; fs = face sensor.
  // calc_slau_kernel.f90.fypp:45-95 (x) and :140-190 (y).
  void face(int dir, int f, int n, const double st[4][6], double fs, double out[4]) const {
    int N = 2;
    if (cfg.order == 6 && 3 <= f && f <= n - 3) N = 6;
    else if (cfg.order >= 4 && 2 <= f && f <= n - 2) N = 4;
    double L[4], Rr[4], hr;
    if (N == 6) {
      for (int v = 0; v < 4; ++v) delta6(c, cfg.tvd, fs, &st[v][0], L[v], Rr[v]);
      hr = wiggle_detector(&st[3][1]);
    } else if (N == 4) {
      for (int v = 0; v < 4; ++v) delta4(c, cfg.tvd, fs, &st[v][1], L[v], Rr[v]);
      hr = wiggle_detector(&st[3][1]);
    } else {
      for (int v = 0; v < 4; ++v) L[v] = st[v][2], Rr[v] = st[v][3];
      hr = 1.0;
    }
    if (cfg.scheme == OX_SCHEME_SLAU) {
      static const double Nx[5] = {0, 1, 0, 0, 0}, Ny[5] = {0, 0, 1, 0, 0};
      double F5[5];
      const double un1 = dir == 0 ? L[1] : L[2], un2 = dir == 0 ? Rr[1] : Rr[2];
      slau_flux(c, cfg.slau_variant, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], 0.0, 0.0, un1, un2, L[3], Rr[3],
                dir == 0 ? Nx : Ny, hr, F5);
      out[0] = F5[0];
      out[1] = F5[1];
      out[2] = F5[2];
      out[3] = F5[4];
    } else {
      double o[5];
      if (dir == 0) {  // local frame (u, v, w)
        roe_flux(c, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], 0.0, 0.0, L[3], Rr[3], o);
        out[0] = o[0];
        out[1] = o[1];
        out[2] = o[2];
        out[3] = o[4];
      } else {  // local frame (v, w, u)
        roe_flux(c, L[0], Rr[0], L[2], Rr[2], 0.0, 0.0, L[1], Rr[1], L[3], Rr[3], o);
        out[0] = o[0];
        out[1] = o[3];
        out[2] = o[1];
        out[3] = o[4];
      }
    }
  }

  // calc_slau_x / calc_slau_y (calc_slau_kernel.f90.fypp:17-200): x faces
  // i = 1..nx-1 of rows j = 2..ny-1, y faces j = 1..ny-1 of columns 2..nx-1
  void conv() {
    const bool sens = cfg.order >= 4;
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        double st[4][6] = {};
        for (int m = 0; m < 6; ++m) {
          const int ii = i - 2 + m;
          if (ii < 1 || ii > nx) continue;
          for (int v = 0; v < 4; ++v) st[v][m] = Qp[iq(ii, v + 1, j)];
        }
        const double fs = sens ? 0.5 * (sensor[i2(i, j)] + sensor[i2(i + 1, j)]) : 0.0;
        double out[4];
        face(0, i, nx, st, fs, out);
        for (int l = 0; l < 4; ++l) E[iE(l + 1, i, j - 1)] = out[l];
      }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        double st[4][6] = {};
        for (int m = 0; m < 6; ++m) {
          const int jj = j - 2 + m;
          if (jj < 1 || jj > ny) continue;
          for (int v = 0; v < 4; ++v) st[v][m] = Qp[iq(i, v + 1, jj)];
        }
        const double fs = sens ? 0.5 * (sensor[i2(i, j)] + sensor[i2(i, j + 1)]) : 0.0;
        double out[4];
        face(1, j, ny, st, fs, out);
        for (int l = 0; l < 4; ++l) F[iF(l + 1, i - 1, j)] = out[l];
      }
  }

  // calc_EF, Euler (calc_flux_base.f90.fypp:87-99 + calc_conv :59-82)
  void calc_EF(const std::vector<double>& q) {
    primitives(q);
    ducros();
    conv();
  }

  // calc_R (calc_steps.f90:8-28): R = cx*dE + cy*dF for the cell (i+1, j+1)
  void calc_R(int i, int j, double c_dtdx, double c_dtdy, double* R) const {
    for (int l = 1; l <= 4; ++l) R[l - 1] = c_dtdy * (-E[iE(l, i, j)] + E[iE(l, i + 1, j)]);
    for (int l = 1; l <= 4; ++l) R[l - 1] = R[l - 1] + c_dtdx * (-F[iF(l, i, j)] + F[iF(l, i, j + 1)]);
  }

  // calc_step1 (calc_steps.f90:31-54)
  void step1(double coef, const std::vector<double>& Q, std::vector<double>& Q2) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i) {
        double R[4];
        calc_R(i, j, coef * dtdx[i - 1], coef * dtdy[j - 1], R);
        for (int l = 1; l <= 4; ++l) Q2[iq(i + 1, l, j + 1)] = Q[iq(i + 1, l, j + 1)] - R[l - 1];
      }
  }

  // calc_step2_3 (calc_steps.f90:89-117)
  void step2_3(double c1, double c2, double c3, double c4inv, const std::vector<double>& Qin, std::vector<double>& Qout) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i) {
        double R[4];
        calc_R(i, j, c3 * dtdx[i - 1], c3 * dtdy[j - 1], R);
        for (int l = 1; l <= 4; ++l) {
          const size_t m = iq(i + 1, l, j + 1);
          Qout[m] = (c1 * Qin[m] + c2 * Qout[m] - R[l - 1]) * c4inv;
        }
      }
  }

  // The conserved boundary value of a primitive state (rho, u, v, p) divided
  // by the Jacobian, in the operation order of OS/set.f90:55-65.
  void bc_state(const double* s, double Jt, double* q) const {
    q[0] = s[0] * Jt;
    q[1] = s[0] * s[1] * Jt;
    q[2] = s[0] * s[2] * Jt;
    q[3] = (s[3] * c.over_gamma_1 + 0.5 * s[0] * (s[1] * s[1] + s[2] * s[2])) * Jt;
  }

  // set_bc (OS/set.f90:48-86): two passes, the second (left/right columns)
  // after the first (top/bottom rows)
  void set_bc(std::vector<double>& Q) const {
    const int split = cfg.top_split;
    for (int i = 1; i <= nx; ++i) {
      const double Jt = 1.0 / Jac[i2(1, ny - 1)];
      double q[4];
      bc_state(i < split ? qin : qpost, Jt, q);
      for (int l = 1; l <= 4; ++l) Q[iq(i, l, ny)] = q[l - 1];
      Q[iq(i, 1, 1)] = Q[iq(i, 1, 2)];
      Q[iq(i, 2, 1)] = Q[iq(i, 2, 2)];
      Q[iq(i, 3, 1)] = -Q[iq(i, 3, 2)];
      Q[iq(i, 4, 1)] = Q[iq(i, 4, 2)];
    }
    for (int j = 1; j <= ny; ++j) {
      const double Jt = 1.0 / Jac[i2(1, j)];
      double q[4];
      bc_state(qin, Jt, q);
      for (int l = 1; l <= 4; ++l) Q[iq(1, l, j)] = q[l - 1];
      for (int l = 1; l <= 4; ++l) Q[iq(nx, l, j)] = Q[iq(nx - 1, l, j)];
    }
  }

  // RungeKutta, RK3 (calc_time_dev.f90.fypp:77-94)
  void rk_step() {
    calc_EF(QJ);
    step1(1.0, QJ, QJ2);
    set_bc(QJ2);
    calc_EF(QJ2);
    step2_3(0.75, 0.25, 0.25, 1.0, QJ, QJ2);
    set_bc(QJ2);
    calc_EF(QJ2);
    step2_3(2.0, 1.0, 2.0, c.one_third, QJ2, QJ);
    set_bc(QJ);
  }
};

}  // namespace oxo

extern "C" {

struct oxo2d_handle {
  oxo::Oracle2D o;
};

static thread_local std::string g_oxo2d_err;

const char* oxo2d_last_error(void) { return g_oxo2d_err.c_str(); }

void* oxo2d_create(const ox2d_config* cfg) {
  auto* h = new oxo2d_handle();
  if (!h->o.init(*cfg)) {
    g_oxo2d_err = h->o.err;
    delete h;
    return nullptr;
  }
  return h;
}

void oxo2d_destroy(void* h) { delete static_cast<oxo2d_handle*>(h); }

// pre_calc (preprocess.f90.fypp:86-92): QJ = Q / J; RungeKutta :61: QJ2 = QJ
int oxo2d_set_state(void* hv, const double* Q) {
  auto& o = static_cast<oxo2d_handle*>(hv)->o;
  for (int j = 1; j <= o.ny; ++j)
    for (int l = 1; l <= 4; ++l)
      for (int i = 1; i <= o.nx; ++i) o.QJ[o.iq(i, l, j)] = Q[o.iq(i, l, j)] / o.Jac[o.i2(i, j)];
  o.QJ2 = o.QJ;
  return 0;
}

// main.f90:80-85: Q = J * QJ
int oxo2d_get_state(void* hv, double* Q) {
  auto& o = static_cast<oxo2d_handle*>(hv)->o;
  for (int j = 1; j <= o.ny; ++j)
    for (int l = 1; l <= 4; ++l)
      for (int i = 1; i <= o.nx; ++i) Q[o.iq(i, l, j)] = o.Jac[o.i2(i, j)] * o.QJ[o.iq(i, l, j)];
  return 0;
}

int oxo2d_step(void* hv, int nsteps) {
  auto& o = static_cast<oxo2d_handle*>(hv)->o;
  for (int t = 0; t < nsteps; ++t) o.rk_step();
  return 0;
}

// Face flux of one x (dir 0) or y (dir 1) face, for the closed-form tests:
// st = [4][6] primitives at points f-2..f+3, n = points along the direction.
void oxo2d_face_flux(const ox2d_config* cfg, int dir, int f, int n, const double* st, double fs, double* out) {
  oxo::Oracle2D o;
  o.cfg = *cfg;
  ox_config c3{};
  c3.gamma = cfg->gamma;
  c3.R = cfg->R;
  c3.Pr = 0.72;
  c3.Prt = 0.9;
  c3.threshold = cfg->threshold;
  o.c.init(c3);
  double s[4][6];
  std::memcpy(s, st, sizeof(s));
  o.face(dir, f, n, s, fs, out);
}

}  // extern "C"
