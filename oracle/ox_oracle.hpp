// ox_oracle.hpp -- FP64 CPU transcription of the reference RHS + Runge-Kutta
// path (the ORACLE).
//
// TEST INFRASTRUCTURE ONLY.  This file is the CPU reference that the HIP
// library is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may build, link or call it; the product never does.
//
// It is a statement-by-statement restatement of the reference's CUDA Fortran
// routines in plain C++17 + OpenMP (FP64).  Fortran 1-based indices and the
// reference's own loop ranges are kept, so each routine can be read side by
// side with the file:line it cites.  Evaluation order follows the Fortran
// source left to right; build with -ffp-contract=off so that no fused
// multiply-add changes a rounding.
//
// Parity pins (tests/test_oracle.py, tests/test_oracle_pins.py,
// tests/test_golden.py):
//   grid / Jacobian / pre_calc   TGV initial KE closed form
//   primitives                   primitives rebuild the conserved state;
//                                T = p/(R rho); Sutherland literals
//   Ducros sensor                u(y),v(x) -> exactly 0; u(x),v(y),w(z) -> 1
//   KEEP2/4/6                    F(q,q) = Euler flux, free stream -> RHS == 0,
//                                periodic conservation, convergence order
//                                (reference thresholds 3.6 / 5.7)
//   MUSCL (all TVD kinds)        exact on linear data, minmod bounds
//   wiggle detector              monotone -> 0, oscillating -> 1
//   SLAU1 / HRSLAU2 / Roe        F(q,q) = Euler flux, mirror symmetry,
//                                supersonic upwinding; hybrid reduces to
//                                KEEP / SLAU(HR=1)
//   central viscous 4/6          discrete symbol on a shear wave (textbook
//                                staggered weights), order on a shear wave
//   RK3 / RK4 steps              Shu-Osher and classical tableaux composed
//                                from the oracle's own right-hand sides;
//                                time order
//   periodic fill                ghost == periodic image, bitwise
//   slip / zero-gradient fill    mirror images by hand; Sod's problem along z
//                                between slip walls against the exact solution
//   LES (SMS)                    off for a single mode, on for a multi-mode
//   everything above             golden digests (tests/golden/)
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/ouxsbli.h"

namespace oxo {

// ---------------------------------------------------------------------------
// Scalar helpers with Fortran semantics.
// ---------------------------------------------------------------------------
// MIN/MAX of two reals: select form, first operand wins ties (the same select
// the HIP kernels use, so signed zeros agree).
template <class R> inline R fmin2(R a, R b) { return (b < a) ? b : a; }
template <class R> inline R fmax2(R a, R b) { return (b > a) ? b : a; }
// SIGN(1.d0, x): IEEE sign of x, -0 -> -1.
inline double fsign1(double x) { return std::copysign(1.0, x); }
inline double fabs_(double x) { return std::fabs(x); }
inline double fsqrt(double x) { return std::sqrt(x); }

// ---------------------------------------------------------------------------
// Constants.  src/mod_constant.f90.fypp:18-35 and the module constants of
// 3D_solver/src/calc_keep_kernel_internal.f90.fypp:7-12.
// ---------------------------------------------------------------------------
struct Consts {
  double gamma, Rg, Pr, Prt, threshold;
  double gamma_1, over_gamma, over_gamma_1, R_over_gamma_1, Cp, over_Cp, Cp_over_Pr;
  double mu0_T0_S_over_T0_2_3;
  double one_third, two_third, one_sixth, one_twelfth, one_sixty, tan10deg2;
  double one_24, one_48, one_60, one_120, one_240, seven_twelfth;
  double gamma_over_gamma_1;  // calc_slau_3d.f90:36 (phi)
  double minmod_b;            // calc_muscl.f90:70
  void init(const ox_config& c) {
    gamma = c.gamma; Rg = c.R; Pr = c.Pr; Prt = c.Prt; threshold = c.threshold;
    gamma_1 = gamma - 1.0;
    over_gamma = 1.0 / gamma;
    over_gamma_1 = 1.0 / (gamma - 1.0);
    R_over_gamma_1 = Rg / (gamma - 1.0);
    Cp = gamma * Rg / (gamma - 1.0);
    over_Cp = 1.0 / Cp;
    Cp_over_Pr = Cp / Pr;
    mu0_T0_S_over_T0_2_3 = 1.716e-5 * 384.2 * std::pow(273.2, -1.5);
    one_third = 1.0 / 3.0;
    two_third = 2.0 / 3.0;
    one_sixth = 1.0 / 6.0;
    one_twelfth = 1.0 / 12.0;
    one_sixty = 1.0 / 60.0;
    const double t = std::tan(10.0 * std::acos(-1.0) / 180.0);
    tan10deg2 = t * t;
    one_24 = 1.0 / 24.0;
    one_48 = 1.0 / 48.0;
    one_60 = 1.0 / 60.0;
    one_120 = 1.0 / 120.0;
    one_240 = 1.0 / 240.0;
    seven_twelfth = 7.0 / 12.0;
    gamma_over_gamma_1 = gamma * over_gamma_1;
    minmod_b = (3.0 - one_third) / (1.0 - one_third);
  }
};

// ---------------------------------------------------------------------------
// KEEP split-form fluxes.  3D_solver/src/calc_keep_3d.f90:1-111.
// Stencil arrays are 0-based here: a[m] is the Fortran a(m+1).
// ---------------------------------------------------------------------------
inline void keep2(const Consts& c, const double* rho, const double* u, const double* v,
                  const double* w, const double* uu, const double* p, const double* T,
                  const double* N, double* F) {
  F[0] = 0.25 * (rho[0] + rho[1]) * (uu[0] + uu[1]);
  F[1] = 0.5 * (F[0] * (u[0] + u[1]) + (p[0] + p[1]) * N[1]);
  F[2] = 0.5 * (F[0] * (v[0] + v[1]) + (p[0] + p[1]) * N[2]);
  F[3] = 0.5 * (F[0] * (w[0] + w[1]) + (p[0] + p[1]) * N[3]);
  F[4] = F[0] * 0.5 * (T[0] + T[1]) * c.R_over_gamma_1;
  F[4] = F[4] + 0.5 * (uu[0] * p[1] + uu[1] * p[0]);
  F[4] = F[4] + 0.5 * F[0] * (u[0] * u[1] + v[0] * v[1] + w[0] * w[1]);
}

inline void keep4(const Consts& c, const double* rho, const double* u, const double* v,
                  const double* w, const double* uu, const double* p, const double* T,
                  const double* N, double* F) {
  double RV1 = (rho[1] + rho[2]) * (uu[1] + uu[2]);
  double RV2 = (rho[1] + rho[3]) * (uu[1] + uu[3]);
  double RV3 = (rho[0] + rho[2]) * (uu[0] + uu[2]);
  F[0] = c.one_third * RV1 - (RV2 + RV3) * c.one_24;
  RV1 = c.one_sixth * RV1;
  RV2 = c.one_48 * RV2;
  RV3 = c.one_48 * RV3;
  const double RV1_RV2 = RV1 - RV2;
  const double RV1_RV3 = RV1 - RV3;
  const double u1 = u[0], u2 = u[1], u3 = u[2], u4 = u[3];
  const double v1 = v[0], v2 = v[1], v3 = v[2], v4 = v[3];
  const double w1 = w[0], w2 = w[1], w3 = w[2], w4 = w[3];
  {
    const double pres = -(c.one_twelfth * p[0]) + c.seven_twelfth * (p[1] + p[2]) - c.one_twelfth * p[3];
    F[1] = -(RV3 * u1) + RV1_RV2 * u2 + RV1_RV3 * u3 - RV2 * u4 + pres * N[1];
    F[2] = -(RV3 * v1) + RV1_RV2 * v2 + RV1_RV3 * v3 - RV2 * v4 + pres * N[2];
    F[3] = -(RV3 * w1) + RV1_RV2 * w2 + RV1_RV3 * w3 - RV2 * w4 + pres * N[3];
  }
  {
    double ene = (-(RV3 * T[0]) + RV1_RV2 * T[1] + RV1_RV3 * T[2] - RV2 * T[3]) * c.R_over_gamma_1;
    ene = ene + ((RV1 * (u2 * u3 + v2 * v3 + w2 * w3)) -
                 (RV2 * (u2 * u4 + v2 * v4 + w2 * w4) + RV3 * (u1 * u3 + v1 * v3 + w1 * w3)));
    F[4] = ene + (c.two_third * (uu[1] * p[2] + uu[2] * p[1]) -
                  c.one_twelfth * (uu[1] * p[3] + uu[3] * p[1] + uu[0] * p[2] + uu[2] * p[0]));
  }
}

inline void keep6(const Consts& c, const double* rho, const double* u, const double* v,
                  const double* w, const double* uu, const double* p, const double* T,
                  const double* N, double* F) {
  double RV1 = (rho[2] + rho[3]) * (uu[2] + uu[3]);
  double RV2 = (rho[2] + rho[4]) * (uu[2] + uu[4]);
  double RV3 = (rho[1] + rho[3]) * (uu[1] + uu[3]);
  double RV4 = (rho[2] + rho[5]) * (uu[2] + uu[5]);
  double RV5 = (rho[1] + rho[4]) * (uu[1] + uu[4]);
  double RV6 = (rho[0] + rho[3]) * (uu[0] + uu[3]);
  F[0] = 0.375 * RV1 - 0.075 * (RV2 + RV3) + (RV4 + RV5 + RV6) * c.one_120;
  RV1 = 0.1875 * RV1;
  RV2 = 0.0375 * RV2;
  RV3 = 0.0375 * RV3;
  RV4 = RV4 * c.one_240;
  RV5 = RV5 * c.one_240;
  RV6 = RV6 * c.one_240;
  const double RV3_RV5 = -RV3 + RV5;
  const double RV1_RV2_RV4 = RV1 - RV2 + RV4;
  const double RV1_RV3_RV6 = RV1 - RV3 + RV6;
  const double RV2_RV5 = -RV2 + RV5;
  {
    const double pres = (p[0] - 8.0 * p[1] + 37.0 * (p[2] + p[3]) - 8.0 * p[4] + p[5]) * c.one_60;
    double ruu = RV6 * u[0] + RV3_RV5 * u[1] + RV1_RV2_RV4 * u[2] + RV1_RV3_RV6 * u[3] + RV2_RV5 * u[4] + RV4 * u[5];
    ruu = ruu + pres * N[1];
    double ruv = RV6 * v[0] + RV3_RV5 * v[1] + RV1_RV2_RV4 * v[2] + RV1_RV3_RV6 * v[3] + RV2_RV5 * v[4] + RV4 * v[5];
    ruv = ruv + pres * N[2];
    double ruw = RV6 * w[0] + RV3_RV5 * w[1] + RV1_RV2_RV4 * w[2] + RV1_RV3_RV6 * w[3] + RV2_RV5 * w[4] + RV4 * w[5];
    ruw = ruw + pres * N[3];
    F[1] = ruu;
    F[2] = ruv;
    F[3] = ruw;
  }
  {
    double ene = (RV6 * T[0] + RV3_RV5 * T[1] + RV1_RV2_RV4 * T[2] + RV1_RV3_RV6 * T[3] + RV2_RV5 * T[4] + RV4 * T[5]) *
                 c.R_over_gamma_1;
    ene = ene + RV1 * (u[2] * u[3] + v[2] * v[3] + w[2] * w[3]) -
          (RV2 * (u[2] * u[4] + v[2] * v[4] + w[2] * w[4]) + RV3 * (u[1] * u[3] + v[1] * v[3] + w[1] * w[3])) +
          (RV4 * (u[2] * u[5] + v[2] * v[5] + w[2] * w[5]) + RV5 * (u[1] * u[4] + v[1] * v[4] + w[1] * w[4]) +
           RV6 * (u[0] * u[3] + v[0] * v[3] + w[0] * w[3]));
    F[4] = ene + (0.75 * (uu[2] * p[3] + uu[3] * p[2]) -
                  0.15 * (uu[2] * p[4] + uu[4] * p[2] + uu[1] * p[3] + uu[3] * p[1]) +
                  (uu[2] * p[5] + uu[5] * p[2] + uu[1] * p[4] + uu[4] * p[1] + uu[0] * p[3] + uu[3] * p[0]) * c.one_60);
  }
}

// ---------------------------------------------------------------------------
// MUSCL reconstruction.  src/calc_muscl.f90:19-169.
// ---------------------------------------------------------------------------
inline double minmod2(double x, double y) {  // :19-24
  const double sgn = fsign1(x);
  return sgn * fmax2(fmin2(fabs_(x), sgn * y), 0.0);
}
inline double minmod3(double x, double y, double z) {  // :27-32
  const double sgn = fsign1(x);
  return sgn * fmax2(fmin2(fmin2(fabs_(x), sgn * y), sgn * z), 0.0);
}
inline double d33(double d1, double d2, double d3) {  // :35-41
  return minmod3(d1, 2.0 * d2, 2.0 * d3) - 2.0 * minmod3(d2, 2.0 * d1, 2.0 * d3) + minmod3(d3, 2.0 * d1, 2.0 * d2);
}
inline void muscl3_nontvd(const Consts& c, double a2, double a3, double d1, double d2, double d3, double& al,
                          double& ar) {  // :44-52
  al = a2 + 0.5 * (d1 + 2.0 * d2) * c.one_third;
  ar = a3 - 0.5 * (d3 + 2.0 * d2) * c.one_third;
}
inline void muscl3_minmod(const Consts& c, double a2, double a3, double d1, double d2, double d3, double& al,
                          double& ar) {  // :55-74
  const double b = c.minmod_b;
  {
    const double dt1 = minmod2(d1, b * d2);
    const double dt2 = minmod2(d2, b * d1);
    al = a2 + 0.5 * (dt1 + 2.0 * dt2) * c.one_third;
  }
  {
    const double dt3 = minmod2(d3, b * d2);
    const double dt4 = minmod2(d2, b * d3);
    ar = a3 - 0.5 * (dt3 + 2.0 * dt4) * c.one_third;
  }
}
inline void muscl4_nontvd(const Consts& c, double a2, double a3, double d1, double d2, double d3, double d4,
                          double d5, double& al, double& ar) {  // :94-102
  al = a2 + (-0.4 * d1 + 2.2 * d2 + 4.8 * d3 - 0.6 * d4) * c.one_twelfth;
  ar = a3 - (-0.6 * d2 + 4.8 * d3 + 2.2 * d4 - 0.4 * d5) * c.one_twelfth;
}
inline void muscl4_tvd(const Consts& c, double a2, double a3, double d1, double d2, double d3, double d4, double d5,
                       double& al, double& ar) {  // :105-124
  const double delta2 = d3 - d33(d2, d3, d4) * c.one_sixth;
  {
    const double delta1 = d2 - d33(d1, d2, d3) * c.one_sixth;
    const double dl = minmod2(delta1, 4.0 * delta2);
    const double dr = minmod2(delta2, 4.0 * delta1);
    al = a2 + (dl + 2.0 * dr) * c.one_sixth;
  }
  {
    const double delta3 = d4 - d33(d3, d4, d5) * c.one_sixth;
    const double dl = minmod2(delta2, 4.0 * delta3);
    const double dr = minmod2(delta3, 4.0 * delta2);
    ar = a3 - (dr + 2.0 * dl) * c.one_sixth;
  }
}
// delta4 / delta6 with the TVD kind of config.fypp TVD (mod_constant.f90.fypp:51-57)
inline void delta4(const Consts& c, int tvd, double sensor, const double* a, double& al, double& ar) {  // :142-153
  const double d1 = -a[0] + a[1];
  const double d2 = -a[1] + a[2];
  const double d3 = -a[2] + a[3];
  if (tvd == OX_TVD_NONE || (tvd == OX_TVD_HYBRID && sensor < c.threshold))
    muscl3_nontvd(c, a[1], a[2], d1, d2, d3, al, ar);
  else
    muscl3_minmod(c, a[1], a[2], d1, d2, d3, al, ar);
}
inline void delta6(const Consts& c, int tvd, double sensor, const double* a, double& al, double& ar) {  // :156-169
  const double d1 = -a[0] + a[1];
  const double d2 = -a[1] + a[2];
  const double d3 = -a[2] + a[3];
  const double d4 = -a[3] + a[4];
  const double d5 = -a[4] + a[5];
  if (tvd == OX_TVD_NONE || (tvd == OX_TVD_HYBRID && sensor < c.threshold))
    muscl4_nontvd(c, a[2], a[3], d1, d2, d3, d4, d5, al, ar);
  else
    muscl4_tvd(c, a[2], a[3], d1, d2, d3, d4, d5, al, ar);
}

// 3D_solver/src/calc_hybrid.f90:75-81
inline double wiggle_detector(const double* phi) {
  const double phi1 = (-phi[0] + phi[1]) * (-phi[1] + phi[2]);
  const double phi2 = (-phi[2] + phi[3]) * (-phi[1] + phi[2]);
  return 0.5 * (1.0 - fsign1(fmin2(phi1, phi2)));
}

// ---------------------------------------------------------------------------
// SLAU family.  3D_solver/src/calc_slau_3d.f90:1-104.
// ---------------------------------------------------------------------------
inline void slau_common(const Consts& c, double rho1, double rho2, double over_rho1, double over_rho2, double un1,
                        double un2, double p1, double p2, double& cc, double& over_c, double& Mp, double& Mm,
                        double& bp, double& bm, double& dp, double& Vtp, double& Vtm) {  // :1-33
  {
    const double cl = fsqrt(c.gamma * p1 * over_rho1);
    const double cr = fsqrt(c.gamma * p2 * over_rho2);
    cc = 0.5 * (cl + cr);
  }
  over_c = 1.0 / cc;
  Mp = un1 * over_c;
  Mm = un2 * over_c;
  {
    const double g = -(fmax2(fmin2(Mp, 0.0), -1.0) * fmin2(fmax2(Mm, 0.0), 1.0));
    const double one_g_Vt = (1.0 - g) * (rho1 * fabs_(un1) + rho2 * fabs_(un2)) / (rho1 + rho2);
    Vtp = one_g_Vt + g * fabs_(un1);
    Vtm = one_g_Vt + g * fabs_(un2);
  }
  bp = (fabs_(Mp) < 1.0) ? 0.25 * (2.0 - Mp) * ((Mp + 1.0) * (Mp + 1.0)) : 0.5 * (1.0 + fsign1(Mp));
  bm = (fabs_(Mm) < 1.0) ? 0.25 * (2.0 + Mm) * ((Mm - 1.0) * (Mm - 1.0)) : 0.5 * (1.0 + fsign1(-Mm));
  dp = -p1 + p2;
}
inline double slau_phi(const Consts& c, double rho, double k, double p, double over_rho) {  // :36-43
  return (p * c.gamma_over_gamma_1 + rho * k) * over_rho;
}
// SLAU1 (:46-72) and HRSLAU2 (:75-104); un = the face-normal velocity
inline void slau_flux(const Consts& c, int variant, double rho1, double rho2, double u1, double u2, double v1,
                      double v2, double w1, double w2, double un1, double un2, double p1, double p2, const double* N,
                      double HR, double* F) {
  double cc, over_c, Mp, Mm, bp, bm, dp, Vtp, Vtm;
  const double over_rho1 = 1.0 / rho1;
  const double over_rho2 = 1.0 / rho2;
  slau_common(c, rho1, rho2, over_rho1, over_rho2, un1, un2, p1, p2, cc, over_c, Mp, Mm, bp, bm, dp, Vtp, Vtm);
  const double k1 = 0.5 * (u1 * u1 + v1 * v1 + w1 * w1);
  const double k2 = 0.5 * (u2 * u2 + v2 * v2 + w2 * w2);
  double mass1, mass2, pres;
  if (variant == OX_SLAU_SLAU1) {
    const double M = fmin2(1.0, fsqrt(k1 + k2) * over_c);
    const double x = (1.0 - M) * (1.0 - M);
    const double mass = 0.25 * (rho1 * (un1 + Vtp) + rho2 * (un2 - Vtm) - x * dp * over_c);
    mass1 = mass + fabs_(mass);
    mass2 = mass - fabs_(mass);
    pres = 0.5 * (p1 + p2 + (bp - bm) * (-dp) + (1.0 - x) * (bp + bm - 1.0) * (p1 + p2));
  } else {
    const double Vec2 = fsqrt(k1 + k2);
    {
      const double M = fmin2(1.0, Vec2 * over_c);
      const double x = (1.0 - M) * (1.0 - M);
      const double mass = 0.25 * (rho1 * (un1 + Vtp) + rho2 * (un2 - Vtm) - x * dp * over_c);
      mass1 = mass + fabs_(mass);
      mass2 = mass - fabs_(mass);
    }
    pres = 0.5 * (p1 + p2 + (bp - bm) * (-dp) + HR * Vec2 * (bp + bm - 1.0) * 0.5 * (rho1 + rho2) * cc);
  }
  F[0] = mass1 + mass2;
  F[1] = mass1 * u1 + mass2 * u2 + pres * N[1];
  F[2] = mass1 * v1 + mass2 * v2 + pres * N[2];
  F[3] = mass1 * w1 + mass2 * w2 + pres * N[3];
  F[4] = mass1 * slau_phi(c, rho1, k1, p1, over_rho1) + mass2 * slau_phi(c, rho2, k2, p2, over_rho2);
}

// ---------------------------------------------------------------------------
// Roe flux.  3D_solver/src/calc_roe_3d.f90:1-88.  Arguments in the face's
// local frame: un = normal velocity, (t1, t2) = the other two components in
// cyclic order; o[5] = (mass, normal, t1, t2, energy) flux components.
// ---------------------------------------------------------------------------
inline void roe_flux(const Consts& c, double rho1, double rho2, double u1, double u2, double v1, double v2,
                     double w1, double w2, double p1, double p2, double* F) {
  const double k1 = 0.5 * (u1 * u1 + v1 * v1 + w1 * w1);
  const double k2 = 0.5 * (u2 * u2 + v2 * v2 + w2 * w2);
  double dQ[5];
  // calc_Roe_common (:57-83)
  const double e1 = p1 * c.over_gamma_1 + rho1 * k1;
  const double e2 = p2 * c.over_gamma_1 + rho2 * k2;
  {
    const double rhou1 = rho1 * u1;
    const double rhou2 = rho2 * u2;
    F[0] = 0.5 * (rhou1 + rhou2);
    F[1] = 0.5 * (rhou1 * u1 + rhou2 * u2 + (p1 + p2));
    F[2] = 0.5 * (rhou1 * v1 + rhou2 * v2);
    F[3] = 0.5 * (rhou1 * w1 + rhou2 * w2);
    F[4] = 0.5 * ((e1 + p1) * u1 + (e2 + p2) * u2);
  }
  dQ[0] = -rho1 + rho2;
  dQ[1] = -(rho1 * u1) + rho2 * u2;
  dQ[2] = -(rho1 * v1) + rho2 * v2;
  dQ[3] = -(rho1 * w1) + rho2 * w2;
  dQ[4] = -e1 + e2;
  // calc_Roe_ave (:39-54)
  const double H1 = c.gamma_over_gamma_1 * p1 / rho1 + k1;
  const double H2 = c.gamma_over_gamma_1 * p2 / rho2 + k2;
  const double rhol = fsqrt(rho1);
  const double rhor = fsqrt(rho2);
  const double rr = 1.0 / (rhol + rhor);
  const double u = (rhol * u1 + rhor * u2) * rr;
  const double v = (rhol * v1 + rhor * v2) * rr;
  const double w = (rhol * w1 + rhor * w2) * rr;
  const double H = (rhol * H1 + rhor * H2) * rr;
  const double cr = fsqrt(c.gamma_1 * (H - 0.5 * (u * u + v * v + w * w)));
  // calc_RARinv (:1-36)
  const double q2 = 0.5 * (u * u + v * v + w * w);
  const double oc = 1.0 / cr;
  const double b2 = c.gamma_1 * oc * oc;
  const double b1 = q2 * b2;
  const double R1 = 0.5 * ((b1 + u * oc) * dQ[0] - (oc + b2 * u) * dQ[1] - b2 * v * dQ[2] - b2 * w * dQ[3] + b2 * dQ[4]);
  const double R2 = oc * (w * dQ[0] - dQ[3]);
  const double R3 = oc * (-(v * dQ[0]) + dQ[2]);
  const double R4 = (1.0 - b1) * dQ[0] + b2 * (u * dQ[1] + v * dQ[2] + w * dQ[3] - dQ[4]);
  const double R5 = 0.5 * ((b1 - u * oc) * dQ[0] + (oc - b2 * u) * dQ[1] - b2 * (v * dQ[2] + w * dQ[3] - dQ[4]));
  const double a1 = fabs_(u - cr);
  const double a2 = fabs_(u);
  const double a3 = fabs_(u + cr);
  const double a2c = a2 * cr;
  const double cu = cr * u;
  F[0] = F[0] - 0.5 * (a1 * R1 + a2 * R4 + a3 * R5);
  F[1] = F[1] - 0.5 * (a1 * (u - cr) * R1 + a2 * u * R4 + a3 * (u + cr) * R5);
  F[2] = F[2] - 0.5 * (a1 * v * R1 + a2c * R3 + a2 * v * R4 + a3 * v * R5);
  F[3] = F[3] - 0.5 * (a1 * w * R1 - a2c * R2 + a2 * w * R4 + a3 * w * R5);
  F[4] = F[4] - 0.5 * (a1 * (H - cu) * R1 + a2 * (-(w * cr * R2) + v * cr * R3 + q2 * R4) + a3 * (H + cu) * R5);
}

// ---------------------------------------------------------------------------
// Central viscous helpers.  3D_solver/src/calc_visc_cent.f90.fypp:2-91.
// ---------------------------------------------------------------------------
inline double vinterp(int N, const double* a) {
  if (N == 4) return 0.0625 * (-a[0] + 9.0 * (a[1] + a[2]) - a[3]);
  return 0.00390625 * (3.0 * a[0] - 25.0 * a[1] + 150.0 * (a[2] + a[3]) - 25.0 * a[4] + 3.0 * a[5]);
}
inline double vdiff(const Consts& c, int N, double d, const double* a) {
  if (N == 4) return (a[0] - 27.0 * (a[1] - a[2]) - a[3]) * c.one_24 * d;
  return (-0.0046875 * a[0] + 0.065104167 * a[1] - 1.171875 * (a[2] - a[3]) - 0.065104167 * a[4] +
          0.0046875 * a[5]) *
         d;
}
inline void tau_straight(const Consts& c, int N, const double* mu, const double* u, const double* vy,
                         const double* wz, double d, double& t11, double& ut11) {
  const double mui = vinterp(N, mu);
  const double vyi = vinterp(N, vy);
  const double wzi = vinterp(N, wz);
  const double ux = vdiff(c, N, d, u);
  t11 = c.two_third * mui * (2.0 * ux - vyi - wzi);
  ut11 = vinterp(N, u) * t11;
}
inline void tau_cross(const Consts& c, int N, const double* mu, const double* v, const double* uy, double d,
                      double& t12, double& vt12) {
  const double mui = vinterp(N, mu);
  const double uyi = vinterp(N, uy);
  const double vx = vdiff(c, N, d, v);
  t12 = mui * (uyi + vx);
  vt12 = vinterp(N, v) * t12;
}
// LES variants (calc_visc_cent.f90.fypp:44-58, :76-91): the work term uses the
// molecular stress, the stress itself 0.5*(mu + mut)
inline void tau_straight_les(const Consts& c, int N, const double* mu, const double* mut, const double* u,
                             const double* vy, const double* wz, double d, double& t11, double& ut11) {
  double mui = vinterp(N, mu);
  const double vyi = vinterp(N, vy);
  const double wzi = vinterp(N, wz);
  const double ux = vdiff(c, N, d, u);
  t11 = c.two_third * mui * (2.0 * ux - vyi - wzi);
  ut11 = vinterp(N, u) * t11;
  mui = 0.5 * (mui + vinterp(N, mut));
  t11 = c.two_third * mui * (2.0 * ux - vyi - wzi);
}
inline void tau_cross_les(const Consts& c, int N, const double* mu, const double* mut, const double* v,
                          const double* uy, double d, double& t12, double& vt12) {
  double mui = vinterp(N, mu);
  const double uyi = vinterp(N, uy);
  const double vx = vdiff(c, N, d, v);
  t12 = mui * (uyi + vx);
  vt12 = vinterp(N, v) * t12;
  mui = 0.5 * (mui + vinterp(N, mut));
  t12 = mui * (uyi + vx);
}

// ---------------------------------------------------------------------------
// ME4Base viscous helpers.  3D_solver/src/calc_visc_me4_base.f90:4-111 and
// the face assembly of calc_visc_high_internal.f90.fypp:146-183: 4th-order
// face quantities at the three faces f-1, f, f+1 of a 6-cell window, combined
// by flux4(a) = (-a1 + 26 a2 - a3)/24.
// ---------------------------------------------------------------------------
inline double flux4(const Consts& c, const double* a) { return (-a[0] + 26.0 * a[1] - a[2]) * c.one_24; }
// 4-point interpolation to face t of a 6-cell window (cells t..t+3)
inline double me4_interp(const double* a, int t) { return 0.0625 * (9.0 * (a[t + 1] + a[t + 2]) - (a[t] + a[t + 3])); }
// the same with the operands in the order of the stress and work terms
inline double me4_interp_w(const double* a, int t) { return 0.0625 * (-a[t] + 9.0 * (a[t + 1] + a[t + 2]) - a[t + 3]); }
inline void tau_straight_me4(const Consts& c, const double* mu3, const double* u, const double* vy, const double* wz,
                             double d, double& t11, double& ut11) {
  double tmp[3], tu[3];
  for (int t = 0; t < 3; ++t)
    tmp[t] = c.two_third * mu3[t] *
             ((2.25 * (-u[t + 1] + u[t + 2]) - (-u[t] + u[t + 3]) * c.one_twelfth) * d - me4_interp_w(vy, t) -
              me4_interp_w(wz, t));
  t11 = flux4(c, tmp);
  for (int t = 0; t < 3; ++t) tu[t] = me4_interp_w(u, t) * tmp[t];
  ut11 = flux4(c, tu);
}
inline void tau_cross_me4(const Consts& c, const double* mu3, const double* v, const double* uy, double d, double& t12,
                          double& vt12) {
  double tmp[3], tv[3];
  for (int t = 0; t < 3; ++t)
    tmp[t] = mu3[t] * ((1.125 * (-v[t + 1] + v[t + 2]) - (-v[t] + v[t + 3]) * c.one_24) * d + me4_interp_w(uy, t));
  t12 = flux4(c, tmp);
  for (int t = 0; t < 3; ++t) tv[t] = me4_interp_w(v, t) * tmp[t];
  vt12 = flux4(c, tv);
}
// LES variants: the work term uses the molecular stress, the stress (mu + mut)
inline void tau_straight_me4_les(const Consts& c, const double* mu3, const double* mut3, const double* u,
                                 const double* vy, const double* wz, double d, double& t11, double& ut11) {
  double t1[3], t2[3];
  for (int t = 0; t < 3; ++t) {
    t1[t] = (2.25 * (-u[t + 1] + u[t + 2]) - (-u[t] + u[t + 3]) * c.one_twelfth) * d;
    t1[t] = t1[t] - me4_interp_w(vy, t);
    t1[t] = t1[t] - me4_interp_w(wz, t);
    t2[t] = c.two_third * mu3[t] * t1[t];
  }
  for (int t = 0; t < 3; ++t) t2[t] = me4_interp_w(u, t) * t2[t];
  ut11 = flux4(c, t2);
  for (int t = 0; t < 3; ++t) t2[t] = c.two_third * (mu3[t] + mut3[t]) * t1[t];
  t11 = flux4(c, t2);
}
inline void tau_cross_me4_les(const Consts& c, const double* mu3, const double* mut3, const double* v,
                              const double* uy, double d, double& t12, double& vt12) {
  double t1[3], t2[3];
  for (int t = 0; t < 3; ++t) {
    t1[t] = (1.125 * (-v[t + 1] + v[t + 2]) - (-v[t] + v[t + 3]) * c.one_24) * d;
    t1[t] = t1[t] + me4_interp_w(uy, t);
    t2[t] = mu3[t] * t1[t];
  }
  for (int t = 0; t < 3; ++t) t2[t] = me4_interp_w(v, t) * t2[t];
  vt12 = flux4(c, t2);
  for (int t = 0; t < 3; ++t) t2[t] = (mu3[t] + mut3[t]) * t1[t];
  t12 = flux4(c, t2);
}

// ---------------------------------------------------------------------------
// The solver state.
// ---------------------------------------------------------------------------
struct Oracle {
  ox_config cfg{};
  Consts c{};
  int nx = 0, ny = 0, nz = 0, order = 0, io = 0, off = 0, vorder = 0;
  bool me4 = false;  // ME4Base viscous stencil (VISC_ORDER 6, VISC_STENCIL 'ME4Base')
  std::vector<double> dx, dy, dz;        // grid spacings (n-1)
  std::vector<double> xix, etay, zetaz;  // 1/dx etc. (n-1)
  std::vector<double> Jac;               // (nx,ny)
  std::vector<double> dtdxdy, dtdydz, dtdzdx;
  std::vector<double> QJ, QJ2, Rs;       // QJ2 is QJs for RK4
  std::vector<double> Qp, T, mu, sensor, mut, qc2;
  double les_delta15 = 0.0;  // delta**(1+alpha) of SMS (calc_les.f90:83-85), grid constant
  std::vector<double> E, F, G;
  std::string err;

  // Fortran-order index helpers (1-based).
  size_t iq(int i, int l, int j, int k) const {
    return (size_t)(i - 1) + (size_t)nx * ((size_t)(l - 1) + 5 * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1)));
  }
  size_t i3(int i, int j, int k) const {
    return (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
  }
  size_t iE(int l, int i, int j, int k) const {
    return (size_t)(l - 1) + 5 * ((size_t)(i - 1) + (size_t)(nx - 1) * ((size_t)(j - 1) + (size_t)(ny - 2) * (size_t)(k - 1)));
  }
  size_t iF(int l, int i, int j, int k) const {
    return (size_t)(l - 1) + 5 * ((size_t)(i - 1) + (size_t)(nx - 2) * ((size_t)(j - 1) + (size_t)(ny - 1) * (size_t)(k - 1)));
  }
  size_t iG(int l, int i, int j, int k) const {
    return (size_t)(l - 1) + 5 * ((size_t)(i - 1) + (size_t)(nx - 2) * ((size_t)(j - 1) + (size_t)(ny - 2) * (size_t)(k - 1)));
  }
  size_t iRs(int i, int l, int j, int k) const {
    return (size_t)(i - 1) + (size_t)(nx - 2) * ((size_t)(l - 1) + 5 * ((size_t)(j - 1) + (size_t)(ny - 2) * (size_t)(k - 1)));
  }
  size_t nstate() const { return (size_t)nx * 5 * ny * nz; }
  size_t ncell() const { return (size_t)nx * ny * nz; }

  bool init(const ox_config& cf);

  // pieces of the path
  void primitives(const std::vector<double>& q);  // calc_quantities[_T]_3D
  void ducros();                                   // calc_Ducros
  void conv_x(); void conv_y(); void conv_z();     // calc_<scheme>_{x,y,z}_in
  void visc_x(); void visc_y(); void visc_z();     // calc_{E,F,G}v{4,6}[_in]
  void calc_mut();                                 // calc_les.f90:96-125 + set_bc_mut_common
  void calc_EFG(const std::vector<double>& q);     // calc_flux_base.f90.fypp:109-198
  void calc_R(int i, int j, int k, double cxy, double cyz, double czx, double* R) const;
  void step1(double coef, const std::vector<double>& Q, std::vector<double>& Q2);
  void step_rk4(double coef1, double coef2, const std::vector<double>& Q, std::vector<double>& Q2);
  void step2_3(double c1, double c2, double c3, double c4inv, const std::vector<double>& Qin,
               std::vector<double>& Qout);
  void step4(std::vector<double>& Q);
  void set_bc(std::vector<double>& Q) const;  // set_bc_cyclic{2,4,6}, or the boundary fill below
  void set_bc_walls(std::vector<double>& Q) const;  // STZ/set.f90:77-138 (slip walls, zero gradient)
  void rk_step();                             // one full RK step
  void rhs(std::vector<double>& R);           // calc_EFG + calc_R, interior only
  void diagnostics(ox_diag& d) const;
};

}  // namespace oxo
