// ox_physics.hpp -- device-side face physics for the gfx950 kernels.
//
// Convective split-form / upwind face fluxes, MUSCL reconstruction and the
// central viscous stress at one face, written for stencils held in registers
// (template-sized arrays, fully unrolled).  Each function cites the reference
// routine whose arithmetic it reproduces; expression trees follow the
// reference's left-to-right evaluation so that, built without FMA contraction,
// the FP64 instantiation is bit-identical to a sequential evaluation of the
// reference formulas.  `R` is the arithmetic type (double for the plain form,
// float for the FP32-bulk form).
#pragma once

#include <hip/hip_runtime.h>

#include "ox_types.hpp"

// Device functions are also host-callable so that the operation-counting
// instantiation (ouxsbli_amd/flops/) can run them on the CPU.
#ifndef OX_HD
#define OX_HD __host__ __device__ __forceinline__
#endif

namespace ox {

// Fortran MIN/MAX as a select: the first operand wins ties, so the sign of a
// zero result is well defined.
template <class R> OX_HD R fmin2(R a, R b) { return (b < a) ? b : a; }
template <class R> OX_HD R fmax2(R a, R b) { return (b > a) ? b : a; }
// SIGN(1, x) with IEEE signed zero.
template <class R> OX_HD R fsign1(R x) { return __builtin_copysign(R(1), x); }
template <class R> OX_HD R fabs_(R x) { return __builtin_fabs(x); }
OX_HD double fsqrt(double x) { return __builtin_sqrt(x); }
OX_HD float fsqrt(float x) { return __builtin_sqrtf(x); }

// Per-case constants in the arithmetic type R (src/mod_constant.f90.fypp:18-35).
template <class R> struct PhysC {
  R gamma, gamma_1, over_gamma_1, R_over_gamma_1, Rgas, Cp_over_Pr, mu0c;
  R one_third, two_third, one_sixth, one_twelfth, one_sixty;
  R one_24, one_48, one_60, one_120, one_240, seven_twelfth;
  R gamma_over_gamma_1, minmod_b, threshold;
  R Cp, Prt, tan10deg2, les_delta15;
};

template <class R> __host__ PhysC<R> make_physc(const HostConsts& h) {
  PhysC<R> c;
  c.gamma = (R)h.gamma; c.gamma_1 = (R)h.gamma_1; c.over_gamma_1 = (R)h.over_gamma_1; c.R_over_gamma_1 = (R)h.R_over_gamma_1; c.Rgas = (R)h.Rgas;
  c.Cp_over_Pr = (R)h.Cp_over_Pr; c.mu0c = (R)h.mu0c;
  c.one_third = (R)h.one_third; c.two_third = (R)h.two_third; c.one_sixth = (R)h.one_sixth;
  c.one_twelfth = (R)h.one_twelfth; c.one_sixty = (R)h.one_sixty;
  c.one_24 = (R)h.one_24; c.one_48 = (R)h.one_48; c.one_60 = (R)h.one_60; c.one_120 = (R)h.one_120;
  c.one_240 = (R)h.one_240; c.seven_twelfth = (R)h.seven_twelfth;
  c.gamma_over_gamma_1 = (R)h.gamma_over_gamma_1; c.minmod_b = (R)h.minmod_b; c.threshold = (R)h.threshold;
  c.Cp = (R)h.Cp; c.Prt = (R)h.Prt; c.tan10deg2 = (R)h.tan10deg2; c.les_delta15 = (R)h.les_delta15;
  return c;
}

// ---------------------------------------------------------------------------
// KEEP split form.  3D_solver/src/calc_keep_3d.f90:1-111.  DIR selects the
// face normal: uu is the normal velocity and the pressure term lands in
// component 1+DIR (the literal `pres * Normal(m)` products are kept so that
// signed zeros match the reference expression).
// ---------------------------------------------------------------------------
template <class R, int DIR>
OX_HD void keep2(const PhysC<R>& c, const R* rho, const R* u, const R* v, const R* w,
                                      const R* p, const R* T, R* F) {
  const R* uu = DIR == 0 ? u : (DIR == 1 ? v : w);
  const R N1 = DIR == 0 ? R(1) : R(0), N2 = DIR == 1 ? R(1) : R(0), N3 = DIR == 2 ? R(1) : R(0);
  F[0] = R(0.25) * (rho[0] + rho[1]) * (uu[0] + uu[1]);
  F[1] = R(0.5) * (F[0] * (u[0] + u[1]) + (p[0] + p[1]) * N1);
  F[2] = R(0.5) * (F[0] * (v[0] + v[1]) + (p[0] + p[1]) * N2);
  F[3] = R(0.5) * (F[0] * (w[0] + w[1]) + (p[0] + p[1]) * N3);
  F[4] = F[0] * R(0.5) * (T[0] + T[1]) * c.R_over_gamma_1;
  F[4] = F[4] + R(0.5) * (uu[0] * p[1] + uu[1] * p[0]);
  F[4] = F[4] + R(0.5) * F[0] * (u[0] * u[1] + v[0] * v[1] + w[0] * w[1]);
}

template <class R, int DIR>
OX_HD void keep4(const PhysC<R>& c, const R* rho, const R* u, const R* v, const R* w,
                                      const R* p, const R* T, R* F) {
  const R* uu = DIR == 0 ? u : (DIR == 1 ? v : w);
  const R N1 = DIR == 0 ? R(1) : R(0), N2 = DIR == 1 ? R(1) : R(0), N3 = DIR == 2 ? R(1) : R(0);
  R RV1 = (rho[1] + rho[2]) * (uu[1] + uu[2]);
  R RV2 = (rho[1] + rho[3]) * (uu[1] + uu[3]);
  R RV3 = (rho[0] + rho[2]) * (uu[0] + uu[2]);
  F[0] = c.one_third * RV1 - (RV2 + RV3) * c.one_24;
  RV1 = c.one_sixth * RV1;
  RV2 = c.one_48 * RV2;
  RV3 = c.one_48 * RV3;
  const R A = RV1 - RV2;  // RV1_RV2
  const R B = RV1 - RV3;  // RV1_RV3
  const R pres = -(c.one_twelfth * p[0]) + c.seven_twelfth * (p[1] + p[2]) - c.one_twelfth * p[3];
  F[1] = -(RV3 * u[0]) + A * u[1] + B * u[2] - RV2 * u[3] + pres * N1;
  F[2] = -(RV3 * v[0]) + A * v[1] + B * v[2] - RV2 * v[3] + pres * N2;
  F[3] = -(RV3 * w[0]) + A * w[1] + B * w[2] - RV2 * w[3] + pres * N3;
  R ene = (-(RV3 * T[0]) + A * T[1] + B * T[2] - RV2 * T[3]) * c.R_over_gamma_1;
  ene = ene + ((RV1 * (u[1] * u[2] + v[1] * v[2] + w[1] * w[2])) -
               (RV2 * (u[1] * u[3] + v[1] * v[3] + w[1] * w[3]) + RV3 * (u[0] * u[2] + v[0] * v[2] + w[0] * w[2])));
  F[4] = ene + (c.two_third * (uu[1] * p[2] + uu[2] * p[1]) -
                c.one_twelfth * (uu[1] * p[3] + uu[3] * p[1] + uu[0] * p[2] + uu[2] * p[0]));
}

template <class R, int DIR>
OX_HD void keep6(const PhysC<R>& c, const R* rho, const R* u, const R* v, const R* w,
                                      const R* p, const R* T, R* F) {
  const R* uu = DIR == 0 ? u : (DIR == 1 ? v : w);
  const R N1 = DIR == 0 ? R(1) : R(0), N2 = DIR == 1 ? R(1) : R(0), N3 = DIR == 2 ? R(1) : R(0);
  R RV1 = (rho[2] + rho[3]) * (uu[2] + uu[3]);
  R RV2 = (rho[2] + rho[4]) * (uu[2] + uu[4]);
  R RV3 = (rho[1] + rho[3]) * (uu[1] + uu[3]);
  R RV4 = (rho[2] + rho[5]) * (uu[2] + uu[5]);
  R RV5 = (rho[1] + rho[4]) * (uu[1] + uu[4]);
  R RV6 = (rho[0] + rho[3]) * (uu[0] + uu[3]);
  F[0] = R(0.375) * RV1 - R(0.075) * (RV2 + RV3) + (RV4 + RV5 + RV6) * c.one_120;
  RV1 = R(0.1875) * RV1;
  RV2 = R(0.0375) * RV2;
  RV3 = R(0.0375) * RV3;
  RV4 = RV4 * c.one_240;
  RV5 = RV5 * c.one_240;
  RV6 = RV6 * c.one_240;
  const R a35 = -RV3 + RV5;
  const R a124 = RV1 - RV2 + RV4;
  const R a136 = RV1 - RV3 + RV6;
  const R a25 = -RV2 + RV5;
  const R pres = (p[0] - R(8) * p[1] + R(37) * (p[2] + p[3]) - R(8) * p[4] + p[5]) * c.one_60;
  R ruu = RV6 * u[0] + a35 * u[1] + a124 * u[2] + a136 * u[3] + a25 * u[4] + RV4 * u[5];
  F[1] = ruu + pres * N1;
  R ruv = RV6 * v[0] + a35 * v[1] + a124 * v[2] + a136 * v[3] + a25 * v[4] + RV4 * v[5];
  F[2] = ruv + pres * N2;
  R ruw = RV6 * w[0] + a35 * w[1] + a124 * w[2] + a136 * w[3] + a25 * w[4] + RV4 * w[5];
  F[3] = ruw + pres * N3;
  R ene = (RV6 * T[0] + a35 * T[1] + a124 * T[2] + a136 * T[3] + a25 * T[4] + RV4 * T[5]) * c.R_over_gamma_1;
  ene = ene + RV1 * (u[2] * u[3] + v[2] * v[3] + w[2] * w[3]) -
        (RV2 * (u[2] * u[4] + v[2] * v[4] + w[2] * w[4]) + RV3 * (u[1] * u[3] + v[1] * v[3] + w[1] * w[3])) +
        (RV4 * (u[2] * u[5] + v[2] * v[5] + w[2] * w[5]) + RV5 * (u[1] * u[4] + v[1] * v[4] + w[1] * w[4]) +
         RV6 * (u[0] * u[3] + v[0] * v[3] + w[0] * w[3]));
  F[4] = ene + (R(0.75) * (uu[2] * p[3] + uu[3] * p[2]) -
                R(0.15) * (uu[2] * p[4] + uu[4] * p[2] + uu[1] * p[3] + uu[3] * p[1]) +
                (uu[2] * p[5] + uu[5] * p[2] + uu[1] * p[4] + uu[4] * p[1] + uu[0] * p[3] + uu[3] * p[0]) * c.one_60);
}

template <class R, int DIR, int N>
OX_HD void keep_face(const PhysC<R>& c, const R* rho, const R* u, const R* v, const R* w, const R* p, const R* T, R* F) {
  if constexpr (N == 2) keep2<R, DIR>(c, rho, u, v, w, p, T, F);
  else if constexpr (N == 4) keep4<R, DIR>(c, rho, u, v, w, p, T, F);
  else keep6<R, DIR>(c, rho, u, v, w, p, T, F);
}

// ---------------------------------------------------------------------------
// MUSCL.  src/calc_muscl.f90:19-169.  tvd: OX_TVD_NONE / TVD / HYBRID.
// ---------------------------------------------------------------------------
// MIN/MAX of the limiters: the hardware min/max (v_min_f64 / v_min_f32, one
// instruction where the compare-and-select form takes three).  For every pair
// of non-NaN operands it returns the value the select form (b < a ? b : a)
// returns; the two differ only in the sign of a zero when the operands are
// +0 and -0.  A limiter's result enters the interface state as an addend
// (a + 0 and a - 0 are a either way) or, through fsign1 in minmod, as the
// factor of a zero, so the sign of a zero never reaches a nonzero value:
// the FP64 results equal the oracle's select-form results value for value.
// Other value types (the operation counter) keep the select form.
template <class R> OX_HD R lmin(R a, R b) { return fmin2(a, b); }
template <class R> OX_HD R lmax(R a, R b) { return fmax2(a, b); }
template <> OX_HD float lmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <> OX_HD float lmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }
template <> OX_HD double lmin<double>(double a, double b) { return __builtin_fmin(a, b); }
template <> OX_HD double lmax<double>(double a, double b) { return __builtin_fmax(a, b); }

template <class R> OX_HD R minmod2(R x, R y) {
  const R sgn = fsign1(x);
  return sgn * lmax(lmin(fabs_(x), sgn * y), R(0));
}
template <class R> OX_HD R minmod3(R x, R y, R z) {
  const R sgn = fsign1(x);
  return sgn * lmax(lmin(lmin(fabs_(x), sgn * y), sgn * z), R(0));
}
template <class R> OX_HD R d33(R d1, R d2, R d3) {
  return minmod3(d1, R(2) * d2, R(2) * d3) - R(2) * minmod3(d2, R(2) * d1, R(2) * d3) +
         minmod3(d3, R(2) * d1, R(2) * d2);
}

// True where the MUSCL interface states take the limited form: TVD kind
// "tvd" everywhere, kind "hybrid" where the face sensor reaches the threshold
// (the MUSCL*Threshold selectors of src/calc_muscl.f90:142-169).  The switch
// is taken once per face, so a wavefront whose faces are all smooth skips the
// limiter arithmetic entirely.
template <class R> OX_HD bool muscl_limited(const PhysC<R>& c, int tvd, R sensor) {
  return !(tvd == 0 || (tvd == 2 && sensor < c.threshold));
}

// delta4 -> MUSCL3rd{nonTVD,Minmod} (:44-91)
template <class R, bool LIM>
OX_HD void delta4_form(const PhysC<R>& c, const R* a, R& al, R& ar) {
  const R d1 = -a[0] + a[1];
  const R d2 = -a[1] + a[2];
  const R d3 = -a[2] + a[3];
  if constexpr (!LIM) {
    al = a[1] + R(0.5) * (d1 + R(2) * d2) * c.one_third;
    ar = a[2] - R(0.5) * (d3 + R(2) * d2) * c.one_third;
  } else {
    const R b = c.minmod_b;
    const R dt1 = minmod2(d1, b * d2);
    const R dt2 = minmod2(d2, b * d1);
    al = a[1] + R(0.5) * (dt1 + R(2) * dt2) * c.one_third;
    const R dt3 = minmod2(d3, b * d2);
    const R dt4 = minmod2(d2, b * d3);
    ar = a[2] - R(0.5) * (dt3 + R(2) * dt4) * c.one_third;
  }
}

// delta6 -> MUSCL4th{nonTVD,TVD} (:94-139)
template <class R, bool LIM>
OX_HD void delta6_form(const PhysC<R>& c, const R* a, R& al, R& ar) {
  const R d1 = -a[0] + a[1];
  const R d2 = -a[1] + a[2];
  const R d3 = -a[2] + a[3];
  const R d4 = -a[3] + a[4];
  const R d5 = -a[4] + a[5];
  if constexpr (!LIM) {
    al = a[2] + (R(-0.4) * d1 + R(2.2) * d2 + R(4.8) * d3 - R(0.6) * d4) * c.one_twelfth;
    ar = a[3] - (R(-0.6) * d2 + R(4.8) * d3 + R(2.2) * d4 - R(0.4) * d5) * c.one_twelfth;
  } else {
    const R delta2 = d3 - d33(d2, d3, d4) * c.one_sixth;
    {
      const R delta1 = d2 - d33(d1, d2, d3) * c.one_sixth;
      const R dl = minmod2(delta1, R(4) * delta2);
      const R dr = minmod2(delta2, R(4) * delta1);
      al = a[2] + (dl + R(2) * dr) * c.one_sixth;
    }
    {
      const R delta3 = d4 - d33(d3, d4, d5) * c.one_sixth;
      const R dl = minmod2(delta2, R(4) * delta3);
      const R dr = minmod2(delta3, R(4) * delta2);
      ar = a[3] - (dr + R(2) * dl) * c.one_sixth;
    }
  }
}

// delta4 / delta6 with the selector (MUSCL3rdThreshold / MUSCL4thThreshold,
// :142-169) for one variable.
template <class R>
OX_HD void delta4(const PhysC<R>& c, int tvd, R sensor, const R* a, R& al, R& ar) {
  if (muscl_limited(c, tvd, sensor)) delta4_form<R, true>(c, a, al, ar);
  else delta4_form<R, false>(c, a, al, ar);
}
template <class R>
OX_HD void delta6(const PhysC<R>& c, int tvd, R sensor, const R* a, R& al, R& ar) {
  if (muscl_limited(c, tvd, sensor)) delta6_form<R, true>(c, a, al, ar);
  else delta6_form<R, false>(c, a, al, ar);
}

// The five primitives of one face in one form.
template <class R, int N, bool LIM>
OX_HD void muscl5(const PhysC<R>& c, const R* rho, const R* u, const R* v, const R* w, const R* p, R* L, R* Rr) {
  if constexpr (N == 4) {
    delta4_form<R, LIM>(c, rho, L[0], Rr[0]);
    delta4_form<R, LIM>(c, u, L[1], Rr[1]);
    delta4_form<R, LIM>(c, v, L[2], Rr[2]);
    delta4_form<R, LIM>(c, w, L[3], Rr[3]);
    delta4_form<R, LIM>(c, p, L[4], Rr[4]);
  } else {
    delta6_form<R, LIM>(c, rho, L[0], Rr[0]);
    delta6_form<R, LIM>(c, u, L[1], Rr[1]);
    delta6_form<R, LIM>(c, v, L[2], Rr[2]);
    delta6_form<R, LIM>(c, w, L[3], Rr[3]);
    delta6_form<R, LIM>(c, p, L[4], Rr[4]);
  }
}

// 3D_solver/src/calc_hybrid.f90:75-81
template <class R> OX_HD R wiggle(const R* phi) {
  const R phi1 = (-phi[0] + phi[1]) * (-phi[1] + phi[2]);
  const R phi2 = (-phi[2] + phi[3]) * (-phi[1] + phi[2]);
  return R(0.5) * (R(1) - fsign1(fmin2(phi1, phi2)));
}

// ---------------------------------------------------------------------------
// SLAU1 / HRSLAU2.  3D_solver/src/calc_slau_3d.f90:1-104.
// ---------------------------------------------------------------------------
template <class R, int DIR>
OX_HD void slau(const PhysC<R>& c, int variant, R rho1, R rho2, R u1, R u2, R v1, R v2, R w1,
                                     R w2, R p1, R p2, R HR, R* F) {
  const R un1 = DIR == 0 ? u1 : (DIR == 1 ? v1 : w1);
  const R un2 = DIR == 0 ? u2 : (DIR == 1 ? v2 : w2);
  const R N1 = DIR == 0 ? R(1) : R(0), N2 = DIR == 1 ? R(1) : R(0), N3 = DIR == 2 ? R(1) : R(0);
  const R over_rho1 = R(1) / rho1;
  const R over_rho2 = R(1) / rho2;
  // SLAU_common
  const R cl = fsqrt(c.gamma * p1 * over_rho1);
  const R cr = fsqrt(c.gamma * p2 * over_rho2);
  const R cc = R(0.5) * (cl + cr);
  const R over_c = R(1) / cc;
  const R Mp = un1 * over_c;
  const R Mm = un2 * over_c;
  const R g = -(fmax2(fmin2(Mp, R(0)), R(-1)) * fmin2(fmax2(Mm, R(0)), R(1)));
  const R one_g_Vt = (R(1) - g) * (rho1 * fabs_(un1) + rho2 * fabs_(un2)) / (rho1 + rho2);
  const R Vtp = one_g_Vt + g * fabs_(un1);
  const R Vtm = one_g_Vt + g * fabs_(un2);
  const R bp = (fabs_(Mp) < R(1)) ? R(0.25) * (R(2) - Mp) * ((Mp + R(1)) * (Mp + R(1))) : R(0.5) * (R(1) + fsign1(Mp));
  const R bm = (fabs_(Mm) < R(1)) ? R(0.25) * (R(2) + Mm) * ((Mm - R(1)) * (Mm - R(1))) : R(0.5) * (R(1) + fsign1(-Mm));
  const R dp = -p1 + p2;
  const R k1 = R(0.5) * (u1 * u1 + v1 * v1 + w1 * w1);
  const R k2 = R(0.5) * (u2 * u2 + v2 * v2 + w2 * w2);
  R mass1, mass2, pres;
  if (variant == 0) {  // SLAU1
    const R M = fmin2(R(1), fsqrt(k1 + k2) * over_c);
    const R x = (R(1) - M) * (R(1) - M);
    const R mass = R(0.25) * (rho1 * (un1 + Vtp) + rho2 * (un2 - Vtm) - x * dp * over_c);
    mass1 = mass + fabs_(mass);
    mass2 = mass - fabs_(mass);
    pres = R(0.5) * (p1 + p2 + (bp - bm) * (-dp) + (R(1) - x) * (bp + bm - R(1)) * (p1 + p2));
  } else {  // HRSLAU2
    const R Vec2 = fsqrt(k1 + k2);
    const R M = fmin2(R(1), Vec2 * over_c);
    const R x = (R(1) - M) * (R(1) - M);
    const R mass = R(0.25) * (rho1 * (un1 + Vtp) + rho2 * (un2 - Vtm) - x * dp * over_c);
    mass1 = mass + fabs_(mass);
    mass2 = mass - fabs_(mass);
    pres = R(0.5) * (p1 + p2 + (bp - bm) * (-dp) + HR * Vec2 * (bp + bm - R(1)) * R(0.5) * (rho1 + rho2) * cc);
  }
  F[0] = mass1 + mass2;
  F[1] = mass1 * u1 + mass2 * u2 + pres * N1;
  F[2] = mass1 * v1 + mass2 * v2 + pres * N2;
  F[3] = mass1 * w1 + mass2 * w2 + pres * N3;
  const R phi1 = (p1 * c.gamma_over_gamma_1 + rho1 * k1) * over_rho1;
  const R phi2 = (p2 * c.gamma_over_gamma_1 + rho2 * k2) * over_rho2;
  F[4] = mass1 * phi1 + mass2 * phi2;
}

// ---------------------------------------------------------------------------
// Roe.  3D_solver/src/calc_roe_3d.f90:1-88, in the face's local frame
// (u = normal velocity); o[5] = (mass, normal, t1, t2, energy).
// ---------------------------------------------------------------------------
template <class R>
OX_HD void roe_local(const PhysC<R>& c, R rho1, R rho2, R u1, R u2, R v1, R v2, R w1, R w2, R p1, R p2, R* F) {
  const R k1 = R(0.5) * (u1 * u1 + v1 * v1 + w1 * w1);
  const R k2 = R(0.5) * (u2 * u2 + v2 * v2 + w2 * w2);
  const R e1 = p1 * c.over_gamma_1 + rho1 * k1;
  const R e2 = p2 * c.over_gamma_1 + rho2 * k2;
  const R rhou1 = rho1 * u1;
  const R rhou2 = rho2 * u2;
  F[0] = R(0.5) * (rhou1 + rhou2);
  F[1] = R(0.5) * (rhou1 * u1 + rhou2 * u2 + (p1 + p2));
  F[2] = R(0.5) * (rhou1 * v1 + rhou2 * v2);
  F[3] = R(0.5) * (rhou1 * w1 + rhou2 * w2);
  F[4] = R(0.5) * ((e1 + p1) * u1 + (e2 + p2) * u2);
  const R dQ0 = -rho1 + rho2;
  const R dQ1 = -(rho1 * u1) + rho2 * u2;
  const R dQ2 = -(rho1 * v1) + rho2 * v2;
  const R dQ3 = -(rho1 * w1) + rho2 * w2;
  const R dQ4 = -e1 + e2;
  const R H1 = c.gamma_over_gamma_1 * p1 / rho1 + k1;
  const R H2 = c.gamma_over_gamma_1 * p2 / rho2 + k2;
  const R rhol = fsqrt(rho1);
  const R rhor = fsqrt(rho2);
  const R rr = R(1) / (rhol + rhor);
  const R u = (rhol * u1 + rhor * u2) * rr;
  const R v = (rhol * v1 + rhor * v2) * rr;
  const R w = (rhol * w1 + rhor * w2) * rr;
  const R H = (rhol * H1 + rhor * H2) * rr;
  const R cr = fsqrt(c.gamma_1 * (H - R(0.5) * (u * u + v * v + w * w)));
  const R q2 = R(0.5) * (u * u + v * v + w * w);
  const R oc = R(1) / cr;
  const R b2 = c.gamma_1 * oc * oc;
  const R b1 = q2 * b2;
  const R R1 = R(0.5) * ((b1 + u * oc) * dQ0 - (oc + b2 * u) * dQ1 - b2 * v * dQ2 - b2 * w * dQ3 + b2 * dQ4);
  const R R2 = oc * (w * dQ0 - dQ3);
  const R R3 = oc * (-(v * dQ0) + dQ2);
  const R R4 = (R(1) - b1) * dQ0 + b2 * (u * dQ1 + v * dQ2 + w * dQ3 - dQ4);
  const R R5 = R(0.5) * ((b1 - u * oc) * dQ0 + (oc - b2 * u) * dQ1 - b2 * (v * dQ2 + w * dQ3 - dQ4));
  const R a1 = fabs_(u - cr);
  const R a2 = fabs_(u);
  const R a3 = fabs_(u + cr);
  const R a2c = a2 * cr;
  const R cu = cr * u;
  F[0] = F[0] - R(0.5) * (a1 * R1 + a2 * R4 + a3 * R5);
  F[1] = F[1] - R(0.5) * (a1 * (u - cr) * R1 + a2 * u * R4 + a3 * (u + cr) * R5);
  F[2] = F[2] - R(0.5) * (a1 * v * R1 + a2c * R3 + a2 * v * R4 + a3 * v * R5);
  F[3] = F[3] - R(0.5) * (a1 * w * R1 - a2c * R2 + a2 * w * R4 + a3 * w * R5);
  F[4] = F[4] - R(0.5) * (a1 * (H - cu) * R1 + a2 * (-(w * cr * R2) + v * cr * R3 + q2 * R4) + a3 * (H + cu) * R5);
}

// roe_local for a face with one tangential velocity (the 2-D solver, where
// the third component is zero): the same operations with the terms that
// multiply or add that exact zero removed.  Removing them changes no value
// (x + 0 and x - 0 are x; only the sign of a zero result can differ), and
// the tangential component may sit in either tangential slot of roe_local:
// with it in the w slot, R2 there is -R3 here exactly, so both face
// directions of the 2-D solver reduce to this one routine.
// o[4] = (mass, normal momentum, tangential momentum, energy).
template <class R>
OX_HD void roe_local2(const PhysC<R>& c, R rho1, R rho2, R un1, R un2, R ut1, R ut2, R p1, R p2, R* o) {
  const R k1 = R(0.5) * (un1 * un1 + ut1 * ut1);
  const R k2 = R(0.5) * (un2 * un2 + ut2 * ut2);
  const R e1 = p1 * c.over_gamma_1 + rho1 * k1;
  const R e2 = p2 * c.over_gamma_1 + rho2 * k2;
  const R rhou1 = rho1 * un1;
  const R rhou2 = rho2 * un2;
  o[0] = R(0.5) * (rhou1 + rhou2);
  o[1] = R(0.5) * (rhou1 * un1 + rhou2 * un2 + (p1 + p2));
  o[2] = R(0.5) * (rhou1 * ut1 + rhou2 * ut2);
  o[3] = R(0.5) * ((e1 + p1) * un1 + (e2 + p2) * un2);
  const R dQ0 = -rho1 + rho2;
  const R dQ1 = -(rho1 * un1) + rho2 * un2;
  const R dQt = -(rho1 * ut1) + rho2 * ut2;
  const R dQ4 = -e1 + e2;
  const R H1 = c.gamma_over_gamma_1 * p1 / rho1 + k1;
  const R H2 = c.gamma_over_gamma_1 * p2 / rho2 + k2;
  const R rhol = fsqrt(rho1);
  const R rhor = fsqrt(rho2);
  const R rr = R(1) / (rhol + rhor);
  const R u = (rhol * un1 + rhor * un2) * rr;
  const R t = (rhol * ut1 + rhor * ut2) * rr;
  const R H = (rhol * H1 + rhor * H2) * rr;
  const R cr = fsqrt(c.gamma_1 * (H - R(0.5) * (u * u + t * t)));
  const R q2 = R(0.5) * (u * u + t * t);
  const R oc = R(1) / cr;
  const R b2 = c.gamma_1 * oc * oc;
  const R b1 = q2 * b2;
  const R R1 = R(0.5) * ((b1 + u * oc) * dQ0 - (oc + b2 * u) * dQ1 - b2 * t * dQt + b2 * dQ4);
  const R R3 = oc * (-(t * dQ0) + dQt);
  const R R4 = (R(1) - b1) * dQ0 + b2 * (u * dQ1 + t * dQt - dQ4);
  const R R5 = R(0.5) * ((b1 - u * oc) * dQ0 + (oc - b2 * u) * dQ1 - b2 * (t * dQt - dQ4));
  const R a1 = fabs_(u - cr);
  const R a2 = fabs_(u);
  const R a3 = fabs_(u + cr);
  const R a2c = a2 * cr;
  const R cu = cr * u;
  o[0] = o[0] - R(0.5) * (a1 * R1 + a2 * R4 + a3 * R5);
  o[1] = o[1] - R(0.5) * (a1 * (u - cr) * R1 + a2 * u * R4 + a3 * (u + cr) * R5);
  o[2] = o[2] - R(0.5) * (a1 * t * R1 + a2c * R3 + a2 * t * R4 + a3 * t * R5);
  o[3] = o[3] - R(0.5) * (a1 * (H - cu) * R1 + a2 * (t * cr * R3 + q2 * R4) + a3 * (H + cu) * R5);
}

// MUSCL left/right states of the five primitives (delta_r2/4/6 of
// calc_roe_kernel_internal.f90.fypp:28-55 = the delta4/delta6 of SLAU).
template <class R, int N>
OX_HD void reconstruct(const PhysC<R>& c, int tvd, R fsens, const R* rho, const R* u, const R* v, const R* w,
                       const R* p, R* L, R* Rr) {
  if constexpr (N == 2) {
    L[0] = rho[0]; Rr[0] = rho[1]; L[1] = u[0]; Rr[1] = u[1]; L[2] = v[0]; Rr[2] = v[1];
    L[3] = w[0]; Rr[3] = w[1]; L[4] = p[0]; Rr[4] = p[1];
  } else {
    if (muscl_limited(c, tvd, fsens)) muscl5<R, N, true>(c, rho, u, v, w, p, L, Rr);
    else muscl5<R, N, false>(c, rho, u, v, w, p, L, Rr);
  }
}

// calc_roe_{x,y,z}_in: MUSCL, then Roe in the local frame (normal, t1, t2)
// with (t1, t2) cyclic after the normal; outputs mapped back.
template <class R, int DIR, int N>
OX_HD void roe_face(const PhysC<R>& c, int tvd, const R* rho, const R* u, const R* v, const R* w, const R* p, R fsens,
                    R* F) {
  R L[5], Rr[5], o[5];
  reconstruct<R, N>(c, tvd, fsens, rho, u, v, w, p, L, Rr);
  constexpr int a = 1 + DIR, b = 1 + (DIR + 1) % 3, d = 1 + (DIR + 2) % 3;
  roe_local<R>(c, L[0], Rr[0], L[a], Rr[a], L[b], Rr[b], L[d], Rr[d], L[4], Rr[4], o);
  F[0] = o[0];
  F[a] = o[1];
  F[b] = o[2];
  F[d] = o[3];
  F[4] = o[4];
}

// SLAU face: interp2/4/6 (calc_slau_kernel_internal.f90.fypp:31-81) then SLAU.
// N = stencil length (2, 4, 6); fsens = face Ducros sensor.
template <class R, int DIR, int N>
OX_HD void slau_face(const PhysC<R>& c, int tvd, int variant, const R* rho, const R* u,
                                          const R* v, const R* w, const R* p, R fsens, R* F) {
  R L[5], Rr[5];
  reconstruct<R, N>(c, tvd, fsens, rho, u, v, w, p, L, Rr);
  R hr;
  if constexpr (N == 2) hr = R(1);
  else if constexpr (N == 4) hr = wiggle(p);
  else hr = wiggle(p + 1);
  slau<R, DIR>(c, variant, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], L[3], Rr[3], L[4], Rr[4], hr, F);
}

// ---------------------------------------------------------------------------
// Central viscous face.  3D_solver/src/calc_visc_cent.f90.fypp:2-91 with the
// face assembly of calc_visc_high.f90.fypp:114-140 (VO=4) and
// calc_visc_high_internal.f90.fypp:185-211 (VO=6) and their F/G counterparts.
// d1..d4 are the cross derivatives the reference stages in LDS:
//   DIR 0 (E): uy, vy, uz, wz     DIR 1 (F): ux, vx, vz, wz
//   DIR 2 (G): ux, wx, vy, wy
// Returns tau[3] (subtracted from components 2..4) and the energy term.
// ---------------------------------------------------------------------------
template <class R, int N> OX_HD R vinterp(const R* a) {
  if constexpr (N == 4) return R(0.0625) * (-a[0] + R(9) * (a[1] + a[2]) - a[3]);
  else return R(0.00390625) * (R(3) * a[0] - R(25) * a[1] + R(150) * (a[2] + a[3]) - R(25) * a[4] + R(3) * a[5]);
}
template <class R, int N> OX_HD R vdiff(const PhysC<R>& c, R d, const R* a) {
  if constexpr (N == 4) return (a[0] - R(27) * (a[1] - a[2]) - a[3]) * c.one_24 * d;
  else
    return (R(-0.0046875) * a[0] + R(0.065104167) * a[1] - R(1.171875) * (a[2] - a[3]) - R(0.065104167) * a[4] +
            R(0.0046875) * a[5]) *
           d;
}

// LES (calc_visc_cent.f90.fypp:44-58, :76-91; calc_visc_high_internal.f90.fypp:196-199):
// mt = rho*nu_t at the stencil, H = Cp*T + |u|^2/2 + qc2; the stresses use
// 0.5*(mu + mut), the work terms the molecular stress, plus the SGS heat flux.
template <class R, int DIR, int N, bool LES = false>
OX_HD void visc_face(const PhysC<R>& c, R d, const R* mu, const R* T, const R* u, const R* v,
                                          const R* w, const R* d1, const R* d2, const R* d3, const R* d4, R* tau,
                                          R& energy, const R* mt = nullptr, const R* H = nullptr) {
  const R mui = vinterp<R, N>(mu);
  const R kT = c.Cp_over_Pr * mui * vdiff<R, N>(c, d, T);
  // straight component: (vel_DIR, a, b); cross components: (vel_m, g_m)
  const R* vs = DIR == 0 ? u : (DIR == 1 ? v : w);
  const R* sa = DIR == 0 ? d2 : (DIR == 1 ? d4 : d1);  // vy | wz | ux
  const R* sb = DIR == 0 ? d4 : (DIR == 1 ? d1 : d3);  // wz | ux | vy
  const R ai = vinterp<R, N>(sa);
  const R bi = vinterp<R, N>(sb);
  const R dn = vdiff<R, N>(c, d, vs);
  R ts = c.two_third * mui * (R(2) * dn - ai - bi);
  const R uts = vinterp<R, N>(vs) * ts;
  // the two cross components in component order
  const R* va = DIR == 0 ? v : u;                      // first cross velocity
  const R* ga = DIR == 0 ? d1 : d2;                    // uy | vx | wx
  const R* vb = DIR == 2 ? v : w;                      // second cross velocity
  const R* gb = DIR == 0 ? d3 : (DIR == 1 ? d3 : d4);  // uz | vz | wy
  const R ga_i = vinterp<R, N>(ga), va_d = vdiff<R, N>(c, d, va);
  const R gb_i = vinterp<R, N>(gb), vb_d = vdiff<R, N>(c, d, vb);
  R ta = mui * (ga_i + va_d);
  const R uta = vinterp<R, N>(va) * ta;
  R tb = mui * (gb_i + vb_d);
  const R utb = vinterp<R, N>(vb) * tb;
  R Hsgs = R(0);
  if constexpr (LES) {
    const R mti = vinterp<R, N>(mt);
    const R m2 = R(0.5) * (mui + mti);
    ts = c.two_third * m2 * (R(2) * dn - ai - bi);
    ta = m2 * (ga_i + va_d);
    tb = m2 * (gb_i + vb_d);
    Hsgs = -(mti * vdiff<R, N>(c, d, H) / c.Prt);
  }
  if constexpr (DIR == 0) {
    tau[0] = ts; tau[1] = ta; tau[2] = tb;
    energy = uts + uta + utb + kT;  // utxx + vtxy + wtxz + kTx
  } else if constexpr (DIR == 1) {
    tau[0] = ta; tau[1] = ts; tau[2] = tb;
    energy = uta + uts + utb + kT;  // utyx + vtyy + wtyz + kTy
  } else {
    tau[0] = ta; tau[1] = tb; tau[2] = ts;
    energy = uta + utb + uts + kT;  // utzx + vtzy + wtzz + kTz
  }
  if constexpr (LES) energy = energy + Hsgs;
}

// ---------------------------------------------------------------------------
// ME4Base viscous face.  3D_solver/src/calc_visc_me4_base.f90:4-111 with the
// face assembly of calc_visc_high_internal.f90.fypp:146-183 (and the F/G
// counterparts): mu (and mut) interpolated to the three faces f-1, f, f+1 of
// the 6-cell window, 4th-order face derivatives there, and the three face
// values combined by flux4(a) = (-a1 + 26 a2 - a3)/24.  The LES stress uses
// (mu + mut) here, the work terms the molecular stress.
// ---------------------------------------------------------------------------
template <class R> OX_HD R flux4(const PhysC<R>& c, R a0, R a1, R a2) { return (-a0 + R(26) * a1 - a2) * c.one_24; }
template <class R> OX_HD R me4_interp(const R* a, int t) {
  return R(0.0625) * (R(9) * (a[t + 1] + a[t + 2]) - (a[t] + a[t + 3]));
}
template <class R> OX_HD R me4_interp_w(const R* a, int t) {
  return R(0.0625) * (-a[t] + R(9) * (a[t + 1] + a[t + 2]) - a[t + 3]);
}
template <class R, bool LES>
OX_HD void me4_straight(const PhysC<R>& c, const R* mu3, const R* mut3, const R* u, const R* sa, const R* sb, R d,
                        R& ts, R& uts) {
  R t1[3], t2[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    if constexpr (!LES) {
      t1[t] = c.two_third * mu3[t] *
              ((R(2.25) * (-u[t + 1] + u[t + 2]) - (-u[t] + u[t + 3]) * c.one_twelfth) * d - me4_interp_w(sa, t) -
               me4_interp_w(sb, t));
    } else {
      t1[t] = (R(2.25) * (-u[t + 1] + u[t + 2]) - (-u[t] + u[t + 3]) * c.one_twelfth) * d;
      t1[t] = t1[t] - me4_interp_w(sa, t);
      t1[t] = t1[t] - me4_interp_w(sb, t);
      t2[t] = c.two_third * mu3[t] * t1[t];
    }
  }
  if constexpr (!LES) {
    ts = flux4(c, t1[0], t1[1], t1[2]);
#pragma unroll
    for (int t = 0; t < 3; ++t) t2[t] = me4_interp_w(u, t) * t1[t];
    uts = flux4(c, t2[0], t2[1], t2[2]);
  } else {
#pragma unroll
    for (int t = 0; t < 3; ++t) t2[t] = me4_interp_w(u, t) * t2[t];
    uts = flux4(c, t2[0], t2[1], t2[2]);
#pragma unroll
    for (int t = 0; t < 3; ++t) t2[t] = c.two_third * (mu3[t] + mut3[t]) * t1[t];
    ts = flux4(c, t2[0], t2[1], t2[2]);
  }
}
template <class R, bool LES>
OX_HD void me4_cross(const PhysC<R>& c, const R* mu3, const R* mut3, const R* v, const R* g, R d, R& t12, R& vt12) {
  R t1[3], t2[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    if constexpr (!LES) {
      t1[t] = mu3[t] * ((R(1.125) * (-v[t + 1] + v[t + 2]) - (-v[t] + v[t + 3]) * c.one_24) * d + me4_interp_w(g, t));
    } else {
      t1[t] = (R(1.125) * (-v[t + 1] + v[t + 2]) - (-v[t] + v[t + 3]) * c.one_24) * d;
      t1[t] = t1[t] + me4_interp_w(g, t);
      t2[t] = mu3[t] * t1[t];
    }
  }
  if constexpr (!LES) {
    t12 = flux4(c, t1[0], t1[1], t1[2]);
#pragma unroll
    for (int t = 0; t < 3; ++t) t2[t] = me4_interp_w(v, t) * t1[t];
    vt12 = flux4(c, t2[0], t2[1], t2[2]);
  } else {
#pragma unroll
    for (int t = 0; t < 3; ++t) t2[t] = me4_interp_w(v, t) * t2[t];
    vt12 = flux4(c, t2[0], t2[1], t2[2]);
#pragma unroll
    for (int t = 0; t < 3; ++t) t2[t] = (mu3[t] + mut3[t]) * t1[t];
    t12 = flux4(c, t2[0], t2[1], t2[2]);
  }
}

template <class R, int DIR, bool LES = false>
OX_HD void visc_face_me4(const PhysC<R>& c, R d, const R* mu, const R* T, const R* u, const R* v, const R* w,
                         const R* d1, const R* d2, const R* d3, const R* d4, R* tau, R& energy,
                         const R* mt = nullptr, const R* H = nullptr) {
  R mu3[3], mut3[3], kT3[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    mu3[t] = me4_interp(mu, t);
    if constexpr (LES) mut3[t] = me4_interp(mt, t);
    kT3[t] = c.Cp_over_Pr * mu3[t] * (R(1.125) * (-T[t + 1] + T[t + 2]) - (-T[t] + T[t + 3]) * c.one_24) * d;
  }
  const R kT = flux4(c, kT3[0], kT3[1], kT3[2]);
  const R* vs = DIR == 0 ? u : (DIR == 1 ? v : w);
  const R* sa = DIR == 0 ? d2 : (DIR == 1 ? d4 : d1);  // vy | wz | ux
  const R* sb = DIR == 0 ? d4 : (DIR == 1 ? d1 : d3);  // wz | ux | vy
  const R* va = DIR == 0 ? v : u;
  const R* ga = DIR == 0 ? d1 : d2;                    // uy | vx | wx
  const R* vb = DIR == 2 ? v : w;
  const R* gb = DIR == 0 ? d3 : (DIR == 1 ? d3 : d4);  // uz | vz | wy
  R ts, uts, ta, uta, tb, utb;
  me4_straight<R, LES>(c, mu3, mut3, vs, sa, sb, d, ts, uts);
  me4_cross<R, LES>(c, mu3, mut3, va, ga, d, ta, uta);
  me4_cross<R, LES>(c, mu3, mut3, vb, gb, d, tb, utb);
  if constexpr (DIR == 0) {
    tau[0] = ts; tau[1] = ta; tau[2] = tb;
    energy = uts + uta + utb + kT;
  } else if constexpr (DIR == 1) {
    tau[0] = ta; tau[1] = ts; tau[2] = tb;
    energy = uta + uts + utb + kT;
  } else {
    tau[0] = ta; tau[1] = tb; tau[2] = ts;
    energy = uta + utb + uts + kT;
  }
  if constexpr (LES) {
    // SGS heat flux over the four central cells of the window
    const R Hsgs = -flux4(c, mut3[0], mut3[1], mut3[2]) * (R(1.125) * (-H[2] + H[3]) - (-H[1] + H[4]) * c.one_24) *
                   d / c.Prt;
    energy = energy + Hsgs;
  }
}

// Cross derivative stencils of the LDS loaders
// (3D_solver/src/load_smem_visc_cent.f90.fypp:52-66): f[m] = value at offset
// m - VO/2 ... (centre excluded is never read).
template <class R, int VO> OX_HD R cross_d(const PhysC<R>& c, R inv, const R* f) {
  // f has 2*(VO/2)+1 entries centred on index VO/2 (VO=4: 5 entries, VO=6: 7)
  if constexpr (VO == 4) {
    return (c.two_third * (-f[1] + f[3]) - c.one_twelfth * (-f[0] + f[4])) * inv;
  } else {
    return (c.one_sixty * (-f[0] + f[6]) + R(0.15) * (f[1] - f[5]) + R(0.75) * (-f[2] + f[4])) * inv;
  }
}

}  // namespace ox
