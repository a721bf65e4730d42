// ox_types.hpp -- host/device shared plain types of the gfx950 library.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/ouxsbli.h"

namespace ox {

// Derived constants of one case, computed once on the host in FP64 exactly
// as src/mod_constant.f90.fypp:18-35 writes them (and the module constants
// of 3D_solver/src/calc_keep_kernel_internal.f90.fypp:7-12).
struct HostConsts {
  double gamma, gamma_1, over_gamma_1, R_over_gamma_1, Rgas, Cp, Cp_over_Pr, Prt, mu0c;
  double one_third, two_third, one_sixth, one_twelfth, one_sixty;
  double one_24, one_48, one_60, one_120, one_240, seven_twelfth;
  double gamma_over_gamma_1, minmod_b, threshold;
  double tan10deg2;   // (tan 10 deg)^2, the SMS angle selector (mod_constant.f90.fypp:34)
  double les_delta15; // SMS filter width to the power 1+alpha (calc_les.f90:83-85), set per grid

  void init(const ox_config& c) {
    gamma = c.gamma;
    gamma_1 = c.gamma - 1.0;
    over_gamma_1 = 1.0 / (c.gamma - 1.0);
    R_over_gamma_1 = c.R / (c.gamma - 1.0);
    Rgas = c.R;
    Cp = c.gamma * c.R / (c.gamma - 1.0);
    Cp_over_Pr = Cp / c.Pr;
    Prt = c.Prt;
    mu0c = 1.716e-5 * 384.2 * std::pow(273.2, -1.5);
    one_third = 1.0 / 3.0;
    two_third = 2.0 / 3.0;
    one_sixth = 1.0 / 6.0;
    one_twelfth = 1.0 / 12.0;
    one_sixty = 1.0 / 60.0;
    one_24 = 1.0 / 24.0;
    one_48 = 1.0 / 48.0;
    one_60 = 1.0 / 60.0;
    one_120 = 1.0 / 120.0;
    one_240 = 1.0 / 240.0;
    seven_twelfth = 7.0 / 12.0;
    gamma_over_gamma_1 = c.gamma * (1.0 / (c.gamma - 1.0));
    minmod_b = (3.0 - one_third) / (1.0 - one_third);
    threshold = c.threshold;
    const double t = std::tan(10.0 * std::acos(-1.0) / 180.0);
    tan10deg2 = t * t;
    // SMS filter width from the inverse spacings of a uniform periodic grid
    const double idx = (double)(c.nx - c.order) / c.Lx, idy = (double)(c.ny - c.order) / c.Ly,
                 idz = (double)(c.nz - c.order) / c.Lz;
    (void)idx; (void)idy; (void)idz;
    les_delta15 = 0.0;
  }
};

// SMS filter width to the power 1+alpha from the three inverse spacings
// (calc_les.f90:83-85), evaluated on the host once per grid.
inline double sms_delta15(double idx, double idy, double idz) {
  const double mn = std::min(idx, std::min(idy, idz));
  const double a1 = mn / idx, a2 = mn / idy;
  const double la1 = std::log(a1), la2 = std::log(a2);
  const double delta =
      std::cosh(std::sqrt(4.0 * (la1 * la1 - la1 * la2 + la2 * la2) / 27.0)) * std::pow(idx * idy * idz, -1.0 / 3.0);
  return std::pow(delta, 1.0 + 0.5);
}

// Grid geometry of a uniform periodic grid (set_grid_cyclic*, all spacings
// equal).  Every metric array of the reference is constant, so the kernels
// take scalars.
struct Geom {
  int nx, ny, nz;        // extents incl. ghosts
  int h;                 // ghost layers per side (= order/2)
  int ix, iy, iz;        // interior sizes
  double jac;            // Jacobian(i,j) = 8/((dx+dx)(dy+dy)*2dz)
  double inv_dx, inv_dy, inv_dz;
  double a_xy, a_yz, a_zx;  // dt * face areas (pre_calc dtdxdy, dtdydz, dtdzdx)
  int bc[3];             // OX_BC_* of x, y, z
};

// Stage epilogue of the last sweep (3D_solver/src/calc_steps.f90):
enum Epilogue : int {
  EP_RK3_S1 = 0,   // calc_step1:   Qout = Qin - R
  EP_RK3_S23 = 1,  // calc_step2_3: Qout = (c1*Qin + c2*Qout - R) * c4inv
  EP_RK4_S = 2,    // calc_step:    Qout = Qin - c1*R ; Rs += c2*R
  EP_RK4_S4 = 3,   // calc_step4:   Q = Q - (Rs + R)*(1/6) ; Rs = 0
  EP_RHS = 4,      // write R (parity hook)
};

struct StageArgs {
  const double* Qin;  // state the RK combination reads as Q^n
  double* Qout;       // state written
  double* Rs;         // RK4 accumulator (state layout)
  double c1, c2, c4inv;
  int ep;
};

}  // namespace ox
