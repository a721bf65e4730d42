// ox_oracle.cpp -- the ORACLE's solver routines and its C entry points.
//
// TEST INFRASTRUCTURE ONLY (see ox_oracle.hpp).  Built into
// oracle/liboxoracle.so by oracle/Makefile; loaded by tests/ and by the
// cpu_baseline leg of bench.py.
#include "ox_oracle.hpp"

#include <cstdio>
#include <cstdlib>

namespace oxo {

// ---------------------------------------------------------------------------
// Setup: set_grid_cyclic{2,4,6}_3D (src/set_coordinate.f90:120-339),
// set_Jacobian_xy3 (:99-117), pre_calc metric arrays
// (3D_solver/src/preprocess.f90.fypp:72-87).
// ---------------------------------------------------------------------------
bool Oracle::init(const ox_config& cf) {
  cfg = cf;
  nx = cf.nx; ny = cf.ny; nz = cf.nz; order = cf.order;
  if (order != 2 && order != 4 && order != 6) { err = "order must be 2, 4 or 6"; return false; }
  io = order / 2 - 1;
  off = order / 2;
  vorder = cf.visc_order ? cf.visc_order : order;
  if (cf.visc != OX_VISC_EULER && vorder != 4 && vorder != 6) { err = "visc_order must be 4 or 6"; return false; }
  if (cf.visc != OX_VISC_EULER && vorder == 6 && order < 6) { err = "visc_order 6 needs order 6"; return false; }
  if (cf.scheme < OX_SCHEME_KEEP || cf.scheme > OX_SCHEME_HYBRID) { err = "unknown scheme"; return false; }
  for (int b : {cf.bc_x, cf.bc_y, cf.bc_z})
    if (b < OX_BC_PERIODIC || b > OX_BC_SLIP) { err = "unknown boundary kind"; return false; }
  if ((cf.bc_x || cf.bc_y || cf.bc_z) && (cf.visc == OX_VISC_LES || cf.scheme == OX_SCHEME_ROE)) {
    err = "non-periodic boundaries: Euler and NS with KEEP, SLAU and Hybrid only";
    return false;
  }
  // ME4Base replaces the central-6 stencil only (calc_visc_high_internal.f90.fypp:17-23);
  // with VISC_ORDER 4 the central-4 stencil is used whatever VISC_STENCIL says
  me4 = cf.visc != OX_VISC_EULER && cf.visc_stencil == OX_STENCIL_ME4BASE && vorder == 6;
  if (nx < 2 * order + 2 || ny < 2 * order + 2 || nz < 2 * order + 2) { err = "grid too small"; return false; }
  c.init(cf);

  // set_grid_cyclic: dx1 = L / (n - ORDER), uniform spacing
  dx.assign(nx - 1, cf.Lx / (double)(nx - order));
  dy.assign(ny - 1, cf.Ly / (double)(ny - order));
  dz.assign(nz - 1, cf.Lz / (double)(nz - order));
  // set_Jacobian_xy3 (interior formula, then edge copies)
  Jac.assign((size_t)nx * ny, 0.0);
  auto J = [&](int i, int j) -> double& { return Jac[(size_t)(i - 1) + (size_t)nx * (j - 1)]; };
  for (int j = 2; j <= ny - 1; ++j)
    for (int i = 2; i <= nx - 1; ++i)
      J(i, j) = 8.0 / ((dx[i - 2] + dx[i - 1]) * (dy[j - 2] + dy[j - 1]) * 2.0 * dz[0]);
  for (int i = 1; i <= nx; ++i) { J(i, 1) = J(i, 2); J(i, ny) = J(i, ny - 1); }
  for (int j = 1; j <= ny; ++j) { J(1, j) = J(2, j); J(nx, j) = J(nx - 1, j); }
  // pre_calc: inverse spacings and dt * face areas
  xix.resize(nx - 1); etay.resize(ny - 1); zetaz.resize(nz - 1);
  for (int i = 0; i < nx - 1; ++i) xix[i] = 1.0 / dx[i];
  for (int j = 0; j < ny - 1; ++j) etay[j] = 1.0 / dy[j];
  for (int k = 0; k < nz - 1; ++k) zetaz[k] = 1.0 / dz[k];
  const double dt = cf.dt;
  dtdxdy.resize((size_t)(nx - 2) * (ny - 2));
  dtdydz.resize((size_t)(ny - 2) * (nz - 2));
  dtdzdx.resize((size_t)(nx - 2) * (nz - 2));
  for (int j = 1; j <= ny - 2; ++j)
    for (int i = 1; i <= nx - 2; ++i)
      dtdxdy[(i - 1) + (size_t)(nx - 2) * (j - 1)] = 0.25 * dt * (dx[i - 1] + dx[i]) * (dy[j - 1] + dy[j]);
  for (int k = 1; k <= nz - 2; ++k)
    for (int j = 1; j <= ny - 2; ++j)
      dtdydz[(j - 1) + (size_t)(ny - 2) * (k - 1)] = 0.25 * dt * (dy[j - 1] + dy[j]) * (dz[k - 1] + dz[k]);
  for (int k = 1; k <= nz - 2; ++k)
    for (int i = 1; i <= nx - 2; ++i)
      dtdzdx[(i - 1) + (size_t)(nx - 2) * (k - 1)] = 0.25 * dt * (dz[k - 1] + dz[k]) * (dx[i - 1] + dx[i]);

  QJ.assign(nstate(), 0.0);
  QJ2.assign(nstate(), 0.0);
  Rs.assign((size_t)(nx - 2) * 5 * (ny - 2) * (nz - 2), 0.0);
  Qp.assign(nstate(), 0.0);
  T.assign(ncell(), 0.0);
  mu.assign(ncell(), 0.0);
  sensor.assign(ncell(), 0.0);
  mut.assign(ncell(), 0.0);
  qc2.assign(ncell(), 0.0);
  {
    // SMS filter width (calc_les.f90:83-85) with dx = 1/spacing; the (1+alpha)
    // power of it is a constant of the grid
    const double idx = xix[0], idy = etay[0], idz = zetaz[0];
    const double mn = std::min(idx, std::min(idy, idz));
    const double a1 = mn / idx, a2 = mn / idy;
    const double la1 = std::log(a1), la2 = std::log(a2);
    const double delta = std::cosh(std::sqrt(4.0 * (la1 * la1 - la1 * la2 + la2 * la2) / 27.0)) *
                         std::pow(idx * idy * idz, -1.0 / 3.0);
    les_delta15 = std::pow(delta, 1.0 + 0.5);
  }
  E.assign((size_t)5 * (nx - 1) * (ny - 2) * (nz - 2), 0.0);
  F.assign((size_t)5 * (nx - 2) * (ny - 1) * (nz - 2), 0.0);
  G.assign((size_t)5 * (nx - 2) * (ny - 2) * (nz - 1), 0.0);
  return true;
}

// ---------------------------------------------------------------------------
// calc_quantities_3D / calc_quantities_T_3D (src/calc_physical_quantities.f90:60-116)
// ---------------------------------------------------------------------------
void Oracle::primitives(const std::vector<double>& q) {
  const bool with_mu = cfg.visc != OX_VISC_EULER;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double Jij = Jac[(size_t)(i - 1) + (size_t)nx * (j - 1)];
        const double over_Q1 = 1.0 / q[iq(i, 1, j, k)];
        const double rho = Jij * q[iq(i, 1, j, k)];
        const double u = q[iq(i, 2, j, k)] * over_Q1;
        const double v = q[iq(i, 3, j, k)] * over_Q1;
        const double w = q[iq(i, 4, j, k)] * over_Q1;
        const double p = c.gamma_1 * (Jij * q[iq(i, 5, j, k)] - 0.5 * rho * (u * u + v * v + w * w));
        Qp[iq(i, 1, j, k)] = rho;
        Qp[iq(i, 2, j, k)] = u;
        Qp[iq(i, 3, j, k)] = v;
        Qp[iq(i, 4, j, k)] = w;
        Qp[iq(i, 5, j, k)] = p;
        const double temp = p / (c.Rg * rho);
        T[i3(i, j, k)] = temp;
        if (with_mu) mu[i3(i, j, k)] = c.mu0_T0_S_over_T0_2_3 / (temp + 111.0) * (temp * fsqrt(temp));
      }
}

// ---------------------------------------------------------------------------
// calc_Ducros (3D_solver/src/calc_hybrid.f90:11-72)
// ---------------------------------------------------------------------------
void Oracle::ducros() {
  const std::vector<double>& Q = Qp;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        const double dx_tmp = 0.25 * (xix[i - 2] + xix[i - 1]);
        const double dy_tmp = 0.25 * (etay[j - 2] + etay[j - 1]);
        const double dz_tmp = 0.25 * (zetaz[k - 2] + zetaz[k - 1]);
        const double dudx = (-Q[iq(i - 1, 2, j, k)] + Q[iq(i + 1, 2, j, k)]) * dx_tmp;
        const double dvdx = (-Q[iq(i - 1, 3, j, k)] + Q[iq(i + 1, 3, j, k)]) * dx_tmp;
        const double dwdx = (-Q[iq(i - 1, 4, j, k)] + Q[iq(i + 1, 4, j, k)]) * dx_tmp;
        const double dudy = (-Q[iq(i, 2, j - 1, k)] + Q[iq(i, 2, j + 1, k)]) * dy_tmp;
        const double dvdy = (-Q[iq(i, 3, j - 1, k)] + Q[iq(i, 3, j + 1, k)]) * dy_tmp;
        const double dwdy = (-Q[iq(i, 4, j - 1, k)] + Q[iq(i, 4, j + 1, k)]) * dy_tmp;
        const double dudz = (-Q[iq(i, 2, j, k - 1)] + Q[iq(i, 2, j, k + 1)]) * dz_tmp;
        const double dvdz = (-Q[iq(i, 3, j, k - 1)] + Q[iq(i, 3, j, k + 1)]) * dz_tmp;
        const double dwdz = (-Q[iq(i, 4, j, k - 1)] + Q[iq(i, 4, j, k + 1)]) * dz_tmp;
        const double div = dudx + dvdy + dwdz;
        const double r1 = dwdy - dvdz;
        const double r2 = dudz - dwdx;
        const double r3 = dvdx - dudy;
        double fd = (div * div) / (div * div + (r1 * r1 + r2 * r2 + r3 * r3) + 1.0e-12);
        fd = fmin2(1.0, fd);
        sensor[i3(i, j, k)] = fd;
      }
  // boundary copies (:54-71)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j) {
      sensor[i3(1, j, k)] = sensor[i3(2, j, k)];
      sensor[i3(nx, j, k)] = sensor[i3(nx - 1, j, k)];
    }
  for (int k = 2; k <= nz - 1; ++k)
    for (int i = 2; i <= nx - 1; ++i) {
      sensor[i3(i, 1, k)] = sensor[i3(i, 2, k)];
      sensor[i3(i, ny, k)] = sensor[i3(i, ny - 1, k)];
    }
  for (int j = 2; j <= ny - 1; ++j)
    for (int i = 2; i <= nx - 1; ++i) {
      sensor[i3(i, j, 1)] = sensor[i3(i, j, 2)];
      sensor[i3(i, j, nz)] = sensor[i3(i, j, nz - 1)];
    }
}

// ---------------------------------------------------------------------------
// Face flux of one convective scheme from a stencil gathered along `dir`.
// KEEP: calc_keep_{x,y,z}_in (3D_solver/src/calc_keep_kernel_internal.f90.fypp:30-153)
// SLAU: calc_slau_{x,y,z}_in (3D_solver/src/calc_slau_kernel_internal.f90.fypp:31-252)
// st[v][m]: v = rho,u,v,w,p,T ; m = stencil position 0..2*io+1
// ---------------------------------------------------------------------------
static const double kNormal[3][5] = {{0, 1, 0, 0, 0}, {0, 0, 1, 0, 0}, {0, 0, 0, 1, 0}};

// MUSCL left/right states of the five primitives: interp2/4/6 of the SLAU
// kernel (calc_slau_kernel_internal.f90.fypp:31-81) and delta_r2/4/6 of the
// Roe and hybrid kernels (calc_roe_kernel_internal.f90.fypp:28-55) -- the same
// reconstruction; only SLAU replaces the sensor by the wiggle flag afterwards.
static void reconstruct(const Oracle& o, int n, double st[6][6], double fsens, double* L, double* Rr) {
  if (n == 2) {
    for (int v = 0; v < 5; ++v) { L[v] = st[v][0]; Rr[v] = st[v][1]; }
  } else if (n == 4) {
    for (int v = 0; v < 5; ++v) delta4(o.c, o.cfg.tvd, fsens, st[v], L[v], Rr[v]);
  } else {
    for (int v = 0; v < 5; ++v) delta6(o.c, o.cfg.tvd, fsens, st[v], L[v], Rr[v]);
  }
}

static void conv_face(const Oracle& o, int dir, double st[6][6], double fsens, double* out) {
  const Consts& c = o.c;
  const double* N = kNormal[dir];
  const int n = 2 * o.io + 2;
  const double* uu = st[1 + dir];
  int scheme = o.cfg.scheme;
  // calc_hybrid_{x,y,z}_in (calc_hybrid_kernel_internal.f90.fypp:96-121): KEEP where
  // the face sensor is <= threshold, else MUSCL + SLAU with HR = 1
  if (scheme == OX_SCHEME_HYBRID) scheme = (fsens <= c.threshold) ? OX_SCHEME_KEEP : -1;
  if (scheme == OX_SCHEME_KEEP) {
    if (n == 2) keep2(c, st[0], st[1], st[2], st[3], uu, st[4], st[5], N, out);
    else if (n == 4) keep4(c, st[0], st[1], st[2], st[3], uu, st[4], st[5], N, out);
    else keep6(c, st[0], st[1], st[2], st[3], uu, st[4], st[5], N, out);
    return;
  }
  double L[5], Rr[5];
  reconstruct(o, n, st, fsens, L, Rr);
  if (scheme == OX_SCHEME_ROE) {
    // calc_roe_{x,y,z}_in: local frame (normal, t1, t2) in cyclic order
    const int a = 1 + dir, b = 1 + (dir + 1) % 3, d = 1 + (dir + 2) % 3;
    double loc[5];
    roe_flux(c, L[0], Rr[0], L[a], Rr[a], L[b], Rr[b], L[d], Rr[d], L[4], Rr[4], loc);
    out[0] = loc[0]; out[a] = loc[1]; out[b] = loc[2]; out[d] = loc[3]; out[4] = loc[4];
    return;
  }
  double HR;
  int variant = o.cfg.slau_variant;
  if (scheme == OX_SCHEME_SLAU) {
    HR = n == 2 ? 1.0 : (n == 4 ? wiggle_detector(&st[4][0]) : wiggle_detector(&st[4][1]));
  } else {  // hybrid, shock branch
    HR = 1.0;
  }
  slau_flux(c, variant, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], L[3], Rr[3], L[1 + dir], Rr[1 + dir], L[4], Rr[4], N,
            HR, out);
}

void Oracle::conv_x() {
  const bool slau = cfg.scheme != OX_SCHEME_KEEP;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = io + 1; i <= nx - (io + 1); ++i) {
        double st[6][6];
        for (int m = 0; m < 2 * io + 2; ++m) {
          const int ii = i - io + m;
          for (int v = 0; v < 5; ++v) st[v][m] = Qp[iq(ii, v + 1, j, k)];
          st[5][m] = T[i3(ii, j, k)];
        }
        const double fs = slau ? 0.5 * (sensor[i3(i, j, k)] + sensor[i3(i + 1, j, k)]) : 0.0;
        double out[5];
        conv_face(*this, 0, st, fs, out);
        for (int l = 1; l <= 5; ++l) E[iE(l, i, j - 1, k - 1)] = out[l - 1];
      }
}

void Oracle::conv_y() {
  const bool slau = cfg.scheme != OX_SCHEME_KEEP;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = io + 1; j <= ny - (io + 1); ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        double st[6][6];
        for (int m = 0; m < 2 * io + 2; ++m) {
          const int jj = j - io + m;
          for (int v = 0; v < 5; ++v) st[v][m] = Qp[iq(i, v + 1, jj, k)];
          st[5][m] = T[i3(i, jj, k)];
        }
        const double fs = slau ? 0.5 * (sensor[i3(i, j, k)] + sensor[i3(i, j + 1, k)]) : 0.0;
        double out[5];
        conv_face(*this, 1, st, fs, out);
        for (int l = 1; l <= 5; ++l) F[iF(l, i - 1, j, k - 1)] = out[l - 1];
      }
}

void Oracle::conv_z() {
  const bool slau = cfg.scheme != OX_SCHEME_KEEP;
#pragma omp parallel for schedule(static)
  for (int k = io + 1; k <= nz - (io + 1); ++k)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        double st[6][6];
        for (int m = 0; m < 2 * io + 2; ++m) {
          const int kk = k - io + m;
          for (int v = 0; v < 5; ++v) st[v][m] = Qp[iq(i, v + 1, j, kk)];
          st[5][m] = T[i3(i, j, kk)];
        }
        const double fs = slau ? 0.5 * (sensor[i3(i, j, k)] + sensor[i3(i, j, k + 1)]) : 0.0;
        double out[5];
        conv_face(*this, 2, st, fs, out);
        for (int l = 1; l <= 5; ++l) G[iG(l, i - 1, j - 1, k)] = out[l - 1];
      }
}

// ---------------------------------------------------------------------------
// Central viscous terms.  Cross derivatives as computed by the LDS loaders
// (3D_solver/src/load_smem_visc_cent.f90.fypp:52-66, :118-132, :184-198);
// face stress and heat flux as calc_{E,F,G}v4 (calc_visc_high.f90.fypp:114-140
// and counterparts) and calc_{E,F,G}v6_in (calc_visc_high_internal.f90.fypp:185-211
// and counterparts).  Only the full-order masks are evaluated: the 2nd-order
// fallback of calc_visc_high.f90.fypp:194 only ever writes faces that feed
// ghost cells, which set_bc overwrites after the stage.
// ---------------------------------------------------------------------------
// d/d(dir) of primitive variable `var` (2=u,3=v,4=w) at cell (i,j,k)
static double cross_deriv(const Oracle& o, int dir, int var, int i, int j, int k) {
  const Consts& c = o.c;
  auto q = [&](int s) {
    if (dir == 0) return o.Qp[o.iq(i + s, var, j, k)];
    if (dir == 1) return o.Qp[o.iq(i, var, j + s, k)];
    return o.Qp[o.iq(i, var, j, k + s)];
  };
  const double inv = dir == 0 ? o.xix[i - 1] : (dir == 1 ? o.etay[j - 1] : o.zetaz[k - 1]);
  // 4th-order central: VISC_ORDER 4, and ME4Base (load_smem_visc_me4_base.f90:67-73)
  if (o.vorder == 4 || o.me4)
    return (c.two_third * (-q(-1) + q(1)) - c.one_twelfth * (-q(-2) + q(2))) * inv;
  return (c.one_sixty * (-q(-3) + q(3)) + 0.15 * (q(-2) - q(2)) + 0.75 * (-q(-1) + q(1))) * inv;
}

// One viscous face in direction `dir` at face index f (the face between
// cells f and f+1 along dir) of row (a,b): returns the 4 corrections.
static void visc_face(const Oracle& o, int dir, int i, int j, int k, double* tau, double& energy) {
  const Consts& c = o.c;
  const int N = o.vorder;
  const int lo = (N == 4) ? -1 : -2;
  const bool les = o.cfg.visc == OX_VISC_LES;
  double mu_[6], T_[6], u_[6], v_[6], w_[6], d1[6], d2[6], d3[6], d4[6], mt[6], H[6];
  for (int m = 0; m < N; ++m) {
    const int s = lo + m;
    const int ii = i + (dir == 0 ? s : 0), jj = j + (dir == 1 ? s : 0), kk = k + (dir == 2 ? s : 0);
    mu_[m] = o.mu[o.i3(ii, jj, kk)];
    T_[m] = o.T[o.i3(ii, jj, kk)];
    u_[m] = o.Qp[o.iq(ii, 2, jj, kk)];
    v_[m] = o.Qp[o.iq(ii, 3, jj, kk)];
    w_[m] = o.Qp[o.iq(ii, 4, jj, kk)];
    if (les) {
      mt[m] = o.mut[o.i3(ii, jj, kk)];
      // H = Cp*T + 0.5*|u|^2 + qc2 (calc_visc_high_internal.f90.fypp:196-198)
      H[m] = c.Cp * T_[m] + 0.5 * (u_[m] * u_[m] + v_[m] * v_[m] + w_[m] * w_[m]) + o.qc2[o.i3(ii, jj, kk)];
    }
    if (dir == 0) {  // uy, vy, uz, wz
      d1[m] = cross_deriv(o, 1, 2, ii, jj, kk);
      d2[m] = cross_deriv(o, 1, 3, ii, jj, kk);
      d3[m] = cross_deriv(o, 2, 2, ii, jj, kk);
      d4[m] = cross_deriv(o, 2, 4, ii, jj, kk);
    } else if (dir == 1) {  // ux, vx, vz, wz
      d1[m] = cross_deriv(o, 0, 2, ii, jj, kk);
      d2[m] = cross_deriv(o, 0, 3, ii, jj, kk);
      d3[m] = cross_deriv(o, 2, 3, ii, jj, kk);
      d4[m] = cross_deriv(o, 2, 4, ii, jj, kk);
    } else {  // ux, wx, vy, wy
      d1[m] = cross_deriv(o, 0, 2, ii, jj, kk);
      d2[m] = cross_deriv(o, 0, 4, ii, jj, kk);
      d3[m] = cross_deriv(o, 1, 3, ii, jj, kk);
      d4[m] = cross_deriv(o, 1, 4, ii, jj, kk);
    }
  }
  const double d = dir == 0 ? o.xix[i - 1] : (dir == 1 ? o.etay[j - 1] : o.zetaz[k - 1]);
  double ta, tb, tc, ua, ub, uc;
  if (o.me4) {
    // calc_visc_high_internal.f90.fypp:146-183 and its F/G counterparts
    double mu3[3], mut3[3], kT3[3];
    for (int t = 0; t < 3; ++t) {
      mu3[t] = me4_interp(mu_, t);
      if (les) mut3[t] = me4_interp(mt, t);
      kT3[t] = c.Cp_over_Pr * mu3[t] * (1.125 * (-T_[t + 1] + T_[t + 2]) - (-T_[t] + T_[t + 3]) * c.one_24) * d;
    }
    const double kT = flux4(c, kT3);
    auto straight = [&](const double* vel, const double* a, const double* b, double& t, double& ut) {
      if (les) tau_straight_me4_les(c, mu3, mut3, vel, a, b, d, t, ut);
      else tau_straight_me4(c, mu3, vel, a, b, d, t, ut);
    };
    auto cross = [&](const double* vel, const double* g, double& t, double& ut) {
      if (les) tau_cross_me4_les(c, mu3, mut3, vel, g, d, t, ut);
      else tau_cross_me4(c, mu3, vel, g, d, t, ut);
    };
    // SGS heat flux over the 4 central cells (H(i-1:i+2))
    const double Hsgs =
        les ? -flux4(c, mut3) * (1.125 * (-H[2] + H[3]) - (-H[1] + H[4]) * c.one_24) * d / c.Prt : 0.0;
    if (dir == 0) {
      straight(u_, d2, d4, ta, ua);
      cross(v_, d1, tb, ub);
      cross(w_, d3, tc, uc);
    } else if (dir == 1) {
      straight(v_, d4, d1, tb, ub);
      cross(u_, d2, ta, ua);
      cross(w_, d3, tc, uc);
    } else {
      straight(w_, d1, d3, tc, uc);
      cross(u_, d2, ta, ua);
      cross(v_, d4, tb, ub);
    }
    tau[0] = ta; tau[1] = tb; tau[2] = tc;
    energy = les ? ua + ub + uc + kT + Hsgs : ua + ub + uc + kT;
    return;
  }
  const double kT = c.Cp_over_Pr * vinterp(N, mu_) * vdiff(c, N, d, T_);
  auto straight = [&](const double* vel, const double* a, const double* b, double& t, double& ut) {
    if (les) tau_straight_les(c, N, mu_, mt, vel, a, b, d, t, ut);
    else tau_straight(c, N, mu_, vel, a, b, d, t, ut);
  };
  auto cross = [&](const double* vel, const double* g, double& t, double& ut) {
    if (les) tau_cross_les(c, N, mu_, mt, vel, g, d, t, ut);
    else tau_cross(c, N, mu_, vel, g, d, t, ut);
  };
  // SGS heat flux -interp(mut)*diff(H)/Prt (calc_visc_high_internal.f90.fypp:199)
  const double Hsgs = les ? -vinterp(N, mt) * vdiff(c, N, d, H) / c.Prt : 0.0;
  if (dir == 0) {
    // txx, txy, txz
    straight(u_, d2 /*vy*/, d4 /*wz*/, ta, ua);
    cross(v_, d1 /*uy*/, tb, ub);
    cross(w_, d3 /*uz*/, tc, uc);
    tau[0] = ta; tau[1] = tb; tau[2] = tc;
  } else if (dir == 1) {
    // tyy (straight), tyx, tyz
    straight(v_, d4 /*wz*/, d1 /*ux*/, tb, ub);
    cross(u_, d2 /*vx*/, ta, ua);
    cross(w_, d3 /*vz*/, tc, uc);
    tau[0] = ta; tau[1] = tb; tau[2] = tc;
  } else {
    // tzz (straight), tzx, tzy
    straight(w_, d1 /*ux*/, d3 /*vy*/, tc, uc);
    cross(u_, d2 /*wx*/, ta, ua);
    cross(v_, d4 /*wy*/, tb, ub);
    tau[0] = ta; tau[1] = tb; tau[2] = tc;
  }
  energy = les ? ua + ub + uc + kT + Hsgs : ua + ub + uc + kT;
}

void Oracle::visc_x() {
  const int ilo = vorder == 4 ? 2 : 3, ihi = vorder == 4 ? nx - 2 : nx - 3;
  const int jlo = vorder == 4 ? 3 : 4, jhi = vorder == 4 ? ny - 2 : ny - 3;
  const int klo = vorder == 4 ? 3 : 4, khi = vorder == 4 ? nz - 2 : nz - 3;
#pragma omp parallel for schedule(static)
  for (int k = klo; k <= khi; ++k)
    for (int j = jlo; j <= jhi; ++j)
      for (int i = ilo; i <= ihi; ++i) {
        double tau[3], en;
        visc_face(*this, 0, i, j, k, tau, en);
        E[iE(2, i, j - 1, k - 1)] = E[iE(2, i, j - 1, k - 1)] - tau[0];
        E[iE(3, i, j - 1, k - 1)] = E[iE(3, i, j - 1, k - 1)] - tau[1];
        E[iE(4, i, j - 1, k - 1)] = E[iE(4, i, j - 1, k - 1)] - tau[2];
        E[iE(5, i, j - 1, k - 1)] = E[iE(5, i, j - 1, k - 1)] - en;
      }
}

void Oracle::visc_y() {
  const int ilo = vorder == 4 ? 3 : 4, ihi = vorder == 4 ? nx - 2 : nx - 3;
  const int jlo = vorder == 4 ? 2 : 3, jhi = vorder == 4 ? ny - 2 : ny - 3;
  const int klo = vorder == 4 ? 3 : 4, khi = vorder == 4 ? nz - 2 : nz - 3;
#pragma omp parallel for schedule(static)
  for (int k = klo; k <= khi; ++k)
    for (int j = jlo; j <= jhi; ++j)
      for (int i = ilo; i <= ihi; ++i) {
        double tau[3], en;
        visc_face(*this, 1, i, j, k, tau, en);
        F[iF(2, i - 1, j, k - 1)] = F[iF(2, i - 1, j, k - 1)] - tau[0];
        F[iF(3, i - 1, j, k - 1)] = F[iF(3, i - 1, j, k - 1)] - tau[1];
        F[iF(4, i - 1, j, k - 1)] = F[iF(4, i - 1, j, k - 1)] - tau[2];
        F[iF(5, i - 1, j, k - 1)] = F[iF(5, i - 1, j, k - 1)] - en;
      }
}

void Oracle::visc_z() {
  const int ilo = vorder == 4 ? 3 : 4, ihi = vorder == 4 ? nx - 2 : nx - 3;
  const int jlo = vorder == 4 ? 3 : 4, jhi = vorder == 4 ? ny - 2 : ny - 3;
  const int klo = vorder == 4 ? 2 : 3, khi = vorder == 4 ? nz - 2 : nz - 3;
#pragma omp parallel for schedule(static)
  for (int k = klo; k <= khi; ++k)
    for (int j = jlo; j <= jhi; ++j)
      for (int i = ilo; i <= ihi; ++i) {
        double tau[3], en;
        visc_face(*this, 2, i, j, k, tau, en);
        G[iG(2, i - 1, j - 1, k)] = G[iG(2, i - 1, j - 1, k)] - tau[0];
        G[iG(3, i - 1, j - 1, k)] = G[iG(3, i - 1, j - 1, k)] - tau[1];
        G[iG(4, i - 1, j - 1, k)] = G[iG(4, i - 1, j - 1, k)] - tau[2];
        G[iG(5, i - 1, j - 1, k)] = G[iG(5, i - 1, j - 1, k)] - en;
      }
}

// ---------------------------------------------------------------------------
// Selective mixed-scale SGS model.  calc_les.f90:12-125; boundary layers by
// set_bc_mut_common (3D_solver/src/set_bc_common.f90:296-335).
// The exponents alpha/2 = (1-alpha)/2 = 1/4 are applied as sqrt(sqrt(.)).
// ---------------------------------------------------------------------------
static double test_filter(const double x[3][3][3]) {  // x[a][b][c] = Fortran x(a+1,b+1,c+1)
  return x[1][1][1] + 0.25 * (x[0][1][1] - 2.0 * x[1][1][1] + x[2][1][1] + x[1][0][1] - 2.0 * x[1][1][1] +
                              x[1][2][1] + x[1][1][0] - 2.0 * x[1][1][1] + x[1][1][2]);
}

static void differential(const double u[3][3][3], const double v[3][3][3], const double w[3][3][3], double dx,
                         double dy, double dz, double* g) {
  g[0] = 0.5 * (-u[0][1][1] + u[2][1][1]) * dx;  // dudx
  g[1] = 0.5 * (-u[1][0][1] + u[1][2][1]) * dy;  // dudy
  g[2] = 0.5 * (-u[1][1][0] + u[1][1][2]) * dz;  // dudz
  g[3] = 0.5 * (-v[0][1][1] + v[2][1][1]) * dx;
  g[4] = 0.5 * (-v[1][0][1] + v[1][2][1]) * dy;
  g[5] = 0.5 * (-v[1][1][0] + v[1][1][2]) * dz;
  g[6] = 0.5 * (-w[0][1][1] + w[2][1][1]) * dx;
  g[7] = 0.5 * (-w[1][0][1] + w[1][2][1]) * dy;
  g[8] = 0.5 * (-w[1][1][0] + w[1][1][2]) * dz;
}

static double sms(const Consts& c, const double u[3][3][3], const double v[3][3][3], const double w[3][3][3],
                  const double ut[3][3][3], const double vt[3][3][3], const double wt[3][3][3], double dx, double dy,
                  double dz, double qc2, double delta15) {
  double g[9], h[9];
  differential(u, v, w, dx, dy, dz, g);
  differential(ut, vt, wt, dx, dy, dz, h);
  const double vor1 = g[7] - g[5], vor2_ = g[2] - g[6], vor3 = g[3] - g[1];
  const double vh1 = h[7] - h[5], vh2 = h[2] - h[6], vh3 = h[3] - h[1];
  const double d1 = vor1 - vh1, d2 = vor2_ - vh2, d3 = vor3 - vh3;
  const double vor2 = vor1 * vor1 + vor2_ * vor2_ + vor3 * vor3;
  const double vorh2 = vh1 * vh1 + vh2 * vh2 + vh3 * vh3;
  const double vdiff2 = d1 * d1 + d2 * d2 + d3 * d3;
  const double svor2 = 2.0 * fsqrt(vor2) * fsqrt(vorh2) + vorh2 + vor2 - vdiff2;
  double tan2;
  if (fabs_(svor2) < 1.0e-30) tan2 = 0.0;
  else tan2 = (2.0 * fsqrt(vorh2) * fsqrt(vor2) - vorh2 - vor2 + vdiff2) / svor2;
  const double rtheta = tan2 / c.tan10deg2;
  const double ftheta = fmin2(1.0, rtheta * rtheta);
  const double S2 = 2.0 * (g[0] * g[0] + g[4] * g[4] + g[8] * g[8]) +
                    ((g[3] + g[1]) * (g[3] + g[1]) + (g[7] + g[5]) * (g[7] + g[5]) + (g[2] + g[6]) * (g[2] + g[6]));
  const double Cm = 0.06;
  return ftheta * Cm * fsqrt(fsqrt(S2)) * fsqrt(fsqrt(qc2)) * delta15;
}

void Oracle::calc_mut() {
  const std::vector<double>& Q = Qp;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        double u3[3][3][3], v3[3][3][3], w3[3][3][3], ut[3][3][3], vt[3][3][3], wt[3][3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int cc = 0; cc < 3; ++cc) {
              u3[a][b][cc] = Q[iq(i - 1 + a, 2, j - 1 + b, k - 1 + cc)];
              v3[a][b][cc] = Q[iq(i - 1 + a, 3, j - 1 + b, k - 1 + cc)];
              w3[a][b][cc] = Q[iq(i - 1 + a, 4, j - 1 + b, k - 1 + cc)];
            }
        if (3 <= i && i <= nx - 2 && 3 <= j && j <= ny - 2 && 3 <= k && k <= nz - 2) {
          // stride_filter of the 5^3 neighbourhood (:88-94)
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              for (int cc = 0; cc < 3; ++cc) {
                double xu[3][3][3], xv[3][3][3], xw[3][3][3];
                for (int p = 0; p < 3; ++p)
                  for (int q = 0; q < 3; ++q)
                    for (int r = 0; r < 3; ++r) {
                      const int ii = i - 2 + a + p, jj = j - 2 + b + q, kk = k - 2 + cc + r;
                      xu[p][q][r] = Q[iq(ii, 2, jj, kk)];
                      xv[p][q][r] = Q[iq(ii, 3, jj, kk)];
                      xw[p][q][r] = Q[iq(ii, 4, jj, kk)];
                    }
                ut[a][b][cc] = test_filter(xu);
                vt[a][b][cc] = test_filter(xv);
                wt[a][b][cc] = test_filter(xw);
              }
        } else {
          std::memcpy(ut, u3, sizeof(u3));
          std::memcpy(vt, v3, sizeof(v3));
          std::memcpy(wt, w3, sizeof(w3));
        }
        const double du = u3[1][1][1] - ut[1][1][1], dv = v3[1][1][1] - vt[1][1][1], dw = w3[1][1][1] - wt[1][1][1];
        const double q2 = 0.5 * (du * du + dv * dv + dw * dw);
        qc2[i3(i, j, k)] = q2;
        mut[i3(i, j, k)] = Q[iq(i, 1, j, k)] *
                           sms(c, u3, v3, w3, ut, vt, wt, xix[i - 1], etay[j - 1], zetaz[k - 1], q2, les_delta15);
      }
  // set_bc_mut_common: x faces, y faces, xy corners, then z planes
  auto bc = [&](std::vector<double>& a) {
    for (int k = 2; k <= nz - 1; ++k)
      for (int j = 2; j <= ny - 1; ++j) {
        a[i3(1, j, k)] = a[i3(nx - 1, j, k)];
        a[i3(nx, j, k)] = a[i3(2, j, k)];
      }
    for (int k = 2; k <= nz - 1; ++k)
      for (int i = 2; i <= nx - 1; ++i) {
        a[i3(i, 1, k)] = a[i3(i, ny - 1, k)];
        a[i3(i, ny, k)] = a[i3(i, 2, k)];
      }
    for (int k = 2; k <= nz - 1; ++k) {
      a[i3(1, 1, k)] = a[i3(nx - 1, ny - 1, k)];
      a[i3(nx, 1, k)] = a[i3(2, ny - 1, k)];
      a[i3(1, ny, k)] = a[i3(nx - 1, 2, k)];
      a[i3(nx, ny, k)] = a[i3(2, 2, k)];
    }
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        a[i3(i, j, 1)] = a[i3(i, j, nz - 1)];
        a[i3(i, j, nz)] = a[i3(i, j, 2)];
      }
  };
  bc(mut);
  bc(qc2);
}

// calc_EFG (3D_solver/src/calc_flux_base.f90.fypp:109-198), periodic, no COMMZ
void Oracle::calc_EFG(const std::vector<double>& q) {
  primitives(q);
  if (cfg.scheme != OX_SCHEME_KEEP) ducros();
  conv_x();
  conv_y();
  conv_z();
  if (cfg.visc == OX_VISC_LES) calc_mut();
  if (cfg.visc != OX_VISC_EULER) {
    visc_x();
    visc_y();
    visc_z();
  }
}

// ---------------------------------------------------------------------------
// calc_R and the stage kernels (3D_solver/src/calc_steps.f90:8-184).
// Loop indices i,j,k = 1..n-2 address cell (i+1,j+1,k+1).
// ---------------------------------------------------------------------------
void Oracle::calc_R(int i, int j, int k, double cxy, double cyz, double czx, double* R) const {
  for (int l = 1; l <= 5; ++l) R[l - 1] = cyz * (-E[iE(l, i, j, k)] + E[iE(l, i + 1, j, k)]);
  for (int l = 1; l <= 5; ++l) R[l - 1] = R[l - 1] + czx * (-F[iF(l, i, j, k)] + F[iF(l, i, j + 1, k)]);
  for (int l = 1; l <= 5; ++l) R[l - 1] = R[l - 1] + cxy * (-G[iG(l, i, j, k)] + G[iG(l, i, j, k + 1)]);
}

#define OX_AREA(i, j, k)                                      \
  const double a_xy = dtdxdy[(i - 1) + (size_t)(nx - 2) * (j - 1)]; \
  const double a_yz = dtdydz[(j - 1) + (size_t)(ny - 2) * (k - 1)]; \
  const double a_zx = dtdzdx[(i - 1) + (size_t)(nx - 2) * (k - 1)];

void Oracle::step1(double coef, const std::vector<double>& Q, std::vector<double>& Q2) {  // :42-71
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz - 2; ++k)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i) {
        OX_AREA(i, j, k);
        double R[5];
        calc_R(i, j, k, coef * a_xy, coef * a_yz, coef * a_zx, R);
        for (int l = 1; l <= 5; ++l) Q2[iq(i + 1, l, j + 1, k + 1)] = Q[iq(i + 1, l, j + 1, k + 1)] - R[l - 1];
      }
}

void Oracle::step_rk4(double coef1, double coef2, const std::vector<double>& Q, std::vector<double>& Q2) {  // :74-107
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz - 2; ++k)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i) {
        OX_AREA(i, j, k);
        double R[5];
        calc_R(i, j, k, a_xy, a_yz, a_zx, R);
        for (int l = 1; l <= 5; ++l) {
          Q2[iq(i + 1, l, j + 1, k + 1)] = Q[iq(i + 1, l, j + 1, k + 1)] - coef1 * R[l - 1];
          Rs[iRs(i, l, j, k)] = Rs[iRs(i, l, j, k)] + coef2 * R[l - 1];
        }
      }
}

void Oracle::step2_3(double c1, double c2, double c3, double c4inv, const std::vector<double>& Qin,
                     std::vector<double>& Qout) {  // :111-146
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz - 2; ++k)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i) {
        OX_AREA(i, j, k);
        double R[5];
        calc_R(i, j, k, c3 * a_xy, c3 * a_yz, c3 * a_zx, R);
        for (int l = 1; l <= 5; ++l) {
          const size_t m = iq(i + 1, l, j + 1, k + 1);
          Qout[m] = (c1 * Qin[m] + c2 * Qout[m] - R[l - 1]) * c4inv;
        }
      }
}

void Oracle::step4(std::vector<double>& Q) {  // :150-184
  const double one_sixth = c.one_sixth;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz - 2; ++k)
    for (int j = 1; j <= ny - 2; ++j)
      for (int i = 1; i <= nx - 2; ++i) {
        OX_AREA(i, j, k);
        double R[5];
        calc_R(i, j, k, a_xy, a_yz, a_zx, R);
        for (int l = 1; l <= 5; ++l) {
          R[l - 1] = Rs[iRs(i, l, j, k)] + R[l - 1];
          Q[iq(i + 1, l, j + 1, k + 1)] = Q[iq(i + 1, l, j + 1, k + 1)] - R[l - 1] * one_sixth;
          Rs[iRs(i, l, j, k)] = 0.0;
        }
      }
}

// set_bc_cyclic{2,4,6} (3D_solver/src/set_bc_common.f90:101-273): x faces over
// interior (j,k), y faces over interior (i,k), xy edges over interior k, then
// whole z planes.
void Oracle::set_bc(std::vector<double>& Q) const {
  if (cfg.bc_x || cfg.bc_y || cfg.bc_z) {
    set_bc_walls(Q);
    return;
  }
  const int h = off;
  const int ix = nx - 2 * h, iy = ny - 2 * h, iz = nz - 2 * h;
  for (int k = h + 1; k <= nz - h; ++k)
    for (int j = h + 1; j <= ny - h; ++j)
      for (int l = 1; l <= 5; ++l)
        for (int g = 1; g <= h; ++g) {
          Q[iq(g, l, j, k)] = Q[iq(g + ix, l, j, k)];
          Q[iq(nx - h + g, l, j, k)] = Q[iq(h + g, l, j, k)];
        }
  for (int k = h + 1; k <= nz - h; ++k)
    for (int i = h + 1; i <= nx - h; ++i)
      for (int l = 1; l <= 5; ++l)
        for (int g = 1; g <= h; ++g) {
          Q[iq(i, l, g, k)] = Q[iq(i, l, g + iy, k)];
          Q[iq(i, l, ny - h + g, k)] = Q[iq(i, l, h + g, k)];
        }
  for (int k = h + 1; k <= nz - h; ++k)
    for (int l = 1; l <= 5; ++l)
      for (int a = 1; a <= h; ++a)
        for (int b = 1; b <= h; ++b) {
          Q[iq(a, l, b, k)] = Q[iq(a + ix, l, b + iy, k)];
          Q[iq(nx - h + a, l, b, k)] = Q[iq(h + a, l, b + iy, k)];
          Q[iq(a, l, ny - h + b, k)] = Q[iq(a + ix, l, h + b, k)];
          Q[iq(nx - h + a, l, ny - h + b, k)] = Q[iq(h + a, l, h + b, k)];
        }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int l = 1; l <= 5; ++l)
        for (int g = 1; g <= h; ++g) {
          Q[iq(i, l, j, g)] = Q[iq(i, l, j, g + iz)];
          Q[iq(i, l, j, nz - h + g)] = Q[iq(i, l, j, h + g)];
        }
}

// A case's set_bc with non-periodic directions, written after
// 3D_solver/STZ/set.f90:77-138: direction by direction in the order x, y, z,
// each over the whole extent of the other two.  Per direction: periodic
// (the image of the interior, as set_bc_cyclic), zero gradient (every ghost
// layer copies the adjacent interior cell, :113-138) or slip wall (ghost
// ovlp-m mirrors ovlp+1+m with the normal momentum negated, :77-112).
void Oracle::set_bc_walls(std::vector<double>& Q) const {
  const int h = off;
  const int kinds[3] = {cfg.bc_x, cfg.bc_y, cfg.bc_z};
  const int n[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) {
    const int N = n[d], m = N - 2 * h;
    // the (1-based) source of ghost index g along d, and the sign of the normal momentum
    auto src = [&](int g, double& sgn) {
      sgn = 1.0;
      if (kinds[d] == OX_BC_PERIODIC) return g <= h ? g + m : g - m;
      if (kinds[d] == OX_BC_EXTRAPOLATE) return g <= h ? h + 1 : N - h;
      sgn = -1.0;
      return g <= h ? 2 * h + 1 - g : 2 * (N - h) + 1 - g;
    };
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const int p = d == 0 ? i : (d == 1 ? j : k);
          if (p > h && p <= N - h) continue;
          double sgn;
          const int s = src(p, sgn);
          const int si = d == 0 ? s : i, sj = d == 1 ? s : j, sk = d == 2 ? s : k;
          for (int l = 1; l <= 5; ++l) {
            const double v = Q[iq(si, l, sj, sk)];
            Q[iq(i, l, j, k)] = (l == 2 + d && sgn < 0.0) ? -v : v;
          }
        }
  }
}

// One step of RungeKutta (3D_solver/src/calc_time_dev.f90.fypp:84-101 for
// RK3, :266-286 for RK4).
void Oracle::rk_step() {
  if (cfg.rk == 3) {
    calc_EFG(QJ);
    step1(1.0, QJ, QJ2);
    set_bc(QJ2);
    calc_EFG(QJ2);
    step2_3(0.75, 0.25, 0.25, 1.0, QJ, QJ2);
    set_bc(QJ2);
    calc_EFG(QJ2);
    step2_3(2.0, 1.0, 2.0, c.one_third, QJ2, QJ);
    set_bc(QJ);
  } else {
    calc_EFG(QJ);
    step_rk4(0.5, 1.0, QJ, QJ2);
    set_bc(QJ2);
    calc_EFG(QJ2);
    step_rk4(0.5, 2.0, QJ, QJ2);
    set_bc(QJ2);
    calc_EFG(QJ2);
    step_rk4(1.0, 2.0, QJ, QJ2);
    set_bc(QJ2);
    calc_EFG(QJ2);
    step4(QJ);
    set_bc(QJ);
  }
}

// RHS on the current state: calc_EFG then calc_R with coefficient 1 at
// interior cells.
void Oracle::rhs(std::vector<double>& R) {
  calc_EFG(QJ);
  std::fill(R.begin(), R.end(), 0.0);
  const int h = off;
  for (int k = h; k <= nz - 1 - h; ++k)
    for (int j = h; j <= ny - 1 - h; ++j)
      for (int i = h; i <= nx - 1 - h; ++i) {
        OX_AREA(i, j, k);
        double r[5];
        calc_R(i, j, k, a_xy, a_yz, a_zx, r);
        for (int l = 1; l <= 5; ++l) R[iq(i + 1, l, j + 1, k + 1)] = r[l - 1];
      }
}

// Interior diagnostics, FP64 with Neumaier summation.  Quantities of
// src/print.f90:42-122 (KE and entropy means), plus enstrophy and the CFL
// wave speed.
struct Neumaier {
  double s = 0.0, c = 0.0;
  void add(double x) {
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  double value() const { return s + c; }
};

void Oracle::diagnostics(ox_diag& d) const {
  const int h = off;
  Neumaier ke, ent, ens, mass, en;
  double wmax = 0.0;
  const double vol_inv_dummy = 0.0;
  (void)vol_inv_dummy;
  for (int k = h + 1; k <= nz - h; ++k)
    for (int j = h + 1; j <= ny - h; ++j)
      for (int i = h + 1; i <= nx - h; ++i) {
        const double Jij = Jac[(size_t)(i - 1) + (size_t)nx * (j - 1)];
        const double q1 = QJ[iq(i, 1, j, k)];
        const double rho = Jij * q1;
        const double u = QJ[iq(i, 2, j, k)] / q1;
        const double v = QJ[iq(i, 3, j, k)] / q1;
        const double w = QJ[iq(i, 4, j, k)] / q1;
        const double p = (c.gamma - 1.0) * (Jij * QJ[iq(i, 5, j, k)] - 0.5 * rho * (u * u + v * v + w * w));
        ke.add(0.5 * rho * (u * u + v * v + w * w));
        ent.add(rho * std::log(p * std::pow(rho, -c.gamma)));
        mass.add(q1);
        en.add(QJ[iq(i, 5, j, k)]);
        const double cs = std::sqrt(c.gamma * p / rho);
        wmax = std::max(wmax, std::max(std::fabs(u), std::max(std::fabs(v), std::fabs(w))) + cs);
        auto vel = [&](int ii, int jj, int kk, int l) { return QJ[iq(ii, l, jj, kk)] / QJ[iq(ii, 1, jj, kk)]; };
        const double hx = 0.5 * xix[i - 1], hy = 0.5 * etay[j - 1], hz = 0.5 * zetaz[k - 1];
        const double dwdy = (vel(i, j + 1, k, 4) - vel(i, j - 1, k, 4)) * hy;
        const double dvdz = (vel(i, j, k + 1, 3) - vel(i, j, k - 1, 3)) * hz;
        const double dudz = (vel(i, j, k + 1, 2) - vel(i, j, k - 1, 2)) * hz;
        const double dwdx = (vel(i + 1, j, k, 4) - vel(i - 1, j, k, 4)) * hx;
        const double dvdx = (vel(i + 1, j, k, 3) - vel(i - 1, j, k, 3)) * hx;
        const double dudy = (vel(i, j + 1, k, 2) - vel(i, j - 1, k, 2)) * hy;
        const double o1 = dwdy - dvdz, o2 = dudz - dwdx, o3 = dvdx - dudy;
        ens.add(0.5 * rho * (o1 * o1 + o2 * o2 + o3 * o3));
      }
  const double n = (double)(nx - 2 * h) * (double)(ny - 2 * h) * (double)(nz - 2 * h);
  d.ke = ke.value() / n;
  d.entropy = ent.value() / n;
  d.enstrophy = ens.value() / n;
  d.max_wave_speed = wmax;
  d.mass = mass.value();
  d.energy = en.value();
}

}  // namespace oxo

// ---------------------------------------------------------------------------
// C entry points (oxo_*), mirroring the product's ox_* ABI.
// ---------------------------------------------------------------------------
extern "C" {

struct oxo_handle {
  oxo::Oracle o;
};

static thread_local std::string g_oxo_err;

const char* oxo_last_error(void) { return g_oxo_err.c_str(); }

void* oxo_create(const ox_config* cfg) {
  auto* h = new oxo_handle();
  if (!h->o.init(*cfg)) {
    g_oxo_err = h->o.err;
    delete h;
    return nullptr;
  }
  return h;
}

void oxo_destroy(void* h) { delete static_cast<oxo_handle*>(h); }

int oxo_set_state(void* hv, const double* Q) {  // pre_calc: QJ = Q / J
  auto& o = static_cast<oxo_handle*>(hv)->o;
  for (int k = 1; k <= o.nz; ++k)
    for (int j = 1; j <= o.ny; ++j)
      for (int l = 1; l <= 5; ++l)
        for (int i = 1; i <= o.nx; ++i) {
          const size_t m = o.iq(i, l, j, k);
          o.QJ[m] = Q[m] / o.Jac[(size_t)(i - 1) + (size_t)o.nx * (j - 1)];
        }
  std::fill(o.Rs.begin(), o.Rs.end(), 0.0);
  return 0;
}

int oxo_get_state(void* hv, double* Q) {  // main.f90:82-93: Q = J * QJ
  auto& o = static_cast<oxo_handle*>(hv)->o;
  for (int k = 1; k <= o.nz; ++k)
    for (int j = 1; j <= o.ny; ++j)
      for (int l = 1; l <= 5; ++l)
        for (int i = 1; i <= o.nx; ++i) {
          const size_t m = o.iq(i, l, j, k);
          Q[m] = o.Jac[(size_t)(i - 1) + (size_t)o.nx * (j - 1)] * o.QJ[m];
        }
  return 0;
}

int oxo_set_qj(void* hv, const double* QJ) {
  auto& o = static_cast<oxo_handle*>(hv)->o;
  std::memcpy(o.QJ.data(), QJ, o.nstate() * sizeof(double));
  std::fill(o.Rs.begin(), o.Rs.end(), 0.0);
  return 0;
}

int oxo_get_qj(void* hv, double* QJ) {
  auto& o = static_cast<oxo_handle*>(hv)->o;
  std::memcpy(QJ, o.QJ.data(), o.nstate() * sizeof(double));
  return 0;
}

// The case's ghost fill (set_bc) applied to the current state, as after a
// stage (test hook).
int oxo_set_bc(void* hv) {
  auto& o = static_cast<oxo_handle*>(hv)->o;
  o.set_bc(o.QJ);
  return 0;
}

int oxo_step(void* hv, int nsteps) {
  auto& o = static_cast<oxo_handle*>(hv)->o;
  for (int t = 0; t < nsteps; ++t) o.rk_step();
  return 0;
}

int oxo_rhs(void* hv, double* R) {
  auto& o = static_cast<oxo_handle*>(hv)->o;
  std::vector<double> r(o.nstate());
  o.rhs(r);
  std::memcpy(R, r.data(), r.size() * sizeof(double));
  return 0;
}

// Primitive fields of the last RHS evaluation: Qp(nx,5,ny,nz), T, mu, sensor.
int oxo_get_primitives(void* hv, double* Qp, double* T, double* mu, double* sensor) {
  auto& o = static_cast<oxo_handle*>(hv)->o;
  if (Qp) std::memcpy(Qp, o.Qp.data(), o.nstate() * sizeof(double));
  if (T) std::memcpy(T, o.T.data(), o.ncell() * sizeof(double));
  if (mu) std::memcpy(mu, o.mu.data(), o.ncell() * sizeof(double));
  if (sensor) std::memcpy(sensor, o.sensor.data(), o.ncell() * sizeof(double));
  return 0;
}

// SGS fields of the last RHS evaluation of an LES case: mut = rho * nu_t and
// qc2 (calc_les.f90:96-125, after set_bc_mut_common), (nx,ny,nz) each.
int oxo_get_les(void* hv, double* mut, double* qc2) {
  auto& o = static_cast<oxo_handle*>(hv)->o;
  if (mut) std::memcpy(mut, o.mut.data(), o.ncell() * sizeof(double));
  if (qc2) std::memcpy(qc2, o.qc2.data(), o.ncell() * sizeof(double));
  return 0;
}

int oxo_diagnostics(void* hv, ox_diag* d) {
  static_cast<oxo_handle*>(hv)->o.diagnostics(*d);
  return 0;
}

// Unit-level hooks for the closed-form tests.
void oxo_keep_flux(const ox_config* cfg, int order, const double* st /*[6][6]*/, int dir, double* out) {
  oxo::Oracle o;
  o.c.init(*cfg);
  o.cfg = *cfg;
  o.cfg.scheme = OX_SCHEME_KEEP;
  o.io = order / 2 - 1;
  double s[6][6];
  std::memcpy(s, st, sizeof(s));
  oxo::conv_face(o, dir, s, 0.0, out);
}

void oxo_slau_flux(const ox_config* cfg, int order, const double* st /*[6][6]*/, int dir, double fsens,
                   double* out) {
  oxo::Oracle o;
  o.c.init(*cfg);
  o.cfg = *cfg;
  o.cfg.scheme = OX_SCHEME_SLAU;
  o.io = order / 2 - 1;
  double s[6][6];
  std::memcpy(s, st, sizeof(s));
  oxo::conv_face(o, dir, s, fsens, out);
}

void oxo_conv_flux(const ox_config* cfg, int order, const double* st /*[6][6]*/, int dir, double fsens,
                   double* out) {
  oxo::Oracle o;
  o.c.init(*cfg);
  o.cfg = *cfg;
  o.io = order / 2 - 1;
  double s[6][6];
  std::memcpy(s, st, sizeof(s));
  oxo::conv_face(o, dir, s, fsens, out);
}

void oxo_muscl(const ox_config* cfg, int order, int tvd, double sensor, const double* a, double* al, double* ar) {
  oxo::Consts c;
  c.init(*cfg);
  if (order == 4) oxo::delta4(c, tvd, sensor, a, *al, *ar);
  else oxo::delta6(c, tvd, sensor, a, *al, *ar);
}

double oxo_wiggle(const double* phi) { return oxo::wiggle_detector(phi); }

}  // extern "C"
