// opcount.cpp -- algorithmic FP operation count per cell and RK stage.
//
// Runs the library's own face functions (ox_physics.hpp, host-callable) on an
// operation-counting scalar.  Because those functions evaluate exactly the
// reference expressions (KEEP, MUSCL, SLAU, central viscous), the count is the
// reference algorithm's count per face.  Counting rule: add, sub, mul, div and
// sqrt are one operation each; negation, abs, copysign, min/max selects and
// comparisons are free.
//
// Per cell and stage (each face shared by two cells, so three faces per cell):
//   primitive recovery + (SLAU) Ducros sensor
//   + 3 x (convective face + viscous face)
//   + the 9 distinct velocity cross derivatives of the viscous terms
//   + calc_R (3 x 5 x (sub, mul, add) minus the first add) + the RK combination.
// 2-D oblique-shock path (ox2d_kernels.hpp): per cell and stage, primitive
// recovery + Ducros sensor + 2 x (MUSCL on 4 primitives + face flux + face
// sensor average) + calc_R (2 x 4 x (sub, mul) + 4 adds) + the RK combination.
// Output: one JSON object per configuration on stdout.
// Build: g++ -std=c++17 -O1 -I/opt/rocm/include -D__HIP_PLATFORM_AMD__ -x c++ opcount.cpp
#include <cstdio>
#include <cstring>

struct Cnt {
  double v;
  Cnt() : v(0) {}
  Cnt(double x) : v(x) {}
};
static long long g_ops = 0;
static inline Cnt operator+(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v + b.v); }
static inline Cnt operator-(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v - b.v); }
static inline Cnt operator*(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v * b.v); }
static inline Cnt operator/(Cnt a, Cnt b) { ++g_ops; return Cnt(a.v / b.v); }
static inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
static inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
static inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
static inline Cnt fsqrt(Cnt a) { ++g_ops; return Cnt(__builtin_sqrt(a.v)); }

#include "../csrc/ox_physics.hpp"

namespace ox {
template <> __host__ __device__ inline Cnt fsign1<Cnt>(Cnt x) { return Cnt(__builtin_copysign(1.0, x.v)); }
template <> __host__ __device__ inline Cnt fabs_<Cnt>(Cnt x) { return Cnt(__builtin_fabs(x.v)); }
}  // namespace ox

using namespace ox;

static PhysC<Cnt> consts(double threshold) {
  ox_config cfg{};
  cfg.gamma = 1.4; cfg.R = 287.03; cfg.Pr = 0.71; cfg.Prt = 0.9; cfg.threshold = threshold;
  HostConsts h;
  h.init(cfg);
  return make_physc<Cnt>(h);
}

// A smooth (monotone) stencil and an oscillating one.
static void fill(Cnt* a, int n, double base, double slope, bool wiggly) {
  for (int m = 0; m < n; ++m) a[m] = Cnt(base + slope * m + (wiggly ? ((m & 1) ? 0.3 : -0.3) * slope : 0.0));
}

template <int N> static long long conv_ops(int scheme, int tvd, int variant, double sensor) {
  PhysC<Cnt> c = consts(0.4);
  Cnt rho[N], u[N], v[N], w[N], p[N], T[N], F[5];
  fill(rho, N, 1.0, 0.01, tvd != 0);
  fill(u, N, 0.3, 0.02, false);
  fill(v, N, -0.2, 0.01, false);
  fill(w, N, 0.1, -0.01, false);
  fill(p, N, 1.0e5, 1.0e2, tvd != 0);
  fill(T, N, 300.0, 1.0, false);
  g_ops = 0;
  if (scheme == 0) {
    if (N == 2) keep2<Cnt, 0>(c, rho, u, v, w, p, T, F);
    else if (N == 4) keep4<Cnt, 0>(c, rho, u, v, w, p, T, F);
    else keep6<Cnt, 0>(c, rho, u, v, w, p, T, F);
  } else {
    Cnt fs = N > 2 ? Cnt(0.5) * (Cnt(sensor) + Cnt(sensor)) : Cnt(0.0);
    slau_face<Cnt, 0, N>(c, tvd, variant, rho, u, v, w, p, fs, F);
  }
  return g_ops;
}

template <int VO> static long long visc_ops() {
  PhysC<Cnt> c = consts(0.4);
  Cnt mu[VO], T[VO], u[VO], v[VO], w[VO], a[VO], b[VO], d[VO], e[VO], tau[3], en;
  fill(mu, VO, 1e-5, 1e-7, false); fill(T, VO, 300, 1, false); fill(u, VO, 1, .1, false);
  fill(v, VO, 1, .1, false); fill(w, VO, 1, .1, false); fill(a, VO, 1, .1, false);
  fill(b, VO, 1, .1, false); fill(d, VO, 1, .1, false); fill(e, VO, 1, .1, false);
  g_ops = 0;
  visc_face<Cnt, 0, VO>(c, Cnt(10.0), mu, T, u, v, w, a, b, d, e, tau, en);
  return g_ops + 4;  // + the 4 subtractions from the convective flux
}

template <int VO> static long long cross_ops() {
  PhysC<Cnt> c = consts(0.4);
  Cnt f[VO + 1];
  fill(f, VO + 1, 1.0, 0.1, false);
  g_ops = 0;
  cross_d<Cnt, VO>(c, Cnt(10.0), f);
  return g_ops;
}

static void report(const char* name, int order, int scheme, int visc, int vo, int rk, int tvd, int variant) {
  long long conv_smooth = 0, conv_tvd = 0;
  if (order == 2) conv_smooth = conv_tvd = conv_ops<2>(scheme, 0, variant, 0.0);
  if (order == 4) { conv_smooth = conv_ops<4>(scheme, 0, variant, 0.0); conv_tvd = conv_ops<4>(scheme, scheme ? (tvd ? 1 : 0) : 0, variant, 1.0); }
  if (order == 6) { conv_smooth = conv_ops<6>(scheme, 0, variant, 0.0); conv_tvd = conv_ops<6>(scheme, scheme ? (tvd ? 1 : 0) : 0, variant, 1.0); }
  if (scheme == 1 && tvd == 1) conv_smooth = conv_tvd;  // pure TVD: always limited
  long long vf = 0, cd = 0;
  if (visc) { vf = vo == 4 ? visc_ops<4>() : visc_ops<6>(); cd = vo == 4 ? cross_ops<4>() : cross_ops<6>(); }
  const long long prim = 6 /*rho,u,v,w: 1 div + 4 mul*/ + 7 /*p*/ + 2 /*T*/ + (visc ? 5 : 0) /*mu*/;
  const long long sensor = (scheme == 1 && order > 2) ? (9 * 2 + 3 /*div*/ + 3 /*rot*/ + 9 /*fd*/ + 3 * 2 /*dxt*/) : 0;
  const long long face_avg = (scheme == 1 && order > 2) ? 2 : 0;  // 0.5*(s_i + s_i+1) per face
  const long long calcR = 3 * 5 * 3 - 5;
  const long long rk_upd = rk == 3 ? 5 * 4 : 5 * 4;  // (c1*Qin + c2*Qout - R)*c4inv ~ 4 ops; RK4 ~ 2+2
  const long long base = prim + sensor + calcR + rk_upd + (visc ? 9 * cd : 0);
  const long long smooth = base + 3 * (conv_smooth + face_avg + vf);
  const long long limited = base + 3 * (conv_tvd + face_avg + vf);
  std::printf("{\"config\": \"%s\", \"order\": %d, \"scheme\": %d, \"visc\": %d, \"visc_order\": %d, \"rk\": %d, "
              "\"conv_face_smooth\": %lld, \"conv_face_limited\": %lld, \"visc_face\": %lld, \"cross_deriv\": %lld, "
              "\"per_cell_stage_smooth\": %lld, \"per_cell_stage_limited\": %lld}\n",
              name, order, scheme, visc, vo, rk, conv_smooth, conv_tvd, vf, cd, smooth, limited);
}

// One x face of the 2-D path: MUSCL (6-point) on rho, u, v, p, then the
// flux of the 3-D routine with w = 0 (ox2d_kernels.hpp, flux2).
static long long face2d_ops(int scheme, bool limited) {
  PhysC<Cnt> c = consts(0.4);
  Cnt st[4][6], L[4], Rr[4], F5[5];
  fill(st[0], 6, 1.0, 0.01, limited);
  fill(st[1], 6, 300.0, 2.0, false);
  fill(st[2], 6, -20.0, 1.0, false);
  fill(st[3], 6, 1.0e5, 1.0e2, limited);
  g_ops = 0;
  for (int v = 0; v < 4; ++v) {
    if (limited) delta6_form<Cnt, true>(c, st[v], L[v], Rr[v]);
    else delta6_form<Cnt, false>(c, st[v], L[v], Rr[v]);
  }
  if (scheme == 2) {
    roe_local<Cnt>(c, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], Cnt(0.0), Cnt(0.0), L[3], Rr[3], F5);
  } else {
    const Cnt hr = wiggle<Cnt>(&st[3][1]);
    slau<Cnt, 0>(c, 1, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], Cnt(0.0), Cnt(0.0), L[3], Rr[3], hr, F5);
  }
  return g_ops;
}

static void report2d(const char* name, int scheme) {
  const long long smooth = face2d_ops(scheme, false), limited = face2d_ops(scheme, true);
  const long long prim = 1 /*1/q0*/ + 1 /*rho*/ + 2 /*u, v*/ + 8 /*p*/;
  const long long sensor = 4 * 2 /*differences x metric*/ + 2 /*div, rot*/ + 6 /*fd*/;
  const long long face_avg = 2;
  const long long calcR = 4 * 5;
  const long long rk_upd = 4 * 4;
  const long long base = prim + sensor + calcR + rk_upd;
  std::printf("{\"config\": \"%s\", \"order\": 6, \"scheme\": %d, \"visc\": 0, \"visc_order\": 0, \"rk\": 3, "
              "\"dim\": 2, \"conv_face_smooth\": %lld, \"conv_face_limited\": %lld, \"per_cell_stage_smooth\": %lld, "
              "\"per_cell_stage_limited\": %lld}\n",
              name, scheme, smooth, limited, base + 2 * (smooth + face_avg), base + 2 * (limited + face_avg));
}

int main() {
  report("keep4_ns4_rk4", 4, 0, 1, 4, 4, 0, 1);
  report("slau6_hrslau2_hybrid_ns6_rk3", 6, 1, 1, 6, 3, 2, 1);
  report("keep6_ns6_rk3", 6, 0, 1, 6, 3, 0, 1);
  report("slau6_euler_rk3", 6, 1, 0, 0, 3, 2, 1);
  report2d("roe6_hybrid_2d_rk3", 2);
  report2d("slau6_hrslau2_hybrid_2d_rk3", 1);
  return 0;
}
