// ox_ic.hip -- initial states built on the device: ox_init_state (3-D),
// ox2d_init_state (2-D oblique shock).
//
// The reference samples the initial field at the interior cell centres on the
// host (3D_solver/NSTGV/set.f90:16-47, ETGV/set.f90:27-40, KHI/set.f90:31-47,
// 2D_solver/OS/set.f90:28-46), fills the
// periodic ghosts (3D_solver/src/set_bc_common.f90:11-97) and divides by the
// Jacobian in pre_calc (3D_solver/src/preprocess.f90.fypp:65-71).  Here one
// kernel writes QJ = Q / J at every interior cell and k_ghost_fill copies the
// periodic images, so no host array exists and nothing crosses the bus.
//
// Cell centres are formed exactly as the grid setup forms them (node
// dx1 * (i - order/2) with dx1 = L / (n - order), centre = the midpoint of two
// nodes, set_grid_cyclic{2,4,6}_3D in src/set_coordinate.f90:120-339), and the
// field expressions keep the operation order of the host restatement in
// ouxsbli_amd/cases.py (ETGV and KHI: that of the reference's set_init), so
// the two differ only where the device's sin, cos and tanh round differently
// from the host's (a few ulp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "ox2d_core.hpp"
#include "ox_solver.hpp"

namespace ox {

// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the output function applied to a counter, so that
// every cell's noise is a pure function of (seed, field, cell index).
__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform noise in [-1, 1) of field f (0, 1, 2 = u, v, w) at linear cell m of
// the padded (nz, ny, nx) array: 53 random bits, scaled and shifted exactly.
__device__ inline double cell_noise(uint64_t seed, long long m, int f) {
  const uint64_t key = (seed + 1ull) * 0x9E3779B97F4A7C15ull + (uint64_t)(3 * m + f);
  const double u01 = (double)(splitmix64(key) >> 11) * 0x1.0p-53;
  return 2.0 * u01 - 1.0;
}

__device__ inline double centre(int c, int off, double dx1) {
  const double a = dx1 * (double)(c - off);
  const double b = dx1 * (double)(c + 1 - off);
  return 0.5 * (a + b);
}

__global__ __launch_bounds__(256) void k_zero(double* __restrict__ a, long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) a[t] = 0.0;
}

struct IcArgs {
  double dx1, dy1, dz1;  // node spacings L / (n - order)
  double Lx, Ly, Lz, gamma, gm1, R, J;
  ox_ic ic;
};

__global__ __launch_bounds__(256) void k_ic(double* __restrict__ QJ, Geom g, IcArgs a) {
  const int i = g.h + blockIdx.x * 64 + threadIdx.x;
  const int j = g.h + blockIdx.y * 4 + threadIdx.y;
  const int k = g.h + blockIdx.z;
  if (i >= g.h + g.ix || j >= g.h + g.iy) return;
  const ox_ic& ic = a.ic;
  double q[5];
  if (ic.kind == OX_IC_TAYLOR_GREEN) {
    const double X = centre(i, g.h, a.dx1), Y = centre(j, g.h, a.dy1), Z = centre(k, g.h, a.dz1);
    const double L0 = ic.L0, V0 = ic.V0;
    const double p = ic.p0 + ic.rho0 * (V0 * V0) * (cos(2.0 * X / L0) + cos(2.0 * Y / L0)) *
                                 (cos(2.0 * Z / L0) + 2.0) / 16.0;
    const double rho = p / (a.R * ic.T0);
    q[0] = rho;
    q[1] = rho * V0 * sin(X / L0) * cos(Y / L0) * cos(Z / L0);
    q[2] = -rho * V0 * cos(X / L0) * sin(Y / L0) * cos(Z / L0);
    q[3] = 0.0;
    q[4] = p / a.gm1 + 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0];
  } else if (ic.kind == OX_IC_ETGV) {
    // ETGV/set.f90:27-40 (the reference's expressions and operation order)
    const double X = centre(i, g.h, a.dx1), Y = centre(j, g.h, a.dy1), Z = centre(k, g.h, a.dz1);
    const double rho0 = ic.rho0, M0 = ic.V0;
    q[0] = rho0;
    q[1] = rho0 * M0 * sin(X) * cos(Y) * cos(Z);
    q[2] = -(rho0 * M0 * cos(X) * sin(Y) * cos(Z));
    q[3] = 0.0;
    q[4] = (1.0 / a.gamma + 0.0625 * rho0 * (M0 * M0) * (cos(2.0 * X) + cos(2.0 * Y)) * (cos(2.0 * Z) + 2.0)) /
               (a.gamma - 1.0) +
           0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0];
  } else if (ic.kind == OX_IC_KHI) {
    // KHI/set.f90:31-47: the stream test uses Lx, as the reference writes it
    const double X = centre(i, g.h, a.dx1), Y = centre(j, g.h, a.dy1), Z = centre(k, g.h, a.dz1);
    const double pi = 3.141592653589793;  // acos(-1.d0)
    const double u1 = ic.V0, rho1 = ic.rho0;
    const double w = ic.amp * u1 * sin(2.0 * pi * Z / a.Lz);
    const double v = ic.amp * u1 * sin(2.0 * pi * X / a.Lx);
    const bool outer = Y > 0.75 * a.Lx || Y < 0.25 * a.Lx;
    const double rho = outer ? rho1 : ic.rho2, u = outer ? u1 : ic.u2;
    q[0] = rho;
    q[1] = rho * u;
    q[2] = rho * v;
    q[3] = rho * w;
    q[4] = ic.p0 / (a.gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
  } else {  // OX_IC_SHEAR_LAYER
    const double y = centre(j, g.h, a.dy1);
    const double U = ic.V0, d = ic.delta;
    const double prof = 0.5 * U * (tanh((y - 0.25 * a.Ly) / d) - tanh((y - 0.75 * a.Ly) / d) - 1.0);
    const long long m = (long long)i + (long long)g.nx * ((long long)j + (long long)g.ny * k);
    const double u = cell_noise(ic.seed, m, 0) * ic.noise + prof;
    const double v = cell_noise(ic.seed, m, 1) * ic.noise;
    const double w = cell_noise(ic.seed, m, 2) * ic.noise;
    const double rho = ic.rho0;
    q[0] = rho;
    q[1] = rho * u;
    q[2] = rho * v;
    q[3] = rho * w;
    const double ke = u * u + v * v + w * w;
    q[4] = ic.p0 / a.gm1 + 0.5 * rho * ke;
  }
#pragma unroll
  for (int l = 0; l < 5; ++l) QJ[i + (size_t)g.nx * (l + 5 * ((size_t)j + (size_t)g.ny * k))] = q[l] / a.J;
}

// ---------------------------------------------------------------------------
// OX_IC_SPECTRAL: the inverse real transform of a band-limited half spectrum,
// done as three pruned partial sums over the mode box (|k| <= K in every
// direction) instead of full-size FFTs:
//   G1[c][kx][ky][z] = sum_kz H[c][kx][ky][kz] e^{2 pi i kz z / n}
//   G2[c][kx][y][z]  = sum_ky G1[c][kx][ky][z] e^{2 pi i ky y / n}
//   u_c(x, y, z)     = (Re G2[c][0][y][z] + 2 sum_{kx>=1} Re(G2[c][kx][y][z] e^{2 pi i kx x / n})) / n^3
// Twiddles come from a table of e^{2 pi i r / n}, r = k m mod n, so every
// angle is reduced exactly.  Two passes over the cells: the first reduces
// sum |u|^2 (per-block Neumaier partials, merged on the host in block order),
// the second recomputes u, scales it to the requested rms and writes QJ.
// ---------------------------------------------------------------------------
struct Cplx {
  double re, im;
};

__global__ __launch_bounds__(256) void k_spec_twiddle(Cplx* __restrict__ tw, int n) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double sn, cs;
  sincospi(2.0 * (double)r / (double)n, &sn, &cs);
  tw[r] = Cplx{cs, sn};
}

__device__ __forceinline__ int tw_index(int k, int m, int n) {
  const long long r = ((long long)k * m) % n;
  return (int)(r < 0 ? r + n : r);
}

// G1: one thread per (z, ky, kx, c); H is [3][K+1][2K+1][2K+1] (c, kx, ky, kz)
__global__ __launch_bounds__(256) void k_spec_z(const Cplx* __restrict__ H, const Cplx* __restrict__ tw,
                                                Cplx* __restrict__ G1, int K, int n) {
  const int nk = 2 * K + 1;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long total = 3LL * (K + 1) * nk * n;
  if (t >= total) return;
  const int z = (int)(t % n);
  const long long row = t / n;  // (c, kx, ky)
  const Cplx* h = H + row * nk;
  double re = 0.0, im = 0.0;
  for (int a = 0; a < nk; ++a) {
    const Cplx e = tw[tw_index(a - K, z, n)];
    re += h[a].re * e.re - h[a].im * e.im;
    im += h[a].re * e.im + h[a].im * e.re;
  }
  G1[t] = Cplx{re, im};  // [c][kx][ky][z]
}

// G2: one thread per (z, y, kx, c)
__global__ __launch_bounds__(256) void k_spec_y(const Cplx* __restrict__ G1, const Cplx* __restrict__ tw,
                                                Cplx* __restrict__ G2, int K, int n) {
  const int nk = 2 * K + 1;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long total = 3LL * (K + 1) * n * n;
  if (t >= total) return;
  const int z = (int)(t % n);
  const int y = (int)((t / n) % n);
  const long long ck = t / ((long long)n * n);  // (c, kx)
  const Cplx* g = G1 + ck * nk * n;
  double re = 0.0, im = 0.0;
  for (int b = 0; b < nk; ++b) {
    const Cplx v = g[(long long)b * n + z];
    const Cplx e = tw[tw_index(b - K, y, n)];
    re += v.re * e.re - v.im * e.im;
    im += v.re * e.im + v.im * e.re;
  }
  G2[t] = Cplx{re, im};  // [c][kx][y][z]
}

// u, v, w at grid point (x, y, z) from G2 (layout [c][kx][y][z])
__device__ __forceinline__ void spec_velocity(const Cplx* __restrict__ G2, const Cplx* __restrict__ tw, int K, int n,
                                              int x, int y, int z, double inv_n3, double* u) {
  const long long plane = (long long)n * n;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const Cplx* g = G2 + (long long)c * (K + 1) * plane + (long long)y * n + z;
    double acc = g[0].re;
    double s = 0.0;
    for (int kx = 1; kx <= K; ++kx) {
      const Cplx v = g[(long long)kx * plane];
      const Cplx e = tw[tw_index(kx, x, n)];
      s += v.re * e.re - v.im * e.im;
    }
    u[c] = (acc + 2.0 * s) * inv_n3;
  }
}

__global__ __launch_bounds__(256) void k_spec_rms(const Cplx* __restrict__ G2, const Cplx* __restrict__ tw, int K,
                                                  int n, double inv_n3, double* __restrict__ partial) {
  __shared__ double ss[256], sc[256];
  const long long N = (long long)n * n * n;
  double sum = 0.0, comp = 0.0;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < N; t += (long long)gridDim.x * 256) {
    const int x = (int)(t % n), y = (int)((t / n) % n), z = (int)(t / ((long long)n * n));
    double u[3];
    spec_velocity(G2, tw, K, n, x, y, z, inv_n3, u);
    const double q = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double tt = sum + q;
    if (__builtin_fabs(sum) >= __builtin_fabs(q)) comp += (sum - tt) + q;
    else comp += (q - tt) + sum;
    sum = tt;
  }
  const int tid = threadIdx.x;
  ss[tid] = sum;
  sc[tid] = comp;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) {
      const double a = ss[tid], b = ss[tid + off];
      const double tt = a + b;
      double c = sc[tid] + sc[tid + off];
      if (__builtin_fabs(a) >= __builtin_fabs(b)) c += (a - tt) + b;
      else c += (b - tt) + a;
      ss[tid] = tt;
      sc[tid] = c;
    }
    __syncthreads();
  }
  if (tid == 0) {
    partial[2 * blockIdx.x] = ss[0];
    partial[2 * blockIdx.x + 1] = sc[0];
  }
}

__global__ __launch_bounds__(256) void k_spec_state(double* __restrict__ QJ, Geom g, const Cplx* __restrict__ G2,
                                                    const Cplx* __restrict__ tw, int K, int n, double inv_n3,
                                                    double scale, double rho, double pe, double J) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= n || y >= n) return;
  double u[3];
  spec_velocity(G2, tw, K, n, x, y, z, inv_n3, u);
  u[0] *= scale;
  u[1] *= scale;
  u[2] *= scale;
  double q[5];
  q[0] = rho;
  q[1] = rho * u[0];
  q[2] = rho * u[1];
  q[3] = rho * u[2];
  q[4] = pe + 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const int i = g.h + x, j = g.h + y, k = g.h + z;
#pragma unroll
  for (int l = 0; l < 5; ++l) QJ[i + (size_t)g.nx * (l + 5 * ((size_t)j + (size_t)g.ny * k))] = q[l] / J;
}

}  // namespace ox

namespace ox2 {
// The OS state: inflow everywhere, the post-shock state on the top row from
// 0-based column top_split - 1 on; QJ(i, l, j) = q / J.
__global__ __launch_bounds__(256) void k_ic_os(double* __restrict__ QJ, int nx, int ny, int split, double qa0,
                                               double qa1, double qa2, double qa3, double qb0, double qb1, double qb2,
                                               double qb3) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nx * ny) return;
  const int i = (int)(t % nx), j = (int)(t / nx);
  const bool post = j == ny - 1 && i >= split - 1;
  const double q[4] = {post ? qb0 : qa0, post ? qb1 : qa1, post ? qb2 : qa2, post ? qb3 : qa3};
#pragma unroll
  for (int l = 0; l < 4; ++l) QJ[i + (size_t)nx * (l + 4 * (size_t)j)] = q[l];
}
}  // namespace ox2

namespace {
// The fresh state's bookkeeping, as after ox_set_qj: RK4 accumulator zero,
// primitives stale (a fill kernel, not a memset, so the call may be captured).
int finish_init(ox_solver* s) {
  ox::Core& C = s->C;
  const size_t n = state_elems(C.g);
  if (s->Rs) {
    ox::k_zero<<<(unsigned)((n + 255) / 256), 256, 0, C.stream>>>(s->Rs, (long long)n);
    OX_HIP(hipGetLastError());
  }
  C.prim_for = nullptr;
  OX_HIP(hipStreamSynchronize(C.stream));
  return OX_OK;
}

int init_spectral(ox_solver* s, const ox_ic* ic) {
  ox::Core& C = s->C;
  const ox::Geom& g = C.g;
  const ox_config& c = C.cfg;
  const int n = g.ix, K = ic->kmax;
  if (g.iy != n || g.iz != n) return ox_fail(OX_EINVAL, "spectral initial state needs a cubic interior");
  if (!ic->modes || K < 1 || 2 * K + 1 > n) return ox_fail(OX_EINVAL, "spectral modes missing or kmax not < n/2");
  if (!(ic->V0 > 0.0) || !(ic->rho0 > 0.0) || !(ic->p0 > 0.0))
    return ox_fail(OX_EINVAL, "spectral initial state needs V0, rho0, p0 > 0");
  OX_HIP(hipSetDevice(c.device));
  const int nk = 2 * K + 1;
  const size_t nH = 3ull * (K + 1) * nk * nk, nG1 = 3ull * (K + 1) * nk * n, nG2 = 3ull * (K + 1) * n * n;
  const int nb = 1024;
  ox::Cplx *H = nullptr, *tw = nullptr, *G1 = nullptr, *G2 = nullptr;
  double* part = nullptr;
  auto release = [&]() {
    (void)hipFree(H);
    (void)hipFree(tw);
    (void)hipFree(G1);
    (void)hipFree(G2);
    (void)hipFree(part);
  };
  if (hipMalloc(&H, nH * sizeof(ox::Cplx)) != hipSuccess || hipMalloc(&tw, n * sizeof(ox::Cplx)) != hipSuccess ||
      hipMalloc(&G1, nG1 * sizeof(ox::Cplx)) != hipSuccess || hipMalloc(&G2, nG2 * sizeof(ox::Cplx)) != hipSuccess ||
      hipMalloc(&part, 2 * nb * sizeof(double)) != hipSuccess) {
    release();
    return ox_fail(OX_ENOMEM, "spectral initial state: device allocation failed");
  }
  int rc = OX_OK;
  std::vector<double> hp(2 * nb);
  do {
    if (hipMemcpyAsync(H, ic->modes, nH * sizeof(ox::Cplx), hipMemcpyHostToDevice, C.stream) != hipSuccess) {
      rc = ox_fail(OX_EDEVICE, "spectral modes upload failed");
      break;
    }
    ox::k_spec_twiddle<<<(n + 255) / 256, 256, 0, C.stream>>>(tw, n);
    ox::k_spec_z<<<(unsigned)((nG1 + 255) / 256), 256, 0, C.stream>>>(H, tw, G1, K, n);
    ox::k_spec_y<<<(unsigned)((nG2 + 255) / 256), 256, 0, C.stream>>>(G1, tw, G2, K, n);
    const double inv_n3 = 1.0 / ((double)n * (double)n * (double)n);
    ox::k_spec_rms<<<nb, 256, 0, C.stream>>>(G2, tw, K, n, inv_n3, part);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, C.stream) != hipSuccess ||
        hipStreamSynchronize(C.stream) != hipSuccess) {
      rc = ox_fail(OX_EDEVICE, "spectral synthesis failed");
      break;
    }
    double sum = 0.0, comp = 0.0;
    for (int b = 0; b < nb; ++b) {
      const double x = hp[2 * b], t = sum + x;
      if (std::fabs(sum) >= std::fabs(x)) comp += (sum - t) + x;
      else comp += (x - t) + sum;
      sum = t;
      comp += hp[2 * b + 1];
    }
    const double mean = (sum + comp) / ((double)n * (double)n * (double)n);
    if (!(mean > 0.0)) {
      rc = ox_fail(OX_EINVAL, "spectral modes give a zero velocity field");
      break;
    }
    const double scale = ic->V0 / std::sqrt(mean);
    const dim3 grid((n + 63) / 64, (n + 3) / 4, n);
    ox::k_spec_state<<<grid, dim3(64, 4, 1), 0, C.stream>>>(s->QJ, g, G2, tw, K, n, inv_n3, scale, ic->rho0,
                                                             ic->p0 / (c.gamma - 1.0), g.jac);
    if (hipGetLastError() != hipSuccess) {
      rc = ox_fail(OX_EDEVICE, "spectral state kernel failed");
      break;
    }
    ox::launch_ghost_fill(C, s->QJ);
    if (hipStreamSynchronize(C.stream) != hipSuccess) rc = ox_fail(OX_EDEVICE, "spectral initial state failed");
  } while (false);
  release();
  if (rc != OX_OK) return rc;
  return finish_init(s);
}
}  // namespace

extern "C" {

int ox_init_state(ox_solver* s, const ox_ic* ic) {
  if (!s || !ic) return ox_fail(OX_EINVAL, "null argument");
  if (ic->kind == OX_IC_SPECTRAL) return init_spectral(s, ic);
  if (ic->kind != OX_IC_TAYLOR_GREEN && ic->kind != OX_IC_SHEAR_LAYER && ic->kind != OX_IC_ETGV &&
      ic->kind != OX_IC_KHI)
    return ox_fail(OX_EINVAL, "unknown initial condition kind");
  if (ic->kind == OX_IC_TAYLOR_GREEN && (ic->L0 <= 0.0 || ic->T0 <= 0.0))
    return ox_fail(OX_EINVAL, "Taylor-Green needs L0 > 0 and T0 > 0");
  if (ic->kind == OX_IC_SHEAR_LAYER && ic->delta <= 0.0) return ox_fail(OX_EINVAL, "shear layer needs delta > 0");
  OX_HIP(hipSetDevice(s->C.cfg.device));
  ox::Core& C = s->C;
  const ox::Geom& g = C.g;
  const ox_config& c = C.cfg;
  ox::IcArgs a{};
  a.dx1 = c.Lx / (double)(c.nx - c.order);
  a.dy1 = c.Ly / (double)(c.ny - c.order);
  a.dz1 = c.Lz / (double)(c.nz - c.order);
  a.Lx = c.Lx;
  a.Ly = c.Ly;
  a.Lz = c.Lz;
  a.gamma = c.gamma;
  a.gm1 = c.gamma - 1.0;
  a.R = c.R;
  a.J = g.jac;
  a.ic = *ic;
  const dim3 grid((g.ix + 63) / 64, (g.iy + 3) / 4, g.iz);
  ox::k_ic<<<grid, dim3(64, 4, 1), 0, C.stream>>>(s->QJ, g, a);
  OX_HIP(hipGetLastError());
  ox::launch_ghost_fill(C, s->QJ);
  OX_HIP(hipGetLastError());
  return finish_init(s);
}

int ox2d_init_state(ox2d_solver* s) {
  if (!s) return ox_fail(OX_EINVAL, "null handle");
  OX_HIP(hipSetDevice(s->cfg.device));
  const ox2d_config& c = s->cfg;
  // conserved values of OS/set.f90:34-45, then / J as pre_calc does
  double qa[4], qb[4];
  const double* pa = c.inflow;
  const double* pb = c.post;
  const double ig = 1.0 / (c.gamma - 1.0);
  const double va[4] = {pa[0], pa[0] * pa[1], pa[0] * pa[2], pa[3] * ig + 0.5 * pa[0] * (pa[1] * pa[1] + pa[2] * pa[2])};
  const double vb[4] = {pb[0], pb[0] * pb[1], pb[0] * pb[2], pb[3] * ig + 0.5 * pb[0] * (pb[1] * pb[1] + pb[2] * pb[2])};
  for (int l = 0; l < 4; ++l) {
    qa[l] = va[l] / s->g.J;
    qb[l] = vb[l] / s->g.J;
  }
  const long long n = (long long)c.nx * c.ny;
  ox2::k_ic_os<<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(s->buf[0], c.nx, c.ny, c.top_split, qa[0], qa[1],
                                                                   qa[2], qa[3], qb[0], qb[1], qb[2], qb[3]);
  OX_HIP(hipGetLastError());
  OX_HIP(hipStreamSynchronize(s->stream));
  return OX_OK;
}

}  // extern "C"
