// ox_diag.hpp -- device reductions: interior diagnostics and the state check.
//
// Reductions are two-level and deterministic: each thread accumulates a
// grid-stride range with Neumaier compensation, each block merges its
// threads' (sum, compensation) pairs in LDS in a fixed tree order and writes
// one partial; the host finishes over the partials in block order.  The
// result therefore does not depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

#include "ox_types.hpp"

namespace ox {

struct Comp {
  double s, c;
};

__device__ __forceinline__ void comp_add(Comp& a, double x) {
  const double t = a.s + x;
  if (__builtin_fabs(a.s) >= __builtin_fabs(x)) a.c += (a.s - t) + x;
  else a.c += (x - t) + a.s;
  a.s = t;
}
__device__ __forceinline__ Comp comp_merge(Comp a, Comp b) {
  Comp r = a;
  comp_add(r, b.s);
  r.c += b.c;
  return r;
}

// max that propagates NaN (fmax drops it): a state with a NaN or negative
// pressure cell reports a NaN wave speed instead of a finite one a caller
// would size dt from.  Once either operand is NaN the result stays NaN.
__host__ __device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

constexpr int DIAG_BLOCK = 256;
constexpr int DIAG_NQ = 5;  // ke, entropy, enstrophy, mass, energy (sums) + max wave speed separately

// QJ(nx,5,ny,nz): interior cells [h, h+n) per axis.
__global__ __launch_bounds__(DIAG_BLOCK) void k_diag(const double* __restrict__ QJ, Geom g, double gamma,
                                                     double* __restrict__ partial) {
  __shared__ double ss[DIAG_NQ][DIAG_BLOCK], sc[DIAG_NQ][DIAG_BLOCK], sm[DIAG_BLOCK];
  Comp acc[DIAG_NQ];
  for (int q = 0; q < DIAG_NQ; ++q) acc[q] = Comp{0.0, 0.0};
  double wmax = 0.0;
  const long long n = (long long)g.ix * g.iy * g.iz;
  const double hx = 0.5 * g.inv_dx, hy = 0.5 * g.inv_dy, hz = 0.5 * g.inv_dz;
  for (long long t = (long long)blockIdx.x * DIAG_BLOCK + threadIdx.x; t < n; t += (long long)gridDim.x * DIAG_BLOCK) {
    const int i = g.h + (int)(t % g.ix);
    const int j = g.h + (int)((t / g.ix) % g.iy);
    const int k = g.h + (int)(t / ((long long)g.ix * g.iy));
    auto Q = [&](int ii, int l, int jj, int kk) { return QJ[ii + (long long)g.nx * (l + 5 * (jj + g.ny * kk))]; };
    auto vel = [&](int ii, int jj, int kk, int l) { return Q(ii, l, jj, kk) / Q(ii, 0, jj, kk); };
    const double q1 = Q(i, 0, j, k);
    const double rho = g.jac * q1;
    const double u = Q(i, 1, j, k) / q1, v = Q(i, 2, j, k) / q1, w = Q(i, 3, j, k) / q1;
    const double p = (gamma - 1.0) * (g.jac * Q(i, 4, j, k) - 0.5 * rho * (u * u + v * v + w * w));
    comp_add(acc[0], 0.5 * rho * (u * u + v * v + w * w));
    comp_add(acc[1], rho * log(p * pow(rho, -gamma)));
    const double dwdy = (vel(i, j + 1, k, 3) - vel(i, j - 1, k, 3)) * hy;
    const double dvdz = (vel(i, j, k + 1, 2) - vel(i, j, k - 1, 2)) * hz;
    const double dudz = (vel(i, j, k + 1, 1) - vel(i, j, k - 1, 1)) * hz;
    const double dwdx = (vel(i + 1, j, k, 3) - vel(i - 1, j, k, 3)) * hx;
    const double dvdx = (vel(i + 1, j, k, 2) - vel(i - 1, j, k, 2)) * hx;
    const double dudy = (vel(i, j + 1, k, 1) - vel(i, j - 1, k, 1)) * hy;
    const double o1 = dwdy - dvdz, o2 = dudz - dwdx, o3 = dvdx - dudy;
    comp_add(acc[2], 0.5 * rho * (o1 * o1 + o2 * o2 + o3 * o3));
    comp_add(acc[3], q1);
    comp_add(acc[4], Q(i, 4, j, k));
    const double cs = sqrt(gamma * p / rho);
    const double um = fmax(fabs(u), fmax(fabs(v), fabs(w))) + cs;
    wmax = nan_max(wmax, um);
  }
  const int tid = threadIdx.x;
  for (int q = 0; q < DIAG_NQ; ++q) {
    ss[q][tid] = acc[q].s;
    sc[q][tid] = acc[q].c;
  }
  sm[tid] = wmax;
  __syncthreads();
  for (int off = DIAG_BLOCK / 2; off > 0; off >>= 1) {
    if (tid < off) {
      for (int q = 0; q < DIAG_NQ; ++q) {
        const Comp m = comp_merge(Comp{ss[q][tid], sc[q][tid]}, Comp{ss[q][tid + off], sc[q][tid + off]});
        ss[q][tid] = m.s;
        sc[q][tid] = m.c;
      }
      sm[tid] = nan_max(sm[tid], sm[tid + off]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    double* out = partial + (size_t)blockIdx.x * (2 * DIAG_NQ + 1);
    for (int q = 0; q < DIAG_NQ; ++q) {
      out[2 * q] = ss[q][0];
      out[2 * q + 1] = sc[q][0];
    }
    out[2 * DIAG_NQ] = sm[0];
  }
}

// Non-physical cell scan: per block, the count of bad interior cells and the
// smallest bad linear cell index (x fastest over nx,ny,nz), -1 if none.
__global__ __launch_bounds__(DIAG_BLOCK) void k_check(const double* __restrict__ QJ, Geom g, double gamma,
                                                      long long* __restrict__ partial) {
  __shared__ long long scount[DIAG_BLOCK], sfirst[DIAG_BLOCK];
  long long count = 0, first = -1;
  const long long n = (long long)g.ix * g.iy * g.iz;
  for (long long t = (long long)blockIdx.x * DIAG_BLOCK + threadIdx.x; t < n; t += (long long)gridDim.x * DIAG_BLOCK) {
    const int i = g.h + (int)(t % g.ix);
    const int j = g.h + (int)((t / g.ix) % g.iy);
    const int k = g.h + (int)(t / ((long long)g.ix * g.iy));
    const double* q = QJ + i + (long long)g.nx * 5 * (j + (long long)g.ny * k);
    const double q1 = q[0];
    const double rho = g.jac * q1;
    const double u = q[g.nx] / q1, v = q[2 * g.nx] / q1, w = q[3 * g.nx] / q1;
    const double p = (gamma - 1.0) * (g.jac * q[4 * g.nx] - 0.5 * rho * (u * u + v * v + w * w));
    const bool bad = !(rho > 0.0) || !(p > 0.0) || !isfinite(rho) || !isfinite(p) || !isfinite(u) || !isfinite(v) ||
                     !isfinite(w);
    if (bad) {
      const long long lin = i + (long long)g.nx * (j + (long long)g.ny * k);
      if (first < 0 || lin < first) first = lin;
      ++count;
    }
  }
  const int tid = threadIdx.x;
  scount[tid] = count;
  sfirst[tid] = first;
  __syncthreads();
  for (int off = DIAG_BLOCK / 2; off > 0; off >>= 1) {
    if (tid < off) {
      scount[tid] += scount[tid + off];
      const long long a = sfirst[tid], b = sfirst[tid + off];
      sfirst[tid] = (a < 0) ? b : ((b < 0) ? a : (a < b ? a : b));
    }
    __syncthreads();
  }
  if (tid == 0) {
    partial[2 * blockIdx.x] = scount[0];
    partial[2 * blockIdx.x + 1] = sfirst[0];
  }
}

}  // namespace ox
