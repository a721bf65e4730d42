// ox2d_capi.hip -- the 2-D entry points of include/ouxsbli.h (ox2d_*).
//
// A 2-D case owns three FP64 state buffers (Q^n and the two RK3 stage
// states) and the descriptors of its three stages, uploaded once at
// creation: the pointers and coefficients of a stage never change, so a step
// is three launches of the fused stage kernel with no host work in between.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ox2d_core.hpp"
#include "ox_solver.hpp"

namespace {

int validate2(const ox2d_config& c) {
  if (c.abi_version != OX_ABI_VERSION) return ox_fail(OX_EINVAL, "abi_version mismatch");
  if (c.nx < 8 || c.ny < 8) return ox_fail(OX_EINVAL, "nx and ny must be >= 8");
  if ((long long)c.nx * c.ny > (1LL << 30)) return ox_fail(OX_EINVAL, "grid too large");
  if (c.order != 2 && c.order != 4 && c.order != 6) return ox_fail(OX_EINVAL, "order must be 2, 4 or 6");
  if (c.scheme != OX_SCHEME_SLAU && c.scheme != OX_SCHEME_ROE)
    return ox_fail(OX_EUNSUPPORTED, "2-D cases support the SLAU and Roe schemes");
  if (c.tvd < OX_TVD_NONE || c.tvd > OX_TVD_HYBRID) return ox_fail(OX_EINVAL, "bad tvd");
  if (c.slau_variant != OX_SLAU_SLAU1 && c.slau_variant != OX_SLAU_HRSLAU2)
    return ox_fail(OX_EINVAL, "bad slau_variant");
  if (c.rk != 3) return ox_fail(OX_EUNSUPPORTED, "2-D cases advance with RK3");
  if (c.precision != OX_PREC_FP64 && c.precision != OX_PREC_FP32_BULK) return ox_fail(OX_EINVAL, "bad precision");
  if (c.bc != OX2D_BC_OBLIQUE_SHOCK) return ox_fail(OX_EUNSUPPORTED, "unknown 2-D boundary kind");
  if (c.top_split < 0 || c.top_split > c.nx + 1) return ox_fail(OX_EINVAL, "top_split out of range");
  if (!(c.gamma > 1.0) || !(c.R > 0.0) || !(c.dt > 0.0) || !(c.Lx > 0.0) || !(c.Ly > 0.0))
    return ox_fail(OX_EINVAL, "gamma, R, dt, Lx, Ly must be positive (gamma > 1)");
  for (int l = 0; l < 4; ++l)
    if (!std::isfinite(c.inflow[l]) || !std::isfinite(c.post[l])) return ox_fail(OX_EINVAL, "non-finite boundary state");
  if (!(c.inflow[0] > 0.0) || !(c.inflow[3] > 0.0) || !(c.post[0] > 0.0) || !(c.post[3] > 0.0))
    return ox_fail(OX_EINVAL, "boundary states need rho > 0 and p > 0");
  return OX_OK;
}

size_t state2_elems(const ox2d_config& c) { return (size_t)c.nx * 4 * (size_t)c.ny; }

}  // namespace

size_t solver2d_footprint(const ox2d_config& c) {
  return 3 * state2_elems(c) * sizeof(double) + 3 * ox2::args2_bytes(c.precision);
}

extern "C" {

int ox2d_footprint(const ox2d_config* cfg, uint64_t* bytes) {
  if (!cfg || !bytes) return ox_fail(OX_EINVAL, "null argument");
  *bytes = solver2d_footprint(*cfg);
  return OX_OK;
}

void ox2d_destroy(ox2d_solver* s) {
  if (!s) return;
  hipSetDevice(s->cfg.device);
  if (s->stream) hipStreamSynchronize(s->stream);
  for (double* b : s->buf)
    if (b) hipFree(b);
  if (s->d_args) hipFree(s->d_args);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->own_stream) hipStreamDestroy(s->own_stream);
  delete s;
}

int ox2d_create(const ox2d_config* cfg, ox2d_solver** out) {
  if (!cfg || !out) return ox_fail(OX_EINVAL, "null argument");
  *out = nullptr;
  int rc = validate2(*cfg);
  if (rc != OX_OK) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ox_fail(OX_EDEVICE, "no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev) return ox_fail(OX_EINVAL, "device ordinal out of range");
  ox2d_solver* s = new (std::nothrow) ox2d_solver();
  if (!s) return ox_fail(OX_ENOMEM, "host allocation failed");
  s->cfg = *cfg;
  const bool ok = cfg->precision == OX_PREC_FP64 ? ox2::find_launch2_fp64(cfg->order, cfg->scheme, s->launch)
                                                 : ox2::find_launch2_fp32(cfg->order, cfg->scheme, s->launch);
  if (!ok) {
    delete s;
    return ox_fail(OX_EUNSUPPORTED, "no 2-D specialisation for this order and scheme");
  }
  s->g.init(*cfg, ox2::tile_rows_of(cfg->precision));
  s->ntiles = s->g.nbx * s->g.nby;
  auto fail = [&](int code, const char* msg) {
    ox2d_destroy(s);
    return ox_fail(code, msg);
  };
  if (hipSetDevice(cfg->device) != hipSuccess) return fail(OX_EDEVICE, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return fail(OX_EDEVICE, "stream/event creation failed");
  s->stream = s->own_stream;
  const size_t bytes = state2_elems(*cfg) * sizeof(double);
  for (double*& b : s->buf) {
    if (hipMalloc(&b, bytes) != hipSuccess) return fail(OX_ENOMEM, "device allocation failed");
    if (hipMemsetAsync(b, 0, bytes, s->stream) != hipSuccess) return fail(OX_EDEVICE, "hipMemsetAsync failed");
  }
  s->args_bytes = ox2::args2_bytes(cfg->precision);
  if (hipMalloc(&s->d_args, 3 * s->args_bytes) != hipSuccess) return fail(OX_ENOMEM, "device allocation failed");
  std::vector<char> h(3 * s->args_bytes);
  for (int st = 0; st < 3; ++st) ox2::write_args2(*cfg, s->g, s->buf, st, 0, h.data() + st * s->args_bytes);
  if (hipMemcpyAsync(s->d_args, h.data(), h.size(), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return fail(OX_EDEVICE, "descriptor upload failed");
  int cap = s->launch[0].max_blocks();
  for (int st = 1; st < 3; ++st) cap = std::min(cap, s->launch[st].max_blocks());
  s->grid = ox2::grid2(s->ntiles, cap);
  *out = s;
  return OX_OK;
}

int ox2d_set_state(ox2d_solver* s, const double* Q) {
  if (!s || !Q) return ox_fail(OX_EINVAL, "null argument");
  OX_HIP(hipSetDevice(s->cfg.device));
  const size_t n = state2_elems(s->cfg);
  std::vector<double> qj(n);
  const double J = s->g.J;
  for (size_t t = 0; t < n; ++t) qj[t] = Q[t] / J;  // pre_calc: QJ = Q / J
  OX_HIP(hipMemcpyAsync(s->buf[0], qj.data(), n * sizeof(double), hipMemcpyHostToDevice, s->stream));
  OX_HIP(hipStreamSynchronize(s->stream));
  return OX_OK;
}

int ox2d_get_state(ox2d_solver* s, double* Q) {
  if (!s || !Q) return ox_fail(OX_EINVAL, "null argument");
  OX_HIP(hipSetDevice(s->cfg.device));
  const size_t n = state2_elems(s->cfg);
  OX_HIP(hipMemcpyAsync(Q, s->buf[0], n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  OX_HIP(hipStreamSynchronize(s->stream));
  const double J = s->g.J;
  for (size_t t = 0; t < n; ++t) Q[t] = J * Q[t];  // main.f90:84: Q = J * QJ
  return OX_OK;
}

int ox2d_step(ox2d_solver* s, int32_t nsteps) {
  if (!s) return ox_fail(OX_EINVAL, "null handle");
  if (nsteps < 0) return ox_fail(OX_EINVAL, "nsteps < 0");
  OX_HIP(hipSetDevice(s->cfg.device));
  OX_HIP(hipEventRecord(s->ev0, s->stream));
  for (int t = 0; t < nsteps; ++t)
    for (int st = 0; st < 3; ++st) s->launch[st].launch(s->d_args + st * s->args_bytes, 1, s->ntiles, s->grid, s->stream);
  OX_HIP(hipGetLastError());
  OX_HIP(hipEventRecord(s->ev1, s->stream));
  OX_HIP(hipEventSynchronize(s->ev1));
  float ms = 0.f;
  OX_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  s->last_ms = ms;
  return OX_OK;
}

int ox2d_last_step_ms(ox2d_solver* s, double* ms) {
  if (!s || !ms) return ox_fail(OX_EINVAL, "null argument");
  *ms = s->last_ms;
  return OX_OK;
}

}  // extern "C"
