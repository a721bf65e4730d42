// ox_solver.hpp -- the handle behind the C ABI's ox_solver (one device-resident
// case) and the helpers the C-ABI units share (ox_capi.hip, ox_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ox_core.hpp"

// Per-kernel event profiling (ox_set_profiling / ox_kernel_times).
struct ProfState {
  std::vector<hipEvent_t> pool;    // events recorded this ox_step call
  std::vector<const char*> names;  // name of the kernel that ends at pool[m] (m >= 1)
  size_t used = 0;
  std::vector<std::string> acc_names;
  std::vector<double> acc_ms;
  std::vector<long long> acc_n;
};

struct ox_solver {
  ox::Core C;
  ProfState prof;
  double* QJ = nullptr;   // state Q^n
  double* QJ2 = nullptr;  // QJ2 (RK3) / QJs (RK4)
  double* Rs = nullptr;   // RK4 accumulator
  double* scratch = nullptr;
  size_t scratch_bytes = 0;
  ox::StageFn stage = nullptr;
  hipStream_t own_stream = nullptr;  // the stream the handle created (C.stream may be borrowed)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
};

// Records the thread's last error message and returns `code`.
int ox_fail(int code, const std::string& msg);

#define OX_HIP(call)                                                                                     \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return ox_fail(OX_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline size_t state_elems(const ox::Geom& g) { return (size_t)g.nx * 5 * g.ny * g.nz; }
inline size_t cell_elems(const ox::Geom& g) { return (size_t)g.nx * g.ny * g.nz; }

void prof_record(ox::Core* C, const char* name);

// Stage `st` (0-based) of one RK step of handle s: source state, epilogue and
// the factor on dt * face area.
void rk_stage(ox_solver* s, int st, const double** Qsrc, ox::StageArgs* sa, double* coef);

// Device bytes a handle with this configuration allocates.
size_t solver_footprint(const ox_config& c);
