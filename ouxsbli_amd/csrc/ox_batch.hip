// ox_batch.hip -- the batch engine of include/ouxsbli.h (ox_batch_*) and the
// segment-length autotuner (ox_autotune).
//
// Every admitted case is an ox_solver whose kernels run on the engine's
// stream.  Cases with the same kernel specialisation, precision and RK scheme
// form a group; one RK stage of a group is three persistent launches (x, y, z
// sweeps; one fused launch for 2-D groups) over the concatenated tiles of its
// cases, each case's tiles reading that case's descriptor.  A 3-D case's
// descriptors depend on its primitive-set parity, which flips every stage,
// so they repeat with period two steps (one for RK4): the engine builds the
// descriptors of the next two steps once per set of cases, keeps both sets
// on the device and captures one HIP graph per set.  Steps then alternate
// between the two graphs with no host work in between, and ox_batch_run
// enqueues many steps back to back before it synchronizes.
// A case's arithmetic is the same code on the same data as its single-case
// launches, so its result does not depend on what else is in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "ox2d_core.hpp"
#include "ox_core.hpp"
#include "ox_kernels.hpp"
#include "ox_solver.hpp"

extern "C" {
int ox_create(const ox_config* cfg, ox_solver** out);
void ox_destroy(ox_solver* s);
int ox_set_state(ox_solver* s, const double* Q);
int ox_get_state(ox_solver* s, double* Q);
int ox_autotune(ox_solver* s, int32_t* lseg_y, int32_t* lseg_z);
int ox2d_create(const ox2d_config* cfg, ox2d_solver** out);
void ox2d_destroy(ox2d_solver* s);
int ox2d_set_state(ox2d_solver* s, const double* Q);
int ox2d_get_state(ox2d_solver* s, double* Q);
}

namespace {

constexpr int kMaxGrid = 1024;  // persistent blocks per launch (4 per CU on 256 CUs)
constexpr size_t kMaxArgs = 2048;
static_assert(sizeof(ox::SweepArgs<double, double>) <= kMaxArgs, "sweep descriptor size");
static_assert(sizeof(ox::SweepArgs<float, float>) <= kMaxArgs, "sweep descriptor size");

// A resident case: a 3-D solver (s) or a 2-D one (s2).
struct BCase {
  int64_t id;
  ox_solver* s = nullptr;
  ox2d_solver* s2 = nullptr;
  int remaining;
  size_t bytes;
  int group;  // index into ox_batch::groups while active

  int dim() const { return s2 ? 2 : 3; }
  double cells() const {
    if (s2) return (double)(s2->cfg.nx - 2) * (s2->cfg.ny - 2);
    const ox::Geom& g = s->C.g;
    return (double)g.ix * g.iy * g.iz;
  }
  int stages() const { return s2 ? 3 : s->C.cfg.rk; }
};

// Cases sharing a kernel specialisation.  3-D groups launch three sweeps per
// stage (LES groups first the SGS viscosity and its two ghost-copy passes);
// 2-D groups one fused stage kernel (launch2[stage]).
struct BGroup {
  int dim, prec, key, rk;
  bool les;
  ox::BatchOps ops;
  ox2::Launch2 launch2[3];
  std::vector<BCase*> members;
};

unsigned cdivu(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

struct ox_batch {
  ox_batch_config cfg{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<std::unique_ptr<BCase>> active;  // admission order
  std::deque<std::unique_ptr<BCase>> done;
  std::vector<BGroup> groups;
  int64_t next_id = 1;
  uint64_t bytes_in_use = 0;
  // per-step launch plan (rebuilt when the set of active cases changes)
  bool plan_valid = false;
  std::vector<ox::BatchGroupLaunch> launches;  // [stage-major: for each stage, for each group]
  std::vector<int> launch_stage;
  std::vector<int> launch_grp;
  std::vector<size_t> args_off;                // per launch: byte offset of its x,y,z descriptor arrays
  // boundary fills of the cases with non-periodic directions, per descriptor
  // set, in launch order: after launch `li`, the fill of one case
  struct BcFill {
    size_t li;
    ox::Geom g;
    void* pn[8];  // the primitive set that launch's epilogue writes
  };
  std::vector<BcFill> bc_fills[2];
  size_t args_bytes = 0;     // one descriptor set (all launches of one step)
  char* d_args = nullptr;    // two sets: set k at d_args + k * args_bytes
  size_t d_args_cap = 0;
  char* h_args = nullptr;    // pinned staging of the two sets
  size_t h_args_cap = 0;
  ox::TileMap* d_map = nullptr;
  size_t d_map_cap = 0;
  int phase = 0;             // the descriptor set the next step reads
  // captured launches of one step per descriptor set for the current cases
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // autotune cache: (precision, stage key, nx, ny, nz) -> (lseg_y, lseg_z)
  std::map<std::tuple<int, int, int, int, int>, std::pair<int, int>> tuned;
  // statistics
  int64_t steps = 0, graph_replays = 0;
  int32_t graph_captures = 0;
  double cell_stages = 0.0, device_ms = 0.0;
};

namespace {

void drop_graphs(ox_batch* b) {
  for (hipGraphExec_t& g : b->graph) {
    if (g) hipGraphExecDestroy(g);
    g = nullptr;
  }
}

void invalidate(ox_batch* b) {
  b->plan_valid = false;
  drop_graphs(b);
}

int ensure_device(char** p, size_t* cap, size_t need) {
  if (need <= *cap) return OX_OK;
  if (*p) hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, need) != hipSuccess) return ox_fail(OX_ENOMEM, "batch descriptor allocation failed");
  *cap = need;
  return OX_OK;
}

// Groups and the per-step launch plan for the current active cases: tile
// maps (uploaded here), descriptor offsets, persistent grid sizes.  The
// descriptors of 2-D groups do not change from step to step and are written
// here; those of 3-D groups are refreshed before every step.
int build_plan(ox_batch* b) {
  b->groups.clear();
  for (auto& up : b->active) {
    BCase* c = up.get();
    const int dim = c->dim();
    int prec, key, rk;
    if (dim == 3) {
      const ox::Core& C = c->s->C;
      prec = C.cfg.precision;
      key = ox::stage_key(C.cfg.order, C.cfg.scheme, C.cfg.visc, C.vorder);
      rk = C.cfg.rk;
    } else {
      const ox2d_config& k2 = c->s2->cfg;
      prec = k2.precision;
      key = ox2::key2(k2.order, k2.scheme);
      rk = 3;
    }
    int gi = -1;
    for (size_t g = 0; g < b->groups.size(); ++g) {
      const BGroup& G = b->groups[g];
      if (G.dim == dim && G.prec == prec && G.key == key && G.rk == rk) gi = (int)g;
    }
    if (gi < 0) {
      BGroup G{};
      G.dim = dim;
      G.prec = prec;
      G.key = key;
      G.rk = rk;
      G.les = dim == 3 && c->s->C.cfg.visc == OX_VISC_LES;
      if (dim == 3) {
        const ox::Core& C = c->s->C;
        const bool ok = prec == OX_PREC_FP64 ? ox::find_batch_fp64(C.cfg.order, C.cfg.scheme, C.cfg.visc, C.vorder, &G.ops)
                                             : ox::find_batch_fp32(C.cfg.order, C.cfg.scheme, C.cfg.visc, C.vorder, &G.ops);
        if (!ok) return ox_fail(OX_EUNSUPPORTED, "no batched specialisation for a case");
      } else {
        for (int st = 0; st < 3; ++st) G.launch2[st] = c->s2->launch[st];
      }
      b->groups.push_back(G);
      gi = (int)b->groups.size() - 1;
    }
    c->group = gi;
    b->groups[gi].members.push_back(c);
  }
  // tile maps: per 3-D group, per direction, one TileMap per member (LES
  // groups: three more, for the SGS viscosity and the two ghost-copy
  // passes); 2-D groups count their tiles (the case lookup reads the
  // descriptors)
  std::vector<ox::TileMap> maps;
  std::vector<size_t> map_off;  // per group: index of its x map (y at +n, z at +2n, LES at +3n..+5n)
  std::vector<std::array<int, 3>> ntiles, ntiles_les;
  for (BGroup& G : b->groups) {
    map_off.push_back(maps.size());
    std::array<int, 3> nt{0, 0, 0};
    std::array<int, 3> nl{0, 0, 0};
    if (G.dim == 2) {
      for (BCase* c : G.members) nt[0] += c->s2->ntiles;
      ntiles.push_back(nt);
      ntiles_les.push_back(nl);
      continue;
    }
    for (int d = 0; d < 3; ++d) {
      for (BCase* c : G.members) {
        const ox::Geom& g = c->s->C.g;
        const int* ls = c->s->C.lseg;
        ox::TileMap m{};
        unsigned nbx, nby, nbz;
        if (d == 0) {
          nbx = cdivu(g.ix, G.ops.xt); nby = cdivu(g.iy, ox::XROWS); nbz = g.iz;
        } else if (d == 1) {
          nbx = cdivu(g.ix, ox::LINE_X); nby = cdivu(g.iz, ox::LINE_Y); nbz = cdivu(g.iy, ls[0]);
        } else {
          nbx = cdivu(g.ix, ox::LINE_X); nby = cdivu(g.iy, ox::LINE_Y); nbz = cdivu(g.iz, ls[1]);
        }
        m.start = nt[d];
        m.nbx = (int)nbx;
        m.nby = (int)nby;
        m.nbz = (int)nbz;
        nt[d] += (int)(nbx * nby * nbz);
        maps.push_back(m);
      }
    }
    ntiles.push_back(nt);
    if (G.les) {
      for (int d = 0; d < 3; ++d) {
        for (BCase* c : G.members) {
          const ox::Geom& g = c->s->C.g;
          ox::TileMap m{};
          if (d == 0) {  // 64 x 4 cells of one plane, as k_les_mut's blocks
            m.nbx = (int)cdivu(g.nx - 2, 64);
            m.nby = (int)cdivu(g.ny - 2, 4);
            m.nbz = g.nz - 2;
          } else {  // 256 items of a ghost-copy pass
            m.nbx = (int)cdivu(ox::les_bc_items(g, d - 1), 256);
            m.nby = 1;
            m.nbz = 1;
          }
          m.start = nl[d];
          nl[d] += m.nbx * m.nby * m.nbz;
          maps.push_back(m);
        }
      }
    }
    ntiles_les.push_back(nl);
  }
  if (!maps.empty()) {
    int rc = ensure_device((char**)&b->d_map, &b->d_map_cap, maps.size() * sizeof(ox::TileMap));
    if (rc != OX_OK) return rc;
    OX_HIP(hipMemcpyAsync(b->d_map, maps.data(), maps.size() * sizeof(ox::TileMap), hipMemcpyHostToDevice, b->stream));
    OX_HIP(hipStreamSynchronize(b->stream));
  }
  // launches: for each stage (up to 4), for each group with that many stages
  b->launches.clear();
  b->launch_stage.clear();
  b->launch_grp.clear();
  b->args_off.clear();
  size_t off = 0;
  for (int st = 0; st < 4; ++st) {
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
      const BGroup& G = b->groups[gi];
      if (st >= G.rk) continue;
      const int n = (int)G.members.size();
      ox::BatchGroupLaunch L{};
      L.stream = b->stream;
      L.ncase = n;
      if (G.dim == 3) {
        for (int d = 0; d < 3; ++d) {
          L.map[d] = b->d_map + map_off[gi] + (size_t)d * n;
          L.ntiles[d] = ntiles[gi][d];
          L.grid[d] = std::max(1, std::min(ntiles[gi][d], kMaxGrid));
          if (G.les) {
            L.map_les[d] = b->d_map + map_off[gi] + (size_t)(3 + d) * n;
            L.ntiles_les[d] = ntiles_les[gi][d];
            L.grid_les[d] = std::max(1, std::min(ntiles_les[gi][d], kMaxGrid));
          }
        }
      } else {
        L.ntiles[0] = ntiles[gi][0];
        L.grid[0] = ox2::grid2(ntiles[gi][0], G.launch2[st].max_blocks());
      }
      b->launches.push_back(L);
      b->launch_stage.push_back(st);
      b->launch_grp.push_back((int)gi);
      b->args_off.push_back(off);
      off += G.dim == 3 ? 3 * (size_t)n * G.ops.args_bytes : (size_t)n * ox2::args2_bytes(G.prec);
      off = (off + 255) & ~(size_t)255;
    }
  }
  b->args_bytes = std::max<size_t>(off, 256);
  int rc = ensure_device(&b->d_args, &b->d_args_cap, 2 * b->args_bytes);
  if (rc != OX_OK) return rc;
  if (2 * b->args_bytes > b->h_args_cap) {
    if (b->h_args) hipHostFree(b->h_args);
    b->h_args = nullptr;
    b->h_args_cap = 0;
    if (hipHostMalloc((void**)&b->h_args, 2 * b->args_bytes, hipHostMallocDefault) != hipSuccess)
      return ox_fail(OX_ENOMEM, "pinned staging allocation failed");
    b->h_args_cap = 2 * b->args_bytes;
  }
  for (size_t li = 0; li < b->launches.size(); ++li) {
    const BGroup& G = b->groups[b->launch_grp[li]];
    const size_t n = G.members.size();
    if (G.dim == 2) {  // the same descriptors in both sets
      const size_t ab = ox2::args2_bytes(G.prec);
      for (int set = 0; set < 2; ++set) {
        int tile0 = 0;
        for (size_t m = 0; m < n; ++m) {
          ox2d_solver* s2 = G.members[m]->s2;
          ox2::write_args2(s2->cfg, s2->g, s2->buf, b->launch_stage[li], tile0,
                           b->h_args + set * b->args_bytes + b->args_off[li] + m * ab);
          tile0 += s2->ntiles;
        }
      }
    }
  }
  // descriptors of the 3-D groups for the next two steps: each case's
  // primitive-set parity advances stage by stage exactly as in ox_step, then
  // returns to where it is now
  std::vector<std::pair<int, const double*>> saved;
  for (auto& up : b->active)
    if (up->s) saved.push_back({up->s->C.cur, up->s->C.prim_for});
  auto restore = [&]() {
    size_t k = 0;
    for (auto& up : b->active)
      if (up->s) {
        up->s->C.cur = saved[k].first;
        up->s->C.prim_for = saved[k].second;
        ++k;
      }
  };
  for (int set = 0; set < 2; ++set) {
    b->bc_fills[set].clear();
    for (size_t li = 0; li < b->launches.size(); ++li) {
      BGroup& G = b->groups[b->launch_grp[li]];
      if (G.dim == 2) continue;
      const int st = b->launch_stage[li];
      const size_t n = G.members.size(), ab = G.ops.args_bytes;
      char* base = b->h_args + set * b->args_bytes + b->args_off[li];
      for (size_t m = 0; m < n; ++m) {
        ox_solver* s = G.members[m]->s;
        const ox::Geom& g = s->C.g;
        const double* Qsrc;
        ox::StageArgs sa;
        double coef;
        rk_stage(s, st, &Qsrc, &sa, &coef);
        if (s->C.prim_for != Qsrc) {
          restore();
          invalidate(b);
          return ox_fail(OX_ESTATE, "batched case without current primitives");
        }
        alignas(16) char tmp[3 * kMaxArgs];
        G.ops.make_args(s->C, sa, coef * g.a_yz, coef * g.a_zx, coef * g.a_xy, tmp);
        for (int d = 0; d < 3; ++d) std::memcpy(base + ((size_t)d * n + m) * ab, tmp + (size_t)d * ab, ab);
        b->launches[li].ep = sa.ep;
        if ((g.bc[0] || g.bc[1] || g.bc[2]) && sa.ep != ox::EP_RHS) {
          ox_batch::BcFill f;
          f.li = li;
          f.g = g;
          for (int m2 = 0; m2 < 8; ++m2) f.pn[m2] = s->C.prim[1 - s->C.cur][m2];
          b->bc_fills[set].push_back(f);
        }
        ox::finish_stage(s->C, sa);
      }
    }
  }
  restore();
  OX_HIP(hipMemcpyAsync(b->d_args, b->h_args, 2 * b->args_bytes, hipMemcpyHostToDevice, b->stream));
  OX_HIP(hipStreamSynchronize(b->stream));
  b->phase = 0;
  b->plan_valid = true;
  return OX_OK;
}

// Advances every 3-D case's host-side stage bookkeeping (primitive-set
// parity) by `steps` RK steps, matching what the device did.
void advance_parity(ox_batch* b, int steps) {
  for (auto& up : b->active) {
    ox_solver* s = up->s;
    if (!s) continue;
    for (int k = 0; k < steps; ++k)
      for (int st = 0; st < s->C.cfg.rk; ++st) {
        const double* Qsrc;
        ox::StageArgs sa;
        double coef;
        rk_stage(s, st, &Qsrc, &sa, &coef);
        ox::finish_stage(s->C, sa);
      }
  }
}

// Segment lengths of a case from the engine's cache, tuning on a miss.
int tune_case(ox_batch* b, ox_solver* s) {
  const ox_config& cfg = s->C.cfg;
  const auto key = std::make_tuple((int)cfg.precision, ox::stage_key(cfg.order, cfg.scheme, cfg.visc, s->C.vorder),
                                   (int)cfg.nx, (int)cfg.ny, (int)cfg.nz);
  auto it = b->tuned.find(key);
  if (it != b->tuned.end()) {
    s->C.lseg[0] = ox::clamp_lseg(s->C.g, 1, it->second.first);
    s->C.lseg[1] = ox::clamp_lseg(s->C.g, 2, it->second.second);
    return OX_OK;
  }
  int32_t ly, lz;
  const int rc = ox_autotune(s, &ly, &lz);  // leaves the state unchanged
  if (rc == OX_OK) b->tuned[key] = {ly, lz};
  return rc;
}

// The launches of one step reading descriptor set `set`.
void enqueue_step(ox_batch* b, int set) {
  size_t fi = 0;
  const std::vector<ox_batch::BcFill>& fills = b->bc_fills[set];
  for (size_t li = 0; li < b->launches.size(); ++li) {
    const BGroup& G = b->groups[b->launch_grp[li]];
    ox::BatchGroupLaunch L = b->launches[li];
    char* base = b->d_args + set * b->args_bytes + b->args_off[li];
    if (G.dim == 3) {
      const size_t n = G.members.size();
      for (int d = 0; d < 3; ++d) L.args[d] = base + (size_t)d * n * G.ops.args_bytes;
      G.ops.launch(L);
      // the boundary fills of this stage's non-periodic cases
      for (; fi < fills.size() && fills[fi].li == li; ++fi) G.ops.bc_after(fills[fi].g, fills[fi].pn, L.stream);
    } else {
      G.launch2[b->launch_stage[li]].launch(base, L.ncase, L.ntiles[0], L.grid[0], L.stream);
    }
  }
}

// Returns a case's solver to its own stream and frees it.
void free_case(BCase* c) {
  if (c->s2) {
    c->s2->stream = c->s2->own_stream;
    ox2d_destroy(c->s2);
  } else {
    c->s->C.stream = c->s->own_stream;
    ox_destroy(c->s);
  }
}

}  // namespace

// One 3-D case joins the batch; `init` gives its new handle the initial
// state (a host upload or a device initial condition).
namespace {
template <class Init>
int admit3d(ox_batch* b, const ox_config* cfg, Init init, int32_t nsteps, int64_t* id) {
  if (nsteps < 1) return ox_fail(OX_EINVAL, "nsteps must be >= 1");
  if (cfg->device != b->cfg.device) return ox_fail(OX_EINVAL, "case device differs from the batch device");
  const size_t need = solver_footprint(*cfg);
  if (need > b->cfg.hbm_budget)
    return ox_fail(OX_EINVAL, "case footprint " + std::to_string(need) + " B exceeds the whole HBM budget " +
                                  std::to_string(b->cfg.hbm_budget) + " B: it can never be admitted");
  if ((int)(b->active.size() + b->done.size()) >= b->cfg.max_cases) return ox_fail(OX_EBUSY, "no free slot");
  if (b->bytes_in_use + need > b->cfg.hbm_budget) return ox_fail(OX_EBUSY, "footprint exceeds the free HBM budget");
  ox_solver* s = nullptr;
  int rc = ox_create(cfg, &s);
  if (rc != OX_OK) return rc;
  rc = init(s);
  if (rc == OX_OK && b->cfg.autotune) rc = tune_case(b, s);
  if (rc != OX_OK) {
    ox_destroy(s);
    return rc;
  }
  // primitives of the initial state on the handle's own stream, then the
  // case joins the engine's stream
  {
    ox::BatchOps ops;
    const bool ok = cfg->precision == OX_PREC_FP64
                        ? ox::find_batch_fp64(cfg->order, cfg->scheme, cfg->visc, s->C.vorder, &ops)
                        : ox::find_batch_fp32(cfg->order, cfg->scheme, cfg->visc, s->C.vorder, &ops);
    if (!ok) {
      ox_destroy(s);
      return ox_fail(OX_EUNSUPPORTED, "no batched specialisation for this case");
    }
    if (s->C.prim_for != s->QJ) ops.prepare(s->C, s->QJ);  // k_prim
    if (hipStreamSynchronize(s->C.stream) != hipSuccess) {
      ox_destroy(s);
      return ox_fail(OX_EDEVICE, "case initialisation failed");
    }
  }
  s->C.stream = b->stream;
  s->C.prof = false;
  auto c = std::make_unique<BCase>();
  c->id = b->next_id++;
  c->s = s;
  c->remaining = nsteps;
  c->bytes = need;
  c->group = -1;
  b->bytes_in_use += need;
  *id = c->id;
  b->active.push_back(std::move(c));
  invalidate(b);
  return OX_OK;
}
}  // namespace

extern "C" {

int ox_footprint(const ox_config* cfg, uint64_t* bytes) {
  if (!cfg || !bytes) return ox_fail(OX_EINVAL, "null argument");
  *bytes = solver_footprint(*cfg);
  return OX_OK;
}

int ox_autotune(ox_solver* s, int32_t* lseg_y, int32_t* lseg_z) {
  if (!s) return ox_fail(OX_EINVAL, "null handle");
  OX_HIP(hipSetDevice(s->C.cfg.device));
  ox::Core& C = s->C;
  const ox::Geom& g = C.g;
  const bool prof = C.prof;
  C.prof = false;
  ox::StageArgs sa{s->QJ, s->QJ2, s->Rs, 0, 0, 0, ox::EP_RHS};
  hipEvent_t e0, e1;
  OX_HIP(hipEventCreate(&e0));
  OX_HIP(hipEventCreate(&e1));
  auto time_rhs = [&](int ly, int lz) -> float {
    C.lseg[0] = ox::clamp_lseg(g, 1, ly);
    C.lseg[1] = ox::clamp_lseg(g, 2, lz);
    s->stage(C, s->QJ, sa, g.a_yz, g.a_zx, g.a_xy);  // warm-up
    hipEventRecord(e0, C.stream);
    for (int r = 0; r < 3; ++r) s->stage(C, s->QJ, sa, g.a_yz, g.a_zx, g.a_xy);
    hipEventRecord(e1, C.stream);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    return ms;
  };
  static const int cand[] = {16, 32, 64, 128};
  int by = ox::LSEG_DEFAULT, bz = ox::LSEG_DEFAULT;
  float best = 1e30f;
  for (int ly : cand) {
    if (ly > 2 * g.iy && ly != cand[0]) continue;
    const float t = time_rhs(ly, bz);
    if (t < best) best = t, by = ly;
  }
  best = 1e30f;
  for (int lz : cand) {
    if (lz > 2 * g.iz && lz != cand[0]) continue;
    const float t = time_rhs(by, lz);
    if (t < best) best = t, bz = lz;
  }
  C.lseg[0] = by = ox::clamp_lseg(g, 1, by);
  C.lseg[1] = bz = ox::clamp_lseg(g, 2, bz);
  C.prof = prof;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  OX_HIP(hipGetLastError());
  if (lseg_y) *lseg_y = by;
  if (lseg_z) *lseg_z = bz;
  return OX_OK;
}

int ox_set_segments(ox_solver* s, int32_t lseg_y, int32_t lseg_z) {
  if (!s) return ox_fail(OX_EINVAL, "null handle");
  if (lseg_y < 1 || lseg_y > 4096 || lseg_z < 1 || lseg_z > 4096)
    return ox_fail(OX_EINVAL, "segment lengths must be in 1..4096");
  s->C.lseg[0] = ox::clamp_lseg(s->C.g, 1, lseg_y);
  s->C.lseg[1] = ox::clamp_lseg(s->C.g, 2, lseg_z);
  return OX_OK;
}

int ox_batch_create(const ox_batch_config* cfg, ox_batch** out) {
  if (!cfg || !out) return ox_fail(OX_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->abi_version != OX_ABI_VERSION) return ox_fail(OX_EINVAL, "abi_version mismatch");
  if (cfg->max_cases < 1) return ox_fail(OX_EINVAL, "max_cases must be >= 1");
  ox_batch* b = new (std::nothrow) ox_batch();
  if (!b) return ox_fail(OX_ENOMEM, "host allocation failed");
  b->cfg = *cfg;
  if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    ox_batch_destroy(b);
    return ox_fail(OX_EDEVICE, "stream/event creation failed");
  }
  *out = b;
  return OX_OK;
}

void ox_batch_destroy(ox_batch* b) {
  if (!b) return;
  hipSetDevice(b->cfg.device);
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (auto& c : b->active) free_case(c.get());
  for (auto& c : b->done) free_case(c.get());
  if (b->d_args) hipFree(b->d_args);
  if (b->d_map) hipFree(b->d_map);
  if (b->h_args) hipHostFree(b->h_args);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

int ox_batch_admit(ox_batch* b, const ox_config* cfg, const double* Q, int32_t nsteps, int64_t* id) {
  if (!b || !cfg || !Q || !id) return ox_fail(OX_EINVAL, "null argument");
  return admit3d(b, cfg, [Q](ox_solver* s) { return ox_set_state(s, Q); }, nsteps, id);
}

int ox_batch_admit_ic(ox_batch* b, const ox_config* cfg, const ox_ic* ic, int32_t nsteps, int64_t* id) {
  if (!b || !cfg || !ic || !id) return ox_fail(OX_EINVAL, "null argument");
  return admit3d(b, cfg, [ic](ox_solver* s) { return ox_init_state(s, ic); }, nsteps, id);
}

int ox_batch_tune(ox_batch* b, const ox_config* cfg, const double* Q) {
  if (!b || !cfg || !Q) return ox_fail(OX_EINVAL, "null argument");
  ox_solver* s = nullptr;
  int rc = ox_create(cfg, &s);
  if (rc != OX_OK) return rc;
  rc = ox_set_state(s, Q);
  if (rc == OX_OK) rc = tune_case(b, s);
  ox_destroy(s);
  return rc;
}

int ox_batch_admit2d(ox_batch* b, const ox2d_config* cfg, const double* Q, int32_t nsteps, int64_t* id) {
  if (!b || !cfg || !id) return ox_fail(OX_EINVAL, "null argument");
  if (nsteps < 1) return ox_fail(OX_EINVAL, "nsteps must be >= 1");
  if (cfg->device != b->cfg.device) return ox_fail(OX_EINVAL, "case device differs from the batch device");
  const size_t need = solver2d_footprint(*cfg);
  if (need > b->cfg.hbm_budget)
    return ox_fail(OX_EINVAL, "case footprint " + std::to_string(need) + " B exceeds the whole HBM budget " +
                                  std::to_string(b->cfg.hbm_budget) + " B: it can never be admitted");
  if ((int)(b->active.size() + b->done.size()) >= b->cfg.max_cases) return ox_fail(OX_EBUSY, "no free slot");
  if (b->bytes_in_use + need > b->cfg.hbm_budget) return ox_fail(OX_EBUSY, "footprint exceeds the free HBM budget");
  ox2d_solver* s = nullptr;
  int rc = ox2d_create(cfg, &s);
  if (rc != OX_OK) return rc;
  rc = Q ? ox2d_set_state(s, Q) : ox2d_init_state(s);
  if (rc != OX_OK) {
    ox2d_destroy(s);
    return rc;
  }
  s->stream = b->stream;
  auto c = std::make_unique<BCase>();
  c->id = b->next_id++;
  c->s2 = s;
  c->remaining = nsteps;
  c->bytes = need;
  c->group = -1;
  b->bytes_in_use += need;
  *id = c->id;
  b->active.push_back(std::move(c));
  invalidate(b);
  return OX_OK;
}

int ox_batch_run(ox_batch* b, int32_t max_steps, int32_t* steps_done, int32_t* active) {
  if (!b) return ox_fail(OX_EINVAL, "null handle");
  if (max_steps < 1) return ox_fail(OX_EINVAL, "max_steps must be >= 1");
  OX_HIP(hipSetDevice(b->cfg.device));
  if (steps_done) *steps_done = 0;
  if (b->active.empty()) {
    if (active) *active = 0;
    return OX_OK;
  }
  if (!b->plan_valid) {
    int rc = build_plan(b);
    if (rc != OX_OK) return rc;
  }
  // steps until the first case finishes (the set of cases changes there)
  int k = max_steps;
  for (auto& c : b->active) k = std::min(k, c->remaining);
  if (b->cfg.use_graph) {
    for (int set = 0; set < 2; ++set) {
      if (b->graph[set]) continue;
      // capture the step's launches once per descriptor set for this set of cases
      hipGraph_t graph = nullptr;
      OX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      enqueue_step(b, set);
      OX_HIP(hipStreamEndCapture(b->stream, &graph));
      const hipError_t e = hipGraphInstantiate(&b->graph[set], graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e != hipSuccess) {
        b->graph[set] = nullptr;
        return ox_fail(OX_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      ++b->graph_captures;
    }
  }
  OX_HIP(hipEventRecord(b->ev0, b->stream));
  for (int t = 0; t < k; ++t) {
    const int set = (b->phase + t) & 1;
    if (b->cfg.use_graph) {
      OX_HIP(hipGraphLaunch(b->graph[set], b->stream));
      ++b->graph_replays;
    } else {
      enqueue_step(b, set);
    }
  }
  OX_HIP(hipGetLastError());
  OX_HIP(hipEventRecord(b->ev1, b->stream));
  OX_HIP(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  OX_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->device_ms += ms;
  b->steps += k;
  b->phase = (b->phase + k) & 1;
  advance_parity(b, k);
  if (steps_done) *steps_done = k;
  // bookkeeping: stage counts, finished cases
  bool changed = false;
  for (size_t a = 0; a < b->active.size();) {
    BCase* c = b->active[a].get();
    b->cell_stages += c->cells() * c->stages() * k;
    c->remaining -= k;
    if (c->remaining == 0) {
      // the result is read with its ghost layers (2-D: written by the last stage)
      if (c->s) ox::launch_ghost_fill(c->s->C, c->s->QJ);
      b->done.push_back(std::move(b->active[a]));
      b->active.erase(b->active.begin() + a);
      changed = true;
    } else {
      ++a;
    }
  }
  if (changed) {
    OX_HIP(hipStreamSynchronize(b->stream));
    invalidate(b);
  }
  if (active) *active = (int32_t)b->active.size();
  return OX_OK;
}

int ox_batch_step(ox_batch* b, int32_t* active) { return ox_batch_run(b, 1, nullptr, active); }

int ox_batch_done(ox_batch* b, int64_t* ids, int32_t cap, int32_t* n) {
  if (!b || !n) return ox_fail(OX_EINVAL, "null argument");
  int32_t k = 0;
  for (auto& c : b->done) {
    if (k < cap && ids) ids[k] = c->id;
    ++k;
  }
  *n = std::min(k, cap);
  return OX_OK;
}

int ox_batch_retire(ox_batch* b, int64_t id, double* Q) {
  if (!b) return ox_fail(OX_EINVAL, "null handle");
  for (auto& c : b->active)
    if (c->id == id) return ox_fail(OX_ESTATE, "case is still stepping");
  for (auto it = b->done.begin(); it != b->done.end(); ++it) {
    if ((*it)->id != id) continue;
    BCase* c = it->get();
    int rc = OX_OK;
    if (Q) rc = c->s2 ? ox2d_get_state(c->s2, Q) : ox_get_state(c->s, Q);
    free_case(c);
    b->bytes_in_use -= c->bytes;
    b->done.erase(it);
    return rc;
  }
  return ox_fail(OX_ENOTFOUND, "unknown case id");
}

int ox_batch_query(ox_batch* b, ox_batch_stats* out) {
  if (!b || !out) return ox_fail(OX_EINVAL, "null argument");
  std::memset(out, 0, sizeof(*out));
  out->active = (int32_t)b->active.size();
  out->done = (int32_t)b->done.size();
  int ng = 0;
  {
    std::vector<std::tuple<int, int, int, int>> seen;
    for (auto& c : b->active) {
      std::tuple<int, int, int, int> t;
      if (c->s2) {
        t = std::make_tuple(2, (int)c->s2->cfg.precision, ox2::key2(c->s2->cfg.order, c->s2->cfg.scheme), 3);
      } else {
        const ox::Core& C = c->s->C;
        t = std::make_tuple(3, (int)C.cfg.precision, ox::stage_key(C.cfg.order, C.cfg.scheme, C.cfg.visc, C.vorder),
                            (int)C.cfg.rk);
      }
      if (std::find(seen.begin(), seen.end(), t) == seen.end()) seen.push_back(t);
    }
    ng = (int)seen.size();
  }
  out->groups = ng;
  out->graph_captures = b->graph_captures;
  out->steps = b->steps;
  out->graph_replays = b->graph_replays;
  out->bytes_in_use = b->bytes_in_use;
  out->cell_stages = b->cell_stages;
  out->device_ms = b->device_ms;
  return OX_OK;
}

}  // extern "C"
