// C ABI of libhns: the SimContext lifecycle and step API that replaces the reference's
// extern "C" HNS_* functions (src/Cuda/HNanoSolver.cu:27-515), the standalone SOP entry points
// (src/Cuda/Advection.cu:169-175, src/Cuda/PressureProjection.cu:127-135), and a device-array
// API over the individual kernels. Errors never cross the boundary as exceptions: each entry
// point returns an hns_status and leaves a message for hns_last_error().
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "hns_adjoint.hpp"
#include "hns_internal.hpp"
#include "hns_mg.hpp"
#include "hns_precond.hpp"
#include "hns_solvers.hpp"

using namespace hns;

struct hns_topology {
  std::shared_ptr<TopoStore> s;
  int64_t builds = 1, reuses = 0;
  // solver scratch of the device-API pressure solves on this topology, made on first use and
  // reused (calls sharing a topology must not overlap)
  mutable std::unique_ptr<Solvers> sol;
  Solvers& solvers() const {
    if (!sol) {
      sol = std::make_unique<Solvers>();
      sol->init(s->L, s->N);
    }
    return *sol;
  }
};
struct hns_precond {
  int device = 0;
  LearnedPrecond net;
  PcWork work;
};
struct hns_diff_tape {
  std::unique_ptr<DiffTape> t;
};
struct hns_payload {
  Payload* p = nullptr;
  ~hns_payload() { destroy_payload(p); }
};

namespace {

thread_local std::string g_err;

hns_status set_err(hns_status c, const std::string& m) {
  g_err = m;
  return c;
}

template <class F>
hns_status guard(F&& f) {
  try {
    const hns_status st = f();
    if (st == HNS_OK) g_err.clear();
    return st;
  } catch (const Error& e) {
    return set_err(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return set_err(HNS_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return set_err(HNS_ERR_INVALID_ARGUMENT, e.what());
  }
}

void require_device() {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) fail(HNS_ERR_NO_DEVICE, "no HIP device is visible");
}

bool starts_with_src(const char* n) { return n && std::strncmp(n, "src_", 4) == 0; }

const hns_value_block* find_block(const hns_grid_data* d, const char* name, int32_t type) {
  for (size_t i = 0; i < d->num_blocks; ++i) {
    const hns_value_block& b = d->blocks[i];
    if (b.name && b.type == type && std::strcmp(b.name, name) == 0) return &b;
  }
  return nullptr;
}
// velocity = the Vec3f block named "velocity", else the first Vec3f block (HNanoSolver.cu:99-110)
const hns_value_block* velocity_block(const hns_grid_data* d) {
  const hns_value_block* first = nullptr;
  for (size_t i = 0; i < d->num_blocks; ++i) {
    const hns_value_block& b = d->blocks[i];
    if (b.type != HNS_VEC3F) continue;
    if (!first) first = &b;
    if (b.name && std::strcmp(b.name, "velocity") == 0) return &b;
  }
  return first;
}
void check_payload(const hns_grid_data* d) {
  if (!d) fail(HNS_ERR_INVALID_ARGUMENT, "payload is null");
  for (size_t i = 0; i < d->num_blocks; ++i) {
    const hns_value_block& b = d->blocks[i];
    if (!b.name) fail(HNS_ERR_INVALID_ARGUMENT, "a value block has no name");
    if (b.type != HNS_FLOAT && b.type != HNS_VEC3F) fail(HNS_ERR_INVALID_ARGUMENT, std::string("block '") + b.name + "' has an unknown type");
  }
}

// A topology built from the same leaf origins (and voxel size, on the current device) as this
// payload: the first coordinate of every leaf compared on the host.
bool same_leaves(const TopoStore& T, const hns_grid_data* d, float voxel_size) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev != T.device) return false;
  if (T.N != d->size || T.voxelSize != voxel_size || !d->coords) return false;
  if (T.host_origins.size() != 3 * (size_t)T.L) return false;
  for (int32_t l = 0; l < T.L; ++l)
    if (std::memcmp(&T.host_origins[3 * (size_t)l], d->coords + 3 * (size_t)l * kLeafVox, 3 * sizeof(int32_t)))
      return false;
  return true;
}

struct Field {
  std::string name;
  DBuf<float> buf;
  bool exists = false;
};

}  // namespace

struct hns_context {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  size_t N = 0;
  int32_t L = 0;
  float voxelSize = 0.f;
  std::shared_ptr<TopoStore> topo;

  DBuf<float> vel, velA, velB, div, p;
  std::vector<Field> state;
  std::map<std::string, DBuf<float>> src;
  std::map<std::string, bool> srcActive;
  DBuf<float> srcVel;
  bool haveSrcVel = false;
  DBuf<float> sdf;
  bool haveSdf = false;
  std::vector<DBuf<float>> advTmp;
  Solvers solvers;
  Timing timing;
  hns_precond* pc = nullptr;  // not owned

  ~hns_context() {
    if (own) (void)hipStreamDestroy(own);
  }
  Field* field(const std::string& n) {
    for (auto& f : state)
      if (f.name == n) return &f;
    return nullptr;
  }
  const float* source(const char* n) const {
    auto it = src.find(n);
    auto a = srcActive.find(n);
    return (it != src.end() && a != srcActive.end() && a->second) ? it->second.p : nullptr;
  }
  DevTopo T() const { return topo->dev(); }
};

namespace {

void upload_velocity_soa(hns_context* C, float* dst, const float* host_aos) {
  DBuf<float> tmp(3 * C->N);
  HNS_HIP(hipMemcpyAsync(tmp.p, host_aos, 3 * C->N * sizeof(float), hipMemcpyHostToDevice, C->stream));
  launch_aos_soa(C->N, tmp.p, dst, true, C->stream);
  HNS_HIP(hipStreamSynchronize(C->stream));
}

void ensure_field_storage(hns_context* C, Field& f) {
  if (!f.buf.p) f.buf.alloc(C->N);
}

// Status of one to-tolerance solve (the fixed RBGS count has none): HNS_ERR_NON_FINITE when a
// residual or ||div|| is not finite, HNS_ERR_NOT_CONVERGED when it ended above tol (iteration cap
// or update breakdown), else HNS_OK. Sets the message of a failing status.
hns_status solve_status(const SolveResult& res, int solver, float tol) {
  if (solver == HNS_SOLVER_RBGS_FIXED) return HNS_OK;
  if (res.end == kLoopNonFinite || std::isnan(res.rel_residual)) {
    g_err = "pressure solve met a residual or right-hand side that is not finite";
    return HNS_ERR_NON_FINITE;
  }
  if (!(res.rel_residual <= (double)tol)) {
    g_err = res.end == kLoopBreakdown
                ? "pressure solve broke down (d . A d <= 0) at iteration " + std::to_string(res.iterations) +
                      " before tol"
                : "pressure solve reached max_iters before tol";
    return HNS_ERR_NOT_CONVERGED;
  }
  return HNS_OK;
}
// the worse of two solve statuses (non-finite before not converged before OK)
hns_status worse(hns_status a, hns_status b) {
  auto rank = [](hns_status x) { return x == HNS_ERR_NON_FINITE ? 2 : (x == HNS_ERR_NOT_CONVERGED ? 1 : 0); };
  return rank(b) > rank(a) ? b : a;
}

// HNS_Advance body for one substep sequence (HNanoSolver.cu:223-456). Returns the worst status of
// the call's to-tolerance solves (every substep's, the disturbance re-solves' included).
hns_status advance(hns_context* C, int32_t iteration, float dt, const hns_params& P, const hns_solve_options* opts,
             hns_solve_stats* stats) {
  const float dx = C->voxelSize;
  const float inv_dx = 1.0f / dx;
  const int SUBS = std::max(1, (int)P.substeps);
  const float dt_sub = dt / float(SUBS);
  const DevTopo T = C->T();
  hipStream_t s = C->stream;
  Timing& tm = C->timing;
  const float* sdf = C->haveSdf ? C->sdf.p : nullptr;

  if (sdf) {
    tm.begin(HNS_K_OTHER, s);
    launch_enforce_collision(T, C->vel.p, sdf, dx, s);
    tm.end(s);
  }
  // advNames snapshot (HNanoSolver.cu:248-250): fields created below are not advected this call
  std::vector<std::string> advNames;
  for (auto& f : C->state)
    if (f.exists) advNames.push_back(f.name);

  const float omega = 2.0f / (1.0f + sinf(3.14159f * dx));  // HNanoSolver.cu:365
  SolveRequest req;
  req.omega = omega;
  req.dx = dx;
  if (opts) {
    req.solver = opts->solver;
    req.tol = opts->tol;
    req.max_iters = opts->max_iters;
    req.check_every = opts->check_every;
  }
  req.want_residual = (stats != nullptr);
  if (C->pc) {
    req.pc = &C->pc->net;
    req.pcw = &C->pc->work;
  }

  hns_status st = HNS_OK;
  std::string msg;
  auto note = [&](const SolveResult& res) {
    const hns_status a = solve_status(res, req.solver, req.tol);
    if (worse(st, a) != st) {
      st = a;
      msg = g_err;
    }
  };
  for (int sub = 0; sub < SUBS; ++sub) {
    // 1) advect velocity u -> A
    tm.begin(HNS_K_ADVECT_VELOCITY, s);
    launch_advect_velocity(T, C->vel.p, C->velA.p, dt_sub, inv_dx, sdf, s);
    tm.end(s);
    // vorticity confinement + velocity sources + gravity, A -> B (skipped when it is the identity:
    // (int)factorScale == 0 makes the confinement force exactly zero, Kernel.cu:707-712)
    const int32_t F = (int32_t)P.factorScale;
    const bool vort = F != 0 && P.vorticityScale != 0.0f;
    const float* srcV = C->haveSrcVel ? C->srcVel.p : nullptr;
    float* cur = C->velA.p;
    if (vort || srcV || P.gravity != 0.0f) {
      tm.begin(HNS_K_VORTICITY_FORCES, s);
      launch_vorticity_forces(T, C->velA.p, C->velB.p, dt_sub, inv_dx, P.vorticityScale, F, vort, srcV, P.gravity, s);
      tm.end(s);
      cur = C->velB.p;
    }
    // scalar sources apply to fields that exist before needOrMake (HNanoSolver.cu:298-311)
    CombustionArgs ca{};
    Field* fT = C->field("temperature");
    Field* fD = C->field("density");
    Field* fF = C->field("fuel");
    ca.src_temp = (fT && fT->exists) ? C->source("src_temperature") : nullptr;
    ca.src_density = (fD && fD->exists) ? C->source("src_density") : nullptr;
    ca.src_fuel = (fF && fF->exists) ? C->source("src_fuel") : nullptr;
    ca.density = (fD && fD->exists) ? fD->buf.p : nullptr;
    if (!ca.density) ca.src_density = nullptr;
    // needOrMake fuel, waste, temperature, flame (HNanoSolver.cu:326-338)
    float* comb[4];
    const char* mk[4] = {"fuel", "waste", "temperature", "flame"};
    for (int q = 0; q < 4; ++q) {
      Field* f = C->field(mk[q]);
      if (!f) {
        C->state.push_back(Field{mk[q], DBuf<float>(), false});
        f = &C->state.back();
        fT = C->field("temperature");
        fD = C->field("density");
        fF = C->field("fuel");
      }
      if (!f->exists) {
        ensure_field_storage(C, *f);
        HNS_HIP(hipMemsetAsync(f->buf.p, 0, C->N * sizeof(float), s));
        f->exists = true;
      }
      comb[q] = f->buf.p;
    }
    if (ca.density) ca.density = C->field("density")->buf.p;
    ca.fuel = comb[0];
    ca.waste = comb[1];
    ca.temp = comb[2];
    ca.flame = comb[3];
    ca.dt = dt_sub;
    ca.temp_gain = P.temperatureRelease;
    ca.expansion = P.expansionRate;
    ca.ambient = P.ambientTemp;
    ca.buoyancy = P.buoyancyStrength;
    float* ustar = (cur == C->velA.p) ? C->velB.p : C->velA.p;
    // 2)+3) divergence, sources, combustion (div += burn * expansion), buoyancy
    tm.begin(HNS_K_DIVERGENCE_COMBUSTION, s);
    launch_divergence(T, cur, C->div.p, inv_dx, true, &ca, ustar, s);
    tm.end(s);
    // 4) pressure from p = 0. Without the disturbance solve, the projection (5) is handed to the
    // solve, which launches it gated behind each read of its stop flag (the device goes on into it
    // while the host reads), and is launched below only when it did not run.
    const bool disturb = P.disturbanceEnable > 0.0f;
    if (!disturb)
      req.on_stop = [&, ustar](const int* gate) {
        launch_subtract_gradient(T, ustar, C->p.p, C->vel.p, inv_dx, sdf, s, gate);
        if (sdf) launch_enforce_collision(T, C->vel.p, sdf, dx, s, gate);
      };
    SolveResult res = C->solvers.solve(T, C->div.p, C->p.p, iteration, req, tm, s);
    req.on_stop = nullptr;
    note(res);
    // 4.5) disturbance: divergence of the buoyant velocity (drops the expansion), max(2, K/2) more
    if (disturb) {
      tm.begin(HNS_K_DIVERGENCE_COMBUSTION, s);
      launch_divergence(T, ustar, C->div.p, inv_dx, true, nullptr, nullptr, s);
      tm.end(s);
      res = C->solvers.solve(T, C->div.p, C->p.p, std::max(2, iteration / 2), req, tm, s);
      note(res);
    }
    if (stats) {
      stats->iterations = res.iterations;
      stats->rel_residual = res.rel_residual;
      stats->solve_ms = res.ms;
      stats->fallback_iterations = res.fallback;
    }
    // 5) projection into the state velocity, then collision
    if (!res.on_stop_ran) {
      tm.begin(HNS_K_GRADIENT, s);
      launch_subtract_gradient(T, ustar, C->p.p, C->vel.p, inv_dx, sdf, s);
      tm.end(s);
      if (sdf) {
        tm.begin(HNS_K_OTHER, s);
        launch_enforce_collision(T, C->vel.p, sdf, dx, s);
        tm.end(s);
      }
    }
    // 6) advect the snapshot's scalars with the projected velocity (misses read element 0)
    for (size_t b = 0; b < advNames.size(); b += FieldSet::kMax) {
      FieldSet fs{};
      const size_t m = std::min<size_t>(FieldSet::kMax, advNames.size() - b);
      if (C->advTmp.size() < advNames.size()) C->advTmp.resize(advNames.size());
      for (size_t q = 0; q < m; ++q) {
        Field* f = C->field(advNames[b + q]);
        if (!C->advTmp[b + q].p) C->advTmp[b + q].alloc(C->N);
        fs.in[q] = f->buf.p;
        fs.out[q] = C->advTmp[b + q].p;
      }
      fs.m = (int32_t)m;
      tm.begin(HNS_K_ADVECT_SCALARS, s);
      launch_advect_scalars(T, C->vel.p, fs, dt_sub, inv_dx, sdf, s);
      tm.end(s);
    }
    for (size_t q = 0; q < advNames.size(); ++q) std::swap(C->field(advNames[q])->buf, C->advTmp[q]);
  }
  HNS_HIP(hipStreamSynchronize(s));
  tm.collect();
  if (st != HNS_OK) g_err = msg;
  return st;
}

}  // namespace

extern "C" {

const char* hns_last_error(void) { return g_err.c_str(); }
int32_t hns_abi_version(void) { return HNS_ABI_VERSION; }

hns_status hns_leaf_order(const int32_t* origins, size_t n_leaves, uint64_t* perm) {
  return guard([&] {
    if ((!origins || !perm) && n_leaves) fail(HNS_ERR_INVALID_ARGUMENT, "hns_leaf_order: null pointer");
    leaf_order_host(origins, n_leaves, perm);
    return HNS_OK;
  });
}

hns_status hns_create_index_grid(const hns_grid_data* data, float voxel_size, hns_topology** out) {
  return guard([&] {
    if (!out) fail(HNS_ERR_INVALID_ARGUMENT, "CreateIndexGrid: output handle is null");
    if (!data) fail(HNS_ERR_INVALID_ARGUMENT, "CreateIndexGrid: payload is null");
    if (data->size == 0) fail(HNS_ERR_EMPTY, "CreateIndexGrid: N == 0");
    require_device();
    if (*out && same_leaves(*(*out)->s, data, voxel_size)) {  // unchanged since the last cook: keep it
      ++(*out)->reuses;
      return HNS_OK;
    }
    // built on the null stream: the caller's later work on any blocking stream is ordered after it
    std::shared_ptr<TopoStore> t = build_topology(data->coords, data->size, voxel_size, nullptr);
    if (*out) delete *out;
    *out = new hns_topology{t};
    return HNS_OK;
  });
}

hns_status hns_topology_builds(const hns_topology* topo, int64_t* builds, int64_t* reuses) {
  return guard([&] {
    if (!topo) fail(HNS_ERR_INVALID_ARGUMENT, "topology handle is null");
    if (builds) *builds = topo->builds;
    if (reuses) *reuses = topo->reuses;
    return HNS_OK;
  });
}

hns_status hns_payload_build(const hns_leaf_grid* grids, size_t num_grids, size_t domain_grid, int32_t padding,
                             hns_payload** out) {
  return guard([&] {
    if (!out) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_build: output handle is null");
    std::unique_ptr<hns_payload> h(new hns_payload());
    h->p = build_payload(grids, num_grids, domain_grid, padding);
    if (*out) delete *out;
    *out = h.release();
    return HNS_OK;
  });
}

hns_status hns_payload_view(hns_payload* payload, hns_grid_data* view, int32_t* pinned) {
  return guard([&] {
    if (!payload || !payload->p || !view) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_view: null argument");
    payload_view(payload->p, view, pinned);
    return HNS_OK;
  });
}

hns_status hns_payload_domain_mask(hns_payload* payload, uint64_t* active, size_t num_leaves) {
  return guard([&] {
    if (!payload || !payload->p) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_domain_mask: null payload");
    payload_domain_mask(payload->p, active, num_leaves);
    return HNS_OK;
  });
}

hns_status hns_write_leaf_grid(const hns_grid_data* data, const char* name, const uint64_t* domain_active,
                               int32_t* origins, float* values, uint64_t* active) {
  return guard([&] {
    write_leaf_grid(data, name, domain_active, origins, values, active);
    return HNS_OK;
  });
}

hns_status hns_payload_destroy(hns_payload** payload) {
  return guard([&] {
    if (payload && *payload) {
      delete *payload;
      *payload = nullptr;
    }
    return HNS_OK;
  });
}

hns_status hns_scatter_leaves(const hns_grid_data* data, const char* name, const uint8_t* mask, float background,
                              int32_t* origins, float* values, uint64_t* active, size_t capacity,
                              size_t* num_leaves) {
  return guard([&] {
    scatter_leaves(data, name, mask, background, origins, values, active, capacity, num_leaves);
    return HNS_OK;
  });
}

hns_status hns_topology_destroy(hns_topology** topo) {
  return guard([&] {
    if (topo && *topo) {
      delete *topo;
      *topo = nullptr;
    }
    return HNS_OK;
  });
}

hns_status hns_topology_info(const hns_topology* topo, size_t* n_voxels, size_t* n_leaves) {
  return guard([&] {
    if (!topo) fail(HNS_ERR_INVALID_ARGUMENT, "topology handle is null");
    if (n_voxels) *n_voxels = topo->s->N;
    if (n_leaves) *n_leaves = (size_t)topo->s->L;
    return HNS_OK;
  });
}

hns_status hns_topology_neighbours(const hns_topology* topo, int32_t* out, size_t capacity) {
  return guard([&] {
    if (!topo || !out) fail(HNS_ERR_INVALID_ARGUMENT, "null argument");
    const size_t L = (size_t)topo->s->L;
    if (capacity < 27 * L) fail(HNS_ERR_INVALID_ARGUMENT, "capacity below 27 x leaves");
    HNS_HIP(hipSetDevice(topo->s->device));
    std::vector<int32_t> h(L * kNbrStride);
    HNS_HIP(hipMemcpy(h.data(), topo->s->nbr.p, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (size_t l = 0; l < L; ++l) std::memcpy(out + 27 * l, h.data() + kNbrStride * l, 27 * sizeof(int32_t));
    return HNS_OK;
  });
}

hns_status hns_destroy_context(hns_context** ctx) {
  return guard([&] {
    if (!ctx || !*ctx) return HNS_OK;
    (void)hipSetDevice((*ctx)->device);
    delete *ctx;
    *ctx = nullptr;
    return HNS_OK;
  });
}

hns_status hns_ensure_context(hns_context** ctx, const hns_grid_data* data, const hns_topology* topo,
                              float voxel_size, int32_t has_collision) {
  return guard([&] {
    if (!ctx) fail(HNS_ERR_INVALID_ARGUMENT, "HNS_EnsureContext: ctx pointer is null");
    check_payload(data);
    const size_t N = data->size;
    if (!N) fail(HNS_ERR_EMPTY, "HNS_EnsureContext: N==0");
    if (!topo) fail(HNS_ERR_INVALID_ARGUMENT, "HNS_EnsureContext: topology handle is null");
    if (topo->s->N != N) fail(HNS_ERR_TOPOLOGY, "HNS_EnsureContext: payload size differs from the topology's");
    HNS_HIP(hipSetDevice(topo->s->device));
    hns_context* C = *ctx;
    const bool needNew = !C || C->N != N || C->voxelSize != voxel_size;
    const hns_value_block* sdfb = find_block(data, "collision_sdf", HNS_FLOAT);
    if (needNew) {
      const hns_value_block* vb = velocity_block(data);
      if (!vb) fail(HNS_ERR_NO_VELOCITY, "No Vec3f blocks in feedback");
      if (!vb->values) fail(HNS_ERR_INVALID_ARGUMENT, "velocity block has no values");
      if (C) {
        delete C;
        *ctx = nullptr;
      }
      std::unique_ptr<hns_context> nc(new hns_context());
      C = nc.get();
      C->device = topo->s->device;
      C->N = N;
      C->L = (int32_t)(N / kLeafVox);
      C->voxelSize = voxel_size;
      C->topo = topo->s;
      HNS_HIP(hipStreamCreate(&C->own));
      C->stream = C->own;
      C->vel.alloc(3 * N);
      C->velA.alloc(3 * N);
      C->velB.alloc(3 * N);
      C->div.alloc(N);
      C->p.alloc(N);
      upload_velocity_soa(C, C->vel.p, static_cast<const float*>(vb->values));
      for (size_t i = 0; i < data->num_blocks; ++i) {
        const hns_value_block& b = data->blocks[i];
        if (b.type != HNS_FLOAT || std::strcmp(b.name, "collision_sdf") == 0 || starts_with_src(b.name)) continue;
        if (!b.values) fail(HNS_ERR_INVALID_ARGUMENT, std::string("block '") + b.name + "' has no values");
        if (C->field(b.name)) continue;
        Field f{b.name, DBuf<float>(N), true};
        HNS_HIP(hipMemcpyAsync(f.buf.p, b.values, N * sizeof(float), hipMemcpyHostToDevice, C->stream));
        C->state.push_back(std::move(f));
      }
      if (has_collision && sdfb && sdfb->values) {
        C->sdf.alloc(N);
        HNS_HIP(hipMemcpyAsync(C->sdf.p, sdfb->values, N * sizeof(float), hipMemcpyHostToDevice, C->stream));
        C->haveSdf = true;
      }
      C->solvers.init(C->L, C->N);
      HNS_HIP(hipStreamSynchronize(C->stream));
      *ctx = nc.release();
      return HNS_OK;
    }
    // reuse: keep device state, refresh topology and collision field (HNanoSolver.cu:139-154)
    C->topo = topo->s;
    if (has_collision && sdfb && sdfb->values) {
      if (!C->sdf.p) C->sdf.alloc(N);
      HNS_HIP(hipMemcpyAsync(C->sdf.p, sdfb->values, N * sizeof(float), hipMemcpyHostToDevice, C->stream));
      HNS_HIP(hipStreamSynchronize(C->stream));
      C->haveSdf = true;
    } else {
      C->sdf.reset();
      C->haveSdf = false;
    }
    return HNS_OK;
  });
}

hns_status hns_upload_sources(hns_context* C, const hns_grid_data* data) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "HNS_UploadSources: null context");
    check_payload(data);
    if (data->size != C->N) fail(HNS_ERR_TOPOLOGY, "HNS_UploadSources: payload size differs from the context's");
    HNS_HIP(hipSetDevice(C->device));
    for (auto& kv : C->srcActive) kv.second = false;
    C->haveSrcVel = false;
    const size_t N = C->N;
    const hns_value_block* v3 = find_block(data, "src_vel", HNS_VEC3F);
    const hns_value_block* sx = find_block(data, "src_vel_x", HNS_FLOAT);
    const hns_value_block* sy = find_block(data, "src_vel_y", HNS_FLOAT);
    const hns_value_block* sz = find_block(data, "src_vel_z", HNS_FLOAT);
    if (v3) {
      if (v3->values) {
        if (!C->srcVel.p) C->srcVel.alloc(3 * N);
        upload_velocity_soa(C, C->srcVel.p, static_cast<const float*>(v3->values));
        C->haveSrcVel = true;
      }
    } else if (sx && sy && sz) {
      if (!C->srcVel.p) C->srcVel.alloc(3 * N);
      HNS_HIP(hipMemsetAsync(C->srcVel.p, 0, 3 * N * sizeof(float), C->stream));
      const hns_value_block* comps[3] = {sx, sy, sz};
      for (int k = 0; k < 3; ++k)
        if (comps[k]->values)
          HNS_HIP(hipMemcpyAsync(C->srcVel.p + k * N, comps[k]->values, N * sizeof(float), hipMemcpyHostToDevice,
                                 C->stream));
      C->haveSrcVel = true;
    }
    for (size_t i = 0; i < data->num_blocks; ++i) {
      const hns_value_block& b = data->blocks[i];
      if (b.type != HNS_FLOAT || !starts_with_src(b.name) || !b.values) continue;
      DBuf<float>& d = C->src[b.name];
      if (!d.p) d.alloc(N);
      HNS_HIP(hipMemcpyAsync(d.p, b.values, N * sizeof(float), hipMemcpyHostToDevice, C->stream));
      C->srcActive[b.name] = true;
    }
    HNS_HIP(hipStreamSynchronize(C->stream));
    return HNS_OK;
  });
}

hns_status hns_advance_ex(hns_context* C, int32_t iteration, float dt, const hns_params* params,
                          const hns_solve_options* opts, hns_solve_stats* stats) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "HNS_Advance: null context");
    if (!params) fail(HNS_ERR_INVALID_ARGUMENT, "HNS_Advance: params is null");
    if (iteration < 0) fail(HNS_ERR_INVALID_ARGUMENT, "HNS_Advance: iteration < 0");
    if (opts && opts->struct_size < sizeof(hns_solve_options))
      fail(HNS_ERR_INVALID_ARGUMENT, "hns_solve_options.struct_size too small");
    if (stats && stats->struct_size < sizeof(hns_solve_stats))
      fail(HNS_ERR_INVALID_ARGUMENT, "hns_solve_stats.struct_size too small");
    HNS_HIP(hipSetDevice(C->device));
    return advance(C, iteration, dt, *params, opts, stats);
  });
}

hns_status hns_advance(hns_context* C, int32_t iteration, float dt, const hns_params* params) {
  return hns_advance_ex(C, iteration, dt, params, nullptr, nullptr);
}

hns_status hns_download_for_output(hns_context* C, hns_grid_data* data) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "HNS_DownloadForOutput: null context");
    check_payload(data);
    if (data->size != C->N) fail(HNS_ERR_TOPOLOGY, "HNS_DownloadForOutput: payload size differs from the context's");
    HNS_HIP(hipSetDevice(C->device));
    const hns_value_block* vb = velocity_block(data);
    if (!vb) return HNS_OK;  // HNanoSolver.cu:478 returns before the scalars
    const size_t N = C->N;
    if (vb->values) {
      launch_aos_soa(N, C->vel.p, C->velA.p, false, C->stream);
      HNS_HIP(hipMemcpyAsync(vb->values, C->velA.p, 3 * N * sizeof(float), hipMemcpyDeviceToHost, C->stream));
    }
    for (auto& f : C->state) {
      if (!f.exists) continue;
      const hns_value_block* b = find_block(data, f.name.c_str(), HNS_FLOAT);
      if (!b || !b->values) continue;
      HNS_HIP(hipMemcpyAsync(b->values, f.buf.p, N * sizeof(float), hipMemcpyDeviceToHost, C->stream));
    }
    HNS_HIP(hipStreamSynchronize(C->stream));
    return HNS_OK;
  });
}

hns_status hns_get_active_mask(hns_context* C, uint8_t* mask, size_t capacity, size_t* n_out) {
  return guard([&] {
    if (!C) {
      if (n_out) *n_out = 0;
      return HNS_OK;
    }
    if (!mask && capacity) fail(HNS_ERR_INVALID_ARGUMENT, "mask is null");
    if (mask && capacity) std::memset(mask, 1, std::min(capacity, C->N));  // build_activity_mask writes 1 (Kernel.cu:74-75)
    if (n_out) *n_out = C->N;
    return HNS_OK;
  });
}

hns_status hns_set_stream(hns_context* C, void* stream) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "null context");
    C->stream = stream ? static_cast<hipStream_t>(stream) : C->own;
    return HNS_OK;
  });
}

hns_status hns_context_info(const hns_context* C, size_t* n_voxels, size_t* n_leaves, int32_t* n_state) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "null context");
    if (n_voxels) *n_voxels = C->N;
    if (n_leaves) *n_leaves = (size_t)C->L;
    if (n_state) {
      int32_t k = 0;
      for (auto& f : C->state) k += f.exists ? 1 : 0;
      *n_state = k;
    }
    return HNS_OK;
  });
}

int32_t hns_has_field(const hns_context* C, const char* name) {
  if (!C || !name) return 0;
  for (auto& f : C->state)
    if (f.exists && f.name == name) return 1;
  return 0;
}

hns_status hns_set_timing(hns_context* C, int32_t enable) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "null context");
    C->timing.enabled = enable != 0;
    return HNS_OK;
  });
}
hns_status hns_get_timing(hns_context* C, int32_t which, double* total_ms, int64_t* launches) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "null context");
    if (which < 0 || which >= HNS_K_COUNT) fail(HNS_ERR_INVALID_ARGUMENT, "kernel class out of range");
    if (total_ms) *total_ms = C->timing.total_ms[which];
    if (launches) *launches = C->timing.launches[which];
    return HNS_OK;
  });
}
hns_status hns_reset_timing(hns_context* C) {
  return guard([&] {
    if (!C) fail(HNS_ERR_NULL_CONTEXT, "null context");
    C->timing.reset();
    return HNS_OK;
  });
}

// ---- standalone entry points --------------------------------------------------------------
namespace {
struct Standalone {
  hipStream_t s = nullptr;
  hipStream_t own = nullptr;
  std::shared_ptr<TopoStore> topo;
  explicit Standalone(const hns_grid_data* data, float voxel_size, void* stream) {
    check_payload(data);
    if (!data->size) fail(HNS_ERR_EMPTY, "N == 0");
    require_device();
    if (stream) s = static_cast<hipStream_t>(stream);
    else {
      HNS_HIP(hipStreamCreate(&own));
      s = own;
    }
    topo = build_topology(data->coords, data->size, voxel_size, s);
  }
  ~Standalone() {
    if (own) (void)hipStreamDestroy(own);
  }
};
const hns_value_block* single_vec3(const hns_grid_data* d) {
  const hns_value_block* v = nullptr;
  int count = 0;
  for (size_t i = 0; i < d->num_blocks; ++i)
    if (d->blocks[i].type == HNS_VEC3F) {
      v = &d->blocks[i];
      ++count;
    }
  if (count != 1) fail(HNS_ERR_NO_VELOCITY, "Expected exactly one Vec3f block (velocity)");
  if (!v->values) fail(HNS_ERR_INVALID_ARGUMENT, "velocity block has no values");
  return v;
}
void to_device_soa(const float* aos, float* soa, size_t n, hipStream_t s) {
  DBuf<float> tmp(3 * n);
  HNS_HIP(hipMemcpyAsync(tmp.p, aos, 3 * n * sizeof(float), hipMemcpyHostToDevice, s));
  launch_aos_soa(n, tmp.p, soa, true, s);
  HNS_HIP(hipStreamSynchronize(s));
}
void from_device_soa(const float* soa, float* aos, size_t n, hipStream_t s) {
  DBuf<float> tmp(3 * n);
  launch_aos_soa(n, soa, tmp.p, false, s);
  HNS_HIP(hipMemcpyAsync(aos, tmp.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, s));
  HNS_HIP(hipStreamSynchronize(s));
}
}  // namespace

hns_status hns_advect_index_grid(hns_grid_data* data, float dt, float voxel_size, void* stream) {
  return guard([&] {
    check_payload(data);
    const hns_value_block* vb = single_vec3(data);
    bool anyFloat = false;
    for (size_t i = 0; i < data->num_blocks; ++i) anyFloat |= data->blocks[i].type == HNS_FLOAT;
    if (!anyFloat) fail(HNS_ERR_NOT_FOUND, "No float blocks found");
    Standalone S(data, voxel_size, stream);
    const size_t N = data->size;
    DBuf<float> vel(3 * N), in(N), out(N);
    to_device_soa(static_cast<const float*>(vb->values), vel.p, N, S.s);
    const DevTopo T = S.topo->dev();
    for (size_t i = 0; i < data->num_blocks; ++i) {
      hns_value_block& b = data->blocks[i];
      if (b.type != HNS_FLOAT || !b.values) continue;
      HNS_HIP(hipMemcpyAsync(in.p, b.values, N * sizeof(float), hipMemcpyHostToDevice, S.s));
      launch_advect_scalar(T, vel.p, in.p, out.p, dt, 1.0f / voxel_size, nullptr, S.s);
      HNS_HIP(hipMemcpyAsync(b.values, out.p, N * sizeof(float), hipMemcpyDeviceToHost, S.s));
    }
    HNS_HIP(hipStreamSynchronize(S.s));
    return HNS_OK;
  });
}

hns_status hns_advect_index_grid_velocity(hns_grid_data* data, float dt, float voxel_size, void* stream) {
  return guard([&] {
    check_payload(data);
    const hns_value_block* vb = single_vec3(data);
    Standalone S(data, voxel_size, stream);
    const size_t N = data->size;
    DBuf<float> vel(3 * N), out(3 * N);
    to_device_soa(static_cast<const float*>(vb->values), vel.p, N, S.s);
    launch_advect_velocity(S.topo->dev(), vel.p, out.p, dt, 1.0f / voxel_size, nullptr, S.s);
    from_device_soa(out.p, static_cast<float*>(vb->values), N, S.s);
    return HNS_OK;
  });
}

hns_status hns_divergence(hns_grid_data* data, float voxel_size, void* stream) {
  return guard([&] {
    check_payload(data);
    const hns_value_block* vb = single_vec3(data);
    const hns_value_block* db = find_block(data, "divergence", HNS_FLOAT);
    if (!db || !db->values) fail(HNS_ERR_NOT_FOUND, "payload has no \"divergence\" float block");
    Standalone S(data, voxel_size, stream);
    const size_t N = data->size;
    DBuf<float> vel(3 * N), div(N);
    to_device_soa(static_cast<const float*>(vb->values), vel.p, N, S.s);
    launch_divergence(S.topo->dev(), vel.p, div.p, 1.0f / voxel_size, false, nullptr, nullptr, S.s);
    HNS_HIP(hipMemcpyAsync(db->values, div.p, N * sizeof(float), hipMemcpyDeviceToHost, S.s));
    HNS_HIP(hipStreamSynchronize(S.s));
    return HNS_OK;
  });
}

hns_status hns_project_non_divergent(hns_grid_data* data, size_t iterations, float voxel_size, void* stream) {
  return guard([&] {
    check_payload(data);
    const hns_value_block* vb = single_vec3(data);
    if (iterations > (size_t)INT32_MAX) fail(HNS_ERR_INVALID_ARGUMENT, "iterations too large");
    Standalone S(data, voxel_size, stream);
    const size_t N = data->size;
    DBuf<float> vel(3 * N), out(3 * N), div(N), p(N);
    to_device_soa(static_cast<const float*>(vb->values), vel.p, N, S.s);
    const DevTopo T = S.topo->dev();
    launch_divergence(T, vel.p, div.p, 1.0f / voxel_size, true, nullptr, nullptr, S.s);
    // PressureProjection.cu:53 evaluates sin in double
    const float omega = (float)(2.0 / (1.0 + std::sin(3.14159 * (double)voxel_size)));
    if (iterations == 0) HNS_HIP(hipMemsetAsync(p.p, 0, N * sizeof(float), S.s));
    for (size_t it = 0; it < iterations; ++it) {
      launch_rbgs(T, div.p, p.p, voxel_size, 0, omega, it == 0, S.s);
      launch_rbgs(T, div.p, p.p, voxel_size, 1, omega, it == 0, S.s);
    }
    launch_subtract_gradient(T, vel.p, p.p, out.p, 1.0f / voxel_size, nullptr, S.s);
    from_device_soa(out.p, static_cast<float*>(vb->values), N, S.s);
    return HNS_OK;
  });
}

// ---- device-array kernel API --------------------------------------------------------------
namespace {
DevTopo topo_of(const hns_topology* t) {
  if (!t) fail(HNS_ERR_INVALID_ARGUMENT, "topology handle is null");
  HNS_HIP(hipSetDevice(t->s->device));
  return t->s->dev();
}
void finish(void* stream, int32_t async) {
  if (!async) HNS_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}
}  // namespace

hns_status hnsk_advect_velocity(const hns_topology* topo, const float* vel, float* out, float dt, float inv_dx,
                                const float* sdf, void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !out) fail(HNS_ERR_INVALID_ARGUMENT, "null array");
    launch_advect_velocity(T, vel, out, dt, inv_dx, sdf, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_advect_scalars(const hns_topology* topo, const float* vel, int32_t m, const float* const* in,
                               float* const* out, float dt, float inv_dx, const float* sdf, void* stream,
                               int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || m < 0 || (m && (!in || !out))) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    for (int32_t b = 0; b < m; b += FieldSet::kMax) {
      FieldSet fs{};
      fs.m = std::min<int32_t>(FieldSet::kMax, m - b);
      for (int32_t q = 0; q < fs.m; ++q) {
        fs.in[q] = in[b + q];
        fs.out[q] = out[b + q];
      }
      launch_advect_scalars(T, vel, fs, dt, inv_dx, sdf, static_cast<hipStream_t>(stream));
    }
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_advect_scalar(const hns_topology* topo, const float* vel, const float* in, float* out, float dt,
                              float inv_dx, const float* sdf, void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !in || !out) fail(HNS_ERR_INVALID_ARGUMENT, "null array");
    launch_advect_scalar(T, vel, in, out, dt, inv_dx, sdf, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_vorticity_forces(const hns_topology* topo, const float* vel, float* out, float dt, float inv_dx,
                                 float vorticity_scale, float factor_scale, const float* src_vel, float gravity,
                                 void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !out || vel == out) fail(HNS_ERR_INVALID_ARGUMENT, "vel and out must be distinct arrays");
    const int32_t F = (int32_t)factor_scale;
    launch_vorticity_forces(T, vel, out, dt, inv_dx, vorticity_scale, F, F != 0 && vorticity_scale != 0.0f, src_vel,
                            gravity, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_divergence(const hns_topology* topo, const float* vel, float* div, float inv_dx,
                           int32_t colour_split, void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !div) fail(HNS_ERR_INVALID_ARGUMENT, "null array");
    launch_divergence(T, vel, div, inv_dx, colour_split != 0, nullptr, nullptr, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_rbgs_sweep(const hns_topology* topo, const float* div_cs, float* p_cs, float dx, int32_t colour,
                           float omega, void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!div_cs || !p_cs || (colour != 0 && colour != 1)) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    launch_rbgs(T, div_cs, p_cs, dx, colour, omega, false, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_residual(const hns_topology* topo, const float* p_cs, const float* div_cs, float dx, double* r2,
                         double* d2, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!p_cs || !div_cs || !r2 || !d2) fail(HNS_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    DBuf<double> partial(2 * (size_t)T.L), out(2);
    launch_residual(T, p_cs, div_cs, dx, partial.p, out.p, s);
    double h[2];
    HNS_HIP(hipStreamSynchronize(s));
    HNS_HIP(hipMemcpy(h, out.p, sizeof(h), hipMemcpyDeviceToHost));
    *r2 = h[0];
    *d2 = h[1];
    return HNS_OK;
  });
}

hns_status hnsk_residual_noalloc(const hns_topology* topo, const float* p_cs, const float* div_cs, float dx,
                                 double* out_dev, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!p_cs || !div_cs || !out_dev) fail(HNS_ERR_INVALID_ARGUMENT, "null argument");
    launch_residual_noalloc(T, p_cs, div_cs, dx, out_dev, static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

hns_status hnsk_pressure_solve(const hns_topology* topo, const float* div_cs, float* p_cs, float dx, float omega,
                               int32_t iterations, const hns_solve_options* opts, hns_solve_stats* stats,
                               void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!div_cs || !p_cs || iterations < 0) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    if (opts && opts->struct_size < sizeof(hns_solve_options))
      fail(HNS_ERR_INVALID_ARGUMENT, "hns_solve_options.struct_size too small");
    if (stats && stats->struct_size < sizeof(hns_solve_stats))
      fail(HNS_ERR_INVALID_ARGUMENT, "hns_solve_stats.struct_size too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    Solvers& sol = topo->solvers();
    Timing tm;
    SolveRequest req;
    req.omega = omega;
    req.dx = dx;
    req.want_residual = true;
    if (opts) {
      req.solver = opts->solver;
      req.tol = opts->tol;
      req.max_iters = opts->max_iters;
      req.check_every = opts->check_every;
    }
    const SolveResult res = sol.solve(T, div_cs, p_cs, iterations, req, tm, s);
    HNS_HIP(hipStreamSynchronize(s));
    if (stats) {
      stats->iterations = res.iterations;
      stats->rel_residual = res.rel_residual;
      stats->solve_ms = res.ms;
      stats->fallback_iterations = res.fallback;
    }
    return opts ? solve_status(res, opts->solver, opts->tol) : HNS_OK;
  });
}

hns_status hnsk_subtract_gradient(const hns_topology* topo, const float* vel, const float* p_cs, float* out,
                                  float inv_dx, const float* sdf, void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !p_cs || !out) fail(HNS_ERR_INVALID_ARGUMENT, "null array");
    launch_subtract_gradient(T, vel, p_cs, out, inv_dx, sdf, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_enforce_collision(const hns_topology* topo, float* vel, const float* sdf, float voxel_size,
                                  void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !sdf) fail(HNS_ERR_INVALID_ARGUMENT, "null array");
    launch_enforce_collision(T, vel, sdf, voxel_size, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_combustion_buoyancy(const hns_topology* topo, float* fuel, float* waste, float* temperature,
                                    float* flame, float* div, const float* vel, float* vel_out, float dt,
                                    float temp_gain, float expansion, float ambient, float buoyancy, void* stream,
                                    int32_t async) {
  return guard([&] {
    if (!topo) fail(HNS_ERR_INVALID_ARGUMENT, "topology handle is null");
    HNS_HIP(hipSetDevice(topo->s->device));
    if (!fuel || !waste || !temperature || !flame || !div || !vel || !vel_out)
      fail(HNS_ERR_INVALID_ARGUMENT, "null array");
    CombustionArgs c{};
    c.fuel = fuel;
    c.waste = waste;
    c.temp = temperature;
    c.flame = flame;
    c.dt = dt;
    c.temp_gain = temp_gain;
    c.expansion = expansion;
    c.ambient = ambient;
    c.buoyancy = buoyancy;
    launch_combustion_buoyancy(topo->s->N, c, div, vel, vel_out, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

hns_status hnsk_to_colour_split(const hns_topology* topo, const float* in, float* out, void* stream, int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!in || !out || in == out) fail(HNS_ERR_INVALID_ARGUMENT, "in and out must be distinct arrays");
    launch_to_colour_split(T, in, out, false, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}
hns_status hnsk_from_colour_split(const hns_topology* topo, const float* in, float* out, void* stream,
                                  int32_t async) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!in || !out || in == out) fail(HNS_ERR_INVALID_ARGUMENT, "in and out must be distinct arrays");
    launch_to_colour_split(T, in, out, true, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}
hns_status hnsk_aos_to_soa(size_t n, const float* in, float* out, void* stream, int32_t async) {
  return guard([&] {
    if ((!in || !out || in == out) && n) fail(HNS_ERR_INVALID_ARGUMENT, "in and out must be distinct arrays");
    launch_aos_soa(n, in, out, true, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}
hns_status hnsk_soa_to_aos(size_t n, const float* in, float* out, void* stream, int32_t async) {
  return guard([&] {
    if ((!in || !out || in == out) && n) fail(HNS_ERR_INVALID_ARGUMENT, "in and out must be distinct arrays");
    launch_aos_soa(n, in, out, false, static_cast<hipStream_t>(stream));
    finish(stream, async);
    return HNS_OK;
  });
}

/* ---- learned preconditioner ------------------------------------------------------------- */

int64_t hns_precond_num_params(void) { return pc_num_params(); }

hns_status hns_precond_create(const float* params, int64_t n, hns_precond** out) {
  return guard([&] {
    if (!out) fail(HNS_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (!params || n != pc_num_params()) fail(HNS_ERR_INVALID_ARGUMENT, "expected hns_precond_num_params() weights");
    require_device();
    auto p = std::make_unique<hns_precond>();
    HNS_HIP(hipGetDevice(&p->device));
    p->net.load_host(params);
    p->net.refresh(nullptr);
    HNS_HIP(hipDeviceSynchronize());
    *out = p.release();
    return HNS_OK;
  });
}

hns_status hns_precond_create_device(float* params_dev, int64_t n, hns_precond** out) {
  return guard([&] {
    if (!out) fail(HNS_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (!params_dev || n != pc_num_params()) fail(HNS_ERR_INVALID_ARGUMENT, "expected hns_precond_num_params() weights");
    require_device();
    auto p = std::make_unique<hns_precond>();
    HNS_HIP(hipGetDevice(&p->device));
    p->net.set_device_params(params_dev);
    p->net.refresh(nullptr);
    HNS_HIP(hipDeviceSynchronize());
    *out = p.release();
    return HNS_OK;
  });
}

hns_status hns_precond_refresh(hns_precond* p, void* stream) {
  return guard([&] {
    if (!p) fail(HNS_ERR_INVALID_ARGUMENT, "null preconditioner");
    p->net.refresh(static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

hns_status hns_precond_destroy(hns_precond** p) {
  return guard([&] {
    if (!p) fail(HNS_ERR_INVALID_ARGUMENT, "null handle pointer");
    if (*p) {
      (void)hipDeviceSynchronize();
      delete *p;
      *p = nullptr;
    }
    return HNS_OK;
  });
}

hns_status hns_set_preconditioner(hns_context* ctx, hns_precond* p) {
  return guard([&] {
    if (!ctx) fail(HNS_ERR_NULL_CONTEXT, "null context");
    ctx->pc = p;
    return HNS_OK;
  });
}

hns_status hnsk_precond_apply(const hns_topology* topo, hns_precond* p, const float* r_cs, float* z_cs, int32_t keep,
                              void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!p || !r_cs || !z_cs) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    p->net.forward(T, p->work, r_cs, z_cs, keep != 0, static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

hns_status hnsk_precond_backward(const hns_topology* topo, hns_precond* p, const float* r_cs, const float* dz_cs,
                                 float* grad, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!p || !r_cs || !dz_cs || !grad) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    p->net.backward(T, p->work, r_cs, dz_cs, grad, static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

hns_status hnsk_precond_adam(int64_t n, float* params, const float* grad, float* m, float* v, float lr, float beta1,
                             float beta2, float eps, int32_t step, void* stream) {
  return guard([&] {
    if (n < 0 || !params || !grad || !m || !v || step < 1) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    launch_pc_adam(n, params, grad, m, v, lr, beta1, beta2, eps, step, static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

hns_status hnsk_poisson_apply(const hns_topology* topo, const float* x_cs, const float* y_cs, float a, float b,
                              float* out_cs, double* work, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!x_cs || !y_cs || !out_cs || !work || out_cs == x_cs) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    launch_pc_poisson(T, x_cs, y_cs, a, b, out_cs, work + 4, work, static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

hns_status hnsk_mg_apply(const hns_topology* topo, const float* r_cs, float* z_cs, int32_t flags, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!r_cs || !z_cs || r_cs == z_cs || (flags & ~(HNS_MG_GRAPH | HNS_MG_COARSE_LAUNCHES | HNS_MG_UNFUSED)))
      fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    const unsigned variant = ((flags & HNS_MG_COARSE_LAUNCHES) ? kMgCoarseLaunches : 0u) |
                             ((flags & HNS_MG_UNFUSED) ? kMgUnfused : 0u);
    topo->solvers().multigrid().apply(T, r_cs, z_cs, static_cast<hipStream_t>(stream), (flags & HNS_MG_GRAPH) != 0,
                                      nullptr, variant);
    return HNS_OK;
  });
}

hns_status hnsk_pressure_solve_learned(const hns_topology* topo, hns_precond* p, const float* div_cs, float* p_cs,
                                       float dx, const hns_solve_options* opts, hns_solve_stats* stats,
                                       void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!p || !div_cs || !p_cs || !opts) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    if (opts->struct_size < sizeof(hns_solve_options))
      fail(HNS_ERR_INVALID_ARGUMENT, "hns_solve_options.struct_size too small");
    if (stats && stats->struct_size < sizeof(hns_solve_stats))
      fail(HNS_ERR_INVALID_ARGUMENT, "hns_solve_stats.struct_size too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    Solvers& sol = topo->solvers();
    Timing tm;
    SolveRequest req;
    req.dx = dx;
    req.omega = 2.0f / (1.0f + sinf(3.14159f * dx));  // the fallback RBGS, HNanoSolver.cu:365
    req.want_residual = true;
    req.solver = HNS_SOLVER_PCG_LEARNED;
    req.tol = opts->tol;
    req.max_iters = opts->max_iters;
    req.pc = &p->net;
    req.pcw = &p->work;
    const SolveResult res = sol.solve(T, div_cs, p_cs, 0, req, tm, s);
    HNS_HIP(hipStreamSynchronize(s));
    if (stats) {
      stats->iterations = res.iterations;
      stats->rel_residual = res.rel_residual;
      stats->solve_ms = res.ms;
      stats->fallback_iterations = res.fallback;
    }
    return solve_status(res, HNS_SOLVER_PCG_LEARNED, opts->tol);
  });
}

/* ---- reverse mode ------------------------------------------------------------------------ */

hns_status hnsk_diff_forward(const hns_topology* topo, const float* vel0, const float* rho0, int32_t steps, float dt,
                             float dx, int32_t iterations, float omega, int32_t checkpoint_every, float* vel_out,
                             float* rho_out, hns_diff_tape** tape, void* stream) {
  return guard([&] {
    if (!tape) fail(HNS_ERR_INVALID_ARGUMENT, "null tape pointer");
    *tape = nullptr;
    (void)topo_of(topo);
    if (!vel0 || !rho0 || steps < 1 || iterations < 0 || !(dx > 0.f))
      fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    auto t = std::make_unique<hns_diff_tape>();
    t->t = std::make_unique<DiffTape>(topo->s, steps, dt, dx, iterations, omega, checkpoint_every);
    t->t->forward(vel0, rho0, vel_out, rho_out, static_cast<hipStream_t>(stream));
    *tape = t.release();
    return HNS_OK;
  });
}

hns_status hnsk_diff_backward(hns_diff_tape* tape, const float* gvel_out, const float* grho_out, float* gvel0,
                              float* grho0, void* stream) {
  return guard([&] {
    if (!tape || !tape->t || !gvel0 || !grho0) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    tape->t->backward(gvel_out, grho_out, gvel0, grho0, static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

hns_status hns_diff_tape_info(const hns_diff_tape* tape, int32_t* checkpoints, int64_t* bytes,
                              int64_t* recompute_mismatches) {
  return guard([&] {
    if (!tape || !tape->t) fail(HNS_ERR_INVALID_ARGUMENT, "null tape");
    if (checkpoints) *checkpoints = tape->t->checkpoints();
    if (bytes) *bytes = tape->t->bytes();
    if (recompute_mismatches) *recompute_mismatches = tape->t->mismatches();
    return HNS_OK;
  });
}

hns_status hns_diff_tape_destroy(hns_diff_tape** tape) {
  return guard([&] {
    if (!tape) fail(HNS_ERR_INVALID_ARGUMENT, "null tape pointer");
    if (*tape) {
      (void)hipDeviceSynchronize();
      delete *tape;
      *tape = nullptr;
    }
    return HNS_OK;
  });
}

hns_status hnsk_advect_scalars_adj(const hns_topology* topo, const float* vel, const float* in, const float* gout,
                                   float* gin, float* gvel, float dt, float inv_dx, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !in || !gout || !gin || !gvel) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    const size_t N = topo->s->N;
    DBuf<float> far(4 * N);
    hipStream_t s = static_cast<hipStream_t>(stream);
    HNS_HIP(hipMemsetAsync(far.p, 0, 4 * N * sizeof(float), s));
    launch_advect_scalars_adj(T, vel, in, gout, gin, gvel, dt, inv_dx, far.p, s);
    HNS_HIP(hipStreamSynchronize(s));
    return HNS_OK;
  });
}

hns_status hnsk_advect_vector_adj(const hns_topology* topo, const float* vel, const float* gout, float* gvel, float dt,
                                  float inv_dx, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!vel || !gout || !gvel) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    const size_t N = topo->s->N;
    DBuf<float> far(3 * N);
    hipStream_t s = static_cast<hipStream_t>(stream);
    HNS_HIP(hipMemsetAsync(far.p, 0, 3 * N * sizeof(float), s));
    launch_advect_vector_adj(T, vel, gout, gvel, dt, inv_dx, far.p, s);
    HNS_HIP(hipStreamSynchronize(s));
    return HNS_OK;
  });
}

hns_status hnsk_rbgs_adj(const hns_topology* topo, float* g_cs, float* gdiv_cs, float dx, int32_t iterations,
                         float omega, void* stream) {
  return guard([&] {
    const DevTopo T = topo_of(topo);
    if (!g_cs || !gdiv_cs || iterations < 0) fail(HNS_ERR_INVALID_ARGUMENT, "bad arguments");
    launch_rbgs_adj(T, g_cs, gdiv_cs, dx, iterations, omega, static_cast<hipStream_t>(stream));
    return HNS_OK;
  });
}

}  // extern "C"
