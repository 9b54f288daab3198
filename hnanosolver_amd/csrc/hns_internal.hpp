// Host-side internals shared by the translation units of libhns.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hns.h"
#include "hns_device.hpp"

namespace hns {

// Error carried to the C ABI boundary, where it becomes a status code + message.
struct Error : std::runtime_error {
  hns_status code;
  Error(hns_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(hns_status c, const std::string& m) { throw Error(c, m); }

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(HNS_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    fail(HNS_ERR_NO_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  fail(HNS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HNS_HIP(call) ::hns::hip_check((call), #call)

// Device buffer with RAII.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { reset(); }
  void alloc(size_t count) {
    reset();
    if (count == 0) return;
    HNS_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Device-resident leaf topology (the role of NanoVDB's IndexGrid in the reference).
struct TopoStore {
  int device = 0;
  size_t N = 0;
  int32_t L = 0;
  float voxelSize = 0.f;
  DBuf<int4> origins;
  DBuf<int32_t> nbr;
  DBuf<unsigned long long> hkeys;
  DBuf<int32_t> hvals;
  uint32_t hmask = 0;
  uint64_t id = 0;  // unique per build in this process (never reused, unlike a freed address)
  std::vector<int32_t> host_origins;  // 3 per leaf: the build's leaf origins, for reuse checks
  DevTopo dev() const {
    DevTopo t;
    t.id = id;
    t.L = L;
    t.hmask = hmask;
    t.origins = origins.p;
    t.nbr = nbr.p;
    t.hkeys = hkeys.p;
    t.hvals = hvals.p;
    return t;
  }
};

// Builds the topology from host coords (3N int32). Throws Error on invalid input.
std::shared_ptr<TopoStore> build_topology(const int32_t* coords, size_t n, float voxelSize, hipStream_t s);
void leaf_order_host(const int32_t* origins, size_t L, uint64_t* perm);

// Host ingestion (ingest.hip).
struct Payload;
Payload* build_payload(const hns_leaf_grid* grids, size_t num_grids, size_t domain_grid, int32_t padding);
void destroy_payload(Payload* p);
void payload_view(Payload* p, hns_grid_data* view, int32_t* pinned);
void payload_domain_mask(Payload* p, uint64_t* active, size_t num_leaves);
void write_leaf_grid(const hns_grid_data* d, const char* name, const uint64_t* domain_active, int32_t* origins,
                     float* values, uint64_t* active);
void scatter_leaves(const hns_grid_data* d, const char* name, const uint8_t* mask, float background, int32_t* origins,
                    float* values, uint64_t* active, size_t capacity, size_t* num_leaves);

// ---- kernel launchers (kernels_step.hip) ------------------------------------------------
void launch_advect_velocity(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, const float* sdf,
                            hipStream_t s);
struct FieldSet {
  static constexpr int kMax = 8;
  const float* in[kMax];
  float* out[kMax];
  int32_t m;
};
void launch_advect_scalars(const DevTopo& T, const float* u, const FieldSet& f, float dt, float inv_dx,
                           const float* sdf, hipStream_t s);
void launch_advect_scalar(const DevTopo& T, const float* u, const float* in, float* out, float dt, float inv_dx,
                          const float* sdf, hipStream_t s);
void launch_vorticity_forces(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, float scale,
                             int32_t F, bool vort, const float* src_vel, float gravity, hipStream_t s);
struct CombustionArgs {
  float* fuel;
  float* waste;
  float* temp;
  float* flame;
  float* density;              // only written when src_density is set
  const float* src_fuel;       // nullable; applied when the field existed before this substep
  const float* src_temp;
  const float* src_density;
  float dt, temp_gain, expansion, ambient, buoyancy;
};
// mode 0: divergence only; mode 1: divergence + sources + combustion + buoyancy
void launch_divergence(const DevTopo& T, const float* u, float* div, float inv_dx, bool colour_split,
                       const CombustionArgs* comb, float* u_out, hipStream_t s);
void launch_combustion_buoyancy(size_t N, const CombustionArgs& c, float* div, const float* u, float* u_out,
                                hipStream_t s);
// gate (non-null): the kernel runs only while the device int *gate is set
void launch_subtract_gradient(const DevTopo& T, const float* u, const float* p_cs, float* out, float inv_dx,
                              const float* sdf, hipStream_t s, const int* gate = nullptr);
void launch_enforce_collision(const DevTopo& T, float* u, const float* sdf, float voxelSize, hipStream_t s,
                              const int* gate = nullptr);
void launch_to_colour_split(const DevTopo& T, const float* in, float* out, bool inverse, hipStream_t s);
void launch_aos_soa(size_t n, const float* in, float* out, bool to_soa, hipStream_t s);

// ---- LDS-staged variants (kernels_advect.hip); HNS_ADVECT_KERNEL=gather disables them ----
bool lds_kernels_enabled();
void launch_advect_velocity_lds(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, const float* sdf,
                                hipStream_t s);
void launch_advect_scalars_lds(const DevTopo& T, const float* u, const FieldSet& f, float dt, float inv_dx,
                               const float* sdf, hipStream_t s);
// F must be +-1 (every stencil inside the 2-voxel halo)
void launch_vorticity_forces_lds(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, float scale,
                                 int32_t F, const float* src_vel, float gravity, hipStream_t s);

// ---- pressure kernels (kernels_pressure.hip) -------------------------------------------
// One colour sweep; first != 0 treats p as zero before the sweep (iteration 0 from p = 0).
void launch_rbgs(const DevTopo& T, const float* div_cs, float* p_cs, float dx, int colour, float omega, bool first,
                 hipStream_t s);
// One full red-black iteration (both colours), temporally blocked in LDS; p_in != p_out.
// first != 0 treats p_in as zero (it is not read).
void launch_rbgs_iteration(const DevTopo& T, const float* div_cs, const float* p_in, float* p_out, float dx,
                           float omega, bool first, hipStream_t s);
// Writes per-workgroup partial sums (2 doubles per leaf) then reduces into out[0..1].
void launch_residual(const DevTopo& T, const float* p_cs, const float* div_cs, float dx, double* partial,
                     double* out, hipStream_t s);
// The same with the partial sums in the library's static device scratch (no allocation; fails for
// more than HNS_NOALLOC_MAX_LEAVES leaves). Calls that overlap on one device race on the scratch.
void launch_residual_noalloc(const DevTopo& T, const float* p_cs, const float* div_cs, float dx, double* out,
                             hipStream_t s);
// Looks up the scratch array's address on the current device once (called when a topology is built).
void noalloc_setup();

}  // namespace hns
