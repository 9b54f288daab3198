// hns.hpp — header-only C++ adapter that restores the reference's C++ launch API on top of the
// C ABI in hns.h, so SOP source written against src/Cuda/HNanoSolver.cu keeps compiling:
//
//   reference (extern "C", C++ references, throws)           this adapter
//   CreateIndexGrid(data, GridHandle&, voxelSize)            CreateIndexGrid(data, hns::TopologyHandle&, voxelSize)
//   HNS_EnsureContext(void**, data, handle, vs, hasColl)     HNS_EnsureContext(void**, data, handle, vs, hasColl)
//   HNS_UploadSources(void*, data)                           HNS_UploadSources(void*, data)
//   HNS_Advance(void*, int iteration, float dt, params)      HNS_Advance(void*, int iteration, float dt, params)
//   HNS_DownloadForOutput(void*, data)                       HNS_DownloadForOutput(void*, data)
//   HNS_GetActiveMask(void*, std::vector<unsigned char>&)    HNS_GetActiveMask(void*, std::vector<unsigned char>&)
//   HNS_DestroyContext(void**)                               HNS_DestroyContext(void**)
//   AdvectIndexGrid / AdvectIndexGridVelocity /              same names; the stream argument is any
//   Divergence / ProjectNonDivergent (stream&)               pointer-sized stream handle (or nullptr)
//   IndexGridBuilder(domain, &data) + addGrid + build        hns::Payload(leaf grids, domain index, padding)
//   scatterScalar / scatterVector (the SOP's output)         hns::ScatterLeaves(data, name, mask, bg, comps)
//   builder.writeIndexGrid<GridT>(name, voxelSize)           hns::WriteLeafGrid(payload, name, comps)
//
// The payload type is any class shaped like HNS::GridIndexedData (src/Utils/GridData.hpp:16-166):
// size(), pCoords() (3 int32 per voxel), getBlocksOfType<T>() and pValues<T>(name). Its Vec3 type
// is HNS_VEC3_TYPE (define it before including, e.g. openvdb::Vec3f; 3 packed floats). The
// GridHandle<cuda::DeviceBuffer> becomes hns::TopologyHandle. CombustionParams may be the SOP's
// own struct: it is copied into hns_params after a size check (72 bytes, Kernels.cuh:5-24).
// Failures throw std::runtime_error, as the reference's CUDA_CHECK does (Utils.cuh:10-18).
#ifndef HNS_HPP_
#define HNS_HPP_

#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "hns.h"

#ifndef HNS_VEC3_TYPE
#error "define HNS_VEC3_TYPE (the payload's Vec3f type, e.g. openvdb::Vec3f) before including hns.hpp"
#endif

namespace hns {

using Vec3 = HNS_VEC3_TYPE;
static_assert(sizeof(Vec3) == 3 * sizeof(float), "HNS_VEC3_TYPE must be three packed floats");

inline void check(hns_status s) {
  if (s != HNS_OK) throw std::runtime_error(std::string(hns_last_error()));
}

// Owns a device topology (the role of nanovdb::GridHandle<nanovdb::cuda::DeviceBuffer>).
class TopologyHandle {
 public:
  TopologyHandle() = default;
  TopologyHandle(const TopologyHandle&) = delete;
  TopologyHandle& operator=(const TopologyHandle&) = delete;
  TopologyHandle(TopologyHandle&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  TopologyHandle& operator=(TopologyHandle&& o) noexcept {
    if (this != &o) {
      reset();
      h_ = o.h_;
      o.h_ = nullptr;
    }
    return *this;
  }
  ~TopologyHandle() { reset(); }
  void reset() {
    if (h_) hns_topology_destroy(&h_);
  }
  hns_topology* get() const { return h_; }
  hns_topology** out() { return &h_; }
  explicit operator bool() const { return h_ != nullptr; }

 private:
  hns_topology* h_ = nullptr;
};

// Flat C view of a GridIndexedData-shaped payload; valid while `data` is alive and unchanged.
template <class GridData>
struct PayloadView {
  std::vector<std::string> names;
  std::vector<hns_value_block> blocks;
  hns_grid_data view{};

  explicit PayloadView(GridData& data) {
    const std::vector<std::string> f = data.template getBlocksOfType<float>();
    const std::vector<std::string> v = data.template getBlocksOfType<Vec3>();
    names.reserve(f.size() + v.size());
    // insertion order is kept per type; the C side only relies on the Vec3 order
    // (velocity = "velocity" else the first Vec3 block, HNanoSolver.cu:99-110)
    for (const auto& n : f) names.push_back(n);
    for (const auto& n : v) names.push_back(n);
    for (size_t i = 0; i < names.size(); ++i) {
      hns_value_block b;
      b.name = names[i].c_str();
      if (i < f.size()) {
        b.type = HNS_FLOAT;
        b.values = static_cast<void*>(data.template pValues<float>(names[i]));
      } else {
        b.type = HNS_VEC3F;
        b.values = static_cast<void*>(data.template pValues<Vec3>(names[i]));
      }
      blocks.push_back(b);
    }
    view.size = data.size();
    view.coords = reinterpret_cast<const int32_t*>(data.pCoords());
    view.num_blocks = blocks.size();
    view.blocks = blocks.empty() ? nullptr : blocks.data();
  }
  hns_grid_data* get() { return &view; }
};

// The leaf-dense payload of one cook built by hns_payload_build (the SOP's domain + IndexGridBuilder,
// SOP_HNanoSolver.cpp:276-301), owned, with HNS::GridIndexedData's interface (size, pCoords,
// getBlocksOfType<T>, pValues<T>) so the calls below take it like the reference's payload.
class Payload {
 public:
  Payload(const std::vector<hns_leaf_grid>& grids, size_t domain_grid, int padding) {
    check(hns_payload_build(grids.data(), grids.size(), domain_grid, padding, &p_));
    int32_t pinned = 0;
    check(hns_payload_view(p_, &view_, &pinned));
    pinned_ = pinned != 0;
  }
  Payload(const Payload&) = delete;
  Payload& operator=(const Payload&) = delete;
  ~Payload() { hns_payload_destroy(&p_); }
  size_t size() const { return view_.size; }
  const int32_t* pCoords() const { return view_.coords; }
  bool pinned() const { return pinned_; }
  const hns_grid_data& view() const { return view_; }
  // the domain's active voxels, 8 words per leaf (hns_payload_domain_mask)
  std::vector<uint64_t> domainMask() const {
    std::vector<uint64_t> m(view_.size / 512 * 8);
    check(hns_payload_domain_mask(p_, m.data(), view_.size / 512));
    return m;
  }
  template <class T>
  std::vector<std::string> getBlocksOfType() const {
    std::vector<std::string> out;
    for (size_t i = 0; i < view_.num_blocks; ++i)
      if (view_.blocks[i].type == type_code<T>()) out.push_back(view_.blocks[i].name);
    return out;
  }
  template <class T>
  T* pValues(const std::string& name) {
    for (size_t i = 0; i < view_.num_blocks; ++i)
      if (view_.blocks[i].type == type_code<T>() && name == view_.blocks[i].name)
        return static_cast<T*>(view_.blocks[i].values);
    return nullptr;
  }

 private:
  template <class T>
  static int32_t type_code() {
    return sizeof(T) == sizeof(Vec3) && !std::is_same<T, float>::value ? HNS_VEC3F : HNS_FLOAT;
  }
  hns_payload* p_ = nullptr;
  hns_grid_data view_{};
  bool pinned_ = false;
};

// The output scatter of SOP_HNanoSolver.cpp:345-388 as leaves: origins, values, active masks.
struct LeafOutput {
  std::vector<int32_t> origins;
  std::vector<float> values;
  std::vector<uint64_t> active;
  size_t leaves = 0;
};
template <class GridData>
LeafOutput ScatterLeaves(GridData& data, const std::string& name, const std::vector<unsigned char>& mask,
                         float background, int comps) {
  PayloadView<GridData> v(data);
  LeafOutput o;
  const uint8_t* m = mask.empty() ? nullptr : mask.data();
  check(hns_scatter_leaves(v.get(), name.c_str(), m, background, nullptr, nullptr, nullptr, 0, &o.leaves));
  o.origins.resize(3 * o.leaves);
  o.values.resize((size_t)comps * 512 * o.leaves);
  o.active.resize(8 * o.leaves);
  check(hns_scatter_leaves(v.get(), name.c_str(), m, background, o.origins.data(), o.values.data(), o.active.data(),
                           o.leaves, &o.leaves));
  return o;
}

// IndexGridBuilder::writeIndexGrid (GridBuilder.hpp:177-226) as leaves: every payload leaf with
// the payload's buffer of `name` and the domain's active voxels
inline LeafOutput WriteLeafGrid(const Payload& payload, const std::string& name, int comps) {
  LeafOutput o;
  o.leaves = payload.size() / 512;
  const std::vector<uint64_t> dm = payload.domainMask();
  o.origins.resize(3 * o.leaves);
  o.values.resize((size_t)comps * 512 * o.leaves);
  o.active.resize(8 * o.leaves);
  check(hns_write_leaf_grid(&payload.view(), name.c_str(), dm.data(), o.origins.data(), o.values.data(),
                            o.active.data()));
  return o;
}

template <class P>
inline hns_params to_params(const P& p) {
  static_assert(sizeof(P) == sizeof(hns_params), "CombustionParams must have the 72-byte layout of Kernels.cuh:5-24");
  hns_params out;
  std::memcpy(&out, &p, sizeof(out));
  return out;
}

}  // namespace hns

// ---- the reference's entry points ------------------------------------------------------------

template <class GridData>
inline void CreateIndexGrid(GridData& data, hns::TopologyHandle& handle, const float voxelSize) {
  hns::PayloadView<GridData> v(data);
  hns::check(hns_create_index_grid(v.get(), voxelSize, handle.out()));
}

template <class GridData>
inline void HNS_EnsureContext(void** ctx, GridData& data, const hns::TopologyHandle& handle, float voxelSize,
                              bool hasCollision) {
  hns::PayloadView<GridData> v(data);
  hns::check(hns_ensure_context(reinterpret_cast<hns_context**>(ctx), v.get(), handle.get(), voxelSize,
                                hasCollision ? 1 : 0));
}

// The reference returns silently on a null context (HNanoSolver.cu:160).
template <class GridData>
inline void HNS_UploadSources(void* ctx, GridData& data) {
  if (!ctx) return;
  hns::PayloadView<GridData> v(data);
  hns::check(hns_upload_sources(static_cast<hns_context*>(ctx), v.get()));
}

// The reference returns silently on a null context (HNanoSolver.cu:225).
template <class Params>
inline void HNS_Advance(void* ctx, int iteration, float dt, const Params& params) {
  if (!ctx) return;
  const hns_params p = hns::to_params(params);
  hns::check(hns_advance(static_cast<hns_context*>(ctx), iteration, dt, &p));
}

// The reference dereferences a null context here (HNanoSolver.cu:472-473); this throws instead.
template <class GridData>
inline void HNS_DownloadForOutput(void* ctx, GridData& data) {
  hns::PayloadView<GridData> v(data);
  hns::check(hns_download_for_output(static_cast<hns_context*>(ctx), v.get()));
}

inline void HNS_GetActiveMask(void* ctx, std::vector<unsigned char>& hostMask) {
  size_t n = 0;
  hns::check(hns_get_active_mask(static_cast<hns_context*>(ctx), nullptr, 0, &n));
  hostMask.resize(n);
  if (n) hns::check(hns_get_active_mask(static_cast<hns_context*>(ctx), hostMask.data(), n, &n));
}

inline void HNS_DestroyContext(void** ctx) { hns::check(hns_destroy_context(reinterpret_cast<hns_context**>(ctx))); }

template <class GridData, class Stream>
inline void AdvectIndexGrid(GridData& data, const float dt, const float voxelSize, const Stream& stream) {
  hns::PayloadView<GridData> v(data);
  hns::check(hns_advect_index_grid(v.get(), dt, voxelSize, reinterpret_cast<void*>(stream)));
}

template <class GridData, class Stream>
inline void AdvectIndexGridVelocity(GridData& data, const float dt, const float voxelSize, const Stream& stream) {
  hns::PayloadView<GridData> v(data);
  hns::check(hns_advect_index_grid_velocity(v.get(), dt, voxelSize, reinterpret_cast<void*>(stream)));
}

template <class GridData, class Stream>
inline void Divergence(GridData& data, const float voxelSize, const Stream& stream) {
  hns::PayloadView<GridData> v(data);
  hns::check(hns_divergence(v.get(), voxelSize, reinterpret_cast<void*>(stream)));
}

template <class GridData, class Stream>
inline void ProjectNonDivergent(GridData& data, const size_t iterations, const float voxelSize, const Stream& stream) {
  hns::PayloadView<GridData> v(data);
  hns::check(hns_project_non_divergent(v.get(), iterations, voxelSize, reinterpret_cast<void*>(stream)));
}

#endif  // HNS_HPP_
