// Host ingestion: the leaf-dense payload the SOP builds from OpenVDB grids, and the scatter of
// results back to leaf grids. Replaces the domain construction of SOP_HNanoSolver.cpp:276-287 (the
// primary velocity's active topology dilated by `padding` voxels over the 26-neighbourhood, united
// with the collision SDF's topology), IndexGridBuilder::computeLeafs/build (GridBuilder.hpp:93-175,
// :231-251: every voxel of every domain leaf, leaves in tree order, each grid's leaf buffer copied
// where the grid has a leaf at the domain leaf's origin, else zeros, and for the SDF the byte
// pattern memset(..., 1, ...) leaves), the output scatter of SOP_HNanoSolver.cpp:345-388, and
// IndexGridBuilder::writeIndexGrid (GridBuilder.hpp:177-226: a grid with the domain's topology whose
// leaf buffers are the payload's).
//
// Input grids come as their leaves, the form OpenVDB and NanoVDB keep them in: per leaf the origin,
// the 512-value buffer in offset order (x << 6 | y << 3 | z) and the 512-bit active mask. Tiles are
// not represented: the reference's domain ignores them too (dilateVoxels with IGNORE_TILES, and the
// LeafManager visits leaves only).
//
// The payload lives in pinned host memory when a HIP device is visible (so the context's uploads
// are DMA transfers, not staged copies of pageable memory) and in ordinary memory otherwise.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "hns_internal.hpp"

namespace hns {
namespace {

constexpr int kWordsPerLeaf = kLeafVox / 64;

struct HostBlock {
  void* p = nullptr;
  bool pinned = false;
  size_t bytes = 0;
  HostBlock() = default;
  HostBlock(const HostBlock&) = delete;
  HostBlock& operator=(const HostBlock&) = delete;
  HostBlock(HostBlock&& o) noexcept : p(o.p), pinned(o.pinned), bytes(o.bytes) { o.p = nullptr; }
  void alloc(size_t n, bool pin) {
    bytes = n;
    if (pin && hipHostMalloc(&p, n, hipHostMallocDefault) == hipSuccess) {
      pinned = true;
      return;
    }
    (void)hipGetLastError();
    p = std::malloc(n ? n : 1);
    if (!p) throw std::bad_alloc();
  }
  ~HostBlock() {
    if (!p) return;
    if (pinned) (void)hipHostFree(p);
    else std::free(p);
  }
};

bool device_visible() {
  int n = 0;
  const bool ok = hipGetDeviceCount(&n) == hipSuccess && n > 0;
  (void)hipGetLastError();
  return ok;
}

uint64_t key_of(const int32_t* o) { return leaf_key(o[0] >> 3, o[1] >> 3, o[2] >> 3); }

// Leaves of a domain grid touched by its active voxels dilated by p (L-infinity, the
// 26-neighbourhood applied p times). Per axis, a voxel at local coordinate x reaches leaf offset d
// (in leaves) iff [x - p, x + p] meets [8 d, 8 d + 7]; a leaf offset is touched iff some active voxel
// lies in the box those per-axis conditions cut out of the leaf, tested as a 512-bit AND.
void dilated_leaves(const hns_leaf_grid& g, int32_t p, std::unordered_set<uint64_t>& out,
                    std::vector<std::array<int32_t, 3>>& origins) {
  const int D = (p + 7) / 8;  // leaf radius
  const int W = 2 * D + 1;
  std::vector<std::array<uint64_t, kWordsPerLeaf>> box(W * W * W);
  std::vector<bool> empty(W * W * W, false);
  for (int a = 0; a < W; ++a)
    for (int b = 0; b < W; ++b)
      for (int c = 0; c < W; ++c) {
        const int d[3] = {a - D, b - D, c - D};
        int lo[3], hi[3];
        bool none = false;
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::max(0, 8 * d[k] - p);
          hi[k] = std::min(7, 8 * d[k] + 7 + p);
          none |= lo[k] > hi[k];
        }
        auto& m = box[(a * W + b) * W + c];
        m.fill(0);
        empty[(a * W + b) * W + c] = none;
        if (none) continue;
        for (int x = lo[0]; x <= hi[0]; ++x)
          for (int y = lo[1]; y <= hi[1]; ++y)
            for (int z = lo[2]; z <= hi[2]; ++z) {
              const int o = (x << 6) | (y << 3) | z;
              m[o >> 6] |= uint64_t(1) << (o & 63);
            }
      }
  for (size_t l = 0; l < g.num_leaves; ++l) {
    const int32_t* org = g.origins + 3 * l;
    const uint64_t* act = g.active ? g.active + (size_t)kWordsPerLeaf * l : nullptr;
    for (int a = 0; a < W; ++a)
      for (int b = 0; b < W; ++b)
        for (int c = 0; c < W; ++c) {
          const int i = (a * W + b) * W + c;
          if (empty[i]) continue;
          bool hit = false;
          for (int w = 0; w < kWordsPerLeaf && !hit; ++w) hit = (box[i][w] & (act ? act[w] : ~uint64_t(0))) != 0;
          if (!hit) continue;
          const std::array<int32_t, 3> o = {org[0] + 8 * (a - D), org[1] + 8 * (b - D), org[2] + 8 * (c - D)};
          if (!leaf_in_range(o[0] >> 3, o[1] >> 3, o[2] >> 3))
            fail(HNS_ERR_TOPOLOGY, "hns_payload_build: the dilated domain leaves [-2^23, 2^23)");
          if (out.insert(key_of(o.data())).second) origins.push_back(o);
        }
  }
}

// The domain's active voxels as 512-bit leaf masks (word x, bit y << 3 | z): the active voxels of
// `g` dilated p times over the 26-neighbourhood (dilateVoxels with NN_FACE_EDGE_VERTEX,
// SOP_HNanoSolver.cpp:280-287), each time one voxel along x, y and z in turn (the 26-neighbourhood
// is the product of the three 3-voxel lines), with what leaves a leaf carried into its neighbour.
struct VoxMasks {
  struct Leaf {
    std::array<int32_t, 3> org;
    std::array<uint64_t, kWordsPerLeaf> m;
  };
  std::unordered_map<uint64_t, Leaf> at;
  void add(const std::array<int32_t, 3>& o, const std::array<uint64_t, kWordsPerLeaf>& m) {
    bool any = false;
    for (uint64_t w : m) any |= w != 0;
    if (!any) return;
    auto it = at.try_emplace(key_of(o.data()), Leaf{o, {}}).first;
    for (int w = 0; w < kWordsPerLeaf; ++w) it->second.m[w] |= m[w];
  }
  // one voxel along `axis` in both directions
  void dilate_axis(int axis) {
    constexpr uint64_t kZ0 = 0x0101010101010101ull, kZ7 = 0x8080808080808080ull;
    std::vector<std::pair<std::array<int32_t, 3>, std::array<uint64_t, kWordsPerLeaf>>> spill;
    for (auto& kv : at) {
      std::array<uint64_t, kWordsPerLeaf>& m = kv.second.m;
      const std::array<uint64_t, kWordsPerLeaf> src = m;
      std::array<uint64_t, kWordsPerLeaf> up{}, down{};  // carried to the +axis and -axis leaves
      for (int x = 0; x < kWordsPerLeaf; ++x) {
        const uint64_t w = src[x];
        if (axis == 2) {
          m[x] |= ((w << 1) & ~kZ0) | ((w >> 1) & ~kZ7);
          up[x] = (w & kZ7) >> 7;
          down[x] = (w & kZ0) << 7;
        } else if (axis == 1) {
          m[x] |= (w << 8) | (w >> 8);
          up[x] = w >> 56;
          down[x] = w << 56;
        } else {
          if (x + 1 < kWordsPerLeaf) m[x + 1] |= w;
          if (x > 0) m[x - 1] |= w;
        }
      }
      if (axis == 0) {
        up[0] = src[kWordsPerLeaf - 1];
        down[kWordsPerLeaf - 1] = src[0];
      }
      std::array<int32_t, 3> o = kv.second.org;
      o[axis] += 8;
      spill.emplace_back(o, up);
      o[axis] -= 16;
      spill.emplace_back(o, down);
    }
    for (auto& sp : spill) {
      bool any = false;
      for (uint64_t w : sp.second) any |= w != 0;
      if (!any) continue;
      if (!leaf_in_range(sp.first[0] >> 3, sp.first[1] >> 3, sp.first[2] >> 3))
        fail(HNS_ERR_TOPOLOGY, "hns_payload_build: the dilated domain leaves [-2^23, 2^23)");
      add(sp.first, sp.second);
    }
  }
  void add_grid(const hns_leaf_grid& g) {
    for (size_t l = 0; l < g.num_leaves; ++l) {
      std::array<uint64_t, kWordsPerLeaf> m;
      for (int w = 0; w < kWordsPerLeaf; ++w) m[w] = g.active ? g.active[(size_t)kWordsPerLeaf * l + w] : ~uint64_t(0);
      add({g.origins[3 * l], g.origins[3 * l + 1], g.origins[3 * l + 2]}, m);
    }
  }
};

void check_grid(const hns_leaf_grid& g, size_t i) {
  const std::string tag = "hns_payload_build: grid " + std::to_string(i);
  if (!g.name) fail(HNS_ERR_INVALID_ARGUMENT, tag + " has no name");
  if (g.type != HNS_FLOAT && g.type != HNS_VEC3F) fail(HNS_ERR_INVALID_ARGUMENT, tag + " has an unknown type");
  if (g.role != HNS_ROLE_VALUE && g.role != HNS_ROLE_SDF) fail(HNS_ERR_INVALID_ARGUMENT, tag + " has an unknown role");
  if (g.role == HNS_ROLE_SDF && g.type != HNS_FLOAT) fail(HNS_ERR_INVALID_ARGUMENT, tag + ": an SDF is a float grid");
  if (g.num_leaves && (!g.origins || !g.values)) fail(HNS_ERR_INVALID_ARGUMENT, tag + " has null leaf arrays");
  for (size_t l = 0; l < g.num_leaves; ++l) {
    const int32_t* o = g.origins + 3 * l;
    if ((o[0] & 7) | (o[1] & 7) | (o[2] & 7)) fail(HNS_ERR_TOPOLOGY, tag + ": a leaf origin is not a multiple of 8");
    if (!leaf_in_range(o[0] >> 3, o[1] >> 3, o[2] >> 3))
      fail(HNS_ERR_TOPOLOGY, tag + ": a leaf origin lies outside [-2^23, 2^23)");
  }
}

// Runs f(begin, end) over [0, n) on up to hardware_concurrency threads.
template <class F>
void parallel_ranges(size_t n, F f) {
  const size_t T = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), 16));
  if (n < 64 || T == 1) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> ts;
  const size_t chunk = (n + T - 1) / T;
  for (size_t t = 0; t < T; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back([=] { f(b, e); });
  }
  for (auto& t : ts) t.join();
}

}  // namespace

struct Payload {
  size_t N = 0, L = 0;
  bool pinned = false;
  HostBlock coords;
  std::vector<std::string> names;
  std::vector<HostBlock> values;
  std::vector<hns_value_block> blocks;
  std::vector<uint64_t> domain;  // the domain's active voxels, 8 words per leaf
};

Payload* build_payload(const hns_leaf_grid* grids, size_t num_grids, size_t domain_grid, int32_t padding) {
  if (!grids || num_grids == 0) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_build: no grids");
  if (domain_grid >= num_grids) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_build: domain grid index out of range");
  if (padding < 0 || padding > 64) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_build: padding outside [0, 64]");
  for (size_t i = 0; i < num_grids; ++i) check_grid(grids[i], i);
  if (grids[domain_grid].role != HNS_ROLE_VALUE)
    fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_build: the domain grid must be a value grid");

  // domain leaves: the dilated active topology of the domain grid, united with the SDF's active topology
  std::unordered_set<uint64_t> keys;
  std::vector<std::array<int32_t, 3>> org;
  dilated_leaves(grids[domain_grid], padding, keys, org);
  for (size_t i = 0; i < num_grids; ++i) {
    if (grids[i].role != HNS_ROLE_SDF) continue;
    dilated_leaves(grids[i], 0, keys, org);
    break;  // the SOP unites the first collision grid only (SOP_HNanoSolver.cpp:262-268)
  }
  const size_t L = org.size();
  if (L == 0) fail(HNS_ERR_EMPTY, "hns_payload_build: the domain has no active voxels");
  if (L * kLeafVox > (size_t)INT32_MAX) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_build: N must be below 2^31");
  // tree order (voxelsToGrid's, PointsToGrid.cuh:596-651; OpenVDB's LeafManager order)
  std::vector<uint64_t> perm(L);
  leaf_order_host(reinterpret_cast<const int32_t*>(org.data()), L, perm.data());
  std::vector<std::array<int32_t, 3>> leaves(L);
  for (size_t r = 0; r < L; ++r) leaves[r] = org[perm[r]];

  auto P = std::make_unique<Payload>();
  P->L = L;
  P->N = L * kLeafVox;
  const bool pin = device_visible();
  P->coords.alloc(3 * P->N * sizeof(int32_t), pin);
  P->pinned = P->coords.pinned;
  int32_t* C = static_cast<int32_t*>(P->coords.p);
  parallel_ranges(L, [&](size_t b, size_t e) {
    for (size_t l = b; l < e; ++l)
      for (int o = 0; o < kLeafVox; ++o) {
        int32_t* c = C + 3 * (l * kLeafVox + o);
        c[0] = leaves[l][0] + (o >> 6);
        c[1] = leaves[l][1] + ((o >> 3) & 7);
        c[2] = leaves[l][2] + (o & 7);
      }
  });

  // the domain's active voxels per leaf (the topology writeIndexGrid gives its grids)
  {
    VoxMasks vm;
    vm.add_grid(grids[domain_grid]);
    for (int32_t k = 0; k < padding; ++k)
      for (int axis = 0; axis < 3; ++axis) vm.dilate_axis(axis);
    for (size_t i = 0; i < num_grids; ++i) {
      if (grids[i].role != HNS_ROLE_SDF) continue;
      vm.add_grid(grids[i]);
      break;
    }
    if (vm.at.size() != L) fail(HNS_ERR_TOPOLOGY, "hns_payload_build: voxel and leaf dilations disagree");
    P->domain.resize(L * kWordsPerLeaf);
    for (size_t l = 0; l < L; ++l) {
      auto it = vm.at.find(key_of(leaves[l].data()));
      if (it == vm.at.end()) fail(HNS_ERR_TOPOLOGY, "hns_payload_build: voxel and leaf dilations disagree");
      std::memcpy(P->domain.data() + l * kWordsPerLeaf, it->second.m.data(), kWordsPerLeaf * sizeof(uint64_t));
    }
  }

  // value blocks in the grids' order; a repeated name keeps the first grid (GridData.hpp:50-53)
  std::vector<size_t> used;
  for (size_t i = 0; i < num_grids; ++i) {
    bool dup = false;
    for (size_t j : used) dup |= std::strcmp(grids[j].name, grids[i].name) == 0;
    if (!dup) used.push_back(i);
  }
  P->names.reserve(used.size());
  P->values.resize(used.size());
  for (size_t b = 0; b < used.size(); ++b) {
    const hns_leaf_grid& g = grids[used[b]];
    const int comps = g.type == HNS_VEC3F ? 3 : 1;
    const size_t leaf_floats = (size_t)comps * kLeafVox;
    P->names.emplace_back(g.name);
    P->values[b].alloc(P->N * comps * sizeof(float), pin);
    P->pinned = P->pinned && P->values[b].pinned;
    std::unordered_map<uint64_t, size_t> at;
    at.reserve(g.num_leaves * 2);
    for (size_t l = 0; l < g.num_leaves; ++l)
      if (!at.emplace(key_of(g.origins + 3 * l), l).second)
        fail(HNS_ERR_TOPOLOGY, "hns_payload_build: grid '" + P->names.back() + "' repeats a leaf");
    float* out = static_cast<float*>(P->values[b].p);
    const int fill = g.role == HNS_ROLE_SDF ? 1 : 0;  // GridBuilder.hpp:116 (SDF) and :136, :157
    parallel_ranges(L, [&](size_t lb, size_t le) {
      for (size_t l = lb; l < le; ++l) {
        float* dst = out + l * leaf_floats;
        auto it = at.find(key_of(leaves[l].data()));
        if (it != at.end()) std::memcpy(dst, g.values + it->second * leaf_floats, leaf_floats * sizeof(float));
        else std::memset(dst, fill, leaf_floats * sizeof(float));
      }
    });
  }
  P->blocks.resize(used.size());
  for (size_t b = 0; b < used.size(); ++b)
    P->blocks[b] = hns_value_block{P->names[b].c_str(), grids[used[b]].type, P->values[b].p};
  return P.release();
}

void destroy_payload(Payload* p) { delete p; }

void payload_domain_mask(Payload* P, uint64_t* active, size_t num_leaves) {
  if (!active) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_domain_mask: null output");
  if (num_leaves != P->L) fail(HNS_ERR_INVALID_ARGUMENT, "hns_payload_domain_mask: num_leaves is not the payload's");
  std::memcpy(active, P->domain.data(), P->domain.size() * sizeof(uint64_t));
}

void payload_view(Payload* P, hns_grid_data* v, int32_t* pinned) {
  v->size = P->N;
  v->coords = static_cast<const int32_t*>(P->coords.p);
  v->num_blocks = P->blocks.size();
  v->blocks = P->blocks.data();
  if (pinned) *pinned = P->pinned ? 1 : 0;
}

// The output scatter (SOP_HNanoSolver.cpp:345-388): a grid of the payload's field with the voxels
// of the mask set on (setValueOn) and every other voxel at the background, as leaves: a payload
// leaf with at least one masked voxel becomes an output leaf (origin, 512 values, active bits).
void scatter_leaves(const hns_grid_data* d, const char* name, const uint8_t* mask, float background, int32_t* origins,
                    float* values, uint64_t* active, size_t capacity, size_t* num_leaves) {
  if (!d || !d->coords || !name || !num_leaves) fail(HNS_ERR_INVALID_ARGUMENT, "hns_scatter_leaves: null argument");
  if (d->size % kLeafVox) fail(HNS_ERR_TOPOLOGY, "hns_scatter_leaves: payload is not leaf-dense");
  const hns_value_block* blk = nullptr;
  for (size_t i = 0; i < d->num_blocks && !blk; ++i)
    if (d->blocks[i].name && std::strcmp(d->blocks[i].name, name) == 0) blk = &d->blocks[i];
  if (!blk || !blk->values) fail(HNS_ERR_NOT_FOUND, std::string("hns_scatter_leaves: no block '") + name + "'");
  const int comps = blk->type == HNS_VEC3F ? 3 : 1;
  const float* src = static_cast<const float*>(blk->values);
  const size_t L = d->size / kLeafVox;
  size_t n = 0;
  for (size_t l = 0; l < L; ++l) {
    bool any = !mask;
    for (int o = 0; o < kLeafVox && !any; ++o) any = mask[l * kLeafVox + o] != 0;
    if (!any) continue;
    if (n < capacity) {
      const int32_t* c = d->coords + 3 * l * kLeafVox;
      if (origins) std::memcpy(origins + 3 * n, c, 3 * sizeof(int32_t));
      for (int o = 0; o < kLeafVox; ++o) {
        const bool on = !mask || mask[l * kLeafVox + o];
        if (values)
          for (int k = 0; k < comps; ++k)
            values[(n * kLeafVox + o) * comps + k] = on ? src[(l * kLeafVox + o) * comps + k] : background;
        if (active) {
          uint64_t& w = active[n * (kLeafVox / 64) + (o >> 6)];
          if ((o & 63) == 0) w = 0;
          if (on) w |= uint64_t(1) << (o & 63);
        }
      }
    }
    ++n;
  }
  *num_leaves = n;
  if (n > capacity && (origins || values || active))
    fail(HNS_ERR_INVALID_ARGUMENT, "hns_scatter_leaves: capacity below the number of output leaves");
}

// writeIndexGrid (GridBuilder.hpp:177-226): the payload field `name` as a grid of every payload
// leaf, its buffer the payload's 512 values (memcpy, active or not) and its active voxels the
// domain's (topologyUnion with the domain grid; null: every voxel).
void write_leaf_grid(const hns_grid_data* d, const char* name, const uint64_t* domain_active, int32_t* origins,
                     float* values, uint64_t* active) {
  if (!d || !d->coords || !name) fail(HNS_ERR_INVALID_ARGUMENT, "hns_write_leaf_grid: null argument");
  if (d->size % kLeafVox) fail(HNS_ERR_TOPOLOGY, "hns_write_leaf_grid: payload is not leaf-dense");
  const hns_value_block* blk = nullptr;
  for (size_t i = 0; i < d->num_blocks && !blk; ++i)
    if (d->blocks[i].name && std::strcmp(d->blocks[i].name, name) == 0) blk = &d->blocks[i];
  if (!blk || !blk->values) fail(HNS_ERR_NOT_FOUND, std::string("hns_write_leaf_grid: no block '") + name + "'");
  const int comps = blk->type == HNS_VEC3F ? 3 : 1;
  const float* src = static_cast<const float*>(blk->values);
  const size_t L = d->size / kLeafVox;
  parallel_ranges(L, [&](size_t b, size_t e) {
    for (size_t l = b; l < e; ++l) {
      if (origins) std::memcpy(origins + 3 * l, d->coords + 3 * l * kLeafVox, 3 * sizeof(int32_t));
      if (values)
        std::memcpy(values + l * kLeafVox * comps, src + l * kLeafVox * comps, kLeafVox * comps * sizeof(float));
      if (active)
        for (int w = 0; w < kWordsPerLeaf; ++w)
          active[l * kWordsPerLeaf + w] = domain_active ? domain_active[l * kWordsPerLeaf + w] : ~uint64_t(0);
    }
  });
}

}  // namespace hns
