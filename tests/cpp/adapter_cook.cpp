// One HNanoSolver cook through include/hns.hpp, written the way SOP_HNanoSolver.cpp:291-335 calls
// the reference: CreateIndexGrid -> HNS_EnsureContext -> HNS_UploadSources -> HNS_Advance ->
// HNS_DownloadForOutput -> HNS_GetActiveMask -> HNS_DestroyContext. The payload class mirrors
// HNS::GridIndexedData's interface (size, pCoords, getBlocksOfType<T>, pValues<T>).
//
//   adapter_cook errors       -> exercises the host-side error paths (no device needed)
//   adapter_cook cook IN OUT  -> two cooks of the leaf grids in IN (hns::Payload, the SOP's ingestion);
//                                writes topology build counts and the scattered output leaves to OUT
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <typeindex>
#include <vector>

struct Vec3f { float x, y, z; };
struct Coord { int32_t x, y, z; };
#define HNS_VEC3_TYPE Vec3f
#include "hns.hpp"

struct CombustionParams {  // the SOP's own copy (SOP_HNanoSolver.hpp:23-43)
  float expansionRate, temperatureRelease, buoyancyStrength, ambientTemp, vorticityScale, factorScale;
  float disturbanceEnable, disturbanceStrength, disturbanceSwirl, disturbanceThreshold, disturbanceGain,
      disturbanceFrequency;
  int substeps;
  float gravity, maskDensityMin, maskDensityMax;
  int downloadVel, simReset;
};

struct GridIndexedData {
  std::vector<Coord> coords;
  struct Block { std::string name; std::type_index type; std::vector<float> values; };
  std::vector<Block> blocks;
  size_t size() const { return coords.size(); }
  Coord* pCoords() { return coords.data(); }
  template <class T> void addValueBlock(const std::string& n) {
    blocks.push_back(Block{n, std::type_index(typeid(T)), std::vector<float>(size() * sizeof(T) / sizeof(float))});
  }
  template <class T> std::vector<std::string> getBlocksOfType() const {
    std::vector<std::string> out;
    for (auto& b : blocks) if (b.type == std::type_index(typeid(T))) out.push_back(b.name);
    return out;
  }
  template <class T> T* pValues(const std::string& n) {
    for (auto& b : blocks) if (b.name == n && b.type == std::type_index(typeid(T))) return reinterpret_cast<T*>(b.values.data());
    return nullptr;
  }
};

static int fails = 0;
#define EXPECT(c) do { if (!(c)) { std::printf("FAILED: %s (line %d)\n", #c, __LINE__); ++fails; } } while (0)

static int errors() {
  // null contexts: Upload and Advance return silently, GetActiveMask clears, Destroy is a no-op
  GridIndexedData d;
  CombustionParams p{};
  HNS_UploadSources(nullptr, d);
  HNS_Advance(nullptr, 10, 0.1f, p);
  std::vector<unsigned char> mask(5, 7);
  HNS_GetActiveMask(nullptr, mask);
  EXPECT(mask.empty());
  void* ctx = nullptr;
  HNS_DestroyContext(&ctx);
  bool threw = false;
  try { HNS_DownloadForOutput(nullptr, d); } catch (const std::runtime_error& e) { threw = std::strstr(e.what(), "null context") != nullptr; }
  EXPECT(threw);
  // empty payload: the reference throws "N==0" (HNanoSolver.cu:61)
  hns::TopologyHandle h;
  threw = false;
  try { CreateIndexGrid(d, h, 0.1f); } catch (const std::runtime_error& e) { threw = std::strstr(e.what(), "N == 0") != nullptr; }
  EXPECT(threw);
  threw = false;
  try { HNS_EnsureContext(&ctx, d, h, 0.1f, false); } catch (const std::runtime_error& e) { threw = std::strstr(e.what(), "N==0") != nullptr; }
  EXPECT(threw);
  std::printf("errors: %s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}

static void rd(void* p, size_t sz, size_t n, FILE* f) {
  if (n && std::fread(p, sz, n, f) != n) throw std::runtime_error("short read");
}
template <class T>
static T rd1(FILE* f) {
  T v;
  rd(&v, sizeof(T), 1, f);
  return v;
}
static void wr(const void* p, size_t sz, size_t n, FILE* f) {
  if (n && std::fwrite(p, sz, n, f) != n) throw std::runtime_error("short write");
}

// Leaf grids of one cook's inputs (what loadGrid hands the SOP, as leaves).
struct InGrid {
  std::string name;
  int32_t type = 0, role = 0;
  std::vector<int32_t> origins;
  std::vector<float> values;
  std::vector<uint64_t> active;
};

// Two cooks of SOP_HNanoSolver.cpp:205-396 on the same inputs, the context and topology cached
// between them as the SOP cache keeps them: payload (domain + IndexGridBuilder) -> CreateIndexGrid ->
// HNS_EnsureContext -> HNS_UploadSources -> HNS_Advance -> HNS_DownloadForOutput ->
// HNS_GetActiveMask -> scatter of every value grid. The second cook keeps the topology (same
// leaves) and advances from the state the first one left.
static int cook(const char* in, const char* out) {
  FILE* f = std::fopen(in, "rb");
  if (!f) return 2;
  const float dx = rd1<float>(f), dt = rd1<float>(f);
  const int32_t K = rd1<int32_t>(f), padding = rd1<int32_t>(f), ng = rd1<int32_t>(f), dom = rd1<int32_t>(f);
  std::vector<InGrid> in_grids(ng);
  for (auto& g : in_grids) {
    g.name.resize(rd1<int32_t>(f));
    rd(&g.name[0], 1, g.name.size(), f);
    g.type = rd1<int32_t>(f);
    g.role = rd1<int32_t>(f);
    const int64_t L = rd1<int64_t>(f);
    const int comps = g.type == HNS_VEC3F ? 3 : 1;
    g.origins.resize(3 * L);
    g.values.resize((size_t)comps * 512 * L);
    g.active.resize(8 * L);
    rd(g.origins.data(), 4, g.origins.size(), f);
    rd(g.values.data(), 4, g.values.size(), f);
    rd(g.active.data(), 8, g.active.size(), f);
  }
  std::fclose(f);
  std::vector<hns_leaf_grid> grids;
  for (auto& g : in_grids)
    grids.push_back(hns_leaf_grid{g.name.c_str(), g.type, g.role, g.origins.size() / 3, g.origins.data(),
                                  g.values.data(), g.active.data()});
  CombustionParams p{0.1f, 0.5f, 1.0f, 23.0f, 1.0f, 0.5f, 0.0f, 0, 0, 0, 0, 0, 1, 0.0f, 0.0f, 1.0f, 1, 0};
  hns::TopologyHandle handle;  // the SOP cache's grid handle and context (SOP_HNanoSolver.hpp:90-98)
  void* ctx = nullptr;
  FILE* o = std::fopen(out, "wb");
  for (int c = 0; c < 2; ++c) {
    hns::Payload data(grids, (size_t)dom, padding);
    CreateIndexGrid(data, handle, dx);
    HNS_EnsureContext(&ctx, data, handle, dx, false);
    HNS_UploadSources(ctx, data);
    HNS_Advance(ctx, K, dt, p);
    HNS_DownloadForOutput(ctx, data);
    std::vector<unsigned char> mask;
    HNS_GetActiveMask(ctx, mask);
    EXPECT(mask.size() == data.size());
    int64_t builds = 0, reuses = 0;
    hns::check(hns_topology_builds(handle.get(), &builds, &reuses));
    const int32_t pinned = data.pinned() ? 1 : 0;
    wr(&builds, 8, 1, o);
    wr(&reuses, 8, 1, o);
    wr(&pinned, 4, 1, o);
    for (auto& g : in_grids) {
      if (g.role != HNS_ROLE_VALUE) continue;
      const int comps = g.type == HNS_VEC3F ? 3 : 1;
      const hns::LeafOutput s = hns::ScatterLeaves(data, g.name, mask, 0.0f, comps);
      const int64_t L = (int64_t)s.leaves;
      wr(&L, 8, 1, o);
      wr(s.origins.data(), 4, s.origins.size(), o);
      wr(s.values.data(), 4, s.values.size(), o);
      wr(s.active.data(), 8, s.active.size(), o);
    }
  }
  HNS_DestroyContext(&ctx);
  EXPECT(ctx == nullptr);
  std::fclose(o);
  std::printf("cook: %s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::strcmp(argv[1], "errors") == 0) return errors();
    if (argc >= 4 && std::strcmp(argv[1], "cook") == 0) return cook(argv[2], argv[3]);
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 3;
  }
  std::printf("usage: adapter_cook errors | cook IN OUT\n");
  return 2;
}
