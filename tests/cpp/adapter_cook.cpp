// One HNanoSolver cook through include/hns.hpp, written the way SOP_HNanoSolver.cpp:291-335 calls
// the reference: CreateIndexGrid -> HNS_EnsureContext -> HNS_UploadSources -> HNS_Advance ->
// HNS_DownloadForOutput -> HNS_GetActiveMask -> HNS_DestroyContext. The payload class mirrors
// HNS::GridIndexedData's interface (size, pCoords, getBlocksOfType<T>, pValues<T>).
//
//   adapter_cook errors            -> exercises the host-side error paths (no device needed)
//   adapter_cook cook IN OUT K DT  -> reads coords/density/velocity from IN, writes results to OUT
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
  if (std::fread(p, sz, n, f) != n) throw std::runtime_error("short read");
}

static int cook(const char* in, const char* out, int K, float dt) {
  FILE* f = std::fopen(in, "rb");
  if (!f) return 2;
  int64_t n = 0;
  float dx = 0;
  rd(&n, 8, 1, f);
  rd(&dx, 4, 1, f);
  GridIndexedData d;
  d.coords.resize(n);
  rd(d.coords.data(), sizeof(Coord), n, f);
  d.addValueBlock<float>("density");
  d.addValueBlock<Vec3f>("velocity");
  rd(d.pValues<float>("density"), 4, n, f);
  rd(d.pValues<Vec3f>("velocity"), 12, n, f);
  std::fclose(f);
  CombustionParams p{0.1f, 0.5f, 1.0f, 23.0f, 1.0f, 0.5f, 0.0f, 0, 0, 0, 0, 0, 1, 0.0f, 0.0f, 1.0f, 1, 0};
  hns::TopologyHandle handle;
  void* ctx = nullptr;
  CreateIndexGrid(d, handle, dx);
  HNS_EnsureContext(&ctx, d, handle, dx, false);
  HNS_UploadSources(ctx, d);
  HNS_Advance(ctx, K, dt, p);
  HNS_DownloadForOutput(ctx, d);
  std::vector<unsigned char> mask;
  HNS_GetActiveMask(ctx, mask);
  EXPECT(mask.size() == (size_t)n);
  HNS_DestroyContext(&ctx);
  EXPECT(ctx == nullptr);
  f = std::fopen(out, "wb");
  std::fwrite(d.pValues<float>("density"), 4, n, f);
  std::fwrite(d.pValues<Vec3f>("velocity"), 12, n, f);
  std::fclose(f);
  std::printf("cook: %s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::strcmp(argv[1], "errors") == 0) return errors();
    if (argc >= 6 && std::strcmp(argv[1], "cook") == 0) return cook(argv[2], argv[3], std::atoi(argv[4]), std::atof(argv[5]));
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 3;
  }
  std::printf("usage: adapter_cook errors | cook IN OUT K DT\n");
  return 2;
}
