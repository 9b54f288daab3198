// TEST INFRASTRUCTURE — checker built against the reference's vendored NanoVDB (header-only,
// externals/nanovdb in the reference repository; NanoVDB v32.7, NanoVDB.h:144-146), never shipped
// in the product. It builds a NanoVDB ValueOnIndex grid from voxel coordinates with NanoVDB's own
// CPU builder (tools::build::Grid + tools::createNanoGrid), the same grid type the reference's
// CreateIndexGrid makes on the device with voxelsToGrid (src/Cuda/HNanoSolver.cu:498-505), and
// writes what the reference's kernels read from it: the index of every input voxel (getValue of
// the IndexGrid, NanoVDB.h:4220-4228; 0 = not active) and the leaf origins in the grid's leaf order.
//
//   nanovdb_index IN OUT   IN: int64 M, then M x (int32 x, y, z)
//                          OUT: int64 L, L x (int32 x, y, z) leaf origins in order, then M x uint64 index
#include <nanovdb/NanoVDB.h>
#include <nanovdb/tools/CreateNanoGrid.h>
#include <nanovdb/tools/GridBuilder.h>

#include <cstdint>
#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: nanovdb_index IN OUT\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t m = 0;
  if (std::fread(&m, 8, 1, f) != 1) return 2;
  std::vector<int32_t> c(3 * (size_t)m);
  if (m && std::fread(c.data(), 4, c.size(), f) != c.size()) return 2;
  std::fclose(f);

  nanovdb::tools::build::Grid<float> src(0.0f);
  auto acc = src.getAccessor();
  for (int64_t i = 0; i < m; ++i) acc.setValue(nanovdb::Coord(c[3 * i], c[3 * i + 1], c[3 * i + 2]), 1.0f);
  // indices as voxelsToGrid assigns them for ValueOnIndex (PointsToGrid.cuh:461-471): background 0, then the
  // active voxels leaf by leaf, no statistics and no tile values
  auto handle = nanovdb::tools::createNanoGrid<nanovdb::tools::build::Grid<float>, nanovdb::ValueOnIndex>(
      src, 0u, /*includeStats=*/false, /*includeTiles=*/false);
  const auto* grid = handle.grid<nanovdb::ValueOnIndex>();
  if (!grid) return 3;
  const auto& tree = grid->tree();
  const uint32_t L = tree.nodeCount(0);
  FILE* o = std::fopen(argv[2], "wb");
  const int64_t nl = L;
  std::fwrite(&nl, 8, 1, o);
  const auto* leaf = tree.getFirstLeaf();
  for (uint32_t l = 0; l < L; ++l) {
    const nanovdb::Coord org = leaf[l].origin();
    const int32_t xyz[3] = {org[0], org[1], org[2]};
    std::fwrite(xyz, 4, 3, o);
  }
  auto gacc = grid->getAccessor();
  for (int64_t i = 0; i < m; ++i) {
    const uint64_t idx = gacc.getValue(nanovdb::Coord(c[3 * i], c[3 * i + 1], c[3 * i + 2]));
    std::fwrite(&idx, 8, 1, o);
  }
  std::fclose(o);
  return 0;
}
