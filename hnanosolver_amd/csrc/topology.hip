// Device-resident leaf topology: replaces CreateIndexGrid -> nanovdb::tools::cuda::voxelsToGrid
// (src/Cuda/HNanoSolver.cu:498-510) and the per-thread NanoVDB tree walks of every kernel.
//
// Input: the leaf-dense coordinate payload the reference's IndexGridBuilder writes
// (src/Utils/GridBuilder.hpp:164-174). The build validates it on the device (each leaf is 512
// consecutive voxels at origin + offset-order local coordinates; leaves strictly increasing in
// the voxelsToGrid key, PointsToGrid.cuh:596-651), then builds a leaf hash and the 27-entry
// neighbour row of every leaf. Only leaf origins stay resident: the kernels derive voxel
// coordinates from (leaf, local offset), so the N x 12 B coordinate array never enters HBM
// after the build.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "hns_internal.hpp"

namespace hns {
namespace {

constexpr int kErrNotLeafDense = 1;
constexpr int kErrOrder = 2;
constexpr int kErrRange = 4;
constexpr int kErrHashFull = 8;

// voxelsToGrid sort key of a leaf origin: root key (PointsToGrid.cuh:596-600) then upper and
// lower child offsets (NanoVDB.h:3568 CoordToOffset at the two internal levels).
struct LeafKey {
  uint64_t root;
  uint32_t local;
};
__host__ __device__ inline LeafKey leaf_sort_key(int32_t x, int32_t y, int32_t z) {
  const int64_t off = int64_t(1) << 31;
  LeafKey k;
  k.root = (uint64_t(uint32_t(int64_t(z) + off) >> 12)) | (uint64_t(uint32_t(int64_t(y) + off) >> 12) << 21) |
           (uint64_t(uint32_t(int64_t(x) + off) >> 12) << 42);
  const uint32_t up = (uint32_t((x & 4095) >> 7) << 10) | (uint32_t((y & 4095) >> 7) << 5) | uint32_t((z & 4095) >> 7);
  const uint32_t lo = (uint32_t((x & 127) >> 3) << 8) | (uint32_t((y & 127) >> 3) << 4) | uint32_t((z & 127) >> 3);
  k.local = (up << 12) | lo;
  return k;
}
__host__ __device__ inline bool key_less(const LeafKey& a, const LeafKey& b) {
  return a.root != b.root ? a.root < b.root : a.local < b.local;
}

__global__ void k_validate(const int32_t* __restrict__ coords, size_t n, int4* __restrict__ origins,
                           int* __restrict__ err) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t l = i >> 9;
  const int32_t nloc = (int32_t)(i & 511);
  const int32_t ox = coords[3 * (l << 9)], oy = coords[3 * (l << 9) + 1], oz = coords[3 * (l << 9) + 2];
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  int e = 0;
  if ((ox & 7) | (oy & 7) | (oz & 7)) e |= kErrNotLeafDense;
  if (x != ox + (nloc >> 6) || y != oy + ((nloc >> 3) & 7) || z != oz + (nloc & 7)) e |= kErrNotLeafDense;
  if (!leaf_in_range(ox >> 3, oy >> 3, oz >> 3)) e |= kErrRange;
  if (e) atomicOr(err, e);
  if (nloc == 0) origins[l] = make_int4(ox, oy, oz, 0);
}

__global__ void k_check_order(const int4* __restrict__ origins, int32_t L, int* __restrict__ err) {
  const int32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l + 1 >= L) return;
  const int4 a = origins[l], b = origins[l + 1];
  if (!key_less(leaf_sort_key(a.x, a.y, a.z), leaf_sort_key(b.x, b.y, b.z))) atomicOr(err, kErrOrder);
}

__global__ void k_hash_insert(const int4* __restrict__ origins, int32_t L, unsigned long long* __restrict__ keys,
                              int32_t* __restrict__ vals, uint32_t hmask, int* __restrict__ err) {
  const int32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= L) return;
  const int4 o = origins[l];
  const unsigned long long k = leaf_key(o.x >> 3, o.y >> 3, o.z >> 3);
  uint32_t h = hash_mix(k) & hmask;
  // keys are unique (strict order was checked first) and the table is at most half full, so a free
  // slot comes within hmask + 1 probes; the bound only turns a broken invariant into an error
  for (uint32_t probe = 0; probe <= hmask; ++probe) {
    const unsigned long long prev = atomicCAS(&keys[h], (unsigned long long)kEmptyKey, k);
    if (prev == kEmptyKey) {
      vals[h] = l;
      return;
    }
    h = (h + 1) & hmask;
  }
  atomicOr(err, kErrHashFull);
}

__global__ void k_neighbours(DevTopo T, int32_t* __restrict__ nbr) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)T.L * kNbrStride) return;
  const int32_t l = (int32_t)(g / kNbrStride), k = (int32_t)(g % kNbrStride);
  if (k >= 27) {
    nbr[g] = -1;
    return;
  }
  const int4 o = T.origins[l];
  const int32_t dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
  nbr[g] = hash_find(T, (o.x >> 3) + dx, (o.y >> 3) + dy, (o.z >> 3) + dz);
}

}  // namespace

std::shared_ptr<TopoStore> build_topology(const int32_t* coords, size_t n, float voxelSize, hipStream_t s) {
  if (n == 0) fail(HNS_ERR_EMPTY, "topology: N == 0");
  if (!coords) fail(HNS_ERR_INVALID_ARGUMENT, "topology: coords is null");
  if (n % kLeafVox) fail(HNS_ERR_TOPOLOGY, "topology: N is not a multiple of 512 (payload is not leaf-dense)");
  if (n > (size_t)INT32_MAX) fail(HNS_ERR_INVALID_ARGUMENT, "topology: N must be below 2^31");
  static std::atomic<uint64_t> next_id{1};
  auto T = std::make_shared<TopoStore>();
  T->id = next_id++;
  HNS_HIP(hipGetDevice(&T->device));
  T->N = n;
  T->L = (int32_t)(n / kLeafVox);
  T->voxelSize = voxelSize;
  const int32_t L = T->L;

  DBuf<int32_t> dcoords(3 * n);
  DBuf<int> derr(1);
  HNS_HIP(hipMemcpyAsync(dcoords.p, coords, 3 * n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HNS_HIP(hipMemsetAsync(derr.p, 0, sizeof(int), s));
  T->origins.alloc(L);
  k_validate<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dcoords.p, n, T->origins.p, derr.p);
  HNS_HIP(hipGetLastError());
  k_check_order<<<(unsigned)((L + 255) / 256), 256, 0, s>>>(T->origins.p, L, derr.p);
  HNS_HIP(hipGetLastError());
  int herr = 0;
  HNS_HIP(hipStreamSynchronize(s));
  HNS_HIP(hipMemcpy(&herr, derr.p, sizeof(int), hipMemcpyDeviceToHost));
  if (herr & kErrNotLeafDense)
    fail(HNS_ERR_TOPOLOGY, "topology: coords are not leaf-dense (each leaf must be 512 consecutive voxels in offset order)");
  if (herr & kErrRange) fail(HNS_ERR_TOPOLOGY, "topology: a coordinate lies outside [-2^23, 2^23)");
  if (herr & kErrOrder)
    fail(HNS_ERR_LEAF_ORDER, "topology: leaves are not in voxelsToGrid order or repeat (use hns_leaf_order)");

  uint32_t cap = 64;
  while (cap < 2u * (uint32_t)L) cap <<= 1;
  T->hmask = cap - 1;
  T->hkeys.alloc(cap);
  T->hvals.alloc(cap);
  HNS_HIP(hipMemsetAsync(T->hkeys.p, 0xFF, cap * sizeof(unsigned long long), s));
  k_hash_insert<<<(unsigned)((L + 255) / 256), 256, 0, s>>>(T->origins.p, L, T->hkeys.p, T->hvals.p, T->hmask,
                                                            derr.p);
  HNS_HIP(hipGetLastError());
  T->nbr.alloc((size_t)L * kNbrStride);
  k_neighbours<<<(unsigned)(((size_t)L * kNbrStride + 255) / 256), 256, 0, s>>>(T->dev(), T->nbr.p);
  HNS_HIP(hipGetLastError());
  HNS_HIP(hipStreamSynchronize(s));
  HNS_HIP(hipMemcpy(&herr, derr.p, sizeof(int), hipMemcpyDeviceToHost));
  if (herr & kErrHashFull) fail(HNS_ERR_TOPOLOGY, "topology: leaf hash table full (internal error)");
  noalloc_setup();  // the allocation-free entry points' scratch address on this device
  T->host_origins.resize(3 * (size_t)L);
  for (int32_t l = 0; l < L; ++l) std::memcpy(&T->host_origins[3 * (size_t)l], coords + 3 * (size_t)l * kLeafVox,
                                              3 * sizeof(int32_t));
  return T;
}

void leaf_order_host(const int32_t* origins, size_t L, uint64_t* perm) {
  std::vector<uint64_t> p(L);
  for (size_t i = 0; i < L; ++i) p[i] = i;
  std::stable_sort(p.begin(), p.end(), [&](uint64_t a, uint64_t b) {
    const int32_t* ca = origins + 3 * a;
    const int32_t* cb = origins + 3 * b;
    return key_less(leaf_sort_key(ca[0], ca[1], ca[2]), leaf_sort_key(cb[0], cb[1], cb[2]));
  });
  std::copy(p.begin(), p.end(), perm);
}

}  // namespace hns
