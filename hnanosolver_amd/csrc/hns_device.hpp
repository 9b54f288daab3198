// Device-side topology and sampling helpers for the gfx950 kernels of the sparse step.
//
// Layout in HBM (all leaf-dense, 512 voxels per 8^3 leaf, leaves in voxelsToGrid order):
//   scalar field      float[N], voxel (leaf l, local x,y,z) at 512*l + (x<<6 | y<<3 | z)
//   velocity          float[3N] SoA planes: x at [0,N), y at [N,2N), z at [2N,3N)
//   pressure/div (RBGS) "colour-split": colour c=(x+y+z)&1 plane at [c*N/2, (c+1)*N/2),
//                     inside it leaf l, local (x,y,z) at 256*l + 32*x + 4*y + (z>>1)
//   topology          int4 origin per leaf; 32-int neighbour row per leaf (27 used,
//                     entry (dx+1)*9+(dy+1)*3+(dz+1), -1 absent); open-addressing leaf hash.
// A voxel lookup replaces NanoVDB's per-thread root->leaf walk (Stencils.hpp:51-71) with one
// LDS read of the neighbour row (|offset| <= 1 leaf) or a hash probe (farther).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hns {

constexpr int32_t kLeafVox = 512;
constexpr int32_t kHalfLeaf = 256;
constexpr int32_t kNbrStride = 32;
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int32_t kLeafBias = 1 << 20;  // leaf coordinates in [-2^20, 2^20) -> voxels in [-2^23, 2^23)

struct DevTopo {
  int32_t L;
  uint32_t hmask;
  const int4* __restrict__ origins;
  const int32_t* __restrict__ nbr;
  const unsigned long long* __restrict__ hkeys;
  const int32_t* __restrict__ hvals;
  uint64_t id;  // host-side identity of the topology build (keys of cached graphs and layouts)
};

__host__ __device__ inline bool leaf_in_range(int32_t lx, int32_t ly, int32_t lz) {
  return (uint32_t)(lx + kLeafBias) < (1u << 21) && (uint32_t)(ly + kLeafBias) < (1u << 21) &&
         (uint32_t)(lz + kLeafBias) < (1u << 21);
}
__host__ __device__ inline uint64_t leaf_key(int32_t lx, int32_t ly, int32_t lz) {
  return ((uint64_t)(uint32_t)(lx + kLeafBias) << 42) | ((uint64_t)(uint32_t)(ly + kLeafBias) << 21) |
         (uint64_t)(uint32_t)(lz + kLeafBias);
}
__host__ __device__ inline uint32_t hash_mix(uint64_t k) {  // MurmurHash3 fmix64
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return (uint32_t)k;
}

__device__ inline int32_t hash_find(const DevTopo& T, int32_t lx, int32_t ly, int32_t lz) {
  if (!leaf_in_range(lx, ly, lz)) return -1;
  const uint64_t k = leaf_key(lx, ly, lz);
  uint32_t h = hash_mix(k) & T.hmask;
  // load factor <= 1/2, so an empty slot ends every probe sequence long before the bound
  for (uint32_t probe = 0; probe <= T.hmask; ++probe) {
    const uint64_t kk = T.hkeys[h];
    if (kk == k) return T.hvals[h];
    if (kk == kEmptyKey) return -1;
    h = (h + 1) & T.hmask;
  }
  return -1;
}

// Workgroup b is dispatched to XCD b % 8. Give each XCD a contiguous run of leaves so that a
// leaf and its +-z/+-y neighbours (adjacent in voxelsToGrid order) share an L2.
__device__ __forceinline__ int32_t leaf_of_block() {
  const uint32_t b = blockIdx.x, G = gridDim.x;
  return (int32_t)((b & 7u) * (G >> 3) + (b >> 3));
}
inline unsigned grid_for_leaves(int32_t L) { return 8u * (unsigned)((L + 7) / 8); }
// The same for a grid of any size G: workgroup b (on XCD b % 8) gets slot b >> 3 of its XCD's
// contiguous run of slots (the first G % 8 XCDs one slot more), a bijection of [0, G), so
// workgroups that handle neighbouring leaves share an L2 and reuse each other's halo cells there.
__device__ __forceinline__ int32_t xcd_slot() {
  const uint32_t b = blockIdx.x, G = gridDim.x, x = b & 7u, q = G >> 3, r = G & 7u;
  return (int32_t)(x * q + (x < r ? x : r) + (b >> 3));
}

// Strided leaf run of a workgroup of a grid smaller than the leaf count (gridDim.x a multiple of
// 8): the leaves are cut into 8 contiguous ranges, one per XCD, and the XCD's workgroups
// interleave inside their range.
struct LeafRun {
  int32_t first, end, step;
};
// (workgroup b of G, G a multiple of 8)
__device__ inline LeafRun leaf_run(int32_t L, int32_t b, int32_t G) {
  const int32_t W = G >> 3;
  const int32_t per = (L + 7) / 8;
  const int32_t lo = (b & 7) * per;
  return LeafRun{lo + (b >> 3), min(L, lo + per), W};
}
__device__ inline LeafRun leaf_run(int32_t L) { return leaf_run(L, (int32_t)blockIdx.x, (int32_t)gridDim.x); }

// float -> int round-down (Stencils.hpp:28-30, __float2int_rd); far/NaN positions are clamped
// to a coordinate outside every domain so they read as misses.
__device__ __forceinline__ int32_t floor_i(float v) {
  float f = floorf(v);
  f = fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
  return (int32_t)f;
}

__device__ __forceinline__ int32_t local_off(int32_t gx, int32_t gy, int32_t gz) {
  return ((gx & 7) << 6) | ((gy & 7) << 3) | (gz & 7);
}
__device__ __forceinline__ int32_t cs_local(int32_t x, int32_t y, int32_t z) {  // x,y,z in [0,8)
  return (x << 5) | (y << 2) | (z >> 1);
}

// Per-workgroup view of one leaf and its 26 neighbours.
struct LeafView {
  const int32_t* nb;  // 27 neighbour leaf ids in LDS
  int32_t ox, oy, oz;
  DevTopo T;

  __device__ __forceinline__ int32_t leaf_at(int32_t gx, int32_t gy, int32_t gz) const {
    const int32_t rx = (gx - ox) >> 3, ry = (gy - oy) >> 3, rz = (gz - oz) >> 3;
    if ((uint32_t)(rx + 1) < 3u && (uint32_t)(ry + 1) < 3u && (uint32_t)(rz + 1) < 3u)
      return nb[(rx + 1) * 9 + (ry + 1) * 3 + (rz + 1)];
    return hash_find(T, gx >> 3, gy >> 3, gz >> 3);
  }
  // leaf-dense index of a voxel, -1 when its leaf is absent
  __device__ __forceinline__ int32_t index(int32_t gx, int32_t gy, int32_t gz) const {
    const int32_t leaf = leaf_at(gx, gy, gz);
    return leaf < 0 ? -1 : leaf * kLeafVox + local_off(gx, gy, gz);
  }
  // colour-split index of a voxel, -1 when absent
  __device__ __forceinline__ int32_t cs_index(int32_t gx, int32_t gy, int32_t gz, int32_t half) const {
    const int32_t leaf = leaf_at(gx, gy, gz);
    if (leaf < 0) return -1;
    const int32_t x = gx & 7, y = gy & 7, z = gz & 7;
    return ((x + y + z) & 1) * half + leaf * kHalfLeaf + cs_local(x, y, z);
  }
};

// Loads the leaf's neighbour row into LDS and returns its view. Call before any early exit
// that is not uniform across the workgroup; ends with a barrier.
__device__ __forceinline__ LeafView load_leaf(const DevTopo& T, int32_t l, int32_t* s_nb) {
  if (threadIdx.x < 27) s_nb[threadIdx.x] = T.nbr[kNbrStride * l + threadIdx.x];
  const int4 o = T.origins[l];
  __syncthreads();
  LeafView v;
  v.nb = s_nb;
  v.ox = o.x;
  v.oy = o.y;
  v.oz = o.z;
  v.T = T;
  return v;
}

// The six face neighbours of a leaf (uniform per workgroup, so they live in SGPRs): the fast
// path of every 7-point stencil (divergence, gradient, residual, CG), no LDS or hash involved.
struct FaceRow {
  int32_t l, xm, xp, ym, yp, zm, zp;
};
__device__ __forceinline__ FaceRow face_row(const DevTopo& T, int32_t l) {
  const int32_t* r = T.nbr + kNbrStride * l;
  return FaceRow{l, r[4], r[22], r[10], r[16], r[12], r[14]};
}
// Leaf holding local coordinate (lx,ly,lz) moved by DIR along axis AX, with the wrapped local
// coordinate in (nx,ny,nz); -1 when that leaf is absent.
template <int AX, int DIR>
__device__ __forceinline__ int32_t face_step(const FaceRow& F, int lx, int ly, int lz, int& nx, int& ny, int& nz) {
  nx = lx;
  ny = ly;
  nz = lz;
  int& c = AX == 0 ? nx : (AX == 1 ? ny : nz);
  c += DIR;
  if ((unsigned)c < 8u) return F.l;
  c &= 7;
  if (AX == 0) return DIR > 0 ? F.xp : F.xm;
  if (AX == 1) return DIR > 0 ? F.yp : F.ym;
  return DIR > 0 ? F.zp : F.zm;
}
template <int AX, int DIR>
__device__ __forceinline__ int32_t face_dense(const FaceRow& F, int lx, int ly, int lz) {
  int nx, ny, nz;
  const int32_t leaf = face_step<AX, DIR>(F, lx, ly, lz, nx, ny, nz);
  return leaf < 0 ? -1 : leaf * kLeafVox + ((nx << 6) | (ny << 3) | nz);
}
template <int AX, int DIR>
__device__ __forceinline__ int32_t face_cs(const FaceRow& F, int lx, int ly, int lz, int32_t half) {
  int nx, ny, nz;
  const int32_t leaf = face_step<AX, DIR>(F, lx, ly, lz, nx, ny, nz);
  return leaf < 0 ? -1 : ((nx + ny + nz) & 1) * half + leaf * kHalfLeaf + cs_local(nx, ny, nz);
}

// Colour-split voxel pairs: a 256-lane workgroup covers a leaf with lane t = (x<<5)|(y<<2)|zz
// owning v0 = (x,y,2zz) in colour c0 = (x+y)&1 and v1 = (x,y,2zz+1) in colour c0^1, both at
// in-plane offset 256*l + t. Their 6-neighbour sums need no per-lane index arithmetic beyond
// the leaf faces: +-x is +-32 (+-224 across a face), +-y is +-4 (+-28), z stays inside the pair
// or moves one slot (3 across a face). get(c, i) returns the value at colour plane c, index i.
// Sums are in the order +x, -x, +y, -y, +z, -z.
template <class T, class G>
__device__ __forceinline__ void pair_face_sums(const FaceRow& F, int32_t t, G get, T v0, T v1, T& s0, T& s1) {
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
  const int c0 = (x + y) & 1, c1 = c0 ^ 1;
  const int32_t b = F.l * kHalfLeaf + t;
  auto at = [&](int c, int32_t leaf, int32_t off) -> T { return leaf < 0 ? T(0) : get(c, leaf * kHalfLeaf + off); };
  const T xp0 = x < 7 ? get(c1, b + 32) : at(c1, F.xp, t - 224);
  const T xm0 = x > 0 ? get(c1, b - 32) : at(c1, F.xm, t + 224);
  const T yp0 = y < 7 ? get(c1, b + 4) : at(c1, F.yp, t - 28);
  const T ym0 = y > 0 ? get(c1, b - 4) : at(c1, F.ym, t + 28);
  const T zm0 = zz > 0 ? get(c1, b - 1) : at(c1, F.zm, t + 3);
  const T xp1 = x < 7 ? get(c0, b + 32) : at(c0, F.xp, t - 224);
  const T xm1 = x > 0 ? get(c0, b - 32) : at(c0, F.xm, t + 224);
  const T yp1 = y < 7 ? get(c0, b + 4) : at(c0, F.yp, t - 28);
  const T ym1 = y > 0 ? get(c0, b - 4) : at(c0, F.ym, t + 28);
  const T zp1 = zz < 3 ? get(c0, b + 1) : at(c0, F.zp, t - 3);
  s0 = xp0 + xm0 + yp0 + ym0 + v1 + zm0;
  s1 = xp1 + xm1 + yp1 + ym1 + zp1 + v0;
}

// LDS image of one leaf of a colour-split vector for 256-lane pair stencils: the two colour
// planes (256 values each, at in-plane offset t), then the 6 face halos -x, +x, -y, +y, -z, +z
// (64 values each, indexed (a<<3)|b by the two in-face coordinates in x, y, z order). Staged
// once per leaf, it serves the 12 neighbour reads of every voxel pair from LDS instead of the
// cache hierarchy.
constexpr int kPairTile = 2 * kHalfLeaf + 6 * 64;
constexpr int kPairHalo = 2 * kHalfLeaf;
// Fills the 6 halos from the face neighbours of leaf F.l (absent leaves read as 0); get(c, i)
// as in pair_face_sums. Every lane of the 256-lane workgroup calls it; the face index is
// uniform per wave.
template <class T, class G>
__device__ __forceinline__ void pair_tile_halo(const FaceRow& F, G get, T* S) {
  for (int e = threadIdx.x; e < 6 * 64; e += 256) {
    const int f = e >> 6, a = (e >> 3) & 7, b = e & 7;
    int vx, vy, vz;
    int32_t nl;
    switch (f) {
      case 0: nl = F.xm; vx = 7; vy = a; vz = b; break;
      case 1: nl = F.xp; vx = 0; vy = a; vz = b; break;
      case 2: nl = F.ym; vx = a; vy = 7; vz = b; break;
      case 3: nl = F.yp; vx = a; vy = 0; vz = b; break;
      case 4: nl = F.zm; vx = a; vy = b; vz = 7; break;
      default: nl = F.zp; vx = a; vy = b; vz = 0; break;
    }
    S[kPairHalo + e] = nl < 0 ? T(0) : get((vx + vy + vz) & 1, nl * kHalfLeaf + cs_local(vx, vy, vz));
  }
}
// Where lane t's pair-tile entries of leaf l come from, as colour-split vector indices, for kernels
// that issue a leaf's loads ahead of the leaf before it. Halo entry t (faces -x, +x, -y, +y, one per
// wave) and entry t + 256 (faces -z, +z; lanes t < 128) of pair_tile_halo's order come from the
// face neighbours nb[0], nb[1] (pair_halo_nbrs). ix[0], ix[1] the lane's voxel pair (colour
// c0 = ((t >> 5) + (t >> 2)) & 1, then the other), ix[2], ix[3] the halo entries; -1 where the
// neighbour leaf is absent or the lane has no second halo entry.
struct PairIdx {
  int32_t ix[4];
};
// nb[0], nb[1] of leaf l for the wave of lane t (the face row's entry -x, +x, -y or +y by wave,
// then -z or +z): scalar loads through the constant address space (the neighbour rows are not
// written while a kernel reads them), which wait on the LDS counter, not on the vector-memory
// counter that the prefetched data loads are tracked by
__device__ __forceinline__ void pair_halo_nbrs(const DevTopo& T, int32_t l, int t, int32_t (&nb)[2]) {
  typedef const __attribute__((address_space(4))) int32_t* ConstRow;
  const int f = __builtin_amdgcn_readfirstlane(t >> 6);
  const ConstRow r = (ConstRow)(T.nbr + kNbrStride * l);
  nb[0] = r[f == 0 ? 4 : f == 1 ? 22 : f == 2 ? 10 : 16];
  nb[1] = r[(f & 1) ? 14 : 12];
}
__device__ __forceinline__ PairIdx pair_tile_index(int32_t l, const int32_t (&nb)[2], int t, int32_t half) {
  const int a = (t >> 3) & 7, b = t & 7, f = t >> 6, c0 = ((t >> 5) + (t >> 2)) & 1;
  PairIdx P;
  P.ix[0] = c0 * half + l * kHalfLeaf + t;
  P.ix[1] = (c0 ^ 1) * half + l * kHalfLeaf + t;
  const int vx = f == 0 ? 7 : (f == 1 ? 0 : a), vy = f < 2 ? a : (f == 2 ? 7 : 0);
  P.ix[2] = nb[0] < 0 ? -1 : ((vx + vy + b) & 1) * half + nb[0] * kHalfLeaf + cs_local(vx, vy, b);
  const int vz = (f & 1) ? 0 : 7;  // entry t + 256: face 4 (-z) for t < 64, 5 (+z) for t < 128
  P.ix[3] = (t >= 128 || nb[1] < 0) ? -1 : ((a + b + vz) & 1) * half + nb[1] * kHalfLeaf + cs_local(a, b, vz);
  return P;
}
// Stores lane t's four pair-tile values (as pair_tile_index orders them) into the tile.
template <class T>
__device__ __forceinline__ void pair_tile_put(T* S, int t, T v0, T v1, T h0, T h1) {
  const int c0 = ((t >> 5) + (t >> 2)) & 1;
  S[c0 * kHalfLeaf + t] = v0;
  S[(c0 ^ 1) * kHalfLeaf + t] = v1;
  S[kPairHalo + t] = h0;
  if (t < 128) S[kPairHalo + 256 + t] = h1;
}
// Same sums as pair_face_sums (same order, same values) from a staged tile.
template <class T>
__device__ __forceinline__ void pair_tile_sums(const T* S, int32_t t, T v0, T v1, T& s0, T& s1) {
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
  const int c0 = (x + y) & 1, c1 = c0 ^ 1;
  const int z0 = 2 * zz, z1 = 2 * zz + 1;
  const int p0 = c0 * kHalfLeaf + t, p1 = c1 * kHalfLeaf + t;
  const T xp0 = S[x < 7 ? p1 + 32 : kPairHalo + 64 + y * 8 + z0];
  const T xm0 = S[x > 0 ? p1 - 32 : kPairHalo + 0 + y * 8 + z0];
  const T yp0 = S[y < 7 ? p1 + 4 : kPairHalo + 192 + x * 8 + z0];
  const T ym0 = S[y > 0 ? p1 - 4 : kPairHalo + 128 + x * 8 + z0];
  const T zm0 = S[zz > 0 ? p1 - 1 : kPairHalo + 256 + x * 8 + y];
  const T xp1 = S[x < 7 ? p0 + 32 : kPairHalo + 64 + y * 8 + z1];
  const T xm1 = S[x > 0 ? p0 - 32 : kPairHalo + 0 + y * 8 + z1];
  const T yp1 = S[y < 7 ? p0 + 4 : kPairHalo + 192 + x * 8 + z1];
  const T ym1 = S[y > 0 ? p0 - 4 : kPairHalo + 128 + x * 8 + z1];
  const T zp1 = S[zz < 3 ? p0 + 1 : kPairHalo + 320 + x * 8 + y];
  s0 = xp0 + xm0 + yp0 + ym0 + v1 + zm0;
  s1 = xp1 + xm1 + yp1 + ym1 + zp1 + v0;
}

struct Vel {
  const float* __restrict__ x;
  const float* __restrict__ y;
  const float* __restrict__ z;
};

__device__ __forceinline__ float3 load_vel(const Vel& u, int32_t idx) {
  return idx < 0 ? make_float3(0.f, 0.f, 0.f) : make_float3(u.x[idx], u.y[idx], u.z[idx]);
}
__device__ __forceinline__ float load_s(const float* __restrict__ f, int32_t idx) { return idx < 0 ? 0.f : f[idx]; }

// TrilinearSampler::sample on Vec3f, device branch (Stencils.hpp:118-153): corners with
// misses as 0, lerp fmaf(w, b - a, a), z then y then x.
__device__ inline float3 trilinear_vec(const LeafView& V, const Vel& u, float px, float py, float pz) {
  const int32_t i = floor_i(px), j = floor_i(py), k = floor_i(pz);
  const float tx = px - (float)i, ty = py - (float)j, tz = pz - (float)k;
  float3 c[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int d = 0; d < 2; ++d) c[a][b][d] = load_vel(u, V.index(i + a, j + b, k + d));
  auto lerp = [](float3 va, float3 vb, float w) {
    return make_float3(fmaf(w, vb.x - va.x, va.x), fmaf(w, vb.y - va.y, va.y), fmaf(w, vb.z - va.z, va.z));
  };
  const float3 z0 = lerp(c[0][0][0], c[0][0][1], tz), z1 = lerp(c[0][1][0], c[0][1][1], tz);
  const float3 z2 = lerp(c[1][0][0], c[1][0][1], tz), z3 = lerp(c[1][1][0], c[1][1][1], tz);
  const float3 y0 = lerp(z0, z1, ty), y1 = lerp(z2, z3, ty);
  return lerp(y0, y1, tx);
}

// TrilinearSampler::sample on float (Stencils.hpp:140: a + w*(b - a)).
__device__ inline float trilinear_s(const LeafView& V, const float* __restrict__ f, float px, float py, float pz) {
  const int32_t i = floor_i(px), j = floor_i(py), k = floor_i(pz);
  const float tx = px - (float)i, ty = py - (float)j, tz = pz - (float)k;
  float c[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int d = 0; d < 2; ++d) c[a][b][d] = load_s(f, V.index(i + a, j + b, k + d));
  auto lerp = [](float va, float vb, float w) { return va + w * (vb - va); };
  const float z0 = lerp(c[0][0][0], c[0][0][1], tz), z1 = lerp(c[0][1][0], c[0][1][1], tz);
  const float z2 = lerp(c[1][0][0], c[1][0][1], tz), z3 = lerp(c[1][1][0], c[1][1][1], tz);
  const float y0 = lerp(z0, z1, ty), y1 = lerp(z2, z3, ty);
  return lerp(y0, y1, tx);
}

// The SOR update of redBlackGaussSeidelUpdate (Kernel.cu:487-488), shared by every red-black
// kernel so they produce identical bits: pGS = (sum6 - div dx^2) / 6 with the reference's
// 0.166666667f.
__device__ __forceinline__ float sor_update(float s, float dv, float pOld, float dx2, float omega) {
  const float pGS = (s - dv * dx2) * 0.166666667f;
  return pOld + omega * (pGS - pOld);
}

// Empty device asm that makes a value opaque to the optimiser, so what is derived from it is
// recomputed where it is used instead of being hoisted into registers across a loop (v: a
// per-lane value, s: a wave-uniform one). The host pass of a template kernel sees nothing.
template <class T>
__device__ __forceinline__ void opaque_v(T& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(v));
#endif
}
template <class T>
__device__ __forceinline__ void opaque_s(T& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+s"(v));
#endif
}

}  // namespace hns
