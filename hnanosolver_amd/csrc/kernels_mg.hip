// Geometric multigrid V-cycle over the leaf hierarchy, the preconditioner of HNS_SOLVER_MGCG (no
// reference counterpart: the reference runs a fixed red-black Gauss-Seidel count,
// src/Cuda/HNanoSolver.cu:364-373).
//
// Levels share the leaf topology: level l has S = 8 >> l cells per leaf axis (l = 0: voxels, in
// the colour-split layout of the CG vectors; l = 1, 2, 3: leaf-dense, cell = leaf * S^3 +
// (x * S + y) * S + z, level 3 one cell per leaf). With piecewise-constant prolongation P (a coarse
// value copied to its 8 children) and restriction R = P^T (the sum of the 8 children), the Galerkin
// operator of A = 6 I - (face neighbours) is R A P = 4 A at the next level: of a child's six faces,
// three stay inside the coarse cell and the four children on a coarse face see the same coarse
// neighbour, so level l carries 4^l (6 u - sum of face neighbours), zero outside active leaves as at
// level 0. Smoothing is red-black Gauss-Seidel on each level's colouring by the parity of the global
// level coordinate (at level 3, the leaf's).
#include <cstdlib>

#include "hns_mg.hpp"

namespace hns {

const MgShape& mg_shape() {
  static const MgShape v = [] {
    MgShape m;
    auto env = [](const char* k, int lo, int hi, int d) {
      const char* e = std::getenv(k);
      const int x = e ? std::atoi(e) : d;
      return x >= lo && x <= hi ? x : d;
    };
    m.coarse = env("HNS_MG_COARSE", 1, 256, m.coarse);
    m.smooth0 = env("HNS_MG_SMOOTH0", 1, 8, m.smooth0);
    m.smooth = env("HNS_MG_SMOOTH", 1, 8, m.smooth);
    m.depth = env("HNS_MG_DEPTH", 1, 8, m.depth);
    m.coarse_omega = 0.01f * (float)env("HNS_MG_COARSE_OMEGA", 100, 199, (int)(100.0f * m.coarse_omega + 0.5f));
    return m;
  }();
  return v;
}

namespace {

constexpr int kTpb = 256;

// Face neighbour of level cell (leaf, x, y, z) along AX by DIR (edge S = 1 << lgS): its index, or -1.
template <int AX, int DIR>
__device__ __forceinline__ int32_t mg_face(const DevTopo& T, int lgS, int32_t leaf, int x, int y, int z) {
  const int S = 1 << lgS;
  int c[3] = {x, y, z};
  c[AX] += DIR;
  int32_t nl = leaf;
  if (c[AX] < 0 || c[AX] >= S) {
    // face slots of the 27-neighbour row: (dx+1)*9 + (dy+1)*3 + (dz+1)
    const int slot = AX == 0 ? (DIR > 0 ? 22 : 4) : (AX == 1 ? (DIR > 0 ? 16 : 10) : (DIR > 0 ? 14 : 12));
    nl = T.nbr[kNbrStride * leaf + slot];
    if (nl < 0) return -1;
    c[AX] &= S - 1;
  }
  return (nl << (3 * lgS)) | (((c[0] << lgS) | c[1]) << lgS) | c[2];
}

__device__ __forceinline__ float mg_at(const float* __restrict__ u, int32_t i) { return i < 0 ? 0.f : u[i]; }

// Sum of the six face neighbours of level cell (leaf, x, y, z), in the order +x, -x, +y, -y, +z, -z.
__device__ __forceinline__ float mg_nsum(const DevTopo& T, int lgS, const float* __restrict__ u, int32_t leaf, int x,
                                         int y, int z) {
  return mg_at(u, mg_face<0, 1>(T, lgS, leaf, x, y, z)) + mg_at(u, mg_face<0, -1>(T, lgS, leaf, x, y, z)) +
         mg_at(u, mg_face<1, 1>(T, lgS, leaf, x, y, z)) + mg_at(u, mg_face<1, -1>(T, lgS, leaf, x, y, z)) +
         mg_at(u, mg_face<2, 1>(T, lgS, leaf, x, y, z)) + mg_at(u, mg_face<2, -1>(T, lgS, leaf, x, y, z));
}

// One SOR colour sweep at level lgS >= 0 (leaf-dense): cells of colour COLOUR move towards the
// solution of s (6 u - sum) = b by omega times the Gauss-Seidel step. FIRST: the level's u starts
// at zero (colour 0 reads no neighbours; both colours take u_old = 0).
template <int COLOUR, bool FIRST>
__global__ __launch_bounds__(kTpb) void k_mg_smooth(DevTopo T, int lgS, const float* __restrict__ b,
                                                    float* __restrict__ u, float inv_s, float omega) {
  const int32_t c = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  const int32_t ncell = T.L << (3 * lgS);
  if (c >= ncell) return;
  const int m = (1 << lgS) - 1;
  const int32_t leaf = c >> (3 * lgS);
  const int x = (c >> (2 * lgS)) & m, y = (c >> lgS) & m, z = c & m;
  // colour of the global level coordinate: the local parity for S even; the leaf's parity at S = 1
  int lp = 0;
  if (lgS == 0) {
    const int4 o = T.origins[leaf];
    lp = ((o.x + o.y + o.z) >> 3) & 1;
  }
  if (((lp + x + y + z) & 1) != COLOUR) {
    if (FIRST && COLOUR == 0) u[c] = 0.f;  // the other colour's first sweep reads these as neighbours
    return;
  }
  const float s = (FIRST && COLOUR == 0) ? 0.f : mg_nsum(T, lgS, u, leaf, x, y, z);
  const float gs = (b[c] * inv_s + s) * (1.0f / 6.0f);
  u[c] = FIRST ? omega * gs : fmaf(omega, gs - u[c], u[c]);  // SOR; omega = 1 is Gauss-Seidel
}

// Restriction of the level-lgS residual b - s (6 u - sum) (leaf-dense) to level lgS + 1: one lane
// per coarse cell sums its 8 children's residuals.
__global__ __launch_bounds__(kTpb) void k_mg_restrict(DevTopo T, int lgS, const float* __restrict__ b,
                                                      const float* __restrict__ u, float s,
                                                      float* __restrict__ bc) {
  const int32_t cc = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  const int lgC = lgS - 1;
  const int32_t ncoarse = T.L << (3 * lgC);
  if (cc >= ncoarse) return;
  const int mc = (1 << lgC) - 1;
  const int32_t leaf = cc >> (3 * lgC);
  const int X = (cc >> (2 * lgC)) & mc, Y = (cc >> lgC) & mc, Z = cc & mc;
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k & 1);
    const int32_t f = (leaf << (3 * lgS)) | (((x << lgS) | y) << lgS) | z;
    sum += b[f] - s * (6.0f * u[f] - mg_nsum(T, lgS, u, leaf, x, y, z));
  }
  bc[cc] = sum;
}

// u (level lgS, leaf-dense) += its parent's value in uc (level lgS + 1)
__global__ __launch_bounds__(kTpb) void k_mg_prolong(int32_t L, int lgS, float* __restrict__ u,
                                                     const float* __restrict__ uc) {
  const int32_t c = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  if (c >= (L << (3 * lgS))) return;
  const int m = (1 << lgS) - 1, lgC = lgS - 1;
  const int32_t leaf = c >> (3 * lgS);
  const int x = (c >> (2 * lgS)) & m, y = (c >> lgS) & m, z = c & m;
  u[c] += uc[(leaf << (3 * lgC)) | ((((x >> 1) << lgC) | (y >> 1)) << lgC) | (z >> 1)];
}

// Level 0 (colour-split): residual b - (6 u - sum) of the leaf's voxel pairs staged in LDS, then
// restricted to the leaf's 64 level-1 cells.
__global__ __launch_bounds__(kTpb) void k_mg0_restrict(DevTopo T, const float* __restrict__ b,
                                                       const float* __restrict__ u, float* __restrict__ b1) {
  __shared__ float s_tile[kPairTile];
  __shared__ float s_r[kLeafVox];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const FaceRow Fr = face_row(T, l);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
  const int c0 = (x + y) & 1;
  const int32_t i0 = c0 * half + l * kHalfLeaf + t, i1 = (c0 ^ 1) * half + l * kHalfLeaf + t;
  auto get = [&](int c, int32_t i) { return u[c * half + i]; };
  const float u0 = u[i0], u1 = u[i1];
  pair_tile_halo(Fr, get, s_tile);
  s_tile[c0 * kHalfLeaf + t] = u0;
  s_tile[(c0 ^ 1) * kHalfLeaf + t] = u1;
  __syncthreads();
  float s0, s1;
  pair_tile_sums(s_tile, t, u0, u1, s0, s1);
  s_r[(x << 6) | (y << 3) | (2 * zz)] = b[i0] - (6.0f * u0 - s0);
  s_r[(x << 6) | (y << 3) | (2 * zz + 1)] = b[i1] - (6.0f * u1 - s1);
  __syncthreads();
  if (t < 64) {
    const int X = t >> 4, Y = (t >> 2) & 3, Z = t & 3;
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      sum += s_r[((2 * X + (k >> 2)) << 6) | ((2 * Y + ((k >> 1) & 1)) << 3) | (2 * Z + (k & 1))];
    b1[l * 64 + t] = sum;
  }
}

// Level 0 (colour-split) u += parent value in u1 (level 1, leaf-dense)
__global__ __launch_bounds__(kTpb) void k_mg0_prolong(int32_t L, float* __restrict__ u, const float* __restrict__ u1) {
  const int32_t l = leaf_of_block();
  if (l >= L) return;
  const int32_t half = L * kHalfLeaf;
  const int t = threadIdx.x;
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int z = 2 * zz + ((c + x + y) & 1);
    u[c * half + l * kHalfLeaf + t] += u1[l * 64 + (((x >> 1) << 4) | ((y >> 1) << 2) | (z >> 1))];
  }
}

inline unsigned cells_grid(int32_t L, int lgS) { return (unsigned)((((int64_t)L << (3 * lgS)) + kTpb - 1) / kTpb); }

void smooth(const DevTopo& T, int lgS, const float* b, float* u, float scale, int colour, bool first, float omega,
            hipStream_t s) {
  const unsigned g = cells_grid(T.L, lgS);
  const float inv = 1.0f / scale;
  if (colour == 0) {
    if (first) k_mg_smooth<0, true><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega);
    else k_mg_smooth<0, false><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega);
  } else {
    if (first) k_mg_smooth<1, true><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega);
    else k_mg_smooth<1, false><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega);
  }
  HNS_HIP(hipGetLastError());
}

}  // namespace

void Multigrid::ensure(int32_t leaves) {
  if (L >= leaves && b[1].p) return;
  if (exec_) {  // the captured V-cycle refers to the buffers about to be replaced
    HNS_HIP(hipGraphExecDestroy(exec_));
    exec_ = nullptr;
  }
  L = leaves;
  for (int l = 1; l < kMgLevels; ++l) {
    const size_t n = (size_t)leaves << (3 * (3 - l));
    b[l].alloc(n);
    u[l].alloc(n);
  }
}

void Multigrid::launch_all(const DevTopo& T, const float* r_cs, float* z_cs, hipStream_t s) {
  const MgShape& M = mg_shape();
  // level 0: z = 0, red-black pre-smoothing (Gauss-Seidel on A z = r)
  for (int k = 0; k < M.smooth0; ++k) {
    launch_rbgs_rhs(T, r_cs, z_cs, 0, k == 0, s);
    launch_rbgs_rhs(T, r_cs, z_cs, 1, k == 0, s);
  }
  k_mg0_restrict<<<grid_for_leaves(T.L), kTpb, 0, s>>>(T, r_cs, z_cs, b[1].p);
  HNS_HIP(hipGetLastError());
  // descend: levels 1, 2 pre-smooth and restrict; level 3 (one cell per leaf) is smoothed M.coarse times
  for (int l = 1; l < kMgLevels; ++l) {
    const int lgS = 3 - l;
    const float scale = (float)(1 << (2 * l));
    const int iters = l == kMgLevels - 1 ? M.coarse : M.smooth;
    const float omega = l == kMgLevels - 1 ? M.coarse_omega : 1.0f;
    for (int k = 0; k < iters; ++k) {
      smooth(T, lgS, b[l].p, u[l].p, scale, 0, k == 0, omega, s);
      smooth(T, lgS, b[l].p, u[l].p, scale, 1, k == 0, omega, s);
    }
    if (l + 1 < kMgLevels) {
      k_mg_restrict<<<cells_grid(T.L, lgS - 1), kTpb, 0, s>>>(T, lgS, b[l].p, u[l].p, scale, b[l + 1].p);
      HNS_HIP(hipGetLastError());
    }
  }
  // ascend: prolong and post-smooth in the reverse colour order
  for (int l = kMgLevels - 2; l >= 1; --l) {
    const int lgS = 3 - l;
    const float scale = (float)(1 << (2 * l));
    k_mg_prolong<<<cells_grid(T.L, lgS), kTpb, 0, s>>>(T.L, lgS, u[l].p, u[l + 1].p);
    HNS_HIP(hipGetLastError());
    for (int k = 0; k < M.smooth; ++k) {
      smooth(T, lgS, b[l].p, u[l].p, scale, 1, false, 1.0f, s);
      smooth(T, lgS, b[l].p, u[l].p, scale, 0, false, 1.0f, s);
    }
  }
  k_mg0_prolong<<<grid_for_leaves(T.L), kTpb, 0, s>>>(T.L, z_cs, u[1].p);
  HNS_HIP(hipGetLastError());
  for (int k = 0; k < M.smooth0; ++k) {
    launch_rbgs_rhs(T, r_cs, z_cs, 1, false, s);
    launch_rbgs_rhs(T, r_cs, z_cs, 0, false, s);
  }
}

void Multigrid::apply(const DevTopo& T, const float* r_cs, float* z_cs, hipStream_t s, bool graph) {
  ensure(T.L);
  if (!graph) {
    launch_all(T, r_cs, z_cs, s);
    return;
  }
  if (!exec_ || key_nbr_ != T.nbr || key_L_ != T.L || key_r_ != r_cs || key_z_ != z_cs) {
    if (exec_) {
      HNS_HIP(hipGraphExecDestroy(exec_));
      exec_ = nullptr;
    }
    if (!cap_) HNS_HIP(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    HNS_HIP(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
    try {
      launch_all(T, r_cs, z_cs, cap_);
    } catch (...) {
      (void)hipStreamEndCapture(cap_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HNS_HIP(hipStreamEndCapture(cap_, &g));
    const hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HNS_HIP(e);
    key_nbr_ = T.nbr;
    key_L_ = T.L;
    key_r_ = r_cs;
    key_z_ = z_cs;
  }
  HNS_HIP(hipGraphLaunch(exec_, s));
}

Multigrid::~Multigrid() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (cap_) (void)hipStreamDestroy(cap_);
}

}  // namespace hns
