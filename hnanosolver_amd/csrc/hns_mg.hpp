// Geometric multigrid V-cycle over the leaf hierarchy (kernels_mg.hip): the preconditioner of
// HNS_SOLVER_MGCG, applied inside the flexible CG of kernels_cg.hip.
#pragma once
#include "hns_internal.hpp"

namespace hns {

// levels 0 (voxels) .. 3 (one cell per leaf)
constexpr int kMgLevels = 4;

// V-cycle shape: red-black SOR iterations on the coarsest level and their over-relaxation, pre- and
// post-smoothing Gauss-Seidel iterations on level 0 and on levels 1-2, and the flexible CG's
// truncation around the V-cycle. Defaults are the measured best at 8 M voxels; HNS_MG_COARSE,
// HNS_MG_COARSE_OMEGA (x100), HNS_MG_SMOOTH0, HNS_MG_SMOOTH and HNS_MG_DEPTH override them (read
// once per process).
struct MgShape {
  int coarse = 12, smooth0 = 1, smooth = 1, depth = 1;
  float coarse_omega = 1.85f;
};
const MgShape& mg_shape();

// The V-cycle is ~35 small launches, most of them on the coarse levels; it is captured once into a
// HIP graph per (topology, r, z) and replayed (the captured kernels take only these pointers, and
// none records a memset), which removes most of the per-launch cost.
struct Multigrid {
  int32_t L = 0;
  DBuf<float> b[kMgLevels], u[kMgLevels];  // levels 1..3 (leaf-dense); level 0 uses the caller's vectors
  void ensure(int32_t leaves);
  // z = V(r) on colour-split level-0 vectors: approximately A^-1 r for A = 6 I - (face neighbours).
  // graph: replay the captured V-cycle (captured on first use for these pointers); otherwise launch
  // the kernels one by one. Both run the same kernels on the same data.
  void apply(const DevTopo& T, const float* r_cs, float* z_cs, hipStream_t s, bool graph = true);
  ~Multigrid();

 private:
  void launch_all(const DevTopo& T, const float* r_cs, float* z_cs, hipStream_t s);
  hipStream_t cap_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  const int32_t* key_nbr_ = nullptr;
  int32_t key_L_ = 0;
  const float* key_r_ = nullptr;
  float* key_z_ = nullptr;
};

// One red-black Gauss-Seidel colour sweep (ω = 1) on A u = b, colour-split (the RBGS kernel with
// the right-hand side taken as b); `first`: u starts at zero.
void launch_rbgs_rhs(const DevTopo& T, const float* b_cs, float* u_cs, int colour, bool first, hipStream_t s);

}  // namespace hns
