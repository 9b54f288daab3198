// Geometric multigrid V-cycle over the leaf hierarchy (kernels_mg.hip): the preconditioner of
// HNS_SOLVER_MGCG, applied inside the flexible CG of kernels_cg.hip.
#pragma once
#include "hns_internal.hpp"

namespace hns {

// levels 0 (voxels) .. 3 (one cell per leaf); Gauss-Seidel iterations on the coarsest level
constexpr int kMgLevels = 4;
constexpr int kMgCoarseIters = 8;
// the flexible CG around the V-cycle keeps this many previous directions
constexpr int kMgFcgDepth = 2;

struct Multigrid {
  int32_t L = 0;
  DBuf<float> b[kMgLevels], u[kMgLevels];  // levels 1..3 (leaf-dense); level 0 uses the caller's vectors
  void ensure(int32_t leaves);
  // z = V(r) on colour-split level-0 vectors: approximately A^-1 r for A = 6 I - (face neighbours)
  void apply(const DevTopo& T, const float* r_cs, float* z_cs, hipStream_t s);
};

// One red-black Gauss-Seidel colour sweep (ω = 1) on A u = b, colour-split (the RBGS kernel with
// the right-hand side taken as b); `first`: u starts at zero.
void launch_rbgs_rhs(const DevTopo& T, const float* b_cs, float* u_cs, int colour, bool first, hipStream_t s);

}  // namespace hns
