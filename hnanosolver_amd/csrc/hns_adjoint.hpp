// Reverse mode of the differentiable step (kernels_adjoint.hip, adjoint.hip).
//
// Differentiable step (the density-control problem): u1 = advect_vector(u), div = divergence(u1),
// p = K red-black Gauss-Seidel iterations from 0, u2 = u1 - grad p, rho' = advect_scalars(rho; u2),
// u' = u2. The adjoint is the exact transpose of each kernel, last kernel first; the projection
// transpose is the same projection with the RBGS iterations transposed (a self-adjoint exact
// solve would transpose to itself).
#pragma once
#include <vector>

#include "hns_internal.hpp"

namespace hns {

// far: 4N floats of zeroed scratch (left zeroed on return)
void launch_advect_scalars_adj(const DevTopo& T, const float* u, const float* in, const float* gout, float* gin,
                               float* gvel, float dt, float inv_dx, float* far, hipStream_t s);
// far: 3N floats of zeroed scratch (left zeroed on return)
void launch_advect_vector_adj(const DevTopo& T, const float* u, const float* gout, float* gvel, float dt, float inv_dx,
                              float* far, hipStream_t s);
// gdiv_cs += (K RBGS iterations from 0)^T g_cs; g_cs is overwritten
void launch_rbgs_adj(const DevTopo& T, float* g_cs, float* gdiv_cs, float dx, int K, float omega, hipStream_t s);
// count += number of elements where a != b (bitwise)
void launch_count_mismatch(int64_t n, const float* a, const float* b, unsigned long long* count, hipStream_t s);

// Forward states at checkpoints every `every` steps; the backward pass recomputes each segment
// from its checkpoint (recomputed states must equal the stored ones bitwise).
class DiffTape {
 public:
  DiffTape(std::shared_ptr<TopoStore> topo, int steps, float dt, float dx, int K, float omega, int every);
  void forward(const float* vel0, const float* rho0, float* vel_out, float* rho_out, hipStream_t s);
  void backward(const float* gvel_out, const float* grho_out, float* gvel0, float* grho0, hipStream_t s);
  int checkpoints() const { return (int)ck_vel_.size(); }
  int64_t bytes() const;
  int64_t mismatches() const { return mismatches_; }

 private:
  void step(const float* vel, const float* rho, float* u1, float* u2, float* rho_out, hipStream_t s);
  std::shared_ptr<TopoStore> topo_;
  int steps_, K_, every_;
  float dt_, dx_, omega_;
  size_t N_;
  std::vector<DBuf<float>> ck_vel_, ck_rho_;  // state at steps 0, every, 2 every, ...
  DBuf<float> out_vel_, out_rho_;             // state after the last step
  std::vector<DBuf<float>> seg_vel_, seg_rho_, seg_u2_;  // one segment's per-step inputs
  DBuf<float> u1_, div_, p_, far_, gu_, gu1_, gp_, gdiv_, grho_in_, gvel_prev_, tmp_vel_, tmp_rho_;
  DBuf<unsigned long long> count_;
  int64_t mismatches_ = 0;
};

}  // namespace hns
