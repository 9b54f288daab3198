// Forward and reverse passes of the differentiable step with checkpointing (no reference
// counterpart). The step is the one described in hns_adjoint.hpp; states are SoA velocity (3N)
// and a density (N), leaf-dense; pressure and divergence colour-split.
#include "hns_adjoint.hpp"

namespace hns {

DiffTape::DiffTape(std::shared_ptr<TopoStore> topo, int steps, float dt, float dx, int K, float omega, int every)
    : topo_(std::move(topo)), steps_(steps), K_(K), every_(every > 0 ? every : 1), dt_(dt), dx_(dx), omega_(omega) {
  N_ = topo_->N;
  const int nck = (steps_ + every_ - 1) / every_;
  ck_vel_.resize(nck);
  ck_rho_.resize(nck);
  for (int i = 0; i < nck; ++i) {
    ck_vel_[i].alloc(3 * N_);
    ck_rho_[i].alloc(N_);
  }
  out_vel_.alloc(3 * N_);
  out_rho_.alloc(N_);
  u1_.alloc(3 * N_);
  div_.alloc(N_);
  p_.alloc(N_);
  far_.alloc(4 * N_);
  HNS_HIP(hipMemset(far_.p, 0, 4 * N_ * sizeof(float)));
  count_.alloc(1);
}

int64_t DiffTape::bytes() const {
  int64_t b = 0;
  for (auto& v : ck_vel_) b += (int64_t)v.bytes();
  for (auto& v : ck_rho_) b += (int64_t)v.bytes();
  for (auto& v : seg_vel_) b += (int64_t)v.bytes();
  for (auto& v : seg_rho_) b += (int64_t)v.bytes();
  for (auto& v : seg_u2_) b += (int64_t)v.bytes();
  const DBuf<float>* bufs[] = {&out_vel_, &out_rho_, &u1_,   &div_,      &p_,         &far_,     &gu_,
                               &gu1_,     &gp_,      &gdiv_, &grho_in_, &gvel_prev_, &tmp_vel_, &tmp_rho_};
  for (auto* d : bufs) b += (int64_t)d->bytes();
  return b;
}

// One step: u1 = advect_vector(vel); div = divergence(u1); p = K RBGS from 0; u2 = u1 - grad p;
// rho_out = advect_scalars(rho; u2).
void DiffTape::step(const float* vel, const float* rho, float* u1, float* u2, float* rho_out, hipStream_t s) {
  const DevTopo T = topo_->dev();
  const float inv_dx = 1.0f / dx_;
  launch_advect_velocity(T, vel, u1, dt_, inv_dx, nullptr, s);
  launch_divergence(T, u1, div_.p, inv_dx, true, nullptr, nullptr, s);
  for (int it = 0; it < K_; ++it) {
    launch_rbgs(T, div_.p, p_.p, dx_, 0, omega_, it == 0, s);
    launch_rbgs(T, div_.p, p_.p, dx_, 1, omega_, it == 0, s);
  }
  if (K_ == 0) HNS_HIP(hipMemsetAsync(p_.p, 0, N_ * sizeof(float), s));
  launch_subtract_gradient(T, u1, p_.p, u2, inv_dx, nullptr, s);
  FieldSet f{};
  f.m = 1;
  f.in[0] = rho;
  f.out[0] = rho_out;
  launch_advect_scalars(T, u2, f, dt_, inv_dx, nullptr, s);
}

void DiffTape::forward(const float* vel0, const float* rho0, float* vel_out, float* rho_out, hipStream_t s) {
  tmp_vel_.alloc(3 * N_);
  tmp_rho_.alloc(N_);
  HNS_HIP(hipMemcpyAsync(out_vel_.p, vel0, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
  HNS_HIP(hipMemcpyAsync(out_rho_.p, rho0, N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
  for (int st = 0; st < steps_; ++st) {
    if (st % every_ == 0) {
      HNS_HIP(hipMemcpyAsync(ck_vel_[st / every_].p, out_vel_.p, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
      HNS_HIP(hipMemcpyAsync(ck_rho_[st / every_].p, out_rho_.p, N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    step(out_vel_.p, out_rho_.p, u1_.p, tmp_vel_.p, tmp_rho_.p, s);
    std::swap(out_vel_, tmp_vel_);
    std::swap(out_rho_, tmp_rho_);
  }
  if (vel_out) HNS_HIP(hipMemcpyAsync(vel_out, out_vel_.p, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (rho_out) HNS_HIP(hipMemcpyAsync(rho_out, out_rho_.p, N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
}

void DiffTape::backward(const float* gvel_out, const float* grho_out, float* gvel0, float* grho0, hipStream_t s) {
  const DevTopo T = topo_->dev();
  const float inv_dx = 1.0f / dx_;
  const int seg = every_;
  if ((int)seg_vel_.size() < seg) {
    seg_vel_.resize(seg);
    seg_rho_.resize(seg);
    seg_u2_.resize(seg);
    for (int i = 0; i < seg; ++i) {
      seg_vel_[i].alloc(3 * N_);
      seg_rho_[i].alloc(N_);
      seg_u2_[i].alloc(3 * N_);
    }
  }
  gu_.alloc(3 * N_);
  gu1_.alloc(3 * N_);
  gp_.alloc(N_);
  gdiv_.alloc(N_);
  grho_in_.alloc(N_);
  gvel_prev_.alloc(3 * N_);
  // running adjoints of the state after the current step
  if (gvel_out)
    HNS_HIP(hipMemcpyAsync(gvel0, gvel_out, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
  else
    HNS_HIP(hipMemsetAsync(gvel0, 0, 3 * N_ * sizeof(float), s));
  if (grho_out)
    HNS_HIP(hipMemcpyAsync(grho0, grho_out, N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
  else
    HNS_HIP(hipMemsetAsync(grho0, 0, N_ * sizeof(float), s));
  HNS_HIP(hipMemsetAsync(count_.p, 0, sizeof(unsigned long long), s));
  for (int c = checkpoints() - 1; c >= 0; --c) {
    const int s0 = c * every_, s1 = std::min(steps_, s0 + every_);
    // recompute the segment's per-step inputs from its checkpoint
    HNS_HIP(hipMemcpyAsync(seg_vel_[0].p, ck_vel_[c].p, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
    HNS_HIP(hipMemcpyAsync(seg_rho_[0].p, ck_rho_[c].p, N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
    for (int st = s0; st < s1; ++st) {
      const int k = st - s0;
      float* nv = k + 1 < s1 - s0 ? seg_vel_[k + 1].p : tmp_vel_.p;
      float* nr = k + 1 < s1 - s0 ? seg_rho_[k + 1].p : tmp_rho_.p;
      step(seg_vel_[k].p, seg_rho_[k].p, u1_.p, seg_u2_[k].p, nr, s);
      HNS_HIP(hipMemcpyAsync(nv, seg_u2_[k].p, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    // the recomputed end state must equal the stored one bitwise
    const float* ev = c + 1 < checkpoints() ? ck_vel_[c + 1].p : out_vel_.p;
    const float* er = c + 1 < checkpoints() ? ck_rho_[c + 1].p : out_rho_.p;
    launch_count_mismatch(3 * (int64_t)N_, tmp_vel_.p, ev, count_.p, s);
    launch_count_mismatch((int64_t)N_, tmp_rho_.p, er, count_.p, s);
    for (int st = s1 - 1; st >= s0; --st) {
      const int k = st - s0;
      // rho' = advect_scalars(rho; u2): grho_in, and d/du2 added to the adjoint of u' = u2
      HNS_HIP(hipMemcpyAsync(gu_.p, gvel0, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
      HNS_HIP(hipMemsetAsync(grho_in_.p, 0, N_ * sizeof(float), s));
      launch_advect_scalars_adj(T, seg_u2_[k].p, seg_rho_[k].p, grho0, grho_in_.p, gu_.p, dt_, inv_dx, far_.p, s);
      // projection: gp = div(gu2), gdiv = (RBGS_K)^T gp, gu1 = gu2 - grad(gdiv)
      launch_divergence(T, gu_.p, gp_.p, inv_dx, true, nullptr, nullptr, s);
      HNS_HIP(hipMemsetAsync(gdiv_.p, 0, N_ * sizeof(float), s));
      launch_rbgs_adj(T, gp_.p, gdiv_.p, dx_, K_, omega_, s);
      launch_subtract_gradient(T, gu_.p, gdiv_.p, gu1_.p, inv_dx, nullptr, s);
      // u1 = advect_vector(vel)
      HNS_HIP(hipMemsetAsync(gvel_prev_.p, 0, 3 * N_ * sizeof(float), s));
      launch_advect_vector_adj(T, seg_vel_[k].p, gu1_.p, gvel_prev_.p, dt_, inv_dx, far_.p, s);
      HNS_HIP(hipMemcpyAsync(gvel0, gvel_prev_.p, 3 * N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
      HNS_HIP(hipMemcpyAsync(grho0, grho_in_.p, N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
  }
  unsigned long long h = 0;
  HNS_HIP(hipStreamSynchronize(s));
  HNS_HIP(hipMemcpy(&h, count_.p, sizeof(h), hipMemcpyDeviceToHost));
  mismatches_ = (int64_t)h;
}

}  // namespace hns
