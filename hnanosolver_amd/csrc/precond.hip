// Host-side sequencing of the learned preconditioner's forward and backward passes (the network
// is described in hns_precond.hpp; no reference counterpart).
#include <cstdlib>

#include "hns_precond.hpp"

namespace hns {

namespace {
constexpr int64_t kWgradPartialFloats = 64LL * 27 * 64 * 64;  // ranges x largest layer
// HNS_PRECOND_UNFUSED=1 runs inference through the layer-by-layer kernels (for comparison)
bool fused_inference() {
  static const bool v = [] {
    const char* e = std::getenv("HNS_PRECOND_UNFUSED");
    return !(e && e[0] == '1');
  }();
  return v;
}
}

// The inference and the training buffers grow separately, each to the largest topology it has
// served: a training pass after a larger inference-only pass must not find gradient buffers sized
// for an earlier, smaller topology.
void PcWork::ensure(int32_t leaves, bool training) {
  if (leaves > L || !h0.p) {
    L = std::max(L, leaves);
    const int64_t n0 = (int64_t)L * 512, n1 = (int64_t)L * 64, n2 = (int64_t)L * 8;
    h0.alloc(n0 * 16);
    p1.alloc(n1 * 16);
    h1.alloc(n1 * 32);
    p2.alloc(n2 * 32);
    h2.alloc(n2 * 64);
    h3.alloc(n2 * 64);
    g2.alloc(n2 * 32);
    s1.alloc(n1 * 32);
    h5.alloc(n1 * 32);
    g1.alloc(n1 * 16);
    s0.alloc(n0 * 16);
  }
  if (training && (leaves > Lt || !h6.p)) {
    Lt = L;
    const int64_t n0 = (int64_t)Lt * 512, n1 = (int64_t)Lt * 64, n2 = (int64_t)Lt * 8;
    h6.alloc(n0 * 16);
    gA0.alloc(n0 * 16);
    gB0.alloc(n0 * 16);
    gA1.alloc(n1 * 32);
    gB1.alloc(n1 * 32);
    gA2.alloc(n2 * 64);
    gB2.alloc(n2 * 64);
    if (!wpart.p) wpart.alloc(kWgradPartialFloats);
  }
}

void LearnedPrecond::load_host(const float* params) {
  const int64_t n = pc_num_params();
  own_.alloc(n);
  HNS_HIP(hipMemcpy(own_.p, params, n * sizeof(float), hipMemcpyHostToDevice));
  params_ = own_.p;
}

void LearnedPrecond::set_device_params(float* params_dev) {
  own_.reset();
  params_ = params_dev;
}

void LearnedPrecond::refresh(hipStream_t s) {
  if (!params_) fail(HNS_ERR_INVALID_ARGUMENT, "preconditioner has no weights");
  const PcOffsets& O = pc_offsets();
  if (!wfwd_.p) wfwd_.alloc(O.fwd[kPcLayers]);
  if (!wbwd_.p) wbwd_.alloc(O.bwd[kPcLayers]);
  launch_pc_swizzle(params_, wfwd_.p, wbwd_.p, s);
}

void LearnedPrecond::forward(const DevTopo& T, PcWork& W, const float* r_cs, float* z_cs, bool keep, hipStream_t s) {
  if (!wfwd_.p) refresh(s);
  W.ensure(T.L, keep);
  const PcOffsets& O = pc_offsets();
  const __bf16* wf = wfwd_.p;
  const int32_t L = T.L;
  // inference runs six leaf-tiled kernels (pool, upsample-add and the 1x1 convs fused into the
  // 3x3x3 ones); training keeps every activation and runs the layers one by one
  if (!keep && fused_inference()) {
    launch_pc_input_pool(T, r_cs, wf + O.fwd[kC0], W.h0.p, W.p1.p, s);
    launch_pc_level1(T, 0, W.p1.p, nullptr, wf + O.fwd[kC1], nullptr, W.h1.p, W.p2.p, s);
    launch_pc_level2(T, 0, W.p2.p, wf + O.fwd[kC2], nullptr, W.h2.p, s);
    launch_pc_level2(T, 1, W.h2.p, wf + O.fwd[kC3], wf + O.fwd[kU2], W.g2.p, s);
    launch_pc_level1(T, 1, W.h1.p, W.g2.p, wf + O.fwd[kC5], wf + O.fwd[kU1], nullptr, W.g1.p, s);
    launch_pc_output_fused(T, W.h0.p, W.g1.p, wf + O.fwd[kC6], params_ + O.param[kOut], z_cs, s);
    return;
  }
  launch_pc_input(T, r_cs, wf + O.fwd[kC0], W.h0.p, s);
  launch_pc_pool(L, 0, 16, W.h0.p, W.p1.p, 0.125f, s);
  launch_pc_conv(T, 1, 27, 16, 32, W.p1.p, wf + O.fwd[kC1], W.h1.p, 1, nullptr, s);
  launch_pc_pool(L, 1, 32, W.h1.p, W.p2.p, 0.125f, s);
  launch_pc_conv(T, 2, 27, 32, 64, W.p2.p, wf + O.fwd[kC2], W.h2.p, 1, nullptr, s);
  launch_pc_conv(T, 2, 27, 64, 64, W.h2.p, wf + O.fwd[kC3], W.h3.p, 1, nullptr, s);
  launch_pc_conv(T, 2, 1, 64, 32, W.h3.p, wf + O.fwd[kU2], W.g2.p, 0, nullptr, s);
  launch_pc_upadd(L, 1, 32, W.h1.p, W.g2.p, W.s1.p, 1.0f, nullptr, s);
  launch_pc_conv(T, 1, 27, 32, 32, W.s1.p, wf + O.fwd[kC5], W.h5.p, 1, nullptr, s);
  launch_pc_conv(T, 1, 1, 32, 16, W.h5.p, wf + O.fwd[kU1], W.g1.p, 0, nullptr, s);
  launch_pc_upadd(L, 0, 16, W.h0.p, W.g1.p, W.s0.p, 1.0f, nullptr, s);
  launch_pc_output(T, W.s0.p, wf + O.fwd[kC6], params_ + O.param[kOut], z_cs, keep ? W.h6.p : nullptr, s);
}

void LearnedPrecond::backward(const DevTopo& T, PcWork& W, const float* r_cs, const float* dz_cs, float* grad,
                              hipStream_t s) {
  if (!W.h6.p || W.Lt < T.L) fail(HNS_ERR_INVALID_ARGUMENT, "backward needs a forward pass that kept activations");
  const PcOffsets& O = pc_offsets();
  const __bf16* wb = wbwd_.p;
  const int32_t L = T.L;
  float* part = W.wpart.p;
  // output projection and c6
  launch_pc_output_bwd(T, dz_cs, W.h6.p, params_ + O.param[kOut], W.gA0.p, s);         // dh6 (masked)
  launch_pc_wgrad_output(T, dz_cs, W.h6.p, part, grad, s);
  launch_pc_wgrad(T, kC6, W.s0.p, W.gA0.p, part, grad, s);
  launch_pc_conv(T, 0, 27, 16, 16, W.gA0.p, wb + O.bwd[kC6], W.gB0.p, 0, nullptr, s);   // ds0
  // s0 = h0 + up(g1), g1 = u1(h5)
  launch_pc_pool(L, 0, 16, W.gB0.p, W.gA1.p, 1.0f, s);                                  // dg1
  launch_pc_wgrad(T, kU1, W.h5.p, W.gA1.p, part, grad, s);
  launch_pc_conv(T, 1, 1, 16, 32, W.gA1.p, wb + O.bwd[kU1], W.gB1.p, 2, W.h5.p, s);     // dh5 (masked)
  launch_pc_wgrad(T, kC5, W.s1.p, W.gB1.p, part, grad, s);
  launch_pc_conv(T, 1, 27, 32, 32, W.gB1.p, wb + O.bwd[kC5], W.gA1.p, 0, nullptr, s);   // ds1
  // s1 = h1 + up(g2), g2 = u2(h3)
  launch_pc_pool(L, 1, 32, W.gA1.p, W.gA2.p, 1.0f, s);                                  // dg2
  launch_pc_wgrad(T, kU2, W.h3.p, W.gA2.p, part, grad, s);
  launch_pc_conv(T, 2, 1, 32, 64, W.gA2.p, wb + O.bwd[kU2], W.gB2.p, 2, W.h3.p, s);     // dh3 (masked)
  launch_pc_wgrad(T, kC3, W.h2.p, W.gB2.p, part, grad, s);
  launch_pc_conv(T, 2, 27, 64, 64, W.gB2.p, wb + O.bwd[kC3], W.gA2.p, 2, W.h2.p, s);    // dh2 (masked)
  launch_pc_wgrad(T, kC2, W.p2.p, W.gA2.p, part, grad, s);
  launch_pc_conv(T, 2, 27, 64, 32, W.gA2.p, wb + O.bwd[kC2], W.gB2.p, 0, nullptr, s);   // dp2
  launch_pc_upadd(L, 1, 32, W.gA1.p, W.gB2.p, W.gB1.p, 0.125f, W.h1.p, s);              // dh1 (masked)
  launch_pc_wgrad(T, kC1, W.p1.p, W.gB1.p, part, grad, s);
  launch_pc_conv(T, 1, 27, 32, 16, W.gB1.p, wb + O.bwd[kC1], W.gA1.p, 0, nullptr, s);   // dp1
  launch_pc_upadd(L, 0, 16, W.gB0.p, W.gA1.p, W.gA0.p, 0.125f, W.h0.p, s);              // dh0 (masked)
  launch_pc_wgrad_input(T, r_cs, W.gA0.p, part, grad, s);
}

}  // namespace hns
