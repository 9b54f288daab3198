// Learned pressure preconditioner: a sparse 3-level 3x3x3 convolutional network over the active
// leaves (BF16 weights and activations, FP32 accumulation on MFMA), its training kernels, and
// the host-side sequencing of forward and backward passes (kernels_precond.hip, precond.hip).
//
// Levels share the leaf topology: level 0 cells are voxels (8^3 per leaf), level 1 cells are
// 2^3 voxels (4^3 per leaf), level 2 cells are 4^3 voxels (2^3 per leaf); a level-l tensor with
// C channels is bf16[L * S^3][C], cell = leaf * S^3 + (cx * S + cy) * S + cz, S = 8 >> l.
//
//   h0 = relu(conv3(r; c0))             1 -> 16  level 0 (r: FP32 colour-split)
//   h1 = relu(conv3(pool(h0); c1))     16 -> 32  level 1
//   h2 = relu(conv3(pool(h1); c2))     32 -> 64  level 2
//   h3 = relu(conv3(h2; c3))           64 -> 64  level 2
//   s1 = h1 + up(conv1(h3; u2))        64 -> 32  level 2 -> 1
//   h5 = relu(conv3(s1; c5))           32 -> 32  level 1
//   s0 = h0 + up(conv1(h5; u1))        32 -> 16  level 1 -> 0
//   z  = relu(conv3(s0; c6)) . w_out   16 -> 16 -> 1, z FP32 colour-split
// No biases: z(a r) = a z(r) for a > 0, so the preconditioner needs no input normalisation.
#pragma once
#include <vector>

#include "hns_internal.hpp"

namespace hns {

struct PcLayer {
  const char* name;
  int taps, cin, cout, level;
};
constexpr int kPcLayers = 9;
extern const PcLayer kPcNet[kPcLayers];
enum PcLayerId { kC0 = 0, kC1, kC2, kC3, kU2, kC5, kU1, kC6, kOut };

// Parameter offsets (FP32 master vector, layer after layer, W[tap][cin][cout]) and offsets of the
// swizzled bf16 MFMA operands (forward and, for backward-data, the mirrored transpose).
struct PcOffsets {
  int64_t param[kPcLayers + 1];
  int64_t fwd[kPcLayers + 1];
  int64_t bwd[kPcLayers + 1];
  PcOffsets();
};
const PcOffsets& pc_offsets();
inline int64_t pc_num_params() { return pc_offsets().param[kPcLayers]; }

// Activations of one forward pass (kept for the backward pass) and gradient scratch, sized for a
// topology of L leaves.
struct PcWork {
  int32_t L = 0;   // leaves the inference activations hold
  int32_t Lt = 0;  // leaves the training activations and gradients (h6, gA*, gB*) hold
  DBuf<__bf16> h0, p1, h1, p2, h2, h3, g2, s1, h5, g1, s0, h6;
  DBuf<__bf16> gA0, gB0, gA1, gB1, gA2, gB2;  // gradient ping-pong per level
  DBuf<float> wpart;                        // weight-gradient partials
  void ensure(int32_t leaves, bool training);
};

class LearnedPrecond {
 public:
  // params: FP32 master weights (pc_num_params()), device memory owned by the caller or by this
  // object when created from host data.
  void load_host(const float* params);
  void set_device_params(float* params_dev);  // shares caller memory (training)
  const float* params() const { return params_; }
  // Rebuilds the bf16 MFMA operands from the FP32 master weights.
  void refresh(hipStream_t s);
  // z = M(r) on colour-split vectors. `keep` stores every activation for a backward pass.
  void forward(const DevTopo& T, PcWork& W, const float* r_cs, float* z_cs, bool keep, hipStream_t s);
  // grad += d/dparams sum(dz . z) for the last forward(keep = true) on the same r.
  void backward(const DevTopo& T, PcWork& W, const float* r_cs, const float* dz_cs, float* grad, hipStream_t s);

 private:
  DBuf<float> own_;
  float* params_ = nullptr;
  DBuf<__bf16> wfwd_, wbwd_;
};

// ---- kernel launchers (kernels_precond.hip) -------------------------------------------
void launch_pc_swizzle(const float* params, __bf16* wfwd, __bf16* wbwd, hipStream_t s);
void launch_pc_input(const DevTopo& T, const float* r_cs, const __bf16* w, __bf16* h0, hipStream_t s);
// y = act(conv(x)), act: 0 none, 1 relu, 2 multiply by (mask > 0) (backward data through a relu)
void launch_pc_conv(const DevTopo& T, int level, int taps, int cin, int cout, const __bf16* x, const __bf16* w,
                    __bf16* y, int act, const __bf16* mask, hipStream_t s);
// z = relu(conv3(x; c6)) . w_out (colour-split FP32); h (nullable) keeps relu(conv3(x; c6)) in bf16
void launch_pc_output(const DevTopo& T, const __bf16* x, const __bf16* w, const float* params_out, float* z_cs,
                      __bf16* h, hipStream_t s);
// Fused level-0 inference (one workgroup per leaf, leaf + halo tile in LDS): h0 = relu(conv3(r; c0))
// and p1 = pool(h0); z = relu(conv3(h0 + up(g1); c6)) . w_out. Same results as the unfused kernels.
void launch_pc_input_pool(const DevTopo& T, const float* r_cs, const __bf16* w, __bf16* h0, __bf16* p1, hipStream_t s);
void launch_pc_output_fused(const DevTopo& T, const __bf16* h0, const __bf16* g1, const __bf16* w,
                            const float* params_out, float* z_cs, hipStream_t s);
// Fused level-1 inference (one wave per leaf, 6^3 tile in LDS). mode 0: ya = h1 = relu(conv3(xa; c1)),
// yb = p2 = pool(h1). mode 1: yb = g1 = u1(relu(conv3(xa + up(xb); c5))) (xa = h1, xb = g2, w2 = u1).
void launch_pc_level1(const DevTopo& T, int mode, const __bf16* xa, const __bf16* xb, const __bf16* w,
                      const __bf16* w2, __bf16* ya, __bf16* yb, hipStream_t s);
// Fused level-2 inference (8 leaves per workgroup, 4^3 tiles in LDS). mode 0: y = h2 = relu(conv3(x; c2)).
// mode 1: y = g2 = u2(relu(conv3(x; c3))) (x = h2, w2 = u2).
void launch_pc_level2(const DevTopo& T, int mode, const __bf16* x, const __bf16* w, const __bf16* w2, __bf16* y,
                      hipStream_t s);
// y_coarse = scale * sum of the 8 children of x (level `level` -> level + 1)
void launch_pc_pool(int32_t L, int level, int c, const __bf16* x, __bf16* y, float scale, hipStream_t s);
// y = a + scale * up(b) at level `level` (b at level + 1); with mask: y = (mask > 0) ? y : 0
void launch_pc_upadd(int32_t L, int level, int c, const __bf16* a, const __bf16* b, __bf16* y, float scale,
                     const __bf16* mask, hipStream_t s);
// dh = w_out * dz masked by (h > 0); wout_grad partials
void launch_pc_output_bwd(const DevTopo& T, const float* dz_cs, const __bf16* h6, const float* params_out,
                          __bf16* dh, hipStream_t s);
// grad[layer] += sum over cells of x[c + d(t)] (x) dy[c] (x bf16 at the layer's level)
void launch_pc_wgrad(const DevTopo& T, int layer, const __bf16* x, const __bf16* dy, float* partial, float* grad,
                     hipStream_t s);
void launch_pc_wgrad_input(const DevTopo& T, const float* r_cs, const __bf16* dh0, float* partial, float* grad,
                           hipStream_t s);
void launch_pc_wgrad_output(const DevTopo& T, const float* dz_cs, const __bf16* h6, float* partial, float* grad,
                            hipStream_t s);
// Adam step on n parameters (FP32), step t >= 1
void launch_pc_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                    int t, hipStream_t s);
// out = a * y + b * (A x) on colour-split vectors (A = 6 I - face neighbours); partials of
// out.out, y.y and out.x per workgroup into part[3 * cg_groups(L)], FP64 sums into sums[3].
void launch_pc_poisson(const DevTopo& T, const float* x, const float* y, float a, float b, float* out, double* part,
                       double* sums, hipStream_t s);

}  // namespace hns
