// Conjugate-gradient kernels on colour-split vectors (kernels_cg.hip).
#pragma once
#include "hns_internal.hpp"

namespace hns {

// Workgroups of every CG launch: each processes a strided run of leaves and leaves one FP64
// partial per dot product, so a consumer kernel can re-reduce the partials itself.
constexpr int kCgMaxGroups = 2048;
inline unsigned cg_groups(int32_t L) {
  const unsigned g = grid_for_leaves(L);
  return g < (unsigned)kCgMaxGroups ? g : (unsigned)kCgMaxGroups;
}

// Device scalars of one solve. rho[k] = r_k . r_k (CG) or sigma_k = d_k . A d_k (flexible CG), alpha[k] the step
// length of iteration k; flag[0] = stop, flag[1] = iterations done when stopped.
struct CgScalars {
  double* pa;     // kCgSlots rows of kCgMaxGroups partials (row = slot), written by init/replace/update and FCG
  double* pb;     // one partial per workgroup (d . q), written by the product kernel
  double* rho;    // max_iters + 1 entries
  double* alpha;  // max_iters + 1 entries
  int* flag;      // 2 entries: stop raised, stop iteration
  int* it;        // flexible CG: the device's iteration counter (0 after init, advanced by the test)
  double tol2;    // stop when r.r <= tol2 * ||b||^2
  double* bb;     // ||b||^2
};
// pa slots: r.r, and for the flexible CG d.r and d.q
enum CgSlot { kSlotRR = 0, kSlotDR = 1, kSlotDQ = 2, kCgSlots = 4 };
// the recurrence residual is replaced by b - A x every this many iterations
constexpr int kResidualReplace = 64;

// r = b = -div dx^2, x = 0; partials of ||b||^2; stop flags cleared.
void launch_cg_init(int32_t L, const float* div, float dx, float* x, float* r, const CgScalars& S, hipStream_t s);
// r = b - A x (residual replacement); partials of r.r. Periodic (force = false): skipped once
// the solve has stopped. Forced (re-testing a stop): always runs and clears the stop flags.
void launch_cg_replace(const DevTopo& T, const float* div, float dx, const float* x, float* r, const CgScalars& S,
                       bool force, hipStream_t s);
// One CG iteration `it` in two launches. Product: d_out = r + beta d_in (beta = 0 when `fresh`),
// q = A d_out. Update: x += alpha d_out, r -= alpha q.
void launch_cg_product(const DevTopo& T, int32_t it, bool fresh, const float* r, const float* d_in, float* d_out,
                       float* q, const CgScalars& S, hipStream_t s);
void launch_cg_update(int32_t L, int32_t it, float* x, float* r, const float* d, const float* q, const CgScalars& S,
                      hipStream_t s);

// Flexible CG with truncation kFcgDepth (Notay's FCG(m)): the directions d_i and products
// q_i = A d_i of the last kFcgDepth iterations stay in the rings Dr, Qr (kFcgDepth vectors of
// L * 512 each, slot i mod kFcgDepth), sigma_i = d_i . q_i in S.rho[i]. One iteration `it`: stop
// test, z = M(r), d = z - sum_i (z.q_i / sigma_i) d_i, q = A d, alpha = (d.r)/(d.q),
// x += alpha d, r -= alpha q. zq: kFcgDepth rows of kCgMaxGroups partials.
constexpr int kFcgDepth = 8;
// iterations between stop-flag reads of a flexible-CG solve (and per replayed multigrid graph)
constexpr int kFcgBatch = 4;
// The flexible-CG kernels take the iteration number from *S.it, so one iteration's launches are the
// same for every iteration (they can be replayed from a graph). Once S.flag is raised, the
// preconditioner's kernels and the rest of the iteration return at once.
// Stop test on the r.r partials (one workgroup): advance = true after an update (the next
// iteration's test; it moves *S.it on), false before iteration 0 (it records ||b||^2) and after a
// residual replacement (the same iteration tested again).
void launch_fcg_test(int32_t L, bool advance, const CgScalars& S, hipStream_t s);
// new direction and its product (three launches), orthogonalised against the last
// min(it, depth) directions (1 <= depth <= kFcgDepth); `it` (the host's count) selects the kernels
void launch_fcg_direction(const DevTopo& T, int32_t it, int depth, const float* r, const float* z, float* Dr, float* Qr,
                          double* zq, const CgScalars& S, hipStream_t s);
// x += alpha d, r -= alpha q
void launch_fcg_update(int32_t L, float* x, float* r, const float* Dr, const float* Qr, const CgScalars& S,
                       hipStream_t s);

}  // namespace hns
