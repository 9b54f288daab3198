// Conjugate-gradient kernels on colour-split vectors (kernels_cg.hip).
#pragma once
#include "hns_internal.hpp"

namespace hns {

enum CgScalarSlot { kBB = 0, kSigma = 1, kScalarSlots = 4 };

// Device scalars of one solve. rho[k] = r_k . r_k; flag[0] = stop, flag[1] = iterations done.
struct CgScalars {
  double* partial;   // one FP64 per leaf
  double* rho;       // max_iters + 1 entries
  double* scal;      // kScalarSlots entries
  int* flag;         // 2 entries
  double tol2;       // stop when rho[k] <= tol2 * ||b||^2
};

// r = b = -div dx^2, x = 0; rho[0] = ||b||^2.
void launch_cg_init(int32_t L, const float* div, float dx, float* x, float* r, const CgScalars& S, hipStream_t s);
// One CG iteration `it`: d_out = r + beta d_in, q = A d_out, x += alpha d_out, r -= alpha q.
void launch_cg_iteration(const DevTopo& T, int32_t it, float* x, float* r, const float* d_in, float* d_out, float* q,
                         const CgScalars& S, hipStream_t s);

}  // namespace hns
