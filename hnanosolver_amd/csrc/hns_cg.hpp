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

// ---- device helpers shared by the CG kernels and the multigrid pass that feeds flexible CG -------
// (256-lane workgroups; partial rows are kCgMaxGroups doubles, one entry per workgroup)
constexpr int kCgTpb = 256;
// Sum of a 256-lane value over the workgroup; the total is returned to every lane.
__device__ inline double block_sum_all(double v, double* s_red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kCgTpb / 64; ++i) t += s_red[i];
  __syncthreads();
  return t;
}

// Sums of N rows of partials (row k at p + k * kCgMaxGroups, G entries each), one pass and one
// barrier pair for all rows, each in the same fixed order as reduce_partials; all lanes get them.
template <int N>
__device__ inline void reduce_partial_rows(const double* __restrict__ p, int G, double* s_red, double (&out)[N]) {
  double v[N];
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = 0.0;
  for (int i = threadIdx.x; i < G; i += kCgTpb)
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += p[k * kCgMaxGroups + i];
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) s_red[k * (kCgTpb / 64) + (threadIdx.x >> 6)] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kCgTpb / 64; ++i) t += s_red[k * (kCgTpb / 64) + i];
    out[k] = t;
  }
  __syncthreads();
}


template <int N>
__device__ inline void block_store_rows(double (&v)[N], double* s_red, double* rows) {
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) s_red[k * (kCgTpb / 64) + (threadIdx.x >> 6)] = v[k];
  __syncthreads();
  if (threadIdx.x < N) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kCgTpb / 64; ++i) t += s_red[threadIdx.x * (kCgTpb / 64) + i];
    rows[threadIdx.x * kCgMaxGroups + blockIdx.x] = t;
  }
}


// Device scalars of one solve. rho[k] = r_k . r_k (CG) or sigma_k = d_k . A d_k (flexible CG), alpha[k] the step
// length of iteration k; flag[0] = stop, flag[1] = iterations done when stopped.
// the flexible CG's largest truncation + 1: the slots of its direction and product rings
constexpr int kFcgDepth = 8;
struct CgScalars {
  double* pa;     // kCgSlots rows of kCgMaxGroups partials (row = slot), written by init/replace/update and FCG
  double* pb;     // one partial per workgroup (d . q), written by the product kernel
  double* rho;    // max_iters + 1 entries
  double* alpha;  // max_iters + 1 entries
  double* rz;     // multigrid solve: z_k . r_k (the residual update's numerator; PCG's beta), max_iters + 1
  int* flag;      // 2 entries: stop raised, stop iteration
  int* it;        // flexible CG: the device's iteration counter (0 after init, advanced by the test)
  double* last;   // flexible CG: r.r and ||b||^2 of the last stop test
  double tol2;    // stop when r.r <= tol2 * ||b||^2
  double* bb;     // ||b||^2
  // flexible CG on the multigrid V-cycle with the x update deferred (null: not deferred): the
  // iteration k whose V-cycle last applied the pending x += alpha_{k-1} d_{k-1} (so every update
  // before k is in x), or the iteration a flush completed x at
  int* xdone = nullptr;
  // flexible CG: the slots of the direction and product rings in use (fcg_slot): kFcgDepth for the
  // general truncation, 2 for the multigrid solve's truncation 1 (the previous pair and the current
  // one), whose 8 M-voxel iterations then cycle 2 x 2 vectors of 33 MB instead of 2 x 8 and keep
  // more of their working set in the 256 MB last-level cache
  int ring = kFcgDepth;
};
// pa slots: r.r, and for the flexible CG d.r and d.q
enum CgSlot { kSlotRR = 0, kSlotDR = 1, kSlotDQ = 2, kCgSlots = 4 };

// Step length of flexible-CG iteration it, alpha = rz[it] / (d_it . q_it), d . q summed from the
// G partials of pa row kSlotDQ; workgroup 0 records S.rho[it] = d . q and S.alpha[it]. A breakdown
// (d . q not positive, which only round-off reaches, A being SPD) raises the stop flag at `it`
// instead, and every workgroup gets false. s_red: kCgTpb / 64 doubles of LDS.
__device__ inline bool fcg_step_length(const CgScalars& S, int32_t it, int G, double* s_red, float& alpha) {
  double dq[1];
  reduce_partial_rows(S.pa + kSlotDQ * kCgMaxGroups, G, s_red, dq);
  const double sigma = dq[0];
  if (!(sigma > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S.flag[1] = it;
      S.flag[0] = 1;
    }
    return false;
  }
  const double alpha_d = S.rz[it] / sigma;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.rho[it] = sigma;
    S.alpha[it] = alpha_d;
  }
  alpha = (float)alpha_d;
  return true;
}
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
// Truncation 1 on the multigrid V-cycle (whose last pass leaves z . q_prev in zq row 0 and z . r in
// pa row kSlotDR): d = z + beta d_prev and q = A d in one launch (first: iteration 0, d = z), partials
// of d . q; the update then takes alpha = (z . r) / (d . q).
// pcg: Fletcher-Reeves beta = (z_k . r_k) / (z_{k-1} . r_{k-1}) from the z . r partials instead
// (the V-cycle then leaves no z . q_prev)
void launch_fcg1_dirspmv(const DevTopo& T, bool first, const float* z, float* Dr, float* Qr, const double* zq,
                         const CgScalars& S, hipStream_t s, bool pcg = false);
// x += alpha d, r -= alpha q. zq: kFcgDepth rows of kCgMaxGroups partials.
// (ring: the slots in use, a divisor of kFcgDepth, so that i + kFcgDepth keeps the slot of i)
__device__ __forceinline__ int fcg_slot(int32_t i, int ring) { return (int)(i % ring); }
// iterations between stop-flag reads of a flexible-CG solve (and per replayed multigrid graph)
constexpr int kFcgBatch = 4;
// The flexible-CG kernels take the iteration number from *S.it, so one iteration's launches are the
// same for every iteration (they can be replayed from a graph). Once S.flag is raised, the
// preconditioner's kernels and the rest of the iteration return at once.
// Stop test on the r.r partials (one workgroup): advance = true after an update (the next
// iteration's test; it moves *S.it on), false before iteration 0 (it records ||b||^2) and after a
// residual replacement (the same iteration tested again).
void launch_fcg_test(int32_t L, bool advance, const CgScalars& S, hipStream_t s);
// If the stop was raised: r = b - A x (FP64-evaluated) and the same iteration tested again on it;
// the flag then stands only if the replaced residual passes. Nothing when no stop was raised.
void launch_fcg_retest(const DevTopo& T, const float* div, float dx, const float* x, float* r, const CgScalars& S,
                       hipStream_t s);
// new direction and its product (three launches), orthogonalised against the last
// min(it, depth) directions (1 <= depth <= kFcgDepth); `it` (the host's count) selects the kernels
void launch_fcg_direction(const DevTopo& T, int32_t it, int depth, const float* r, const float* z, float* Dr, float* Qr,
                          double* zq, const CgScalars& S, hipStream_t s);
// x += alpha d, r -= alpha q (x null: r only, the x update deferred to the next V-cycle)
void launch_fcg_update(int32_t L, float* x, float* r, const float* Dr, const float* Qr, const CgScalars& S,
                       hipStream_t s);
// Deferred x update (S.xdone set): iteration k's V-cycle applies x += alpha_{k-1} d_{k-1} on the
// CUs its one-workgroup coarse solve leaves idle, and its product kernel then records *S.xdone = k.
// Where x must be complete before that, the flush applies the pending update (k = *S.it, when
// k >= 1 and *S.xdone < k); the next advance-free stop test then records *S.xdone = k, so the next
// V-cycle does not apply it again. when: kFlushRunning (a periodic residual replacement: only while
// no stop is raised; its test follows), kFlushStopped (before a stop re-test: only once a stop is
// raised; the re-test follows), kFlushEnd (after the solve: always; nothing follows).
enum FcgFlush { kFlushRunning = 0, kFlushStopped = 1, kFlushEnd = 2 };
void launch_fcg_xflush(int32_t L, float* x, const float* Dr, const CgScalars& S, FcgFlush when, hipStream_t s);

}  // namespace hns
