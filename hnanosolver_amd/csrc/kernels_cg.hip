// Conjugate gradients for the pressure Poisson problem L p = div (HNS_SOLVER_CG and the flexible
// preconditioned variant behind HNS_SOLVER_PCG_LEARNED; no reference counterpart: the reference
// runs a fixed RBGS count, src/Cuda/HNanoSolver.cu:366-373).
//
// Operator: A = -L scaled by dx^2, (A p)_i = 6 p_i - sum of the 6 neighbours, missing neighbours
// are 0 (the same Dirichlet boundary as RBGS, Kernel.cu:477-482); right-hand side b = -div dx^2.
// Vectors are colour-split FP32 (the RBGS layout); dot products are FP64.
//
// Two launches per iteration and no separate reduction kernel: every 256-lane workgroup walks a
// strided run of leaves (lane t owns in-plane offset 256*l + t of both colour planes) and leaves
// one FP64 partial per dot product (at most 2048 partials); the next kernel's workgroups each
// re-reduce all partials in the same fixed order, so every workgroup holds bitwise the same
// alpha, beta and stop decision, and the result is deterministic. The stop test
// (||r||^2 <= tol^2 ||b||^2) runs on the device; the host synchronises per batch. The recurrence
// residual is replaced by b - A x every kResidualReplace iterations and before a stop is
// accepted, so FP32 drift of the recurrence cannot end the solve early.
//
// The preconditioned solve is flexible CG with truncation kFcgDepth (Notay's FCG(m); below), since
// the learned preconditioner is nonlinear.
//
// Per CG iteration and voxel: product reads r, d (8 B) and writes d, q (8 B); update reads
// x, d, r, q (16 B) and writes x, r (8 B): 40 B.
#include <algorithm>

#include "hns_cg.hpp"

namespace hns {
namespace {

constexpr int kTpb = 256;

// Sum of the G partials p[0..G), fixed order, returned to all lanes.
__device__ inline double reduce_partials(const double* __restrict__ p, int G, double* s_red) {
  double v = 0.0;
  for (int i = threadIdx.x; i < G; i += kTpb) v += p[i];
  return block_sum_all(v, s_red);
}

__device__ inline void clear_flags(const CgScalars& S) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.flag[0] = 0;
    S.flag[1] = 0;
  }
}

// r = b = -div dx^2, x = 0; partials of ||b||^2 into pa slot kSlotRR; stop flags cleared
__global__ __launch_bounds__(kTpb) void k_cg_init(int32_t L, const float* __restrict__ div, float dx2,
                                                  float* __restrict__ x, float* __restrict__ r, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  const LeafRun R = leaf_run(L);
  const int32_t half = L * kHalfLeaf;
  double acc = 0.0;
  for (int32_t l = R.first; l < R.end; l += R.step) {
    const int32_t i0 = l * kHalfLeaf + threadIdx.x, i1 = half + i0;
    const float b0 = -div[i0] * dx2, b1 = -div[i1] * dx2;
    x[i0] = 0.f;
    x[i1] = 0.f;
    r[i0] = b0;
    r[i1] = b1;
    acc += (double)b0 * (double)b0 + (double)b1 * (double)b1;
  }
  const double t = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotRR * kCgMaxGroups + blockIdx.x] = t;
  clear_flags(S);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *S.it = 0;
    if (S.xdone) *S.xdone = 0;
  }
}

// Residual replacement: r = b - A x (evaluated in FP64, stored FP32); partials of r.r into pa
// slot kSlotRR. A periodic replacement (force == 0) is skipped once the solve has stopped; a
// forced one (re-testing a stop) runs and clears the stop flags.
__global__ __launch_bounds__(kTpb) void k_cg_replace(DevTopo T, const float* __restrict__ div, float dx2,
                                                     const float* __restrict__ x, float* __restrict__ r,
                                                     CgScalars S, int force) {
  __shared__ double s_red[kTpb / 64];
  __shared__ double s_tile[2][kPairTile];
  if (force == 0 && S.flag[0]) return;  // periodic: not once the solve has stopped
  if (force == 2 && !S.flag[0]) return;  // re-test of a stop: only when one was raised
  const LeafRun R = leaf_run(T.L);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const int c0 = ((t >> 5) + (t >> 2)) & 1;
  auto get = [&](int c, int32_t i) { return (double)x[c * half + i]; };
  double acc = 0.0;
  int buf = 0;
  for (int32_t l = R.first; l < R.end; l += R.step, buf ^= 1) {
    double* St = s_tile[buf];
    const FaceRow Fr = face_row(T, l);
    const int32_t i0 = c0 * half + l * kHalfLeaf + t, i1 = (c0 ^ 1) * half + l * kHalfLeaf + t;
    const double x0 = (double)x[i0], x1 = (double)x[i1];
    pair_tile_halo(Fr, get, St);
    St[c0 * kHalfLeaf + t] = x0;
    St[(c0 ^ 1) * kHalfLeaf + t] = x1;
    __syncthreads();
    double s0, s1;
    pair_tile_sums(St, t, x0, x1, s0, s1);
    const float r0 = (float)(-(double)div[i0] * (double)dx2 - (6.0 * x0 - s0));
    const float r1 = (float)(-(double)div[i1] * (double)dx2 - (6.0 * x1 - s1));
    r[i0] = r0;
    r[i1] = r1;
    acc += (double)r0 * (double)r0 + (double)r1 * (double)r1;
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotRR * kCgMaxGroups + blockIdx.x] = tt;
  if (force == 1) clear_flags(S);
}

// Iteration `it` (0-based), product half: reduce r.r from pa; stop when converged;
// d_out = r + beta d_in at the voxel pair and its neighbours;
// q = A d_out; partials of d_out . q into pb. `fresh` restarts the direction (beta = 0) and, at
// it == 0, records ||b||^2. d_in and d_out are distinct buffers, so no workgroup reads a value
// another one has already replaced.
__global__ __launch_bounds__(kTpb) void k_cg_spmv(DevTopo T, int32_t it, int fresh, const float* __restrict__ src,
                                                  const float* __restrict__ d_in, float* __restrict__ d_out,
                                                  float* __restrict__ q, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  __shared__ float s_tile[2][kPairTile];
  if (S.flag[0]) return;
  const int G = (int)gridDim.x;
  double rows[1];
  reduce_partial_rows(S.pa + kSlotRR * kCgMaxGroups, G, s_red, rows);
  const double rr = rows[0];
  const double bb = it == 0 ? rr : S.bb[0];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (rr <= S.tol2 * bb) {  // identical decision in every workgroup
    if (lead) {
      if (it == 0) S.bb[0] = bb;
      S.flag[1] = it;
      S.flag[0] = 1;
    }
    return;
  }
  const double beta = fresh ? 0.0 : rr / S.rho[it - 1];
  if (lead) {
    if (it == 0) S.bb[0] = bb;
    S.rho[it] = rr;
  }
  const float bf = (float)beta;
  const LeafRun R = leaf_run(T.L);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const int c0 = ((t >> 5) + (t >> 2)) & 1;
  double acc = 0.0;
  int buf = 0;
  // d at the leaf and its face halos is staged in LDS (double-buffered: one barrier per leaf)
  for (int32_t l = R.first; l < R.end; l += R.step, buf ^= 1) {
    float* St = s_tile[buf];
    const FaceRow Fr = face_row(T, l);
    const int32_t i0 = c0 * half + l * kHalfLeaf + t, i1 = (c0 ^ 1) * half + l * kHalfLeaf + t;
    float d0, d1;
    if (fresh) {
      auto get = [&](int c, int32_t i) { return src[c * half + i]; };
      d0 = src[i0];
      d1 = src[i1];
      pair_tile_halo(Fr, get, St);
    } else {
      auto get = [&](int c, int32_t i) { return fmaf(bf, d_in[c * half + i], src[c * half + i]); };
      d0 = fmaf(bf, d_in[i0], src[i0]);
      d1 = fmaf(bf, d_in[i1], src[i1]);
      pair_tile_halo(Fr, get, St);
    }
    St[c0 * kHalfLeaf + t] = d0;
    St[(c0 ^ 1) * kHalfLeaf + t] = d1;
    __syncthreads();
    float s0, s1;
    pair_tile_sums(St, t, d0, d1, s0, s1);
    const float q0 = 6.0f * d0 - s0, q1 = 6.0f * d1 - s1;
    d_out[i0] = d0;
    d_out[i1] = d1;
    q[i0] = q0;
    q[i1] = q1;
    acc += (double)d0 * (double)q0 + (double)d1 * (double)q1;
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pb[blockIdx.x] = tt;
}

// Update half: sigma = d.q from pb; alpha = rho[it]/sigma; x += alpha d; r -= alpha q;
// partials of r.r into pa.
__global__ __launch_bounds__(kTpb) void k_cg_update(int32_t L, int32_t it, float* __restrict__ x,
                                                    float* __restrict__ r, const float* __restrict__ d,
                                                    const float* __restrict__ q, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  if (S.flag[0]) return;
  const double sigma = reduce_partials(S.pb, (int)gridDim.x, s_red);
  if (!(sigma > 0.0)) {  // breakdown (A is SPD; only round-off gets here); same in every workgroup
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S.flag[1] = it;
      S.flag[0] = 1;
    }
    return;
  }
  const double alpha_d = S.rho[it] / sigma;
  if (blockIdx.x == 0 && threadIdx.x == 0) S.alpha[it] = alpha_d;
  const float alpha = (float)alpha_d;
  const LeafRun R = leaf_run(L);
  const int32_t half = L * kHalfLeaf;
  double acc = 0.0;
  for (int32_t l = R.first; l < R.end; l += R.step) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int32_t i = c * half + l * kHalfLeaf + threadIdx.x;
      x[i] = fmaf(alpha, d[i], x[i]);
      const float rv = fmaf(-alpha, q[i], r[i]);
      r[i] = rv;
      acc += (double)rv * (double)rv;
    }
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotRR * kCgMaxGroups + blockIdx.x] = tt;
}

// ---- flexible CG with truncation m (Notay's FCG(m)), the learned-preconditioner solve ----------
// Iteration k, after the host has applied z = M(r): the ring keeps the last m directions d_i and
// products q_i = A d_i (slot i mod m) with sigma_i = d_i . q_i in S.rho[i].
//   dots:   stop test on r . r (as the CG product); partials of z . q_i for the k' = min(k, m)
//           ring entries (i = k-1, ..., k-k'), one row of zq per entry
//   dir:    d_k = z - sum_i (z . q_i / sigma_i) d_i into slot k mod m; partials of d_k . r
//   spmv:   q_k = A d_k into slot k mod m; partials of d_k . q_k
//   update: sigma_k = d_k . q_k, alpha_k = (d_k . r) / sigma_k; x += alpha_k d_k, r -= alpha_k q_k;
//           partials of r . r
// The d_k . r and d_k . q_k partials go to the pa rows kSlotDR and kSlotDQ.
// Stop test of flexible-CG iteration it, before its preconditioner runs: one workgroup reduces the
// G partials of r.r (G = the CG grid) and raises the stop flag, on which every later kernel of the
// iteration, the preconditioner's included, returns at once. advance: first move the device
// iteration counter on (the test that follows an update); iteration 0 records ||b||^2.
// retest: only after a stop was raised (and r replaced by b - A x since): the stop is lowered and
// the same iteration tested again on the replaced residual.
__global__ __launch_bounds__(kTpb) void k_fcg_test(int G, int advance, int retest, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  if (retest ? !S.flag[0] : S.flag[0]) return;
  const int32_t it = *S.it + advance;
  double rr1[1];
  reduce_partial_rows(S.pa + kSlotRR * kCgMaxGroups, G, s_red, rr1);  // every lane has read *S.it
  const double rr = rr1[0];
  const double bb = it == 0 ? rr : S.bb[0];
  if (threadIdx.x == 0) {
    if (retest) S.flag[0] = 0;
    S.last[0] = rr;
    S.last[1] = bb;
    if (advance) *S.it = it;
    else if (S.xdone) *S.xdone = it;  // after a flush (or before iteration 0): x holds every update
    if (it == 0) S.bb[0] = bb;
    if (rr <= S.tol2 * bb) {
      S.flag[1] = it;
      S.flag[0] = 1;
    }
  }
}

// NP = min(it, kFcgDepth) ring entries in use (a template argument, so every load is unconditional)
template <int NP>
__global__ __launch_bounds__(kTpb) void k_fcg_dots(int32_t L, int64_t n, const float* __restrict__ z,
                                                   const float* __restrict__ Qr, double* __restrict__ zq,
                                                   CgScalars S) {
  __shared__ double s_red[kFcgDepth * (kTpb / 64)];
  if (NP == 0 || S.flag[0]) return;
  const int32_t it = *S.it;
  const float* q[kFcgDepth];
#pragma unroll
  for (int k = 0; k < kFcgDepth; ++k) q[k] = Qr + (int64_t)fcg_slot(it - 1 - k + kFcgDepth, S.ring) * n;
  double acc[kFcgDepth];
#pragma unroll
  for (int k = 0; k < kFcgDepth; ++k) acc[k] = 0.0;
  const LeafRun R = leaf_run(L);
  const int32_t half = L * kHalfLeaf;
  for (int32_t l = R.first; l < R.end; l += R.step) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int32_t i = c * half + l * kHalfLeaf + threadIdx.x;
      const double zv = (double)z[i];
#pragma unroll
      for (int k = 0; k < NP; ++k) acc[k] += zv * (double)q[k][i];
    }
  }
  block_store_rows(acc, s_red, zq);
}

template <int NP>
__global__ __launch_bounds__(kTpb) void k_fcg_dir(int32_t L, int64_t n, const float* __restrict__ z,
                                                  const float* __restrict__ r, float* __restrict__ Dr,
                                                  const double* __restrict__ zq, CgScalars S) {
  __shared__ double s_red[kFcgDepth * (kTpb / 64)];
  if (S.flag[0]) return;
  const int32_t it = *S.it;
  const int G = (int)gridDim.x;
  double zqk[NP > 0 ? NP : 1];
  if (NP > 0) reduce_partial_rows(zq, G, s_red, zqk);
  float bf[NP > 0 ? NP : 1];
  const float* dk[NP > 0 ? NP : 1];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    bf[k] = (float)(-zqk[k] / S.rho[it - 1 - k]);
    dk[k] = Dr + (int64_t)fcg_slot(it - 1 - k + kFcgDepth, S.ring) * n;
  }
  float* dout = Dr + (int64_t)fcg_slot(it, S.ring) * n;  // the slot of d_{it-m}, read first by the same lane
  const LeafRun R = leaf_run(L);
  const int32_t half = L * kHalfLeaf;
  double acc = 0.0;
  for (int32_t l = R.first; l < R.end; l += R.step) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int32_t i = c * half + l * kHalfLeaf + threadIdx.x;
      float d = z[i];
#pragma unroll
      for (int k = 0; k < NP; ++k) d = fmaf(bf[k], dk[k][i], d);
      dout[i] = d;
      acc += (double)d * (double)r[i];
    }
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotDR * kCgMaxGroups + blockIdx.x] = tt;
}

__global__ __launch_bounds__(kTpb) void k_fcg_spmv(DevTopo T, int64_t n, const float* __restrict__ Dr,
                                                   float* __restrict__ Qr, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  __shared__ float s_tile[2][kPairTile];
  if (S.flag[0]) return;
  const int32_t it = *S.it;
  const float* __restrict__ d = Dr + (int64_t)fcg_slot(it, S.ring) * n;
  float* __restrict__ q = Qr + (int64_t)fcg_slot(it, S.ring) * n;
  const LeafRun R = leaf_run(T.L);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const int c0 = ((t >> 5) + (t >> 2)) & 1;
  auto get = [&](int c, int32_t i) { return d[c * half + i]; };
  double acc = 0.0;
  int buf = 0;
  for (int32_t l = R.first; l < R.end; l += R.step, buf ^= 1) {
    float* St = s_tile[buf];
    const FaceRow Fr = face_row(T, l);
    const int32_t i0 = c0 * half + l * kHalfLeaf + t, i1 = (c0 ^ 1) * half + l * kHalfLeaf + t;
    const float d0 = d[i0], d1 = d[i1];
    pair_tile_halo(Fr, get, St);
    St[c0 * kHalfLeaf + t] = d0;
    St[(c0 ^ 1) * kHalfLeaf + t] = d1;
    __syncthreads();
    float s0, s1;
    pair_tile_sums(St, t, d0, d1, s0, s1);
    const float q0 = 6.0f * d0 - s0, q1 = 6.0f * d1 - s1;
    q[i0] = q0;
    q[i1] = q1;
    acc += (double)d0 * (double)q0 + (double)d1 * (double)q1;
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotDQ * kCgMaxGroups + blockIdx.x] = tt;
}

// Flexible CG with truncation 1 on the multigrid V-cycle, direction and product in one pass:
// beta = -(z . q_prev) / sigma_prev from the partials the V-cycle's last pass left in zq row 0
// (NP = 1; NP = 0 at iteration 0: d = z), d = z + beta d_prev formed for the leaf's voxel pairs and
// its face halos (the same fmaf as k_fcg_dir), q = A d, d and q into their ring slots, partials of
// d . q. The numerator of alpha, z . r, is also left by the V-cycle (pa row kSlotDR). NP = 1 also
// forms d = z when the device counter is at iteration 0, so one launch serves every iteration. A
// lane's loads for the next leaf of the run are issued before the current leaf is staged (the first
// leaf's before the partials reduce), so every workgroup keeps one leaf of loads in flight behind
// the one it computes.
template <int NP, bool PCG>
__global__ __launch_bounds__(kTpb) void k_fcg1_dirspmv(DevTopo T, int64_t n, const float* __restrict__ z,
                                                       float* __restrict__ Dr, float* __restrict__ Qr,
                                                       const double* __restrict__ zq, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  __shared__ float s_tile[2][kPairTile];
  if (S.flag[0]) return;
  const int32_t it = *S.it;
  const bool chain = NP > 0 && it > 0;  // d = z + beta d_prev; otherwise d = z
  const float* __restrict__ dp = chain ? Dr + (int64_t)fcg_slot(it - 1 + kFcgDepth, S.ring) * n : z;
  float* __restrict__ d = Dr + (int64_t)fcg_slot(it, S.ring) * n;
  float* __restrict__ q = Qr + (int64_t)fcg_slot(it, S.ring) * n;
  const LeafRun R = leaf_run(T.L);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  // lane t's pair-tile values of one leaf (pair_tile_index order): z, d_prev, and whether each
  // halo entry exists (an absent one loads element 0 and stages 0)
  struct Leaf {
    float z[4], p[4];
    bool h2, h3;
  };
  auto fetch = [&](int32_t l, const int32_t (&nb)[2], Leaf& v) {
    const PairIdx P = pair_tile_index(l, nb, t, half);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int32_t j = P.ix[k] < 0 ? 0 : P.ix[k];
      v.z[k] = z[j];
      if (NP > 0) v.p[k] = dp[j];
    }
    v.h2 = P.ix[2] >= 0;
    v.h3 = P.ix[3] >= 0;
  };
  // the halo neighbours of the next leaf (n1) are loaded one leaf earlier still (n2); past the run's
  // end both stand for leaves of the run, so every load stays in bounds and goes unused
  Leaf cur, nxt;
  int32_t n1[2], n2[2];
  if (R.first < R.end) {
    pair_halo_nbrs(T, R.first, t, n1);
    fetch(R.first, n1, cur);
    pair_halo_nbrs(T, R.first + R.step < R.end ? R.first + R.step : R.first, t, n1);
  }
  // this iteration's V-cycle has applied the deferred x update of iteration it - 1
  if (S.xdone && blockIdx.x == 0 && threadIdx.x == 0) *S.xdone = it;
  float beta = 0.f;
  if (PCG) {
    double rz1[1];
    reduce_partial_rows(S.pa + kSlotDR * kCgMaxGroups, (int)gridDim.x, s_red, rz1);
    if (blockIdx.x == 0 && threadIdx.x == 0) S.rz[it] = rz1[0];
    if (chain) beta = (float)(rz1[0] / S.rz[it - 1]);
  } else {
    if (chain) {
      double zq1[1];
      reduce_partial_rows(zq, (int)gridDim.x, s_red, zq1);
      beta = (float)(-zq1[0] / S.rho[it - 1]);
    }
    if (blockIdx.x == 0) {  // z . r once for the update (its numerator), as the update would sum it
      double zr1[1];
      reduce_partial_rows(S.pa + kSlotDR * kCgMaxGroups, (int)gridDim.x, s_red, zr1);
      if (threadIdx.x == 0) S.rz[it] = zr1[0];
    }
  }
  double acc = 0.0;
  // one leaf: the loads of the leaf after it (halo neighbours nb_next, loaded one leaf earlier) into
  // nx and of the halo neighbours of the leaf after that into nb_after, then the leaf itself from c
  auto leaf = [&](int32_t l, const Leaf& c, Leaf& nx, const int32_t (&nb_next)[2], int32_t (&nb_after)[2],
                  float* St) {
    pair_halo_nbrs(T, l + 2 * R.step < R.end ? l + 2 * R.step : l, t, nb_after);
    fetch(l + R.step < R.end ? l + R.step : l, nb_next, nx);
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = chain ? fmaf(beta, c.p[k], c.z[k]) : c.z[k];
    if (!c.h2) v[2] = 0.f;
    if (!c.h3) v[3] = 0.f;
    pair_tile_put(St, t, v[0], v[1], v[2], v[3]);
    __syncthreads();
    float s0, s1;
    pair_tile_sums(St, t, v[0], v[1], s0, s1);
    const float q0 = 6.0f * v[0] - s0, q1 = 6.0f * v[1] - s1;
    const int c0 = ((t >> 5) + (t >> 2)) & 1;
    const int32_t i0 = c0 * half + l * kHalfLeaf + t, i1 = (c0 ^ 1) * half + l * kHalfLeaf + t;
    d[i0] = v[0];
    d[i1] = v[1];
    q[i0] = q0;
    q[i1] = q1;
    acc += (double)v[0] * (double)q0 + (double)v[1] * (double)q1;
  };
  // two leaves per trip, the register sets swapping roles, so no value is copied between them (a
  // copy would wait for the loads in flight); the tile alternates as in the one-leaf loop
  for (int32_t l = R.first; l < R.end; l += 2 * R.step) {
    leaf(l, cur, nxt, n1, n2, s_tile[0]);
    if (l + R.step >= R.end) break;
    leaf(l + R.step, nxt, cur, n2, n1, s_tile[1]);
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotDQ * kCgMaxGroups + blockIdx.x] = tt;
}

__global__ __launch_bounds__(kTpb) void k_fcg_update(int32_t L, int64_t n, float* __restrict__ x,
                                                     float* __restrict__ r, const float* __restrict__ Dr,
                                                     const float* __restrict__ Qr, CgScalars S) {
  __shared__ double s_red[2 * (kTpb / 64)];
  if (S.flag[0]) return;
  const int32_t it = *S.it;
  double rows[2];  // d . r, d . q
  reduce_partial_rows(S.pa + kSlotDR * kCgMaxGroups, (int)gridDim.x, s_red, rows);
  const double sigma = rows[1];
  if (!(sigma > 0.0)) {  // breakdown (A is SPD; only round-off gets here); same in every workgroup
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S.flag[1] = it;
      S.flag[0] = 1;
    }
    return;
  }
  const double alpha_d = rows[0] / sigma;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.rho[it] = sigma;
    S.alpha[it] = alpha_d;
  }
  const float alpha = (float)alpha_d;
  const float* __restrict__ d = Dr + (int64_t)fcg_slot(it, S.ring) * n;
  const float* __restrict__ q = Qr + (int64_t)fcg_slot(it, S.ring) * n;
  const LeafRun R = leaf_run(L);
  const int32_t half = L * kHalfLeaf;
  double acc = 0.0;
  for (int32_t l = R.first; l < R.end; l += R.step) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int32_t i = c * half + l * kHalfLeaf + threadIdx.x;
      x[i] = fmaf(alpha, d[i], x[i]);
      const float rv = fmaf(-alpha, q[i], r[i]);
      r[i] = rv;
      acc += (double)rv * (double)rv;
    }
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotRR * kCgMaxGroups + blockIdx.x] = tt;
}

// k_fcg_update without the x update (deferred: S.alpha[it] and the direction ring keep what it
// needs): r -= alpha q, partials of r . r. With half the loads of the full update, the loads of two
// leaves are kept in flight: the first pair's before the partials reduce, the next pair's while a
// pair is stored. The arithmetic and the order of the r . r sums are the full update's.
__global__ __launch_bounds__(kTpb) void k_fcg_rupdate(int32_t L, int64_t n, float* __restrict__ r,
                                                      const float* __restrict__ Qr, CgScalars S) {
  __shared__ double s_red[2 * (kTpb / 64)];
  if (S.flag[0]) return;
  const int32_t it = *S.it;
  const float* __restrict__ q = Qr + (int64_t)fcg_slot(it, S.ring) * n;
  const LeafRun R = leaf_run(L);
  const int32_t half = L * kHalfLeaf;
  auto at = [&](int32_t leaf, int c) { return c * half + leaf * kHalfLeaf + (int32_t)threadIdx.x; };
  // leaves l and l + R.step (those below R.end): element 2 j + c is leaf l + j R.step, colour half c
  auto load = [&](int32_t l, float* rv, float* qv) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int32_t lj = l + j * R.step;
      if (lj < R.end) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          rv[2 * j + c] = r[at(lj, c)];
          qv[2 * j + c] = q[at(lj, c)];
        }
      }
    }
  };
  float r0[4], q0[4];
  if (R.first < R.end) load(R.first, r0, q0);
  // d . q reduced here; the numerator z . r was summed once by the product kernel into S.rz[it]
  float alpha;
  if (!fcg_step_length(S, it, (int)gridDim.x, s_red, alpha)) return;
  double acc = 0.0;
  for (int32_t l = R.first; l < R.end; l += 2 * R.step) {
    float r1[4], q1[4];
    if (l + 2 * R.step < R.end) load(l + 2 * R.step, r1, q1);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int32_t lj = l + j * R.step;
      if (lj < R.end) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const float rv = fmaf(-alpha, q0[2 * j + c], r0[2 * j + c]);
          r[at(lj, c)] = rv;
          acc += (double)rv * (double)rv;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      r0[e] = r1[e];
      q0[e] = q1[e];
    }
  }
  const double tt = block_sum_all(acc, s_red);
  if (threadIdx.x == 0) S.pa[kSlotRR * kCgMaxGroups + blockIdx.x] = tt;
}

// The pending deferred x update (`when`: an FcgFlush): x += alpha_{k-1} d_{k-1}, k = *S.it, when
// k >= 1 and *S.xdone < k, with k_fcg_update's arithmetic
__global__ __launch_bounds__(kTpb) void k_fcg_xflush(int32_t L, int64_t n, float* __restrict__ x,
                                                     const float* __restrict__ Dr, CgScalars S, int when) {
  if (when == kFlushRunning && S.flag[0]) return;
  if (when == kFlushStopped && !S.flag[0]) return;
  const int32_t k = *S.it;
  if (k < 1 || *S.xdone >= k) return;
  const float alpha = (float)S.alpha[k - 1];
  const float* __restrict__ d = Dr + (int64_t)fcg_slot(k - 1, S.ring) * n;
  const LeafRun R = leaf_run(L);
  const int32_t half = L * kHalfLeaf;
  for (int32_t l = R.first; l < R.end; l += R.step) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int32_t i = c * half + l * kHalfLeaf + threadIdx.x;
      x[i] = fmaf(alpha, d[i], x[i]);
    }
  }
}

}  // namespace

void launch_cg_init(int32_t L, const float* div, float dx, float* x, float* r, const CgScalars& S, hipStream_t s) {
  k_cg_init<<<cg_groups(L), kTpb, 0, s>>>(L, div, dx * dx, x, r, S);
  HNS_HIP(hipGetLastError());
}
void launch_cg_replace(const DevTopo& T, const float* div, float dx, const float* x, float* r, const CgScalars& S,
                       bool force, hipStream_t s) {
  k_cg_replace<<<cg_groups(T.L), kTpb, 0, s>>>(T, div, dx * dx, x, r, S, force ? 1 : 0);
  HNS_HIP(hipGetLastError());
}
void launch_cg_product(const DevTopo& T, int32_t it, bool fresh, const float* r, const float* d_in, float* d_out,
                       float* q, const CgScalars& S, hipStream_t s) {
  k_cg_spmv<<<cg_groups(T.L), kTpb, 0, s>>>(T, it, fresh ? 1 : 0, r, d_in, d_out, q, S);
  HNS_HIP(hipGetLastError());
}
void launch_cg_update(int32_t L, int32_t it, float* x, float* r, const float* d, const float* q, const CgScalars& S,
                      hipStream_t s) {
  k_cg_update<<<cg_groups(L), kTpb, 0, s>>>(L, it, x, r, d, q, S);
  HNS_HIP(hipGetLastError());
}
void launch_fcg_test(int32_t L, bool advance, const CgScalars& S, hipStream_t s) {
  k_fcg_test<<<1, kTpb, 0, s>>>((int)cg_groups(L), advance ? 1 : 0, 0, S);
  HNS_HIP(hipGetLastError());
}
void launch_fcg_retest(const DevTopo& T, const float* div, float dx, const float* x, float* r, const CgScalars& S,
                       hipStream_t s) {
  k_cg_replace<<<cg_groups(T.L), kTpb, 0, s>>>(T, div, dx * dx, x, r, S, 2);
  HNS_HIP(hipGetLastError());
  k_fcg_test<<<1, kTpb, 0, s>>>((int)cg_groups(T.L), 0, 1, S);
  HNS_HIP(hipGetLastError());
}
void launch_fcg_direction(const DevTopo& T, int32_t it, int depth, const float* r, const float* z, float* Dr, float* Qr,
                          double* zq, const CgScalars& S, hipStream_t s) {
  if (depth < 1 || depth > kFcgDepth) fail(HNS_ERR_INVALID_ARGUMENT, "flexible CG truncation out of range");
  const unsigned g = cg_groups(T.L);
  const int64_t n = (int64_t)T.L * kLeafVox;
  static_assert(kFcgDepth == 8, "dispatch below covers ring fills 0..8");
  switch (std::min<int32_t>(it, depth)) {
#define HNS_FCG_CASE(NP)                                                   \
  case NP:                                                                 \
    k_fcg_dots<NP><<<g, kTpb, 0, s>>>(T.L, n, z, Qr, zq, S);               \
    HNS_HIP(hipGetLastError());                                            \
    k_fcg_dir<NP><<<g, kTpb, 0, s>>>(T.L, n, z, r, Dr, zq, S);             \
    HNS_HIP(hipGetLastError());                                            \
    break;
    HNS_FCG_CASE(0)
    HNS_FCG_CASE(1)
    HNS_FCG_CASE(2)
    HNS_FCG_CASE(3)
    HNS_FCG_CASE(4)
    HNS_FCG_CASE(5)
    HNS_FCG_CASE(6)
    HNS_FCG_CASE(7)
    HNS_FCG_CASE(8)
#undef HNS_FCG_CASE
  }
  k_fcg_spmv<<<g, kTpb, 0, s>>>(T, n, Dr, Qr, S);
  HNS_HIP(hipGetLastError());
}
void launch_fcg1_dirspmv(const DevTopo& T, bool first, const float* z, float* Dr, float* Qr, const double* zq,
                         const CgScalars& S, hipStream_t s, bool pcg) {
  const unsigned g = cg_groups(T.L);
  const int64_t n = (int64_t)T.L * kLeafVox;
  if (pcg) {
    if (first) k_fcg1_dirspmv<0, true><<<g, kTpb, 0, s>>>(T, n, z, Dr, Qr, zq, S);
    else k_fcg1_dirspmv<1, true><<<g, kTpb, 0, s>>>(T, n, z, Dr, Qr, zq, S);
  } else {
    if (first) k_fcg1_dirspmv<0, false><<<g, kTpb, 0, s>>>(T, n, z, Dr, Qr, zq, S);
    else k_fcg1_dirspmv<1, false><<<g, kTpb, 0, s>>>(T, n, z, Dr, Qr, zq, S);
  }
  HNS_HIP(hipGetLastError());
}
void launch_fcg_update(int32_t L, float* x, float* r, const float* Dr, const float* Qr, const CgScalars& S,
                       hipStream_t s) {
  if (x) k_fcg_update<<<cg_groups(L), kTpb, 0, s>>>(L, (int64_t)L * kLeafVox, x, r, Dr, Qr, S);
  else k_fcg_rupdate<<<cg_groups(L), kTpb, 0, s>>>(L, (int64_t)L * kLeafVox, r, Qr, S);
  HNS_HIP(hipGetLastError());
}
void launch_fcg_xflush(int32_t L, float* x, const float* Dr, const CgScalars& S, FcgFlush when, hipStream_t s) {
  if (!S.xdone) fail(HNS_ERR_INVALID_ARGUMENT, "x flush without a deferred x update");
  k_fcg_xflush<<<cg_groups(L), kTpb, 0, s>>>(L, (int64_t)L * kLeafVox, x, Dr, S, (int)when);
  HNS_HIP(hipGetLastError());
}

}  // namespace hns
