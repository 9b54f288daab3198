// Conjugate gradients for the pressure Poisson problem L p = div (HNS_SOLVER_CG; no reference
// counterpart: the reference runs a fixed RBGS count, src/Cuda/HNanoSolver.cu:366-373).
//
// Operator: A = -L scaled by dx^2, (A p)_i = 6 p_i - sum of the 6 neighbours, missing neighbours
// are 0 (the same Dirichlet boundary as RBGS, Kernel.cu:477-482); right-hand side b = -div dx^2.
// Vectors are colour-split FP32 (the RBGS layout); dot products are FP64, reduced per leaf and
// then by a one-workgroup reduce kernel in a fixed order (deterministic). alpha and beta are
// formed on the device from the reduced scalars, and a device flag stops the remaining
// iterations of a batch once ||r||^2 <= tol^2 ||b||^2, so the host only synchronises per batch.
//
// Per iteration and voxel: spmv reads r, d (8 B) and writes d, q (8 B); update reads x, d, r, q
// (16 B) and writes x, r (8 B): 40 B.
#include "hns_cg.hpp"

namespace hns {
namespace {

constexpr int kTpb = 512;

// Block-wide FP64 sum over 512 lanes (8 wave64): shuffles, then LDS.
__device__ inline double block_sum(double v, double* s_red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kTpb / 64; ++i) t += s_red[i];
  }
  __syncthreads();
  return t;  // valid in lane 0
}

// One workgroup sums the per-leaf partials in a fixed order (thread t takes t, t+1024, ...,
// then a wave/LDS tree), so every reduction is deterministic. A single-address atomic counter
// across all leaf workgroups would serialise 16k device-scope atomics per reduction.
constexpr int kRed = 1024;
__device__ inline double reduce_partials(const double* __restrict__ partial, int32_t L) {
  __shared__ double s_red[kRed / 64];
  double v = 0.0;
  for (int32_t i = threadIdx.x; i < L; i += kRed) v += partial[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kRed / 64; ++i) t += s_red[i];
  return t;  // valid in thread 0
}

// mode 0: rho[0] = ||b||^2 = scal[kBB], flags reset (stop at once when b == 0)
// mode 1: scal[kSigma] = d.q (stop on breakdown)
// mode 2: rho[it+1] = r.r, iterations = it+1, stop when rho[it+1] <= tol2 ||b||^2
__global__ __launch_bounds__(kRed) void k_cg_reduce(const double* __restrict__ partial, int32_t L, int mode, int32_t it,
                                                    CgScalars S) {
  if (mode != 0 && S.flag[0]) return;
  const double t = reduce_partials(partial, L);
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    S.rho[0] = t;
    S.scal[kBB] = t;
    S.flag[0] = t == 0.0 ? 1 : 0;
    S.flag[1] = 0;
  } else if (mode == 1) {
    S.scal[kSigma] = t;
    if (!(t > 0.0)) S.flag[0] = 1;  // breakdown (A is SPD; only round-off gets here)
  } else {
    S.rho[it + 1] = t;
    S.flag[1] = it + 1;
    if (t <= S.tol2 * S.scal[kBB]) S.flag[0] = 1;
  }
}

// r = b = -div dx^2, x = 0; rho[0] = ||b||^2 = scal[kBB]
__global__ __launch_bounds__(kTpb) void k_cg_init(int32_t L, const float* __restrict__ div, float dx2,
                                                  float* __restrict__ x, float* __restrict__ r, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  const int32_t l = leaf_of_block();
  if (l >= L) return;
  const int32_t n = threadIdx.x;
  const int32_t half = L * kHalfLeaf;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t i = ((lx + ly + lz) & 1) * half + l * kHalfLeaf + cs_local(lx, ly, lz);
  const float b = -div[i] * dx2;
  x[i] = 0.f;
  r[i] = b;
  const double part = block_sum((double)b * (double)b, s_red);
  if (threadIdx.x == 0) S.partial[l] = part;
}

// Iteration `it` (0-based): d_out = r + beta d_in with beta = rho[it]/rho[it-1] (d_out = r at it 0),
// evaluated at this voxel and its 6 neighbours; q = A d_out; scal[kSigma] = d_out . q. d_in and
// d_out are distinct buffers, so no workgroup reads a value another one has already replaced.
__global__ __launch_bounds__(kTpb) void k_cg_spmv(DevTopo T, int32_t it, const float* __restrict__ d_in,
                                                  const float* __restrict__ r, float* __restrict__ d_out,
                                                  float* __restrict__ q, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  if (S.flag[0]) return;
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const FaceRow Fr = face_row(T, l);
  const int32_t half = T.L * kHalfLeaf;
  const int32_t n = threadIdx.x;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t i = ((lx + ly + lz) & 1) * half + l * kHalfLeaf + cs_local(lx, ly, lz);
  const int32_t nb[6] = {face_cs<0, 1>(Fr, lx, ly, lz, half), face_cs<0, -1>(Fr, lx, ly, lz, half),
                         face_cs<1, 1>(Fr, lx, ly, lz, half), face_cs<1, -1>(Fr, lx, ly, lz, half),
                         face_cs<2, 1>(Fr, lx, ly, lz, half), face_cs<2, -1>(Fr, lx, ly, lz, half)};
  float s = 0.f, dn;
  if (it == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (nb[k] >= 0) s += r[nb[k]];
    dn = r[i];
  } else {
    const float beta = (float)(S.rho[it] / S.rho[it - 1]);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (nb[k] >= 0) s += fmaf(beta, d_in[nb[k]], r[nb[k]]);
    dn = fmaf(beta, d_in[i], r[i]);
  }
  const float qv = 6.0f * dn - s;
  d_out[i] = dn;
  q[i] = qv;
  const double part = block_sum((double)dn * (double)qv, s_red);
  if (threadIdx.x == 0) S.partial[l] = part;
}

// alpha = rho[it]/sigma; x += alpha d; r -= alpha q; rho[it+1] = r.r; convergence flag
__global__ __launch_bounds__(kTpb) void k_cg_update(int32_t L, int32_t it, float* __restrict__ x,
                                                    float* __restrict__ r, const float* __restrict__ d,
                                                    const float* __restrict__ q, CgScalars S) {
  __shared__ double s_red[kTpb / 64];
  if (S.flag[0]) return;
  const int32_t l = leaf_of_block();
  if (l >= L) return;
  const int32_t n = threadIdx.x;
  const int32_t half = L * kHalfLeaf;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t i = ((lx + ly + lz) & 1) * half + l * kHalfLeaf + cs_local(lx, ly, lz);
  const float alpha = (float)(S.rho[it] / S.scal[kSigma]);
  const float dv = d[i];
  x[i] = fmaf(alpha, dv, x[i]);
  const float rv = fmaf(-alpha, q[i], r[i]);
  r[i] = rv;
  const double part = block_sum((double)rv * (double)rv, s_red);
  if (threadIdx.x == 0) S.partial[l] = part;
}

}  // namespace

void launch_cg_init(int32_t L, const float* div, float dx, float* x, float* r, const CgScalars& S, hipStream_t s) {
  k_cg_init<<<grid_for_leaves(L), kTpb, 0, s>>>(L, div, dx * dx, x, r, S);
  HNS_HIP(hipGetLastError());
  k_cg_reduce<<<1, kRed, 0, s>>>(S.partial, L, 0, 0, S);
  HNS_HIP(hipGetLastError());
}
void launch_cg_iteration(const DevTopo& T, int32_t it, float* x, float* r, const float* d_in, float* d_out, float* q,
                         const CgScalars& S, hipStream_t s) {
  k_cg_spmv<<<grid_for_leaves(T.L), kTpb, 0, s>>>(T, it, d_in, r, d_out, q, S);
  HNS_HIP(hipGetLastError());
  k_cg_reduce<<<1, kRed, 0, s>>>(S.partial, T.L, 1, it, S);
  HNS_HIP(hipGetLastError());
  k_cg_update<<<grid_for_leaves(T.L), kTpb, 0, s>>>(T.L, it, x, r, d_out, q, S);
  HNS_HIP(hipGetLastError());
  k_cg_reduce<<<1, kRed, 0, s>>>(S.partial, T.L, 2, it, S);
  HNS_HIP(hipGetLastError());
}

}  // namespace hns
