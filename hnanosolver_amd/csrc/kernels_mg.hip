// Geometric multigrid V-cycle over the leaf hierarchy, the preconditioner of HNS_SOLVER_MGCG (no
// reference counterpart: the reference runs a fixed red-black Gauss-Seidel count,
// src/Cuda/HNanoSolver.cu:364-373).
//
// Levels share the leaf topology: level l has S = 8 >> l cells per leaf axis (l = 0: voxels, in
// the colour-split layout of the CG vectors; l = 1, 2, 3: leaf-dense, cell = leaf * S^3 +
// (x * S + y) * S + z, level 3 one cell per leaf). With piecewise-constant prolongation P (a coarse
// value copied to its 8 children) and restriction R = P^T (the sum of the 8 children), the Galerkin
// operator of A = 6 I - (face neighbours) is R A P = 4 A at the next level: of a child's six faces,
// three stay inside the coarse cell and the four children on a coarse face see the same coarse
// neighbour, so level l carries 4^l (6 u - sum of face neighbours), zero outside active leaves as at
// level 0. Smoothing is red-black Gauss-Seidel on each level's colouring by the parity of the global
// level coordinate (at level 3, the leaf's).
#include <algorithm>
#include <array>
#include <unordered_map>
#include <vector>

#include "hns_cg.hpp"
#include "hns_mg.hpp"

namespace hns {

const MgShape& mg_shape() {
  static const MgShape v{};
  return v;
}

namespace {

constexpr int kTpb = 256;



// Face neighbour of level cell (leaf, x, y, z) along AX by DIR (edge S = 1 << lgS): its index, or -1.
template <int AX, int DIR>
__device__ __forceinline__ int32_t mg_face(const DevTopo& T, int lgS, int32_t leaf, int x, int y, int z) {
  const int S = 1 << lgS;
  int c[3] = {x, y, z};
  c[AX] += DIR;
  int32_t nl = leaf;
  if (c[AX] < 0 || c[AX] >= S) {
    // face slots of the 27-neighbour row: (dx+1)*9 + (dy+1)*3 + (dz+1)
    const int slot = AX == 0 ? (DIR > 0 ? 22 : 4) : (AX == 1 ? (DIR > 0 ? 16 : 10) : (DIR > 0 ? 14 : 12));
    nl = T.nbr[kNbrStride * leaf + slot];
    if (nl < 0) return -1;
    c[AX] &= S - 1;
  }
  return (nl << (3 * lgS)) | (((c[0] << lgS) | c[1]) << lgS) | c[2];
}

__device__ __forceinline__ float mg_at(const float* __restrict__ u, int32_t i) { return i < 0 ? 0.f : u[i]; }

// Sum of the six face neighbours of level cell (leaf, x, y, z), in the order +x, -x, +y, -y, +z, -z.
__device__ __forceinline__ float mg_nsum(const DevTopo& T, int lgS, const float* __restrict__ u, int32_t leaf, int x,
                                         int y, int z) {
  return mg_at(u, mg_face<0, 1>(T, lgS, leaf, x, y, z)) + mg_at(u, mg_face<0, -1>(T, lgS, leaf, x, y, z)) +
         mg_at(u, mg_face<1, 1>(T, lgS, leaf, x, y, z)) + mg_at(u, mg_face<1, -1>(T, lgS, leaf, x, y, z)) +
         mg_at(u, mg_face<2, 1>(T, lgS, leaf, x, y, z)) + mg_at(u, mg_face<2, -1>(T, lgS, leaf, x, y, z));
}

// One SOR colour sweep at level lgS >= 1 (leaf-dense, S = 1 << lgS even): cells of colour COLOUR
// move towards the solution of s (6 u - sum) = b by omega times the Gauss-Seidel step. One lane per
// z-pair of cells, updating the pair's cell of this colour (the colour is the parity of the local
// coordinate: leaf origins are even at every level). FIRST: the level's u starts at zero (colour 0
// reads no neighbours and also zeroes the pair's colour-1 cell; both colours take u_old = 0).
// stop: the solve's stop flag (null: none); once raised, the kernel returns at once.
template <int COLOUR, bool FIRST>
__global__ __launch_bounds__(kTpb) void k_mg_smooth(DevTopo T, int lgS, const float* __restrict__ b,
                                                    float* __restrict__ u, float inv_s, float omega,
                                                    const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int32_t h = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  if (h >= ((T.L << (3 * lgS)) >> 1)) return;
  const int m = (1 << lgS) - 1;
  const int32_t c0 = h << 1;
  const int32_t leaf = c0 >> (3 * lgS);
  const int x = (c0 >> (2 * lgS)) & m, y = (c0 >> lgS) & m;
  const int par = (x + y + COLOUR) & 1;
  const int z = (c0 & m) | par;
  const int32_t c = c0 | par;
  if (FIRST && COLOUR == 0) u[c ^ 1] = 0.f;  // the colour-1 cell, read as a neighbour by nobody yet
  const float s = (FIRST && COLOUR == 0) ? 0.f : mg_nsum(T, lgS, u, leaf, x, y, z);
  const float gs = (b[c] * inv_s + s) * (1.0f / 6.0f);
  u[c] = FIRST ? omega * gs : fmaf(omega, gs - u[c], u[c]);  // SOR; omega = 1 is Gauss-Seidel
}

// The same SOR colour sweep on level 3 (one cell per leaf; the colour is the leaf's parity), one
// launch per colour: the path for topologies too large for k_mg_coarse.
template <int COLOUR, bool FIRST>
__global__ __launch_bounds__(kTpb) void k_mg_leaf_sweep(DevTopo T, const float* __restrict__ b, float* __restrict__ u,
                                                        float inv_s, float omega, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int32_t c = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  if (c >= T.L) return;
  const int4 o = T.origins[c];
  if ((((o.x + o.y + o.z) >> 3) & 1) != COLOUR) {
    if (FIRST && COLOUR == 0) u[c] = 0.f;
    return;
  }
  const float s = (FIRST && COLOUR == 0) ? 0.f : mg_nsum(T, 0, u, c, 0, 0, 0);
  const float gs = (b[c] * inv_s + s) * (1.0f / 6.0f);
  u[c] = FIRST ? omega * gs : fmaf(omega, gs - u[c], u[c]);
}

// Level 3 solved by `iters` red-black SOR iterations in one workgroup. The leaves are laid out in
// colour-sorted slots (Multigrid::prepare_coarse): colour 0 in slots [0, H), colour 1 in [H, 2H),
// H = kCoarseTpb * CH, unused slots padding; slot 2H is a zero that absent neighbours point at. b and
// u are in slot order too (the level-2 passes place each leaf's value at its slot), so the solve's
// loads and stores are coalesced: from leaf-ordered vectors, gathering b and scattering u through
// the slot map cost one workgroup ~16 us of single-CU address work at 8 M voxels. Padding slots of
// u receive values that nothing reads.
// u lives in LDS by slot; lane t owns slots t + k * kCoarseTpb (k < CH: colour 0, else colour 1)
// and keeps their scaled right-hand sides, old values and face-neighbour slots (16-bit pairs) in
// registers, so a colour sweep is 6 * CH LDS reads per lane, a barrier, and no branches (a padding
// slot reads only the zero slot and stays 0). Identical arithmetic, in the same order, to `iters`
// pairs of k_mg_leaf_sweep launches.
constexpr int kCoarseTpb = 1024;
constexpr int kCoarseMaxCH = 8;
// at most this many extra workgroups carry the deferred x update beside the coarse solve (each on a
// CU of its own: it holds the solve's LDS allocation), after a sleep that lets the solve's initial
// loads meet an idle memory system. Measured at 8 M voxels (the update 96 MB): beside the one-level
// solve (~50 us) 96 workgroups sleeping CH / 2 * 127 * 64 cycles (~13 us) beat 255 starting at once
// by 2 % of the fire step; beside the two-level solve (~31 us) 128 sleeping 2 * 127 * 64 cycles
// (~7 us) beat 255 starting at once by 1.5 % (128-192 workgroups and 1-2 sleeps within noise).
constexpr int kCoarseSideGroups = 96, kCoarseSideGroups2 = 128, kCoarseSleeps2 = 2;

// Workgroup g of G extra workgroups of the coarse launch: the flexible CG's deferred
// x += alpha_{k-1} d_{k-1} (k = *it; McgXUpdate) over n floats, k_fcg_update's fmaf, after
// sleeps * 127 * 64 idle cycles. Four float4 of each vector in flight per lane (n is a multiple of
// 512; a vector not 16-byte aligned takes the same loop a float at a time): with one float per load
// the update of the 8 M-voxel solve (96 MB) took ~45 us on 255 CUs and outlasted the coarse solve.
__device__ void coarse_side_xupdate_n(const McgXUpdate& X, const int* __restrict__ it, int64_t n, int g, int G,
                                      int sleeps) {
  for (int j = 0; j < sleeps; ++j) __builtin_amdgcn_s_sleep(127);
  const int32_t k = *it;
  if (k < 1 || *X.xdone >= k) return;
  const float alpha = (float)X.alpha[k - 1];
  const float* __restrict__ d = X.Dr + (int64_t)fcg_slot(k - 1, X.ring) * n;
  float* __restrict__ x = X.x;
  const int64_t stride = (int64_t)G * kCoarseTpb;
  int64_t i = (int64_t)g * kCoarseTpb + threadIdx.x;
  if (((uintptr_t)d | (uintptr_t)x) & 15) {
    for (; i < n; i += stride) x[i] = fmaf(alpha, d[i], x[i]);
    return;
  }
  const float4* __restrict__ d4 = reinterpret_cast<const float4*>(d);
  float4* __restrict__ x4 = reinterpret_cast<float4*>(x);
  const int64_t n4 = n >> 2;
  auto upd = [alpha](float4 a, float4 b) {
    return make_float4(fmaf(alpha, a.x, b.x), fmaf(alpha, a.y, b.y), fmaf(alpha, a.z, b.z), fmaf(alpha, a.w, b.w));
  };
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const float4 a0 = d4[i], a1 = d4[i + stride], a2 = d4[i + 2 * stride], a3 = d4[i + 3 * stride];
    const float4 b0 = x4[i], b1 = x4[i + stride], b2 = x4[i + 2 * stride], b3 = x4[i + 3 * stride];
    x4[i] = upd(a0, b0);
    x4[i + stride] = upd(a1, b1);
    x4[i + 2 * stride] = upd(a2, b2);
    x4[i + 3 * stride] = upd(a3, b3);
  }
  for (; i < n4; i += stride) x4[i] = upd(d4[i], x4[i]);
}

// Workgroup 0 solves the coarsest level; workgroups 1.. (launched only with X.x set) apply the
// deferred x update, on CUs the solve leaves idle.
template <int CH>
__global__ __launch_bounds__(kCoarseTpb) void k_mg_coarse(const uint32_t* __restrict__ nbp, const float* __restrict__ b,
                                                          float* __restrict__ u, float inv_s, float omega, int iters,
                                                          const int* __restrict__ stop, McgXUpdate X,
                                                          const int* __restrict__ it_dev, int64_t n, int sleeps) {
  constexpr int C = 2 * CH, H = kCoarseTpb * CH;
  __shared__ float s_u[2 * H + 1];
  __shared__ float s_bi[2 * H];  // scaled right-hand sides by slot (in LDS to leave registers for loads)
  if (stop && *stop) return;
  if (blockIdx.x > 0) {
    coarse_side_xupdate_n(X, it_dev, n, (int)blockIdx.x - 1, (int)gridDim.x - 1, sleeps);
    return;
  }
  const int t = threadIdx.x;
  float uo[C];
  uint32_t nb[C][3];
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int j = t + k * kCoarseTpb;
    nb[k][0] = nbp[j];          // +x | -x << 16
    nb[k][1] = nbp[2 * H + j];  // +y | -y << 16
    nb[k][2] = nbp[4 * H + j];  // +z | -z << 16
    s_bi[j] = b[j] * inv_s;
    uo[k] = 0.f;
  }
  for (int j = t; j <= 2 * H; j += kCoarseTpb) s_u[j] = 0.f;
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int colour = 0; colour < 2; ++colour) {
      // new values of this colour's slots first (they read only the other colour), then the stores;
      // the reads of NG slots are issued before their sums
      constexpr int NG = CH < 2 ? CH : 2;
      float v[CH];
#pragma unroll
      for (int kk = 0; kk < CH; kk += NG) {
        float a[NG][6], bb[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const int k = colour * CH + kk + g;
          bb[g] = s_bi[t + k * kCoarseTpb];
          if (!(it == 0 && colour == 0)) {
            uint32_t n0 = nb[k][0], n1 = nb[k][1], n2 = nb[k][2];
            opaque_v(n0);  // unpack here, not once for all iterations into registers
            opaque_v(n1);
            opaque_v(n2);
            a[g][0] = s_u[n0 & 0xffff];
            a[g][1] = s_u[n0 >> 16];
            a[g][2] = s_u[n1 & 0xffff];
            a[g][3] = s_u[n1 >> 16];
            a[g][4] = s_u[n2 & 0xffff];
            a[g][5] = s_u[n2 >> 16];
          }
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const int k = colour * CH + kk + g;
          float sum = 0.f;
          if (!(it == 0 && colour == 0)) sum = a[g][0] + a[g][1] + a[g][2] + a[g][3] + a[g][4] + a[g][5];
          const float gs = (bb[g] + sum) * (1.0f / 6.0f);
          v[kk + g] = it == 0 ? omega * gs : fmaf(omega, gs - uo[k], uo[k]);
        }
      }
#pragma unroll
      for (int kk = 0; kk < CH; ++kk) {
        const int k = colour * CH + kk;
        uo[k] = v[kk];
        s_u[t + k * kCoarseTpb] = v[kk];
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int k = 0; k < C; ++k) u[t + k * kCoarseTpb] = uo[k];
}

// ---------------------------------------------------------------------------------------------
// Two-level coarse solve. Level 4 groups the leaves in blocks of 2^3 (block coordinate = leaf
// coordinate >> 1, colour = the block coordinate's parity). Its operator is the Galerkin product
// P^T A3 P of level 3's A3 = 6 I - (face neighbours) under piecewise-constant P: block B carries
// D_B = the sum over its leaves of (6 - their face neighbours inside B) on the diagonal and -n_BB'
// (the leaf face pairs between B and B') towards each face-adjacent block: 24 and -4 for full
// blocks (4 A3 at the block scale), the exact weights for partial ones. The coarse solve runs nu1
// red-black Gauss-Seidel iterations (relaxation omega3) on level 3 from zero, restricts the
// residual (a block's leaves summed, times over4), runs it4 red-black SOR iterations (omega4) on
// level 4 from zero, adds each block's correction to its leaves and runs nu2 level-3 iterations in
// the reverse colour order. On the 8 M-voxel fire topology (16320 leaves, 2368 blocks) one
// iteration before and after with 6 on level 4 leaves a smaller energy-norm error than 12 SOR
// iterations (1.85) on level 3 alone, with 4 level-3 colour sweeps instead of 24.
// Level-4 slots are colour-sorted like level 3's (slot t + k * kCoarseTpb of colour k / CH4 for
// lane t, slot 2 H4 a zero); a lane keeps each slot's packed face-neighbour slots, weights n_BB'
// (3 bits each in the order +x, -x, +y, -y, +z, -z), 1 / D_B and right-hand side in registers.
constexpr int kC4MaxCH = 4;       // level-4 slots per lane and colour (the one-workgroup level-4 solve)
constexpr int kC4MaxFusedCH = 2;  // the same inside the fused two-level kernel (its registers)

struct C4Dev {
  const uint32_t* nb;    // 3 planes of 2 H4 words: slot across +a | slot across -a << 16
  const uint32_t* w;     // 2 H4 words: the six face weights
  const float* invd;     // 2 H4: 1 / D_B (0 for padding slots)
  const uint32_t* ch3;   // fused: 4 planes of 2 H4 words: level-3 slots of children 2p | 2p + 1 << 16
  const uint32_t* par3;  // fused: H words: level-4 slot of level-3 slot j | of slot j + H << 16
  const int32_t* chl;    // launches: 8 planes of 2 H4 leaves (-1: absent), child q = 4 dx + 2 dy + dz
  const int32_t* parl;   // launches: level-4 slot by leaf
  float* b4;             // launches: 2 H4 right-hand sides
  float* u4;             // launches: 2 H4 corrections
  int it4, nu1, nu2;
  float omega4, over4, omega3, omega3p;
};

template <int CH4>
struct L4Regs {
  uint32_t nb[2 * CH4][3];
  uint32_t w[2 * CH4];
  float invd[2 * CH4], b[2 * CH4], eo[2 * CH4];
};
template <int CH4>
__device__ __forceinline__ void l4_load(const C4Dev& C, L4Regs<CH4>& R) {
  constexpr int H4 = CH4 * kCoarseTpb;
#pragma unroll
  for (int k = 0; k < 2 * CH4; ++k) {
    const int j = (int)threadIdx.x + k * kCoarseTpb;
#pragma unroll
    for (int a = 0; a < 3; ++a) R.nb[k][a] = C.nb[2 * a * H4 + j];
    R.w[k] = C.w[j];
    R.invd[k] = C.invd[j];
  }
}
// iters (>= 1) red-black SOR iterations on level 4 from zero: e (LDS, by slot) with e[2 H4] = 0
// read before it is first written only by the colour-1 sweep (the caller zeroes it before the
// colour-0 sweep's barrier); ends with a barrier, the corrections in e and in R.eo
template <int CH4>
__device__ __forceinline__ void l4_sweeps(float* s_e, L4Regs<CH4>& R, int iters, float omega) {
  const int t = threadIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int colour = 0; colour < 2; ++colour) {
      float v[CH4];
#pragma unroll
      for (int kk = 0; kk < CH4; ++kk) {
        const int k = colour * CH4 + kk;
        float acc = R.b[k];
        if (!(it == 0 && colour == 0)) {
          uint32_t n0 = R.nb[k][0], n1 = R.nb[k][1], n2 = R.nb[k][2];
          opaque_v(n0);
          opaque_v(n1);
          opaque_v(n2);
          const float a0 = s_e[n0 & 0xffff], a1 = s_e[n0 >> 16], a2 = s_e[n1 & 0xffff];
          const float a3 = s_e[n1 >> 16], a4 = s_e[n2 & 0xffff], a5 = s_e[n2 >> 16];
          const uint32_t w = R.w[k];
          acc = fmaf((float)(w & 7u), a0, acc);
          acc = fmaf((float)((w >> 3) & 7u), a1, acc);
          acc = fmaf((float)((w >> 6) & 7u), a2, acc);
          acc = fmaf((float)((w >> 9) & 7u), a3, acc);
          acc = fmaf((float)((w >> 12) & 7u), a4, acc);
          acc = fmaf((float)((w >> 15) & 7u), a5, acc);
        }
        const float gs = acc * R.invd[k];
        v[kk] = it == 0 ? omega * gs : fmaf(omega, gs - R.eo[k], R.eo[k]);
      }
#pragma unroll
      for (int kk = 0; kk < CH4; ++kk) {
        const int k = colour * CH4 + kk;
        R.eo[k] = v[kk];
        s_e[t + k * kCoarseTpb] = v[kk];
      }
      __syncthreads();
    }
  }
}

// Workgroup 0 runs the two-level coarse solve with level 3 (CH slots per lane and colour) and
// level 4 (CH4) in LDS; workgroups 1.. apply the deferred x update (as k_mg_coarse's). Identical
// arithmetic, in the same order, to the launches of Multigrid's unfused two-level path. The level-3
// iterations before and after level 4 run as one loop whose body sweeps colour A, then colour B;
// the level-4 phase runs inside it after the last pre-smoothing iteration and then swaps A and B
// (registers and slot bases), so the post-smoothing's reverse colour order reuses the same code: a
// straight-line version (every sweep inlined where it runs) took ~150 instruction-cache misses per
// launch against ~70 for the loop of k_mg_coarse, ~14 us more of fixed cost at 8 M voxels.
template <int CH, int CH4>
__global__ __launch_bounds__(kCoarseTpb) void k_mg_coarse2(const uint32_t* __restrict__ nbp,
                                                           const float* __restrict__ b, float* __restrict__ u,
                                                           float inv_s, C4Dev C, const int* __restrict__ stop,
                                                           McgXUpdate X, const int* __restrict__ it_dev, int64_t n,
                                                           int sleeps) {
  constexpr int H = kCoarseTpb * CH, H4 = kCoarseTpb * CH4;
  static_assert(H4 <= H, "level 4 lives in level 3's right-hand-side array");
  __shared__ float s_u[2 * H + 1];
  __shared__ float s_x[2 * H + 1];  // level-3 right-hand sides, then the residual, then level 4's e
  if (stop && *stop) return;
  if (blockIdx.x > 0) {
    coarse_side_xupdate_n(X, it_dev, n, (int)blockIdx.x - 1, (int)gridDim.x - 1, sleeps);
    return;
  }
  const int t = threadIdx.x;
  // A: the colour swept first in an iteration (0 before level 4, 1 after), B: the other; lane t
  // holds slots base + t + kk * kCoarseTpb of each
  float ua[CH], ub[CH];
  uint32_t na[CH][3], nb[CH][3];
  int base_a = 0, base_b = H;
#pragma unroll
  for (int kk = 0; kk < CH; ++kk) {
    const int ja = t + kk * kCoarseTpb, jb = H + ja;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      na[kk][a] = nbp[2 * a * H + ja];
      nb[kk][a] = nbp[2 * a * H + jb];
    }
    s_x[ja] = b[ja] * inv_s;
    s_x[jb] = b[jb] * inv_s;
    ua[kk] = 0.f;
    ub[kk] = 0.f;
  }
  for (int j = t; j <= 2 * H; j += kCoarseTpb) s_u[j] = 0.f;
  if (t == 0) s_x[2 * H] = 0.f;
  __syncthreads();
  // one level-3 colour sweep of the slots base + t + kk * kCoarseTpb; noread: the colour-0 sweep
  // from zero (no neighbour reads); zero_old: u was zero; keep: also leave the residual in s_x (each
  // lane reads and writes only its own slots of s_x)
  auto sweep3 = [&](float(&uo)[CH], uint32_t(&nbr)[CH][3], int base, bool noread, bool zero_old, bool keep,
                    float omega) {
    constexpr int NG = CH < 2 ? CH : 2;
    float v[CH];
#pragma unroll
    for (int kk = 0; kk < CH; kk += NG) {
      float a[NG][6], bb[NG];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        bb[g] = s_x[base + t + (kk + g) * kCoarseTpb];
        if (!noread) {
          uint32_t n0 = nbr[kk + g][0], n1 = nbr[kk + g][1], n2 = nbr[kk + g][2];
          opaque_v(n0);
          opaque_v(n1);
          opaque_v(n2);
          a[g][0] = s_u[n0 & 0xffff];
          a[g][1] = s_u[n0 >> 16];
          a[g][2] = s_u[n1 & 0xffff];
          a[g][3] = s_u[n1 >> 16];
          a[g][4] = s_u[n2 & 0xffff];
          a[g][5] = s_u[n2 >> 16];
        }
      }
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int k = kk + g;
        float sum = 0.f;
        if (!noread) sum = a[g][0] + a[g][1] + a[g][2] + a[g][3] + a[g][4] + a[g][5];
        const float gs = (bb[g] + sum) * (1.0f / 6.0f);
        v[k] = zero_old ? omega * gs : fmaf(omega, gs - uo[k], uo[k]);
        if (keep) s_x[base + t + k * kCoarseTpb] = fmaf(-6.0f, v[k], bb[g] + sum);
      }
    }
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) {
      uo[kk] = v[kk];
      s_u[base + t + kk * kCoarseTpb] = v[kk];
    }
    __syncthreads();
  };
  const int nit = C.nu1 + C.nu2;
#pragma nounroll
  for (int it = 0; it < nit; ++it) {
    const float om = it < C.nu1 ? C.omega3 : C.omega3p;
    sweep3(ua, na, base_a, it == 0, it == 0, false, om);
    sweep3(ub, nb, base_b, false, it == 0, it == C.nu1 - 1, om);
    if (it != C.nu1 - 1) continue;
    // colour 0's residual (colour 1 left its own in the last sweep)
    constexpr int NG3 = CH < 2 ? CH : 2;
#pragma unroll
    for (int kk = 0; kk < CH; kk += NG3) {
      float a[NG3][6], bb[NG3];
#pragma unroll
      for (int g = 0; g < NG3; ++g) {
        const int k = kk + g;
        bb[g] = s_x[t + k * kCoarseTpb];
        uint32_t n0 = na[k][0], n1 = na[k][1], n2 = na[k][2];
        opaque_v(n0);
        opaque_v(n1);
        opaque_v(n2);
        a[g][0] = s_u[n0 & 0xffff];
        a[g][1] = s_u[n0 >> 16];
        a[g][2] = s_u[n1 & 0xffff];
        a[g][3] = s_u[n1 >> 16];
        a[g][4] = s_u[n2 & 0xffff];
        a[g][5] = s_u[n2 >> 16];
      }
#pragma unroll
      for (int g = 0; g < NG3; ++g) {
        const int k = kk + g;
        const float sum = a[g][0] + a[g][1] + a[g][2] + a[g][3] + a[g][4] + a[g][5];
        s_x[t + k * kCoarseTpb] = fmaf(-6.0f, ua[k], bb[g] + sum);
      }
    }
    // level 4's arrays through pointers made opaque here: loads at fixed addresses would otherwise
    // be hoisted out of the loop, their registers live through every sweep (spilled)
    C4Dev Cl = C;
    opaque_s(Cl.nb);
    opaque_s(Cl.w);
    opaque_s(Cl.invd);
    opaque_s(Cl.ch3);
    opaque_s(Cl.par3);
    L4Regs<CH4> R;
    l4_load(Cl, R);
    uint32_t ch[2 * CH4][4];
#pragma unroll
    for (int k = 0; k < 2 * CH4; ++k)
#pragma unroll
      for (int p = 0; p < 4; ++p) ch[k][p] = Cl.ch3[2 * p * H4 + t + k * kCoarseTpb];
    __syncthreads();
    // restriction: each block's children in the order q = 4 dx + 2 dy + dz (absent: the zero slot)
#pragma unroll
    for (int k = 0; k < 2 * CH4; ++k) {
      float acc = 0.f;
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        acc += s_x[ch[k][p] & 0xffff];
        acc += s_x[ch[k][p] >> 16];
      }
      R.b[k] = C.over4 * acc;
    }
    __syncthreads();  // every residual read before s_x holds level 4
    if (t == 0) s_x[2 * H4] = 0.f;
    l4_sweeps<CH4>(s_x, R, C.it4, C.omega4);
    uint32_t par[CH];
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) par[kk] = Cl.par3[t + kk * kCoarseTpb];
    // prolongation: each leaf adds its block's correction
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) {
      ua[kk] = s_u[t + kk * kCoarseTpb] + s_x[par[kk] & 0xffff];
      ub[kk] = s_u[H + t + kk * kCoarseTpb] + s_x[par[kk] >> 16];
    }
    // the right-hand sides again, as at the start (reloaded: registers are short here)
    float ba[CH], bb2[CH];
    const float* bl = b;
    opaque_s(bl);
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) {
      ba[kk] = bl[t + kk * kCoarseTpb] * inv_s;
      bb2[kk] = bl[H + t + kk * kCoarseTpb] * inv_s;
    }
    __syncthreads();  // every read of level 4 before s_x holds the right-hand sides again
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) {
      s_u[t + kk * kCoarseTpb] = ua[kk];
      s_u[H + t + kk * kCoarseTpb] = ub[kk];
      s_x[t + kk * kCoarseTpb] = ba[kk];
      s_x[H + t + kk * kCoarseTpb] = bb2[kk];
    }
    __syncthreads();
    // the post-smoothing sweeps colour 1 first
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) {
      const float tu = ua[kk];
      ua[kk] = ub[kk];
      ub[kk] = tu;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const uint32_t tn = na[kk][a];
        na[kk][a] = nb[kk][a];
        nb[kk][a] = tn;
      }
    }
    base_a = H;
    base_b = 0;
  }
#pragma unroll
  for (int kk = 0; kk < CH; ++kk) {
    u[base_a + t + kk * kCoarseTpb] = ua[kk];
    u[base_b + t + kk * kCoarseTpb] = ub[kk];
  }
}

// The two-level solve's launches path (topologies whose level 3 does not fit one workgroup, and the
// unfused variant): the level-3 residual (b / s + sum) - 6 u by leaf ...
__global__ __launch_bounds__(kTpb) void k_mg_leaf_residual(DevTopo T, const float* __restrict__ b,
                                                           const float* __restrict__ u, float* __restrict__ rho,
                                                           float inv_s, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int32_t c = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  if (c >= T.L) return;
  rho[c] = fmaf(-6.0f, u[c], b[c] * inv_s + mg_nsum(T, 0, u, c, 0, 0, 0));
}
// ... restricted to level 4 by slot (n4 = 2 H4 slots) ...
__global__ __launch_bounds__(kTpb) void k_mg4_restrict(C4Dev C, int32_t n4, const float* __restrict__ rho,
                                                       const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int32_t j = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  if (j >= n4) return;
  float acc = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int32_t l = C.chl[(int64_t)q * n4 + j];
    acc += l < 0 ? 0.f : rho[l];
  }
  C.b4[j] = C.over4 * acc;
}
// ... level 4 solved by one workgroup ...
template <int CH4>
__global__ __launch_bounds__(kCoarseTpb) void k_mg4_solve(C4Dev C, const int* __restrict__ stop) {
  constexpr int H4 = CH4 * kCoarseTpb;
  __shared__ float s_e[2 * H4 + 1];
  if (stop && *stop) return;
  const int t = threadIdx.x;
  L4Regs<CH4> R;
  l4_load(C, R);
#pragma unroll
  for (int k = 0; k < 2 * CH4; ++k) R.b[k] = C.b4[t + k * kCoarseTpb];
  if (t == 0) s_e[2 * H4] = 0.f;
  l4_sweeps<CH4>(s_e, R, C.it4, C.omega4);
#pragma unroll
  for (int k = 0; k < 2 * CH4; ++k) C.u4[t + k * kCoarseTpb] = R.eo[k];
}
// ... and its correction added to each leaf
__global__ __launch_bounds__(kTpb) void k_mg4_prolong(int32_t L, const int32_t* __restrict__ parl,
                                                      const float* __restrict__ u4, float* __restrict__ u,
                                                      const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int32_t c = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  if (c >= L) return;
  u[c] = u[c] + u4[parl[c]];
}

// Restriction of the level-lgS residual b - s (6 u - sum) (leaf-dense, lgS = 1 or 2) to level
// lgS + 1: one lane per fine cell forms its residual in LDS, then one lane per coarse cell sums
// its 8 children (a workgroup holds whole leaves: 256 is a multiple of S^3). map3: where level 3
// keeps a leaf's value (null: at the leaf index; the one-workgroup coarse solve's slot order).
__global__ __launch_bounds__(kTpb) void k_mg_restrict(DevTopo T, int lgS, const float* __restrict__ b,
                                                      const float* __restrict__ u, float s, float over,
                                                      float* __restrict__ bc, const int32_t* __restrict__ map3,
                                                      const int* __restrict__ stop) {
  __shared__ float s_r[kTpb];
  if (stop && *stop) return;
  const int32_t ncell = T.L << (3 * lgS);
  const int32_t f = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  const int m = (1 << lgS) - 1;
  if (f < ncell) {
    const int32_t leaf = f >> (3 * lgS);
    const int x = (f >> (2 * lgS)) & m, y = (f >> lgS) & m, z = f & m;
    s_r[threadIdx.x] = b[f] - s * (6.0f * u[f] - mg_nsum(T, lgS, u, leaf, x, y, z));
  }
  __syncthreads();
  const int lgC = lgS - 1, mc = (1 << lgC) - 1;
  const int32_t cc = (int32_t)(blockIdx.x * (kTpb / 8) + threadIdx.x);
  if (threadIdx.x >= kTpb / 8 || cc >= (T.L << (3 * lgC))) return;
  const int lb = threadIdx.x >> (3 * lgC);  // leaf within the workgroup
  const int X = (cc >> (2 * lgC)) & mc, Y = (cc >> lgC) & mc, Z = cc & mc;
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k & 1);
    sum += s_r[(lb << (3 * lgS)) | (((x << lgS) | y) << lgS) | z];
  }
  bc[lgC == 0 && map3 ? map3[cc] : cc] = over * sum;
}

// u (level lgS, leaf-dense) += its parent's value in uc (level lgS + 1; map3 as k_mg_restrict's)
__global__ __launch_bounds__(kTpb) void k_mg_prolong(int32_t L, int lgS, float* __restrict__ u,
                                                     const float* __restrict__ uc, const int32_t* __restrict__ map3,
                                                     const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int32_t c = (int32_t)(blockIdx.x * kTpb + threadIdx.x);
  if (c >= (L << (3 * lgS))) return;
  const int m = (1 << lgS) - 1, lgC = lgS - 1;
  const int32_t leaf = c >> (3 * lgS);
  const int x = (c >> (2 * lgS)) & m, y = (c >> lgS) & m, z = c & m;
  u[c] += uc[lgC == 0 && map3 ? map3[leaf] : (leaf << (3 * lgC)) | ((((x >> 1) << lgC) | (y >> 1)) << lgC) | (z >> 1)];
}

// Level 0 (colour-split): residual b - (6 u - sum) of the leaf's voxel pairs staged in LDS, then
// restricted to the leaf's 64 level-1 cells.
__global__ __launch_bounds__(kTpb) void k_mg0_restrict(DevTopo T, const float* __restrict__ b,
                                                       const float* __restrict__ u, float over,
                                                       float* __restrict__ b1, const int* __restrict__ stop) {
  __shared__ float s_tile[kPairTile];
  __shared__ float s_r[kLeafVox];
  if (stop && *stop) return;
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const FaceRow Fr = face_row(T, l);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
  const int c0 = (x + y) & 1;
  const int32_t i0 = c0 * half + l * kHalfLeaf + t, i1 = (c0 ^ 1) * half + l * kHalfLeaf + t;
  auto get = [&](int c, int32_t i) { return u[c * half + i]; };
  const float u0 = u[i0], u1 = u[i1];
  pair_tile_halo(Fr, get, s_tile);
  s_tile[c0 * kHalfLeaf + t] = u0;
  s_tile[(c0 ^ 1) * kHalfLeaf + t] = u1;
  __syncthreads();
  float s0, s1;
  pair_tile_sums(s_tile, t, u0, u1, s0, s1);
  s_r[(x << 6) | (y << 3) | (2 * zz)] = b[i0] - (6.0f * u0 - s0);
  s_r[(x << 6) | (y << 3) | (2 * zz + 1)] = b[i1] - (6.0f * u1 - s1);
  __syncthreads();
  if (t < 64) {
    const int X = t >> 4, Y = (t >> 2) & 3, Z = t & 3;
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      sum += s_r[((2 * X + (k >> 2)) << 6) | ((2 * Y + ((k >> 1) & 1)) << 3) | (2 * Z + (k & 1))];
    b1[l * 64 + t] = over * sum;
  }
}

// Level 0 (colour-split) u += parent value in u1 (level 1, leaf-dense)
__global__ __launch_bounds__(kTpb) void k_mg0_prolong(int32_t L, float* __restrict__ u, const float* __restrict__ u1,
                                                      const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int32_t l = leaf_of_block();
  if (l >= L) return;
  const int32_t half = L * kHalfLeaf;
  const int t = threadIdx.x;
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int z = 2 * zz + ((c + x + y) & 1);
    u[c * half + l * kHalfLeaf + t] += u1[l * 64 + (((x >> 1) << 4) | ((y >> 1) << 2) | (z >> 1))];
  }
}

// ---- level 0 fused: one workgroup per leaf on a 12^3 LDS tile ------------------------------
// Tile coordinates 0..11 are the leaf's local coordinates -2..9, index x * 145 + y * 12 + z (a plane
// every 145 floats, not 144: lanes staging a z face span four x planes, which 144 = 16 mod 32 put on
// the same LDS banks, four-way; 145 spreads them over all 32, and the core's reads, one x plane per
// half-wave, do not depend on it); the colour of a tile position is (x + y + z) & 1 (leaf origins are multiples of 8). A red-black
// iteration of the core needs, of the ring around it, the face slabs at depth 1 (both colours) and
// depth 2 and the depth-1 edges (colour 0 of those suffices; both are staged); corners and deeper
// edges are never read and are not staged. Values of absent leaves are 0 (the Dirichlet boundary).
constexpr int kT12 = 12, kT12P = 145, kT12Size = kT12P * (kT12 - 1) + kT12 * kT12;  // (tile floats)
__device__ __forceinline__ int t12(int x, int y, int z) { return x * kT12P + y * kT12 + z; }

struct TilePos {
  int32_t nl;      // leaf holding the position (-1: absent)
  int vx, vy, vz;  // local coordinates in that leaf
  int tx, ty, tz;  // tile coordinates
};
// face slab entry e in [0, 768): face e >> 7 (-x, +x, -y, +y, -z, +z), depth ((e >> 6) & 1) + 1,
// in-face coordinates ((e >> 3) & 7, e & 7) in x, y, z order
__device__ __forceinline__ TilePos face_pos(const FaceRow& F, int e) {
  // selects, not a switch: the staging loops stay one basic block, so all their loads overlap
  const int f = e >> 7, ax = f >> 1, pos = f & 1, d = (e >> 6) & 1, a = (e >> 3) & 7, b = e & 7;
  const int vd = pos ? d : 7 - d, td = pos ? 10 + d : 1 - d;
  TilePos p;
  p.nl = f == 0 ? F.xm : f == 1 ? F.xp : f == 2 ? F.ym : f == 3 ? F.yp : f == 4 ? F.zm : F.zp;
  p.vx = ax == 0 ? vd : a;
  p.vy = ax == 0 ? a : (ax == 1 ? vd : b);
  p.vz = ax == 2 ? vd : b;
  p.tx = ax == 0 ? td : a + 2;
  p.ty = ax == 0 ? a + 2 : (ax == 1 ? td : b + 2);
  p.tz = ax == 2 ? td : b + 2;
  return p;
}
// depth-1 edge entry e in [0, 96): edge g = e >> 3 (0-3 along z, 4-7 along y, 8-11 along x; bits
// 0 and 1 of g the signs of the two stepped axes), position e & 7 along it
__device__ __forceinline__ TilePos edge_pos(const DevTopo& T, int32_t l, int e) {
  const int g = e >> 3, w = e & 7, s0 = (g & 1) ? 1 : -1, s1 = (g & 2) ? 1 : -1;
  int d[3];
  if (g < 4) {
    d[0] = s0; d[1] = s1; d[2] = 0;
  } else if (g < 8) {
    d[0] = s0; d[1] = 0; d[2] = s1;
  } else {
    d[0] = 0; d[1] = s0; d[2] = s1;
  }
  int v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = d[k] == 0 ? w : (d[k] > 0 ? 0 : 7);
  TilePos p;
  p.nl = T.nbr[kNbrStride * l + (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)];
  p.vx = v[0]; p.vy = v[1]; p.vz = v[2];
  p.tx = v[0] + 2 + 8 * d[0]; p.ty = v[1] + 2 + 8 * d[1]; p.tz = v[2] + 2 + 8 * d[2];
  return p;
}
__device__ __forceinline__ float t12_sum(const float* S, int x, int y, int z) {  // +x, -x, +y, -y, +z, -z
  return S[t12(x + 1, y, z)] + S[t12(x - 1, y, z)] + S[t12(x, y + 1, z)] + S[t12(x, y - 1, z)] +
         S[t12(x, y, z + 1)] + S[t12(x, y, z - 1)];
}
__device__ __forceinline__ float t12_sumi(const float* S, int i) {  // the same sum at tile index i
  return S[i + kT12P] + S[i - kT12P] + S[i + kT12] + S[i - kT12] + S[i + 1] + S[i - 1];
}
// colour-1 positions updated ahead of the core's colour 0: the core's (one per lane) and those of
// the depth-1 face slabs (lanes < 192: face t >> 5, 32 positions each)
__device__ __forceinline__ bool ring_c1(const FaceRow& F, int t, int& tx, int& ty, int& tz) {
  const int f = t >> 5, a = (t >> 2) & 7, bh = t & 3;
  const TilePos p = face_pos(F, (f << 7) | (a << 3) | (2 * bh));
  tx = p.tx; ty = p.ty; tz = p.tz;
  if (((tx + ty + tz) & 1) == 0) {  // step to the colour-1 neighbour along the slab's second axis
    if (f < 4) ++tz; else ++ty;
  }
  return p.nl >= 0;
}

// Leaf-independent staging positions of the level-0 tile, one entry per lane slot, built at compile
// time so the kernels decode a table word instead of recomputing positions (the tile passes are
// bound by integer instruction issue). Face entry e (as face_pos): bits 0-10 tile index, 11-18
// colour-split local index in the neighbour, 19 its colour, 20-25 its level-1 parent cell. Edge
// entry (as edge_pos): the same, plus the 27-slot neighbour index in bits 26-30. Ring entry (as
// ring_c1, lanes < 192): tile index of the colour-1 slab position.
struct Tab12 {
  uint32_t face[768];
  uint32_t edge[96];
  uint32_t ring[192];
};
constexpr uint32_t tab_word(int tx, int ty, int tz, int vx, int vy, int vz) {
  return (uint32_t)(tx * kT12P + ty * kT12 + tz) | ((uint32_t)((vx << 5) | (vy << 2) | (vz >> 1)) << 11) |
         ((uint32_t)((vx + vy + vz) & 1) << 19) | ((uint32_t)(((vx >> 1) << 4) | ((vy >> 1) << 2) | (vz >> 1)) << 20);
}
constexpr Tab12 make_tab12() {
  Tab12 T{};
  for (int e = 0; e < 768; ++e) {
    const int f = e >> 7, ax = f >> 1, pos = f & 1, d = (e >> 6) & 1, a = (e >> 3) & 7, b = e & 7;
    const int vd = pos ? d : 7 - d, td = pos ? 10 + d : 1 - d;
    const int vx = ax == 0 ? vd : a, vy = ax == 0 ? a : (ax == 1 ? vd : b), vz = ax == 2 ? vd : b;
    const int tx = ax == 0 ? td : a + 2, ty = ax == 0 ? a + 2 : (ax == 1 ? td : b + 2), tz = ax == 2 ? td : b + 2;
    T.face[e] = tab_word(tx, ty, tz, vx, vy, vz);
    if (d == 0 && (b & 1) == 0 && (e & 127) < 64) {  // ring: the colour-1 position of the z (or y) pair
      const int r = (f << 5) | (a << 2) | (b >> 1);
      int rz = tz, ry = ty;
      if (((tx + ty + tz) & 1) == 0) {
        if (f < 4) ++rz; else ++ry;
      }
      T.ring[r] = (uint32_t)(tx * kT12P + ry * kT12 + rz);
    }
  }
  for (int e = 0; e < 96; ++e) {
    const int g = e >> 3, w = e & 7, s0 = (g & 1) ? 1 : -1, s1 = (g & 2) ? 1 : -1;
    const int d0 = g < 8 ? s0 : 0, d1 = g < 4 ? s1 : (g < 8 ? 0 : s0), d2 = g < 4 ? 0 : s1;
    const int vx = d0 == 0 ? w : (d0 > 0 ? 0 : 7), vy = d1 == 0 ? w : (d1 > 0 ? 0 : 7), vz = d2 == 0 ? w : (d2 > 0 ? 0 : 7);
    T.edge[e] = tab_word(vx + 2 + 8 * d0, vy + 2 + 8 * d1, vz + 2 + 8 * d2, vx, vy, vz) |
                ((uint32_t)((d0 + 1) * 9 + (d1 + 1) * 3 + (d2 + 1)) << 26);
  }
  return T;
}
__constant__ Tab12 kTab12 = make_tab12();
// table word -> tile index / colour-split index in leaf nl / its level-1 parent in leaf nl
static_assert(kT12Size <= 2048, "a tile index must fit the table word's 11 bits");
__device__ __forceinline__ int tw_tile(uint32_t w) { return (int)(w & 2047u); }
__device__ __forceinline__ uint32_t tw_cs(uint32_t w, uint32_t nl, uint32_t half) {
  return ((w >> 19) & 1u) * half + nl * (uint32_t)kHalfLeaf + ((w >> 11) & 255u);
}
__device__ __forceinline__ uint32_t tw_parent(uint32_t w, uint32_t nl) { return nl * 64u + ((w >> 20) & 63u); }
// element i of a device vector through a 32-bit byte offset (every vector here is below 4 GiB)
__device__ __forceinline__ float ld32(const float* __restrict__ p, uint32_t i) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + (i << 2));
}
// face f (uniform per wave) of the face row
__device__ __forceinline__ int32_t face_nl(const FaceRow& F, int f) {
  return f == 0 ? F.xm : f == 1 ? F.xp : f == 2 ? F.ym : f == 3 ? F.yp : f == 4 ? F.zm : F.zp;
}

// Pre-smoothing and restriction of the V-cycle's level 0 in one pass per leaf: the pointwise
// (Jacobi) pre-smoothing step from z = 0, z' = c r with c = jacobi weight * 0.166666667f, the
// residual r - A z' (A = 6 I - neighbours, in the order +x, -x, +y, -y, +z, -z) and its
// restriction to level 1 (times `over`). z' is a function of r alone, so only the level-1
// right-hand side is written: the post pass forms z' again from r. Each leaf's z' and its six face
// halos are staged in LDS (the pair tile), then the residual tile: two barriers per leaf,
// persistent over a strided leaf run. A lane's loads for the next leaf of the run are issued before
// the current leaf is staged, from halo neighbours loaded one leaf earlier still (pair_halo_nbrs),
// two leaves per trip with the register sets swapping roles. The arithmetic of a pointwise scaling
// followed by k_mg0_restrict, so the unfused V-cycle gives the same bits.
// TEST: the grid has 8 more workgroups in front, and workgroup 0 runs the flexible CG's stop test
// after the previous update beside the pass (k_fcg_test with advance, on the G = cg_groups(L)
// partials of r . r); the later kernels of the iteration read its verdict. On a stop this pass's
// output goes unused.
template <bool TEST>
__global__ __launch_bounds__(kTpb) void k_mg0_pre(DevTopo T, const float* __restrict__ b, float c, float over,
                                                  float* __restrict__ b1, const int* __restrict__ stop, CgScalars S) {
  __shared__ float s_t[2][kPairTile];
  __shared__ float s_r[2][kLeafVox];
  __shared__ double s_red[kTpb / 64];
  if (stop && *stop) return;
  if (TEST && blockIdx.x < 8) {
    if (blockIdx.x == 0) {
      double rr1[1];
      reduce_partial_rows(S.pa + kSlotRR * kCgMaxGroups, (int)gridDim.x - 8, s_red, rr1);
      if (threadIdx.x == 0) {
        const double rr = rr1[0], bb = S.bb[0];
        const int32_t it = *S.it + 1;
        S.last[0] = rr;
        S.last[1] = bb;
        *S.it = it;
        if (rr <= S.tol2 * bb) {
          S.flag[1] = it;
          S.flag[0] = 1;
        }
      }
    }
    return;
  }
  const LeafRun R = TEST ? leaf_run(T.L, (int32_t)blockIdx.x - 8, (int32_t)gridDim.x - 8) : leaf_run(T.L);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
  // lane t's r values of a leaf's pair tile (pair_tile_index order) and whether each halo entry
  // exists (an absent one loads element 0 and stages 0)
  struct Leaf {
    float r[4];
    bool h2, h3;
  };
  auto fetch = [&](int32_t l, const int32_t (&nb)[2], Leaf& v) {
    const PairIdx P = pair_tile_index(l, nb, t, half);
#pragma unroll
    for (int k = 0; k < 4; ++k) v.r[k] = b[P.ix[k] < 0 ? 0 : P.ix[k]];
    v.h2 = P.ix[2] >= 0;
    v.h3 = P.ix[3] >= 0;
  };
  auto leaf = [&](int32_t l, const Leaf& cl, Leaf& nx, const int32_t (&nb_next)[2], int32_t (&nb_after)[2],
                  float* St, float* Sr) {
    pair_halo_nbrs(T, l + 2 * R.step < R.end ? l + 2 * R.step : l, t, nb_after);
    fetch(l + R.step < R.end ? l + R.step : l, nb_next, nx);
    const float r0 = cl.r[0], r1 = cl.r[1];
    const float u0 = c * r0, u1 = c * r1;
    pair_tile_put(St, t, u0, u1, cl.h2 ? c * cl.r[2] : 0.f, cl.h3 ? c * cl.r[3] : 0.f);
    __syncthreads();
    float s0, s1;
    pair_tile_sums(St, t, u0, u1, s0, s1);
    Sr[(x << 6) | (y << 3) | (2 * zz)] = r0 - (6.0f * u0 - s0);
    Sr[(x << 6) | (y << 3) | (2 * zz + 1)] = r1 - (6.0f * u1 - s1);
    __syncthreads();
    if (t < 64) {
      const int X = t >> 4, Y = (t >> 2) & 3, Z = t & 3;
      float sum = 0.f;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        sum += Sr[((2 * X + (k >> 2)) << 6) | ((2 * Y + ((k >> 1) & 1)) << 3) | (2 * Z + (k & 1))];
      b1[l * 64 + t] = over * sum;
    }
  };
  if (R.first >= R.end) return;
  Leaf A, B;
  int32_t n1[2], n2[2];
  pair_halo_nbrs(T, R.first, t, n1);
  fetch(R.first, n1, A);
  pair_halo_nbrs(T, R.first + R.step < R.end ? R.first + R.step : R.first, t, n1);
  for (int32_t l = R.first; l < R.end; l += 2 * R.step) {
    leaf(l, A, B, n1, n2, s_t[0], s_r[0]);
    if (l + R.step >= R.end) break;
    leaf(l + R.step, B, A, n2, n1, s_t[1], s_r[1]);
  }
}

// z = c r on a colour-split level-0 vector (the unfused V-cycle's pointwise pre-smoothing)
__global__ __launch_bounds__(kTpb) void k_mg0_jacobi(int64_t n, const float* __restrict__ r, float c,
                                                     float* __restrict__ z, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int64_t i = (int64_t)blockIdx.x * kTpb + threadIdx.x;
  if (i < n) z[i] = c * r[i];
}

// Prolongation and post-smoothing of the V-cycle's level 0 for leaf l: the pre-smoothed z' = c b
// plus the parent value in u1 staged on the tile, then one red-black iteration on A z = b in the
// reverse order (colour 1 on the core and the depth-1 slabs, then colour 0 on the core), the core
// written to zout. Bitwise the result of k_mg0_jacobi, k_mg0_prolong and two k_rbgs sweeps.
// mg0_post_nbrs loads the neighbour leaves the lane's staging reads of a leaf come from (a leaf ahead
// of its data, so the data loads do not wait on them); mg0_post_load issues the leaf's global reads;
// mg0_post_compute stages and updates the tile and returns the lane's final colour-0 / colour-1
// values (v0, v1) and their right-hand sides (r0, r1).
struct Post0Nbr {
  int32_t fnl[3];  // the lane's face-slab entries: faces (t >> 7) + 2 k (-x, -y, -z below lane 128)
  int32_t enl;     // its depth-1 edge entry
  int32_t rnl;     // lanes t < 192: face t >> 5, whose depth-1 slab positions it updates
};
__device__ __forceinline__ void mg0_post_nbrs(const DevTopo& T, int32_t l, uint32_t ew, Post0Nbr& N) {
  // the face entries are uniform per wave: scalar loads through the constant address space (the
  // neighbour rows are not written while a V-cycle runs)
  typedef const __attribute__((address_space(4))) int32_t* ConstRow;
  const int t = threadIdx.x;
  const int hi = __builtin_amdgcn_readfirstlane(t >> 7);
  const ConstRow row = (ConstRow)(T.nbr + kNbrStride * l);
  N.fnl[0] = row[hi ? 22 : 4];
  N.fnl[1] = row[hi ? 16 : 10];
  N.fnl[2] = row[hi ? 14 : 12];
  const int32_t* r = T.nbr + kNbrStride * l;
  N.enl = r[(int)(ew >> 26)];
  // face f = t >> 5 (-x, +x, -y, +y, -z, +z) at row entry 13 -+ 9, 3, 1 (4, 22, 10, 16, 12, 14)
  const int f = t >> 5, d = f < 2 ? 9 : (f < 4 ? 3 : 1);
  N.rnl = r[f < 6 ? 13 + ((f & 1) ? d : -d) : 13];
}
struct Post0In {
  int32_t fnl[3], enl, rnl;
  float fu[3], fb[3], eb, eu, up, b0, b1;
  float q0, q1;  // Q: the previous product at the lane's pair (z . q_prev)
};
template <bool Q>
__device__ __forceinline__ void mg0_post_load(const DevTopo& T, int32_t l, const float* __restrict__ b,
                                              const float* __restrict__ u1, const uint32_t (&fw)[3], uint32_t ew,
                                              const Post0Nbr& N, const float* __restrict__ qp, Post0In& P) {
  const int t = threadIdx.x;
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
  const int c0 = (x + y) & 1;
  const uint32_t uh = (uint32_t)(T.L * kHalfLeaf), ul = (uint32_t)l;
  P.enl = N.enl;
  P.rnl = N.rnl;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    P.fnl[k] = N.fnl[k];
    const uint32_t src = P.fnl[k] < 0 ? ul : (uint32_t)P.fnl[k];
    P.fu[k] = ld32(u1, tw_parent(fw[k], src));
    P.fb[k] = ld32(b, tw_cs(fw[k], src, uh));
  }
  P.up = ld32(u1, ul * 64u + (uint32_t)(((x >> 1) << 4) + ((y >> 1) << 2) + zz));  // the pair's parent
  P.b0 = ld32(b, (uint32_t)c0 * uh + ul * kHalfLeaf + t);
  P.b1 = ld32(b, (uint32_t)(c0 ^ 1) * uh + ul * kHalfLeaf + t);
  const uint32_t esrc = P.enl < 0 ? ul : (uint32_t)P.enl;
  P.eb = ld32(b, tw_cs(ew, esrc, uh));
  P.eu = ld32(u1, tw_parent(ew, esrc));
  if (Q) {
    P.q0 = ld32(qp, ul * kHalfLeaf + t);
    P.q1 = ld32(qp, uh + ul * kHalfLeaf + t);
  }
}
__device__ __forceinline__ void mg0_post_compute(const DevTopo& T, int32_t l, const Post0In& P,
                                                 const uint32_t (&fw)[3], uint32_t ew, uint32_t rw,
                                                 float* __restrict__ zout,
                                                 float* s_b, float* s_z, float c, float w, float& v0, float& v1,
                                                 float& r0, float& r1) {
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const float dx2 = -1.0f;
  const int x = t >> 5, y = (t >> 2) & 7, zz = t & 3;
  const int c0 = (x + y) & 1;
  {
    // the pair in colour order (z parity alternating with y across a half-wave: conflict-free banks)
    const int i0 = t12(x + 2, y + 2, 2 * zz + 2);
    const float z0 = c * P.b0 + P.up, z1 = c * P.b1 + P.up;
    s_z[i0 + c0] = c0 ? z1 : z0;
    s_z[i0 + (c0 ^ 1)] = c0 ? z0 : z1;
    s_b[i0 + c0] = c0 ? P.b1 : P.b0;
    s_b[i0 + (c0 ^ 1)] = c0 ? P.b0 : P.b1;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int i = tw_tile(fw[k]);
    s_z[i] = P.fnl[k] < 0 ? 0.f : c * P.fb[k] + P.fu[k];
    s_b[i] = P.fnl[k] < 0 ? 0.f : P.fb[k];  // read at depth 1 only
  }
  // lanes past the 96 edge entries store into the pad slot past the tile (no branch around the
  // edge loads, so they go out with the others)
  s_z[t < 96 ? tw_tile(ew) : kT12Size] = P.enl < 0 ? 0.f : c * P.eb + P.eu;
  __syncthreads();
  // colour 1 on the core and the depth-1 slabs
  {
    const int zc1 = 2 * zz + (c0 ^ 1);
    const int i = t12(x + 2, y + 2, zc1 + 2);
    s_z[i] = sor_update(t12_sum(s_z, x + 2, y + 2, zc1 + 2), s_b[i], s_z[i], dx2, w);
  }
  if (t < 192 && P.rnl >= 0) {
    const int i = (int)rw;
    s_z[i] = sor_update(t12_sumi(s_z, i), s_b[i], s_z[i], dx2, w);
  }
  __syncthreads();
  // colour 0 on the core, then the core out
  const int zc0 = 2 * zz + c0;
  const int i = t12(x + 2, y + 2, zc0 + 2), i1 = t12(x + 2, y + 2, 2 * zz + (c0 ^ 1) + 2);
  v0 = sor_update(t12_sum(s_z, x + 2, y + 2, zc0 + 2), s_b[i], s_z[i], dx2, w);
  v1 = s_z[i1];
  r0 = s_b[i];
  r1 = s_b[i1];
  zout[l * kHalfLeaf + t] = v0;         // colour 0 plane: the lane's colour-0 voxel
  zout[half + l * kHalfLeaf + t] = v1;  // colour 1 plane
}
__device__ __forceinline__ void lane_tables(uint32_t (&fw)[3], uint32_t& ew, uint32_t& rw) {
  const int t = threadIdx.x;
  ew = kTab12.edge[t < 96 ? t : 0];
  rw = kTab12.ring[t < 192 ? t : 0];
#pragma unroll
  for (int k = 0; k < 3; ++k) fw[k] = kTab12.face[t + k * kTpb];
}

// One workgroup per leaf (the V-cycle on its own).
__global__ __launch_bounds__(kTpb) void k_mg0_post(DevTopo T, const float* __restrict__ b,
                                                   const float* __restrict__ u1, float* __restrict__ zout, float c,
                                                   float w, const int* __restrict__ stop) {
  __shared__ float s_b[kT12Size];
  __shared__ float s_z[kT12Size + 1];  // + the pad slot of lanes without an edge entry
  if (stop && *stop) return;
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  uint32_t fw[3], ew, rw;
  lane_tables(fw, ew, rw);
  Post0Nbr N;
  mg0_post_nbrs(T, l, ew, N);
  Post0In P;
  mg0_post_load<false>(T, l, b, u1, fw, ew, N, nullptr, P);
  float v0, v1, r0, r1;
  mg0_post_compute(T, l, P, fw, ew, rw, zout, s_b, s_z, c, w, v0, v1, r0, r1);
}

// Inside flexible CG (truncation 1): the same pass over a strided leaf run per workgroup
// (cg_groups(L) of them), also leaving the partials of z . r (b = r) and, from iteration 1, of
// z . q_prev for the direction and step length.
__global__ __launch_bounds__(kTpb) void k_mg0_post_fcg(DevTopo T, const float* __restrict__ b,
                                                       const float* __restrict__ u1, float* __restrict__ zout,
                                                       McgHook H, float c, float w, const int* __restrict__ stop) {
  __shared__ float s_b[kT12Size];
  __shared__ float s_z[kT12Size + 1];  // + the pad slot of lanes without an edge entry
  __shared__ double s_red[kTpb / 64];
  if (stop && *stop) return;
  const int32_t it = *H.it;
  const int32_t half = T.L * kHalfLeaf;
  const float* __restrict__ qp = H.Qr + (int64_t)fcg_slot(it - 1 + kFcgDepth, H.ring) * H.n;
  const LeafRun R = leaf_run(T.L);
  double zr = 0.0, zq = 0.0;
  uint32_t fw[3], ew, rw;
  lane_tables(fw, ew, rw);
  // the next leaf's neighbour indices are loaded while the current leaf computes, and the current
  // leaf's reads (the previous product's included) go out together (loading a whole leaf of data
  // ahead costs more in occupancy than it hides); past the run's end the indices of a leaf of the
  // run stand in, loaded and unused
  Post0Nbr nc, nn;
  if (R.first < R.end) mg0_post_nbrs(T, R.first, ew, nc);
  for (int32_t l = R.first; l < R.end; l += R.step) {
    mg0_post_nbrs(T, l + R.step < R.end ? l + R.step : l, ew, nn);
    Post0In cur;
    mg0_post_load<true>(T, l, b, u1, fw, ew, nc, qp, cur);
    float v0, v1, r0, r1;
    mg0_post_compute(T, l, cur, fw, ew, rw, zout, s_b, s_z, c, w, v0, v1, r0, r1);
    zr += (double)v0 * (double)r0 + (double)v1 * (double)r1;
    if (H.zq && it > 0) zq += (double)v0 * (double)cur.q0 + (double)v1 * (double)cur.q1;
    __syncthreads();  // the tile is restaged for the next leaf
    nc = nn;
  }
  const double tr = block_sum_all(zr, s_red);
  const double tq = block_sum_all(zq, s_red);
  if (threadIdx.x == 0) {
    H.zr[blockIdx.x] = tr;
    if (H.zq) H.zq[blockIdx.x] = tq;
  }
}

// ---- levels 1 and 2 fused: the same two passes on leaf-dense levels ---------------------------
// S = 1 << LG cells per leaf axis (4 at level 1, 2 at level 2), one lane per cell, 256 / S^3 leaves
// per workgroup, each on its own (S + 4)^3 tile (local coordinates -2..S+1). Arithmetic of
// k_mg_smooth (omega = 1), k_mg_restrict and k_mg_prolong, in the same order.
template <int LG>
struct LTile {
  static constexpr int S = 1 << LG, E = S + 4, N = E * E * E, C = S * S * S, LPW = kTpb / C;
  static constexpr int NF = 12 * S * S / C, NE = (12 * S + C - 1) / C;  // face / edge entries per lane
  __device__ static int at(int x, int y, int z) { return (x * E + y) * E + z; }
  __device__ static int32_t cell(int32_t leaf, int x, int y, int z) {
    return (leaf << (3 * LG)) | (((x << LG) | y) << LG) | z;
  }
  // face slab entry e in [0, 12 S^2): face e >> (2 LG + 1), depth ((e >> 2 LG) & 1) + 1
  __device__ static TilePos face(const FaceRow& F, int e) {
    const int f = e >> (2 * LG + 1), ax = f >> 1, pos = f & 1, d = (e >> (2 * LG)) & 1;
    const int a = (e >> LG) & (S - 1), b = e & (S - 1);
    const int vd = pos ? d : S - 1 - d, td = pos ? S + 2 + d : 1 - d;
    TilePos p;
    p.nl = f == 0 ? F.xm : f == 1 ? F.xp : f == 2 ? F.ym : f == 3 ? F.yp : f == 4 ? F.zm : F.zp;
    p.vx = ax == 0 ? vd : a;
    p.vy = ax == 0 ? a : (ax == 1 ? vd : b);
    p.vz = ax == 2 ? vd : b;
    p.tx = ax == 0 ? td : a + 2;
    p.ty = ax == 0 ? a + 2 : (ax == 1 ? td : b + 2);
    p.tz = ax == 2 ? td : b + 2;
    return p;
  }
  // depth-1 edge entry e in [0, 12 S), as edge_pos
  __device__ static TilePos edge(const DevTopo& T, int32_t l, int e) {
    const int g = e >> LG, w = e & (S - 1), s0 = (g & 1) ? 1 : -1, s1 = (g & 2) ? 1 : -1;
    int d[3];
    if (g < 4) {
      d[0] = s0; d[1] = s1; d[2] = 0;
    } else if (g < 8) {
      d[0] = s0; d[1] = 0; d[2] = s1;
    } else {
      d[0] = 0; d[1] = s0; d[2] = s1;
    }
    int v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = d[k] == 0 ? w : (d[k] > 0 ? 0 : S - 1);
    TilePos p;
    p.nl = T.nbr[kNbrStride * l + (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)];
    p.vx = v[0]; p.vy = v[1]; p.vz = v[2];
    p.tx = v[0] + 2 + S * d[0]; p.ty = v[1] + 2 + S * d[1]; p.tz = v[2] + 2 + S * d[2];
    return p;
  }
  __device__ static float nsum(const float* A, int x, int y, int z) {  // +x, -x, +y, -y, +z, -z
    return A[at(x + 1, y, z)] + A[at(x - 1, y, z)] + A[at(x, y + 1, z)] + A[at(x, y - 1, z)] + A[at(x, y, z + 1)] +
           A[at(x, y, z - 1)];
  }
};

// Level LG's colour-1 positions ahead of the core's colour 0: the core's (the lane's cell, if of
// colour 1) and the depth-1 face slabs' (6 S^2 / 2 of them, entries lane, lane + C, ...).
template <int LG, class F>
__device__ __forceinline__ void ltile_c1(const FaceRow& Fr, int lane, int cx, int cy, int cz, F&& update) {
  using Tl = LTile<LG>;
  if (((cx + cy + cz) & 1) == 1) update(cx + 2, cy + 2, cz + 2);
  for (int e = lane; e < 3 * Tl::S * Tl::S; e += Tl::C) {
    const int f = e / (Tl::S * Tl::S / 2), j = e % (Tl::S * Tl::S / 2);
    const int a = j / (Tl::S / 2), bh = j % (Tl::S / 2);
    const TilePos p = Tl::face(Fr, (f << (2 * LG + 1)) | (a << LG) | (2 * bh));
    if (p.nl < 0) continue;
    int tx = p.tx, ty = p.ty, tz = p.tz;
    if (((tx + ty + tz) & 1) == 0) {
      if (f < 4) ++tz; else ++ty;
    }
    update(tx, ty, tz);
  }
}

// Level 1 (S = 4, one wave per leaf) from compile-time staging tables, as kTab12 serves level 0:
// the passes were bound by VALU issue (~330 instructions a wave, most of them the generic tile
// addressing of LTile). Per lane (the leaf's cell lane = (x << 4) | (y << 2) | z):
//   face[k * 64 + lane]: face-slab entry lane + 64 k of LTile<2>::face (face 2 k + (lane >> 5)):
//       tile index (bits 0-8), source cell in the neighbour (9-14), its level-2 parent (15-17)
//   edge[lane]: depth-1 edge entry min(lane, 47) of LTile<2>::edge: the same fields (lanes 48-63
//       repeat entry 47, so that every lane stores: the same value to the same place)
//   ring[lane] (lanes < 48): the colour-1 slab position ltile_c1<2> updates, tile index (bits 0-8)
//       on face lane >> 3
struct TabL1 {
  uint32_t face[192];
  uint32_t edge[64];
  uint32_t ring[64];
};
constexpr uint32_t tl1_word(int tx, int ty, int tz, int vx, int vy, int vz) {
  return (uint32_t)((tx * 8 + ty) * 8 + tz) | ((uint32_t)((((vx << 2) | vy) << 2) | vz) << 9) |
         ((uint32_t)((((vx >> 1) << 1) | (vy >> 1)) << 1 | (vz >> 1)) << 15);
}
constexpr TabL1 make_tabl1() {
  TabL1 T{};
  for (int e = 0; e < 192; ++e) {  // LTile<2>::face
    const int f = e >> 5, ax = f >> 1, pos = f & 1, d = (e >> 4) & 1, a = (e >> 2) & 3, b = e & 3;
    const int vd = pos ? d : 3 - d, td = pos ? 6 + d : 1 - d;
    const int vx = ax == 0 ? vd : a, vy = ax == 0 ? a : (ax == 1 ? vd : b), vz = ax == 2 ? vd : b;
    const int tx = ax == 0 ? td : a + 2, ty = ax == 0 ? a + 2 : (ax == 1 ? td : b + 2), tz = ax == 2 ? td : b + 2;
    T.face[e] = tl1_word(tx, ty, tz, vx, vy, vz);
  }
  for (int lane = 0; lane < 64; ++lane) {  // LTile<2>::edge
    const int e = lane < 48 ? lane : 47;
    const int g = e >> 2, w = e & 3, s0 = (g & 1) ? 1 : -1, s1 = (g & 2) ? 1 : -1;
    const int d0 = g < 8 ? s0 : 0, d1 = g < 4 ? s1 : (g < 8 ? 0 : s0), d2 = g < 4 ? 0 : s1;
    const int vx = d0 == 0 ? w : (d0 > 0 ? 0 : 3), vy = d1 == 0 ? w : (d1 > 0 ? 0 : 3), vz = d2 == 0 ? w : (d2 > 0 ? 0 : 3);
    T.edge[lane] = tl1_word(vx + 2 + 4 * d0, vy + 2 + 4 * d1, vz + 2 + 4 * d2, vx, vy, vz);
  }
  for (int lane = 0; lane < 48; ++lane) {  // ltile_c1<2>'s slab entry
    const int f = lane / 8, j = lane % 8, a = j / 2, bh = j % 2;
    const int e = (f << 5) | (a << 2) | (2 * bh);
    const int ax = f >> 1, pos = f & 1, d = (e >> 4) & 1, aa = (e >> 2) & 3, b = e & 3;
    const int td = pos ? 6 + d : 1 - d;
    int tx = ax == 0 ? td : aa + 2, ty = ax == 0 ? aa + 2 : (ax == 1 ? td : b + 2), tz = ax == 2 ? td : b + 2;
    if (((tx + ty + tz) & 1) == 0) {
      if (f < 4) ++tz;
      else ++ty;
    }
    T.ring[lane] = (uint32_t)((tx * 8 + ty) * 8 + tz);
  }
  return T;
}
__constant__ TabL1 kTabL1 = make_tabl1();

// face_row of a wave-uniform leaf through scalar loads (the constant address space: the neighbour
// rows are not written while a kernel reads them)
__device__ __forceinline__ FaceRow face_row_s(const DevTopo& T, int32_t l) {
  typedef const __attribute__((address_space(4))) int32_t* ConstRow;
  const ConstRow r = (ConstRow)(T.nbr + kNbrStride * l);
  return FaceRow{l, r[4], r[22], r[10], r[16], r[12], r[14]};
}
// the 27-slot neighbour index of LTile<2>::edge's entry group g = e >> 2, five bits per group
constexpr uint64_t make_edge_slots() {
  uint64_t k = 0;
  for (int g = 0; g < 12; ++g) {
    const int s0 = (g & 1) ? 1 : -1, s1 = (g & 2) ? 1 : -1;
    const int d0 = g < 8 ? s0 : 0, d1 = g < 4 ? s1 : (g < 8 ? 0 : s0), d2 = g < 4 ? 0 : s1;
    k |= (uint64_t)((d0 + 1) * 9 + (d1 + 1) * 3 + (d2 + 1)) << (5 * g);
  }
  return k;
}
constexpr uint64_t kEdgeSlots = make_edge_slots();
__device__ __forceinline__ int edge_slot(int lane) { return (int)((kEdgeSlots >> (5 * (min(lane, 47) >> 2))) & 31u); }

// the six neighbours' sum at tile index i of an 8^3 tile, +x, -x, +y, -y, +z, -z (LTile<2>::nsum)
__device__ __forceinline__ float tl1_nsum(const float* A, int i) {
  return A[i + 64] + A[i - 64] + A[i + 8] + A[i - 8] + A[i + 1] + A[i - 1];
}

// Level 1 leaves per wave: the wave's leaves l0, l0 + 4, ... (l0 = 4 NL slot + wave) have all their
// loads issued up front, and the wave then stages and smooths them one after the other in its one
// tile. At 8M voxels one leaf per wave asks for twice the chip's 8,192 wave slots, so half the waves
// wait for the first half to finish and the chain of dependent loads is paid twice; with two leaves
// per wave every wave is resident at once.
constexpr int kL1Leaves = 2;

// bit f set where face f (-x, +x, -y, +y, -z, +z) of the leaf has a neighbour
__device__ __forceinline__ uint32_t face_mask(const FaceRow& Fr) {
  return (uint32_t)(Fr.xm >= 0) | (uint32_t)(Fr.xp >= 0) << 1 | (uint32_t)(Fr.ym >= 0) << 2 |
         (uint32_t)(Fr.yp >= 0) << 3 | (uint32_t)(Fr.zm >= 0) << 4 | (uint32_t)(Fr.zp >= 0) << 5;
}

// k_mgl_pre<2> (level 1: one red-black iteration from zero, the residual and its restriction to
// level 2) with the staging positions from kTabL1: the same values in the same order.
template <int NL>
__global__ __launch_bounds__(kTpb) void k_mgl1_pre(DevTopo T, const float* __restrict__ b, float* __restrict__ zout,
                                                   float scale, float over, float* __restrict__ bc, float w,
                                                   const int* __restrict__ stop) {
  __shared__ float s_b[4][512];
  __shared__ float s_z[4][512];
  __shared__ float s_r[kTpb];
  if (stop && *stop) return;
  const int lw = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int32_t l0 = xcd_slot() * (4 * NL) + lw;
  const float inv_s = 1.0f / scale;
  const int cx = lane >> 4, cy = (lane >> 2) & 3, cz = lane & 3;
  const int ic = ((cx + 2) * 8 + cy + 2) * 8 + cz + 2;  // the lane's cell in the tile
  float* B = s_b[lw];
  float* Z = s_z[lw];
  uint32_t fw[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) fw[k] = kTabL1.face[lane + 64 * k];
  const uint32_t ew = kTabL1.edge[lane], rw = kTabL1.ring[lane];
  const int es = edge_slot(lane);
  // every load of the wave's leaves issued before the first LDS store, the neighbour indices of
  // every leaf ahead of every data load: the vector-memory counter completes in order (absent
  // leaves load the own leaf, then read 0)
  FaceRow Fr[NL];
  float fv[NL][3], bv[NL], ev[NL];
  int32_t nl[NL][3], enl[NL];
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const int32_t l = l0 + 4 * n;
    if (l >= T.L) break;
    Fr[n] = face_row_s(T, l);
    enl[n] = T.nbr[kNbrStride * l + es];
  }
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const int32_t l = l0 + 4 * n;
    if (l >= T.L) break;
    const bool hi = lane & 32;  // face 2 k + 1 of entry k, else face 2 k
    nl[n][0] = hi ? Fr[n].xp : Fr[n].xm;
    nl[n][1] = hi ? Fr[n].yp : Fr[n].ym;
    nl[n][2] = hi ? Fr[n].zp : Fr[n].zm;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      fv[n][k] = b[((nl[n][k] < 0 ? l : nl[n][k]) << 6) | (int32_t)((fw[k] >> 9) & 63u)];
    }
    bv[n] = b[(l << 6) | lane];
    ev[n] = b[((enl[n] < 0 ? l : enl[n]) << 6) | (int32_t)((ew >> 9) & 63u)];
  }
  __builtin_amdgcn_sched_barrier(0);  // no use of a loaded value scheduled among the loads
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const int32_t l = l0 + 4 * n;
    const bool live = l < T.L;  // wave-uniform
    if (n > 0) __syncthreads();  // the tile's previous leaf fully read
    if (live) {
      auto put = [&](int i, float v) {
        B[i] = v;
        Z[i] = (((i >> 6) + (i >> 3) + i) & 1) == 0 ? w * ((v * inv_s + 0.f) * (1.0f / 6.0f)) : 0.f;
      };
      put(ic, bv[n]);
#pragma unroll
      for (int k = 0; k < 3; ++k) put((int)(fw[k] & 511u), nl[n][k] < 0 ? 0.f : fv[n][k]);
      put((int)(ew & 511u), enl[n] < 0 ? 0.f : ev[n]);  // lanes 48-63 repeat lane 47's entry
    }
    __syncthreads();
    if (live) {
      auto update = [&](int i) {
        const float gs = (B[i] * inv_s + tl1_nsum(Z, i)) * (1.0f / 6.0f);
        Z[i] = w * gs;
      };
      if (((cx + cy + cz) & 1) == 1) update(ic);
      if (lane < 48 && ((face_mask(Fr[n]) >> (lane >> 3)) & 1u)) update((int)rw);
    }
    __syncthreads();
    if (live) {
      const float u = Z[ic];
      s_r[threadIdx.x] = B[ic] - scale * (6.0f * u - tl1_nsum(Z, ic));
      zout[(l << 6) | lane] = u;
    }
    __syncthreads();
    if (live && lane < 8) {
      const int X = lane >> 2, Y = (lane >> 1) & 1, Z0 = lane & 1;
      float sum = 0.f;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z0 + (k & 1);
        sum += s_r[(lw << 6) | (((x << 2) | y) << 2) | z];
      }
      bc[((int32_t)l << 3) | lane] = over * sum;
    }
  }
}

// k_mgl_post<2> (level 1: the prolongation from level 2 and one red-black iteration in the reverse
// colour order) with the staging positions from kTabL1: the same values in the same order.
template <int NL>
__global__ __launch_bounds__(kTpb) void k_mgl1_post(DevTopo T, const float* __restrict__ b,
                                                    const float* __restrict__ zin, const float* __restrict__ uc,
                                                    float* __restrict__ zout, float scale, float w,
                                                    const int* __restrict__ stop) {
  __shared__ float s_b[4][512];
  __shared__ float s_z[4][512];
  if (stop && *stop) return;
  const int lw = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int32_t l0 = xcd_slot() * (4 * NL) + lw;
  const float inv_s = 1.0f / scale;
  const int cx = lane >> 4, cy = (lane >> 2) & 3, cz = lane & 3;
  const int ic = ((cx + 2) * 8 + cy + 2) * 8 + cz + 2;
  float* B = s_b[lw];
  float* Z = s_z[lw];
  uint32_t fw[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) fw[k] = kTabL1.face[lane + 64 * k];
  const uint32_t ew = kTabL1.edge[lane], rw = kTabL1.ring[lane];
  const int es = edge_slot(lane);
  const uint32_t cw = (uint32_t)(((((cx >> 1) << 1) | (cy >> 1)) << 1 | (cz >> 1)) << 15) | ((uint32_t)lane << 9);
  // the prolonged value of cell `word` in leaf nl of own leaf l (absent leaves load the own leaf,
  // then read 0)
  auto zval = [&](int32_t l, int32_t nl, uint32_t word) {
    const int32_t src = nl < 0 ? l : nl;
    const float v = zin[(src << 6) | (int32_t)((word >> 9) & 63u)] + uc[(src << 3) | (int32_t)((word >> 15) & 7u)];
    return nl < 0 ? 0.f : v;
  };
  // every load of the wave's leaves issued before the first LDS store, the neighbour indices of
  // every leaf ahead of every data load
  FaceRow Fr[NL];
  float fz[NL][3], fb[NL][3], zc[NL], bc0[NL], ez[NL];
  int32_t nl[NL][3], enl[NL];
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const int32_t l = l0 + 4 * n;
    if (l >= T.L) break;
    Fr[n] = face_row_s(T, l);
    enl[n] = T.nbr[kNbrStride * l + es];
  }
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const int32_t l = l0 + 4 * n;
    if (l >= T.L) break;
    const bool hi = lane & 32;  // face 2 k + 1 of entry k, else face 2 k
    nl[n][0] = hi ? Fr[n].xp : Fr[n].xm;
    nl[n][1] = hi ? Fr[n].yp : Fr[n].ym;
    nl[n][2] = hi ? Fr[n].zp : Fr[n].zm;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      fz[n][k] = zval(l, nl[n][k], fw[k]);
      fb[n][k] = b[((nl[n][k] < 0 ? l : nl[n][k]) << 6) | (int32_t)((fw[k] >> 9) & 63u)];
    }
    zc[n] = zval(l, l, cw);
    bc0[n] = b[(l << 6) | lane];
    ez[n] = zval(l, enl[n], ew);
  }
  __builtin_amdgcn_sched_barrier(0);
  auto sweep = [&](int i) {
    const float gs = (B[i] * inv_s + tl1_nsum(Z, i)) * (1.0f / 6.0f);
    Z[i] = fmaf(w, gs - Z[i], Z[i]);
  };
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const int32_t l = l0 + 4 * n;
    const bool live = l < T.L;  // wave-uniform
    if (n > 0) __syncthreads();  // the tile's previous leaf fully read
    if (live) {
      Z[ic] = zc[n];
      B[ic] = bc0[n];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int j = (int)(fw[k] & 511u);
        Z[j] = fz[n][k];
        B[j] = nl[n][k] < 0 ? 0.f : fb[n][k];  // read at depth 1 only
      }
      Z[(int)(ew & 511u)] = ez[n];  // lanes 48-63 repeat lane 47's entry
    }
    __syncthreads();
    if (live) {
      if (((cx + cy + cz) & 1) == 1) sweep(ic);
      if (lane < 48 && ((face_mask(Fr[n]) >> (lane >> 3)) & 1u)) sweep((int)rw);
    }
    __syncthreads();
    if (live) {
      if (((cx + cy + cz) & 1) == 0) sweep(ic);
      zout[(l << 6) | lane] = Z[ic];
    }
  }
}

// map3 (LG = 1, restricting to level 3): as k_mg_restrict's
template <int LG>
__global__ __launch_bounds__(kTpb) void k_mgl_pre(DevTopo T, const float* __restrict__ b, float* __restrict__ zout,
                                                  float scale, float over, float* __restrict__ bc, float w,
                                                  const int32_t* __restrict__ map3, const int* __restrict__ stop) {
  using Tl = LTile<LG>;
  __shared__ float s_b[Tl::LPW][Tl::N];
  __shared__ float s_z[Tl::LPW][Tl::N];
  __shared__ float s_r[kTpb];
  if (stop && *stop) return;
  const int lw = threadIdx.x >> (3 * LG), lane = threadIdx.x & (Tl::C - 1);
  const int32_t l = xcd_slot() * Tl::LPW + lw;
  const bool live = l < T.L;
  const float inv_s = 1.0f / scale;
  const int cx = lane >> (2 * LG), cy = (lane >> LG) & (Tl::S - 1), cz = lane & (Tl::S - 1);
  float* B = s_b[lw];
  float* Z = s_z[lw];
  // restricting to level 3: the leaf's slot there, loaded with the leaf's data (the store address at
  // the end would otherwise wait for it after the last barrier)
  const int32_t slot3 = LG == 1 && map3 && live ? map3[l] : 0;
  FaceRow Fr{};
  if (live) {
    Fr = face_row(T, l);
    auto put = [&](int tx, int ty, int tz, float bv) {
      const int i = Tl::at(tx, ty, tz);
      B[i] = bv;
      Z[i] = ((tx + ty + tz) & 1) == 0 ? w * ((bv * inv_s + 0.f) * (1.0f / 6.0f)) : 0.f;
    };
    // every load issued before the first LDS store (absent leaves load the own leaf, then read 0)
    TilePos fp[Tl::NF], ep[Tl::NE];
    float fv[Tl::NF], ev[Tl::NE];
#pragma unroll
    for (int k = 0; k < Tl::NE; ++k) ep[k] = Tl::edge(T, l, std::min(lane + k * Tl::C, 12 * Tl::S - 1));
#pragma unroll
    for (int k = 0; k < Tl::NF; ++k) {
      fp[k] = Tl::face(Fr, lane + k * Tl::C);
      fv[k] = b[Tl::cell(fp[k].nl < 0 ? l : fp[k].nl, fp[k].vx, fp[k].vy, fp[k].vz)];
    }
    const float bv = b[Tl::cell(l, cx, cy, cz)];
#pragma unroll
    for (int k = 0; k < Tl::NE; ++k) ev[k] = b[Tl::cell(ep[k].nl < 0 ? l : ep[k].nl, ep[k].vx, ep[k].vy, ep[k].vz)];
    put(cx + 2, cy + 2, cz + 2, bv);
#pragma unroll
    for (int k = 0; k < Tl::NF; ++k) put(fp[k].tx, fp[k].ty, fp[k].tz, fp[k].nl < 0 ? 0.f : fv[k]);
#pragma unroll
    for (int k = 0; k < Tl::NE; ++k)
      if (lane + k * Tl::C < 12 * Tl::S) put(ep[k].tx, ep[k].ty, ep[k].tz, ep[k].nl < 0 ? 0.f : ev[k]);
  }
  __syncthreads();
  if (live)
    ltile_c1<LG>(Fr, lane, cx, cy, cz, [&](int tx, int ty, int tz) {
      const int i = Tl::at(tx, ty, tz);
      const float gs = (B[i] * inv_s + Tl::nsum(Z, tx, ty, tz)) * (1.0f / 6.0f);
      Z[i] = w * gs;
    });
  __syncthreads();
  if (live) {
    const int i = Tl::at(cx + 2, cy + 2, cz + 2);
    const float u = Z[i];
    s_r[threadIdx.x] = B[i] - scale * (6.0f * u - Tl::nsum(Z, cx + 2, cy + 2, cz + 2));
    zout[Tl::cell(l, cx, cy, cz)] = u;
  }
  __syncthreads();
  // one lane per coarse cell: the leaf's S^3 / 8 children sums
  constexpr int CC = Tl::C / 8;
  if (!live || lane >= CC) return;
  const int lgC = LG - 1, mc = (1 << lgC) - 1;
  const int X = lane >> (2 * lgC), Y = (lane >> lgC) & mc, Z0 = lane & mc;
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z0 + (k & 1);
    sum += s_r[(lw << (3 * LG)) | (((x << LG) | y) << LG) | z];
  }
  bc[LG == 1 && map3 ? slot3 : ((int32_t)l << (3 * lgC)) | lane] = over * sum;
}

// map3 (LG = 1, prolonging from level 3): as k_mg_restrict's. PRO = false: one more red-black
// iteration on zin alone (no prolongation; uc and map3 unused), the post-smoothing's next iteration.
template <int LG, bool PRO = true>
__global__ __launch_bounds__(kTpb) void k_mgl_post(DevTopo T, const float* __restrict__ b,
                                                   const float* __restrict__ zin, const float* __restrict__ uc,
                                                   float* __restrict__ zout, float scale, float w,
                                                   const int32_t* __restrict__ map3, const int* __restrict__ stop) {
  using Tl = LTile<LG>;
  __shared__ float s_b[Tl::LPW][Tl::N];
  __shared__ float s_z[Tl::LPW][Tl::N];
  if (stop && *stop) return;
  const int lw = threadIdx.x >> (3 * LG), lane = threadIdx.x & (Tl::C - 1);
  const int32_t l = xcd_slot() * Tl::LPW + lw;
  const bool live = l < T.L;
  const float inv_s = 1.0f / scale;
  const int cx = lane >> (2 * LG), cy = (lane >> LG) & (Tl::S - 1), cz = lane & (Tl::S - 1);
  constexpr int lgC = LG - 1;
  float* B = s_b[lw];
  float* Z = s_z[lw];
  auto zval = [&](int32_t nl, int vx, int vy, int vz) {  // absent leaves load the own leaf, then read 0
    const int32_t src = nl < 0 ? l : nl;
    float v = zin[Tl::cell(src, vx, vy, vz)];
    if (PRO)
      v = v + uc[LG == 1 && map3 ? map3[src]
                                 : (src << (3 * lgC)) | ((((vx >> 1) << lgC) | (vy >> 1)) << lgC) | (vz >> 1)];
    return nl < 0 ? 0.f : v;
  };
  FaceRow Fr{};
  if (live) {
    Fr = face_row(T, l);
    // every load issued before the first LDS store
    TilePos fp[Tl::NF], ep[Tl::NE];
    float fz[Tl::NF], fb[Tl::NF], ez[Tl::NE];
#pragma unroll
    for (int k = 0; k < Tl::NE; ++k) ep[k] = Tl::edge(T, l, std::min(lane + k * Tl::C, 12 * Tl::S - 1));
#pragma unroll
    for (int k = 0; k < Tl::NF; ++k) {
      fp[k] = Tl::face(Fr, lane + k * Tl::C);
      fz[k] = zval(fp[k].nl, fp[k].vx, fp[k].vy, fp[k].vz);
      fb[k] = b[Tl::cell(fp[k].nl < 0 ? l : fp[k].nl, fp[k].vx, fp[k].vy, fp[k].vz)];
    }
    const float zc = zval(l, cx, cy, cz), bc0 = b[Tl::cell(l, cx, cy, cz)];
#pragma unroll
    for (int k = 0; k < Tl::NE; ++k) ez[k] = zval(ep[k].nl, ep[k].vx, ep[k].vy, ep[k].vz);
    const int i = Tl::at(cx + 2, cy + 2, cz + 2);
    Z[i] = zc;
    B[i] = bc0;
#pragma unroll
    for (int k = 0; k < Tl::NF; ++k) {
      const int j = Tl::at(fp[k].tx, fp[k].ty, fp[k].tz);
      Z[j] = fz[k];
      B[j] = fp[k].nl < 0 ? 0.f : fb[k];  // read at depth 1 only
    }
#pragma unroll
    for (int k = 0; k < Tl::NE; ++k)
      if (lane + k * Tl::C < 12 * Tl::S) Z[Tl::at(ep[k].tx, ep[k].ty, ep[k].tz)] = ez[k];
  }
  __syncthreads();
  auto sweep = [&](int tx, int ty, int tz) {
    const int i = Tl::at(tx, ty, tz);
    const float gs = (B[i] * inv_s + Tl::nsum(Z, tx, ty, tz)) * (1.0f / 6.0f);
    Z[i] = fmaf(w, gs - Z[i], Z[i]);
  };
  if (live) ltile_c1<LG>(Fr, lane, cx, cy, cz, sweep);
  __syncthreads();
  if (!live) return;
  if (((cx + cy + cz) & 1) == 0) sweep(cx + 2, cy + 2, cz + 2);
  zout[Tl::cell(l, cx, cy, cz)] = Z[Tl::at(cx + 2, cy + 2, cz + 2)];
}

inline unsigned cells_grid(int32_t L, int lgS) { return (unsigned)((((int64_t)L << (3 * lgS)) + kTpb - 1) / kTpb); }

// one colour sweep on level lgS >= 1 (one lane per z-pair) or on level 3 (lgS = 0, one lane per leaf)
void smooth(const DevTopo& T, int lgS, const float* b, float* u, float scale, int colour, bool first, float omega,
            const int* stop, hipStream_t s) {
  const float inv = 1.0f / scale;
  if (lgS == 0) {
    const unsigned g = cells_grid(T.L, 0);
    if (colour == 0) {
      if (first) k_mg_leaf_sweep<0, true><<<g, kTpb, 0, s>>>(T, b, u, inv, omega, stop);
      else k_mg_leaf_sweep<0, false><<<g, kTpb, 0, s>>>(T, b, u, inv, omega, stop);
    } else {
      if (first) k_mg_leaf_sweep<1, true><<<g, kTpb, 0, s>>>(T, b, u, inv, omega, stop);
      else k_mg_leaf_sweep<1, false><<<g, kTpb, 0, s>>>(T, b, u, inv, omega, stop);
    }
  } else {
    const unsigned g = (unsigned)(((((int64_t)T.L << (3 * lgS)) >> 1) + kTpb - 1) / kTpb);
    if (colour == 0) {
      if (first) k_mg_smooth<0, true><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega, stop);
      else k_mg_smooth<0, false><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega, stop);
    } else {
      if (first) k_mg_smooth<1, true><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega, stop);
      else k_mg_smooth<1, false><<<g, kTpb, 0, s>>>(T, lgS, b, u, inv, omega, stop);
    }
  }
  HNS_HIP(hipGetLastError());
}

// The coarsest level's layouts: level 3's colour-sorted slots for one workgroup (ch 0: none) and
// level 4 of the two-level solve (ch4 0: none; fused4: both fit the fused two-level kernel)
struct CoarseLayout {
  int ch;
  const uint32_t* nbp;
  int ch4;
  bool fused4;
  C4Dev c4;
  float* rho;
};

template <int CH>
void launch_coarse2(int ch4, unsigned grid, const CoarseLayout& Ly, const float* b, float* u, float inv,
                    const int* stop, const McgXUpdate& X, const int* it_dev, int64_t n, int sleeps, hipStream_t s) {
  if (ch4 == 1)
    k_mg_coarse2<CH, 1><<<grid, kCoarseTpb, 0, s>>>(Ly.nbp, b, u, inv, Ly.c4, stop, X, it_dev, n, sleeps);
  else if constexpr (CH >= 2)
    k_mg_coarse2<CH, 2><<<grid, kCoarseTpb, 0, s>>>(Ly.nbp, b, u, inv, Ly.c4, stop, X, it_dev, n, sleeps);
}

// Level 3 solved: by the two-level solve when the shape asks for it and level 4 has a layout, else
// by `coarse` SOR iterations; in one workgroup when the layouts fit it (and `launches` is not
// asked), else by launches with the same arithmetic
void coarse_solve(const DevTopo& T, const CoarseLayout& Ly, const float* b, float* u, float scale, const int* stop,
                  bool launches, const McgHook* hook, hipStream_t s) {
  const MgShape& M = mg_shape();
  const float inv = 1.0f / scale;
  const bool two = M.c4 && Ly.ch4 > 0;
  const bool fused = !launches && (two ? Ly.fused4 : Ly.ch > 0);
  const McgXUpdate X = hook ? hook->xu : McgXUpdate{nullptr, nullptr, nullptr, nullptr, kFcgDepth};
  if (X.x && !fused) fail(HNS_ERR_INVALID_ARGUMENT, "the deferred x update needs the one-workgroup coarse solve");
  const int64_t n = hook ? hook->n : 0;
  const int* it_dev = hook ? hook->it : nullptr;
  if (!fused) {
    if (!two) {
      for (int k = 0; k < M.coarse; ++k) {
        smooth(T, 0, b, u, scale, 0, k == 0, M.coarse_omega, stop, s);
        smooth(T, 0, b, u, scale, 1, k == 0, M.coarse_omega, stop, s);
      }
      return;
    }
    for (int k = 0; k < M.c3_pre; ++k) {
      smooth(T, 0, b, u, scale, 0, k == 0, M.c3_omega, stop, s);
      smooth(T, 0, b, u, scale, 1, k == 0, M.c3_omega, stop, s);
    }
    k_mg_leaf_residual<<<cells_grid(T.L, 0), kTpb, 0, s>>>(T, b, u, Ly.rho, inv, stop);
    HNS_HIP(hipGetLastError());
    const int32_t n4 = 2 * Ly.ch4 * kCoarseTpb;
    k_mg4_restrict<<<(unsigned)((n4 + kTpb - 1) / kTpb), kTpb, 0, s>>>(Ly.c4, n4, Ly.rho, stop);
    HNS_HIP(hipGetLastError());
    switch (Ly.ch4) {
      case 1: k_mg4_solve<1><<<1, kCoarseTpb, 0, s>>>(Ly.c4, stop); break;
      case 2: k_mg4_solve<2><<<1, kCoarseTpb, 0, s>>>(Ly.c4, stop); break;
      default: k_mg4_solve<4><<<1, kCoarseTpb, 0, s>>>(Ly.c4, stop); break;
    }
    HNS_HIP(hipGetLastError());
    k_mg4_prolong<<<cells_grid(T.L, 0), kTpb, 0, s>>>(T.L, Ly.c4.parl, Ly.c4.u4, u, stop);
    HNS_HIP(hipGetLastError());
    for (int k = 0; k < M.c3_post; ++k) {
      smooth(T, 0, b, u, scale, 1, false, M.c3_omega_post, stop, s);
      smooth(T, 0, b, u, scale, 0, false, M.c3_omega_post, stop, s);
    }
    return;
  }
  const int max_side = M.side_groups > 0 ? M.side_groups : (two ? kCoarseSideGroups2 : kCoarseSideGroups);
  const int side = X.x ? (int)std::min<int64_t>(max_side, (n + 4 * kCoarseTpb - 1) / (4 * kCoarseTpb)) : 0;
  const unsigned grid = 1u + (unsigned)side;
  if (two) {
    const int sleeps = M.side_sleeps >= 0 ? M.side_sleeps : kCoarseSleeps2;
    switch (Ly.ch) {
      case 1: launch_coarse2<1>(Ly.ch4, grid, Ly, b, u, inv, stop, X, it_dev, n, sleeps, s); break;
      case 2: launch_coarse2<2>(Ly.ch4, grid, Ly, b, u, inv, stop, X, it_dev, n, sleeps, s); break;
      case 4: launch_coarse2<4>(Ly.ch4, grid, Ly, b, u, inv, stop, X, it_dev, n, sleeps, s); break;
      default: launch_coarse2<8>(Ly.ch4, grid, Ly, b, u, inv, stop, X, it_dev, n, sleeps, s); break;
    }
    HNS_HIP(hipGetLastError());
    return;
  }
  const int iters = M.coarse;
  const float omega = M.coarse_omega;
  const uint32_t* nbp = Ly.nbp;
  // the solve's length grows with CH, and so does the sleep
  const int sl = M.side_sleeps >= 0 ? M.side_sleeps : Ly.ch / 2;
  switch (Ly.ch) {
    case 1: k_mg_coarse<1><<<grid, kCoarseTpb, 0, s>>>(nbp, b, u, inv, omega, iters, stop, X, it_dev, n, sl); break;
    case 2: k_mg_coarse<2><<<grid, kCoarseTpb, 0, s>>>(nbp, b, u, inv, omega, iters, stop, X, it_dev, n, sl); break;
    case 4: k_mg_coarse<4><<<grid, kCoarseTpb, 0, s>>>(nbp, b, u, inv, omega, iters, stop, X, it_dev, n, sl); break;
    default: k_mg_coarse<8><<<grid, kCoarseTpb, 0, s>>>(nbp, b, u, inv, omega, iters, stop, X, it_dev, n, sl); break;
  }
  HNS_HIP(hipGetLastError());
}

}  // namespace

void Multigrid::ensure(int32_t leaves) {
  if (L >= leaves && b[1].p) return;
  if (exec_) {  // the captured V-cycle refers to the buffers about to be replaced
    HNS_HIP(hipGraphExecDestroy(exec_));
    exec_ = nullptr;
  }
  L = leaves;
  ++generation;
  for (int l = 1; l < kMgLevels; ++l) {
    const size_t n = (size_t)leaves << (3 * (3 - l));
    b[l].alloc(n);
    u[l].alloc(n);
    if (l < kMgLevels - 1) ut[l].alloc(n);
  }
}

namespace {

// Colour-sorted slots: the items of colour c (c = 0, 1) take slots [c H, c H + count), sorted
// lexicographically by coordinate with the fastest-varying axis chosen (of x, y, z) for the fewest
// LDS bank conflicts in the one-workgroup solve's neighbour reads: a wave reads one neighbour of
// each of 64 consecutive slots at a time, and such a read takes as many cycles as the most distinct
// words falling in one of the 64 four-byte banks (counted here over the whole solve). At 8 M voxels
// in a y-cylinder, y fastest counted 18 % fewer such cycles than the topology's order on level 3
// and took its solve from 53.0 to 51.8 us: it waits more on the latency of its dependent reads and
// barriers than on LDS throughput. The order does not change the result: a colour's update reads
// only the other colour. nb[i][q]: item i's neighbour across face q (+x, -x, +y, -y, +z, -z), -1
// for none.
std::vector<int32_t> colour_slots(const std::vector<int4>& xyz, const std::vector<int>& colour,
                                  const std::vector<std::array<int32_t, 6>>& nb, int32_t H) {
  const size_t n = xyz.size();
  std::vector<int32_t> lv[2];
  for (size_t i = 0; i < n; ++i) lv[colour[i]].push_back((int32_t)i);
  std::vector<int32_t> slot(n), best;
  int64_t best_cost = -1;
  for (int fast = 0; fast < 3; ++fast) {
    const int s0 = fast == 0 ? 1 : 0, s1 = fast == 2 ? 1 : 2;  // the slower axes, in x, y, z order
    auto ax = [&](int32_t i, int a) { return a == 0 ? xyz[i].x : (a == 1 ? xyz[i].y : xyz[i].z); };
    for (int c = 0; c < 2; ++c) {
      std::sort(lv[c].begin(), lv[c].end(), [&](int32_t p, int32_t q) {
        if (ax(p, s0) != ax(q, s0)) return ax(p, s0) < ax(q, s0);
        if (ax(p, s1) != ax(q, s1)) return ax(p, s1) < ax(q, s1);
        return ax(p, fast) < ax(q, fast);
      });
      for (size_t i = 0; i < lv[c].size(); ++i) slot[lv[c][i]] = c * H + (int32_t)i;
    }
    int64_t cost = 0;
    uint32_t addr[64];
    for (int c = 0; c < 2; ++c)
      for (size_t g = 0; g < lv[c].size(); g += 64) {
        const int m = (int)std::min<size_t>(64, lv[c].size() - g);
        for (int q = 0; q < 6; ++q) {
          for (int i = 0; i < 64; ++i) {
            const int32_t nl = i < m ? nb[lv[c][g + i]][q] : -1;
            addr[i] = nl < 0 ? (uint32_t)(2 * H) : (uint32_t)slot[nl];
          }
          std::sort(addr, addr + 64);
          int per_bank[64] = {0}, worst = 0;
          for (int i = 0; i < 64; ++i)
            if (i == 0 || addr[i] != addr[i - 1]) worst = std::max(worst, ++per_bank[addr[i] & 63]);
          cost += worst;
        }
      }
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = slot;
    }
  }
  return best;
}

// slots per lane and colour for a colour's `nmax` items (a power of two up to cap), 0: too many
int slots_per_lane(size_t nmax, int cap) {
  int ch = 1;
  while (ch < cap && (size_t)ch * kCoarseTpb < nmax) ch *= 2;
  return (size_t)ch * kCoarseTpb < nmax ? 0 : ch;
}

template <class T>
void upload(DBuf<T>& d, const std::vector<T>& h) {
  d.alloc(h.size());
  HNS_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace

bool Multigrid::coarse_fused() const { return mg_shape().c4 && c4_ch_ > 0 ? c4_fused_ : coarse_ch_ > 0; }

void Multigrid::prepare_coarse(const DevTopo& T, hipStream_t s) {
  if (ckey_id_ == T.id && ckey_L_ == T.L) return;
  if (exec_) {  // a captured V-cycle may refer to the layout about to be replaced
    HNS_HIP(hipGraphExecDestroy(exec_));
    exec_ = nullptr;
  }
  coarse_ch_ = 0;
  c4_ch_ = 0;
  c4_fused_ = false;
  ++generation;
  ckey_id_ = T.id;
  ckey_L_ = T.L;
  const int32_t L = T.L;
  if (L <= 0) return;
  std::vector<int4> org((size_t)L);
  std::vector<int32_t> nbr((size_t)L * kNbrStride);
  HNS_HIP(hipStreamSynchronize(s));
  HNS_HIP(hipMemcpy(org.data(), T.origins, org.size() * sizeof(int4), hipMemcpyDeviceToHost));
  HNS_HIP(hipMemcpy(nbr.data(), T.nbr, nbr.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  static const int kFace[6] = {22, 4, 16, 10, 14, 12};  // +x, -x, +y, -y, +z, -z
  std::vector<std::array<int32_t, 6>> lnb((size_t)L);
  std::vector<int> lcol((size_t)L);
  for (int32_t l = 0; l < L; ++l) {
    for (int q = 0; q < 6; ++q) lnb[l][q] = nbr[(size_t)l * kNbrStride + kFace[q]];
    lcol[l] = ((org[l].x + org[l].y + org[l].z) >> 3) & 1;
  }

  // level 3: one workgroup when each colour fits kCoarseMaxCH slots per lane
  size_t ncol[2] = {0, 0};
  for (int32_t l = 0; l < L; ++l) ++ncol[lcol[l]];
  const int ch = slots_per_lane(std::max(ncol[0], ncol[1]), kCoarseMaxCH);
  const int32_t H = ch * kCoarseTpb;
  std::vector<int32_t> slot3;
  if (ch > 0) {
    slot3 = colour_slots(org, lcol, lnb, H);
    std::vector<uint32_t> nbp((size_t)6 * H, (uint32_t)(2 * H) | ((uint32_t)(2 * H) << 16));
    for (int32_t l = 0; l < L; ++l) {
      const int32_t j = slot3[l];
      uint32_t f[6];
      for (int q = 0; q < 6; ++q) f[q] = lnb[l][q] < 0 ? (uint32_t)(2 * H) : (uint32_t)slot3[lnb[l][q]];
      for (int a = 0; a < 3; ++a) nbp[(size_t)2 * a * H + j] = f[2 * a] | (f[2 * a + 1] << 16);
    }
    upload(cmap_, slot3);
    upload(cnb_, nbp);
    // level 3's vectors in slot order (2 H entries; padding slots of b stay zero)
    const size_t n3 = std::max<size_t>((size_t)L, (size_t)2 * H);
    if (b[kMgLevels - 1].n < n3) {
      b[kMgLevels - 1].alloc(n3);
      u[kMgLevels - 1].alloc(n3);
    }
    HNS_HIP(hipMemset(b[kMgLevels - 1].p, 0, b[kMgLevels - 1].bytes()));
    HNS_HIP(hipMemset(u[kMgLevels - 1].p, 0, u[kMgLevels - 1].bytes()));
    coarse_ch_ = ch;
  }

  // level 4: blocks of 2^3 leaves (block coordinate = leaf coordinate >> 1), child q = 4 dx + 2 dy + dz
  std::unordered_map<uint64_t, int32_t> bid;
  std::vector<int4> bxyz;
  std::vector<std::array<int32_t, 8>> child;
  std::vector<int32_t> blk((size_t)L);
  auto bkey = [](int x, int y, int z) {
    return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) |
           (uint64_t)(uint32_t)(z + (1 << 20));
  };
  for (int32_t l = 0; l < L; ++l) {
    const int lx = org[l].x >> 3, ly = org[l].y >> 3, lz = org[l].z >> 3;
    const int bx = lx >> 1, by = ly >> 1, bz = lz >> 1;
    auto it = bid.emplace(bkey(bx, by, bz), (int32_t)bxyz.size());
    if (it.second) {
      bxyz.push_back(make_int4(bx, by, bz, 0));
      child.emplace_back();
      child.back().fill(-1);
    }
    blk[l] = it.first->second;
    child[blk[l]][((lx & 1) << 2) | ((ly & 1) << 1) | (lz & 1)] = l;
  }
  const int32_t B = (int32_t)bxyz.size();
  std::vector<int> bcol((size_t)B);
  size_t bncol[2] = {0, 0};
  for (int32_t i = 0; i < B; ++i) ++bncol[bcol[i] = (bxyz[i].x + bxyz[i].y + bxyz[i].z) & 1];
  const int ch4 = slots_per_lane(std::max(bncol[0], bncol[1]), kC4MaxCH);
  if (ch4 == 0) return;
  const int32_t H4 = ch4 * kCoarseTpb;
  // Galerkin weights: D_B = sum over B's leaves of (6 - their face neighbours inside B); n_BB' = the
  // leaf face pairs between B and the block across face q
  std::vector<std::array<int32_t, 6>> bnb((size_t)B);
  std::vector<std::array<int, 6>> cnt((size_t)B);
  std::vector<int> D((size_t)B, 0);
  for (int32_t i = 0; i < B; ++i) {
    bnb[i].fill(-1);
    cnt[i].fill(0);
  }
  for (int32_t l = 0; l < L; ++l) {
    const int32_t b = blk[l];
    int inside = 0;
    for (int q = 0; q < 6; ++q) {
      const int32_t nl = lnb[l][q];
      if (nl < 0) continue;
      if (blk[nl] == b) {
        ++inside;
      } else {
        ++cnt[b][q];
        bnb[b][q] = blk[nl];
      }
    }
    D[b] += 6 - inside;
  }
  const std::vector<int32_t> slot4 = colour_slots(bxyz, bcol, bnb, H4);
  const uint32_t z4 = (uint32_t)(2 * H4);
  std::vector<uint32_t> nb4((size_t)6 * H4, z4 | (z4 << 16)), w4((size_t)2 * H4, 0u);
  std::vector<float> invd((size_t)2 * H4, 0.f);
  std::vector<int32_t> chl((size_t)16 * H4, -1), parl((size_t)L);
  for (int32_t i = 0; i < B; ++i) {
    const int32_t j = slot4[i];
    uint32_t f[6], w = 0;
    for (int q = 0; q < 6; ++q) {
      f[q] = bnb[i][q] < 0 ? z4 : (uint32_t)slot4[bnb[i][q]];
      w |= (uint32_t)cnt[i][q] << (3 * q);
    }
    for (int a = 0; a < 3; ++a) nb4[(size_t)2 * a * H4 + j] = f[2 * a] | (f[2 * a + 1] << 16);
    w4[j] = w;
    invd[j] = 1.0f / (float)D[i];
    for (int q = 0; q < 8; ++q) chl[(size_t)q * 2 * H4 + j] = child[i][q];
  }
  for (int32_t l = 0; l < L; ++l) parl[l] = slot4[blk[l]];
  upload(c4nb_, nb4);
  upload(c4w_, w4);
  upload(c4invd_, invd);
  upload(c4chl_, chl);
  upload(c4parl_, parl);
  c4b_.alloc((size_t)2 * H4);
  c4u_.alloc((size_t)2 * H4);
  rho3_.alloc((size_t)L);
  c4_ch_ = ch4;
  if (ch > 0 && ch4 <= std::min(ch, kC4MaxFusedCH)) {
    // the fused kernel's children (level-3 slots, absent: the zero slot 2 H) and parents (level-4
    // slots, padding: the zero slot 2 H4)
    const uint32_t z3 = (uint32_t)(2 * H);
    std::vector<uint32_t> ch3((size_t)8 * H4, z3 | (z3 << 16)), par3((size_t)H, z4 | (z4 << 16));
    for (int32_t i = 0; i < B; ++i)
      for (int p = 0; p < 4; ++p) {
        const int32_t c0 = child[i][2 * p], c1 = child[i][2 * p + 1];
        const uint32_t s0 = c0 < 0 ? z3 : (uint32_t)slot3[c0], s1 = c1 < 0 ? z3 : (uint32_t)slot3[c1];
        ch3[(size_t)p * 2 * H4 + slot4[i]] = s0 | (s1 << 16);
      }
    for (int32_t l = 0; l < L; ++l) {
      const int32_t j = slot3[l];
      const uint32_t p = (uint32_t)slot4[blk[l]];
      uint32_t& word = par3[j < H ? j : j - H];
      word = j < H ? (word & 0xffff0000u) | p : (word & 0xffffu) | (p << 16);
    }
    upload(c4ch3_, ch3);
    upload(c4par3_, par3);
    c4_fused_ = true;
  }
}

void Multigrid::launch_all(const DevTopo& T, const float* r_cs, float* z_cs, const int* stop, unsigned variant,
                           hipStream_t s, const McgHook* hook, const CgScalars* test) {
  const MgShape& M = mg_shape();
  const bool fused = M.smooth0 == 1 && !(variant & kMgUnfused);
  if (hook && !fused) fail(HNS_ERR_INVALID_ARGUMENT, "the flexible-CG pass needs the fused level-0 V-cycle");
  if (test && (!hook || stop != test->flag))
    fail(HNS_ERR_INVALID_ARGUMENT, "the folded stop test needs the flexible-CG pass and the solve's stop flag");
  // level 3 in the one-workgroup solve's slot order when that solve runs
  const int32_t* map3 = !(variant & kMgCoarseLaunches) && coarse_fused() ? cmap_.p : nullptr;
  // level 0's pointwise pre-smoothing step z' = c r from z = 0 (c: the Jacobi weight over the
  // diagonal, as the red-black sweep's 0.166666667f)
  const float c0 = M.jacobi0 * 0.166666667f;
  if (fused) {
    // pre-smoothing and restriction in one pass (z' is formed again by the post pass)
    if (test) k_mg0_pre<true><<<cg_groups(T.L) + 8, kTpb, 0, s>>>(T, r_cs, c0, M.over, b[1].p, stop, *test);
    else k_mg0_pre<false><<<cg_groups(T.L), kTpb, 0, s>>>(T, r_cs, c0, M.over, b[1].p, stop, CgScalars{});
    HNS_HIP(hipGetLastError());
  } else {
    const int64_t n = (int64_t)T.L * kLeafVox;
    k_mg0_jacobi<<<(unsigned)((n + kTpb - 1) / kTpb), kTpb, 0, s>>>(n, r_cs, c0, z_cs, stop);
    HNS_HIP(hipGetLastError());
    k_mg0_restrict<<<grid_for_leaves(T.L), kTpb, 0, s>>>(T, r_cs, z_cs, M.over, b[1].p, stop);
    HNS_HIP(hipGetLastError());
  }
  // descend: levels 1, 2 pre-smooth and restrict
  const bool fused_l = M.smooth == 1 && !(variant & kMgUnfused);
  for (int l = 1; l < kMgLevels - 1; ++l) {
    const int lgS = 3 - l;
    const float scale = (float)(1 << (2 * l));
    if (fused_l) {
      const unsigned g = (unsigned)(((int64_t)T.L * (1 << (3 * lgS)) + kTpb - 1) / kTpb);
      if (lgS == 2)
        k_mgl1_pre<kL1Leaves><<<(unsigned)((T.L + 4 * kL1Leaves - 1) / (4 * kL1Leaves)), kTpb, 0, s>>>(
            T, b[l].p, ut[l].p, scale, M.over_deep, b[l + 1].p, M.omega_l, stop);
      else
        k_mgl_pre<1><<<g, kTpb, 0, s>>>(T, b[l].p, ut[l].p, scale, M.over3, b[l + 1].p, M.omega_l, map3, stop);
      HNS_HIP(hipGetLastError());
      continue;
    }
    for (int k = 0; k < M.smooth; ++k) {
      smooth(T, lgS, b[l].p, u[l].p, scale, 0, k == 0, M.omega_l, stop, s);
      smooth(T, lgS, b[l].p, u[l].p, scale, 1, k == 0, M.omega_l, stop, s);
    }
    k_mg_restrict<<<cells_grid(T.L, lgS), kTpb, 0, s>>>(T, lgS, b[l].p, u[l].p, scale, l == 2 ? M.over3 : M.over_deep,
                                                        b[l + 1].p, map3, stop);
    HNS_HIP(hipGetLastError());
  }
  // level 3 (one cell per leaf): the two-level solve, or M.coarse over-relaxed red-black iterations
  CoarseLayout Ly{coarse_ch_, cnb_.p, c4_ch_, c4_fused_, C4Dev{}, rho3_.p};
  Ly.c4 = C4Dev{c4nb_.p, c4w_.p, c4invd_.p, c4ch3_.p, c4par3_.p, c4chl_.p, c4parl_.p, c4b_.p, c4u_.p,
                M.c4_iters, M.c3_pre, M.c3_post, M.c4_omega, M.c4_over, M.c3_omega, M.c3_omega_post};
  coarse_solve(T, Ly, b[kMgLevels - 1].p, u[kMgLevels - 1].p, (float)(1 << (2 * (kMgLevels - 1))), stop,
               (variant & kMgCoarseLaunches) != 0, hook, s);
  // ascend: prolong and post-smooth in the reverse colour order (level 2: M.post2 iterations, the
  // fused path's later ones alternating between u[2] and ut[2]; uc: the level below's result)
  const float* uc = u[kMgLevels - 1].p;
  for (int l = kMgLevels - 2; l >= 1; --l) {
    const int lgS = 3 - l;
    const float scale = (float)(1 << (2 * l));
    const int npost = l == 2 ? M.post2 : M.smooth;
    if (fused_l) {
      const unsigned g = (unsigned)(((int64_t)T.L * (1 << (3 * lgS)) + kTpb - 1) / kTpb);
      if (lgS == 2) {
        k_mgl1_post<kL1Leaves><<<(unsigned)((T.L + 4 * kL1Leaves - 1) / (4 * kL1Leaves)), kTpb, 0, s>>>(
            T, b[l].p, ut[l].p, uc, u[l].p, scale, M.omega_l, stop);
        HNS_HIP(hipGetLastError());  // (fused_l: M.smooth = 1 post iteration on level 1)
        uc = u[l].p;
        continue;
      }
      k_mgl_post<1><<<g, kTpb, 0, s>>>(T, b[l].p, ut[l].p, uc, u[l].p, scale, M.omega_l, map3, stop);
      HNS_HIP(hipGetLastError());
      float* cur = u[l].p;
      float* oth = ut[l].p;  // its pre-smoothed values are read by the first post pass only
      for (int k = 1; k < npost; ++k) {
        k_mgl_post<1, false><<<g, kTpb, 0, s>>>(T, b[l].p, cur, nullptr, oth, scale, M.omega_l, nullptr, stop);
        HNS_HIP(hipGetLastError());
        std::swap(cur, oth);
      }
      uc = cur;
      continue;
    }
    k_mg_prolong<<<cells_grid(T.L, lgS), kTpb, 0, s>>>(T.L, lgS, u[l].p, uc, map3, stop);
    HNS_HIP(hipGetLastError());
    uc = u[l].p;
    for (int k = 0; k < npost; ++k) {
      smooth(T, lgS, b[l].p, u[l].p, scale, 1, false, M.omega_l, stop, s);
      smooth(T, lgS, b[l].p, u[l].p, scale, 0, false, M.omega_l, stop, s);
    }
  }
  if (fused) {
    if (hook) k_mg0_post_fcg<<<cg_groups(T.L), kTpb, 0, s>>>(T, r_cs, u[1].p, z_cs, *hook, c0, M.omega0, stop);
    else k_mg0_post<<<grid_for_leaves(T.L), kTpb, 0, s>>>(T, r_cs, u[1].p, z_cs, c0, M.omega0, stop);
    HNS_HIP(hipGetLastError());
    return;
  }
  k_mg0_prolong<<<grid_for_leaves(T.L), kTpb, 0, s>>>(T.L, z_cs, u[1].p, stop);
  HNS_HIP(hipGetLastError());
  for (int k = 0; k < M.smooth0; ++k) {
    launch_rbgs_rhs(T, r_cs, z_cs, 1, false, M.omega0, stop, s);
    launch_rbgs_rhs(T, r_cs, z_cs, 0, false, M.omega0, stop, s);
  }
}

void Multigrid::apply(const DevTopo& T, const float* r_cs, float* z_cs, hipStream_t s, bool graph, const int* stop,
                      unsigned variant) {
  ensure(T.L);
  prepare_coarse(T, s);
  if (!graph) {
    launch_all(T, r_cs, z_cs, stop, variant, s);
    return;
  }
  if (!exec_ || key_id_ != T.id || key_L_ != T.L || key_r_ != r_cs || key_z_ != z_cs || key_stop_ != stop ||
      key_variant_ != variant) {
    if (exec_) {
      HNS_HIP(hipGraphExecDestroy(exec_));
      exec_ = nullptr;
    }
    if (!cap_) HNS_HIP(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    HNS_HIP(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
    try {
      launch_all(T, r_cs, z_cs, stop, variant, cap_);
    } catch (...) {
      (void)hipStreamEndCapture(cap_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HNS_HIP(hipStreamEndCapture(cap_, &g));
    const hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HNS_HIP(e);
    key_id_ = T.id;
    key_L_ = T.L;
    key_r_ = r_cs;
    key_z_ = z_cs;
    key_stop_ = stop;
    key_variant_ = variant;
  }
  HNS_HIP(hipGraphLaunch(exec_, s));
}

Multigrid::~Multigrid() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (cap_) (void)hipStreamDestroy(cap_);
}

}  // namespace hns
