// Reverse-mode (adjoint) kernels of the differentiable step: advect velocity, project (divergence,
// K red-black Gauss-Seidel iterations, gradient subtraction), advect density with the projected
// velocity. No reference counterpart (the reference is forward only); each kernel is the exact
// transpose of the forward kernel it names, recomputing that kernel's intermediates with the same
// operation order so the clamp branches of MacCormack (Kernel.cu:264-276, :398-407) resolve the
// same way.
//
// Advection transposes scatter: every voxel sends its gradient to the trilinear corners of its
// back- and forward-trace, to itself and to the neighbour that won the clamp. Each workgroup
// accumulates its scatter in LDS over its leaf plus a 2-voxel halo (12^3 cells per component) and
// flushes it once: the leaf's inner 4^3 voxels, which no other workgroup's tile reaches, with
// plain read-modify-writes; the rest (the leaf's 2-voxel border band and the halo) with global
// atomics. Targets farther than the halo (CFL > 2) go to a separate buffer with atomics and are
// added afterwards; advect_scalars' misses read element 0 (Kernel.cu:187-188), so their gradient
// goes to element 0, summed per workgroup first.
#include "hns_adjoint.hpp"

namespace hns {
namespace {

constexpr int kH = 2, kE = 8 + 2 * kH, kE3 = kE * kE * kE;

struct Tri {
  int32_t i0, j0, k0;
  float w[8], dw[8][3];
};
// Corner q = (a, b, c) = (q & 1, (q >> 1) & 1, q >> 2) of the cell holding p; weights as the
// forward's setup_interp forms them ((x-factor * y-factor) * z-factor) and their derivatives.
__device__ inline Tri tri_cell(float px, float py, float pz) {
  Tri c;
  c.i0 = floor_i(px);
  c.j0 = floor_i(py);
  c.k0 = floor_i(pz);
  const float t[3] = {px - (float)c.i0, py - (float)c.j0, pz - (float)c.k0};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int b[3] = {q & 1, (q >> 1) & 1, q >> 2};
    float f[3], df[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      f[a] = b[a] ? t[a] : 1.0f - t[a];
      df[a] = b[a] ? 1.0f : -1.0f;
    }
    c.w[q] = f[0] * f[1] * f[2];
    c.dw[q][0] = df[0] * f[1] * f[2];
    c.dw[q][1] = f[0] * df[1] * f[2];
    c.dw[q][2] = f[0] * f[1] * df[2];
  }
  return c;
}

// Scatter target bookkeeping of one workgroup.
struct Scatter {
  float* tile;        // ncomp * kE3 LDS accumulators
  float* e0;          // ncomp LDS accumulators for element 0
  int ox, oy, oz;     // leaf origin
  float* far[4];      // far targets per component (global, atomics)
  const LeafView* V;
  // add v to component k of the voxel at global coordinates (x,y,z); `miss0` routes an absent
  // voxel to element 0 (advect_scalars' miss rule) instead of dropping it
  __device__ inline void add(int k, int32_t x, int32_t y, int32_t z, float v, bool miss0) const {
    if (v == 0.f) return;
    const int tx = x - ox + kH, ty = y - oy + kH, tz = z - oz + kH;
    const int32_t idx = V->index(x, y, z);
    if (idx < 0) {
      if (miss0) atomicAdd(&e0[k], v);
      return;
    }
    if ((unsigned)tx < (unsigned)kE && (unsigned)ty < (unsigned)kE && (unsigned)tz < (unsigned)kE)
      atomicAdd(&tile[k * kE3 + (tx * kE + ty) * kE + tz], v);
    else
      atomicAdd(&far[k][idx], v);
  }
};

// Writes the LDS tile into the global arrays dst[k] (+=): plain RMW on the leaf's inner 4^3,
// atomics elsewhere; element-0 sums with one atomic per component.
__device__ inline void flush(const LeafView& V, const float* tile, const float* e0, float* const* dst, int ncomp) {
  for (int c = threadIdx.x; c < kE3; c += blockDim.x) {
    const int tx = c / (kE * kE), ty = (c / kE) % kE, tz = c % kE;
    const int lx = tx - kH, ly = ty - kH, lz = tz - kH;  // leaf-local, -2 .. 9
    const bool inner = lx >= 2 && lx < 6 && ly >= 2 && ly < 6 && lz >= 2 && lz < 6;
    int32_t idx = -2;
    for (int k = 0; k < ncomp; ++k) {
      const float v = tile[k * kE3 + c];
      if (v == 0.f) continue;
      if (idx == -2) idx = V.index(V.ox + lx, V.oy + ly, V.oz + lz);
      if (idx < 0) continue;  // never scattered to (absent voxels route elsewhere)
      if (inner)
        dst[k][idx] += v;
      else
        atomicAdd(&dst[k][idx], v);
    }
  }
  if (threadIdx.x < ncomp && e0[threadIdx.x] != 0.f) atomicAdd(&dst[threadIdx.x][0], e0[threadIdx.x]);
}

// ---- advect_scalars transpose (one field, no collision SDF) ----------------------------------
__global__ __launch_bounds__(512) void k_advect_scalars_adj(DevTopo T, const float* __restrict__ u,
                                                            const float* __restrict__ in,
                                                            const float* __restrict__ gout, float* gin, float* gvel,
                                                            float sdt, float* far_in, float* far_vel) {
  __shared__ int32_t s_nb[27];
  __shared__ float s_tile[4 * kE3];
  __shared__ float s_e0[4];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  for (int c = threadIdx.x; c < 4 * kE3; c += blockDim.x) s_tile[c] = 0.f;
  if (threadIdx.x < 4) s_e0[threadIdx.x] = 0.f;
  __syncthreads();
  const Vel U{u, u + N, u + 2 * N};
  // component 0: gin, 1..3: gvel x, y, z
  Scatter S{s_tile, s_e0, V.ox, V.oy, V.oz, {far_in, far_vel, far_vel + N, far_vel + 2 * N}, &V};
  const int32_t n = threadIdx.x;
  const int32_t x = V.ox + (n >> 6), y = V.oy + ((n >> 3) & 7), z = V.oz + (n & 7);
  const int32_t idx = l * kLeafVox + n;
  const float g = gout[idx];
  if (g != 0.f) {
    const float px = (float)x, py = (float)y, pz = (float)z;
    const float vcx = U.x[idx], vcy = U.y[idx], vcz = U.z[idx];
    const float bx = px - sdt * vcx, by = py - sdt * vcy, bz = pz - sdt * vcz;
    const Tri B = tri_cell(bx, by, bz);
    float vfx = 0.f, vfy = 0.f, vfz = 0.f, phiF = 0.f;
    float inB[8], vB[8][3];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int32_t j = max(V.index(B.i0 + (q & 1), B.j0 + ((q >> 1) & 1), B.k0 + (q >> 2)), 0);
      vB[q][0] = U.x[j];
      vB[q][1] = U.y[j];
      vB[q][2] = U.z[j];
      inB[q] = in[j];
      vfx = fmaf(B.w[q], vB[q][0], vfx);
      vfy = fmaf(B.w[q], vB[q][1], vfy);
      vfz = fmaf(B.w[q], vB[q][2], vfz);
      phiF = fmaf(inB[q], B.w[q], phiF);
    }
    const float fx = bx + sdt * vfx, fy = by + sdt * vfy, fz = bz + sdt * vfz;
    const Tri F = tri_cell(fx, fy, fz);
    float inF[8], phiB = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      inF[q] = in[max(V.index(F.i0 + (q & 1), F.j0 + ((q >> 1) & 1), F.k0 + (q >> 2)), 0)];
      phiB = fmaf(inF[q], F.w[q], phiB);
    }
    const float phiOrig = in[idx];
    const float corr = fmaf(0.5f, phiOrig - phiB, phiF);
    // clamp argument: 0..5 neighbour (-x,+x,-y,+y,-z,+z), 6 the voxel, 7 phiF
    const int ox6[6] = {-1, 1, 0, 0, 0, 0}, oy6[6] = {0, 0, -1, 1, 0, 0}, oz6[6] = {0, 0, 0, 0, -1, 1};
    float mn = phiOrig, mx = phiOrig;
    int amn = 6, amx = 6;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const float v = in[max(V.index(x + ox6[q], y + oy6[q], z + oz6[q]), 0)];
      if (v < mn) { mn = v; amn = q; }
      if (v > mx) { mx = v; amx = q; }
    }
    if (phiF < mn) { mn = phiF; amn = 7; }
    if (phiF > mx) { mx = phiF; amx = 7; }
    const float c1 = fminf(corr, mx);
    int arg = -1;  // -1: the corrected value
    if (mn > c1) arg = amn;
    else if (corr > mx) arg = amx;
    float gcorr = 0.f, gphiF = 0.f;
    if (arg < 0) gcorr = g;
    else if (arg == 7) gphiF = g;
    else if (arg == 6) S.add(0, x, y, z, g, true);
    else S.add(0, x + ox6[arg], y + oy6[arg], z + oz6[arg], g, true);
    gphiF += gcorr;
    S.add(0, x, y, z, 0.5f * gcorr, true);
    const float gphiB = -0.5f * gcorr;
    float gfp[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      S.add(0, F.i0 + (q & 1), F.j0 + ((q >> 1) & 1), F.k0 + (q >> 2), F.w[q] * gphiB, true);
#pragma unroll
      for (int k = 0; k < 3; ++k) gfp[k] += F.dw[q][k] * inF[q] * gphiB;
    }
    float gbp[3] = {gfp[0], gfp[1], gfp[2]};
    const float gvF[3] = {gfp[0] * sdt, gfp[1] * sdt, gfp[2] * sdt};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int32_t cx = B.i0 + (q & 1), cy = B.j0 + ((q >> 1) & 1), cz = B.k0 + (q >> 2);
#pragma unroll
      for (int k = 0; k < 3; ++k) S.add(1 + k, cx, cy, cz, gvF[k] * B.w[q], true);
      S.add(0, cx, cy, cz, B.w[q] * gphiF, true);
      const float vdot = vB[q][0] * gvF[0] + vB[q][1] * gvF[1] + vB[q][2] * gvF[2];
#pragma unroll
      for (int k = 0; k < 3; ++k) gbp[k] += B.dw[q][k] * (vdot + inB[q] * gphiF);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) S.add(1 + k, x, y, z, -gbp[k] * sdt, true);
  }
  __syncthreads();
  float* dst[4] = {gin, gvel, gvel + N, gvel + 2 * N};
  flush(V, s_tile, s_e0, dst, 4);
}

// ---- advect_vector transpose (no collision SDF); misses read 0, so their gradient is dropped ----
__global__ __launch_bounds__(512) void k_advect_vector_adj(DevTopo T, const float* __restrict__ u,
                                                           const float* __restrict__ gout, float* gvel, float sdt,
                                                           float* far_vel) {
  __shared__ int32_t s_nb[27];
  __shared__ float s_tile[3 * kE3];
  __shared__ float s_e0[4];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  for (int c = threadIdx.x; c < 3 * kE3; c += blockDim.x) s_tile[c] = 0.f;
  if (threadIdx.x < 4) s_e0[threadIdx.x] = 0.f;
  __syncthreads();
  const Vel U{u, u + N, u + 2 * N};
  Scatter S{s_tile, s_e0, V.ox, V.oy, V.oz, {far_vel, far_vel + N, far_vel + 2 * N, nullptr}, &V};
  const int32_t n = threadIdx.x;
  const int32_t x = V.ox + (n >> 6), y = V.oy + ((n >> 3) & 7), z = V.oz + (n & 7);
  const int32_t idx = l * kLeafVox + n;
  const float g[3] = {gout[idx], gout[N + idx], gout[2 * N + idx]};
  if (g[0] != 0.f || g[1] != 0.f || g[2] != 0.f) {
    const float px = (float)x, py = (float)y, pz = (float)z;
    const float vo[3] = {U.x[idx], U.y[idx], U.z[idx]};
    const float bx = px - sdt * vo[0], by = py - sdt * vo[1], bz = pz - sdt * vo[2];
    const float3 vf3 = trilinear_vec(V, U, bx, by, bz);
    const float vf[3] = {vf3.x, vf3.y, vf3.z};
    const float fx = bx + sdt * vf[0], fy = by + sdt * vf[1], fz = bz + sdt * vf[2];
    const float3 vb3 = trilinear_vec(V, U, fx, fy, fz);
    const float vb[3] = {vb3.x, vb3.y, vb3.z};
    const Tri B = tri_cell(bx, by, bz), F = tri_cell(fx, fy, fz);
    const int ox6[6] = {-1, 1, 0, 0, 0, 0}, oy6[6] = {0, 0, -1, 1, 0, 0}, oz6[6] = {0, 0, 0, 0, -1, 1};
    int32_t nidx[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) nidx[q] = V.index(x + ox6[q], y + oy6[q], z + oz6[q]);
    float gvo[3] = {0.f, 0.f, 0.f}, gvf[3] = {0.f, 0.f, 0.f}, gcorr[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float c = vf[k] + 0.5f * (vo[k] - vb[k]);
      // argument: 0..5 neighbour, 6 vo, 7 vf, 8 missing neighbour (reads 0, gradient dropped)
      float mn = vo[k], mx = vo[k];
      int amn = 6, amx = 6;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const float v = nidx[q] < 0 ? 0.f : (k == 0 ? U.x : (k == 1 ? U.y : U.z))[nidx[q]];
        const int id = nidx[q] < 0 ? 8 : q;
        if (v < mn) { mn = v; amn = id; }
        if (v > mx) { mx = v; amx = id; }
      }
      if (vf[k] < mn) { mn = vf[k]; amn = 7; }
      if (vf[k] > mx) { mx = vf[k]; amx = 7; }
      const float c1 = fminf(c, mx);
      int arg = -1;
      if (mn > c1) arg = amn;
      else if (c > mx) arg = amx;
      if (arg < 0) gcorr[k] = g[k];
      else if (arg == 6) gvo[k] += g[k];
      else if (arg == 7) gvf[k] += g[k];
      else if (arg < 6) S.add(k, x + ox6[arg], y + oy6[arg], z + oz6[arg], g[k], false);
    }
    float gvb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gvf[k] += gcorr[k];
      gvo[k] += 0.5f * gcorr[k];
      gvb[k] = -0.5f * gcorr[k];
    }
    float gfp[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int32_t cx = F.i0 + (q & 1), cy = F.j0 + ((q >> 1) & 1), cz = F.k0 + (q >> 2);
#pragma unroll
      for (int k = 0; k < 3; ++k) S.add(k, cx, cy, cz, gvb[k] * F.w[q], false);
      const float3 v = load_vel(U, V.index(cx, cy, cz));
      const float d = v.x * gvb[0] + v.y * gvb[1] + v.z * gvb[2];
#pragma unroll
      for (int k = 0; k < 3; ++k) gfp[k] += F.dw[q][k] * d;
    }
    float gbp[3] = {gfp[0], gfp[1], gfp[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) gvf[k] += gfp[k] * sdt;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int32_t cx = B.i0 + (q & 1), cy = B.j0 + ((q >> 1) & 1), cz = B.k0 + (q >> 2);
#pragma unroll
      for (int k = 0; k < 3; ++k) S.add(k, cx, cy, cz, gvf[k] * B.w[q], false);
      const float3 v = load_vel(U, V.index(cx, cy, cz));
      const float d = v.x * gvf[0] + v.y * gvf[1] + v.z * gvf[2];
#pragma unroll
      for (int k = 0; k < 3; ++k) gbp[k] += B.dw[q][k] * d;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) S.add(k, x, y, z, gvo[k] - gbp[k] * sdt, false);
  }
  __syncthreads();
  float* dst[3] = {gvel, gvel + N, gvel + 2 * N};
  flush(V, s_tile, s_e0, dst, 3);
}

// ---- transposed red-black Gauss-Seidel ---------------------------------------------------------
// Forward sweep of colour c: p_c <- (1-w) p_c + w inv6 (sum of the other colour's 6 neighbours -
// dx^2 div_c) (Kernel.cu:487-488). Its transpose sends w inv6 g_c to the other colour's
// neighbours, -w inv6 dx^2 g_c to div_c, and keeps (1-w) g_c. Run last sweep first, the gather of
// one transposed sweep and the keep/div part of the next fuse into one pass over one colour:
//   MODE 0 (gather only):  g_X += w inv6 sum g_nbr
//   MODE 1 (fused):        gdiv_X += -w inv6 dx^2 g_X;  g_X = (1-w) g_X + w inv6 sum g_nbr
//   MODE 2 (div only):     gdiv_X += -w inv6 dx^2 g_X
template <int COLOUR, int MODE>
__global__ __launch_bounds__(256) void k_rbgs_adj(DevTopo T, float* __restrict__ g, float* __restrict__ gdiv,
                                                  float dx2, float omega) {
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const int32_t t = threadIdx.x;
  const int32_t x = t >> 5, y = (t >> 2) & 7, zh = t & 3;
  const int32_t z = (zh << 1) | ((x + y + COLOUR) & 1);
  const int32_t half = T.L * kHalfLeaf;
  const float* __restrict__ go = g + (1 - COLOUR) * half;
  float* __restrict__ gc = g + COLOUR * half;
  const int32_t base = l * kHalfLeaf, i = base + t;
  const float wi = omega * 0.166666667f;
  const float gold = gc[i];
  if (MODE != 0) gdiv[COLOUR * half + i] += -wi * dx2 * gold;
  if (MODE == 2) return;
  const FaceRow Fr = face_row(T, l);
  auto ld = [&](int32_t leaf, int xx, int yy, int zz) {
    return leaf < 0 ? 0.f : go[leaf * kHalfLeaf + cs_local(xx, yy, zz)];
  };
  const float s = (x < 7 ? go[base + cs_local(x + 1, y, z)] : ld(Fr.xp, 0, y, z)) +
                  (x > 0 ? go[base + cs_local(x - 1, y, z)] : ld(Fr.xm, 7, y, z)) +
                  (y < 7 ? go[base + cs_local(x, y + 1, z)] : ld(Fr.yp, x, 0, z)) +
                  (y > 0 ? go[base + cs_local(x, y - 1, z)] : ld(Fr.ym, x, 7, z)) +
                  (z < 7 ? go[base + cs_local(x, y, z + 1)] : ld(Fr.zp, x, y, 0)) +
                  (z > 0 ? go[base + cs_local(x, y, z - 1)] : ld(Fr.zm, x, y, 7));
  gc[i] = (MODE == 0 ? gold : (1.0f - omega) * gold) + wi * s;
}

__global__ void k_add_into(int64_t n, float* __restrict__ dst, float* __restrict__ src) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float v = src[e];
  if (v != 0.f) {
    dst[e] += v;
    src[e] = 0.f;  // leaves the far buffer cleared for the next use
  }
}

__global__ void k_count_mismatch(int64_t n, const float* __restrict__ a, const float* __restrict__ b,
                                 unsigned long long* count) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool diff = e < n && __float_as_uint(a[e]) != __float_as_uint(b[e]);
  const unsigned long long m = __ballot(diff);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
}

}  // namespace

void launch_count_mismatch(int64_t n, const float* a, const float* b, unsigned long long* count, hipStream_t s) {
  if (n <= 0) return;
  k_count_mismatch<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, a, b, count);
  HNS_HIP(hipGetLastError());
}

void launch_advect_scalars_adj(const DevTopo& T, const float* u, const float* in, const float* gout, float* gin,
                               float* gvel, float dt, float inv_dx, float* far, hipStream_t s) {
  const int64_t N = (int64_t)T.L * kLeafVox;
  k_advect_scalars_adj<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, in, gout, gin, gvel, dt * inv_dx, far, far + N);
  HNS_HIP(hipGetLastError());
  k_add_into<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(N, gin, far);
  HNS_HIP(hipGetLastError());
  k_add_into<<<(unsigned)((3 * N + 255) / 256), 256, 0, s>>>(3 * N, gvel, far + N);
  HNS_HIP(hipGetLastError());
}

void launch_advect_vector_adj(const DevTopo& T, const float* u, const float* gout, float* gvel, float dt, float inv_dx,
                              float* far, hipStream_t s) {
  const int64_t N = (int64_t)T.L * kLeafVox;
  k_advect_vector_adj<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, gout, gvel, dt * inv_dx, far);
  HNS_HIP(hipGetLastError());
  k_add_into<<<(unsigned)((3 * N + 255) / 256), 256, 0, s>>>(3 * N, gvel, far);
  HNS_HIP(hipGetLastError());
}

void launch_rbgs_adj(const DevTopo& T, float* g_cs, float* gdiv_cs, float dx, int K, float omega, hipStream_t s) {
  if (K <= 0) return;
  const float dx2 = dx * dx;
  const unsigned G = grid_for_leaves(T.L);
  // forward: (sweep 0, sweep 1) x K; transpose: (sweep 1^T, sweep 0^T) x K with
  // sweep c^T = gather into colour 1-c from c, then keep/div on colour c
  k_rbgs_adj<0, 0><<<G, 256, 0, s>>>(T, g_cs, gdiv_cs, dx2, omega);  // gather of sweep 1^T
  HNS_HIP(hipGetLastError());
  for (int k = 0; k < 2 * K - 1; ++k) {
    if ((k & 1) == 0)
      k_rbgs_adj<1, 1><<<G, 256, 0, s>>>(T, g_cs, gdiv_cs, dx2, omega);  // keep/div of 1 + gather of 0^T
    else
      k_rbgs_adj<0, 1><<<G, 256, 0, s>>>(T, g_cs, gdiv_cs, dx2, omega);  // keep/div of 0 + gather of 1^T
    HNS_HIP(hipGetLastError());
  }
  k_rbgs_adj<0, 2><<<G, 256, 0, s>>>(T, g_cs, gdiv_cs, dx2, omega);  // keep/div of sweep 0^T (first)
  HNS_HIP(hipGetLastError());
}

}  // namespace hns
