// LDS-staged semi-Lagrangian kernels for gfx950: advect_vector (Kernel.cu:349-426),
// advect_scalars (Kernel.cu:163-289) and vorticityConfinement + forces (Kernel.cu:682-727,
// :739-755, :834-838), with the same arithmetic as the gather kernels of kernels_step.hip.
//
// Each 512-lane workgroup stages its 8^3 leaf plus a 2-voxel halo (a 12^3 tile built from the
// leaf's 27-entry neighbour row) in LDS. Every trilinear corner, clamp neighbour and curl stencil
// inside the tile is then an LDS read instead of a global gather; a lane whose trace leaves the
// tile (CFL above ~1 in the fire scenes) takes the global path for that voxel. Tile cells of
// absent leaves hold the reference's miss value: 0 for advect_vector and the curl stencils
// (Stencils.hpp:83), element 0 of the field for advect_scalars (Kernel.cu:187-188, :229-230).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "hns_internal.hpp"

namespace hns {
namespace {

constexpr int kH = 2;            // halo
constexpr int kE = 8 + 2 * kH;   // tile edge
constexpr int kE3 = kE * kE * kE;
// The scalar advection's tile has a 1-voxel halo: its two traces and the clamp's neighbours stay
// inside it while a trace is under one voxel (CFL < 1), and a sample outside reads the field itself
// (the same values), so the tile only changes where the values come from. At 8 M voxels this took
// the 5-field kernel from 271 to 196 us (the velocity kernel, not bound by its fill, was slower so).
constexpr int kSH = 1, kSE = 8 + 2 * kSH, kSE3 = kSE * kSE * kSE;
// Its tiles interleave four fields per row: row (x, y) of group g holds slots 4g..4g+3, each as the
// row's 10 z values, a row every 40 floats and a plane every 400. A half-wave (4 y-rows of 8 lanes
// in z) then reads y*40 + z + c, 32 distinct banks (40 = 8 mod 32); a row every 10 floats put two
// lanes on one bank in most reads (the kernel was bound by those conflicts, not by its VALU).
constexpr int kSRow = 4 * kSE, kSPlane = kSE * kSRow, kSGroup = kSE * kSPlane;
__device__ __forceinline__ constexpr int stix(int x, int y, int z) { return x * kSPlane + y * kSRow + z; }
// offset of field slot f (0..2 the velocity components, 3 + s scalar field s)
__device__ __forceinline__ constexpr int sslot(int f) { return (f >> 2) * kSGroup + (f & 3) * kSE; }

// The 12^3 tiles (velocity advection, vorticity) keep a cell's three components in one row: row
// (x, y) holds the x, y and z components of its 12 z values at offsets 0, 12 and 24, a row every 40
// floats (4 unused) and a plane every 480, so their half-wave reads are conflict-free as well.
constexpr int kRow = 40, kPlane = kE * kRow, kTile = kE * kPlane;
__device__ __forceinline__ constexpr int tix(int x, int y, int z) { return x * kPlane + y * kRow + z; }

// Tile cell c -> its tile address tix() << 14 | (27-slot of the leaf holding it) << 9 | its
// leaf-dense local index, built at compile time so the fill decodes a table word instead of
// dividing c (the fills were a large share of the kernels' integer instructions).
struct FillTab {
  uint32_t w[kE3];
};
constexpr FillTab make_fill_tab() {
  FillTab t{};
  for (int c = 0; c < kE3; ++c) {
    const int tx = c / (kE * kE), ty = (c / kE) % kE, tz = c % kE;
    const int lx = tx - kH, ly = ty - kH, lz = tz - kH;
    const int slot = ((lx >> 3) + 1) * 9 + ((ly >> 3) + 1) * 3 + ((lz >> 3) + 1);
    t.w[c] = ((uint32_t)tix(tx, ty, tz) << 14) | ((uint32_t)slot << 9) |
             (uint32_t)(((lx & 7) << 6) | ((ly & 7) << 3) | (lz & 7));
  }
  return t;
}
static_assert(tix(kE - 1, kE - 1, kE - 1) < (1 << 13), "tile address must fit the fill word");
__constant__ FillTab kFill = make_fill_tab();
// the same for the scalar advection's 10^3 tile (1-voxel halo), with the cell's interleaved tile
// address stix() in bits 14..25
struct FillTabS {
  uint32_t w[kSE3];
};
constexpr FillTabS make_fill_tab_s() {
  FillTabS t{};
  for (int c = 0; c < kSE3; ++c) {
    const int tx = c / (kSE * kSE), ty = (c / kSE) % kSE, tz = c % kSE;
    const int lx = tx - kSH, ly = ty - kSH, lz = tz - kSH;
    const int slot = ((lx >> 3) + 1) * 9 + ((ly >> 3) + 1) * 3 + ((lz >> 3) + 1);
    t.w[c] = ((uint32_t)stix(tx, ty, tz) << 14) | ((uint32_t)slot << 9) |
             (uint32_t)(((lx & 7) << 6) | ((ly & 7) << 3) | (lz & 7));
  }
  return t;
}
static_assert(stix(kSE - 1, kSE - 1, kSE - 1) < (1 << 12), "tile address must fit the fill word");
__constant__ FillTabS kFillS = make_fill_tab_s();
// element i of a device vector through a 32-bit byte offset (every vector here is below 4 GiB)
__device__ __forceinline__ float ld32(const float* __restrict__ p, uint32_t i) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + (i << 2));
}

// Fills NC (at most 3) components of the tile from a leaf-dense field with plane stride `stride`;
// absent cells get miss[c]. Ends with a barrier.
template <int NC>
__device__ inline void fill_tile(float* __restrict__ s, const float* __restrict__ f, int32_t stride,
                                 const int32_t* s_nb, int l, const float* miss) {
  // every load issued before the first LDS store (an absent cell loads element 0, then takes miss)
  constexpr int kIt = (kE3 + 511) / 512;
  static_assert(NC <= 3, "a tile row holds three components");
  float v[kIt][NC];
  bool hit[kIt];
  int a[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const uint32_t w = kFill.w[min((int)threadIdx.x + it * 512, kE3 - 1)];
    a[it] = (int)(w >> 14);
    const int32_t nl = s_nb[(w >> 9) & 31];
    hit[it] = nl >= 0;
    const uint32_t g = nl < 0 ? 0u : ((uint32_t)nl << 9) + (w & 511u);
#pragma unroll
    for (int k = 0; k < NC; ++k) v[it][k] = ld32(f + (int64_t)k * stride, g);
  }
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int c = (int)threadIdx.x + it * 512;
    if (c < kE3) {
#pragma unroll
      for (int k = 0; k < NC; ++k) s[k * kE + a[it]] = hit[it] ? v[it][k] : miss[k];
    }
  }
  (void)l;
  __syncthreads();
}

__device__ __forceinline__ float3 f3(float a, float b, float c) { return make_float3(a, b, c); }
__device__ __forceinline__ float3 tvec(const float* s, int c) { return f3(s[c], s[kE + c], s[2 * kE + c]); }

// tile corner base of a sample position (tile coordinates), or -1 when the 2x2x2 stencil is not
// inside the tile
// (H: the tile's halo)
template <int H = kH>
__device__ __forceinline__ bool tile_base(const LeafView& V, int32_t i, int32_t j, int32_t k, int& bi, int& bj,
                                          int& bk) {
  constexpr int E = 8 + 2 * H;
  bi = i - (V.ox - H);
  bj = j - (V.oy - H);
  bk = k - (V.oz - H);
  return (unsigned)bi <= (unsigned)(E - 2) && (unsigned)bj <= (unsigned)(E - 2) && (unsigned)bk <= (unsigned)(E - 2);
}

// TrilinearSampler on Vec3f from the tile (same operations as trilinear_vec).
__device__ inline float3 tile_trilinear_vec(const float* s, int bi, int bj, int bk, float tx, float ty, float tz) {
  float3 c[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int d = 0; d < 2; ++d) c[a][b][d] = tvec(s, tix(bi + a, bj + b, bk + d));
  auto lerp = [](float3 va, float3 vb, float w) {
    return f3(fmaf(w, vb.x - va.x, va.x), fmaf(w, vb.y - va.y, va.y), fmaf(w, vb.z - va.z, va.z));
  };
  const float3 z0 = lerp(c[0][0][0], c[0][0][1], tz), z1 = lerp(c[0][1][0], c[0][1][1], tz);
  const float3 z2 = lerp(c[1][0][0], c[1][0][1], tz), z3 = lerp(c[1][1][0], c[1][1][1], tz);
  const float3 y0 = lerp(z0, z1, ty), y1 = lerp(z2, z3, ty);
  return lerp(y0, y1, tx);
}

__device__ inline float3 sample_vec(const LeafView& V, const Vel& U, const float* s, float px, float py, float pz) {
  const int32_t i = floor_i(px), j = floor_i(py), k = floor_i(pz);
  int bi, bj, bk;
  if (tile_base(V, i, j, k, bi, bj, bk))
    return tile_trilinear_vec(s, bi, bj, bk, px - (float)i, py - (float)j, pz - (float)k);
  return trilinear_vec(V, U, px, py, pz);
}

// same helpers as kernels_step.hip (getSDFNormal + applyNoSlipBoundary, Kernel.cu:88-127)
__device__ inline float3 sdf_normal_g(const LeafView& V, const float* __restrict__ sdf, int32_t x, int32_t y,
                                      int32_t z, float inv_dx) {
  const float r = load_s(sdf, V.index(x + 1, y, z)), le = load_s(sdf, V.index(x - 1, y, z));
  const float t = load_s(sdf, V.index(x, y + 1, z)), b = load_s(sdf, V.index(x, y - 1, z));
  const float f = load_s(sdf, V.index(x, y, z + 1)), k = load_s(sdf, V.index(x, y, z - 1));
  const float s = 0.5f * inv_dx;
  const float gx = s * (r - le), gy = s * (t - b), gz = s * (f - k);
  const float len = sqrtf(gx * gx + gy * gy + gz * gz);
  if (!(len > 1e-6f)) return f3(0.f, 0.f, 0.f);
  const float il = 1.0f / len;
  return f3(il * gx, il * gy, il * gz);
}
__device__ inline float3 no_slip_blend_g(float3 v, float3 n, float blend) {
  const float vdotn = v.x * n.x + v.y * n.y + v.z * n.z;
  const float3 m = f3(v.x - vdotn * n.x, v.y - vdotn * n.y, v.z - vdotn * n.z);
  const float a = 1.0f - blend;
  return f3(a * v.x + blend * m.x, a * v.y + blend * m.y, a * v.z + blend * m.z);
}

// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_advect_velocity_lds(DevTopo T, const float* __restrict__ u,
                                                             float* __restrict__ out, float sdt, float inv_dx,
                                                             const float* __restrict__ sdf) {
  __shared__ int32_t s_nb[27];
  __shared__ float s_u[kTile];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const Vel U{u, u + N, u + 2 * N};
  const float zero[3] = {0.f, 0.f, 0.f};
  fill_tile<3>(s_u, u, N, s_nb, l, zero);
  const int32_t n = threadIdx.x;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t x = V.ox + lx, y = V.oy + ly, z = V.oz + lz;
  const int32_t idx = l * kLeafVox + n;
  const int ct = tix(lx + kH, ly + kH, lz + kH);
  const float3 vo = tvec(s_u, ct);
  const float px = (float)x, py = (float)y, pz = (float)z;
  float bx = px - sdt * vo.x, by = py - sdt * vo.y, bz = pz - sdt * vo.z;
  if (sdf && trilinear_s(V, sdf, bx, by, bz) < 0.f) { bx = px; by = py; bz = pz; }
  const float3 vf = sample_vec(V, U, s_u, bx, by, bz);
  float fx = bx + sdt * vf.x, fy = by + sdt * vf.y, fz = bz + sdt * vf.z;
  if (sdf && trilinear_s(V, sdf, fx, fy, fz) < 0.f) { fx = bx; fy = by; fz = bz; }
  const float3 vb = sample_vec(V, U, s_u, fx, fy, fz);
  float3 c = f3(vf.x + 0.5f * (vo.x - vb.x), vf.y + 0.5f * (vo.y - vb.y), vf.z + 0.5f * (vo.z - vb.z));
  float3 mn = vo, mx = vo;
  const int nct[6] = {ct - kPlane, ct + kPlane, ct - kRow, ct + kRow, ct - 1, ct + 1};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float3 v = tvec(s_u, nct[q]);
    mn = f3(fminf(mn.x, v.x), fminf(mn.y, v.y), fminf(mn.z, v.z));
    mx = f3(fmaxf(mx.x, v.x), fmaxf(mx.y, v.y), fmaxf(mx.z, v.z));
  }
  mn = f3(fminf(mn.x, vf.x), fminf(mn.y, vf.y), fminf(mn.z, vf.z));
  mx = f3(fmaxf(mx.x, vf.x), fmaxf(mx.y, vf.y), fmaxf(mx.z, vf.z));
  c = f3(fmaxf(mn.x, fminf(c.x, mx.x)), fmaxf(mn.y, fminf(c.y, mx.y)), fmaxf(mn.z, fminf(c.z, mx.z)));
  if (sdf) {
    const float sv = sdf[idx];
    if (sv < 0.f) {
      c = f3(0.f, 0.f, 0.f);
    } else if (sv < 0.1f) {
      c = no_slip_blend_g(c, sdf_normal_g(V, sdf, x, y, z, inv_dx), 1.0f - (sv / 1.5f));
    }
  }
  out[idx] = c.x;
  out[N + idx] = c.y;
  out[2 * N + idx] = c.z;
}

// ---------------------------------------------------------------------------------------
// advect_scalars: the velocity tile serves the back-trace; one scalar tile is refilled per field.
struct Interp8T {
  int32_t idx[8];  // tile address or leaf-dense index
};
__device__ inline void interp_weights(float x, float y, float z, int32_t& i0, int32_t& j0, int32_t& k0, float* w) {
  i0 = floor_i(x);
  j0 = floor_i(y);
  k0 = floor_i(z);
  const float tx = x - (float)i0, ty = y - (float)j0, tz = z - (float)k0;
  const float itx = 1.0f - tx, ity = 1.0f - ty, itz = 1.0f - tz;
  const float w00 = itx * ity, w10 = tx * ity, w01 = itx * ty, w11 = tx * ty;
  w[0] = w00 * itz; w[1] = w10 * itz; w[2] = w01 * itz; w[3] = w11 * itz;
  w[4] = w00 * tz;  w[5] = w10 * tz;  w[6] = w01 * tz;  w[7] = w11 * tz;
}
// The explicit-weight stencil (corner q at (i0 + (q&1), j0 + ((q>>1)&1), k0 + (q>>2))) as weights,
// the interleaved tile address of corner 0 and whether the stencil lies in the scalar tile; corner q
// is then base + corner_off(q), a compile-time offset (an LDS immediate)
struct Interp8B {
  int32_t i0, j0, k0, base;
  bool in_tile;
  float w[8];
};
__device__ __forceinline__ constexpr int corner_off(int q) {
  return (q & 1) * kSPlane + ((q >> 1) & 1) * kSRow + (q >> 2);
}
__device__ __forceinline__ void setup_interp_base(const LeafView& V, float x, float y, float z, Interp8B& d) {
  interp_weights(x, y, z, d.i0, d.j0, d.k0, d.w);
  int bi, bj, bk;
  d.in_tile = tile_base<kSH>(V, d.i0, d.j0, d.k0, bi, bj, bk);
  d.base = stix(bi, bj, bk);
}
// corner q's index for a lane of a wave that left the tile somewhere: the tile address, or the
// leaf-dense index (a miss reads element 0)
__device__ __forceinline__ int32_t corner_idx(const LeafView& V, const Interp8B& d, int q) {
  return d.in_tile ? d.base + corner_off(q) : max(V.index(d.i0 + (q & 1), d.j0 + ((q >> 1) & 1), d.k0 + (q >> 2)), 0);
}
__device__ __forceinline__ float3 svec(const float* s, int c) { return f3(s[c], s[sslot(1) + c], s[sslot(2) + c]); }

// The velocity tile and all M field tiles are filled in one pass (dynamic LDS: ceil((3 + M) / 4)
// groups of 4000 floats), so the loads of every tile are in flight together and there is one barrier.
// M = F.m fields (a template argument: every field loop unrolls without run-time checks)
template <int M>
__global__ __launch_bounds__(512) void k_advect_scalars_lds(DevTopo T, const float* __restrict__ u, FieldSet F,
                                                            float sdt, const float* __restrict__ sdf) {
  __shared__ int32_t s_nb[27];
  extern __shared__ float s_dyn[];
  float* s_u = s_dyn;
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const Vel U{u, u + N, u + 2 * N};
  {
    // advect_scalars reads element 0 on a miss (velocity and fields alike): an absent cell loads
    // element 0 of each plane. Every load of the tile is issued before the first LDS store.
    constexpr int kIt = (kSE3 + 511) / 512;
    float v[kIt][3 + M];
    int a[kIt];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const uint32_t w = kFillS.w[min((int)threadIdx.x + it * 512, kSE3 - 1)];
      a[it] = (int)(w >> 14);
      const int32_t nl = s_nb[(w >> 9) & 31];
      const int32_t g = nl < 0 ? 0 : (int32_t)(((uint32_t)nl << 9) + (w & 511u));
      v[it][0] = u[g];
      v[it][1] = u[N + g];
      v[it][2] = u[2 * N + g];
#pragma unroll
      for (int s = 0; s < M; ++s) v[it][3 + s] = F.in[s][g];
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int c = (int)threadIdx.x + it * 512;
      if (c < kSE3) {
#pragma unroll
        for (int k = 0; k < 3 + M; ++k) s_u[sslot(k) + a[it]] = v[it][k];
      }
    }
    __syncthreads();
  }
  const int32_t n = threadIdx.x;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t x = V.ox + lx, y = V.oy + ly, z = V.oz + lz;
  const int32_t idx = l * kLeafVox + n;
  const int ct = stix(lx + kSH, ly + kSH, lz + kSH);
  const float px = (float)x, py = (float)y, pz = (float)z;
  const float3 vc = svec(s_u, ct);
  float bx = px - sdt * vc.x, by = py - sdt * vc.y, bz = pz - sdt * vc.z;
  if (sdf && trilinear_s(V, sdf, bx, by, bz) < 0.f) { bx = px; by = py; bz = pz; }
  Interp8B bd;
  setup_interp_base(V, bx, by, bz, bd);
  const bool bt = bd.in_tile;
  float vfx = 0.f, vfy = 0.f, vfz = 0.f;
  // waves whose traces all stay in the tile (the common case) take a path without global reads,
  // addressing the corners from one base with compile-time offsets
  if (__all(bt)) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float3 v = svec(s_u, bd.base + corner_off(q));
      vfx = fmaf(bd.w[q], v.x, vfx);
      vfy = fmaf(bd.w[q], v.y, vfy);
      vfz = fmaf(bd.w[q], v.z, vfz);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int32_t i = corner_idx(V, bd, q);
      const float3 v = bt ? svec(s_u, i) : f3(U.x[i], U.y[i], U.z[i]);
      vfx = fmaf(bd.w[q], v.x, vfx);
      vfy = fmaf(bd.w[q], v.y, vfy);
      vfz = fmaf(bd.w[q], v.z, vfz);
    }
  }
  float fx = bx + sdt * vfx, fy = by + sdt * vfy, fz = bz + sdt * vfz;
  if (sdf && trilinear_s(V, sdf, fx, fy, fz) < 0.f) { fx = bx; fy = by; fz = bz; }
  Interp8B fd;
  setup_interp_base(V, fx, fy, fz, fd);
  const bool ft = fd.in_tile;
  const int nct[6] = {ct - kSPlane, ct + kSPlane, ct - kSRow, ct + kSRow, ct - 1, ct + 1};
  const bool wave_tile = __all(bt && ft);
  // a wave that left the tile somewhere: each lane's corner indices, tile or leaf-dense, once
  Interp8T bg, fg;
  if (!wave_tile) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      bg.idx[q] = corner_idx(V, bd, q);
      fg.idx[q] = corner_idx(V, fd, q);
    }
  }
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const float* __restrict__ d = F.in[s];
    const float* s_f = s_dyn + sslot(3 + s);
    const float phiOrig = s_f[ct];
    float phiF = 0.f, phiB = 0.f;
    if (wave_tile) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        phiF = fmaf(s_f[bd.base + corner_off(q)], bd.w[q], phiF);
        phiB = fmaf(s_f[fd.base + corner_off(q)], fd.w[q], phiB);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        phiF = fmaf(bt ? s_f[bg.idx[q]] : d[bg.idx[q]], bd.w[q], phiF);
        phiB = fmaf(ft ? s_f[fg.idx[q]] : d[fg.idx[q]], fd.w[q], phiB);
      }
    }
    const float corr = fmaf(0.5f, phiOrig - phiB, phiF);
    float mn = phiOrig, mx = phiOrig;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const float v = s_f[nct[q]];
      mn = fminf(mn, v);
      mx = fmaxf(mx, v);
    }
    mn = fminf(mn, phiF);
    mx = fmaxf(mx, phiF);
    F.out[s][idx] = fmaxf(mn, fminf(corr, mx));
  }
}

// ---------------------------------------------------------------------------------------
// vorticity confinement + sources + gravity with |factorScale| truncating to 1 (the fire
// scenes): every stencil (curl at +-1, |curl| at +-1 +-1) lies inside the 2-voxel halo.
__device__ inline float3 tile_curl(const float* s, int c, float factor) {
  const float3 pX = tvec(s, c + kPlane), mX = tvec(s, c - kPlane);
  const float3 pY = tvec(s, c + kRow), mY = tvec(s, c - kRow);
  const float3 pZ = tvec(s, c + 1), mZ = tvec(s, c - 1);
  return f3(((pY.z - mY.z) - (pZ.y - mZ.y)) * factor, ((pZ.x - mZ.x) - (pX.z - mX.z)) * factor,
            ((pX.y - mX.y) - (pY.x - mY.x)) * factor);
}
__device__ inline float tile_curl_mag(const float* s, int c, float factor) {
  const float3 w = tile_curl(s, c, factor);
  // correctly rounded, as the reference's __fsqrt_rn (Utils.cuh:242); on gfx950 HIP's __fsqrt_rn
  // is a bare v_sqrt_f32 while sqrtf carries the rounding correction
  return sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
}

__global__ __launch_bounds__(512) void k_vorticity_forces_lds(DevTopo T, const float* __restrict__ u,
                                                              float* __restrict__ out, float dt, float inv_dx,
                                                              float scale, int32_t F, const float* __restrict__ src,
                                                              float gravity) {
  __shared__ int32_t s_nb[27];
  __shared__ float s_u[kTile];
  // |curl| on the tile's inner 10^3 (the leaf and a 1-voxel ring), a row every 24 floats and a plane
  // every 240 (24 = -8 mod 32: the half-wave reads are conflict-free)
  constexpr int kMRow = 24, kMPlane = 10 * kMRow;
  __shared__ float s_mag[10 * kMPlane];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const float zero[3] = {0.f, 0.f, 0.f};
  fill_tile<3>(s_u, u, N, s_nb, l, zero);
  const float factor = 0.5f * inv_dx;
  // each |curl| once (the gradient below reads six of them per voxel); the same function on the same
  // tile values as evaluating it at the neighbour, so the bits do not change
  for (int m = threadIdx.x; m < 1000; m += 512) {
    const int mx = m / 100, my = (m / 10) % 10, mz = m % 10;
    s_mag[mx * kMPlane + my * kMRow + mz] = tile_curl_mag(s_u, tix(mx + 1, my + 1, mz + 1), factor);
  }
  __syncthreads();
  const int32_t n = threadIdx.x;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t idx = l * kLeafVox + n;
  const int ct = tix(lx + kH, ly + kH, lz + kH);
  float3 v = tvec(s_u, ct);
  const float3 w = tile_curl(s_u, ct, factor);
  const int cm = (lx + 1) * kMPlane + (ly + 1) * kMRow + (lz + 1);
  const int sx = F * kMPlane, sy = F * kMRow, sz = F;  // F is +-1 here
  const float mpX = s_mag[cm + sx], mmX = s_mag[cm - sx];
  const float mpY = s_mag[cm + sy], mmY = s_mag[cm - sy];
  const float mpZ = s_mag[cm + sz], mmZ = s_mag[cm - sz];
  const float gx = (mpX - mmX) * 0.5f * inv_dx;
  const float gy = (mpY - mmY) * 0.5f * inv_dx;
  const float gz = (mpZ - mmZ) * 0.5f * inv_dx;
  const float gl = sqrtf(gx * gx + gy * gy + gz * gz) + 1e-5f;
  const float Nx = gx / gl, Ny = gy / gl, Nz = gz / gl;
  const float fx = scale * (Ny * w.z - Nz * w.y);
  const float fy = scale * (Nz * w.x - Nx * w.z);
  const float fz = scale * (Nx * w.y - Ny * w.x);
  v = f3(v.x + dt * fx, v.y + dt * fy, v.z + dt * fz);
  if (src) v = f3(v.x + dt * src[idx], v.y + dt * src[N + idx], v.z + dt * src[2 * N + idx]);
  if (gravity != 0.0f) v.y -= gravity * dt;
  out[idx] = v.x;
  out[N + idx] = v.y;
  out[2 * N + idx] = v.z;
}

}  // namespace

bool lds_kernels_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("HNS_ADVECT_KERNEL");
    return !(e && std::strcmp(e, "gather") == 0);
  }();
  return on;
}

void launch_advect_velocity_lds(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, const float* sdf,
                                hipStream_t s) {
  k_advect_velocity_lds<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, out, dt * inv_dx, inv_dx, sdf);
  HNS_HIP(hipGetLastError());
}
void launch_advect_scalars_lds(const DevTopo& T, const float* u, const FieldSet& f, float dt, float inv_dx,
                               const float* sdf, hipStream_t s) {
  // at most 5 field tiles per launch: (3 + 5) 12^3 floats = 55 KiB of LDS, under the 64 KiB a
  // workgroup can always allocate; more fields take further launches (the trace is recomputed)
  constexpr int kPerLaunch = 5;
  for (int b = 0; b < f.m; b += kPerLaunch) {
    FieldSet g{};
    g.m = std::min(kPerLaunch, f.m - b);
    for (int q = 0; q < g.m; ++q) {
      g.in[q] = f.in[b + q];
      g.out[q] = f.out[b + q];
    }
    const size_t lds = (size_t)((3 + g.m + 3) / 4) * kSGroup * sizeof(float);
    const unsigned grid = grid_for_leaves(T.L);
    const float sdt = dt * inv_dx;
    switch (g.m) {
      case 1: k_advect_scalars_lds<1><<<grid, 512, lds, s>>>(T, u, g, sdt, sdf); break;
      case 2: k_advect_scalars_lds<2><<<grid, 512, lds, s>>>(T, u, g, sdt, sdf); break;
      case 3: k_advect_scalars_lds<3><<<grid, 512, lds, s>>>(T, u, g, sdt, sdf); break;
      case 4: k_advect_scalars_lds<4><<<grid, 512, lds, s>>>(T, u, g, sdt, sdf); break;
      default: k_advect_scalars_lds<5><<<grid, 512, lds, s>>>(T, u, g, sdt, sdf); break;
    }
    HNS_HIP(hipGetLastError());
  }
}
void launch_vorticity_forces_lds(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, float scale,
                                 int32_t F, const float* src_vel, float gravity, hipStream_t s) {
  k_vorticity_forces_lds<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, out, dt, inv_dx, scale, F, src_vel, gravity);
  HNS_HIP(hipGetLastError());
}

}  // namespace hns
