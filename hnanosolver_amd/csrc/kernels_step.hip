// Step kernels of the sparse smoke/fire solver for gfx950 (CDNA4). All are HBM/latency-bound
// stencils with no matrix work: one 512-thread workgroup (8 wave64) per 8^3 leaf, one voxel per
// lane, neighbour leaves resolved through the leaf's 27-entry row held in LDS, velocity in SoA
// planes so every plane access of a leaf is a contiguous 2 KiB run.
//
// Each kernel restates a reference kernel of src/Cuda/Kernel.cu (cited per kernel) with the
// same floating-point operation order; where the reference runs in place over a stencil
// (vorticityConfinement, HNanoSolver.cu:279-283) these kernels are out of place.
#include "hns_internal.hpp"

namespace hns {
namespace {

__device__ __forceinline__ float3 f3(float a, float b, float c) { return make_float3(a, b, c); }

// getSDFNormal + applyNoSlipBoundary (Kernel.cu:88-127): nearest samples at Coord offsets.
__device__ inline float3 sdf_normal(const LeafView& V, const float* __restrict__ sdf, int32_t x, int32_t y, int32_t z,
                                    float inv_dx) {
  const float r = load_s(sdf, V.index(x + 1, y, z)), le = load_s(sdf, V.index(x - 1, y, z));
  const float t = load_s(sdf, V.index(x, y + 1, z)), b = load_s(sdf, V.index(x, y - 1, z));
  const float f = load_s(sdf, V.index(x, y, z + 1)), k = load_s(sdf, V.index(x, y, z - 1));
  const float s = 0.5f * inv_dx;
  const float gx = s * (r - le), gy = s * (t - b), gz = s * (f - k);
  const float len = sqrtf(gx * gx + gy * gy + gz * gz);
  if (!(len > 1e-6f)) return f3(0.f, 0.f, 0.f);
  const float il = 1.0f / len;  // Vec3 / s == (1/s) * v (nanovdb Math.h:650)
  return f3(il * gx, il * gy, il * gz);
}
__device__ inline float3 no_slip_blend(float3 v, float3 n, float blend) {
  const float vdotn = v.x * n.x + v.y * n.y + v.z * n.z;
  const float3 m = f3(v.x - vdotn * n.x, v.y - vdotn * n.y, v.z - vdotn * n.z);
  const float a = 1.0f - blend;
  return f3(a * v.x + blend * m.x, a * v.y + blend * m.y, a * v.z + blend * m.z);
}

// ---------------------------------------------------------------------------------------
// advect_vector, Kernel.cu:349-426: MacCormack with per-component clamp.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_advect_velocity(DevTopo T, const float* __restrict__ u,
                                                         float* __restrict__ out, float sdt, float inv_dx,
                                                         const float* __restrict__ sdf) {
  __shared__ int32_t s_nb[27];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const Vel U{u, u + N, u + 2 * N};
  const int32_t n = threadIdx.x;
  const int32_t x = V.ox + (n >> 6), y = V.oy + ((n >> 3) & 7), z = V.oz + (n & 7);
  const int32_t idx = l * kLeafVox + n;
  const float3 vo = f3(U.x[idx], U.y[idx], U.z[idx]);
  const float px = (float)x, py = (float)y, pz = (float)z;
  float bx = px - sdt * vo.x, by = py - sdt * vo.y, bz = pz - sdt * vo.z;
  if (sdf && trilinear_s(V, sdf, bx, by, bz) < 0.f) { bx = px; by = py; bz = pz; }
  const float3 vf = trilinear_vec(V, U, bx, by, bz);
  float fx = bx + sdt * vf.x, fy = by + sdt * vf.y, fz = bz + sdt * vf.z;
  if (sdf && trilinear_s(V, sdf, fx, fy, fz) < 0.f) { fx = bx; fy = by; fz = bz; }
  const float3 vb = trilinear_vec(V, U, fx, fy, fz);
  float3 c = f3(vf.x + 0.5f * (vo.x - vb.x), vf.y + 0.5f * (vo.y - vb.y), vf.z + 0.5f * (vo.z - vb.z));
  float3 mn = vo, mx = vo;
  const int32_t nidx[6] = {V.index(x - 1, y, z), V.index(x + 1, y, z), V.index(x, y - 1, z),
                           V.index(x, y + 1, z), V.index(x, y, z - 1), V.index(x, y, z + 1)};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float3 v = load_vel(U, nidx[q]);
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
      const float3 nrm = sdf_normal(V, sdf, x, y, z, inv_dx);
      c = no_slip_blend(c, nrm, 1.0f - (sv / 1.5f));  // Kernel.cu:419 divides by 1.5
    }
  }
  out[idx] = c.x;
  out[N + idx] = c.y;
  out[2 * N + idx] = c.z;
}

// ---------------------------------------------------------------------------------------
// advect_scalars, Kernel.cu:163-289: M fields share the back-trace and interpolation
// weights; misses read element 0 (off == 0 ? 0 : off - 1).
// ---------------------------------------------------------------------------------------
struct Interp8 {
  int32_t idx[8];
  float w[8];
};
__device__ inline Interp8 setup_interp(const LeafView& V, float x, float y, float z) {
  const int32_t i0 = floor_i(x), i1 = i0 + 1, j0 = floor_i(y), j1 = j0 + 1, k0 = floor_i(z), k1 = k0 + 1;
  const float tx = x - (float)i0, ty = y - (float)j0, tz = z - (float)k0;
  const float itx = 1.0f - tx, ity = 1.0f - ty, itz = 1.0f - tz;
  const float w00 = itx * ity, w10 = tx * ity, w01 = itx * ty, w11 = tx * ty;
  Interp8 d;
  d.w[0] = w00 * itz; d.w[1] = w10 * itz; d.w[2] = w01 * itz; d.w[3] = w11 * itz;
  d.w[4] = w00 * tz;  d.w[5] = w10 * tz;  d.w[6] = w01 * tz;  d.w[7] = w11 * tz;
  const int32_t cx[8] = {i0, i1, i0, i1, i0, i1, i0, i1};
  const int32_t cy[8] = {j0, j0, j1, j1, j0, j0, j1, j1};
  const int32_t cz[8] = {k0, k0, k0, k0, k1, k1, k1, k1};
#pragma unroll
  for (int q = 0; q < 8; ++q) d.idx[q] = max(V.index(cx[q], cy[q], cz[q]), 0);
  return d;
}

__global__ __launch_bounds__(512) void k_advect_scalars(DevTopo T, const float* __restrict__ u, FieldSet F,
                                                        float sdt, const float* __restrict__ sdf) {
  __shared__ int32_t s_nb[27];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const Vel U{u, u + N, u + 2 * N};
  const int32_t n = threadIdx.x;
  const int32_t x = V.ox + (n >> 6), y = V.oy + ((n >> 3) & 7), z = V.oz + (n & 7);
  const int32_t idx = l * kLeafVox + n;
  const float px = (float)x, py = (float)y, pz = (float)z;
  const float3 vc = f3(U.x[idx], U.y[idx], U.z[idx]);
  float bx = px - sdt * vc.x, by = py - sdt * vc.y, bz = pz - sdt * vc.z;
  if (sdf && trilinear_s(V, sdf, bx, by, bz) < 0.f) { bx = px; by = py; bz = pz; }
  const Interp8 bd = setup_interp(V, bx, by, bz);
  float vfx = 0.f, vfy = 0.f, vfz = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int32_t j = bd.idx[q];
    vfx = fmaf(bd.w[q], U.x[j], vfx);
    vfy = fmaf(bd.w[q], U.y[j], vfy);
    vfz = fmaf(bd.w[q], U.z[j], vfz);
  }
  float fx = bx + sdt * vfx, fy = by + sdt * vfy, fz = bz + sdt * vfz;
  if (sdf && trilinear_s(V, sdf, fx, fy, fz) < 0.f) { fx = bx; fy = by; fz = bz; }
  const Interp8 fd = setup_interp(V, fx, fy, fz);
  const int32_t nb[6] = {max(V.index(x - 1, y, z), 0), max(V.index(x + 1, y, z), 0), max(V.index(x, y - 1, z), 0),
                         max(V.index(x, y + 1, z), 0), max(V.index(x, y, z - 1), 0), max(V.index(x, y, z + 1), 0)};
  for (int s = 0; s < F.m; ++s) {
    const float* __restrict__ d = F.in[s];
    const float phiOrig = d[idx];
    float phiF = 0.f, phiB = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      phiF = fmaf(d[bd.idx[q]], bd.w[q], phiF);
      phiB = fmaf(d[fd.idx[q]], fd.w[q], phiB);
    }
    const float corr = fmaf(0.5f, phiOrig - phiB, phiF);
    float mn = phiOrig, mx = phiOrig;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const float v = d[nb[q]];
      mn = fminf(mn, v);
      mx = fmaxf(mx, v);
    }
    mn = fminf(mn, phiF);
    mx = fmaxf(mx, phiF);
    F.out[s][idx] = fmaxf(mn, fminf(corr, mx));
  }
}

// advect_scalar, Kernel.cu:294-344 (standalone AdvectIndexGrid path; misses read 0).
__global__ __launch_bounds__(512) void k_advect_scalar(DevTopo T, const float* __restrict__ u,
                                                       const float* __restrict__ in, float* __restrict__ out,
                                                       float sdt, const float* __restrict__ sdf) {
  __shared__ int32_t s_nb[27];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const Vel U{u, u + N, u + 2 * N};
  const int32_t n = threadIdx.x;
  const int32_t x = V.ox + (n >> 6), y = V.oy + ((n >> 3) & 7), z = V.oz + (n & 7);
  const int32_t idx = l * kLeafVox + n;
  const float px = (float)x, py = (float)y, pz = (float)z;
  const float phiOrig = in[idx];
  const float3 vc = f3(U.x[idx], U.y[idx], U.z[idx]);
  float bx = px - sdt * vc.x, by = py - sdt * vc.y, bz = pz - sdt * vc.z;
  if (sdf && trilinear_s(V, sdf, bx, by, bz) < 0.f) { bx = px; by = py; bz = pz; }
  const float phiF = trilinear_s(V, in, bx, by, bz);
  const float3 vf = trilinear_vec(V, U, bx, by, bz);
  float fx = bx + sdt * vf.x, fy = by + sdt * vf.y, fz = bz + sdt * vf.z;
  if (sdf && trilinear_s(V, sdf, fx, fy, fz) < 0.f) { fx = bx; fy = by; fz = bz; }
  const float phiB = trilinear_s(V, in, fx, fy, fz);
  const float corr = phiF + 0.5f * (phiOrig - phiB);
  float mn = phiOrig, mx = phiOrig;
  const int32_t nidx[6] = {V.index(x - 1, y, z), V.index(x + 1, y, z), V.index(x, y - 1, z),
                           V.index(x, y + 1, z), V.index(x, y, z - 1), V.index(x, y, z + 1)};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float v = load_s(in, nidx[q]);
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  mn = fminf(mn, phiF);
  mx = fmaxf(mx, phiF);
  out[idx] = fmaxf(mn, fminf(corr, mx));
}

// ---------------------------------------------------------------------------------------
// vorticityConfinement (Kernel.cu:682-727, Utils.cuh:226-243) + add_velocity_source(_xyz)
// (Kernel.cu:739-755) + add_gravity (Kernel.cu:834-838), fused: the three read and write
// the same voxel's velocity in that order. Out of place (the reference's in-place run races).
// ---------------------------------------------------------------------------------------
__device__ inline float3 curl_at(const LeafView& V, const Vel& U, int32_t x, int32_t y, int32_t z, float factor) {
  const float3 pX = load_vel(U, V.index(x + 1, y, z)), mX = load_vel(U, V.index(x - 1, y, z));
  const float3 pY = load_vel(U, V.index(x, y + 1, z)), mY = load_vel(U, V.index(x, y - 1, z));
  const float3 pZ = load_vel(U, V.index(x, y, z + 1)), mZ = load_vel(U, V.index(x, y, z - 1));
  return f3(((pY.z - mY.z) - (pZ.y - mZ.y)) * factor, ((pZ.x - mZ.x) - (pX.z - mX.z)) * factor,
            ((pX.y - mX.y) - (pY.x - mY.x)) * factor);
}
__device__ inline float curl_mag(const LeafView& V, const Vel& U, int32_t x, int32_t y, int32_t z, float factor) {
  const float3 w = curl_at(V, U, x, y, z, factor);
  // correctly rounded, as the reference's __fsqrt_rn (Utils.cuh:242); on gfx950 HIP's __fsqrt_rn
  // is a bare v_sqrt_f32 while sqrtf carries the rounding correction
  return sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
}

template <bool VORT>
__global__ __launch_bounds__(512) void k_vorticity_forces(DevTopo T, const float* __restrict__ u,
                                                          float* __restrict__ out, float dt, float inv_dx,
                                                          float scale, int32_t F, const float* __restrict__ src,
                                                          float gravity) {
  __shared__ int32_t s_nb[27];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const Vel U{u, u + N, u + 2 * N};
  const int32_t n = threadIdx.x;
  const int32_t idx = l * kLeafVox + n;
  float3 v = f3(U.x[idx], U.y[idx], U.z[idx]);
  if (VORT) {
    const int32_t x = V.ox + (n >> 6), y = V.oy + ((n >> 3) & 7), z = V.oz + (n & 7);
    const float factor = 0.5f * inv_dx;
    const float3 w = curl_at(V, U, x, y, z, factor);
    const float mpX = curl_mag(V, U, x + F, y, z, factor), mmX = curl_mag(V, U, x - F, y, z, factor);
    const float mpY = curl_mag(V, U, x, y + F, z, factor), mmY = curl_mag(V, U, x, y - F, z, factor);
    const float mpZ = curl_mag(V, U, x, y, z + F, factor), mmZ = curl_mag(V, U, x, y, z - F, factor);
    const float gx = (mpX - mmX) * 0.5f * inv_dx;
    const float gy = (mpY - mmY) * 0.5f * inv_dx;
    const float gz = (mpZ - mmZ) * 0.5f * inv_dx;
    const float gl = sqrtf(gx * gx + gy * gy + gz * gz) + 1e-5f;
    const float Nx = gx / gl, Ny = gy / gl, Nz = gz / gl;
    const float fx = scale * (Ny * w.z - Nz * w.y);
    const float fy = scale * (Nz * w.x - Nx * w.z);
    const float fz = scale * (Nx * w.y - Ny * w.x);
    v = f3(v.x + dt * fx, v.y + dt * fy, v.z + dt * fz);
  }
  if (src) v = f3(v.x + dt * src[idx], v.y + dt * src[N + idx], v.z + dt * src[2 * N + idx]);
  if (gravity != 0.0f) v.y -= gravity * dt;
  out[idx] = v.x;
  out[N + idx] = v.y;
  out[2 * N + idx] = v.z;
}

// ---------------------------------------------------------------------------------------
// divergence (Kernel.cu:432-453), optionally fused with the scalar sources (Kernel.cu:732-737),
// combustion_oxygen (Kernel.cu:641-677) and temperature_buoyancy (Kernel.cu:546-563): the
// reference runs them back to back on the same voxel (HNanoSolver.cu:298-355); buoyancy reads
// the post-combustion temperature and writes a separate velocity buffer, since neighbours
// read the pre-buoyancy velocity for their divergence.
// ---------------------------------------------------------------------------------------
template <bool COMB, bool CS>
__global__ __launch_bounds__(512) void k_divergence(DevTopo T, const float* __restrict__ u, float* __restrict__ div,
                                                    float inv_dx, CombustionArgs C, float* __restrict__ u_out) {
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const FaceRow Fr = face_row(T, l);
  const int32_t N = T.L * kLeafVox;
  const Vel U{u, u + N, u + 2 * N};
  const int32_t n = threadIdx.x;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t idx = l * kLeafVox + n;
  const float cx = U.x[idx], cy = U.y[idx], cz = U.z[idx];
  const float xp = (cx + load_s(U.x, face_dense<0, 1>(Fr, lx, ly, lz))) * 0.5f;
  const float xm = (cx + load_s(U.x, face_dense<0, -1>(Fr, lx, ly, lz))) * 0.5f;
  const float yp = (cy + load_s(U.y, face_dense<1, 1>(Fr, lx, ly, lz))) * 0.5f;
  const float ym = (cy + load_s(U.y, face_dense<1, -1>(Fr, lx, ly, lz))) * 0.5f;
  const float zp = (cz + load_s(U.z, face_dense<2, 1>(Fr, lx, ly, lz))) * 0.5f;
  const float zm = (cz + load_s(U.z, face_dense<2, -1>(Fr, lx, ly, lz))) * 0.5f;
  float d = (xp - xm + yp - ym + zp - zm) * inv_dx;
  if (COMB) {
    float fuel = C.fuel[idx];
    float temp = C.temp[idx];
    if (C.src_temp) temp += C.dt * C.src_temp[idx];
    if (C.src_density) C.density[idx] += C.dt * C.src_density[idx];
    if (C.src_fuel) fuel += C.dt * C.src_fuel[idx];
    const float waste = C.waste[idx], flame = C.flame[idx];
    if (fuel < 0.001f) fuel = 0.0f;
    const float oxygen = 1.0f - fuel - waste;
    if (oxygen < 0.0f) {
      C.fuel[idx] = fuel;
      C.temp[idx] = temp;
    } else {
      const float burn = fminf(oxygen, fuel);
      C.fuel[idx] = fuel - burn;
      C.waste[idx] = waste + burn * 2.0f;
      C.flame[idx] = fmaxf(flame, fminf(1.0f, burn * 10.0f));
      temp = temp + burn * C.temp_gain;
      C.temp[idx] = temp;
      d += burn * C.expansion;
    }
    float ox = cx, oy = cy, oz = cz;
    if (temp > C.ambient) {
      const float f = fmaxf(0.0f, (temp - C.ambient) * C.buoyancy);
      ox = cx + C.dt * 0.0f;
      oy = cy + C.dt * f;
      oz = cz + C.dt * 0.0f;
    }
    u_out[idx] = ox;
    u_out[N + idx] = oy;
    u_out[2 * N + idx] = oz;
  }
  if (CS) {
    const int32_t half = N >> 1;
    div[((lx + ly + lz) & 1) * half + l * kHalfLeaf + cs_local(lx, ly, lz)] = d;
  } else {
    div[idx] = d;
  }
}

// combustion_oxygen + temperature_buoyancy alone (leaf-dense div, pointwise).
__global__ void k_combustion_buoyancy(int32_t N, CombustionArgs C, float* __restrict__ div,
                                      const float* __restrict__ u, float* __restrict__ u_out) {
  const int32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  float fuel = C.fuel[idx];
  float temp = C.temp[idx];
  const float waste = C.waste[idx], flame = C.flame[idx];
  if (fuel < 0.001f) fuel = 0.0f;
  const float oxygen = 1.0f - fuel - waste;
  if (oxygen < 0.0f) {
    C.fuel[idx] = fuel;
  } else {
    const float burn = fminf(oxygen, fuel);
    C.fuel[idx] = fuel - burn;
    C.waste[idx] = waste + burn * 2.0f;
    C.flame[idx] = fmaxf(flame, fminf(1.0f, burn * 10.0f));
    temp = temp + burn * C.temp_gain;
    C.temp[idx] = temp;
    div[idx] += burn * C.expansion;
  }
  const float cx = u[idx], cy = u[N + idx], cz = u[2 * N + idx];
  float ox = cx, oy = cy, oz = cz;
  if (temp > C.ambient) {
    const float f = fmaxf(0.0f, (temp - C.ambient) * C.buoyancy);
    ox = cx + C.dt * 0.0f;
    oy = cy + C.dt * f;
    oz = cz + C.dt * 0.0f;
  }
  u_out[idx] = ox;
  u_out[N + idx] = oy;
  u_out[2 * N + idx] = oz;
}

// ---------------------------------------------------------------------------------------
// subtractPressureGradient, Kernel.cu:495-541 (p colour-split).
// ---------------------------------------------------------------------------------------
// gate (non-null): runs only while *gate is set (launched ahead of the host's read of a solve's stop
// flag; returns at once when the solve has not ended)
__global__ __launch_bounds__(512) void k_subtract_gradient(DevTopo T, const float* __restrict__ u,
                                                           const float* __restrict__ p, float* __restrict__ out,
                                                           float inv_dx, const float* __restrict__ sdf,
                                                           const int* __restrict__ gate) {
  __shared__ int32_t s_nb[27];
  __shared__ float s_p[10 * 10 * 10];  // p of the leaf (local -1..8 at 0..9) with its six face halos
  if (gate && !*gate) return;
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  // the neighbour row in LDS (and its barrier) only for the collision path's normal
  LeafView V{};
  if (sdf) V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox, half = N >> 1;
  const int32_t n = threadIdx.x;
  const int32_t lx = n >> 6, ly = (n >> 3) & 7, lz = n & 7;
  const int32_t x = V.ox + lx, y = V.oy + ly, z = V.oz + lz;
  const int32_t idx = l * kLeafVox + n;
  const FaceRow Fr = face_row(T, l);
  auto at = [](int a, int b, int c) { return (a * 10 + b) * 10 + c; };
  // the velocity's loads are issued with the pressure's, before the barrier (the compiler does not
  // move loads across it)
  const float u0 = u[idx], u1 = u[N + idx], u2 = u[2 * N + idx];
  {
    // the lane's own voxel, and lanes < 384 one face-halo voxel each (absent leaves read 0)
    const float own = p[((lx + ly + lz) & 1) * half + l * kHalfLeaf + cs_local(lx, ly, lz)];
    float hv = 0.f;
    int hi = 0;
    if (n < 384) {
      const int f = n >> 6, a = (n >> 3) & 7, b = n & 7;
      int32_t nl;
      int vx, vy, vz;
      switch (f) {
        case 0: nl = Fr.xm; vx = 7; vy = a; vz = b; hi = at(0, a + 1, b + 1); break;
        case 1: nl = Fr.xp; vx = 0; vy = a; vz = b; hi = at(9, a + 1, b + 1); break;
        case 2: nl = Fr.ym; vx = a; vy = 7; vz = b; hi = at(a + 1, 0, b + 1); break;
        case 3: nl = Fr.yp; vx = a; vy = 0; vz = b; hi = at(a + 1, 9, b + 1); break;
        case 4: nl = Fr.zm; vx = a; vy = b; vz = 7; hi = at(a + 1, b + 1, 0); break;
        default: nl = Fr.zp; vx = a; vy = b; vz = 0; hi = at(a + 1, b + 1, 9); break;
      }
      if (nl >= 0) hv = p[((vx + vy + vz) & 1) * half + nl * kHalfLeaf + cs_local(vx, vy, vz)];
    }
    s_p[at(lx + 1, ly + 1, lz + 1)] = own;
    if (n < 384) s_p[hi] = hv;
  }
  __syncthreads();
  const int c = at(lx + 1, ly + 1, lz + 1);
  const float pxp = s_p[c + 100], pxm = s_p[c - 100];
  const float pyp = s_p[c + 10], pym = s_p[c - 10];
  const float pzp = s_p[c + 1], pzm = s_p[c - 1];
  const float s = 0.5f * inv_dx;
  float3 v = f3(u0 - s * (pxp - pxm), u1 - s * (pyp - pym), u2 - s * (pzp - pzm));
  if (sdf) {
    const float sv = sdf[idx];
    if (sv < 0.f) {
      v = f3(0.f, 0.f, 0.f);
    } else if (sv < 0.1f) {
      const float3 nrm = sdf_normal(V, sdf, x, y, z, inv_dx);
      v = no_slip_blend(v, nrm, 1.0f - (sv / 0.1f));
    }
  }
  out[idx] = v.x;
  out[N + idx] = v.y;
  out[2 * N + idx] = v.z;
}

// enforceCollisionBoundaries, Kernel.cu:132-158 (in place; each lane reads only the SDF of
// its neighbours and writes its own velocity).
__global__ __launch_bounds__(512) void k_enforce_collision(DevTopo T, float* __restrict__ u,
                                                           const float* __restrict__ sdf, float voxelSize,
                                                           const int* __restrict__ gate) {
  __shared__ int32_t s_nb[27];
  if (gate && !*gate) return;  // as k_subtract_gradient's
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const LeafView V = load_leaf(T, l, s_nb);
  const int32_t N = T.L * kLeafVox;
  const int32_t n = threadIdx.x;
  const int32_t idx = l * kLeafVox + n;
  const float sv = sdf[idx];
  if (sv < 0.f) {
    u[idx] = 0.f;
    u[N + idx] = 0.f;
    u[2 * N + idx] = 0.f;
    return;
  }
  if (sv < 0.1f) {
    const int32_t x = V.ox + (n >> 6), y = V.oy + ((n >> 3) & 7), z = V.oz + (n & 7);
    const float3 nrm = sdf_normal(V, sdf, x, y, z, 1.0f / voxelSize);
    const float3 v = no_slip_blend(f3(u[idx], u[N + idx], u[2 * N + idx]), nrm, 1.0f - (sv / 0.1f));
    u[idx] = v.x;
    u[N + idx] = v.y;
    u[2 * N + idx] = v.z;
  }
}

// ---------------------------------------------------------------------------------------
// layout converters
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_colour_split(int32_t L, const float* __restrict__ in, float* __restrict__ out,
                                                      int inverse) {
  const int32_t l = blockIdx.x;
  const int32_t n = threadIdx.x;
  const int32_t x = n >> 6, y = (n >> 3) & 7, z = n & 7;
  const int32_t half = L * kHalfLeaf;
  const int32_t cs = ((x + y + z) & 1) * half + l * kHalfLeaf + cs_local(x, y, z);
  const int32_t d = l * kLeafVox + n;
  if (inverse) out[d] = in[cs];
  else out[cs] = in[d];
}

__global__ void k_aos_soa(int64_t n, const float* __restrict__ in, float* __restrict__ out, int to_soa) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (to_soa) {
    out[i] = in[3 * i];
    out[n + i] = in[3 * i + 1];
    out[2 * n + i] = in[3 * i + 2];
  } else {
    out[3 * i] = in[i];
    out[3 * i + 1] = in[n + i];
    out[3 * i + 2] = in[2 * n + i];
  }
}

}  // namespace

// ---- launchers ---------------------------------------------------------------------------
void launch_advect_velocity(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, const float* sdf,
                            hipStream_t s) {
  if (lds_kernels_enabled()) return launch_advect_velocity_lds(T, u, out, dt, inv_dx, sdf, s);
  k_advect_velocity<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, out, dt * inv_dx, inv_dx, sdf);
  HNS_HIP(hipGetLastError());
}
void launch_advect_scalars(const DevTopo& T, const float* u, const FieldSet& f, float dt, float inv_dx,
                           const float* sdf, hipStream_t s) {
  if (lds_kernels_enabled()) return launch_advect_scalars_lds(T, u, f, dt, inv_dx, sdf, s);
  k_advect_scalars<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, f, dt * inv_dx, sdf);
  HNS_HIP(hipGetLastError());
}
void launch_advect_scalar(const DevTopo& T, const float* u, const float* in, float* out, float dt, float inv_dx,
                          const float* sdf, hipStream_t s) {
  k_advect_scalar<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, in, out, dt * inv_dx, sdf);
  HNS_HIP(hipGetLastError());
}
void launch_vorticity_forces(const DevTopo& T, const float* u, float* out, float dt, float inv_dx, float scale,
                             int32_t F, bool vort, const float* src_vel, float gravity, hipStream_t s) {
  if (vort && (F == 1 || F == -1) && lds_kernels_enabled())
    return launch_vorticity_forces_lds(T, u, out, dt, inv_dx, scale, F, src_vel, gravity, s);
  if (vort)
    k_vorticity_forces<true><<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, out, dt, inv_dx, scale, F, src_vel, gravity);
  else
    k_vorticity_forces<false><<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, out, dt, inv_dx, scale, F, src_vel, gravity);
  HNS_HIP(hipGetLastError());
}
void launch_divergence(const DevTopo& T, const float* u, float* div, float inv_dx, bool colour_split,
                       const CombustionArgs* comb, float* u_out, hipStream_t s) {
  const unsigned g = grid_for_leaves(T.L);
  CombustionArgs c{};
  if (comb) c = *comb;
  if (comb) {
    if (colour_split) k_divergence<true, true><<<g, 512, 0, s>>>(T, u, div, inv_dx, c, u_out);
    else k_divergence<true, false><<<g, 512, 0, s>>>(T, u, div, inv_dx, c, u_out);
  } else {
    if (colour_split) k_divergence<false, true><<<g, 512, 0, s>>>(T, u, div, inv_dx, c, u_out);
    else k_divergence<false, false><<<g, 512, 0, s>>>(T, u, div, inv_dx, c, u_out);
  }
  HNS_HIP(hipGetLastError());
}
void launch_combustion_buoyancy(size_t N, const CombustionArgs& c, float* div, const float* u, float* u_out,
                                hipStream_t s) {
  k_combustion_buoyancy<<<(unsigned)((N + 255) / 256), 256, 0, s>>>((int32_t)N, c, div, u, u_out);
  HNS_HIP(hipGetLastError());
}
void launch_subtract_gradient(const DevTopo& T, const float* u, const float* p_cs, float* out, float inv_dx,
                              const float* sdf, hipStream_t s, const int* gate) {
  k_subtract_gradient<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, p_cs, out, inv_dx, sdf, gate);
  HNS_HIP(hipGetLastError());
}
void launch_enforce_collision(const DevTopo& T, float* u, const float* sdf, float voxelSize, hipStream_t s,
                              const int* gate) {
  k_enforce_collision<<<grid_for_leaves(T.L), 512, 0, s>>>(T, u, sdf, voxelSize, gate);
  HNS_HIP(hipGetLastError());
}
void launch_to_colour_split(const DevTopo& T, const float* in, float* out, bool inverse, hipStream_t s) {
  k_colour_split<<<(unsigned)T.L, 512, 0, s>>>(T.L, in, out, inverse ? 1 : 0);
  HNS_HIP(hipGetLastError());
}
void launch_aos_soa(size_t n, const float* in, float* out, bool to_soa, hipStream_t s) {
  if (n == 0) return;
  k_aos_soa<<<(unsigned)((n + 255) / 256), 256, 0, s>>>((int64_t)n, in, out, to_soa ? 1 : 0);
  HNS_HIP(hipGetLastError());
}

}  // namespace hns
