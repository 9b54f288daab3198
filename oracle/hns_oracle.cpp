// TEST INFRASTRUCTURE — CPU oracle for the sparse smoke/fire step.
//
// This library is the correctness checker, never the product. Only tests/, the
// smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py may load
// it. The shipped path (hnanosolver_amd/libhns.so) neither links nor calls it.
//
// It is a from-scratch C++17 restatement of the reference algorithm, templated on the
// arithmetic type R (float = the reference's FP32 path, double = the FP64 host oracle
// that the FP32 results are also measured against). Every function cites the
// reference file:line it follows (paths relative to the reference repository root).
//
// Index semantics (the NanoVDB ValueOnIndex grid built by voxelsToGrid):
//   offset(ijk) = 0 when the 8^3 leaf holding ijk is absent, else
//   1 + 512 * rank(leaf) + ((x&7)<<6 | (y&7)<<3 | (z&7)),
//   where rank is the leaf's position in voxelsToGrid's sort order
//   (externals/nanovdb/nanovdb/tools/cuda/PointsToGrid.cuh:596-651, leaf prefix sums
//   :454-477, LeafData<ValueOnIndex>::getValue NanoVDB.h:4220-4228).
// Voxel arrays are indexed by the caller's order (the reference's `tid`), neighbour
// reads by offset-1, exactly as the reference kernels do. The lookup here is a hash
// of leaf origins with a one-leaf cache, independent of the GPU path's neighbour tables.
//
// Parity pins for each function are in tests/test_oracle_*.py (closed forms, invariants,
// brute force and the known answers of the reference's Tests/IndexGrid.cpp).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle_mg {
struct MG;
MG* create(const int32_t* origins, int64_t L);
void destroy(MG* g);
int mgcg(MG& g, const float* div, float dx, double tol, int max_iters, float* p, double* rel);
}  // namespace oracle_mg

namespace oracle {

// ---------------------------------------------------------------------------------------
// Vec3 with the component-wise semantics of nanovdb::math::Vec3 (math/Math.h:630-760).
// ---------------------------------------------------------------------------------------
template <typename R>
struct V3 {
  R v[3];
  V3() : v{R(0), R(0), R(0)} {}
  V3(R a, R b, R c) : v{a, b, c} {}
  R& operator[](int i) { return v[i]; }
  const R& operator[](int i) const { return v[i]; }
  V3 operator+(const V3& o) const { return V3(v[0] + o[0], v[1] + o[1], v[2] + o[2]); }
  V3 operator-(const V3& o) const { return V3(v[0] - o[0], v[1] - o[1], v[2] - o[2]); }
  V3 operator*(R s) const { return V3(s * v[0], s * v[1], s * v[2]); }  // Math.h:649
  R dot(const V3& o) const { return v[0] * o[0] + v[1] * o[1] + v[2] * o[2]; }
  R length() const { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
};
template <typename R>
V3<R> operator*(R s, const V3<R>& a) { return V3<R>(s * a[0], s * a[1], s * a[2]); }  // Math.h:746

// ---------------------------------------------------------------------------------------
// Topology: leaf-dense coordinates + NanoVDB-order leaf ranks.
// ---------------------------------------------------------------------------------------
static inline uint64_t packLeaf(int32_t x, int32_t y, int32_t z) {
  // leaf coordinates (ijk >> 3), 21 bits each; valid for |ijk| < 2^23
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)(uint32_t)(x >> 3) & m) << 42 | ((uint64_t)(uint32_t)(y >> 3) & m) << 21 |
         ((uint64_t)(uint32_t)(z >> 3) & m);
}

// PointsToGrid.cuh:596-600 root key: (ijk + 2^31) >> 12 per axis, x in the top bits.
static inline uint64_t rootKey(int32_t x, int32_t y, int32_t z) {
  const int64_t off = int64_t(1) << 31;
  return (uint64_t(uint32_t(int64_t(z) + off) >> 12)) | (uint64_t(uint32_t(int64_t(y) + off) >> 12) << 21) |
         (uint64_t(uint32_t(int64_t(x) + off) >> 12) << 42);
}
// PointsToGrid.cuh:640-645 voxel key inside a root tile: upper offset (NanoVDB.h:3568, 32^3
// children of 128^3), lower offset (16^3 children of 8^3); the 9-bit voxel part is dropped.
static inline uint32_t tileLocalKey(int32_t x, int32_t y, int32_t z) {
  const uint32_t up = (uint32_t((x & 4095) >> 7) << 10) | (uint32_t((y & 4095) >> 7) << 5) | uint32_t((z & 4095) >> 7);
  const uint32_t lo = (uint32_t((x & 127) >> 3) << 8) | (uint32_t((y & 127) >> 3) << 4) | uint32_t((z & 127) >> 3);
  return (up << 12) | lo;
}

struct Topo {
  int64_t N = 0, L = 0;
  std::vector<int32_t> coords;                     // 3N, caller order
  std::unordered_map<uint64_t, int64_t> rankOf;    // packed leaf -> rank in NanoVDB order
};

// One-leaf cache, the role ReadAccessor's cache plays in Stencils.hpp:51-71.
struct Acc {
  const Topo* t;
  uint64_t lastKey = ~0ull;
  int64_t lastBase = -1;
  explicit Acc(const Topo* tp) : t(tp) {}
  int64_t offset(int32_t x, int32_t y, int32_t z) {
    const uint64_t k = packLeaf(x, y, z);
    if (k != lastKey) {
      auto it = t->rankOf.find(k);
      lastKey = k;
      lastBase = (it == t->rankOf.end()) ? -1 : it->second * 512;
    }
    if (lastBase < 0) return 0;
    return 1 + lastBase + (((x & 7) << 6) | ((y & 7) << 3) | (z & 7));
  }
};

// IndexSampler<T,0> (Stencils.hpp:81-89): miss -> value 0.
template <typename T>
static inline T nearest0(Acc& a, const T* data, int32_t x, int32_t y, int32_t z) {
  const int64_t off = a.offset(x, y, z);
  return off == 0 ? T() : data[off - 1];
}

// Floor with fractional remainder (Stencils.hpp:25-43); float->int round-down.
template <typename R>
static inline int32_t floorI(R v) {
  const R f = std::floor(v);
  if (!(f >= R(-2147483648.0))) return INT32_MIN;  // also catches NaN, as a defined value
  if (f > R(2147483520.0)) return 2147483520;
  return int32_t(f);
}

// TrilinearSampler::sample for scalars (Stencils.hpp:118-153): lerp z, then y, then x,
// a + w*(b-a) (the non-Vec3 branch, :140).
template <typename R>
static R trilinearScalar(Acc& a, const R* data, const V3<R>& p) {
  const int32_t i = floorI(p[0]), j = floorI(p[1]), k = floorI(p[2]);
  const R tx = p[0] - R(i), ty = p[1] - R(j), tz = p[2] - R(k);
  R v[2][2][2];
  for (int di = 0; di < 2; ++di)
    for (int dj = 0; dj < 2; ++dj)
      for (int dk = 0; dk < 2; ++dk) v[di][dj][dk] = nearest0(a, data, i + di, j + dj, k + dk);
  auto lerp = [](R va, R vb, R w) { return va + w * (vb - va); };
  const R z0 = lerp(v[0][0][0], v[0][0][1], tz), z1 = lerp(v[0][1][0], v[0][1][1], tz);
  const R z2 = lerp(v[1][0][0], v[1][0][1], tz), z3 = lerp(v[1][1][0], v[1][1][1], tz);
  const R y0 = lerp(z0, z1, ty), y1 = lerp(z2, z3, ty);
  return lerp(y0, y1, tx);
}

// TrilinearSampler::sample for Vec3f on the device (Stencils.hpp:129-135): fmaf(w, b-a, a)
// per component, z then y then x.
template <typename R>
static V3<R> trilinearVec(Acc& a, const V3<R>* data, const V3<R>& p) {
  const int32_t i = floorI(p[0]), j = floorI(p[1]), k = floorI(p[2]);
  const R tx = p[0] - R(i), ty = p[1] - R(j), tz = p[2] - R(k);
  V3<R> v[2][2][2];
  for (int di = 0; di < 2; ++di)
    for (int dj = 0; dj < 2; ++dj)
      for (int dk = 0; dk < 2; ++dk) v[di][dj][dk] = nearest0(a, data, i + di, j + dj, k + dk);
  auto lerp = [](const V3<R>& va, const V3<R>& vb, R w) {
    return V3<R>(std::fma(w, vb[0] - va[0], va[0]), std::fma(w, vb[1] - va[1], va[1]), std::fma(w, vb[2] - va[2], va[2]));
  };
  const V3<R> z0 = lerp(v[0][0][0], v[0][0][1], tz), z1 = lerp(v[0][1][0], v[0][1][1], tz);
  const V3<R> z2 = lerp(v[1][0][0], v[1][0][1], tz), z3 = lerp(v[1][1][0], v[1][1][1], tz);
  const V3<R> y0 = lerp(z0, z1, ty), y1 = lerp(z2, z3, ty);
  return lerp(y0, y1, tx);
}

template <typename R>
static inline V3<R> posOf(const Topo& t, int64_t idx) {
  // Coord::asVec3s (Math.h:758): float(i); the FP64 oracle uses double(i).
  return V3<R>(R(t.coords[3 * idx]), R(t.coords[3 * idx + 1]), R(t.coords[3 * idx + 2]));
}

// ---------------------------------------------------------------------------------------
// SDF helpers (Kernel.cu:81-127).
// ---------------------------------------------------------------------------------------
template <typename R>
static V3<R> sdfNormalAt(Acc& a, const R* sdf, int32_t x, int32_t y, int32_t z, R inv_dx) {
  // getSDFNormal -> gradientSDF with Coord offsets (nearest samples), Kernel.cu:88-115
  const R right = nearest0(a, sdf, x + 1, y, z), left = nearest0(a, sdf, x - 1, y, z);
  const R top = nearest0(a, sdf, x, y + 1, z), bottom = nearest0(a, sdf, x, y - 1, z);
  const R front = nearest0(a, sdf, x, y, z + 1), back = nearest0(a, sdf, x, y, z - 1);
  const V3<R> g = V3<R>(right - left, top - bottom, front - back) * (R(0.5) * inv_dx);
  const R len = g.length();
  return len > R(1e-6f) ? (R(1) / len) * g : V3<R>();  // Vec3 / s == (1/s) * v (Math.h:650)
}
template <typename R>
static V3<R> noSlip(const V3<R>& v, const V3<R>& n) {  // applyNoSlipBoundary, Kernel.cu:122-127
  const R vdotn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
  return v - n * vdotn;
}

// ---------------------------------------------------------------------------------------
// Kernels (all out-of-place unless noted; OpenMP over voxels).
// ---------------------------------------------------------------------------------------

// enforceCollisionBoundaries, Kernel.cu:132-158 (in place on vel; pointwise).
template <typename R>
void enforceCollision(const Topo& t, V3<R>* vel, const R* sdf, R voxelSize) {
  if (!sdf) return;
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t* c = &t.coords[3 * idx];
      const R s = nearest0(a, sdf, c[0], c[1], c[2]);
      if (s < R(0)) {
        vel[idx] = V3<R>();
        continue;
      }
      const R margin = R(0.1f);
      if (s < margin) {
        const V3<R> n = sdfNormalAt(a, sdf, c[0], c[1], c[2], R(1) / voxelSize);
        const R blend = R(1) - (s / margin);
        const V3<R> v = vel[idx];
        const V3<R> m = noSlip(v, n);
        vel[idx] = v * (R(1) - blend) + m * blend;
      }
    }
  }
}

// advect_vector, Kernel.cu:349-426.
template <typename R>
void advectVector(const Topo& t, const V3<R>* vel, V3<R>* out, R dt, R inv_dx, const R* sdf) {
  const R sdt = dt * inv_dx;
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t* c = &t.coords[3 * idx];
      const V3<R> pos = posOf<R>(t, idx);
      const V3<R> velOrig = nearest0(a, vel, c[0], c[1], c[2]);
      V3<R> backPos = pos - velOrig * sdt;
      if (sdf && trilinearScalar(a, sdf, backPos) < R(0)) backPos = pos;
      const V3<R> velForward = trilinearVec(a, vel, backPos);
      V3<R> fwdPos2 = backPos + velForward * sdt;
      if (sdf && trilinearScalar(a, sdf, fwdPos2) < R(0)) fwdPos2 = backPos;
      const V3<R> velBackward = trilinearVec(a, vel, fwdPos2);
      const V3<R> err = velOrig - velBackward;
      V3<R> corr = velForward + R(0.5f) * err;
      V3<R> mn = velOrig, mx = velOrig;
      for (int dim = 0; dim < 3; ++dim)
        for (int off = -1; off <= 1; off += 2) {
          int32_t n[3] = {c[0], c[1], c[2]};
          n[dim] += off;
          const V3<R> vn = nearest0(a, vel, n[0], n[1], n[2]);
          for (int k = 0; k < 3; ++k) {
            mn[k] = std::fmin(mn[k], vn[k]);
            mx[k] = std::fmax(mx[k], vn[k]);
          }
        }
      for (int k = 0; k < 3; ++k) {
        mn[k] = std::fmin(mn[k], velForward[k]);
        mx[k] = std::fmax(mx[k], velForward[k]);
        corr[k] = std::fmax(mn[k], std::fmin(corr[k], mx[k]));
      }
      if (sdf) {
        const R s = nearest0(a, sdf, c[0], c[1], c[2]);
        if (s < R(0)) {
          corr = V3<R>();
        } else if (s < R(0.1f)) {
          const V3<R> n = sdfNormalAt(a, sdf, c[0], c[1], c[2], inv_dx);
          const R blend = R(1) - (s / R(1.5f));  // Kernel.cu:419 uses 1.5, not the 0.1 margin
          const V3<R> ns = noSlip(corr, n);
          corr = corr * (R(1) - blend) + ns * blend;
        }
      }
      out[idx] = corr;
    }
  }
}

// setupInterpolation of advect_scalars, Kernel.cu:204-233: explicit weights; a miss maps
// to element 0 (off==0 ? 0 : off-1).
template <typename R>
struct Interp {
  int64_t idx[8];
  R w[8];
};
template <typename R>
static Interp<R> setupInterp(Acc& a, const V3<R>& p) {
  const int32_t i0 = floorI(p[0]), i1 = i0 + 1, j0 = floorI(p[1]), j1 = j0 + 1, k0 = floorI(p[2]), k1 = k0 + 1;
  const R tx = p[0] - R(i0), ty = p[1] - R(j0), tz = p[2] - R(k0);
  const R itx = R(1) - tx, ity = R(1) - ty, itz = R(1) - tz;
  const R w00 = itx * ity, w10 = tx * ity, w01 = itx * ty, w11 = tx * ty;
  Interp<R> d;
  d.w[0] = w00 * itz; d.w[1] = w10 * itz; d.w[2] = w01 * itz; d.w[3] = w11 * itz;
  d.w[4] = w00 * tz;  d.w[5] = w10 * tz;  d.w[6] = w01 * tz;  d.w[7] = w11 * tz;
  const int32_t cc[8][3] = {{i0, j0, k0}, {i1, j0, k0}, {i0, j1, k0}, {i1, j1, k0},
                            {i0, j0, k1}, {i1, j0, k1}, {i0, j1, k1}, {i1, j1, k1}};
  for (int q = 0; q < 8; ++q) {
    const int64_t off = a.offset(cc[q][0], cc[q][1], cc[q][2]);
    d.idx[q] = off == 0 ? 0 : off - 1;
  }
  return d;
}

// advect_scalars, Kernel.cu:163-289 (M fields sharing one back-trace; misses -> element 0).
template <typename R>
void advectScalars(const Topo& t, const V3<R>* vel, int m, const R* const* in, R* const* out, R dt, R inv_dx,
                   const R* sdf) {
  const R sdt = dt * inv_dx;
  static const int offs[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t* c = &t.coords[3 * idx];
      int64_t orig = a.offset(c[0], c[1], c[2]);
      orig = orig == 0 ? 0 : orig - 1;
      const V3<R> pos = posOf<R>(t, idx);
      const V3<R> velC = vel[orig];
      V3<R> backPos = pos - velC * sdt;
      if (sdf && trilinearScalar(a, sdf, backPos) < R(0)) backPos = pos;
      const Interp<R> bd = setupInterp(a, backPos);
      V3<R> velF;
      for (int q = 0; q < 8; ++q) {
        const V3<R>& v = vel[bd.idx[q]];
        // velF = velF + v * w (contracted to fma on the device)
        velF = V3<R>(std::fma(bd.w[q], v[0], velF[0]), std::fma(bd.w[q], v[1], velF[1]), std::fma(bd.w[q], v[2], velF[2]));
      }
      V3<R> fwdPos2 = backPos + velF * sdt;
      if (sdf && trilinearScalar(a, sdf, fwdPos2) < R(0)) fwdPos2 = backPos;
      const Interp<R> fd = setupInterp(a, fwdPos2);
      int64_t nb[6];
      for (int n = 0; n < 6; ++n) {
        const int64_t off = a.offset(c[0] + offs[n][0], c[1] + offs[n][1], c[2] + offs[n][2]);
        nb[n] = off == 0 ? 0 : off - 1;
      }
      for (int s = 0; s < m; ++s) {
        const R* d = in[s];
        const R phiOrig = d[orig];
        R phiF = R(0), phiB = R(0);
        for (int q = 0; q < 8; ++q) {
          phiF = std::fma(d[bd.idx[q]], bd.w[q], phiF);
          phiB = std::fma(d[fd.idx[q]], fd.w[q], phiB);
        }
        const R error = phiOrig - phiB;
        const R corr = std::fma(R(0.5f), error, phiF);
        R mn = phiOrig, mx = phiOrig;
        for (int n = 0; n < 6; ++n) {
          mn = std::fmin(mn, d[nb[n]]);
          mx = std::fmax(mx, d[nb[n]]);
        }
        mn = std::fmin(mn, phiF);
        mx = std::fmax(mx, phiF);
        out[s][idx] = std::fmax(mn, std::fmin(corr, mx));
      }
    }
  }
}

// advect_scalar, Kernel.cu:294-344 (single field, TrilinearSampler, misses -> 0).
template <typename R>
void advectScalar(const Topo& t, const V3<R>* vel, const R* in, R* out, R dt, R inv_dx, const R* sdf) {
  const R sdt = dt * inv_dx;
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t* c = &t.coords[3 * idx];
      const V3<R> pos = posOf<R>(t, idx);
      const R phiOrig = nearest0(a, in, c[0], c[1], c[2]);
      const V3<R> velC = nearest0(a, vel, c[0], c[1], c[2]);
      V3<R> backPos = pos - velC * sdt;
      if (sdf && trilinearScalar(a, sdf, backPos) < R(0)) backPos = pos;
      const R phiF = trilinearScalar(a, in, backPos);
      const V3<R> velF = trilinearVec(a, vel, backPos);
      V3<R> fwdPos2 = backPos + velF * sdt;
      if (sdf && trilinearScalar(a, sdf, fwdPos2) < R(0)) fwdPos2 = backPos;
      const R phiB = trilinearScalar(a, in, fwdPos2);
      const R error = phiOrig - phiB;
      const R corr = phiF + R(0.5f) * error;
      R mn = phiOrig, mx = phiOrig;
      for (int dim = 0; dim < 3; ++dim)
        for (int off = -1; off <= 1; off += 2) {
          int32_t n[3] = {c[0], c[1], c[2]};
          n[dim] += off;
          const R v = nearest0(a, in, n[0], n[1], n[2]);
          mn = std::fmin(mn, v);
          mx = std::fmax(mx, v);
        }
      mn = std::fmin(mn, phiF);
      mx = std::fmax(mx, phiF);
      out[idx] = std::fmax(mn, std::fmin(corr, mx));
    }
  }
}

// computeVorticityMag, Utils.cuh:226-243.
template <typename R>
static R vortMag(Acc& a, const V3<R>* vel, int32_t x, int32_t y, int32_t z, R factor) {
  const V3<R> pX = nearest0(a, vel, x + 1, y, z), mX = nearest0(a, vel, x - 1, y, z);
  const V3<R> pY = nearest0(a, vel, x, y + 1, z), mY = nearest0(a, vel, x, y - 1, z);
  const V3<R> pZ = nearest0(a, vel, x, y, z + 1), mZ = nearest0(a, vel, x, y, z - 1);
  const R ox = ((pY[2] - mY[2]) - (pZ[1] - mZ[1])) * factor;
  const R oy = ((pZ[0] - mZ[0]) - (pX[2] - mX[2])) * factor;
  const R oz = ((pX[1] - mX[1]) - (pY[0] - mY[0])) * factor;
  return std::sqrt(ox * ox + oy * oy + oz * oz);
}

// vorticityConfinement, Kernel.cu:682-727. The reference runs it in place
// (HNanoSolver.cu:279-283), which races when (int)factorScale != 0; the oracle defines the
// out-of-place result (every read sees the input field).
template <typename R>
void vorticity(const Topo& t, const V3<R>* vel, V3<R>* out, R dt, R inv_dx, R scale, float factorScale) {
  const int32_t F = int32_t(factorScale);  // nanovdb::Coord(float,...) truncates (Kernel.cu:707)
  const R factor = R(0.5f) * inv_dx;
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t x = t.coords[3 * idx], y = t.coords[3 * idx + 1], z = t.coords[3 * idx + 2];
      const V3<R> pX = nearest0(a, vel, x + 1, y, z), mX = nearest0(a, vel, x - 1, y, z);
      const V3<R> pY = nearest0(a, vel, x, y + 1, z), mY = nearest0(a, vel, x, y - 1, z);
      const V3<R> pZ = nearest0(a, vel, x, y, z + 1), mZ = nearest0(a, vel, x, y, z - 1);
      const R ox = ((pY[2] - mY[2]) - (pZ[1] - mZ[1])) * factor;
      const R oy = ((pZ[0] - mZ[0]) - (pX[2] - mX[2])) * factor;
      const R oz = ((pX[1] - mX[1]) - (pY[0] - mY[0])) * factor;
      const R mpX = vortMag(a, vel, x + F, y, z, factor), mmX = vortMag(a, vel, x - F, y, z, factor);
      const R mpY = vortMag(a, vel, x, y + F, z, factor), mmY = vortMag(a, vel, x, y - F, z, factor);
      const R mpZ = vortMag(a, vel, x, y, z + F, factor), mmZ = vortMag(a, vel, x, y, z - F, factor);
      const R gx = (mpX - mmX) * R(0.5f) * inv_dx;
      const R gy = (mpY - mmY) * R(0.5f) * inv_dx;
      const R gz = (mpZ - mmZ) * R(0.5f) * inv_dx;
      const R gl = std::sqrt(gx * gx + gy * gy + gz * gz) + R(1e-5f);
      const R Nx = gx / gl, Ny = gy / gl, Nz = gz / gl;
      const V3<R> f(scale * (Ny * oz - Nz * oy), scale * (Nz * ox - Nx * oz), scale * (Nx * oy - Ny * ox));
      out[idx] = vel[idx] + f * dt;
    }
  }
}

// divergence, Kernel.cu:432-453 (face averages; a missing neighbour reads as 0).
template <typename R>
void divergence(const Topo& t, const V3<R>* vel, R* div, R inv_dx) {
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t x = t.coords[3 * idx], y = t.coords[3 * idx + 1], z = t.coords[3 * idx + 2];
      const V3<R> c = vel[idx];
      const R xp = (c + nearest0(a, vel, x + 1, y, z))[0] * R(0.5f);
      const R xm = (c + nearest0(a, vel, x - 1, y, z))[0] * R(0.5f);
      const R yp = (c + nearest0(a, vel, x, y + 1, z))[1] * R(0.5f);
      const R ym = (c + nearest0(a, vel, x, y - 1, z))[1] * R(0.5f);
      const R zp = (c + nearest0(a, vel, x, y, z + 1))[2] * R(0.5f);
      const R zm = (c + nearest0(a, vel, x, y, z - 1))[2] * R(0.5f);
      div[idx] = (xp - xm + yp - ym + zp - zm) * inv_dx;
    }
  }
}

// redBlackGaussSeidelUpdate, Kernel.cu:459-489 (one colour, in place).
template <typename R>
void rbgsColour(const Topo& t, const R* div, R* p, R dx, int color, R omega) {
  const R dx2 = dx * dx;
  const R inv6 = R(0.166666667f);
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t i = t.coords[3 * idx], j = t.coords[3 * idx + 1], k = t.coords[3 * idx + 2];
      if (((i + j + k) & 1) != color) continue;
      const R s = nearest0(a, (const R*)p, i + 1, j, k) + nearest0(a, (const R*)p, i - 1, j, k) +
                  nearest0(a, (const R*)p, i, j + 1, k) + nearest0(a, (const R*)p, i, j - 1, k) +
                  nearest0(a, (const R*)p, i, j, k + 1) + nearest0(a, (const R*)p, i, j, k - 1);
      const R pOld = p[idx];
      const R pGS = (s - div[idx] * dx2) * inv6;
      p[idx] = pOld + omega * (pGS - pOld);
    }
  }
}

// Poisson residual (new; compute_residual is declared but never defined, Kernels.cuh:59-60):
// r = div - (sum6(p_nbr) - 6 p) / dx^2; returns sum r^2 and sum div^2 in double.
template <typename R>
void residual(const Topo& t, const R* p, const R* div, R dx, double* r2, double* d2) {
  double sr = 0.0, sd = 0.0;
  const double inv_dx2 = 1.0 / (double(dx) * double(dx));
#pragma omp parallel reduction(+ : sr, sd)
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t i = t.coords[3 * idx], j = t.coords[3 * idx + 1], k = t.coords[3 * idx + 2];
      const double s = double(nearest0(a, p, i + 1, j, k)) + double(nearest0(a, p, i - 1, j, k)) +
                       double(nearest0(a, p, i, j + 1, k)) + double(nearest0(a, p, i, j - 1, k)) +
                       double(nearest0(a, p, i, j, k + 1)) + double(nearest0(a, p, i, j, k - 1));
      const double r = double(div[idx]) - (s - 6.0 * double(p[idx])) * inv_dx2;
      sr += r * r;
      sd += double(div[idx]) * double(div[idx]);
    }
  }
  *r2 = sr;
  *d2 = sd;
}

// Conjugate gradients on A p = b, A = (6 p - sum6 p_nbr) (the Dirichlet-0 operator of the RBGS
// sweep scaled by dx^2), b = -div dx^2, from p = 0 until ||r|| <= tol ||b|| (textbook CG,
// Hestenes & Stiefel 1952), in the arithmetic type R with FP64 dot products. New: the reference
// has no to-tolerance solver.
template <typename R>
int conjugateGradient(const Topo& t, const R* div, R dx, double tol, int maxIters, R* p, double* relOut) {
  const int64_t N = t.N;
  std::vector<R> r(N), d(N), q(N);
  const R dx2 = dx * dx;
  double rr = 0.0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    p[i] = R(0);
    r[i] = -div[i] * dx2;
    d[i] = r[i];
    rr += double(r[i]) * double(r[i]);
  }
  const double bb = rr;
  int it = 0;
  while (it < maxIters && rr > tol * tol * bb) {
    double dq = 0.0;
#pragma omp parallel reduction(+ : dq)
    {
      Acc a(&t);
#pragma omp for schedule(static)
      for (int64_t idx = 0; idx < N; ++idx) {
        const int32_t x = t.coords[3 * idx], y = t.coords[3 * idx + 1], z = t.coords[3 * idx + 2];
        const R s = nearest0(a, d.data(), x + 1, y, z) + nearest0(a, d.data(), x - 1, y, z) +
                    nearest0(a, d.data(), x, y + 1, z) + nearest0(a, d.data(), x, y - 1, z) +
                    nearest0(a, d.data(), x, y, z + 1) + nearest0(a, d.data(), x, y, z - 1);
        q[idx] = R(6) * d[idx] - s;
        dq += double(d[idx]) * double(q[idx]);
      }
    }
    if (!(dq > 0.0)) break;
    const R alpha = R(rr / dq);
    double rr2 = 0.0;
#pragma omp parallel for reduction(+ : rr2) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      p[i] += alpha * d[i];
      r[i] -= alpha * q[i];
      rr2 += double(r[i]) * double(r[i]);
    }
    const R beta = R(rr2 / rr);
    rr = rr2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) d[i] = r[i] + beta * d[i];
    ++it;
  }
  *relOut = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
  return it;
}

// subtractPressureGradient, Kernel.cu:495-541.
template <typename R>
void subtractGradient(const Topo& t, const V3<R>* vel, const R* p, V3<R>* out, R inv_dx, const R* sdf) {
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t x = t.coords[3 * idx], y = t.coords[3 * idx + 1], z = t.coords[3 * idx + 2];
      const R pxp = nearest0(a, p, x + 1, y, z), pxm = nearest0(a, p, x - 1, y, z);
      const R pyp = nearest0(a, p, x, y + 1, z), pym = nearest0(a, p, x, y - 1, z);
      const R pzp = nearest0(a, p, x, y, z + 1), pzm = nearest0(a, p, x, y, z - 1);
      const V3<R> g = V3<R>(pxp - pxm, pyp - pym, pzp - pzm) * (R(0.5f) * inv_dx);
      V3<R> u = vel[idx] - g;
      if (sdf) {
        const R s = nearest0(a, sdf, x, y, z);
        if (s < R(0)) {
          u = V3<R>();
        } else if (s < R(0.1f)) {
          const V3<R> n = sdfNormalAt(a, sdf, x, y, z, inv_dx);
          const R blend = R(1) - (s / R(0.1f));
          const V3<R> ns = noSlip(u, n);
          u = u * (R(1) - blend) + ns * blend;
        }
      }
      out[idx] = u;
    }
  }
}

// combustion_oxygen, Kernel.cu:641-677 (pointwise; div read-modify-write).
template <typename R>
void combustionOxygen(int64_t N, R* fuelA, R* wasteA, R* tempA, R* div, R* flameA, R temp_gain, R expansion) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    R fuel = fuelA[i];
    const R waste = wasteA[i], temperature = tempA[i], flame = flameA[i];
    if (fuel < R(0.001f)) fuel = R(0);
    const R oxygen = R(1) - fuel - waste;
    if (oxygen < R(0)) {
      fuelA[i] = fuel;  // copy-through keeps the snapped fuel (Kernel.cu:657-662)
      continue;
    }
    const R burn = std::fmin(oxygen, fuel);
    fuelA[i] = fuel - burn;
    wasteA[i] = waste + burn * R(2);
    flameA[i] = std::fmax(flame, std::fmin(R(1), burn * R(10)));
    tempA[i] = temperature + burn * temp_gain;
    div[i] += burn * expansion;
  }
}

// temperature_buoyancy, Kernel.cu:546-563 (pointwise).
template <typename R>
void buoyancy(int64_t N, V3<R>* vel, const R* temp, R dt, R ambient, R beta) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const R T = temp[i];
    if (T <= ambient) continue;
    const V3<R> f(R(0), std::fmax(R(0), (T - ambient) * beta), R(0));
    vel[i] = vel[i] + f * dt;
  }
}

// ---------------------------------------------------------------------------------------
// Reverse mode of the differentiable smoke step (new; the reference has no adjoint). Forward:
//   u1 = advect_vector(u), div = divergence(u1), p = K RBGS iterations from 0 (Kernel.cu:459-489),
//   u2 = u1 - grad p, rho' = advect_scalars(rho; u2).
// Each adjoint below is the exact transpose of the forward operation as written (trilinear
// weights differentiated inside the cell, the MacCormack clamp through the branch it took, the
// truncated RBGS sweeps transposed in reverse order, misses scattered where they gathered).
// Adjoints accumulate (+=) into their outputs. Serial loops keep the scatters deterministic.
// ---------------------------------------------------------------------------------------

// weights and their gradient for one trilinear cell: corner q = (q&1, (q>>1)&1, q>>2)
template <typename R>
struct TriCell {
  int32_t i0, j0, k0;
  R w[8];
  R dw[8][3];
};
template <typename R>
static TriCell<R> triCell(const V3<R>& p) {
  TriCell<R> c;
  c.i0 = floorI(p[0]);
  c.j0 = floorI(p[1]);
  c.k0 = floorI(p[2]);
  const R t[3] = {p[0] - R(c.i0), p[1] - R(c.j0), p[2] - R(c.k0)};
  for (int q = 0; q < 8; ++q) {
    const int b[3] = {q & 1, (q >> 1) & 1, q >> 2};
    R f[3], df[3];
    for (int a = 0; a < 3; ++a) {
      f[a] = b[a] ? t[a] : R(1) - t[a];
      df[a] = b[a] ? R(1) : R(-1);
    }
    c.w[q] = f[0] * f[1] * f[2];
    c.dw[q][0] = df[0] * f[1] * f[2];
    c.dw[q][1] = f[0] * df[1] * f[2];
    c.dw[q][2] = f[0] * f[1] * df[2];
  }
  return c;
}

// Adjoint of advect_scalars for one field (Kernel.cu:163-289): gin += d/d(in), gvel += d/d(vel).
template <typename R>
void advectScalarsAdjoint(const Topo& t, const V3<R>* vel, const R* in, const R* gout, R* gin, V3<R>* gvel, R dt,
                          R inv_dx) {
  const R sdt = dt * inv_dx;
  static const int offs[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  Acc a(&t);
  auto idxOf = [&](int32_t x, int32_t y, int32_t z) {
    const int64_t off = a.offset(x, y, z);
    return off == 0 ? int64_t(0) : off - 1;  // miss -> element 0
  };
  for (int64_t idx = 0; idx < t.N; ++idx) {
    const R g = gout[idx];
    if (g == R(0)) continue;
    const int32_t* c = &t.coords[3 * idx];
    const int64_t orig = idxOf(c[0], c[1], c[2]);
    const V3<R> pos = posOf<R>(t, idx);
    const V3<R> vc = vel[orig];
    const V3<R> bp = pos - vc * sdt;
    const TriCell<R> B = triCell(bp);
    int64_t bi[8];
    V3<R> velF;
    R phiF = R(0);
    // recomputed exactly as advectScalars does, so the clamp takes the same branch
    for (int q = 0; q < 8; ++q) {
      bi[q] = idxOf(B.i0 + (q & 1), B.j0 + ((q >> 1) & 1), B.k0 + (q >> 2));
      const V3<R>& v = vel[bi[q]];
      velF = V3<R>(std::fma(B.w[q], v[0], velF[0]), std::fma(B.w[q], v[1], velF[1]), std::fma(B.w[q], v[2], velF[2]));
      phiF = std::fma(in[bi[q]], B.w[q], phiF);
    }
    const V3<R> fp = bp + velF * sdt;
    const TriCell<R> F = triCell(fp);
    int64_t fi[8];
    R phiB = R(0);
    for (int q = 0; q < 8; ++q) {
      fi[q] = idxOf(F.i0 + (q & 1), F.j0 + ((q >> 1) & 1), F.k0 + (q >> 2));
      phiB = std::fma(in[fi[q]], F.w[q], phiB);
    }
    const R corr = std::fma(R(0.5f), in[orig] - phiB, phiF);
    // clamp: which argument produced the output (first minimum / maximum in evaluation order)
    int64_t nb[6];
    for (int n = 0; n < 6; ++n) nb[n] = idxOf(c[0] + offs[n][0], c[1] + offs[n][1], c[2] + offs[n][2]);
    R mn = in[orig], mx = in[orig];
    int64_t amn = orig, amx = orig;  // -1 marks phiF
    for (int n = 0; n < 6; ++n) {
      if (in[nb[n]] < mn) { mn = in[nb[n]]; amn = nb[n]; }
      if (in[nb[n]] > mx) { mx = in[nb[n]]; amx = nb[n]; }
    }
    if (phiF < mn) { mn = phiF; amn = -1; }
    if (phiF > mx) { mx = phiF; amx = -1; }
    R gcorr = R(0), gphiF = R(0);
    const R c1 = std::fmin(corr, mx);
    if (mn > c1) {
      if (amn < 0) gphiF += g; else gin[amn] += g;
    } else if (corr > mx) {
      if (amx < 0) gphiF += g; else gin[amx] += g;
    } else {
      gcorr = g;
    }
    gphiF += gcorr;
    gin[orig] += R(0.5) * gcorr;
    const R gphiB = R(-0.5) * gcorr;
    V3<R> gfp;
    for (int q = 0; q < 8; ++q) {
      gin[fi[q]] += F.w[q] * gphiB;
      for (int k = 0; k < 3; ++k) gfp[k] += F.dw[q][k] * in[fi[q]] * gphiB;
    }
    V3<R> gbp = gfp;
    const V3<R> gvelF = gfp * sdt;
    for (int q = 0; q < 8; ++q) {
      gvel[bi[q]] = gvel[bi[q]] + gvelF * B.w[q];
      gin[bi[q]] += B.w[q] * gphiF;
      const R vdot = vel[bi[q]].dot(gvelF);
      for (int k = 0; k < 3; ++k) gbp[k] += B.dw[q][k] * (vdot + in[bi[q]] * gphiF);
    }
    gvel[orig] = gvel[orig] - gbp * sdt;
  }
}

// Adjoint of advect_vector (Kernel.cu:349-426), no collision: gvel += d/d(vel) of out.
template <typename R>
void advectVectorAdjoint(const Topo& t, const V3<R>* vel, const V3<R>* gout, V3<R>* gvel, R dt, R inv_dx) {
  const R sdt = dt * inv_dx;
  Acc a(&t);
  auto idxOf = [&](int32_t x, int32_t y, int32_t z) { return a.offset(x, y, z) - 1; };  // -1: miss (reads 0)
  auto at = [&](int64_t i) { return i < 0 ? V3<R>() : vel[i]; };
  for (int64_t idx = 0; idx < t.N; ++idx) {
    const V3<R> g = gout[idx];
    if (g[0] == R(0) && g[1] == R(0) && g[2] == R(0)) continue;
    const int32_t* c = &t.coords[3 * idx];
    const V3<R> pos = posOf<R>(t, idx);
    const V3<R> vo = vel[idx];
    const V3<R> bp = pos - vo * sdt;
    const TriCell<R> B = triCell(bp);
    int64_t bi[8];
    for (int q = 0; q < 8; ++q) bi[q] = idxOf(B.i0 + (q & 1), B.j0 + ((q >> 1) & 1), B.k0 + (q >> 2));
    const V3<R> vf = trilinearVec(a, vel, bp);  // as advectVector computes it
    const V3<R> fp = bp + vf * sdt;
    const TriCell<R> F = triCell(fp);
    int64_t fi[8];
    for (int q = 0; q < 8; ++q) fi[q] = idxOf(F.i0 + (q & 1), F.j0 + ((q >> 1) & 1), F.k0 + (q >> 2));
    const V3<R> vb = trilinearVec(a, vel, fp);
    const V3<R> corr = vf + R(0.5f) * (vo - vb);
    int64_t nb[6];
    const int offs[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int n = 0; n < 6; ++n) nb[n] = idxOf(c[0] + offs[n][0], c[1] + offs[n][1], c[2] + offs[n][2]);
    V3<R> gvo, gvf, gcorr;
    for (int k = 0; k < 3; ++k) {
      // argument ids: -2 = vo, -3 = vf, -4 = a missing neighbour (value 0), >= 0 neighbour index
      R mn = vo[k], mx = vo[k];
      int64_t amn = -2, amx = -2;
      for (int n = 0; n < 6; ++n) {
        const R v = nb[n] < 0 ? R(0) : vel[nb[n]][k];
        const int64_t id = nb[n] < 0 ? -4 : nb[n];
        if (v < mn) { mn = v; amn = id; }
        if (v > mx) { mx = v; amx = id; }
      }
      if (vf[k] < mn) { mn = vf[k]; amn = -3; }
      if (vf[k] > mx) { mx = vf[k]; amx = -3; }
      const R c1 = std::fmin(corr[k], mx);
      int64_t arg = -5;  // -5: the corrected value
      if (mn > c1) arg = amn;
      else if (corr[k] > mx) arg = amx;
      if (arg == -5) gcorr[k] = g[k];
      else if (arg == -2) gvo[k] += g[k];
      else if (arg == -3) gvf[k] += g[k];
      else if (arg >= 0) gvel[arg][k] += g[k];
    }
    gvf = gvf + gcorr;
    gvo = gvo + gcorr * R(0.5);
    const V3<R> gvb = gcorr * R(-0.5);
    V3<R> gfp;
    for (int q = 0; q < 8; ++q) {
      if (fi[q] >= 0) gvel[fi[q]] = gvel[fi[q]] + gvb * F.w[q];
      const R d = at(fi[q]).dot(gvb);
      for (int k = 0; k < 3; ++k) gfp[k] += F.dw[q][k] * d;
    }
    V3<R> gbp = gfp;
    gvf = gvf + gfp * sdt;
    for (int q = 0; q < 8; ++q) {
      if (bi[q] >= 0) gvel[bi[q]] = gvel[bi[q]] + gvf * B.w[q];
      const R d = at(bi[q]).dot(gvf);
      for (int k = 0; k < 3; ++k) gbp[k] += B.dw[q][k] * d;
    }
    gvo = gvo - gbp * sdt;
    gvel[idx] = gvel[idx] + gvo;
  }
}

// Adjoint of divergence (Kernel.cu:432-453): div = 0.5 inv_dx sum_axes (u(x+e) - u(x-e)), the centre
// terms cancelling; gvel_a(y) += 0.5 inv_dx (gdiv(y - e_a) - gdiv(y + e_a)) over present voxels.
template <typename R>
void divergenceAdjoint(const Topo& t, const R* gdiv, V3<R>* gvel, R inv_dx) {
  const R h = R(0.5) * inv_dx;
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t* c = &t.coords[3 * idx];
      for (int k = 0; k < 3; ++k) {
        int32_t m[3] = {c[0], c[1], c[2]}, p[3] = {c[0], c[1], c[2]};
        m[k] -= 1;
        p[k] += 1;
        gvel[idx][k] += h * (nearest0(a, gdiv, m[0], m[1], m[2]) - nearest0(a, gdiv, p[0], p[1], p[2]));
      }
    }
  }
}

// Adjoint of the gradient subtraction (Kernel.cu:495-541) w.r.t. p: u2 = u1 - 0.5 inv_dx (p(x+e) - p(x-e)),
// so d/dp of sum(gu2 . u2) is gp(y) += 0.5 inv_dx sum_axes (gu2_a(y + e_a) - gu2_a(y - e_a)) (the minus
// sign of the subtraction included). d/du1 is the identity.
template <typename R>
void gradientAdjoint(const Topo& t, const V3<R>* gu2, R* gp, R inv_dx) {
  const R h = R(0.5) * inv_dx;
#pragma omp parallel
  {
    Acc a(&t);
#pragma omp for schedule(static)
    for (int64_t idx = 0; idx < t.N; ++idx) {
      const int32_t* c = &t.coords[3 * idx];
      R s = R(0);
      for (int k = 0; k < 3; ++k) {
        int32_t m[3] = {c[0], c[1], c[2]}, p[3] = {c[0], c[1], c[2]};
        m[k] -= 1;
        p[k] += 1;
        s += nearest0(a, gu2, p[0], p[1], p[2])[k] - nearest0(a, gu2, m[0], m[1], m[2])[k];
      }
      gp[idx] += h * s;
    }
  }
}

// Adjoint of K RBGS iterations from p = 0 (div -> p is linear): gdiv += R_K^T gp. Each colour sweep
// p_c <- (1-w) p_c + w inv6 (sum nbr p_(1-c) - dx^2 div_c) is transposed, last sweep first:
// the other colour gathers w inv6 lambda_c from its neighbours, then lambda_c feeds div_c with
// -w inv6 dx^2 and keeps (1-w) of itself.
template <typename R>
void rbgsAdjoint(const Topo& t, const R* gp_in, R* gdiv, R dx, int K, R omega) {
  const int64_t N = t.N;
  const R inv6 = R(0.166666667f), dx2 = dx * dx;
  std::vector<R> lam(gp_in, gp_in + N), add(N);
  for (int it = K - 1; it >= 0; --it) {
    for (int c = 1; c >= 0; --c) {
#pragma omp parallel
      {
        Acc a(&t);
#pragma omp for schedule(static)
        for (int64_t idx = 0; idx < N; ++idx) {
          const int32_t i = t.coords[3 * idx], j = t.coords[3 * idx + 1], k = t.coords[3 * idx + 2];
          add[idx] = R(0);
          if (((i + j + k) & 1) == c) continue;
          const R s = nearest0(a, lam.data(), i + 1, j, k) + nearest0(a, lam.data(), i - 1, j, k) +
                      nearest0(a, lam.data(), i, j + 1, k) + nearest0(a, lam.data(), i, j - 1, k) +
                      nearest0(a, lam.data(), i, j, k + 1) + nearest0(a, lam.data(), i, j, k - 1);
          add[idx] = omega * inv6 * s;
        }
      }
#pragma omp parallel for schedule(static)
      for (int64_t idx = 0; idx < N; ++idx) {
        const int32_t i = t.coords[3 * idx], j = t.coords[3 * idx + 1], k = t.coords[3 * idx + 2];
        if (((i + j + k) & 1) == c) {
          gdiv[idx] += -omega * inv6 * dx2 * lam[idx];
          lam[idx] *= (R(1) - omega);
        } else {
          lam[idx] += add[idx];
        }
      }
    }
  }
  // the first iteration starts from p = 0: lambda left on p0 has nowhere to go
}

// Forward of the differentiable step on (vel, rho), recording the intermediates its adjoint needs.
template <typename R>
void diffStep(const Topo& t, std::vector<V3<R>>& vel, std::vector<R>& rho, R dt, R dx, int K, R omega,
              std::vector<V3<R>>* u2Out) {
  const int64_t N = t.N;
  const R inv_dx = R(1) / dx;
  std::vector<V3<R>> u1(N), u2(N);
  std::vector<R> div(N), p(N, R(0)), out(N);
  advectVector(t, vel.data(), u1.data(), dt, inv_dx, (const R*)nullptr);
  divergence(t, u1.data(), div.data(), inv_dx);
  for (int it = 0; it < K; ++it) {
    rbgsColour(t, div.data(), p.data(), dx, 0, omega);
    rbgsColour(t, div.data(), p.data(), dx, 1, omega);
  }
  subtractGradient(t, u1.data(), p.data(), u2.data(), inv_dx, (const R*)nullptr);
  const R* ins[1] = {rho.data()};
  R* outs[1] = {out.data()};
  advectScalars(t, u2.data(), 1, ins, outs, dt, inv_dx, (const R*)nullptr);
  if (u2Out) *u2Out = u2;
  vel = u2;
  rho = out;
}

// Loss L = sum (rho_S - target)^2 after S steps from (vel0, rho0) and its gradient w.r.t. vel0 and
// rho0 (the FP64 host tape stores every step's inputs).
template <typename R>
R diffLossGrad(const Topo& t, const R* vel0, const R* rho0, const R* target, int steps, R dt, R dx, int K, R omega,
               R* gvel0, R* grho0) {
  const int64_t N = t.N;
  const R inv_dx = R(1) / dx;
  std::vector<std::vector<V3<R>>> tapeU(steps), tapeU2(steps);
  std::vector<std::vector<R>> tapeRho(steps);
  std::vector<V3<R>> vel(reinterpret_cast<const V3<R>*>(vel0), reinterpret_cast<const V3<R>*>(vel0) + N);
  std::vector<R> rho(rho0, rho0 + N);
  for (int s = 0; s < steps; ++s) {
    tapeU[s] = vel;
    tapeRho[s] = rho;
    diffStep(t, vel, rho, dt, dx, K, omega, &tapeU2[s]);
  }
  R loss = R(0);
  std::vector<R> grho(N);
  for (int64_t i = 0; i < N; ++i) {
    const R d = rho[i] - target[i];
    loss += d * d;
    grho[i] = R(2) * d;
  }
  std::vector<V3<R>> gvel(N);
  for (int s = steps - 1; s >= 0; --s) {
    std::vector<R> grhoIn(N, R(0));
    std::vector<V3<R>> gu2 = gvel;  // adjoint of this step's output velocity
    advectScalarsAdjoint(t, tapeU2[s].data(), tapeRho[s].data(), grho.data(), grhoIn.data(), gu2.data(), dt, inv_dx);
    std::vector<R> gp(N, R(0)), gdiv(N, R(0));
    gradientAdjoint(t, gu2.data(), gp.data(), inv_dx);  // already d/dp of u2 = u1 - grad p
    rbgsAdjoint(t, gp.data(), gdiv.data(), dx, K, omega);
    std::vector<V3<R>> gu1 = gu2;
    divergenceAdjoint(t, gdiv.data(), gu1.data(), inv_dx);
    std::vector<V3<R>> gu(N);
    advectVectorAdjoint(t, tapeU[s].data(), gu1.data(), gu.data(), dt, inv_dx);
    gvel = gu;
    grho = grhoIn;
  }
  for (int64_t i = 0; i < N; ++i) {
    gvel0[3 * i] = gvel[i][0];
    gvel0[3 * i + 1] = gvel[i][1];
    gvel0[3 * i + 2] = gvel[i][2];
    grho0[i] = grho[i];
  }
  return loss;
}

// ---------------------------------------------------------------------------------------
// HNS_Advance sequence, HNanoSolver.cu:223-456.
// ---------------------------------------------------------------------------------------
struct Params {  // CombustionParams, Kernels.cuh:5-24 (72 bytes)
  float expansionRate, temperatureRelease, buoyancyStrength, ambientTemp, vorticityScale, factorScale;
  float disturbanceEnable, disturbanceStrength, disturbanceSwirl, disturbanceThreshold, disturbanceGain, disturbanceFrequency;
  int32_t substeps;
  float gravity, maskDensityMin, maskDensityMax;
  int32_t downloadVel, simReset;
};
static_assert(sizeof(Params) == 72, "CombustionParams layout");

template <typename R>
int advance(const Topo& t, float voxelSize, int iterations, float dt, const Params& P, R* velFlat, int nState,
            const char* const* names, R* const* fields, const int32_t* present, int nSrc, const char* const* srcNames,
            const R* const* srcs, const R* srcVelV3, const R* sdf,
            const std::function<void(const R* div, R* p)>* solve = nullptr) {
  const int64_t N = t.N;
  if (N == 0) return -2;
  V3<R>* vel = reinterpret_cast<V3<R>*>(velFlat);
  const float dxf = voxelSize;
  const float inv_dxf = 1.0f / dxf;
  const R dx = R(dxf), inv_dx = R(inv_dxf);
  const int SUBS = std::max(1, int(P.substeps));
  const float dt_subf = dt / float(SUBS);
  const R dt_sub = R(dt_subf);

  std::unordered_map<std::string, int> stateIdx;
  std::vector<char> exists(nState, 0);
  for (int i = 0; i < nState; ++i) {
    stateIdx[names[i]] = i;
    exists[i] = present[i] ? 1 : 0;
  }
  auto findState = [&](const char* n) -> int {
    auto it = stateIdx.find(n);
    return (it == stateIdx.end() || !exists[it->second]) ? -1 : it->second;
  };
  auto findSrc = [&](const char* n) -> const R* {
    for (int i = 0; i < nSrc; ++i)
      if (std::strcmp(srcNames[i], n) == 0) return srcs[i];
    return nullptr;
  };
  const R* sx = findSrc("src_vel_x");
  const R* sy = findSrc("src_vel_y");
  const R* sz = findSrc("src_vel_z");
  const bool haveXYZ = !srcVelV3 && sx && sy && sz;

  std::vector<V3<R>> adv(N), tmpV(N);
  std::vector<R> divg(N, R(0)), pres(N, R(0));

  if (sdf) enforceCollision(t, vel, sdf, dx);

  // advNames snapshot before the substep loop (HNanoSolver.cu:248-250)
  std::vector<int> advList;
  for (int i = 0; i < nState; ++i)
    if (exists[i]) advList.push_back(i);

  const float omegaf = 2.0f / (1.0f + std::sin(3.14159f * dxf));  // HNanoSolver.cu:365 (sinf)
  const R omega = R(omegaf);

  for (int s = 0; s < SUBS; ++s) {
    // 1) advect velocity (mask is identically 1, Kernel.cu:74-75)
    advectVector(t, vel, adv.data(), dt_sub, inv_dx, sdf);
    // vorticity confinement, out of place
    if (int32_t(P.factorScale) != 0 && P.vorticityScale != 0.0f) {
      vorticity(t, adv.data(), tmpV.data(), dt_sub, inv_dx, R(P.vorticityScale), P.factorScale);
      std::swap(adv, tmpV);
    }
    // velocity sources (HNanoSolver.cu:286-295)
    if (srcVelV3) {
      const V3<R>* sv = reinterpret_cast<const V3<R>*>(srcVelV3);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) adv[i] = adv[i] + sv[i] * dt_sub;
    } else if (haveXYZ) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) {
        adv[i][0] += dt_sub * sx[i];
        adv[i][1] += dt_sub * sy[i];
        adv[i][2] += dt_sub * sz[i];
      }
    }
    // scalar sources (HNanoSolver.cu:298-311)
    static const char* pairs[3][2] = {{"src_temperature", "temperature"}, {"src_density", "density"}, {"src_fuel", "fuel"}};
    for (auto& pr : pairs) {
      const R* src = findSrc(pr[0]);
      const int d = findState(pr[1]);
      if (src && d >= 0) {
        R* f = fields[d];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) f[i] += dt_sub * src[i];
      }
    }
    // gravity (HNanoSolver.cu:314-318)
    if (P.gravity != 0.0f) {
      const R g = R(P.gravity);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) adv[i][1] -= g * dt_sub;
    }
    // 2) divergence
    divergence(t, adv.data(), divg.data(), inv_dx);
    // 3) needOrMake + combustion + buoyancy (HNanoSolver.cu:326-355)
    int idxs[4];
    const char* mk[4] = {"fuel", "waste", "temperature", "flame"};
    for (int q = 0; q < 4; ++q) {
      auto it = stateIdx.find(mk[q]);
      if (it == stateIdx.end()) return -9;  // caller must provide storage for the created fields
      if (!exists[it->second]) {
        std::fill(fields[it->second], fields[it->second] + N, R(0));
        exists[it->second] = 1;
      }
      idxs[q] = it->second;
    }
    combustionOxygen(N, fields[idxs[0]], fields[idxs[1]], fields[idxs[2]], divg.data(), fields[idxs[3]],
                     R(P.temperatureRelease), R(P.expansionRate));
    buoyancy(N, adv.data(), fields[idxs[2]], dt_sub, R(P.ambientTemp), R(P.buoyancyStrength));
    // 4) pressure, p = 0 then K iterations of colour 0 then colour 1 (or the given solve)
    std::fill(pres.begin(), pres.end(), R(0));
    if (solve) (*solve)(divg.data(), pres.data());
    for (int it = 0; !solve && it < iterations; ++it) {
      rbgsColour(t, divg.data(), pres.data(), dx, 0, omega);
      rbgsColour(t, divg.data(), pres.data(), dx, 1, omega);
    }
    // 4.5) disturbance: recompute divergence (drops the expansion), p = 0, max(2, K/2) more
    if (P.disturbanceEnable > 0.0f) {
      divergence(t, adv.data(), divg.data(), inv_dx);
      std::fill(pres.begin(), pres.end(), R(0));
      if (solve) (*solve)(divg.data(), pres.data());
      const int it2 = std::max(2, iterations / 2);
      for (int it = 0; !solve && it < it2; ++it) {
        rbgsColour(t, divg.data(), pres.data(), dx, 0, omega);
        rbgsColour(t, divg.data(), pres.data(), dx, 1, omega);
      }
    }
    // 5) projection
    subtractGradient(t, adv.data(), pres.data(), vel, inv_dx, sdf);
    if (sdf) enforceCollision(t, vel, sdf, dx);
    // 6) scalar advection of the snapshot with the projected velocity
    if (!advList.empty()) {
      const int m = int(advList.size());
      std::vector<std::vector<R>> outs(m, std::vector<R>(N));
      std::vector<const R*> ins(m);
      std::vector<R*> outp(m);
      for (int q = 0; q < m; ++q) {
        ins[q] = fields[advList[q]];
        outp[q] = outs[q].data();
      }
      advectScalars(t, vel, m, ins.data(), outp.data(), dt_sub, inv_dx, sdf);
      for (int q = 0; q < m; ++q) std::copy(outs[q].begin(), outs[q].end(), fields[advList[q]]);
    }
  }
  return 0;
}

}  // namespace oracle

// =========================================================================================
// C API (ctypes). Arrays are leaf-dense, length N (Vec3 arrays are AoS, length 3N).
// =========================================================================================
using oracle::Topo;
using oracle::V3;

extern "C" {

// Builds the topology. Returns null when n is not a positive multiple of 512 or when a leaf
// appears twice (the reference builder never produces either).
void* oracle_topo_create(const int32_t* coords, int64_t n) {
  if (n <= 0 || (n % 512) != 0) return nullptr;
  auto* t = new Topo();
  t->N = n;
  t->L = n / 512;
  t->coords.assign(coords, coords + 3 * n);
  struct K {
    uint64_t root;
    uint32_t local;
    int64_t leaf;
  };
  std::vector<K> keys(t->L);
  for (int64_t l = 0; l < t->L; ++l) {
    const int32_t* c = &coords[3 * 512 * l];
    keys[l] = {oracle::rootKey(c[0], c[1], c[2]), oracle::tileLocalKey(c[0], c[1], c[2]), l};
  }
  std::sort(keys.begin(), keys.end(), [](const K& a, const K& b) {
    return a.root != b.root ? a.root < b.root : a.local < b.local;
  });
  for (int64_t r = 0; r < t->L; ++r) {
    const int32_t* c = &coords[3 * 512 * keys[r].leaf];
    if (!t->rankOf.emplace(oracle::packLeaf(c[0], c[1], c[2]), r).second) {
      delete t;
      return nullptr;
    }
  }
  return t;
}
void oracle_topo_destroy(void* t) { delete static_cast<Topo*>(t); }

// NanoVDB-semantic offset of a voxel (0 = miss).
int64_t oracle_offset(void* t, int32_t x, int32_t y, int32_t z) {
  oracle::Acc a(static_cast<Topo*>(t));
  return a.offset(x, y, z);
}

// Canonical leaf order (voxelsToGrid's sort) of L leaf origins: perm[r] = input index of rank r.
void oracle_leaf_order(const int32_t* origins, int64_t L, int64_t* perm) {
  std::vector<int64_t> p(L);
  for (int64_t i = 0; i < L; ++i) p[i] = i;
  std::stable_sort(p.begin(), p.end(), [&](int64_t a, int64_t b) {
    const int32_t* ca = &origins[3 * a];
    const int32_t* cb = &origins[3 * b];
    const uint64_t ra = oracle::rootKey(ca[0], ca[1], ca[2]), rb = oracle::rootKey(cb[0], cb[1], cb[2]);
    if (ra != rb) return ra < rb;
    return oracle::tileLocalKey(ca[0], ca[1], ca[2]) < oracle::tileLocalKey(cb[0], cb[1], cb[2]);
  });
  std::copy(p.begin(), p.end(), perm);
}

// Samplers, exposed for the known-answer tests of Tests/IndexGrid.cpp:138-142, :213-223, :278-281.
float oracle_sample_nearest_f32(void* t, const float* data, int32_t x, int32_t y, int32_t z) {
  oracle::Acc a(static_cast<Topo*>(t));
  return oracle::nearest0(a, data, x, y, z);
}
float oracle_sample_trilinear_f32(void* t, const float* data, float x, float y, float z) {
  oracle::Acc a(static_cast<Topo*>(t));
  return oracle::trilinearScalar<float>(a, data, V3<float>(x, y, z));
}
void oracle_sample_trilinear_vec_f32(void* t, const float* data, float x, float y, float z, float* out3) {
  oracle::Acc a(static_cast<Topo*>(t));
  const V3<float> v = oracle::trilinearVec<float>(a, reinterpret_cast<const V3<float>*>(data), V3<float>(x, y, z));
  out3[0] = v[0];
  out3[1] = v[1];
  out3[2] = v[2];
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

#define ORACLE_KERNELS(SUF, R)                                                                                       \
  void oracle_advect_vector_##SUF(void* t, const R* vel, R* out, float dt, float inv_dx, const R* sdf) {              \
    oracle::advectVector<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel), reinterpret_cast<V3<R>*>(out), \
                            R(dt), R(inv_dx), sdf);                                                                  \
  }                                                                                                                  \
  void oracle_advect_scalars_##SUF(void* t, const R* vel, int m, const R* const* in, R* const* out, float dt,         \
                                   float inv_dx, const R* sdf) {                                                     \
    oracle::advectScalars<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel), m, in, out, R(dt), R(inv_dx), \
                             sdf);                                                                                   \
  }                                                                                                                  \
  void oracle_advect_scalar_##SUF(void* t, const R* vel, const R* in, R* out, float dt, float inv_dx, const R* sdf) {  \
    oracle::advectScalar<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel), in, out, R(dt), R(inv_dx), sdf); \
  }                                                                                                                  \
  void oracle_vorticity_##SUF(void* t, const R* vel, R* out, float dt, float inv_dx, float scale, float factorScale) {  \
    oracle::vorticity<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel), reinterpret_cast<V3<R>*>(out),    \
                         R(dt), R(inv_dx), R(scale), factorScale);                                                   \
  }                                                                                                                  \
  void oracle_divergence_##SUF(void* t, const R* vel, R* div, float inv_dx) {                                         \
    oracle::divergence<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel), div, R(inv_dx));               \
  }                                                                                                                  \
  void oracle_rbgs_##SUF(void* t, const R* div, R* p, float dx, int color, float omega) {                             \
    oracle::rbgsColour<R>(*static_cast<Topo*>(t), div, p, R(dx), color, R(omega));                                   \
  }                                                                                                                  \
  void oracle_residual_##SUF(void* t, const R* p, const R* div, float dx, double* r2, double* d2) {                   \
    oracle::residual<R>(*static_cast<Topo*>(t), p, div, R(dx), r2, d2);                                              \
  }                                                                                                                  \
  int oracle_cg_##SUF(void* t, const R* div, float dx, double tol, int maxIters, R* p, double* rel) {              \
    return oracle::conjugateGradient<R>(*static_cast<Topo*>(t), div, R(dx), tol, maxIters, p, rel);                 \
  }                                                                                                                  \
  void oracle_subtract_gradient_##SUF(void* t, const R* vel, const R* p, R* out, float inv_dx, const R* sdf) {        \
    oracle::subtractGradient<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel), p,                       \
                                reinterpret_cast<V3<R>*>(out), R(inv_dx), sdf);                                      \
  }                                                                                                                  \
  void oracle_enforce_collision_##SUF(void* t, R* vel, const R* sdf, float voxelSize) {                               \
    oracle::enforceCollision<R>(*static_cast<Topo*>(t), reinterpret_cast<V3<R>*>(vel), sdf, R(voxelSize));           \
  }                                                                                                                  \
  void oracle_combustion_##SUF(int64_t n, R* fuel, R* waste, R* temp, R* div, R* flame, float gain, float expansion) { \
    oracle::combustionOxygen<R>(n, fuel, waste, temp, div, flame, R(gain), R(expansion));                            \
  }                                                                                                                  \
  void oracle_buoyancy_##SUF(int64_t n, R* vel, const R* temp, float dt, float ambient, float beta) {                  \
    oracle::buoyancy<R>(n, reinterpret_cast<V3<R>*>(vel), temp, R(dt), R(ambient), R(beta));                         \
  }                                                                                                                  \
  int oracle_advance_##SUF(void* t, float voxelSize, int iterations, float dt, const oracle::Params* P, R* vel,     \
                           int nState, const char* const* names, R* const* fields, const int32_t* present, int nSrc, \
                           const char* const* srcNames, const R* const* srcs, const R* srcVelV3, const R* sdf) {     \
    return oracle::advance<R>(*static_cast<Topo*>(t), voxelSize, iterations, dt, *P, vel, nState, names, fields,     \
                              present, nSrc, srcNames, srcs, srcVelV3, sdf);                                         \
  }

#define ORACLE_ADJOINTS(SUF, R)                                                                                      \
  void oracle_advect_scalars_adj_##SUF(void* t, const R* vel, const R* in, const R* gout, R* gin, R* gvel, float dt,  \
                                       float inv_dx) {                                                               \
    oracle::advectScalarsAdjoint<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel), in, gout, gin,      \
                                    reinterpret_cast<V3<R>*>(gvel), R(dt), R(inv_dx));                               \
  }                                                                                                                  \
  void oracle_advect_vector_adj_##SUF(void* t, const R* vel, const R* gout, R* gvel, float dt, float inv_dx) {        \
    oracle::advectVectorAdjoint<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(vel),                      \
                                   reinterpret_cast<const V3<R>*>(gout), reinterpret_cast<V3<R>*>(gvel), R(dt),      \
                                   R(inv_dx));                                                                       \
  }                                                                                                                  \
  void oracle_divergence_adj_##SUF(void* t, const R* gdiv, R* gvel, float inv_dx) {                                  \
    oracle::divergenceAdjoint<R>(*static_cast<Topo*>(t), gdiv, reinterpret_cast<V3<R>*>(gvel), R(inv_dx));          \
  }                                                                                                                  \
  void oracle_gradient_adj_##SUF(void* t, const R* gu2, R* gp, float inv_dx) {                                       \
    oracle::gradientAdjoint<R>(*static_cast<Topo*>(t), reinterpret_cast<const V3<R>*>(gu2), gp, R(inv_dx));         \
  }                                                                                                                  \
  void oracle_rbgs_adj_##SUF(void* t, const R* gp, R* gdiv, float dx, int K, float omega) {                          \
    oracle::rbgsAdjoint<R>(*static_cast<Topo*>(t), gp, gdiv, R(dx), K, R(omega));                                   \
  }                                                                                                                  \
  void oracle_diff_forward_##SUF(void* t, R* vel, R* rho, int steps, float dt, float dx, int K, float omega) {       \
    const Topo& T = *static_cast<Topo*>(t);                                                                          \
    std::vector<V3<R>> v(reinterpret_cast<V3<R>*>(vel), reinterpret_cast<V3<R>*>(vel) + T.N);                        \
    std::vector<R> r(rho, rho + T.N);                                                                                \
    for (int s = 0; s < steps; ++s) oracle::diffStep<R>(T, v, r, R(dt), R(dx), K, R(omega), nullptr);               \
    std::copy(v.begin(), v.end(), reinterpret_cast<V3<R>*>(vel));                                                    \
    std::copy(r.begin(), r.end(), rho);                                                                              \
  }                                                                                                                  \
  double oracle_diff_loss_grad_##SUF(void* t, const R* vel0, const R* rho0, const R* target, int steps, float dt,    \
                                     float dx, int K, float omega, R* gvel0, R* grho0) {                             \
    return double(oracle::diffLossGrad<R>(*static_cast<Topo*>(t), vel0, rho0, target, steps, R(dt), R(dx), K,       \
                                          R(omega), gvel0, grho0));                                                  \
  }

ORACLE_KERNELS(f32, float)
ORACLE_KERNELS(f64, double)

// HNS_Advance with every pressure solve by the multigrid-preconditioned CG to tol (the library's
// HNS_SOLVER_MGCG) instead of the fixed RBGS count; iterations[k] receives the k-th solve's count
// (up to max_solves entries). FP32 only.
int oracle_advance_mgcg_f32(void* t, float voxelSize, float tol, float dt, const oracle::Params* P, float* vel,
                            int nState, const char* const* names, float* const* fields, const int32_t* present,
                            int nSrc, const char* const* srcNames, const float* const* srcs, const float* srcVelV3,
                            const float* sdf, int32_t* iterations, int max_solves) {
  const Topo& topo = *static_cast<Topo*>(t);
  const int64_t L = topo.N / 512;
  std::vector<int32_t> org(3 * (size_t)L);
  for (int64_t l = 0; l < L; ++l)
    for (int a = 0; a < 3; ++a) org[3 * l + a] = topo.coords[3 * (l * 512) + a];
  oracle_mg::MG* mg = oracle_mg::create(org.data(), L);
  int k = 0;
  const std::function<void(const float*, float*)> solve = [&](const float* div, float* p) {
    double rel = 0.0;
    const int it = oracle_mg::mgcg(*mg, div, voxelSize, (double)tol, 1000, p, &rel);
    if (k < max_solves) iterations[k] = it;
    ++k;
  };
  const int st = oracle::advance<float>(topo, voxelSize, 0, dt, *P, vel, nState, names, fields, present, nSrc,
                                        srcNames, srcs, srcVelV3, sdf, &solve);
  oracle_mg::destroy(mg);
  return st;
}
ORACLE_ADJOINTS(f32, float)
ORACLE_ADJOINTS(f64, double)

}  // extern "C"
