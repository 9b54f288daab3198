// Pressure kernels: red-black Gauss-Seidel SOR sweeps (redBlackGaussSeidelUpdate,
// src/Cuda/Kernel.cu:459-489) and the Poisson residual (new; compute_residual is declared but
// never defined in the reference, src/Cuda/Kernels.cuh:59-60).
//
// Pressure and divergence are stored colour-split: all colour-0 voxels ((x+y+z)&1 == 0) of
// every leaf, then all colour-1 voxels; inside a colour, leaf l's 256 voxels are contiguous at
// 256*l + 32*x + 4*y + (z>>1). A colour sweep therefore launches 256 lanes per leaf (no idle
// half-wave as in the reference's all-N launch with early exit), reads its own colour's div and
// p and writes p as fully coalesced 1 KiB runs, and gathers the other colour's 6 neighbours
// from the same leaf or a face-neighbour leaf. Per sweep and voxel of the domain that is
// 2 B div + 2 B p + 2 B p written + 2 B of the other colour = 8 B; 16 B per iteration.
#include <cstdlib>
#include <mutex>

#include "hns_internal.hpp"
#include "hns_mg.hpp"

namespace hns {
namespace {

// Scratch of the allocation-free residual (hnsk_residual_noalloc): the per-leaf partial sums
// (2 doubles per leaf) of k_residual for up to kNoAllocMaxLeaves leaves, 2 MiB per device, held
// until the process ends. Host code never passes `scratch` itself to a kernel (in host code the
// name is only a stand-in for the symbol): it passes the device address hipGetSymbolAddress gave
// for the current device, looked up once per device.
static __device__ __align__(256) double scratch[2 * (size_t)HNS_NOALLOC_MAX_LEAVES];

constexpr int kNbXm = 4, kNbXp = 22, kNbYm = 10, kNbYp = 16, kNbZm = 12, kNbZp = 14;  // (dx+1)*9+(dy+1)*3+(dz+1)

// LPW leaves per workgroup (256 lanes each); the workgroup group index is XCD-mapped. The other
// colour's plane of the leaf and the 6 half-faces of it that the sweep reads across the leaf
// faces (32 voxels each: x faces indexed (y, z>>1), y faces (x, z>>1), z faces (x, y>>1)) are
// staged in LDS first, so the 6 neighbour reads of every voxel come from LDS.
constexpr int kRbTile = kHalfLeaf + 6 * 32;
template <int COLOUR, bool FIRST, int LPW>
__global__ __launch_bounds__(256 * LPW) void k_rbgs(DevTopo T, const float* __restrict__ div, float* __restrict__ p,
                                                    float dx2, float omega, const int* __restrict__ stop) {
  __shared__ float s_o[LPW][kRbTile];
  if (stop && *stop) return;  // a multigrid V-cycle of a solve that has stopped
  const int32_t lw = LPW > 1 ? (int32_t)(threadIdx.x >> 8) : 0;
  const int32_t l = leaf_of_block() * LPW + lw;
  const bool live = l < T.L;  // a partially filled workgroup still reaches the barrier
  const int32_t t = threadIdx.x & 255;
  const int32_t x = t >> 5, y = (t >> 2) & 7, zh = t & 3;
  const int32_t z = (zh << 1) | ((x + y + COLOUR) & 1);
  const int32_t half = T.L * kHalfLeaf;
  const float* __restrict__ po = p + (1 - COLOUR) * half;
  float* __restrict__ pc = p + COLOUR * half;
  const int32_t base = l * kHalfLeaf;
  float* S = s_o[lw];
  constexpr bool kNbrs = !(FIRST && COLOUR == 0);  // the first colour-0 sweep sees p = 0
  if (kNbrs) {
    if (live) {
      S[t] = po[base + t];
      if (t < 192) {
        const int f = t >> 5, a = (t >> 2) & 7, bh = t & 3;
        constexpr int o = 1 - COLOUR;  // parity offset of the other colour
        int vx, vy, vz, k;
        switch (f) {
          case 0: k = kNbXm; vx = 7; vy = a; vz = 2 * bh + ((a + 7 + o) & 1); break;
          case 1: k = kNbXp; vx = 0; vy = a; vz = 2 * bh + ((a + o) & 1); break;
          case 2: k = kNbYm; vx = a; vy = 7; vz = 2 * bh + ((a + 7 + o) & 1); break;
          case 3: k = kNbYp; vx = a; vy = 0; vz = 2 * bh + ((a + o) & 1); break;
          case 4: k = kNbZm; vx = a; vy = 2 * bh + ((a + 7 + o) & 1); vz = 7; break;
          default: k = kNbZp; vx = a; vy = 2 * bh + ((a + o) & 1); vz = 0; break;
        }
        const int32_t nl = T.nbr[kNbrStride * l + k];
        S[kHalfLeaf + t] = nl < 0 ? 0.f : po[nl * kHalfLeaf + cs_local(vx, vy, vz)];
      }
    }
    __syncthreads();
  }
  if (!live) return;
  float s = 0.f;
  if (kNbrs) {
    const float pxp = S[x < 7 ? cs_local(x + 1, y, z) : kHalfLeaf + 32 + y * 4 + (z >> 1)];
    const float pxm = S[x > 0 ? cs_local(x - 1, y, z) : kHalfLeaf + 0 + y * 4 + (z >> 1)];
    const float pyp = S[y < 7 ? cs_local(x, y + 1, z) : kHalfLeaf + 96 + x * 4 + (z >> 1)];
    const float pym = S[y > 0 ? cs_local(x, y - 1, z) : kHalfLeaf + 64 + x * 4 + (z >> 1)];
    const float pzp = S[z < 7 ? cs_local(x, y, z + 1) : kHalfLeaf + 160 + x * 4 + (y >> 1)];
    const float pzm = S[z > 0 ? cs_local(x, y, z - 1) : kHalfLeaf + 128 + x * 4 + (y >> 1)];
    s = pxp + pxm + pyp + pym + pzp + pzm;
  }
  const float dv = div[COLOUR * half + base + t];
  const float pOld = FIRST ? 0.f : pc[base + t];
  pc[base + t] = sor_update(s, dv, pOld, dx2, omega);
}

// One full red-black iteration per launch (colour 0 then colour 1), temporally blocked in LDS.
// The workgroup stages its leaf plus a 2-voxel halo of p (12^3) and the divergence of the inner
// 10^3, updates colour 0 on the 10^3 (its halo ring included, so colour 1 of the 8^3 core sees
// exactly the values the global colour-0 sweep would produce), then colour 1 on the 8^3 core, and
// writes the core. Cells of absent leaves stay 0 and are never updated (the Dirichlet boundary of
// Kernel.cu:477-482). Reads p_in, writes p_out (distinct: neighbours read the halo from p_in while
// this workgroup writes). HBM per voxel and iteration: 4 B p + 4 B div read, 4 B p written.
constexpr int kT = 12;  // tile edge: 8 + 2 * 2
__device__ __forceinline__ int tile_idx(int x, int y, int z) { return (x * kT + y) * kT + z; }

template <bool FIRST>
__global__ __launch_bounds__(512) void k_rbgs_iter(DevTopo T, const float* __restrict__ div,
                                                   const float* __restrict__ p_in, float* __restrict__ p_out,
                                                   float dx2, float omega) {
  __shared__ int32_t s_nb[27];
  __shared__ float s_p[kT * kT * kT];
  __shared__ float s_d[10 * 10 * 10];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const int32_t tid = threadIdx.x;
  if (tid < 27) s_nb[tid] = T.nbr[kNbrStride * l + tid];
  __syncthreads();
  const int32_t half = T.L * kHalfLeaf;
  // tile coordinate t in [0,12) is local coordinate t-2; its leaf offset is (t-2)>>3 in {-1,0,1}
  for (int c = tid; c < kT * kT * kT; c += 512) {
    const int tx = c / (kT * kT), ty = (c / kT) % kT, tz = c % kT;
    float v = 0.f;
    if (!FIRST) {
      const int lx = tx - 2, ly = ty - 2, lz = tz - 2;
      const int32_t nl = s_nb[((lx >> 3) + 1) * 9 + ((ly >> 3) + 1) * 3 + ((lz >> 3) + 1)];
      if (nl >= 0) {
        const int x = lx & 7, y = ly & 7, z = lz & 7;
        v = p_in[((x + y + z) & 1) * half + nl * kHalfLeaf + cs_local(x, y, z)];
      }
    }
    s_p[c] = v;
  }
  for (int c = tid; c < 1000; c += 512) {
    const int tx = c / 100 + 1, ty = (c / 10) % 10 + 1, tz = c % 10 + 1;
    const int lx = tx - 2, ly = ty - 2, lz = tz - 2;
    const int32_t nl = s_nb[((lx >> 3) + 1) * 9 + ((ly >> 3) + 1) * 3 + ((lz >> 3) + 1)];
    float v = 0.f;
    if (nl >= 0) {
      const int x = lx & 7, y = ly & 7, z = lz & 7;
      v = div[((x + y + z) & 1) * half + nl * kHalfLeaf + cs_local(x, y, z)];
    }
    s_d[c] = v;
  }
  __syncthreads();
  // colour 0 on the 10^3 region [1,11)^3: 100 rows of 5 cells
  if (tid < 500) {
    const int row = tid / 5, j = tid % 5;
    const int tx = 1 + row / 10, ty = 1 + row % 10;
    const int tz = 1 + ((tx + ty + 1) & 1) + 2 * j;  // (tx+ty+tz) even: colour 0
    const int lx = tx - 2, ly = ty - 2, lz = tz - 2;
    if (s_nb[((lx >> 3) + 1) * 9 + ((ly >> 3) + 1) * 3 + ((lz >> 3) + 1)] >= 0) {
      const float s = s_p[tile_idx(tx + 1, ty, tz)] + s_p[tile_idx(tx - 1, ty, tz)] + s_p[tile_idx(tx, ty + 1, tz)] +
                      s_p[tile_idx(tx, ty - 1, tz)] + s_p[tile_idx(tx, ty, tz + 1)] + s_p[tile_idx(tx, ty, tz - 1)];
      const int c = tile_idx(tx, ty, tz);
      s_p[c] = sor_update(s, s_d[((tx - 1) * 10 + (ty - 1)) * 10 + (tz - 1)], s_p[c], dx2, omega);
    }
  }
  __syncthreads();
  // colour 1 on the 8^3 core [2,10)^3: 64 rows of 4 cells (always present)
  if (tid < 256) {
    const int row = tid >> 2, j = tid & 3;
    const int tx = 2 + (row >> 3), ty = 2 + (row & 7);
    const int tz = 2 + ((tx + ty + 1) & 1) + 2 * j;  // (tx+ty+tz) odd: colour 1
    const float s = s_p[tile_idx(tx + 1, ty, tz)] + s_p[tile_idx(tx - 1, ty, tz)] + s_p[tile_idx(tx, ty + 1, tz)] +
                    s_p[tile_idx(tx, ty - 1, tz)] + s_p[tile_idx(tx, ty, tz + 1)] + s_p[tile_idx(tx, ty, tz - 1)];
    const int c = tile_idx(tx, ty, tz);
    s_p[c] = sor_update(s, s_d[((tx - 1) * 10 + (ty - 1)) * 10 + (tz - 1)], s_p[c], dx2, omega);
  }
  __syncthreads();
  // write the core: lane = colour-split local index within each colour
  {
    const int colour = tid >> 8, t = tid & 255;
    const int x = t >> 5, y = (t >> 2) & 7, z = ((t & 3) << 1) | ((x + y + colour) & 1);
    p_out[colour * half + l * kHalfLeaf + t] = s_p[tile_idx(x + 2, y + 2, z + 2)];
  }
}

// Residual r = div - (sum6 p_nbr - 6 p) / dx^2 in FP64; per-leaf partial sums of r^2 and div^2.
// 256 lanes per leaf, each owning a colour-split voxel pair.
__global__ __launch_bounds__(256) void k_residual(DevTopo T, const float* __restrict__ p,
                                                  const float* __restrict__ div, double inv_dx2,
                                                  double* __restrict__ partial) {
  __shared__ double s_red[2][4];
  __shared__ double s_tile[kPairTile];
  const int32_t l = leaf_of_block();
  if (l >= T.L) return;
  const FaceRow Fr = face_row(T, l);
  const int32_t half = T.L * kHalfLeaf;
  const int32_t n = threadIdx.x;
  const int c0 = ((n >> 5) + (n >> 2)) & 1;
  const int32_t i0 = c0 * half + l * kHalfLeaf + n, i1 = (c0 ^ 1) * half + l * kHalfLeaf + n;
  auto get = [&](int c, int32_t i) { return (double)p[c * half + i]; };
  const double p0 = (double)p[i0], p1 = (double)p[i1];
  pair_tile_halo(Fr, get, s_tile);
  s_tile[c0 * kHalfLeaf + n] = p0;
  s_tile[(c0 ^ 1) * kHalfLeaf + n] = p1;
  __syncthreads();
  double s0, s1;
  pair_tile_sums(s_tile, n, p0, p1, s0, s1);
  const double d0 = (double)div[i0], d1 = (double)div[i1];
  const double r0 = d0 - (s0 - 6.0 * p0) * inv_dx2, r1 = d1 - (s1 - 6.0 * p1) * inv_dx2;
  double a = r0 * r0 + r1 * r1, b = d0 * d0 + d1 * d1;
  // wave64 reduction with DPP/permute shuffles, then across the 4 waves through LDS
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int w = n >> 6;
  if ((n & 63) == 0) {
    s_red[0][w] = a;
    s_red[1][w] = b;
  }
  __syncthreads();
  if (n == 0) {
    double ta = 0.0, tb = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ta += s_red[0][i];
      tb += s_red[1][i];
    }
    partial[2 * l] = ta;
    partial[2 * l + 1] = tb;
  }
}

// Deterministic final reduction: one workgroup, fixed summation order.
__global__ __launch_bounds__(1024) void k_reduce_pairs(const double* __restrict__ partial, int32_t L,
                                                       double* __restrict__ out) {
  __shared__ double s_red[2][16];
  double a = 0.0, b = 0.0;
  for (int32_t i = threadIdx.x; i < L; i += blockDim.x) {
    a += partial[2 * i];
    b += partial[2 * i + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[0][w] = a;
    s_red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tb = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) {
      ta += s_red[0][i];
      tb += s_red[1][i];
    }
    out[0] = ta;
    out[1] = tb;
  }
}

}  // namespace

namespace {
template <int LPW>
void launch_rbgs_lpw(const DevTopo& T, const float* div_cs, float* p_cs, float dx2, int colour, float omega, bool first,
                     hipStream_t s, const int* stop = nullptr) {
  const unsigned g = grid_for_leaves((T.L + LPW - 1) / LPW);
  const unsigned b = 256 * LPW;
  if (colour == 0) {
    if (first) k_rbgs<0, true, LPW><<<g, b, 0, s>>>(T, div_cs, p_cs, dx2, omega, stop);
    else k_rbgs<0, false, LPW><<<g, b, 0, s>>>(T, div_cs, p_cs, dx2, omega, stop);
  } else {
    if (first) k_rbgs<1, true, LPW><<<g, b, 0, s>>>(T, div_cs, p_cs, dx2, omega, stop);
    else k_rbgs<1, false, LPW><<<g, b, 0, s>>>(T, div_cs, p_cs, dx2, omega, stop);
  }
}
int rbgs_lpw() {
  static const int v = [] {
    const char* e = std::getenv("HNS_RBGS_LPW");
    const int k = e ? std::atoi(e) : 1;
    return (k == 2 || k == 4) ? k : 1;
  }();
  return v;
}
}  // namespace

// The sweep on A u = b with A = 6 I - (face neighbours): div = b and dx^2 = -1 make the SOR target
// (sum - div dx^2) / 6 exactly (sum + b) / 6; omega = 1 (Gauss-Seidel).
void launch_rbgs_rhs(const DevTopo& T, const float* b_cs, float* u_cs, int colour, bool first, float omega,
                     const int* stop, hipStream_t s) {
  launch_rbgs_lpw<1>(T, b_cs, u_cs, -1.0f, colour, omega, first, s, stop);
  HNS_HIP(hipGetLastError());
}

void launch_rbgs(const DevTopo& T, const float* div_cs, float* p_cs, float dx, int colour, float omega, bool first,
                 hipStream_t s) {
  const float dx2 = dx * dx;
  switch (rbgs_lpw()) {
    case 2: launch_rbgs_lpw<2>(T, div_cs, p_cs, dx2, colour, omega, first, s); break;
    case 4: launch_rbgs_lpw<4>(T, div_cs, p_cs, dx2, colour, omega, first, s); break;
    default: launch_rbgs_lpw<1>(T, div_cs, p_cs, dx2, colour, omega, first, s); break;
  }
  HNS_HIP(hipGetLastError());
}

void launch_rbgs_iteration(const DevTopo& T, const float* div_cs, const float* p_in, float* p_out, float dx,
                           float omega, bool first, hipStream_t s) {
  const unsigned g = grid_for_leaves(T.L);
  const float dx2 = dx * dx;
  if (first) k_rbgs_iter<true><<<g, 512, 0, s>>>(T, div_cs, p_in, p_out, dx2, omega);
  else k_rbgs_iter<false><<<g, 512, 0, s>>>(T, div_cs, p_in, p_out, dx2, omega);
  HNS_HIP(hipGetLastError());
}

void launch_residual(const DevTopo& T, const float* p_cs, const float* div_cs, float dx, double* partial,
                     double* out, hipStream_t s) {
  const double inv_dx2 = 1.0 / ((double)dx * (double)dx);
  k_residual<<<grid_for_leaves(T.L), 256, 0, s>>>(T, p_cs, div_cs, inv_dx2, partial);
  HNS_HIP(hipGetLastError());
  k_reduce_pairs<<<1, 1024, 0, s>>>(partial, T.L, out);
  HNS_HIP(hipGetLastError());
}

// Device address of `scratch` on the current device, looked up once per device and kept (each
// device holds its own copy of the array at its own address). The lookup runs at set-up, when a
// topology is built on the device (noalloc_setup); the runtime may back the array then.
static double* scratch_ptr() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* ptrs[kMaxDevices] = {};
  int dev = 0;
  HNS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(HNS_ERR_INVALID_ARGUMENT, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    HNS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptrs[dev] = static_cast<double*>(p);
  }
  return ptrs[dev];
}

void noalloc_setup() { (void)scratch_ptr(); }

void launch_residual_noalloc(const DevTopo& T, const float* p_cs, const float* div_cs, float dx, double* out,
                             hipStream_t s) {
  if (T.L > HNS_NOALLOC_MAX_LEAVES)
    fail(HNS_ERR_INVALID_ARGUMENT, "hnsk_residual_noalloc: topology has " + std::to_string(T.L) +
                                       " leaves, above HNS_NOALLOC_MAX_LEAVES (" +
                                       std::to_string(HNS_NOALLOC_MAX_LEAVES) + ")");
  launch_residual(T, p_cs, div_cs, dx, scratch_ptr(), out, s);
}

}  // namespace hns
