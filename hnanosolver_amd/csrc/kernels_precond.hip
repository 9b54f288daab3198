// Kernels of the learned pressure preconditioner (no reference counterpart; the reference runs a
// fixed red-black Gauss-Seidel count, src/Cuda/HNanoSolver.cu:364-373).
//
// Convolutions are implicit GEMMs on gfx950 MFMA (v_mfma_f32_16x16x32_bf16, BF16 operands, FP32
// accumulation), written transposed so the output is channel-contiguous per lane:
//   D[cout][cell] = sum_k W^T[cout][k] X[k][cell],  k = (tap, cin) flattened tap-major,
// one MFMA per 16 output channels x 16 cells x 32 k. The B operand of lane l is 8 consecutive
// input channels (16 B) of the neighbour of cell (l & 15) at the lane group's tap, gathered
// straight from the level's channel-contiguous bf16 tensor through the leaf neighbour table; the
// A operand is the weight tile, pre-swizzled so each lane loads its 16 B contiguously. Lane l
// ends with channels 4*(l>>4) .. +3 of cell (l & 15), stored as one 8 B write.
// Backward-data is the same kernel on the mirrored, transposed weights; weight gradients are
// MFMAs over cells (k = cell) on tiles staged and transposed through LDS, one workgroup per
// (tap, cell range), partials reduced in a fixed order.
#include <algorithm>
#include <cmath>

#include "hns_cg.hpp"
#include "hns_precond.hpp"

namespace hns {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

const PcLayer kPcNet[kPcLayers] = {{"c0", 27, 1, 16, 0},  {"c1", 27, 16, 32, 1}, {"c2", 27, 32, 64, 2},
                                   {"c3", 27, 64, 64, 2}, {"u2", 1, 64, 32, 2},  {"c5", 27, 32, 32, 1},
                                   {"u1", 1, 32, 16, 1},  {"c6", 27, 16, 16, 0}, {"out", 1, 16, 1, 0}};

namespace {
inline int64_t swz_size(int taps, int cin, int cout) {
  if (cout < 16 || cin == 0) return 0;
  const int64_t nk = (taps * cin + 31) / 32;
  return nk * (cout / 16) * 64 * 8;
}
}  // namespace

PcOffsets::PcOffsets() {
  param[0] = fwd[0] = bwd[0] = 0;
  for (int i = 0; i < kPcLayers; ++i) {
    const PcLayer& Ly = kPcNet[i];
    param[i + 1] = param[i] + (int64_t)Ly.taps * Ly.cin * Ly.cout;
    fwd[i + 1] = fwd[i] + swz_size(Ly.taps, Ly.cin, Ly.cout);
    // backward data is not needed for the input layer (r has no gradient) nor the fused output
    bwd[i + 1] = bwd[i] + ((i == kC0 || i == kOut) ? 0 : swz_size(Ly.taps, Ly.cout, Ly.cin));
  }
}
const PcOffsets& pc_offsets() {
  static const PcOffsets o;
  return o;
}

namespace {

constexpr int kTpb = 256;

__device__ __forceinline__ floatx4 mfma(const bf16x8& a, const bf16x8& b, const floatx4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int v = 0; v < 8; ++v) z[v] = (__bf16)0.0f;
  return z;
}

// Neighbour of level cell `c` (coordinates leaf, cx, cy, cz at edge 2^lgS) at offset (dx,dy,dz)
// in {-1,0,1}^3; -1 when the neighbour's leaf is absent.
__device__ __forceinline__ int32_t nbr_cell(const DevTopo& T, int lgS, int32_t leaf, int cx, int cy, int cz, int dx,
                                            int dy, int dz) {
  const int Sm = (1 << lgS) - 1;
  int nx = cx + dx, ny = cy + dy, nz = cz + dz;
  const int ox = nx >> lgS, oy = ny >> lgS, oz = nz >> lgS;  // -1, 0 or 1
  int32_t nl = leaf;
  if (ox | oy | oz) nl = T.nbr[kNbrStride * leaf + (ox + 1) * 9 + (oy + 1) * 3 + (oz + 1)];
  if (nl < 0) return -1;
  nx &= Sm;
  ny &= Sm;
  nz &= Sm;
  return (nl << (3 * lgS)) | (nx << (2 * lgS)) | (ny << lgS) | nz;
}

struct CellPos {
  int32_t c, leaf;
  int x, y, z;
  bool ok;
};
__device__ __forceinline__ CellPos cell_pos(int32_t c, int32_t ncell, int lgS) {
  CellPos p;
  p.c = c;
  p.ok = c < ncell;
  p.leaf = c >> (3 * lgS);
  const int32_t loc = c & ((1 << (3 * lgS)) - 1);
  p.x = loc >> (2 * lgS);
  p.y = (loc >> lgS) & ((1 << lgS) - 1);
  p.z = loc & ((1 << lgS) - 1);
  return p;
}

// ---- weight swizzle: FP32 master W[tap][cin][cout] -> per-lane MFMA A operands -------------
// TRANS: the operand of the transposed conv with mirrored taps (backward data).
__global__ void k_pc_swizzle(const float* __restrict__ W, __bf16* __restrict__ dst, int taps, int cin, int cout,
                             int trans, int64_t count) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const int Ci = trans ? cout : cin, Co = trans ? cin : cout;
  const int MT = Co / 16;
  const int v = (int)(e & 7), l = (int)((e >> 3) & 63);
  const int64_t tile = e >> 9;
  const int mt = (int)(tile % MT), kc = (int)(tile / MT);
  const int kb = 32 * kc + 8 * (l >> 4);
  int tap, ci;
  if (Ci == 1) {
    tap = kb + v;
    ci = 0;
  } else {
    tap = kb / Ci;
    ci = kb % Ci + v;
  }
  const int co = 16 * mt + (l & 15);
  float w = 0.f;
  if (tap < taps) {
    if (trans) {
      const int t0 = taps == 27 ? 26 - tap : 0;
      w = W[((int64_t)t0 * cin + co) * cout + ci];
    } else {
      w = W[((int64_t)tap * cin + ci) * cout + co];
    }
  }
  dst[e] = (__bf16)w;
}

// ---- input layer: h0 = relu(conv3(r; c0)), r FP32 colour-split, K = 27 taps padded to 32 ------
template <int NT>
__global__ __launch_bounds__(kTpb) void k_pc_input(DevTopo T, const float* __restrict__ r,
                                                   const bf16x8* __restrict__ w, __bf16* __restrict__ h0) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, j = lane & 15;
  const int32_t ncell = T.L * kLeafVox;
  const int32_t half = T.L * kHalfLeaf;
  const int32_t cell0 = ((int32_t)blockIdx.x * 4 + wave) * (16 * NT);
  if (cell0 >= ncell) return;
  const bf16x8 a = w[lane];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const CellPos P = cell_pos(cell0 + 16 * t + j, ncell, 3);
    bf16x8 b = zero8();
    if (P.ok) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int tap = 8 * g + v;
        if (tap < 27) {
          const int32_t nc = nbr_cell(T, 3, P.leaf, P.x, P.y, P.z, tap / 9 - 1, (tap / 3) % 3 - 1, tap % 3 - 1);
          if (nc >= 0) {
            const int32_t nl = nc >> 9, x = (nc >> 6) & 7, y = (nc >> 3) & 7, z = nc & 7;
            b[v] = (__bf16)r[((x + y + z) & 1) * half + nl * kHalfLeaf + cs_local(x, y, z)];
          }
        }
      }
    }
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(a, b, acc);
    if (P.ok) {
      bf16x4 o;
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = (__bf16)fmaxf(acc[v], 0.f);
      *reinterpret_cast<bf16x4*>(h0 + (int64_t)P.c * 16 + 4 * g) = o;
    }
  }
}

// ---- generic sparse conv (3x3x3 or 1x1) on a level -------------------------------------------
// ACT 0: none; 1: relu; 2: keep where mask > 0 (relu backward); 3: relu, then the fused 16 -> 1
// output projection into z (FP32 colour-split) with the relu'd values optionally kept in y.
template <int TAPS, int CIN, int COUT, int ACT, int NT>
__global__ __launch_bounds__(kTpb) void k_pc_conv(DevTopo T, int lgS, const __bf16* __restrict__ x,
                                                  const bf16x8* __restrict__ w, __bf16* __restrict__ y,
                                                  const __bf16* __restrict__ mask, const float* __restrict__ wout,
                                                  float* __restrict__ z) {
  constexpr int NK = (TAPS * CIN + 31) / 32, MT = COUT / 16;
  static_assert(CIN % 8 == 0 && COUT % 16 == 0, "channel counts");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, j = lane & 15;
  const int32_t ncell = T.L << (3 * lgS);
  const int32_t cell0 = ((int32_t)blockIdx.x * 4 + wave) * (16 * NT);
  if (cell0 >= ncell) return;  // wave-uniform; the kernel has no barrier
  CellPos P[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) P[t] = cell_pos(cell0 + 16 * t + j, ncell, lgS);
  floatx4 acc[NT][MT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[t][m] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < NK; ++kc) {
    const int kb = 32 * kc + 8 * g;
    const int tap = kb / CIN, ci0 = kb % CIN;
    bf16x8 b[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      b[t] = zero8();
      if (P[t].ok && tap < TAPS) {
        const int32_t nc = TAPS == 1 ? P[t].c
                                     : nbr_cell(T, lgS, P[t].leaf, P[t].x, P[t].y, P[t].z, tap / 9 - 1,
                                                (tap / 3) % 3 - 1, tap % 3 - 1);
        if (nc >= 0) b[t] = *reinterpret_cast<const bf16x8*>(x + (int64_t)nc * CIN + ci0);
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const bf16x8 a = w[(kc * MT + m) * 64 + lane];
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t][m] = mfma(a, b[t], acc[t][m]);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (ACT == 3) {
      // relu, then z = sum_co w_out[co] h[co]: 4 channels per lane, summed over the 4 lane groups
      float part = 0.f;
      bf16x4 o;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        o[v] = (__bf16)fmaxf(acc[t][0][v], 0.f);
        part += (float)o[v] * (float)(__bf16)wout[4 * g + v];
      }
      if (y && P[t].ok) *reinterpret_cast<bf16x4*>(y + (int64_t)P[t].c * 16 + 4 * g) = o;
      part += __shfl_xor(part, 16, 64);
      part += __shfl_xor(part, 32, 64);
      if (g == 0 && P[t].ok) {
        const int32_t half = T.L * kHalfLeaf;
        z[((P[t].x + P[t].y + P[t].z) & 1) * half + P[t].leaf * kHalfLeaf + cs_local(P[t].x, P[t].y, P[t].z)] = part;
      }
      continue;
    }
    if (!P[t].ok) continue;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int64_t o0 = (int64_t)P[t].c * COUT + 16 * m + 4 * g;
      bf16x4 o;
      if (ACT == 2) {
        const bf16x4 mk = *reinterpret_cast<const bf16x4*>(mask + o0);
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = (__bf16)((float)mk[v] > 0.f ? acc[t][m][v] : 0.f);
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = (__bf16)(ACT == 1 ? fmaxf(acc[t][m][v], 0.f) : acc[t][m][v]);
      }
      *reinterpret_cast<bf16x4*>(y + o0) = o;
    }
  }
}

// ---- fused level-0 inference kernels: one workgroup per leaf, a 10^3 tile (leaf + 1-voxel halo)
// staged in LDS once instead of 27 gathers per voxel from the cache hierarchy -------------------
constexpr int kT0 = 10, kT03 = kT0 * kT0 * kT0;
__device__ __forceinline__ int tile0(int x, int y, int z) { return (x * kT0 + y) * kT0 + z; }
// tile cell c (coordinates t-1 in {-1..8}^3 relative to the leaf) -> its leaf (or -1) and local offset
__device__ __forceinline__ int32_t tile_leaf(const int32_t* s_nb, int c, int& lx, int& ly, int& lz) {
  const int tx = c / (kT0 * kT0), ty = (c / kT0) % kT0, tz = c % kT0;
  const int x = tx - 1, y = ty - 1, z = tz - 1;
  const int ox = x < 0 ? -1 : (x > 7 ? 1 : 0), oy = y < 0 ? -1 : (y > 7 ? 1 : 0), oz = z < 0 ? -1 : (z > 7 ? 1 : 0);
  lx = x & 7;
  ly = y & 7;
  lz = z & 7;
  return s_nb[(ox + 1) * 9 + (oy + 1) * 3 + (oz + 1)];
}

// h0 = relu(conv3(r; c0)) written for the leaf, and p1 = pool(h0) from the same workgroup.
__device__ __attribute__((aligned(16))) __bf16 g_pc_zero[8];  // zero-initialised
typedef __attribute__((address_space(3))) void lds_void;
// 16 (4) bytes per lane from global memory straight into LDS at dst + 16 (4) * lane; dst is
// wave-uniform. The loads count against vmcnt; an LDS access issued behind one waits for it.
__device__ __forceinline__ void lds_load16(const void* src, void* dst) {
  __builtin_amdgcn_global_load_lds(src, (lds_void*)dst, 16, 0, 0);
}
__device__ __forceinline__ void lds_load4(const void* src, void* dst) {
  __builtin_amdgcn_global_load_lds(src, (lds_void*)dst, 4, 0, 0);
}
// Same arithmetic as k_pc_input followed by k_pc_pool (same MFMA K order, same child order).
__global__ __launch_bounds__(kTpb) void k_pc_input_pool(DevTopo T, const float* __restrict__ r,
                                                        const bf16x8* __restrict__ w, __bf16* __restrict__ h0,
                                                        __bf16* __restrict__ p1) {
  __shared__ int32_t s_nb[2][27];
  __shared__ __attribute__((aligned(16))) float s_r[kT03 + 24];  // 1000 values, padded to whole waves
  __shared__ __attribute__((aligned(16))) __bf16 s_h[kLeafVox * 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, j = lane & 15;
  const LeafRun R = leaf_run(T.L);
  if (R.first >= R.end) return;  // uniform per workgroup
  const int32_t half = T.L * kHalfLeaf;
  // r tile cells e = tid + 256 it: neighbour slot (bits 0-4), colour (5), colour-split local offset (6-13)
  constexpr int kIt = (kT03 + kTpb - 1) / kTpb;
  int desc[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int c = min(tid + kTpb * it, kT03 - 1);
    const int x = c / 100 - 1, y = (c / 10) % 10 - 1, z = c % 10 - 1;
    const int slot = ((x >> 3) + 1) * 9 + ((y >> 3) + 1) * 3 + ((z >> 3) + 1);
    const int lx = x & 7, ly = y & 7, lz = z & 7;
    desc[it] = slot | (((lx + ly + lz) & 1) << 5) | (cs_local(lx, ly, lz) << 6);
  }
  const bf16x8 a = w[lane];
  // taps 8 g + v of this lane: tile-cell offsets relative to the output cell's (0,0,0) neighbour
  int off[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int tap = min(8 * g + v, 26);
    off[v] = tile0(tap / 9, (tap / 3) % 3, tap % 3);
  }
  const int nvalid = min(8, 27 - 8 * g);  // taps beyond 26 are padding (zero)
  if (tid < 27) s_nb[0][tid] = T.nbr[kNbrStride * R.first + tid];
  int nb = 0;
  for (int32_t l = R.first; l < R.end; l += R.step, nb ^= 1) {
    const int32_t ln = l + R.step;
    __syncthreads();  // neighbour row visible; the previous leaf's tiles are no longer read
    const float* src[kIt];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      int d = desc[it];
      opaque_v(d);
      const int32_t nl = s_nb[nb][d & 31];
      src[it] = nl < 0 ? reinterpret_cast<const float*>(g_pc_zero)
                       : r + ((d >> 5) & 1) * half + nl * kHalfLeaf + ((d >> 6) & 255);
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it)
      if (kTpb * it + 64 * wave < kT03) lds_load4(src[it], s_r + kTpb * it + 64 * wave);
    if (tid < 27 && ln < R.end) s_nb[nb ^ 1][tid] = T.nbr[kNbrStride * ln + tid];
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int tile = wave; tile < kLeafVox / 16; tile += 4) {
      const int cell = 16 * tile + j;
      const int base = tile0(cell >> 6, (cell >> 3) & 7, cell & 7);
      bf16x8 b;
#pragma unroll
      for (int v = 0; v < 8; ++v) b[v] = v < nvalid ? (__bf16)s_r[base + off[v]] : (__bf16)0.0f;
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(a, b, acc);
      bf16x4 o;
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = (__bf16)fmaxf(acc[v], 0.f);
      *reinterpret_cast<bf16x4*>(h0 + ((int64_t)l * kLeafVox + cell) * 16 + 4 * g) = o;
      *reinterpret_cast<bf16x4*>(s_h + cell * 16 + 4 * g) = o;
    }
    __syncthreads();
    // p1: 64 coarse cells x 4 groups of 4 channels, one per lane, children in k_pc_pool's order
    {
      const int cc = tid >> 2, q4 = (tid & 3) * 4;
      const int cx = cc >> 4, cy = (cc >> 2) & 3, cz = cc & 3;
      float sum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int fc = ((2 * cx + (k >> 2)) << 6) | ((2 * cy + ((k >> 1) & 1)) << 3) | (2 * cz + (k & 1));
        const bf16x4 v = *reinterpret_cast<const bf16x4*>(s_h + fc * 16 + q4);
#pragma unroll
        for (int q = 0; q < 4; ++q) sum[q] += (float)v[q];
      }
      bf16x4 o;
#pragma unroll
      for (int q = 0; q < 4; ++q) o[q] = (__bf16)(sum[q] * 0.125f);
      *reinterpret_cast<bf16x4*>(p1 + ((int64_t)l * 64 + cc) * 16 + q4) = o;
    }
  }
}

// z = relu(conv3(s0; c6)) . w_out with s0 = h0 + up(g1) formed in LDS over the leaf and its halo.
// Same arithmetic as k_pc_upadd at level 0 followed by the ACT 3 output conv.
//
// Persistent workgroups, four per compute unit, walk a strided run of leaves with one LDS buffer
// each. A leaf's h0 halo tile and g1 parent tile arrive by direct global-to-LDS loads (no
// registers held), s0 = h0 + up(g1) is formed in place, then the MFMAs run; while one workgroup
// waits for its loads the others on the compute unit keep the matrix cores busy. The neighbour row
// of the next leaf is staged in LDS one leaf ahead, so no tile load waits on a lookup. Absent
// neighbour leaves load a zero chunk.
constexpr int kOutTileB = kT03 * 32;                     // s0 tile: 1000 cells x 16 bf16
constexpr int kOutParB = 216 * 32;                       // g1 parent tile: 6^3 level-1 cells x 16 bf16
constexpr int kOutBufB = kOutTileB + kOutParB;
constexpr int kOutChunks = kT03 * 2, kOutParChunks = 216 * 2;
constexpr int kOutIt = (kOutChunks + kTpb - 1) / kTpb;   // 8 tile chunks per lane
constexpr int kOutParIt = (kOutParChunks + kTpb - 1) / kTpb;  // 2 parent chunks per lane
// element offset (bf16) of tap t's cell row relative to the output cell's row in the 10^3 tile
constexpr int out_tap_off(int t) { return t < 27 ? ((t / 9) * 100 + ((t / 3) % 3) * 10 + t % 3) * 16 : 0; }

__global__ __launch_bounds__(kTpb, 4) void k_pc_output_fused(DevTopo T, const __bf16* __restrict__ h0,
                                                             const __bf16* __restrict__ g1,
                                                             const bf16x8* __restrict__ w,
                                                             const float* __restrict__ wout, float* __restrict__ z) {
  __shared__ __attribute__((aligned(16))) unsigned char s_buf[kOutBufB];
  __shared__ int32_t s_nb[2][27];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, j = lane & 15;
  const LeafRun R = leaf_run(T.L);
  if (R.first >= R.end) return;  // uniform per workgroup
  constexpr int NK = (27 * 16 + 31) / 32;
  bf16x8 a[NK];
#pragma unroll
  for (int kc = 0; kc < NK; ++kc) a[kc] = w[kc * 64 + lane];
  const bool hi = g >= 2;                             // lanes of the odd tap of each k-step
  const bool pad = hi && 2 * (NK - 1) + 1 >= 27;      // the last step's odd tap is padding
  // lane j's cell in tile t is (t, 2 wave + (j >> 3), j & 7); its tile row is rowj + 1600 t
  const int rowj = tile0(0, 2 * wave + (j >> 3), j & 7) * 16 + 8 * (g & 1);
  // the odd tap of a k-step lies 16, 128 or 1248 elements after the even one
  const int rb16 = rowj + (hi ? 16 : 0), rb128 = rowj + (hi ? 128 : 0), rb1248 = rowj + (hi ? 1248 : 0);
  float wo[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) wo[v] = (float)(__bf16)wout[4 * g + v];
  const int32_t half = T.L * kHalfLeaf;
  __bf16* S = reinterpret_cast<__bf16*>(s_buf);
  const __bf16* P = reinterpret_cast<const __bf16*>(s_buf + kOutTileB);
  // Chunk descriptors, the same for every leaf: neighbour slot (bits 0-4), local voxel or level-1
  // cell (bits 5-13), parent cell in the parent tile (bits 14-21, tile chunks only), data half
  // (bit 22).
  int desc[kOutIt + kOutParIt];
#pragma unroll
  for (int it = 0; it < kOutIt; ++it) {
    const int e = min(tid + kTpb * it, kOutChunks - 1), c = e >> 1;
    const int x = c / 100 - 1, y = (c / 10) % 10 - 1, zz = c % 10 - 1;
    const int slot = ((x >> 3) + 1) * 9 + ((y >> 3) + 1) * 3 + ((zz >> 3) + 1);
    const int loc = ((x & 7) << 6) | ((y & 7) << 3) | (zz & 7);
    const int par = (((x >> 1) + 1) * 6 + ((y >> 1) + 1)) * 6 + ((zz >> 1) + 1);
    desc[it] = slot | (loc << 5) | (par << 14) | ((e & 1) << 22);
  }
#pragma unroll
  for (int it = 0; it < kOutParIt; ++it) {
    const int e = min(tid + kTpb * it, kOutParChunks - 1), pc = e >> 1;
    const int x = pc / 36 - 1, y = (pc / 6) % 6 - 1, zz = pc % 6 - 1;
    const int slot = ((x >> 2) + 1) * 9 + ((y >> 2) + 1) * 3 + ((zz >> 2) + 1);
    const int loc = ((x & 3) << 4) | ((y & 3) << 2) | (zz & 3);
    desc[kOutIt + it] = slot | (loc << 5) | ((e & 1) << 22);
  }
  if (tid < 27) s_nb[0][tid] = T.nbr[kNbrStride * R.first + tid];
  int nb = 0;
  for (int32_t l = R.first; l < R.end; l += R.step, nb ^= 1) {
    const int32_t ln = l + R.step;
    __syncthreads();  // this leaf's neighbour row is visible; the previous leaf's MFMAs are done
    // Every source address is formed (reading the neighbour row from LDS) before the first tile
    // load is issued: an LDS read issued behind an LDS-bound load waits for that load.
    const __bf16* src[kOutIt + kOutParIt];
#pragma unroll
    for (int it = 0; it < kOutIt + kOutParIt; ++it) {
      int d = desc[it];
      opaque_v(d);  // keep derived values per leaf, not in registers across the loop
      const int32_t nl = s_nb[nb][d & 31];
      const int loc = (d >> 5) & 511, hf = (d >> 22) & 1;
      src[it] = nl < 0 ? g_pc_zero
                       : (it < kOutIt ? h0 + ((int64_t)nl * kLeafVox + loc) * 16 : g1 + ((int64_t)nl * 64 + loc) * 16) + 8 * hf;
    }
    const int tv = tid;
#pragma unroll
    for (int it = 0; it < kOutIt; ++it)
      if (tv + kTpb * it < kOutChunks)
        lds_load16(src[it], s_buf + (kTpb * it + 64 * wave) * 16);
#pragma unroll
    for (int it = 0; it < kOutParIt; ++it)
      if (tv + kTpb * it < kOutParChunks)
        lds_load16(src[kOutIt + it], s_buf + kOutTileB + (kTpb * it + 64 * wave) * 16);
    if (tid < 27 && ln < R.end) s_nb[nb ^ 1][tid] = T.nbr[kNbrStride * ln + tid];
    __builtin_amdgcn_s_waitcnt(0);  // this lane's tile loads have landed
    __syncthreads();                // and everyone's
    // s0 = h0 + up(g1) in place
#pragma unroll 2
    for (int it = 0; it < kOutIt; ++it) {
      const int e = tv + kTpb * it;
      if (e < kOutChunks) {
        const int p = (desc[it] >> 14) & 255;
        const bf16x8 va = *reinterpret_cast<const bf16x8*>(S + e * 8);
        const bf16x8 vb = *reinterpret_cast<const bf16x8*>(P + p * 16 + 8 * (e & 1));
        bf16x8 o;
#pragma unroll
        for (int q = 0; q < 8; ++q) o[q] = (__bf16)((float)va[q] + 1.0f * (float)vb[q]);
        *reinterpret_cast<bf16x8*>(S + e * 8) = o;
      }
    }
    __syncthreads();
    // Tiles go in pairs with two independent accumulators; the 8 x NK (tile, k-step) MFMAs run as
    // one unrolled sequence with a ring of kRing B operands read ahead of their MFMA. Step st:
    // pair st / (2 NK), k-step (st / 2) % NK, tile 2 pair + (st & 1).
    constexpr int kRing = 4, kSteps = 8 * NK;
    auto bsrc = [&](int st) {
      const int t = 2 * (st / (2 * NK)) + (st & 1), kc = (st >> 1) % NK;
      const int d = out_tap_off(2 * kc + 1) - out_tap_off(2 * kc);  // 16, 128, 1248 (or padding)
      const int rb = d == 16 ? rb16 : (d == 128 ? rb128 : (d == 1248 ? rb1248 : rowj));
      return reinterpret_cast<const bf16x8*>(S + rb + out_tap_off(2 * kc) + 1600 * t);
    };
    bf16x8 ring[kRing];
#pragma unroll
    for (int st = 0; st < kRing; ++st) ring[st] = *bsrc(st);
    floatx4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int st = 0; st < kSteps; ++st) {
      const int u = st & 1, kc = (st >> 1) % NK, t = 2 * (st / (2 * NK)) + u;
      const bf16x8 bv = ring[st % kRing];
      if (st + kRing < kSteps) ring[st % kRing] = *bsrc(st + kRing);
      acc[u] = mfma(a[kc], (kc == NK - 1 && pad) ? zero8() : bv, acc[u]);
      if (kc == NK - 1) {
        const int y = 2 * wave + (j >> 3), zz = j & 7;
        float part = 0.f;
#pragma unroll
        for (int v = 0; v < 4; ++v) part += (float)(__bf16)fmaxf(acc[u][v], 0.f) * wo[v];
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (g == 0) z[((t + y + zz) & 1) * half + l * kHalfLeaf + cs_local(t, y, zz)] = part;
        acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
      }
    }
  }
}

// ---- fused level-1 and level-2 inference kernels ----------------------------------------------
// Each stages a halo tile of its input in LDS (absent leaves as 0, rows padded by 16 bytes so
// consecutive rows start in different banks) and runs the same MFMA sequence per cell as
// k_pc_conv (same K order), so results are bitwise those of the unfused kernels.

// Leaf holding the tile cell (tx,ty,tz) of a tile of edge S+2 around a leaf of edge S = 1 << lgS
// (tile coordinate t is level coordinate t-1), or -1, with the local coordinates in that leaf.
__device__ __forceinline__ int32_t halo_leaf(const int32_t* __restrict__ nrow, int lgS, int tx, int ty, int tz,
                                             int& lx, int& ly, int& lz) {
  const int x = tx - 1, y = ty - 1, z = tz - 1, m = (1 << lgS) - 1;
  lx = x & m;
  ly = y & m;
  lz = z & m;
  return nrow[((x >> lgS) + 1) * 9 + ((y >> lgS) + 1) * 3 + ((z >> lgS) + 1)];
}

// Level 1 (4^3 cells per leaf): one wave per leaf. Persistent workgroups walk a strided run of
// groups of four leaves, one per wave; each wave's 6^3 tile (and, for c5, the 4^3 tile of g2
// parents) arrives by direct global-to-LDS loads, every address formed before the first load is
// issued, and the wave computes all 4 cell tiles x both output-channel tiles, so every B operand
// is read once from LDS and every A operand once per leaf.
//   MODE 0 (c1):      tile = p1 (16 ch); h1 = relu(conv) to ya; p2 = pool(h1) to yb (level 2)
//   MODE 1 (c5, u1):  tile = h1 + up(g2) (xa = h1, xb = g2); h5 = relu(conv) kept in LDS;
//                     g1 = u1(h5) to yb (16 ch)
constexpr int kT1 = 6, kT13 = kT1 * kT1 * kT1;
// tile-cell offset of tap t (0..26; offsets (t/9, t/3%3, t%3) - 1) in a 6^3 tile; 0 for padding
constexpr int l1_tap(int t) { return t < 27 ? ((t / 9) * kT1 + (t / 3) % 3) * kT1 + t % 3 : 0; }
template <int MODE>
struct L1Shape {
  static constexpr int CIN = MODE == 0 ? 16 : 32, CH = CIN / 8;   // 16-byte chunks per cell
  static constexpr int NCH = kT13 * CH, IT = (NCH + 63) / 64;     // tile chunks, per lane
  static constexpr int NPCH = MODE == 0 ? 0 : 64 * CH, PIT = NPCH / 64;  // parent chunks, per lane
  static constexpr int TILE = kT13 * CIN, WAVE = TILE + 64 * CIN * (MODE == 0 ? 0 : 1);  // bf16 per wave
};
template <int MODE>
__global__ __launch_bounds__(kTpb, 2) void k_pc_level1(DevTopo T, const __bf16* __restrict__ xa,
                                                    const __bf16* __restrict__ xb, const bf16x8* __restrict__ w,
                                                    const bf16x8* __restrict__ w2, __bf16* __restrict__ ya,
                                                    __bf16* __restrict__ yb) {
  using Sh = L1Shape<MODE>;
  constexpr int CIN = Sh::CIN, CH = Sh::CH, MT = 2, NK = (27 * CIN + 31) / 32;
  __shared__ __attribute__((aligned(16))) __bf16 s_t[4][Sh::WAVE];
  __shared__ int32_t s_nb[2][4][27];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, j = lane & 15;
  const int32_t ngroups = (T.L + 3) / 4;
  const LeafRun R = leaf_run(ngroups);
  if (R.first >= R.end) return;  // uniform per workgroup
  __bf16* S = s_t[wave];
  const __bf16* P = s_t[wave] + Sh::TILE;
  // chunk descriptors (the same for every leaf): slot (bits 0-4), local level-1 cell (5-10) or
  // level-2 cell (parents), parent tile cell (11-16), chunk part (17-18)
  int desc[Sh::IT + Sh::PIT];
#pragma unroll
  for (int it = 0; it < Sh::IT; ++it) {
    const int e = min(lane + 64 * it, Sh::NCH - 1), c = e / CH;
    const int x = c / 36 - 1, y = (c / 6) % 6 - 1, z = c % 6 - 1;
    const int slot = ((x >> 2) + 1) * 9 + ((y >> 2) + 1) * 3 + ((z >> 2) + 1);
    const int loc = ((x & 3) << 4) | ((y & 3) << 2) | (z & 3);
    const int par = (((x >> 1) + 1) * 4 + ((y >> 1) + 1)) * 4 + ((z >> 1) + 1);
    desc[it] = slot | (loc << 5) | (par << 11) | ((e % CH) << 17);
  }
#pragma unroll
  for (int it = 0; it < Sh::PIT; ++it) {
    const int e = lane + 64 * it, pc = e / CH;
    const int x = (pc >> 4) - 1, y = ((pc >> 2) & 3) - 1, z = (pc & 3) - 1;
    const int slot = ((x >> 1) + 1) * 9 + ((y >> 1) + 1) * 3 + ((z >> 1) + 1);
    const int loc = ((x & 1) << 2) | ((y & 1) << 1) | (z & 1);
    desc[Sh::IT + it] = slot | (loc << 5) | ((e % CH) << 17);
  }
  const int rowc = ((j >> 2) * kT1 + (j & 3)) * CIN;  // cell 16 t + j = (t, j >> 2, j & 3)
  if (lane < 27 && R.first * 4 + wave < T.L) s_nb[0][wave][lane] = T.nbr[kNbrStride * (R.first * 4 + wave) + lane];
  int nb = 0;
  for (int32_t gi = R.first; gi < R.end; gi += R.step, nb ^= 1) {
    const int32_t l = gi * 4 + wave, gn = gi + R.step;
    const bool live = l < T.L;
    __syncthreads();  // neighbour rows visible; the previous group's outputs are read
    if (live) {
      const __bf16* src[Sh::IT + Sh::PIT];
#pragma unroll
      for (int it = 0; it < Sh::IT + Sh::PIT; ++it) {
        int d = desc[it];
        opaque_v(d);  // derived values per leaf, not held across the loop
        const int32_t nl = s_nb[nb][wave][d & 31];
        const int loc = (d >> 5) & 63, part = (d >> 17) & 3;
        src[it] = nl < 0 ? g_pc_zero
                         : (it < Sh::IT ? xa + ((int64_t)nl * 64 + loc) * CIN : xb + ((int64_t)nl * 8 + loc) * CIN) +
                               8 * part;
      }
#pragma unroll
      for (int it = 0; it < Sh::IT; ++it)
        if (lane + 64 * it < Sh::NCH)
          lds_load16(src[it], S + 512 * it);
#pragma unroll
      for (int it = 0; it < Sh::PIT; ++it)
        lds_load16(src[Sh::IT + it], S + Sh::TILE + 512 * it);
    }
    if (lane < 27 && gn < R.end && gn * 4 + wave < T.L)
      s_nb[nb ^ 1][wave][lane] = T.nbr[kNbrStride * (gn * 4 + wave) + lane];
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (MODE == 1 && live) {
      // s1 = h1 + up(g2) in place
#pragma unroll 2
      for (int it = 0; it < Sh::IT; ++it) {
        const int e = lane + 64 * it;
        if (e < Sh::NCH) {
          const int par = (desc[it] >> 11) & 63;
          const bf16x8 va = *reinterpret_cast<const bf16x8*>(S + e * 8);
          const bf16x8 vb = *reinterpret_cast<const bf16x8*>(P + par * CIN + 8 * (e % CH));
          bf16x8 o;
#pragma unroll
          for (int q = 0; q < 8; ++q) o[q] = (__bf16)((float)va[q] + 1.0f * (float)vb[q]);
          *reinterpret_cast<bf16x8*>(S + e * 8) = o;
        }
      }
    }
    if (MODE == 1) __syncthreads();
    floatx4 acc[4][MT];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[t][m] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (live) {
      const bf16x8* wl = w;
      opaque_s(wl);  // weights are re-read from the cache per leaf, not held in registers
#pragma unroll
      for (int kc = 0; kc < NK; ++kc) {
        // CIN 32: k-step kc is tap kc, channels 8 g..; CIN 16: taps 2 kc (g < 2) and 2 kc + 1 (g >= 2)
        const int tlo = CIN == 32 ? kc : 2 * kc, thi = CIN == 32 ? kc : 2 * kc + 1;
        const int off = (g < 2 ? l1_tap(tlo) : l1_tap(thi)) * CIN + (CIN == 32 ? 8 * g : 8 * (g & 1));
        const bool padk = thi >= 27 && g >= 2;
        bf16x8 a[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) a[m] = wl[(kc * MT + m) * 64 + lane];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          bf16x8 b = *reinterpret_cast<const bf16x8*>(S + rowc + 36 * CIN * t + (padk ? 0 : off));
          if (padk) b = zero8();
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[t][m] = mfma(a[m], b, acc[t][m]);
        }
      }
    }
    __syncthreads();  // the tile region now holds the 64 x 32 relu'd outputs (cell * 32 + channel)
    if (live) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          bf16x4 o;
#pragma unroll
          for (int v = 0; v < 4; ++v) o[v] = (__bf16)fmaxf(acc[t][m][v], 0.f);
          const int cell = 16 * t + j;
          if (MODE == 0) *reinterpret_cast<bf16x4*>(ya + ((int64_t)l * 64 + cell) * 32 + 16 * m + 4 * g) = o;
          *reinterpret_cast<bf16x4*>(S + cell * 32 + 16 * m + 4 * g) = o;
        }
    }
    __syncthreads();
    if (!live) continue;
    if (MODE == 0) {
      // p2: 8 coarse cells x 8 groups of 4 channels, children in k_pc_pool's order
      const int cc = lane >> 3, q4 = (lane & 7) * 4;
      const int cx = cc >> 2, ccy = (cc >> 1) & 1, ccz = cc & 1;
      float sum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int fc = ((2 * cx + (k >> 2)) << 4) | ((2 * ccy + ((k >> 1) & 1)) << 2) | (2 * ccz + (k & 1));
        const bf16x4 v = *reinterpret_cast<const bf16x4*>(S + fc * 32 + q4);
#pragma unroll
        for (int q = 0; q < 4; ++q) sum[q] += (float)v[q];
      }
      bf16x4 o;
#pragma unroll
      for (int q = 0; q < 4; ++q) o[q] = (__bf16)(sum[q] * 0.125f);
      *reinterpret_cast<bf16x4*>(yb + ((int64_t)l * 8 + cc) * 32 + q4) = o;
    } else {
      // g1 = u1(h5): K = 32 (one MFMA step), 16 output channels
      const bf16x8 a = w2[lane];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(S + (16 * t + j) * 32 + 8 * g);
        const floatx4 r = mfma(a, b, floatx4{0.f, 0.f, 0.f, 0.f});
        bf16x4 o;
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = (__bf16)r[v];
        *reinterpret_cast<bf16x4*>(yb + ((int64_t)l * 64 + 16 * t + j) * 16 + 4 * g) = o;
      }
    }
  }
}

// Level 2 (2^3 cells per leaf): persistent workgroups walk a strided run of groups of 8
// consecutive leaves; a group's 4^3 tiles arrive by direct global-to-LDS loads (addresses formed
// before the first load is issued); wave w computes output-channel tile w for all 64 cells
// (4 cell tiles of 2 leaves each), so the weights of a layer are read once per 8 leaves.
//   MODE 0 (c2):      tile = p2 (32 ch); h2 = relu(conv) to y (64 ch)
//   MODE 1 (c3, u2):  tile = h2 (64 ch); h3 = relu(conv) kept in LDS; g2 = u2(h3) to y (32 ch)
constexpr int kT2 = 4, kT23 = kT2 * kT2 * kT2;
template <int MODE>
__global__ __launch_bounds__(kTpb, 2) void k_pc_level2(DevTopo T, const __bf16* __restrict__ x,
                                                       const bf16x8* __restrict__ w, const bf16x8* __restrict__ w2,
                                                       __bf16* __restrict__ y) {
  constexpr int CIN = MODE == 0 ? 32 : 64, CH = CIN / 8, MT = 4, NK = (27 * CIN + 31) / 32, R = CIN;
  constexpr int NCH = 8 * kT23 * CH, IT = NCH / kTpb;  // 16-byte chunks per group, per lane
  __shared__ __attribute__((aligned(16))) __bf16 S[8 * kT23 * R];
  __shared__ int32_t s_nb[2][8 * 27];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, j = lane & 15;
  const int32_t ngroups = (T.L + 7) / 8;
  const LeafRun RG = leaf_run(ngroups);
  if (RG.first >= RG.end) return;  // uniform per workgroup
  // chunk e = tid + 256 it: leaf slot (bits 0-2), neighbour slot (3-7), local cell (8-10), part (11-13)
  int desc[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + kTpb * it;
    const int ls = e / (kT23 * CH), rem = e % (kT23 * CH), c = rem / CH, part = rem % CH;
    const int xx = (c >> 4) - 1, yy = ((c >> 2) & 3) - 1, zz = (c & 3) - 1;
    const int slot = ((xx >> 1) + 1) * 9 + ((yy >> 1) + 1) * 3 + ((zz >> 1) + 1);
    const int loc = ((xx & 1) << 2) | ((yy & 1) << 1) | (zz & 1);
    desc[it] = ls | (slot << 3) | (loc << 8) | (part << 11);
  }
  auto load_rows = [&](int32_t gi, int b) {
    if (tid < 8 * 27) {
      const int32_t l = gi * 8 + tid / 27;
      s_nb[b][tid] = l < T.L ? T.nbr[kNbrStride * l + tid % 27] : -1;  // slots past the last leaf stage zeros
    }
  };
  load_rows(RG.first, 0);
  int nb = 0;
  for (int32_t gi = RG.first; gi < RG.end; gi += RG.step, nb ^= 1) {
    const int32_t l0 = gi * 8, gn = gi + RG.step;
    const int nlive = min(8, T.L - l0);
    __syncthreads();  // neighbour rows visible; the previous group's tiles are no longer read
    {
      const __bf16* src[IT];
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        int d = desc[it];
        opaque_v(d);
        const int32_t nl = s_nb[nb][(d & 7) * 27 + ((d >> 3) & 31)];
        src[it] = nl < 0 ? g_pc_zero : x + ((int64_t)nl * 8 + ((d >> 8) & 7)) * CIN + 8 * ((d >> 11) & 7);
      }
#pragma unroll
      for (int it = 0; it < IT; ++it) lds_load16(src[it], S + (kTpb * it + 64 * wave) * 8);
    }
    if (gn < RG.end) load_rows(gn, nb ^ 1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // lane j of cell tile t: leaf slot 2 t + (j >> 3), local cell j & 7
    const int sj = j >> 3, cj = j & 7, cx = cj >> 2, cy = (cj >> 1) & 1, cz = cj & 1;
    const int m = wave;
    floatx4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    // lane's tile row for cell tile t: rowc + 2 kT23 t; a k-step covers one tap (CIN 32) or half a tap (CIN 64)
    const int rowc = (sj * kT23 + (cx * kT2 + cy) * kT2 + cz) * R + 8 * g;
    const bf16x8* wl = w;
    opaque_s(wl);  // weights are re-read from the cache per group, not held in registers
#pragma unroll 6
    for (int kc = 0; kc < NK; ++kc) {
      const int tap = CIN == 32 ? kc : kc >> 1, ci0 = CIN == 32 ? 0 : 32 * (kc & 1);
      const int off = ((tap / 9 * kT2 + (tap / 3) % 3) * kT2 + tap % 3) * R + ci0;
      const bf16x8 a = wl[(kc * MT + m) * 64 + lane];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(S + rowc + 2 * kT23 * R * t + off);
        acc[t] = mfma(a, b, acc[t]);
      }
    }
    if (MODE == 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (2 * t + sj >= nlive) continue;
        bf16x4 o;
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = (__bf16)fmaxf(acc[t][v], 0.f);
        *reinterpret_cast<bf16x4*>(y + ((int64_t)(l0 + 2 * t + sj) * 8 + cj) * 64 + 16 * m + 4 * g) = o;
      }
      continue;
    }
    __syncthreads();  // the tile region now holds h3 for the 64 cells (cell * 64 + channel)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      bf16x4 o;
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = (__bf16)fmaxf(acc[t][v], 0.f);
      *reinterpret_cast<bf16x4*>(S + ((2 * t + sj) * 8 + cj) * 64 + 16 * m + 4 * g) = o;
    }
    __syncthreads();
    // g2 = u2(h3): K = 64 (two MFMA steps), 32 output channels; wave w: output tile w & 1, cell tiles 2 (w >> 1) + {0, 1}
    const int m2 = wave & 1;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const int t = 2 * (wave >> 1) + tt;
      floatx4 r = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < 2; ++kc) {
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(S + ((2 * t + sj) * 8 + cj) * 64 + 32 * kc + 8 * g);
        r = mfma(w2[(kc * 2 + m2) * 64 + lane], b, r);
      }
      if (2 * t + sj >= nlive) continue;
      bf16x4 o;
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = (__bf16)r[v];
      *reinterpret_cast<bf16x4*>(y + ((int64_t)(l0 + 2 * t + sj) * 8 + cj) * 32 + 16 * m2 + 4 * g) = o;
    }
  }
}

// ---- pooling (children sum times scale) and upsample-add -------------------------------------
__global__ void k_pc_pool(int32_t L, int lgSc, int c8, const __bf16* __restrict__ x, __bf16* __restrict__ y,
                          float scale) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ncell = (int64_t)L << (3 * lgSc);
  if (e >= ncell * c8) return;
  const int32_t cc = (int32_t)(e / c8);
  const int part = (int)(e % c8);
  const CellPos P = cell_pos(cc, (int32_t)ncell, lgSc);
  const int lgSf = lgSc + 1, C = 8 * c8;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int32_t fc = (P.leaf << (3 * lgSf)) | ((2 * P.x + (k >> 2)) << (2 * lgSf)) |
                       ((2 * P.y + ((k >> 1) & 1)) << lgSf) | (2 * P.z + (k & 1));
    const bf16x8 v = *reinterpret_cast<const bf16x8*>(x + (int64_t)fc * C + 8 * part);
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] += (float)v[q];
  }
  bf16x8 o;
#pragma unroll
  for (int q = 0; q < 8; ++q) o[q] = (__bf16)(s[q] * scale);
  *reinterpret_cast<bf16x8*>(y + (int64_t)cc * C + 8 * part) = o;
}

template <bool MASK>
__global__ void k_pc_upadd(int32_t L, int lgSf, int c8, const __bf16* __restrict__ a, const __bf16* __restrict__ b,
                           __bf16* __restrict__ y, float scale, const __bf16* __restrict__ mask) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ncell = (int64_t)L << (3 * lgSf);
  if (e >= ncell * c8) return;
  const int32_t fc = (int32_t)(e / c8);
  const int part = (int)(e % c8);
  const CellPos P = cell_pos(fc, (int32_t)ncell, lgSf);
  const int lgSc = lgSf - 1, C = 8 * c8;
  const int32_t pc = (P.leaf << (3 * lgSc)) | ((P.x >> 1) << (2 * lgSc)) | ((P.y >> 1) << lgSc) | (P.z >> 1);
  const int64_t o0 = (int64_t)fc * C + 8 * part;
  const bf16x8 va = *reinterpret_cast<const bf16x8*>(a + o0);
  const bf16x8 vb = *reinterpret_cast<const bf16x8*>(b + (int64_t)pc * C + 8 * part);
  bf16x8 mk;
  if (MASK) mk = *reinterpret_cast<const bf16x8*>(mask + o0);
  bf16x8 o;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    float v = (float)va[q] + scale * (float)vb[q];
    if (MASK && !((float)mk[q] > 0.f)) v = 0.f;
    o[q] = (__bf16)v;
  }
  *reinterpret_cast<bf16x8*>(y + o0) = o;
}

// ---- output backward: dh6 = w_out * dz where h6 > 0 ------------------------------------------
__global__ void k_pc_output_bwd(int32_t L, const float* __restrict__ dz, const __bf16* __restrict__ h6,
                                const float* __restrict__ wout, __bf16* __restrict__ dh) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t ncell = L * kLeafVox;
  if (e >= (int64_t)ncell * 2) return;
  const int32_t c = (int32_t)(e >> 1);
  const int part = (int)(e & 1);
  const CellPos P = cell_pos(c, ncell, 3);
  const float d = dz[((P.x + P.y + P.z) & 1) * (L * kHalfLeaf) + P.leaf * kHalfLeaf + cs_local(P.x, P.y, P.z)];
  const int64_t o0 = (int64_t)c * 16 + 8 * part;
  const bf16x8 h = *reinterpret_cast<const bf16x8*>(h6 + o0);
  bf16x8 o;
#pragma unroll
  for (int q = 0; q < 8; ++q) o[q] = (__bf16)((float)h[q] > 0.f ? (float)(__bf16)wout[8 * part + q] * d : 0.f);
  *reinterpret_cast<bf16x8*>(dh + o0) = o;
}

// ---- weight gradients ------------------------------------------------------------------------
// Workgroup (tap, range): partial[range][tap][ci][co] = sum over the range's cells of
// x[nbr(c, tap)][ci] dy[c][co]. 64-cell blocks of x (gathered) and dy are staged in LDS; MFMA
// operands take 8 consecutive cells of one channel (k = cell).
constexpr int kWgBlock = 64;
template <int TAPS, int CIN, int COUT>
__global__ __launch_bounds__(kTpb) void k_pc_wgrad(DevTopo T, int lgS, int R, const __bf16* __restrict__ x,
                                                   const __bf16* __restrict__ dy, float* __restrict__ partial) {
  constexpr int XS = CIN + 8, YS = COUT + 8;  // padded LDS rows
  constexpr int NI = CIN / 16, NJ = COUT / 16, NTILE = NI * NJ, TPW = (NTILE + 3) / 4;
  __shared__ __bf16 sx[kWgBlock * XS];
  __shared__ __bf16 sy[kWgBlock * YS];
  const int tap = (int)(blockIdx.x % TAPS), rg = (int)(blockIdx.x / TAPS);
  const int dxx = TAPS == 1 ? 0 : tap / 9 - 1, dyy = TAPS == 1 ? 0 : (tap / 3) % 3 - 1,
            dzz = TAPS == 1 ? 0 : tap % 3 - 1;
  const int32_t ncell = T.L << (3 * lgS);
  const int32_t per = ((ncell + R - 1) / R + kWgBlock - 1) / kWgBlock * kWgBlock;
  const int32_t cb0 = rg * per, ce = min(ncell, cb0 + per);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, j = lane & 15;
  floatx4 acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int32_t cb = cb0; cb < ce; cb += kWgBlock) {
    for (int e = threadIdx.x; e < kWgBlock * (CIN / 8); e += kTpb) {
      const int row = e / (CIN / 8), part = e % (CIN / 8);
      const int32_t c = cb + row;
      bf16x8 v = zero8();
      if (c < ce) {
        const CellPos P = cell_pos(c, ncell, lgS);
        const int32_t nc = TAPS == 1 ? c : nbr_cell(T, lgS, P.leaf, P.x, P.y, P.z, dxx, dyy, dzz);
        if (nc >= 0) v = *reinterpret_cast<const bf16x8*>(x + (int64_t)nc * CIN + 8 * part);
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) sx[row * XS + 8 * part + q] = v[q];
    }
    for (int e = threadIdx.x; e < kWgBlock * (COUT / 8); e += kTpb) {
      const int row = e / (COUT / 8), part = e % (COUT / 8);
      const int32_t c = cb + row;
      bf16x8 v = zero8();
      if (c < ce) v = *reinterpret_cast<const bf16x8*>(dy + (int64_t)c * COUT + 8 * part);
#pragma unroll
      for (int q = 0; q < 8; ++q) sy[row * YS + 8 * part + q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kWgBlock / 32; ++kk) {
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        const int tile = wave + 4 * q;
        if (tile < NTILE) {
          const int it = tile / NJ, jt = tile % NJ;
          bf16x8 a, b;
#pragma unroll
          for (int v = 0; v < 8; ++v) {
            const int row = 32 * kk + 8 * g + v;
            a[v] = sx[row * XS + 16 * it + j];
            b[v] = sy[row * YS + 16 * jt + j];
          }
          acc[q] = mfma(a, b, acc[q]);
        }
      }
    }
    __syncthreads();
  }
  float* out = partial + ((int64_t)rg * TAPS + tap) * CIN * COUT;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int tile = wave + 4 * q;
    if (tile < NTILE) {
      const int it = tile / NJ, jt = tile % NJ;
#pragma unroll
      for (int v = 0; v < 4; ++v) out[(16 * it + 4 * g + v) * COUT + 16 * jt + j] = acc[q][v];
    }
  }
}

// Input layer (cin = 1, FP32 source rounded to bf16 as in the forward): VALU, 16 accumulators.
__global__ __launch_bounds__(kTpb) void k_pc_wgrad_input(DevTopo T, int R, const float* __restrict__ r,
                                                         const __bf16* __restrict__ dh0,
                                                         float* __restrict__ partial) {
  __shared__ float s_red[4][16];
  const int tap = (int)(blockIdx.x % 27), rg = (int)(blockIdx.x / 27);
  const int32_t ncell = T.L * kLeafVox, half = T.L * kHalfLeaf;
  const int32_t per = (ncell + R - 1) / R;
  const int32_t cb0 = rg * per, ce = min(ncell, cb0 + per);
  float acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.f;
  for (int32_t c = cb0 + (int32_t)threadIdx.x; c < ce; c += kTpb) {
    const CellPos P = cell_pos(c, ncell, 3);
    const int32_t nc = nbr_cell(T, 3, P.leaf, P.x, P.y, P.z, tap / 9 - 1, (tap / 3) % 3 - 1, tap % 3 - 1);
    if (nc < 0) continue;
    const int32_t nl = nc >> 9, x = (nc >> 6) & 7, y = (nc >> 3) & 7, z = nc & 7;
    const float rv = (float)(__bf16)r[((x + y + z) & 1) * half + nl * kHalfLeaf + cs_local(x, y, z)];
    const bf16x8 d0 = *reinterpret_cast<const bf16x8*>(dh0 + (int64_t)c * 16);
    const bf16x8 d1 = *reinterpret_cast<const bf16x8*>(dh0 + (int64_t)c * 16 + 8);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      acc[q] += rv * (float)d0[q];
      acc[8 + q] += rv * (float)d1[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    float v = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int q = threadIdx.x;
    partial[((int64_t)rg * 27 + tap) * 16 + q] = s_red[0][q] + s_red[1][q] + s_red[2][q] + s_red[3][q];
  }
}

// Output projection: partial[range][co] = sum_c h6[c][co] dz[c].
__global__ __launch_bounds__(kTpb) void k_pc_wgrad_output(DevTopo T, int R, const float* __restrict__ dz,
                                                          const __bf16* __restrict__ h6,
                                                          float* __restrict__ partial) {
  __shared__ float s_red[4][16];
  const int rg = (int)blockIdx.x;
  const int32_t ncell = T.L * kLeafVox, half = T.L * kHalfLeaf;
  const int32_t per = (ncell + R - 1) / R;
  const int32_t cb0 = rg * per, ce = min(ncell, cb0 + per);
  float acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.f;
  for (int32_t c = cb0 + (int32_t)threadIdx.x; c < ce; c += kTpb) {
    const CellPos P = cell_pos(c, ncell, 3);
    const float d = dz[((P.x + P.y + P.z) & 1) * half + P.leaf * kHalfLeaf + cs_local(P.x, P.y, P.z)];
    const bf16x8 h0 = *reinterpret_cast<const bf16x8*>(h6 + (int64_t)c * 16);
    const bf16x8 h1 = *reinterpret_cast<const bf16x8*>(h6 + (int64_t)c * 16 + 8);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      acc[q] += d * (float)h0[q];
      acc[8 + q] += d * (float)h1[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    float v = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int q = threadIdx.x;
    partial[(int64_t)rg * 16 + q] = s_red[0][q] + s_red[1][q] + s_red[2][q] + s_red[3][q];
  }
}

// grad[e] += sum_r partial[r][e], fixed order
__global__ void k_pc_reduce(int64_t n, int R, const float* __restrict__ partial, float* __restrict__ grad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int r = 0; r < R; ++r) s += partial[(int64_t)r * n + e];
  grad[e] += s;
}

__global__ void k_pc_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gv = g[e];
  const float mv = b1 * m[e] + (1.f - b1) * gv;
  const float vv = b2 * v[e] + (1.f - b2) * gv * gv;
  m[e] = mv;
  v[e] = vv;
  p[e] -= lr * (mv / c1) / (sqrtf(vv / c2) + eps);
}

// out = a*y + b*(A x) on colour-split vectors; per-workgroup partials of out.out, y.y, out.x
__device__ inline double block_sum_256(double v, double* s_red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = s_red[0] + s_red[1] + s_red[2] + s_red[3];
  __syncthreads();
  return t;
}
__global__ __launch_bounds__(kTpb) void k_pc_poisson(DevTopo T, const float* __restrict__ x,
                                                     const float* __restrict__ y, float a, float b,
                                                     float* __restrict__ out, double* __restrict__ part) {
  __shared__ double s_red[4];
  const LeafRun R = leaf_run(T.L);
  const int32_t half = T.L * kHalfLeaf;
  const int t = threadIdx.x;
  const int c0 = ((t >> 5) + (t >> 2)) & 1;
  auto get = [&](int c, int32_t i) { return x[c * half + i]; };
  double oo = 0.0, yy = 0.0, ox = 0.0;
  for (int32_t l = R.first; l < R.end; l += R.step) {
    const FaceRow Fr = face_row(T, l);
    const int32_t i0 = c0 * half + l * kHalfLeaf + t, i1 = (c0 ^ 1) * half + l * kHalfLeaf + t;
    const float x0 = x[i0], x1 = x[i1];
    float s0, s1;
    pair_face_sums(Fr, t, get, x0, x1, s0, s1);
    const float y0 = y[i0], y1 = y[i1];
    const float o0 = a * y0 + b * (6.0f * x0 - s0), o1 = a * y1 + b * (6.0f * x1 - s1);
    out[i0] = o0;
    out[i1] = o1;
    oo += (double)o0 * o0 + (double)o1 * o1;
    yy += (double)y0 * y0 + (double)y1 * y1;
    ox += (double)o0 * x0 + (double)o1 * x1;
  }
  const double to = block_sum_256(oo, s_red);
  const double ty = block_sum_256(yy, s_red);
  const double tx = block_sum_256(ox, s_red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = to;
    part[3 * blockIdx.x + 1] = ty;
    part[3 * blockIdx.x + 2] = tx;
  }
}
__global__ __launch_bounds__(kTpb) void k_pc_sum3(const double* __restrict__ part, int G, double* __restrict__ sums) {
  __shared__ double s_red[4];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < G; i += kTpb) {
    a += part[3 * i];
    b += part[3 * i + 1];
    c += part[3 * i + 2];
  }
  const double ta = block_sum_256(a, s_red);
  const double tb = block_sum_256(b, s_red);
  const double tc = block_sum_256(c, s_red);
  if (threadIdx.x == 0) {
    sums[0] = ta;
    sums[1] = tb;
    sums[2] = tc;
  }
}

inline unsigned blocks_for(int64_t n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }
// Workgroups of a persistent leaf-run kernel: `per_cu` resident workgroups on each of the
// device's compute units (a multiple of 8, the XCD count), at most one per leaf.
unsigned persistent_groups(int32_t L, int per_cu) {
  static const int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 256;
    return n;
  }();
  const unsigned want = 8u * (unsigned)((cus * per_cu + 7) / 8);
  return std::min(want, grid_for_leaves(L));
}
constexpr int kNT = 2;  // 16-cell tiles per wave

// number of cell ranges of a weight-gradient launch
inline int wgrad_ranges(int64_t ncell) {
  const int64_t r = (ncell + 2047) / 2048;
  return (int)std::max<int64_t>(1, std::min<int64_t>(64, r));
}

}  // namespace

void launch_pc_swizzle(const float* params, __bf16* wfwd, __bf16* wbwd, hipStream_t s) {
  const PcOffsets& O = pc_offsets();
  for (int i = 0; i < kPcLayers; ++i) {
    const PcLayer& Ly = kPcNet[i];
    const int64_t nf = O.fwd[i + 1] - O.fwd[i], nb = O.bwd[i + 1] - O.bwd[i];
    if (nf > 0) {
      k_pc_swizzle<<<blocks_for(nf, 256), 256, 0, s>>>(params + O.param[i], wfwd + O.fwd[i], Ly.taps, Ly.cin, Ly.cout,
                                                      0, nf);
      HNS_HIP(hipGetLastError());
    }
    if (nb > 0) {
      k_pc_swizzle<<<blocks_for(nb, 256), 256, 0, s>>>(params + O.param[i], wbwd + O.bwd[i], Ly.taps, Ly.cin, Ly.cout,
                                                      1, nb);
      HNS_HIP(hipGetLastError());
    }
  }
}

void launch_pc_input(const DevTopo& T, const float* r_cs, const __bf16* w, __bf16* h0, hipStream_t s) {
  const int64_t ncell = (int64_t)T.L * kLeafVox;
  k_pc_input<kNT><<<blocks_for(ncell, 64 * kNT), kTpb, 0, s>>>(T, r_cs, reinterpret_cast<const bf16x8*>(w), h0);
  HNS_HIP(hipGetLastError());
}

#define HNS_PC_CONV(TAPS, CIN, COUT)                                                                     \
  if (taps == TAPS && cin == CIN && cout == COUT) {                                                     \
    if (act == 0)                                                                                       \
      k_pc_conv<TAPS, CIN, COUT, 0, kNT><<<g, kTpb, 0, s>>>(T, lgS, x, wv, y, mask, nullptr, nullptr);  \
    else if (act == 1)                                                                                  \
      k_pc_conv<TAPS, CIN, COUT, 1, kNT><<<g, kTpb, 0, s>>>(T, lgS, x, wv, y, mask, nullptr, nullptr);  \
    else                                                                                                \
      k_pc_conv<TAPS, CIN, COUT, 2, kNT><<<g, kTpb, 0, s>>>(T, lgS, x, wv, y, mask, nullptr, nullptr);  \
    HNS_HIP(hipGetLastError());                                                                         \
    return;                                                                                             \
  }

void launch_pc_conv(const DevTopo& T, int level, int taps, int cin, int cout, const __bf16* x, const __bf16* w,
                    __bf16* y, int act, const __bf16* mask, hipStream_t s) {
  const int lgS = 3 - level;
  const int64_t ncell = (int64_t)T.L << (3 * lgS);
  const unsigned g = blocks_for(ncell, 64 * kNT);
  const bf16x8* wv = reinterpret_cast<const bf16x8*>(w);
  // forward layers and their backward-data transposes
  HNS_PC_CONV(27, 16, 32)
  HNS_PC_CONV(27, 32, 64)
  HNS_PC_CONV(27, 64, 64)
  HNS_PC_CONV(1, 64, 32)
  HNS_PC_CONV(27, 32, 32)
  HNS_PC_CONV(1, 32, 16)
  HNS_PC_CONV(27, 16, 16)
  HNS_PC_CONV(27, 32, 16)
  HNS_PC_CONV(27, 64, 32)
  HNS_PC_CONV(1, 32, 64)
  HNS_PC_CONV(1, 16, 32)
  fail(HNS_ERR_INVALID_ARGUMENT, "no preconditioner conv kernel for this shape");
}
#undef HNS_PC_CONV

void launch_pc_output(const DevTopo& T, const __bf16* x, const __bf16* w, const float* params_out, float* z_cs,
                      __bf16* h, hipStream_t s) {
  const int64_t ncell = (int64_t)T.L * kLeafVox;
  k_pc_conv<27, 16, 16, 3, kNT><<<blocks_for(ncell, 64 * kNT), kTpb, 0, s>>>(
      T, 3, x, reinterpret_cast<const bf16x8*>(w), h, nullptr, params_out, z_cs);
  HNS_HIP(hipGetLastError());
}

void launch_pc_input_pool(const DevTopo& T, const float* r_cs, const __bf16* w, __bf16* h0, __bf16* p1, hipStream_t s) {
  k_pc_input_pool<<<persistent_groups(T.L, 7), kTpb, 0, s>>>(T, r_cs, reinterpret_cast<const bf16x8*>(w), h0, p1);
  HNS_HIP(hipGetLastError());
}

void launch_pc_output_fused(const DevTopo& T, const __bf16* h0, const __bf16* g1, const __bf16* w,
                            const float* params_out, float* z_cs, hipStream_t s) {
  k_pc_output_fused<<<persistent_groups(T.L, 4), kTpb, 0, s>>>(T, h0, g1, reinterpret_cast<const bf16x8*>(w), params_out,
                                                               z_cs);
  HNS_HIP(hipGetLastError());
}

void launch_pc_level1(const DevTopo& T, int mode, const __bf16* xa, const __bf16* xb, const __bf16* w,
                      const __bf16* w2, __bf16* ya, __bf16* yb, hipStream_t s) {
  const int32_t groups = (T.L + 3) / 4;  // four leaves per workgroup pass
  const bf16x8* wv = reinterpret_cast<const bf16x8*>(w);
  const bf16x8* w2v = reinterpret_cast<const bf16x8*>(w2);
  if (mode == 0)
    k_pc_level1<0><<<persistent_groups(groups, 4), kTpb, 0, s>>>(T, xa, xb, wv, w2v, ya, yb);
  else
    k_pc_level1<1><<<persistent_groups(groups, 2), kTpb, 0, s>>>(T, xa, xb, wv, w2v, ya, yb);
  HNS_HIP(hipGetLastError());
}

void launch_pc_level2(const DevTopo& T, int mode, const __bf16* x, const __bf16* w, const __bf16* w2, __bf16* y,
                      hipStream_t s) {
  const int32_t groups = (T.L + 7) / 8;  // eight leaves per workgroup pass
  const bf16x8* wv = reinterpret_cast<const bf16x8*>(w);
  const bf16x8* w2v = reinterpret_cast<const bf16x8*>(w2);
  if (mode == 0)
    k_pc_level2<0><<<persistent_groups(groups, 4), kTpb, 0, s>>>(T, x, wv, w2v, y);
  else
    k_pc_level2<1><<<persistent_groups(groups, 2), kTpb, 0, s>>>(T, x, wv, w2v, y);
  HNS_HIP(hipGetLastError());
}

void launch_pc_pool(int32_t L, int level, int c, const __bf16* x, __bf16* y, float scale, hipStream_t s) {
  const int lgSc = 3 - (level + 1);
  const int64_t n = ((int64_t)L << (3 * lgSc)) * (c / 8);
  k_pc_pool<<<blocks_for(n, 256), 256, 0, s>>>(L, lgSc, c / 8, x, y, scale);
  HNS_HIP(hipGetLastError());
}

void launch_pc_upadd(int32_t L, int level, int c, const __bf16* a, const __bf16* b, __bf16* y, float scale,
                     const __bf16* mask, hipStream_t s) {
  const int lgSf = 3 - level;
  const int64_t n = ((int64_t)L << (3 * lgSf)) * (c / 8);
  if (mask)
    k_pc_upadd<true><<<blocks_for(n, 256), 256, 0, s>>>(L, lgSf, c / 8, a, b, y, scale, mask);
  else
    k_pc_upadd<false><<<blocks_for(n, 256), 256, 0, s>>>(L, lgSf, c / 8, a, b, y, scale, nullptr);
  HNS_HIP(hipGetLastError());
}

void launch_pc_output_bwd(const DevTopo& T, const float* dz_cs, const __bf16* h6, const float* params_out,
                          __bf16* dh, hipStream_t s) {
  const int64_t n = (int64_t)T.L * kLeafVox * 2;
  k_pc_output_bwd<<<blocks_for(n, 256), 256, 0, s>>>(T.L, dz_cs, h6, params_out, dh);
  HNS_HIP(hipGetLastError());
}

#define HNS_PC_WGRAD(TAPS, CIN, COUT)                                                             \
  if (Ly.taps == TAPS && Ly.cin == CIN && Ly.cout == COUT) {                                      \
    k_pc_wgrad<TAPS, CIN, COUT><<<(unsigned)(R * TAPS), kTpb, 0, s>>>(T, lgS, R, x, dy, partial); \
    HNS_HIP(hipGetLastError());                                                                   \
    done = true;                                                                                  \
  }

void launch_pc_wgrad(const DevTopo& T, int layer, const __bf16* x, const __bf16* dy, float* partial, float* grad,
                     hipStream_t s) {
  const PcLayer& Ly = kPcNet[layer];
  const int lgS = 3 - Ly.level;
  const int R = wgrad_ranges((int64_t)T.L << (3 * lgS));
  bool done = false;
  HNS_PC_WGRAD(27, 16, 32)
  HNS_PC_WGRAD(27, 32, 64)
  HNS_PC_WGRAD(27, 64, 64)
  HNS_PC_WGRAD(1, 64, 32)
  HNS_PC_WGRAD(27, 32, 32)
  HNS_PC_WGRAD(1, 32, 16)
  HNS_PC_WGRAD(27, 16, 16)
  if (!done) fail(HNS_ERR_INVALID_ARGUMENT, "no weight-gradient kernel for this layer");
  const int64_t n = (int64_t)Ly.taps * Ly.cin * Ly.cout;
  k_pc_reduce<<<blocks_for(n, 256), 256, 0, s>>>(n, R, partial, grad + pc_offsets().param[layer]);
  HNS_HIP(hipGetLastError());
}
#undef HNS_PC_WGRAD

void launch_pc_wgrad_input(const DevTopo& T, const float* r_cs, const __bf16* dh0, float* partial, float* grad,
                           hipStream_t s) {
  const int R = wgrad_ranges((int64_t)T.L * kLeafVox);
  k_pc_wgrad_input<<<(unsigned)(R * 27), kTpb, 0, s>>>(T, R, r_cs, dh0, partial);
  HNS_HIP(hipGetLastError());
  k_pc_reduce<<<blocks_for(27 * 16, 256), 256, 0, s>>>(27 * 16, R, partial, grad + pc_offsets().param[kC0]);
  HNS_HIP(hipGetLastError());
}

void launch_pc_wgrad_output(const DevTopo& T, const float* dz_cs, const __bf16* h6, float* partial, float* grad,
                            hipStream_t s) {
  const int R = wgrad_ranges((int64_t)T.L * kLeafVox);
  k_pc_wgrad_output<<<(unsigned)R, kTpb, 0, s>>>(T, R, dz_cs, h6, partial);
  HNS_HIP(hipGetLastError());
  k_pc_reduce<<<1, 256, 0, s>>>(16, R, partial, grad + pc_offsets().param[kOut]);
  HNS_HIP(hipGetLastError());
}

void launch_pc_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                    int t, hipStream_t s) {
  const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
  k_pc_adam<<<blocks_for(n, 256), 256, 0, s>>>(n, p, g, m, v, lr, b1, b2, eps, c1, c2);
  HNS_HIP(hipGetLastError());
}

void launch_pc_poisson(const DevTopo& T, const float* x, const float* y, float a, float b, float* out, double* part,
                       double* sums, hipStream_t s) {
  const unsigned G = cg_groups(T.L);
  k_pc_poisson<<<G, kTpb, 0, s>>>(T, x, y, a, b, out, part);
  HNS_HIP(hipGetLastError());
  k_pc_sum3<<<1, kTpb, 0, s>>>(part, (int)G, sums);
  HNS_HIP(hipGetLastError());
}

}  // namespace hns
