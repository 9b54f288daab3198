// Geometric multigrid V-cycle over the leaf hierarchy (kernels_mg.hip): the preconditioner of
// HNS_SOLVER_MGCG, applied inside the flexible CG of kernels_cg.hip.
#pragma once
#include "hns_cg.hpp"
#include "hns_internal.hpp"

namespace hns {

// levels 0 (voxels) .. 3 (one cell per leaf)
constexpr int kMgLevels = 4;

// V-cycle shape: red-black SOR iterations on the coarsest level (one workgroup for up to 8192
// leaves of each colour, else one launch per colour sweep) and their over-relaxation, pre- and
// post-smoothing Gauss-Seidel iterations on level 0 and on levels 1-2, and the flexible CG's
// truncation around the V-cycle, and the scale of every coarse-grid correction (piecewise-constant
// prolongation under-corrects smooth errors; over-correcting by 1.6 took the 8 M-voxel fire solve
// from 16 to 12 iterations). The values are the measured best at 8 M voxels, fixed at build time:
// they are the library's solver, not per-process settings.
struct MgShape {
  int coarse = 12, smooth0 = 1, smooth = 1, depth = 1, pcg = 0;
  float coarse_omega = 1.85f;
  // The two-level coarse solve (c4 = 1; 0: `coarse` SOR iterations on level 3 alone): c3_pre and
  // c3_post red-black iterations on level 3 (relaxations c3_omega and c3_omega_post) around c4_iters
  // SOR iterations (c4_omega) on level 4 (2^3-leaf blocks, Galerkin operator), the restricted
  // residual scaled by c4_over. side_groups workgroups carry the deferred x update beside the
  // one-workgroup coarse solve after side_sleeps * 127 * 64 idle cycles (-1: 96 and CH / 2 beside
  // the one-level solve, 128 and 2 beside the two-level one).
  int c4 = 1, c3_pre = 1, c3_post = 3, c4_iters = 6;
  float c3_omega = 1.1f, c3_omega_post = 1.9f, c4_omega = 1.7f, c4_over = 1.6f;
  int side_groups = -1, side_sleeps = -1;
  float over = 1.6f;  // coarse-correction scale (applied to each restricted residual)
  float over_deep = 1.6f;  // the same for the restriction from level 1 to level 2
  float over3 = 1.8f;      // and from level 2 to level 3
  // Level 0 is pre-smoothed by one pointwise (Jacobi) step from zero, z' = jacobi0 r / 6, and
  // post-smoothed by one red-black Gauss-Seidel iteration (relaxation omega0); levels 1-2 by one
  // red-black SOR iteration (omega_l) on the way down, and level 1 by one, level 2 by post2 on the
  // way up. Chosen with the CPU oracle's multigrid-preconditioned CG on the smoke, control and fire
  // scenes (up to 14 M voxels) and on 14 consecutive steps of the 8 M-voxel fire scene: every scene
  // converges to 1e-5 in 8 iterations (9 with one post iteration on level 2, level 2 restricted by
  // 1.6 and two level-3 iterations at 1.8 after level 4; the largest residual after 8 iterations over
  // the 14 steps 8.9e-6, against 3.0e-5), and the level-0 pre pass writes only the level-1
  // right-hand side.
  float jacobi0 = 1.75f;
  float omega0 = 1.0f, omega_l = 1.5f;
  int post2 = 2;
};
const MgShape& mg_shape();

// The default V-cycle is 7 launches (one fused pre- and post-smoothing pass per level 0-2 and the
// one-workgroup coarse solve); the unfused variant ~35, most of them on the coarse levels. It is
// captured once into a HIP graph per (topology, r, z, stop, variant) and replayed (the captured
// kernels take only these pointers, and none records a memset), which removes most of the
// per-launch cost.
// V-cycle variants that compute the same bits as the default (Multigrid::apply's `variant` bits)
constexpr unsigned kMgCoarseLaunches = 1;  // coarsest level as one launch per colour sweep
constexpr unsigned kMgUnfused = 2;         // levels 0-2 as separate sweep, restriction and prolongation launches

// The deferred x update of the flexible CG (x null: none): the coarse-level launch's extra
// workgroups apply x += alpha_{k-1} d_{k-1} (alpha[k - 1], direction ring slot (k - 1) mod
// kFcgDepth of Dr, n floats each) when k = *it >= 1 and *xdone < k, while its first workgroup
// solves the coarsest level. Needs the one-workgroup coarse solve (Multigrid::coarse_fused).
struct McgXUpdate {
  float* x;
  const float* Dr;
  const double* alpha;
  const int* xdone;
  int ring;  // the direction ring's slots in use (CgScalars::ring)
};
// What the last level-0 pass leaves for flexible CG with truncation 1 (kernels_cg.hip): per
// workgroup (cg_groups(L) of them) partials of z . r into zr and of z . q_prev into zq, q_prev the
// product of the previous iteration in the ring Qr (slot (*it - 1) mod ring, n floats each).
// xu: the deferred x update, if any.
struct McgHook {
  double* zr;
  double* zq;  // null: no z . q_prev (the PCG variant)
  const float* Qr;
  int64_t n;
  const int* it;
  McgXUpdate xu;
  int ring;  // the product ring's slots in use (CgScalars::ring)
};

struct Multigrid {
  int32_t L = 0;
  DBuf<float> b[kMgLevels], u[kMgLevels];  // levels 1..3 (leaf-dense); level 0 uses the caller's vectors
  DBuf<float> ut[kMgLevels];               // levels 1..2: pre-smoothed u of the fused path
  void ensure(int32_t leaves);
  // buffers and coarse layout for topology T (synchronises s when the layout is built)
  void prepare(const DevTopo& T, hipStream_t s) {
    ensure(T.L);
    prepare_coarse(T, s);
  }
  // the V-cycle's kernels launched on s (for capture into a caller's graph; prepare first)
  // hook: also leave the flexible-CG partials (null: none); test: the level-0 pre pass first runs
  // the flexible CG's stop test that follows the previous update (k_fcg_test's advancing test, on
  // the same partials, in every workgroup; null: none)
  void enqueue(const DevTopo& T, const float* r_cs, float* z_cs, const int* stop, hipStream_t s,
               const McgHook* hook = nullptr, const CgScalars* test = nullptr) {
    launch_all(T, r_cs, z_cs, stop, 0u, s, hook, test);
  }
  uint64_t generation = 0;  // advanced whenever buffers or the coarse layout change
  // the coarsest level is solved by one workgroup (after prepare for the topology)
  bool coarse_fused() const;
  // z = V(r) on colour-split level-0 vectors: approximately A^-1 r for A = 6 I - (face neighbours).
  // graph: replay the captured V-cycle (captured on first use for these pointers); otherwise launch
  // the kernels one by one. Both run the same kernels on the same data. stop: a device flag (null:
  // none); every kernel of the V-cycle returns at once while it is non-zero (the solve has ended).
  // variant: kMg* bits selecting an equivalent kernel sequence (same bits).
  void apply(const DevTopo& T, const float* r_cs, float* z_cs, hipStream_t s, bool graph = true,
             const int* stop = nullptr, unsigned variant = 0);
  ~Multigrid();

 private:
  void launch_all(const DevTopo& T, const float* r_cs, float* z_cs, const int* stop, unsigned variant,
                  hipStream_t s, const McgHook* hook = nullptr, const CgScalars* test = nullptr);
  // colour-sorted slot layout of the coarsest level for its one-workgroup solve, built on the host
  // once per topology (synchronises s); coarse_ch_ = 0 when the leaves do not fit one workgroup
  void prepare_coarse(const DevTopo& T, hipStream_t s);
  DBuf<int32_t> cmap_;   // leaf -> slot
  DBuf<uint32_t> cnb_;   // three planes of packed 16-bit face-neighbour slots
  int coarse_ch_ = 0;
  // level 4 of the two-level coarse solve (c4_ch_ = 0: none; kernels_mg.hip's C4Dev for the arrays)
  DBuf<uint32_t> c4nb_, c4w_, c4ch3_, c4par3_;
  DBuf<float> c4invd_, c4b_, c4u_, rho3_;
  DBuf<int32_t> c4chl_, c4parl_;
  int c4_ch_ = 0;
  bool c4_fused_ = false;  // levels 3 and 4 fit the fused two-level kernel
  uint64_t ckey_id_ = 0;  // topology the coarse layout was made for
  int32_t ckey_L_ = 0;
  hipStream_t cap_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  uint64_t key_id_ = 0;
  int32_t key_L_ = 0;
  const float* key_r_ = nullptr;
  float* key_z_ = nullptr;
  const int* key_stop_ = nullptr;
  unsigned key_variant_ = 0;
};

// One red-black SOR colour sweep (relaxation omega) on A u = b, colour-split (the RBGS kernel with
// the right-hand side taken as b); `first`: u starts at zero; returns at once while *stop != 0.
void launch_rbgs_rhs(const DevTopo& T, const float* b_cs, float* u_cs, int colour, bool first, float omega,
                     const int* stop, hipStream_t s);

}  // namespace hns
