// Pressure-solve driver: the reference's fixed-count red-black Gauss-Seidel loop
// (src/Cuda/HNanoSolver.cu:364-373), and the to-tolerance variants selected through
// hns_solve_options (no reference counterpart).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "hns_cg.hpp"
#include "hns_mg.hpp"
#include "hns_precond.hpp"
#include "hns_solvers.hpp"

namespace hns {

struct CgState {
  DBuf<float> r, d0, d1, q, z;
  DBuf<float> dring, qring;  // flexible CG: kFcgDepth directions and products
  DBuf<double> rho, alpha, rz, pa, pb, bb, zq, last;
  // the solve's control words, read back in one copy: ints 0-3 (stop raised, stop iteration,
  // device iteration counter, deferred-x-update mark), then doubles 2-3 (the last stop test's r . r
  // and ||b||^2)
  DBuf<double> ctl;
  double* h_ctl = nullptr;
  int* flag() const { return reinterpret_cast<int*>(ctl.p); }
  const int* h_flag() const { return reinterpret_cast<const int*>(h_ctl); }
  int cap = 0;
  // runs of multigrid-preconditioned iterations captured as graphs, one per run length (at most
  // kBatchGraphs lengths: the first read's prediction, kFcgBatch, the runs cut at replacements),
  // and what they were made for
  static constexpr int kBatchGraphs = 6;
  hipStream_t cap_stream = nullptr;
  hipEvent_t read_ev = nullptr;  // the stop-flag read, for the host to wait on alone (on_stop)
  int batch_len[kBatchGraphs] = {};
  hipGraphExec_t batch_exec[kBatchGraphs] = {};
  int batch_next = 0;  // the slot a new length replaces when every slot is taken
  void drop_batches() {
    for (int i = 0; i < kBatchGraphs; ++i) {
      if (batch_exec[i]) (void)hipGraphExecDestroy(batch_exec[i]);
      batch_exec[i] = nullptr;
      batch_len[i] = 0;
    }
    batch_next = 0;
  }
  uint64_t key_topo = 0;
  int32_t key_L = 0;
  const float* key_p = nullptr;
  uint64_t key_gen = 0;
  int key_depth = 0;
  // the captured kernels take CgScalars by value: its per-iteration arrays (rho, alpha, rz:
  // reallocated together when a solve asks for more iterations than any before, which advances
  // arrays_gen) and its tolerance are part of what a graph was made for
  uint64_t arrays_gen = 0, key_arrays = 0;
  double key_tol2 = -1.0;
  ~CgState() {
    drop_batches();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    if (read_ev) (void)hipEventDestroy(read_ev);
    if (h_ctl) (void)hipHostFree(h_ctl);
  }
};

// One flexible-CG iteration on the multigrid V-cycle, launched on s (it: the host's count, which
// only selects kernels). Truncation 1 (the default): the V-cycle's last pass leaves z . r and
// z . q_prev, then direction and product in one launch; otherwise the general FCG(depth) kernels.
// fold_test: the stop test after the previous iteration's update runs inside the V-cycle's first
// pass (its separate launch was left out); own_test: this iteration's update is followed by its
// stop test launch (otherwise the next iteration folds it).
void Solvers::mg_iteration(const DevTopo& T, float* p, const CgScalars& S, int depth, int it, hipStream_t s,
                           Timing* tm, bool fold_test, bool own_test) {
  CgState& G = *cg_;
  auto begin = [&](int k) {
    if (tm) tm->begin(k, s);
  };
  auto end = [&] {
    if (tm) tm->end(s);
  };
  if (depth == 1 && mg_shape().smooth0 == 1) {
    const bool pcg = mg_shape().pcg != 0;
    // with S.xdone set, x += alpha d of the previous iteration rides on the coarse-level launch
    const McgXUpdate X{S.xdone ? p : nullptr, G.dring.p, S.alpha, S.xdone, S.ring};
    const McgHook H{S.pa + kSlotDR * kCgMaxGroups, pcg ? nullptr : G.zq.p, G.qring.p, (int64_t)N_, S.it, X, S.ring};
    begin(HNS_K_PRECOND);
    mg_->enqueue(T, G.r.p, G.z.p, S.flag, s, &H, fold_test ? &S : nullptr);
    end();
    begin(HNS_K_CG_PRODUCT);
    launch_fcg1_dirspmv(T, it == 0, G.z.p, G.dring.p, G.qring.p, G.zq.p, S, s, pcg);
    end();
  } else {
    if (fold_test || !own_test) fail(HNS_ERR_INVALID_ARGUMENT, "the stop test folds only into the fused V-cycle");
    begin(HNS_K_PRECOND);
    mg_->enqueue(T, G.r.p, G.z.p, S.flag, s);
    end();
    begin(HNS_K_CG_PRODUCT);
    launch_fcg_direction(T, it, depth, G.r.p, G.z.p, G.dring.p, G.qring.p, G.zq.p, S, s);
    end();
  }
  begin(HNS_K_CG_UPDATE);
  launch_fcg_update(T.L, S.xdone ? nullptr : p, G.r.p, G.dring.p, G.qring.p, S, s);
  end();
  if (own_test) {
    begin(HNS_K_OTHER);
    launch_fcg_test(T.L, true, S, s);
    end();
  }
}

// n >= 2 flexible-CG iterations on the multigrid V-cycle, replayed from a graph made on first use
// for these buffers and this run length. Every kernel takes the iteration number from the device
// counter (the product kernel forms d = z at iteration 0), so one graph serves a run of n from any
// iteration (from iteration depth on with truncation above 1).
void Solvers::mg_batch(const DevTopo& T, float* p, const CgScalars& S, int depth, int n, hipStream_t s) {
  CgState& G = *cg_;
  Multigrid& M = *mg_;
  M.prepare(T, s);
  if (G.key_topo != T.id || G.key_L != T.L || G.key_p != p || G.key_gen != M.generation ||
      G.key_depth != depth || G.key_arrays != G.arrays_gen || G.key_tol2 != S.tol2) {
    G.drop_batches();
    G.key_topo = T.id;
    G.key_L = T.L;
    G.key_p = p;
    G.key_gen = M.generation;
    G.key_depth = depth;
    G.key_arrays = G.arrays_gen;
    G.key_tol2 = S.tol2;
  }
  int slot = -1;
  for (int i = 0; i < CgState::kBatchGraphs; ++i)
    if (G.batch_exec[i] && G.batch_len[i] == n) slot = i;
  if (slot < 0) {
    slot = G.batch_next;
    G.batch_next = (G.batch_next + 1) % CgState::kBatchGraphs;
    if (G.batch_exec[slot]) {
      HNS_HIP(hipGraphExecDestroy(G.batch_exec[slot]));
      G.batch_exec[slot] = nullptr;
      G.batch_len[slot] = 0;
    }
    if (!G.cap_stream) HNS_HIP(hipStreamCreateWithFlags(&G.cap_stream, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    HNS_HIP(hipStreamBeginCapture(G.cap_stream, hipStreamCaptureModeThreadLocal));
    try {
      // with truncation 1, every iteration but the run's first takes the stop test of the update
      // before it inside its V-cycle's first pass (one launch fewer per iteration; the run's last
      // update keeps its own test, which the host's read follows)
      const bool fold = depth == 1 && mg_shape().smooth0 == 1;
      for (int k = 0; k < n; ++k)
        mg_iteration(T, p, S, depth, depth, G.cap_stream, nullptr, fold && k > 0, !fold || k == n - 1);
    } catch (...) {
      (void)hipStreamEndCapture(G.cap_stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HNS_HIP(hipStreamEndCapture(G.cap_stream, &g));
    const hipError_t e = hipGraphInstantiate(&G.batch_exec[slot], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HNS_HIP(e);
    G.batch_len[slot] = n;
  }
  HNS_HIP(hipGraphLaunch(G.batch_exec[slot], s));
}

Solvers::Solvers() = default;

Multigrid& Solvers::multigrid() {
  if (!mg_) mg_ = new Multigrid();
  return *mg_;
}

// the device time of a to-tolerance solve ends at each stop-flag read (the last one ends it); the
// read then waits for the stream, so the event is complete when the solve returns
void Solvers::mark_solve_end(const SolveRequest& r, hipStream_t s) {
  if (!r.want_residual) return;
  HNS_HIP(hipEventRecord(e1_, s));
  e1_marked_ = true;
}

Solvers::~Solvers() {
  if (h_red_) (void)hipHostFree(h_red_);
  if (e0_) (void)hipEventDestroy(e0_);
  if (e1_) (void)hipEventDestroy(e1_);
  delete cg_;
  delete mg_;
}

void Solvers::init(int32_t L, size_t N) {
  L_ = L;
  N_ = N;
  // One launch per colour (the reference's structure, measured faster on MI355X) unless
  // HNS_RBGS_KERNEL=blocked selects the LDS temporally blocked full iteration; both give
  // identical results.
  const char* k = std::getenv("HNS_RBGS_KERNEL");
  per_colour_ = !(k && std::strcmp(k, "blocked") == 0);
  pb_.alloc(N);
  partial_.alloc(2 * (size_t)L);
  red_.alloc(2);
  if (!h_red_) HNS_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_red_), 2 * sizeof(double), hipHostMallocDefault));
  if (!e0_) HNS_HIP(hipEventCreate(&e0_));
  if (!e1_) HNS_HIP(hipEventCreate(&e1_));
}

// n iterations continuing from p (or from p = 0 when first_it == 0); the result ends in p.
void Solvers::rbgs_iters(const DevTopo& T, const float* div, float* p, int first_it, int n, const SolveRequest& r,
                         Timing& tm, hipStream_t s) {
  if (per_colour_) {
    for (int it = first_it; it < first_it + n; ++it) {
      tm.begin(HNS_K_RBGS, s);
      launch_rbgs(T, div, p, r.dx, 0, r.omega, it == 0, s);
      tm.end(s);
      tm.begin(HNS_K_RBGS, s);
      launch_rbgs(T, div, p, r.dx, 1, r.omega, it == 0, s);
      tm.end(s);
    }
    return;
  }
  // ping-pong between p and pb_: a first iteration does not read its input, so it may write p
  // directly; choose the first destination so that the last one is p
  float* src = p;
  float* dst = ((n & 1) == (first_it == 0 ? 1 : 0)) ? p : pb_.p;
  if (first_it != 0 && dst == p) dst = pb_.p;  // odd n from a live p: finish with a copy
  for (int k = 0; k < n; ++k) {
    const int it = first_it + k;
    tm.begin(HNS_K_RBGS, s);
    launch_rbgs_iteration(T, div, src, dst, r.dx, r.omega, it == 0, s);
    tm.end(s);
    src = dst;
    dst = (dst == p) ? pb_.p : p;
  }
  if (n > 0 && src != p) HNS_HIP(hipMemcpyAsync(p, src, N_ * sizeof(float), hipMemcpyDeviceToDevice, s));
}

double Solvers::rel_residual(const DevTopo& T, const float* div, const float* p, float dx, Timing& tm,
                             hipStream_t s) {
  tm.begin(HNS_K_RESIDUAL, s);
  launch_residual(T, p, div, dx, partial_.p, red_.p, s);
  tm.end(s);
  HNS_HIP(hipMemcpyAsync(h_red_, red_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HNS_HIP(hipStreamSynchronize(s));
  const double r2 = h_red_[0], d2 = h_red_[1];
  if (!std::isfinite(r2) || !std::isfinite(d2)) return NAN;  // a NaN or Inf in p or div
  return d2 > 0.0 ? std::sqrt(r2 / d2) : (r2 > 0.0 ? INFINITY : 0.0);
}

SolveResult Solvers::solve(const DevTopo& T, const float* div, float* p, int K, const SolveRequest& r, Timing& tm,
                           hipStream_t s) {
  SolveResult res;
  e1_marked_ = false;
  if (r.want_residual) HNS_HIP(hipEventRecord(e0_, s));
  switch (r.solver) {
    case HNS_SOLVER_RBGS_FIXED: {
      if (K == 0) HNS_HIP(hipMemsetAsync(p, 0, N_ * sizeof(float), s));
      rbgs_iters(T, div, p, 0, K, r, tm, s);
      res.iterations = K;
      if (r.want_residual) res.rel_residual = rel_residual(T, div, p, r.dx, tm, s);
      break;
    }
    case HNS_SOLVER_RBGS_TOL: {
      const int chunk = r.check_every > 0 ? r.check_every : 50;
      const int cap = r.max_iters > 0 ? r.max_iters : 100000;
      HNS_HIP(hipMemsetAsync(p, 0, N_ * sizeof(float), s));
      RbgsDev d{this, T, div, p, r, tm, s};
      const LoopResult lr = rbgs_tol_loop(d, (double)r.tol, cap, chunk, 0, rel_residual(T, div, p, r.dx, tm, s),
                                          &res.rel_residual);
      res.iterations = lr.iterations;
      res.end = lr.end;
      break;
    }
    case HNS_SOLVER_CG:
    case HNS_SOLVER_PCG_LEARNED:
    case HNS_SOLVER_MGCG:
      res = solve_cg(T, div, p, r, tm, s);
      break;
    default:
      fail(HNS_ERR_INVALID_ARGUMENT, "unknown pressure solver");
  }
  if (r.want_residual) {
    // a to-tolerance CG solve's time ends at the stop-flag read that ended it, whose event the
    // host has already waited for: no second wait on the stream here
    if (!e1_marked_) {
      HNS_HIP(hipEventRecord(e1_, s));
      HNS_HIP(hipEventSynchronize(e1_));
    }
    float ms = 0.f;
    HNS_HIP(hipEventElapsedTime(&ms, e0_, e1_));
    res.ms = ms;
  }
  return res;
}

SolveResult Solvers::solve_cg(const DevTopo& T, const float* div, float* p, const SolveRequest& r, Timing& tm,
                              hipStream_t s) {
  // flexible CG on a preconditioner: the learned network (truncation kFcgDepth) or the multigrid
  // V-cycle (linear and nearly symmetric: a short truncation suffices)
  const bool learned = r.solver == HNS_SOLVER_PCG_LEARNED, mg = r.solver == HNS_SOLVER_MGCG;
  const bool pre = learned || mg;
  const int depth = learned ? kFcgDepth : mg_shape().depth;
  if (learned && (!r.pc || !r.pcw)) fail(HNS_ERR_INVALID_ARGUMENT, "no learned preconditioner is set");
  if (mg) multigrid();
  const int cap = r.max_iters > 0 ? r.max_iters : 100000;
  if (!cg_) cg_ = new CgState();
  CgState& G = *cg_;
  if (!G.r.p) {
    G.r.alloc(N_);
    G.d0.alloc(N_);
    G.d1.alloc(N_);
    G.q.alloc(N_);
    HNS_HIP(hipMemsetAsync(G.q.p, 0, N_ * sizeof(float), s));
    G.pa.alloc((size_t)kCgMaxGroups * kCgSlots);
    G.pb.alloc(kCgMaxGroups);
    G.bb.alloc(1);
    G.ctl.alloc(4);
    HNS_HIP(hipHostMalloc(reinterpret_cast<void**>(&G.h_ctl), 4 * sizeof(double), hipHostMallocDefault));
  }
  if (G.cap < cap) {
    G.rho.alloc((size_t)cap + 1);
    G.alpha.alloc((size_t)cap + 1);
    G.rz.alloc((size_t)cap + 1);
    G.cap = cap;
    ++G.arrays_gen;
  }
  CgScalars S;
  S.pa = G.pa.p;
  S.pb = G.pb.p;
  S.rho = G.rho.p;
  S.alpha = G.alpha.p;
  S.rz = G.rz.p;
  S.flag = G.flag();
  S.it = G.flag() + 2;
  S.last = G.ctl.p + 2;
  S.bb = G.bb.p;
  S.tol2 = (double)r.tol * (double)r.tol;
  if (mg && depth == 1) S.ring = 2;
  // FCG(1) on the fused V-cycle with a one-workgroup coarse solve: the x update is deferred into
  // the next V-cycle's coarse-level launch (flushed where x must be complete)
  if (mg && depth == 1 && mg_shape().smooth0 == 1) {
    mg_->prepare(T, s);
    if (mg_->coarse_fused()) S.xdone = G.flag() + 3;
  }
  tm.begin(HNS_K_OTHER, s);
  launch_cg_init(T.L, div, r.dx, p, G.r.p, S, s);
  tm.end(s);
  // The device stops at the first iteration whose recurrence residual passes the test. The stop
  // is accepted only when that residual was just replaced by b - A x; otherwise the residual is
  // replaced and the same iteration is tested again (and continued if it fails).
  if (pre && !G.z.p) {
    G.z.alloc(N_);
    G.dring.alloc((size_t)kFcgDepth * N_);
    G.qring.alloc((size_t)kFcgDepth * N_);
    G.zq.alloc((size_t)kFcgDepth * kCgMaxGroups);
  }
  // a preconditioned iteration costs far more than a sync: check the stop flag every kFcgBatch
  // iterations (the multigrid solve replays each run between two reads from one graph)
  // (the multigrid solve's first read comes after 2 batches: it rarely stops before, and iterations
  // past a stop return at once)
  // (with per-kernel timing on, the flag is read after every iteration, so no launch past the stop
  // returning at once is timed: the timed launches are the ones that do the work, and the device
  // computes the same as with batches)
  // (the multigrid solve reads first after the previous solve's count when it has one)
  const int mg_first = mg_last_iters_ > 0 ? std::max(mg_last_iters_, kFcgBatch) : 2 * kFcgBatch;
  const int batch0 = tm.enabled ? 1 : (mg ? mg_first : (pre ? kFcgBatch : 8));
  const int batch_max = tm.enabled ? 1 : (pre ? kFcgBatch : 128);
  LoopResult lr;
  bool continued = false;  // the reads launched r.on_stop
  if (pre) {
    FcgDev d{this, T, div, p, r, S, depth, mg, tm, s};
    lr = fcg_loop(d, cap, batch0, batch_max, kResidualReplace);
    continued = d.continues();
  } else {
    CgDev d{this, T, div, p, r, S, tm, s};
    lr = cg_loop(d, cap, batch0, batch_max, kResidualReplace);
  }
  // x complete before anything reads it; after a stop nothing is pending (the flush before the
  // stop's re-test applied it), so the launch is left out there
  if (S.xdone && lr.end != kLoopStopped) {
    tm.begin(HNS_K_OTHER, s);
    launch_fcg_xflush(T.L, p, G.dring.p, S, kFlushEnd, s);
    tm.end(s);
  }
  SolveResult res;
  res.iterations = lr.iterations;
  res.end = lr.end;
  if (mg) mg_last_iters_ = lr.end == kLoopStopped ? lr.iterations : 0;
  // the read that ended the solve found the stop standing, which its on_stop work was gated on
  res.on_stop_ran = continued && lr.end == kLoopStopped;
  if (pre && lr.end == kLoopStopped) {
    // the stop stands on r = b - A x evaluated in FP64 (its r.r and ||b||^2 came with the flag)
    const double rr = lr.rr, bb = lr.bb;
    res.rel_residual = bb > 0.0 ? std::sqrt(rr / bb) : (rr > 0.0 ? INFINITY : 0.0);
  } else {
    res.rel_residual = rel_residual(T, div, p, r.dx, tm, s);
  }
  if (!std::isfinite(res.rel_residual)) res.end = kLoopNonFinite;
  // plain RBGS is the fallback of the learned preconditioner (an untrained network can stall the
  // solve, a breakdown can end it): continue from the PCG iterate while the residual is finite.
  // The multigrid solve stops at max_iters like CG.
  if (learned && res.end != kLoopNonFinite && !(res.rel_residual <= (double)r.tol)) {
    RbgsDev d{this, T, div, p, r, tm, s};
    const LoopResult fb = rbgs_tol_loop(d, (double)r.tol, 200000, 50, 1, res.rel_residual, &res.rel_residual);
    res.fallback = fb.iterations;
    res.end = fb.end;
  }
  return res;
}

// ---- device operations of the solve loops (hns_solve_loop.hpp) ------------------------------------

void Solvers::RbgsDev::sweeps(int first_it, int n) { sol->rbgs_iters(T, div, p, first_it, n, r, tm, s); }
double Solvers::RbgsDev::residual() { return sol->rel_residual(T, div, p, r.dx, tm, s); }

void Solvers::CgDev::replace_periodic(int) {
  tm.begin(HNS_K_OTHER, s);
  launch_cg_replace(T, div, r.dx, p, sol->cg_->r.p, S, false, s);
  tm.end(s);
}
void Solvers::CgDev::iteration(int it) {
  CgState& G = *sol->cg_;
  float* d_in = (it & 1) ? G.d1.p : G.d0.p;
  float* d_out = (it & 1) ? G.d0.p : G.d1.p;
  tm.begin(HNS_K_CG_PRODUCT, s);
  launch_cg_product(T, it, it == 0, G.r.p, d_in, d_out, G.q.p, S, s);
  tm.end(s);
  tm.begin(HNS_K_CG_UPDATE, s);
  launch_cg_update(T.L, it, p, G.r.p, d_out, G.q.p, S, s);
  tm.end(s);
}
CgReport Solvers::CgDev::read() {
  CgState& G = *sol->cg_;
  sol->mark_solve_end(r, s);
  HNS_HIP(hipMemcpyAsync(G.h_ctl, G.ctl.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  HNS_HIP(hipStreamSynchronize(s));
  return CgReport{G.h_flag()[0], G.h_flag()[1]};
}
void Solvers::CgDev::replace_forced() {
  tm.begin(HNS_K_OTHER, s);
  launch_cg_replace(T, div, r.dx, p, sol->cg_->r.p, S, true, s);
  tm.end(s);
}

void Solvers::FcgDev::replace_periodic(int) {
  CgState& G = *sol->cg_;
  tm.begin(HNS_K_OTHER, s);
  if (S.xdone) launch_fcg_xflush(T.L, p, G.dring.p, S, kFlushRunning, s);
  launch_cg_replace(T, div, r.dx, p, G.r.p, S, false, s);
  launch_fcg_test(T.L, false, S, s);  // test the residual just replaced
  tm.end(s);
}
// (with per-kernel timing on, the iterations are launched one by one so events can bracket them)
// (from iteration 0 on with the fused truncation-1 iteration, whose product kernel forms d = z
// there; the general FCG(depth) kernels are chosen by the host's count, so from depth on)
int Solvers::FcgDev::batchable(int it, int left) {
  const bool any_it = depth == 1 && mg_shape().smooth0 == 1;
  return mg && !tm.enabled && (any_it || it >= depth) ? left : 0;
}
void Solvers::FcgDev::batch(int, int n) {
  tm.begin(HNS_K_PRECOND, s);
  sol->mg_batch(T, p, S, depth, n, s);
  tm.end(s);
}
void Solvers::FcgDev::first_test() {
  tm.begin(HNS_K_OTHER, s);
  launch_fcg_test(T.L, false, S, s);
  tm.end(s);
}
// flexible CG on z = M(r); the stop test of iteration it > 0 ran after the last update
void Solvers::FcgDev::iteration(int it) {
  CgState& G = *sol->cg_;
  if (mg) {
    sol->mg_->prepare(T, s);
    sol->mg_iteration(T, p, S, depth, it, s, &tm);
    return;
  }
  tm.begin(HNS_K_PRECOND, s);
  r.pc->forward(T, *r.pcw, G.r.p, G.z.p, false, s);
  tm.end(s);
  tm.begin(HNS_K_CG_PRODUCT, s);
  launch_fcg_direction(T, it, depth, G.r.p, G.z.p, G.dring.p, G.qring.p, G.zq.p, S, s);
  tm.end(s);
  tm.begin(HNS_K_CG_UPDATE, s);
  launch_fcg_update(T.L, p, G.r.p, G.dring.p, G.qring.p, S, s);
  launch_fcg_test(T.L, true, S, s);
  tm.end(s);
}
// a stop raised in the batch is re-tested on r = b - A x before the host looks, so one read
// decides: stopped, or continue from the device's iteration
FcgReport Solvers::FcgDev::end_batch() {
  CgState& G = *sol->cg_;
  tm.begin(HNS_K_OTHER, s);
  if (S.xdone) launch_fcg_xflush(T.L, p, G.dring.p, S, kFlushStopped, s);
  launch_fcg_retest(T, div, r.dx, p, G.r.p, S, s);
  tm.end(s);
  sol->mark_solve_end(r, s);
  HNS_HIP(hipMemcpyAsync(G.h_ctl, G.ctl.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  if (continues()) {
    if (!G.read_ev) HNS_HIP(hipEventCreateWithFlags(&G.read_ev, hipEventDisableTiming));
    HNS_HIP(hipEventRecord(G.read_ev, s));
    r.on_stop(S.flag);
    HNS_HIP(hipEventSynchronize(G.read_ev));
  } else {
    HNS_HIP(hipStreamSynchronize(s));
  }
  const int* f = G.h_flag();
  return FcgReport{f[0], f[1], f[2], G.h_ctl[2], G.h_ctl[3]};
}

}  // namespace hns
