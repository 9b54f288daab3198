// Pressure-solve driver and per-kernel event timing of a context.
#pragma once
#include <functional>
#include <vector>

#include "hns_internal.hpp"
#include "hns_solve_loop.hpp"

namespace hns {

// HIP-event timing of kernel classes (HNS_K_*), enabled per context for measurement runs.
struct Timing {
  bool enabled = false;
  double total_ms[HNS_K_COUNT] = {};
  int64_t launches[HNS_K_COUNT] = {};

  void begin(int k, hipStream_t s) {
    if (!enabled) return;
    cur_ = k;
    a_ = event();
    HNS_HIP(hipEventRecord(a_, s));
  }
  void end(hipStream_t s) {
    if (!enabled) return;
    hipEvent_t b = event();
    HNS_HIP(hipEventRecord(b, s));
    pending_.push_back(Rec{cur_, a_, b});
  }
  // call after the stream has been synchronised
  void collect() {
    for (const Rec& r : pending_) {
      float ms = 0.f;
      HNS_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      total_ms[r.k] += ms;
      launches[r.k] += 1;
    }
    pending_.clear();
    used_ = 0;
  }
  void reset() {
    for (int k = 0; k < HNS_K_COUNT; ++k) {
      total_ms[k] = 0.0;
      launches[k] = 0;
    }
  }
  ~Timing() {
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }

 private:
  struct Rec {
    int k;
    hipEvent_t a, b;
  };
  hipEvent_t event() {
    if (used_ == pool_.size()) {
      hipEvent_t e;
      HNS_HIP(hipEventCreate(&e));
      pool_.push_back(e);
    }
    return pool_[used_++];
  }
  std::vector<hipEvent_t> pool_;
  size_t used_ = 0;
  std::vector<Rec> pending_;
  int cur_ = 0;
  hipEvent_t a_ = nullptr;
};

class LearnedPrecond;
struct PcWork;

struct SolveRequest {
  LearnedPrecond* pc = nullptr;  // HNS_SOLVER_PCG_LEARNED
  PcWork* pcw = nullptr;
  int solver = HNS_SOLVER_RBGS_FIXED;
  float tol = 1e-5f;
  int max_iters = 100000;
  int check_every = 0;
  float omega = 1.f;
  float dx = 1.f;
  bool want_residual = false;
  // The work that follows a converged multigrid-preconditioned solve, launched on the solve's stream
  // after each read of the stop flag is enqueued and gated by that device flag (its kernels return
  // at once unless the read finds the solve stopped); the host then waits for the read alone, so
  // the device goes on into this work while the host takes the verdict. Empty: none.
  std::function<void(const int* gate)> on_stop;
};
struct SolveResult {
  int iterations = 0;
  int fallback = 0;  // RBGS iterations after a learned PCG that missed tol
  double rel_residual = -1.0;
  double ms = 0.0;
  LoopEnd end = kLoopStopped;  // how a to-tolerance solve ended (kLoopStopped for the fixed count)
  bool on_stop_ran = false;    // the request's on_stop work ran on the device (the caller skips it)
};

struct CgState;

// Pressure solves on colour-split p and div (p starts from 0 in every solve, HNanoSolver.cu:364).
class Solvers {
 public:
  Solvers();
  ~Solvers();
  Solvers(const Solvers&) = delete;
  Solvers& operator=(const Solvers&) = delete;
  void init(int32_t L, size_t N);
  SolveResult solve(const DevTopo& T, const float* div, float* p, int K, const SolveRequest& r, Timing& tm,
                    hipStream_t s);
  // ||div - L p|| / ||div|| (0 when div == 0); synchronises s.
  double rel_residual(const DevTopo& T, const float* div, const float* p, float dx, Timing& tm, hipStream_t s);
  // the multigrid V-cycle of the HNS_SOLVER_MGCG solves (made on first use)
  struct Multigrid& multigrid();

 private:
  int32_t L_ = 0;
  size_t N_ = 0;
  bool per_colour_ = false;
  DBuf<float> pb_;  // ping-pong partner of p for the temporally blocked iteration
  DBuf<double> partial_, red_;
  double* h_red_ = nullptr;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  bool e1_marked_ = false;  // e1_ was recorded before a stop-flag read this solve waited for
  void mark_solve_end(const SolveRequest& r, hipStream_t s);
  CgState* cg_ = nullptr;
  struct Multigrid* mg_ = nullptr;
  // iterations the last multigrid-preconditioned solve on these solvers took to its stop (0: none
  // yet): the next solve reads its stop flag first right after that many (consecutive steps of a
  // simulation take about the same count), so no batch runs launches past a predicted stop
  int mg_last_iters_ = 0;
  void rbgs_iters(const DevTopo& T, const float* div, float* p, int first_it, int n, const SolveRequest& r, Timing& tm,
                  hipStream_t s);
  SolveResult solve_cg(const DevTopo& T, const float* div, float* p, const SolveRequest& r, Timing& tm,
                       hipStream_t s);
  void mg_batch(const DevTopo& T, float* p, const struct CgScalars& S, int depth, int n, hipStream_t s);
  void mg_iteration(const DevTopo& T, float* p, const struct CgScalars& S, int depth, int it, hipStream_t s,
                    Timing* tm = nullptr, bool fold_test = false, bool own_test = true);

  // The device operations the loops of hns_solve_loop.hpp drive (one solve's arrays and stream).
  struct RbgsDev {
    Solvers* sol;
    const DevTopo& T;
    const float* div;
    float* p;
    const SolveRequest& r;
    Timing& tm;
    hipStream_t s;
    void sweeps(int first_it, int n);
    double residual();
  };
  struct CgDev {
    Solvers* sol;
    const DevTopo& T;
    const float* div;
    float* p;
    const SolveRequest& r;
    const struct CgScalars& S;
    Timing& tm;
    hipStream_t s;
    void replace_periodic(int it);
    void iteration(int it);
    CgReport read();
    void replace_forced();
  };
  struct FcgDev {
    Solvers* sol;
    const DevTopo& T;
    const float* div;
    float* p;
    const SolveRequest& r;
    const struct CgScalars& S;
    int depth;
    bool mg;
    Timing& tm;
    hipStream_t s;
    void replace_periodic(int it);
    int batchable(int it, int left);
    void batch(int it, int n);
    void first_test();
    void iteration(int it);
    FcgReport end_batch();
    // the reads launch the request's on_stop work (the multigrid solve without per-kernel timing)
    bool continues() const { return mg && !tm.enabled && static_cast<bool>(r.on_stop); }
  };
};

}  // namespace hns
