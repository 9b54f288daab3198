// Pressure-solve driver: the reference's fixed-count red-black Gauss-Seidel loop
// (src/Cuda/HNanoSolver.cu:364-373), and the to-tolerance variants selected through
// hns_solve_options (no reference counterpart).
#include <algorithm>
#include <cmath>

#include "hns_solvers.hpp"

namespace hns {

struct CgState {};

Solvers::Solvers() = default;

Solvers::~Solvers() {
  if (h_red_) (void)hipHostFree(h_red_);
  if (e0_) (void)hipEventDestroy(e0_);
  if (e1_) (void)hipEventDestroy(e1_);
  delete cg_;
}

void Solvers::init(int32_t L, size_t N) {
  L_ = L;
  N_ = N;
  partial_.alloc(2 * (size_t)L);
  red_.alloc(2);
  if (!h_red_) HNS_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_red_), 2 * sizeof(double), hipHostMallocDefault));
  if (!e0_) HNS_HIP(hipEventCreate(&e0_));
  if (!e1_) HNS_HIP(hipEventCreate(&e1_));
}

void Solvers::rbgs_iters(const DevTopo& T, const float* div, float* p, int first_it, int n, const SolveRequest& r,
                         Timing& tm, hipStream_t s) {
  for (int it = first_it; it < first_it + n; ++it) {
    tm.begin(HNS_K_RBGS, s);
    launch_rbgs(T, div, p, r.dx, 0, r.omega, it == 0, s);
    tm.end(s);
    tm.begin(HNS_K_RBGS, s);
    launch_rbgs(T, div, p, r.dx, 1, r.omega, it == 0, s);
    tm.end(s);
  }
}

double Solvers::rel_residual(const DevTopo& T, const float* div, const float* p, float dx, Timing& tm,
                             hipStream_t s) {
  tm.begin(HNS_K_RESIDUAL, s);
  launch_residual(T, p, div, dx, partial_.p, red_.p, s);
  tm.end(s);
  HNS_HIP(hipMemcpyAsync(h_red_, red_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HNS_HIP(hipStreamSynchronize(s));
  const double r2 = h_red_[0], d2 = h_red_[1];
  return d2 > 0.0 ? std::sqrt(r2 / d2) : (r2 > 0.0 ? INFINITY : 0.0);
}

SolveResult Solvers::solve(const DevTopo& T, const float* div, float* p, int K, const SolveRequest& r, Timing& tm,
                           hipStream_t s) {
  SolveResult res;
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
      int done = 0;
      HNS_HIP(hipMemsetAsync(p, 0, N_ * sizeof(float), s));
      double rel = rel_residual(T, div, p, r.dx, tm, s);
      while (rel > (double)r.tol && done < cap) {
        const int n = std::min(chunk, cap - done);
        rbgs_iters(T, div, p, done, n, r, tm, s);
        done += n;
        rel = rel_residual(T, div, p, r.dx, tm, s);
      }
      res.iterations = done;
      res.rel_residual = rel;
      break;
    }
    case HNS_SOLVER_CG:
    case HNS_SOLVER_PCG_LEARNED:
      res = solve_cg(T, div, p, r, tm, s);
      break;
    default:
      fail(HNS_ERR_INVALID_ARGUMENT, "unknown pressure solver");
  }
  if (r.want_residual) {
    HNS_HIP(hipEventRecord(e1_, s));
    HNS_HIP(hipEventSynchronize(e1_));
    float ms = 0.f;
    HNS_HIP(hipEventElapsedTime(&ms, e0_, e1_));
    res.ms = ms;
  }
  return res;
}

SolveResult Solvers::solve_cg(const DevTopo&, const float*, float*, const SolveRequest&, Timing&, hipStream_t) {
  fail(HNS_ERR_INVALID_ARGUMENT, "conjugate-gradient solvers are not built yet");
}

}  // namespace hns
