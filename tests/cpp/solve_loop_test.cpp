// CPU test of the pressure solves' host control (hnanosolver_amd/csrc/hns_solve_loop.hpp): the
// same loop templates the library runs, driven by simulated devices that follow the kernels' flag
// and counter rules (kernels_cg.hip): every kernel of an iteration returns at once while the stop
// flag is up; the flexible-CG update raises the flag without advancing the counter when d . A d is
// not positive; the re-test replaces r by b - A x and lowers the flag unless the replaced residual
// passes. Each simulated device counts its launches and aborts the run past a budget, so a loop that
// does not end fails here instead of hanging. Prints "solve loops: ok" on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <limits>
#include <string>

#include "../../hnanosolver_amd/csrc/hns_solve_loop.hpp"

using namespace hns;

namespace {

int g_failures = 0;
void check(bool ok, const std::string& what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what.c_str());
    ++g_failures;
  }
}

// Launch budget of one run: past the limit the simulated device throws, which ends a loop that
// would not end by itself (the run then reports it as blown).
struct BudgetBlown {};
struct Budget {
  long used = 0, limit = 0;
  bool blown = false;
  bool take() {
    if (++used > limit) {
      blown = true;
      throw BudgetBlown{};
    }
    return true;
  }
};

const double kNaN = std::numeric_limits<double>::quiet_NaN();
const double kInf = std::numeric_limits<double>::infinity();

// A residual history: rr_rec(k) is the recurrence r . r after k updates, rr_true(k) what b - A x
// gives at that point; breakdown(k): the update of iteration k finds d . A d <= 0 (or NaN).
struct Problem {
  double bb = 1.0;
  double tol2 = 1e-10;
  std::function<double(int)> rr_rec = [](int k) { return std::pow(0.5, k); };
  std::function<double(int)> rr_true = [](int k) { return std::pow(0.5, k); };
  std::function<bool(int)> breakdown = [](int) { return false; };
  bool stuck = false;  // every kernel returns at once without touching the flag or the counter
};

// ---- flexible CG device ---------------------------------------------------------------------------
struct FakeFcg {
  const Problem& P;
  Budget budget;
  int it = 0;  // the device iteration counter
  int flag = 0, stop_it = 0;
  double rr = 0.0, last_rr = 0.0, last_bb = 0.0;
  FakeFcg(const Problem& p, long limit) : P(p) { budget.limit = limit; }

  void test(bool advance) {  // k_fcg_test
    if (flag || P.stuck) return;
    if (advance) ++it;
    last_rr = rr;
    last_bb = P.bb;
    if (rr <= P.tol2 * P.bb) {
      flag = 1;
      stop_it = it;
    }
  }
  std::string replaced;  // host iterations of the periodic replacements, "64 128 ..."
  bool batch_over_replacement = false;
  void replace_periodic(int h) {
    budget.take();
    replaced += std::to_string(h) + " ";
    if (P.stuck || flag) return;
    rr = P.rr_true(it);
    test(false);
  }
  int batchable(int, int left) { return left; }  // as the multigrid solve: any run of iterations
  void batch(int h, int n) {
    for (int k = 1; k < n; ++k)  // a replacement point strictly inside a replayed batch
      if ((h + k) % 64 == 0) batch_over_replacement = true;
    for (int k = 0; k < n; ++k) iteration(h + k);
  }
  void first_test() {
    rr = P.rr_rec(0);
    test(false);
  }
  void iteration(int) {
    budget.take();
    if (P.stuck || flag) return;
    if (P.breakdown(it)) {  // k_fcg_update: sigma <= 0 raises the flag, counter unchanged
      flag = 1;
      stop_it = it;
      return;
    }
    rr = P.rr_rec(it + 1);
    test(true);
  }
  FcgReport end_batch() {
    if (flag && !P.stuck) {  // re-test on b - A x
      rr = P.rr_true(it);
      flag = 0;
      test(false);
    }
    return FcgReport{flag, stop_it, it, last_rr, last_bb};
  }
};

LoopResult run_fcg(const Problem& P, int cap, long limit, bool* blown, std::string* replaced = nullptr,
                   int batch0 = 8) {
  FakeFcg d(P, limit);
  LoopResult r;
  try {
    r = fcg_loop(d, cap, batch0, 4, 64);
  } catch (const BudgetBlown&) {
  }
  *blown = d.budget.blown;
  check(!d.batch_over_replacement, "fcg: no replayed batch contains a replacement point");
  if (replaced) *replaced = d.replaced;
  return r;
}

// ---- CG device ------------------------------------------------------------------------------------
struct FakeCg {
  const Problem& P;
  Budget budget;
  int flag = 0, stop_it = 0;
  double rr = 1.0;
  bool replaced = false;
  int rec_base = 0;
  FakeCg(const Problem& p, long limit) : P(p) {
    budget.limit = limit;
    rr = P.rr_rec(0);
  }
  void replace_periodic(int it) {
    budget.take();
    if (flag) return;
    rr = P.rr_true(it);
  }
  void iteration(int it) {  // k_cg_spmv: stop test on r . r, then k_cg_update: breakdown or step
    budget.take();
    if (flag) return;
    if (rr <= P.tol2 * P.bb) {
      flag = 1;
      stop_it = it;
      return;
    }
    if (P.breakdown(it)) {
      flag = 1;
      stop_it = it;
      return;
    }
    rr = P.rr_rec(it + 1);
  }
  CgReport read() { return CgReport{flag, stop_it}; }
  void replace_forced() {
    budget.take();
    rr = P.rr_true(stop_it);
    flag = 0;
  }
};

LoopResult run_cg(const Problem& P, int cap, long limit, bool* blown) {
  FakeCg d(P, limit);
  LoopResult r;
  try {
    r = cg_loop(d, cap, 8, 128, 64);
  } catch (const BudgetBlown&) {
  }
  *blown = d.budget.blown;
  return r;
}
template <class Dev>
LoopResult run_rbgs(Dev& d, double tol, int cap, int chunk, int done, double rel0, double* rel) {
  LoopResult r;
  try {
    r = rbgs_tol_loop(d, tol, cap, chunk, done, rel0, rel);
  } catch (const BudgetBlown&) {
  }
  return r;
}

// ---- RBGS to tol ----------------------------------------------------------------------------------
struct FakeRbgs {
  std::function<double(int)> rel;
  Budget budget;
  int done = 0;
  void sweeps(int, int n) {
    budget.take();
    done += n;
  }
  double residual() { return rel(done); }
};

void fcg_cases() {
  bool blown = false;
  {  // converges: r.r = 2^-k passes tol2 = 1e-10 first at k = 34
    Problem P;
    const LoopResult r = run_fcg(P, 1000, 10000, &blown);
    check(!blown && r.end == kLoopStopped && r.iterations == 34, "fcg: converges at 34");
  }
  {  // the recurrence residual drifts below the true one: stops are re-tested and rejected until
     // the true residual passes (k = 40)
    Problem P;
    P.rr_rec = [](int k) { return std::pow(0.5, k); };
    P.rr_true = [](int k) { return std::pow(0.5, k) * 64.0; };
    const LoopResult r = run_fcg(P, 1000, 10000, &blown);
    check(!blown && r.end == kLoopStopped && r.iterations >= 40, "fcg: drifted recurrence stops on b - A x");
  }
  {  // a breakdown at iteration 7 that every retry repeats (a degenerate preconditioner output):
     // the loop ends after one batch without progress, well inside the budget
    Problem P;
    P.breakdown = [](int k) { return k == 7; };
    const LoopResult r = run_fcg(P, 100000, 200, &blown);
    check(!blown && r.end == kLoopBreakdown && r.iterations == 7, "fcg: repeated breakdown ends the solve");
  }
  {  // NaN from the 5th update on: r.r is NaN, every update breaks down, the last test reads NaN
    Problem P;
    P.rr_rec = [](int k) { return k >= 5 ? kNaN : std::pow(0.5, k); };
    P.rr_true = P.rr_rec;
    P.breakdown = [](int k) { return k >= 5; };
    const LoopResult r = run_fcg(P, 100000, 200, &blown);
    check(!blown && r.end == kLoopNonFinite, "fcg: NaN residual ends the solve");
  }
  {  // NaN right-hand side: ||b||^2 is NaN from the first test
    Problem P;
    P.bb = kNaN;
    P.rr_rec = [](int) { return kNaN; };
    P.rr_true = P.rr_rec;
    P.breakdown = [](int) { return true; };
    const LoopResult r = run_fcg(P, 100000, 200, &blown);
    check(!blown && r.end == kLoopNonFinite && r.iterations == 0, "fcg: NaN right-hand side");
  }
  {  // an infinite right-hand side: r.r <= tol2 ||b||^2 holds as inf <= inf, so the device stops;
     // the stop's r.r / ||b||^2 is not finite and the solver reports it (solvers.hip)
    Problem P;
    P.bb = kInf;
    P.rr_rec = [](int) { return kInf; };
    P.rr_true = P.rr_rec;
    const LoopResult r = run_fcg(P, 100000, 200, &blown);
    check(!blown && r.end == kLoopStopped && !std::isfinite(r.rr / r.bb), "fcg: Inf right-hand side");
  }
  {  // a device whose counter never moves and whose flag stays down
    Problem P;
    P.stuck = true;
    const LoopResult r = run_fcg(P, 100000, 200, &blown);
    check(!blown && r.end == kLoopBreakdown && r.iterations == 0, "fcg: stuck counter ends the solve");
  }
  {  // slow convergence: the cap ends it
    Problem P;
    P.rr_rec = [](int k) { return 1.0 / (1.0 + k); };
    P.rr_true = P.rr_rec;
    const LoopResult r = run_fcg(P, 50, 1000, &blown);
    check(!blown && r.end == kLoopCapped && r.iterations == 50, "fcg: capped at 50");
  }
  {  // a long solve with first batches of odd lengths: replaced before iterations 64, 128 and 192,
     // never inside a batch, capped at 200
    for (int b0 : {4, 9, 13, 63}) {
      Problem P;
      P.rr_rec = [](int k) { return 1.0 / (1.0 + k); };
      P.rr_true = P.rr_rec;
      std::string replaced;
      const LoopResult r = run_fcg(P, 200, 10000, &blown, &replaced, b0);
      check(!blown && r.end == kLoopCapped && r.iterations == 200 && replaced == "64 128 192 ",
            "fcg: periodic replacements at 64, 128, 192 with first batch " + std::to_string(b0) + " (got " +
                replaced + ")");
    }
  }
  {  // zero right-hand side: stops before the first iteration
    Problem P;
    P.bb = 0.0;
    P.rr_rec = [](int) { return 0.0; };
    P.rr_true = P.rr_rec;
    const LoopResult r = run_fcg(P, 1000, 100, &blown);
    check(!blown && r.end == kLoopStopped && r.iterations == 0, "fcg: zero right-hand side");
  }
  {  // breakdowns at 3 and 9 that the replaced residual clears once (round-off): progress resumes
    Problem P;
    bool seen3 = false, seen9 = false;
    P.breakdown = [&](int k) {
      bool& seen = k == 3 ? seen3 : seen9;
      if ((k == 3 || k == 9) && !seen) return seen = true;
      return false;
    };
    const LoopResult r = run_fcg(P, 1000, 10000, &blown);
    check(!blown && r.end == kLoopStopped && r.iterations == 34, "fcg: one-off breakdowns");
  }
}

void cg_cases() {
  bool blown = false;
  {
    Problem P;
    const LoopResult r = run_cg(P, 1000, 10000, &blown);
    check(!blown && r.end == kLoopStopped && r.iterations == 34, "cg: converges at 34");
  }
  {  // NaN from iteration 5: breakdown stops, accepted once the replaced residual repeats it
    Problem P;
    P.rr_rec = [](int k) { return k >= 5 ? kNaN : std::pow(0.5, k); };
    P.rr_true = P.rr_rec;
    P.breakdown = [](int k) { return k >= 5; };
    const LoopResult r = run_cg(P, 100000, 1000, &blown);
    check(!blown && r.end == kLoopStopped && r.iterations == 5, "cg: NaN ends at the breakdown");
  }
  {
    Problem P;
    P.rr_rec = [](int k) { return std::pow(0.5, k); };
    P.rr_true = [](int k) { return std::pow(0.5, k) * 64.0; };
    const LoopResult r = run_cg(P, 1000, 10000, &blown);
    check(!blown && r.end == kLoopStopped && r.iterations >= 40, "cg: drifted recurrence");
  }
  {
    Problem P;
    P.rr_rec = [](int k) { return 1.0 / (1.0 + k); };
    P.rr_true = P.rr_rec;
    const LoopResult r = run_cg(P, 300, 10000, &blown);
    check(!blown && r.end == kLoopCapped && r.iterations == 300, "cg: capped at 300");
  }
}

void rbgs_cases() {
  {
    FakeRbgs d{[](int k) { return std::pow(0.9, k); }, Budget{0, 100000}};
    double rel = 0.0;
    const LoopResult r = run_rbgs(d, 1e-5, 100000, 50, 0, d.residual(), &rel);
    check(r.end == kLoopStopped && rel <= 1e-5 && r.iterations % 50 == 0, "rbgs: converges");
  }
  {  // a NaN residual ends the loop at once instead of running to the cap
    FakeRbgs d{[](int) { return kNaN; }, Budget{0, 10}};
    double rel = 0.0;
    const LoopResult r = run_rbgs(d, 1e-5, 200000, 50, 1, kNaN, &rel);
    check(!d.budget.blown && r.end == kLoopNonFinite && r.iterations == 0, "rbgs: NaN residual");
  }
  {
    FakeRbgs d{[](int k) { return k < 100 ? 1.0 : kNaN; }, Budget{0, 10}};
    double rel = 0.0;
    const LoopResult r = run_rbgs(d, 1e-5, 200000, 50, 1, 1.0, &rel);
    check(!d.budget.blown && r.end == kLoopNonFinite && r.iterations == 100, "rbgs: NaN after 100");
  }
  {
    FakeRbgs d{[](int) { return 0.5; }, Budget{0, 1000}};
    double rel = 0.0;
    const LoopResult r = run_rbgs(d, 1e-5, 120, 50, 0, 0.5, &rel);
    check(r.end == kLoopCapped && r.iterations == 120, "rbgs: capped");
  }
}

}  // namespace

int main() {
  fcg_cases();
  cg_cases();
  rbgs_cases();
  if (g_failures) {
    std::printf("%d failure(s)\n", g_failures);
    return 1;
  }
  std::printf("solve loops: ok\n");
  return 0;
}
