// Host control of the to-tolerance pressure solves: when to launch, when to read the device's stop
// flag, and when a solve ends. No reference counterpart (the reference runs a fixed red-black
// Gauss-Seidel count, src/Cuda/HNanoSolver.cu:364-373).
//
// Host-only and free of HIP types: each loop is a template over a device-operations policy, so the
// solver launches the real kernels through it and a CPU test drives the same loops with a simulated
// device (stuck iteration counters, repeated breakdowns, non-finite residuals). Every loop ends:
//   - CG: the host iteration count only grows between stop reports, and a stop reported at the
//     iteration whose residual was just replaced is final;
//   - flexible CG: a batch after which the device iteration counter has not moved past the batch's
//     first iteration is a breakdown of that iteration (the update found d . A d <= 0 or not a
//     number, and the replaced residual did not pass), which the same launches would repeat; a
//     non-finite r . r or ||b||^2 ends the solve at once;
//   - RBGS to a tolerance: stops at the cap or on a non-finite residual.
#pragma once
#include <algorithm>
#include <cmath>

namespace hns {

// How a loop ended.
enum LoopEnd {
  kLoopStopped = 0,    // the device accepted a stop (a converged residual, or CG's final breakdown stop)
  kLoopCapped = 1,     // the iteration cap was reached
  kLoopBreakdown = 2,  // flexible CG: the device counter made no progress over a batch
  kLoopNonFinite = 3,  // a residual norm or ||b||^2 is not finite
};

struct LoopResult {
  int iterations = 0;
  LoopEnd end = kLoopCapped;
  double rr = -1.0, bb = -1.0;  // flexible CG: r . r and ||b||^2 of the device's last stop test
};

// ---- conjugate gradients ------------------------------------------------------------------------
// What the host reads after a batch of CG iterations: the stop flag and the iteration it was
// raised at (a converged recurrence residual or an update breakdown).
struct CgReport {
  int stop = 0;
  int stop_it = 0;
};

// Dev provides
//   void replace_periodic(int it);  residual replacement r = b - A x before iteration it
//   void iteration(int it);         one CG iteration (product + update)
//   CgReport read();                synchronises and reads the stop flag
//   void replace_forced();          r = b - A x and the stop flag cleared (re-testing a stop)
// batch0/batch_max: iterations between reads (the batch doubles up to batch_max while no stop is
// raised). A stop is accepted only at the iteration whose residual was just replaced.
template <class Dev>
LoopResult cg_loop(Dev& d, int cap, int batch0, int batch_max, int replace_every) {
  LoopResult out;
  int it = 0, batch = std::max(1, batch0), replaced_at = 0;
  while (it < cap) {
    const int nb = std::min(batch, cap - it);
    for (int k = 0; k < nb; ++k, ++it) {
      if (it > 0 && it % replace_every == 0 && replaced_at != it) {
        d.replace_periodic(it);
        replaced_at = it;
      }
      d.iteration(it);
    }
    const CgReport r = d.read();
    if (r.stop) {
      // the device stops at an iteration at or after the batch's first: `it` never decreases
      it = std::max(r.stop_it, it - nb);
      if (it == replaced_at) {
        out.iterations = it;
        out.end = kLoopStopped;
        return out;
      }
      d.replace_forced();
      replaced_at = it;
      batch = std::max(1, batch0);
      continue;
    }
    batch = std::min(batch * 2, std::max(1, batch_max));
  }
  out.iterations = it;
  out.end = kLoopCapped;
  return out;
}

// ---- flexible (preconditioned) CG ---------------------------------------------------------------
// What the host reads after a batch: the stop flag (standing after its re-test on b - A x), the
// iteration it stopped at, the device iteration counter, and the last test's r . r and ||b||^2.
struct FcgReport {
  int stop = 0;
  int stop_it = 0;
  int dev_it = 0;
  double rr = 0.0, bb = 0.0;
};

// Dev provides
//   void replace_periodic(int it);     residual replacement before iteration it, and its stop test
//   int batchable(int it, int left);   how many of the next `left` iterations from it go as one
//                                      replayed batch (below 2: they are launched one by one)
//   void batch(int it, int n);         n iterations from it as one batch
//   void first_test();                 the stop test before iteration 0 (records ||b||^2)
//   void iteration(int it);            one iteration (preconditioner, direction, product, update,
//                                      stop test advancing the device counter)
//   FcgReport end_batch();             re-test of a raised stop, then synchronise and read
// Between two reads the loop runs `batch` iterations (batch0 first, then batch_max), cut at the
// periodic replacement points, so a replayed batch never contains one.
template <class Dev>
LoopResult fcg_loop(Dev& d, int cap, int batch0, int batch_max, int replace_every) {
  LoopResult out;
  int it = 0, batch = std::max(1, batch0), replaced_at = 0;
  while (it < cap) {
    const int start = it;
    const int nb = std::min(batch, cap - it);
    for (int k = 0; k < nb;) {
      if (it == 0) d.first_test();
      if (it > 0 && it % replace_every == 0 && replaced_at != it) {
        d.replace_periodic(it);
        replaced_at = it;
      }
      const int n = std::min(d.batchable(it, std::min(nb - k, replace_every - it % replace_every)), nb - k);
      if (n >= 2) {
        d.batch(it, n);
        it += n;
        k += n;
        continue;
      }
      d.iteration(it);
      ++it;
      ++k;
    }
    const FcgReport r = d.end_batch();
    out.rr = r.rr;
    out.bb = r.bb;
    if (r.stop) {
      out.iterations = r.stop_it;
      out.end = kLoopStopped;
      return out;
    }
    if (!std::isfinite(r.rr) || !std::isfinite(r.bb)) {
      out.iterations = r.dev_it;
      out.end = kLoopNonFinite;
      return out;
    }
    if (r.dev_it <= start) {
      // no update completed in the whole batch: iteration `start` broke down on the residual the
      // re-test just replaced, and the same deterministic launches would break down again
      out.iterations = r.dev_it;
      out.end = kLoopBreakdown;
      return out;
    }
    it = std::min(r.dev_it, cap);
    batch = std::max(1, batch_max);
  }
  out.iterations = it;
  out.end = kLoopCapped;
  return out;
}

// ---- red-black Gauss-Seidel to a tolerance --------------------------------------------------------
// Dev provides
//   void sweeps(int first_it, int n);  n iterations continuing from iteration first_it
//   double residual();                 relative residual (synchronises)
// `done` iterations have already run; the loop continues in chunks until rel <= tol, the cap, or a
// non-finite residual.
template <class Dev>
LoopResult rbgs_tol_loop(Dev& d, double tol, int cap, int chunk, int done, double rel, double* rel_out) {
  LoopResult out;
  chunk = std::max(1, chunk);
  int n_run = 0;
  while (std::isfinite(rel) && rel > tol && n_run < cap) {
    const int n = std::min(chunk, cap - n_run);
    d.sweeps(done + n_run, n);
    n_run += n;
    rel = d.residual();
  }
  *rel_out = rel;
  out.iterations = n_run;
  out.end = !std::isfinite(rel) ? kLoopNonFinite : (rel <= tol ? kLoopStopped : kLoopCapped);
  return out;
}

}  // namespace hns
