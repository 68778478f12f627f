// Control logic of the minimizers, host-only and independent of HIP: every
// vector operation goes through a backend `Ops`, which the library implements
// on the device (deft_amd.hip) and the tests implement on the host
// (tests/host/minimizer_harness.cpp).  The backend provides
//
//   int  plan_prec() const;                       FFT precision of the plan
//   int  energy(int prec, double *E);             energy at the current point
//                                                 (pending moves applied)
//   int  gradient(int prec, bool have_old, int pgrad_kind,
//                 double *E, double *pgg, double *pgo);
//                                                 grad and pgrad at the current
//                                                 point with pg.g and pg.oldgrad
//   int  cg_update(double beta, bool have_dir, double *gd);
//                                                 d = -pg + beta d (or -pg),
//                                                 oldgrad = grad, gd = grad.d
//   int  cg_reset(double *gd);                    d = -grad, gd = grad.d
//   void move(double s);                          x += s d
//   double newton_scale() const;                  kT/dV (see grad_error below)
//   bool sk_available();                          whether gradient() takes MIN_PGRAD_SK
//
// Two minimizers live here:
//   - the new API's Minimize (src/new/Minimize.cpp:9-313, src/new/Minimize.h:13-199)
//     with this build's FP32 -> FP64 precision ladder around it;
//   - the old API's ConjugateGradient / PreconditionedConjugateGradient with
//     QuadraticLineMinimizer, wrapped by Precision and MaxIter
//     (src/ConjugateGradient.cpp:59-171, src/QuadraticLineMinimizer.cpp:27-215,
//     src/Precision.cpp:23-156, src/Minimizer.cpp, src/Minimizer.h:16-81).
// The reference's old API recomputes energy() on every call; the energy of a
// point is deterministic, so both minimizers here evaluate a point once and
// reuse the value until the point moves.  The decisions, and so the iterates,
// are unchanged by that.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>

#include "../../include/deft_amd.h"

namespace deft {

#define DEFT_MIN_TRY(expr)           \
  do {                               \
    const int rc_ = (expr);          \
    if (rc_ != DEFT_OK) return rc_;  \
  } while (0)

// kinds of preconditioned gradient the backend computes (the device's PgradKind)
enum MinPgrad : int { MIN_PGRAD_NONE = 0, MIN_PGRAD_NEW = 1, MIN_PGRAD_OLD = 2, MIN_PGRAD_SK = 3 };

inline bool min_better(double a, double b) { return a < b || b != b; }

// ===========================================================================
// new API Minimize + the precision ladder
// ===========================================================================
constexpr int kConfirmIterations = 3;
constexpr int kMaxPhaseHalvings = 12;
constexpr int kMaxCollapseRestarts = 3;
constexpr int kMaxGradientRestarts = 4;
// Along FP64 minimizations of hard walls and solutes (32^3 to 64^3, eta 0.3
// and 0.4) the Newton estimate below lay between 1.2 and 10 times the energy
// still to be gained, never under it: a factor of 10 bounds the remaining
// error by 10x the target precision and passes every ordinary stop, whose
// true error is far below the target.
constexpr double kDefaultGradCheck = 10.0;

struct MinState {
  int iter = 0, maxiter = 10000000, miniter = 6;
  double precision = 0.0, relative_precision = 1e-15;
  bool use_precond = false;
  int precond_kind = MIN_PGRAD_NEW;  // with use_precond: grad/n (the reference's) or MIN_PGRAD_SK
  double step = 0.1, orig_stepsize = 0.1, oldgradsqr = 1.0;
  double deltaE = 0.0, dEdn = 0.0, log_avg = 0.0, error_estimate = 0.0;
  bool have_dir = false, have_old = false;
  bool E_valid = false;
  double E = 0.0;
  bool G_valid = false;
  double pgg = 0.0, pgo = 0.0;
  int num_energy = 0, num_grad = 0;
  int stop = DEFT_STOP_NONE;
  // mixed precision: run FP32 FFTs until this relative precision (or until the
  // FP32 line search stalls), then continue with FP64 FFTs; 0 disables
  double switch_rp = 0.0;
  int polish_grad_prec = DEFT_FFT_FP32;  // gradient precision after the switch
  int e_prec = -1, g_prec = -1;          // evaluation precisions (-1: the plan's)
  int fp32_iters = -1;      // iterations done in phase 0 (-1: no switch yet)
  int iter_at_phase2 = -1;  // iteration count when phase 2 began (-1: not reached)
  int phase = 0;            // precision phase of the mixed ladder
  int confirm_until = 0;    // the final phase runs at least until this iteration
  // fused line search: the trial point s2 of the quadratic step gets energy and
  // gradient from one evaluation (when both use one FFT precision), so an
  // accepted s2 needs no second evaluation at the start of the next iteration;
  // -1: on exactly when the mixed-precision ladder is on
  int fuse_ls = -1;
  bool s2_accepted = true;  // the last quadratic step kept its trial point s2
  double good_step = 0.0;   // step of the last iteration that clearly lowered the energy
  int collapse_restarts = 0;  // final-phase restarts after a collapsed last step
  bool restart_met = false;   // the iteration before a collapse restart met the precision
  // gradient confirmation of the ladder's final verdict: the stop is accepted
  // only if the Newton estimate of the remaining energy error from the final
  // gradient, grad_error = (kT/dV) pg.g / 2 (pgrad = grad/n), is below
  // grad_check times the target precision; 0 disables, < 0 means
  // kDefaultGradCheck
  double grad_check = -1.0;
  int grad_restarts = 0;      // direction restarts after a failed gradient confirmation
  double last_grad_error = -1.0;  // grad_error of the last confirmation (-1: none)
};

inline int min_energy_prec(const MinState &m, int plan_prec) { return m.e_prec >= 0 ? m.e_prec : plan_prec; }
inline int min_grad_prec(const MinState &m, int plan_prec) { return m.g_prec >= 0 ? m.g_prec : plan_prec; }

inline void min_invalidate(MinState &m) {
  m.E_valid = false;
  m.G_valid = false;
}

template <class Ops>
int min_energy(MinState &m, Ops &ops, double *E) {
  if (!m.E_valid) {
    DEFT_MIN_TRY(ops.energy(min_energy_prec(m, ops.plan_prec()), &m.E));
    m.E_valid = true;
    m.num_energy++;
  }
  *E = m.E;
  return DEFT_OK;
}

// Minimize::pgrad (src/new/Minimize.h:136-165): with preconditioning the
// energy, grad and pgrad come from one evaluation; without, pgrad = grad.
// When gradients run at a lower FFT precision than energies, the evaluation's
// energy is not used: the cached energy of the same Veff is kept (or computed
// at the energy precision), so the line search only ever compares energies of
// one precision.
template <class Ops>
int min_pgrad(MinState &m, Ops &ops) {
  if (m.G_valid) return DEFT_OK;
  const int pp = ops.plan_prec();
  const bool same = min_energy_prec(m, pp) == min_grad_prec(m, pp);
  const double Ekeep = m.E;
  const bool keep = !same && m.E_valid;
  if (m.use_precond) min_invalidate(m);
  double Eg;
  DEFT_MIN_TRY(ops.gradient(min_grad_prec(m, pp), m.have_old, m.use_precond ? m.precond_kind : MIN_PGRAD_NONE, &Eg,
                            &m.pgg, &m.pgo));
  if (same && (m.use_precond || !m.E_valid)) {
    m.E = Eg;
    m.E_valid = true;
  } else if (keep) {
    m.E = Ekeep;
    m.E_valid = true;
  }
  if (m.use_precond && same) m.num_energy++;
  m.num_grad++;
  m.G_valid = true;
  return DEFT_OK;
}

inline int min_stop_with(MinState &m, int reason, int *more) {
  m.stop = reason;
  *more = 0;
  return DEFT_OK;
}

// One Minimize::improve_energy (src/new/Minimize.cpp:9-313).
template <class Ops>
int improve_energy(MinState &m, Ops &ops, int verbose, int *more) {
  m.iter++;
  if (m.iter >= m.maxiter) return min_stop_with(m, DEFT_STOP_MAXITER, more);
  double E0;
  DEFT_MIN_TRY(min_energy(m, ops, &E0));
  const double old_deltaE = m.deltaE;
  if (E0 != E0) return min_stop_with(m, DEFT_STOP_NAN_ENERGY, more);

  // Polak-Ribiere direction with automatic reset
  DEFT_MIN_TRY(min_pgrad(m, ops));
  min_invalidate(m);
  double beta = m.have_old ? (m.pgg - m.pgo) / m.oldgradsqr : m.pgg / m.oldgradsqr;
  if (beta < 0 || beta != beta || m.oldgradsqr == 0) beta = 0;
  m.oldgradsqr = m.pgg;
  double gdotd;
  DEFT_MIN_TRY(ops.cg_update(beta, m.have_dir, &gdotd));
  m.have_dir = true;
  m.have_old = true;
  if (gdotd > 0) {
    DEFT_MIN_TRY(ops.cg_reset(&gdotd));
    if (verbose) std::printf("\t\treset to gradient\n");
  }

  // quadratic line minimization
  const double slope = -gdotd;
  if (E0 >= std::numeric_limits<double>::max() || E0 <= -std::numeric_limits<double>::max())
    return min_stop_with(m, DEFT_STOP_INFINITE_ENERGY, more);
  if (slope != slope) return min_stop_with(m, DEFT_STOP_NAN_SLOPE, more);
  if (slope == 0) return min_stop_with(m, DEFT_STOP_ZERO_SLOPE, more);
  if (m.step == m.orig_stepsize && std::fabs(E0) > 1e-200 && std::fabs(slope) > 1e-100)
    m.step = 0.001 * std::fabs(E0) / slope;
  if (slope * m.step < 0) m.step *= -1;
  double step1 = m.step;
  ops.move(step1);
  double Etried = -137, E;
  // Below the last precision phase the halving is capped: a line search that
  // needs more than kMaxPhaseHalvings halvings has reached the round-off of the
  // phase's energies, and the phase ends where the iteration began.
  const bool capped = m.switch_rp > 0 && m.phase < 2;
  int halvings = 0;
  while (true) {
    step1 *= 0.5;
    ops.move(-step1);
    min_invalidate(m);
    DEFT_MIN_TRY(min_energy(m, ops, &E));
    if (E == Etried) break;
    Etried = E;
    if (!min_better(E0, E)) break;
    if (capped && ++halvings > kMaxPhaseHalvings) {
      ops.move(-step1);
      min_invalidate(m);
      m.deltaE = 0;
      return min_stop_with(m, DEFT_STOP_NO_CHANGE, more);
    }
  }
  double E1;
  DEFT_MIN_TRY(min_energy(m, ops, &E1));
  const double curvature = 2.0 * (E1 - E0 + step1 * slope) / (step1 * step1);
  if (curvature <= 0.0) {
    if (min_better(E0, E1)) {
      min_invalidate(m);
      ops.move(-step1);
      return min_stop_with(m, DEFT_STOP_NEGATIVE_CURVATURE, more);
    }
    double Ebest = E0;
    while (true) {
      DEFT_MIN_TRY(min_energy(m, ops, &E));
      if (!(E <= Ebest)) break;
      Ebest = E;
      ops.move(step1);
      min_invalidate(m);
      step1 *= 2;
    }
    step1 *= 0.5;
    ops.move(-step1);
    m.step = step1;
    min_invalidate(m);
  } else if (E1 == E0) {
    do {
      min_invalidate(m);
      ops.move(-step1);
      step1 *= 2;
      ops.move(step1);
      DEFT_MIN_TRY(min_energy(m, ops, &E));
      // (the reference's loop has no exit when step1 is 0: a zero step never
      // changes the energy; stop doubling there, and at an infinite step)
    } while (E == E0 && step1 != 0 && std::isfinite(step1));
    if (E > E0) {
      ops.move(-step1);
      min_invalidate(m);
      step1 = 0;
    } else if (E != E) {
      ops.move(-step1);
      min_invalidate(m);
      step1 = 0;
    }
  } else {
    const double step2 = slope / curvature;
    m.step = step2;
    min_invalidate(m);
    ops.move(step2 - step1);
    double E2;
    const int pp = ops.plan_prec();
    // fused only while trial points are being kept: a rejected s2 wastes its gradient
    const bool fuse = (m.fuse_ls > 0 || (m.fuse_ls < 0 && m.switch_rp > 0)) && m.s2_accepted &&
                      min_energy_prec(m, pp) == min_grad_prec(m, pp);
    if (fuse) {
      // E(s2) from the evaluation that also yields the gradient the next
      // iteration starts from; the energies and the trajectory are unchanged
      const int ne = m.num_energy;
      DEFT_MIN_TRY(min_pgrad(m, ops));
      if (m.num_energy == ne) m.num_energy++;
      E2 = m.E;
    } else {
      DEFT_MIN_TRY(min_energy(m, ops, &E2));
    }
    m.s2_accepted = !(min_better(E1, E2) && min_better(E1, E0));
    if (!m.s2_accepted) {
      min_invalidate(m);
      ops.move(-(step2 - step1));
      m.step = step1;
    } else if (E2 > E0) {
      min_invalidate(m);
      ops.move(-step2);
      return min_stop_with(m, DEFT_STOP_ENERGY_ROSE, more);
    }
  }

  // convergence estimate (src/new/Minimize.cpp:263-312)
  double newE;
  DEFT_MIN_TRY(min_energy(m, ops, &newE));
  m.deltaE = newE - E0;
  const double w = 0.01;
  m.dEdn = (1 - w) * m.dEdn + w * m.deltaE;
  m.dEdn = std::max(m.deltaE, old_deltaE);
  const double new_log = (m.deltaE != 0 && old_deltaE != 0) ? std::log(std::fabs(m.deltaE / old_deltaE)) : 0.0;
  m.log_avg = (1 - w) * m.log_avg + w * new_log;
  const double ratio_avg = std::exp(m.log_avg);
  double error_guess = std::fabs(std::max(m.dEdn, m.deltaE) / m.log_avg);
  if (ratio_avg >= 1) error_guess = m.precision + std::fabs(newE);
  m.error_estimate = 2 * error_guess;
  // the step of the last iteration whose decrease stands clear of the round-off
  // of its energies: the next precision phase starts from it
  if (m.deltaE < -(min_energy_prec(m, ops.plan_prec()) == DEFT_FFT_FP32 ? 1e-6 : 1e-13) * std::fabs(newE))
    m.good_step = m.step;
  if (verbose)
    std::printf("iter %4d  E = %.15g  dE = %.3e  err = %.3e  step = %.3e  (#E %d, #G %d)\n", m.iter, newE, m.deltaE,
                m.error_estimate, m.step, m.num_energy, m.num_grad);
  const int miniter = std::max(m.miniter, m.confirm_until);
  if (m.deltaE == 0 && old_deltaE == 0) {
    m.stop = DEFT_STOP_NO_CHANGE;
    *more = m.iter < miniter;
    return DEFT_OK;
  }
  const double rp = (m.switch_rp > 0 && m.phase == 0) ? m.switch_rp : m.relative_precision;
  if (!(m.error_estimate < m.precision) && !(m.error_estimate < rp * std::fabs(newE))) {
    m.stop = DEFT_STOP_NONE;
    *more = 1;
    return DEFT_OK;
  }
  m.stop = DEFT_STOP_CONVERGED;
  *more = m.iter < miniter;
  return DEFT_OK;
}

// Restarts the CG direction from steepest descent at the step of the last
// iteration that clearly lowered the energy; without one (a start so close to
// the minimum that no iteration's decrease stood clear of the round-off), at
// the reference's first step 0.001 |E0| / slope (orig_stepsize selects it):
// the step the line searches left behind has collapsed to round-off size.
inline void min_restart_direction(MinState &m) {
  m.have_dir = false;
  m.have_old = false;
  m.oldgradsqr = 1.0;
  m.step = m.good_step > 0 ? m.good_step : m.orig_stepsize;
}

// Newton estimate of the energy still to be gained from the current point:
// with pgrad = grad/n and the ideal-gas Hessian n dV/kT of Veff near the
// minimum, grad . H^-1 . grad = (kT/dV) pg.g, and a quadratic model lies
// half of that above its minimum.
template <class Ops>
double grad_error(const MinState &m, const Ops &ops) {
  return 0.5 * ops.newton_scale() * m.pgg;
}

inline bool min_grad_check_on(const MinState &m) {
  const double factor = m.grad_check < 0 ? kDefaultGradCheck : m.grad_check;
  return factor > 0 && m.use_precond;
}

// The ladder's final verdict: a stop that the ladder would report as
// converged is confirmed by the gradient at the final point.  Returns 1 when
// accepted, 0 when the minimizer has to go on (the direction restarts), -1
// when the confirmation failed too often (the stop stays unconverged).
template <class Ops>
int min_confirm_by_gradient(MinState &m, Ops &ops, int verbose, int *verdict) {
  *verdict = 1;
  const double factor = m.grad_check < 0 ? kDefaultGradCheck : m.grad_check;
  if (!(factor > 0) || !m.use_precond) return DEFT_OK;
  const double keepE = m.E;
  const bool keepEv = m.E_valid;
  double ge;
  if (m.precond_kind != MIN_PGRAD_SK && ops.sk_available()) {
    // The structure-factor pgrad gives the Newton estimate with the bulk
    // fluid's Hessian (exact in the uniform fluid); grad/n's ideal-gas Hessian
    // is 1/S(0) times too soft at long wavelengths (15x at eta = 0.35, 24x at
    // 0.40), which exceeds the factor below at minima the energy has reached.
    double Eg, pgg, pgo;
    DEFT_MIN_TRY(ops.gradient(min_grad_prec(m, ops.plan_prec()), false, MIN_PGRAD_SK, &Eg, &pgg, &pgo));
    m.G_valid = false;  // the backend's pgrad is the structure-factor one now
    ge = 0.5 * ops.newton_scale() * pgg;
  } else {
    DEFT_MIN_TRY(min_pgrad(m, ops));  // cached when the fused line search evaluated it
    ge = grad_error(m, ops);
  }
  if (keepEv) {
    m.E = keepE;  // the verdict's energy (same point, same precision: bit-identical)
    m.E_valid = true;
  }
  m.last_grad_error = ge;
  const double target = std::max(m.precision, m.relative_precision * std::fabs(m.E));
  if (ge <= factor * target) return DEFT_OK;
  if (verbose) std::printf("gradient confirmation failed: %.3e > %g x %.3e\n", ge, factor, target);
  if (m.grad_restarts >= kMaxGradientRestarts) {
    *verdict = -1;
    return DEFT_OK;
  }
  m.grad_restarts++;
  min_restart_direction(m);
  *verdict = 0;
  return DEFT_OK;
}

// improve_energy plus the FP32 -> FP64 precision ladder:
//   phase 0: energies and gradients with FP32 FFTs;
//   phase 1: line-search energies with FP64 FFTs, gradients still FP32 (the
//            CG direction and convergence history carry over);
//   phase 2: everything FP64 (the CG direction restarts).
// A phase ends whenever the reference algorithm would stop (converged at the
// phase's precision, or a stalled line search); NaN, infinity and the
// iteration cap stop the minimizer at any phase.  Only phase 2 may end the
// minimization, and it ends it as converged only when the gradient at the
// final point confirms the energy-based verdict (min_confirm_by_gradient).
// Phase 1 is skipped when the polishing gradient precision is FP64.
template <class Ops>
int improve_energy_mixed(MinState &m, Ops &ops, int verbose, int *more) {
  // CONVERGED here means the last iteration met the precision and the final
  // phase is confirming it
  // (a no-change verdict still inside the confirmation window counts as met:
  // two line searches in a row returned to their start point)
  const bool met = m.switch_rp > 0 && m.phase >= 2 &&
                   (m.stop == DEFT_STOP_CONVERGED || m.stop == DEFT_STOP_NO_CHANGE || m.restart_met);
  m.restart_met = false;
  DEFT_MIN_TRY(improve_energy(m, ops, verbose, more));
  const bool final_phase = m.switch_rp > 0 && m.phase >= 2;
  // A confirming iteration whose line search runs into the round-off of the
  // FP64 energies (no lower energy found, the point is left unchanged) keeps
  // the verdict of the iteration before it, as the reference would have
  // stopped there; so does the first iteration after a collapse restart.
  // Either way the gradient has the last word below.
  // A final-phase line search that can no longer lower the FP64 energy (it
  // rose, did not change, or met negative curvature at round-off level) also
  // goes to the gradient when the previous iteration had not met the
  // precision yet: the FP32 iterates can reach the round-off floor one
  // iteration before the energy-based estimate does.  Without the gradient
  // confirmation such a stop stays unconverged.
  bool candidate = false;  // a final-phase stop the ladder reports as converged
  const bool stalled = m.stop == DEFT_STOP_ENERGY_ROSE || m.stop == DEFT_STOP_NO_CHANGE ||
                       m.stop == DEFT_STOP_NEGATIVE_CURVATURE;
  if (final_phase && !*more && stalled && (met || min_grad_check_on(m))) {
    m.stop = DEFT_STOP_CONVERGED;
    candidate = true;
  } else if (!*more && final_phase && m.stop == DEFT_STOP_CONVERGED) {
    // The reference's error estimate trusts one iteration's energy change, so
    // a line search that collapsed to a vanishing step reads as convergence;
    // in the final phase of the ladder such a verdict restarts the CG
    // direction from the last good step instead (at most kMaxCollapseRestarts
    // times).
    if (m.good_step > 0 && std::fabs(m.step) < 1e-3 * std::fabs(m.good_step) &&
        m.collapse_restarts < kMaxCollapseRestarts) {
      m.collapse_restarts++;
      m.restart_met = true;
      min_restart_direction(m);
      m.stop = DEFT_STOP_NONE;
      if (verbose) std::printf("collapsed step: restarting the direction\n");
      *more = 1;
      return DEFT_OK;
    }
    candidate = true;
  }
  if (candidate) {
    int verdict;
    DEFT_MIN_TRY(min_confirm_by_gradient(m, ops, verbose, &verdict));
    if (verdict == 0) {
      m.stop = DEFT_STOP_NONE;
      *more = 1;
    } else if (verdict < 0) {
      m.stop = DEFT_STOP_GRADIENT;
    }
    return DEFT_OK;
  }
  if (*more || m.switch_rp <= 0 || m.phase >= 2) return DEFT_OK;
  const int s = m.stop;
  if (s != DEFT_STOP_CONVERGED && s != DEFT_STOP_NEGATIVE_CURVATURE && s != DEFT_STOP_ENERGY_ROSE &&
      s != DEFT_STOP_NO_CHANGE)
    return DEFT_OK;
  const int pp = ops.plan_prec();
  const int old_g = min_grad_prec(m, pp);
  if (m.phase == 0) {
    m.fp32_iters = m.iter;
    m.phase = (m.polish_grad_prec == DEFT_FFT_FP32) ? 1 : 2;
  } else {
    m.phase = 2;
  }
  if (m.phase == 2) {
    m.iter_at_phase2 = m.iter;
    // a phase may hand over on one spuriously small energy change (the error
    // estimate trusts a single iteration); the all-FP64 phase has to confirm
    // convergence over a few iterations of its own
    m.confirm_until = m.iter + kConfirmIterations;
  }
  m.e_prec = DEFT_FFT_FP64;
  m.g_prec = (m.phase == 1) ? DEFT_FFT_FP32 : DEFT_FFT_FP64;
  min_invalidate(m);
  // A phase usually ends on a line search that ran into the round-off of its
  // energies and shrank the step to nothing; the next phase starts from the
  // step of the last iteration that clearly lowered the energy.
  if (m.good_step > 0 && std::fabs(m.step) < 1e-3 * std::fabs(m.good_step))
    m.step = m.good_step;
  else if (m.good_step <= 0)
    m.step = m.orig_stepsize;  // no clear decrease yet: the reference's first step
  if (m.g_prec != old_g) {
    m.have_dir = false;
    m.have_old = false;
    m.oldgradsqr = 1.0;
  }
  m.stop = DEFT_STOP_NONE;
  if (verbose) std::printf("precision phase %d after %d iterations\n", m.phase, m.iter);
  *more = 1;
  return DEFT_OK;
}

// ===========================================================================
// old API: ConjugateGradient / PreconditionedConjugateGradient with
// QuadraticLineMinimizer, Precision and MaxIter
// ===========================================================================
struct OldCgState {
  bool preconditioned = false;  // PreconditionedConjugateGradientType (pgrad (dF/dn)/(Veff - kT))
  int iter = 0;                 // MinimizerInterface::iter of the conjugate gradient
  double step = 0.1, orig_step = 0.1;
  double oldgradsqr = 0.0;
  bool have_dir = false, have_old = false;  // direction and oldgrad start at zero
  bool E_valid = false;
  double E = 0.0;
  bool G_valid = false;
  double pgg = 0.0, pgo = 0.0;
  int num_energy = 0, num_grad = 0;
  int prec = -1;  // FFT precision of the evaluations (-1: the plan's)
  int stop = DEFT_STOP_NONE;
  // Precision(precision, m) around the conjugate gradient; 0: no wrapper
  double precision = 0.0;
  double dEdn = 0.0, log_avg = 0.0, error_estimate = 1.0, deltaE = 0.0;
  int wrap_iter = 0;  // MinimizerModifier::iter of the wrappers
  // MaxIter(maxiter, m); < 0: no wrapper
  int maxiter = -1;
  int linmin_reason = DEFT_STOP_NONE;  // why the last line minimization returned
};

// MinimizerInterface::minimize(f, gd, x) of the conjugate gradient
// (src/ConjugateGradient.cpp:34-44): iteration counts (the wrappers' too,
// src/Minimizer.h:34-40), step, direction and old gradient start over; the
// Precision wrapper keeps its energy history, as the reference's
// MinimizerModifier::minimize does.
inline void old_restart(OldCgState &m) {
  m.iter = 0;
  m.wrap_iter = 0;
  m.step = m.orig_step;
  m.oldgradsqr = 0.0;
  m.have_dir = false;
  m.have_old = false;
  m.E_valid = false;
  m.G_valid = false;
}

inline int old_prec(const OldCgState &m, int plan_prec) { return m.prec >= 0 ? m.prec : plan_prec; }

template <class Ops>
int old_energy(OldCgState &m, Ops &ops, double *E) {
  if (!m.E_valid) {
    DEFT_MIN_TRY(ops.energy(old_prec(m, ops.plan_prec()), &m.E));
    m.E_valid = true;
    m.num_energy++;
  }
  *E = m.E;
  return DEFT_OK;
}

// grad() or pgrad() of MinimizerInterface (src/Minimizer.h:53-78): the
// preconditioned form computes grad and pgrad in one integralgrad.
template <class Ops>
int old_grad(OldCgState &m, Ops &ops) {
  if (m.G_valid) return DEFT_OK;
  double Eg;
  DEFT_MIN_TRY(ops.gradient(old_prec(m, ops.plan_prec()), m.have_old,
                            m.preconditioned ? MIN_PGRAD_OLD : MIN_PGRAD_NONE, &Eg, &m.pgg, &m.pgo));
  if (!m.E_valid) {
    m.E = Eg;
    m.E_valid = true;
  }
  m.num_grad++;
  m.G_valid = true;
  return DEFT_OK;
}

inline void old_moved(OldCgState &m) {
  m.E_valid = false;
  m.G_valid = false;
}

// QuadraticLineMinimizerType::improve_energy (src/QuadraticLineMinimizer.cpp:27-198)
// along the plan's direction d with slope = grad.d; *step is the conjugate
// gradient's remembered step.  Always returns "no more" (the reference's
// return value is false on every path).
template <class Ops>
int quadratic_line_minimize(OldCgState &m, Ops &ops, double slope, int verbose) {
  double E0;
  DEFT_MIN_TRY(old_energy(m, ops, &E0));
  m.linmin_reason = DEFT_STOP_NONE;
  if (E0 != E0) {
    m.linmin_reason = DEFT_STOP_NAN_ENERGY;
    return DEFT_OK;
  }
  if (E0 > 1e307 || E0 < -1e307) {
    m.linmin_reason = DEFT_STOP_INFINITE_ENERGY;
    return DEFT_OK;
  }
  if (slope != slope) {
    m.linmin_reason = DEFT_STOP_NAN_SLOPE;
    return DEFT_OK;
  }
  if (slope == 0) {
    m.linmin_reason = DEFT_STOP_ZERO_SLOPE;
    return DEFT_OK;
  }
  if (slope * m.step > 0) m.step *= -1;
  double step1 = m.step;
  ops.move(step1);
  old_moved(m);
  double Etried = -137, E;
  do {
    step1 *= 0.5;
    ops.move(-step1);
    old_moved(m);
    DEFT_MIN_TRY(old_energy(m, ops, &E));
    if (E == Etried) break;
    Etried = E;
  } while (min_better(E0, E));
  double E1;
  DEFT_MIN_TRY(old_energy(m, ops, &E1));
  // parabolic extrapolation from E0, E1 and the slope
  const double curvature = 2.0 * (E1 - E0 - step1 * slope) / (step1 * step1);
  const double step2 = -slope / curvature;
  if (verbose) std::printf("\t\tQuad: E0 = %.15g  E1 = %.15g  slope = %.7g  curvature = %.7g\n", E0, E1, slope, curvature);
  if (curvature <= 0.0) {
    if (min_better(E0, E1)) {
      ops.move(-step1);
      old_moved(m);
      m.linmin_reason = DEFT_STOP_NEGATIVE_CURVATURE;
      return DEFT_OK;
    }
    // not moving far enough: keep doubling the step
    double Ebest = E0;
    while (true) {
      DEFT_MIN_TRY(old_energy(m, ops, &E));
      if (!(E <= Ebest)) break;
      Ebest = E;
      ops.move(step1);
      old_moved(m);
      step1 *= 2;
    }
    step1 *= 0.5;
    ops.move(-step1);
    old_moved(m);
    m.step = step1;
  } else if (E1 == E0) {
    do {
      ops.move(-step1);
      step1 *= 2;
      ops.move(step1);
      old_moved(m);
      DEFT_MIN_TRY(old_energy(m, ops, &E));
    } while (E == E0);
    if (E > E0) {
      ops.move(-step1);
      old_moved(m);
      step1 = 0;
      if (verbose) std::printf("\t\tQuad: failed to find any improvement!  :(\n");
    } else if (E != E) {
      ops.move(-step1);
      old_moved(m);
      step1 = 0;
      if (verbose) std::printf("\t\tQuad: found NaN with smallest possible step!!!\n");
    }
  } else {
    m.step = step2;  // remembered for the next line minimization
    ops.move(step2 - step1);
    old_moved(m);
    double E2;
    DEFT_MIN_TRY(old_energy(m, ops, &E2));
    if (min_better(E1, E2) && min_better(E1, E0)) {
      // the first try was better
      ops.move(-(step2 - step1));
      old_moved(m);
    } else if (E2 > E0) {
      ops.move(-step2);
      old_moved(m);
      m.linmin_reason = DEFT_STOP_ENERGY_ROSE;
      return DEFT_OK;
    }
  }
  return DEFT_OK;
}

// ConjugateGradientType::improve_energy (src/ConjugateGradient.cpp:59-108) or
// PreconditionedConjugateGradientType::improve_energy (:115-160).
template <class Ops>
int old_cg_improve(OldCgState &m, Ops &ops, int verbose, bool *improved) {
  m.iter++;
  double E0;
  DEFT_MIN_TRY(old_energy(m, ops, &E0));
  if (E0 != E0) {
    m.stop = DEFT_STOP_NAN_ENERGY;
    *improved = false;
    return DEFT_OK;
  }
  DEFT_MIN_TRY(old_grad(m, ops));
  m.G_valid = false;  // invalidate_cache(): the gradient is used once
  // Polak-Ribiere: g = -grad (plain) with beta = g.(g - oldgrad)/|oldgrad|^2,
  // or beta = -pgrad.(-grad - oldgrad)/(-pgrad.oldgrad) (preconditioned); both
  // are (pg.grad - pg.grad_old)/oldgradsqr with pg = grad for the plain form
  double beta = (m.pgg - m.pgo) / m.oldgradsqr;
  if (beta < 0 || beta != beta || m.oldgradsqr == 0) beta = 0;
  m.oldgradsqr = m.pgg;
  double gd;  // grad.direction of the new direction -pg + beta direction
  DEFT_MIN_TRY(ops.cg_update(beta, m.have_dir, &gd));
  m.have_dir = true;
  m.have_old = true;
  if (!m.preconditioned && gd > 0) {
    // the direction is uphill: reset to the gradient
    DEFT_MIN_TRY(ops.cg_reset(&gd));
    if (verbose) std::printf("reset to gradient...\n");
  }
  if (verbose) std::printf("beta = %g\n", beta);
  // one QuadraticLineMinimizer call (its improve_energy always returns false,
  // so the reference's loop of up to 100 calls runs it once)
  DEFT_MIN_TRY(quadratic_line_minimize(m, ops, gd, verbose));
  double E;
  DEFT_MIN_TRY(old_energy(m, ops, &E));
  if (verbose) std::printf("%sEnergy = %.16g\tstep = %g\n", m.preconditioned ? "PCG " : "CG ", E, m.step);
  *improved = m.preconditioned ? (E < E0 || beta != 0) : (E < E0);
  return DEFT_OK;
}

// The minimizer stack: MaxIter (if set) around Precision (if set) around the
// conjugate gradient.  *more as the outermost improve_energy returns.
template <class Ops>
int old_improve_energy(OldCgState &m, Ops &ops, int verbose, int *more) {
  // MaxIterType::improve_energy (src/Minimizer.cpp): counts its own calls
  if (m.maxiter >= 0 && m.wrap_iter > m.maxiter) {
    m.stop = DEFT_STOP_MAXITER;
    *more = 0;
    return DEFT_OK;
  }
  m.wrap_iter++;
  if (!(m.precision > 0)) {
    bool improved;
    DEFT_MIN_TRY(old_cg_improve(m, ops, verbose, &improved));
    if (m.stop == DEFT_STOP_NAN_ENERGY) {
      *more = 0;
      return DEFT_OK;
    }
    m.stop = improved ? DEFT_STOP_NONE : DEFT_STOP_NOT_IMPROVED;
    *more = improved ? 1 : 0;
    return DEFT_OK;
  }
  // PrecisionMinimizer::improve_energy (src/Precision.cpp:54-154)
  const double w = 0.1;
  const double old_deltaE = m.deltaE;
  double old_E, new_E;
  DEFT_MIN_TRY(old_energy(m, ops, &old_E));
  bool not_done;
  DEFT_MIN_TRY(old_cg_improve(m, ops, verbose, &not_done));
  if (m.stop == DEFT_STOP_NAN_ENERGY) not_done = false;
  DEFT_MIN_TRY(old_energy(m, ops, &new_E));
  m.deltaE = old_E - new_E;
  m.dEdn = (1 - w) * m.dEdn + w * m.deltaE;
  m.dEdn = std::max(m.deltaE, old_deltaE);
  const double new_log = (m.deltaE != 0 && old_deltaE != 0) ? std::log(std::fabs(m.deltaE / old_deltaE)) : 0.0;
  m.log_avg = (1 - w) * m.log_avg + w * new_log;
  const double ratio_avg = std::exp(m.log_avg);
  double error_guess = std::fabs(std::max(m.dEdn, m.deltaE) / m.log_avg);
  if (ratio_avg >= 1) error_guess = m.precision + std::fabs(new_E);
  m.error_estimate = 2 * error_guess;
  if (m.deltaE == 0 && old_deltaE == 0) {
    if (verbose) std::printf("We got no change twice in a row, so we're done!\n");
    m.stop = DEFT_STOP_NO_CHANGE;
    *more = 0;
    return DEFT_OK;
  }
  if (not_done && !(m.error_estimate < m.precision)) {
    if (verbose) std::printf("Error estimate is %10g\n", m.error_estimate);
    m.stop = DEFT_STOP_NONE;
    *more = 1;
    return DEFT_OK;
  }
  if (verbose) std::printf("Converged with precision of %g!\n", m.error_estimate);
  if (m.stop != DEFT_STOP_NAN_ENERGY) m.stop = not_done ? DEFT_STOP_CONVERGED : DEFT_STOP_NOT_IMPROVED;
  *more = 0;
  return DEFT_OK;
}

}  // namespace deft
