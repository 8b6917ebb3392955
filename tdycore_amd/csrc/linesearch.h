// Decisions of the backtracking line search used by the Newton solver
// (PETSc 3.17 SNESLINESEARCHBT with its defaults, the line search of
// SNESNEWTONLS that TDycore's SNES and TS solves run under, src/tdyti.c:135,188):
// the sufficient-decrease test, the quadratic first and cubic later step
// lengths, and the end of the search.  Plain arithmetic, usable from host code
// and from the device Newton kernels, and compiled on its own by the CPU
// tests.
//
// A NaN or Inf function norm ends the search as a failure
// (SNES_LINESEARCH_FAILED_NANORINF): it never passes the decrease test and
// never yields a step length.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define TDY_HD __host__ __device__
#else
#define TDY_HD
#endif

namespace tdyhip {

struct BTParams {
  double alpha = 1e-4;      // sufficient-decrease constant
  double minlambda = 1e-12; // smallest step length tried
};

// 0.5 g^2 <= 0.5 f^2 + lambda alpha initslope, false for a non-finite g
TDY_HD inline bool bt_sufficient(double fnorm, double gnorm, double lambda, double initslope, double alpha) {
  if (!std::isfinite(gnorm)) return false;
  return 0.5 * gnorm * gnorm <= 0.5 * fnorm * fnorm + lambda * alpha * initslope;
}

// Step length after the full step failed: minimiser of the quadratic through
// f(0), f'(0) and f(lambda), kept in [0.1 lambda, 0.5 lambda].
TDY_HD inline double bt_quadratic(double fnorm, double gnorm, double lambda, double initslope) {
  double lt = -initslope / (gnorm * gnorm - fnorm * fnorm - 2.0 * lambda * initslope);
  if (!(lt <= 0.5 * lambda)) lt = 0.5 * lambda;  // also catches a NaN quotient
  return (lt <= 0.1 * lambda) ? 0.1 * lambda : lt;
}

// Later step lengths: minimiser of the cubic through the last two trials,
// kept in [0.1 lambda, 0.5 lambda].
TDY_HD inline double bt_cubic(double fnorm, double gnorm, double lambda, double gnormprev, double lambdaprev,
                              double initslope) {
  const double t1 = 0.5 * (gnorm * gnorm - fnorm * fnorm) - lambda * initslope;
  const double t2 = 0.5 * (gnormprev * gnormprev - fnorm * fnorm) - lambdaprev * initslope;
  const double a = (t1 / (lambda * lambda) - t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
  const double b = (-lambdaprev * t1 / (lambda * lambda) + lambda * t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
  double d = b * b - 3 * a * initslope;
  if (d < 0.0) d = 0.0;
  double lt = (a == 0.0) ? -initslope / (2.0 * b) : (-b + std::sqrt(d)) / (3.0 * a);
  if (!(lt <= 0.5 * lambda)) lt = 0.5 * lambda;
  return (lt <= 0.1 * lambda) ? 0.1 * lambda : lt;
}

// The search gives up once the step length is at most minlambda (also for a
// NaN step length, so the loop always ends).
TDY_HD inline bool bt_exhausted(double lambda, double minlambda) { return !(lambda > minlambda); }

// initslope = -(F, J Y) with PETSc's sign and zero guards.
TDY_HD inline double bt_initslope(double dot_f_jy) {
  double s = dot_f_jy;
  if (s > 0.0) s = -s;
  if (s == 0.0) s = -1.0;
  return s;
}

// Outcome of one search.
struct BTResult {
  double lambda = 1.0;   // last step length tried
  double gnorm = 0.0;    // residual norm at it
  int accepted = 0;      // 1: sufficient decrease at lambda
  int nonfinite = 0;     // 1: ended by a NaN or Inf residual norm
  int trials = 0;        // residual evaluations
};

// The whole search from the full step.  trial(lambda, &gnorm) evaluates the
// residual norm at X - lambda Y and returns nonzero on an error (handed back
// at once); stop() is polled before every backtracking trial and a nonzero
// value is handed back too.  With have_gnorm1 the caller already has the norm
// at lambda = 1 (gnorm1; then trial is not called for the full step).
template <class Trial, class Stop>
inline int bt_search(double fnorm, double initslope, const BTParams &p, bool have_gnorm1, double gnorm1, Trial trial,
                     Stop stop, BTResult &r) {
  int rc;
  r = BTResult();
  r.lambda = 1.0;
  if (have_gnorm1) {
    r.gnorm = gnorm1;
  } else {
    if ((rc = trial(r.lambda, &r.gnorm))) return rc;
  }
  r.trials = 1;
  if (bt_sufficient(fnorm, r.gnorm, r.lambda, initslope, p.alpha)) { r.accepted = 1; return 0; }
  if (!std::isfinite(r.gnorm)) { r.nonfinite = 1; return 0; }
  double lambdaprev = r.lambda, gnormprev = r.gnorm;
  r.lambda = bt_quadratic(fnorm, r.gnorm, r.lambda, initslope);
  if ((rc = stop())) return rc;
  if ((rc = trial(r.lambda, &r.gnorm))) return rc;
  r.trials++;
  while (!bt_sufficient(fnorm, r.gnorm, r.lambda, initslope, p.alpha)) {
    if (!std::isfinite(r.gnorm)) { r.nonfinite = 1; return 0; }
    if (bt_exhausted(r.lambda, p.minlambda)) return 0;
    const double lt = bt_cubic(fnorm, r.gnorm, r.lambda, gnormprev, lambdaprev, initslope);
    lambdaprev = r.lambda;
    gnormprev = r.gnorm;
    r.lambda = lt;
    if ((rc = stop())) return rc;
    if ((rc = trial(r.lambda, &r.gnorm))) return rc;
    r.trials++;
  }
  r.accepted = 1;
  return 0;
}

}  // namespace tdyhip
