// CPU check of the backtracking line search (tdycore_amd/csrc/linesearch.h):
// runs bt_search on scripted residual norms and prints one line per case,
//   <case> rc accepted nonfinite trials lambda
// which tests/test_linesearch.py compares with what the search must do.
#include <cmath>
#include <cstdio>
#include <functional>
#include <limits>

#include "../../tdycore_amd/csrc/linesearch.h"

using tdyhip::BTParams;
using tdyhip::BTResult;
using tdyhip::bt_search;

static void run(const char *name, double fnorm, double initslope, std::function<double(double, int)> g,
                int stop_at = -1) {
  BTParams p;
  BTResult r;
  int calls = 0, polls = 0;
  auto trial = [&](double lam, double *gn) -> int {
    *gn = g(lam, calls++);
    return 0;
  };
  auto stop = [&]() -> int { return (stop_at >= 0 && ++polls >= stop_at) ? 6 : 0; };
  const int rc = bt_search(fnorm, initslope, p, false, 0.0, trial, stop, r);
  std::printf("%s %d %d %d %d %.17g %.17g\n", name, rc, r.accepted, r.nonfinite, r.trials, r.lambda, r.gnorm);
}

int main() {
  const double nan = std::numeric_limits<double>::quiet_NaN(), inf = std::numeric_limits<double>::infinity();
  // exact Newton direction: ||F(x - lam y)|| = (1 - lam) f
  run("newton", 2.0, -4.0, [](double lam, int) { return 2.0 * std::fabs(1.0 - lam); });
  // full step overshoots: g(lam) = f |1 - lam + 3 lam^2|
  run("overshoot", 1.0, -1.0, [](double lam, int) { return std::fabs(1.0 - lam + 3.0 * lam * lam); });
  // NaN at the full step, Inf at the first backtrack
  run("nan_full", 1.0, -1.0, [nan](double, int) { return nan; });
  run("inf_second", 1.0, -1.0, [inf](double, int k) { return k == 0 ? 5.0 : inf; });
  // NaN at a later trial
  run("nan_third", 1.0, -1.0, [nan](double, int k) { return k < 2 ? 5.0 : nan; });
  // never a decrease: the step shrinks until minlambda
  run("never", 1.0, -1.0, [](double, int) { return 2.0; });
  // a residual that rises after the first backtrack: the cubic fits must keep shrinking the step
  run("rising", 1.0, -0.5, [](double lam, int k) { return k == 0 ? 1.0 : 2.0; });
  // cancelled at the second backtracking trial
  run("cancel", 1.0, -1.0, [](double, int) { return 2.0; }, 2);
  return 0;
}
