// Host build of the launch plan (gaussium_amd/csrc/gx_plan.h) for the tests:
// one C entry that returns the plan as (combo, X, Y, side) quadruples.
#include "gx_plan.h"

extern "C" int plan_case(int ncls, const int* n_items, int nitems, const double* tms, int ntms, int mode,
                         int nside, int* out) {
  const std::vector<int> n(n_items, n_items + nitems);
  const std::vector<double> t(tms, tms + ntms);
  int k = 0;
  for (const gx::PlanStep& p : gx::launch_plan(ncls, n, t, mode, nside)) {
    out[4 * k] = p.combo;
    out[4 * k + 1] = p.X;
    out[4 * k + 2] = p.Y;
    out[4 * k + 3] = p.side;
    ++k;
  }
  return k;
}
