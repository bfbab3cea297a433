// Launch plan of one J/K, two-density or gradient build (host code, no HIP):
// which class combinations to launch, in what order, on which stream.
#pragma once
#include <algorithm>
#include <cstddef>
#include <vector>

namespace gx {

struct PlanStep {
  int combo, X, Y, side;  // combo = X (X + 1) / 2 + Y; side -1 = the context stream
};

// (combo, X, Y, side stream) per class combination with work (n_items > 0),
// over `ncls` pair classes.  With side streams (nside > 1), once the
// autotuner has timed every such combination (`tms[combo + mode * ncombo]`
// > 0): longest first, each onto the stream with the least time queued so
// far (the nside side streams and the context stream, in that order on
// ties), so that the streams end together; otherwise class order, round
// robin over the side streams (all on the context stream when nside = 0).
inline std::vector<PlanStep> launch_plan(int ncls, const std::vector<int>& n_items, const std::vector<double>& tms,
                                         int mode, int nside) {
  const int ncombo = ncls * (ncls + 1) / 2;
  std::vector<PlanStep> plan;
  int combo = 0;
  for (int X = 0; X < ncls; ++X)
    for (int Y = 0; Y <= X; ++Y, ++combo)
      if ((size_t)combo < n_items.size() && n_items[combo] > 0) plan.push_back({combo, X, Y, -1});
  auto ms = [&](const PlanStep& p) {
    const size_t i = (size_t)(p.combo + mode * ncombo);
    return i < tms.size() ? tms[i] : 0.0;
  };
  bool timed = nside > 1;
  for (const PlanStep& p : plan) timed = timed && ms(p) > 0.0;
  if (!timed) {
    for (size_t i = 0; i < plan.size(); ++i) plan[i].side = nside > 0 ? (int)(i % nside) : -1;
    return plan;
  }
  std::stable_sort(plan.begin(), plan.end(), [&](const PlanStep& a, const PlanStep& b) { return ms(a) > ms(b); });
  std::vector<double> queued(nside + 1, 0.0);  // the side streams, then the context stream
  for (PlanStep& p : plan) {
    const int k = (int)(std::min_element(queued.begin(), queued.end()) - queued.begin());
    queued[k] += ms(p);
    p.side = k < nside ? k : -1;
  }
  return plan;
}

}  // namespace gx
