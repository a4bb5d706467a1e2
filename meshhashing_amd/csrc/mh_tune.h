// The shape choice of the online marching-cubes tuner: host code only (no
// device calls), kept apart so that the rule is tested on the CPU.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace mh {

// Trial p of an online tuning over `shapes` launch shapes runs shape 0 for
// even p and candidate 1 + ((p - 1) / 2) % (shapes - 1) for odd p, so that
// every candidate trial sits between two trials of shape 0 on the frames
// just before and after.
inline int mc_trial_shape(uint32_t p, int shapes) {
  return (p & 1u) ? 1 + (int)(((p - 1u) / 2u) % (uint32_t)(shapes - 1)) : 0;
}

struct McPick {
  int shape;          // the chosen shape
  double default_ms;  // shape 0's median trial
  double best_ms;     // default_ms x the chosen candidate's worst ratio (default_ms for shape 0)
};

// Frames differ in content as well as size, and a trial's time drifts with
// them (a first frame, where every block is new, costs several times a later
// one): each candidate trial is compared with the faster of the shape-0
// trials on either side of it, and a candidate counts with its worst ratio
// over its trials (a trial without a time, or without a shape-0 trial after
// it, rules the candidate out).  Shape 0 stays unless a candidate's worst
// ratio is at most `threshold`; of those, the least wins.
inline McPick mc_pick(const std::vector<double>& tp, int shapes, double threshold = 0.97) {
  std::vector<double> worst((size_t)shapes, 0.0);
  std::vector<double> dflt;
  for (size_t p = 0; p < tp.size(); ++p) {
    if (!(p & 1)) {
      if (std::isfinite(tp[p])) dflt.push_back(tp[p]);
      continue;
    }
    const double r = p + 1 < tp.size() ? tp[p] / std::fmin(tp[p - 1], tp[p + 1]) : (double)INFINITY;
    const int c = mc_trial_shape((uint32_t)p, shapes);
    worst[(size_t)c] = std::isfinite(r) ? std::fmax(worst[(size_t)c], r) : (double)INFINITY;
  }
  McPick out{0, INFINITY, INFINITY};
  double best_r = threshold;
  for (int k = 1; k < shapes; ++k)
    if (worst[(size_t)k] > 0.0 && worst[(size_t)k] <= best_r) {
      best_r = worst[(size_t)k];
      out.shape = k;
    }
  std::sort(dflt.begin(), dflt.end());
  out.default_ms = dflt.empty() ? (double)INFINITY : dflt[dflt.size() / 2];
  out.best_ms = out.shape ? out.default_ms * best_r : out.default_ms;
  return out;
}

}  // namespace mh
