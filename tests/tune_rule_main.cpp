// Runs the marching-cubes tuner's shape rule (mh_tune.h) on trial times read
// from stdin: "shapes n t0 t1 ... t(n-1)" ("nan" for a trial without a time);
// prints "shape default_ms best_ms" and, per trial, the shape it ran.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mh_tune.h"

int main() {
  int shapes = 0, n = 0;
  if (std::scanf("%d %d", &shapes, &n) != 2) return 2;
  std::vector<double> tp((size_t)n);
  for (int i = 0; i < n; ++i) {
    char buf[64];
    if (std::scanf("%63s", buf) != 1) return 2;
    tp[(size_t)i] = std::strtod(buf, nullptr);
  }
  const mh::McPick p = mh::mc_pick(tp, shapes);
  std::printf("%d %.9g %.9g\n", p.shape, p.default_ms, p.best_ms);
  for (int i = 0; i < n; ++i) std::printf("%d ", mh::mc_trial_shape((uint32_t)i, shapes));
  std::printf("\n");
  return 0;
}
