// chooser.h -- the learned parameter chooser (oversampling, support, tile).
#pragma once
#include <cstddef>

namespace dmx {

struct LearnedChoice {
  int kidx = -1;
  int log2tile = 0;
  size_t nover[3] = {1, 1, 1};
};

// Picks (kernel row, oversampled grid, tile) with the trained network.
// Returns false when no network is loaded or the prediction cannot be
// mapped onto a tabulated kernel reaching eps; the caller then falls back to
// the analytic choice.
bool learned_choose(int ndim, const size_t *nuni, size_t M, double eps,
                    double smin, double smax, bool fp64, bool gridding,
                    LearnedChoice &out);

}  // namespace dmx
