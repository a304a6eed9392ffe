// chooser.h -- the learned parameter chooser (oversampling, support, tile).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace dmx {

struct LearnedChoice {
  int kidx = -1;
  int log2tile = 0;
  size_t nover[3] = {1, 1, 1};
  int verified = 0;  // the accuracy spot-check passed for this signature
};

// Picks (kernel row, oversampled grid, tile) with the trained network.
// Returns false when no network is loaded or the prediction cannot be
// mapped onto a tabulated kernel reaching eps; the caller then falls back to
// the analytic choice.
bool learned_choose(int ndim, const size_t *nuni, size_t M, double eps,
                    double smin, double smax, bool fp64, bool gridding,
                    LearnedChoice &out);
// result of the accuracy spot-check of a learned choice: keep it, or use
// the analytic choice for this signature from now on
void chooser_verdict(int ndim, const size_t *nuni, size_t M, double eps,
                     double smin, double smax, bool fp64, bool gridding,
                     bool ok);
// raw network outputs for given features; false without loaded weights
bool chooser_predict(const double *feat, double *out, hipStream_t s);
// the 6 input features of a request (ndim, mean log2 size, log10 M,
// -log10 eps, fp64, type 1)
void chooser_features(int ndim, const size_t *nuni, size_t M, double eps,
                      bool fp64, bool gridding, double *f);
// weights (MLP_NPARAM floats) and normalisation (feature mean[6], std[6],
// target mean[3], std[3])
void chooser_set(const float *w, const double *norm);
bool chooser_loaded();

}  // namespace dmx
