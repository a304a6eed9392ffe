// chooser.hip -- the learned parameter chooser: features of a request, the
// network's forward pass on the GPU (mlp.hip), and the mapping of its three
// outputs (oversampling factor, support, tile) onto a tabulated kernel that
// reaches the requested accuracy.  Decisions are cached per request
// signature, so the network runs once per new kind of request.
#include "chooser.h"

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "dmx.h"
#include "mlp.h"
#include "params.h"
#include "runtime.h"

namespace dmx {

namespace {
struct Net {
  bool loaded = false;
  std::vector<float> w;       // MLP_NPARAM
  double norm[18];            // feature mean/std (6+6), target mean/std (3+3)
  int device = -1;
  Buf dw;                     // device copy of the weights
  std::map<std::string, LearnedChoice> cache;
  std::mutex mu;
};
Net &net() {
  static Net *n = new Net();
  return *n;
}
}  // namespace

void chooser_features(int ndim, const size_t *nuni, size_t M, double eps,
                      bool fp64, bool gridding, double *f) {
  double l2 = 0;
  for (int d = 0; d < ndim; ++d) l2 += std::log2((double)nuni[d]);
  f[0] = ndim;
  f[1] = l2 / ndim;
  f[2] = std::log10((double)std::max<size_t>(M, 1));
  f[3] = -std::log10(eps);
  f[4] = fp64 ? 1. : 0.;
  f[5] = gridding ? 1. : 0.;
}

void chooser_set(const float *w, const double *norm) {
  Net &n = net();
  std::lock_guard<std::mutex> lk(n.mu);
  n.w.assign(w, w + MLP_NPARAM);
  for (int i = 0; i < 18; ++i) n.norm[i] = norm[i];
  n.loaded = true;
  n.device = -1;
  n.cache.clear();
}

bool chooser_loaded() { return net().loaded; }

// raw network outputs (denormalised): sigma, W, log2tile
bool chooser_predict(const double *feat, double *out, hipStream_t s) {
  Net &n = net();
  std::lock_guard<std::mutex> lk(n.mu);
  if (!n.loaded) return false;
  int dev = 0;
  DMX_HIP(hipGetDevice(&dev));
  if (n.device != dev) {
    n.dw.reset(MLP_NPARAM * 4);
    DMX_HIP(hipMemcpy(n.dw.p, n.w.data(), MLP_NPARAM * 4,
                      hipMemcpyHostToDevice));
    n.device = dev;
  }
  float x[MLP_IN];
  for (int i = 0; i < MLP_IN; ++i)
    x[i] = (float)((feat[i] - n.norm[i]) / n.norm[6 + i]);
  Buf io(64);
  DMX_HIP(hipMemcpyAsync(io.p, x, sizeof(x), hipMemcpyHostToDevice, s));
  mlp_forward(n.dw.as<float>(), io.as<float>(), 1, nullptr,
              io.as<float>() + 8, s);
  float y[MLP_OUT];
  DMX_HIP(hipMemcpyAsync(y, io.as<float>() + 8, sizeof(y),
                         hipMemcpyDeviceToHost, s));
  DMX_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < MLP_OUT; ++k)
    out[k] = y[k] * n.norm[12 + 3 + k] + n.norm[12 + k];
  return true;
}

// Map (sigma, W, log2tile) onto the kernel table: the support nearest the
// prediction that has a row reaching eps within [smin, smax], and on that
// support the row whose oversampling factor is nearest the prediction.
bool map_choice(int ndim, const size_t *nuni, double eps, double smin,
                double smax, bool fp64, const double *pred,
                LearnedChoice &out) {
  const auto &db = kernel_db();
  const int Wmax = fp64 ? 16 : 8;
  int Wp = (int)std::lround(pred[1]);
  Wp = std::max(4, std::min(Wmax, Wp));
  for (int dist = 0; dist <= 12; ++dist) {
    for (int sgn : {1, -1}) {
      int W = Wp + sgn * dist;
      if (dist == 0 && sgn < 0) continue;
      if (W < 4 || W > Wmax) continue;
      int best = -1;
      double bd = 1e30;
      for (size_t i = 0; i < db.size(); ++i) {
        const KRow &k = db[i];
        if (k.ndim != ndim || k.single != (int)!fp64 || k.W != W) continue;
        if (k.epsilon > eps || k.ofactor < smin || k.ofactor > smax) continue;
        double dd = std::fabs(k.ofactor - pred[0]);
        if (dd < bd) { bd = dd; best = (int)i; }
      }
      if (best >= 0) {
        out.kidx = best;
        for (int d = 0; d < ndim; ++d)
          out.nover[d] = oversampled_length(nuni[d], db[best].ofactor);
        int l2 = (int)std::lround(pred[2]);
        out.log2tile = std::max(2, std::min(ndim == 1 ? 10 : 6, l2));
        if (ndim == 1) out.log2tile = std::max(out.log2tile, 6);
        return true;
      }
    }
  }
  return false;
}

// The log2 tile the network's third head is an offset from: the tiles of
// its training set (chooser_train.py tile_anchor), kept apart from the
// library's default tile so that a change of the default does not move the
// network's choices.
static int tile_anchor(int ndim, bool fp64) {
  if (ndim == 1) return 9;
  if (ndim == 2) return fp64 ? 4 : 5;
  return 3;
}

static std::string sig_key(int ndim, const size_t *nuni, size_t M,
                           double eps, double smin, double smax, bool fp64,
                           bool gridding) {
  std::string key = std::to_string(ndim) + ":" + std::to_string(M) + ":" +
                    std::to_string(eps) + ":" + std::to_string(smin) + ":" +
                    std::to_string(smax) + ":" + (fp64 ? "d" : "s") +
                    (gridding ? "1" : "2");
  for (int d = 0; d < ndim; ++d) key += ":" + std::to_string(nuni[d]);
  return key;
}

bool learned_choose(int ndim, const size_t *nuni, size_t M, double eps,
                    double smin, double smax, bool fp64, bool gridding,
                    LearnedChoice &out) {
  Net &n = net();
  if (!n.loaded) return false;
  const std::string key =
      sig_key(ndim, nuni, M, eps, smin, smax, fp64, gridding);
  {
    std::lock_guard<std::mutex> lk(n.mu);
    auto it = n.cache.find(key);
    if (it != n.cache.end()) {
      out = it->second;
      return out.kidx >= 0;
    }
  }
  double f[MLP_IN], pred[MLP_OUT];
  chooser_features(ndim, nuni, M, eps, fp64, gridding, f);
  hipStream_t s;
  DMX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  bool ok = chooser_predict(f, pred, s);
  (void)hipStreamDestroy(s);
  // the heads are offsets from the analytic choice (oversampling factor,
  // support, log2 tile): where the network has learned nothing better it
  // outputs about zero and the mapping lands on the analytic setting
  if (ok) {
    size_t nover_an[3];
    const int kan = find_params(eps, smin, smax, ndim, nuni, M, gridding,
                                !fp64, 8, 8, nover_an);
    if (kan < 0) {
      ok = false;
    } else {
      pred[0] += kernel_db()[kan].ofactor;
      pred[1] += kernel_db()[kan].W;
      pred[2] += tile_anchor(ndim, fp64);
    }
  }
  LearnedChoice lc;
  if (ok) ok = map_choice(ndim, nuni, eps, smin, smax, fp64, pred, lc);
  if (!ok) lc.kidx = -1;
  {
    std::lock_guard<std::mutex> lk(n.mu);
    n.cache[key] = lc;
  }
  out = lc;
  return ok;
}

void chooser_verdict(int ndim, const size_t *nuni, size_t M, double eps,
                     double smin, double smax, bool fp64, bool gridding,
                     bool ok) {
  // after the spot-check: keep the learned choice (verified) or pin this
  // signature to the analytic choice
  Net &n = net();
  const std::string key =
      sig_key(ndim, nuni, M, eps, smin, smax, fp64, gridding);
  std::lock_guard<std::mutex> lk(n.mu);
  if (ok)
    n.cache[key].verified = 1;
  else
    n.cache[key] = LearnedChoice();
}

}  // namespace dmx
