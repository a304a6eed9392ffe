// wgridder.h -- radio-interferometric (de)gridding on the NUFFT
// (wgridder.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "grid.h"
#include "runtime.h"

namespace dmx {

// geometry and inputs of one call (device pointers)
struct WgArgs {
  const double *uvw;   // (nrow, 3), contiguous
  const double *fovc;  // nchan: freq / c
  const void *vis;     // gridding: (nrow, nchan) complex, strides vst0/1
  int vis_f64;
  const void *wgt;     // (nrow, nchan) real or NULL, strides wst0/1
  int wgt_f64;
  const uint8_t *mask; // (nrow, nchan) or NULL, strides mst0/1
  int64_t vst0, vst1, wst0, wst1, mst0, mst1;
  size_t nrow, nchan;
  int nx, ny;
  double px, py;
  double ufac, vfac, wfac;  // -1 where flipped
  double lshift, mshift, nshift;
  int gridding, wgrid, divide_by_n, allow_nshift;
};

// per-visibility arrays (row-major over (row, chan)): NUFFT coordinates
// (M, 2), w, activity, and (gridding) value * wgt * shift phase (FP64)
struct WgPoints {
  double *coord;
  double *w;
  uint8_t *active;
  void *val;
  int val_f64;
  unsigned long long *wminmax;  // ordered-key min, max of w; active count
};

// w planes: w_p = wmin + p dw, kernel support W, FP64 coefficients
struct WgPlanes {
  double wmin, dw;
  int nplanes, W;
  const double *coef;
};

struct WgJob {
  WgArgs a{};
  WgPoints o{}, so{};
  WgPlanes pl{};
  Params p;
  CorfacArgs cw{};
  double dw = 0, nm1min = 0, nm1max = 0, wmin_d = 0, wmax_d = 0;
  size_t nvis = 0, M = 0;
  hipStream_t s = nullptr;
  // checks, frequencies, n - 1 range and nshift (a must be filled)
  void setup(const double *freq_h, size_t nchan, size_t nrow, hipStream_t s);
  // coordinates, w, activity, values; w range and active count
  void prep(const void *vis, bool vis_f64, int64_t vst0, int64_t vst1);
  // NUFFT parameters (and w planes) for accuracy eps
  void choose(double eps, double smin, double smax, bool fp64_values,
              const dmx_options &opt);
  void grid(void *dirty, bool dirty_f64, int64_t ds0, int64_t ds1);
  void degrid(const void *dirty, bool dirty_f64, int64_t ds0, int64_t ds1,
              void *vis, bool vis_f64, int64_t vs0, int64_t vs1);
  void zero_dirty(void *dirty, bool dirty_f64, int64_t ds0, int64_t ds1);

 private:
  Buf fovc_, pts_, val_, mm_, keys_, skeys_, idx_, sorted_, sval_, wcoef_;
  std::vector<size_t> first_;
  void sort_planes();
  void plane_range(int p, size_t &lo, size_t &hi) const;
};

}  // namespace dmx
