// grid.h -- argument blocks of the grid-side kernels in grid.hip.
#pragma once
#include <cstdint>

#include "dmx.h"

namespace dmx {

struct CorfacArgs {
  int n, nq, W, pad_;
  double dx;
  double x[32], wpsi[32];
};

// uniform <-> oversampled grid copy with correction.  Axes padded to 3 with
// n = N = 1; cf[d] points at n/2+1 correction factors (a single 1.0 for
// padded axes); ust: strides of the uniform array in elements.  full[d]
// (pad only): cover all N[d] grid positions of axis d, writing zero where
// the mode is not retained, instead of the n[d] retained ones.
struct GridCopyArgs {
  void *grid;
  void *uni;
  int u_f64, fft_order;
  int n[3], N[3];
  int full[3];
  int64_t ust[3];
  const double *cf[3];
};

struct NudftArgs {
  int type, ndim, forward, fft_order;
  int64_t M, nsel;
  int64_t n[3];
  double period[3];
  const double *coord;  // (M, ndim)
  const double *vals;   // complex128: M points (type 1) or prod(n) modes
  const int64_t *sel;
  double *out;          // complex128 nsel
};

void launch_corfac(const CorfacArgs &a, double *out, hipStream_t s);
// dst[i] = vals[i], i < n <= CONST_MAX, written by a kernel whose argument
// carries the values (no host-to-device copy); as float when as_f32
constexpr int CONST_MAX = 448;
void launch_store_consts(void *dst, const double *vals, int n, bool as_f32,
                         hipStream_t s);
void launch_grid_copy(bool fp64, bool pad, const GridCopyArgs &a,
                      hipStream_t s);
void launch_nudft(const NudftArgs &a, hipStream_t s);
// dst[i] = src[idx[i]] for complex128 arrays
void launch_gather_c128(const void *src, const int64_t *idx, size_t n,
                        void *dst, hipStream_t s);

}  // namespace dmx
