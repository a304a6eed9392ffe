// fft.h -- launch descriptor of the Stockham FFT over a line set.
#pragma once
#include <cstdint>
#include <vector>

#include "dmx.h"

namespace dmx {

#define FFT_THREADS 256
#define FFT_MAXEL32 8192  // 64 KiB of complex float per workgroup
#define FFT_MAXEL64 4096  // 64 KiB of complex double per workgroup
#define FFT_KM(T, R)                                                   \
  (((sizeof(T) == 8 ? FFT_MAXEL64 : FFT_MAXEL32) + (R)*FFT_THREADS - 1) / \
   ((R)*FFT_THREADS))

// Lines along one axis of a row-major array.  Line (u, v) starts at
// U(u)*st_u + V(v)*st_v, element k at + k*st_ax, where the virtual indices
// map to actual ones through at most two ranges:
// U(u) = u < un1 ? ua1 + u : ua2 + (u - un1)   (same for V).
struct FftLaunch {
  void *data;
  const void *tw;  // exp(-2 pi i m / L), m < L, compute type
  int L, LP, Bl, npass, sign;
  int radix[24];
  int64_t st_ax, st_u, st_v;
  int nu, nv;
  int un1, ua1, ua2, vn1, va1, va2;
};

void fft_twiddles(int L, bool fp64, std::vector<unsigned char> &out);
// Runs the FFT over every line of the set (in place).  `scratch` must be as
// large as the array when L exceeds the LDS capacity.
void fft_lineset(const FftLaunch &p, bool fp64, void *scratch, hipStream_t s);

}  // namespace dmx
