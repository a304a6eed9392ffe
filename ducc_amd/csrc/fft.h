// fft.h -- launch descriptor of the Stockham FFT over a line set.
#pragma once
#include <cstdint>
#include <vector>

#include "dmx.h"

namespace dmx {

#define FFT_THREADS 256
#define FFT_MAXEL32 8192  // 64 KiB of complex float per workgroup
#define FFT_MAXEL64 4096  // 64 KiB of complex double per workgroup
// butterflies per thread and pass for workgroups of at most MAXEL elements
#define FFT_KM_EL(MAXEL, R) (((MAXEL) + (R)*FFT_THREADS - 1) / ((R)*FFT_THREADS))
#define FFT_KM(T, R) FFT_KM_EL(sizeof(T) == 8 ? FFT_MAXEL64 : FFT_MAXEL32, R)

// Constants of one Stockham pass, computed on the host: radix, butterflies
// per line (L/R), sub-transform length before the pass (Ns), twiddle stride
// L/(Ns*R), and float reciprocals for the division-free index math.
struct FftPass {
  int radix, nbf, ns, twstep;
  float inv_nbf, inv_ns;
};
// fills pass[0..npass) for a line of length L; returns npass
int fft_plan_passes(int L, FftPass *pass, int maxpass);

// Lines along one axis of a row-major array.  Line (u, v) starts at
// U(u)*st_u + V(v)*st_v, element k at + k*st_ax, where the virtual indices
// map to actual ones through at most two ranges:
// U(u) = u < un1 ? ua1 + u : ua2 + (u - un1)   (same for V).
struct FftLaunch {
  void *data;
  const void *tw;  // exp(-2 pi i m / L), m < L, compute type
  int L, LP, Bl, npass, sign;
  FftPass pass[24];
  int64_t st_ax, st_u, st_v;
  int nu, nv;
  int un1, ua1, ua2, vn1, va1, va2;
  // elements kn1 <= k < ka2 are read as zero (the input is known to vanish
  // there)
  int kprune, kn1, ka2;
  // elements sn1 <= k < sa2 of the result are not stored (no later pass
  // reads them)
  int sprune, sn1, sa2;
};

// Crop + deconvolution of type 1 fused into the last column pass of a 2-D
// transform: the retained mode at grid (row, column) goes to the uniform
// array (strides ust0, ust1 in elements; complex128 if u_f64), times the
// FP64 correction factors cf0[|n0/2 - i0|] * cf1[|n1/2 - i1|]
// (comp_indices, nufft.h:142-151).
struct FftCrop {
  void *uni;
  const double *cf0, *cf1;
  int64_t ust0, ust1;
  int n0, N0, n1, N1;
  int fft_order, u_f64;
};

// One half of the two-pass FFT along a strided axis (see k_fft_cols).
struct FftCols {
  const void *src;
  void *dst;
  const void *tw_sub;   // twiddles of length Lsub
  const void *tw_full;  // twiddles of length L (inter-pass factors)
  int L, Lsub, npass, sign, twiddle, Bl;
  FftPass pass[24];
  int64_t ia, ib, oa, ob;  // row maps of input / output (in rows)
  int64_t rs;              // row stride (elements)
  int ncol, vn1, va1, va2; // columns (virtual -> actual, stride 1)
  int nlines;
  int nouter, on1, oa1, oa2;
  int64_t st_o;
  int sprune, sn1, sa2;  // output rows sn1 <= row < sa2 are not stored
  int kprune, kn1, ka2;  // input rows kn1 <= row < ka2 are read as zero
  int crop;              // store through `cr` instead of dst (pass B)
  FftCrop cr;
};

// FFT of length L = L1*L2 along a strided axis, as two coalesced passes:
// A: length-L1 FFTs over rows n1*L2+n2, times w_L^{n2 k1}, in place (or,
// with `a_dst`, from p.data into a_dst at the same positions);
// B: length-L2 FFTs over rows k1*L2+n2 of A's output, written to row
// k1 + L1*k2 of dst, i.e. in natural order (dst may be p.data when a_dst is
// given: the result then replaces the input; with `crop`, pass B writes the
// retained modes to the uniform array instead, see FftCrop).  Returns false
// when L has no split worth using.
bool fft_columns_2pass(const FftLaunch &p, int L1, void *dst, bool fp64,
                       const void *tw_L1, const void *tw_L2,
                       const void *tw_L, hipStream_t s,
                       void *a_dst = nullptr, const FftCrop *crop = nullptr);
// a divisor of L near sqrt(L) with both factors <= 128 (0: none)
int fft_split(int L);

void fft_twiddles(int L, bool fp64, std::vector<unsigned char> &out);
// the same table computed on the device into dst (L complex values of the
// compute type), for callers that make no host-to-device copies
void launch_fft_twiddles(void *dst, int L, bool fp64, hipStream_t s);
// Runs the FFT over every line of the set (in place).  `scratch` must be as
// large as the array when L exceeds the LDS capacity.
void fft_lineset(const FftLaunch &p, bool fp64, void *scratch, hipStream_t s);

}  // namespace dmx
