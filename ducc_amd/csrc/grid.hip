// grid.hip -- correction factors, the fused deconvolve + zero-pad of type 2
// and the fused crop + deconvolve of type 1 (nufft.h:1299-1312, 1331-1344,
// comp_indices at nufft.h:142-151), plus the direct FP64 NUDFT used as the
// spot-check of the parameter chooser (the sum of test_nufft.py:32-43).
#include <algorithm>
#include <type_traits>

#include "dmx.h"
#include "grid.h"

namespace dmx {

namespace {

// corfac[k] = 1 / sum_i wpsi_i cos(pi W x_i k dx), k < n  (FP64 on the GPU;
// gridding_kernel.h:77-109)
__global__ void k_corfac(CorfacArgs a, double *out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const double pi = 3.141592653589793238462643383279502884197;
  double v = (double)k * a.dx, s = 0;
  for (int i = 0; i < a.nq; ++i) s += a.wpsi[i] * cos((pi * a.W * a.x[i]) * v);
  out[k] = 1.0 / s;
}

struct ConstPayload {
  double v[CONST_MAX];
};
__global__ void k_store_consts(void *dst, ConstPayload pl, int n,
                               int as_f32) {
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i >= n) return;
  if (as_f32)
    ((float *)dst)[i] = (float)pl.v[i];
  else
    ((double *)dst)[i] = pl.v[i];
}

// Position p along one axis of the copied region -> correction index icf,
// index into the uniform array iu, grid index ig (comp_indices,
// nufft.h:142-151: mode idx sits at grid index idx - n/2 mod N and at
// uniform index idx, or idx + n - n/2 mod n in FFT order).  With `full`, p
// is the grid index itself and the result says whether a mode sits there.
__device__ __forceinline__ bool axis_pos(int p, int n, int N, int full,
                                         int fft_order, int &icf, int &iu,
                                         int &ig) {
  const int h = n / 2;
  int idx;
  if (full) {
    ig = p;
    if (p < n - h) idx = p + h;
    else if (p >= N - h) idx = p - (N - h);
    else {
      // no mode here: in-range indices all the same, as callers may still
      // form their addresses
      icf = 0;
      iu = 0;
      return false;
    }
  } else {
    idx = p;
    ig = idx + N - h;
    if (ig >= N) ig -= N;
  }
  icf = h - idx;
  if (icf < 0) icf = -icf;
  iu = fft_order ? idx + n - h : idx;
  if (iu >= n) iu -= n;
  return true;
}

constexpr int GC_EL = 4;  // elements per lane, 256 apart

// uniform element <-> compute-type complex
template <typename T, typename U>
__device__ __forceinline__ void load_uni(const void *uni, int64_t i, T &re,
                                         T &im) {
  const U x = ((const U *)uni)[i];
  re = (T)x.x;
  im = (T)x.y;
}

// mode: 0 = crop (grid -> uniform), 1 = pad (uniform -> grid).  U: the
// uniform array's complex type.  One workgroup per (row of the first two
// axes, run of 256 * GC_EL along the last axis) of the region: the first two
// axes' index math is done once per row, lanes take consecutive elements
// of the last axis, and each lane issues its GC_EL loads before its stores.
// Axes are right-aligned (padded axes first).
template <typename T, typename U, int MODE>
__global__ __launch_bounds__(256) void k_grid_copy(GridCopyArgs a,
                                                   uint32_t nblocks,
                                                   uint32_t nchunk) {
  using C = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
  const uint32_t R1 = (uint32_t)(a.full[1] ? a.N[1] : a.n[1]);
  const int R2 = a.full[2] ? a.N[2] : a.n[2];
  C *grid = (C *)a.grid;
  for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const uint32_t row = b / nchunk;
    const int base = (int)(b - row * nchunk) * (256 * GC_EL) + (int)threadIdx.x;
    const uint32_t p0 = row / R1, p1 = row - p0 * R1;
    int c0, u0, g0, c1, u1, g1;
    bool keep01 =
        axis_pos((int)p0, a.n[0], a.N[0], a.full[0], a.fft_order, c0, u0, g0);
    keep01 &=
        axis_pos((int)p1, a.n[1], a.N[1], a.full[1], a.fft_order, c1, u1, g1);
    const double f01 = a.cf[0][c0] * a.cf[1][c1];  // c0, c1 always set
    const int64_t growb = ((int64_t)g0 * a.N[1] + g1) * a.N[2];
    const int64_t urow = (int64_t)u0 * a.ust[0] + (int64_t)u1 * a.ust[1];
    T re[GC_EL], im[GC_EL];
    int64_t gi[GC_EL];
    bool ok[GC_EL], keep[GC_EL];
#pragma unroll
    for (int e = 0; e < GC_EL; ++e) {
      const int p2 = base + e * 256;
      ok[e] = p2 < R2;
      keep[e] = false;
      gi[e] = 0;
      re[e] = im[e] = T(0);
      if (!ok[e]) continue;
      int c2, u2, g2;
      // (evaluated even where keep01 is false: g2 is needed for the zero)
      const bool k2 =
          axis_pos(p2, a.n[2], a.N[2], a.full[2], a.fft_order, c2, u2, g2);
      keep[e] = keep01 && k2;
      gi[e] = growb + g2;
      if (!keep[e]) continue;
      const T f = (T)(f01 * a.cf[2][c2]);
      const int64_t ui = urow + (int64_t)u2 * a.ust[2];
      if (MODE == 0) {
        const C g = grid[gi[e]];
        re[e] = g.x * f;
        im[e] = g.y * f;
        gi[e] = ui;  // now the destination
      } else {
        load_uni<T, U>(a.uni, ui, re[e], im[e]);
        re[e] *= f;
        im[e] *= f;
      }
    }
#pragma unroll
    for (int e = 0; e < GC_EL; ++e) {
      if (!ok[e]) continue;
      if (MODE == 0) {
        U o;
        o.x = re[e];
        o.y = im[e];
        ((U *)a.uni)[gi[e]] = o;
      } else {
        C g;  // zero where no mode sits (re = im = 0)
        g.x = re[e];
        g.y = im[e];
        grid[gi[e]] = g;
      }
    }
  }
}

__global__ void k_nudft(NudftArgs a) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double sh[256 * 5];
  const double pi = 3.141592653589793238462643383279502884197;
  const double sg = a.forward ? -1.0 : 1.0;
  double accr = 0, acci = 0;
  if (a.type == 1) {
    // one thread per selected mode, loop over all points
    double kk[3] = {0, 0, 0};
    if (q < a.nsel) {
      int64_t flat = a.sel[q];
      int64_t n1 = a.n[1], n2 = a.n[2];
      int64_t uu[3] = {flat / (n1 * n2), (flat / n2) % n1, flat % n2};
      for (int d = 0; d < a.ndim; ++d) {
        int64_t n = a.n[d];
        int64_t idx = a.fft_order ? (uu[d] + n / 2) % n : uu[d];
        kk[d] = (double)(idx - n / 2) * (2 * pi / a.period[d]);
      }
    }
    // this block's share of the points (blockIdx.y of gridDim.y chunks)
    const int64_t chunk = (a.M + gridDim.y - 1) / gridDim.y;
    const int64_t jlo = (int64_t)blockIdx.y * chunk;
    const int64_t jhi = min(a.M, jlo + chunk);
    for (int64_t base = jlo; base < jhi; base += 256) {
      __syncthreads();
      int64_t jp = base + threadIdx.x;
      if (jp < jhi) {
        for (int d = 0; d < a.ndim; ++d)
          sh[threadIdx.x * 5 + d] = a.coord[jp * a.ndim + d];
        sh[threadIdx.x * 5 + 3] = a.vals[2 * jp];
        sh[threadIdx.x * 5 + 4] = a.vals[2 * jp + 1];
      }
      __syncthreads();
      int cnt = (int)min((int64_t)256, jhi - base);
      if (q < a.nsel)
        for (int t = 0; t < cnt; ++t) {
          double ph = 0;
          for (int d = 0; d < a.ndim; ++d) ph += kk[d] * sh[t * 5 + d];
          double s, c;
          sincos(sg * ph, &s, &c);
          double vr = sh[t * 5 + 3], vi = sh[t * 5 + 4];
          accr += vr * c - vi * s;
          acci += vr * s + vi * c;
        }
    }
  } else {
    // one thread per selected point, loop over all modes
    double x[3] = {0, 0, 0};
    if (q < a.nsel) {
      int64_t j = a.sel[q];
      for (int d = 0; d < a.ndim; ++d)
        x[d] = a.coord[j * a.ndim + d] * (2 * pi / a.period[d]);
    }
    int64_t ntot = (int64_t)a.n[0] * a.n[1] * a.n[2];
    int64_t n1 = a.n[1], n2 = a.n[2];
    const int64_t chunk = (ntot + gridDim.y - 1) / gridDim.y;
    const int64_t flo = (int64_t)blockIdx.y * chunk;
    const int64_t fhi = min(ntot, flo + chunk);
    if (q < a.nsel)
      for (int64_t flat = flo; flat < fhi; ++flat) {
        int64_t uu[3] = {flat / (n1 * n2), (flat / n2) % n1, flat % n2};
        double ph = 0;
        for (int d = 0; d < a.ndim; ++d) {
          int64_t n = a.n[d];
          int64_t idx = a.fft_order ? (uu[d] + n / 2) % n : uu[d];
          ph += (double)(idx - n / 2) * x[d];
        }
        double s, c;
        sincos(sg * ph, &s, &c);
        double vr = a.vals[2 * flat], vi = a.vals[2 * flat + 1];
        accr += vr * c - vi * s;
        acci += vr * s + vi * c;
      }
  }
  if (q < a.nsel) {
    if (gridDim.y == 1) {
      a.out[2 * q] = accr;
      a.out[2 * q + 1] = acci;
    } else {  // partial sums of the chunks (out zeroed by the launcher)
      atomicAdd(&a.out[2 * q], accr);
      atomicAdd(&a.out[2 * q + 1], acci);
    }
  }
}

__global__ void k_gather_c128(const double2 *src, const int64_t *idx,
                              size_t n, double2 *dst) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

}  // namespace

void launch_gather_c128(const void *src, const int64_t *idx, size_t n,
                        void *dst, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_c128, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, (const double2 *)src, idx, n,
                     (double2 *)dst);
  DMX_HIP(hipGetLastError());
}

void launch_corfac(const CorfacArgs &a, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_corfac, dim3((a.n + 255) / 256), dim3(256), 0, s, a,
                     out);
  DMX_HIP(hipGetLastError());
}

void launch_store_consts(void *dst, const double *vals, int n, bool as_f32,
                         hipStream_t s) {
  DMX_ASSERT(n >= 0 && n <= CONST_MAX, "too many constants for one launch");
  if (!n) return;
  ConstPayload pl{};
  for (int i = 0; i < n; ++i) pl.v[i] = vals[i];
  hipLaunchKernelGGL(k_store_consts, dim3((n + 255) / 256), dim3(256), 0, s,
                     dst, pl, n, as_f32 ? 1 : 0);
  DMX_HIP(hipGetLastError());
}

void launch_grid_copy(bool fp64, bool pad, const GridCopyArgs &a0,
                      hipStream_t s) {
  // right-align the axes: padded axes (n = N = 1) come last in the
  // arguments, the kernel wants the contiguous one last
  GridCopyArgs a = a0;
  int shift = 0;
  while (shift < 2 && a0.N[2 - shift] == 1 && a0.n[2 - shift] == 1) ++shift;
  for (int d = 0; d < 3; ++d) {
    const int e = d - shift;
    a.n[d] = e >= 0 ? a0.n[e] : 1;
    a.N[d] = e >= 0 ? a0.N[e] : 1;
    a.full[d] = e >= 0 ? (pad ? a0.full[e] : 0) : 0;
    a.ust[d] = e >= 0 ? a0.ust[e] : 0;
    a.cf[d] = e >= 0 ? a0.cf[e] : a0.cf[2];  // a padded axis's single 1.0
  }
  uint64_t R[3];
  for (int d = 0; d < 3; ++d) R[d] = (uint64_t)(a.full[d] ? a.N[d] : a.n[d]);
  if (!R[0] || !R[1] || !R[2]) return;
  const uint64_t nchunk = (R[2] + 256 * GC_EL - 1) / (256 * GC_EL);
  const uint64_t nb64 = R[0] * R[1] * nchunk;
  DMX_ASSERT(nb64 < (1ull << 32), "grid too large for the copy kernel");
  const uint32_t nblocks = (uint32_t)nb64;
  const unsigned grid = std::min<uint32_t>(nblocks, 1u << 20);
#define DMX_GC(T, U, M)                                                    \
  hipLaunchKernelGGL((k_grid_copy<T, U, M>), dim3(grid), dim3(256), 0, s, a, \
                     nblocks, (uint32_t)nchunk)
  if (fp64) {
    if (a.u_f64) { if (pad) DMX_GC(double, double2, 1); else DMX_GC(double, double2, 0); }
    else { if (pad) DMX_GC(double, float2, 1); else DMX_GC(double, float2, 0); }
  } else {
    if (a.u_f64) { if (pad) DMX_GC(float, double2, 1); else DMX_GC(float, double2, 0); }
    else { if (pad) DMX_GC(float, float2, 1); else DMX_GC(float, float2, 0); }
  }
#undef DMX_GC
  DMX_HIP(hipGetLastError());
}

void launch_nudft(const NudftArgs &a, hipStream_t s) {
  if (!a.nsel) return;
  const unsigned bx = (unsigned)((a.nsel + 255) / 256);
  // split the inner sum (chunks of at least 256 terms) so that about 4096
  // workgroups run
  const int64_t inner = a.type == 1 ? a.M : a.n[0] * a.n[1] * a.n[2];
  int64_t ny = std::max<int64_t>(1, std::min<int64_t>(4096 / bx, inner / 256));
  ny = std::min<int64_t>(ny, 65535);
  if (ny > 1) launch_zero(a.out, (size_t)a.nsel * 16, s);
  hipLaunchKernelGGL(k_nudft, dim3(bx, (unsigned)ny), dim3(256), 0, s, a);
  DMX_HIP(hipGetLastError());
}

}  // namespace dmx
