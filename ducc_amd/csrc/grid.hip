// grid.hip -- correction factors, the fused deconvolve + zero-pad of type 2
// and the fused crop + deconvolve of type 1 (nufft.h:1299-1312, 1331-1344,
// comp_indices at nufft.h:142-151), plus the direct FP64 NUDFT used as the
// spot-check of the parameter chooser (the sum of test_nufft.py:32-43).
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

struct AxisMap {
  int n, N, fft_order;
};

__device__ __forceinline__ void comp_idx(int idx, const AxisMap &m, int &icf,
                                         int &iuser, int &igrid) {
  int h = m.n / 2;
  icf = h - idx;
  if (icf < 0) icf = -icf;
  int i1 = m.fft_order ? m.n - h + idx : idx;
  if (i1 >= m.n) i1 -= m.n;
  iuser = i1;
  int i2 = m.N - h + idx;
  if (i2 >= m.N) i2 -= m.N;
  igrid = i2;
}

// mode: 0 = crop (grid -> uniform), 1 = pad (uniform -> grid)
template <typename T, int MODE>
__global__ void k_grid_copy(GridCopyArgs a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n1 = a.n[1], n2 = a.n[2];
  if (q >= (int64_t)a.n[0] * n1 * n2) return;
  const int i = (int)(q / (n1 * n2));
  const int j = (int)((q / n2) % n1);
  const int k = (int)(q % n2);
  int c[3], u[3], g[3];
  AxisMap m0{a.n[0], a.N[0], a.fft_order}, m1{a.n[1], a.N[1], a.fft_order},
      m2{a.n[2], a.N[2], a.fft_order};
  comp_idx(i, m0, c[0], u[0], g[0]);
  comp_idx(j, m1, c[1], u[1], g[1]);
  comp_idx(k, m2, c[2], u[2], g[2]);
  double f = a.cf0[c[0]] * a.cf1[c[1]] * a.cf2[c[2]];
  const int64_t gi = ((int64_t)g[0] * a.N[1] + g[1]) * a.N[2] + g[2];
  const int64_t ui = (int64_t)u[0] * a.ust[0] + (int64_t)u[1] * a.ust[1] +
                     (int64_t)u[2] * a.ust[2];
  T *grid = (T *)a.grid;
  if (MODE == 0) {
    T re = grid[2 * gi] * (T)f, im = grid[2 * gi + 1] * (T)f;
    if (a.u_f64)
      ((double2 *)a.uni)[ui] = make_double2((double)re, (double)im);
    else
      ((float2 *)a.uni)[ui] = make_float2((float)re, (float)im);
  } else {
    T re, im;
    if (a.u_f64) {
      double2 x = ((const double2 *)a.uni)[ui];
      re = (T)x.x; im = (T)x.y;
    } else {
      float2 x = ((const float2 *)a.uni)[ui];
      re = (T)x.x; im = (T)x.y;
    }
    grid[2 * gi] = re * (T)f;
    grid[2 * gi + 1] = im * (T)f;
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

void launch_grid_copy(bool fp64, bool pad, const GridCopyArgs &a,
                      hipStream_t s) {
  int64_t n = (int64_t)a.n[0] * a.n[1] * a.n[2];
  if (!n) return;
  unsigned nb = (unsigned)((n + 255) / 256);
  if (fp64) {
    if (pad) hipLaunchKernelGGL((k_grid_copy<double, 1>), dim3(nb), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_grid_copy<double, 0>), dim3(nb), dim3(256), 0, s, a);
  } else {
    if (pad) hipLaunchKernelGGL((k_grid_copy<float, 1>), dim3(nb), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_grid_copy<float, 0>), dim3(nb), dim3(256), 0, s, a);
  }
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
