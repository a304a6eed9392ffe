// wgridder.hip -- radio-interferometric (de)gridding with the w-term
// (role of ducc0.wgridder vis2dirty / dirty2vis, src/ducc0/wgridder/
// wgridder.h:333-1700), built on the NUFFT.
//
// Visibility (row, chan) sits at (u, v, w) = uvw[row] * freq[chan] / c
// (signs flipped by flip_u/v/w); pixel (i, j) at x = lshift + (i -
// nx/2) px, y = mshift + (j - ny/2) py with n = sqrt(1 - x^2 - y^2).
//   vis2dirty: dirty = Re(sum vis wgt e^{2 pi i (x u + y v - w (n - 1))}) / n
//   dirty2vis: vis = wgt sum dirty / n e^{-2 pi i (x u + y v - w (n - 1))}
// Without w-gridding the w term and the division by n are absent.
//
// The phase centre shift and nshift = -(min + max)/2 of n - 1 over the
// image move into a per-visibility phase e^{2 pi i (lshift u + mshift v +
// nshift w)} (wgridder.h:1121-1133), so that with
// m(i, j) = n - 1 + nshift,
//   x u + y v - w (n - 1) = (i - nx/2) px u + (j - ny/2) py v - w m + shift.
// The x / y part is a 2-D NUFFT with modes i - nx/2, j - ny/2 at the
// coordinates (2 pi px u, 2 pi py v): type 1 for gridding, type 2 for
// degridding.  The w part is w-stacking with the same ES kernel psi along w
// (wgridder.h:1429-1459, 1465-1496): planes w_p = wmin + p dw,
//   e^{-2 pi i w m} ~= corfunc(m dw) sum_p psi((w - w_p) / dw) e^{-2 pi i w_p m}
// (Poisson summation; corfunc = 1 / Fourier transform of psi, the kernel's
// correction function, wgridder.h:1361-1371), dw = 0.5 / ofactor /
// max|m|.  Each visibility touches the W planes of its kernel support:
// they are sorted once by first plane (stable radix sort), so every plane's
// visibilities are one contiguous range, and each plane is one 2-D NUFFT of
// that range with the visibilities weighted by psi, whose image is
// multiplied by the plane's w-screen e^{-2 pi i w_p m} and accumulated
// (gridding); degridding runs the same steps transposed.  The kernel row
// comes from ducc0's 3-D table (as wgridder.h:1522 chooses for w-gridding):
// one row for both the uv grid and the w planes.
#include <cmath>
#include <cstring>
#include <vector>

#include "grid.h"
#include "params.h"
#include "runtime.h"
#include "wgridder.h"

namespace dmx {

namespace {

constexpr double SPEED_OF_LIGHT = 299792458.;

__device__ __forceinline__ unsigned long long ordered_key(double x) {
  const long long b = __double_as_longlong(x);
  return b < 0 ? ~(unsigned long long)b
               : ((unsigned long long)b | 0x8000000000000000ull);
}
__device__ __forceinline__ void wave_minmax(double &lo, double &hi) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o));
    hi = fmax(hi, __shfl_xor(hi, o));
  }
}

// e^{2 pi i phi}, phi reduced to [0, 1) first (exact in double)
__device__ __forceinline__ double2 expi2pi(double phi) {
  const double r = phi - floor(phi);
  double s, c;
  sincospi(2.0 * r, &s, &c);
  return make_double2(c, s);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// n - 1 of pixel (i, j) (accurate near zero, wgridder.h:1363-1364; beyond
// the horizon -sqrt(x^2 + y^2 - 1) - 1) and whether n > 0
__device__ __forceinline__ double pixel_nm1(const WgArgs &a, int i, int j,
                                            bool &inside) {
  const double x = a.lshift + (i - a.nx / 2) * a.px;
  const double y = a.mshift + (j - a.ny / 2) * a.py;
  const double r2 = x * x + y * y, t = 1.0 - r2;
  inside = t >= 0.0;
  return inside ? -r2 / (sqrt(t) + 1.0) : -sqrt(-t) - 1.0;
}

// 1 / Fourier transform of the kernel at frequency v (cycles per cell)
__device__ __forceinline__ double corfunc(const CorfacArgs &c, double v) {
  const double pi = 3.141592653589793238462643383279502884197;
  double s = 0.0;
  for (int k = 0; k < c.nq; ++k) s += c.wpsi[k] * cos((pi * c.W * c.x[k]) * v);
  return 1.0 / s;
}

// per visibility: NUFFT coordinates, w, activity, and (gridding) the value
// times weight and shift phase; min / max of w over the active ones and
// their count (grid-stride, reduced per workgroup: one atomic per
// workgroup and quantity)
constexpr int PREP_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_wg_prep(WgArgs a, WgPoints o) {
  __shared__ double slo[4], shi[4];
  __shared__ unsigned long long scnt[4];
  const size_t M = a.nrow * a.nchan;
  double wlo = 1e300, whi = -1e300;
  unsigned long long cnt = 0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < M;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t r = q / a.nchan, c = q - r * a.nchan;
    const double f = a.fovc[c];
    const double u = a.uvw[3 * r] * a.ufac * f, v = a.uvw[3 * r + 1] * a.vfac * f,
                 w = a.uvw[3 * r + 2] * a.wfac * f;
    bool active = !a.mask || a.mask[r * a.mst0 + c * a.mst1] != 0;
    double wt = 1.0;
    if (a.wgt) {
      wt = a.wgt_f64 ? ((const double *)a.wgt)[r * a.wst0 + c * a.wst1]
                     : (double)((const float *)a.wgt)[r * a.wst0 + c * a.wst1];
      active = active && wt != 0.0;
    }
    double2 val = make_double2(0.0, 0.0);
    if (a.gridding) {
      const int64_t vi = (int64_t)r * a.vst0 + (int64_t)c * a.vst1;
      if (a.vis_f64) {
        val = ((const double2 *)a.vis)[vi];
      } else {
        const float2 z = ((const float2 *)a.vis)[vi];
        val = make_double2(z.x, z.y);
      }
      active = active && (val.x != 0.0 || val.y != 0.0);
      const double2 ph =
          expi2pi(a.lshift * u + a.mshift * v + a.nshift * w);
      val = cmul(make_double2(val.x * wt, val.y * wt), ph);
    }
    o.coord[2 * q] = 2.0 * M_PI * a.px * u;
    o.coord[2 * q + 1] = 2.0 * M_PI * a.py * v;
    o.w[q] = w;
    o.active[q] = active ? 1 : 0;
    if (!active) val = make_double2(0.0, 0.0);
    if (a.gridding) ((double2 *)o.val)[q] = val;
    if (active) {
      wlo = fmin(wlo, w);
      whi = fmax(whi, w);
      ++cnt;
    }
  }
  wave_minmax(wlo, whi);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    slo[wave] = wlo;
    shi[wave] = whi;
    scnt[wave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      wlo = fmin(wlo, slo[k]);
      whi = fmax(whi, shi[k]);
      cnt += scnt[k];
    }
    if (cnt) {
      atomicMin(&o.wminmax[0], ordered_key(wlo));
      atomicMax(&o.wminmax[1], ordered_key(whi));
      atomicAdd(&o.wminmax[2], cnt);
    }
  }
}

// first w-plane of every visibility (inactive ones: nplanes, sorted last)
__global__ void k_wg_keys(const WgPoints o, size_t M, WgPlanes pl,
                          uint32_t *keys) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= M) return;
  uint32_t k = (uint32_t)pl.nplanes;
  if (o.active[q]) {
    const double uw = (o.w[q] - pl.wmin) / pl.dw;
    int i0 = (int)floor(uw - 0.5 * pl.W + 1.0);
    i0 = i0 < 0 ? 0 : (i0 > pl.nplanes - pl.W ? pl.nplanes - pl.W : i0);
    k = (uint32_t)i0;
  }
  keys[q] = k;
}

// the visibilities in first-plane order
__global__ void k_wg_gather(const WgPoints o, const uint32_t *idx, size_t M,
                            WgPoints s) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const size_t q = idx[i];
  s.coord[2 * i] = o.coord[2 * q];
  s.coord[2 * i + 1] = o.coord[2 * q + 1];
  s.w[i] = o.w[q];
  if (o.val) ((double2 *)s.val)[i] = ((const double2 *)o.val)[q];
}

// psi weight of plane p for the visibility at w with first plane i0
__device__ __forceinline__ double plane_weight(const WgPlanes &pl, double w,
                                               uint32_t i0, int p) {
  const double uw = (w - pl.wmin) / pl.dw;
  const double frac = uw - (double)i0;
  return piece_weight<double>((double)(pl.W - 1) - 2.0 * frac, pl.coef, pl.W,
                              p - (int)i0);
}

// gridding, plane p: vals[i - lo] = val_s[i] * psi(w_i - w_p), in the
// compute type (out_f64)
__global__ void k_wg_plane_vals(const WgPoints s, const uint32_t *keys,
                                size_t lo, size_t hi, WgPlanes pl, int p,
                                void *vals, int out_f64) {
  const size_t i = lo + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const double wt = plane_weight(pl, s.w[i], keys[i], p);
  const double2 v = ((const double2 *)s.val)[i];
  if (out_f64)
    ((double2 *)vals)[i - lo] = make_double2(v.x * wt, v.y * wt);
  else
    ((float2 *)vals)[i - lo] =
        make_float2((float)(v.x * wt), (float)(v.y * wt));
}

// gridding, plane p: acc += image_p * e^{-2 pi i w_p m}
__global__ void k_wg_accum_image(WgArgs a, const void *img, int img_f64,
                                 double wp, double2 *acc) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ny) return;
  bool inside;
  const double m = pixel_nm1(a, i, j, inside) + a.nshift;
  const size_t k = (size_t)i * a.ny + j;
  double2 z;
  if (img_f64) {
    z = ((const double2 *)img)[k];
  } else {
    const float2 f = ((const float2 *)img)[k];
    z = make_double2(f.x, f.y);
  }
  const double2 r = cmul(z, expi2pi(-wp * m));
  acc[k].x += r.x;
  acc[k].y += r.y;
}

// gridding, last step: dirty = Re(acc) * corfunc(m dw) / n (w-gridding) or
// Re(image) (without)
__global__ void k_wg_finish_dirty(WgArgs a, const void *src, int src_f64,
                                  int wgrid, CorfacArgs cw, double dw,
                                  void *dirty, int dirty_f64, int64_t dst0,
                                  int64_t dst1) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ny) return;
  const size_t k = (size_t)i * a.ny + j;
  double re = src_f64 ? ((const double2 *)src)[k].x
                      : (double)((const float2 *)src)[k].x;
  if (wgrid) {
    bool inside;
    const double nm1 = pixel_nm1(a, i, j, inside);
    double fct = 0.0;
    if (inside) {
      fct = corfunc(cw, (nm1 + a.nshift) * dw);
      if (a.divide_by_n) fct /= nm1 + 1.0;
    } else if (!a.divide_by_n) {
      fct = corfunc(cw, (nm1 + a.nshift) * dw);
    }
    re *= fct;
  }
  const int64_t di = (int64_t)i * dst0 + (int64_t)j * dst1;
  if (dirty_f64)
    ((double *)dirty)[di] = re;
  else
    ((float *)dirty)[di] = (float)re;
}

// degridding, first step: the corrected image (FP64, contiguous):
// dirty * corfunc(m dw) / n (w-gridding) or dirty (without)
__global__ void k_wg_prep_image(WgArgs a, const void *dirty, int dirty_f64,
                                int64_t ds0, int64_t ds1, int wgrid,
                                CorfacArgs cw, double dw, double *img) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ny) return;
  const int64_t di = (int64_t)i * ds0 + (int64_t)j * ds1;
  double d = dirty_f64 ? ((const double *)dirty)[di]
                       : (double)((const float *)dirty)[di];
  if (wgrid) {
    bool inside;
    const double nm1 = pixel_nm1(a, i, j, inside);
    double fct = 0.0;
    if (inside) {
      fct = corfunc(cw, (nm1 + a.nshift) * dw);
      if (a.divide_by_n) fct /= nm1 + 1.0;
    } else if (!a.divide_by_n) {
      fct = corfunc(cw, (nm1 + a.nshift) * dw);
    }
    d *= fct;
  }
  img[(size_t)i * a.ny + j] = d;
}

// degridding, plane p (wp; or the whole image with wgrid == 0): complex
// image in the compute type, img * e^{+2 pi i w_p m}
__global__ void k_wg_plane_image(WgArgs a, const double *img, int wgrid,
                                 double wp, void *out, int out_f64) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ny) return;
  const size_t k = (size_t)i * a.ny + j;
  double2 z = make_double2(img[k], 0.0);
  if (wgrid) {
    bool inside;
    const double m = pixel_nm1(a, i, j, inside) + a.nshift;
    const double2 e = expi2pi(wp * m);
    z = make_double2(img[k] * e.x, img[k] * e.y);
  }
  if (out_f64)
    ((double2 *)out)[k] = z;
  else
    ((float2 *)out)[k] = make_float2((float)z.x, (float)z.y);
}

// degridding, plane p: acc_s[i] += res[i - lo] * psi(w_i - w_p)
__global__ void k_wg_plane_accum_vis(const WgPoints s, const uint32_t *keys,
                                     size_t lo, size_t hi, WgPlanes pl, int p,
                                     const void *res, int res_f64,
                                     double2 *acc) {
  const size_t i = lo + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const double wt = plane_weight(pl, s.w[i], keys[i], p);
  double2 z;
  if (res_f64) {
    z = ((const double2 *)res)[i - lo];
  } else {
    const float2 f = ((const float2 *)res)[i - lo];
    z = make_double2(f.x, f.y);
  }
  acc[i].x += z.x * wt;
  acc[i].y += z.y * wt;
}

// degridding, last step: vis[row, chan] = value * wgt * e^{-2 pi i (lshift
// u + mshift v + nshift w)} (active), 0 (inactive); value from the sorted
// accumulator through the permutation (w-gridding) or in input order
__global__ void k_wg_finish_vis(WgArgs a, const WgPoints o, const void *res,
                                int res_f64, const uint32_t *idx, size_t M,
                                void *vis, int vis_f64, int64_t vs0,
                                int64_t vs1) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const size_t q = idx ? idx[i] : i;
  const size_t r = q / a.nchan, c = q - r * a.nchan;
  double2 z = make_double2(0.0, 0.0);
  if (o.active[q]) {
    if (res_f64) {
      z = ((const double2 *)res)[i];
    } else {
      const float2 f = ((const float2 *)res)[i];
      z = make_double2(f.x, f.y);
    }
    const double f = a.fovc[c];
    const double u = a.uvw[3 * r] * a.ufac * f, v = a.uvw[3 * r + 1] * a.vfac * f,
                 w = a.uvw[3 * r + 2] * a.wfac * f;
    z = cmul(z, expi2pi(-(a.lshift * u + a.mshift * v + a.nshift * w)));
    if (a.wgt) {
      const double wt =
          a.wgt_f64 ? ((const double *)a.wgt)[r * a.wst0 + c * a.wst1]
                    : (double)((const float *)a.wgt)[r * a.wst0 + c * a.wst1];
      z.x *= wt;
      z.y *= wt;
    }
  }
  const int64_t vi = (int64_t)r * vs0 + (int64_t)c * vs1;
  if (vis_f64)
    ((double2 *)vis)[vi] = z;
  else
    ((float2 *)vis)[vi] = make_float2((float)z.x, (float)z.y);
}

unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
double from_key(unsigned long long k) {
  const long long b = (k & 0x8000000000000000ull)
                          ? (long long)(k & 0x7FFFFFFFFFFFFFFFull)
                          : (long long)~k;
  double x;
  std::memcpy(&x, &b, 8);
  return x;
}

// n - 1 over the image extent (wgridder.h:1503-1520)
void nm1_range(const WgArgs &a, double &lo, double &hi) {
  const double xmin = a.lshift - 0.5 * a.nx * a.px,
               xmax = xmin + (a.nx - 1) * a.px,
               ymin = a.mshift - 0.5 * a.ny * a.py,
               ymax = ymin + (a.ny - 1) * a.py;
  std::vector<double> xe{xmin, xmax}, ye{ymin, ymax};
  if (xmin * xmax < 0) xe.push_back(0);
  if (ymin * ymax < 0) ye.push_back(0);
  lo = 1e300;
  hi = -1e300;
  for (double x : xe)
    for (double y : ye) {
      const double t = x * x + y * y;
      const double v = t <= 1. ? std::sqrt(1. - t) - 1. : -std::sqrt(t - 1.) - 1.;
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
}

}  // namespace

void WgJob::setup(const double *freq_h, size_t nchan_, size_t nrow_,
                  hipStream_t s_) {
  s = s_;
  a.nrow = nrow_;
  a.nchan = nchan_;
  M = nrow_ * nchan_;
  std::vector<double> fovc(nchan_);
  for (size_t i = 0; i < nchan_; ++i) {
    DMX_ASSERT(freq_h[i] > 0, "negative channel frequency encountered");
    if (i > 0)
      DMX_ASSERT(freq_h[i] >= freq_h[i - 1],
                 "channel frequencies must be sorted in ascending order");
    fovc[i] = freq_h[i] / SPEED_OF_LIGHT;
  }
  DMX_ASSERT(nrow_ < ((size_t)1 << 32), "too many rows in the MS");
  DMX_ASSERT(nchan_ < ((size_t)1 << 16), "too many channels in the MS");
  DMX_ASSERT(M < 0xFFFFFFFFull, "too many visibilities");
  DMX_ASSERT(a.nx % 2 == 0, "nx_dirty must be even");
  DMX_ASSERT(a.ny % 2 == 0, "ny_dirty must be even");
  DMX_ASSERT(a.px > 0, "pixsize_x must be positive");
  DMX_ASSERT(a.py > 0, "pixsize_y must be positive");
  fovc_.stream = s;
  fovc_.reset(std::max<size_t>(nchan_, 1) * 8);
  DMX_HIP(hipMemcpyAsync(fovc_.p, fovc.data(), nchan_ * 8,
                         hipMemcpyHostToDevice, s));
  DMX_HIP(hipStreamSynchronize(s));
  a.fovc = fovc_.as<double>();
  if (a.wgrid) {
    double lo, hi;
    nm1_range(a, lo, hi);
    nm1min = lo;
    nm1max = hi;
    a.nshift = a.allow_nshift ? -0.5 * (hi + lo) : 0.0;
  }
}

void WgJob::prep(const void *vis, bool vis_f64, int64_t vst0, int64_t vst1) {
  pts_.stream = s;
  pts_.reset(M * (16 + 8 + 1) + 256);
  char *b = (char *)pts_.p;
  o.coord = (double *)b;
  o.w = (double *)(b + M * 16);
  o.active = (uint8_t *)(b + M * 24);
  o.val = nullptr;
  o.val_f64 = 1;  // FP64 until the planes convert to the compute type
  if (a.gridding) {
    val_.stream = s;
    val_.reset(M * 16);
    o.val = val_.p;
  }
  mm_.stream = s;
  mm_.reset(3 * 8);
  const unsigned long long init[3] = {~0ull, 0ull, 0ull};
  DMX_HIP(hipMemcpyAsync(mm_.p, init, sizeof(init), hipMemcpyHostToDevice, s));
  DMX_HIP(hipStreamSynchronize(s));
  o.wminmax = mm_.as<unsigned long long>();
  a.vis = vis;
  a.vis_f64 = vis_f64 ? 1 : 0;
  a.vst0 = vst0;
  a.vst1 = vst1;
  if (M)
    hipLaunchKernelGGL(k_wg_prep,
                       dim3(std::min<unsigned>(nblk(M, 256), PREP_BLOCKS)),
                       dim3(256), 0, s, a, o);
  DMX_HIP(hipGetLastError());
  unsigned long long h[3];
  DMX_HIP(hipMemcpyAsync(h, mm_.p, sizeof(h), hipMemcpyDeviceToHost, s));
  DMX_HIP(hipStreamSynchronize(s));
  nvis = (size_t)h[2];
  if (nvis) {
    wmin_d = from_key(h[0]);
    wmax_d = from_key(h[1]);
  }
}

void WgJob::choose(double eps, double smin, double smax, bool fp64_values,
                   const dmx_options &opt) {
  const size_t nuni[3] = {(size_t)a.nx, (size_t)a.ny, 1};
  const double per[2] = {2 * M_PI, 2 * M_PI};
  if (!a.wgrid) {
    p = choose_params(2, nuni, std::max<size_t>(M, 1), true, eps, smin, smax,
                      per, false, fp64_values, opt);
    return;
  }
  // a row of the 3-D table (wgridder.h:1522) for the uv grid and the w
  // planes: per candidate, plane count and a cost of the per-plane FFTs
  // plus the W^3 taps per visibility (GPU constants: 6.3e-13 s per grid
  // cell and log2 step of the FFT, 1e-12 s per tap, 3e-5 s per plane)
  const double mx = std::max(std::fabs(nm1max + a.nshift),
                             std::fabs(nm1min + a.nshift));
  bool fp64 = fp64_values || eps < opt.fp64_below;
  auto rows = available_kernels(eps, 3, !fp64, smin, smax);
  if (rows.empty() && !fp64) {
    fp64 = true;
    rows = available_kernels(eps, 3, false, smin, smax);
  }
  DMX_ASSERT(!rows.empty(), "no appropriate kernel found");
  const auto &db = kernel_db();
  double best = 1e300, bnp = 1;
  int bi = -1;
  for (int r : rows) {
    const KRow &k = db[r];
    const double nu = (double)oversampled_length(a.nx, k.ofactor),
                 nv = (double)oversampled_length(a.ny, k.ofactor);
    const double dwk = mx > 0 ? 0.5 / k.ofactor / mx : 1.0;
    const double np = std::floor((wmax_d - wmin_d) / dwk + k.W);
    const double cost =
        np * (6.3e-13 * nu * nv * std::log2(nu * nv) + 3e-5) +
        1e-12 * (double)nvis * k.W * k.W * k.W;
    if (cost < best) {
      best = cost;
      bi = r;
      bnp = np;
    }
  }
  const KRow &k = db[bi];
  // each visibility is spread onto W of the bnp planes
  p = kernel_params(2, nuni, per, false, fp64, eps, bi, k.ofactor, opt,
                    (double)nvis * k.W / std::max(bnp, 1.0));
  pl.W = k.W;
  dw = mx > 0 ? 0.5 / k.ofactor / mx : 1.0;
  // wgridder.h:683-688
  pl.nplanes = (int)std::floor((wmax_d - wmin_d) / dw + k.W);
  DMX_ASSERT(pl.nplanes < (1 << 16), "too many w planes");
  pl.wmin = 0.5 * (wmin_d + wmax_d) - 0.5 * (pl.nplanes - 1) * dw;
  pl.dw = dw;
  auto cd = device_coeffs(k.W, k.beta, k.e0);
  wcoef_.stream = s;
  wcoef_.reset(cd.size() * 8);
  DMX_HIP(hipMemcpyAsync(wcoef_.p, cd.data(), cd.size() * 8,
                         hipMemcpyHostToDevice, s));
  DMX_HIP(hipStreamSynchronize(s));
  pl.coef = wcoef_.as<double>();
  cw = CorfacArgs{};
  cw.nq = corr_quadrature(k.W, k.beta * k.W, k.e0, cw.x, cw.wpsi);
  cw.W = k.W;
}

void WgJob::sort_planes() {
  // first plane per visibility, stable sort, plane ranges
  keys_.stream = s;
  keys_.reset(M * 4);
  skeys_.stream = s;
  skeys_.reset(M * 4);
  idx_.stream = s;
  idx_.reset(M * 4);
  hipLaunchKernelGGL(k_wg_keys, dim3(nblk(M, 256)), dim3(256), 0, s, o, M, pl,
                     keys_.as<uint32_t>());
  int kb = 1;
  while ((1 << kb) <= pl.nplanes) ++kb;
  const size_t ws = radix_sort_workspace(M);
  Buf work(ws, s);
  radix_sort_pairs(keys_.as<uint32_t>(), M, kb, idx_.as<uint32_t>(),
                   skeys_.as<uint32_t>(), work.p, ws, s);
  sorted_.stream = s;
  sorted_.reset(M * 24 + 256);
  so = o;
  so.coord = sorted_.as<double>();
  so.w = so.coord + 2 * M;
  so.active = nullptr;
  so.val = nullptr;
  if (a.gridding) {
    sval_.stream = s;
    sval_.reset(M * 16);
    so.val = sval_.p;
  }
  hipLaunchKernelGGL(k_wg_gather, dim3(nblk(M, 256)), dim3(256), 0, s, o,
                     (const uint32_t *)idx_.p, M, so);
  DMX_HIP(hipGetLastError());
  // visibilities per first plane -> contiguous ranges
  std::vector<uint32_t> sk(M);
  DMX_HIP(hipMemcpyAsync(sk.data(), skeys_.p, M * 4, hipMemcpyDeviceToHost,
                         s));
  DMX_HIP(hipStreamSynchronize(s));
  first_.assign(pl.nplanes + 2, 0);
  // first_[k] = first sorted position with key >= k
  size_t pos = 0;
  for (int k = 0; k <= pl.nplanes + 1; ++k) {
    while (pos < M && sk[pos] < (uint32_t)k) ++pos;
    first_[k] = pos;
  }
}

void WgJob::plane_range(int p_, size_t &lo, size_t &hi) const {
  const int k0 = std::max(0, p_ - pl.W + 1);
  lo = first_[k0];
  hi = first_[std::min(p_ + 1, pl.nplanes)];
}

void WgJob::grid(void *dirty, bool dirty_f64, int64_t ds0, int64_t ds1) {
  const dim3 g2(nblk((size_t)a.ny, 256), (unsigned)a.nx);
  const size_t nimg = (size_t)a.nx * a.ny;
  const int64_t ust[2] = {a.ny, 1};
  const size_t es = p.fp64 ? 16 : 8;
  if (!a.wgrid) {
    Buf img(nimg * es, s);
    Plan pln(p, s);
    pln.build_index(o.coord, false, M, false);
    pln.nu2u(false, o.val, true, img.p, p.fp64, ust);
    hipLaunchKernelGGL(k_wg_finish_dirty, g2, dim3(256), 0, s, a,
                       (const void *)img.p, p.fp64 ? 1 : 0, 0, cw, 0.0, dirty,
                       dirty_f64 ? 1 : 0, ds0, ds1);
    DMX_HIP(hipGetLastError());
    DMX_HIP(hipStreamSynchronize(s));
    pln.mark_done();
    return;
  }
  sort_planes();
  Buf acc(nimg * 16, s), img(nimg * es, s), vals(std::max<size_t>(M, 1) * es, s);
  launch_zero(acc.p, acc.bytes, s);
  // one plan for all planes: each plane re-sorts its range into the plan's
  // buffers, which are reused in the order of the one stream
  Plan pln(p, s);
  for (int pp = 0; pp < pl.nplanes; ++pp) {
    size_t lo, hi;
    plane_range(pp, lo, hi);
    if (hi <= lo) continue;
    hipLaunchKernelGGL(k_wg_plane_vals, dim3(nblk(hi - lo, 256)), dim3(256), 0,
                       s, so, (const uint32_t *)skeys_.p, lo, hi, pl, pp,
                       vals.p, p.fp64 ? 1 : 0);
    pln.build_index(so.coord + 2 * lo, false, hi - lo, false);
    pln.nu2u(false, vals.p, p.fp64, img.p, p.fp64, ust);
    hipLaunchKernelGGL(k_wg_accum_image, g2, dim3(256), 0, s, a,
                       (const void *)img.p, p.fp64 ? 1 : 0,
                       pl.wmin + pp * dw, acc.as<double2>());
    DMX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_wg_finish_dirty, g2, dim3(256), 0, s, a,
                     (const void *)acc.p, 1, 1, cw, dw, dirty,
                     dirty_f64 ? 1 : 0, ds0, ds1);
  DMX_HIP(hipGetLastError());
  DMX_HIP(hipStreamSynchronize(s));
  pln.mark_done();
}

void WgJob::degrid(const void *dirty, bool dirty_f64, int64_t ds0,
                   int64_t ds1, void *vis, bool vis_f64, int64_t vs0,
                   int64_t vs1) {
  const dim3 g2(nblk((size_t)a.ny, 256), (unsigned)a.nx);
  const size_t nimg = (size_t)a.nx * a.ny;
  const int64_t ust[2] = {a.ny, 1};
  const size_t es = p.fp64 ? 16 : 8;
  Buf imgd(nimg * 8, s), img(nimg * es, s);
  hipLaunchKernelGGL(k_wg_prep_image, g2, dim3(256), 0, s, a, dirty,
                     dirty_f64 ? 1 : 0, ds0, ds1, a.wgrid, cw, dw,
                     imgd.as<double>());
  if (!a.wgrid) {
    hipLaunchKernelGGL(k_wg_plane_image, g2, dim3(256), 0, s, a,
                       (const double *)imgd.p, 0, 0.0, img.p, p.fp64 ? 1 : 0);
    Buf res(std::max<size_t>(M, 1) * es, s);
    Plan pln(p, s);
    pln.build_index(o.coord, false, M, false);
    pln.u2nu(true, img.p, p.fp64, ust, res.p, p.fp64);
    hipLaunchKernelGGL(k_wg_finish_vis, dim3(nblk(M, 256)), dim3(256), 0, s,
                       a, o, (const void *)res.p, p.fp64 ? 1 : 0, nullptr, M,
                       vis, vis_f64 ? 1 : 0, vs0, vs1);
    DMX_HIP(hipGetLastError());
    DMX_HIP(hipStreamSynchronize(s));
    pln.mark_done();
    return;
  }
  sort_planes();
  Buf acc(std::max<size_t>(M, 1) * 16, s), res(std::max<size_t>(M, 1) * es, s);
  launch_zero(acc.p, acc.bytes, s);
  Plan pln(p, s);  // one plan for all planes, as in grid()
  for (int pp = 0; pp < pl.nplanes; ++pp) {
    size_t lo, hi;
    plane_range(pp, lo, hi);
    if (hi <= lo) continue;
    hipLaunchKernelGGL(k_wg_plane_image, g2, dim3(256), 0, s, a,
                       (const double *)imgd.p, 1, pl.wmin + pp * dw, img.p,
                       p.fp64 ? 1 : 0);
    pln.build_index(so.coord + 2 * lo, false, hi - lo, false);
    pln.u2nu(true, img.p, p.fp64, ust, res.p, p.fp64);
    hipLaunchKernelGGL(k_wg_plane_accum_vis, dim3(nblk(hi - lo, 256)),
                       dim3(256), 0, s, so, (const uint32_t *)skeys_.p, lo, hi,
                       pl, pp, (const void *)res.p, p.fp64 ? 1 : 0,
                       acc.as<double2>());
    DMX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_wg_finish_vis, dim3(nblk(M, 256)), dim3(256), 0, s, a,
                     o, (const void *)acc.p, 1, (const uint32_t *)idx_.p, M,
                     vis, vis_f64 ? 1 : 0, vs0, vs1);
  DMX_HIP(hipGetLastError());
  DMX_HIP(hipStreamSynchronize(s));
  pln.mark_done();
}

void WgJob::zero_dirty(void *dirty, bool dirty_f64, int64_t ds0,
                       int64_t ds1) {
  // an image with no active visibility (wgridder.h:1659-1662)
  const dim3 g2(nblk((size_t)a.ny, 256), (unsigned)a.nx);
  Buf z((size_t)a.nx * a.ny * 16, s);
  launch_zero(z.p, z.bytes, s);
  hipLaunchKernelGGL(k_wg_finish_dirty, g2, dim3(256), 0, s, a,
                     (const void *)z.p, 1, 0, cw, 0.0, dirty,
                     dirty_f64 ? 1 : 0, ds0, ds1);
  DMX_HIP(hipGetLastError());
  DMX_HIP(hipStreamSynchronize(s));
}

}  // namespace dmx
