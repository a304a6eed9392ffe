// sht.hip -- spherical harmonic synthesis at arbitrary points and its
// adjoint (role of ducc0.sht.synthesis_general / adjoint_synthesis_general,
// src/ducc0/sht/sht.cc:2926-2993 with SphereInterpol,
// src/ducc0/sht/sphere_interpol.h), spin 0, built on the NUFFT.
//
// f(theta, phi) = sum_{l,m} a_lm Y_lm(theta, phi) for a real map, i.e.
// sum_{m>=0} w_m Re(leg_m(theta) e^{i m phi}), w_0 = 1, w_m = 2, with
// leg_m(theta) = sum_l a_lm lambda_lm(cos theta) (orthonormal Legendre
// functions with the Condon-Shortley phase, sht.cc:134-136, 552).
//
//  1. k_alm2leg: leg_m on N Clenshaw-Curtis rings theta_j = j pi/(N-1)
//     (both poles included, N-1 >= lmax+1), written straight into the
//     doubled sphere: the function on theta in [0, 2 pi) with
//     f(2 pi - theta, phi) = f(theta, phi + pi), i.e. row 2N-2-j holds
//     (-1)^m leg_m(theta_j), and column -m holds conj(leg_m) (a real map).
//     The rows are samples of a trigonometric polynomial of degree lmax
//     in theta, so
//  2. a c2c FFT over the 2N-2 rows (scaled by 1/(2N-2)) gives its exact
//     Fourier coefficients C[k, m] (k in FFT order), and
//  3. f at the points is the type-2 NUFFT of C (u2nu, sign +, periodic in
//     both coordinates with period 2 pi, FFT order), whose real part is
//     the map.  This is the same decomposition as SphereInterpol (theta
//     resampling by FFT, phi FFT, interpolation with the ES kernel on the
//     bordered planes), with the borders replaced by periodic wrap on the
//     doubled sphere.
// The adjoint runs the same steps transposed: type-1 NUFFT of the map
// values (sign -) onto the same grid, inverse FFT over the rows, the
// doubled rows folded back, and k_leg2alm: a_lm = sum_j g_m(theta_j)
// lambda_lm(theta_j).
//
// lambda_lm by the three-term recursion in l (orthonormal form,
// sht.cc:204-218): lambda_{l+1} = (x lambda_l - eps_l lambda_{l-1}) /
// eps_{l+1}, eps_l = sqrt((l^2 - m^2) / (4 l^2 - 1)), started from
// lambda_mm = (-1)^m mfac_m sin^m theta.  sin^m theta underflows for large
// m near the poles; the recursion then runs on values scaled by 2^(512 k)
// (k < 0) until they reach the normal range, and adds nothing before
// (role of the scaled recursion of sht.cc:77-82).  All in FP64.
#include <cmath>
#include <vector>

#include "c2c.h"
#include "runtime.h"
#include "sht.h"

namespace dmx {

namespace {

constexpr int SHT_THREADS = 256;
constexpr int SHT_R = 4;       // ring pairs per thread
constexpr int SHT_CH = 512;    // l values per LDS chunk
constexpr double SHT_BIG = 0x1p256, SHT_DOWN = 0x1p-512;
constexpr int SHT_SB = 512;

// (-1)^m mfac[m] sin^m as value * 2^(512 k), |value| in (2^-512, 2]
__device__ __forceinline__ void lambda_mm(double st, int m, double mf,
                                          double &v, int &k) {
  if (m == 0) {
    v = mf;
    k = 0;
    return;
  }
  if (st == 0.0) {  // pole: lambda_lm = 0 for m > 0
    v = 0.0;
    k = 0;
    return;
  }
  // sin^m by binary powering on (mantissa, exponent) pairs
  int be;
  double b = frexp(st, &be);
  double pm = 1.0;
  int pe = 0;
  unsigned e = (unsigned)m;
  while (e) {
    if (e & 1u) {
      int t;
      pm = frexp(pm * b, &t);
      pe += be + t;
    }
    e >>= 1;
    if (e) {
      int t;
      b = frexp(b * b, &t);
      be = 2 * be + t;
    }
  }
  // pe <= 0 (st <= 1): k = ceil(pe / 512) <= 0, pe - 512 k in (-512, 0]
  k = -((-pe) / SHT_SB);
  v = ldexp(((m & 1) ? -mf : mf) * pm, pe - SHT_SB * k);
}

// LDS chunk of one m: a_lm (complex FP64) and the recursion coefficients
// A_l = 1 / eps_{l+1}, B_l = eps_l / eps_{l+1}
struct ShtChunk {
  double2 a[SHT_CH];
  double2 ab[SHT_CH];
};

__device__ __forceinline__ double eps_lm(int l, int m) {
  if (l <= m) return 0.0;
  const double L = l, Mm = m;
  return sqrt((L * L - Mm * Mm) / (4.0 * L * L - 1.0));
}

__device__ void load_chunk(ShtChunk &c, int l0, int l1, int m,
                           const ShtArgs &g, bool with_alm) {
  for (int i = threadIdx.x; i < l1 - l0; i += blockDim.x) {
    const int l = l0 + i;
    const double e1 = eps_lm(l + 1, m);
    c.ab[i] = make_double2(1.0 / e1, eps_lm(l, m) / e1);
    if (with_alm) {
      const int64_t ix = g.mstart[m] + (int64_t)l * g.lstride;
      if (g.alm_f64) {
        c.a[i] = ((const double2 *)g.alm)[ix];
      } else {
        const float2 z = ((const float2 *)g.alm)[ix];
        c.a[i] = make_double2(z.x, z.y);
      }
    }
  }
}

// ring pair q: ring j = q and its mirror about the equator N-1-j
__device__ __forceinline__ void ring_xy(int j, int N, double &x, double &st) {
  double s, c;
  sincospi((double)j / (double)(N - 1), &s, &c);
  x = c;
  st = s;
}

__global__ __launch_bounds__(SHT_THREADS) void k_alm2leg(ShtArgs g,
                                                         double2 *G) {
  __shared__ ShtChunk ch;
  const int m = blockIdx.y;
  const int N = g.N, nphi = g.nphi, npairs = (N + 1) / 2;
  const int q0 = (blockIdx.x * SHT_THREADS + threadIdx.x) * SHT_R;
  double x[SHT_R], lam[SHT_R], lamp[SHT_R];
  int k[SHT_R];
  double2 se[SHT_R], so[SHT_R];
  const double mf = g.mfac[m];
#pragma unroll
  for (int r = 0; r < SHT_R; ++r) {
    const int j = min(q0 + r, npairs - 1);
    double st;
    ring_xy(j, N, x[r], st);
    lambda_mm(st, m, mf, lam[r], k[r]);
    lamp[r] = 0.0;
    se[r] = so[r] = make_double2(0.0, 0.0);
  }
  for (int l0 = m; l0 <= g.lmax; l0 += SHT_CH) {
    const int l1 = min(l0 + SHT_CH, g.lmax + 1);
    __syncthreads();
    load_chunk(ch, l0, l1, m, g, true);
    __syncthreads();
    for (int i = 0; i < l1 - l0; ++i) {
      const double2 a = ch.a[i];
      const double2 ab = ch.ab[i];
      const bool odd = ((l0 + i - m) & 1) != 0;
#pragma unroll
      for (int r = 0; r < SHT_R; ++r) {
        const double w = k[r] == 0 ? lam[r] : 0.0;
        if (odd) {
          so[r].x = fma(a.x, w, so[r].x);
          so[r].y = fma(a.y, w, so[r].y);
        } else {
          se[r].x = fma(a.x, w, se[r].x);
          se[r].y = fma(a.y, w, se[r].y);
        }
        const double nl = fma(ab.x * x[r], lam[r], -ab.y * lamp[r]);
        lamp[r] = lam[r];
        lam[r] = nl;
        if (k[r] < 0 && fabs(nl) > SHT_BIG) {
          lam[r] *= SHT_DOWN;
          lamp[r] *= SHT_DOWN;
          ++k[r];
        }
      }
    }
  }
  const double sg = (m & 1) ? -1.0 : 1.0;
#pragma unroll
  for (int r = 0; r < SHT_R; ++r) {
    const int j = q0 + r;
    if (j >= npairs) break;
    const int jm = N - 1 - j;
    const double2 vn = make_double2(se[r].x + so[r].x, se[r].y + so[r].y);
    const double2 vs = make_double2(se[r].x - so[r].x, se[r].y - so[r].y);
    // row, value: ring j, its mirror N-1-j, and their doubled images
    auto put = [&](int row, double2 v) {
      G[(size_t)row * nphi + m] = v;
      if (m > 0) G[(size_t)row * nphi + (nphi - m)] = make_double2(v.x, -v.y);
    };
    put(j, vn);
    if (j > 0 && j < N - 1) put(2 * N - 2 - j, make_double2(sg * vn.x, sg * vn.y));
    if (jm != j) {
      put(jm, vs);
      if (jm > 0 && jm < N - 1)
        put(2 * N - 2 - jm, make_double2(sg * vs.x, sg * vs.y));
    }
  }
}

// adjoint: a_lm = sum_j g_m(j) lambda_lm(theta_j), g_m(j) = Gt[j][m] +
// (-1)^m Gt[2N-2-j][m] for 0 < j < N-1 (the doubled rows folded back)
__global__ __launch_bounds__(SHT_THREADS) void k_leg2alm(ShtArgs g,
                                                         const double2 *Gt,
                                                         double2 *acc) {
  __shared__ ShtChunk ch;
  __shared__ double2 red[SHT_THREADS / 64][SHT_CH];
  const int m = blockIdx.y;
  const int N = g.N, nphi = g.nphi, npairs = (N + 1) / 2;
  const int q0 = (blockIdx.x * SHT_THREADS + threadIdx.x) * SHT_R;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double sg = (m & 1) ? -1.0 : 1.0;
  double x[SHT_R], lam[SHT_R], lamp[SHT_R];
  int k[SHT_R];
  double2 ge[SHT_R], go[SHT_R];
  const double mf = g.mfac[m];
  auto gval = [&](int j) {
    double2 v = Gt[(size_t)j * nphi + m];
    if (j > 0 && j < N - 1) {
      const double2 w = Gt[(size_t)(2 * N - 2 - j) * nphi + m];
      v.x += sg * w.x;
      v.y += sg * w.y;
    }
    return v;
  };
#pragma unroll
  for (int r = 0; r < SHT_R; ++r) {
    const int j = q0 + r;
    double st;
    ring_xy(min(j, npairs - 1), N, x[r], st);
    lambda_mm(st, m, mf, lam[r], k[r]);
    lamp[r] = 0.0;
    ge[r] = go[r] = make_double2(0.0, 0.0);
    if (j < npairs) {
      const int jm = N - 1 - j;
      const double2 vn = gval(j);
      if (jm == j) {
        ge[r] = go[r] = vn;
      } else {
        const double2 vs = gval(jm);
        ge[r] = make_double2(vn.x + vs.x, vn.y + vs.y);
        go[r] = make_double2(vn.x - vs.x, vn.y - vs.y);
      }
    }
  }
  for (int l0 = m; l0 <= g.lmax; l0 += SHT_CH) {
    const int l1 = min(l0 + SHT_CH, g.lmax + 1);
    __syncthreads();
    load_chunk(ch, l0, l1, m, g, false);
    __syncthreads();
    for (int i = 0; i < l1 - l0; ++i) {
      const double2 ab = ch.ab[i];
      const bool odd = ((l0 + i - m) & 1) != 0;
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int r = 0; r < SHT_R; ++r) {
        const double w = k[r] == 0 ? lam[r] : 0.0;
        const double2 v = odd ? go[r] : ge[r];
        sr = fma(w, v.x, sr);
        si = fma(w, v.y, si);
        const double nl = fma(ab.x * x[r], lam[r], -ab.y * lamp[r]);
        lamp[r] = lam[r];
        lam[r] = nl;
        if (k[r] < 0 && fabs(nl) > SHT_BIG) {
          lam[r] *= SHT_DOWN;
          lamp[r] *= SHT_DOWN;
          ++k[r];
        }
      }
      // sum over the wave's rings
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
      }
      if (lane == 0) red[wave][i] = make_double2(sr, si);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < l1 - l0; i += blockDim.x) {
      double2 s = red[0][i];
#pragma unroll
      for (int w = 1; w < SHT_THREADS / 64; ++w) {
        s.x += red[w][i].x;
        s.y += red[w][i].y;
      }
      double2 *dst = acc + g.accofs[m] + (l0 + i - m);
      if (gridDim.x == 1) {
        *dst = s;
      } else {
        atomicAdd(&dst->x, s.x);
        atomicAdd(&dst->y, s.y);
      }
    }
  }
}

// complex results -> real map (real part), strided
template <typename TO>
__global__ void k_real_part(const double2 *z, size_t M, TO *map,
                            int64_t mstride) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) map[(int64_t)i * mstride] = (TO)z[i].x;
}

// real map -> complex values (imaginary part zero)
template <typename TI>
__global__ void k_to_complex(const TI *map, int64_t mstride, size_t M,
                             double2 *z) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) z[i] = make_double2((double)map[(int64_t)i * mstride], 0.0);
}

// accumulated a_lm -> output array (m = 0: the real part, as the real
// field's a_l0 are real)
template <typename TO>
__global__ void k_store_alm(const double2 *acc, ShtArgs g, TO *alm) {
  const int m = blockIdx.y;
  const int l = m + blockIdx.x * blockDim.x + threadIdx.x;
  if (l > g.lmax) return;
  double2 v = acc[g.accofs[m] + (l - m)];
  if (m == 0) v.y = 0.0;
  const int64_t ix = g.mstart[m] + (int64_t)l * g.lstride;
  alm[2 * ix] = (TO)v.x;
  alm[2 * ix + 1] = (TO)v.y;
}

unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

ShtGeometry sht_geometry(size_t lmax, size_t mmax) {
  DMX_ASSERT(mmax <= lmax, "mmax must not exceed lmax");
  DMX_ASSERT(lmax < (1u << 20), "lmax too large");
  ShtGeometry s;
  s.lmax = (int)lmax;
  s.mmax = (int)mmax;
  // N - 1 >= lmax + 1 rings spacings: the 2N-2 doubled rows sample a
  // trigonometric polynomial of degree lmax exactly; smooth for the FFT
  s.N = (int)smooth_length((int64_t)lmax + 1) + 1;
  s.nphi = 2 * (int)mmax + 1;
  return s;
}

void sht_mfac(size_t mmax, std::vector<double> &mf) {
  // sht.cc:134-136: mfac_0 = 1/sqrt(4 pi), mfac_i = mfac_{i-1} *
  // sqrt((2i+1)/(2i))
  mf.resize(mmax + 1);
  mf[0] = 1.0 / std::sqrt(4.0 * M_PI);
  for (size_t i = 1; i <= mmax; ++i)
    mf[i] = mf[i - 1] * std::sqrt((2.0 * i + 1.0) / (2.0 * i));
}

void launch_alm2leg(const ShtArgs &g, double2 *G, hipStream_t s) {
  const int npairs = (g.N + 1) / 2;
  const dim3 grid(nblk((size_t)npairs, SHT_THREADS * SHT_R),
                  (unsigned)(g.mmax + 1));
  hipLaunchKernelGGL(k_alm2leg, grid, dim3(SHT_THREADS), 0, s, g, G);
  DMX_HIP(hipGetLastError());
}

void launch_leg2alm(const ShtArgs &g, const double2 *Gt, double2 *acc,
                    hipStream_t s) {
  const int npairs = (g.N + 1) / 2;
  const dim3 grid(nblk((size_t)npairs, SHT_THREADS * SHT_R),
                  (unsigned)(g.mmax + 1));
  hipLaunchKernelGGL(k_leg2alm, grid, dim3(SHT_THREADS), 0, s, g, Gt, acc);
  DMX_HIP(hipGetLastError());
}

bool leg2alm_needs_zero(const ShtArgs &g) {
  const int npairs = (g.N + 1) / 2;
  return nblk((size_t)npairs, SHT_THREADS * SHT_R) > 1;
}

void launch_real_part(const double2 *z, size_t M, void *map, bool map_f64,
                      int64_t mstride, hipStream_t s) {
  if (!M) return;
  if (map_f64)
    hipLaunchKernelGGL(k_real_part<double>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, z, M, (double *)map, mstride);
  else
    hipLaunchKernelGGL(k_real_part<float>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, z, M, (float *)map, mstride);
  DMX_HIP(hipGetLastError());
}

void launch_to_complex(const void *map, bool map_f64, int64_t mstride,
                       size_t M, double2 *z, hipStream_t s) {
  if (!M) return;
  if (map_f64)
    hipLaunchKernelGGL(k_to_complex<double>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, (const double *)map, mstride, M, z);
  else
    hipLaunchKernelGGL(k_to_complex<float>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, (const float *)map, mstride, M, z);
  DMX_HIP(hipGetLastError());
}

void launch_store_alm(const double2 *acc, const ShtArgs &g, void *alm,
                      bool alm_f64, hipStream_t s) {
  const dim3 grid(nblk((size_t)g.lmax + 1, 256), (unsigned)(g.mmax + 1));
  if (alm_f64)
    hipLaunchKernelGGL(k_store_alm<double>, grid, dim3(256), 0, s, acc, g,
                       (double *)alm);
  else
    hipLaunchKernelGGL(k_store_alm<float>, grid, dim3(256), 0, s, acc, g,
                       (float *)alm);
  DMX_HIP(hipGetLastError());
}

}  // namespace dmx
