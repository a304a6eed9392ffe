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
//  2. a c2c FFT over the 2N-2 rows (scaled by 1/(2N-2): the factor is
//     applied to the map, or in the adjoint to the a_lm, at the end) gives
//     its exact Fourier coefficients C[k, m] (k in FFT order), and
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
#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
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

// Sum over the 64 lanes of eight values per lane: on return lane L holds,
// in out, the total of v[idx] with idx = ((L >> 5) & 1) * 4 + ((L >> 4) & 1)
// * 2 + ((L >> 3) & 1), in every lane of its group of 8.  Three halving
// exchanges (32, 16, 8) then a butterfly over 4, 2, 1: 10 shuffles of NV
// doubles for 8 sums instead of 48.
template <int NV>
__device__ __forceinline__ void reduce8(double (&v)[8][NV], double (&out)[NV],
                                        int lane) {
  const bool h1 = (lane >> 5) & 1, h2 = (lane >> 4) & 1, h3 = (lane >> 3) & 1;
  double u[4][NV], w[2][NV];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const double keep = h1 ? v[i + 4][c] : v[i][c];
      const double send = h1 ? v[i][c] : v[i + 4][c];
      u[i][c] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const double keep = h2 ? u[i + 2][c] : u[i][c];
      const double send = h2 ? u[i][c] : u[i + 2][c];
      w[i][c] = keep + __shfl_xor(send, 16);
    }
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const double keep = h3 ? w[1][c] : w[0][c];
    const double send = h3 ? w[0][c] : w[1][c];
    double z = keep + __shfl_xor(send, 8);
    z += __shfl_xor(z, 4);
    z += __shfl_xor(z, 2);
    z += __shfl_xor(z, 1);
    out[c] = z;
  }
}

// The same for four values per lane: lane L holds the total of v[idx],
// idx = ((L >> 5) & 1) * 2 + ((L >> 4) & 1), in every lane of its group of
// 16; 7 shuffles of NV doubles for 4 sums instead of 24.
template <int NV>
__device__ __forceinline__ void reduce4(double (&v)[4][NV], double (&out)[NV],
                                        int lane) {
  const bool h1 = (lane >> 5) & 1, h2 = (lane >> 4) & 1;
  double u[2][NV];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const double keep = h1 ? v[i + 2][c] : v[i][c];
      const double send = h1 ? v[i][c] : v[i + 2][c];
      u[i][c] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const double keep = h2 ? u[1][c] : u[0][c];
    const double send = h2 ? u[0][c] : u[1][c];
    double z = keep + __shfl_xor(send, 16);
    z += __shfl_xor(z, 8);
    z += __shfl_xor(z, 4);
    z += __shfl_xor(z, 2);
    z += __shfl_xor(z, 1);
    out[c] = z;
  }
}

// The scaled recursions are rescaled once per eight steps: one step
// multiplies max(|lambda_l|, |lambda_{l-1}|) by at most 1.5 / eps_{l+1} <
// 1.5 sqrt(2 l + 3) < 2^12 for l < 2^20, so eight steps stay below 2^96 of
// growth, far inside the 2^768 left above SHT_BIG
__device__ __forceinline__ void rescale8(double &lam, double &lamp, int &k) {
  if (k < 0 && fabs(lam) > SHT_BIG) {
    lam *= SHT_DOWN;
    lamp *= SHT_DOWN;
    ++k;
  }
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
  // entries up to the next multiple of 8 zero
  const int n = l1 - l0, npad = (n + 7) & ~7;
  for (int i = n + threadIdx.x; i < npad; i += blockDim.x) {
    c.ab[i] = make_double2(0.0, 0.0);
    c.a[i] = make_double2(0.0, 0.0);
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
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
  const int q0 = (blockIdx.x * blockDim.x + threadIdx.x) * SHT_R;
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
    // eight l at a time (zero-padded; l0 - m is even, so u's parity is
    // that of l - m)
    for (int i0 = 0; i0 < l1 - l0; i0 += 8) {
#pragma unroll 2
      for (int u = 0; u < 8; ++u) {
        const double2 a = ch.a[i0 + u];
        const double2 ab = ch.ab[i0 + u];
#pragma unroll
        for (int r = 0; r < SHT_R; ++r) {
          const double w = k[r] == 0 ? lam[r] : 0.0;
          double2 &acc = (u & 1) ? so[r] : se[r];
          acc.x = fma(a.x, w, acc.x);
          acc.y = fma(a.y, w, acc.y);
          const double nl = fma(ab.x * x[r], lam[r], -ab.y * lamp[r]);
          lamp[r] = lam[r];
          lam[r] = nl;
        }
      }
#pragma unroll
      for (int r = 0; r < SHT_R; ++r) rescale8(lam[r], lamp[r], k[r]);
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
  const int q0 = (blockIdx.x * blockDim.x + threadIdx.x) * SHT_R;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
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
    // eight l at a time (l0 - m is even: the parities alternate from even),
    // summed over the wave's rings with reduce8
    for (int i0 = 0; i0 < l1 - l0; i0 += 8) {
      double v[8][2];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 ab = ch.ab[i0 + u];
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int r = 0; r < SHT_R; ++r) {
          const double w = k[r] == 0 ? lam[r] : 0.0;
          const double2 gv = (u & 1) ? go[r] : ge[r];
          sr = fma(w, gv.x, sr);
          si = fma(w, gv.y, si);
          const double nl = fma(ab.x * x[r], lam[r], -ab.y * lamp[r]);
          lamp[r] = lam[r];
          lam[r] = nl;
        }
        v[u][0] = sr;
        v[u][1] = si;
      }
#pragma unroll
      for (int r = 0; r < SHT_R; ++r) rescale8(lam[r], lamp[r], k[r]);
      double o[2];
      reduce8<2>(v, o, lane);
      if ((lane & 7) == 0) {
        const int idx = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 +
                        ((lane >> 3) & 1);
        red[wave][i0 + idx] = make_double2(o[0], o[1]);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < l1 - l0; i += blockDim.x) {
      double2 s = red[0][i];
      for (int w = 1; w < nwaves; ++w) {
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

// complex results times fct -> real map: the real parts, and with cstride
// != 0 the imaginary parts as the second component; strided
template <typename TO>
__global__ void k_real_part(const double2 *z, size_t M, TO *map,
                            int64_t mstride, int64_t cstride, double fct) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double2 v = z[i];
  map[(int64_t)i * mstride] = (TO)(v.x * fct);
  if (cstride) map[cstride + (int64_t)i * mstride] = (TO)(v.y * fct);
}

// real map -> complex values: the second component (cstride != 0) as the
// imaginary part, else zero
template <typename TI>
__global__ void k_to_complex(const TI *map, int64_t mstride, int64_t cstride,
                             size_t M, double2 *z) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double re = (double)map[(int64_t)i * mstride];
  const double im = cstride ? (double)map[cstride + (int64_t)i * mstride] : 0.0;
  z[i] = make_double2(re, im);
}

// accumulated a_lm times fct -> output array (m = 0: the real part, as the
// real field's a_l0 are real)
template <typename TO>
__global__ void k_store_alm(const double2 *acc, ShtArgs g, TO *alm,
                            double fct) {
  const int m = blockIdx.y;
  const int l = m + blockIdx.x * blockDim.x + threadIdx.x;
  if (l > g.lmax) return;
  double2 v = acc[g.accofs[m] + (l - m)];
  v.x *= fct;
  v.y *= fct;
  if (m == 0) v.y = 0.0;
  const int64_t ix = g.mstart[m] + (int64_t)l * g.lstride;
  alm[2 * ix] = (TO)v.x;
  alm[2 * ix + 1] = (TO)v.y;
}


// ------------------------------------------------------------- spin > 0
// Spin-s fields Q + i U = -sum_{l, |m| <= l} (E_lm + i B_lm) sY_lm with
// sY_lm = lambda^s_lm(theta) e^{i m phi} and E_{l,-m} = (-1)^m conj(E_lm)
// (B likewise): the convention of sht.cc's spin transforms (its get_norm,
// sht.cc:85-101, and the E / B combinations of its inner loops).  Per m >= 0
// the two columns of the field on ring theta are
//   G_m(theta)  = -sum_l a_l lambda^{+s}_lm(theta),      a_l = E + i B,
//   G_-m(theta) = -(-1)^s sum_l abar_l lambda^{-s}_lm(theta),
//                                     abar_l = conj(E) + i conj(B)   (m > 0),
// using lambda^s_{l,-m} = (-1)^(s+m) lambda^{-s}_lm.  lambda^{+-s}_lm =
// (-1)^s sqrt((2l+1)/(4 pi)) d^l_{m, -+s}(theta) follows the three-term
// recursion of the Wigner functions in l,
//   lambda_{l+1} = (A_l cos theta +- B_l) lambda_l - C_l lambda_{l-1},
//   R_l = sqrt((l^2 - m^2)(l^2 - s^2)),
//   A_l = sqrt((2l+3)/(2l+1)) (2l+1)(l+1) / R_{l+1},
//   B_l = sqrt((2l+3)/(2l+1)) (2l+1) m s / (l R_{l+1}),
//   C_l = sqrt((2l+3)/(2l-1)) (l+1) R_l / (l R_{l+1}),
// from the closed form at l0 = max(m, s) (SpinStart), both chains in the
// scaled representation of lambda_mm above.  The ring pair (theta, pi -
// theta) shares the chains: lambda^{+-s}_lm(pi - theta) = (-1)^(l+m)
// lambda^{-+s}_lm(theta).  On the doubled sphere row 2N-2-j holds (-1)^(m+s)
// times row j (d^l_{m,m'}(-theta) = (-1)^(m-m') d^l_{m,m'}(theta)).
constexpr int SP_CH = 256;  // l values per LDS chunk
constexpr int SP_R = 4;     // ring pairs per thread

struct SpinChunk {
  double2 a[SP_CH];   // E + i B (synthesis)
  double2 ab[SP_CH];  // conj(E) + i conj(B)
  double A[SP_CH], B[SP_CH], C[SP_CH];
};

// base^n as pm * 2^pe, pm in [0.5, 1) (or 0); base in [0, 1]
__device__ __forceinline__ void pow_me(double base, int n, double &pm,
                                       int &pe) {
  pm = 1.0;
  pe = 0;
  if (n == 0) return;
  if (base == 0.0) {
    pm = 0.0;
    return;
  }
  int be;
  double b = frexp(base, &be);
  unsigned e = (unsigned)n;
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
}

// sgn * mant * 2^e * c^na * s^nb as v * 2^(512 k)
__device__ __forceinline__ void spin_start(double mant, int e, int sgn,
                                           double c, int na, double s, int nb,
                                           double &v, int &k) {
  double pc, ps;
  int ec, es;
  pow_me(c, na, pc, ec);
  pow_me(s, nb, ps, es);
  int t;
  const double pm = frexp(mant * pc * ps, &t);
  if (pm == 0.0) {
    v = 0.0;
    k = 0;
    return;
  }
  const int E = e + ec + es + t;
  k = E >= 0 ? 0 : -((-E) / SHT_SB);
  v = ldexp(sgn < 0 ? -pm : pm, E - SHT_SB * k);
}

// one step of a spin recursion; rescaled by rescale8 every eight steps: for
// l >= max(m, s), R_{l+1} >= 2l + 1, so |A_l| <= 1.3 (l + 1), |B_l| <= 1.3 l
// and |C_l| < 2.6, and one step grows max(|lambda_l|, |lambda_{l-1}|) by
// less than 2.6 l + 4 < 2^23 for l < 2^20: eight steps stay below 2^184,
// far inside the 2^768 left above SHT_BIG
__device__ __forceinline__ void spin_step(double &lam, double &lamp,
                                          double axb, double C) {
  const double nl = fma(axb, lam, -C * lamp);
  lamp = lam;
  lam = nl;
}

// recursion coefficients of chunk [lb, le) and, for the synthesis, the
// (weighted) a_lm; entries up to the next multiple of 8 zero
__device__ void load_spin_chunk(SpinChunk &c, int lb, int le, int m,
                                const ShtArgs &g, bool with_alm) {
  const int s = g.spin;
  const int l0 = max(m, s);
  const int n = le - lb, npad = (n + 7) & ~7;
  for (int i = threadIdx.x; i < npad; i += blockDim.x) {
    double A = 0.0, B = 0.0, C = 0.0;
    double2 a = make_double2(0.0, 0.0), ab = a;
    if (i < n) {
      const int l = lb + i;
      const double L = l, M2 = (double)m * m, S2 = (double)s * s;
      const double L1 = L + 1.0;
      const double r1 = sqrt((L1 * L1 - M2) * (L1 * L1 - S2));
      const double n1 = sqrt((2.0 * L + 3.0) / (2.0 * L + 1.0));
      A = n1 * (2.0 * L + 1.0) * L1 / r1;
      B = n1 * (2.0 * L + 1.0) * (double)m * (double)s / (L * r1);
      if (l > l0) {
        const double rl = sqrt((L * L - M2) * (L * L - S2));
        C = sqrt((2.0 * L + 3.0) / (2.0 * L - 1.0)) * L1 * rl / (L * r1);
      }
      if (with_alm) {
        const int64_t ix = g.mstart[m] + (int64_t)l * g.lstride;
        double2 e, b = make_double2(0.0, 0.0);
        if (g.alm_f64) {
          const double2 *p = (const double2 *)g.alm;
          e = p[ix];
          if (g.nalm_c > 1) b = p[g.acomp + ix];
        } else {
          const float2 *p = (const float2 *)g.alm;
          const float2 ef = p[ix];
          e = make_double2(ef.x, ef.y);
          if (g.nalm_c > 1) {
            const float2 bf = p[g.acomp + ix];
            b = make_double2(bf.x, bf.y);
          }
        }
        if (g.deriv1) {
          const double w = sqrt(L * L1);
          e.x *= w;
          e.y *= w;
        }
        a = make_double2(e.x - b.y, e.y + b.x);
        ab = make_double2(e.x + b.y, b.x - e.y);
      }
    }
    c.A[i] = A;
    c.B[i] = B;
    c.C[i] = C;
    c.a[i] = a;
    c.ab[i] = ab;
  }
}

// ring j: cos theta, cos(theta/2), sin(theta/2), theta = j pi/(N-1)
__device__ __forceinline__ void ring_half(int j, int N, double &x, double &c,
                                          double &sn) {
  double s1, c1;
  sincospi((double)j / (double)(N - 1), &s1, &c1);
  x = c1;
  sincospi((double)j / (double)(2 * (N - 1)), &sn, &c);
}

__global__ __launch_bounds__(SHT_THREADS) void k_alm2leg_spin(ShtArgs g,
                                                              double2 *G) {
  __shared__ SpinChunk ch;
  const int m = blockIdx.y, s = g.spin;
  const int N = g.N, nphi = g.nphi, npairs = (N + 1) / 2;
  const int q0 = (blockIdx.x * blockDim.x + threadIdx.x) * SP_R;
  const SpinStart ss = g.sstart[m];
  double x[SP_R], lp[SP_R], lpp[SP_R], lm[SP_R], lmp[SP_R];
  int kp[SP_R], km[SP_R];
  double2 nA[SP_R], sA[SP_R], nB[SP_R], sB[SP_R];
#pragma unroll
  for (int r = 0; r < SP_R; ++r) {
    const int j = min(q0 + r, npairs - 1);
    double c, sn;
    ring_half(j, N, x[r], c, sn);
    spin_start(ss.mant, ss.exp, ss.sgnp, c, ss.a, sn, ss.b, lp[r], kp[r]);
    spin_start(ss.mant, ss.exp, ss.sgnm, c, ss.b, sn, ss.a, lm[r], km[r]);
    lpp[r] = lmp[r] = 0.0;
    nA[r] = sA[r] = nB[r] = sB[r] = make_double2(0.0, 0.0);
  }
  for (int lb = max(m, s); lb <= g.lmax; lb += SP_CH) {
    const int le = min(lb + SP_CH, g.lmax + 1);
    __syncthreads();
    load_spin_chunk(ch, lb, le, m, g, true);
    __syncthreads();
    // eight l at a time (zero-padded); the parity of l + m alternates from
    // that of lb + m
    const int par0 = (lb + m) & 1;
    for (int i0 = 0; i0 < le - lb; i0 += 8) {
#pragma unroll 2
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u;
        const double2 a = ch.a[i], ab = ch.ab[i];
        const double A = ch.A[i], B = ch.B[i], C = ch.C[i];
        const bool odd = ((u & 1) ^ par0) != 0;
#pragma unroll
        for (int r = 0; r < SP_R; ++r) {
          const double wp = kp[r] == 0 ? lp[r] : 0.0;
          const double wm = km[r] == 0 ? lm[r] : 0.0;
          const double wps = odd ? -wp : wp, wms = odd ? -wm : wm;
          nA[r].x = fma(a.x, wp, nA[r].x);
          nA[r].y = fma(a.y, wp, nA[r].y);
          sA[r].x = fma(a.x, wms, sA[r].x);
          sA[r].y = fma(a.y, wms, sA[r].y);
          nB[r].x = fma(ab.x, wm, nB[r].x);
          nB[r].y = fma(ab.y, wm, nB[r].y);
          sB[r].x = fma(ab.x, wps, sB[r].x);
          sB[r].y = fma(ab.y, wps, sB[r].y);
          const double ax = A * x[r];
          spin_step(lp[r], lpp[r], ax + B, C);
          spin_step(lm[r], lmp[r], ax - B, C);
        }
      }
#pragma unroll
      for (int r = 0; r < SP_R; ++r) {
        rescale8(lp[r], lpp[r], kp[r]);
        rescale8(lm[r], lmp[r], km[r]);
      }
    }
  }
  const double sgs = (s & 1) ? 1.0 : -1.0;      // -(-1)^s
  const double sgd = ((m + s) & 1) ? -1.0 : 1.0;  // doubled rows
#pragma unroll
  for (int r = 0; r < SP_R; ++r) {
    const int j = q0 + r;
    if (j >= npairs) break;
    const int jm = N - 1 - j;
    auto put = [&](int row, int col, double2 v) {
      G[(size_t)row * nphi + col] = v;
      if (row > 0 && row < N - 1)
        G[(size_t)(2 * N - 2 - row) * nphi + col] =
            make_double2(sgd * v.x, sgd * v.y);
    };
    put(j, m, make_double2(-nA[r].x, -nA[r].y));
    if (jm != j) put(jm, m, make_double2(-sA[r].x, -sA[r].y));
    if (m > 0) {
      put(j, nphi - m, make_double2(sgs * nB[r].x, sgs * nB[r].y));
      if (jm != j) put(jm, nphi - m, make_double2(sgs * sB[r].x, sgs * sB[r].y));
    }
  }
}

// adjoint: H+_l = sum_rings g_m lambda^{+s}_l, H-_l = sum_rings g_-m
// lambda^{-s}_l (g the doubled rows folded back, sign (-1)^(m+s)), into
// acc[accofs[m] + l - m] and acc[nacc + accofs[m] + l - m]
__global__ __launch_bounds__(SHT_THREADS) void k_leg2alm_spin(
    ShtArgs g, const double2 *Gt, double2 *acc) {
  __shared__ SpinChunk ch;
  __shared__ double4 red[SHT_THREADS / 64][SP_CH];
  const int m = blockIdx.y, s = g.spin;
  const int N = g.N, nphi = g.nphi, npairs = (N + 1) / 2;
  const int q0 = (blockIdx.x * blockDim.x + threadIdx.x) * SP_R;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const double sgd = ((m + s) & 1) ? -1.0 : 1.0;
  const int colm = m, coln = m ? nphi - m : 0;
  const SpinStart ss = g.sstart[m];
  double x[SP_R], lp[SP_R], lpp[SP_R], lm[SP_R], lmp[SP_R];
  int kp[SP_R], km[SP_R];
  double2 nm_[SP_R], sm_[SP_R], nn_[SP_R], sn_[SP_R];
  auto gval = [&](int j, int col) {
    double2 v = Gt[(size_t)j * nphi + col];
    if (j > 0 && j < N - 1) {
      const double2 w = Gt[(size_t)(2 * N - 2 - j) * nphi + col];
      v.x += sgd * w.x;
      v.y += sgd * w.y;
    }
    return v;
  };
#pragma unroll
  for (int r = 0; r < SP_R; ++r) {
    const int j = q0 + r;
    double c, sn;
    ring_half(min(j, npairs - 1), N, x[r], c, sn);
    spin_start(ss.mant, ss.exp, ss.sgnp, c, ss.a, sn, ss.b, lp[r], kp[r]);
    spin_start(ss.mant, ss.exp, ss.sgnm, c, ss.b, sn, ss.a, lm[r], km[r]);
    lpp[r] = lmp[r] = 0.0;
    const double2 z = make_double2(0.0, 0.0);
    nm_[r] = sm_[r] = nn_[r] = sn_[r] = z;
    if (j < npairs) {
      const int jm = N - 1 - j;
      nm_[r] = gval(j, colm);
      nn_[r] = gval(j, coln);
      if (jm != j) {
        sm_[r] = gval(jm, colm);
        sn_[r] = gval(jm, coln);
      }
    }
  }
  for (int lb = max(m, s); lb <= g.lmax; lb += SP_CH) {
    const int le = min(lb + SP_CH, g.lmax + 1);
    __syncthreads();
    load_spin_chunk(ch, lb, le, m, g, false);
    __syncthreads();
    const int par0 = (lb + m) & 1;
    for (int i0 = 0; i0 < le - lb; i0 += 4) {
      double v[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u;
        const double A = ch.A[i], B = ch.B[i], C = ch.C[i];
        const bool odd = ((u & 1) ^ par0) != 0;
        double hpx = 0.0, hpy = 0.0, hmx = 0.0, hmy = 0.0;
#pragma unroll
        for (int r = 0; r < SP_R; ++r) {
          const double wp = kp[r] == 0 ? lp[r] : 0.0;
          const double wm = km[r] == 0 ? lm[r] : 0.0;
          const double wps = odd ? -wp : wp, wms = odd ? -wm : wm;
          hpx = fma(wp, nm_[r].x, fma(wms, sm_[r].x, hpx));
          hpy = fma(wp, nm_[r].y, fma(wms, sm_[r].y, hpy));
          hmx = fma(wm, nn_[r].x, fma(wps, sn_[r].x, hmx));
          hmy = fma(wm, nn_[r].y, fma(wps, sn_[r].y, hmy));
          const double ax = A * x[r];
          spin_step(lp[r], lpp[r], ax + B, C);
          spin_step(lm[r], lmp[r], ax - B, C);
        }
        v[u][0] = hpx;
        v[u][1] = hpy;
        v[u][2] = hmx;
        v[u][3] = hmy;
      }
      if (i0 & 4) {  // every eight steps
#pragma unroll
        for (int r = 0; r < SP_R; ++r) {
          rescale8(lp[r], lpp[r], kp[r]);
          rescale8(lm[r], lmp[r], km[r]);
        }
      }
      double o[4];
      reduce4<4>(v, o, lane);
      if ((lane & 15) == 0) {
        const int idx = ((lane >> 5) & 1) * 2 + ((lane >> 4) & 1);
        red[wave][i0 + idx] = make_double4(o[0], o[1], o[2], o[3]);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < le - lb; i += blockDim.x) {
      double4 t = red[0][i];
      for (int w = 1; w < nwaves; ++w) {
        const double4 q = red[w][i];
        t.x += q.x;
        t.y += q.y;
        t.z += q.z;
        t.w += q.w;
      }
      double2 *dp = acc + g.accofs[m] + (lb + i - m);
      double2 *dm = dp + g.nacc;
      if (gridDim.x == 1) {
        *dp = make_double2(t.x, t.y);
        *dm = make_double2(t.z, t.w);
      } else {
        atomicAdd(&dp->x, t.x);
        atomicAdd(&dp->y, t.y);
        atomicAdd(&dm->x, t.z);
        atomicAdd(&dm->y, t.w);
      }
    }
  }
}

// (H+, H-) -> E = -(H+ + (-1)^s conj(H-)) / 2, B = i (H+ - (-1)^s conj(H-))
// / 2 (zero for l < max(m, s); m = 0 real; deriv1: E weighted by
// sqrt(l (l + 1)))
template <typename TO>
__global__ void k_store_alm_spin(const double2 *acc, ShtArgs g, TO *alm,
                                 double fct) {
  const int m = blockIdx.y, s = g.spin;
  const int l = m + blockIdx.x * blockDim.x + threadIdx.x;
  if (l > g.lmax) return;
  double2 e = make_double2(0.0, 0.0), b = e;
  if (l >= max(m, s)) {
    const size_t o = g.accofs[m] + (l - m);
    const double2 ha = acc[o], hb = acc[g.nacc + o];
    const double2 hp = make_double2(ha.x * fct, ha.y * fct),
                  hm = make_double2(hb.x * fct, hb.y * fct);
    const double sg = (s & 1) ? -1.0 : 1.0;
    const double cx = sg * hm.x, cy = -sg * hm.y;  // (-1)^s conj(H-)
    e = make_double2(-0.5 * (hp.x + cx), -0.5 * (hp.y + cy));
    b = make_double2(-0.5 * (hp.y - cy), 0.5 * (hp.x - cx));
    if (g.deriv1) {
      const double w = sqrt((double)l * (l + 1.0));
      e.x *= w;
      e.y *= w;
    }
    if (m == 0) e.y = b.y = 0.0;
  }
  const int64_t ix = g.mstart[m] + (int64_t)l * g.lstride;
  alm[2 * ix] = (TO)e.x;
  alm[2 * ix + 1] = (TO)e.y;
  if (g.nalm_c > 1) {
    alm[2 * (g.acomp + ix)] = (TO)b.x;
    alm[2 * (g.acomp + ix) + 1] = (TO)b.y;
  }
}

// ---------------------------------------------------------- LSMR helpers
// y = a x + b y over n reals of type T (x may alias y)
template <typename T>
__global__ void k_axpby(size_t n, double a, const T *x, double b, T *y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

// block sum of v into *out (FP64 atomic per block)
__device__ __forceinline__ void block_sum_to(double v, double *out) {
  __shared__ double part[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
    atomicAdd(out, t);
  }
}

// sum of squares of n reals
template <typename T>
__global__ __launch_bounds__(256) void k_sumsq(size_t n, const T *x,
                                               double *out) {
  double v = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double t = (double)x[i];
    v += t * t;
  }
  block_sum_to(v, out);
}

// sum over the layout's entries of w_m |a_lm|^2, w_0 = 1, w_m = 2, over
// ncomp components at stride acomp (complex elements)
template <typename T>
__global__ __launch_bounds__(256) void k_alm_sumsq(ShtArgs g, const T *alm,
                                                   int ncomp, int64_t acomp,
                                                   double *out) {
  const int m = blockIdx.y;
  const int l = m + blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (l <= g.lmax) {
    const int64_t ix = g.mstart[m] + (int64_t)l * g.lstride;
    for (int c = 0; c < ncomp; ++c) {
      const double re = (double)alm[2 * (c * acomp + ix)],
                   im = (double)alm[2 * (c * acomp + ix) + 1];
      v += re * re + im * im;
    }
    if (m > 0) v *= 2.0;
  }
  block_sum_to(v, out);
}

// component c, point i: src[c * s0 + i * s1] -> dst[c * M + i]
template <typename T>
__global__ void k_copy_map(int ncomp, size_t M, const T *src, int64_t s0,
                           int64_t s1, T *dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  for (int c = 0; c < ncomp; ++c)
    dst[c * M + i] = src[(int64_t)c * s0 + (int64_t)i * s1];
}

unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// threads per block of the Legendre kernels: the ring pairs over R per
// thread, in whole waves, at most SHT_THREADS
unsigned ring_threads(int npairs, int R) {
  const unsigned t = nblk((size_t)npairs, (unsigned)R);
  return std::min<unsigned>(SHT_THREADS, (t + 63) / 64 * 64);
}

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
  const unsigned th = ring_threads(npairs, SHT_R);
  const dim3 grid(nblk((size_t)npairs, th * SHT_R), (unsigned)(g.mmax + 1));
  hipLaunchKernelGGL(k_alm2leg, grid, dim3(th), 0, s, g, G);
  DMX_HIP(hipGetLastError());
}

void launch_leg2alm(const ShtArgs &g, const double2 *Gt, double2 *acc,
                    hipStream_t s) {
  const int npairs = (g.N + 1) / 2;
  const unsigned th = ring_threads(npairs, SHT_R);
  const dim3 grid(nblk((size_t)npairs, th * SHT_R), (unsigned)(g.mmax + 1));
  hipLaunchKernelGGL(k_leg2alm, grid, dim3(th), 0, s, g, Gt, acc);
  DMX_HIP(hipGetLastError());
}

bool leg2alm_needs_zero(const ShtArgs &g) {
  const int npairs = (g.N + 1) / 2;
  const int R = g.spin > 0 ? SP_R : SHT_R;
  return nblk((size_t)npairs, ring_threads(npairs, R) * R) > 1;
}

// x * 2^e -> the same with x in [0.5, 1)
static void normalize_ld(long double &x, long &e) {
  int t;
  x = frexpl(x, &t);
  e += t;
}

void sht_spin_start(size_t mmax, int spin, std::vector<SpinStart> &out) {
  out.resize(mmax + 1);
  const int s = spin;
  // binomial(2 l0, l0 - min(m, s)) as bm * 2^e in long double, by its ratio
  // from one m to the next: for m < s (l0 = s) binomial(2s, s - m) from
  // binomial(2s, s) down, for m >= s (l0 = m) binomial(2m, m - s) up from
  // binomial(2s, 0) = 1
  long double bm = 1.0L;
  long e = 0;
  auto norm = [&] { normalize_ld(bm, e); };
  for (int i = 1; i <= s; ++i) {  // binomial(2s, s): the entry of m = 0
    bm *= (long double)(s + i) / (long double)i;
    norm();
  }
  for (size_t mm = 0; mm <= mmax; ++mm) {
    const int m = (int)mm;
    const int l0 = std::max(m, s);
    if (m > 0) {
      if (m < s) {  // binomial(2s, s-m) = binomial(2s, s-m+1) (s-m+1)/(s+m)
        bm *= (long double)(s - m + 1) / (long double)(s + m);
      } else if (m == s) {  // binomial(2s, 0)
        bm = 1.0L;
        e = 0;
      } else {  // binomial(2m, m-s) from binomial(2m-2, m-1-s)
        bm *= (long double)(2 * m - 1) * (long double)(2 * m) /
              ((long double)(m - s) * (long double)(m + s));
      }
      norm();
    }
    long double b = bm;
    long ee = e;
    if (ee & 1) {
      b *= 2.0L;
      ee -= 1;
    }
    b = std::sqrt(b) * std::sqrt((2.0L * l0 + 1.0L) / (4.0L * (long double)M_PI));
    int t;
    SpinStart st;
    st.mant = std::frexp((double)b, &t);
    st.exp = (int)(ee / 2) + t;
    st.a = std::abs(m - s);
    st.b = m + s;
    st.sgnp = (m & 1) ? -1 : 1;
    st.sgnm = (m >= s) ? st.sgnp : ((s & 1) ? -1 : 1);
    st.pad = 0;
    out[mm] = st;
  }
}

void launch_alm2leg_spin(const ShtArgs &g, double2 *G, hipStream_t s) {
  const int npairs = (g.N + 1) / 2;
  const unsigned th = ring_threads(npairs, SP_R);
  const dim3 grid(nblk((size_t)npairs, th * SP_R), (unsigned)(g.mmax + 1));
  hipLaunchKernelGGL(k_alm2leg_spin, grid, dim3(th), 0, s, g, G);
  DMX_HIP(hipGetLastError());
}

void launch_leg2alm_spin(const ShtArgs &g, const double2 *Gt, double2 *acc,
                         hipStream_t s) {
  const int npairs = (g.N + 1) / 2;
  const unsigned th = ring_threads(npairs, SP_R);
  const dim3 grid(nblk((size_t)npairs, th * SP_R), (unsigned)(g.mmax + 1));
  hipLaunchKernelGGL(k_leg2alm_spin, grid, dim3(th), 0, s, g, Gt, acc);
  DMX_HIP(hipGetLastError());
}

void launch_store_alm_spin(const double2 *acc, const ShtArgs &g, void *alm,
                           bool alm_f64, double fct, hipStream_t s) {
  const dim3 grid(nblk((size_t)g.lmax + 1, 256), (unsigned)(g.mmax + 1));
  if (alm_f64)
    hipLaunchKernelGGL(k_store_alm_spin<double>, grid, dim3(256), 0, s, acc,
                       g, (double *)alm, fct);
  else
    hipLaunchKernelGGL(k_store_alm_spin<float>, grid, dim3(256), 0, s, acc, g,
                       (float *)alm, fct);
  DMX_HIP(hipGetLastError());
}

void launch_real_part(const double2 *z, size_t M, void *map, bool map_f64,
                      int64_t mstride, int64_t cstride, double fct,
                      hipStream_t s) {
  if (!M) return;
  if (map_f64)
    hipLaunchKernelGGL(k_real_part<double>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, z, M, (double *)map, mstride, cstride, fct);
  else
    hipLaunchKernelGGL(k_real_part<float>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, z, M, (float *)map, mstride, cstride, fct);
  DMX_HIP(hipGetLastError());
}

void launch_to_complex(const void *map, bool map_f64, int64_t mstride,
                       int64_t cstride, size_t M, double2 *z, hipStream_t s) {
  if (!M) return;
  if (map_f64)
    hipLaunchKernelGGL(k_to_complex<double>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, (const double *)map, mstride, cstride, M, z);
  else
    hipLaunchKernelGGL(k_to_complex<float>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, (const float *)map, mstride, cstride, M, z);
  DMX_HIP(hipGetLastError());
}

void launch_store_alm(const double2 *acc, const ShtArgs &g, void *alm,
                      bool alm_f64, double fct, hipStream_t s) {
  const dim3 grid(nblk((size_t)g.lmax + 1, 256), (unsigned)(g.mmax + 1));
  if (alm_f64)
    hipLaunchKernelGGL(k_store_alm<double>, grid, dim3(256), 0, s, acc, g,
                       (double *)alm, fct);
  else
    hipLaunchKernelGGL(k_store_alm<float>, grid, dim3(256), 0, s, acc, g,
                       (float *)alm, fct);
  DMX_HIP(hipGetLastError());
}


void lsmr_axpby(bool f64, size_t n, double a, const void *x, double b,
                void *y, hipStream_t s) {
  if (!n) return;
  const unsigned nb = (unsigned)std::min<size_t>(nblk(n, 256), 8192);
  if (f64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(nb), dim3(256), 0, s, n, a,
                       (const double *)x, b, (double *)y);
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3(nb), dim3(256), 0, s, n, a,
                       (const float *)x, b, (float *)y);
  DMX_HIP(hipGetLastError());
}

void lsmr_sumsq(bool f64, size_t n, const void *x, double *out,
                hipStream_t s) {
  if (!n) return;
  const unsigned nb = (unsigned)std::min<size_t>(nblk(n, 256), 2048);
  if (f64)
    hipLaunchKernelGGL(k_sumsq<double>, dim3(nb), dim3(256), 0, s, n,
                       (const double *)x, out);
  else
    hipLaunchKernelGGL(k_sumsq<float>, dim3(nb), dim3(256), 0, s, n,
                       (const float *)x, out);
  DMX_HIP(hipGetLastError());
}

void lsmr_alm_sumsq(const ShtArgs &g, bool f64, const void *alm, int ncomp,
                    int64_t acomp, double *out, hipStream_t s) {
  const dim3 grid(nblk((size_t)g.lmax + 1, 256), (unsigned)(g.mmax + 1));
  if (f64)
    hipLaunchKernelGGL(k_alm_sumsq<double>, grid, dim3(256), 0, s, g,
                       (const double *)alm, ncomp, acomp, out);
  else
    hipLaunchKernelGGL(k_alm_sumsq<float>, grid, dim3(256), 0, s, g,
                       (const float *)alm, ncomp, acomp, out);
  DMX_HIP(hipGetLastError());
}


void lsmr_copy_map(bool f64, int ncomp, size_t M, const void *src,
                   int64_t s0, int64_t s1, void *dst, hipStream_t s) {
  if (!M) return;
  if (f64)
    hipLaunchKernelGGL(k_copy_map<double>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, ncomp, M, (const double *)src, s0, s1,
                       (double *)dst);
  else
    hipLaunchKernelGGL(k_copy_map<float>, dim3(nblk(M, 256)), dim3(256), 0,
                       s, ncomp, M, (const float *)src, s0, s1, (float *)dst);
  DMX_HIP(hipGetLastError());
}

// LSMR (D. C.-L. Fong and M. A. Saunders, "LSMR: an iterative algorithm
// for sparse least-squares problems", SIAM J. Sci. Comput. 33 (2011)
// 2950-2971), without damping, from x0 = 0: the solver of
// pseudo_analysis_general (sht.cc:3050-3056 calls ducc0's with atol =
// 1e-14 sqrt(npix), btol = epsilon, conlim = 1e8).  The vectors live on the
// device; the scalar recurrences run on the host in FP64, and the norms
// (nx / nb: the x- and b-space norms) return to the host each iteration.
LsmrResult lsmr_solve(size_t nx, size_t nb, bool f64,
                      const std::function<void(const void *, void *)> &A,
                      const std::function<void(const void *, void *)> &AT,
                      const std::function<double(const void *)> &normx,
                      const std::function<double(const void *)> &normb,
                      void *x, void *v, void *h, void *hbar, void *u,
                      void *w, void *t, double atol, double btol,
                      double conlim, size_t maxiter, hipStream_t s) {
  LsmrResult r{0, 0, 0.0, 0.0};
  auto copy = [&](size_t n, const void *from, void *to) {
    lsmr_axpby(f64, n, 1.0, from, 0.0, to, s);
  };
  auto scale = [&](size_t n, double a, void *y) {
    lsmr_axpby(f64, n, 0.0, y, a, y, s);
  };
  // beta u = b, alpha v = A^T u
  double beta = normb(u);
  const double bnorm = beta;
  if (beta > 0) scale(nb, 1.0 / beta, u);
  AT(u, t);
  copy(nx, t, v);
  double alpha = normx(v);
  if (alpha > 0) scale(nx, 1.0 / alpha, v);
  double zetabar = alpha * beta, alphabar = alpha, rho = 1.0, rhobar = 1.0,
         cbar = 1.0, sbar = 0.0;
  copy(nx, v, h);
  double betadd = beta, betad = 0.0, rhodold = 1.0, tautildeold = 0.0,
         thetatilde = 0.0, zeta = 0.0, d = 0.0;
  double normA2 = alpha * alpha, maxrbar = 0.0, minrbar = 1e100;
  double normA = std::sqrt(normA2), normr = beta, normar = alpha * beta;
  const double ctol = conlim > 0 ? 1.0 / conlim : 0.0;
  if (normar == 0.0) {
    r.rnorm = bnorm > 0 ? normr / bnorm : 0.0;
    return r;
  }
  size_t itn = 0, istop = 0;
  while (itn < maxiter) {
    ++itn;
    // beta u = A v - alpha u, alpha v = A^T u - beta v
    A(v, w);
    lsmr_axpby(f64, nb, 1.0, w, -alpha, u, s);
    beta = normb(u);
    if (beta > 0) scale(nb, 1.0 / beta, u);
    AT(u, t);
    lsmr_axpby(f64, nx, 1.0, t, -beta, v, s);
    alpha = normx(v);
    if (alpha > 0) scale(nx, 1.0 / alpha, v);
    // rotations (no damping: alphahat = alphabar)
    const double rhoold = rho;
    rho = std::hypot(alphabar, beta);
    const double c = alphabar / rho, sn = beta / rho;
    const double thetanew = sn * alpha;
    alphabar = c * alpha;
    const double rhobarold = rhobar, zetaold = zeta;
    const double thetabar = sbar * rho, rhotemp = cbar * rho;
    rhobar = std::hypot(cbar * rho, thetanew);
    cbar = cbar * rho / rhobar;
    sbar = thetanew / rhobar;
    zeta = cbar * zetabar;
    zetabar = -sbar * zetabar;
    // hbar, x, h
    lsmr_axpby(f64, nx, 1.0, h, -(thetabar * rho / (rhoold * rhobarold)),
               hbar, s);
    lsmr_axpby(f64, nx, zeta / (rho * rhobar), hbar, 1.0, x, s);
    lsmr_axpby(f64, nx, 1.0, v, -(thetanew / rho), h, s);
    // the residual norm estimate
    const double betaacute = betadd, betahat = c * betaacute;
    betadd = -sn * betaacute;
    const double thetatildeold = thetatilde;
    const double rhotildeold = std::hypot(rhodold, thetabar);
    const double ctildeold = rhodold / rhotildeold,
                 stildeold = thetabar / rhotildeold;
    thetatilde = stildeold * rhobar;
    rhodold = ctildeold * rhobar;
    betad = -stildeold * betad + ctildeold * betahat;
    tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold;
    const double taud = (zeta - thetatilde * tautildeold) / rhodold;
    normr = std::sqrt(d + (betad - taud) * (betad - taud) + betadd * betadd);
    // norm and condition estimates of A
    normA2 += beta * beta;
    normA = std::sqrt(normA2);
    normA2 += alpha * alpha;
    maxrbar = std::max(maxrbar, rhobarold);
    if (itn > 1) minrbar = std::min(minrbar, rhobarold);
    const double condA =
        std::max(maxrbar, rhotemp) / std::min(minrbar, rhotemp);
    // stopping rules
    normar = std::fabs(zetabar);
    const double xnorm = normx(x);
    const double test1 = normr / bnorm;
    const double test2 = normA * normr != 0.0
                             ? normar / (normA * normr)
                             : std::numeric_limits<double>::infinity();
    const double test3 = 1.0 / condA;
    const double t1 = test1 / (1.0 + normA * xnorm / bnorm);
    const double rtol = btol + atol * normA * xnorm / bnorm;
    if (itn >= maxiter) istop = 7;
    if (1.0 + test3 <= 1.0) istop = 6;
    if (1.0 + test2 <= 1.0) istop = 5;
    if (1.0 + t1 <= 1.0) istop = 4;
    if (test3 <= ctol) istop = 3;
    if (test2 <= atol) istop = 2;
    if (test1 <= rtol) istop = 1;
    if (istop > 0) break;
  }
  r.istop = istop;
  r.itn = itn;
  r.rnorm = normr / bnorm;
  r.arnorm = normA * normr != 0.0 ? normar / (normA * normr) : 0.0;
  return r;
}

}  // namespace dmx
