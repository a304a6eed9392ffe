// sht.h -- spherical harmonic synthesis at arbitrary points and its adjoint
// on the NUFFT (sht.hip).
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "dmx.h"

namespace dmx {

// Ring / mode geometry: N Clenshaw-Curtis rings (theta_j = j pi/(N-1)),
// doubled to 2N-2 rows of a periodic theta axis; nphi = 2 mmax + 1 modes in
// phi (FFT order)
struct ShtGeometry {
  int lmax, mmax, N, nphi;
};
ShtGeometry sht_geometry(size_t lmax, size_t mmax);
// (sht.cc:134-136) the normalisation of lambda_mm without its sin^m factor
void sht_mfac(size_t mmax, std::vector<double> &mf);

// Start of the spin-s recursions of one m at l0 = max(m, s):
//   lambda^{+s}_{l0,m}(theta) = sgnp * mant 2^exp * cos^a(theta/2) sin^b(theta/2)
//   lambda^{-s}_{l0,m}(theta) = sgnm * mant 2^exp * cos^b(theta/2) sin^a(theta/2)
// with a = |m - s|, b = m + s and mant 2^exp = sqrt((2 l0 + 1) / (4 pi))
// sqrt(binomial(2 l0, l0 - min(m, s))) (the closed forms of the Wigner
// functions d^{l0}_{m, -+s})
struct SpinStart {
  double mant;
  int exp, sgnp, sgnm, a, b, pad;
};
void sht_spin_start(size_t mmax, int spin, std::vector<SpinStart> &out);

// device arguments of the Legendre kernels
struct ShtArgs {
  const void *alm;  // complex a_lm, index mstart[m] + l * lstride
  int alm_f64;
  const int64_t *mstart;
  int64_t lstride;
  int lmax, mmax, N, nphi;
  const double *mfac;     // mmax + 1
  const int64_t *accofs;  // adjoint: offset of a_mm in the FP64 accumulator
  // spin > 0: the a_lm of component c (E, then B) at c * acomp + index;
  // nalm_c components (2, or 1 when B is absent); deriv1: a_lm weighted by
  // sqrt(l (l + 1)); the adjoint's second accumulator at acc + nacc
  int spin, nalm_c, deriv1;
  int64_t acomp, nacc;
  const SpinStart *sstart;  // mmax + 1
};

// a_lm -> the doubled-sphere rows G[(2N-2) x nphi] (complex FP64)
void launch_alm2leg(const ShtArgs &g, double2 *G, hipStream_t s);
// doubled rows Gt -> a_lm accumulated in acc (complex FP64, per m the
// l = m..lmax run at accofs[m]); acc must be zero when
// leg2alm_needs_zero(g)
void launch_leg2alm(const ShtArgs &g, const double2 *Gt, double2 *acc,
                    hipStream_t s);
bool leg2alm_needs_zero(const ShtArgs &g);
// complex values times fct -> the map: component 0 the real parts, and for
// cstride != 0 component 1 (at map + cstride) the imaginary parts
void launch_real_part(const double2 *z, size_t M, void *map, bool map_f64,
                      int64_t mstride, int64_t cstride, double fct,
                      hipStream_t s);
// the map -> complex values: real part component 0, imaginary part
// component 1 when cstride != 0, else zero
void launch_to_complex(const void *map, bool map_f64, int64_t mstride,
                       int64_t cstride, size_t M, double2 *z, hipStream_t s);
// acc times fct -> the a_lm output
void launch_store_alm(const double2 *acc, const ShtArgs &g, void *alm,
                      bool alm_f64, double fct, hipStream_t s);
// spin > 0 (g.spin): the same steps with the pairs of spin-weighted
// functions; G / Gt hold the complex field Q + i U, acc 2 * nacc entries
void launch_alm2leg_spin(const ShtArgs &g, double2 *G, hipStream_t s);
void launch_leg2alm_spin(const ShtArgs &g, const double2 *Gt, double2 *acc,
                         hipStream_t s);
void launch_store_alm_spin(const double2 *acc, const ShtArgs &g, void *alm,
                           bool alm_f64, double fct, hipStream_t s);

// vector steps of the LSMR solver of pseudo_analysis_general: y = a x + b y
// over n reals; *out += the sum of squares of n reals; *out += the a_lm
// inner product of alm with itself (m > 0 entries twice) over ncomp
// components at stride acomp
void lsmr_axpby(bool f64, size_t n, double a, const void *x, double b,
                void *y, hipStream_t s);
void lsmr_sumsq(bool f64, size_t n, const void *x, double *out,
                hipStream_t s);
void lsmr_alm_sumsq(const ShtArgs &g, bool f64, const void *alm, int ncomp,
                    int64_t acomp, double *out, hipStream_t s);
// ncomp components of M points at strides (s0, s1) -> contiguous
void lsmr_copy_map(bool f64, int ncomp, size_t M, const void *src,
                   int64_t s0, int64_t s1, void *dst, hipStream_t s);
// LSMR from x0 = 0 for the operator A (x-space -> b-space) and its adjoint
// AT, with b in u on entry; nx / nb reals per vector; v, h, hbar, t
// (x-space) and w (b-space) are scratch; x holds the solution on return
struct LsmrResult {
  size_t istop, itn;
  double rnorm, arnorm;  // |r| / |b|, |A^T r| / (|A| |r|)
};
LsmrResult lsmr_solve(size_t nx, size_t nb, bool f64,
                      const std::function<void(const void *, void *)> &A,
                      const std::function<void(const void *, void *)> &AT,
                      const std::function<double(const void *)> &normx,
                      const std::function<double(const void *)> &normb,
                      void *x, void *v, void *h, void *hbar, void *u,
                      void *w, void *t, double atol, double btol,
                      double conlim, size_t maxiter, hipStream_t s);

}  // namespace dmx
