// sht.h -- spherical harmonic synthesis at arbitrary points and its adjoint
// on the NUFFT (sht.hip).
#pragma once
#include <cstdint>
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

// device arguments of the Legendre kernels
struct ShtArgs {
  const void *alm;  // complex a_lm, index mstart[m] + l * lstride
  int alm_f64;
  const int64_t *mstart;
  int64_t lstride;
  int lmax, mmax, N, nphi;
  const double *mfac;     // mmax + 1
  const int64_t *accofs;  // adjoint: offset of a_mm in the FP64 accumulator
};

// a_lm -> the doubled-sphere rows G[(2N-2) x nphi] (complex FP64)
void launch_alm2leg(const ShtArgs &g, double2 *G, hipStream_t s);
// doubled rows Gt -> a_lm accumulated in acc (complex FP64, per m the
// l = m..lmax run at accofs[m]); acc must be zero when
// leg2alm_needs_zero(g)
void launch_leg2alm(const ShtArgs &g, const double2 *Gt, double2 *acc,
                    hipStream_t s);
bool leg2alm_needs_zero(const ShtArgs &g);
void launch_real_part(const double2 *z, size_t M, void *map, bool map_f64,
                      int64_t mstride, hipStream_t s);
void launch_to_complex(const void *map, bool map_f64, int64_t mstride,
                       size_t M, double2 *z, hipStream_t s);
void launch_store_alm(const double2 *acc, const ShtArgs &g, void *alm,
                      bool alm_f64, hipStream_t s);

}  // namespace dmx
