// ref_driver.cc -- thin extern "C" shim over the reference ducc0 NUFFT,
// compiled from the ducc0 sources in place (see oracle/Makefile, target
// `ref`) into oracle/_ref/libducc_ref.so.
//
// TEST INFRASTRUCTURE ONLY: it pins the CPU restatement in oracle.c, writes
// golden fixtures, and is the timed CPU baseline.  Nothing in the product
// links it.  Entry points mirror the ducc0 calls they wrap:
//   nu2u / u2nu free functions      src/ducc0/nufft/nufft.h:1896-1958
//   findNufftParameters             src/ducc0/nufft/nufft.h:155-208
//   selectKernel / Coeff / corfunc  src/ducc0/math/gridding_kernel.{h,cc}
//   bestEpsilon                     src/ducc0/math/gridding_kernel.cc:1562
//   build_index (coord_idx order)   src/ducc0/nufft/nufft.h:892-907,1364,1862
//   c2c                             src/ducc0/fft/fft.h:848-850
//   plan (Nufft class)              src/ducc0/nufft/nufft.h:437-503
//   synthesis_general / adjoint     src/ducc0/sht/sht.cc:2926-2993
//   ms2dirty / dirty2ms (general)   src/ducc0/wgridder/wgridder.h:1700-1735
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "ducc0/nufft/nufft.h"
#include "ducc0/fft/fftnd_impl.h"
#include "ducc0/sht/sht.h"
#include "ducc0/wgridder/wgridder.h"

// sht.cc calls the real-FFT passes whose definitions live in fft1d_impl.h;
// the library build instantiates them in the translation unit that includes
// that header, which here is this one
template class ducc0::detail_fft::rfftpass<float>;
template class ducc0::detail_fft::rfftpass<double>;

using namespace ducc0;
using std::complex;
using std::size_t;
using std::vector;

namespace {
thread_local char g_err[1024];
template <typename F> int guard(F &&f) {
  try {
    f();
    return 0;
  } catch (std::exception &e) {
    std::snprintf(g_err, sizeof(g_err), "%s", e.what());
    return 1;
  }
}

// exposes the protected coord_idx / nover / supp of a reference plan
template <typename T, typename Tc, size_t nd>
struct Peek : public Nufft<T, T, Tc, nd> {
  using Base = Nufft<T, T, Tc, nd>;
  using Anc = detail_nufft::Nufft_ancestor<T, T, nd>;
  using Base::Base;
  void get(uint32_t *idx, size_t *nover, int *supp) const {
    const auto &ci = this->Anc::coord_idx;
    for (size_t i = 0; i < ci.size(); ++i) idx[i] = ci[i];
    for (size_t d = 0; d < nd; ++d) nover[d] = this->Anc::nover[d];
    *supp = int(this->Anc::supp);
  }
};

template <typename T, typename Tc, size_t nd>
void sort_index(size_t M, const Tc *coords, const size_t *nuni,
                const double *period, double eps, double smin, double smax,
                int gridding, size_t nthreads, uint32_t *idx, size_t *nover,
                int *supp) {
  cmav<Tc, 2> c(coords, {M, nd});
  std::array<size_t, nd> shp;
  for (size_t d = 0; d < nd; ++d) shp[d] = nuni[d];
  vector<double> per(period, period + nd);
  Peek<T, Tc, nd> p(gridding != 0, c, shp, eps, nthreads, smin, smax, per,
                    false);
  p.get(idx, nover, supp);
}

template <typename T, typename Tc>
void do_nu2u(int ndim, size_t M, const Tc *coords, const complex<T> *pts,
             int forward, double eps, size_t nthreads, const size_t *nuni,
             complex<T> *out, double smin, double smax, const double *period,
             int fft_order) {
  cmav<Tc, 2> c(coords, {M, size_t(ndim)});
  cmav<complex<T>, 1> p(pts, {M});
  vector<size_t> shp(nuni, nuni + ndim);
  vfmav<complex<T>> o(out, shp);
  vector<double> per(period, period + ndim);
  nu2u<T, T>(c, p, forward != 0, eps, nthreads, o, 0, smin, smax, per,
             fft_order != 0);
}

template <typename T, typename Tc>
void do_u2nu(int ndim, size_t M, const Tc *coords, const complex<T> *grid,
             int forward, double eps, size_t nthreads, const size_t *nuni,
             complex<T> *out, double smin, double smax, const double *period,
             int fft_order) {
  cmav<Tc, 2> c(coords, {M, size_t(ndim)});
  vector<size_t> shp(nuni, nuni + ndim);
  cfmav<complex<T>> g(grid, shp);
  vmav<complex<T>, 1> o(out, {M});
  vector<double> per(period, period + ndim);
  u2nu<T, T>(c, g, forward != 0, eps, nthreads, o, 0, smin, smax, per,
             fft_order != 0);
}
// a ducc0 plan (Nufft<T, T, T, nd>: values and coordinates in one
// precision, as ducc0.nufft.plan requires, nufft_pymod.cc:145-150)
struct RefPlanBase {
  virtual ~RefPlanBase() = default;
  virtual void nu2u(int forward, const void *pts, void *out) = 0;
  virtual void u2nu(int forward, const void *grid, void *out) = 0;
};
template <typename T, size_t nd> struct RefPlan : RefPlanBase {
  size_t M;
  std::array<size_t, nd> shp;
  Nufft<T, T, T, nd> plan;
  RefPlan(bool gridding, size_t M_, const T *coords,
          const std::array<size_t, nd> &shp_, double eps, size_t nthreads,
          double smin, double smax, const vector<double> &per, bool fo)
      : M(M_), shp(shp_),
        plan(gridding, cmav<T, 2>(coords, {M_, nd}), shp_, eps, nthreads,
             smin, smax, per, fo) {}
  void nu2u(int forward, const void *pts, void *out) override {
    cmav<complex<T>, 1> p((const complex<T> *)pts, {M});
    vmav<complex<T>, nd> o((complex<T> *)out, shp);
    plan.nu2u(forward != 0, 0, p, o);
  }
  void u2nu(int forward, const void *grid, void *out) override {
    cmav<complex<T>, nd> g((const complex<T> *)grid, shp);
    vmav<complex<T>, 1> o((complex<T> *)out, {M});
    plan.u2nu(forward != 0, 0, g, o);
  }
};
template <typename T, size_t nd>
RefPlanBase *make_plan(int nu2u, size_t M, const void *coords,
                       const size_t *nuni, double eps, size_t nthreads,
                       double smin, double smax, const double *period,
                       int fft_order) {
  std::array<size_t, nd> shp;
  for (size_t d = 0; d < nd; ++d) shp[d] = nuni[d];
  vector<double> per(period, period + nd);
  return new RefPlan<T, nd>(nu2u != 0, M, (const T *)coords, shp, eps,
                            nthreads, smin, smax, per, fft_order != 0);
}
}  // namespace

extern "C" {

// ducc0.nufft.plan (nufft_pymod.cc:139-261): coords (M, ndim) and values in
// one precision (single != 0: float32 / complex64)
int ref_plan_create(void **plan, int nu2u, int ndim, size_t M,
                    const void *coords, int single, const size_t *nuni,
                    double eps, size_t nthreads, double smin, double smax,
                    const double *period, int fft_order) {
  return guard([&] {
    *plan = nullptr;
#define REF_MK(T, ND)                                                        \
  *plan = make_plan<T, ND>(nu2u, M, coords, nuni, eps, nthreads, smin, smax, \
                           period, fft_order)
    if (single) {
      if (ndim == 1) REF_MK(float, 1);
      else if (ndim == 2) REF_MK(float, 2);
      else REF_MK(float, 3);
    } else {
      if (ndim == 1) REF_MK(double, 1);
      else if (ndim == 2) REF_MK(double, 2);
      else REF_MK(double, 3);
    }
#undef REF_MK
  });
}
int ref_plan_nu2u(void *plan, int forward, const void *points, void *out) {
  return guard([&] { ((RefPlanBase *)plan)->nu2u(forward, points, out); });
}
int ref_plan_u2nu(void *plan, int forward, const void *grid, void *out) {
  return guard([&] { ((RefPlanBase *)plan)->u2nu(forward, grid, out); });
}
void ref_plan_destroy(void *plan) { delete (RefPlanBase *)plan; }

const char *ref_last_error() { return g_err; }

int ref_vlen(int single) {
  return single ? int(detail_nufft::mysimd<float>::size())
                : int(detail_nufft::mysimd<double>::size());
}

double ref_best_epsilon(int ndim, int single, double smin, double smax) {
  double r = -1;
  if (guard([&] { r = bestEpsilon(size_t(ndim), single != 0, smin, smax); }))
    return -1;
  return r;
}

int ref_find_params(double eps, double smin, double smax, int ndim,
                    const size_t *dims, size_t npoints, int gridding,
                    int single, size_t nthreads, size_t *nover) {
  int res = -1;
  int rc = guard([&] {
    vector<size_t> d(dims, dims + ndim);
    size_t kidx;
    vector<size_t> bd;
    if (single) {
      auto t = detail_nufft::findNufftParameters<float, float>(
          eps, smin, smax, d, npoints, gridding != 0, nthreads);
      kidx = std::get<0>(t); bd = std::get<1>(t);
    } else {
      auto t = detail_nufft::findNufftParameters<double, double>(
          eps, smin, smax, d, npoints, gridding != 0, nthreads);
      kidx = std::get<0>(t); bd = std::get<1>(t);
    }
    for (int i = 0; i < ndim; ++i) nover[i] = bd[i];
    res = int(kidx);
  });
  return rc ? -1 : res;
}

// coefficients of KernelDB row kidx: W*(D+1) doubles, D returned
int ref_kernel_coeffs(int kidx, double *coeff, int *D) {
  return guard([&] {
    auto k = selectKernel(size_t(kidx));
    const auto &c = k->Coeff();
    for (size_t i = 0; i < c.size(); ++i) coeff[i] = c[i];
    *D = int(k->degree());
  });
}

int ref_corfac(int kidx, size_t n, double dx, double *out) {
  return guard([&] {
    auto k = selectKernel(size_t(kidx));
    auto v = k->corfunc(n, dx, 1);
    for (size_t i = 0; i < n; ++i) out[i] = v[i];
  });
}

// processing order of a reference plan (coords float64, values per `single`)
int ref_sort_index(int ndim, size_t M, const double *coords,
                   const size_t *nuni, const double *period, double eps,
                   double smin, double smax, int gridding, int single,
                   size_t nthreads, uint32_t *idx, size_t *nover, int *supp) {
  return guard([&] {
#define REF_SORT(T, ND) \
  sort_index<T, double, ND>(M, coords, nuni, period, eps, smin, smax, \
                            gridding, nthreads, idx, nover, supp)
    if (single) {
      if (ndim == 1) REF_SORT(float, 1);
      else if (ndim == 2) REF_SORT(float, 2);
      else REF_SORT(float, 3);
    } else {
      if (ndim == 1) REF_SORT(double, 1);
      else if (ndim == 2) REF_SORT(double, 2);
      else REF_SORT(double, 3);
    }
#undef REF_SORT
  });
}

// value dtype: single!=0 -> complex64, else complex128;
// coord dtype: coord_single!=0 -> float32, else float64.
int ref_nu2u(int ndim, size_t M, const void *coords, int coord_single,
             const void *points, int single, int forward, double eps,
             size_t nthreads, const size_t *nuni, void *out, double smin,
             double smax, const double *period, int fft_order) {
  return guard([&] {
#define REF_CALL(T, TC)                                                       \
  do_nu2u<T, TC>(ndim, M, (const TC *)coords, (const complex<T> *)points,     \
                 forward, eps, nthreads, nuni, (complex<T> *)out, smin, smax, \
                 period, fft_order)
    if (single) {
      if (coord_single) REF_CALL(float, float); else REF_CALL(float, double);
    } else {
      if (coord_single) REF_CALL(double, float); else REF_CALL(double, double);
    }
#undef REF_CALL
  });
}

int ref_u2nu(int ndim, size_t M, const void *coords, int coord_single,
             const void *grid, int single, int forward, double eps,
             size_t nthreads, const size_t *nuni, void *out, double smin,
             double smax, const double *period, int fft_order) {
  return guard([&] {
#define REF_CALL(T, TC)                                                       \
  do_u2nu<T, TC>(ndim, M, (const TC *)coords, (const complex<T> *)grid,       \
                 forward, eps, nthreads, nuni, (complex<T> *)out, smin, smax, \
                 period, fft_order)
    if (single) {
      if (coord_single) REF_CALL(float, float); else REF_CALL(float, double);
    } else {
      if (coord_single) REF_CALL(double, float); else REF_CALL(double, double);
    }
#undef REF_CALL
  });
}

// contiguous complex arrays of `shape`; single!=0 -> complex64
int ref_c2c(int single, int ndim, const int64_t *shape, const void *in,
            void *out, const int64_t *axes, int naxes, int forward,
            double fct, size_t nthreads) {
  return guard([&] {
    std::vector<size_t> shp(shape, shape + ndim), ax(axes, axes + naxes);
    if (single) {
      cfmav<complex<float>> a((const complex<float> *)in, shp);
      vfmav<complex<float>> o((complex<float> *)out, shp);
      c2c(a, o, ax, forward != 0, (float)fct, nthreads);
    } else {
      cfmav<complex<double>> a((const complex<double> *)in, shp);
      vfmav<complex<double>> o((complex<double> *)out, shp);
      c2c(a, o, ax, forward != 0, fct, nthreads);
    }
  });
}

// ducc0.sht.synthesis_general (adjoint = 0) or adjoint_synthesis_general
// (adjoint = 1): alm (nalm_c, nalm) complex, map (nmap, npt) real, loc
// (npt, 2) float64; mode 0 STANDARD, 1 GRAD_ONLY, 2 DERIV1; nalm_c = 1 for
// spin 0 or mode GRAD_ONLY / DERIV1, else 2;
// nmap = 1 for spin 0, else 2; single != 0 -> complex64 / float32
int ref_sht_general(int single, int adjoint, void *alm, size_t nalm,
                    void *map, size_t lmax, const size_t *mstart, size_t nm,
                    ptrdiff_t lstride, const double *loc, size_t npt,
                    double eps, double smin, double smax, size_t nthreads,
                    size_t spin, int mode) {
  return guard([&] {
    cmav<size_t, 1> ms(mstart, {nm});
    cmav<double, 2> lc(loc, {npt, 2});
    const auto md = mode == 1 ? GRAD_ONLY : (mode == 2 ? DERIV1 : STANDARD);
    const size_t na = (spin == 0 || mode != 0) ? 1 : 2,
                 nmp = spin == 0 ? 1 : 2;
    auto run = [&](auto tag) {
      using T = decltype(tag);
      if (adjoint)
        adjoint_synthesis_general(
            vmav<complex<T>, 2>((complex<T> *)alm, {na, nalm}),
            cmav<T, 2>((const T *)map, {nmp, npt}), spin, lmax, ms, lstride,
            lc, eps, smin, smax, nthreads, md, false);
      else
        synthesis_general(cmav<complex<T>, 2>((const complex<T> *)alm,
                                              {na, nalm}),
                          vmav<T, 2>((T *)map, {nmp, npt}), spin, lmax, ms,
                          lstride, lc, eps, smin, smax, nthreads, md, false);
    };
    if (single)
      run(float(0));
    else
      run(double(0));
  });
}

// ducc0.sht.pseudo_analysis_general: map (nmap, npt) real -> alm (na,
// nalm) complex (na = nmap = 1 for spin 0, else 2); res[0..3] = istop,
// itn, |r| / |b|, |A^T r| / (|A| |r|)
int ref_sht_pseudo_analysis(int single, void *alm, size_t nalm,
                            const void *map, size_t lmax, const size_t *mstart,
                            size_t nm, ptrdiff_t lstride, const double *loc,
                            size_t npt, double smin, double smax,
                            size_t nthreads, size_t maxiter, double eps,
                            size_t spin, double *res) {
  return guard([&] {
    cmav<size_t, 1> ms(mstart, {nm});
    cmav<double, 2> lc(loc, {npt, 2});
    const size_t na = spin == 0 ? 1 : 2;
    auto run = [&](auto tag) {
      using T = decltype(tag);
      auto r = pseudo_analysis_general(
          vmav<complex<T>, 2>((complex<T> *)alm, {na, nalm}),
          cmav<T, 2>((const T *)map, {na, npt}), spin, lmax, ms, lstride,
          lc, smin, smax, nthreads, maxiter, eps, false);
      res[0] = (double)std::get<0>(r);
      res[1] = (double)std::get<1>(r);
      res[2] = std::get<2>(r);
      res[3] = std::get<3>(r);
    };
    if (single)
      run(float(0));
    else
      run(double(0));
  });
}

// the w-gridder (vis2dirty when adjoint == 0, dirty2vis when 1): uvw
// (nrow, 3), freq (nchan), vis (nrow, nchan) complex, wgt (nrow, nchan)
// real or NULL, mask (nrow, nchan) uint8 or NULL, dirty (nx, ny) real;
// single != 0 -> float32 / complex64 throughout
int ref_wgridder(int single, int adjoint, const double *uvw, size_t nrow,
                 const double *freq, size_t nchan, void *vis,
                 const void *wgt, const uint8_t *mask, void *dirty, size_t nx,
                 size_t ny, double px, double py, double eps, int do_wg,
                 int flip_u, int flip_v, int flip_w, int divide_by_n,
                 double smin, double smax, double cx, double cy,
                 size_t nthreads) {
  return guard([&] {
    cmav<double, 2> uv(uvw, {nrow, 3});
    cmav<double, 1> fr(freq, {nchan});
    // an empty array means "absent" to the reference
    auto mk = mask ? cmav<uint8_t, 2>(mask, {nrow, nchan})
                   : cmav<uint8_t, 2>(nullptr, {0, 0});
    auto run = [&](auto tag) {
      using T = decltype(tag);
      auto wg = wgt ? cmav<T, 2>((const T *)wgt, {nrow, nchan})
                    : cmav<T, 2>(nullptr, {0, 0});
      if (adjoint) {
        cmav<T, 2> d((const T *)dirty, {nx, ny});
        vmav<complex<T>, 2> v((complex<T> *)vis, {nrow, nchan});
        dirty2ms<T, T>(uv, fr, d, wg, mk, px, py, eps, do_wg != 0, nthreads,
                       v, 0, flip_u != 0, flip_v != 0, flip_w != 0,
                       divide_by_n != 0, smin, smax, cx, cy, true);
      } else {
        cmav<complex<T>, 2> v((const complex<T> *)vis, {nrow, nchan});
        vmav<T, 2> d((T *)dirty, {nx, ny});
        ms2dirty<T, T>(uv, fr, v, wg, mk, px, py, eps, do_wg != 0, nthreads,
                       d, 0, flip_u != 0, flip_v != 0, flip_w != 0,
                       divide_by_n != 0, smin, smax, cx, cy, true);
      }
    };
    if (single)
      run(float(0));
    else
      run(double(0));
  });
}

}  // extern "C"
