// params.cpp -- kernel database, ducc0's analytic parameter choice and the
// construction of the piecewise-polynomial ES kernel and its correction
// quadrature (host side, FP64 / 80-bit).
#include "params.h"

#include <cmath>
#include <stdexcept>

namespace dmx {

const std::vector<KRow> &kernel_db() {
  // rows of ducc0's KernelDB, generated from ducc_amd/data/kernel_db.csv
  static const std::vector<KRow> db = {
#include "_gen/kernel_db.inc"
  };
  return db;
}

// smallest 2,3,5,7,11-smooth number >= n (n <= 12 returns n); same result
// as fft.h:93-119.
size_t good_size_cmplx(size_t n) {
  if (n <= 12) return n;
  size_t best = 2 * n;
  for (size_t a = 1; a < best; a *= 11)
    for (size_t b = a; b < best; b *= 7)
      for (size_t c = b; c < best; c *= 5)
        for (size_t d = c; d < best; d *= 3) {
          size_t x = d;
          while (x < n) x *= 2;
          if (x < best) best = x;
        }
  return best;
}

size_t oversampled_length(size_t n, double ofactor) {
  size_t v = 2 * good_size_cmplx((size_t)(n * ofactor * 0.5) + 1);
  return v < 16 ? 16 : v;
}

// per support W, the row with the smallest oversampling factor in
// [smin, smax] reaching eps (gridding_kernel.h:420-444)
std::vector<int> available_kernels(double eps, int ndim, bool single,
                                   double smin, double smax) {
  const auto &db = kernel_db();
  double ofc[20];
  int id[20];
  const int n = (int)db.size();
  for (int i = 0; i < 20; ++i) { ofc[i] = smax; id[i] = n; }
  const int Wlim = single ? 8 : 16;
  for (int i = 0; i < n; ++i) {
    const KRow &k = db[i];
    if (k.ndim == ndim && k.single == (int)single && k.W <= Wlim &&
        k.epsilon <= eps && k.ofactor <= ofc[k.W] && k.ofactor >= smin) {
      ofc[k.W] = k.ofactor;
      id[k.W] = i;
    }
  }
  std::vector<int> res;
  for (int i = 0; i < 20; ++i)
    if (id[i] < n) res.push_back(id[i]);
  return res;
}

double best_epsilon(int ndim, bool single, double smin, double smax) {
  double res = 1000.;
  for (const auto &k : kernel_db())
    if (k.ndim == ndim && k.single == (int)single && k.epsilon <= res &&
        k.ofactor <= smax && k.ofactor >= smin)
      res = k.epsilon;
  return res < 1000. ? res : -1.;
}

// FFT term ~ N log N (divided by a sigmoid in the thread count) plus a
// spreading term ~ npoints * taps; argmin over the available kernels
// (nufft.h:155-208).
int find_params(double eps, double smin, double smax, int ndim,
                const size_t *dims, size_t npoints, bool gridding,
                bool single, int vlen, size_t nthreads, size_t *nover) {
  auto idx = available_kernels(eps, ndim, single, smin, smax);
  if (idx.empty()) return -1;
  const auto &db = kernel_db();
  const double nref = 2048, costref = 0.0693;
  double mincost = 1e300;
  int best = -1;
  for (int i : idx) {
    const KRow &k = db[i];
    size_t W = (size_t)k.W;
    size_t nvec = (W + vlen - 1) / vlen;
    size_t lb[3];
    double gsz = 1;
    for (int d = 0; d < ndim; ++d) {
      lb[d] = oversampled_length(dims[d], k.ofactor);
      gsz *= (double)lb[d];
    }
    double fftcost = gsz / (nref * nref) * (std::log(gsz) / std::log(nref * nref)) *
                     costref;
    size_t kp = nvec * vlen;
    for (int d = 0; d + 1 < ndim; ++d) kp *= W;
    double gridcost = 2.2e-10 * (double)npoints *
                      ((double)kp + (double)(ndim * nvec * (W + 3) * vlen));
    (void)gridding;  // sizeof(Tacc)/sizeof(Tcalc) == 1 here
    gridcost /= (double)nthreads;
    double x2 = (double)nthreads - 1, m2 = 6. - 1;
    fftcost /= 1. + x2 / std::pow(1. + std::pow(x2 / m2, 2.), 0.5);
    double cost = fftcost + gridcost;
    if (cost < mincost) {
      mincost = cost;
      best = i;
      for (int d = 0; d < ndim; ++d) nover[d] = lb[d];
    }
  }
  return best;
}

double es_kernel(double v, double beta, double e0) {
  double t = (1 - v) * (1 + v);
  if (t < 0) return 0.;
  return std::exp(beta * (std::pow(t, e0) - 1));
}

// Chebyshev interpolation of psi on each of the W pieces at D+1 nodes,
// converted to monomials of the local coordinate (gridding_kernel.cc:28-76).
void kernel_coeffs(int W, int D, double beta, double e0, double *coeff) {
  const double pi = 3.141592653589793238462643383279502884197;
  const int n = D + 1;
  std::vector<double> root(n), y(n), cf(n), mono(n), T((size_t)n * n, 0.);
  for (int i = 0; i < n; ++i) root[i] = std::cos((2 * i + 1.) * pi / (2 * D + 2));
  T[0] = 1.;
  if (n > 1) T[n + 1] = 1.;
  for (int j = 2; j < n; ++j) {
    T[(size_t)j * n] = -T[(size_t)(j - 2) * n];
    for (int k = 1; k <= j; ++k)
      T[(size_t)j * n + k] = 2 * T[(size_t)(j - 1) * n + k - 1] - T[(size_t)(j - 2) * n + k];
  }
  for (int p = 0; p < W; ++p) {
    double l = -1 + 2. * p / W, r = -1 + 2. * (p + 1) / W;
    double avg = 0;
    for (int j = 0; j < n; ++j) {
      y[j] = es_kernel(root[j] * (r - l) * 0.5 + (r + l) * 0.5, beta, e0);
      avg += y[j];
    }
    avg /= n;
    for (int j = 0; j < n; ++j) y[j] -= avg;
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < n; ++k)
        s += 2. / n * y[k] * std::cos(j * (2 * k + 1) * pi / (2 * D + 2));
      cf[j] = s;
    }
    cf[0] *= 0.5;
    for (int k = 0; k < n; ++k) mono[k] = 0;
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) mono[k] += T[(size_t)j * n + k] * cf[j];
    mono[0] += avg;
    for (int j = 0; j < n; ++j) coeff[j * W + p] = mono[D - j];
  }
}

std::vector<double> device_coeffs(int W, double beta_row, double e0) {
  const int D = W + 3;
  std::vector<double> c((size_t)W * (D + 1));
  kernel_coeffs(W, D, beta_row * W, e0, c.data());
  const int DP = W + 3 + (W & 1), NE = (DP + 1) / 2, NH = (W + 1) / 2;
  // coefficient of x^pw of piece i (0 beyond the computed degree)
  auto coef = [&](int pw, int i) -> double {
    if (pw > D) return 0.;
    return c[(size_t)(D - pw) * W + i];
  };
  std::vector<double> out((size_t)2 * NE * NH);
  for (int s = 0; s < NE; ++s)
    for (int i = 0; i < NH; ++i) {
      out[(size_t)(2 * s) * NH + i] = coef(DP - 2 * s, i);
      out[(size_t)(2 * s + 1) * NH + i] = coef(DP - 1 - 2 * s, i);
    }
  return out;
}

// Newton iteration on the Legendre recurrence in long double for the
// (n+1)/2 non-negative roots, ordered from the smallest
// (gl_integrator.cc:38-77, 220-235).
void gl_nodes(int n, double *x, double *w) {
  const long double pi = 3.141592653589793238462643383279502884197L;
  const int m = (n + 1) >> 1;
  for (int ii = 0; ii < m; ++ii) {
    const int i = m - ii;
    const long double dn = n;
    const long double t0 = 1 - (1 - 1 / dn) / (8 * dn * dn);
    const long double t1 = 1 / (4 * dn + 2);
    long double x0 = std::cos((double)(pi * (4 * i - 1) * t1)) * t0;
    long double dpdx = 0;
    bool stop = false;
    for (int it = 0; it < 100; ++it) {
      long double P1 = 1, P0 = x0;
      for (int k = 2; k <= n; ++k) {
        long double P2 = P1;
        P1 = P0;
        P0 = x0 * P1 + ((long double)k - 1) / k * (x0 * P1 - P2);
      }
      long double ax = x0 < 0 ? -x0 : x0;
      long double om = ax > 0.1L ? (1 + ax) * (1 - ax) : 1 - x0 * x0;
      dpdx = (P1 - x0 * P0) * n / om;
      long double x1 = x0 - P0 / dpdx;
      long double dx = x0 - x1;
      x0 = x1;
      if (stop) break;
      if ((dx < 0 ? -dx : dx) <= 3e-14L) stop = true;
    }
    long double ax = x0 < 0 ? -x0 : x0;
    long double om = ax > 0.1L ? (1 + ax) * (1 - ax) : 1 - x0 * x0;
    x[ii] = (double)x0;
    w[ii] = (double)(2. / (om * dpdx * dpdx));
  }
}

// 2p-point rule with p = floor(1.5W)+2; wpsi_i = 2 w_i psi(x_i) W/2 with the
// exact ES kernel (gridding_kernel.h:116-129).
int corr_quadrature(int W, double beta, double e0, double *x, double *wpsi) {
  const int p = (int)(1.5 * W) + 2, nq = 2 * p, m = (nq + 1) >> 1;
  double w[64];
  gl_nodes(nq, x, w);
  for (int i = 0; i < m; ++i)
    wpsi[i] = 2. * w[i] * es_kernel(x[i], beta, e0) * W * 0.5;
  return m;
}

}  // namespace dmx
