// params.h -- host-side parameter choice and kernel construction.
#pragma once
#include <cstddef>
#include <vector>

namespace dmx {

// One row of the tabulated ES-kernel database (gridding_kernel.cc:78-1537).
struct KRow {
  int W;
  double ofactor, epsilon, beta, e0;
  int ndim, single;
};
const std::vector<KRow> &kernel_db();

size_t good_size_cmplx(size_t n);
std::vector<int> available_kernels(double eps, int ndim, bool single,
                                   double smin, double smax);
double best_epsilon(int ndim, bool single, double smin, double smax);
// ducc0's cost model; returns the row index (or -1) and the oversampled
// shape.  vlen / nthreads are the host SIMD width and thread count that the
// model assumes.
// tile side (log2) of the spreading/interpolation subproblems before the
// LDS limit is applied: 2^9 (1-D), 2^5 / 2^4 (2-D FP32 / FP64), 2^3 (3-D)
int default_log2tile(int ndim, bool fp64);
int find_params(double eps, double smin, double smax, int ndim,
                const size_t *dims, size_t npoints, bool gridding,
                bool single, int vlen, size_t nthreads, size_t *nover);
// oversampled length for a given oversampling factor (nufft.h:176-177)
size_t oversampled_length(size_t n, double ofactor);

// Piecewise polynomial coefficients in the reference layout:
// coeff[j*W+i] multiplies x^(D-j) of piece i.
void kernel_coeffs(int W, int D, double beta, double e0, double *coeff);
// Device layout for kernel_weights<T,W> (see dmx.h), degree raised to
// DP = W+3+(W&1); returns NE*2*NH values.
std::vector<double> device_coeffs(int W, double beta_row, double e0);

// Gauss-Legendre non-negative nodes and weights of an n-point rule.
void gl_nodes(int n, double *x, double *w);
// Nodes x_i and weights wpsi_i of the correction-factor quadrature.
int corr_quadrature(int W, double beta, double e0, double *x, double *wpsi);
double es_kernel(double v, double beta, double e0);

}  // namespace dmx
