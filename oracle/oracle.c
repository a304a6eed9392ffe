/*
 * oracle.c -- CPU restatement of ducc0's NUFFT type 1/2 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in
 * ducc_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, and only to compare against.  Nothing in the product links
 * or calls it.
 *
 * Every function cites the reference lines (paths relative to the ducc0
 * source tree) whose behaviour it restates.  How each one is pinned:
 *   orc_good_size_cmplx   brute-force smooth-number search (tests) + ducc0 build
 *   orc_find_params       compiled ducc0 findNufftParameters (oracle/_ref)
 *   orc_kernel_coeffs     numpy Chebyshev interpolation + ducc0 PolynomialKernel
 *   orc_gl_nodes          numpy.polynomial.legendre.leggauss
 *   orc_corfac            direct quadrature in numpy + ducc0 corfunc
 *   orc_fold / orc_keys   closed-form grid-point cases + ducc0 coord_idx order
 *   orc_sort_stable       numpy argsort(kind='stable') + ducc0 bucket_sort2
 *   orc_fft_c2c           numpy.fft
 *   orc_nu2u / orc_u2nu   direct NUDFT (this file, itself pinned by the
 *                         single-point closed form) + ducc0 nu2u/u2nu outputs
 *   orc_l2error           closed form on small vectors
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------ */
/* kernel database (rows of ducc0's KernelDB, loaded from kernel_db.csv) */
/* ------------------------------------------------------------------ */
typedef struct {
  int W;
  double ofactor, epsilon, beta, e0;
  int ndim, single;
} KRow;

static KRow *g_db = NULL;
static int g_ndb = 0;

int orc_set_db(int n, const int *W, const double *of, const double *eps,
               const double *beta, const double *e0, const int *ndim,
               const int *single) {
  free(g_db);
  g_db = (KRow *)malloc(sizeof(KRow) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    g_db[i].W = W[i];
    g_db[i].ofactor = of[i];
    g_db[i].epsilon = eps[i];
    g_db[i].beta = beta[i];
    g_db[i].e0 = e0[i];
    g_db[i].ndim = ndim[i];
    g_db[i].single = single[i];
  }
  g_ndb = n;
  return 0;
}

/* smallest 2,3,5,7,11-smooth number >= n; n<=12 returns n.
   Follows src/ducc0/fft/fft.h:93-119 (same result, own enumeration). */
size_t orc_good_size_cmplx(size_t n) {
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

/* Wmax: 8 for single, 16 for double (gridding_kernel.h:402-403) */
static int wmax(int single) { return single ? 8 : 16; }

/* For every support W, the row with the smallest oversampling factor in
   [smin, smax] whose tabulated error is <= eps.
   Follows gridding_kernel.h:420-444.  Returns the count written to idx. */
int orc_available_kernels(double eps, int ndim, int single, double smin,
                          double smax, int *idx) {
  double ofc[20];
  int id[20];
  for (int i = 0; i < 20; ++i) { ofc[i] = smax; id[i] = g_ndb; }
  int Wlim = wmax(single);
  for (int i = 0; i < g_ndb; ++i) {
    const KRow *k = &g_db[i];
    if (k->ndim == ndim && k->single == single && k->W <= Wlim &&
        k->epsilon <= eps && k->ofactor <= ofc[k->W] && k->ofactor >= smin) {
      ofc[k->W] = k->ofactor;
      id[k->W] = i;
    }
  }
  int n = 0;
  for (int i = 0; i < 20; ++i)
    if (id[i] < g_ndb) idx[n++] = id[i];
  return n;
}

/* Minimum tabulated epsilon.  Follows gridding_kernel.cc:1562-1574. */
double orc_best_epsilon(int ndim, int single, double smin, double smax) {
  double res = 1000.;
  for (int i = 0; i < g_ndb; ++i) {
    const KRow *k = &g_db[i];
    if (k->ndim == ndim && k->single == single && k->epsilon <= res &&
        k->ofactor <= smax && k->ofactor >= smin)
      res = k->epsilon;
  }
  return res < 1000. ? res : -1.;
}

void orc_db_row(int idx, int *W, double *of, double *eps, double *beta,
                double *e0) {
  *W = g_db[idx].W; *of = g_db[idx].ofactor; *eps = g_db[idx].epsilon;
  *beta = g_db[idx].beta; *e0 = g_db[idx].e0;
}

/* ducc0's analytic cost model over the available kernels.
   Follows nufft.h:155-208: FFT cost ~ N log N scaled by a sigmoid in the
   thread count, spreading cost ~ npoints * (kernel taps + evaluation).
   vlen is the SIMD width ducc0 would use (min(8, native)); sizeratio is
   sizeof(Tacc)/sizeof(Tcalc) (1 for all Python-reachable instantiations).
   Returns the KernelDB row index, or -1 when no kernel qualifies. */
int orc_find_params(double eps, double smin, double smax, int ndim,
                    const size_t *dims, size_t npoints, int gridding,
                    int single, int vlen, size_t nthreads, double sizeratio,
                    size_t *nover) {
  int idx[20];
  int n = orc_available_kernels(eps, ndim, single, smin, smax, idx);
  if (n == 0) return -1;
  const double nref = 2048, costref = 0.0693;
  double mincost = 1e300;
  int best = -1;
  for (int i = 0; i < n; ++i) {
    const KRow *k = &g_db[idx[i]];
    size_t W = (size_t)k->W;
    size_t nvec = (W + vlen - 1) / vlen;
    size_t lb[3];
    double gsz = 1;
    for (int d = 0; d < ndim; ++d) {
      lb[d] = 2 * orc_good_size_cmplx((size_t)(dims[d] * k->ofactor * 0.5) + 1);
      if (lb[d] < 16) lb[d] = 16;
      gsz *= (double)lb[d];
    }
    double logterm = log(gsz) / log(nref * nref);
    double fftcost = gsz / (nref * nref) * logterm * costref;
    size_t kp = nvec * vlen;
    for (int d = 0; d + 1 < ndim; ++d) kp *= W;
    double gridcost = 2.2e-10 * (double)npoints *
                      ((double)kp + (double)(ndim * nvec * (W + 3) * vlen));
    if (gridding) gridcost *= sizeratio;
    gridcost /= (double)nthreads;
    double x2 = (double)nthreads - 1, m2 = 6. - 1;
    double sig = 1. + x2 / pow(1. + pow(x2 / m2, 2.), 0.5);
    fftcost /= sig;
    double cost = fftcost + gridcost;
    if (cost < mincost) {
      mincost = cost;
      best = idx[i];
      for (int d = 0; d < ndim; ++d) nover[d] = lb[d];
    }
  }
  return best;
}

/* ES kernel psi(v) = exp(beta*(((1-v)(1+v))^e0 - 1)) on [-1,1].
   Follows gridding_kernel.cc:1539-1544 (beta already multiplied by W). */
double orc_esk(double v, double beta, double e0) {
  double t = (1 - v) * (1 + v);
  if (t < 0) return 0.;
  return exp(beta * (pow(t, e0) - 1));
}

/* Piecewise polynomial approximation of psi: W pieces on [-1,1], each of
   degree D from Chebyshev interpolation at D+1 nodes, converted to
   monomials of the local coordinate in [-1,1].  coeff[j*W+i] multiplies
   x^(D-j) of piece i.  Follows gridding_kernel.cc:28-76 and :1552-1560
   (beta_row * W, D = W+3). */
void orc_kernel_coeffs(int W, int D, double beta, double e0, double *coeff) {
  const double pi = 3.141592653589793238462643383279502884197;
  int n = D + 1;
  double *root = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
  double *cf = malloc(sizeof(double) * n), *mono = malloc(sizeof(double) * n);
  double *T = calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; ++i) root[i] = cos((2 * i + 1.) * pi / (2 * D + 2));
  /* T[j][k]: coefficient of x^k in Chebyshev polynomial T_j */
  T[0] = 1.;
  if (n > 1) T[1 * n + 1] = 1.;
  for (int j = 2; j < n; ++j) {
    T[j * n] = -T[(j - 2) * n];
    for (int k = 1; k <= j; ++k)
      T[j * n + k] = 2 * T[(j - 1) * n + k - 1] - T[(j - 2) * n + k];
  }
  for (int p = 0; p < W; ++p) {
    double l = -1 + 2. * p / W, r = -1 + 2. * (p + 1) / W;
    double avg = 0;
    for (int j = 0; j < n; ++j) {
      y[j] = orc_esk(root[j] * (r - l) * 0.5 + (r + l) * 0.5, beta, e0);
      avg += y[j];
    }
    avg /= n;
    for (int j = 0; j < n; ++j) y[j] -= avg;
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < n; ++k)
        s += 2. / n * y[k] * cos(j * (2 * k + 1) * pi / (2 * D + 2));
      cf[j] = s;
    }
    cf[0] *= 0.5;
    for (int k = 0; k < n; ++k) mono[k] = 0;
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) mono[k] += T[j * n + k] * cf[j];
    mono[0] += avg;
    for (int j = 0; j < n; ++j) coeff[j * W + p] = mono[D - j];
  }
  free(root); free(y); free(cf); free(mono); free(T);
}

/* Gauss-Legendre: the (n+1)/2 non-negative nodes in decreasing-index order
   of ducc0 (x[0] is the smallest non-negative node) and their weights.
   Newton iteration on the Legendre recurrence in long double.
   Follows gl_integrator.cc:38-77 and :220-235. */
void orc_gl_nodes(int n, double *x, double *w) {
  const long double pi = 3.141592653589793238462643383279502884197L;
  int m = (n + 1) >> 1;
  for (int ii = 0; ii < m; ++ii) {
    int i = m - ii; /* 1-based root counted from the largest */
    long double dn = n;
    long double t0 = 1 - (1 - 1 / dn) / (8 * dn * dn);
    long double t1 = 1 / (4 * dn + 2);
    long double x0 = cos((double)(pi * (4 * i - 1) * t1)) * t0;
    long double dpdx = 0;
    int stop = 0;
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
      if ((dx < 0 ? -dx : dx) <= 3e-14L) stop = 1;
    }
    long double ax = x0 < 0 ? -x0 : x0;
    long double om = ax > 0.1L ? (1 + ax) * (1 - ax) : 1 - x0 * x0;
    x[ii] = (double)x0;
    w[ii] = (double)(2. / (om * dpdx * dpdx));
  }
}

/* Correction factors corfac[k] = 1/sum_i wpsi_i cos(pi*W*x_i*k*dx),
   k = 0..n-1, with x_i the non-negative nodes of a 2p-point Gauss-Legendre
   rule, p = floor(1.5W)+2, and wpsi_i = w_i^sym * psi(x_i) * W/2 using the
   exact ES kernel.  Follows gridding_kernel.h:68-129. */
void orc_corfac(int W, double beta, double e0, size_t n, double dx,
                double *out) {
  const double pi = 3.141592653589793238462643383279502884197;
  int p = (int)(1.5 * W) + 2;
  int nq = 2 * p, m = (nq + 1) >> 1;
  double x[64], w[64];
  orc_gl_nodes(nq, x, w);
  double wpsi[64];
  for (int i = 0; i < m; ++i) {
    double ws = w[i];
    if (nq & 1 && i == 0) ws *= 0.5;
    ws *= 2;
    wpsi[i] = ws * orc_esk(x[i], beta, e0) * W * 0.5;
  }
  for (size_t k = 0; k < n; ++k) {
    double v = (double)k * dx, s = 0;
    for (int i = 0; i < m; ++i) s += wpsi[i] * cos((pi * W * x[i]) * v);
    out[k] = 1. / s;
  }
}

/* ------------------------------------------------------------------ */
/* folding and tile keys                                               */
/* ------------------------------------------------------------------ */
typedef struct {
  int ndim, W, nsafe;
  size_t nover[3];
  double coordfct[3], shift[3];
  int maxi0[3];
  int coord_double; /* 1: range reduction in long double (Tcoord=double) */
} Geo;

static void geo_init(Geo *g, int ndim, int W, const size_t *nover,
                     const double *period, int coord_double) {
  g->ndim = ndim;
  g->W = W;
  g->nsafe = (W + 1) / 2;
  g->coord_double = coord_double;
  for (int d = 0; d < ndim; ++d) {
    g->nover[d] = nover[d];
    g->coordfct[d] = 1. / period[d];
    g->shift[d] = W * (-0.5) + 1 + (double)nover[d];
    g->maxi0[d] = (int)(nover[d] + g->nsafe) - W;
  }
}

/* Fold one coordinate: first touched grid index i0 and the distance frac
   from i0 to the point in grid units.  Follows nufft.h:285-298. */
static void fold1(const Geo *g, int d, double x, int *i0, double *frac) {
  double tmp = x * g->coordfct[d];
  if (g->coord_double) {
    long double t2 = (long double)(tmp - floor(tmp)) * (long double)g->nover[d];
    int v = (int)(t2 + (long double)g->shift[d]) - (int)g->nover[d];
    if (v > g->maxi0[d]) v = g->maxi0[d];
    *i0 = v;
    *frac = (double)(t2 - v);
  } else {
    double t2 = (tmp - floor(tmp)) * (double)g->nover[d];
    int v = (int)(t2 + g->shift[d]) - (int)g->nover[d];
    if (v > g->maxi0[d]) v = g->maxi0[d];
    *i0 = v;
    *frac = t2 - v;
  }
}

void orc_fold(int ndim, int W, const size_t *nover, const double *period,
              int coord_double, size_t M, const double *coords, int *i0,
              double *frac) {
  Geo g;
  geo_init(&g, ndim, W, nover, period, coord_double);
  for (size_t j = 0; j < M; ++j)
    for (int d = 0; d < ndim; ++d)
      fold1(&g, d, coords[j * ndim + d], &i0[j * ndim + d],
            &frac[j * ndim + d]);
}

/* Tile key of each point.  1D/2D: linear tile index with
   ntiles = (nover >> log2tile) + 3 per axis (nufft.h:892-907, 1364-1381).
   3D: when lsq2 < log2tile, ducc0's refined key of nufft.h:1862-1892;
   lsq2 == log2tile gives the plain linear tile index.  A negative lsq2 asks
   for ducc0's own refinement rule (2^28 bins). */
void orc_keys(int ndim, int W, const size_t *nover, const double *period,
              int coord_double, int log2tile, int lsq2, size_t M,
              const double *coords, uint32_t *keys) {
  Geo g;
  geo_init(&g, ndim, W, nover, period, coord_double);
  size_t nt[3];
  for (int d = 0; d < ndim; ++d) nt[d] = (nover[d] >> log2tile) + 3;
  if (ndim == 3 && lsq2 < 0) {
    lsq2 = log2tile;
    while (lsq2 >= 1 &&
           ((nt[0] * nt[1] * nt[2]) << (3 * (log2tile - lsq2))) <
               ((size_t)1 << 28))
      --lsq2;
  }
  if (ndim != 3) lsq2 = log2tile;
  int ss = log2tile - lsq2;
  uint32_t msk = (1u << ss) - 1;
  for (size_t j = 0; j < M; ++j) {
    uint32_t t[3];
    for (int d = 0; d < ndim; ++d) {
      int i0;
      double fr;
      fold1(&g, d, coords[j * ndim + d], &i0, &fr);
      t[d] = (uint32_t)((i0 + g.nsafe) >> lsq2);
    }
    if (ndim == 1)
      keys[j] = t[0];
    else if (ndim == 2)
      keys[j] = (uint32_t)(t[0] * nt[1] + t[1]);
    else {
      uint32_t lo = ((t[0] & msk) << (2 * ss)) | ((t[1] & msk) << ss) |
                    (t[2] & msk);
      uint32_t hi = (uint32_t)((t[0] >> ss) * nt[1] * nt[2] +
                               (t[1] >> ss) * nt[2] + (t[2] >> ss));
      keys[j] = (hi << (3 * ss)) | lo;
    }
  }
}

/* Stable ascending order of keys (equal keys keep input order): the order
   ducc0's bucket_sort2 produces (bucket_sort.h:170-354).  Plain counting
   sort over [0, maxkey]. */
void orc_sort_stable(size_t M, const uint32_t *keys, uint32_t maxkey,
                     uint32_t *idx) {
  size_t nb = (size_t)maxkey + 2;
  size_t *cnt = calloc(nb, sizeof(size_t));
  for (size_t j = 0; j < M; ++j) cnt[keys[j] + 1]++;
  for (size_t b = 1; b < nb; ++b) cnt[b] += cnt[b - 1];
  for (size_t j = 0; j < M; ++j) idx[cnt[keys[j]]++] = (uint32_t)j;
  free(cnt);
}

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed radix, unnormalised, exponent sign -1 if forward */
/* ------------------------------------------------------------------ */
static void fft_rec(cplx *x, size_t n, size_t stride, cplx *out, cplx *tmp,
                    int fwd) {
  const double pi = 3.141592653589793238462643383279502884197;
  if (n == 1) { out[0] = x[0]; return; }
  size_t p = 0;
  for (size_t f = 2; f * f <= n; ++f)
    if (n % f == 0) { p = f; break; }
  if (p == 0) p = n;
  size_t m = n / p;
  double sg = fwd ? -1. : 1.;
  if (m == 1) { /* plain DFT of prime length */
    for (size_t k = 0; k < n; ++k) {
      cplx s = 0;
      for (size_t j = 0; j < n; ++j) {
        double ang = sg * 2 * pi * (double)((j * k) % n) / (double)n;
        s += x[j * stride] * (cos(ang) + I * sin(ang));
      }
      out[k] = s;
    }
    return;
  }
  /* decimation in time: p sub-transforms of length m */
  for (size_t r = 0; r < p; ++r)
    fft_rec(x + r * stride, m, stride * p, tmp + r * m, out, fwd);
  for (size_t k = 0; k < m; ++k)
    for (size_t q = 0; q < p; ++q) {
      size_t kk = k + q * m;
      cplx s = 0;
      for (size_t r = 0; r < p; ++r) {
        double ang = sg * 2 * pi * (double)((r * kk) % n) / (double)n;
        s += tmp[r * m + k] * (cos(ang) + I * sin(ang));
      }
      out[kk] = s;
    }
}

/* In-place c2c along one axis of a C-ordered array of shape dims[0..ndim). */
static void fft_axis(cplx *a, int ndim, const size_t *dims, int axis,
                     int fwd) {
  size_t n = dims[axis], inner = 1, outer = 1;
  for (int d = axis + 1; d < ndim; ++d) inner *= dims[d];
  for (int d = 0; d < axis; ++d) outer *= dims[d];
  cplx *line = malloc(sizeof(cplx) * n), *out = malloc(sizeof(cplx) * n);
  cplx *tmp = malloc(sizeof(cplx) * n);
  for (size_t o = 0; o < outer; ++o)
    for (size_t i = 0; i < inner; ++i) {
      cplx *base = a + o * n * inner + i;
      for (size_t k = 0; k < n; ++k) line[k] = base[k * inner];
      fft_rec(line, n, 1, out, tmp, fwd);
      for (size_t k = 0; k < n; ++k) base[k * inner] = out[k];
    }
  free(line); free(out); free(tmp);
}

void orc_fft_c2c(int ndim, const size_t *dims, double *data, int fwd) {
  for (int ax = 0; ax < ndim; ++ax)
    fft_axis((cplx *)data, ndim, dims, ax, fwd);
}

/* ------------------------------------------------------------------ */
/* the NUFFT itself                                                    */
/* ------------------------------------------------------------------ */
/* comp_indices of nufft.h:142-151: corfac index, user index, grid index */
static void comp_idx(size_t idx, size_t nuni, size_t nbig, int fft_order,
                     size_t *icf, size_t *iuser, size_t *igrid) {
  long a = (long)(nuni / 2) - (long)idx;
  *icf = (size_t)(a < 0 ? -a : a);
  size_t i1 = fft_order ? nuni - nuni / 2 + idx : idx;
  if (i1 >= nuni) i1 -= nuni;
  *iuser = i1;
  size_t i2 = nbig - nuni / 2 + idx;
  if (i2 >= nbig) i2 -= nbig;
  *igrid = i2;
}

typedef struct {
  Geo g;
  int D;
  double *coeff;      /* W*(D+1) */
  double *corfac[3];
  size_t nuni[3];
  size_t ntot_over;
} Plan;

/* kernel weights of the W grid points i0..i0+W-1 for local coordinate
   x = W-1-2*frac (nufft.h:976-981; gridding_kernel.h:134-174) */
static void kweights(const Plan *P, double frac, double *w) {
  int W = P->g.W, D = P->D;
  double x = -frac * 2 + (W - 1);
  for (int i = 0; i < W; ++i) {
    double r = P->coeff[i];
    for (int j = 1; j <= D; ++j) r = r * x + P->coeff[j * W + i];
    w[i] = r;
  }
}

static int plan_init(Plan *P, int ndim, const size_t *nuni, size_t M,
                     double eps, double smin, double smax,
                     const double *period, int gridding, int single, int vlen,
                     size_t nthreads, int coord_double, int *kidx_out,
                     size_t *nover_out) {
  size_t nover[3];
  int kidx = orc_find_params(eps, smin, smax, ndim, nuni, M, gridding, single,
                             vlen, nthreads, 1., nover);
  if (kidx < 0) return -1;
  const KRow *k = &g_db[kidx];
  geo_init(&P->g, ndim, k->W, nover, period, coord_double);
  P->D = k->W + 3;
  P->coeff = malloc(sizeof(double) * k->W * (P->D + 1));
  orc_kernel_coeffs(k->W, P->D, k->beta * k->W, k->e0, P->coeff);
  P->ntot_over = 1;
  for (int d = 0; d < ndim; ++d) {
    P->nuni[d] = nuni[d];
    P->corfac[d] = malloc(sizeof(double) * (nuni[d] / 2 + 1));
    orc_corfac(k->W, k->beta * k->W, k->e0, nuni[d] / 2 + 1,
               1. / (double)nover[d], P->corfac[d]);
    P->ntot_over *= nover[d];
  }
  if (kidx_out) *kidx_out = kidx;
  if (nover_out)
    for (int d = 0; d < ndim; ++d) nover_out[d] = nover[d];
  return 0;
}

static void plan_free(Plan *P) {
  free(P->coeff);
  for (int d = 0; d < P->g.ndim; ++d) free(P->corfac[d]);
}

/* Type 1 (nonuniform -> uniform).  Spread with the piecewise polynomial
   kernel onto the oversampled grid (periodic wrap), c2c FFT, crop the
   central modes and multiply by the correction factors.
   Follows nufft.h:1274-1315 (2D), 836-862 (1D), 1731-1794 (3D).
   coords: (M, ndim) float64 row-major; points: M complex128;
   out: C-ordered nuni complex128.  Returns 0, or -1 when no kernel. */
int orc_nu2u(int ndim, size_t M, const double *coords, const double *points,
             int forward, double eps, double smin, double smax,
             const double *period, int fft_order, const size_t *nuni,
             int single, int vlen, size_t nthreads, int coord_double,
             double *out, int *kidx_out, size_t *nover_out) {
  Plan P;
  if (plan_init(&P, ndim, nuni, M, eps, smin, smax, period, 1, single, vlen,
                nthreads, coord_double, kidx_out, nover_out))
    return -1;
  const Geo *g = &P.g;
  int W = g->W;
  cplx *grid = calloc(P.ntot_over, sizeof(cplx));
  const cplx *pts = (const cplx *)points;
  double wk[3][32];
  for (size_t j = 0; j < M; ++j) {
    int i0[3];
    for (int d = 0; d < ndim; ++d) {
      double fr;
      fold1(g, d, coords[j * ndim + d], &i0[d], &fr);
      kweights(&P, fr, wk[d]);
    }
    cplx v = pts[j];
    if (ndim == 1) {
      for (int a = 0; a < W; ++a) {
        size_t ia = (size_t)((i0[0] + a + (int)g->nover[0]) % (int)g->nover[0]);
        grid[ia] += v * wk[0][a];
      }
    } else if (ndim == 2) {
      for (int a = 0; a < W; ++a) {
        size_t ia = (size_t)((i0[0] + a + (int)g->nover[0]) % (int)g->nover[0]);
        for (int b = 0; b < W; ++b) {
          size_t ib = (size_t)((i0[1] + b + (int)g->nover[1]) % (int)g->nover[1]);
          grid[ia * g->nover[1] + ib] += v * (wk[0][a] * wk[1][b]);
        }
      }
    } else {
      for (int a = 0; a < W; ++a) {
        size_t ia = (size_t)((i0[0] + a + (int)g->nover[0]) % (int)g->nover[0]);
        for (int b = 0; b < W; ++b) {
          size_t ib = (size_t)((i0[1] + b + (int)g->nover[1]) % (int)g->nover[1]);
          for (int c = 0; c < W; ++c) {
            size_t ic =
                (size_t)((i0[2] + c + (int)g->nover[2]) % (int)g->nover[2]);
            grid[(ia * g->nover[1] + ib) * g->nover[2] + ic] +=
                v * (wk[0][a] * wk[1][b] * wk[2][c]);
          }
        }
      }
    }
  }
  orc_fft_c2c(ndim, g->nover, (double *)grid, forward);
  cplx *o = (cplx *)out;
  size_t n0 = nuni[0], n1 = ndim > 1 ? nuni[1] : 1, n2 = ndim > 2 ? nuni[2] : 1;
  size_t N1 = ndim > 1 ? g->nover[1] : 1, N2 = ndim > 2 ? g->nover[2] : 1;
  for (size_t i = 0; i < n0; ++i) {
    size_t ci, ui, gi;
    comp_idx(i, n0, g->nover[0], fft_order, &ci, &ui, &gi);
    for (size_t j = 0; j < n1; ++j) {
      size_t cj = 0, uj = 0, gj = 0;
      if (ndim > 1) comp_idx(j, n1, g->nover[1], fft_order, &cj, &uj, &gj);
      for (size_t k = 0; k < n2; ++k) {
        size_t ck = 0, uk = 0, gk = 0;
        if (ndim > 2) comp_idx(k, n2, g->nover[2], fft_order, &ck, &uk, &gk);
        double f = P.corfac[0][ci];
        if (ndim > 1) f *= P.corfac[1][cj];
        if (ndim > 2) f *= P.corfac[2][ck];
        o[(ui * n1 + uj) * n2 + uk] = grid[(gi * N1 + gj) * N2 + gk] * f;
      }
    }
  }
  free(grid);
  plan_free(&P);
  return 0;
}

/* Type 2 (uniform -> nonuniform): multiply by the correction factors and
   zero-pad onto the oversampled grid, c2c FFT, interpolate with the
   polynomial kernel.  Follows nufft.h:1317-1362 (2D), 864-890 (1D),
   1796-1860 (3D). */
int orc_u2nu(int ndim, size_t M, const double *coords, const double *grid_in,
             int forward, double eps, double smin, double smax,
             const double *period, int fft_order, const size_t *nuni,
             int single, int vlen, size_t nthreads, int coord_double,
             double *out, int *kidx_out, size_t *nover_out) {
  Plan P;
  if (plan_init(&P, ndim, nuni, M, eps, smin, smax, period, 0, single, vlen,
                nthreads, coord_double, kidx_out, nover_out))
    return -1;
  const Geo *g = &P.g;
  int W = g->W;
  cplx *grid = calloc(P.ntot_over, sizeof(cplx));
  const cplx *u = (const cplx *)grid_in;
  size_t n0 = nuni[0], n1 = ndim > 1 ? nuni[1] : 1, n2 = ndim > 2 ? nuni[2] : 1;
  size_t N1 = ndim > 1 ? g->nover[1] : 1, N2 = ndim > 2 ? g->nover[2] : 1;
  for (size_t i = 0; i < n0; ++i) {
    size_t ci, ui, gi;
    comp_idx(i, n0, g->nover[0], fft_order, &ci, &ui, &gi);
    for (size_t j = 0; j < n1; ++j) {
      size_t cj = 0, uj = 0, gj = 0;
      if (ndim > 1) comp_idx(j, n1, g->nover[1], fft_order, &cj, &uj, &gj);
      for (size_t k = 0; k < n2; ++k) {
        size_t ck = 0, uk = 0, gk = 0;
        if (ndim > 2) comp_idx(k, n2, g->nover[2], fft_order, &ck, &uk, &gk);
        double f = P.corfac[0][ci];
        if (ndim > 1) f *= P.corfac[1][cj];
        if (ndim > 2) f *= P.corfac[2][ck];
        grid[(gi * N1 + gj) * N2 + gk] = u[(ui * n1 + uj) * n2 + uk] * f;
      }
    }
  }
  orc_fft_c2c(ndim, g->nover, (double *)grid, forward);
  cplx *o = (cplx *)out;
  double wk[3][32];
  for (size_t j = 0; j < M; ++j) {
    int i0[3];
    for (int d = 0; d < ndim; ++d) {
      double fr;
      fold1(g, d, coords[j * ndim + d], &i0[d], &fr);
      kweights(&P, fr, wk[d]);
    }
    cplx s = 0;
    if (ndim == 1) {
      for (int a = 0; a < W; ++a) {
        size_t ia = (size_t)((i0[0] + a + (int)g->nover[0]) % (int)g->nover[0]);
        s += grid[ia] * wk[0][a];
      }
    } else if (ndim == 2) {
      for (int a = 0; a < W; ++a) {
        size_t ia = (size_t)((i0[0] + a + (int)g->nover[0]) % (int)g->nover[0]);
        cplx r = 0;
        for (int b = 0; b < W; ++b) {
          size_t ib = (size_t)((i0[1] + b + (int)g->nover[1]) % (int)g->nover[1]);
          r += grid[ia * N1 + ib] * wk[1][b];
        }
        s += r * wk[0][a];
      }
    } else {
      for (int a = 0; a < W; ++a) {
        size_t ia = (size_t)((i0[0] + a + (int)g->nover[0]) % (int)g->nover[0]);
        for (int b = 0; b < W; ++b) {
          size_t ib = (size_t)((i0[1] + b + (int)g->nover[1]) % (int)g->nover[1]);
          cplx r = 0;
          for (int c = 0; c < W; ++c) {
            size_t ic =
                (size_t)((i0[2] + c + (int)g->nover[2]) % (int)g->nover[2]);
            r += grid[(ia * N1 + ib) * N2 + ic] * wk[2][c];
          }
          s += r * (wk[0][a] * wk[1][b]);
        }
      }
    }
    o[j] = s;
  }
  free(grid);
  plan_free(&P);
  return 0;
}

/* ------------------------------------------------------------------ */
/* direct NUDFT (the brute-force reference of test_nufft.py:32-43)     */
/* ------------------------------------------------------------------ */
/* Mode m along an axis of length n is k = m - n/2; user index follows
   fft_order (ifftshift) as in comp_indices.  Phase sign -1 if forward.
   Type 1: out[k] = sum_j c_j exp(isign*i*2pi/P * k.x_j), over the
   modes listed in `modes` (flat user indices), or all when modes==NULL. */
static void mode_of_user_index(size_t u, size_t n, int fft_order, long *k) {
  /* inverse of comp_indices' user index */
  size_t idx = fft_order ? (u + n / 2) % n : u;
  *k = (long)idx - (long)(n / 2);
}

void orc_nudft_type1(int ndim, size_t M, const double *coords,
                     const double *points, int forward, const double *period,
                     int fft_order, const size_t *nuni, size_t nmodes,
                     const int64_t *modes, double *out) {
  const double pi = 3.141592653589793238462643383279502884197;
  size_t n1 = ndim > 1 ? nuni[1] : 1, n2 = ndim > 2 ? nuni[2] : 1;
  size_t ntot = nuni[0] * n1 * n2;
  if (!modes) nmodes = ntot;
  const cplx *c = (const cplx *)points;
  cplx *o = (cplx *)out;
  double sg = forward ? -1. : 1.;
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t q = 0; q < nmodes; ++q) {
    size_t flat = modes ? (size_t)modes[q] : q;
    size_t u[3] = {flat / (n1 * n2), (flat / n2) % n1, flat % n2};
    long k[3] = {0, 0, 0};
    for (int d = 0; d < ndim; ++d)
      mode_of_user_index(u[d], nuni[d], fft_order, &k[d]);
    cplx s = 0;
    for (size_t j = 0; j < M; ++j) {
      double ph = 0;
      for (int d = 0; d < ndim; ++d)
        ph += (double)k[d] * (2 * pi / period[d]) * coords[j * ndim + d];
      s += c[j] * (cos(sg * ph) + I * sin(sg * ph));
    }
    o[q] = s;
  }
}

/* Type 2: out[j] = sum_k g[k] exp(isign*i*2pi/P * k.x_j) over the points
   listed in `pidx` (or all). */
void orc_nudft_type2(int ndim, size_t M, const double *coords,
                     const double *grid_in, int forward, const double *period,
                     int fft_order, const size_t *nuni, size_t npts,
                     const int64_t *pidx, double *out) {
  const double pi = 3.141592653589793238462643383279502884197;
  size_t n1 = ndim > 1 ? nuni[1] : 1, n2 = ndim > 2 ? nuni[2] : 1;
  size_t ntot = nuni[0] * n1 * n2;
  if (!pidx) npts = M;
  const cplx *gr = (const cplx *)grid_in;
  cplx *o = (cplx *)out;
  double sg = forward ? -1. : 1.;
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t q = 0; q < npts; ++q) {
    size_t j = pidx ? (size_t)pidx[q] : q;
    cplx s = 0;
    for (size_t flat = 0; flat < ntot; ++flat) {
      size_t u[3] = {flat / (n1 * n2), (flat / n2) % n1, flat % n2};
      double ph = 0;
      for (int d = 0; d < ndim; ++d) {
        long k;
        mode_of_user_index(u[d], nuni[d], fft_order, &k);
        ph += (double)k * (2 * pi / period[d]) * coords[j * ndim + d];
      }
      s += gr[flat] * (cos(sg * ph) + I * sin(sg * ph));
    }
    o[q] = s;
  }
}

/* sqrt(sum|a-b|^2 / max(sum|a|^2, sum|b|^2)) accumulated in long double.
   Follows python/misc_pymod.cc:129-170. */
double orc_l2error(size_t n, const double *a, const double *b) {
  long double sa = 0, sb = 0, sd = 0;
  for (size_t i = 0; i < n; ++i) {
    long double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i],
                bi = b[2 * i + 1];
    sa += ar * ar + ai * ai;
    sb += br * br + bi * bi;
    sd += (ar - br) * (ar - br) + (ai - bi) * (ai - bi);
  }
  long double m = sa > sb ? sa : sb;
  if (m == 0) return 0.;
  return (double)sqrtl(sd / m);
}
