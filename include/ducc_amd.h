/*
 * ducc_amd.h -- C ABI of the MI355X-native NUFFT (types 1 and 2, 1-3 D).
 *
 * Drop-in boundary for ducc0.nufft (python/nufft_pymod.cc:445-471): every
 * entry point is extern "C", noexcept in effect, returns 0 on success and a
 * nonzero code on failure, with the message available from dmx_last_error()
 * (thread-local).  Arguments are validated on the host before any GPU work;
 * messages follow ducc0's MR_assert texts where ducc0 has one.
 *
 * Arrays are passed as dmx_array descriptors (role of ducc0's
 * ArrayDescriptor, src/ducc0/bindings/array_descriptor.h:34-47), in C
 * (row-major) axis order, strides in elements.  `device` = 0 means host
 * memory (copied in/out by the library), 1 means device memory on the
 * current HIP device (used in place).  The caller owns every buffer.
 */
#ifndef DUCC_AMD_H
#define DUCC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* typecodes (the values of src/ducc0/bindings/typecode.h:44-80: element
   size - 1, plus 32 for unsigned integers, plus 64 for complex) */
enum {
  DMX_F32 = 3,
  DMX_F64 = 7,
  DMX_C64 = 67,
  DMX_C128 = 71,
  DMX_U8 = 32,
  DMX_U32 = 35
};

typedef struct dmx_array {
  void *data;
  int64_t ndim;
  int64_t shape[10];
  int64_t stride[10]; /* in elements */
  int32_t typecode;
  int32_t device; /* 0 host, 1 device */
} dmx_array;

/* Options with no ducc0 counterpart (precision policy, parameter chooser,
   stream).  A NULL pointer means all defaults. */
typedef struct dmx_options {
  int32_t precision; /* 0 auto (FP32 when eps >= fp64_below and the
                        single-precision kernel table reaches eps), 1 FP32,
                        2 FP64 */
  int32_t chooser;   /* 0 analytic (ducc0's cost model), 1 learned */
  int32_t log2tile;  /* 0 automatic, else tile side 2^log2tile */
  int32_t kernel_idx;/* -1 automatic, else force this kernel table row */
  double fp64_below; /* auto policy switch, default 1e-6 */
  void *stream;      /* hipStream_t to run on; NULL: library stream (the
                        default stream when DMX_STREAM_GIVEN is set) */
  double sigma;      /* >0: force this oversampling factor (with kernel_idx) */
  int32_t spot_check;/* learned choices: verify accuracy once per request
                        signature with an FP64 direct sum, fall back to the
                        analytic choice on failure (default 1; 0 off; 2
                        treats the check as failed, for tests) */
  int32_t flags;     /* DMX_STREAM_GIVEN | DMX_ASYNC */
  int32_t deterministic; /* 1: type-1 results identical bit for bit from
                            run to run (no floating-point atomics in the
                            spreading; slower, and per-subproblem FP64
                            regions in device memory); type 2 is always
                            deterministic */
  int32_t reserved[5];
} dmx_options;

/* dmx_options.flags */
#define DMX_STREAM_GIVEN 1 /* run on `stream` even when it is NULL */
#define DMX_ASYNC 2        /* when every array is in device memory, return
                              once the work is queued on the stream (results
                              are ordered on that stream, as torch
                              operations are); host arrays always wait */
#define DMX_POINT_RECORDS 4 /* unplanned transforms sort the points' records
                              (tile-local offsets, kernel arguments and
                              value or input index) instead of a
                              permutation, so that spreading and
                              interpolation read them in order; same results
                              (type 2 bit for bit).  Off by default: on
                              MI355X the record sort moves 2.5x the bytes
                              of the index sort and costs more than the
                              gathers it saves */
#define DMX_SPREAD_FP64_ACC 16 /* type-1 spreading of the FP32 path
                              accumulates each tile region in FP64 (two LDS
                              float atomics per tap) also where it would
                              use packed 32.32 fixed point (one 64-bit
                              integer atomic per tap; eps >= 1e-5) */
#define DMX_CHECK_OFFSETS 8 /* debugging aid: spreading and interpolation
                              count the points whose fold lands outside the
                              tile the index build put them in (the kernels
                              clamp such offsets), and the call fails with
                              an error when any did; one device counter and
                              a wait per call */

typedef struct dmx_plan dmx_plan;

/* Parameters a plan or call actually used. */
typedef struct dmx_plan_info {
  int32_t ndim, W, kernel_idx, fp64, log2tile, chooser_used;
  int64_t nuni[3], nover[3];
  double ofactor, eps_tab, epsilon;
  int64_t npoints;
} dmx_plan_info;

/* Last error message of the calling thread ("" when none). */
const char *dmx_last_error(void);

/* Fill `o` with the defaults. */
void dmx_default_options(dmx_options *o);

/* ducc0.nufft.bestEpsilon (gridding_kernel.cc:1562-1574); negative on error. */
double dmx_best_epsilon(size_t ndim, int singleprec, double sigma_min,
                        double sigma_max);

/* ducc0.nufft.u2nu (python/nufft_pymod.cc:52-94; nufft.h:1929-1958).
   grid: complex nuni array; coord: (M, ndim) real; out: (M,) complex.
   periodicity: ndim doubles.  `info` (may be NULL) receives the parameters
   used.  verbosity > 0 prints the request and a phase-timing table to
   stdout (nufft.h:371-380, 469-471) and makes the call wait for its work;
   nthreads is accepted and ignored. */
int dmx_u2nu(const dmx_array *grid, const dmx_array *coord, int forward,
             double epsilon, size_t nthreads, dmx_array *out,
             size_t verbosity, double sigma_min, double sigma_max,
             const double *periodicity, int fft_order,
             const dmx_options *opt, dmx_plan_info *info);

/* ducc0.nufft.nu2u (python/nufft_pymod.cc:96-137; nufft.h:1897-1927). */
int dmx_nu2u(const dmx_array *points, const dmx_array *coord, int forward,
             double epsilon, size_t nthreads, dmx_array *out,
             size_t verbosity, double sigma_min, double sigma_max,
             const double *periodicity, int fft_order,
             const dmx_options *opt, dmx_plan_info *info);

/* ducc0.nufft.plan (python/nufft_pymod.cc:139-261; nufft.h:437-503):
   sorts the points once; the plan keeps the sorted index and a sorted copy
   of the coordinates on the device. */
int dmx_plan_create(dmx_plan **plan, int nu2u, const dmx_array *coord,
                    const int64_t *grid_shape, size_t ndim, double epsilon,
                    size_t nthreads, double sigma_min, double sigma_max,
                    const double *periodicity, int fft_order,
                    const dmx_options *opt);
/* opt (may be NULL): the stream and DMX_ASYNC flag of this call (the
   other fields are ignored); NULL runs on the plan's creation stream with
   its creation flags.  A call on another stream than the plan's previous
   one first waits there for the work the plan queued before (the plan
   reuses its buffers).  verbosity > 0 prints the request and a table of
   phase timings to stdout (as ducc0 does, nufft.h:469-471) and makes the
   call wait for its work. */
int dmx_plan_nu2u(dmx_plan *plan, int forward, size_t verbosity,
                  const dmx_array *points, dmx_array *out,
                  const dmx_options *opt);
int dmx_plan_u2nu(dmx_plan *plan, int forward, size_t verbosity,
                  const dmx_array *grid, dmx_array *out,
                  const dmx_options *opt);
int dmx_plan_get_info(const dmx_plan *plan, dmx_plan_info *info);
int dmx_plan_destroy(dmx_plan *plan);

/* Gradient of u2nu with respect to the point coordinates (no ducc0
   counterpart; used by the autograd wrapper): for every point j and axis d,
   out[j*ndim+d] = d u2nu(grid)_j / d coord[j,d], computed by interpolating
   with the derivative of the spreading kernel.  Same arguments as dmx_u2nu;
   out is an (M, ndim) complex array of the grid's precision. */
int dmx_u2nu_coord_grad(const dmx_array *grid, const dmx_array *coord,
                        int forward, double epsilon, dmx_array *out,
                        double sigma_min, double sigma_max,
                        const double *periodicity, int fft_order,
                        const dmx_options *opt);

/* Direct FP64 NUDFT on the device, the spot-check used by the parameter
   chooser and the error reports (sum of test_nufft.py:32-43).
   type 1: out[q] = sum_j pts_j exp(isign i 2pi/P k(q).x_j) for the modes
   listed in `sel` (flat user indices, int64); type 2: out[q] = sum_k
   grid_k exp(isign i 2pi/P k.x_{sel[q]}).  All arrays device or host,
   complex128 / float64. */
int dmx_nudft(int type, const dmx_array *coord, const dmx_array *values,
              const int64_t *shape, size_t ndim, int forward,
              const double *periodicity, int fft_order, const int64_t *sel,
              size_t nsel, dmx_array *out);

/* Stable sort of points into tiles (the index build of nufft.h:1364-1381):
   returns the tile keys and the permutation for given geometry.  Exposed for
   bit-exact parity tests.  keys/idx: device or host uint32 arrays of M. */
int dmx_sort_points(const dmx_array *coord, const int64_t *nover,
                    size_t ndim, int W, int log2tile,
                    const double *periodicity, dmx_array *keys,
                    dmx_array *idx);

/* The same sort carrying the points' records (the index build of unplanned
   transforms): for the sorted position p, words[p] = tile << (ndim*log2tile)
   | tile-local offsets (log2tile bits per axis, axis d at bit d*log2tile),
   idx[p] = the point's input index (the permutation of dmx_sort_points) and,
   when xarg is not NULL, xarg[p, d] = W-1-2*frac of axis d (float64
   compute type).  Exposed for bit-exact parity tests.  words/idx: uint32
   arrays of M, xarg: (M, ndim) float64, device or host.  Fails where the
   tile key and offsets need more than 32 bits. */
int dmx_sort_records(const dmx_array *coord, const int64_t *nover,
                     size_t ndim, int W, int log2tile,
                     const double *periodicity, dmx_array *words,
                     dmx_array *idx, dmx_array *xarg);

/* Complex FFT along chosen axes: replaces ducc0::c2c(cfmav in, vfmav out,
   shape_t axes, bool forward, T fct, size_t nthreads)
   (src/ducc0/fft/fft.h:848-850), as bound by ducc0.fft.c2c
   (python/fft_pymod.cc:122-172, 164-172).  out = fct * DFT(in) over
   axes[0..naxes) (negative axes count from the end; an axis named twice is
   transformed twice), with exponent sign -1 if forward.  in: complex64 /
   complex128, or float32 / float64 (read as complex with zero imaginary
   part); out: complex of the same precision and shape, any strides, host or
   device; out may be in itself (identical strides).  Any length: 2,3,5,7,
   11-smooth lengths run as Stockham passes, others by Bluestein's algorithm.
   nthreads is accepted and ignored.  opt: stream and DMX_ASYNC as for the
   transforms (NULL for defaults). */
int dmx_c2c(const dmx_array *in, const dmx_array *out, const int64_t *axes,
            size_t naxes, int forward, double fct, size_t nthreads,
            const dmx_options *opt);

/* ---- learned parameter chooser (no ducc0 counterpart; its fallback and
   the reference for its labels is ducc0's findNufftParameters,
   nufft.h:155-208) ----
   Network: 6 features -> 64 -> 64 -> 64 (ReLU) -> 3 outputs (oversampling
   factor, support W, log2 tile), float32, MLP_NPARAM = 8963 parameters
   (W1 64x6, b1, W2 64x64, b2, W3 64x64, b3, W4 3x64, b4).  All arrays host
   memory; the work runs on the GPU. */
size_t dmx_mlp_nparam(void);
/* Y[B][3] = net(X[B][6]) */
int dmx_mlp_forward(const float *params, const float *X, int B, float *Y);
/* loss = sum_b sum_k w_k (y_bk - T_bk)^2 / (3B) and its gradient */
int dmx_mlp_grad(const float *params, const float *X, const float *T,
                 const float *w, int B, float *grad, float *loss);
/* `steps` full-batch Adam steps in place on params; loss per step into
   `losses` (may be NULL) */
int dmx_mlp_train(float *params, const float *X, const float *T,
                  const float *w, int B, int steps, float lr,
                  float weight_decay, float *losses);
/* install trained weights and the normalisation (feature mean[6], std[6],
   target mean[3], std[3]) used by chooser = learned */
int dmx_chooser_set(const float *params, const double *norm);
/* the 6 features of a request */
int dmx_chooser_features(size_t ndim, const int64_t *shape, size_t npoints,
                         double epsilon, int fp64, int nu2u, double *f);
/* the parameters a call would use, without running it (info->chooser_used
   tells whether the network's choice was taken) */
int dmx_choose(size_t ndim, const int64_t *shape, size_t npoints,
               double epsilon, int values_f64, int nu2u, double sigma_min,
               double sigma_max, const dmx_options *opt,
               dmx_plan_info *info);

/* Phase timing (role of ducc0's TimerHierarchy / verbosity output,
   src/ducc0/infra/timers.h:69-262, nufft.h:469-471): when enabled, every
   transform records HIP events around its phases ("building index",
   "zeroing grid", "spreading", "FFT axis k", "grid correction",
   "interpolation") on its stream; totals accumulate per process.
   dmx_profile_read writes lines "name\ttotal_ms\tcount\n". */
int dmx_profile_enable(int on);
int dmx_profile_reset(void);
int dmx_profile_read(char *buf, size_t buflen);

/* Allocation-free type 1 / type 2 transform (no counterpart in the
   reference; for callers that cannot allocate: code under HIP graph
   capture, tight serving loops, harnesses that forbid allocation).  A
   variant of dmx_nu2u (type 1: points -> uniform) and dmx_u2nu (type 2:
   uniform -> points) with the same argument meaning and validation, except:
   every array must be in device memory (coord and points C-contiguous);
   the analytic parameter choice only (opt->chooser must be 0); it runs on
   opt->stream (NULL: the null stream) and waits for its work unless
   DMX_ASYNC is set.  All temporaries come from one static device array of
   DMX_NOALLOC_SCRATCH_BYTES per GPU: a request needing more is rejected
   with an error before any work is queued.  Calls that overlap on one GPU
   share that array and race. */
#define DMX_NOALLOC_SCRATCH_BYTES (64u << 20)
int dmx_nufft_noalloc(int type, const dmx_array *uniform,
                      const dmx_array *coord, const dmx_array *points,
                      int forward, double epsilon, double sigma_min,
                      double sigma_max, const double *periodicity,
                      int fft_order, const dmx_options *opt,
                      dmx_plan_info *info);

/* ---- spherical harmonic synthesis at arbitrary points ----
   ducc0.sht.synthesis_general (src/ducc0/sht/sht.cc:2926-2960, bound at
   python/sht_pymod.cc:805-845, 2299), at the points theta = loc[i, 0] in
   [0, pi], phi = loc[i, 1]:
     spin 0:  map[0, i] = sum_{l, m} a_lm Y_lm  for the real field whose
              a_lm (m >= 0) are given;
     spin s > 0: map[0, i] + i map[1, i] = -sum_{l, |m| <= l} (E_lm + i B_lm)
              sY_lm, E_{l,-m} = (-1)^m conj(E_lm) (B likewise), with
              alm[0] = E and alm[1] = B (mode 0, STANDARD), alm[0] = E and
              B = 0 (mode 1, GRAD_ONLY), or alm[0] = a_lm of a spin-0 field
              with E_lm = sqrt(l (l + 1)) a_lm, B = 0 and s = 1 (mode 2,
              DERIV1: the map is its gradient (d/dtheta, d/dphi / sin
              theta)).
   alm: (ncomp, nalm) complex64 / complex128 (ncomp 1 for spin 0 or modes 1
   and 2, else 2), a_lm at index mstart[m] + l * lstride (m = 0 .. nm-1 =
   mmax, l = m .. lmax; entries with l < spin are ignored); map: (1, M) for
   spin 0, else (2, M), float32 / float64 matching the a_lm precision (any
   strides); loc: (M, 2) float64.  The points are interpolated from the
   doubled sphere's Fourier coefficients with the NUFFT to accuracy epsilon
   (sigma_min / sigma_max as for the NUFFT; ducc0's defaults are 1.1 /
   2.6).  Modes 1 and 2 need spin > 0, mode 2 spin 1 (ducc0 computes mode 2
   as spin 1 whatever spin it is given).  nthreads is accepted and ignored;
   verbosity > 0 prints the NUFFT parameters. */
int dmx_sht_synthesis_general(const dmx_array *alm, dmx_array *map,
                              size_t spin, size_t lmax,
                              const int64_t *mstart, size_t nm,
                              int64_t lstride, const dmx_array *loc,
                              double epsilon, double sigma_min,
                              double sigma_max, size_t nthreads, int mode,
                              size_t verbosity, const dmx_options *opt);
/* ducc0.sht.adjoint_synthesis_general (src/ducc0/sht/sht.cc:2962-2993,
   python/sht_pymod.cc:847-870, 2300): the adjoint of the above with respect
   to the real inner products of the map and of the a_lm (m > 0 entries
   counted twice), written into alm at the (l, m) positions of the layout
   (zero for l < spin; the m = 0 entries real); for host arrays the other
   entries come back zero, device arrays keep them.  Same arguments. */
int dmx_sht_adjoint_synthesis_general(dmx_array *alm, const dmx_array *map,
                                      size_t spin, size_t lmax,
                                      const int64_t *mstart, size_t nm,
                                      int64_t lstride, const dmx_array *loc,
                                      double epsilon, double sigma_min,
                                      double sigma_max, size_t nthreads,
                                      int mode, size_t verbosity,
                                      const dmx_options *opt);

/* ducc0.sht.pseudo_analysis_general (src/ducc0/sht/sht.cc:3003-3058,
   python/sht_pymod.cc:885-927, 2301): a_lm whose synthesis_general best
   fits the map values at loc in the least-squares sense, by LSMR from zero
   (atol 1e-14 sqrt(npoints), btol epsilon, conlim 1e8, at most maxiter
   iterations), the synthesis and its adjoint in mode STANDARD at accuracy
   0.1 epsilon; the a_lm norm counts m > 0 entries twice.  alm (output):
   (ncomp, nalm) complex as for the synthesis, every entry written (zero
   outside the layout); map: (nmaps, npoints) real of the same precision.
   istop (1: solution found, 2: least-squares solution found, 3: condition
   too large, 7: maxiter reached, as in ducc0), itn, |r| / |b| and
   |A^T r| / (|A| |r|) are returned through the pointers (each may be
   NULL).  The points are sorted once for all iterations. */
int dmx_sht_pseudo_analysis_general(dmx_array *alm, const dmx_array *map,
                                    size_t spin, size_t lmax,
                                    const int64_t *mstart, size_t nm,
                                    int64_t lstride, const dmx_array *loc,
                                    double sigma_min, double sigma_max,
                                    size_t nthreads, size_t maxiter,
                                    double epsilon, size_t verbosity,
                                    size_t *istop, size_t *itn,
                                    double *rnorm, double *arnorm,
                                    const dmx_options *opt);

/* ---- radio-interferometric gridding with the w-term ----
   ducc0.wgridder.vis2dirty (src/ducc0/wgridder/wgridder.h:1626-1714, bound
   at python/wgridder_pymod.cc:561-640; ms2dirty :483-493 is the same call
   with fixed options): for pixel (i, j) at x = cx + (i - nx/2) pixsize_x,
   y = cy + (j - ny/2) pixsize_y, n = sqrt(1 - x^2 - y^2),
     dirty(i, j) = sum_{row, chan} Re(vis wgt e^{2 pi i (x u + y v - w (n-1))}) / n
   with (u, v, w) = uvw[row] * freq[chan] / c (each sign flipped by flip_u /
   v / w, cx / cy flipped with u / v); without do_wgridding the w term and
   the division are absent; divide_by_n = 0 skips the division.  Masked
   (mask == 0) or zero-weight visibilities are skipped.
   uvw: (nrow, 3) float64; freq: (nchan,) float64, positive, ascending;
   vis: (nrow, nchan) complex64 / complex128 (the precision of the call);
   wgt: (nrow, nchan) real of that precision or NULL; mask: (nrow, nchan)
   uint8 or NULL; dirty: (npix_x, npix_y) real of that precision, both even.
   allow_nshift = 1 lets the w planes be centred on the image's n - 1
   range (wgridder.h:1520), as ducc0 does by default.  Accuracy epsilon;
   sigma_min / sigma_max bound the oversampling (ducc0's defaults 1.1 /
   2.6).  nthreads and verbosity are accepted and ignored. */
int dmx_vis2dirty(const dmx_array *uvw, const dmx_array *freq,
                  const dmx_array *vis, const dmx_array *wgt,
                  const dmx_array *mask, double pixsize_x, double pixsize_y,
                  double epsilon, int do_wgridding, int flip_u, int flip_v,
                  int flip_w, int divide_by_n, double sigma_min,
                  double sigma_max, double center_x, double center_y,
                  int allow_nshift, size_t nthreads, size_t verbosity,
                  dmx_array *dirty, const dmx_options *opt);
/* ducc0.wgridder.dirty2vis (wgridder.h:1716-1735, wgridder_pymod.cc:
   594-620; dirty2ms :544-550): the adjoint,
     vis(row, chan) = wgt sum_{i, j} dirty(i, j) / n e^{-2 pi i (x u + y v - w (n-1))},
   zero for masked or zero-weight visibilities.  Same arguments; `vis` is
   the output. */
int dmx_dirty2vis(const dmx_array *uvw, const dmx_array *freq,
                  const dmx_array *dirty, const dmx_array *wgt,
                  const dmx_array *mask, double pixsize_x, double pixsize_y,
                  double epsilon, int do_wgridding, int flip_u, int flip_v,
                  int flip_w, int divide_by_n, double sigma_min,
                  double sigma_max, double center_x, double center_y,
                  int allow_nshift, size_t nthreads, size_t verbosity,
                  dmx_array *vis, const dmx_options *opt);

/* Number of HIP devices visible (0 when none). */
int dmx_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
