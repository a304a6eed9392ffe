// dmx.h -- internal declarations shared by the host runtime and the gfx950
// kernels of the NUFFT.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmx {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
#define DMX_ASSERT(cond, msg) \
  do {                        \
    if (!(cond)) throw ::dmx::Error(msg); \
  } while (0)
#define DMX_HIP(call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess)                                               \
      throw ::dmx::Error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                         " at " #call);                                 \
  } while (0)

// ------------------------------------------------------------- geometry
// Fold / tile geometry of one transform, passed by value to kernels.
// nover: oversampled grid; shift = nover - W/2 + 1; maxi0 = nover+nsafe-W
// (nufft.h:415-421); ntiles = (nover >> log2tile) + 3 (nufft.h:1366-1367).
struct Geo {
  int ndim, W, nsafe, log2tile;
  int nover[3], ntiles[3], maxi0[3];
  double coordfct[3], shift[3];
  int coord_f32;  // 1: coordinates stored as float32
  int coord_v2;   // 1: 2-D float64 coordinates 16-byte aligned (one load)
};

// One subproblem: a run of at most MAXSUB sorted points of one tile; rank:
// its place among the tile's subproblems (0 for the first).
struct Sub {
  uint32_t tile, start, end, rank;
};

constexpr int MAXSUB = 1024;

inline size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }
// LDS one workgroup may allocate on gfx950 (160 KiB per CU, all of it
// available to a single workgroup once the kernel opts in)
constexpr size_t LDS_MAX_BYTES = 163840;

#ifdef __HIPCC__
// lets `kernel` take more than the default 64 KiB of dynamic LDS
template <typename K>
inline void allow_big_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return;
  DMX_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)LDS_MAX_BYTES));
}
#endif
// points per LDS staging round of the spreader: a whole subproblem in 1-D
// and 2-D (FP32), half of one in FP64, 256 in 3-D where the region itself
// needs most of the LDS
__host__ __device__ constexpr int spread_chunk(int ndim, bool fp64) {
  return ndim == 3 ? 256 : (fp64 ? MAXSUB / 2 : MAXSUB);
}
// LDS of one spreading workgroup: re/im planes of FP64 accumulators over the
// region of `cells` cells plus the staging of one chunk of points (values,
// kernel arguments, offsets, class counts)
inline constexpr size_t spread_lds_bytes(size_t cells, int ndim, bool fp64) {
  return 2 * cells * sizeof(double) +
         (((size_t)spread_chunk(ndim, fp64) *
               (2 * (fp64 ? 8 : 4) + ndim * (fp64 ? 8 : 4) + 4) +
           32 * 4 + 15) &
          ~(size_t)15);
}

// ----------------------------------------------------------- device helpers
#ifdef __HIPCC__

// Fold one coordinate into (first grid index i0, local kernel coordinate).
// For float64 coordinates the product (t - floor t) * nover is formed exactly
// as a double-double, so i0 agrees with an 80-bit range reduction
// (nufft.h:285-298) except when that product lies within 2^-64 (relative) of
// an integer.  For float32 coordinates the reference reduces in double, which
// is reproduced operation for operation (no contraction).
__device__ __forceinline__ int fold_axis(const Geo &g, int d, double x,
                                         double &frac) {
#pragma clang fp contract(off)
  const double n = (double)g.nover[d];
  double tmp = x * g.coordfct[d];
  double a = tmp - floor(tmp);
  int i;
  if (!g.coord_f32) {
    double hi = a * n;
    double lo = fma(a, n, -hi);
    double s = hi + g.shift[d];
    double bv = s - hi;
    double err = (hi - (s - bv)) + (g.shift[d] - bv);
    double fs = floor(s);
    double r = (s - fs) + (err + lo);
    i = (int)fs;
    if (r < 0.) i -= 1;
    else if (r >= 1.) i += 1;
    i -= g.nover[d];
    if (i > g.maxi0[d]) i = g.maxi0[d];
    frac = (hi - (double)i) + lo;
  } else {
    double t2 = a * n;
    i = (int)(t2 + g.shift[d]) - g.nover[d];
    if (i > g.maxi0[d]) i = g.maxi0[d];
    frac = t2 - (double)i;
  }
  // finite coordinates always give 1-nsafe <= i <= maxi0; the clamp keeps
  // non-finite ones (whose results are meaningless) inside the tile range
  if (!(i >= 1 - g.nsafe)) i = 1 - g.nsafe;
  return i;
}

__device__ __forceinline__ double load_coord(const void *c, size_t j, int ndim,
                                             int d, int f32) {
  return f32 ? (double)((const float *)c)[j * ndim + d]
             : ((const double *)c)[j * ndim + d];
}

// all coordinates of point j (one 16-byte load for aligned 2-D float64)
__device__ __forceinline__ void load_point(const Geo &g, const void *c,
                                           size_t j, int ndim, double *x) {
  if (ndim == 2 && g.coord_v2) {
    const double2 v = ((const double2 *)c)[j];
    x[0] = v.x;
    x[1] = v.y;
  } else {
    for (int d = 0; d < ndim; ++d) x[d] = load_coord(c, j, ndim, d, g.coord_f32);
  }
}

// Values of the W kernel pieces at local coordinate x (x = W-1-2*frac),
// i.e. the weights of grid points i0..i0+W-1.  Horner in x^2 on the odd and
// even coefficient halves; piece W-1-i is piece i mirrored
// (gridding_kernel.h:300-318).  cf layout: for s = 0..NE-1,
// cf[(2s)*NH+i] = coefficient of x^(DP-2s), cf[(2s+1)*NH+i] = of x^(DP-1-2s).
template <typename T, int W>
__device__ __forceinline__ void kernel_weights(T x, const T *__restrict__ cf,
                                               T *w) {
  constexpr int DP = W + 3 + (W & 1);
  constexpr int NE = (DP + 1) / 2;
  constexpr int NH = (W + 1) / 2;
  const T x2 = x * x;
#pragma unroll
  for (int i = 0; i < NH; ++i) {
    T o = cf[i], e = cf[NH + i];
#pragma unroll
    for (int s = 1; s < NE; ++s) {
      o = o * x2 + cf[(2 * s) * NH + i];
      e = e * x2 + cf[(2 * s + 1) * NH + i];
    }
    w[i] = e + x * o;
    if (W - 1 - i != i) w[W - 1 - i] = e - x * o;
  }
}

// Weights and their derivatives d/dx of the W pieces at x.
template <typename T, int W>
__device__ __forceinline__ void kernel_weights_deriv(T x,
                                                     const T *__restrict__ cf,
                                                     T *w, T *dw) {
  constexpr int DP = W + 3 + (W & 1);
  constexpr int NE = (DP + 1) / 2;
  constexpr int NH = (W + 1) / 2;
  const T x2 = x * x;
#pragma unroll
  for (int i = 0; i < NH; ++i) {
    // p(x) = E(x^2) + x O(x^2); p'(x) = 2x E'(x^2) + O(x^2) + 2x^2 O'(x^2)
    T o = cf[i], e = cf[NH + i], od = T(0), ed = T(0);
#pragma unroll
    for (int s = 1; s < NE; ++s) {
      od = od * x2 + o;
      ed = ed * x2 + e;
      o = o * x2 + cf[(2 * s) * NH + i];
      e = e * x2 + cf[(2 * s + 1) * NH + i];
    }
    T dE = T(2) * x * ed;           // d/dx E(x^2)
    T dO = o + T(2) * x2 * od;      // d/dx (x O(x^2))
    w[i] = e + x * o;
    dw[i] = dE + dO;
    if (W - 1 - i != i) {
      w[W - 1 - i] = e - x * o;
      dw[W - 1 - i] = dE - dO;  // d/dx p(-x) = -p'(-x); E part even, O part
    }
  }
}

// weight of kernel piece q alone at local coordinate x (the same even / odd
// Horner in x^2 as kernel_weights, for a run-time support W)
template <typename T>
__device__ __forceinline__ T piece_weight(T x, const T *__restrict__ cf, int W,
                                          int q) {
  const int DP = W + 3 + (W & 1), NE = (DP + 1) / 2, NH = (W + 1) / 2;
  const int i = q < NH ? q : W - 1 - q;
  const T x2 = x * x;
  T o = cf[i], e = cf[NH + i];
  for (int s = 1; s < NE; ++s) {
    o = o * x2 + cf[(2 * s) * NH + i];
    e = e * x2 + cf[(2 * s + 1) * NH + i];
  }
  return q < NH ? e + x * o : e - x * o;
}

__device__ __forceinline__ int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}
#endif  // __HIPCC__

// --------------------------------------------------------- host plumbing
struct FftAxisPlan;  // defined in fft.hip

// Launch wrappers (defined in the .hip files).
// tile key of every point; hist0 (optional): the first sort pass's
// histogram, as radix_sort_hist0 locates it in the sort workspace
void launch_fold_keys(const Geo &g, const void *coords, size_t M,
                      uint32_t *keys, uint32_t *hist0, hipStream_t s);
// stable LSD sort of keys < 2^keybits; hist0_ready: the first pass's
// histogram was already written by launch_fold_keys; keys_out may be NULL.
// tb_start_c / tb_end (zeroed, one entry per key value, optional): the
// final pass records each key's [start, end) in the sorted order, start
// complemented (~start)
void radix_sort_pairs(const uint32_t *keys_in, size_t M, int keybits,
                      uint32_t *idx_out, uint32_t *keys_out, void *work,
                      size_t work_bytes, hipStream_t s,
                      bool hist0_ready = false, uint32_t *tb_start_c = nullptr,
                      uint32_t *tb_end = nullptr);
size_t radix_sort_workspace(size_t M);
// offsets of one radix pass: the per-block histogram of the digit
// (keys >> shift) & 255 (unless hist_ready), its digit-major per-chunk scans
// (offs) and the scan of the chunk sums (bsum); a block's global start of
// digit d is offs[d*nblocks+b] + bsum[(d*nblocks+b) / 2048]
void radix_pass_offsets(const uint32_t *keys, size_t M, int shift,
                        uint32_t *hist, bool hist_ready, uint32_t *offs,
                        uint32_t *bsum, hipStream_t s);
uint32_t *radix_sort_hist0(void *work, size_t M);
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n,
                        uint32_t *total, void *work, hipStream_t s);
size_t scan_workspace(size_t n);
// from the tile ranges recorded by the sort (start complemented, end 0 for
// empty tiles): per-tile start (in place) and point counts, and the
// subproblem list (runs of at most MAXSUB points of one tile)
void launch_tile_subs(size_t ntiles, uint32_t *counts, uint32_t *tstart,
                      const uint32_t *tend, uint32_t *nsub,
                      uint32_t *sub_start, Sub *subs, uint32_t *nsub_total,
                      void *scanwork, hipStream_t s);
void launch_gather_coords(const void *coords, int coord_f32, int ndim,
                          const uint32_t *idx, size_t M, void *out,
                          hipStream_t s);
void launch_zero(void *p, size_t bytes, hipStream_t s);

// point records in tile order (records.hip): structure of arrays
struct RecBufs {
  uint32_t *word;  // tile << offbits | tile-local offsets (log2tile bits
                   // per axis, axis d at bit d*log2tile)
  void *x[3];      // kernel argument W-1-2*frac per axis, compute type
  void *v;         // type 1: value, complex compute type
  uint32_t *j;     // type 2: input index of the point
};
size_t rec_bytes_per_point(int ndim, bool fp64, bool type1);
size_t rec_carve_bytes(size_t M, int ndim, bool fp64, bool type1);
RecBufs rec_carve(void *p, size_t M, int ndim, bool fp64, bool type1);
size_t rec_sort_workspace(size_t M);
// stable sort of the points by tile, carrying their records: `out` receives
// them in tile order (the order of radix_sort_pairs' permutation), `tmp` is
// the ping-pong buffer; tb_start_c / tb_end as for radix_sort_pairs
void rec_sort(const Geo &g, int offbits, int keybits, bool fp64, bool type1,
              const void *coords, const void *values, int val_f64, size_t M,
              const RecBufs &out, const RecBufs &tmp, void *work,
              size_t work_bytes, uint32_t *tb_start_c, uint32_t *tb_end,
              hipStream_t s);

// spreading / interpolation (spread.hip, interp.hip)
struct SpreadArgs {
  Geo g;
  const Sub *subs;
  const uint32_t *nsub_total;
  const uint32_t *idx;      // sorted order; NULL: identity (coords sorted)
  const void *coords;       // (M, ndim) float32/float64
  const void *values;       // (M,) c64 or c128
  int val_f64;
  int values_sorted;        // 1: values are stored in sorted order
  void *grid;               // oversampled grid, complex of compute type
  const void *coef;         // kernel coefficients of compute type
  size_t nsub_max;
  int lds_side;             // tile + W - 1
  int use_lds;
  int coords_sorted;        // 1: coords are stored in sorted order
  // deterministic mode: per-subproblem FP64 regions, and each tile's first
  // subproblem and subproblem count (index build tables)
  void *slab;
  const uint32_t *tile_sub0, *tile_nsub;
  // 1: the points are records in tile order (rec_sort): spreading reads
  // rec.word / rec.x / rec.v, interpolation rec.word / rec.x / rec.j, at
  // the sorted position, instead of idx / coords / values
  int rec;
  RecBufs recs;
  // non-NULL: incremented for every point whose fold lands outside its
  // tile (the offset is clamped either way)
  uint32_t *offcheck;
  // 1: FP32 spreading accumulates the region in packed fixed point
  // (spread_impl.h, SpreadStash<..., FIX>)
  int fixacc;
  // 1: the finished region of a tile, complex of the compute type, goes to
  // a.slab + tile * cells (plain stores; one workgroup takes all of the
  // tile's subproblems and adds the later ones to the stored region)
  // instead of into the grid with atomics; launch_fold_regions then writes
  // every grid cell
  int slab_out;
};
void launch_spread(bool fp64, const SpreadArgs &a, hipStream_t s);
void launch_interp(bool fp64, const SpreadArgs &a, void *out, int out_f64,
                   hipStream_t s);
void launch_interp_grad(bool fp64, const SpreadArgs &a, void *out,
                        int out_f64, hipStream_t s);
// deterministic type-1 spreading (spread_det.hip): writes every cell of
// a.grid (no zeroing needed); a.slab holds spread_det_slab_bytes(a)
void launch_spread_det(bool fp64, const SpreadArgs &a, hipStream_t s);
size_t spread_det_slab_bytes(const SpreadArgs &a);
// after launch_spread with a.slab_out: every grid cell = the sum of the
// region cells covering it (spread_det.hip); writes all of a.grid
void launch_fold_regions(bool fp64, const SpreadArgs &a, hipStream_t s);
// bytes of a.slab for launch_spread with a.slab_out (one region per tile)
size_t spread_region_slab_bytes(const SpreadArgs &a, bool fp64);

}  // namespace dmx
