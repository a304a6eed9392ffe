// spread_det.hip -- type-1 spreading in deterministic mode: results
// identical bit for bit from run to run, whatever the scheduling.
//
// The default spreader (spread_impl.h) accumulates with LDS and global
// float atomics, whose order varies between runs, so the last bits of a
// type-1 result vary (as ducc0's do with nthreads > 1: its tile buffers are
// added under a lock in thread order, nufft.h:937-953).  Here no floating-
// point atomic is used:
//  (1) k_spread_det: one wave per subproblem adds the W^d taps of each of
//      its points, point after point in sorted order, into an FP64 region
//      in LDS (each lane owns distinct taps of the point, so no two lanes
//      write one cell at once), and writes the region to that subproblem's
//      slab in HBM;
//  (2) k_fold_slabs: every cell of the oversampled grid is the sum of the
//      slab cells that cover it, taken in a fixed order (periodic images,
//      then tiles, then the tile's subproblems in sorted order), rounded
//      once to the compute type.
// The sorted order itself is the stable tile order of the index build, a
// function of the coordinates only.  Slabs take nsub_max * (tile + W - 1)^d
// complex FP64 cells.
// The fold also serves the default spreader when its subproblems store
// their regions (complex of the compute type) instead of adding them to the
// grid with atomics (launch_fold_regions).
#include <type_traits>

#include "dmx.h"

namespace dmx {

namespace {

template <typename T, int NDIM>
__global__ __launch_bounds__(64) void k_spread_det(SpreadArgs a) {
  extern __shared__ unsigned char smem_raw[];
  const uint32_t total = *a.nsub_total;
  const int S = a.lds_side, W = a.g.W, tile = 1 << a.g.log2tile;
  int ncell = S, ntap = W;
  if (NDIM > 1) { ncell *= S; ntap *= W; }
  if (NDIM > 2) { ncell *= S; ntap *= W; }
  double *sre = (double *)smem_raw;
  double *sim = sre + ncell;
  T *wts = (T *)(sim + ncell);  // NDIM * W weights of the current point
  const T *cf = (const T *)a.coef;
  const int lane = threadIdx.x;
  double2 *slab = (double2 *)a.slab;
  for (uint32_t sidx = blockIdx.x; sidx < total; sidx += gridDim.x) {
    const Sub sb = a.subs[sidx];
    int b0[3];
    {
      uint32_t t[3] = {0, 0, 0};
      uint32_t tl = sb.tile;
      for (int d = NDIM - 1; d >= 0; --d) {
        t[d] = tl % (uint32_t)a.g.ntiles[d];
        tl /= (uint32_t)a.g.ntiles[d];
      }
      for (int d = 0; d < 3; ++d)
        b0[d] = ((int)t[d] << a.g.log2tile) - a.g.nsafe;
    }
    for (int c = lane; c < ncell; c += 64) {
      sre[c] = 0.0;
      sim[c] = 0.0;
    }
    for (uint32_t p = sb.start; p < sb.end; ++p) {
      // every lane reads the same point (broadcast loads)
      const uint32_t j = (a.coords_sorted && a.values_sorted) ? p : a.idx[p];
      double xc[NDIM];
      load_point(a.g, a.coords, a.coords_sorted ? (size_t)p : (size_t)j, NDIM,
                 xc);
      int o[NDIM];
      T xk[NDIM];
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        double frac;
        const int i0 = fold_axis(a.g, d, xc[d], frac);
        xk[d] = (T)((double)(W - 1) - 2.0 * frac);
        const int od = i0 - b0[d];
        if (od < 0 || od >= tile) {
          if (a.offcheck && lane == 0) atomicAdd(a.offcheck, 1u);
          o[d] = od < 0 ? 0 : tile - 1;
        } else {
          o[d] = od;
        }
      }
      T vr, vi;
      {
        const size_t jv = a.values_sorted ? (size_t)p : (size_t)j;
        if (a.val_f64) {
          const double2 v = ((const double2 *)a.values)[jv];
          vr = (T)v.x;
          vi = (T)v.y;
        } else {
          const float2 v = ((const float2 *)a.values)[jv];
          vr = (T)v.x;
          vi = (T)v.y;
        }
      }
      // the previous point's taps are done before its weights are replaced
      __syncthreads();
      if (lane < NDIM * W) {
        const int d = lane / W;
        T xd = xk[0];
#pragma unroll
        for (int e = 1; e < NDIM; ++e)
          if (e == d) xd = xk[e];
        wts[lane] = piece_weight<T>(xd, cf, W, lane - d * W);
      }
      __syncthreads();
      for (int tap = lane; tap < ntap; tap += 64) {
        int q[3] = {0, 0, 0}, r = tap;
        for (int d = NDIM - 1; d >= 0; --d) {
          q[d] = r % W;
          r /= W;
        }
        int cell = o[0] + q[0];
        T f = vr * wts[q[0]], g = vi * wts[q[0]];
        if (NDIM > 1) {
          cell = cell * S + o[1] + q[1];
          f = f * wts[W + q[1]];
          g = g * wts[W + q[1]];
        }
        if (NDIM > 2) {
          cell = cell * S + o[2] + q[2];
          f = f * wts[2 * W + q[2]];
          g = g * wts[2 * W + q[2]];
        }
        sre[cell] += (double)f;
        sim[cell] += (double)g;
      }
    }
    __syncthreads();
    double2 *dst = slab + (size_t)sidx * ncell;
    for (int c = lane; c < ncell; c += 64)
      dst[c] = make_double2(sre[c], sim[c]);
    __syncthreads();
  }
}

// floor(a / 2^L)
__device__ __forceinline__ int floorshift(int a, int L) {
  return a >= 0 ? a >> L : -((-a + (1 << L) - 1) >> L);
}

// Along one axis, the (tile t, region offset o) pairs covering grid index
// g: for each periodic image u = g + k*n (k = -1, 0, 1, ...) the tiles
// whose region [t*tile - nsafe, t*tile - nsafe + S) holds u, in increasing
// (k, t).  A small-grid region may wrap several times: no pair is skipped.
struct AxisPairs {
  int g, n, L, nsafe, S, ntiles, umax;
  int k, u, t, thi;
  __device__ bool first(int g_, int n_, int L_, int nsafe_, int S_,
                        int ntiles_) {
    g = g_; n = n_; L = L_; nsafe = nsafe_; S = S_; ntiles = ntiles_;
    umax = (ntiles << L) + S;
    k = -2;
    t = 0;
    thi = -1;
    return next();
  }
  // advances to the next pair; false when there is none
  __device__ bool next() {
    ++t;
    while (t > thi) {
      ++k;
      u = g + k * n;
      if (u >= umax) return false;
      if (u < -nsafe) continue;
      // ceil((u + nsafe - S + 1) / tile) .. floor((u + nsafe) / tile)
      t = floorshift(u + nsafe - S + (1 << L), L);
      if (t < 0) t = 0;
      thi = floorshift(u + nsafe, L);
      if (thi > ntiles - 1) thi = ntiles - 1;
    }
    return true;
  }
  __device__ int off() const { return u + nsafe - (t << L); }
};

// grid cell -> sum of the slab cells covering it, in the fixed order of
// AxisPairs on axis 0, then 1, then 2, and the tile's subproblems in sorted
// order, accumulated in FP64 and rounded once to the compute type T.  ST:
// the slab element (double2 for the deterministic slabs, the compute
// type's complex for regions of the default spreader).  The last axis runs
// over the threads of a block, the others over blockIdx.y (and z).
template <typename T, typename ST, int NDIM>
__global__ __launch_bounds__(256) void k_fold_slabs(SpreadArgs a) {
  const int S = a.lds_side, L = a.g.log2tile, nsafe = a.g.nsafe;
  int ncell = S;
  if (NDIM > 1) ncell *= S;
  if (NDIM > 2) ncell *= S;
  const ST *slab = (const ST *)a.slab;
  T *grid = (T *)a.grid;
  const int n0 = a.g.nover[0], n1 = NDIM > 1 ? a.g.nover[1] : 1,
            n2 = NDIM > 2 ? a.g.nover[2] : 1;
  const int nl = NDIM == 1 ? n0 : (NDIM == 2 ? n1 : n2);
  const int ny = NDIM == 1 ? 1 : (NDIM == 2 ? n0 : n1);
  const int nz = NDIM == 3 ? n0 : 1;
  for (int gz = blockIdx.z; gz < nz; gz += gridDim.z)
    for (int gy = blockIdx.y; gy < ny; gy += gridDim.y)
      for (int gx = blockIdx.x * blockDim.x + threadIdx.x; gx < nl;
           gx += gridDim.x * blockDim.x) {
        int g[3];
        if (NDIM == 1) {
          g[0] = gx;
        } else if (NDIM == 2) {
          g[0] = gy;
          g[1] = gx;
        } else {
          g[0] = gz;
          g[1] = gy;
          g[2] = gx;
        }
        double re = 0.0, im = 0.0;
        AxisPairs A0, A1, A2;
        for (bool h0 = A0.first(g[0], n0, L, nsafe, S, a.g.ntiles[0]); h0;
             h0 = A0.next()) {
          bool h1 = true;
          if (NDIM > 1) h1 = A1.first(g[1], n1, L, nsafe, S, a.g.ntiles[1]);
          for (; h1; h1 = NDIM > 1 ? A1.next() : false) {
            bool h2 = true;
            if (NDIM > 2)
              h2 = A2.first(g[2], n2, L, nsafe, S, a.g.ntiles[2]);
            for (; h2; h2 = NDIM > 2 ? A2.next() : false) {
              uint32_t tl = (uint32_t)A0.t;
              int cell = A0.off();
              if (NDIM > 1) {
                tl = tl * (uint32_t)a.g.ntiles[1] + (uint32_t)A1.t;
                cell = cell * S + A1.off();
              }
              if (NDIM > 2) {
                tl = tl * (uint32_t)a.g.ntiles[2] + (uint32_t)A2.t;
                cell = cell * S + A2.off();
              }
              const uint32_t s0 = a.tile_sub0[tl], ns = a.tile_nsub[tl];
              for (uint32_t q = s0; q < s0 + ns; ++q) {
                const ST v = slab[(size_t)q * ncell + cell];
                re += (double)v.x;
                im += (double)v.y;
              }
            }
          }
        }
        size_t gi = (size_t)g[0];
        if (NDIM > 1) gi = gi * n1 + g[1];
        if (NDIM > 2) gi = gi * n2 + g[2];
        grid[2 * gi] = (T)re;
        grid[2 * gi + 1] = (T)im;
      }
}

// The (tile, region offset) pairs of grid index g along one axis, without
// loops: the periodic images g - n, g, g + n and at most two tiles each
// (the caller has checked that images g +- 2n cover nothing and that a
// region spans at most two tiles, see fold_regions).  Returns the number of
// pairs, 3 standing for "more than two" (cells at the periodic seam of
// small grids), whose pairs are not stored.
__device__ __forceinline__ int axis_pairs2(int g, int n, int L, int nsafe,
                                           int S, int ntiles, int (&t)[2],
                                           int (&o)[2]) {
  const int umax = (ntiles << L) + S;
  int c = 0;
  t[0] = t[1] = o[0] = o[1] = 0;
#pragma unroll
  for (int k = -1; k <= 1; ++k) {
    const int u = g + k * n;
    int lo = floorshift(u + nsafe - S + (1 << L), L);
    int hi = floorshift(u + nsafe, L);
    lo = lo < 0 ? 0 : lo;
    hi = hi > ntiles - 1 ? ntiles - 1 : hi;
    if (u < -nsafe || u >= umax) hi = lo - 1;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int tt = lo + j;
      if (tt <= hi) {
        if (c == 0) {
          t[0] = tt;
          o[0] = u + nsafe - (tt << L);
        } else if (c == 1) {
          t[1] = tt;
          o[1] = u + nsafe - (tt << L);
        }
        ++c;
      }
    }
  }
  return c > 2 ? 3 : c;
}

// rows of the grid (along the second-to-last axis) per workgroup of the
// region folds: consecutive rows read consecutive rows of the same regions,
// so a workgroup reuses the cache lines it fetched
constexpr int FOLD_BAND = 16;

// Sum of the region cells covering grid cell g (regions of the default
// spreader, one per tile), by the AxisPairs walk; FP64 sums.  An empty
// tile's (never written) region is ignored.
template <typename C, int NDIM>
__device__ void fold_cell_walk(const SpreadArgs &a, const int (&g)[3],
                               double &re, double &im) {
  const int S = a.lds_side, L = a.g.log2tile, nsafe = a.g.nsafe;
  int ncell = S;
  if (NDIM > 1) ncell *= S;
  if (NDIM > 2) ncell *= S;
  const C *slab = (const C *)a.slab;
  re = 0.0;
  im = 0.0;
  AxisPairs A0, A1, A2;
  for (bool h0 = A0.first(g[0], a.g.nover[0], L, nsafe, S, a.g.ntiles[0]);
       h0; h0 = A0.next()) {
    bool h1 = true;
    if (NDIM > 1)
      h1 = A1.first(g[1], a.g.nover[1], L, nsafe, S, a.g.ntiles[1]);
    for (; h1; h1 = NDIM > 1 ? A1.next() : false) {
      bool h2 = true;
      if (NDIM > 2)
        h2 = A2.first(g[2], a.g.nover[2], L, nsafe, S, a.g.ntiles[2]);
      for (; h2; h2 = NDIM > 2 ? A2.next() : false) {
        uint32_t tl = (uint32_t)A0.t;
        int cell = A0.off();
        if (NDIM > 1) {
          tl = tl * (uint32_t)a.g.ntiles[1] + (uint32_t)A1.t;
          cell = cell * S + A1.off();
        }
        if (NDIM > 2) {
          tl = tl * (uint32_t)a.g.ntiles[2] + (uint32_t)A2.t;
          cell = cell * S + A2.off();
        }
        if (a.tile_nsub[tl]) {
          const C v = slab[(size_t)tl * ncell + cell];
          re += (double)v.x;
          im += (double)v.y;
        }
      }
    }
  }
}

// Regions of the default spreader (one per tile, SpreadArgs::slab_out):
// grid cell -> sum of the region cells covering it, for geometries where
// axis_pairs2 applies (fold_regions checks).  A thread keeps one index gx
// of the last axis for the whole launch (its pairs are found once); the
// pairs of the other axes are found once per row.  A cell's (up to 2^d)
// region cells and tile counts are loaded together; an empty tile's region
// is ignored.  Cells with more than two pairs on an axis (next to the
// periodic seam) take fold_cell_walk.  Region indices fit 32 bits
// (SLAB_MAX_BYTES).  Sums in the compute type.
template <typename T, int NDIM>
__global__ __launch_bounds__(256) void k_fold_regions(SpreadArgs a) {
  using C = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const int S = a.lds_side, L = a.g.log2tile, nsafe = a.g.nsafe;
  uint32_t ncell = (uint32_t)S;
  if (NDIM > 1) ncell *= S;
  if (NDIM > 2) ncell *= S;
  const C *slab = (const C *)a.slab;
  const uint32_t *tnsub = a.tile_nsub;
  C *grid = (C *)a.grid;
  const int n0 = a.g.nover[0], n1 = NDIM > 1 ? a.g.nover[1] : 1,
            n2 = NDIM > 2 ? a.g.nover[2] : 1;
  const int nl = NDIM == 1 ? n0 : (NDIM == 2 ? n1 : n2);
  const int ny = NDIM == 1 ? 1 : (NDIM == 2 ? n0 : n1);
  const int nz = NDIM == 3 ? n0 : 1;
  const uint32_t ntl = (uint32_t)a.g.ntiles[NDIM - 1];
  for (int gx = blockIdx.x * blockDim.x + threadIdx.x; gx < nl;
       gx += gridDim.x * blockDim.x) {
    int tx[2], ox[2];
    const int cx = axis_pairs2(gx, nl, L, nsafe, S, (int)ntl, tx, ox);
    for (int gz = blockIdx.z; gz < nz; gz += gridDim.z) {
      for (int yb = blockIdx.y; yb * FOLD_BAND < ny; yb += gridDim.y) {
        const int gy1 = min((yb + 1) * FOLD_BAND, ny);
        for (int gy = yb * FOLD_BAND; gy < gy1; ++gy) {
          // the outer axes: up to 2 x 2 (tile, offset) pairs of this row
          int g[3] = {gx, 0, 0};
          if (NDIM == 2) {
            g[0] = gy;
            g[1] = gx;
          }
          if (NDIM == 3) {
            g[0] = gz;
            g[1] = gy;
            g[2] = gx;
          }
          int ta[2] = {0, 0}, oa[2] = {0, 0}, tb[2] = {0, 0}, ob[2] = {0, 0};
          int ca = 1, cb = 1;
          if (NDIM > 1)
            ca = axis_pairs2(g[0], n0, L, nsafe, S, a.g.ntiles[0], ta, oa);
          if (NDIM > 2)
            cb = axis_pairs2(g[1], n1, L, nsafe, S, a.g.ntiles[1], tb, ob);
          size_t gi = (size_t)g[0];
          if (NDIM > 1) gi = gi * n1 + g[1];
          if (NDIM > 2) gi = gi * n2 + g[2];
          C o;
          if (ca > 2 || cb > 2 || cx > 2) {
            double re, im;
            fold_cell_walk<C, NDIM>(a, g, re, im);
            o.x = (T)re;
            o.y = (T)im;
            grid[gi] = o;
            continue;
          }
          T re = T(0), im = T(0);
          C v[2][2][2];
          uint32_t ns[2][2][2];
          bool on_[2][2][2];
#pragma unroll
          for (int i0 = 0; i0 < (NDIM > 1 ? 2 : 1); ++i0)
#pragma unroll
            for (int i1 = 0; i1 < (NDIM > 2 ? 2 : 1); ++i1)
#pragma unroll
              for (int i2 = 0; i2 < 2; ++i2) {
                bool on = i2 < cx;
                uint32_t tl = 0, cell = 0;
                if (NDIM > 1) {
                  on = on && i0 < ca;
                  tl = (uint32_t)ta[i0];
                  cell = (uint32_t)oa[i0];
                }
                if (NDIM > 2) {
                  on = on && i1 < cb;
                  tl = tl * (uint32_t)a.g.ntiles[1] + (uint32_t)tb[i1];
                  cell = cell * S + (uint32_t)ob[i1];
                }
                tl = tl * ntl + (uint32_t)tx[i2];
                cell = cell * S + (uint32_t)ox[i2];
                // unconditional loads (an unused pair reads cell 0 of tile
                // 0), so that all of them are in flight at once
                tl = on ? tl : 0u;
                cell = on ? cell : 0u;
                on_[i0][i1][i2] = on;
                ns[i0][i1][i2] = tnsub[tl];
                v[i0][i1][i2] = slab[tl * ncell + cell];
              }
#pragma unroll
          for (int i0 = 0; i0 < (NDIM > 1 ? 2 : 1); ++i0)
#pragma unroll
            for (int i1 = 0; i1 < (NDIM > 2 ? 2 : 1); ++i1)
#pragma unroll
              for (int i2 = 0; i2 < 2; ++i2) {
                const bool use = on_[i0][i1][i2] && ns[i0][i1][i2] != 0;
                re += use ? (T)v[i0][i1][i2].x : T(0);
                im += use ? (T)v[i0][i1][i2].y : T(0);
              }
          o.x = re;
          o.y = im;
          grid[gi] = o;
        }
      }
    }
  }
}

// every cell by fold_cell_walk (geometries where axis_pairs2 does not
// apply: small grids, regions wider than two tiles)
template <typename T, int NDIM>
__global__ __launch_bounds__(256) void k_fold_regions_walk(SpreadArgs a) {
  using C = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  C *grid = (C *)a.grid;
  const int n0 = a.g.nover[0], n1 = NDIM > 1 ? a.g.nover[1] : 1,
            n2 = NDIM > 2 ? a.g.nover[2] : 1;
  const int nl = NDIM == 1 ? n0 : (NDIM == 2 ? n1 : n2);
  const int ny = NDIM == 1 ? 1 : (NDIM == 2 ? n0 : n1);
  const int nz = NDIM == 3 ? n0 : 1;
  for (int gz = blockIdx.z; gz < nz; gz += gridDim.z)
    for (int gy = blockIdx.y; gy < ny; gy += gridDim.y)
      for (int gx = blockIdx.x * blockDim.x + threadIdx.x; gx < nl;
           gx += gridDim.x * blockDim.x) {
        int g[3] = {gx, 0, 0};
        if (NDIM == 2) {
          g[0] = gy;
          g[1] = gx;
        } else if (NDIM == 3) {
          g[0] = gz;
          g[1] = gy;
          g[2] = gx;
        }
        double re, im;
        fold_cell_walk<C, NDIM>(a, g, re, im);
        size_t gi = (size_t)g[0];
        if (NDIM > 1) gi = gi * n1 + g[1];
        if (NDIM > 2) gi = gi * n2 + g[2];
        C o;
        o.x = (T)re;
        o.y = (T)im;
        grid[gi] = o;
      }
}

template <int NDIM>
dim3 fold_grid(const Geo &g) {
  const int n0 = g.nover[0], n1 = NDIM > 1 ? g.nover[1] : 1,
            n2 = NDIM > 2 ? g.nover[2] : 1;
  const int nl = NDIM == 1 ? n0 : (NDIM == 2 ? n1 : n2);
  const int ny = NDIM == 1 ? 1 : (NDIM == 2 ? n0 : n1);
  const int nz = NDIM == 3 ? n0 : 1;
  return dim3((unsigned)std::min((nl + 255) / 256, 1 << 20),
              (unsigned)std::min(ny, 65535), (unsigned)std::min(nz, 65535));
}

template <typename T, typename ST, int NDIM>
void fold(const SpreadArgs &a, hipStream_t s) {
  hipLaunchKernelGGL((k_fold_slabs<T, ST, NDIM>), fold_grid<NDIM>(a.g),
                     dim3(256), 0, s, a);
  DMX_HIP(hipGetLastError());
}

template <int NDIM>
dim3 fold_regions_grid(const Geo &g) {
  dim3 gr = fold_grid<NDIM>(g);
  const int ny = NDIM == 1 ? 1 : (NDIM == 2 ? g.nover[0] : g.nover[1]);
  gr.y = (unsigned)std::min((ny + FOLD_BAND - 1) / FOLD_BAND, 65535);
  return gr;
}

template <typename T, int NDIM>
void fold_regions(const SpreadArgs &a, hipStream_t s) {
  // axis_pairs2 applies when a region spans at most two tiles and the
  // images g +- 2n of every grid index lie outside all regions
  const int S = a.lds_side, L = a.g.log2tile;
  bool fast = S <= 2 << L;
  for (int d = 0; d < NDIM; ++d)
    fast = fast &&
           2 * (int64_t)a.g.nover[d] >= ((int64_t)a.g.ntiles[d] << L) + S &&
           a.g.nover[d] > a.g.nsafe;
  if (fast)
    hipLaunchKernelGGL((k_fold_regions<T, NDIM>), fold_regions_grid<NDIM>(a.g),
                       dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_fold_regions_walk<T, NDIM>), fold_grid<NDIM>(a.g),
                       dim3(256), 0, s, a);
  DMX_HIP(hipGetLastError());
}

template <typename T, int NDIM>
void go(const SpreadArgs &a, hipStream_t s) {
  size_t ncell = a.lds_side;
  for (int d = 1; d < NDIM; ++d) ncell *= a.lds_side;
  const unsigned nb = (unsigned)std::min<size_t>(a.nsub_max, 8192);
  if (nb) {
    const size_t sh = 16 * ncell + NDIM * 16 * sizeof(T);
    allow_big_lds(k_spread_det<T, NDIM>, sh);
    hipLaunchKernelGGL((k_spread_det<T, NDIM>), dim3(nb), dim3(64), sh, s, a);
  }
  fold<T, double2, NDIM>(a, s);
}

}  // namespace

size_t spread_det_slab_bytes(const SpreadArgs &a) {
  size_t ncell = a.lds_side;
  for (int d = 1; d < a.g.ndim; ++d) ncell *= a.lds_side;
  return a.nsub_max * ncell * 16;
}

size_t spread_region_slab_bytes(const SpreadArgs &a, bool fp64) {
  size_t cells = 1, ntiles = 1;
  for (int d = 0; d < a.g.ndim; ++d) {
    cells *= (size_t)a.lds_side;
    ntiles *= (size_t)a.g.ntiles[d];
  }
  return ntiles * cells * (fp64 ? 16 : 8);
}

void launch_fold_regions(bool fp64, const SpreadArgs &a, hipStream_t s) {
  DMX_ASSERT(a.slab && a.tile_nsub, "internal: region fold arguments");
  switch (a.g.ndim) {
    case 1:
      return fp64 ? fold_regions<double, 1>(a, s)
                  : fold_regions<float, 1>(a, s);
    case 2:
      return fp64 ? fold_regions<double, 2>(a, s)
                  : fold_regions<float, 2>(a, s);
    case 3:
      return fp64 ? fold_regions<double, 3>(a, s)
                  : fold_regions<float, 3>(a, s);
  }
  throw Error("transform must be 1D/2D/3D");
}

void launch_spread_det(bool fp64, const SpreadArgs &a, hipStream_t s) {
  DMX_ASSERT(a.use_lds, "deterministic mode needs the tile region in LDS");
  DMX_ASSERT(a.g.W <= 16 && a.slab && a.tile_sub0 && a.tile_nsub,
             "internal: deterministic spreading arguments");
  switch (a.g.ndim) {
    case 1: return fp64 ? go<double, 1>(a, s) : go<float, 1>(a, s);
    case 2: return fp64 ? go<double, 2>(a, s) : go<float, 2>(a, s);
    case 3: return fp64 ? go<double, 3>(a, s) : go<float, 3>(a, s);
  }
  throw Error("transform must be 1D/2D/3D");
}

}  // namespace dmx
