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
      for (int d = 0; d < 3; ++d) b0[d] = ((int)t[d] << a.g.log2tile) - a.g.nsafe;
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
    for (int c = lane; c < ncell; c += 64) dst[c] = make_double2(sre[c], sim[c]);
    __syncthreads();
  }
}

__device__ __forceinline__ int floordiv(int a, int b) {
  return a >= 0 ? a / b : -((-a + b - 1) / b);
}

// Along one axis, the (tile t, region offset o) pairs covering grid index
// g: for each periodic image u = g + k*n (k = -1, 0, 1, ...) the tiles
// whose region [t*tile - nsafe, t*tile - nsafe + S) holds u, in increasing
// (k, t).  A small-grid region may wrap several times: no pair is skipped.
struct AxisPairs {
  int g, n, tile, nsafe, S, ntiles, umax;
  int k, u, t, thi;
  __device__ bool first(int g_, int n_, int tile_, int nsafe_, int S_,
                        int ntiles_) {
    g = g_; n = n_; tile = tile_; nsafe = nsafe_; S = S_; ntiles = ntiles_;
    umax = ntiles * tile + S;
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
      t = floordiv(u + nsafe - S + tile, tile);  // ceil((u+nsafe-S+1)/tile)
      if (t < 0) t = 0;
      thi = floordiv(u + nsafe, tile);
      if (thi > ntiles - 1) thi = ntiles - 1;
    }
    return true;
  }
  __device__ int off() const { return u + nsafe - t * tile; }
};

// grid cell -> sum of the slab cells covering it, in the fixed order of
// AxisPairs on axis 0, then 1, then 2, and the tile's subproblems in sorted
// order; rounded once to the compute type
template <typename T, int NDIM>
__global__ __launch_bounds__(256) void k_fold_slabs(SpreadArgs a,
                                                    size_t ncells_grid) {
  const int S = a.lds_side, tile = 1 << a.g.log2tile, nsafe = a.g.nsafe;
  int ncell = S;
  if (NDIM > 1) ncell *= S;
  if (NDIM > 2) ncell *= S;
  const double2 *slab = (const double2 *)a.slab;
  T *grid = (T *)a.grid;
  for (size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
       gi < ncells_grid; gi += (size_t)gridDim.x * blockDim.x) {
    int g[3] = {0, 0, 0};
    size_t r = gi;
    for (int d = NDIM - 1; d >= 0; --d) {
      g[d] = (int)(r % (size_t)a.g.nover[d]);
      r /= (size_t)a.g.nover[d];
    }
    double re = 0.0, im = 0.0;
    AxisPairs A0, A1, A2;
    for (bool h0 = A0.first(g[0], a.g.nover[0], tile, nsafe, S,
                            a.g.ntiles[0]);
         h0; h0 = A0.next()) {
      bool h1 = true;
      if (NDIM > 1)
        h1 = A1.first(g[1], a.g.nover[1], tile, nsafe, S, a.g.ntiles[1]);
      for (; h1; h1 = NDIM > 1 ? A1.next() : false) {
        bool h2 = true;
        if (NDIM > 2)
          h2 = A2.first(g[2], a.g.nover[2], tile, nsafe, S, a.g.ntiles[2]);
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
            const double2 v = slab[(size_t)q * ncell + cell];
            re += v.x;
            im += v.y;
          }
        }
      }
    }
    grid[2 * gi] = (T)re;
    grid[2 * gi + 1] = (T)im;
  }
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
  size_t ng = 1;
  for (int d = 0; d < NDIM; ++d) ng *= (size_t)a.g.nover[d];
  const unsigned gb = (unsigned)std::min<size_t>((ng + 255) / 256, 1u << 16);
  hipLaunchKernelGGL((k_fold_slabs<T, NDIM>), dim3(gb), dim3(256), 0, s, a,
                     ng);
  DMX_HIP(hipGetLastError());
}

}  // namespace

size_t spread_det_slab_bytes(const SpreadArgs &a) {
  size_t ncell = a.lds_side;
  for (int d = 1; d < a.g.ndim; ++d) ncell *= a.lds_side;
  return a.nsub_max * ncell * 16;
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
