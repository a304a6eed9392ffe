// spread_impl.h -- type-1 spreading and type-2 interpolation kernels.
//
// Role of spreading_helper / interpolation_helper (nufft.h:639-834,
// 1141-1272, 1589-1729), designed for the GPU: one workgroup per
// subproblem (at most MAXSUB sorted points of one tile).  The spreader
// accumulates the tile plus its kernel halo, (2^log2tile + W - 1)^d cells,
// in LDS with LDS float atomics and flushes the non-zero cells to the
// oversampled grid in HBM with hardware float atomics (periodic wrap).  The
// interpolator stages the same region from HBM into LDS once and reduces the
// W^d taps per point from LDS.  A no-LDS variant (direct HBM atomics /
// loads) serves 3-D FP64 kernels whose region would not fit in 64 KiB.
//
// Per point: FP64 fold (fold_axis), kernel argument x = W-1-2*frac cast to
// the compute type, W weights per axis by even/odd Horner (kernel_weights).
#pragma once
#include <type_traits>

#include "dmx.h"

namespace dmx {

template <typename T> struct cplx2;
template <> struct cplx2<float> { using type = float2; };
template <> struct cplx2<double> { using type = double2; };

template <typename T>
__device__ __forceinline__ void load_value(const void *v, size_t j, int f64,
                                           T &re, T &im) {
  if (f64) {
    double2 x = ((const double2 *)v)[j];
    re = (T)x.x;
    im = (T)x.y;
  } else {
    float2 x = ((const float2 *)v)[j];
    re = (T)x.x;
    im = (T)x.y;
  }
}

template <typename T>
__device__ __forceinline__ void store_value(void *v, size_t j, int f64, T re,
                                            T im) {
  if (f64)
    ((double2 *)v)[j] = make_double2((double)re, (double)im);
  else
    ((float2 *)v)[j] = make_float2((float)re, (float)im);
}

// tile index -> per-axis tile coordinates -> region origin b0
__device__ __forceinline__ void tile_origin(const Geo &g, uint32_t tile,
                                            int *b0) {
  uint32_t t[3] = {0, 0, 0};
  if (g.ndim == 1) {
    t[0] = tile;
  } else if (g.ndim == 2) {
    t[0] = tile / (uint32_t)g.ntiles[1];
    t[1] = tile % (uint32_t)g.ntiles[1];
  } else {
    uint32_t n12 = (uint32_t)g.ntiles[1] * (uint32_t)g.ntiles[2];
    t[0] = tile / n12;
    t[1] = (tile / (uint32_t)g.ntiles[2]) % (uint32_t)g.ntiles[1];
    t[2] = tile % (uint32_t)g.ntiles[2];
  }
  for (int d = 0; d < 3; ++d) b0[d] = ((int)t[d] << g.log2tile) - g.nsafe;
}

// A lane's walk over the cells c, c + step, c + 2 step, ... of a region
// with sides (S, S, inner) (the last NDIM of them): c = (u * S + v) *
// inner + w.  The first cell is split once; each step adds constant
// increments with carries, so no cell needs an integer division.
template <int NDIM>
struct RegionWalk {
  int u = 0, v = 0, w = 0, du = 0, dv = 0, dw = 0, S = 1, inner = 1;
  __device__ RegionWalk(int c, int step, int S_, int inner_)
      : S(S_), inner(inner_) {
    w = c;
    dw = step;
    if (NDIM > 1) {
      w = c % inner;
      dw = step % inner;
      const int r = c / inner, q = step / inner;
      if (NDIM == 2) {
        u = r;
        du = q;
      } else {
        v = r % S;
        u = r / S;
        dv = q % S;
        du = q / S;
      }
    }
  }
  __device__ __forceinline__ void next() {
    w += dw;
    if (NDIM > 1 && w >= inner) {
      w -= inner;
      if (NDIM == 2) ++u;
      else ++v;
    }
    if (NDIM == 3) {
      v += dv;
      if (v >= S) {
        v -= S;
        ++u;
      }
    }
    u += du;
  }
};

// grid index of a region coordinate: one compare inside the period
__device__ __forceinline__ int wrapg(int x, int n) {
  return (unsigned)x < (unsigned)n ? x : wrapi(x, n);
}

// grid element of the region cell at walk position rw (origin b0)
template <int NDIM>
__device__ __forceinline__ size_t region_grid_index(const RegionWalk<NDIM> &rw,
                                                    const int *b0, int n0,
                                                    int n1, int n2) {
  if constexpr (NDIM == 1) return (size_t)wrapg(b0[0] + rw.w, n0);
  else if constexpr (NDIM == 2)
    return (size_t)wrapg(b0[0] + rw.u, n0) * n1 + wrapg(b0[1] + rw.w, n1);
  else
    return ((size_t)wrapg(b0[0] + rw.u, n0) * n1 + wrapg(b0[1] + rw.v, n1)) *
               n2 +
           wrapg(b0[2] + rw.w, n2);
}

template <typename T, int NDIM, int W>
__device__ __forceinline__ void point_setup(const SpreadArgs &a, uint32_t p,
                                            uint32_t &j, int *i0, T (*w)[W]) {
  // the original index is only needed for data still in input order
  j = (a.coords_sorted && a.values_sorted) ? p : a.idx[p];
  const size_t cj = a.coords_sorted ? (size_t)p : (size_t)j;
  const T *cf = (const T *)a.coef;
  double x_[NDIM];
  load_point(a.g, a.coords, cj, NDIM, x_);
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    double frac;
    i0[d] = fold_axis(a.g, d, x_[d], frac);
    T x = (T)((double)(W - 1) - 2.0 * frac);
    kernel_weights<T, W>(x, cf, w[d]);
  }
}

// LDS accumulation is always in FP64: ds_add_f64 runs at close to the rate
// of plain LDS stores on gfx950 while ds_add_f32 is several times slower
// under contention (measured: 2-D, M=1e7, W=7: 0.71 ms vs 5.1 ms), and the
// wider accumulator only improves accuracy.
using Acc = double;

template <typename T, int W>
__device__ __forceinline__ void spread_row3(Acc *sre, Acc *sim, int row,
                                            T fr, T fi, const T *w2) {
#pragma unroll
  for (int c2 = 0; c2 < W; ++c2) {
    atomicAdd(&sre[row + c2], (Acc)(fr * w2[c2]));
    atomicAdd(&sim[row + c2], (Acc)(fi * w2[c2]));
  }
}

// Points of a subproblem are spread in chunks (spread_chunk: the whole
// subproblem in 1-D/2-D FP32).  ds_add_f64 resolves bank conflicts per
// quarter-wave: 16 lanes whose cells are distinct mod 16 run at ~7 lane-
// atomics/CU/clock, random cells at ~2.5 (measured on gfx950), and these
// atomics bound the spreader.  So:
// phase 1 (thread t <-> points t, t+256, ... of the chunk, all loads in
//   flight first): fold, kernel argument, region offset and value; a rank
//   inside the point's class (first cell mod 16); after a prefix over the 16
//   classes each point has a position in the class-sorted order, which is
//   split into 64 runs of K = ceil(n/64) positions.  The point is written to
//   stash slot j*64 + r (run r, place j in the run);
// phase 2: lane l takes run r(l) = 4*(l mod 16) + l/16, so the 16 lanes of
//   a quarter-wave hold 16 consecutive runs, i.e. (nearly) all 16 classes,
//   and all W^d taps of the point, which shift every lane by the same
//   offset, stay conflict-free; stash slots j*64 + r are consecutive across
//   the wave, so the stash reads are conflict-free as well.
// FIX (FP32 only): the region accumulates in fixed point, one 64-bit LDS
// integer atomic per tap instead of two FP64 ones.  A tap adds the 64-bit
// integer  re_i + 2^32 im_i  of its rounded, scaled contribution (the real
// part sign-extended); as long as every cell's sums of re_i and of im_i
// stay below 2^31 in magnitude, a cell's total T holds re = the
// sign-extended low half of T and im = (T - re) / 2^32, both exact integer
// sums (independent of the order of the atomics).  ds_add_u64 runs faster
// than ds_add_f64 on gfx950 (about 8.2 vs 6.6 lane-atomics per CU and clock
// on distinct cells, 5.3 vs 2.9 on random ones), and there is one per tap
// instead of two.
// The scale of a subproblem comes from a bound on its cells' sums: the
// region is cut into blocks of 8 cells per axis, each point adds
// ceil(2^20 max(|re|, |im|) / vmax) (vmax: the subproblem's largest) to the
// blocks its W^d footprint touches (u32 LDS atomics), and a cell's sums are
// at most the largest block's total * vmax * 2^-20 * max(w) (the points
// reaching a cell all touch its block).  The values are scaled so that this
// bound is 0.999 * 2^31.  Regions of more than 128 blocks take the sum over
// all the subproblem's points as the bound.  The rounding to integers then
// costs about 1e-7 of the output's relative L2 error on dense subproblems,
// so the mode serves the FP32 path at eps >= 1e-5 (Params::spread_fix).
// The whole subproblem is one chunk (the scale must be known before the
// first tap).
template <typename T, int NDIM, bool FIX = false>
struct SpreadStash {
  static constexpr int SC = FIX ? MAXSUB : spread_chunk(NDIM, sizeof(T) == 8);
  typename cplx2<T>::type v[SC];
  T x[NDIM][SC];
  uint32_t o[SC];  // region offsets, 10 bits per axis
  uint32_t cnt[16], off[16];
};
// FIX: the scale's inputs and result, beside the region
constexpr int FIX_BLK = 8, FIX_MAXBLK = 128;
struct FixShared {
  uint32_t vmaxb;  // bits of the largest max(|re|, |im|) (positive float)
  float vsum;      // sum of max(|re|, |im|)
  float scale, inv;
  uint32_t blk[FIX_MAXBLK];
};
__host__ __device__ inline size_t fix_region_bytes(size_t cells) {
  return (cells * 8 + 15) / 16 * 16 + (sizeof(FixShared) + 15) / 16 * 16;
}

// bound on the product of a point's kernel weights (the piece polynomials
// approximate psi <= 1 to within the kernel's epsilon)
constexpr float FIX_WMAX = 1.0625f;

// one fixed-point tap: re and im rounded to int32, packed as re + 2^32 im
__device__ __forceinline__ unsigned long long fix_pack(float re, float im) {
  const int a = __float2int_rn(re), b = __float2int_rn(im);
  return (unsigned long long)(unsigned)a |
         ((unsigned long long)(unsigned)(b + (a >> 31)) << 32);
}
// a cell's total -> (re, im) integers
__device__ __forceinline__ void fix_unpack(unsigned long long t, int &re,
                                           int &im) {
  re = (int)(unsigned)t;
  im = (int)((long long)(t - (unsigned long long)(long long)re) >> 32);
}

template <typename T, int NDIM>
__device__ __forceinline__ int region_cell(uint32_t o, int S) {
  const int o0 = o & 1023, o1 = (o >> 10) & 1023, o2 = (o >> 20) & 1023;
  if constexpr (NDIM == 1) return o0;
  else if constexpr (NDIM == 2) return o0 * S + o1;
  else return (o0 * S + o1) * S + o2;
}

template <typename T, int NDIM, int W, bool LDS, bool FIX = false>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs a) {
  static_assert(!FIX || (sizeof(T) == 4 && LDS), "fixed point: FP32 in LDS");
  extern __shared__ unsigned char smem_raw[];
  const uint32_t total = *a.nsub_total;
  const int S = a.lds_side;
  const int tile = 1 << a.g.log2tile;
  int ncell = S;
  if (NDIM > 1) ncell *= S;
  if (NDIM > 2) ncell *= S;
  Acc *sre = (Acc *)smem_raw;
  Acc *sim = sre + ncell;
  unsigned long long *sfx = (unsigned long long *)smem_raw;  // FIX cells
  FixShared &fx =
      *(FixShared *)(smem_raw + ((size_t)ncell * 8 + 15) / 16 * 16);
  SpreadStash<T, NDIM, FIX> &st = *(SpreadStash<T, NDIM, FIX> *)(
      smem_raw + (FIX ? fix_region_bytes(ncell)
                      : 2 * (size_t)ncell * sizeof(Acc)));
  // FIX: blocks of FIX_BLK cells per axis over the region
  const int nbx = (S + FIX_BLK - 1) / FIX_BLK;
  int nblk = nbx;
  if (NDIM > 1) nblk *= nbx;
  if (NDIM > 2) nblk *= nbx;
  T *grid = (T *)a.grid;
  const T *cf = (const T *)a.coef;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // the flush's walk over the region (sides S), split once per lane
  const RegionWalk<NDIM> flush_walk0(tid, (int)blockDim.x, S, S);
  for (uint32_t sidx0 = blockIdx.x; sidx0 < total; sidx0 += gridDim.x) {
    // stored regions: a tile's first subproblem brings all of the tile's
    // subproblems to this workgroup, the others are skipped
    uint32_t nq = 1;
    if (LDS && a.slab_out) {
      const Sub s0 = a.subs[sidx0];
      if (s0.rank != 0) continue;
      nq = a.tile_nsub[s0.tile];
    }
    for (uint32_t q = 0; q < nq; ++q) {
      const uint32_t sidx = sidx0 + q;
      const Sub sb = a.subs[sidx];
      int b0[3];
      tile_origin(a.g, sb.tile, b0);
      if constexpr (LDS) {
        for (int c = tid; c < ncell; c += blockDim.x) {
          if constexpr (FIX) {
            sfx[c] = 0ull;
          } else {
            sre[c] = Acc(0);
            sim[c] = Acc(0);
          }
        }
        float fscale = 1.f, finv = 1.f;  // FIX: this subproblem's scale
        bool fbad = false;  // FIX: a value of the subproblem is not finite
        constexpr int SC = SpreadStash<T, NDIM, FIX>::SC, PT = SC / 256;
        for (uint32_t c0 = sb.start; c0 < sb.end; c0 += SC) {
          const int nc = (int)min((uint32_t)SC, sb.end - c0);
          const int K = (nc + 63) >> 6;            // run length
          // pos / K == pos * kmul >> 16
          const uint32_t kmul = (65536u + K - 1) / K;
          if (tid < 16) st.cnt[tid] = 0;
          if constexpr (FIX) {
            if (tid == 0) {
              fx.vmaxb = 0u;
              fx.vsum = 0.f;
            }
            if (tid < FIX_MAXBLK) fx.blk[tid] = 0u;
          }
          // the previous chunk's phase 2 is done with the stash (and the
          // accumulators are zero before the first chunk)
          __syncthreads();
          T xs[PT][NDIM];
          uint32_t os[PT], rnk[PT];
          int cls[PT];
          T vr_[PT], vi_[PT];
          if (a.rec) {
            // records in tile order: offsets, kernel arguments and values
            // read in sequence
            const int L = a.g.log2tile;
            const uint32_t tm = (1u << L) - 1u;
#pragma unroll
            for (int k = 0; k < PT; ++k) {
              const int i = tid + 256 * k;
              os[k] = 0;
              if (i < nc) {
                const uint32_t p = c0 + i;
                const uint32_t w = a.recs.word[p];
                uint32_t o = 0;
#pragma unroll
                for (int d = 0; d < NDIM; ++d) {
                  o |= ((w >> (d * L)) & tm) << (10 * d);
                  xs[k][d] = ((const T *)a.recs.x[d])[p];
                }
                os[k] = o;
                const typename cplx2<T>::type v =
                    ((const typename cplx2<T>::type *)a.recs.v)[p];
                vr_[k] = v.x;
                vi_[k] = v.y;
              }
            }
          } else {
            // loads of all this thread's points first (memory-level
            // parallelism)
            uint32_t jj[PT];
            double cc[PT][NDIM];
#pragma unroll
            for (int k = 0; k < PT; ++k) {
              const int i = tid + 256 * k;
              jj[k] = i < nc ? a.idx[c0 + i] : 0u;
            }
#pragma unroll
            for (int k = 0; k < PT; ++k) {
              const int i = tid + 256 * k;
              if (i < nc) {
                const uint32_t p = c0 + i;
                const size_t cj = a.coords_sorted ? (size_t)p : (size_t)jj[k];
                load_point(a.g, a.coords, cj, NDIM, cc[k]);
                load_value<T>(a.values,
                              a.values_sorted ? (size_t)p : (size_t)jj[k],
                              a.val_f64, vr_[k], vi_[k]);
              }
            }
#pragma unroll
            for (int k = 0; k < PT; ++k) {
              const int i = tid + 256 * k;
              os[k] = 0;
              if (i < nc) {
                uint32_t o = 0;
#pragma unroll
                for (int d = 0; d < NDIM; ++d) {
                  double frac;
                  const int i0 = fold_axis(a.g, d, cc[k][d], frac);
                  xs[k][d] = (T)((double)(W - 1) - 2.0 * frac);
                  int od = i0 - b0[d];
                  if (od < 0 || od >= tile) {
                    if (a.offcheck) atomicAdd(a.offcheck, 1u);
                    od = od < 0 ? 0 : tile - 1;
                  }
                  o |= (uint32_t)od << (10 * d);
                }
                os[k] = o;
              }
            }
          }
#pragma unroll
          for (int k = 0; k < PT; ++k) {
            const int i = tid + 256 * k;
            cls[k] = 0;
            rnk[k] = 0;
            if (i < nc) {
              cls[k] = region_cell<T, NDIM>(os[k], S) & 15;
              rnk[k] = atomicAdd(&st.cnt[cls[k]], 1u);
            }
          }
          float amax[FIX ? PT : 1];
          if constexpr (FIX) {
            float vs = 0.f, vm = 0.f;
#pragma unroll
            for (int k = 0; k < PT; ++k) {
              amax[k] = 0.f;
              if (tid + 256 * k < nc) {
                const float ar = fabsf((float)vr_[k]),
                            ai = fabsf((float)vi_[k]);
                // fmaxf ignores NaN: a NaN part counts as infinite
                amax[k] = (ar != ar || ai != ai) ? __builtin_inff()
                                                 : fmaxf(ar, ai);
              }
              vs += amax[k];
              vm = fmaxf(vm, amax[k]);
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
              vs += __shfl_xor(vs, o);
              vm = fmaxf(vm, __shfl_xor(vm, o));
            }
            if (lane == 0) {
              atomicAdd(&fx.vsum, vs);
              atomicMax(&fx.vmaxb, __float_as_uint(vm));
            }
          }
          __syncthreads();
          if constexpr (FIX) {
            // each point's quantized magnitude into the blocks it touches
            const float vmax = __uint_as_float(fx.vmaxb);
            if (nblk <= FIX_MAXBLK && vmax > 0.f) {
#pragma unroll
              for (int k = 0; k < PT; ++k) {
                if (tid + 256 * k >= nc) continue;
                const uint32_t q =
                    (uint32_t)ceilf(amax[k] / vmax * 1048576.f);
                int blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
#pragma unroll
                for (int d = 0; d < NDIM; ++d) {
                  const int od = (int)((os[k] >> (10 * d)) & 1023u);
                  blo[d] = od / FIX_BLK;
                  bhi[d] = (od + W - 1) / FIX_BLK;
                }
                for (int b0_ = blo[0]; b0_ <= bhi[0]; ++b0_)
                  for (int b1_ = blo[1]; b1_ <= bhi[1]; ++b1_)
                    for (int b2_ = blo[2]; b2_ <= bhi[2]; ++b2_) {
                      int bi = b0_;
                      if (NDIM > 1) bi = bi * nbx + b1_;
                      if (NDIM > 2) bi = bi * nbx + b2_;
                      atomicAdd(&fx.blk[bi], q);
                    }
              }
            }
          }
          if (tid < 16) {
            uint32_t run = 0;
            for (int k = 0; k < tid; ++k) run += st.cnt[k];
            st.off[tid] = run;
          }
          __syncthreads();
          if constexpr (FIX) {
            // wave 0: the bound, and the scale that maps it to 0.999 * 2^31
            // (published by the barrier before phase 2)
            if (wave == 0) {
              const float vmax = __uint_as_float(fx.vmaxb);
              double bound = (double)fx.vsum;
              if (nblk <= FIX_MAXBLK) {
                uint32_t m = 0u;
                for (int b = lane; b < nblk; b += 64) m = max(m, fx.blk[b]);
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
                bound = (double)m * (double)vmax * (1.0 / 1048576.0);
              }
              bound *= (double)FIX_WMAX;
              if (lane == 0) {
                const double sc = bound > 0.0 ? 0.999 * 2147483648.0 / bound
                                              : 1.0;
                fx.scale = (float)fmin(sc, 1e30);
                fx.inv = (float)(1.0 / (double)fx.scale);
              }
            }
          }
#pragma unroll
          for (int k = 0; k < PT; ++k) {
            const int i = tid + 256 * k;
            if (i < nc) {
              const uint32_t pos = st.off[cls[k]] + rnk[k];
              const uint32_t r = (pos * kmul) >> 16, jr = pos - r * K;
              const int slot = (int)(jr * 64 + r);
#pragma unroll
              for (int d = 0; d < NDIM; ++d) st.x[d][slot] = xs[k][d];
              st.o[slot] = os[k];
              st.v[slot].x = vr_[k];
              st.v[slot].y = vi_[k];
            }
          }
          __syncthreads();
          if constexpr (FIX) {
            fscale = fx.scale;
            finv = fx.inv;
            fbad = !(__uint_as_float(fx.vmaxb) <= 3.4028235e38f);
          }
          const int run = ((lane & 15) << 2) | (lane >> 4);
          // 3-D tiles of few points (runs shorter than the four waves): every
          // wave takes every run and a quarter of each point's W planes, so
          // all waves work; the planes shift all lanes alike, which keeps
          // the lanes of a quarter-wave on distinct classes
          const bool split = NDIM == 3 && K < 4;
          const int jr0 = split ? 0 : wave, jstep = split ? 1 : 4;
          const int q00 = split ? wave : 0, qstep = split ? 4 : 1;
          for (int jr = jr0; jr < K; jr += jstep) {
            if (run * K + jr >= nc) continue;
            const int i = jr * 64 + run;
            const uint32_t o = st.o[i];
            T vr = st.v[i].x, vi = st.v[i].y;
            T w[NDIM][W];
#pragma unroll
            for (int d = 0; d < NDIM; ++d)
              kernel_weights<T, W>(st.x[d][i], cf, w[d]);
            const int cell0 = region_cell<T, NDIM>(o, S);
            if constexpr (FIX) {
              vr *= fscale;
              vi *= fscale;
              if constexpr (NDIM == 1) {
#pragma unroll
                for (int q0 = 0; q0 < W; ++q0)
                  atomicAdd(&sfx[cell0 + q0],
                            fix_pack(vr * w[0][q0], vi * w[0][q0]));
              } else if constexpr (NDIM == 2) {
#pragma unroll
                for (int q0 = 0; q0 < W; ++q0) {
                  const T fr = vr * w[0][q0], fi = vi * w[0][q0];
                  const int row = cell0 + q0 * S;
#pragma unroll
                  for (int q1 = 0; q1 < W; ++q1)
                    atomicAdd(&sfx[row + q1],
                              fix_pack(fr * w[1][q1], fi * w[1][q1]));
                }
              } else {
#pragma unroll 1
                for (int q0 = q00; q0 < W; q0 += qstep) {
                  T w0 = w[0][0];
#pragma unroll
                  for (int k = 1; k < W; ++k)
                    if (k == q0) w0 = w[0][k];
#pragma unroll
                  for (int q1 = 0; q1 < W; ++q1) {
                    const T f01 = w0 * w[1][q1];
                    const T fr = vr * f01, fi = vi * f01;
                    const int row = cell0 + (q0 * S + q1) * S;
#pragma unroll
                    for (int q2 = 0; q2 < W; ++q2)
                      atomicAdd(&sfx[row + q2],
                                fix_pack(fr * w[2][q2], fi * w[2][q2]));
                  }
                }
              }
            } else if constexpr (NDIM == 1) {
#pragma unroll
              for (int q0 = 0; q0 < W; ++q0) {
                atomicAdd(&sre[cell0 + q0], (Acc)(vr * w[0][q0]));
                atomicAdd(&sim[cell0 + q0], (Acc)(vi * w[0][q0]));
              }
            } else if constexpr (NDIM == 2) {
#pragma unroll
              for (int q0 = 0; q0 < W; ++q0) {
                const T fr = vr * w[0][q0], fi = vi * w[0][q0];
                const int row = cell0 + q0 * S;
#pragma unroll
                for (int q1 = 0; q1 < W; ++q1) {
                  atomicAdd(&sre[row + q1], (Acc)(fr * w[1][q1]));
                  atomicAdd(&sim[row + q1], (Acc)(fi * w[1][q1]));
                }
              }
            } else {
              if (W <= 8 && qstep == 1) {
#pragma unroll
                for (int q0 = 0; q0 < W; ++q0)
#pragma unroll
                  for (int q1 = 0; q1 < W; ++q1) {
                    const T f01 = w[0][q0] * w[1][q1];
                    spread_row3<T, W>(sre, sim, cell0 + (q0 * S + q1) * S,
                                      vr * f01, vi * f01, w[2]);
                  }
              } else {
#pragma unroll 1
                for (int q0 = q00; q0 < W; q0 += qstep) {
                  T w0 = w[0][0];
#pragma unroll
                  for (int k = 1; k < W; ++k)
                    if (k == q0) w0 = w[0][k];
#pragma unroll
                  for (int q1 = 0; q1 < W; ++q1) {
                    const T f01 = w0 * w[1][q1];
                    spread_row3<T, W>(sre, sim, cell0 + (q0 * S + q1) * S,
                                      vr * f01, vi * f01, w[2]);
                  }
                }
              }
            }
          }
        }
        __syncthreads();
        if (a.slab_out) {
          // the whole region to the tile's slab (the first subproblem stores
          // it, later ones add to it: each thread revisits the cells it
          // wrote); the fold adds the overlaps of neighbouring tiles.  A
          // non-finite value makes every cell NaN.
          typename cplx2<T>::type *dst =
              (typename cplx2<T>::type *)a.slab + (size_t)sb.tile * ncell;
          for (int c = tid; c < ncell; c += blockDim.x) {
            T re, im;
            if constexpr (FIX) {
              int ri, ii;
              fix_unpack(sfx[c], ri, ii);
              re = fbad ? (T)__builtin_nanf("") : (float)ri * finv;
              im = fbad ? (T)__builtin_nanf("") : (float)ii * finv;
            } else {
              re = (T)sre[c];
              im = (T)sim[c];
            }
            typename cplx2<T>::type v;
            if (sb.rank == 0) {
              v.x = re;
              v.y = im;
            } else {
              v = dst[c];
              v.x += re;
              v.y += im;
            }
            dst[c] = v;
          }
          __syncthreads();
          continue;
        }
        const int n0 = a.g.nover[0], n1 = NDIM > 1 ? a.g.nover[1] : 1,
                  n2 = NDIM > 2 ? a.g.nover[2] : 1;
        if (FIX && fbad) {
          // a NaN or infinite value: the integer sums mean nothing, and the
          // FFT turns one NaN grid cell into NaN everywhere, as FP64
          // accumulation of such a value would
          if (tid == 0) {
            size_t gi = (size_t)wrapi(b0[0], n0);
            if (NDIM > 1) gi = gi * n1 + wrapi(b0[1], n1);
            if (NDIM > 2) gi = gi * n2 + wrapi(b0[2], n2);
            atomicAdd(&grid[2 * gi], (T)__builtin_nanf(""));
          }
        }
        RegionWalk<NDIM> rw = flush_walk0;
        for (int c = tid; c < ncell && !(FIX && fbad); c += blockDim.x) {
          // this cell's grid element, then the lane's next cell
          const size_t gi = region_grid_index(rw, b0, n0, n1, n2);
          rw.next();
          T re, im;
          if constexpr (FIX) {
            const unsigned long long t = sfx[c];
            if (t == 0ull) continue;
            int ri, ii;
            fix_unpack(t, ri, ii);
            re = (float)ri * finv;
            im = (float)ii * finv;
          } else {
            re = (T)sre[c];
            im = (T)sim[c];
          }
          if (re == T(0) && im == T(0)) continue;
          atomicAdd(&grid[2 * gi], re);
          atomicAdd(&grid[2 * gi + 1], im);
        }
        __syncthreads();
      } else {
        // direct HBM accumulation with periodic wrap (regions beyond LDS)
        const int n0 = a.g.nover[0], n1 = NDIM > 1 ? a.g.nover[1] : 1,
                  n2 = NDIM > 2 ? a.g.nover[2] : 1;
        for (uint32_t p = sb.start + tid; p < sb.end; p += blockDim.x) {
          uint32_t j;
          int i0[3];
          T w[NDIM][W];
          point_setup<T, NDIM, W>(a, p, j, i0, w);
          T vr, vi;
          load_value<T>(a.values, a.values_sorted ? (size_t)p : (size_t)j,
                        a.val_f64, vr, vi);
#pragma unroll 1
          for (int c0 = 0; c0 < W; ++c0) {
            const int g0 = wrapi(i0[0] + c0, n0);
            if constexpr (NDIM == 1) {
              atomicAdd(&grid[2 * (size_t)g0], vr * w[0][c0]);
              atomicAdd(&grid[2 * (size_t)g0 + 1], vi * w[0][c0]);
            } else {
#pragma unroll 1
              for (int c1 = 0; c1 < W; ++c1) {
                const int g1 = wrapi(i0[1] + c1, n1);
                const T f01 = w[0][c0] * w[1][c1];
                if constexpr (NDIM == 2) {
                  size_t gi = (size_t)g0 * n1 + g1;
                  atomicAdd(&grid[2 * gi], vr * f01);
                  atomicAdd(&grid[2 * gi + 1], vi * f01);
                } else {
                  const size_t rb = ((size_t)g0 * n1 + g1) * n2;
#pragma unroll
                  for (int c2 = 0; c2 < W; ++c2) {
                    const size_t gi = rb + wrapi(i0[2] + c2, n2);
                    atomicAdd(&grid[2 * gi], vr * f01 * w[2][c2]);
                    atomicAdd(&grid[2 * gi + 1], vi * f01 * w[2][c2]);
                  }
                }
              }
            }
          }
        }
      }
    }
  }
}

// Padded inner side of the interpolation's LDS region: even (16-byte
// aligned rows of complex float) and two cells longer than the region, so
// the aligned pair reads of the FP32 2-D path stay inside the row.
__host__ __device__ inline int interp_inner_pad(int S) { return (S + 3) & ~1; }

// MODE 0: values (out[j] complex); MODE 1: coordinate gradient
// (out[j*NDIM+d] complex = d/dcoord_d of the interpolated value).
// The region is staged in LDS as interleaved complex cells (one 8/16-byte
// read per tap); 2-D FP32 reads aligned pairs of cells (16 bytes) and picks
// the taps by the parity of the first column.
template <typename T, int NDIM, int W, bool LDS, int MODE>
__global__ __launch_bounds__(256) void k_interp(SpreadArgs a, void *out,
                                                int out_f64) {
  using C2 = typename cplx2<T>::type;
  extern __shared__ unsigned char smem_raw[];
  const uint32_t total = *a.nsub_total;
  const int S = a.lds_side;
  const int SP = interp_inner_pad(S);
  const int tile = 1 << a.g.log2tile;
  int ncell = SP;
  if (NDIM > 1) ncell *= S;
  if (NDIM > 2) ncell *= S;
  C2 *buf = (C2 *)smem_raw;
  const T *grid = (const T *)a.grid;
  const C2 *gridc = (const C2 *)a.grid;
  const int n0 = a.g.nover[0], n1 = NDIM > 1 ? a.g.nover[1] : 1,
            n2 = NDIM > 2 ? a.g.nover[2] : 1;
  // region staging: the inner index runs over the padded side (the pad
  // cells hold copies of in-range cells, never used in arithmetic); a
  // batch of STG loads of a lane is in flight before its LDS writes (two
  // for FP64: eight double2 registers push the wide FP64 kernels past 2
  // waves per SIMD)
  constexpr int STG = sizeof(T) == 8 ? 2 : 8;
  const RegionWalk<NDIM> walk0((int)threadIdx.x, (int)blockDim.x, S, SP);
  for (uint32_t sidx = blockIdx.x; sidx < total; sidx += gridDim.x) {
    const Sub sb = a.subs[sidx];
    int b0[3];
    tile_origin(a.g, sb.tile, b0);
    if (LDS) {
      RegionWalk<NDIM> rw = walk0;
      for (int cb = 0; cb < ncell; cb += STG * (int)blockDim.x) {
        C2 v[STG];
#pragma unroll
        for (int k = 0; k < STG; ++k) {
          const bool in = cb + k * (int)blockDim.x + (int)threadIdx.x < ncell;
          const size_t gi = region_grid_index(rw, b0, n0, n1, n2);
          v[k] = gridc[in ? gi : 0];
          rw.next();
        }
#pragma unroll
        for (int k = 0; k < STG; ++k) {
          const int c = cb + k * (int)blockDim.x + (int)threadIdx.x;
          if (c < ncell) buf[c] = v[k];
        }
      }
      __syncthreads();
    }
    // 3-D tiles of few points: tpp adjacent lanes share a point, each
    // summing a 1/tpp share of its W planes (combined by lane shuffles), so
    // that more of the workgroup works
    const uint32_t nc = sb.end - sb.start;
    auto point_loop = [&](auto tpp_c) {
    constexpr int tpp = decltype(tpp_c)::value;
    const uint32_t lp = threadIdx.x / tpp;
    const int part = (int)(threadIdx.x % tpp);
    const uint32_t pstep = blockDim.x / tpp;
    // one point per lane and round (tpp == 1): the indices and coordinates
    // of all of a lane's rounds (at most MAXSUB / 256) are loaded up front,
    // so that their latencies overlap
    // (records: index, word and kernel arguments; the arguments are kept
    // as doubles, exact for either compute type)
    constexpr int PF = tpp == 1 ? MAXSUB / 256 : 1;
    uint32_t jpf[PF], wpf[PF];
    double xpf[PF][NDIM];
    if constexpr (tpp == 1) {
      if (a.rec) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
          const uint32_t q = lp + (uint32_t)k * pstep;
          jpf[k] = wpf[k] = 0u;
          if (q < nc) {
            const uint32_t p = sb.start + q;
            jpf[k] = a.recs.j[p];
            wpf[k] = a.recs.word[p];
#pragma unroll
            for (int d = 0; d < NDIM; ++d)
              xpf[k][d] = (double)((const T *)a.recs.x[d])[p];
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
          const uint32_t q = lp + (uint32_t)k * pstep;
          jpf[k] = q < nc ? a.idx[sb.start + q] : 0u;
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) {
          const uint32_t q = lp + (uint32_t)k * pstep;
          if (q < nc)
            load_point(a.g, a.coords,
                       a.coords_sorted ? (size_t)(sb.start + q)
                                       : (size_t)jpf[k],
                       NDIM, xpf[k]);
        }
      }
    }
    for (uint32_t pb = 0; pb < nc; pb += pstep) {
      const bool valid = tpp == 1 || pb + lp < nc;
      if (tpp == 1 && pb + lp >= nc) break;
      const uint32_t p = sb.start + pb + lp;
      uint32_t j = 0;
      T sr = 0, si = 0;  // MODE 0: this lane's share of the value
      if (valid) {
      const T *cf = (const T *)a.coef;
      int i0[3];
      T w[NDIM][W];
      T dw[MODE ? NDIM : 1][W];
      double x_[NDIM];
      uint32_t wd = 0;
      if constexpr (tpp == 1) {
        // round pb / pstep (< PF: nc <= MAXSUB, pstep = 256), selected
        // with constant indices so the arrays stay in registers
        const int kk = (int)(pb / pstep);
#pragma unroll
        for (int k = 0; k < PF; ++k)
          if (k == kk) {
            j = jpf[k];
            wd = wpf[k];
#pragma unroll
            for (int d = 0; d < NDIM; ++d) x_[d] = xpf[k][d];
          }
      } else if (a.rec) {
        j = a.recs.j[p];
        wd = a.recs.word[p];
#pragma unroll
        for (int d = 0; d < NDIM; ++d)
          x_[d] = (double)((const T *)a.recs.x[d])[p];
      } else {
        j = a.idx[p];
        load_point(a.g, a.coords, a.coords_sorted ? (size_t)p : (size_t)j,
                   NDIM, x_);
      }
      int o[3];
      T xk[NDIM];
      if (a.rec) {
        // the record's tile-local offsets and kernel arguments
        const int L = a.g.log2tile;
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          o[d] = (int)((wd >> (d * L)) & ((1u << L) - 1u));
          i0[d] = b0[d] + o[d];
          xk[d] = (T)x_[d];
        }
      } else {
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          double frac;
          i0[d] = fold_axis(a.g, d, x_[d], frac);
          xk[d] = (T)((double)(W - 1) - 2.0 * frac);
          o[d] = i0[d] - b0[d];
          if (o[d] < 0 || o[d] >= tile) {
            if (a.offcheck) atomicAdd(a.offcheck, 1u);
            o[d] = o[d] < 0 ? 0 : tile - 1;
          }
        }
      }
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        if constexpr (MODE == 0)
          kernel_weights<T, W>(xk[d], cf, w[d]);
        else
          kernel_weights_deriv<T, W>(xk[d], cf, w[d], dw[d]);
      }
      // cell value accessor
      auto cell = [&](int c0, int c1, int c2, T &re, T &im) {
        if (LDS) {
          int c;
          if constexpr (NDIM == 1) c = o[0] + c0;
          else if constexpr (NDIM == 2) c = (o[0] + c0) * SP + o[1] + c1;
          else c = ((o[0] + c0) * S + o[1] + c1) * SP + o[2] + c2;
          const C2 v = buf[c];
          re = v.x;
          im = v.y;
        } else {
          size_t gi;
          if constexpr (NDIM == 1) gi = (size_t)wrapi(i0[0] + c0, n0);
          else if constexpr (NDIM == 2)
            gi = (size_t)wrapi(i0[0] + c0, n0) * n1 + wrapi(i0[1] + c1, n1);
          else
            gi = ((size_t)wrapi(i0[0] + c0, n0) * n1 + wrapi(i0[1] + c1, n1)) *
                     n2 +
                 wrapi(i0[2] + c2, n2);
          re = grid[2 * gi];
          im = grid[2 * gi + 1];
        }
      };
      if constexpr (MODE == 0) {
        T ar = 0, ai = 0;
        if constexpr (NDIM == 1) {
#pragma unroll
          for (int c0 = 0; c0 < W; ++c0) {
            T re, im;
            cell(c0, 0, 0, re, im);
            ar += re * w[0][c0];
            ai += im * w[0][c0];
          }
        } else if constexpr (NDIM == 2 && LDS && sizeof(T) == 4) {
          constexpr int P = (W + 2) / 2;  // aligned cell pairs per row
          const bool odd = o[1] & 1;
#pragma unroll
          for (int c0 = 0; c0 < W; ++c0) {
            const float4 *q =
                (const float4 *)(buf + (o[0] + c0) * SP + (o[1] & ~1));
            float4 v[P];
#pragma unroll
            for (int m = 0; m < P; ++m) v[m] = q[m];
            T rr = 0, ri = 0;
#pragma unroll
            for (int c1 = 0; c1 < W; ++c1) {
              // tap c1 is cell c1 (even start) or c1+1 (odd start)
              const float4 ve = v[c1 >> 1], vo = v[(c1 + 1) >> 1];
              const T er = (c1 & 1) ? ve.z : ve.x, ei = (c1 & 1) ? ve.w : ve.y;
              const T orr = (c1 & 1) ? vo.x : vo.z, oi = (c1 & 1) ? vo.y : vo.w;
              const T re = odd ? orr : er, im = odd ? oi : ei;
              rr += re * w[1][c1];
              ri += im * w[1][c1];
            }
            ar += rr * w[0][c0];
            ai += ri * w[0][c0];
          }
        } else if constexpr (NDIM == 2) {
#pragma unroll
          for (int c0 = 0; c0 < W; ++c0) {
            T rr = 0, ri = 0;
#pragma unroll
            for (int c1 = 0; c1 < W; ++c1) {
              T re, im;
              cell(c0, c1, 0, re, im);
              rr += re * w[1][c1];
              ri += im * w[1][c1];
            }
            ar += rr * w[0][c0];
            ai += ri * w[0][c0];
          }
        } else {
#pragma unroll 1
          for (int c0 = part; c0 < W; c0 += tpp) {
            T pr = 0, pi = 0;
#pragma unroll
            for (int c1 = 0; c1 < W; ++c1) {
              T rr = 0, ri = 0;
#pragma unroll
              for (int c2 = 0; c2 < W; ++c2) {
                T re, im;
                cell(c0, c1, c2, re, im);
                rr += re * w[2][c2];
                ri += im * w[2][c2];
              }
              pr += rr * w[1][c1];
              pi += ri * w[1][c1];
            }
            const T w0 = w[0][c0];
            ar += pr * w0;
            ai += pi * w0;
          }
        }
        sr = ar;
        si = ai;
      } else {
        // gradient: sum over cells of g * dw_d * prod_{e != d} w_e, times
        // dx/dcoord_d = -2 * nover_d / period_d
        T gr[NDIM], gi_[NDIM];
#pragma unroll
        for (int d = 0; d < NDIM; ++d) gr[d] = gi_[d] = T(0);
#pragma unroll 1
        for (int c0 = 0; c0 < W; ++c0) {
#pragma unroll 1
          for (int c1 = 0; c1 < (NDIM > 1 ? W : 1); ++c1) {
#pragma unroll
            for (int c2 = 0; c2 < (NDIM > 2 ? W : 1); ++c2) {
              T re, im;
              cell(c0, c1, c2, re, im);
              T f[NDIM];
#pragma unroll
              for (int d = 0; d < NDIM; ++d) {
                T prod = T(1);
#pragma unroll
                for (int e = 0; e < NDIM; ++e) {
                  const int ce = e == 0 ? c0 : (e == 1 ? c1 : c2);
                  prod *= (e == d) ? dw[e][ce] : w[e][ce];
                }
                f[d] = prod;
              }
#pragma unroll
              for (int d = 0; d < NDIM; ++d) {
                gr[d] += re * f[d];
                gi_[d] += im * f[d];
              }
            }
          }
        }
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
          const T s = (T)(-2.0 * (double)a.g.nover[d] * a.g.coordfct[d]);
          store_value<T>(out, (size_t)j * NDIM + d, out_f64, gr[d] * s,
                         gi_[d] * s);
        }
      }
      }  // valid
      if constexpr (MODE == 0) {
        // every lane takes part in the shuffles (tpp consecutive lanes)
        for (int off = 1; off < tpp; off <<= 1) {
          sr += __shfl_xor(sr, off);
          si += __shfl_xor(si, off);
        }
        if (valid && part == 0) store_value<T>(out, j, out_f64, sr, si);
      }
    }
    };
    // (only for W >= 8: the extra loop variants cost registers, and smaller
    // supports come with larger oversampling, i.e. more points per tile)
    if constexpr (NDIM == 3 && MODE == 0 && W >= 8) {
      if (nc <= 64)
        point_loop(std::integral_constant<int, 4>{});
      else if (nc <= 128)
        point_loop(std::integral_constant<int, 2>{});
      else
        point_loop(std::integral_constant<int, 1>{});
    } else {
      point_loop(std::integral_constant<int, 1>{});
    }
    if (LDS) __syncthreads();
  }
}

}  // namespace dmx
