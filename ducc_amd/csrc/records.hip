// records.hip -- the points themselves sorted into tile order (index build
// of unplanned transforms).
//
// The index sort (sort.hip) returns a permutation; spreading and
// interpolation then gather every point's coordinates (16 bytes in 2-D) and
// value (8 or 16 bytes) through it, and each such gather pulls a whole
// cache line.  Here the stable LSD radix sort carries the point's record
// instead of its index, so that the consumers read records in order:
//
//   word  uint32   tile << offbits | tile-local offset of each axis
//                  (log2tile bits per axis, axis d at bit d*log2tile)
//   x[d]  T        kernel argument W-1-2*frac of axis d in the compute type
//   v     T2       type 1: the point's value in the compute type
//   j     uint32   type 2: the point's input index (where its result goes)
//
// (structure of arrays, one array per field).  The first pass folds the
// coordinates itself (the same FP64 fold as everywhere, fold_axis), so no
// key array is written; the radix digits are taken from the tile bits of
// the word, i.e. the order is the stable tile order of the index sort
// (nufft.h:1364-1381 with bucket_sort.h:170-354): the record at sorted
// position p is the point idx[p] of the index sort.
#include "radix.h"

namespace dmx {

namespace {
using namespace radix;

template <typename T> struct c2t;
template <> struct c2t<float> { using type = float2; };
template <> struct c2t<double> { using type = double2; };

struct RecPassArgs {
  Geo g;
  size_t M;
  int offbits, shift;  // digit = (word >> (offbits + shift)) & 255
  const void *coords;  // first pass: (M, ndim) float32/float64
  const void *values;  // first pass, type 1: (M,) c64 / c128
  int val_f64;
  RecBufs in, out;
  const uint32_t *offs, *offs_chunk;
  int nblocks;
  uint32_t *tb_start_c, *tb_end;  // last pass: tile ranges
};

// Moves one field of the block through LDS: item r of this thread goes to
// local position lpos[r]; position q is then stored at out[gpos(q)*ostride]
// by thread q % 256, so consecutive lanes store consecutive positions of a
// digit run.
template <typename E>
__device__ __forceinline__ void stage_move(E *stage, const E (&val)[RITEMS],
                                           const uint32_t (&lpos)[RITEMS],
                                           const bool (&valid)[RITEMS],
                                           const uint32_t (&gpos)[RITEMS],
                                           int nvalid, E *out,
                                           size_t ostride) {
#pragma unroll
  for (int r = 0; r < RITEMS; ++r)
    if (valid[r]) stage[lpos[r]] = val[r];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RITEMS; ++i) {
    const int q = i * RT + (int)threadIdx.x;
    if (q < nvalid) out[(size_t)gpos[i] * ostride] = stage[q];
  }
  __syncthreads();
}

// One pass of the record sort over a block of RTILE points.  FIRST: the
// records are made from the coordinates (and values) in input order;
// otherwise they are read from the previous pass's arrays.
template <typename T, int NDIM, bool TYPE1, bool FIRST>
__global__ __launch_bounds__(RT) void k_rec_scatter(RecPassArgs a) {
  using T2 = typename c2t<T>::type;
  __shared__ RankShared rs;
  __shared__ uint2 stage8[RTILE];
  uint32_t *stage4 = (uint32_t *)stage8;
  const int tid = threadIdx.x, wave = tid >> 6;
  const unsigned lane = lane_id();
  const size_t M = a.M;
  const size_t blk0 = (size_t)blockIdx.x * RTILE;
  const size_t c0 = blk0 + (size_t)wave * (RTILE / NW);
  const int dsh = a.offbits + a.shift;
  // every field of this thread's items is loaded before the ranking, so
  // that all of the block's global loads are in flight together
  uint32_t word[RITEMS], dig[RITEMS];
  bool valid[RITEMS];
  T xk[NDIM][RITEMS];
  T vr[TYPE1 ? RITEMS : 1], vi[TYPE1 ? RITEMS : 1];
  uint32_t jv[TYPE1 ? 1 : RITEMS];
  auto load_value = [&](int r, size_t j) {
    if constexpr (TYPE1) {
      vr[r] = vi[r] = T(0);
      if (!valid[r]) return;
      if constexpr (FIRST) {
        if (a.val_f64) {
          const double2 z = ((const double2 *)a.values)[j];
          vr[r] = (T)z.x;
          vi[r] = (T)z.y;
        } else {
          const float2 z = ((const float2 *)a.values)[j];
          vr[r] = (T)z.x;
          vi[r] = (T)z.y;
        }
      } else {
        const T2 z = ((const T2 *)a.in.v)[j];
        vr[r] = z.x;
        vi[r] = z.y;
      }
    } else {
      if constexpr (FIRST)
        jv[r] = (uint32_t)j;
      else
        jv[r] = valid[r] ? a.in.j[j] : 0u;
    }
  };
  if constexpr (FIRST) {
    const int L = a.g.log2tile;
    const int tmask = (1 << L) - 1;
    double xc[RITEMS][NDIM];
#pragma unroll
    for (int r = 0; r < RITEMS; ++r) {
      const size_t j = c0 + (size_t)r * 64 + lane;
      valid[r] = j < M;
      if (valid[r]) load_point(a.g, a.coords, j, NDIM, xc[r]);
    }
#pragma unroll
    for (int r = 0; r < RITEMS; ++r)
      load_value(r, c0 + (size_t)r * 64 + lane);
#pragma unroll
    for (int r = 0; r < RITEMS; ++r) {
      uint32_t key = 0, off = 0;
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        double frac = 0.0;
        const int i0 = valid[r] ? fold_axis(a.g, d, xc[r][d], frac) : 0;
        xk[d][r] = (T)((double)(a.g.W - 1) - 2.0 * frac);
        // i0 >= 1 - nsafe, so i0 + nsafe >= 1
        const uint32_t u = (uint32_t)(i0 + a.g.nsafe);
        key = (d ? key * (uint32_t)a.g.ntiles[d] : 0u) + (u >> L);
        off |= (u & (uint32_t)tmask) << (d * L);
      }
      word[r] = valid[r] ? (key << a.offbits) | off : 0xFFFFFFFFu;
    }
  } else {
#pragma unroll
    for (int r = 0; r < RITEMS; ++r) {
      const size_t j = c0 + (size_t)r * 64 + lane;
      valid[r] = j < M;
      word[r] = valid[r] ? a.in.word[j] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
#pragma unroll
      for (int r = 0; r < RITEMS; ++r) {
        const size_t j = c0 + (size_t)r * 64 + lane;
        xk[d][r] = valid[r] ? ((const T *)a.in.x[d])[j] : T(0);
      }
#pragma unroll
    for (int r = 0; r < RITEMS; ++r)
      load_value(r, c0 + (size_t)r * 64 + lane);
  }
#pragma unroll
  for (int r = 0; r < RITEMS; ++r) dig[r] = (word[r] >> dsh) & 0xFFu;
  const size_t og = (size_t)tid * a.nblocks + blockIdx.x;
  const uint32_t gofs = a.offs[og] + a.offs_chunk[og / STILE];
  uint32_t lpos[RITEMS];
  rank_block(dig, valid, lpos, rs, gofs);
  const int nvalid = (int)min((size_t)RTILE, M - blk0);

  // words: staged, then each position's global place from its digit
#pragma unroll
  for (int r = 0; r < RITEMS; ++r)
    if (valid[r]) stage4[lpos[r]] = word[r];
  __syncthreads();
  uint32_t gpos[RITEMS];
#pragma unroll
  for (int i = 0; i < RITEMS; ++i) {
    const int q = i * RT + tid;
    gpos[i] = 0;
    if (q < nvalid) {
      const uint32_t w = stage4[q];
      gpos[i] = rs.gbase[(w >> dsh) & 0xFFu] + (uint32_t)q;
      a.out.word[gpos[i]] = w;
      if (a.tb_end) {
        // last pass: equal tiles are contiguous here and globally, so the
        // block's first / last point of a tile bound the tile's range
        // (start kept complemented, so both arrays start at zero)
        const uint32_t t = w >> a.offbits;
        if (q == 0 || (stage4[q - 1] >> a.offbits) != t)
          atomicMax(&a.tb_start_c[t], ~gpos[i]);
        if (q == nvalid - 1 || (stage4[q + 1] >> a.offbits) != t)
          atomicMax(&a.tb_end[t], gpos[i] + 1);
      }
    }
  }
  __syncthreads();

  // kernel arguments, one axis at a time
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    stage_move<T>((T *)stage8, xk[d], lpos, valid, gpos, nvalid,
                  (T *)a.out.x[d], 1);

  if constexpr (TYPE1) {
    // the value, in the compute type; 16-byte values move as halves
    if constexpr (sizeof(T) == 4) {
      uint2 v[RITEMS];
#pragma unroll
      for (int r = 0; r < RITEMS; ++r)
        v[r] = make_uint2(__float_as_uint((float)vr[r]),
                          __float_as_uint((float)vi[r]));
      stage_move<uint2>(stage8, v, lpos, valid, gpos, nvalid,
                        (uint2 *)a.out.v, 1);
    } else {
      stage_move<T>((T *)stage8, vr, lpos, valid, gpos, nvalid,
                    (T *)a.out.v, 2);
      stage_move<T>((T *)stage8, vi, lpos, valid, gpos, nvalid,
                    (T *)a.out.v + 1, 2);
    }
  } else {
    stage_move<uint32_t>(stage4, jv, lpos, valid, gpos, nvalid, a.out.j, 1);
  }
}

template <typename T, int NDIM, bool TYPE1>
void launch_pass(bool first, const RecPassArgs &a, hipStream_t s) {
  const dim3 grid((unsigned)a.nblocks), block(RT);
  if (first)
    hipLaunchKernelGGL((k_rec_scatter<T, NDIM, TYPE1, true>), grid, block, 0,
                       s, a);
  else
    hipLaunchKernelGGL((k_rec_scatter<T, NDIM, TYPE1, false>), grid, block,
                       0, s, a);
}

template <typename T, int NDIM>
void launch_pass_t(bool type1, bool first, const RecPassArgs &a,
                   hipStream_t s) {
  if (type1)
    launch_pass<T, NDIM, true>(first, a, s);
  else
    launch_pass<T, NDIM, false>(first, a, s);
}

}  // namespace

size_t rec_bytes_per_point(int ndim, bool fp64, bool type1) {
  const size_t t = fp64 ? 8 : 4;
  return 4 + ndim * t + (type1 ? 2 * t : 4);
}

RecBufs rec_carve(void *p, size_t M, int ndim, bool fp64, bool type1) {
  const size_t t = fp64 ? 8 : 4;
  char *c = (char *)p;
  RecBufs r{};
  // the value array first: 16-byte elements stay 16-byte aligned
  if (type1) {
    r.v = c;
    c += round256(M * 2 * t);
  }
  for (int d = 0; d < ndim; ++d) {
    r.x[d] = c;
    c += round256(M * t);
  }
  r.word = (uint32_t *)c;
  c += round256(M * 4);
  if (!type1) {
    r.j = (uint32_t *)c;
    c += round256(M * 4);
  }
  return r;
}

size_t rec_carve_bytes(size_t M, int ndim, bool fp64, bool type1) {
  const size_t t = fp64 ? 8 : 4;
  return (type1 ? round256(M * 2 * t) : round256(M * 4)) +
         ndim * round256(M * t) + round256(M * 4);
}

size_t rec_sort_workspace(size_t M) {
  const size_t nblocks = (M + RTILE - 1) / RTILE;
  const size_t hist = (size_t)RB * nblocks;
  return 2 * hist * sizeof(uint32_t) + scan_workspace(hist) + 256;
}

void rec_sort(const Geo &g, int offbits, int keybits, bool fp64, bool type1,
              const void *coords, const void *values, int val_f64, size_t M,
              const RecBufs &out, const RecBufs &tmp, void *work,
              size_t work_bytes, uint32_t *tb_start_c, uint32_t *tb_end,
              hipStream_t s) {
  if (!M) return;
  DMX_ASSERT(work_bytes >= rec_sort_workspace(M),
             "record sort workspace too small");
  DMX_ASSERT(offbits + keybits <= 32, "internal: record word overflow");
  DMX_ASSERT(M < 0xFFFFFFFFull, "too many points for one transform");
  const size_t nblocks = (M + RTILE - 1) / RTILE;
  const size_t hist_n = (size_t)RB * nblocks;
  uint32_t *hist = (uint32_t *)work;
  uint32_t *offs = hist + hist_n;
  uint32_t *bsum = offs + hist_n;
  const int npass = keybits <= 0 ? 1 : (keybits + 7) / 8;
  // the first pass's digit histogram comes from a fold of the coordinates
  // (tile keys only, nothing written but the histogram)
  launch_fold_keys(g, coords, M, nullptr, hist, s);
  const RecBufs *src = nullptr;
  for (int p = 0; p < npass; ++p) {
    const bool last = p == npass - 1;
    // ping-pong so that the last pass writes `out`
    const RecBufs *dst = ((npass - 1 - p) % 2 == 0) ? &out : &tmp;
    if (p > 0)
      radix_pass_offsets(src->word, M, offbits + 8 * p, hist, false, offs,
                         bsum, s);
    else
      radix_pass_offsets(nullptr, M, 0, hist, true, offs, bsum, s);
    RecPassArgs a{};
    a.g = g;
    a.M = M;
    a.offbits = offbits;
    a.shift = 8 * p;
    a.coords = coords;
    a.values = values;
    a.val_f64 = val_f64;
    if (src) a.in = *src;
    a.out = *dst;
    a.offs = offs;
    a.offs_chunk = bsum;
    a.nblocks = (int)nblocks;
    a.tb_start_c = last ? tb_start_c : nullptr;
    a.tb_end = last ? tb_end : nullptr;
    switch (g.ndim) {
      case 1:
        fp64 ? launch_pass_t<double, 1>(type1, p == 0, a, s)
             : launch_pass_t<float, 1>(type1, p == 0, a, s);
        break;
      case 2:
        fp64 ? launch_pass_t<double, 2>(type1, p == 0, a, s)
             : launch_pass_t<float, 2>(type1, p == 0, a, s);
        break;
      case 3:
        fp64 ? launch_pass_t<double, 3>(type1, p == 0, a, s)
             : launch_pass_t<float, 3>(type1, p == 0, a, s);
        break;
      default:
        throw Error("transform must be 1D/2D/3D");
    }
    DMX_HIP(hipGetLastError());
    src = dst;
  }
}

}  // namespace dmx
