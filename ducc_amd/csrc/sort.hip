// sort.hip -- folding points into tiles and the stable counting sort of the
// index build (role of nufft.h:892-907, 1364-1381, 1862-1892 and
// bucket_sort.h:170-354).  The order produced is the stable ascending order
// of the tile keys: ties keep input order, which is the unique permutation a
// stable sort gives and the one ducc0's bucket_sort2 returns.
//
// Layout: keys / indices are uint32 arrays of M in HBM.  The sort is an LSD
// radix sort with 8-bit digits: per pass one histogram kernel (one LDS
// histogram per 4096-key block), one exclusive scan of the digit-major
// block table, and one scatter kernel that ranks equal digits inside a
// wavefront with ballots (stable) and across the block's four wavefronts in
// wave order, so every pass is stable; the ranked block is staged in LDS and
// written out in digit runs.
#include "radix.h"

namespace dmx {

namespace {
using namespace radix;

// Tile key per point (nufft.h:1364-1381 role); one 4096-point block per
// workgroup, so that the first sort pass's per-block digit histogram comes
// out of the same sweep.  Each lane loads a batch of FB points before it
// folds them, so the loads of a batch are in flight together.
constexpr int FB = 16;
template <int NDIM>
__global__ __launch_bounds__(RT) void k_fold_keys(Geo g, const void *coords,
                                                  size_t M, uint32_t *keys,
                                                  uint32_t *hist0,
                                                  int nblocks) {
  __shared__ uint32_t h[RB];
  const int tid = threadIdx.x;
  if (hist0) h[tid] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * RTILE;
  for (int i0 = 0; i0 < RITEMS; i0 += FB) {
    double x[FB][NDIM];
#pragma unroll
    for (int q = 0; q < FB; ++q) {
      const size_t j = base + (size_t)(i0 + q) * RT + tid;
      if (j < M) load_point(g, coords, j, NDIM, x[q]);
    }
#pragma unroll
    for (int q = 0; q < FB; ++q) {
      const size_t j = base + (size_t)(i0 + q) * RT + tid;
      if (j >= M) continue;
      uint32_t t[NDIM];
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        double fr;
        const int c0 = fold_axis(g, d, x[q][d], fr);
        t[d] = (uint32_t)((c0 + g.nsafe) >> g.log2tile);
      }
      uint32_t key = t[0];
#pragma unroll
      for (int d = 1; d < NDIM; ++d) key = key * (uint32_t)g.ntiles[d] + t[d];
      if (keys) keys[j] = key;
      if (hist0) atomicAdd(&h[key & 0xFFu], 1u);
    }
  }
  if (hist0) {
    __syncthreads();
    hist0[(size_t)tid * nblocks + blockIdx.x] = h[tid];
  }
}

__global__ __launch_bounds__(RT) void k_radix_hist(const uint32_t *keys,
                                                   size_t M, int shift,
                                                   uint32_t *hist,
                                                   int nblocks) {
  __shared__ uint32_t h[RB];
  h[threadIdx.x] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * RTILE;
  // all of a lane's key loads in flight before its LDS atomics
  uint32_t kv[RITEMS];
#pragma unroll
  for (int i = 0; i < RITEMS; ++i) {
    size_t j = base + (size_t)i * RT + threadIdx.x;
    kv[i] = j < M ? keys[j] : 0u;
  }
#pragma unroll
  for (int i = 0; i < RITEMS; ++i) {
    size_t j = base + (size_t)i * RT + threadIdx.x;
    if (j < M) atomicAdd(&h[(kv[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one 4096-key block, staged through LDS so that the
// global writes go out in digit runs (consecutive lanes -> consecutive
// addresses):
//  (1) wave w ranks its 1024 keys [w*1024, (w+1)*1024) in order, 64 at a
//      time: ballot peers give the rank inside the 64, a per-wave per-digit
//      counter in LDS (touched only by the group leader) the rank before;
//  (2) one prefix over (digit, wave) gives each key its position in the
//      block's locally sorted order, and each digit its global base;
//  (3) keys, then indices, are written to LDS in local order and read back
//      by position, so lane t of round i stores element i*256+t of the
//      sorted block at global base[digit] + (position - local start).
__global__ __launch_bounds__(RT) void k_radix_scatter(
    const uint32_t *keys_in, const uint32_t *idx_in, size_t M, int shift,
    const uint32_t *offs, const uint32_t *offs_chunk, int nblocks,
    uint32_t *keys_out, uint32_t *idx_out, uint32_t *tb_start_c,
    uint32_t *tb_end) {
  __shared__ RankShared rs;
  __shared__ uint32_t stage[RTILE];
  const int tid = threadIdx.x, wave = tid >> 6;
  const unsigned lane = lane_id();
  const size_t blk0 = (size_t)blockIdx.x * RTILE;
  const size_t c0 = blk0 + (size_t)wave * (RTILE / NW);
  uint32_t key[RITEMS], ival[RITEMS], dig[RITEMS];
  bool valid[RITEMS];
#pragma unroll
  for (int r = 0; r < RITEMS; ++r) {
    const size_t j = c0 + (size_t)r * 64 + lane;
    valid[r] = j < M;
    key[r] = valid[r] ? keys_in[j] : 0xFFFFFFFFu;
  }
  // this block's global base of digit tid (chunk scan + scanned chunk
  // sums), and the indices that travel with the keys: loaded now, so that
  // their latency overlaps the ranking
  const size_t og = (size_t)tid * nblocks + blockIdx.x;
  const uint32_t gofs = offs[og] + offs_chunk[og / STILE];
#pragma unroll
  for (int r = 0; r < RITEMS; ++r) {
    const size_t j = c0 + (size_t)r * 64 + lane;
    ival[r] = idx_in ? (valid[r] ? idx_in[j] : 0u) : (uint32_t)j;
    dig[r] = (key[r] >> shift) & 0xFFu;
  }
  uint32_t lpos[RITEMS];
  rank_block(dig, valid, lpos, rs, gofs);
  const uint32_t *gbase = rs.gbase;
  const int nvalid = (int)min((size_t)RTILE, M - blk0);
#pragma unroll
  for (int r = 0; r < RITEMS; ++r)
    if (valid[r]) stage[lpos[r]] = key[r];
  __syncthreads();
  uint32_t gpos[RITEMS];
#pragma unroll
  for (int i = 0; i < RITEMS; ++i) {
    const int q = i * RT + tid;
    gpos[i] = 0;
    if (q < nvalid) {
      const uint32_t k = stage[q];
      gpos[i] = gbase[(k >> shift) & 0xFFu] + (uint32_t)q;
      if (keys_out) keys_out[gpos[i]] = k;
      if (tb_end) {
        // final pass: equal keys are contiguous here and globally, so the
        // block's first / last element of a key bound the tile's range
        // (start kept complemented, so both arrays start at zero)
        if (q == 0 || stage[q - 1] != k) atomicMax(&tb_start_c[k], ~gpos[i]);
        if (q == nvalid - 1 || stage[q + 1] != k)
          atomicMax(&tb_end[k], gpos[i] + 1);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RITEMS; ++r) {
    const size_t j = c0 + (size_t)r * 64 + lane;
    if (j < M) stage[lpos[r]] = ival[r];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RITEMS; ++i) {
    const int q = i * RT + tid;
    if (q < nvalid) idx_out[gpos[i]] = stage[q];
  }
}

__global__ __launch_bounds__(ST) void k_scan_blocks(const uint32_t *in,
                                                    uint32_t *out, size_t n,
                                                    uint32_t *bsum) {
  __shared__ uint32_t sh[ST];
  size_t base = (size_t)blockIdx.x * STILE + (size_t)threadIdx.x * SITEMS;
  uint32_t v[SITEMS], s = 0;
#pragma unroll
  for (int i = 0; i < SITEMS; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t ex = block_exclusive_scan(s, sh, tot);
#pragma unroll
  for (int i = 0; i < SITEMS; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// one workgroup scans the whole array (in may equal out); used for the
// block sums of the multi-block scan and for short arrays in one launch
__global__ __launch_bounds__(ST) void k_scan_single(const uint32_t *in,
                                                    uint32_t *out, size_t n,
                                                    uint32_t *total) {
  __shared__ uint32_t sh[ST];
  uint32_t carry = 0;
  for (size_t c = 0; c < n; c += STILE) {
    size_t base = c + (size_t)threadIdx.x * SITEMS;
    uint32_t v[SITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SITEMS; ++i) {
      v[i] = (base + i < n) ? in[base + i] : 0u;
      s += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, sh, tot) + carry;
#pragma unroll
    for (int i = 0; i < SITEMS; ++i) {
      if (base + i < n) out[base + i] = ex;
      ex += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(ST) void k_scan_add(uint32_t *out, size_t n,
                                                 const uint32_t *bsum) {
  size_t base = (size_t)blockIdx.x * STILE;
  uint32_t a = bsum[blockIdx.x];
  for (int i = threadIdx.x; i < STILE; i += ST)
    if (base + i < n) out[base + i] += a;
}

// ------------------------------------------------------------ subproblems
// tile [start, end) from the sort's final pass (start stored complemented;
// empty tiles have end 0), points per tile and subproblems per tile (runs
// of at most MAXSUB points)
__global__ void k_counts_nsub(uint32_t *tstart, const uint32_t *tend,
                              size_t ntiles, uint32_t *counts,
                              uint32_t *nsub) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t e = tend[t];
  const uint32_t st = e ? ~tstart[t] : 0u;
  tstart[t] = st;
  const uint32_t c = e - st;
  counts[t] = c;
  nsub[t] = (c + MAXSUB - 1) / MAXSUB;
}

__global__ void k_fill_subs(const uint32_t *counts, const uint32_t *tstart,
                            const uint32_t *nsub, const uint32_t *sstart,
                            size_t ntiles, Sub *subs) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t n = nsub[t], s0 = sstart[t], ts = tstart[t], te = ts + counts[t];
  for (uint32_t s = 0; s < n; ++s) {
    Sub sb;
    sb.tile = (uint32_t)t;
    sb.start = ts + s * MAXSUB;
    sb.end = min(te, sb.start + (uint32_t)MAXSUB);
    sb.rank = s;
    subs[s0 + s] = sb;
  }
}

// The three steps above in one workgroup, for tile tables of at most 4096
// tiles (one launch instead of four; longer tables take the launches above,
// which measured faster from about 10^4 tiles on): per 4096-tile chunk, start / count / subproblem
// count of each tile, a block prefix of the subproblem counts (with the
// carry of the chunks before) and the subproblems themselves.  The scan
// wants TS_ITEMS consecutive tiles per lane, global memory wants
// consecutive tiles per wave: the chunk's tables pass through LDS (one pad
// word per TS_ITEMS entries keeps both access patterns free of bank
// conflicts), so every global load and store of the tables is coalesced.
constexpr int TS_ITEMS = 16;
constexpr int TS_CH = ST * TS_ITEMS, TS_CHP = TS_CH + TS_CH / TS_ITEMS;
__device__ __forceinline__ int ts_slot(int q) { return q + q / TS_ITEMS; }
__global__ __launch_bounds__(ST) void k_tile_subs_one(
    uint32_t *tstart, const uint32_t *tend, size_t ntiles, uint32_t *counts,
    uint32_t *nsub, uint32_t *sstart, Sub *subs, uint32_t *nsub_total) {
  __shared__ uint32_t sh[ST];
  __shared__ uint32_t la[TS_CHP], lb[TS_CHP];
  const int tid = threadIdx.x;
  uint32_t carry = 0;
  for (size_t c = 0; c < ntiles; c += TS_CH) {
    const int nc = (int)min((size_t)TS_CH, ntiles - c);
    // coalesced loads, all in flight (clamped indices past the chunk's end
    // read its last tile; those entries are never used)
    uint32_t ev[TS_ITEMS], sv[TS_ITEMS];
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      const size_t t = c + (size_t)min(k * ST + tid, nc - 1);
      ev[k] = tend[t];
      sv[k] = tstart[t];
    }
    __syncthreads();  // the previous chunk's LDS readers are done
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      la[ts_slot(k * ST + tid)] = ev[k];
      lb[ts_slot(k * ST + tid)] = sv[k];
    }
    __syncthreads();
    const int q0 = tid * TS_ITEMS;
    uint32_t st[TS_ITEMS], cn[TS_ITEMS], ns[TS_ITEMS], sum = 0;
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      st[k] = cn[k] = ns[k] = 0;
      if (q0 + k < nc) {
        const uint32_t e = la[ts_slot(q0 + k)];
        st[k] = e ? ~lb[ts_slot(q0 + k)] : 0u;
        cn[k] = e - st[k];
        ns[k] = (cn[k] + MAXSUB - 1) / MAXSUB;
        sum += ns[k];
      }
    }
    uint32_t tot;
    const uint32_t ex0 = block_exclusive_scan(sum, sh, tot) + carry;
    uint32_t ex = ex0;
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      if (q0 + k >= nc) continue;
      for (uint32_t q = 0; q < ns[k]; ++q) {
        Sub sb;
        sb.tile = (uint32_t)(c + q0 + k);
        sb.start = st[k] + q * MAXSUB;
        sb.end = min(st[k] + cn[k], sb.start + (uint32_t)MAXSUB);
        sb.rank = q;
        subs[ex + q] = sb;
      }
      ex += ns[k];
    }
    // starts and counts, then subproblem counts and first subproblems,
    // back through LDS to coalesced stores
    __syncthreads();  // all lanes have read la / lb
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      la[ts_slot(q0 + k)] = st[k];
      lb[ts_slot(q0 + k)] = cn[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      const int q = k * ST + tid;
      if (q < nc) {
        tstart[c + q] = la[ts_slot(q)];
        counts[c + q] = lb[ts_slot(q)];
      }
    }
    __syncthreads();
    ex = ex0;
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      la[ts_slot(q0 + k)] = ns[k];
      lb[ts_slot(q0 + k)] = ex;
      ex += ns[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TS_ITEMS; ++k) {
      const int q = k * ST + tid;
      if (q < nc) {
        nsub[c + q] = la[ts_slot(q)];
        sstart[c + q] = lb[ts_slot(q)];
      }
    }
    carry += tot;
  }
  if (tid == 0) *nsub_total = carry;
}

template <typename T>
__global__ void k_gather_coords(const T *c, int ndim, const uint32_t *idx,
                                size_t M, T *out) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  size_t s = idx[j];
  for (int d = 0; d < ndim; ++d) out[j * ndim + d] = c[s * ndim + d];
}

__global__ void k_zero16(uint4 *p, size_t n16) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}
__global__ void k_zero1(unsigned char *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

inline unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
}  // namespace

void launch_fold_keys(const Geo &g, const void *coords, size_t M,
                      uint32_t *keys, uint32_t *hist0, hipStream_t s) {
  if (!M) return;
  const size_t nblocks = (M + RTILE - 1) / RTILE;
  const dim3 grid((unsigned)nblocks), block(RT);
  if (g.ndim == 1)
    hipLaunchKernelGGL(k_fold_keys<1>, grid, block, 0, s, g, coords, M, keys,
                       hist0, (int)nblocks);
  else if (g.ndim == 2)
    hipLaunchKernelGGL(k_fold_keys<2>, grid, block, 0, s, g, coords, M, keys,
                       hist0, (int)nblocks);
  else
    hipLaunchKernelGGL(k_fold_keys<3>, grid, block, 0, s, g, coords, M, keys,
                       hist0, (int)nblocks);
  DMX_HIP(hipGetLastError());
}

size_t scan_workspace(size_t n) {
  size_t nb = (n + STILE - 1) / STILE;
  return (nb + 64) * sizeof(uint32_t);
}

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n,
                        uint32_t *total, void *work, hipStream_t s) {
  if (!n) return;
  size_t nb = (n + STILE - 1) / STILE;
  // one workgroup walks the chunks one after another (a load latency each):
  // only for one or two chunks; longer arrays take the three-launch scan
  if (nb <= 2) {
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(ST), 0, s, in, out, n,
                       total);
    DMX_HIP(hipGetLastError());
    return;
  }
  uint32_t *bsum = (uint32_t *)work;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(ST), 0, s, in,
                     out, n, bsum);
  hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(ST), 0, s,
                     (const uint32_t *)bsum, bsum, nb, total);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(ST), 0, s, out, n,
                     (const uint32_t *)bsum);
  DMX_HIP(hipGetLastError());
}

size_t radix_sort_workspace(size_t M) {
  size_t nblocks = (M + RTILE - 1) / RTILE;
  size_t hist = (size_t)RB * nblocks;
  return 4 * M * sizeof(uint32_t) + 2 * hist * sizeof(uint32_t) +
         scan_workspace(hist) + 256;
}

uint32_t *radix_sort_hist0(void *work, size_t M) {
  return (uint32_t *)work + 4 * M;
}

void radix_sort_pairs(const uint32_t *keys_in, size_t M, int keybits,
                      uint32_t *idx_out, uint32_t *keys_out, void *work,
                      size_t work_bytes, hipStream_t s, bool hist0_ready,
                      uint32_t *tb_start_c, uint32_t *tb_end) {
  if (!M) return;
  DMX_ASSERT(work_bytes >= radix_sort_workspace(M), "sort workspace too small");
  size_t nblocks = (M + RTILE - 1) / RTILE;
  size_t hist_n = (size_t)RB * nblocks;
  uint32_t *kbuf[2] = {(uint32_t *)work, (uint32_t *)work + M};
  uint32_t *ibuf[2] = {(uint32_t *)work + 2 * M, (uint32_t *)work + 3 * M};
  uint32_t *hist = (uint32_t *)work + 4 * M;
  uint32_t *offs = hist + hist_n;
  void *sw = offs + hist_n;
  int npass = keybits <= 0 ? 1 : (keybits + 7) / 8;
  const uint32_t *kin = keys_in;
  const uint32_t *iin = nullptr;  // identity permutation on the first pass
  for (int p = 0; p < npass; ++p) {
    bool last = p == npass - 1;
    uint32_t *kout = last ? keys_out : kbuf[p & 1];
    uint32_t *iout = last ? idx_out : ibuf[p & 1];
    uint32_t *bsum = (uint32_t *)sw;
    radix_pass_offsets(kin, M, 8 * p, hist, p == 0 && hist0_ready, offs,
                       bsum, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nblocks), dim3(RT), 0,
                       s, kin, iin, M, 8 * p, (const uint32_t *)offs,
                       (const uint32_t *)bsum,
                       (int)nblocks, kout, iout, last ? tb_start_c : nullptr,
                       last ? tb_end : nullptr);
    DMX_HIP(hipGetLastError());
    kin = kout;
    iin = iout;
  }
}

void radix_pass_offsets(const uint32_t *keys, size_t M, int shift,
                        uint32_t *hist, bool hist_ready, uint32_t *offs,
                        uint32_t *bsum, hipStream_t s) {
  const size_t nblocks = (M + RTILE - 1) / RTILE;
  const size_t hist_n = (size_t)RB * nblocks;
  if (!hist_ready)
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nblocks), dim3(RT), 0, s,
                       keys, M, shift, hist, (int)nblocks);
  // digit-major offsets: per-chunk scans and the scan of the chunk sums;
  // the scatter adds the two
  const size_t nch = (hist_n + STILE - 1) / STILE;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nch), dim3(ST), 0, s,
                     (const uint32_t *)hist, offs, hist_n, bsum);
  hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(ST), 0, s,
                     (const uint32_t *)bsum, bsum, nch, nullptr);
  DMX_HIP(hipGetLastError());
}

void launch_tile_subs(size_t ntiles, uint32_t *counts, uint32_t *tstart,
                      const uint32_t *tend, uint32_t *nsub,
                      uint32_t *sub_start, Sub *subs, uint32_t *nsub_total,
                      void *scanwork, hipStream_t s) {
  DMX_ASSERT(ntiles > 0, "internal: no tiles");
  if (ntiles <= (size_t)ST * TS_ITEMS) {
    hipLaunchKernelGGL(k_tile_subs_one, dim3(1), dim3(ST), 0, s, tstart, tend,
                       ntiles, counts, nsub, sub_start, subs, nsub_total);
    DMX_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_counts_nsub, dim3(nblk(ntiles, 256)), dim3(256), 0, s,
                     tstart, tend, ntiles, counts, nsub);
  exclusive_scan_u32(nsub, sub_start, ntiles, nsub_total, scanwork, s);
  hipLaunchKernelGGL(k_fill_subs, dim3(nblk(ntiles, 256)), dim3(256), 0, s,
                     counts, tstart, (const uint32_t *)nsub,
                     (const uint32_t *)sub_start, ntiles, subs);
  DMX_HIP(hipGetLastError());
}

void launch_gather_coords(const void *coords, int coord_f32, int ndim,
                          const uint32_t *idx, size_t M, void *out,
                          hipStream_t s) {
  if (!M) return;
  if (coord_f32)
    hipLaunchKernelGGL(k_gather_coords<float>, dim3(nblk(M, 256)), dim3(256),
                       0, s, (const float *)coords, ndim, idx, M,
                       (float *)out);
  else
    hipLaunchKernelGGL(k_gather_coords<double>, dim3(nblk(M, 256)), dim3(256),
                       0, s, (const double *)coords, ndim, idx, M,
                       (double *)out);
  DMX_HIP(hipGetLastError());
}

void launch_zero(void *p, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  size_t n16 = bytes / 16, rest = bytes - n16 * 16;
  if (n16) {
    unsigned nb = (unsigned)std::min<size_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zero16, dim3(nb), dim3(256), 0, s, (uint4 *)p, n16);
  }
  if (rest)
    hipLaunchKernelGGL(k_zero1, dim3(1), dim3(256), 0, s,
                       (unsigned char *)p + n16 * 16, rest);
  DMX_HIP(hipGetLastError());
}

}  // namespace dmx
