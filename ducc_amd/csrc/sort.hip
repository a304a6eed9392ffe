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
// wave order, so every pass is stable.
#include "dmx.h"

namespace dmx {

namespace {
constexpr int RB = 256;         // buckets per pass
constexpr int RT = 256;         // threads per block
constexpr int RITEMS = 16;      // keys per thread
constexpr int RTILE = RT * RITEMS;

__global__ __launch_bounds__(256) void k_fold_keys(Geo g, const void *coords,
                                                   size_t M, uint32_t *keys,
                                                   uint32_t *counts) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  uint32_t t[3] = {0, 0, 0};
  for (int d = 0; d < g.ndim; ++d) {
    double fr;
    int i0 = fold_axis(g, d, load_coord(coords, j, g.ndim, d, g.coord_f32), fr);
    t[d] = (uint32_t)((i0 + g.nsafe) >> g.log2tile);
  }
  uint32_t key;
  if (g.ndim == 1)
    key = t[0];
  else if (g.ndim == 2)
    key = t[0] * (uint32_t)g.ntiles[1] + t[1];
  else
    key = (t[0] * (uint32_t)g.ntiles[1] + t[1]) * (uint32_t)g.ntiles[2] + t[2];
  keys[j] = key;
  if (counts) atomicAdd(&counts[key], 1u);
}

__global__ __launch_bounds__(RT) void k_radix_hist(const uint32_t *keys,
                                                   size_t M, int shift,
                                                   uint32_t *hist,
                                                   int nblocks) {
  __shared__ uint32_t h[RB];
  h[threadIdx.x] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * RTILE;
#pragma unroll 4
  for (int i = 0; i < RITEMS; ++i) {
    size_t j = base + (size_t)i * RT + threadIdx.x;
    if (j < M) atomicAdd(&h[(keys[j] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__device__ __forceinline__ unsigned lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// lanes of the wave holding the same 8-bit digit (ballot per bit)
__device__ __forceinline__ unsigned long long digit_peers(uint32_t d,
                                                          bool valid) {
  unsigned long long peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int bit = (d >> b) & 1;
    const unsigned long long bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  return valid ? peers : 0ull;
}

// Stable scatter of one 4096-key block.  Wave w owns keys
// [w*1024, (w+1)*1024) of the block: (1) each wave counts its digits,
// (2) one prefix over the four waves gives every wave its output base per
// digit, (3) each wave ranks its keys in order (ballot peers + a running
// per-digit counter that only its own lanes touch), with no further
// workgroup barrier.
__global__ __launch_bounds__(RT) void k_radix_scatter(
    const uint32_t *keys_in, const uint32_t *idx_in, size_t M, int shift,
    const uint32_t *offs, int nblocks, uint32_t *keys_out,
    uint32_t *idx_out) {
  constexpr int NW = RT / 64, WCH = RTILE / NW, WR = WCH / 64;
  __shared__ uint32_t cnt[NW][RB];
  const int tid = threadIdx.x, wave = tid >> 6;
  const unsigned lane = lane_id();
#pragma unroll
  for (int w = 0; w < NW; ++w) cnt[w][tid] = 0;
  __syncthreads();
  const size_t c0 = (size_t)blockIdx.x * RTILE + (size_t)wave * WCH;
  const unsigned long long below_mask =
      lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int r = 0; r < WR; ++r) {
    const size_t j = c0 + (size_t)r * 64 + lane;
    const bool valid = j < M;
    const uint32_t d = valid ? (keys_in[j] >> shift) & 0xFFu : 0u;
    const unsigned long long peers = digit_peers(d, valid);
    if (valid && (peers & below_mask) == 0)
      cnt[wave][d] += (uint32_t)__popcll(peers);
  }
  __syncthreads();
  {
    uint32_t run = offs[(size_t)tid * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t c = cnt[w][tid];
      cnt[w][tid] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int r = 0; r < WR; ++r) {
    const size_t j = c0 + (size_t)r * 64 + lane;
    const bool valid = j < M;
    const uint32_t k = valid ? keys_in[j] : 0u;
    const uint32_t d = (k >> shift) & 0xFFu;
    const unsigned long long peers = digit_peers(d, valid);
    const unsigned long long below = peers & below_mask;
    if (valid) {
      const uint32_t pos = cnt[wave][d] + (uint32_t)__popcll(below);
      if (keys_out) keys_out[pos] = k;
      idx_out[pos] = idx_in ? idx_in[j] : (uint32_t)j;
    }
    // every lane has read its base before the group leader advances it
    __builtin_amdgcn_wave_barrier();
    if (valid && below == 0) cnt[wave][d] += (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
}

// ------------------------------------------------------------------ scan
constexpr int ST = 256, SITEMS = 8, STILE = ST * SITEMS;

__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh,
                                         uint32_t &total) {
  // Hillis-Steele over 256 values in LDS
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < ST; o <<= 1) {
    uint32_t x = t >= o ? sh[t - o] : 0u;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  uint32_t incl = sh[t];
  total = sh[ST - 1];
  __syncthreads();
  return incl - v;
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

__global__ __launch_bounds__(ST) void k_scan_top(uint32_t *bsum, size_t nb,
                                                 uint32_t *total) {
  __shared__ uint32_t sh[ST];
  uint32_t carry = 0;
  for (size_t c = 0; c < nb; c += STILE) {
    size_t base = c + (size_t)threadIdx.x * SITEMS;
    uint32_t v[SITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SITEMS; ++i) {
      v[i] = (base + i < nb) ? bsum[base + i] : 0u;
      s += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, sh, tot) + carry;
#pragma unroll
    for (int i = 0; i < SITEMS; ++i) {
      if (base + i < nb) bsum[base + i] = ex;
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

// ------------------------------------------------------------ tile bounds
// first and one-past-last sorted position of every non-empty tile, from the
// sorted keys (no atomics: hot tiles of clustered points would serialise a
// histogram); empty tiles keep start = end = 0.
__global__ void k_tile_bounds(const uint32_t *skeys, size_t M,
                              uint32_t *tstart, uint32_t *tend) {
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  const uint32_t k = skeys[p];
  if (p == 0 || skeys[p - 1] != k) tstart[k] = (uint32_t)p;
  if (p == M - 1 || skeys[p + 1] != k) tend[k] = (uint32_t)(p + 1);
}

__global__ void k_counts(const uint32_t *tstart, const uint32_t *tend,
                         size_t ntiles, uint32_t *counts) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) counts[t] = tend[t] - tstart[t];
}

// ------------------------------------------------------------ subproblems
__global__ void k_nsub(const uint32_t *counts, size_t ntiles, uint32_t *nsub) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) nsub[t] = (counts[t] + MAXSUB - 1) / MAXSUB;
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
    sb.pad_ = 0;
    subs[s0 + s] = sb;
  }
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
                      uint32_t *keys, uint32_t *counts, hipStream_t s) {
  if (!M) return;
  hipLaunchKernelGGL(k_fold_keys, dim3(nblk(M, 256)), dim3(256), 0, s, g,
                     coords, M, keys, counts);
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
  uint32_t *bsum = (uint32_t *)work;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(ST), 0, s, in,
                     out, n, bsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(ST), 0, s, bsum, nb, total);
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

void radix_sort_pairs(const uint32_t *keys_in, size_t M, int keybits,
                      uint32_t *idx_out, uint32_t *keys_out, void *work,
                      size_t work_bytes, hipStream_t s) {
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
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nblocks), dim3(RT), 0, s,
                       kin, M, 8 * p, hist, (int)nblocks);
    exclusive_scan_u32(hist, offs, hist_n, nullptr, sw, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nblocks), dim3(RT), 0,
                       s, kin, iin, M, 8 * p, (const uint32_t *)offs,
                       (int)nblocks, kout, iout);
    DMX_HIP(hipGetLastError());
    kin = kout;
    iin = iout;
  }
}

void launch_build_subs(const uint32_t *counts, const uint32_t *tile_start,
                       size_t ntiles, uint32_t *nsub, uint32_t *sub_start,
                       Sub *subs, uint32_t *nsub_total, void *scanwork,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_nsub, dim3(nblk(ntiles, 256)), dim3(256), 0, s, counts,
                     ntiles, nsub);
  exclusive_scan_u32(nsub, sub_start, ntiles, nsub_total, scanwork, s);
  hipLaunchKernelGGL(k_fill_subs, dim3(nblk(ntiles, 256)), dim3(256), 0, s,
                     counts, tile_start, (const uint32_t *)nsub,
                     (const uint32_t *)sub_start, ntiles, subs);
  DMX_HIP(hipGetLastError());
}

void launch_tile_bounds(const uint32_t *skeys, size_t M, size_t ntiles,
                        uint32_t *tstart, uint32_t *tend, uint32_t *counts,
                        hipStream_t s) {
  launch_zero(tstart, ntiles * 4, s);
  launch_zero(tend, ntiles * 4, s);
  if (M)
    hipLaunchKernelGGL(k_tile_bounds, dim3(nblk(M, 256)), dim3(256), 0, s,
                       skeys, M, tstart, tend);
  hipLaunchKernelGGL(k_counts, dim3(nblk(ntiles, 256)), dim3(256), 0, s,
                     (const uint32_t *)tstart, (const uint32_t *)tend, ntiles,
                     counts);
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
