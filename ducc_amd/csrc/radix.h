// radix.h -- block-level pieces of the stable LSD radix sort shared by the
// index sort (sort.hip) and the record sort (records.hip): block geometry,
// wave ballots that rank equal 8-bit digits, the block prefix sum, and the
// per-block ranking step.
#pragma once
#include "dmx.h"

namespace dmx {
namespace radix {

constexpr int RB = 256;         // buckets per pass
constexpr int RT = 256;         // threads per block
constexpr int RITEMS = 16;      // keys per thread
constexpr int RTILE = RT * RITEMS;
constexpr int NW = RT / 64;     // waves per block

// scan blocks
constexpr int ST = 256, SITEMS = 8, STILE = ST * SITEMS;

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

// exclusive prefix of one value per thread over a 256-thread block; total
// receives the block sum.  Inclusive scan inside each wave by lane shifts,
// then the four wave totals through LDS: two barriers (a Hillis-Steele
// scan in LDS takes seventeen).
static_assert(ST % 64 == 0 && ST / 64 <= 16 && RT == ST,
              "block scan: whole waves, the ranking's block size");
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v,
                                                         uint32_t *sh,
                                                         uint32_t &total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, (unsigned)o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < ST / 64; ++i) {
    const uint32_t s = sh[i];
    before += i < w ? s : 0u;
    tot += s;
  }
  total = tot;
  // sh is reused by the next call
  __syncthreads();
  return before + x - v;
}

// LDS of the ranking step
struct RankShared {
  uint32_t cnt[NW][RB];
  uint32_t gbase[RB];
  uint32_t scan_sh[RB];
};

// Stable rank of the block's keys by digit.  Wave w holds the keys
// [w*1024, (w+1)*1024) of the block, item r of lane l being key
// w*1024 + r*64 + l.  On return lpos[r] is the item's position in the
// block's stably sorted order and rs.gbase[d] + lpos is its global position
// (gofs: this block's global start of digit threadIdx.x, from the scanned
// digit-major table).  Ends with a block barrier.
__device__ __forceinline__ void rank_block(const uint32_t (&dig)[RITEMS],
                                           const bool (&valid)[RITEMS],
                                           uint32_t (&lpos)[RITEMS],
                                           RankShared &rs, uint32_t gofs) {
  const int tid = threadIdx.x, wave = tid >> 6;
  const unsigned lane = lane_id();
#pragma unroll
  for (int w = 0; w < NW; ++w) rs.cnt[w][tid] = 0;
  __syncthreads();
  const unsigned long long below_mask =
      lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t rank[RITEMS];
#pragma unroll
  for (int r = 0; r < RITEMS; ++r) {
    const uint32_t d = dig[r];
    const unsigned long long peers = digit_peers(d, valid[r]);
    const unsigned long long below = peers & below_mask;
    const uint32_t old = rs.cnt[wave][d];
    // every lane has read the counter before the group leader advances it
    __builtin_amdgcn_wave_barrier();
    if (valid[r] && below == 0)
      rs.cnt[wave][d] = old + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rank[r] = old + (uint32_t)__popcll(below);
  }
  __syncthreads();
  {
    // thread tid owns digit tid: wave prefix, block-local start, global base
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t c = rs.cnt[w][tid];
      rs.cnt[w][tid] = run;
      run += c;
    }
    uint32_t tot;
    const uint32_t start = block_exclusive_scan(run, rs.scan_sh, tot);
#pragma unroll
    for (int w = 0; w < NW; ++w) rs.cnt[w][tid] += start;
    rs.gbase[tid] = gofs - start;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RITEMS; ++r) lpos[r] = rs.cnt[wave][dig[r]] + rank[r];
  __syncthreads();
}

}  // namespace radix
}  // namespace dmx
