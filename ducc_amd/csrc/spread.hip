// spread.hip -- instantiation and launch of the type-1 spreading kernels.
#include "spread_impl.h"

namespace dmx {

namespace {
constexpr unsigned SPREAD_MAX_BLOCKS = 8192;

template <typename T, int NDIM, int W>
void go(const SpreadArgs &a, hipStream_t s) {
  static_assert(sizeof(SpreadStash<T, NDIM>) <=
                    spread_lds_bytes(0, NDIM, sizeof(T) == 8),
                "LDS staging size");
  static_assert(SpreadStash<T, NDIM>::SC % 256 == 0 &&
                    SpreadStash<T, NDIM>::SC <= 65536,
                "staging chunk: whole rounds of 256 threads, 16-bit list");
  unsigned nb = (unsigned)std::min<size_t>(a.nsub_max, SPREAD_MAX_BLOCKS);
  if (nb == 0) return;
  if (a.use_lds) {
    size_t cells = a.lds_side;
    for (int d = 1; d < NDIM; ++d) cells *= a.lds_side;
    if constexpr (sizeof(T) == 4) {
      // packed fixed point: one 8-byte cell, the whole subproblem staged
      const size_t shf =
          fix_region_bytes(cells) + sizeof(SpreadStash<T, NDIM, true>);
      if (a.fixacc && shf <= LDS_MAX_BYTES) {
        allow_big_lds(k_spread<T, NDIM, W, true, true>, shf);
        hipLaunchKernelGGL((k_spread<T, NDIM, W, true, true>), dim3(nb),
                           dim3(256), shf, s, a);
        DMX_HIP(hipGetLastError());
        return;
      }
    }
    size_t sh = spread_lds_bytes(cells, NDIM, sizeof(T) == 8);
    allow_big_lds(k_spread<T, NDIM, W, true>, sh);
    hipLaunchKernelGGL((k_spread<T, NDIM, W, true>), dim3(nb), dim3(256), sh,
                       s, a);
  } else {
    if constexpr (NDIM == 3)
      hipLaunchKernelGGL((k_spread<T, NDIM, W, false>), dim3(nb), dim3(256), 0,
                         s, a);
    else
      throw Error("internal: no-LDS spreading only for 3D");
  }
  DMX_HIP(hipGetLastError());
}

template <typename T, int NDIM>
void by_w(const SpreadArgs &a, hipStream_t s) {
  switch (a.g.W) {
    case 4: return go<T, NDIM, 4>(a, s);
    case 5: return go<T, NDIM, 5>(a, s);
    case 6: return go<T, NDIM, 6>(a, s);
    case 7: return go<T, NDIM, 7>(a, s);
    case 8: return go<T, NDIM, 8>(a, s);
    default: break;
  }
  if constexpr (sizeof(T) == 8) {
    switch (a.g.W) {
      case 9: return go<T, NDIM, 9>(a, s);
      case 10: return go<T, NDIM, 10>(a, s);
      case 11: return go<T, NDIM, 11>(a, s);
      case 12: return go<T, NDIM, 12>(a, s);
      case 13: return go<T, NDIM, 13>(a, s);
      case 14: return go<T, NDIM, 14>(a, s);
      case 15: return go<T, NDIM, 15>(a, s);
      case 16: return go<T, NDIM, 16>(a, s);
      default: break;
    }
  }
  throw Error("requested support out of range");
}
}  // namespace

void launch_spread(bool fp64, const SpreadArgs &a, hipStream_t s) {
  switch (a.g.ndim) {
    case 1: return fp64 ? by_w<double, 1>(a, s) : by_w<float, 1>(a, s);
    case 2: return fp64 ? by_w<double, 2>(a, s) : by_w<float, 2>(a, s);
    case 3: return fp64 ? by_w<double, 3>(a, s) : by_w<float, 3>(a, s);
  }
  throw Error("transform must be 1D/2D/3D");
}

}  // namespace dmx
