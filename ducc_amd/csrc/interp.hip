// interp.hip -- instantiation and launch of the type-2 interpolation kernels
// and of the kernel-derivative interpolation used for coordinate gradients.
#include "spread_impl.h"

namespace dmx {

namespace {
constexpr unsigned INTERP_MAX_BLOCKS = 8192;

template <typename T, int NDIM, int W, int MODE>
void go(const SpreadArgs &a, void *out, int out_f64, hipStream_t s) {
  unsigned nb = (unsigned)std::min<size_t>(a.nsub_max, INTERP_MAX_BLOCKS);
  if (nb == 0) return;
  if (a.use_lds) {
    size_t cells = interp_inner_pad(a.lds_side);
    for (int d = 1; d < NDIM; ++d) cells *= a.lds_side;
    size_t sh = 2 * cells * sizeof(T);
    allow_big_lds(k_interp<T, NDIM, W, true, MODE>, sh);
    hipLaunchKernelGGL((k_interp<T, NDIM, W, true, MODE>), dim3(nb), dim3(256),
                       sh, s, a, out, out_f64);
  } else {
    if constexpr (NDIM == 3)
      hipLaunchKernelGGL((k_interp<T, NDIM, W, false, MODE>), dim3(nb),
                         dim3(256), 0, s, a, out, out_f64);
    else
      throw Error("internal: no-LDS interpolation only for 3D");
  }
  DMX_HIP(hipGetLastError());
}

template <typename T, int NDIM, int MODE>
void by_w(const SpreadArgs &a, void *out, int f64, hipStream_t s) {
  switch (a.g.W) {
    case 4: return go<T, NDIM, 4, MODE>(a, out, f64, s);
    case 5: return go<T, NDIM, 5, MODE>(a, out, f64, s);
    case 6: return go<T, NDIM, 6, MODE>(a, out, f64, s);
    case 7: return go<T, NDIM, 7, MODE>(a, out, f64, s);
    case 8: return go<T, NDIM, 8, MODE>(a, out, f64, s);
    default: break;
  }
  if constexpr (sizeof(T) == 8) {
    switch (a.g.W) {
      case 9: return go<T, NDIM, 9, MODE>(a, out, f64, s);
      case 10: return go<T, NDIM, 10, MODE>(a, out, f64, s);
      case 11: return go<T, NDIM, 11, MODE>(a, out, f64, s);
      case 12: return go<T, NDIM, 12, MODE>(a, out, f64, s);
      case 13: return go<T, NDIM, 13, MODE>(a, out, f64, s);
      case 14: return go<T, NDIM, 14, MODE>(a, out, f64, s);
      case 15: return go<T, NDIM, 15, MODE>(a, out, f64, s);
      case 16: return go<T, NDIM, 16, MODE>(a, out, f64, s);
      default: break;
    }
  }
  throw Error("requested support out of range");
}

template <int MODE>
void dispatch(bool fp64, const SpreadArgs &a, void *out, int f64,
              hipStream_t s) {
  switch (a.g.ndim) {
    case 1:
      return fp64 ? by_w<double, 1, MODE>(a, out, f64, s)
                  : by_w<float, 1, MODE>(a, out, f64, s);
    case 2:
      return fp64 ? by_w<double, 2, MODE>(a, out, f64, s)
                  : by_w<float, 2, MODE>(a, out, f64, s);
    case 3:
      return fp64 ? by_w<double, 3, MODE>(a, out, f64, s)
                  : by_w<float, 3, MODE>(a, out, f64, s);
  }
  throw Error("transform must be 1D/2D/3D");
}
}  // namespace

void launch_interp(bool fp64, const SpreadArgs &a, void *out, int out_f64,
                   hipStream_t s) {
  dispatch<0>(fp64, a, out, out_f64, s);
}

void launch_interp_grad(bool fp64, const SpreadArgs &a, void *out,
                        int out_f64, hipStream_t s) {
  dispatch<1>(fp64, a, out, out_f64, s);
}

}  // namespace dmx
