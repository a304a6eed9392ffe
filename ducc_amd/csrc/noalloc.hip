// noalloc.hip -- dmx_nufft_noalloc: a type 1 / type 2 transform that
// allocates nothing (no hipMalloc, hipFree, stream or event creation, no
// host-to-device copy), for callers that cannot allocate: code under HIP
// graph capture, tight serving loops, harnesses that forbid allocation.
//
// Every temporary of the transform (kernel constants, index build, the
// oversampled grid and its second buffer, FFT twiddles) is carved from one
// file-scope device array, `scratch`, of DMX_NOALLOC_SCRATCH_BYTES per GPU.
// The same kernels as dmx_nu2u / dmx_u2nu run; the constants are written by
// kernels (values passed as kernel arguments, twiddles computed on the
// device) instead of being copied from the host.  Each GPU holds its own
// copy of the array at its own address, looked up once per device with
// hipGetSymbolAddress; host code uses only that address, never the array's
// name.  Calls that overlap on one GPU share the array and race; every
// call re-initialises what it uses (the sort's tables and the grid are
// cleared by fill kernels).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "ducc_amd.h"
#include "runtime.h"

namespace dmx {
namespace {

static __device__ __align__(256) unsigned char
    scratch[DMX_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_scratch_ptr[MAX_DEVICES] = {nullptr};

// this device's address of `scratch` (the current device must be `dev`)
void *scratch_ptr(int dev) {
  DMX_ASSERT(dev >= 0 && dev < MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    DMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DMX_ASSERT(p != nullptr, "no device address for the static scratch");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

bool contiguous_c(const dmx_array &a) {
  int64_t s = 1;
  for (int i = (int)a.ndim - 1; i >= 0; --i) {
    if (a.shape[i] != 1 && a.stride[i] != s) return false;
    s *= a.shape[i];
  }
  return true;
}

}  // namespace
}  // namespace dmx

using namespace dmx;

extern "C" int dmx_nufft_noalloc(int type, const dmx_array *uniform,
                                 const dmx_array *coord,
                                 const dmx_array *points, int forward,
                                 double epsilon, double sigma_min,
                                 double sigma_max, const double *periodicity,
                                 int fft_order, const dmx_options *opt_in,
                                 dmx_plan_info *info) {
  try {
    DMX_ASSERT(uniform && coord && points && periodicity, "null argument");
    DMX_ASSERT(type == 1 || type == 2, "type must be 1 or 2");
    dmx_options opt;
    dmx_default_options(&opt);
    if (opt_in) opt = *opt_in;
    DMX_ASSERT(opt.chooser == 0,
               "the learned chooser is not available without allocation");
    DMX_ASSERT(opt.deterministic == 0,
               "deterministic mode is not available without allocation");
    DMX_ASSERT(!(opt.flags & DMX_CHECK_OFFSETS),
               "offset checks are not available without allocation");
    // validation (same rules and messages as dmx_nu2u / dmx_u2nu)
    DMX_ASSERT(coord->ndim == 2, "coord must be a 2D array");
    DMX_ASSERT(coord->typecode == DMX_F32 || coord->typecode == DMX_F64,
               "not yet supported");
    const bool cplx = (uniform->typecode == DMX_C64 ||
                       uniform->typecode == DMX_C128) &&
                      points->typecode == uniform->typecode;
    DMX_ASSERT(cplx, "not yet supported");
    const int ndim = (int)uniform->ndim;
    DMX_ASSERT(ndim >= 1 && ndim <= 3, "transform must be 1D/2D/3D");
    DMX_ASSERT(coord->shape[1] == ndim, "dimensionality mismatch");
    DMX_ASSERT(points->ndim == 1, "points must be a 1D array");
    DMX_ASSERT(points->shape[0] == coord->shape[0],
               "number of points mismatch");
    for (int d = 0; d < ndim; ++d)
      DMX_ASSERT(uniform->shape[d] >= 1, "uniform grid dimensions mismatch");
    DMX_ASSERT(uniform->device && coord->device && points->device,
               "dmx_nufft_noalloc takes device arrays only");
    DMX_ASSERT(contiguous_c(*coord) && contiguous_c(*points),
               "device arrays must be C-contiguous here");
    const size_t M = (size_t)coord->shape[0];
    size_t nuni[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) nuni[d] = (size_t)uniform->shape[d];
    const bool vf64 = uniform->typecode == DMX_C128;
    Params p = choose_params(ndim, nuni, M, type == 1, epsilon, sigma_min,
                             sigma_max, periodicity, fft_order != 0, vf64,
                             opt);
    // regions flushed with atomics (storing them would need far more
    // scratch); the whole requirement is checked before anything is queued
    p.slab_flush = false;
    const size_t need = Plan::scratch_bytes(p, M);
    if (need > (size_t)DMX_NOALLOC_SCRATCH_BYTES)
      throw Error("dmx_nufft_noalloc: this request needs " +
                  std::to_string((need + (1 << 20) - 1) >> 20) +
                  " MiB of scratch, more than the " +
                  std::to_string((size_t)DMX_NOALLOC_SCRATCH_BYTES >> 20) +
                  " MiB static scratch holds; use dmx_nu2u / dmx_u2nu");
    int dev = 0;
    DMX_HIP(hipGetDevice(&dev));
    Arena arena;
    arena.base = (char *)scratch_ptr(dev);
    arena.size = DMX_NOALLOC_SCRATCH_BYTES;
    // the caller's stream, else the null stream (no stream is created)
    const hipStream_t s = (hipStream_t)opt.stream;
    int64_t ust[3] = {0, 0, 0};
    for (int d = 0; d < ndim; ++d) ust[d] = uniform->stride[d];
    {
      ArenaScope scope(&arena);
      Plan plan(p, s);
      plan.build_index(coord->data, coord->typecode == DMX_F32, M, false);
      if (type == 1)
        plan.nu2u(forward != 0, points->data, vf64, uniform->data, vf64, ust);
      else
        plan.u2nu(forward != 0, uniform->data, vf64, ust, points->data, vf64);
      if (info) plan.info(info);
      // the plan's buffers are the static scratch: nothing to keep alive
      if (!(opt.flags & DMX_ASYNC)) DMX_HIP(hipStreamSynchronize(s));
      plan.mark_done();
    }
    set_error("");
    return 0;
  } catch (std::exception &e) {
    set_error(e.what());
    (void)hipGetLastError();
    return 1;
  } catch (...) {
    set_error("unknown error");
    return 1;
  }
}
