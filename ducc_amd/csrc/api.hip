// api.hip -- the extern "C" boundary (include/ducc_amd.h).  Validates every
// argument on the host before any GPU work (messages follow ducc0's
// MR_assert texts, nufft.h:342-343, 400, 409, 1904-1905,
// gridding_kernel.h:442, nufft_pymod.cc:93), stages host arrays through the
// device, runs the plan on one HIP stream and returns synchronously.
#include <hip/hip_runtime.h>

#include <cstring>
#include <deque>
#include <mutex>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ducc_amd.h"
#include <cmath>

#include "chooser.h"
#include "grid.h"
#include "c2c.h"
#include "mlp.h"
#include "params.h"
#include "runtime.h"
#include "sht.h"
#include "wgridder.h"

namespace dmx {

namespace {
thread_local std::string g_err;

template <typename F> int guard(F &&f) {
  try {
    f();
    g_err.clear();
    return 0;
  } catch (std::exception &e) {
    g_err = e.what();
    (void)hipGetLastError();
    return 1;
  } catch (...) {
    g_err = "unknown error";
    return 1;
  }
}

size_t elsize(int tc) {
  switch (tc) {
    case DMX_F32: return 4;
    case DMX_F64: return 8;
    case DMX_C64: return 8;
    case DMX_C128: return 16;
    case DMX_U32: return 4;
    case DMX_U8: return 1;
  }
  throw Error("not yet supported");
}

int64_t nelem(const dmx_array &a) {
  int64_t n = 1;
  for (int i = 0; i < a.ndim; ++i) n *= a.shape[i];
  return n;
}

bool contiguous(const dmx_array &a) {
  int64_t s = 1;
  for (int i = (int)a.ndim - 1; i >= 0; --i) {
    if (a.shape[i] != 1 && a.stride[i] != s) return false;
    s *= a.shape[i];
  }
  return true;
}

// copy a strided host array into a contiguous host buffer (and back)
void pack(const dmx_array &a, void *dst) {
  size_t es = elsize(a.typecode);
  int64_t n = nelem(a);
  if (contiguous(a)) {
    std::memcpy(dst, a.data, (size_t)n * es);
    return;
  }
  std::vector<int64_t> idx(a.ndim, 0);
  const char *src = (const char *)a.data;
  char *d = (char *)dst;
  for (int64_t q = 0; q < n; ++q) {
    int64_t off = 0;
    for (int i = 0; i < a.ndim; ++i) off += idx[i] * a.stride[i];
    std::memcpy(d + q * es, src + off * (int64_t)es, es);
    for (int i = (int)a.ndim - 1; i >= 0; --i) {
      if (++idx[i] < a.shape[i]) break;
      idx[i] = 0;
    }
  }
}
void unpack(const void *srcv, const dmx_array &a) {
  size_t es = elsize(a.typecode);
  int64_t n = nelem(a);
  if (contiguous(a)) {
    std::memcpy(a.data, srcv, (size_t)n * es);
    return;
  }
  std::vector<int64_t> idx(a.ndim, 0);
  const char *s = (const char *)srcv;
  char *dst = (char *)a.data;
  for (int64_t q = 0; q < n; ++q) {
    int64_t off = 0;
    for (int i = 0; i < a.ndim; ++i) off += idx[i] * a.stride[i];
    std::memcpy(dst + off * (int64_t)es, s + q * es, es);
    for (int i = (int)a.ndim - 1; i >= 0; --i) {
      if (++idx[i] < a.shape[i]) break;
      idx[i] = 0;
    }
  }
}

// An input array on the device.  Host arrays are packed and copied in;
// device arrays are used in place (contiguity required unless `strided_ok`).
struct DevIn {
  const void *p = nullptr;
  Buf tmp;
  int64_t st[10] = {0};
  DevIn(const dmx_array &a, hipStream_t s, bool strided_ok) {
    size_t es = elsize(a.typecode);
    int64_t n = nelem(a);
    if (a.device) {
      p = a.data;
      if (!strided_ok)
        DMX_ASSERT(contiguous(a), "device arrays must be C-contiguous here");
      for (int i = 0; i < a.ndim; ++i) st[i] = a.stride[i];
      return;
    }
    int64_t s1 = 1;
    for (int i = (int)a.ndim - 1; i >= 0; --i) {
      st[i] = s1;
      s1 *= a.shape[i];
    }
    if (n == 0) return;
    std::vector<char> h((size_t)n * es);
    pack(a, h.data());
    tmp.stream = s;  // released stream-ordered: kernels on s read it
    tmp.reset((size_t)n * es);
    DMX_HIP(hipMemcpyAsync(tmp.p, h.data(), h.size(), hipMemcpyHostToDevice, s));
    DMX_HIP(hipStreamSynchronize(s));
    p = tmp.p;
  }
};

// An output array on the device; host outputs are written back by finish().
struct DevOut {
  void *p = nullptr;
  Buf tmp;
  const dmx_array *a;
  int64_t st[10] = {0};
  DevOut(const dmx_array &arr, hipStream_t s, bool strided_ok) : a(&arr) {
    tmp.stream = s;  // released stream-ordered: kernels on s write it
    size_t es = elsize(arr.typecode);
    int64_t n = nelem(arr);
    if (arr.device) {
      p = arr.data;
      if (!strided_ok)
        DMX_ASSERT(contiguous(arr), "device arrays must be C-contiguous here");
      for (int i = 0; i < arr.ndim; ++i) st[i] = arr.stride[i];
      return;
    }
    int64_t s1 = 1;
    for (int i = (int)arr.ndim - 1; i >= 0; --i) {
      st[i] = s1;
      s1 *= arr.shape[i];
    }
    if (n) {
      tmp.reset((size_t)n * es);
      p = tmp.p;
    }
  }
  // wait = false: device outputs are left to complete in stream order
  void finish(hipStream_t s, bool wait = true) {
    if (a->device && !wait) return;
    DMX_HIP(hipStreamSynchronize(s));
    if (a->device || !tmp.p) return;
    std::vector<char> h(tmp.bytes);
    DMX_HIP(hipMemcpy(h.data(), tmp.p, tmp.bytes, hipMemcpyDeviceToHost));
    unpack(h.data(), *a);
  }
};

hipStream_t library_stream() {
  thread_local hipStream_t s = nullptr;
  thread_local int dev = -1;
  int cur = 0;
  DMX_HIP(hipGetDevice(&cur));
  if (!s || dev != cur) {
    DMX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    dev = cur;
  }
  return s;
}

// A second stream per thread and device (non-blocking, default priority)
// for the part of a call that can run beside the main stream, with the
// events that fork work to it and join it back.  Re-recording an event is
// safe: a stream wait uses the record that precedes it.
struct Aux {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
Aux &aux() {
  thread_local std::map<int, Aux> per_dev;
  int cur = 0;
  DMX_HIP(hipGetDevice(&cur));
  Aux &a = per_dev[cur];
  if (!a.s) {
    DMX_HIP(hipStreamCreateWithFlags(&a.s, hipStreamNonBlocking));
    DMX_HIP(hipEventCreateWithFlags(&a.fork, hipEventDisableTiming));
    DMX_HIP(hipEventCreateWithFlags(&a.join, hipEventDisableTiming));
  }
  return a;
}

dmx_options defaults() {
  dmx_options o;
  dmx_default_options(&o);
  return o;
}

bool is_complex(int tc) { return tc == DMX_C64 || tc == DMX_C128; }
bool is_real(int tc) { return tc == DMX_F32 || tc == DMX_F64; }

void check_coord(const dmx_array &c) {
  DMX_ASSERT(c.ndim == 2, "coord must be a 2D array");
  DMX_ASSERT(is_real(c.typecode), "not yet supported");
}

// shared request validation of u2nu / nu2u; returns ndim
int validate(const dmx_array &uni, const dmx_array &coord,
             const dmx_array &pts) {
  check_coord(coord);
  DMX_ASSERT(is_complex(uni.typecode) && is_complex(pts.typecode),
             "not yet supported");
  DMX_ASSERT(uni.typecode == pts.typecode, "not yet supported");
  const int ndim = (int)uni.ndim;
  DMX_ASSERT(ndim >= 1 && ndim <= 3, "transform must be 1D/2D/3D");
  DMX_ASSERT(coord.shape[1] == ndim, "dimensionality mismatch");
  DMX_ASSERT(pts.ndim == 1, "points must be a 1D array");
  DMX_ASSERT(pts.shape[0] == coord.shape[0], "number of points mismatch");
  for (int d = 0; d < ndim; ++d)
    DMX_ASSERT(uni.shape[d] >= 1, "uniform grid dimensions mismatch");
  return ndim;
}

hipStream_t pick_stream(const dmx_options &o) {
  if (o.flags & DMX_STREAM_GIVEN) return (hipStream_t)o.stream;
  return o.stream ? (hipStream_t)o.stream : library_stream();
}

// Plans of calls that returned before their work finished (DMX_ASYNC): kept,
// with an event recorded after their last kernel, until that event has
// completed, so that their buffers go back to the pool only when no kernel
// can still touch them.  At most MAX_RETIRED are kept; beyond that the
// oldest is waited for.
struct Retired {
  hipEvent_t ev;
  std::unique_ptr<Plan> plan;
  std::shared_ptr<void> keep;  // buffers of a call without a plan
};
std::mutex g_retire_mu;
// never destroyed: plans must not be torn down after the HIP runtime at exit
std::deque<Retired> &g_retired = *new std::deque<Retired>();
constexpr size_t MAX_RETIRED = 4;

// frees retired plans whose work has finished; wait_all: wait for all
void reap(bool wait_all) {
  std::vector<Retired> done;
  {
    std::lock_guard<std::mutex> lk(g_retire_mu);
    while (!g_retired.empty()) {
      Retired &r = g_retired.front();
      const bool wait = wait_all || g_retired.size() > MAX_RETIRED;
      hipError_t q = wait ? hipEventSynchronize(r.ev) : hipEventQuery(r.ev);
      if (q == hipErrorNotReady) {
        (void)hipGetLastError();  // "not ready" is not an error here
        break;
      }
      if (r.plan) r.plan->mark_done();
      done.push_back(std::move(r));
      g_retired.pop_front();
    }
  }
  for (auto &r : done) {
    r.plan.reset();
    r.keep.reset();
    (void)hipEventDestroy(r.ev);
  }
}

void retire(std::unique_ptr<Plan> plan, hipStream_t s,
            std::shared_ptr<void> keep = nullptr) {
  hipEvent_t ev;
  DMX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  DMX_HIP(hipEventRecord(ev, s));
  {
    std::lock_guard<std::mutex> lk(g_retire_mu);
    g_retired.push_back({ev, std::move(plan), std::move(keep)});
  }
  reap(false);
}

// Accuracy spot-check of a parameter set: a type-1 transform of random
// complex128 values at up to 4096 of the request's points (evenly strided
// sample), compared on 32 random modes with the FP64 direct sum on the GPU.
// Passes when the relative l2 error is within the reference tests' bound
// (4/2/1 eps in 1/2/3-D, test_nufft.py:79,140,203).
bool spot_check(const Params &p, const void *coords, bool coord_f32, size_t M,
                hipStream_t s) {
  const int nd = p.ndim;
  const size_t S = std::min<size_t>(M, 4096);
  if (S == 0) return true;
  std::vector<uint32_t> sel(S);
  for (size_t i = 0; i < S; ++i) sel[i] = (uint32_t)((i * M) / S);
  Buf dsel(S * 4), dsub(S * nd * 8);
  DMX_HIP(hipMemcpyAsync(dsel.p, sel.data(), S * 4, hipMemcpyHostToDevice, s));
  launch_gather_coords(coords, coord_f32, nd, dsel.as<uint32_t>(), S, dsub.p,
                       s);
  std::vector<double> c64(S * nd);
  if (coord_f32) {
    std::vector<float> c32(S * nd);
    DMX_HIP(hipMemcpyAsync(c32.data(), dsub.p, S * nd * 4,
                           hipMemcpyDeviceToHost, s));
    DMX_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < S * nd; ++i) c64[i] = c32[i];
  } else {
    DMX_HIP(hipMemcpyAsync(c64.data(), dsub.p, S * nd * 8,
                           hipMemcpyDeviceToHost, s));
    DMX_HIP(hipStreamSynchronize(s));
  }
  // values: a fixed pseudo-random sequence
  std::vector<double> v(2 * S);
  uint64_t st = 0x9E3779B97F4A7C15ull;
  for (auto &x : v) {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    x = (double)(st >> 11) / 9007199254740992.0 - 0.5;
  }
  size_t N = 1;
  for (int d = 0; d < nd; ++d) N *= p.nuni[d];
  const size_t nm = std::min<size_t>(N, 32);
  std::vector<int64_t> modes(nm);
  for (size_t i = 0; i < nm; ++i) {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    modes[i] = nm == N ? (int64_t)i : (int64_t)(st % N);
  }
  Buf dvals(2 * S * 8), dc64(S * nd * 8), dout(N * 16), dmodes(nm * 8),
      dref(nm * 16), dgot(nm * 16);
  DMX_HIP(hipMemcpyAsync(dvals.p, v.data(), 2 * S * 8, hipMemcpyHostToDevice, s));
  DMX_HIP(hipMemcpyAsync(dc64.p, c64.data(), S * nd * 8, hipMemcpyHostToDevice, s));
  DMX_HIP(hipMemcpyAsync(dmodes.p, modes.data(), nm * 8, hipMemcpyHostToDevice, s));
  {
    Plan sp(p, s);
    sp.build_index(dsub.p, coord_f32, S, false);
    int64_t ust[3] = {0, 0, 0};
    int64_t acc = 1;
    for (int d = nd - 1; d >= 0; --d) {
      ust[d] = acc;
      acc *= (int64_t)p.nuni[d];
    }
    sp.nu2u(true, dvals.p, true, dout.p, true, ust);
  }
  launch_gather_c128(dout.p, dmodes.as<int64_t>(), nm, dgot.p, s);
  NudftArgs a{};
  a.type = 1;
  a.ndim = nd;
  a.forward = 1;
  a.fft_order = p.fft_order ? 1 : 0;
  a.M = (int64_t)S;
  a.nsel = (int64_t)nm;
  for (int d = 0; d < 3; ++d) {
    a.n[d] = d < nd ? (int64_t)p.nuni[d] : 1;
    a.period[d] = d < nd ? p.period[d] : 1.;
  }
  a.coord = dc64.as<double>();
  a.vals = dvals.as<double>();
  a.sel = dmodes.as<int64_t>();
  a.out = dref.as<double>();
  launch_nudft(a, s);
  std::vector<double> got(2 * nm), ref(2 * nm);
  DMX_HIP(hipMemcpyAsync(got.data(), dgot.p, nm * 16, hipMemcpyDeviceToHost, s));
  DMX_HIP(hipMemcpyAsync(ref.data(), dref.p, nm * 16, hipMemcpyDeviceToHost, s));
  DMX_HIP(hipStreamSynchronize(s));
  double num = 0, den = 0;
  for (size_t i = 0; i < 2 * nm; ++i) {
    num += (got[i] - ref[i]) * (got[i] - ref[i]);
    den += ref[i] * ref[i];
  }
  const double mul = nd == 1 ? 4. : (nd == 2 ? 2. : 1.);
  return den == 0 || std::sqrt(num / den) <= mul * p.eps;
}

int run_unplanned(bool type1, const dmx_array *uni, const dmx_array *coord,
                  const dmx_array *pts, int forward, double epsilon,
                  double smin, double smax, const double *per, int fft_order,
                  const dmx_options *opt_in, dmx_plan_info *info,
                  size_t verbosity = 0, bool grad = false,
                  const dmx_array *gradout = nullptr) {
  return guard([&] {
    DMX_ASSERT(uni && coord && pts && per, "null argument");
    reap(false);
    dmx_options opt = opt_in ? *opt_in : defaults();
    const int ndim = validate(*uni, *coord, *pts);
    const size_t M = (size_t)coord->shape[0];
    size_t nuni[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) nuni[d] = (size_t)uni->shape[d];
    const bool vf64 = uni->typecode == DMX_C128;
    Params p = choose_params(ndim, nuni, M, type1, epsilon, smin, smax, per,
                             fft_order != 0, vf64, opt);
    hipStream_t s = pick_stream(opt);
    DevIn cin(*coord, s, false);
    const bool cf32 = coord->typecode == DMX_F32;
    if (p.chooser_used && !p.chooser_verified && opt.spot_check) {
      // spot_check == 2 forces the failure branch (tests of the fallback)
      const bool ok = opt.spot_check == 2 ? false
                                          : spot_check(p, cin.p, cf32, M, s);
      chooser_verdict(ndim, nuni, M, epsilon, smin, smax, p.fp64, type1, ok);
      if (!ok) {
        dmx_options o2 = opt;
        o2.chooser = 0;
        p = choose_params(ndim, nuni, M, type1, epsilon, smin, smax, per,
                          fft_order != 0, vf64, o2);
      }
    }
    // return before the work finishes only when no host array is staged
    // (staging buffers die with this call)
    const bool async = (opt.flags & DMX_ASYNC) && coord->device &&
                       uni->device && pts->device &&
                       (!grad || gradout->device);
    auto plan = std::make_unique<Plan>(p, s);
    if (verbosity > 0) {
      // ducc0 prints the request before the work and the phase timings
      // after it (nufft.h:488-491, 499-502); the timings need the results,
      // so a verbose call waits for its work
      plan->report_params(type1);
      plan->start_timing();
    }
    std::unique_ptr<DevIn> ui;  // type 2: the uniform array, on the device
    if (!type1) ui = std::make_unique<DevIn>(*uni, s, true);
    // The grid side (type 1: zeroing; type 2: pad and FFT) does not depend
    // on the points: for large requests it runs on the second stream while
    // the points are sorted on s, and s waits for it before the points are
    // processed (small ones would only pay the cross-stream latency).
    size_t nover_total = 1;
    for (int d = 0; d < ndim; ++d) nover_total *= p.nover[d];
    const bool fork = M >= ((size_t)1 << 20) && nover_total >= ((size_t)1 << 20);
    if (fork) {
      Aux &ax = aux();
      DMX_HIP(hipEventRecord(ax.fork, s));
      DMX_HIP(hipStreamWaitEvent(ax.s, ax.fork, 0));
      plan->set_stream(ax.s);
      if (type1)
        plan->nu2u_grid();
      else
        plan->u2nu_grid(forward != 0, ui->p, vf64, ui->st, grad);
      DMX_HIP(hipEventRecord(ax.join, ax.s));
      plan->set_stream(s);
    }
    // on request, the points' records are sorted (type 1 with their
    // values) where the request fits the record layout
    std::unique_ptr<DevIn> pin;
    if (type1) pin = std::make_unique<DevIn>(*pts, s, false);
    const bool recs =
        !grad && (opt.flags & DMX_POINT_RECORDS) &&
        plan->build_records(cin.p, cf32, M, type1, type1 ? pin->p : nullptr,
                            vf64);
    if (!recs) plan->build_index(cin.p, cf32, M, false);
    if (fork) DMX_HIP(hipStreamWaitEvent(s, aux().join, 0));
    if (type1) {
      DevOut uo(*uni, s, true);
      plan->nu2u(forward != 0, pin->p, vf64, uo.p, vf64, uo.st);
      uo.finish(s, !async);
    } else if (!grad) {
      DevOut po(*pts, s, false);
      plan->u2nu(forward != 0, ui->p, vf64, ui->st, po.p, vf64);
      po.finish(s, !async);
    } else {
      DevOut go(*gradout, s, false);
      plan->u2nu_grad(forward != 0, ui->p, vf64, ui->st, go.p, vf64);
      go.finish(s, !async);
    }
    if (info) plan->info(info);
    if (verbosity > 0)
      plan->report_timings(type1 ? "nu2u" : (grad ? "u2nu_coord_grad" : "u2nu"));
    if (async) retire(std::move(plan), s);
  });
}
}  // namespace

void set_error(const char *msg) { g_err = msg; }

}  // namespace dmx

struct dmx_plan {
  std::unique_ptr<dmx::Plan> plan;
  bool nu2u;
  bool async;  // DMX_ASYNC at creation
  int ndim;
  int64_t shape[3];
  hipStream_t stream;  // creation stream: calls without options run here
};

using namespace dmx;

namespace {
// the stream and async mode of one planned call: the caller's options, else
// the plan's creation settings; the plan's work moves to that stream in
// order after what it queued before
hipStream_t plan_stream(dmx_plan *plan, const dmx_options *opt, bool &async) {
  hipStream_t s = plan->stream;
  async = plan->async;
  if (opt) {
    s = pick_stream(*opt);
    async = (opt->flags & DMX_ASYNC) != 0;
  }
  plan->plan->use_stream(s);
  return s;
}
}  // namespace

extern "C" {

const char *dmx_last_error(void) { return g_err.c_str(); }

void dmx_default_options(dmx_options *o) {
  std::memset(o, 0, sizeof(*o));
  o->precision = 0;
  o->chooser = 0;
  o->log2tile = 0;
  o->kernel_idx = -1;
  o->fp64_below = 1e-6;
  o->stream = nullptr;
  o->sigma = 0;
  o->spot_check = 1;
}

size_t dmx_mlp_nparam(void) { return MLP_NPARAM; }

int dmx_mlp_forward(const float *params, const float *X, int B, float *Y) {
  return guard([&] {
    DMX_ASSERT(params && X && Y && B > 0, "bad arguments");
    hipStream_t s = library_stream();
    Buf dp(MLP_NPARAM * 4), dx((size_t)B * MLP_IN * 4),
        dy((size_t)B * MLP_OUT * 4);
    DMX_HIP(hipMemcpyAsync(dp.p, params, MLP_NPARAM * 4, hipMemcpyHostToDevice, s));
    DMX_HIP(hipMemcpyAsync(dx.p, X, (size_t)B * MLP_IN * 4, hipMemcpyHostToDevice, s));
    mlp_forward(dp.as<float>(), dx.as<float>(), B, nullptr, dy.as<float>(), s);
    DMX_HIP(hipMemcpyAsync(Y, dy.p, (size_t)B * MLP_OUT * 4, hipMemcpyDeviceToHost, s));
    DMX_HIP(hipStreamSynchronize(s));
  });
}

namespace {
struct MlpState {
  Buf P, X, T, w, act, dz, lossb, G, loss, m, v;
  MlpState(const float *params, const float *X_, const float *T_,
           const float *w_, int B, hipStream_t s)
      : P(MLP_NPARAM * 4), X((size_t)B * MLP_IN * 4),
        T((size_t)B * MLP_OUT * 4), w(MLP_OUT * 4),
        act(mlp_act_floats(B) * 4), dz((mlp_act_floats(B) + B * MLP_OUT) * 4),
        lossb((size_t)B * MLP_OUT * 4), G(MLP_NPARAM * 4), loss(4),
        m(MLP_NPARAM * 4), v(MLP_NPARAM * 4) {
    DMX_HIP(hipMemcpyAsync(P.p, params, MLP_NPARAM * 4, hipMemcpyHostToDevice, s));
    DMX_HIP(hipMemcpyAsync(X.p, X_, (size_t)B * MLP_IN * 4, hipMemcpyHostToDevice, s));
    DMX_HIP(hipMemcpyAsync(T.p, T_, (size_t)B * MLP_OUT * 4, hipMemcpyHostToDevice, s));
    static const float ones[MLP_OUT] = {1.f, 1.f, 1.f};
    DMX_HIP(hipMemcpyAsync(w.p, w_ ? w_ : ones, MLP_OUT * 4, hipMemcpyHostToDevice, s));
    launch_zero(m.p, MLP_NPARAM * 4, s);
    launch_zero(v.p, MLP_NPARAM * 4, s);
    DMX_HIP(hipStreamSynchronize(s));
  }
  void grad(int B, hipStream_t s) {
    mlp_grad(P.as<float>(), X.as<float>(), T.as<float>(), w.as<float>(), B,
             act.as<float>(), dz.as<float>(), lossb.as<float>(), G.as<float>(),
             loss.as<float>(), s);
  }
};
}  // namespace

int dmx_mlp_grad(const float *params, const float *X, const float *T,
                 const float *w, int B, float *grad, float *loss) {
  return guard([&] {
    DMX_ASSERT(params && X && T && grad && B > 0, "bad arguments");
    hipStream_t s = library_stream();
    MlpState st(params, X, T, w, B, s);
    st.grad(B, s);
    DMX_HIP(hipMemcpyAsync(grad, st.G.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost, s));
    if (loss)
      DMX_HIP(hipMemcpyAsync(loss, st.loss.p, 4, hipMemcpyDeviceToHost, s));
    DMX_HIP(hipStreamSynchronize(s));
  });
}

int dmx_mlp_train(float *params, const float *X, const float *T,
                  const float *w, int B, int steps, float lr,
                  float weight_decay, float *losses) {
  return guard([&] {
    DMX_ASSERT(params && X && T && B > 0 && steps >= 0, "bad arguments");
    hipStream_t s = library_stream();
    MlpState st(params, X, T, w, B, s);
    Buf lh((size_t)std::max(steps, 1) * 4);
    for (int t = 1; t <= steps; ++t) {
      st.grad(B, s);
      if (losses)
        DMX_HIP(hipMemcpyAsync(lh.as<float>() + (t - 1), st.loss.p, 4,
                               hipMemcpyDeviceToDevice, s));
      mlp_adam_step(st.P.as<float>(), st.G.as<float>(), st.m.as<float>(),
                    st.v.as<float>(), lr, t, weight_decay, s);
    }
    DMX_HIP(hipMemcpyAsync(params, st.P.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost, s));
    if (losses && steps)
      DMX_HIP(hipMemcpyAsync(losses, lh.p, (size_t)steps * 4, hipMemcpyDeviceToHost, s));
    DMX_HIP(hipStreamSynchronize(s));
  });
}

int dmx_chooser_set(const float *params, const double *norm) {
  return guard([&] {
    DMX_ASSERT(params && norm, "bad arguments");
    for (int i = 6; i < 12; ++i) DMX_ASSERT(norm[i] > 0, "bad normalisation");
    for (int i = 15; i < 18; ++i) DMX_ASSERT(norm[i] > 0, "bad normalisation");
    chooser_set(params, norm);
  });
}

int dmx_chooser_features(size_t ndim, const int64_t *shape, size_t npoints,
                         double epsilon, int fp64, int nu2u, double *f) {
  return guard([&] {
    DMX_ASSERT(ndim >= 1 && ndim <= 3 && shape && f && epsilon > 0,
               "bad arguments");
    size_t n[3] = {1, 1, 1};
    for (size_t d = 0; d < ndim; ++d) n[d] = (size_t)shape[d];
    chooser_features((int)ndim, n, npoints, epsilon, fp64 != 0, nu2u != 0, f);
  });
}

int dmx_choose(size_t ndim, const int64_t *shape, size_t npoints,
               double epsilon, int values_f64, int nu2u, double sigma_min,
               double sigma_max, const dmx_options *opt_in,
               dmx_plan_info *info) {
  return guard([&] {
    DMX_ASSERT(shape && info, "null argument");
    DMX_ASSERT(ndim >= 1 && ndim <= 3, "transform must be 1D/2D/3D");
    dmx_options opt = opt_in ? *opt_in : defaults();
    size_t n[3] = {1, 1, 1};
    double per[3] = {2 * M_PI, 2 * M_PI, 2 * M_PI};
    for (size_t d = 0; d < ndim; ++d) n[d] = (size_t)shape[d];
    Params p = choose_params((int)ndim, n, npoints, nu2u != 0, epsilon,
                             sigma_min, sigma_max, per, false, values_f64 != 0,
                             opt);
    std::memset(info, 0, sizeof(*info));
    info->ndim = p.ndim;
    info->W = p.W;
    info->kernel_idx = p.kidx;
    info->fp64 = p.fp64;
    info->log2tile = p.log2tile;
    info->chooser_used = p.chooser_used;
    for (int d = 0; d < 3; ++d) {
      info->nuni[d] = d < p.ndim ? (int64_t)p.nuni[d] : 1;
      info->nover[d] = d < p.ndim ? (int64_t)p.nover[d] : 1;
    }
    info->ofactor = p.ofactor;
    info->eps_tab = p.eps_tab;
    info->epsilon = p.eps;
    info->npoints = (int64_t)npoints;
  });
}

double dmx_best_epsilon(size_t ndim, int singleprec, double sigma_min,
                        double sigma_max) {
  double r = -1;
  int rc = guard([&] {
    DMX_ASSERT(ndim >= 1 && ndim <= 3, "bad dimensionality");
    r = best_epsilon((int)ndim, singleprec != 0, sigma_min, sigma_max);
    DMX_ASSERT(r > 0, "no appropriate kernel found");
  });
  return rc ? -1. : r;
}

int dmx_u2nu(const dmx_array *grid, const dmx_array *coord, int forward,
             double epsilon, size_t nthreads, dmx_array *out,
             size_t verbosity, double sigma_min, double sigma_max,
             const double *periodicity, int fft_order,
             const dmx_options *opt, dmx_plan_info *info) {
  (void)nthreads;
  return run_unplanned(false, grid, coord, out, forward, epsilon, sigma_min,
                       sigma_max, periodicity, fft_order, opt, info,
                       verbosity);
}

int dmx_nu2u(const dmx_array *points, const dmx_array *coord, int forward,
             double epsilon, size_t nthreads, dmx_array *out,
             size_t verbosity, double sigma_min, double sigma_max,
             const double *periodicity, int fft_order,
             const dmx_options *opt, dmx_plan_info *info) {
  (void)nthreads;
  return run_unplanned(true, out, coord, points, forward, epsilon, sigma_min,
                       sigma_max, periodicity, fft_order, opt, info,
                       verbosity);
}

int dmx_u2nu_coord_grad(const dmx_array *grid, const dmx_array *coord,
                        int forward, double epsilon, dmx_array *out,
                        double sigma_min, double sigma_max,
                        const double *periodicity, int fft_order,
                        const dmx_options *opt) {
  return guard([&] {
    DMX_ASSERT(grid && coord && out && periodicity, "null argument");
    check_coord(*coord);
    DMX_ASSERT(out->ndim == 2 && out->shape[0] == coord->shape[0] &&
                   out->shape[1] == coord->shape[1],
               "gradient output must have the shape of coord");
    DMX_ASSERT(out->typecode == grid->typecode, "not yet supported");
    // view the (M, ndim) output as the points array for validation
    dmx_array pts = *out;
    pts.ndim = 1;
    pts.shape[0] = coord->shape[0];
    pts.stride[0] = 1;
    int rc = run_unplanned(false, grid, coord, &pts, forward, epsilon,
                           sigma_min, sigma_max, periodicity, fft_order, opt,
                           nullptr, 0, true, out);
    if (rc) throw Error(g_err);
  });
}

int dmx_plan_create(dmx_plan **plan, int nu2u, const dmx_array *coord,
                    const int64_t *grid_shape, size_t ndim, double epsilon,
                    size_t nthreads, double sigma_min, double sigma_max,
                    const double *periodicity, int fft_order,
                    const dmx_options *opt_in) {
  (void)nthreads;
  return guard([&] {
    DMX_ASSERT(plan && coord && grid_shape && periodicity, "null argument");
    *plan = nullptr;
    check_coord(*coord);
    DMX_ASSERT(ndim >= 1 && ndim <= 3, "transform must be 1D/2D/3D");
    DMX_ASSERT(coord->shape[1] == (int64_t)ndim, "dimensionality mismatch");
    dmx_options opt = opt_in ? *opt_in : defaults();
    size_t nuni[3] = {1, 1, 1};
    for (size_t d = 0; d < ndim; ++d) {
      DMX_ASSERT(grid_shape[d] >= 1, "uniform grid dimensions mismatch");
      nuni[d] = (size_t)grid_shape[d];
    }
    const size_t M = (size_t)coord->shape[0];
    // a plan computes in the precision of its coordinates unless told
    // otherwise (nufft_pymod.cc:145-150)
    Params p = choose_params((int)ndim, nuni, M, nu2u != 0, epsilon,
                             sigma_min, sigma_max, periodicity,
                             fft_order != 0, coord->typecode == DMX_F64, opt);
    auto h = std::make_unique<dmx_plan>();
    h->stream = pick_stream(opt);
    h->async = (opt.flags & DMX_ASYNC) != 0;
    h->nu2u = nu2u != 0;
    h->ndim = (int)ndim;
    for (size_t d = 0; d < 3; ++d) h->shape[d] = d < ndim ? grid_shape[d] : 1;
    h->plan = std::make_unique<Plan>(p, h->stream);
    // host coordinates are staged in a buffer that returns to the pool only
    // once the index build queued on the plan's stream has read it
    DevIn cin(*coord, h->stream, false);
    h->plan->build_index(cin.p, coord->typecode == DMX_F32, M, true);
    *plan = h.release();
  });
}

int dmx_plan_nu2u(dmx_plan *plan, int forward, size_t verbosity,
                  const dmx_array *points, dmx_array *out,
                  const dmx_options *opt) {
  return guard([&] {
    DMX_ASSERT(plan && points && out, "null argument");
    DMX_ASSERT(is_complex(points->typecode) && points->typecode == out->typecode,
               "not yet supported");
    DMX_ASSERT(points->ndim == 1 &&
                   (size_t)points->shape[0] == plan->plan->npoints(),
               "number of points mismatch");
    DMX_ASSERT(out->ndim == plan->ndim, "uniform grid dimensions mismatch");
    for (int d = 0; d < plan->ndim; ++d)
      DMX_ASSERT(out->shape[d] == plan->shape[d],
                 "uniform grid dimensions mismatch");
    bool async = false;
    hipStream_t s = plan_stream(plan, opt, async);
    const bool f64 = points->typecode == DMX_C128;
    if (verbosity > 0) {
      plan->plan->report_params(true);
      plan->plan->start_timing();
    }
    DevIn pin(*points, s, false);
    DevOut uo(*out, s, true);
    plan->plan->nu2u(forward != 0, pin.p, f64, uo.p, f64, uo.st);
    uo.finish(s, !(async && points->device && out->device));
    if (verbosity > 0) plan->plan->report_timings("nu2u");
    plan->plan->flush_profile(false);
  });
}

int dmx_plan_u2nu(dmx_plan *plan, int forward, size_t verbosity,
                  const dmx_array *grid, dmx_array *out,
                  const dmx_options *opt) {
  return guard([&] {
    DMX_ASSERT(plan && grid && out, "null argument");
    DMX_ASSERT(is_complex(grid->typecode) && grid->typecode == out->typecode,
               "not yet supported");
    DMX_ASSERT(out->ndim == 1 && (size_t)out->shape[0] == plan->plan->npoints(),
               "number of points mismatch");
    DMX_ASSERT(grid->ndim == plan->ndim, "uniform grid dimensions mismatch");
    for (int d = 0; d < plan->ndim; ++d)
      DMX_ASSERT(grid->shape[d] == plan->shape[d],
                 "uniform grid dimensions mismatch");
    bool async = false;
    hipStream_t s = plan_stream(plan, opt, async);
    const bool f64 = grid->typecode == DMX_C128;
    if (verbosity > 0) {
      plan->plan->report_params(false);
      plan->plan->start_timing();
    }
    DevIn gin(*grid, s, true);
    DevOut po(*out, s, false);
    plan->plan->u2nu(forward != 0, gin.p, f64, gin.st, po.p, f64);
    po.finish(s, !(async && grid->device && out->device));
    if (verbosity > 0) plan->plan->report_timings("u2nu");
    plan->plan->flush_profile(false);
  });
}

int dmx_plan_get_info(const dmx_plan *plan, dmx_plan_info *info) {
  return guard([&] {
    DMX_ASSERT(plan && info, "null argument");
    plan->plan->info(info);
  });
}

int dmx_plan_destroy(dmx_plan *plan) {
  return guard([&] { delete plan; });
}

namespace {
// Shared set-up of the general synthesis and its adjoint: validation
// (sht_pymod.cc:805-870, sht.cc:2926-2937), the ring / mode geometry, the
// device tables and the NUFFT parameters of the doubled-sphere grid
// (SphereInterpol's kernel choice, sphere_interpol.h:426-428: the gridding
// cost model on the grid the points are interpolated from).
struct ShtJob {
  ShtGeometry geo;
  Params p;
  size_t M = 0;
  bool vf64 = true;
  int spin = 0;
  Buf mfac, mstart, accofs, sstart;
  size_t nacc = 0;
  ShtArgs args{};
  ShtJob(const dmx_array &alm, const dmx_array &map, size_t spin_,
         size_t lmax, const int64_t *mstart_h, size_t nm, int64_t lstride,
         const dmx_array &loc, double epsilon, double smin, double smax,
         int mode, const dmx_options &opt) {
    DMX_ASSERT(loc.ndim == 2 && loc.shape[1] == 2,
               "last dimension of loc must have size 2");
    DMX_ASSERT(loc.typecode == DMX_F64, "unsupported combination of data types");
    DMX_ASSERT(is_complex(alm.typecode), "unsupported combination of data types");
    vf64 = alm.typecode == DMX_C128;
    DMX_ASSERT(map.typecode == (vf64 ? DMX_F64 : DMX_F32),
               "unsupported combination of data types");
    DMX_ASSERT(mode >= 0 && mode <= 2, "unknown SHT mode");
    // sht.cc:2931-2935, 1857-1868: GRAD_ONLY / DERIV1 need spin > 0 (DERIV1
    // is the gradient of a spin-0 field: spin 1)
    if (mode != 0) DMX_ASSERT(spin_ > 0, "DERIV and GRAD_ONLY modes require spin>0");
    if (mode == 2) DMX_ASSERT(spin_ == 1, "DERIV1 mode requires spin 1");
    DMX_ASSERT(spin_ < (1u << 20), "spin too large");
    spin = (int)spin_;
    DMX_ASSERT(nm > 0 && mstart_h, "need at least m=0");
    DMX_ASSERT(nm <= lmax + 1, "mmax>lmax");
    const int64_t nalm_c = spin == 0 ? 1 : (mode == 0 ? 2 : 1),
                  nmaps = spin == 0 ? 1 : 2;
    DMX_ASSERT(alm.ndim == 2 && alm.shape[0] == nalm_c,
               "number of components mismatch in alm");
    DMX_ASSERT(map.ndim == 2 && map.shape[0] == nmaps,
               "number of components mismatch in map");
    M = (size_t)loc.shape[0];
    DMX_ASSERT(map.shape[1] == (int64_t)M, "array shape mismatch");
    // every a_lm index inside the array (sht_pymod.cc:168-180)
    int64_t need = 0;
    for (size_t m = 0; m < nm; ++m) {
      const int64_t lo = mstart_h[m] + (int64_t)m * lstride,
                    hi = mstart_h[m] + (int64_t)lmax * lstride;
      DMX_ASSERT(lo >= 0 && hi >= 0, "impossible a_lm memory layout");
      need = std::max(need, std::max(lo, hi) + 1);
    }
    DMX_ASSERT(alm.shape[1] >= need, "a_lm array too small");
    geo = sht_geometry(lmax, nm - 1);
    size_t nuni[3] = {(size_t)(2 * geo.N - 2), (size_t)geo.nphi, 1};
    const double per[2] = {2 * M_PI, 2 * M_PI};
    p = choose_params(2, nuni, M, true, epsilon, smin, smax, per, true, vf64,
                      opt);
    std::vector<double> mf;
    sht_mfac(nm - 1, mf);
    std::vector<int64_t> ao(nm);
    for (size_t m = 0; m < nm; ++m) {
      ao[m] = (int64_t)nacc;
      nacc += lmax + 1 - m;
    }
    mfac.reset(nm * 8);
    mstart.reset(nm * 8);
    accofs.reset(nm * 8);
    DMX_HIP(hipMemcpy(mfac.p, mf.data(), nm * 8, hipMemcpyHostToDevice));
    DMX_HIP(hipMemcpy(mstart.p, mstart_h, nm * 8, hipMemcpyHostToDevice));
    DMX_HIP(hipMemcpy(accofs.p, ao.data(), nm * 8, hipMemcpyHostToDevice));
    if (spin > 0) {
      std::vector<SpinStart> ss;
      sht_spin_start(nm - 1, spin, ss);
      sstart.reset(nm * sizeof(SpinStart));
      DMX_HIP(hipMemcpy(sstart.p, ss.data(), nm * sizeof(SpinStart),
                        hipMemcpyHostToDevice));
      args.sstart = sstart.as<SpinStart>();
    }
    args.alm_f64 = vf64 ? 1 : 0;
    args.mstart = mstart.as<int64_t>();
    args.lstride = lstride;
    args.lmax = geo.lmax;
    args.mmax = geo.mmax;
    args.N = geo.N;
    args.nphi = geo.nphi;
    args.mfac = mfac.as<double>();
    args.accofs = accofs.as<int64_t>();
    args.spin = spin;
    args.nalm_c = (int)nalm_c;
    args.deriv1 = mode == 2 ? 1 : 0;
    args.nacc = (int64_t)nacc;
  }
  size_t rows() const { return (size_t)(2 * geo.N - 2); }
  // the doubled-sphere grid's c2c over its rows (theta), in place and
  // unscaled: its factor 1/(2N-2) is applied to the map or the a_lm at the
  // end (theta_scale), every step in between being linear, so that the
  // 90 MB grid is not rewritten for it
  std::shared_ptr<void> theta_fft(void *G, bool forward, hipStream_t s) {
    const int64_t st[2] = {geo.nphi, 1};
    const int64_t shape[2] = {(int64_t)rows(), geo.nphi};
    const int axes[1] = {0};
    return c2c_exec(G, DMX_C128, st, G, st, 2, shape, axes, 1, forward, 1.0,
                    s);
  }
  double theta_scale() const { return 1.0 / (double)rows(); }
  // the points sorted for the NUFFT of this grid
  std::unique_ptr<Plan> plan(const void *loc, hipStream_t s) {
    auto pl = std::make_unique<Plan>(p, s);
    pl->build_index(loc, false, M, false);
    return pl;
  }
};
}  // namespace

namespace {
// The synthesis and its adjoint on device data, with the points sorted
// once for any number of applications (the index of one NUFFT plan serves
// both directions).  alm: complex, component c at c * acomp + index; map:
// real, component c / point i at c * ms0 + i * ms1 (ms0 unused for spin 0).
struct ShtOp {
  ShtJob &job;
  hipStream_t s;
  std::unique_ptr<Plan> pl;
  Buf G, z, acc;
  ShtOp(ShtJob &j, const void *loc_dev, hipStream_t s_) : job(j), s(s_) {
    G.stream = z.stream = acc.stream = s;
    G.reset(job.rows() * job.geo.nphi * 16);
    z.reset(std::max<size_t>(job.M, 1) * 16);
    acc.reset(job.nacc * 16 * (job.spin > 0 ? 2 : 1));
    if (job.M) pl = job.plan(loc_dev, s);
  }
  ~ShtOp() {
    if (pl) pl->mark_done();
  }
  void synth(const void *alm, int64_t acomp, void *map, int64_t ms0,
             int64_t ms1) {
    job.args.alm = alm;
    job.args.acomp = acomp;
    if (job.spin > 0)
      launch_alm2leg_spin(job.args, G.as<double2>(), s);
    else
      launch_alm2leg(job.args, G.as<double2>(), s);
    auto keep = job.theta_fft(G.p, true, s);
    if (!job.M) return;
    const int64_t ust[2] = {job.geo.nphi, 1};
    pl->u2nu(false, G.p, true, ust, z.p, true);
    launch_real_part(z.as<double2>(), job.M, map, job.vf64, ms1,
                     job.spin > 0 ? ms0 : 0, job.theta_scale(), s);
  }
  void adjoint(const void *map, int64_t ms0, int64_t ms1, void *alm,
               int64_t acomp) {
    job.args.acomp = acomp;
    if (job.M) {
      launch_to_complex(map, job.vf64, ms1, job.spin > 0 ? ms0 : 0, job.M,
                        z.as<double2>(), s);
      const int64_t ust[2] = {job.geo.nphi, 1};
      pl->nu2u(true, z.p, true, G.p, true, ust);
    } else {
      launch_zero(G.p, G.bytes, s);
    }
    auto keep = job.theta_fft(G.p, false, s);
    if (leg2alm_needs_zero(job.args)) launch_zero(acc.p, acc.bytes, s);
    if (job.spin > 0) {
      launch_leg2alm_spin(job.args, G.as<double2>(), acc.as<double2>(), s);
      launch_store_alm_spin(acc.as<double2>(), job.args, alm, job.vf64,
                            job.theta_scale(), s);
    } else {
      launch_leg2alm(job.args, G.as<double2>(), acc.as<double2>(), s);
      launch_store_alm(acc.as<double2>(), job.args, alm, job.vf64,
                       job.theta_scale(), s);
    }
  }
};
}  // namespace

int dmx_sht_synthesis_general(const dmx_array *alm, dmx_array *map,
                              size_t spin, size_t lmax,
                              const int64_t *mstart, size_t nm,
                              int64_t lstride, const dmx_array *loc,
                              double epsilon, double sigma_min,
                              double sigma_max, size_t nthreads, int mode,
                              size_t verbosity, const dmx_options *opt_in) {
  (void)nthreads;
  return guard([&] {
    DMX_ASSERT(alm && map && loc, "null argument");
    reap(false);
    dmx_options opt = opt_in ? *opt_in : defaults();
    ShtJob job(*alm, *map, spin, lmax, mstart, nm, lstride, *loc, epsilon,
               sigma_min, sigma_max, mode, opt);
    hipStream_t s = pick_stream(opt);
    DevIn ain(*alm, s, false), lin(*loc, s, false);
    DevOut mo(*map, s, true);
    ShtOp op(job, lin.p, s);
    op.synth(ain.p, ain.st[0], mo.p, mo.st[0], mo.st[1]);
    if (verbosity > 0 && op.pl) op.pl->report_params(false);
    mo.finish(s);
  });
}

int dmx_sht_adjoint_synthesis_general(dmx_array *alm, const dmx_array *map,
                                      size_t spin, size_t lmax,
                                      const int64_t *mstart, size_t nm,
                                      int64_t lstride, const dmx_array *loc,
                                      double epsilon, double sigma_min,
                                      double sigma_max, size_t nthreads,
                                      int mode, size_t verbosity,
                                      const dmx_options *opt_in) {
  (void)nthreads;
  return guard([&] {
    DMX_ASSERT(alm && map && loc, "null argument");
    reap(false);
    dmx_options opt = opt_in ? *opt_in : defaults();
    ShtJob job(*alm, *map, spin, lmax, mstart, nm, lstride, *loc, epsilon,
               sigma_min, sigma_max, mode, opt);
    hipStream_t s = pick_stream(opt);
    DevIn min_(*map, s, true), lin(*loc, s, false);
    DevOut ao(*alm, s, false);
    // entries outside the (l, m) layout of a host array come back zero
    if (!alm->device) launch_zero(ao.p, (size_t)nelem(*alm) * elsize(alm->typecode), s);
    ShtOp op(job, lin.p, s);
    op.adjoint(min_.p, min_.st[0], min_.st[1], ao.p, ao.st[0]);
    if (verbosity > 0 && op.pl) op.pl->report_params(true);
    ao.finish(s);
  });
}

int dmx_sht_pseudo_analysis_general(dmx_array *alm, const dmx_array *map,
                                    size_t spin, size_t lmax,
                                    const int64_t *mstart, size_t nm,
                                    int64_t lstride, const dmx_array *loc,
                                    double sigma_min, double sigma_max,
                                    size_t nthreads, size_t maxiter,
                                    double epsilon, size_t verbosity,
                                    size_t *istop_out, size_t *itn_out,
                                    double *rnorm_out, double *arnorm_out,
                                    const dmx_options *opt_in) {
  (void)nthreads;
  return guard([&] {
    DMX_ASSERT(alm && map && loc, "null argument");
    DMX_ASSERT(epsilon > 0, "epsilon must be positive");
    reap(false);
    dmx_options opt = opt_in ? *opt_in : defaults();
    // sht.cc:3003-3058: LSMR on the synthesis (STANDARD mode) with its
    // adjoint, both at 0.1 epsilon
    ShtJob job(*alm, *map, spin, lmax, mstart, nm, lstride, *loc,
               0.1 * epsilon, sigma_min, sigma_max, 0, opt);
    hipStream_t s = pick_stream(opt);
    DevIn min_(*map, s, true), lin(*loc, s, false);
    DevOut ao(*alm, s, false);
    const bool f64 = job.vf64;
    const int ncomp = job.spin > 0 ? 2 : 1;
    const size_t es = f64 ? 8 : 4;
    const int64_t nalm = alm->shape[1];
    const size_t na = (size_t)ncomp * nalm * 2, nmp = (size_t)ncomp * job.M;
    // LSMR vectors: x, v, h, hbar (a_lm, contiguous, entries outside the
    // layout stay zero), u, w (map, contiguous); b: the map
    Buf x(na * es, s), v(na * es, s), h(na * es, s), hb(na * es, s),
        u(std::max<size_t>(nmp, 1) * es, s), w(std::max<size_t>(nmp, 1) * es, s),
        nrm(64, s);
    for (Buf *b : {&x, &v, &h, &hb}) launch_zero(b->p, b->bytes, s);
    // u = b: the map, made contiguous
    lsmr_copy_map(f64, ncomp, job.M, min_.p, min_.st[0], min_.st[1], u.p, s);
    ShtOp op(job, lin.p, s);
    const int64_t acomp = nalm;
    auto alm_norm = [&](const void *a) {
      launch_zero(nrm.p, 8, s);
      lsmr_alm_sumsq(job.args, f64, a, ncomp, acomp, nrm.as<double>(), s);
      double r = 0;
      DMX_HIP(hipMemcpyAsync(&r, nrm.p, 8, hipMemcpyDeviceToHost, s));
      DMX_HIP(hipStreamSynchronize(s));
      return std::sqrt(r);
    };
    auto map_norm = [&](const void *m) {
      launch_zero(nrm.p, 8, s);
      lsmr_sumsq(f64, nmp, m, nrm.as<double>(), s);
      double r = 0;
      DMX_HIP(hipMemcpyAsync(&r, nrm.p, 8, hipMemcpyDeviceToHost, s));
      DMX_HIP(hipStreamSynchronize(s));
      return std::sqrt(r);
    };
    // A^T u lands in a scratch a_lm vector (the adjoint writes only the
    // layout's entries; the others stay zero)
    Buf at(na * es, s);
    launch_zero(at.p, at.bytes, s);
    auto applyA = [&](const void *vin, void *mout) {
      op.synth(vin, acomp, mout, (int64_t)job.M, 1);
    };
    auto applyAT = [&](const void *min, void *aout) {
      op.adjoint(min, (int64_t)job.M, 1, aout, acomp);
    };
    LsmrResult res = lsmr_solve(
        na, nmp, f64, applyA, applyAT, alm_norm, map_norm, x.p, v.p, h.p,
        hb.p, u.p, w.p, at.p, 1e-14 * std::sqrt((double)job.M), epsilon, 1e8,
        maxiter, s);
    DMX_HIP(hipMemcpyAsync(ao.p, x.p, na * es, hipMemcpyDeviceToDevice, s));
    ao.finish(s);
    if (verbosity > 0)
      std::printf("pseudo_analysis_general: istop %zu, %zu iterations, "
                  "|r|/|b| %.3e\n", res.istop, res.itn, res.rnorm);
    if (istop_out) *istop_out = res.istop;
    if (itn_out) *itn_out = res.itn;
    if (rnorm_out) *rnorm_out = res.rnorm;
    if (arnorm_out) *arnorm_out = res.arnorm;
  });
}

namespace {
// shared validation and set-up of vis2dirty / dirty2vis
// (wgridder_pymod.cc:483-560, wgridder.h:1626-1700)
void wg_common(WgJob &job, const dmx_array &uvw, const dmx_array &freq,
               const dmx_array *wgt, const dmx_array *mask,
               const dmx_array &dirty, int64_t nrow, int64_t nchan,
               double pixsize_x, double pixsize_y, int do_wgridding,
               int flip_u, int flip_v, int flip_w, int divide_by_n,
               double center_x, double center_y, int allow_nshift,
               bool single, hipStream_t s, std::unique_ptr<DevIn> &uin,
               std::unique_ptr<DevIn> &win, std::unique_ptr<DevIn> &min_) {
  DMX_ASSERT(uvw.ndim == 2 && uvw.shape[1] == 3 && uvw.typecode == DMX_F64,
             "uvw must be a (nrow, 3) float64 array");
  DMX_ASSERT(freq.ndim == 1 && freq.typecode == DMX_F64,
             "freq must be a float64 vector");
  DMX_ASSERT(uvw.shape[0] == nrow && freq.shape[0] == nchan,
             "dimension mismatch");
  DMX_ASSERT(dirty.ndim == 2, "dirty must be a 2D array");
  DMX_ASSERT(dirty.typecode == (single ? DMX_F32 : DMX_F64),
             "unsupported combination of data types");
  if (wgt) {
    DMX_ASSERT(wgt->ndim == 2 && wgt->shape[0] == nrow &&
                   wgt->shape[1] == nchan,
               "wgt: dimension mismatch");
    DMX_ASSERT(wgt->typecode == (single ? DMX_F32 : DMX_F64),
               "unsupported combination of data types");
  }
  if (mask) {
    DMX_ASSERT(mask->ndim == 2 && mask->shape[0] == nrow &&
                   mask->shape[1] == nchan,
               "mask: dimension mismatch");
    DMX_ASSERT(mask->typecode == DMX_U8, "mask must be uint8");
  }
  // frequencies on the host for the checks and f / c
  std::vector<double> fh((size_t)nchan);
  {
    dmx_array f = freq;
    if (freq.device) {
      DMX_ASSERT(contiguous(freq), "device arrays must be C-contiguous here");
      if (nchan)
        DMX_HIP(hipMemcpy(fh.data(), freq.data, (size_t)nchan * 8,
                          hipMemcpyDeviceToHost));
    } else {
      pack(f, fh.data());
    }
  }
  uin = std::make_unique<DevIn>(uvw, s, false);
  if (wgt) win = std::make_unique<DevIn>(*wgt, s, true);
  if (mask) min_ = std::make_unique<DevIn>(*mask, s, true);
  WgArgs &a = job.a;
  a.uvw = (const double *)uin->p;
  a.wgt = wgt ? win->p : nullptr;
  a.wgt_f64 = single ? 0 : 1;
  if (wgt) {
    a.wst0 = win->st[0];
    a.wst1 = win->st[1];
  }
  a.mask = mask ? (const uint8_t *)min_->p : nullptr;
  if (mask) {
    a.mst0 = min_->st[0];
    a.mst1 = min_->st[1];
  }
  a.nx = (int)dirty.shape[0];
  a.ny = (int)dirty.shape[1];
  a.px = pixsize_x;
  a.py = pixsize_y;
  a.ufac = flip_u ? -1.0 : 1.0;
  a.vfac = flip_v ? -1.0 : 1.0;
  a.wfac = flip_w ? -1.0 : 1.0;
  a.lshift = flip_u ? -center_x : center_x;
  a.mshift = flip_v ? -center_y : center_y;
  a.wgrid = do_wgridding ? 1 : 0;
  a.divide_by_n = divide_by_n ? 1 : 0;
  a.allow_nshift = allow_nshift ? 1 : 0;
  job.setup(fh.data(), (size_t)nchan, (size_t)nrow, s);
}
}  // namespace

int dmx_vis2dirty(const dmx_array *uvw, const dmx_array *freq,
                  const dmx_array *vis, const dmx_array *wgt,
                  const dmx_array *mask, double pixsize_x, double pixsize_y,
                  double epsilon, int do_wgridding, int flip_u, int flip_v,
                  int flip_w, int divide_by_n, double sigma_min,
                  double sigma_max, double center_x, double center_y,
                  int allow_nshift, size_t nthreads, size_t verbosity,
                  dmx_array *dirty, const dmx_options *opt_in) {
  (void)nthreads;
  (void)verbosity;
  return guard([&] {
    DMX_ASSERT(uvw && freq && vis && dirty, "null argument");
    DMX_ASSERT(epsilon > 0, "epsilon must be positive");
    reap(false);
    dmx_options opt = opt_in ? *opt_in : defaults();
    DMX_ASSERT(vis->ndim == 2 && is_complex(vis->typecode),
               "vis must be a complex (nrow, nchan) array");
    const bool single = vis->typecode == DMX_C64;
    hipStream_t s = pick_stream(opt);
    WgJob job;
    job.a.gridding = 1;
    std::unique_ptr<DevIn> uin, win, min_;
    wg_common(job, *uvw, *freq, wgt, mask, *dirty, vis->shape[0],
              vis->shape[1], pixsize_x, pixsize_y, do_wgridding, flip_u,
              flip_v, flip_w, divide_by_n, center_x, center_y, allow_nshift,
              single, s, uin, win, min_);
    DevIn vin(*vis, s, true);
    DevOut dout(*dirty, s, true);
    job.prep(vin.p, !single, vin.st[0], vin.st[1]);
    if (job.nvis == 0) {
      job.zero_dirty(dout.p, !single, dout.st[0], dout.st[1]);
    } else {
      job.choose(epsilon, sigma_min, sigma_max, !single, opt);
      job.grid(dout.p, !single, dout.st[0], dout.st[1]);
    }
    dout.finish(s);
  });
}

int dmx_dirty2vis(const dmx_array *uvw, const dmx_array *freq,
                  const dmx_array *dirty, const dmx_array *wgt,
                  const dmx_array *mask, double pixsize_x, double pixsize_y,
                  double epsilon, int do_wgridding, int flip_u, int flip_v,
                  int flip_w, int divide_by_n, double sigma_min,
                  double sigma_max, double center_x, double center_y,
                  int allow_nshift, size_t nthreads, size_t verbosity,
                  dmx_array *vis, const dmx_options *opt_in) {
  (void)nthreads;
  (void)verbosity;
  return guard([&] {
    DMX_ASSERT(uvw && freq && vis && dirty, "null argument");
    DMX_ASSERT(epsilon > 0, "epsilon must be positive");
    reap(false);
    dmx_options opt = opt_in ? *opt_in : defaults();
    DMX_ASSERT(vis->ndim == 2 && is_complex(vis->typecode),
               "vis must be a complex (nrow, nchan) array");
    const bool single = vis->typecode == DMX_C64;
    hipStream_t s = pick_stream(opt);
    WgJob job;
    job.a.gridding = 0;
    std::unique_ptr<DevIn> uin, win, min_;
    wg_common(job, *uvw, *freq, wgt, mask, *dirty, vis->shape[0],
              vis->shape[1], pixsize_x, pixsize_y, do_wgridding, flip_u,
              flip_v, flip_w, divide_by_n, center_x, center_y, allow_nshift,
              single, s, uin, win, min_);
    DevIn din(*dirty, s, true);
    DevOut vout(*vis, s, true);
    job.prep(nullptr, !single, 0, 0);
    if (job.M) {
      // (inactive visibilities, all of them when nvis == 0, come back zero)
      job.choose(epsilon, sigma_min, sigma_max, !single, opt);
      job.degrid(din.p, !single, din.st[0], din.st[1], vout.p, !single,
                 vout.st[0], vout.st[1]);
    }
    vout.finish(s);
  });
}

int dmx_nudft(int type, const dmx_array *coord, const dmx_array *values,
              const int64_t *shape, size_t ndim, int forward,
              const double *periodicity, int fft_order, const int64_t *sel,
              size_t nsel, dmx_array *out) {
  return guard([&] {
    DMX_ASSERT(coord && values && shape && periodicity && out, "null argument");
    DMX_ASSERT(type == 1 || type == 2, "type must be 1 or 2");
    DMX_ASSERT(coord->typecode == DMX_F64 && values->typecode == DMX_C128 &&
                   out->typecode == DMX_C128,
               "direct NUDFT works on float64 / complex128");
    DMX_ASSERT(ndim >= 1 && ndim <= 3 && coord->ndim == 2 &&
                   coord->shape[1] == (int64_t)ndim,
               "dimensionality mismatch");
    DMX_ASSERT(nelem(*out) == (int64_t)nsel, "output size mismatch");
    hipStream_t s = library_stream();
    DevIn c(*coord, s, false), v(*values, s, false);
    DevOut o(*out, s, false);
    Buf selb(nsel * 8);
    if (nsel) {
      DMX_HIP(hipMemcpyAsync(selb.p, sel, nsel * 8, hipMemcpyHostToDevice, s));
      DMX_HIP(hipStreamSynchronize(s));
    }
    NudftArgs a{};
    a.type = type;
    a.ndim = (int)ndim;
    a.forward = forward;
    a.fft_order = fft_order;
    a.M = coord->shape[0];
    a.nsel = (int64_t)nsel;
    for (int d = 0; d < 3; ++d) {
      a.n[d] = d < (int)ndim ? shape[d] : 1;
      a.period[d] = d < (int)ndim ? periodicity[d] : 1.;
    }
    a.coord = (const double *)c.p;
    a.vals = (const double *)v.p;
    a.sel = selb.as<int64_t>();
    a.out = (double *)o.p;
    launch_nudft(a, s);
    o.finish(s);
  });
}

namespace {
// parameters of a stand-alone sort with the given geometry (the sort never
// evaluates the kernel; any kernel-table row of matching W serves)
Params sort_params(const dmx_array *coord, const int64_t *nover, size_t ndim,
                   int W, int log2tile, const double *periodicity) {
  DMX_ASSERT(coord && nover && periodicity, "null argument");
  check_coord(*coord);
  DMX_ASSERT(ndim >= 1 && ndim <= 3 && coord->shape[1] == (int64_t)ndim,
             "dimensionality mismatch");
  DMX_ASSERT(W >= 1 && W <= 16 && log2tile >= 0 && log2tile <= 12,
             "bad sort geometry");
  Params p;
  p.ndim = (int)ndim;
  p.W = W;
  p.log2tile = log2tile;
  for (size_t d = 0; d < ndim; ++d) {
    p.nover[d] = (size_t)nover[d];
    p.nuni[d] = (size_t)nover[d] / 2;
    p.period[d] = periodicity[d];
    DMX_ASSERT(nover[d] >= 2 * ((W + 1) / 2), "oversampled length too small");
  }
  const auto &db = kernel_db();
  for (size_t i = 0; i < db.size(); ++i)
    if (db[i].W == W) {
      p.beta = db[i].beta;
      p.e0 = db[i].e0;
      break;
    }
  DMX_ASSERT(p.beta > 0, "no kernel of this support");
  return p;
}
}  // namespace

int dmx_sort_points(const dmx_array *coord, const int64_t *nover,
                    size_t ndim, int W, int log2tile,
                    const double *periodicity, dmx_array *keys,
                    dmx_array *idx) {
  return guard([&] {
    DMX_ASSERT(keys && idx, "null argument");
    Params p = sort_params(coord, nover, ndim, W, log2tile, periodicity);
    DMX_ASSERT(keys->typecode == DMX_U32 && idx->typecode == DMX_U32,
               "keys and idx must be uint32");
    const size_t M = (size_t)coord->shape[0];
    DMX_ASSERT(nelem(*keys) == (int64_t)M && nelem(*idx) == (int64_t)M,
               "number of points mismatch");
    hipStream_t s = library_stream();
    Plan plan(p, s);
    DevIn cin(*coord, s, false);
    plan.build_index(cin.p, coord->typecode == DMX_F32, M, false);
    DevOut ko(*keys, s, false), io(*idx, s, false);
    if (M) {
      plan.keys_copy_to((uint32_t *)ko.p);
      DMX_HIP(hipMemcpyAsync(io.p, plan.idx(), M * 4, hipMemcpyDeviceToDevice,
                             s));
    }
    ko.finish(s);
    io.finish(s);
  });
}

int dmx_sort_records(const dmx_array *coord, const int64_t *nover,
                     size_t ndim, int W, int log2tile,
                     const double *periodicity, dmx_array *words,
                     dmx_array *idx, dmx_array *xarg) {
  return guard([&] {
    DMX_ASSERT(words && idx, "null argument");
    Params p = sort_params(coord, nover, ndim, W, log2tile, periodicity);
    p.fp64 = true;
    p.use_lds = true;
    DMX_ASSERT(words->typecode == DMX_U32 && idx->typecode == DMX_U32,
               "words and idx must be uint32");
    const size_t M = (size_t)coord->shape[0];
    DMX_ASSERT(nelem(*words) == (int64_t)M && nelem(*idx) == (int64_t)M,
               "number of points mismatch");
    if (xarg)
      DMX_ASSERT(xarg->typecode == DMX_F64 && xarg->ndim == 2 &&
                     xarg->shape[0] == (int64_t)M &&
                     xarg->shape[1] == (int64_t)ndim,
                 "xarg must be a (M, ndim) float64 array");
    hipStream_t s = library_stream();
    Plan plan(p, s);
    DevIn cin(*coord, s, false);
    DevOut wo(*words, s, false), io(*idx, s, false);
    std::unique_ptr<DevOut> xo;
    if (xarg) xo = std::make_unique<DevOut>(*xarg, s, false);
    if (M) {
      DMX_ASSERT(plan.build_records(cin.p, coord->typecode == DMX_F32, M,
                                    false, nullptr, false),
                 "the record layout does not apply to this geometry");
      const RecBufs &r = plan.records();
      DMX_HIP(hipMemcpyAsync(wo.p, r.word, M * 4, hipMemcpyDeviceToDevice, s));
      DMX_HIP(hipMemcpyAsync(io.p, r.j, M * 4, hipMemcpyDeviceToDevice, s));
      if (xo)
        for (size_t d = 0; d < ndim; ++d)
          DMX_HIP(hipMemcpy2DAsync((char *)xo->p + 8 * d, 8 * ndim, r.x[d], 8,
                                   8, M, hipMemcpyDeviceToDevice, s));
    }
    wo.finish(s);
    io.finish(s);
    if (xo) xo->finish(s);
  });
}

int dmx_profile_enable(int on) {
  return guard([&] { Profiler::get().enable(on != 0); });
}

int dmx_profile_reset(void) {
  return guard([&] {
    reap(true);
    Profiler::get().reset();
  });
}

int dmx_profile_read(char *buf, size_t buflen) {
  return guard([&] {
    DMX_ASSERT(buf && buflen, "null argument");
    reap(true);
    std::string r = Profiler::get().report();
    DMX_ASSERT(r.size() + 1 <= buflen, "profile buffer too small");
    std::memcpy(buf, r.c_str(), r.size() + 1);
  });
}

int dmx_c2c(const dmx_array *in, const dmx_array *out, const int64_t *axes,
            size_t naxes, int forward, double fct, size_t nthreads,
            const dmx_options *opt_in) {
  (void)nthreads;
  return guard([&] {
    DMX_ASSERT(in && out && (axes || !naxes), "null argument");
    reap(false);
    dmx_options opt = opt_in ? *opt_in : defaults();
    const int ndim = (int)in->ndim;
    DMX_ASSERT(ndim >= 1 && ndim <= 10, "bad array dimensionality");
    DMX_ASSERT(out->ndim == ndim, "dimension mismatch");
    for (int d = 0; d < ndim; ++d)
      DMX_ASSERT(out->shape[d] == in->shape[d], "dimension mismatch");
    const bool fp64 = out->typecode == DMX_C128;
    DMX_ASSERT(fp64 || out->typecode == DMX_C64, "incorrect data type");
    DMX_ASSERT(fp64 ? (in->typecode == DMX_C128 || in->typecode == DMX_F64)
                    : (in->typecode == DMX_C64 || in->typecode == DMX_F32),
               "incorrect data type");
    // makeaxes (python/fft_pymod.cc:73-94)
    DMX_ASSERT(naxes >= 1 && naxes <= (size_t)ndim, "bad axes argument");
    int ax[10];
    for (size_t q = 0; q < naxes; ++q) {
      int64_t a = axes[q];
      if (a < 0) a += ndim;
      DMX_ASSERT(a >= 0 && a < ndim, "axes exceeds dimensionality of output");
      ax[q] = (int)a;
    }
    hipStream_t s = pick_stream(opt);
    const bool async = (opt.flags & DMX_ASYNC) && in->device && out->device;
    DevIn din(*in, s, true);
    DevOut dout(*out, s, true);
    auto job = c2c_exec(din.p, in->typecode, din.st, dout.p, dout.st, ndim,
                        in->shape, ax, (int)naxes, forward != 0, fct, s);
    dout.finish(s, !async);
    if (async) retire(nullptr, s, job);
  });
}

int dmx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

}  // extern "C"
