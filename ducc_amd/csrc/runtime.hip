// runtime.hip -- host runtime of the NUFFT: memory pool, parameter choice,
// and the plan executor that sequences the gfx950 kernels.
//
// Type 1 (nu2u):  zero grid -> spread (LDS tiles, atomic flush) -> pruned
//                 FFT (last axis over all lines, then each earlier axis over
//                 the kept bands only) -> crop + deconvolve.
// Type 2 (u2nu):  deconvolve + pad (zeros only where an FFT reads whole
//                 lines) -> pruned FFT in reverse order -> interpolate (LDS
//                 tiles).
// (nufft.h:1274-1362, 836-890, 1731-1860.)
#include "runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "chooser.h"
#include "grid.h"
#include "params.h"

namespace dmx {

// ----------------------------------------------------------------- arena
namespace {
thread_local Arena *t_arena = nullptr;
}
void *Arena::take(size_t bytes) {
  const size_t n = round256(bytes);
  if (used + n > size)
    throw Error("scratch exhausted: the request needs more than the " +
                std::to_string(size >> 20) + " MiB of static scratch");
  void *p = base + used;
  used += n;
  return p;
}
ArenaScope::ArenaScope(Arena *a) : prev_(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev_; }
Arena *current_arena() { return t_arena; }

// ------------------------------------------------------------------ pool
Pool &Pool::get() {
  static Pool *p = new Pool();  // never destroyed: outlives static teardown
  return *p;
}

void Pool::collect_locked() {
  size_t k = 0;
  for (size_t i = 0; i < pending_.size(); ++i) {
    Pending &q = pending_[i];
    const hipError_t e = hipEventQuery(q.ev);
    if (e == hipErrorNotReady) {
      (void)hipGetLastError();  // "not ready" is not an error here
      pending_[k++] = q;
      continue;
    }
    auto it = live_.find(q.p);
    if (it != live_.end()) {
      free_.insert({it->second, q.p});
      live_.erase(it);
    }
    spare_events_.push_back(q.ev);
  }
  pending_.resize(k);
}

void *Pool::alloc(size_t bytes) {
  if (Arena *a = t_arena) return a->take(bytes);
  int dev = 0;
  DMX_HIP(hipGetDevice(&dev));
  bytes = (bytes + 255) & ~(size_t)255;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (!pending_.empty()) collect_locked();
    auto it = free_.lower_bound({dev, bytes});
    if (it != free_.end() && it->first.first == dev &&
        it->first.second <= 2 * bytes) {
      void *p = it->second;
      live_[p] = it->first;
      free_.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // release cached blocks and retry once
    trim();
    (void)hipGetLastError();
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess)
    throw Error(std::string("device memory allocation of ") +
                std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(mu_);
  live_[p] = {dev, bytes};
  return p;
}

void Pool::release(void *p, hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  if (s) {
    // the block is used by work queued on s: it becomes free once an event
    // recorded now has completed (the block stays in live_ until then)
    hipEvent_t ev = nullptr;
    if (!spare_events_.empty()) {
      ev = spare_events_.back();
      spare_events_.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) !=
               hipSuccess) {
      ev = nullptr;
    }
    if (ev && hipEventRecord(ev, s) == hipSuccess) {
      pending_.push_back({ev, p});
      return;
    }
    // no event: wait for the stream instead
    (void)hipGetLastError();
    if (ev) spare_events_.push_back(ev);
    (void)hipStreamSynchronize(s);
  }
  free_.insert({it->second, p});
  live_.erase(it);
}

void Pool::trim() {
  std::lock_guard<std::mutex> lk(mu_);
  (void)hipDeviceSynchronize();
  collect_locked();
  for (auto &kv : free_) (void)hipFree(kv.second);
  free_.clear();
}

// -------------------------------------------------------------- profiler
Profiler &Profiler::get() {
  static Profiler *p = new Profiler();
  return *p;
}
void Profiler::add(const std::string &name, double ms) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = acc_.find(name);
  if (it == acc_.end()) {
    order_.push_back(name);
    acc_[name] = {ms, 1};
  } else {
    it->second.first += ms;
    it->second.second += 1;
  }
}
void Profiler::reset() {
  std::lock_guard<std::mutex> lk(mu_);
  acc_.clear();
  order_.clear();
}
std::string Profiler::report() {
  std::lock_guard<std::mutex> lk(mu_);
  std::string s;
  for (auto &n : order_) {
    auto &v = acc_[n];
    s += n + "\t" + std::to_string(v.first) + "\t" + std::to_string(v.second) +
         "\n";
  }
  return s;
}

Plan::Phase::Phase(Plan *pl, std::string n) : p(pl) {
  // no events in arena mode: that caller allocates nothing
  if ((!Profiler::get().enabled() && !p->timing_) || current_arena()) return;
  name = p->tag_ + "/" + n;
  DMX_HIP(hipEventCreate(&a));
  DMX_HIP(hipEventCreate(&b));
  DMX_HIP(hipEventRecord(a, p->s_));
}
Plan::Phase::~Phase() {
  if (!a) return;
  (void)hipEventRecord(b, p->s_);
  p->pend_.push_back({name, a, b});
}

void Plan::flush_profile(bool wait) {
  size_t n = 0;
  for (; n < pend_.size(); ++n) {
    auto &q = pend_[n];
    if (!wait && hipEventQuery(q.b) != hipSuccess) {
      (void)hipGetLastError();  // "not ready" is not an error here
      break;
    }
    float ms = 0;
    if (hipEventSynchronize(q.b) == hipSuccess &&
        hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) {
      if (Profiler::get().enabled()) Profiler::get().add(q.name, ms);
      if (timing_) timed_.push_back({q.name, ms});
    }
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  pend_.erase(pend_.begin(), pend_.begin() + n);
}

void Plan::use_stream(hipStream_t s) {
  if (s == s_) return;
  if (!last_) DMX_HIP(hipEventCreateWithFlags(&last_, hipEventDisableTiming));
  DMX_HIP(hipEventRecord(last_, s_));
  DMX_HIP(hipStreamWaitEvent(s, last_, 0));
  s_ = s;
}

void Plan::start_timing() {
  flush_profile();  // earlier phases are not part of this call
  timing_ = true;
  timed_.clear();
  if (!t0_) DMX_HIP(hipEventCreate(&t0_));
  DMX_HIP(hipEventRecord(t0_, s_));
}

namespace {
std::string dims(const size_t *n, int nd) {
  std::string s;
  for (int d = 0; d < nd; ++d) s += (d ? ", " : "") + std::to_string(n[d]);
  return s;
}
}  // namespace

void Plan::report_params(bool type1) const {
  size_t ngrid = 1;
  for (int d = 0; d < p_.ndim; ++d) ngrid *= p_.nover[d];
  std::printf("%s:\n  device, grid=(%s), oversampled grid=(%s), supp=%d, "
              "eps=%g\n  npoints=%zu\n  memory overhead: %gGB (index) + "
              "%gGB (oversampled grid)\n  %s path, kernel row %d, sigma=%g, "
              "tile=%d^%d\n",
              type1 ? "Nu2u" : "U2nu", dims(p_.nuni, p_.ndim).c_str(),
              dims(p_.nover, p_.ndim).c_str(), p_.W, p_.eps, M_,
              M_ * 4.0 / (1 << 30),
              ngrid * (p_.fp64 ? 16.0 : 8.0) / (1 << 30),
              p_.fp64 ? "FP64" : "FP32", p_.kidx, p_.ofactor,
              1 << p_.log2tile, p_.ndim);
  std::fflush(stdout);
}

void Plan::report_timings(const char *what) {
  hipEvent_t t1;
  DMX_HIP(hipEventCreate(&t1));
  DMX_HIP(hipEventRecord(t1, s_));
  DMX_HIP(hipEventSynchronize(t1));
  float total_ms = 0;
  (void)hipEventElapsedTime(&total_ms, t0_, t1);
  (void)hipEventDestroy(t1);
  flush_profile();
  timing_ = false;
  // phases by decreasing time; the remainder (launch gaps, staging, work
  // overlapped on the second stream) as "<unaccounted>"
  std::map<std::string, double> acc;
  for (auto &ph : timed_) {
    const size_t cut = ph.first.find('/');
    acc[cut == std::string::npos ? ph.first : ph.first.substr(cut + 1)] +=
        ph.second * 1e-3;
  }
  std::vector<std::pair<std::string, double>> v(acc.begin(), acc.end());
  std::sort(v.begin(), v.end(),
            [](const auto &a, const auto &b) { return a.second > b.second; });
  const double total = total_ms * 1e-3;
  size_t w = 13;  // length of "<unaccounted>"
  for (auto &e : v) w = std::max(w, e.first.size());
  std::printf("\nTotal wall clock time for %s: %.4gs\n|\n", what, total);
  double sum = 0;
  auto line = [&](const std::string &n, double t) {
    std::printf("+- %-*s: %6.2f%% (%.4fs)\n", (int)w, n.c_str(),
                total > 0 ? 100 * t / total : 0., t);
  };
  for (auto &e : v) {
    line(e.first, e.second);
    sum += e.second;
  }
  if (sum < 0.999 * total) line("<unaccounted>", total - sum);
  std::fflush(stdout);
  timed_.clear();
}

// -------------------------------------------------------- parameter choice
int default_log2tile(int ndim, bool fp64) {
  (void)fp64;
  if (ndim == 1) return 9;
  // 2-D: 32 x 32 in both precisions (FP64 at 16 x 16 flushed and staged
  // regions of 3x the tile's cells at W = 13, 1.9x at 32 x 32)
  if (ndim == 2) return 5;
  return 3;
}

namespace {
// LDS of the tile region: FP64 accumulators and point staging (spreading;
// interpolation stages the same region in the compute type)
size_t lds_bytes(int ndim, int tile, int W, bool fp64) {
  size_t S = (size_t)tile + W - 1, c = S, ci = (S + 3) & ~(size_t)1;
  for (int d = 1; d < ndim; ++d) {
    c *= S;
    ci *= S;
  }
  return std::max(spread_lds_bytes(c, ndim, fp64), ci * (fp64 ? 16 : 8));
}
constexpr size_t LDS_LIMIT = LDS_MAX_BYTES;

// reference limits: nufft.h:225-228, 400, 409 with its own tile sizes
void reference_limits(const Params &p) {
  int l2 = p.ndim == 1 ? 9 : (p.ndim == 2 ? (p.fp64 ? 4 : 5) : 4);
  double mx = p.ndim == 1 ? 4294967285.0 : (p.ndim == 2 ? 65526.0 : 1014.0);
  for (int d = 0; d < p.ndim; ++d)
    DMX_ASSERT((double)(p.nover[d] >> l2) <= mx, "oversampled grid too large");
}
}  // namespace

Params choose_params(int ndim, const size_t *nuni, size_t M, bool gridding,
                     double eps, double smin, double smax,
                     const double *period, bool fft_order, bool values_f64,
                     const dmx_options &opt) {
  DMX_ASSERT(ndim >= 1 && ndim <= 3, "transform must be 1D/2D/3D");
  DMX_ASSERT(M <= 0xFFFFFFFFull, "too many nonuniform points");
  DMX_ASSERT(eps > 0, "epsilon must be positive");
  Params p;
  p.ndim = ndim;
  for (int d = 0; d < ndim; ++d) {
    p.nuni[d] = nuni[d];
    p.period[d] = period[d];
    DMX_ASSERT(period[d] > 0, "periodicity must be positive");
  }
  p.fft_order = fft_order;
  p.eps = eps;
  p.deterministic = opt.deterministic != 0;
  if (opt.precision == 2)
    p.fp64 = true;
  else if (opt.precision == 1)
    p.fp64 = false;
  else  // the value precision, raised to FP64 for tight tolerances
    p.fp64 = values_f64 || eps < opt.fp64_below ||
             available_kernels(eps, ndim, true, smin, smax).empty();
  const auto &db = kernel_db();
  int kidx = -1;
  int l2t = 0;
  if (opt.kernel_idx >= 0) {
    DMX_ASSERT(opt.kernel_idx < (int)db.size(), "no appropriate kernel found");
    kidx = opt.kernel_idx;
    const KRow &k = db[kidx];
    DMX_ASSERT(k.ndim == ndim, "kernel table row has the wrong dimensionality");
    DMX_ASSERT(p.fp64 || k.W <= 8, "FP32 path supports W <= 8");
    double of = opt.sigma > 0 ? opt.sigma : k.ofactor;
    for (int d = 0; d < ndim; ++d) p.nover[d] = oversampled_length(nuni[d], of);
  } else if (opt.chooser == 1) {
    LearnedChoice lc;
    if (learned_choose(ndim, nuni, M, eps, smin, smax, p.fp64, gridding, lc)) {
      kidx = lc.kidx;
      l2t = lc.log2tile;
      for (int d = 0; d < ndim; ++d) p.nover[d] = lc.nover[d];
      p.chooser_used = 1;
      p.chooser_verified = lc.verified;
    }
  }
  if (kidx < 0) {
    kidx = find_params(eps, smin, smax, ndim, nuni, M, gridding, !p.fp64, 8, 8,
                       p.nover);
    DMX_ASSERT(kidx >= 0, "no appropriate kernel found");
  }
  finish_params(p, kidx, l2t, opt, (double)M);
  return p;
}

void finish_params(Params &p, int kidx, int l2t, const dmx_options &opt,
                   double npts) {
  const int ndim = p.ndim;
  const KRow &k = kernel_db()[kidx];
  p.kidx = kidx;
  p.check_offsets = (opt.flags & DMX_CHECK_OFFSETS) != 0;
  p.spread_fix = !p.fp64 && p.eps >= FIX_MIN_EPS &&
                 !(opt.flags & DMX_SPREAD_FP64_ACC);
  p.W = k.W;
  p.beta = k.beta;
  p.e0 = k.e0;
  p.ofactor = k.ofactor;
  p.eps_tab = k.epsilon;
  reference_limits(p);
  for (int d = 0; d < ndim; ++d)
    DMX_ASSERT(p.nover[d] >= 2 * (size_t)((p.W + 1) / 2),
               "oversampled length too small");
  // tile: requested, else the default shrunk until the LDS region fits
  if (opt.log2tile > 0) l2t = opt.log2tile;
  int def = default_log2tile(ndim, p.fp64);
  if (l2t <= 0) {
    l2t = def;
    // sparse 2-D FP32: below one point per sixteen oversampled cells a
    // 64 x 64 tile mostly stays within one subproblem (MAXSUB points) and
    // its region's halo is a smaller share of the cells flushed and staged
    // (denser sets split such tiles and flush their 70^2 regions several
    // times over: clustered points at one per ten cells on average already
    // do; FP64 regions of 64^2 leave one workgroup per CU)
    if (ndim == 2 && !p.fp64 && npts > 0 &&
        16 * npts < (double)p.nover[0] * (double)p.nover[1])
      l2t = 6;
  }
  DMX_ASSERT(l2t >= 2 && l2t <= 10, "log2tile out of range");
  p.use_lds = true;
  while (lds_bytes(ndim, 1 << l2t, p.W, p.fp64) > LDS_LIMIT && l2t > 2) --l2t;
  if (lds_bytes(ndim, 1 << l2t, p.W, p.fp64) > LDS_LIMIT) {
    DMX_ASSERT(ndim == 3, "internal: LDS region too large");
    p.use_lds = false;
  }
  p.log2tile = l2t;
  size_t ntot = 1;
  for (int d = 0; d < ndim; ++d) ntot *= ((p.nover[d] >> l2t) + 3);
  DMX_ASSERT(ntot < 0xFFFFFFFFull, "oversampled grid too large");
  // 3-D regions are (tile + W - 1)^3 cells around tile^3: flushing them
  // with atomics costs about half of the spreading, storing them and
  // folding the overlaps costs a fraction of that
  p.slab_flush = ndim == 3 && p.use_lds && !p.deterministic &&
                 !opt.deterministic;
}

Params kernel_params(int ndim, const size_t *nuni, const double *period,
                     bool fft_order, bool fp64, double eps, int kidx,
                     double ofactor, const dmx_options &opt, double npts) {
  DMX_ASSERT(ndim >= 1 && ndim <= 3, "transform must be 1D/2D/3D");
  DMX_ASSERT(kidx >= 0 && kidx < (int)kernel_db().size(),
             "no appropriate kernel found");
  DMX_ASSERT(fp64 || kernel_db()[kidx].W <= 8, "FP32 path supports W <= 8");
  Params p;
  p.ndim = ndim;
  for (int d = 0; d < ndim; ++d) {
    p.nuni[d] = nuni[d];
    p.period[d] = period[d];
    p.nover[d] = oversampled_length(nuni[d], ofactor);
  }
  p.fft_order = fft_order;
  p.eps = eps;
  p.fp64 = fp64;
  finish_params(p, kidx, 0, opt, npts);
  p.ofactor = ofactor;
  return p;
}

// ------------------------------------------------------------------ plan
Plan::Plan(const Params &p, hipStream_t s) : p_(p), s_(s) {
  g_.ndim = p.ndim;
  g_.W = p.W;
  g_.nsafe = (p.W + 1) / 2;
  g_.log2tile = p.log2tile;
  for (int d = 0; d < 3; ++d) {
    if (d < p.ndim) {
      g_.nover[d] = (int)p.nover[d];
      g_.ntiles[d] = (int)((p.nover[d] >> p.log2tile) + 3);
      g_.maxi0[d] = (int)(p.nover[d] + g_.nsafe) - p.W;
      g_.coordfct[d] = 1. / p.period[d];
      g_.shift[d] = p.W * (-0.5) + 1 + (double)p.nover[d];
    } else {
      g_.nover[d] = 1;
      g_.ntiles[d] = 1;
      g_.maxi0[d] = 0;
      g_.coordfct[d] = 1.;
      g_.shift[d] = 0.;
    }
  }
  lds_side_ = (1 << p.log2tile) + p.W - 1;
  ntiles_ = (size_t)g_.ntiles[0] * g_.ntiles[1] * g_.ntiles[2];
  su_ = current_arena() ? arena_setup(p, s) : setup_for(p);
  if (p.check_offsets) {
    DMX_ASSERT(!current_arena(),
               "offset checks are not available without allocation");
    offcheck_.stream = s;
    offcheck_.reset(4);
    launch_zero(offcheck_.p, 4, s);
  }
}

Plan::~Plan() {
  if (!done_) (void)hipStreamSynchronize(s_);
  flush_profile();
  if (t0_) (void)hipEventDestroy(t0_);
  if (last_) (void)hipEventDestroy(last_);
}

namespace {
// Device constants of a parameter set (kernel coefficients in the compute
// type, FP64 correction factors per axis) and FFT twiddles per length,
// built once and kept in small per-process caches: a repeated request
// touches no host-device copy and waits on no stream.
struct SetupCache {
  std::mutex mu;
  std::map<std::string, std::shared_ptr<Setup>> m;
  std::vector<std::string> order;
  std::map<std::string, std::shared_ptr<Buf>> tw;
};
SetupCache &setup_cache() {
  static SetupCache *c = new SetupCache();
  return *c;
}
constexpr size_t SETUP_CACHE_MAX = 64;
}  // namespace

std::shared_ptr<Setup> Plan::setup_for(const Params &p) {
  int dev = 0;
  DMX_HIP(hipGetDevice(&dev));
  std::string key = std::to_string(dev) + ":" + std::to_string(p.kidx) + ":" +
                    std::to_string(p.W) + ":" + (p.fp64 ? "d" : "s");
  char b[64];
  std::snprintf(b, sizeof(b), ":%.17g:%.17g", p.beta, p.e0);
  key += b;
  for (int d = 0; d < p.ndim; ++d)
    key += ":" + std::to_string(p.nuni[d]) + "/" + std::to_string(p.nover[d]);
  auto &C = setup_cache();
  std::lock_guard<std::mutex> lk(C.mu);
  auto it = C.m.find(key);
  if (it != C.m.end()) return it->second;
  auto su = std::make_shared<Setup>();
  // kernel coefficients in the compute type
  auto c = device_coeffs(p.W, p.beta, p.e0);
  if (p.fp64) {
    su->coef.reset(c.size() * 8);
    DMX_HIP(hipMemcpy(su->coef.p, c.data(), c.size() * 8,
                      hipMemcpyHostToDevice));
  } else {
    std::vector<float> cf(c.begin(), c.end());
    su->coef.reset(cf.size() * 4);
    DMX_HIP(hipMemcpy(su->coef.p, cf.data(), cf.size() * 4,
                      hipMemcpyHostToDevice));
  }
  // correction factors, FP64 on the device
  CorfacArgs ca{};
  ca.nq = corr_quadrature(p.W, p.beta * p.W, p.e0, ca.x, ca.wpsi);
  ca.W = p.W;
  for (int d = 0; d < p.ndim; ++d) {
    bool reuse = d > 0 && p.nuni[d] == p.nuni[d - 1] &&
                 p.nover[d] == p.nover[d - 1];
    if (reuse) continue;
    ca.n = (int)(p.nuni[d] / 2 + 1);
    ca.dx = 1. / (double)p.nover[d];
    su->corfac[d].reset((size_t)ca.n * 8);
    launch_corfac(ca, su->corfac[d].as<double>(), nullptr);
  }
  static const double one = 1.0;
  su->one.reset(8);
  DMX_HIP(hipMemcpy(su->one.p, &one, 8, hipMemcpyHostToDevice));
  DMX_HIP(hipDeviceSynchronize());
  C.m[key] = su;
  C.order.push_back(key);
  if (C.order.size() > SETUP_CACHE_MAX) {
    C.m.erase(C.order.front());  // plans in flight keep their shared_ptr
    C.order.erase(C.order.begin());
  }
  return su;
}

std::shared_ptr<Setup> Plan::arena_setup(const Params &p, hipStream_t s) {
  auto su = std::make_shared<Setup>();
  auto c = device_coeffs(p.W, p.beta, p.e0);
  su->coef.reset(c.size() * (p.fp64 ? 8 : 4));
  launch_store_consts(su->coef.p, c.data(), (int)c.size(), !p.fp64, s);
  CorfacArgs ca{};
  ca.nq = corr_quadrature(p.W, p.beta * p.W, p.e0, ca.x, ca.wpsi);
  ca.W = p.W;
  for (int d = 0; d < p.ndim; ++d) {
    if (d > 0 && p.nuni[d] == p.nuni[d - 1] && p.nover[d] == p.nover[d - 1])
      continue;
    ca.n = (int)(p.nuni[d] / 2 + 1);
    ca.dx = 1. / (double)p.nover[d];
    su->corfac[d].reset((size_t)ca.n * 8);
    launch_corfac(ca, su->corfac[d].as<double>(), s);
  }
  const double one = 1.0;
  su->one.reset(8);
  launch_store_consts(su->one.p, &one, 1, false, s);
  return su;
}

namespace {
// the distinct FFT lengths whose twiddles a transform uses: each axis's
// length, plus the two factors of the two-pass column FFT
void fft_lengths(const Params &p, bool has_scratch, std::vector<int> &Ls) {
  const int maxel = p.fp64 ? FFT_MAXEL64 : FFT_MAXEL32;
  auto add = [&](int L) {
    if (std::find(Ls.begin(), Ls.end(), L) == Ls.end()) Ls.push_back(L);
  };
  for (int d = 0; d < p.ndim; ++d) {
    const int L = (int)p.nover[d];
    if (L <= 1) continue;
    if (p.ndim == 2 && d == 0 && maxel / L < 16 && fft_split(L) > 0 &&
        has_scratch) {
      add(fft_split(L));
      add(L / fft_split(L));
    }
    add(L);
  }
}
bool needs_scratch_grid(const Params &p) {
  bool need = p.ndim == 2;
  for (int d = 0; d < p.ndim; ++d)
    if ((int)p.nover[d] > (p.fp64 ? FFT_MAXEL64 : FFT_MAXEL32)) need = true;
  return need;
}
}  // namespace

size_t Plan::scratch_bytes(const Params &p, size_t M) {
  const size_t esz = p.fp64 ? 16 : 8;
  size_t b = 0;
  // constants (arena_setup)
  b += round256(device_coeffs(p.W, p.beta, p.e0).size() * (p.fp64 ? 8 : 4));
  for (int d = 0; d < p.ndim; ++d) {
    if (d > 0 && p.nuni[d] == p.nuni[d - 1] && p.nover[d] == p.nover[d - 1])
      continue;
    b += round256((p.nuni[d] / 2 + 1) * 8);
  }
  b += round256(8);
  // index build (build_index)
  size_t ntiles = 1;
  for (int d = 0; d < p.ndim; ++d) ntiles *= (p.nover[d] >> p.log2tile) + 3;
  if (M) {
    const size_t nsub_max = std::min(ntiles, M) + (M + MAXSUB - 1) / MAXSUB;
    b += round256(M * 4) + round256(ntiles * 4 * 5) +
         round256(std::max(radix_sort_workspace(M), scan_workspace(ntiles))) +
         round256(M * 4) + round256(nsub_max * sizeof(Sub)) + round256(4);
  }
  // grid side (prepare_grid_side)
  size_t n = 1;
  for (int d = 0; d < p.ndim; ++d) n *= p.nover[d];
  const bool scr = needs_scratch_grid(p);
  b += round256(n * esz) * (scr ? 2 : 1);
  // twiddles (twiddle_for)
  std::vector<int> Ls;
  fft_lengths(p, scr, Ls);
  for (int L : Ls) b += round256((size_t)L * esz);
  return b;
}

const void *Plan::twiddle_for(int L) {
  if (current_arena()) {
    auto it = arena_tw_.find(L);
    if (it != arena_tw_.end()) return it->second;
    auto buf = std::make_shared<Buf>((size_t)L * (p_.fp64 ? 16 : 8));
    launch_fft_twiddles(buf->p, L, p_.fp64, s_);
    tw_keep_.push_back(buf);
    arena_tw_[L] = buf->p;
    return buf->p;
  }
  int dev = 0;
  DMX_HIP(hipGetDevice(&dev));
  const std::string key = std::to_string(dev) + ":" + std::to_string(L) +
                          (p_.fp64 ? "d" : "s");
  auto &C = setup_cache();
  std::lock_guard<std::mutex> lk(C.mu);
  auto it = C.tw.find(key);
  if (it != C.tw.end()) {
    tw_keep_.push_back(it->second);
    return it->second->p;
  }
  std::vector<unsigned char> h;
  fft_twiddles(L, p_.fp64, h);
  auto buf = std::make_shared<Buf>(h.size());
  DMX_HIP(hipMemcpy(buf->p, h.data(), h.size(), hipMemcpyHostToDevice));
  if (C.tw.size() > 4 * SETUP_CACHE_MAX) C.tw.clear();
  C.tw[key] = buf;
  tw_keep_.push_back(buf);
  return buf->p;
}

void Plan::build_index(const void *coords, bool coord_f32, size_t M,
                       bool keep_sorted_coords) {
  tag_ = "index";
  M_ = M;
  coord_f32_ = coord_f32;
  g_.coord_f32 = coord_f32 ? 1 : 0;
  coords_ = coords;
  user_coords_ = coords;
  coords_sorted_ = false;
  g_.coord_v2 = (p_.ndim == 2 && !coord_f32 &&
                 reinterpret_cast<uintptr_t>(coords) % 16 == 0)
                    ? 1
                    : 0;
  if (M == 0) return;
  int keybits = 1;
  while (((size_t)1 << keybits) < ntiles_) ++keybits;
  // temporaries live as long as the plan, so nothing here waits for the
  // stream (the pool may hand freed blocks to other streams)
  size_t ws = std::max(radix_sort_workspace(M), scan_workspace(ntiles_));
  tmp_keys_.stream = s_;
  tmp_keys_.reset(M * 4);
  tmp_tiles_.stream = s_;
  tmp_tiles_.reset(ntiles_ * 4 * 5);
  tmp_work_.stream = s_;
  tmp_work_.reset(ws);
  uint32_t *counts = tmp_tiles_.as<uint32_t>(), *tstart = counts + ntiles_,
           *tend = tstart + ntiles_;
  {
    Phase ph(this, "building index");
    launch_zero(tstart, ntiles_ * 8, s_);  // tstart and tend, adjacent
    launch_fold_keys(g_, coords, M, tmp_keys_.as<uint32_t>(),
                     radix_sort_hist0(tmp_work_.p, M), s_);
    idx_.stream = s_;
    idx_.reset(M * 4);
    radix_sort_pairs(tmp_keys_.as<uint32_t>(), M, keybits, idx_.as<uint32_t>(),
                     nullptr, tmp_work_.p, ws, s_, true, tstart, tend);
    finish_index(counts);
  }
  if (keep_sorted_coords) {
    // a random gather of 1e7 points costs about as much as the gathers it
    // saves in one spreading plus one interpolation, so only plans keep it
    Phase ps(this, "sorting coords");
    sorted_coords_.stream = s_;
    sorted_coords_.reset(M * p_.ndim * (coord_f32 ? 4 : 8));
    launch_gather_coords(coords, coord_f32, p_.ndim, idx_.as<uint32_t>(), M,
                         sorted_coords_.p, s_);
    coords_ = sorted_coords_.p;
    coords_sorted_ = true;
    g_.coord_v2 = (p_.ndim == 2 && !coord_f32) ? 1 : 0;  // pool: aligned
  }
}

void Plan::finish_index(uint32_t *counts) {
  uint32_t *tstart = counts + ntiles_, *tend = tstart + ntiles_,
           *nsub = tend + ntiles_, *sstart = nsub + ntiles_;
  nsub_max_ = std::min(ntiles_, M_) + (M_ + MAXSUB - 1) / MAXSUB;
  subs_.stream = s_;
  subs_.reset(nsub_max_ * sizeof(Sub));
  nsub_total_.stream = s_;
  nsub_total_.reset(4);
  launch_tile_subs(ntiles_, counts, tstart, tend, nsub, sstart,
                   subs_.as<Sub>(), nsub_total_.as<uint32_t>(), tmp_work_.p,
                   s_);
  tile_nsub_ = nsub;
  tile_sub0_ = sstart;
}

bool Plan::build_records(const void *coords, bool coord_f32, size_t M,
                         bool type1, const void *values, bool values_f64) {
  if (!p_.use_lds || p_.deterministic || M == 0) return false;
  int keybits = 1;
  while (((size_t)1 << keybits) < ntiles_) ++keybits;
  const int offbits = p_.ndim * p_.log2tile;
  if (keybits + offbits > 32 || p_.log2tile > 10) return false;
  tag_ = "index";
  M_ = M;
  coord_f32_ = coord_f32;
  g_.coord_f32 = coord_f32 ? 1 : 0;
  coords_ = nullptr;  // the records replace the coordinates
  user_coords_ = coords;
  coords_sorted_ = false;
  g_.coord_v2 = (p_.ndim == 2 && !coord_f32 &&
                 reinterpret_cast<uintptr_t>(coords) % 16 == 0)
                    ? 1
                    : 0;
  // the record arrays and temporaries live as long as the plan, so nothing
  // here waits for the stream
  const size_t rb = rec_carve_bytes(M, p_.ndim, p_.fp64, type1);
  rec_buf_.stream = s_;
  rec_buf_.reset(rb);
  rec_tmp_.stream = s_;
  rec_tmp_.reset(rb);
  recs_ = rec_carve(rec_buf_.p, M, p_.ndim, p_.fp64, type1);
  const RecBufs tmp = rec_carve(rec_tmp_.p, M, p_.ndim, p_.fp64, type1);
  const size_t ws = std::max(rec_sort_workspace(M), scan_workspace(ntiles_));
  tmp_tiles_.stream = s_;
  tmp_tiles_.reset(ntiles_ * 4 * 5);
  tmp_work_.stream = s_;
  tmp_work_.reset(ws);
  uint32_t *counts = tmp_tiles_.as<uint32_t>(), *tstart = counts + ntiles_,
           *tend = tstart + ntiles_;
  {
    Phase ph(this, "building index");
    launch_zero(tstart, ntiles_ * 8, s_);  // tstart and tend, adjacent
    rec_sort(g_, offbits, keybits, p_.fp64, type1, coords, values,
             values_f64 ? 1 : 0, M, recs_, tmp, tmp_work_.p, ws, tstart, tend,
             s_);
    finish_index(counts);
  }
  rec_ = true;
  rec_type1_ = type1;
  return true;
}

uint32_t *Plan::keys_copy_to(uint32_t *dev_keys_out) {
  Geo g = g_;
  g.coord_v2 = (p_.ndim == 2 && !coord_f32_ &&
                reinterpret_cast<uintptr_t>(user_coords_) % 16 == 0)
                   ? 1
                   : 0;
  launch_fold_keys(g, user_coords_, M_, dev_keys_out, nullptr, s_);
  return dev_keys_out;
}

void Plan::fill_spread_args(SpreadArgs &a) const {
  a.g = g_;
  a.subs = subs_.as<Sub>();
  a.nsub_total = nsub_total_.as<uint32_t>();
  a.idx = idx_.as<uint32_t>();
  a.coords = coords_;
  a.coef = su_->coef.p;
  a.nsub_max = nsub_max_;
  a.lds_side = lds_side_;
  a.use_lds = p_.use_lds ? 1 : 0;
  a.coords_sorted = coords_sorted_ ? 1 : 0;
  a.tile_nsub = tile_nsub_;
  a.tile_sub0 = tile_sub0_;
  a.rec = rec_ ? 1 : 0;
  a.recs = recs_;
  a.offcheck = p_.check_offsets ? offcheck_.as<uint32_t>() : nullptr;
  a.fixacc = p_.spread_fix ? 1 : 0;
}

void Plan::check_offsets() {
  if (!p_.check_offsets) return;
  uint32_t n = 0;
  DMX_HIP(hipMemcpyAsync(&n, offcheck_.p, 4, hipMemcpyDeviceToHost, s_));
  DMX_HIP(hipStreamSynchronize(s_));
  if (n) {
    launch_zero(offcheck_.p, 4, s_);
    throw Error("internal: " + std::to_string(n) +
                " point offsets fell outside their tile");
  }
}

void Plan::prepare_grid_side() {
  size_t n = 1;
  for (int d = 0; d < p_.ndim; ++d) n *= p_.nover[d];
  size_t bytes = n * (p_.fp64 ? 16 : 8);
  grid_.stream = s_;
  if (grid_.bytes != bytes) grid_.reset(bytes);
  // the second buffer serves lines longer than LDS and the out-of-place
  // half of the two-pass column FFT (2-D)
  scratch_.stream = s_;
  if (needs_scratch_grid(p_) && scratch_.bytes != bytes) scratch_.reset(bytes);
  cur_ = grid_.p;
}

GridCopyArgs Plan::copy_args(const void *uni, bool uni_f64,
                             const int64_t *ust) const {
  GridCopyArgs c{};
  c.uni = const_cast<void *>(uni);
  c.u_f64 = uni_f64 ? 1 : 0;
  c.fft_order = p_.fft_order ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    c.n[d] = d < p_.ndim ? (int)p_.nuni[d] : 1;
    c.N[d] = d < p_.ndim ? (int)p_.nover[d] : 1;
    c.ust[d] = d < p_.ndim ? ust[d] : 0;
    if (d >= p_.ndim) {
      c.cf[d] = su_->one.as<double>();
      continue;
    }
    int e = d;
    while (!su_->corfac[e].p) --e;  // reused from the previous axis
    c.cf[d] = su_->corfac[e].as<double>();
  }
  return c;
}

Plan::FftRoute Plan::fft_route(int d) const {
  const int L = (int)p_.nover[d];
  if (L <= 1) return FftRoute::none;
  const int maxel = p_.fp64 ? FFT_MAXEL64 : FFT_MAXEL32;
  // 2-D columns too long for >= 16 per workgroup: two coalesced passes
  if (p_.ndim == 2 && d == 0 && maxel / L < 16 && fft_split(L) > 0 &&
      scratch_.p)
    return FftRoute::columns;
  return L <= maxel ? FftRoute::lds : FftRoute::global;
}

void Plan::pad_grid(const void *uni, bool uni_f64, const int64_t *ust) {
  GridCopyArgs c = copy_args(uni, uni_f64, ust);
  c.grid = grid_.p;
  // An FFT in LDS (rows, or the two-pass columns) reads the retained modes
  // of its axis only (the rest is known to be zero), so along such an axis
  // the retained rows suffice.  An FFT of whole lines (global passes) reads
  // every position: up to the last axis done that way, every grid position
  // is written, zero outside the modes.
  int last_whole = -1;
  for (int d = 0; d < p_.ndim; ++d)
    if (fft_route(d) == FftRoute::global) last_whole = d;
  for (int d = 0; d <= last_whole; ++d) c.full[d] = 1;
  Phase ph(this, "grid correction");
  launch_grid_copy(p_.fp64, true, c, s_);
}

bool Plan::run_fft(bool forward, bool type1, const GridCopyArgs *crop) {
  bool cropped = false;
  const int nd = p_.ndim, f = 3 - nd;
  int64_t E[3], st[3];
  int n_act[3], N_act[3];
  for (int a = 0; a < 3; ++a) {
    if (a < f) {
      E[a] = 1; n_act[a] = 1; N_act[a] = 1;
    } else {
      E[a] = (int64_t)p_.nover[a - f];
      n_act[a] = (int)p_.nuni[a - f];
      N_act[a] = (int)p_.nover[a - f];
    }
  }
  st[2] = 1;
  st[1] = E[2];
  st[0] = E[1] * E[2];
  auto range = [&](int ax, bool restrict_, int &cnt, int &n1, int &a1,
                   int &a2) {
    if (!restrict_ || ax < f) {
      cnt = (int)E[ax]; n1 = (int)E[ax]; a1 = 0; a2 = 0;
    } else {
      int n = n_act[ax], N = N_act[ax];
      n1 = (n + 1) / 2; a1 = 0; a2 = N - n / 2; cnt = n1 + n / 2;
    }
  };
  std::vector<int> order;
  if (type1) for (int a = 2; a >= f; --a) order.push_back(a);
  else for (int a = f; a <= 2; ++a) order.push_back(a);
  const int maxel = p_.fp64 ? FFT_MAXEL64 : FFT_MAXEL32;
  for (int a : order) {
    int pax = a == 0 ? 1 : 0, qax = a == 2 ? 1 : 2;
    FftLaunch L{};
    L.data = cur_;
    L.L = (int)E[a];
    L.st_ax = st[a];
    L.st_u = st[pax];
    L.st_v = st[qax];
    L.sign = forward ? -1 : 1;
    range(pax, pax > a, L.nu, L.un1, L.ua1, L.ua2);
    range(qax, qax > a, L.nv, L.vn1, L.va1, L.va2);
    if (L.L <= 1) continue;
    if (type1) {
      // type 1: later passes and the crop read only the retained modes of
      // this axis
      L.sprune = 1;
      L.sn1 = (n_act[a] + 1) / 2;
      L.sa2 = N_act[a] - n_act[a] / 2;
    }
    Phase ph(this, std::string("FFT axis ") + std::to_string(a - f));
    const FftRoute route = fft_route(a - f);
    if (!type1 && route != FftRoute::global) {
      // type 2: along this axis only the retained modes were written (by
      // the pad or by the FFTs of the earlier axes); the rest is zero
      L.kprune = 1;
      L.kn1 = (n_act[a] + 1) / 2;
      L.ka2 = N_act[a] - n_act[a] / 2;
    }
    if (route == FftRoute::columns) {
      // two coalesced passes (64 columns per workgroup), the second writing
      // natural order into the other buffer
      void *other = cur_ == grid_.p ? scratch_.p : grid_.p;
      // rows longer than LDS cannot skip the unwritten columns: clear them
      if (!type1 && E[2] > maxel) launch_zero(other, grid_.bytes, s_);
      const int L1 = fft_split(L.L);
      // type 1, 2-D: the column pass is the last one; it writes the
      // retained modes to the uniform array with their correction factors
      FftCrop fc{};
      const bool fuse = type1 && crop && nd == 2 && a == f;
      if (fuse) {
        fc.uni = crop->uni;
        fc.cf0 = crop->cf[0];
        fc.cf1 = crop->cf[1];
        fc.ust0 = crop->ust[0];
        fc.ust1 = crop->ust[1];
        fc.n0 = crop->n[0];
        fc.N0 = crop->N[0];
        fc.n1 = crop->n[1];
        fc.N1 = crop->N[1];
        fc.fft_order = crop->fft_order;
        fc.u_f64 = crop->u_f64;
      }
      fft_columns_2pass(L, L1, other, p_.fp64, twiddle_for(L1),
                        twiddle_for(L.L / L1), twiddle_for(L.L), s_, nullptr,
                        fuse ? &fc : nullptr);
      cur_ = other;
      cropped = cropped || fuse;
      continue;
    }
    L.tw = twiddle_for(L.L);
    fft_lineset(L, p_.fp64, cur_ == grid_.p ? scratch_.p : grid_.p, s_);
  }
  return cropped;
}

void Plan::nu2u_grid() {
  tag_ = "nu2u";
  prepare_grid_side();
  // with stored regions the fold writes every grid cell
  grid_zeroed_ = !p_.slab_flush;
  if (grid_zeroed_) {
    Phase ph(this, "zeroing grid");
    launch_zero(grid_.p, grid_.bytes, s_);
  }
  grid_ready_ = true;
}

void Plan::nu2u(bool forward, const void *points, bool pts_f64, void *uni,
                bool uni_f64, const int64_t *ust) {
  if (!grid_ready_) nu2u_grid();
  grid_ready_ = false;
  tag_ = "nu2u";
  DMX_ASSERT(!rec_ || rec_type1_, "internal: type-2 records for type 1");
  if (M_ > 0) {
    Phase ph(this, "spreading");
    SpreadArgs a{};
    fill_spread_args(a);
    a.values = points;
    a.val_f64 = pts_f64 ? 1 : 0;
    a.grid = grid_.p;
    const size_t rb =
        p_.slab_flush ? spread_region_slab_bytes(a, p_.fp64) : 0;
    if (p_.deterministic) {
      const size_t sb = spread_det_slab_bytes(a);
      slab_.stream = s_;
      if (slab_.bytes != sb) slab_.reset(sb);
      a.slab = slab_.p;
      launch_spread_det(p_.fp64, a, s_);  // writes every grid cell
    } else if (p_.slab_flush && rb <= SLAB_MAX_BYTES) {
      slab_.stream = s_;
      if (slab_.bytes != rb) slab_.reset(rb);
      a.slab = slab_.p;
      a.slab_out = 1;
      launch_spread(p_.fp64, a, s_);
      launch_fold_regions(p_.fp64, a, s_);  // writes every grid cell
    } else {
      if (!grid_zeroed_) launch_zero(grid_.p, grid_.bytes, s_);
      launch_spread(p_.fp64, a, s_);
    }
    check_offsets();
  } else if (!grid_zeroed_) {
    launch_zero(grid_.p, grid_.bytes, s_);
  }
  GridCopyArgs c = copy_args(uni, uni_f64, ust);
  if (run_fft(forward, true, &c)) return;  // the last FFT pass cropped
  c.grid = cur_;
  Phase ph(this, "grid correction");
  launch_grid_copy(p_.fp64, false, c, s_);
}

void Plan::u2nu_grid(bool forward, const void *uni, bool uni_f64,
                     const int64_t *ust, bool grad) {
  tag_ = grad ? "u2nu_grad" : "u2nu";
  prepare_grid_side();
  pad_grid(uni, uni_f64, ust);
  run_fft(forward, false);
  grid_ready_ = true;
}

void Plan::u2nu(bool forward, const void *uni, bool uni_f64,
                const int64_t *ust, void *out, bool out_f64) {
  if (!grid_ready_) u2nu_grid(forward, uni, uni_f64, ust);
  grid_ready_ = false;
  tag_ = "u2nu";
  DMX_ASSERT(!rec_ || !rec_type1_, "internal: type-1 records for type 2");
  if (M_ > 0) {
    SpreadArgs a{};
    fill_spread_args(a);
    a.grid = cur_;
    Phase ph(this, "interpolation");
    launch_interp(p_.fp64, a, out, out_f64 ? 1 : 0, s_);
    check_offsets();
  }
}

void Plan::u2nu_grad(bool forward, const void *uni, bool uni_f64,
                     const int64_t *ust, void *out, bool out_f64) {
  if (!grid_ready_) u2nu_grid(forward, uni, uni_f64, ust, true);
  grid_ready_ = false;
  tag_ = "u2nu_grad";
  DMX_ASSERT(!rec_, "internal: coordinate gradient needs the point index");
  if (M_ > 0) {
    SpreadArgs a{};
    fill_spread_args(a);
    a.grid = cur_;
    Phase ph(this, "interpolation (derivative)");
    launch_interp_grad(p_.fp64, a, out, out_f64 ? 1 : 0, s_);
    check_offsets();
  }
}

void Plan::info(dmx_plan_info *o) const {
  std::memset(o, 0, sizeof(*o));
  o->ndim = p_.ndim;
  o->W = p_.W;
  o->kernel_idx = p_.kidx;
  o->fp64 = p_.fp64;
  o->log2tile = p_.log2tile;
  o->chooser_used = p_.chooser_used;
  for (int d = 0; d < 3; ++d) {
    o->nuni[d] = d < p_.ndim ? (int64_t)p_.nuni[d] : 1;
    o->nover[d] = d < p_.ndim ? (int64_t)p_.nover[d] : 1;
  }
  o->ofactor = p_.ofactor;
  o->eps_tab = p_.eps_tab;
  o->epsilon = p_.eps;
  o->npoints = (int64_t)M_;
}

}  // namespace dmx
