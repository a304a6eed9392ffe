// runtime.h -- device memory pool, per-call staging of host/device arrays,
// and the NUFFT plan (sorted index, kernel, correction factors, grid).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "ducc_amd.h"
#include "dmx.h"
#include "fft.h"
#include "grid.h"

namespace dmx {

// Bump allocation from caller-provided device memory (the static scratch of
// the allocation-free entry point): while an ArenaScope is alive on a
// thread, every pool allocation of that thread is carved from
// [base, base + size) in 256-byte units, and releases are no-ops.
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  void *take(size_t bytes);
};
class ArenaScope {
 public:
  explicit ArenaScope(Arena *a);
  ~ArenaScope();
  ArenaScope(const ArenaScope &) = delete;
  ArenaScope &operator=(const ArenaScope &) = delete;

 private:
  Arena *prev_;
};
Arena *current_arena();

// Caching allocator per device: freed blocks are kept and reused for
// requests of at most their size (and at least half of it).  A block
// released with a stream is stream-ordered: it is handed out again only
// after an event recorded on that stream at release time has completed, so
// work still queued there can never see the block reused by another stream
// or thread.
class Pool {
 public:
  static Pool &get();
  void *alloc(size_t bytes);
  // s == nullptr: no queued work can touch p any more
  void release(void *p, hipStream_t s = nullptr);
  void trim();

 private:
  struct Pending {
    hipEvent_t ev;
    void *p;
  };
  void collect_locked();  // completed pending blocks -> free list
  std::mutex mu_;
  std::map<void *, std::pair<int, size_t>> live_;
  std::multimap<std::pair<int, size_t>, void *> free_;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> spare_events_;
};

struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
  // the stream whose queued work may still use the block when it is freed
  // (nullptr: the owner has waited for all such work)
  hipStream_t stream = nullptr;
  Buf() = default;
  explicit Buf(size_t n, hipStream_t s = nullptr) : stream(s) { reset(n); }
  ~Buf() { free(); }
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  Buf(Buf &&o) noexcept : p(o.p), bytes(o.bytes), stream(o.stream) {
    o.p = nullptr;
    o.bytes = 0;
  }
  Buf &operator=(Buf &&o) noexcept {
    if (this != &o) {
      free();
      p = o.p;
      bytes = o.bytes;
      stream = o.stream;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  void reset(size_t n) {
    free();
    if (n) p = Pool::get().alloc(n);
    bytes = n;
  }
  void free() {
    if (p) Pool::get().release(p, stream);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return (T *)p; }
};

// Phase timing with HIP events recorded on the plan's stream, under the
// phase names of the reference's TimerHierarchy (nufft.h:840-861,
// 1326-1358).  Off unless enabled; totals accumulate per process.
class Profiler {
 public:
  static Profiler &get();
  bool enabled() const { return on_; }
  void enable(bool on) { on_ = on; }
  void add(const std::string &name, double ms);
  void reset();
  std::string report();

 private:
  bool on_ = false;
  std::mutex mu_;
  std::vector<std::string> order_;
  std::map<std::string, std::pair<double, long>> acc_;
};

struct Params {
  int ndim = 0;
  size_t nuni[3] = {1, 1, 1};
  size_t nover[3] = {1, 1, 1};
  double period[3] = {1, 1, 1};
  bool fft_order = false;
  double eps = 0;
  int kidx = -1, W = 0;
  double beta = 0, e0 = 0, ofactor = 0, eps_tab = 0;
  bool fp64 = true;
  int log2tile = 4;
  bool use_lds = true;
  int chooser_used = 0;
  int chooser_verified = 0;
  bool deterministic = false;  // type 1 without floating-point atomics
  bool check_offsets = false;  // DMX_CHECK_OFFSETS
  // FP32 spreading in packed fixed point (eps >= FIX_MIN_EPS, no
  // DMX_SPREAD_FP64_ACC)
  bool spread_fix = false;
  // type 1 stores each subproblem's region and folds the overlaps into the
  // grid (no grid zeroing, no global atomics); 3-D
  bool slab_flush = false;
};
// the largest region store the default spreader takes; larger requests
// flush with atomics
constexpr size_t SLAB_MAX_BYTES = (size_t)16 << 30;
// the smallest eps for which FP32 spreading accumulates in fixed point
constexpr double FIX_MIN_EPS = 1e-5;

// device constants shared by all plans of one parameter set
struct Setup {
  Buf coef, corfac[3], one;
};

// choose precision, kernel, oversampled grid and tile for a request
Params choose_params(int ndim, const size_t *nuni, size_t M, bool gridding,
                     double eps, double smin, double smax,
                     const double *period, bool fft_order, bool values_f64,
                     const dmx_options &opt);

// the parameters for kernel-table row kidx (of any dimensionality) at
// oversampling factor ofactor, with the tile logic of choose_params; npts:
// the points one spreading / interpolation call processes (0: unknown)
Params kernel_params(int ndim, const size_t *nuni, const double *period,
                     bool fft_order, bool fp64, double eps, int kidx,
                     double ofactor, const dmx_options &opt, double npts);
// choose_params' common tail: kernel row, limits, tile (l2t <= 0: the
// default for the dimensionality, precision and npts points per call)
void finish_params(Params &p, int kidx, int l2t, const dmx_options &opt,
                   double npts);

class Plan {
 public:
  Plan(const Params &p, hipStream_t s);
  ~Plan();
  // sorts the points; coords: device (M, ndim), f32 or f64, contiguous,
  // alive while the plan executes unless keep_sorted_coords.
  // keep_sorted_coords: store a copy in sorted order (for plans executed
  // several times: one random gather instead of one per execution).
  void build_index(const void *coords, bool coord_f32, size_t M,
                   bool keep_sorted_coords);
  // For a single transform: sorts the points' records (rec_sort) instead of
  // an index, so that spreading / interpolation read them in order.
  // values: type 1, the device values (M complex, c64 / c128), alive until
  // the index build has run; nullptr for type 2.  Returns false, with
  // nothing queued, where the record layout does not apply (no LDS region,
  // deterministic mode, tile key and offsets wider than 32 bits): then
  // build_index is the route.
  bool build_records(const void *coords, bool coord_f32, size_t M,
                     bool type1, const void *values, bool values_f64);
  bool uses_records() const { return rec_; }
  const RecBufs &records() const { return recs_; }
  // The grid-side part of a transform that does not depend on the points
  // (type 1: the zeroed grid; type 2: pad and FFT), enqueued ahead of the
  // rest, e.g. on another stream while build_index runs.  The next nu2u /
  // u2nu / u2nu_grad then does only the remainder (same arguments).
  void nu2u_grid();
  void u2nu_grid(bool forward, const void *uni, bool uni_f64,
                 const int64_t *ust, bool grad = false);
  // the stream the plan enqueues on from now on (no ordering with the
  // previous stream: for forks whose join the caller arranges)
  void set_stream(hipStream_t s) { s_ = s; }
  // the stream the plan enqueues on from now on; s first waits for the work
  // already queued on the previous stream (the plan's buffers are reused)
  void use_stream(hipStream_t s);
  // record phase timings of the following work for report_timings (the
  // verbosity output), whether or not the process-wide profiler is on
  void start_timing();
  // ducc0's verbosity report (nufft.h:371-380): request and parameters
  void report_params(bool type1) const;
  // waits for the stream and prints the phase table of the work since
  // start_timing (role of TimerHierarchy::report, timers.h:117-160)
  void report_timings(const char *what);
  // type 1: points (device, M complex, c64/c128) -> uniform (device, strided)
  void nu2u(bool forward, const void *points, bool pts_f64, void *uni,
            bool uni_f64, const int64_t *ust);
  // type 2: uniform (device, strided) -> points (device, M complex)
  void u2nu(bool forward, const void *uni, bool uni_f64, const int64_t *ust,
            void *out, bool out_f64);
  // d u2nu / d coord -> out (M, ndim) complex
  void u2nu_grad(bool forward, const void *uni, bool uni_f64,
                 const int64_t *ust, void *out, bool out_f64);
  // the sort results of build_index (device)
  const uint32_t *idx() const { return idx_.as<uint32_t>(); }
  uint32_t *keys_copy_to(uint32_t *dev_keys_out);  // re-fold keys
  void info(dmx_plan_info *o) const;
  // device bytes a transform of M points with parameters p allocates (the
  // plan's constants, index build, grid buffers and FFT twiddles), each
  // block rounded to 256 bytes as the pool and the arena hand them out
  static size_t scratch_bytes(const Params &p, size_t M);
  // the caller has seen all work queued so far complete (the destructor then
  // does not wait for the stream)
  void mark_done() { done_ = true; }
  const Params &params() const { return p_; }
  hipStream_t stream() const { return s_; }
  size_t npoints() const { return M_; }
  Geo geo() const { return g_; }
  // resolves recorded phase events (after the stream has been synchronised)
  // moves finished phase timings to the profiler; wait = false stops at
  // the first phase still running instead of waiting for it
  void flush_profile(bool wait = true);

  struct Phase {
    Plan *p;
    std::string name;
    hipEvent_t a = nullptr, b = nullptr;
    Phase(Plan *pl, std::string n);
    ~Phase();
  };

 private:
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pend_;
  bool timing_ = false;                  // start_timing() was called
  hipEvent_t t0_ = nullptr;              // start of the timed call
  hipEvent_t last_ = nullptr;            // use_stream's ordering event
  std::vector<std::pair<std::string, double>> timed_;  // phases, this call
  bool done_ = false;
  bool grid_ready_ = false;  // a *_grid call enqueued the grid side
  bool grid_zeroed_ = true;  // nu2u_grid zeroed the grid
  std::string tag_ = "plan";  // "index", "nu2u" or "u2nu": phase prefix
  void prepare_grid_side();
  // how the FFT along real axis d runs: in LDS, as the two-pass 2-D column
  // FFT, or as global passes (lines longer than LDS)
  enum class FftRoute { none, lds, columns, global };
  FftRoute fft_route(int d) const;
  GridCopyArgs copy_args(const void *uni, bool uni_f64,
                         const int64_t *ust) const;
  // zero-padded, corrected uniform array -> grid, type 2
  void pad_grid(const void *uni, bool uni_f64, const int64_t *ust);
  // FFT passes of the oversampled grid; with `crop` (type 1), the last pass
  // may also write the corrected retained modes to the uniform array:
  // returns true when it did
  bool run_fft(bool forward, bool type1, const GridCopyArgs *crop = nullptr);
  void fill_spread_args(SpreadArgs &a) const;
  // DMX_CHECK_OFFSETS: waits for the kernels and fails if any point's
  // offset was clamped
  void check_offsets();
  Buf offcheck_;  // the clamp counter (4 bytes, zeroed at creation)

  Params p_;
  hipStream_t s_;
  Geo g_{};
  size_t M_ = 0;
  size_t ntiles_ = 0;
  size_t nsub_max_ = 0;
  int lds_side_ = 0;
  const void *coords_ = nullptr;  // user coords or the sorted copy
  const void *user_coords_ = nullptr;  // caller's coordinates (input order)
  bool coords_sorted_ = false;
  bool coord_f32_ = false;
  std::shared_ptr<Setup> su_;
  std::vector<std::shared_ptr<Buf>> tw_keep_;
  static std::shared_ptr<Setup> setup_for(const Params &p);
  // the constants built in the current arena by kernels on s (no cache, no
  // host-to-device copy)
  static std::shared_ptr<Setup> arena_setup(const Params &p, hipStream_t s);
  std::map<int, const void *> arena_tw_;  // twiddles built in the arena
  Buf idx_, subs_, nsub_total_, sorted_coords_;
  bool rec_ = false, rec_type1_ = false;
  RecBufs recs_{};
  Buf rec_buf_, rec_tmp_;
  // the tile tables and subproblem list from the sort's tile ranges
  void finish_index(uint32_t *tables);
  Buf tmp_keys_, tmp_tiles_, tmp_work_;
  Buf grid_, scratch_, slab_;
  const uint32_t *tile_nsub_ = nullptr, *tile_sub0_ = nullptr;
  void *cur_ = nullptr;  // buffer holding the grid after the FFT passes
  const void *twiddle_for(int L);
};

// error channel of the C ABI
void set_error(const char *msg);

}  // namespace dmx
