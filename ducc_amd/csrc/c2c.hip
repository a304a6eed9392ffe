// c2c.hip -- complex FFT of an N-d array along chosen axes on the Stockham
// kernels of fft.hip: the drop-in for ducc0::c2c (src/ducc0/fft/fft.h:
// 848-850; Python binding python/fft_pymod.cc:122-172).
//
// The input is copied (scaled by fct, real input widened to complex) into a
// contiguous work array, the output itself when that is contiguous.  Each
// axis is then transformed in place, the array seen as (outer, L, inner):
//  * L 2,3,5,7,11-smooth: Stockham passes, in LDS when the line fits,
//    otherwise out-of-place global passes;
//  * any other L: Bluestein's chirp-z algorithm (X_m = w_m sum_k x_k w_k
//    conj(w_{m-k}), w_k = exp(-/+ i pi k^2 / L)) as a circular convolution
//    of smooth length n2 >= 2L-1, done with two FFTs of length n2 on an
//    (outer, n2, inner) buffer and a product with the precomputed FFT of the
//    chirp.
// The chirp and its transform are computed once per (device, L, precision)
// in FP64 and cached (a bounded cache).
#include "c2c.h"

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "fft.h"
#include "runtime.h"

namespace dmx {

namespace {

template <typename T>
using Cx = typename std::conditional<sizeof(T) == 4, float2, double2>::type;

template <typename T>
__device__ __forceinline__ Cx<T> cx_mul(Cx<T> a, Cx<T> b) {
  Cx<T> r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// element loads of the input types as complex<T>
template <typename T>
__device__ __forceinline__ Cx<T> ld(const float *p, int64_t i) {
  Cx<T> r;
  r.x = (T)p[i];
  r.y = T(0);
  return r;
}
template <typename T>
__device__ __forceinline__ Cx<T> ld(const double *p, int64_t i) {
  Cx<T> r;
  r.x = (T)p[i];
  r.y = T(0);
  return r;
}
template <typename T>
__device__ __forceinline__ Cx<T> ld(const float2 *p, int64_t i) {
  const float2 v = p[i];
  Cx<T> r;
  r.x = (T)v.x;
  r.y = (T)v.y;
  return r;
}
template <typename T>
__device__ __forceinline__ Cx<T> ld(const double2 *p, int64_t i) {
  const double2 v = p[i];
  Cx<T> r;
  r.x = (T)v.x;
  r.y = (T)v.y;
  return r;
}

struct NdCopy {
  int ndim, linear;
  int64_t n;
  int64_t shape[10];
  int64_t sst[10], dst[10];  // strides in elements
  double fct;
};

// dst = fct * src over an N-d index space (both strided, or both linear)
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_nd_copy(const S *src, Cx<T> *dst,
                                                 NdCopy a) {
  const T f = (T)a.fct;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < a.n;
       q += (int64_t)gridDim.x * 256) {
    int64_t so = q, doff = q;
    if (!a.linear) {
      int64_t r = q;
      so = 0;
      doff = 0;
      for (int d = a.ndim - 1; d >= 0; --d) {
        const int64_t i = r % a.shape[d];
        r /= a.shape[d];
        so += i * a.sst[d];
        doff += i * a.dst[d];
      }
    }
    Cx<T> v = ld<T>(src, so);
    v.x *= f;
    v.y *= f;
    dst[doff] = v;
  }
}

// q -> (o, k, i) over (outer, K, inner); 32-bit division when it fits
__device__ __forceinline__ void split3(int64_t q, int64_t K, int64_t inner,
                                       bool small, int64_t &o, int64_t &k,
                                       int64_t &i) {
  if (small) {
    const uint32_t uq = (uint32_t)q, ui = (uint32_t)inner, uk = (uint32_t)K;
    const uint32_t t = uq / ui;
    i = uq - t * ui;
    const uint32_t oo = t / uk;
    k = t - oo * uk;
    o = oo;
  } else {
    const int64_t t = q / inner;
    i = q - t * inner;
    o = t / K;
    k = t - o * K;
  }
}

struct BlueArgs {
  int64_t outer, inner, n;  // n: elements of the index space of the launch
  int L, n2, conj_w, small;
};

// MODE 0: a[o,k,i] = k < L ? x[o,k,i] w'[k] : 0        over (outer, n2, inner)
// MODE 1: a[o,k,i] *= B[k]                             over (outer, n2, inner)
// MODE 2: x[o,m,i] = w'[m] a[o,m,i]                    over (outer, L, inner)
// w' = w or conj(w) (the transform's sign)
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_bluestein(Cx<T> *x, Cx<T> *a,
                                                   const Cx<T> *tab,
                                                   BlueArgs b) {
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < b.n;
       q += (int64_t)gridDim.x * 256) {
    int64_t o, k, i;
    split3(q, MODE == 2 ? b.L : b.n2, b.inner, b.small != 0, o, k, i);
    const int64_t ia = (o * b.n2 + k) * b.inner + i;
    if (MODE == 1) {
      a[ia] = cx_mul<T>(a[ia], tab[k]);
      continue;
    }
    Cx<T> w;
    if (k < b.L) {
      w = tab[k];
      if (b.conj_w) w.y = -w.y;
    }
    const int64_t ix = (o * b.L + k) * b.inner + i;
    if (MODE == 0) {
      Cx<T> v;
      if (k < b.L) {
        v = cx_mul<T>(x[ix], w);
      } else {
        v.x = T(0);
        v.y = T(0);
      }
      a[ia] = v;
    } else {
      x[ix] = cx_mul<T>(a[ia], w);
    }
  }
}

unsigned grid_for(int64_t n) {
  const int64_t nb = (n + 255) / 256;
  return (unsigned)std::min<int64_t>(std::max<int64_t>(nb, 1), 1 << 20);
}

bool is_smooth(int64_t n) {
  for (int r : {2, 3, 5, 7, 11})
    while (n % r == 0) n /= r;
  return n == 1;
}

// ------------------------------------------------------------ table cache
// Twiddles and Bluestein tables per (device, length, precision).  Bounded:
// past TABLES_MAX entries the cache starts over; calls in flight keep their
// tables through their own references (Job::keep).
constexpr size_t TABLES_MAX = 512;
struct Tables {
  std::mutex mu;
  std::map<std::string, std::shared_ptr<Buf>> m;
  void put(const std::string &k, std::shared_ptr<Buf> b) {
    if (m.size() >= TABLES_MAX) m.clear();
    m[k] = std::move(b);
  }
};
Tables &tables() {
  static Tables *t = new Tables();  // never destroyed (buffers outlive exit)
  return *t;
}
std::string key_of(const char *what, int64_t a, int64_t b, bool fp64) {
  int dev = 0;
  DMX_HIP(hipGetDevice(&dev));
  return std::string(what) + ":" + std::to_string(dev) + ":" +
         std::to_string(a) + ":" + std::to_string(b) + (fp64 ? "d" : "s");
}

std::shared_ptr<Buf> twiddles(int L, bool fp64) {
  auto &T = tables();
  const std::string k = key_of("tw", L, 0, fp64);
  std::lock_guard<std::mutex> lk(T.mu);
  auto it = T.m.find(k);
  if (it != T.m.end()) return it->second;
  std::vector<unsigned char> h;
  fft_twiddles(L, fp64, h);
  auto buf = std::make_shared<Buf>(h.size());
  DMX_HIP(hipMemcpy(buf->p, h.data(), h.size(), hipMemcpyHostToDevice));
  T.put(k, buf);
  return buf;
}

// w_k = exp(-i pi k^2 / L), k < L, from the exact residue k^2 mod 2L
void chirp_host(int L, std::vector<double> &w) {
  const double pi = 3.141592653589793238462643383279502884197;
  w.resize(2 * (size_t)L);
  for (int64_t k = 0; k < L; ++k) {
    const uint64_t r = ((uint64_t)k * (uint64_t)k) % (2 * (uint64_t)L);
    const double ang = pi * (double)r / (double)L;
    w[2 * k] = std::cos(ang);
    w[2 * k + 1] = -std::sin(ang);
  }
}

void run_lines(void *data, int64_t outer, int L, int64_t inner, bool fp64,
               int sign, const void *tw, void *scratch, hipStream_t s);

// Bluestein tables for length L (convolution length n2): the chirp w and
// B = FFT_n2(b~) / n2 with b~[j] = conj(w_j) (j < L), b~[n2-j] = conj(w_j)
// (0 < j < L), else 0, for the forward sign; the backward transform uses
// conj(w) and the transform of conj(b~) (cached under its own key).
struct Blue {
  std::shared_ptr<Buf> w, B;
};
Blue bluestein_tables(int L, int n2, bool fp64, int sign, hipStream_t s) {
  auto &T = tables();
  const std::string kw = key_of("bw", L, 0, fp64);
  const std::string kb = key_of(sign < 0 ? "bBf" : "bBb", L, n2, fp64);
  {
    std::lock_guard<std::mutex> lk(T.mu);
    auto iw = T.m.find(kw), ib = T.m.find(kb);
    if (iw != T.m.end() && ib != T.m.end()) return {iw->second, ib->second};
  }
  std::vector<double> w;
  chirp_host(L, w);
  const size_t esz = fp64 ? 16 : 8;
  auto wb = std::make_shared<Buf>((size_t)L * esz);
  if (fp64) {
    DMX_HIP(hipMemcpy(wb->p, w.data(), (size_t)L * 16, hipMemcpyHostToDevice));
  } else {
    std::vector<float> wf(w.begin(), w.end());
    DMX_HIP(hipMemcpy(wb->p, wf.data(), (size_t)L * 8, hipMemcpyHostToDevice));
  }
  // b~ in FP64, scaled by 1/n2, transformed with the FP64 kernels
  std::vector<double> bt(2 * (size_t)n2, 0.0);
  const double inv = 1.0 / (double)n2;
  const double cs = sign < 0 ? 1.0 : -1.0;  // conj(w) for the forward sign
  for (int64_t j = 0; j < L; ++j) {
    const double re = w[2 * j] * inv, im = -w[2 * j + 1] * inv * cs;
    bt[2 * j] = re;
    bt[2 * j + 1] = im;
    if (j > 0) {
      bt[2 * (n2 - j)] = re;
      bt[2 * (n2 - j) + 1] = im;
    }
  }
  Buf b64((size_t)n2 * 16), sc;
  DMX_HIP(hipMemcpy(b64.p, bt.data(), (size_t)n2 * 16, hipMemcpyHostToDevice));
  if (n2 > FFT_MAXEL64) sc.reset((size_t)n2 * 16);
  auto tw64 = twiddles(n2, true);
  run_lines(b64.p, 1, n2, 1, true, -1, tw64->p, sc.p, s);
  auto Bb = std::make_shared<Buf>((size_t)n2 * esz);
  if (fp64) {
    DMX_HIP(hipMemcpyAsync(Bb->p, b64.p, (size_t)n2 * 16,
                           hipMemcpyDeviceToDevice, s));
  } else {
    NdCopy c{};
    c.ndim = 1;
    c.linear = 1;
    c.n = n2;
    c.fct = 1.0;
    hipLaunchKernelGGL((k_nd_copy<double2, float>), dim3(grid_for(n2)),
                       dim3(256), 0, s, (const double2 *)b64.p,
                       (float2 *)Bb->p, c);
    DMX_HIP(hipGetLastError());
  }
  // tables are shared across streams: complete before publishing
  DMX_HIP(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> lk(T.mu);
  T.put(kw, wb);
  T.put(kb, Bb);
  return {wb, Bb};
}

// Stockham FFT of every line of the contiguous array (outer, L, inner) along
// its middle axis, in place (scratch: as large as the array when L is longer
// than a workgroup's LDS holds).
void run_lines(void *data, int64_t outer, int L, int64_t inner, bool fp64,
               int sign, const void *tw, void *scratch, hipStream_t s) {
  const size_t esz = fp64 ? 16 : 8;
  FftLaunch p{};
  p.L = L;
  p.tw = tw;
  p.sign = sign;
  if (inner == 1) {
    // contiguous lines: all of them along the second launch index
    DMX_ASSERT(outer <= 0x7fffffff, "too many FFT lines");
    p.data = data;
    p.st_ax = 1;
    p.nu = 1;
    p.un1 = 1;
    p.st_u = 0;
    p.nv = (int)outer;
    p.vn1 = p.nv;
    p.st_v = L;
    fft_lineset(p, fp64, scratch, s);
    return;
  }
  DMX_ASSERT(inner <= 0x7fffffff, "FFT line set too large");
  const int64_t chunk = 65535;  // lines along the first launch index
  for (int64_t u0 = 0; u0 < outer; u0 += chunk) {
    const int64_t off = u0 * (int64_t)L * inner;
    p.data = (char *)data + off * esz;
    p.st_ax = inner;
    p.nu = (int)std::min<int64_t>(chunk, outer - u0);
    p.un1 = p.nu;
    p.st_u = (int64_t)L * inner;
    p.nv = (int)inner;
    p.vn1 = p.nv;
    p.st_v = 1;
    fft_lineset(p, fp64, scratch ? (char *)scratch + off * esz : nullptr, s);
  }
}

// the two-pass column FFT of (outer, L, inner) along its middle axis:
// pass A data -> scratch, pass B scratch -> data (same layout)
void run_columns(void *data, int64_t outer, int L, int L1, int64_t inner,
                 bool fp64, int sign, const void *tw1, const void *tw2,
                 const void *tw, void *scratch, hipStream_t s) {
  const size_t esz = fp64 ? 16 : 8;
  DMX_ASSERT(inner <= 0x7fffffff, "FFT line set too large");
  const int64_t chunk = 65535;
  for (int64_t u0 = 0; u0 < outer; u0 += chunk) {
    const int64_t off = u0 * (int64_t)L * inner;
    FftLaunch p{};
    p.L = L;
    p.sign = sign;
    p.data = (char *)data + off * esz;
    p.st_ax = inner;
    p.nu = (int)std::min<int64_t>(chunk, outer - u0);
    p.un1 = p.nu;
    p.st_u = (int64_t)L * inner;
    p.nv = (int)inner;
    p.vn1 = p.nv;
    p.st_v = 1;
    const bool ok = fft_columns_2pass(p, L1, p.data, fp64, tw1, tw2, tw, s,
                                      (char *)scratch + off * esz);
    DMX_ASSERT(ok, "internal: column FFT split rejected");
  }
}

struct Job {
  std::vector<std::shared_ptr<Buf>> keep;
  Buf work, scratch, blue, blue_scratch;
};

// Phase timers (profiling only): HIP events on the call's stream, read back
// at the end of the call.
struct Timers {
  bool on;
  hipStream_t s;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev;
  explicit Timers(hipStream_t st) : on(Profiler::get().enabled()), s(st) {}
  void begin(const std::string &name) {
    if (!on) return;
    hipEvent_t a, b;
    DMX_HIP(hipEventCreate(&a));
    DMX_HIP(hipEventCreate(&b));
    DMX_HIP(hipEventRecord(a, s));
    ev.push_back({name, {a, b}});
  }
  void end() {
    if (on) DMX_HIP(hipEventRecord(ev.back().second.second, s));
  }
  ~Timers() {
    for (auto &e : ev) {
      float ms = 0;
      if (hipEventSynchronize(e.second.second) == hipSuccess &&
          hipEventElapsedTime(&ms, e.second.first, e.second.second) ==
              hipSuccess)
        Profiler::get().add("c2c/" + e.first, ms);
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
  }
};

template <typename T>
void copy_nd(const void *src, int src_tc, const int64_t *sst, void *dst,
             const int64_t *dstst, int ndim, const int64_t *shape, double fct,
             hipStream_t s) {
  NdCopy c{};
  c.ndim = ndim;
  c.n = 1;
  for (int d = 0; d < ndim; ++d) {
    c.shape[d] = shape[d];
    c.sst[d] = sst[d];
    c.dst[d] = dstst[d];
    c.n *= shape[d];
  }
  // linear when both arrays are contiguous (strides of size-1 axes do not
  // matter)
  bool lin = true;
  int64_t st = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    if (shape[d] != 1 && (sst[d] != st || dstst[d] != st)) lin = false;
    st *= shape[d];
  }
  c.linear = lin ? 1 : 0;
  c.fct = fct;
  if (!c.n) return;
  const dim3 g(grid_for(c.n)), b(256);
  Cx<T> *o = (Cx<T> *)dst;
  switch (src_tc) {
    case DMX_F32:
      hipLaunchKernelGGL((k_nd_copy<float, T>), g, b, 0, s, (const float *)src, o, c);
      break;
    case DMX_F64:
      hipLaunchKernelGGL((k_nd_copy<double, T>), g, b, 0, s, (const double *)src, o, c);
      break;
    case DMX_C64:
      hipLaunchKernelGGL((k_nd_copy<float2, T>), g, b, 0, s, (const float2 *)src, o, c);
      break;
    case DMX_C128:
      hipLaunchKernelGGL((k_nd_copy<double2, T>), g, b, 0, s, (const double2 *)src, o, c);
      break;
    default:
      throw Error("unsupported data type");
  }
  DMX_HIP(hipGetLastError());
}

template <typename T, int MODE>
void launch_blue(void *x, void *a, const void *tab, const BlueArgs &b,
                 hipStream_t s) {
  hipLaunchKernelGGL((k_bluestein<T, MODE>), dim3(grid_for(b.n)), dim3(256), 0,
                     s, (Cx<T> *)x, (Cx<T> *)a, (const Cx<T> *)tab, b);
  DMX_HIP(hipGetLastError());
}

}  // namespace

int64_t smooth_length(int64_t n) {
  if (n <= 1) return 1;
  while (!is_smooth(n)) ++n;
  return n;
}

std::shared_ptr<void> c2c_exec(const void *in, int in_tc, const int64_t *in_st,
                               void *out, const int64_t *out_st, int ndim,
                               const int64_t *shape, const int *axes,
                               int naxes, bool forward, double fct,
                               hipStream_t s) {
  auto job = std::make_shared<Job>();
  const bool fp64 = in_tc == DMX_C128 || in_tc == DMX_F64;
  const size_t esz = fp64 ? 16 : 8;
  int64_t n = 1;
  for (int d = 0; d < ndim; ++d) n *= shape[d];
  if (n == 0) return job;
  int64_t cst[10];
  {
    int64_t st = 1;
    for (int d = ndim - 1; d >= 0; --d) {
      cst[d] = st;
      st *= shape[d];
    }
  }
  bool out_contig = true;
  for (int d = 0; d < ndim; ++d)
    if (shape[d] != 1 && out_st[d] != cst[d]) out_contig = false;
  void *work = out;
  if (!out_contig) {
    job->work.reset((size_t)n * esz);
    work = job->work.p;
  }
  Timers tm(s);
  // in place on a contiguous array without scaling: nothing to copy
  bool in_contig = in == work && in_tc == (fp64 ? DMX_C128 : DMX_C64);
  for (int d = 0; d < ndim; ++d)
    if (shape[d] != 1 && in_st[d] != cst[d]) in_contig = false;
  if (!(in_contig && fct == 1.0)) {
    tm.begin("copy in");
    if (fp64)
      copy_nd<double>(in, in_tc, in_st, work, cst, ndim, shape, fct, s);
    else
      copy_nd<float>(in, in_tc, in_st, work, cst, ndim, shape, fct, s);
    tm.end();
  }
  const int sign = forward ? -1 : 1;
  const int maxel = fp64 ? FFT_MAXEL64 : FFT_MAXEL32;
  for (int q = 0; q < naxes; ++q) {
    const int ax = axes[q];
    const int64_t L64 = shape[ax];
    if (L64 <= 1) continue;
    DMX_ASSERT(L64 < (1ll << 30), "FFT length too large");
    const int L = (int)L64;
    int64_t outer = 1, inner = 1;
    for (int d = 0; d < ax; ++d) outer *= shape[d];
    for (int d = ax + 1; d < ndim; ++d) inner *= shape[d];
    if (is_smooth(L)) {
      auto tw = twiddles(L, fp64);
      job->keep.push_back(tw);
      tm.begin("FFT axis " + std::to_string(ax));
      const int L1 = fft_split(L);
      if (inner > 1 && maxel / L < 16 && L1 > 0) {
        // strided lines too long for >= 16 per workgroup: the two coalesced
        // column passes (64 columns per workgroup), through the scratch
        // array and back
        if (!job->scratch.p) job->scratch.reset((size_t)n * esz);
        auto t1 = twiddles(L1, fp64), t2 = twiddles(L / L1, fp64);
        job->keep.push_back(t1);
        job->keep.push_back(t2);
        run_columns(work, outer, L, L1, inner, fp64, sign, t1->p, t2->p,
                    tw->p, job->scratch.p, s);
      } else {
        if (L > maxel && !job->scratch.p) job->scratch.reset((size_t)n * esz);
        run_lines(work, outer, L, inner, fp64, sign, tw->p, job->scratch.p,
                  s);
      }
      tm.end();
      continue;
    }
    const int n2 = (int)smooth_length(2 * L64 - 1);
    Blue bt = bluestein_tables(L, n2, fp64, sign, s);
    auto tw2 = twiddles(n2, fp64);
    job->keep.push_back(bt.w);
    job->keep.push_back(bt.B);
    job->keep.push_back(tw2);
    const int64_t nb = outer * (int64_t)n2 * inner;
    if (job->blue.bytes < (size_t)nb * esz) job->blue.reset((size_t)nb * esz);
    if (n2 > maxel && job->blue_scratch.bytes < (size_t)nb * esz)
      job->blue_scratch.reset((size_t)nb * esz);
    BlueArgs b{};
    b.outer = outer;
    b.inner = inner;
    b.L = L;
    b.n2 = n2;
    b.conj_w = sign > 0 ? 1 : 0;
    b.n = nb;
    b.small = nb < (1ll << 32) ? 1 : 0;
    void *A = job->blue.p;
    tm.begin("Bluestein axis " + std::to_string(ax));
    if (fp64) launch_blue<double, 0>(work, A, bt.w->p, b, s);
    else launch_blue<float, 0>(work, A, bt.w->p, b, s);
    run_lines(A, outer, n2, inner, fp64, -1, tw2->p, job->blue_scratch.p, s);
    if (fp64) launch_blue<double, 1>(work, A, bt.B->p, b, s);
    else launch_blue<float, 1>(work, A, bt.B->p, b, s);
    run_lines(A, outer, n2, inner, fp64, 1, tw2->p, job->blue_scratch.p, s);
    b.n = outer * (int64_t)L * inner;
    if (fp64) launch_blue<double, 2>(work, A, bt.w->p, b, s);
    else launch_blue<float, 2>(work, A, bt.w->p, b, s);
    tm.end();
  }
  if (!out_contig) {
    tm.begin("copy out");
    if (fp64)
      copy_nd<double>(work, DMX_C128, cst, out, out_st, ndim, shape, 1.0, s);
    else
      copy_nd<float>(work, DMX_C64, cst, out, out_st, ndim, shape, 1.0, s);
    tm.end();
  }
  return job;
}

}  // namespace dmx
