// fft.hip -- hand-written mixed-radix Stockham c2c FFT over the oversampled
// grid (role of c2c / general_nd / cfftp passes, fft.h:848-850,
// fftnd_impl.h:593-, fft1d_impl.h:194-825).  Unnormalised; forward means
// exponent -1.  Grid sizes are 2*(2,3,5,7,11-smooth), so the radices are
// 8, 4, 2, 3, 5, 7, 11 (radix 8, then 4, first, as the reference factorises
// complex lengths, fft.h:195-217).
//
// A "line set" is every line along one axis of a row-major 3-D view
// (1-D and 2-D grids get leading unit axes); the two other axes index the
// lines and each may be restricted to the kept bands [0,(n+1)/2) and
// [N-n/2, N), which is the reference's pruned FFT (nufft.h:1287-1298,
// 1345-1356, 1749-1781, 1815-1851).
//
// LDS kernel: a workgroup stages Bl whole lines in LDS (coalesced: along
// rows for the last axis, across Bl adjacent columns otherwise), runs all
// Stockham passes in LDS with register staging (one LDS buffer, read all
// inputs -> barrier -> write all outputs) and writes the lines back.
// Lines too long for LDS use global-memory Stockham passes (out of place).
#include <cmath>
#include <vector>

#include "dmx.h"
#include "fft.h"

namespace dmx {

namespace {

template <typename T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) {
  C r; r.x = a.x + b.x; r.y = a.y + b.y; return r;
}
template <typename C> __device__ __forceinline__ C csub(C a, C b) {
  C r; r.x = a.x - b.x; r.y = a.y - b.y; return r;
}
// multiply by sign*i
template <typename C> __device__ __forceinline__ C cmuli(C a, int sign) {
  C r;
  if (sign < 0) { r.x = a.y; r.y = -a.x; }
  else { r.x = -a.y; r.y = a.x; }
  return r;
}

// cos/sin(2 pi k / R) for the odd radices
template <int R> struct Tw;
template <> struct Tw<3> {
  static constexpr double c[2] = {-0.5, -0.5};
  static constexpr double s[2] = {0.86602540378443864676, -0.86602540378443864676};
};
template <> struct Tw<5> {
  static constexpr double c[4] = {0.30901699437494742410, -0.80901699437494742410,
                                  -0.80901699437494742410, 0.30901699437494742410};
  static constexpr double s[4] = {0.95105651629515357212, 0.58778525229247312917,
                                  -0.58778525229247312917, -0.95105651629515357212};
};
template <> struct Tw<7> {
  static constexpr double c[6] = {0.62348980185873353053, -0.22252093395631440429,
                                  -0.90096886790241912624, -0.90096886790241912624,
                                  -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[6] = {0.78183148246802980871, 0.97492791218182360702,
                                  0.43388373911755812048, -0.43388373911755812048,
                                  -0.97492791218182360702, -0.78183148246802980871};
};
template <> struct Tw<11> {
  static constexpr double c[10] = {
      0.84125353283118116886, 0.41541501300188642553, -0.14231483827328514044,
      -0.65486073394528506406, -0.95949297361449738989, -0.95949297361449738989,
      -0.65486073394528506406, -0.14231483827328514044, 0.41541501300188642553,
      0.84125353283118116886};
  static constexpr double s[10] = {
      0.54064081745559758211, 0.90963199535451837141, 0.98982144188093273238,
      0.75574957435425828377, 0.28173255684142969771, -0.28173255684142969771,
      -0.75574957435425828377, -0.98982144188093273238, -0.90963199535451837141,
      -0.54064081745559758211};
};

// in-register DFT of length R, exponent sign*2*pi*i*r*k/R
template <typename T, int R, typename C>
__device__ __forceinline__ void dft(C *v, int sign) {
  if constexpr (R == 1) {
    (void)v; (void)sign;
  } else if constexpr (R == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    C a2 = cadd(v[1], v[3]), a3 = cmuli(csub(v[1], v[3]), sign);
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
  } else if constexpr (R == 8) {
    // two length-4 DFTs (even / odd inputs), odd outputs times W^k with
    // W = exp(sign 2 pi i / 8), then the length-2 combination
    C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft<T, 4>(e, sign);
    dft<T, 4>(o, sign);
    const T h = (T)0.70710678118654752440;
    const T sg = (T)sign;
    C w1, w3;
    w1.x = (o[1].x - sg * o[1].y) * h;  // (1 + sign i) / sqrt 2
    w1.y = (o[1].y + sg * o[1].x) * h;
    w3.x = (-o[3].x - sg * o[3].y) * h;  // (-1 + sign i) / sqrt 2
    w3.y = (-o[3].y + sg * o[3].x) * h;
    const C w2 = cmuli(o[2], sign);
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  } else {
    // odd R: pair r with R-r
    constexpr int H = (R - 1) / 2;
    C sp[H], sm[H];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      sp[r - 1] = cadd(v[r], v[R - r]);
      sm[r - 1] = csub(v[r], v[R - r]);
    }
    C y0 = v[0];
#pragma unroll
    for (int r = 0; r < H; ++r) y0 = cadd(y0, sp[r]);
    C out[R];
    out[0] = y0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      C a = v[0], b;
      b.x = T(0);
      b.y = T(0);
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const int m = (r * k) % R;  // m in 1..R-1
        const T cc = (T)Tw<R>::c[m - 1], ss = (T)Tw<R>::s[m - 1];
        a.x += sp[r - 1].x * cc;
        a.y += sp[r - 1].y * cc;
        b.x += sm[r - 1].x * ss;
        b.y += sm[r - 1].y * ss;
      }
      C ib = cmuli(b, sign);
      out[k] = cadd(a, ib);
      out[R - k] = csub(a, ib);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
}

// Quotient and remainder of x / d for 0 <= x < 2^22 and 0 < d, from a
// float reciprocal and one correction step: the estimate is within one of
// the true quotient because x * inv carries a relative error below 2^-22.
// Integer division compiles to a long VALU sequence, which dominated the
// butterflies before.
__device__ __forceinline__ int fdivq(int x, int d, float inv, int &r) {
  int q = (int)((float)x * inv);
  r = x - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
  return q;
}

// One radix-R Stockham pass over Bl LDS-resident lines (stride LP), with
// the transform sign SIGN known at compile time.  The inputs of all
// butterflies a thread owns are read and twiddled before the barrier (tw: L
// twiddles), so after it only the DFTs and the writes remain.
template <typename T, int R, int KM, int SIGN>
__device__ __forceinline__ void lds_pass(typename V2<T>::t *buf, int LP,
                                         int Bl, const FftPass &ps,
                                         const typename V2<T>::t *tw) {
  using C = typename V2<T>::t;
  const int nbf = ps.nbf, Ns = ps.ns, twstep = ps.twstep;
  const int total = nbf * Bl;
  C v[KM][R];
  int obase[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int q = threadIdx.x + k * blockDim.x;
    if (q < total) {
      int jj;
      const int b = Bl == 1 ? (jj = q, 0) : fdivq(q, nbf, ps.inv_nbf, jj);
#pragma unroll
      for (int r = 0; r < R; ++r) v[k][r] = buf[b * LP + jj + r * nbf];
      int kk;
      const int jq = Ns == 1 ? (kk = 0, jj) : fdivq(jj, Ns, ps.inv_ns, kk);
      if (Ns > 1) {
        const int t1 = kk * twstep;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C w = tw[r * t1];
          if (SIGN > 0) w.y = -w.y;
          v[k][r] = cmul(v[k][r], w);
        }
      }
      obase[k] = b * LP + jq * Ns * R + kk;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int q = threadIdx.x + k * blockDim.x;
    if (q < total) {
      dft<T, R>(v[k], SIGN);
#pragma unroll
      for (int r = 0; r < R; ++r) buf[obase[k] + r * ps.ns] = v[k][r];
    }
  }
  __syncthreads();
}

// all passes of one line set staged in LDS
template <typename T, int MAXEL, int SIGN>
__device__ __forceinline__ void lds_passes(typename V2<T>::t *buf, int LP,
                                           int Bl, int npass,
                                           const FftPass *pass,
                                           const typename V2<T>::t *tw) {
  for (int s = 0; s < npass; ++s) {
    const FftPass ps = pass[s];
    switch (ps.radix) {
      case 8: lds_pass<T, 8, FFT_KM_EL(MAXEL, 8), SIGN>(buf, LP, Bl, ps, tw); break;
      case 4: lds_pass<T, 4, FFT_KM_EL(MAXEL, 4), SIGN>(buf, LP, Bl, ps, tw); break;
      case 2: lds_pass<T, 2, FFT_KM_EL(MAXEL, 2), SIGN>(buf, LP, Bl, ps, tw); break;
      case 3: lds_pass<T, 3, FFT_KM_EL(MAXEL, 3), SIGN>(buf, LP, Bl, ps, tw); break;
      case 5: lds_pass<T, 5, FFT_KM_EL(MAXEL, 5), SIGN>(buf, LP, Bl, ps, tw); break;
      case 7: lds_pass<T, 7, FFT_KM_EL(MAXEL, 7), SIGN>(buf, LP, Bl, ps, tw); break;
      case 11: lds_pass<T, 11, FFT_KM_EL(MAXEL, 11), SIGN>(buf, LP, Bl, ps, tw); break;
      default: break;
    }
  }
}

// One radix-R Stockham pass over a single line (one line per workgroup)
// whose input (IN_G) or output (OUT_G) is the line in global memory instead
// of LDS: the first pass reads its butterflies' elements jj + r * nbf
// directly (consecutive lanes read consecutive elements), the last one
// writes its results obase + r * ns directly; the pruned ranges of the
// launch read as zero / are not stored.  A lane past the last butterfly
// loads the last butterfly's elements (unconditional loads, all in flight
// at once) and writes nothing.  No barrier is needed between reading
// global input and writing LDS, nor after writing global output.
template <typename T, int R, int KM, int SIGN, bool IN_G, bool OUT_G>
__device__ __forceinline__ void row_pass(typename V2<T>::t *buf,
                                         const FftPass &ps,
                                         const typename V2<T>::t *tw,
                                         typename V2<T>::t *data,
                                         int64_t lbase, const FftLaunch &p) {
  using C = typename V2<T>::t;
  const int nbf = ps.nbf, Ns = ps.ns, twstep = ps.twstep;
  C v[KM][R];
  int obase[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int q = threadIdx.x + k * blockDim.x;
    const int jj = IN_G ? min(q, nbf - 1) : q;
    if (IN_G || q < nbf) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = jj + r * nbf;
        if constexpr (IN_G) {
          // an element known to be zero is not fetched: its load repeats
          // the butterfly's first element and the value is replaced
          const bool zero = p.kprune && e >= p.kn1 && e < p.ka2;
          C x = data[lbase + (int64_t)(zero ? jj : e) * p.st_ax];
          if (zero) {
            x.x = T(0);
            x.y = T(0);
          }
          v[k][r] = x;
        } else {
          v[k][r] = buf[e];
        }
      }
      int kk;
      const int jq = Ns == 1 ? (kk = 0, jj) : fdivq(jj, Ns, ps.inv_ns, kk);
      if (Ns > 1) {
        const int t1 = kk * twstep;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C w = tw[r * t1];
          if (SIGN > 0) w.y = -w.y;
          v[k][r] = cmul(v[k][r], w);
        }
      }
      obase[k] = jq * Ns * R + kk;
    }
  }
  if (!IN_G) __syncthreads();
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int q = threadIdx.x + k * blockDim.x;
    if (q < nbf) {
      dft<T, R>(v[k], SIGN);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = obase[k] + r * ps.ns;
        if constexpr (OUT_G) {
          if (!(p.sprune && e >= p.sn1 && e < p.sa2))
            data[lbase + (int64_t)e * p.st_ax] = v[k][r];
        } else {
          buf[e] = v[k][r];
        }
      }
    }
  }
  if (!OUT_G) __syncthreads();
}

template <typename T, int MAXEL, int SIGN, bool IN_G, bool OUT_G>
__device__ __forceinline__ void row_pass_radix(typename V2<T>::t *buf,
                                               const FftPass &ps,
                                               const typename V2<T>::t *tw,
                                               typename V2<T>::t *data,
                                               int64_t lbase,
                                               const FftLaunch &p) {
#define DMX_ROWPASS(R) \
  row_pass<T, R, FFT_KM_EL(MAXEL, R), SIGN, IN_G, OUT_G>(buf, ps, tw, data, \
                                                        lbase, p)
  switch (ps.radix) {
    case 8: DMX_ROWPASS(8); break;
    case 4: DMX_ROWPASS(4); break;
    case 2: DMX_ROWPASS(2); break;
    case 3: DMX_ROWPASS(3); break;
    case 5: DMX_ROWPASS(5); break;
    case 7: DMX_ROWPASS(7); break;
    case 11: DMX_ROWPASS(11); break;
    default: break;
  }
#undef DMX_ROWPASS
}

// all passes of a single line: the first reads the line from global
// memory, the last writes it back, the others run in LDS
template <typename T, int MAXEL, int SIGN>
__device__ __forceinline__ void row_passes(typename V2<T>::t *buf,
                                           const FftLaunch &p,
                                           const typename V2<T>::t *tw,
                                           typename V2<T>::t *data,
                                           int64_t lbase) {
  const int np = p.npass;
  if (np == 1) {
    row_pass_radix<T, MAXEL, SIGN, true, true>(buf, p.pass[0], tw, data,
                                               lbase, p);
    return;
  }
  row_pass_radix<T, MAXEL, SIGN, true, false>(buf, p.pass[0], tw, data, lbase,
                                              p);
  if (np > 2)
    lds_passes<T, MAXEL, SIGN>(buf, p.L, 1, np - 2, p.pass + 1, tw);
  row_pass_radix<T, MAXEL, SIGN, false, true>(buf, p.pass[np - 1], tw, data,
                                              lbase, p);
}

__device__ __forceinline__ int64_t map_range(int v, int n1, int a1, int a2) {
  return v < n1 ? (int64_t)(a1 + v) : (int64_t)(a2 + (v - n1));
}

// MAXEL: the most elements (lines x length) the launch puts in one
// workgroup; it sizes the register staging of the passes, and the half-size
// variant leaves room for more workgroups per CU
template <typename T, int MAXEL, int SIGN>
__global__ __launch_bounds__(FFT_THREADS) __attribute__((
    amdgpu_waves_per_eu(sizeof(T) == 4 && MAXEL <= 4096 ? 5 : 1))) void
k_fft_lds(FftLaunch p) {
  using C = typename V2<T>::t;
  extern __shared__ unsigned char smem_raw[];
  C *buf = (C *)smem_raw;
  C *data = (C *)p.data;
  const int L = p.L, LP = p.LP, Bl = p.Bl;
  const C *tw = (const C *)p.tw;
  const int u = blockIdx.y;
  const int v0 = blockIdx.x * Bl;
  const int64_t ubase = map_range(u, p.un1, p.ua1, p.ua2) * p.st_u;
  const int nlines = min(Bl, p.nv - v0);
  const int nel = nlines * L;
  const float inv_L = 1.0f / (float)L, inv_nl = 1.0f / (float)nlines;
  // load (unrolled: several loads of a lane in flight before their LDS
  // writes); one line per workgroup needs no element -> (line, index)
  // division
  const int64_t lbase =
      ubase + map_range(v0, p.vn1, p.va1, p.va2) * p.st_v;
  if (nlines == 1 && p.npass >= 1) {
    // one line: the first and last passes read and write global memory
    row_passes<T, MAXEL, SIGN>(buf, p, tw, data, lbase);
    return;
  }
  {
#pragma unroll 4
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
      int b, k;
      if (p.st_ax == 1) b = fdivq(e, L, inv_L, k);
      else k = fdivq(e, nlines, inv_nl, b);
      C x;
      if (p.kprune && k >= p.kn1 && k < p.ka2) {
        x.x = T(0);
        x.y = T(0);
      } else {
        const int64_t addr = ubase +
                             map_range(v0 + b, p.vn1, p.va1, p.va2) * p.st_v +
                             k * p.st_ax;
        x = data[addr];
      }
      buf[b * LP + k] = x;
    }
  }
  __syncthreads();
  lds_passes<T, MAXEL, SIGN>(buf, LP, nlines, p.npass, p.pass, tw);
  // store
  if (nlines == 1) {
#pragma unroll 4
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
      if (p.sprune && k >= p.sn1 && k < p.sa2) continue;
      data[lbase + (int64_t)k * p.st_ax] = buf[k];
    }
  } else {
#pragma unroll 4
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
      int b, k;
      if (p.st_ax == 1) b = fdivq(e, L, inv_L, k);
      else k = fdivq(e, nlines, inv_nl, b);
      if (p.sprune && k >= p.sn1 && k < p.sa2) continue;
      const int64_t addr = ubase +
                           map_range(v0 + b, p.vn1, p.va1, p.va2) * p.st_v +
                           k * p.st_ax;
      data[addr] = buf[b * LP + k];
    }
  }
}

// Sub-FFTs over rows of Bl adjacent columns (one half of the two-pass
// column FFT, see fft_columns_2pass): line u, element e sits at row
// u*ia + e*ib of the source; output element k goes to row u*oa + k*ob of
// the destination, optionally times w_L^{(k*u) mod L}.  Loads and stores
// move Bl contiguous columns per row (coalesced).
template <typename T, int MAXEL, int SIGN>
__global__ __launch_bounds__(FFT_THREADS) __attribute__((
    amdgpu_waves_per_eu(sizeof(T) == 4 && MAXEL <= 4096 ? 5 : 1))) void
k_fft_cols(FftCols p) {
  using C = typename V2<T>::t;
  extern __shared__ unsigned char smem_raw[];
  C *buf = (C *)smem_raw;
  const C *src = (const C *)p.src;
  C *dst = (C *)p.dst;
  const int Ls = p.Lsub, LP = p.Lsub + 1;
  // the Lsub twiddles of the sub-FFTs sit behind the columns in LDS
  C *tws = buf + (size_t)p.Bl * LP;
  for (int i = threadIdx.x; i < Ls; i += blockDim.x)
    tws[i] = ((const C *)p.tw_sub)[i];
  const int u = blockIdx.y;
  const int v0 = blockIdx.x * p.Bl;
  const int nb = min(p.Bl, p.ncol - v0);
  const int64_t obase =
      map_range(blockIdx.z, p.on1, p.oa1, p.oa2) * p.st_o;
  const int nel = nb * Ls;
  const float inv_nb = 1.0f / (float)nb, inv_L = 1.0f / (float)p.L;
  // element i = e * nb + b (row e, column b): a lane's first (e, b) from one
  // division, the next ones (i + FFT_THREADS) by a constant step
  int sb_;
  const int se = fdivq(FFT_THREADS, nb, inv_nb, sb_);
  int b_first;
  const int e_first = fdivq((int)threadIdx.x, nb, inv_nb, b_first);
  // a lane's global loads are all issued before its LDS writes (nel < MAXEL)
  constexpr int NPER = (MAXEL + FFT_THREADS - 1) / FFT_THREADS;
  {
    C tmp[NPER];
    int slot[NPER];
    int e = e_first, b = b_first;
#pragma unroll
    for (int j = 0; j < NPER; ++j) {
      const int i = (int)threadIdx.x + j * FFT_THREADS;
      slot[j] = -1;
      if (i < nel) {
        const int64_t row = (int64_t)u * p.ia + (int64_t)e * p.ib;
        slot[j] = b * LP + e;
        if (p.kprune && row >= p.kn1 && row < p.ka2) {
          tmp[j].x = T(0);
          tmp[j].y = T(0);
        } else {
          tmp[j] = src[obase + row * p.rs +
                       map_range(v0 + b, p.vn1, p.va1, p.va2)];
        }
      }
      b += sb_;
      e += se;
      if (b >= nb) {
        b -= nb;
        ++e;
      }
    }
#pragma unroll
    for (int j = 0; j < NPER; ++j)
      if (slot[j] >= 0) buf[slot[j]] = tmp[j];
  }
  __syncthreads();
  lds_passes<T, MAXEL, SIGN>(buf, LP, nb, p.npass, p.pass, tws);
  const C *twf = (const C *)p.tw_full;
  int kn = e_first, bn = b_first;
  for (int i = threadIdx.x; i < nel; i += FFT_THREADS) {
    // this element (k, b), then the lane's next one
    const int k = kn, b = bn;
    bn += sb_;
    kn += se;
    if (bn >= nb) {
      bn -= nb;
      ++kn;
    }
    C x = buf[b * LP + k];
    if (p.twiddle && u > 0 && k > 0) {
      int m;
      fdivq(k * u, p.L, inv_L, m);  // k * u < L1 * L2 = L
      C w = twf[m];
      if (SIGN > 0) w.y = -w.y;
      x = cmul(x, w);
    }
    const int64_t row = (int64_t)u * p.oa + (int64_t)k * p.ob;
    if (p.sprune && row >= p.sn1 && row < p.sa2) continue;
    const int col = (int)map_range(v0 + b, p.vn1, p.va1, p.va2);
    if (p.crop) {
      // grid index -> mode index (retained rows and columns only reach
      // here), correction index and index into the uniform array
      const FftCrop &q = p.cr;
      const int g0 = (int)row, h0 = q.n0 / 2, h1 = q.n1 / 2;
      const int i0 = g0 < q.n0 - h0 ? g0 + h0 : g0 - (q.N0 - h0);
      const int i1 = col < q.n1 - h1 ? col + h1 : col - (q.N1 - h1);
      const int c0 = h0 > i0 ? h0 - i0 : i0 - h0;
      const int c1 = h1 > i1 ? h1 - i1 : i1 - h1;
      int u0 = q.fft_order ? i0 + q.n0 - h0 : i0;
      if (u0 >= q.n0) u0 -= q.n0;
      int u1 = q.fft_order ? i1 + q.n1 - h1 : i1;
      if (u1 >= q.n1) u1 -= q.n1;
      const T fct = (T)(q.cf0[c0] * q.cf1[c1]);
      const int64_t ui = (int64_t)u0 * q.ust0 + (int64_t)u1 * q.ust1;
      if (q.u_f64)
        ((double2 *)q.uni)[ui] =
            make_double2((double)(x.x * fct), (double)(x.y * fct));
      else
        ((float2 *)q.uni)[ui] =
            make_float2((float)(x.x * fct), (float)(x.y * fct));
      continue;
    }
    dst[obase + row * p.rs + col] = x;
  }
}

// One out-of-place radix-R pass in global memory over all lines of the set.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_fft_global_pass(FftLaunch p,
                                                         const void *src_,
                                                         void *dst_, int Ns) {
  using C = typename V2<T>::t;
  const C *src = (const C *)src_;
  C *dst = (C *)dst_;
  const C *tw = (const C *)p.tw;
  const int L = p.L;
  const int nbf = L / R;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nl = (int64_t)p.nu * p.nv;
  if (q >= nl * nbf) return;
  const int64_t line = q / nbf;
  const int jj = (int)(q - line * nbf);
  const int u = (int)(line / p.nv), v = (int)(line - (int64_t)u * p.nv);
  const int64_t base = map_range(u, p.un1, p.ua1, p.ua2) * p.st_u +
                       map_range(v, p.vn1, p.va1, p.va2) * p.st_v;
  C x[R];
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = src[base + (int64_t)(jj + r * nbf) * p.st_ax];
  if (R > 1) {
    const int kk = jj % Ns;
    const int twstep = L / (Ns * R);
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = tw[(int64_t)r * kk * twstep];
        if (p.sign > 0) w.y = -w.y;
        x[r] = cmul(x[r], w);
      }
    }
    dft<T, R>(x, p.sign);
    const int ob = (jj / Ns) * Ns * R + kk;
#pragma unroll
    for (int r = 0; r < R; ++r)
      dst[base + (int64_t)(ob + r * Ns) * p.st_ax] = x[r];
  } else {
    dst[base + (int64_t)jj * p.st_ax] = x[0];
  }
}

std::vector<int> factorize(int L);
}  // namespace

int fft_plan_passes(int L, FftPass *pass, int maxpass) {
  const auto f = factorize(L);
  DMX_ASSERT((int)f.size() <= maxpass, "too many FFT passes");
  int Ns = 1;
  for (size_t i = 0; i < f.size(); ++i) {
    FftPass &q = pass[i];
    q.radix = f[i];
    q.nbf = L / f[i];
    q.ns = Ns;
    q.twstep = L / (Ns * f[i]);
    q.inv_nbf = 1.0f / (float)q.nbf;
    q.inv_ns = 1.0f / (float)Ns;
    Ns *= f[i];
  }
  return (int)f.size();
}

namespace {
std::vector<int> factorize(int L) {
  std::vector<int> f;
  for (int r : {8, 4, 2, 3, 5, 7, 11})
    while (L % r == 0 && L > 1) {
      f.push_back(r);
      L /= r;
    }
  DMX_ASSERT(L == 1, "FFT length is not 2,3,5,7,11-smooth");
  return f;
}

}  // namespace

void fft_twiddles(int L, bool fp64, std::vector<unsigned char> &out) {
  // exp(-2 pi i m / L), m = 0..L-1, computed in double
  const double pi = 3.141592653589793238462643383279502884197;
  size_t esz = fp64 ? 16 : 8;
  out.resize((size_t)L * esz);
  for (int m = 0; m < L; ++m) {
    // reduce the angle to the first octant-symmetric form for accuracy
    double c, s;
    long mm = m;
    // use sincospi-like evaluation on 2m/L in [0,2)
    double x = 2.0 * (double)mm / (double)L;
    if (x > 1.0) x -= 2.0;  // (-1, 1]
    c = std::cos(pi * x);
    s = -std::sin(pi * x);
    if (4 * (long)m == (long)L) { c = 0; s = -1; }
    if (2 * (long)m == (long)L) { c = -1; s = 0; }
    if (4 * (long)m == 3 * (long)L) { c = 0; s = 1; }
    if (fp64) {
      double *p = (double *)out.data() + 2 * (size_t)m;
      p[0] = c; p[1] = s;
    } else {
      float *p = (float *)out.data() + 2 * (size_t)m;
      p[0] = (float)c; p[1] = (float)s;
    }
  }
}

namespace {
__global__ void k_twiddles(void *out, int L, int fp64) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= L) return;
  double x = 2.0 * (double)m / (double)L;
  if (x > 1.0) x -= 2.0;  // (-1, 1]
  double sn, c;
  sincospi(x, &sn, &c);
  double s = -sn;
  if (4 * (long)m == (long)L) { c = 0; s = -1; }
  if (2 * (long)m == (long)L) { c = -1; s = 0; }
  if (4 * (long)m == 3 * (long)L) { c = 0; s = 1; }
  if (fp64)
    ((double2 *)out)[m] = make_double2(c, s);
  else
    ((float2 *)out)[m] = make_float2((float)c, (float)s);
}
}  // namespace

void launch_fft_twiddles(void *dst, int L, bool fp64, hipStream_t s) {
  if (L <= 0) return;
  hipLaunchKernelGGL(k_twiddles, dim3((L + 255) / 256), dim3(256), 0, s, dst,
                     L, fp64 ? 1 : 0);
  DMX_HIP(hipGetLastError());
}

int fft_split(int L) {
  int best = 0;
  for (int d = 2; d * d <= L; ++d)
    if (L % d == 0 && L / d <= 128) best = d;
  return best;
}

bool fft_columns_2pass(const FftLaunch &p, int L1, void *dst, bool fp64,
                       const void *tw_L1, const void *tw_L2,
                       const void *tw_L, hipStream_t s, void *a_dst,
                       const FftCrop *crop) {
  const int L = p.L, L2 = L / L1;
  if (L1 < 2 || L2 < 2 || L1 * L2 != L || p.st_v != 1) return false;
  DMX_ASSERT(p.nu <= 65535 && L1 <= 65535 && L2 <= 65535,
             "FFT line set too large");
  const int maxel = fp64 ? FFT_MAXEL64 : FFT_MAXEL32;
  // columns per workgroup: at least 32 (coalesced row segments), within
  // half the LDS budget when that allows 32 (more workgroups per CU)
  const int Lm = std::max(L1, L2);
  int Bl = (maxel / 2 - Lm) / (Lm + 1);
  const bool half = Bl >= 32;
  if (!half) Bl = (maxel - Lm) / (Lm + 1);
  Bl = std::min(Bl, fp64 ? 32 : 64);
  Bl = std::max(1, std::min(Bl, p.nv));
  const size_t esz = fp64 ? 16 : 8;
  FftCols c{};
  c.L = L;
  c.sign = p.sign;
  c.Bl = Bl;
  c.rs = p.st_ax;
  c.ncol = p.nv;
  c.vn1 = p.vn1; c.va1 = p.va1; c.va2 = p.va2;
  c.nouter = p.nu;
  c.on1 = p.un1; c.oa1 = p.ua1; c.oa2 = p.ua2;
  c.st_o = p.st_u;
  c.tw_full = tw_L;
  auto launch = [&](int Lsub) {
    c.Lsub = Lsub;
    c.npass = fft_plan_passes(Lsub, c.pass, 24);
    size_t sh = ((size_t)Bl * (Lsub + 1) + Lsub) * esz;
    dim3 grid((unsigned)((p.nv + Bl - 1) / Bl), (unsigned)c.nlines,
              (unsigned)p.nu);
#define DMX_COLS(TT, ME, SG) \
  hipLaunchKernelGGL((k_fft_cols<TT, ME, SG>), grid, dim3(FFT_THREADS), sh, s, c)
    const bool neg = c.sign < 0;
    if (fp64) {
      if (half) { if (neg) DMX_COLS(double, FFT_MAXEL64 / 2, -1); else DMX_COLS(double, FFT_MAXEL64 / 2, 1); }
      else { if (neg) DMX_COLS(double, FFT_MAXEL64, -1); else DMX_COLS(double, FFT_MAXEL64, 1); }
    } else {
      if (half) { if (neg) DMX_COLS(float, FFT_MAXEL32 / 2, -1); else DMX_COLS(float, FFT_MAXEL32 / 2, 1); }
      else { if (neg) DMX_COLS(float, FFT_MAXEL32, -1); else DMX_COLS(float, FFT_MAXEL32, 1); }
    }
#undef DMX_COLS
    DMX_HIP(hipGetLastError());
  };
  // pass A: length L1 over rows n1*L2+n2, twiddled (in place unless a_dst)
  void *amid = a_dst ? a_dst : p.data;
  c.src = p.data;
  c.dst = amid;
  c.tw_sub = tw_L1;
  c.twiddle = 1;
  // pass A reads natural order: known-zero input rows as asked
  c.kprune = p.kprune;
  c.kn1 = p.kn1;
  c.ka2 = p.ka2;
  c.ia = 1; c.ib = L2; c.oa = 1; c.ob = L2;
  c.nlines = L2;
  launch(L1);
  // pass B: length L2 over rows k1*L2+n2, to row k1 + L1*k2 of dst
  c.src = amid;
  c.dst = dst;
  c.tw_sub = tw_L2;
  c.twiddle = 0;
  c.ia = L2; c.ib = 1; c.oa = 1; c.ob = L1;
  c.nlines = L1;
  c.kprune = 0;
  if (crop) {
    c.crop = 1;
    c.cr = *crop;
  }
  c.sprune = p.sprune;  // pass B writes natural order: prune as asked
  c.sn1 = p.sn1;
  c.sa2 = p.sa2;
  launch(L2);
  return true;
}

void fft_lineset(const FftLaunch &base, bool fp64, void *scratch,
                 hipStream_t s) {
  FftLaunch p = base;
  if (p.nu <= 0 || p.nv <= 0 || p.L <= 1) return;
  p.npass = fft_plan_passes(p.L, p.pass, 24);
  const size_t esz = fp64 ? 16 : 8;
  const int maxel = fp64 ? FFT_MAXEL64 : FFT_MAXEL32;
  if (p.L <= maxel) {
    int Bl = maxel / p.L;
    // contiguous lines: about 32 KiB of LDS per workgroup, so that several
    // workgroups per CU overlap their loads with the passes of the others
    if (p.st_ax == 1) Bl = std::min(Bl, std::max(1, (maxel / 2) / p.L));
    Bl = std::max(1, std::min(Bl, p.nv));
    p.Bl = Bl;
    p.LP = p.L;
    size_t sh = (size_t)Bl * p.LP * esz;
    dim3 grid((unsigned)((p.nv + Bl - 1) / Bl), (unsigned)p.nu);
    DMX_ASSERT(p.nu <= 65535, "FFT line set too large");
    const bool half = (size_t)Bl * p.L <= (size_t)maxel / 2;
#define DMX_LDS(TT, ME, SG) \
  hipLaunchKernelGGL((k_fft_lds<TT, ME, SG>), grid, dim3(FFT_THREADS), sh, s, p)
    const bool neg = p.sign < 0;
    if (fp64) {
      if (half) { if (neg) DMX_LDS(double, FFT_MAXEL64 / 2, -1); else DMX_LDS(double, FFT_MAXEL64 / 2, 1); }
      else { if (neg) DMX_LDS(double, FFT_MAXEL64, -1); else DMX_LDS(double, FFT_MAXEL64, 1); }
    } else {
      if (half) { if (neg) DMX_LDS(float, FFT_MAXEL32 / 2, -1); else DMX_LDS(float, FFT_MAXEL32 / 2, 1); }
      else { if (neg) DMX_LDS(float, FFT_MAXEL32, -1); else DMX_LDS(float, FFT_MAXEL32, 1); }
    }
#undef DMX_LDS
    DMX_HIP(hipGetLastError());
    return;
  }
  // long lines: global Stockham passes, ping-pong between data and scratch
  DMX_ASSERT(scratch != nullptr, "internal: FFT scratch missing");
  void *src = p.data, *dst = scratch;
  int Ns = 1;
  auto run = [&](int R) {
    int64_t nthr = (int64_t)p.nu * p.nv * (p.L / R);
    unsigned nb = (unsigned)((nthr + 255) / 256);
#define DMX_GP(TT, RR) \
  hipLaunchKernelGGL((k_fft_global_pass<TT, RR>), dim3(nb), dim3(256), 0, s, p, src, dst, Ns)
    if (fp64) {
      switch (R) {
        case 1: DMX_GP(double, 1); break; case 2: DMX_GP(double, 2); break;
        case 3: DMX_GP(double, 3); break; case 4: DMX_GP(double, 4); break;
        case 8: DMX_GP(double, 8); break;
        case 5: DMX_GP(double, 5); break; case 7: DMX_GP(double, 7); break;
        case 11: DMX_GP(double, 11); break;
      }
    } else {
      switch (R) {
        case 1: DMX_GP(float, 1); break; case 2: DMX_GP(float, 2); break;
        case 3: DMX_GP(float, 3); break; case 4: DMX_GP(float, 4); break;
        case 8: DMX_GP(float, 8); break;
        case 5: DMX_GP(float, 5); break; case 7: DMX_GP(float, 7); break;
        case 11: DMX_GP(float, 11); break;
      }
    }
#undef DMX_GP
    DMX_HIP(hipGetLastError());
    std::swap(src, dst);
  };
  for (int i = 0; i < p.npass; ++i) {
    run(p.pass[i].radix);
    Ns *= p.pass[i].radix;
  }
  if (src != p.data) {
    // odd number of passes: copy back with an identity pass
    Ns = 1;
    run(1);
  }
}

}  // namespace dmx
