// ptycho_gen.hip — multislice forward / loss / adjoint for patterns of any size Nr x Nc (2..512 each)
// whose prime factors are 2, 3, 5 and 7, for the sizes the power-of-two kernels do not take (non-square,
// 96 x 96, 192 x 160, 16 x 16, ...).
//
// The chain has the shape of the 256 x 256 one (ptycho256.hip): each workgroup owns one (pattern, mode)
// item and keeps its wave in a per-item buffer W in the row-transformed domain (r, kc), and a slice
// costs one row pass and one column pass over W:
//
//   row pass     chunks of rows: inverse row DFT -> natural-order wave -> elementwise work (incident-wave
//                store, transmission, descan, the adjoint's gradient patch) -> forward row DFT -> W
//   column pass  chunks of columns: column DFT -> 2-D spectrum -> propagator (far field, probe-gradient
//                epilogue) -> inverse column DFT -> W
//
// A chunk is staged in LDS and transformed there by a Stockham autosort FFT (mixed radix 8, 4, 2, 3, 5, 7,
// ping-pong between two LDS buffers, twiddles from a per-length table built in double precision).  A
// butterfly of radix R at a stage with stride s reads a_j = x[w + j n/R] (w = p s + q) and writes
// y[q + s (R p + k)] = W_n^{p s k} sum_j a_j W_R^{j k}, the decimation-in-frequency form that leaves the
// result in natural order after the last stage.  Lines are padded to an odd stride so that the column
// chunks' transposed LDS accesses spread over the banks.  Everything follows the reference lines that
// ptycho.hip cites; fftfreq values are rounded to float32 from double as the reference builds them
// (ptycho_utils.py:167-168), patch offsets are fftfreq(n, 1/n) per axis (dataset_models.py:519-520) and
// the intensities are fftshifted by n // 2 per axis (detector_models.py:29).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "common.h"
#include "ptycho_common.h"

namespace qem {
namespace gen {

constexpr int T = 256;      // threads per workgroup
constexpr int EB = 2048;    // complex values per LDS chunk buffer
constexpr int NMAX = 512;   // longest axis
constexpr int MAXST = 12;   // Stockham stages (4 bits each in the packed radix code)

struct Smem {
  cf buf[2][EB];
  cf tw_r[NMAX];    // exp(-2 pi i k / Nr): column transforms
  cf tw_c[NMAX];    // exp(-2 pi i k / Nc): row transforms
  float fq_r[NMAX]; // fftfreq(Nr)[k] in float32
  float fq_c[NMAX]; // fftfreq(Nc)[k] in float32
  cf aux[NMAX];     // per-item column factors exp(-2 pi i fftfreq(Nc)[c] d_c) of the descan ramp
  cf aux_r[NMAX];   // per-item row factors exp(-2 pi i fftfreq(Nr)[r] d_r) of the descan ramp
  double red[T / 64];
};

// Radices of a length-n transform, 4 bits per stage, in stage order (8s first, then 4, 2, 3, 5, 7).
struct Axis {
  unsigned long long code;
  int ns;
};
__device__ __forceinline__ Axis make_axis(int n) {
  Axis a{0ull, 0};
  const int rads[6] = {8, 4, 2, 3, 5, 7};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const int r = rads[i];
    while (n % r == 0 && n > 1 && a.ns < MAXST) {
      a.code |= (unsigned long long)r << (4 * a.ns);
      ++a.ns;
      n /= r;
    }
  }
  return a;
}

struct Geo {
  int NR, NC;
  Axis ar, ac;  // column transforms (length NR), row transforms (length NC)
};

// fftfreq-order offset of index k on a length-n axis: k below ceil(n / 2), k - n from there.
__device__ __forceinline__ int foff(int k, int n) { return k < (n + 1) / 2 ? k : k - n; }

// numpy's fftfreq(n, d=1)[k] (a float64 quotient) rounded to float32.
__device__ __forceinline__ float fftfreq_f32(int k, int n) { return (float)((double)foff(k, n) / (double)n); }

__device__ __forceinline__ void init_tables(Smem& sm, int NR, int NC) {
  for (int k = threadIdx.x; k < NR; k += T) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)NR, &s, &c);
    sm.tw_r[k] = mk((float)c, (float)s);
    sm.fq_r[k] = fftfreq_f32(k, NR);
  }
  for (int k = threadIdx.x; k < NC; k += T) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)NC, &s, &c);
    sm.tw_c[k] = mk((float)c, (float)s);
    sm.fq_c[k] = fftfreq_f32(k, NC);
  }
  __syncthreads();
}

__device__ __forceinline__ Geo make_geo(const Params& p) {
  Geo g;
  g.NR = p.NR;
  g.NC = p.NC;
  g.ar = make_axis(p.NR);
  g.ac = make_axis(p.NC);
  return g;
}

// 4-point DFT (unnormalised; conjugate roots if INV).
template <bool INV>
__device__ __forceinline__ void dft4(cf a0, cf a1, cf a2, cf a3, cf& o0, cf& o1, cf& o2, cf& o3) {
  const cf t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), d = csub(a1, a3);
  const cf t3 = INV ? mk(-d.y, d.x) : mk(d.y, -d.x);  // +i d (inverse) or -i d (forward)
  o0 = cadd(t0, t2);
  o2 = csub(t0, t2);
  o1 = cadd(t1, t3);
  o3 = csub(t1, t3);
}

// cos and sin of 2 pi m / R for the odd radices (m = 1 .. (R - 1) / 2).
template <int R>
struct OddRoots;
template <>
struct OddRoots<3> {
  static constexpr float c[2] = {1.0f, -0.5f};
  static constexpr float s[2] = {0.0f, 0.866025403784438647f};
};
template <>
struct OddRoots<5> {
  static constexpr float c[3] = {1.0f, 0.309016994374947424f, -0.809016994374947424f};
  static constexpr float s[3] = {0.0f, 0.951056516295153572f, 0.587785252292473129f};
};
template <>
struct OddRoots<7> {
  static constexpr float c[4] = {1.0f, 0.623489801858733531f, -0.222520933956314404f, -0.900968867902419126f};
  static constexpr float s[4] = {0.0f, 0.781831482468029809f, 0.974927912181823607f, 0.433883739117558120f};
};

// Length-R DFT of a (unnormalised; conjugate roots if INV).  Odd R: with s_j = a_j + a_{R-j} and
// d_j = a_j - a_{R-j} (j = 1 .. h, h = (R - 1) / 2),
//   X_k = a_0 + sum_j s_j cos(2 pi j k / R) -+ i sum_j d_j sin(2 pi j k / R),  X_{R-k} the same with +-,
// the upper sign for the forward transform.
template <int R, bool INV>
__device__ __forceinline__ void dft(const cf (&a)[R], cf (&o)[R]) {
  if constexpr (R == 2) {
    o[0] = cadd(a[0], a[1]);
    o[1] = csub(a[0], a[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(a[0], a[1], a[2], a[3], o[0], o[1], o[2], o[3]);
  } else if constexpr (R == 8) {
    // two 4-point DFTs of the even and odd inputs, then X_k = E_k + W8^k O_k, X_{k+4} = E_k - W8^k O_k
    cf e0, e1, e2, e3, q0, q1, q2, q3;
    dft4<INV>(a[0], a[2], a[4], a[6], e0, e1, e2, e3);
    dft4<INV>(a[1], a[3], a[5], a[7], q0, q1, q2, q3);
    constexpr float h = 0.70710678118654752f;
    const cf w1 = INV ? mk(h * (q1.x - q1.y), h * (q1.x + q1.y)) : mk(h * (q1.x + q1.y), h * (q1.y - q1.x));
    const cf w2 = INV ? mk(-q2.y, q2.x) : mk(q2.y, -q2.x);
    const cf w3 = INV ? mk(-h * (q3.x + q3.y), h * (q3.x - q3.y)) : mk(h * (q3.y - q3.x), -h * (q3.x + q3.y));
    o[0] = cadd(e0, q0);
    o[4] = csub(e0, q0);
    o[1] = cadd(e1, w1);
    o[5] = csub(e1, w1);
    o[2] = cadd(e2, w2);
    o[6] = csub(e2, w2);
    o[3] = cadd(e3, w3);
    o[7] = csub(e3, w3);
  } else {
    constexpr int h = (R - 1) / 2;
    cf sj[h + 1], dj[h + 1];
    cf dc = a[0];
#pragma unroll
    for (int j = 1; j <= h; ++j) {
      sj[j] = cadd(a[j], a[R - j]);
      dj[j] = csub(a[j], a[R - j]);
      dc = cadd(dc, sj[j]);
    }
    o[0] = dc;
#pragma unroll
    for (int k = 1; k <= h; ++k) {
      cf re = a[0], im = mk(0.0f, 0.0f);
#pragma unroll
      for (int j = 1; j <= h; ++j) {
        int m = (j * k) % R;
        const float sg = m <= h ? 1.0f : -1.0f;  // sin(2 pi (R - m) / R) = -sin(2 pi m / R)
        m = m <= h ? m : R - m;
        re = mk(re.x + OddRoots<R>::c[m] * sj[j].x, re.y + OddRoots<R>::c[m] * sj[j].y);
        im = mk(im.x + sg * OddRoots<R>::s[m] * dj[j].x, im.y + sg * OddRoots<R>::s[m] * dj[j].y);
      }
      // forward: X_k = re - i im, X_{R-k} = re + i im; inverse: the other way round
      const cf mi = mk(im.y, -im.x);  // -i im
      o[k] = INV ? csub(re, mi) : cadd(re, mi);
      o[R - k] = INV ? cadd(re, mi) : csub(re, mi);
    }
  }
}

// Exact quotient x / d for 0 <= x < 2^22 and 1 <= d < 2^12 by a float reciprocal: (x + 1/2) / d lies at
// least 1/(2d) from an integer and the two roundings move it by less than (x + 1/2) 2^-23 / d.
struct Div {
  float r;
  __device__ __forceinline__ explicit Div(int d) : r(1.0f / (float)d) {}
  __device__ __forceinline__ int operator()(int x) const { return (int)(((float)x + 0.5f) * r); }
};

// One Stockham stage of radix R over lc lines of length n (line stride LS) from x to y.
template <int R, bool INV>
__device__ __forceinline__ void stage(const cf* x, cf* y, int lc, int n, int LS, int s, const cf* tw) {
  const int nq = n / R;  // butterflies per line
  const int nb = lc * nq;
  const Div dq(nq), dsd(s);
  for (int bi = threadIdx.x; bi < nb; bi += T) {
    const int l = dq(bi), w = bi - l * nq;
    const int p = dsd(w), q = w - p * s;
    const int ps = p * s;
    const cf* xi = x + l * LS + w;
    cf a[R], o[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = xi[j * nq];
    dft<R, INV>(a, o);
    cf* yo = y + l * LS + q + ps * R;
    yo[0] = o[0];
#pragma unroll
    for (int k = 1; k < R; ++k) {
      const cf t = tw[ps * k];  // W_n^{p s k}; p s k < n
      yo[k * s] = INV ? cmulc(o[k], t) : cmul(o[k], t);
    }
  }
}

// Unnormalised DFT (inverse: conjugate twiddles) of lc lines in LDS; x and y are the two buffers, the
// result is in the returned one.  Ends with a barrier.
template <bool INV>
__device__ cf* lds_fft(cf* x, cf* y, int lc, int n, const Axis& ax, const cf* tw) {
  const int LS = n | 1;
  int s = 1;
  for (int st = 0; st < ax.ns; ++st) {
    const int r = (int)((ax.code >> (4 * st)) & 15ull);
    switch (r) {
      case 8: stage<8, INV>(x, y, lc, n, LS, s, tw); break;
      case 4: stage<4, INV>(x, y, lc, n, LS, s, tw); break;
      case 2: stage<2, INV>(x, y, lc, n, LS, s, tw); break;
      case 3: stage<3, INV>(x, y, lc, n, LS, s, tw); break;
      case 5: stage<5, INV>(x, y, lc, n, LS, s, tw); break;
      default: stage<7, INV>(x, y, lc, n, LS, s, tw); break;
    }
    __syncthreads();
    cf* t = x;
    x = y;
    y = t;
    s *= r;
  }
  return x;
}

enum : int { NONE = 0, DFT_FWD = 1, DFT_INV = 2 };

constexpr int KPT = EB / T;  // chunk elements per thread

// Elementwise work of a pass, in two parts so that every thread issues all of its global loads before
// it uses any: fetch(r, c) loads what pixel (r, c) needs into an Aux, apply(r, c, v, aux) returns the new
// value (and may store).
struct Aux {
  cf a, b;
};
template <typename F, typename A>
struct Op2 {
  F fetch;
  A apply;
};
template <typename F, typename A>
__device__ __forceinline__ Op2<F, A> op2(F f, A a) {
  return Op2<F, A>{f, a};
}
struct NoFetch {
  __device__ __forceinline__ Aux operator()(int, int) const { return Aux{}; }
};
struct Keep {
  __device__ __forceinline__ cf operator()(int, int, cf v, const Aux&) const { return v; }
};
__device__ __forceinline__ Op2<NoFetch, Keep> no_op() { return Op2<NoFetch, Keep>{NoFetch(), Keep()}; }

// A pass over the lines of one item: rows (COLS false, lines of length NC) or columns (lines of length
// NR), in chunks that fit an LDS buffer.  Each chunk: load src -> PRE transform -> apply the op to every
// element -> POST transform -> store to dst (if not null).  Chunks touch disjoint lines and read all of
// theirs before writing, so src may equal dst.  Element e of a chunk is held by thread e % T; column
// chunks map consecutive elements to consecutive columns of a row (coalesced), row chunks to
// consecutive columns of one row.
template <int PRE, int POST, bool COLS, typename Op>
__device__ void pass(const cf* src, cf* dst, Smem& sm, const Geo& g, Op&& op) {
  const int n = COLS ? g.NR : g.NC;
  const int nl = COLS ? g.NC : g.NR;
  const int LS = n | 1;
  const int L = EB / LS;  // >= 3 for n <= 512
  const Axis& ax = COLS ? g.ar : g.ac;
  const cf* tw = COLS ? sm.tw_r : sm.tw_c;
  const int NC = g.NC;
  const Div dn(n);
  // element e = threadIdx.x + k T (k < KPT) of the chunk of lines [l0, l0 + lc): line l, index i
  auto split = [&](int l0, int lc, const Div& dl, int e, int& l, int& i) {
    if constexpr (COLS) {
      i = dl(e);
      l = e - i * lc;
    } else {
      // through the absolute element index: the split then depends on l0, so the compiler cannot
      // hoist every chunk's (l, i) out of the chunk and slice loops and keep them live in registers
      const int ea = e + l0 * n;
      const int la = dn(ea);
      i = ea - la * n;
      l = la - l0;
    }
  };
  auto gidx = [&](int l0, int l, int i) { return COLS ? i * NC + l0 + l : (l0 + l) * NC + i; };
  // the chunk's values are loaded one chunk ahead (before the previous chunk's transforms), and the
  // op's own global loads before the chunk's first transform: both latencies hide behind LDS work
  cf v[KPT];
  {
    const int lc = min(L, nl);
    const Div dl(lc);
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int e = threadIdx.x + k * T;
      int l, i;
      split(0, lc, dl, e, l, i);
      if (e < lc * n) v[k] = src[gidx(0, l, i)];
    }
  }
  for (int l0 = 0; l0 < nl; l0 += L) {
    const int lc = min(L, nl - l0);
    const int E = lc * n;
    const Div dl(lc);
    cf* A = sm.buf[0];
    cf* Bf = sm.buf[1];
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int e = threadIdx.x + k * T;
      int l, i;
      split(l0, lc, dl, e, l, i);
      if (e < E) A[l * LS + i] = v[k];
    }
    __syncthreads();
    Aux x[KPT];
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int e = threadIdx.x + k * T;
      int l, i;
      split(l0, lc, dl, e, l, i);
      if (e < E) x[k] = op.fetch(COLS ? i : l0 + l, COLS ? l0 + l : i);
    }
    if (l0 + L < nl) {  // the next chunk's lines are disjoint from this one's: safe to read before the stores
      const int l1 = l0 + L, lc1 = min(L, nl - l1);
      const Div dl1(lc1);
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const int e = threadIdx.x + k * T;
        int l, i;
        split(l1, lc1, dl1, e, l, i);
        if (e < lc1 * n) v[k] = src[gidx(l1, l, i)];
      }
    }
    cf* R = A;
    if constexpr (PRE != NONE) R = lds_fft<PRE == DFT_INV>(A, Bf, lc, n, ax, tw);
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int e = threadIdx.x + k * T;
      int l, i;
      split(l0, lc, dl, e, l, i);
      if (e < E) R[l * LS + i] = op.apply(COLS ? i : l0 + l, COLS ? l0 + l : i, R[l * LS + i], x[k]);
    }
    if constexpr (POST != NONE) {
      __syncthreads();
      R = lds_fft<POST == DFT_INV>(R, R == A ? Bf : A, lc, n, ax, tw);
    }
    if (dst != nullptr) {
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const int e = threadIdx.x + k * T;
        int l, i;
        split(l0, lc, dl, e, l, i);
        if (e < E) dst[gidx(l0, l, i)] = R[l * LS + i];
      }
    }
    __syncthreads();  // the chunk is stored before the next one overwrites the buffers
  }
}

// Object element of patch pixel (r, c) for a patch with origin (r0m, c0m) reduced modulo (H, W):
// row (r0 + fftfreq offset of r) mod H, likewise the column (H >= Nr, W >= Nc: one wrap at most).
__device__ __forceinline__ cf obj_patch(const cf* O, int r0m, int c0m, int r, int c, const Geo& g, int H, int W) {
  int y = r0m + foff(r, g.NR);
  y += y < 0 ? H : 0;
  y -= y >= H ? H : 0;
  int x = c0m + foff(c, g.NC);
  x += x < 0 ? W : 0;
  x -= x >= W ? W : 0;
  return O[(size_t)y * W + x];
}

// The descan ramp's row and column factors of pattern b into sm.aux_r / sm.aux (the ramp at (r, c) is
// aux_r[r] aux[c], ptychography_base.py:903-905 with ptycho_utils.py:156-176).
__device__ __forceinline__ void descan_tables(Smem& sm, const Geo& g, const float* descan, int b) {
  const float dr = descan[2 * b], dc = descan[2 * b + 1];
  for (int k = threadIdx.x; k < g.NR; k += T) sm.aux_r[k] = expm2pi(sm.fq_r[k] * dr);
  for (int k = threadIdx.x; k < g.NC; k += T) sm.aux[k] = expm2pi(sm.fq_c[k] * dc);
  __syncthreads();
}

// pk[m] = fft2(probe[m]); the blocks past the modes fill the shift-ramp table
// ramps[b][k] = exp(-2 pi i fftfreq(Nr)[k] dr_b) for k < Nr, then exp(-2 pi i fftfreq(Nc)[k] dc_b).
__global__ __launch_bounds__(T) void k_spectrum_gen(Params p) {
  const int NR = p.NR, NC = p.NC;
  if ((int)blockIdx.x >= p.M) {
    const int idx = ((int)blockIdx.x - p.M) * T + (int)threadIdx.x;
    if (idx >= p.B * (NR + NC)) return;
    const int b = idx / (NR + NC), k = idx - b * (NR + NC);
    int r0, c0;
    float dr, dc;
    origin_and_shift<0>(p.pos, b, r0, c0, dr, dc);
    p.ramps[idx] = k < NR ? expm2pi(fftfreq_f32(k, NR) * dr) : expm2pi(fftfreq_f32(k - NR, NC) * dc);
    return;
  }
  __shared__ Smem sm;
  const Geo g = make_geo(p);
  init_tables(sm, NR, NC);
  const size_t N2 = (size_t)NR * NC;
  const int m = blockIdx.x;
  cf* Wb = p.scratch + (size_t)m * N2;
  pass<NONE, DFT_FWD, false>(p.probe + (size_t)m * N2, Wb, sm, g, no_op());
  pass<DFT_FWD, NONE, true>(Wb, p.pk + (size_t)m * N2, sm, g, no_op());
}

// Forward multislice of one (pattern, mode); p.fwd_mode as in the other chains.  W holds the
// row-transformed wave between passes (unnormalised row DFTs; the column pass scales by 1 / (Nr Nc)).
template <bool DESCAN, bool PG>
__global__ __launch_bounds__(T) void k_forward_gen(Params p) {
  __shared__ Smem sm;
  const Geo g = make_geo(p);
  init_tables(sm, g.NR, g.NC);
  const int NR = g.NR, NC = g.NC;
  const size_t N2 = (size_t)NR * NC;
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<0>(p.pos, b, r0, c0, dr, dc);
  const int r0m = pmod(r0, p.H), c0m = pmod(c0, p.W);
  cf* Wb = p.scratch + (size_t)item * N2;
  const int mode = p.fwd_mode;
  const bool far = mode != FWD_RECOMPUTE;
  const float inv_n2 = (float)(1.0 / (double)N2);
  const float inv_sqrt = (float)(1.0 / sqrt((double)N2));
  const size_t HW = (size_t)p.H * p.W;
  auto slot = [&](int s) -> cf* {
    if (mode == FWD_STORE_ALL || mode == FWD_RECOMPUTE) return p.psi + ((size_t)item * p.wslots + (s - p.s0)) * N2;
    if (mode == FWD_CHECKPOINT && s % p.ck == 0) return p.ckpt + ((size_t)item * p.nck + s / p.ck) * N2;
    return nullptr;
  };

  if (far) {
    // psi_0 = ifft2(fft2(P_m) ramp_b): inverse column DFT here, inverse row DFT in the first row pass
    const cf* rr = p.ramps + (size_t)b * (NR + NC);
    pass<NONE, DFT_INV, true>(p.pk + (size_t)m * N2, Wb, sm, g,
                              op2([&](int r, int c) { return Aux{rr[r], rr[NR + c]}; },
                                  [&](int, int, cf v, const Aux& x) { return cscale(cmul(v, cmul(x.a, x.b)), inv_n2); }));
  }
  if (DESCAN && far) descan_tables(sm, g, p.descan, b);
  for (int s = p.s0; s < p.s1; ++s) {
    const cf* O = p.obj + (size_t)s * HW;
    cf* keep = slot(s);
    const bool transmit = far || s + 1 < p.s1;  // a recomputed segment needs only its incident waves
    const bool exit_ramp = DESCAN && far && s == p.S - 1;
    auto op = op2(
        [&](int r, int c) {
          Aux x;
          if (transmit) x.a = obj_patch(O, r0m, c0m, r, c, g, p.H, p.W);
          return x;
        },
        [&](int r, int c, cf v, const Aux& x) {
          if (keep != nullptr) st_stream(keep + (size_t)r * NC + c, v);
          if (transmit) {
            v = cmul(x.a, v);
            if (exit_ramp) v = cmul(v, cmul(sm.aux_r[r], sm.aux[c]));
          }
          return v;
        });
    if (mode == FWD_RECOMPUTE && s == p.s0) {
      // a segment of an earlier forward starts from its checkpoint wave (natural order)
      const cf* ck = p.ckpt + ((size_t)item * p.nck + p.s0 / p.ck) * N2;
      if (transmit)
        pass<NONE, DFT_FWD, false>(ck, Wb, sm, g, op);
      else
        pass<NONE, NONE, false>(ck, nullptr, sm, g, op);
    } else if (transmit) {
      pass<DFT_INV, DFT_FWD, false>(Wb, Wb, sm, g, op);
    } else {
      pass<DFT_INV, NONE, false>(Wb, nullptr, sm, g, op);
    }
    if (s + 1 < p.s1) {
      const cf* pr = p.prop + (size_t)s * N2;
      cf* xs = PG ? p.pspec + ((size_t)item * (p.S - 1) + s) * N2 : nullptr;
      pass<DFT_FWD, DFT_INV, true>(Wb, Wb, sm, g,
                                   op2([&](int r, int c) { return Aux{pr[(size_t)r * NC + c], cf{}}; },
                                       [&](int r, int c, cf v, const Aux& x) {
                                         // fft2 of the transmitted wave, for the propagator gradient
                                         if constexpr (PG) st_stream(xs + (size_t)r * NC + c, v);
                                         return cscale(cmul(v, x.a), inv_n2);
                                       }));
    } else if (far) {
      pass<DFT_FWD, NONE, true>(Wb, p.X + (size_t)item * N2, sm, g,
                                op2(NoFetch(), [&](int, int, cf v, const Aux&) { return cscale(v, inv_sqrt); }));
    }
  }
}

// Loss over the batch (ptychography_base.py:908-932) from the far fields in X (natural order):
// I = sum_m |X_m|^2 written fftshifted, the loss terms against the fftshifted targets and mask, dL/dI
// (unshifted) into gI; per-block partial sums in fixed order, kLossTiles tiles of 256 per block.
constexpr int kTiles = 8;
__global__ __launch_bounds__(256) void k_loss_gen(Params p, int ntiles) {
  __shared__ double red[4];
  const int NR = p.NR, NC = p.NC;
  const int N2 = NR * NC;
  const size_t total = (size_t)p.B * N2;
  const bool amp = (p.loss_type == QEM_LOSS_L2_AMPLITUDE || p.loss_type == QEM_LOSS_L1_AMPLITUDE);
  const float scale = (p.loss_type == QEM_LOSS_POISSON) ? 1.0f / p.mean_I : (float)p.P / ((float)p.B * p.mean_I);
  double acc = 0.0;
  for (int tile = blockIdx.x * kTiles; tile < ntiles && tile < (blockIdx.x + 1) * kTiles; ++tile) {
    const size_t idx = (size_t)tile * 256 + threadIdx.x;
    if (idx >= total) continue;
    const int b = (int)(idx / N2);
    const int k = (int)(idx - (size_t)b * N2);
    const int kr = k / NC, kc = k - kr * NC;
    float I = 0.0f;
    for (int m = 0; m < p.M; ++m) I += cabs2(p.X[((size_t)b * p.M + m) * N2 + k]);
    int sr = kr + NR / 2, sc = kc + NC / 2;
    sr -= sr >= NR ? NR : 0;
    sc -= sc >= NC ? NC : 0;
    const int ks = sr * NC + sc;
    if (p.intens_out != nullptr) p.intens_out[(size_t)b * N2 + ks] = I;
    const float t = p.targets[(size_t)b * N2 + ks];
    const float pred = amp ? sqrtf(I + 1e-9f) : I;
    float term, dpred;
    if (p.loss_type == QEM_LOSS_POISSON) {
      term = pred - t * logf(pred + 1e-6f);
      dpred = 1.0f - t / (pred + 1e-6f);
    } else {
      const float mk_ = p.mask[ks];
      const float d = pred * mk_ - t * mk_;
      if (p.loss_type == QEM_LOSS_L2_AMPLITUDE || p.loss_type == QEM_LOSS_L2_INTENSITY) {
        term = d * d;
        dpred = 2.0f * d * mk_;
      } else {
        term = fabsf(d);
        dpred = (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f)) * mk_;
      }
    }
    const float dI = amp ? dpred * 0.5f / pred : dpred;
    p.gI[(size_t)b * N2 + k] = scale * dI;
    acc += (double)term;
  }
  const double tot = block_sum<256>(acc, red);
  if (threadIdx.x == 0) p.loss_part[blockIdx.x] = tot;
}

// Adjoint of one (pattern, mode) over the slice segment [s0, s1), as k_adjoint256: the top segment
// starts from the far field, a lower one from the row-transformed g the segment above left in X; the
// bottom segment ends with G = fft2(g_psi_0): probe-gradient spectrum conj(ramp) G over X and the
// position-gradient partials.
template <bool DESCAN, bool PG>
__global__ __launch_bounds__(T) void k_adjoint_gen(Params p) {
  __shared__ Smem sm;
  const Geo g = make_geo(p);
  init_tables(sm, g.NR, g.NC);
  const int NR = g.NR, NC = g.NC;
  const size_t N2 = (size_t)NR * NC;
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<0>(p.pos, b, r0, c0, dr, dc);
  const int r0m = pmod(r0, p.H), c0m = pmod(c0, p.W);
  cf* Wb = p.scratch + (size_t)item * N2;
  cf* Xs = p.X + (size_t)item * N2;
  const float inv_n2 = (float)(1.0 / (double)N2);
  const size_t HW = (size_t)p.H * p.W;
  const bool top = p.s1 == p.S;

  if (top) {
    // g_x = F_ortho^H (2 dL/dI X): inverse column DFT here, inverse row DFT in the first row pass
    const float* gI = p.gI + (size_t)b * N2;
    const float sc = (float)(2.0 * (double)p.dloss[0] / sqrt((double)N2));
    pass<NONE, DFT_INV, true>(Xs, Wb, sm, g,
                              op2([&](int r, int c) { return Aux{mk(gI[(size_t)r * NC + c], 0.0f), cf{}}; },
                                  [&](int, int, cf v, const Aux& x) { return cscale(v, sc * x.a.x); }));
  }

  double dar = 0.0, dac = 0.0;
  if (DESCAN && top) {
    // descan (applied after the last slice in the forward), a pass of its own: its gradient
    // 2 pi sum f Im(conj(g) x_out) with x_out = ramp O_{S-1} psi_{S-1}, and g <- conj(ramp) g
    const cf* O = p.obj + (size_t)(p.S - 1) * HW;
    const cf* A = p.psi + ((size_t)item * p.wslots + (p.S - 1 - p.s0)) * N2;
    descan_tables(sm, g, p.descan, b);
    const float inv_nc = (float)(1.0 / (double)NC);
    pass<DFT_INV, DFT_FWD, false>(Wb, Wb, sm, g, op2([&](int r, int c) {
      return Aux{obj_patch(O, r0m, c0m, r, c, g, p.H, p.W), A[(size_t)r * NC + c]};
    }, [&](int r, int c, cf v, const Aux& x) {
      const cf d = cmul(sm.aux_r[r], sm.aux[c]);
      const cf xo = cmul(cmul(x.a, x.b), d);
      const float im = cconjmul(v, xo).y;
      dar += (double)(sm.fq_r[r] * im);
      dac += (double)(sm.fq_c[c] * im);
      // 1 / Nc: the forward row DFT is not followed by a column pass, so it leaves R g / Nc, the form
      // every row pass expects (R^H R = Nc for the unnormalised row DFT R)
      return cscale(cmulc(v, d), inv_nc);
    }));
  }
  for (int s = p.s1 - 1; s >= p.s0; --s) {
    const cf* O = p.obj + (size_t)s * HW;
    cf* A = p.psi + ((size_t)item * p.wslots + (s - p.s0)) * N2;
    // natural-order g_x of slice s: object-gradient patch conj(psi_s) g over psi_s, g_psi = conj(O_s) g
    pass<DFT_INV, DFT_FWD, false>((!top && s == p.s1 - 1) ? (const cf*)Xs : (const cf*)Wb, Wb, sm, g,
                                  op2(
                                      [&](int r, int c) {
                                        return Aux{obj_patch(O, r0m, c0m, r, c, g, p.H, p.W),
                                                   ld_stream(A + (size_t)r * NC + c)};
                                      },
                                      [&](int r, int c, cf v, const Aux& x) {
                                        st_stream(A + (size_t)r * NC + c, cconjmul(x.b, v));
                                        return cconjmul(x.a, v);
                                      }));
    if (s > 0) {
      // g_x <- ifft2(fft2(g_psi) conj(prop_{s-1})); at a segment bottom it goes to X for the segment below
      const cf* pr = p.prop + (size_t)(s - 1) * N2;
      cf* xs = PG ? p.pspec + ((size_t)item * (p.S - 1) + (s - 1)) * N2 : nullptr;
      pass<DFT_FWD, DFT_INV, true>(Wb, s == p.s0 ? Xs : Wb, sm, g, op2([&](int r, int c) {
        const size_t k = (size_t)r * NC + c;
        Aux x;
        x.a = pr[k];
        if constexpr (PG) x.b = ld_stream(xs + k);
        return x;
      }, [&](int r, int c, cf v, const Aux& x) {
        // this item's dL/dP_{s-1}: conj(fft2(x_{s-1})) fft2(g_psi_s) / (Nr Nc)
        if constexpr (PG) st_stream(xs + (size_t)r * NC + c, cscale(cconjmul(x.b, v), inv_n2));
        return cscale(cmulc(v, x.a), inv_n2);
      }));
    }
  }
  if constexpr (DESCAN) {
    if (top && (p.flags & QEM_FLAG_GRAD_DESCAN)) {
      const double tr = block_sum<T>(dar, sm.red);
      const double tc = block_sum<T>(dac, sm.red);
      if (threadIdx.x == 0) {
        p.gdesc_part[2 * item] = (float)(2.0 * M_PI * tr);
        p.gdesc_part[2 * item + 1] = (float)(2.0 * M_PI * tc);
      }
    }
  }
  if (p.s0 > 0) return;
  // G = fft2(g_psi_0); probe-gradient spectrum conj(ramp_b) G over X, position partials
  // (2 pi / (Nr Nc)) sum f Im(conj(G) P_m ramp_b)
  const cf* pk = p.pk + (size_t)m * N2;
  const cf* rr = p.ramps + (size_t)b * (NR + NC);
  double ar = 0.0, ac = 0.0;
  pass<DFT_FWD, NONE, true>(Wb, Xs, sm, g, op2([&](int r, int c) {
    return Aux{pk[(size_t)r * NC + c], cmul(rr[r], rr[NR + c])};
  }, [&](int r, int c, cf v, const Aux& x) {
    const float im = cconjmul(v, cmul(x.a, x.b)).y;
    ar += (double)(sm.fq_r[r] * im);
    ac += (double)(sm.fq_c[c] * im);
    return cmulc(v, x.b);
  }));
  if (p.flags & QEM_FLAG_GRAD_POSITIONS) {
    const double tr = block_sum<T>(ar, sm.red);
    const double tc = block_sum<T>(ac, sm.red);
    if (threadIdx.x == 0) {
      p.gpos_part[2 * item] = (float)(2.0 * M_PI * tr / (double)N2);
      p.gpos_part[2 * item + 1] = (float)(2.0 * M_PI * tc / (double)N2);
    }
  }
}

// Deterministic object-gradient scatter-add over 32 x 32 tiles of one slice: each block sums, in
// increasing (b, m) order, the per-mode patch gradients (natural order in psi) that land on its tile.
// Object row y lies in the patch of origin r0 iff u = (y - r0 + Nr // 2) mod H < Nr; its patch row is
// (u - Nr // 2) mod Nr (fftfreq order); likewise the columns.
__global__ __launch_bounds__(256) void k_obj_reduce_gen(Params p) {
  __shared__ int su0[256], sv0[256];
  const int NR = p.NR, NC = p.NC, hr = NR / 2, hc = NC / 2;
  const size_t N2 = (size_t)NR * NC;
  const int s = p.s0 + blockIdx.z;
  const int X0 = blockIdx.x * 32, Y0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int x = X0 + tx;
  cf acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = mk(0.0f, 0.0f);
  const size_t slab = (size_t)p.wslots * N2;  // stride between modes of one pattern
  for (int b0 = 0; b0 < p.B; b0 += 256) {
    const int nb = min(256, p.B - b0);
    if ((int)threadIdx.x < nb) {
      const int b = b0 + threadIdx.x;
      const int r0 = (int)rintf(p.pos[2 * b]), c0 = (int)rintf(p.pos[2 * b + 1]);
      const int u0 = pmod(Y0 - r0 + hr, p.H), v0 = pmod(X0 - c0 + hc, p.W);
      const bool hit = (u0 < NR || u0 + 31 >= p.H) && (v0 < NC || v0 + 31 >= p.W);
      su0[threadIdx.x] = hit ? u0 : -1;
      sv0[threadIdx.x] = v0;
    }
    __syncthreads();
    for (int i = 0; i < nb; ++i) {
      const int u0 = su0[i];
      if (u0 < 0) continue;
      int ux = sv0[i] + tx;
      if (ux >= p.W) ux -= p.W;
      const bool xok = x < p.W && ux < NC;
      int col = ux - hc;
      col += col < 0 ? NC : 0;
      const cf* gp = p.psi + ((size_t)(b0 + i) * p.M * p.wslots + blockIdx.z) * N2 + (xok ? col : 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int dy = ty + 8 * j;
        int uy = u0 + dy;
        if (uy >= p.H) uy -= p.H;
        if (!(xok && (Y0 + dy < p.H) && uy < NR)) continue;
        int row = uy - hr;
        row += row < 0 ? NR : 0;
        const cf* rp = gp + (size_t)row * NC;
        for (int mm = 0; mm < p.M; ++mm) {
          const cf t = ld_stream(rp + (size_t)mm * slab);
          acc[j].x += t.x;
          acc[j].y += t.y;
        }
      }
    }
    __syncthreads();
  }
  if (x >= p.W) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int y = Y0 + ty + 8 * j;
    if (y < p.H) p.grad_obj[(size_t)s * p.H * p.W + (size_t)y * p.W + x] = acc[j];
  }
}

// Deterministic object-gradient scatter-add, one object row per block (used when the row fits the
// LDS), one thread per patch column (blockDim.x = Nc rounded up to whole waves): the block lists, in
// batch order, the patterns whose patch covers row y (ballot compaction); for each listed pattern
// every thread sums the modes of its column of the covering patch row in order and adds the sum into
// the W-wide LDS row.  One pattern's columns land on distinct object columns (W >= Nc), consecutive
// patterns' may coincide, so a barrier separates the adds of consecutive patterns: every object pixel
// sums its patterns in increasing b.  K patterns' loads (MC modes each) are issued together.
template <int MC>
__global__ __launch_bounds__(512) void k_obj_reduce_rows_gen(Params p) {
  extern __shared__ cf acc_row[];  // p.W entries
  __shared__ int lb[512], lrow[512], lc0[512];
  __shared__ int wcount[8];
  constexpr int K = 4;
  const int t = threadIdx.x, TB = blockDim.x, nw = TB >> 6;
  const int NR = p.NR, NC = p.NC, hr = NR / 2;
  const size_t N2 = (size_t)NR * NC;
  const int y = blockIdx.x, slot = blockIdx.y, s = p.s0 + slot;
  const size_t slab = (size_t)p.wslots * N2;  // stride between modes of one pattern
  const bool col_ok = t < NC;
  const int ic = col_ok ? t : NC - 1;  // lanes past Nc read the row's last element and add nothing
  const int xo = foff(ic, NC);         // the column's offset from the patch origin
  for (int x = t; x < p.W; x += TB) acc_row[x] = mk(0.0f, 0.0f);
  for (int b0 = 0; b0 < p.B; b0 += TB) {
    const int b = b0 + t;
    int u = NR, c0m = 0;
    if (b < p.B) {
      const int r0 = (int)rintf(p.pos[2 * b]), c0 = (int)rintf(p.pos[2 * b + 1]);
      u = pmod(y - r0 + hr, p.H);
      c0m = pmod(c0, p.W);
    }
    const bool hit = u < NR;
    const unsigned long long bal = __ballot(hit);
    const int lane = t & 63, wave = t >> 6;
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();  // also orders the zeroing / the previous chunk's reads of the list
    int off = 0, cnt = 0;
    for (int w = 0; w < nw; ++w) {
      off += w < wave ? wcount[w] : 0;
      cnt += wcount[w];
    }
    if (hit) {
      int row = u - hr;
      row += row < 0 ? NR : 0;
      lb[off + pre] = b;
      lrow[off + pre] = row;
      lc0[off + pre] = c0m;
    }
    __syncthreads();
    for (int i0 = 0; i0 < cnt; i0 += K) {
      cf val[K];
      int xk[K];
      size_t base[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int i = i0 + k < cnt ? i0 + k : i0;  // past the list's end: loaded, unused
        base[k] = ((size_t)lb[i] * p.M * p.wslots + slot) * N2 + (size_t)lrow[i] * NC + ic;
        int x = lc0[i] + xo;  // one wrap each way
        x += x < 0 ? p.W : 0;
        x -= x >= p.W ? p.W : 0;
        xk[k] = x;
        val[k] = mk(0.0f, 0.0f);
      }
      // the modes MC at a time: every load of a group is issued before its first add (mode indices
      // past M - 1 re-read mode M - 1 and are not added), and the adds run in mode order
      for (int m0 = 0; m0 < p.M; m0 += MC) {
        cf v[K][MC];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int j = 0; j < MC; ++j) {
            const int m = m0 + j < p.M ? m0 + j : p.M - 1;
            v[k][j] = ld_stream(p.psi + base[k] + (size_t)m * slab);
          }
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int j = 0; j < MC; ++j) {
            if (m0 + j < p.M) {
              val[k].x += v[k][j].x;
              val[k].y += v[k][j].y;
            }
          }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (i0 + k < cnt) {  // uniform over the block
          if (col_ok) {
            acc_row[xk[k]].x += val[k].x;
            acc_row[xk[k]].y += val[k].y;
          }
          __syncthreads();
        }
      }
    }
    __syncthreads();  // the list is rewritten by the next chunk
  }
  cf* out = p.grad_obj + ((size_t)s * p.H + y) * p.W;
  for (int x = t; x < p.W; x += TB) out[x] = acc_row[x];
}

// gsum[m][k] = sum_b X[b][m][k] in a fixed order (4 lanes per frequency, every 4th pattern each,
// combined in lane order).
__global__ __launch_bounds__(256) void k_probe_reduce_gen(Params p) {
  __shared__ cf part[4][64];
  const size_t N2 = (size_t)p.NR * p.NC;
  const int kk = threadIdx.x & 63, gq = threadIdx.x >> 6;
  const size_t idx = (size_t)blockIdx.x * 64 + kk;  // over M * N2
  const bool ok = idx < (size_t)p.M * N2;
  cf acc = mk(0.0f, 0.0f);
  if (ok) {
    const size_t m = idx / N2, k = idx - m * N2;
    const cf* src = p.X + m * N2 + k;
    const size_t stride = (size_t)p.M * N2;
    for (int b = gq; b < p.B; b += 4) {
      const cf t = src[(size_t)b * stride];
      acc.x += t.x;
      acc.y += t.y;
    }
  }
  part[gq][kk] = acc;
  __syncthreads();
  if (gq == 0 && ok) {
    cf t = part[0][kk];
    for (int i = 1; i < 4; ++i) {
      t.x += part[i][kk].x;
      t.y += part[i][kk].y;
    }
    p.gsum[idx] = t;
  }
}

// grad_probe[m] = ifft2(gsum[m]) (1 / (Nr Nc) included).
__global__ __launch_bounds__(T) void k_probe_grad_gen(Params p) {
  __shared__ Smem sm;
  const Geo g = make_geo(p);
  init_tables(sm, g.NR, g.NC);
  const size_t N2 = (size_t)g.NR * g.NC;
  const int m = blockIdx.x;
  cf* Wb = p.scratch + (size_t)m * N2;
  const float inv_n2 = (float)(1.0 / (double)N2);
  pass<DFT_INV, NONE, true>(p.gsum + (size_t)m * N2, Wb, sm, g, no_op());
  pass<DFT_INV, NONE, false>(Wb, p.grad_probe + (size_t)m * N2, sm, g,
                             op2(NoFetch(), [&](int, int, cf v, const Aux&) { return cscale(v, inv_n2); }));
}

// grad_prop[s][k] = sum over the items of their contributions in p.pspec, fixed order.
__global__ __launch_bounds__(256) void k_prop_reduce_gen(Params p, cf* grad_prop) {
  __shared__ cf part[4][64];
  const size_t N2 = (size_t)p.NR * p.NC;
  const int kk = threadIdx.x & 63, gq = threadIdx.x >> 6;
  const size_t idx = (size_t)blockIdx.x * 64 + kk;  // over (S-1) * N2
  const size_t slot = (size_t)(p.S - 1) * N2;
  const bool ok = idx < slot;
  const int items = p.B * p.M;
  cf acc = mk(0.0f, 0.0f);
  if (ok) {
    for (int it = gq; it < items; it += 4) {
      const cf t = ld_stream(p.pspec + (size_t)it * slot + idx);
      acc.x += t.x;
      acc.y += t.y;
    }
  }
  part[gq][kk] = acc;
  __syncthreads();
  if (gq == 0 && ok) {
    cf t = part[0][kk];
    for (int i = 1; i < 4; ++i) {
      t.x += part[i][kk].x;
      t.y += part[i][kk].y;
    }
    grad_prop[idx] = t;
  }
}

}  // namespace gen

unsigned long long params_signature_gen() { return params_signature_local(); }

// Whether the generic kernels take an Nr x Nc pattern: 2 <= Nr, Nc <= 512 with prime factors 2, 3, 5, 7.
bool gen_supported(int nr, int nc) {
  auto smooth = [](int n) {
    if (n < 2 || n > gen::NMAX) return false;
    for (int r : {2, 3, 5, 7})
      while (n % r == 0) n /= r;
    return n == 1;
  };
  return smooth(nr) && smooth(nc);
}

int launch_gen_forward_kernel(const Params& p, hipStream_t st) {
  const dim3 grid(p.B * p.M), blk(gen::T);
  if (p.pspec != nullptr && p.descan != nullptr)
    hipLaunchKernelGGL((gen::k_forward_gen<true, true>), grid, blk, 0, st, p);
  else if (p.pspec != nullptr)
    hipLaunchKernelGGL((gen::k_forward_gen<false, true>), grid, blk, 0, st, p);
  else if (p.descan != nullptr && p.fwd_mode != FWD_RECOMPUTE)
    hipLaunchKernelGGL((gen::k_forward_gen<true, false>), grid, blk, 0, st, p);
  else
    hipLaunchKernelGGL((gen::k_forward_gen<false, false>), grid, blk, 0, st, p);
  return check_last_launch("forward (generic size)");
}

int launch_gen_forward(const Params& p, hipStream_t st) {
  const int ramp_blocks = (p.B * (p.NR + p.NC) + gen::T - 1) / gen::T;
  hipLaunchKernelGGL(gen::k_spectrum_gen, dim3(p.M + ramp_blocks), dim3(gen::T), 0, st, p);
  if (int rc = check_last_launch("probe spectrum (generic size)")) return rc;
  return launch_gen_forward_kernel(p, st);
}

int launch_gen_loss(const Params& p, int ntiles, hipStream_t st) {
  const int nparts = (ntiles + gen::kTiles - 1) / gen::kTiles;
  hipLaunchKernelGGL(gen::k_loss_gen, dim3(nparts), dim3(256), 0, st, p, ntiles);
  return check_last_launch("loss (generic size)");
}

int gen_loss_parts(int ntiles) { return (ntiles + gen::kTiles - 1) / gen::kTiles; }

int launch_gen_adjoint(const Params& p, hipStream_t st) {
  const dim3 grid(p.B * p.M), blk(gen::T);
  if (p.pspec != nullptr && p.descan != nullptr)
    hipLaunchKernelGGL((gen::k_adjoint_gen<true, true>), grid, blk, 0, st, p);
  else if (p.pspec != nullptr)
    hipLaunchKernelGGL((gen::k_adjoint_gen<false, true>), grid, blk, 0, st, p);
  else if (p.descan != nullptr)
    hipLaunchKernelGGL((gen::k_adjoint_gen<true, false>), grid, blk, 0, st, p);
  else
    hipLaunchKernelGGL((gen::k_adjoint_gen<false, false>), grid, blk, 0, st, p);
  return check_last_launch("adjoint (generic size)");
}

int launch_gen_obj_reduce(const Params& q, hipStream_t st) {
  const size_t row_bytes = (size_t)q.W * sizeof(cf);
  if (row_bytes <= 48 * 1024) {
    const dim3 g(q.H, q.s1 - q.s0);
    const dim3 blk((unsigned)((q.NC + 63) / 64 * 64));  // one thread per patch column, whole waves (<= 512)
    if (q.M >= 4)
      hipLaunchKernelGGL(gen::k_obj_reduce_rows_gen<4>, g, blk, row_bytes, st, q);
    else if (q.M >= 2)
      hipLaunchKernelGGL(gen::k_obj_reduce_rows_gen<2>, g, blk, row_bytes, st, q);
    else
      hipLaunchKernelGGL(gen::k_obj_reduce_rows_gen<1>, g, blk, row_bytes, st, q);
  } else {
    const dim3 g((q.W + 31) / 32, (q.H + 31) / 32, q.s1 - q.s0);
    hipLaunchKernelGGL(gen::k_obj_reduce_gen, g, dim3(256), 0, st, q);
  }
  return check_last_launch("object reduce (generic size)");
}

int launch_gen_probe_grad(const Params& p, hipStream_t st) {
  const size_t n = (size_t)p.M * p.NR * p.NC;
  hipLaunchKernelGGL(gen::k_probe_reduce_gen, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, p);
  if (int rc = check_last_launch("probe reduce (generic size)")) return rc;
  hipLaunchKernelGGL(gen::k_probe_grad_gen, dim3(p.M), dim3(gen::T), 0, st, p);
  return check_last_launch("probe gradient (generic size)");
}

int launch_gen_prop_reduce(const Params& p, cf* grad_prop, hipStream_t st) {
  const size_t n = (size_t)(p.S - 1) * p.NR * p.NC;
  hipLaunchKernelGGL(gen::k_prop_reduce_gen, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, p, grad_prop);
  return check_last_launch("propagator gradient (generic size)");
}

}  // namespace qem
