// fft2.h — register/LDS 2-D complex64 FFT for one N x N array per workgroup (gfx950).
//
// An N x N array (N = 32, 64, 128) lives in the registers of one workgroup: T = N*N/PTS threads,
// PTS complex values each.  A 2-D transform is three register passes separated by two LDS
// exchanges (one N*N*8-byte buffer, 128 KiB at N = 128, which fits the 160 KiB LDS of a gfx950 CU):
//
//   columns (index c):  N = RC1 * RC2      c = RC2*c1 + c0          kc = kc0 + RC1*kc1
//   rows    (index r):  N = RR1 * RR2 * RR3 r = NRP*r1 + RR3*r2 + r3 kr = k1 + RR1*(k2 + RR2*k3)
//
//   pass 1 (layout P1: thread = (rp, c0), regs (r1, c1)):  DFT_RR1 over r1, twiddle W_N^(rp*k1),
//                                                          DFT_RC1 over c1, twiddle W_N^(c0*kc0)
//   pass 2 (layout P2: thread = (k1, r3, kc0), regs (r2, c0)): DFT_RC2 over c0, DFT_RR2 over r2,
//                                                          twiddle W_N^(RR1*r3*k2)
//   pass 3 (layout P3: thread = (kc, q), regs (h, r3)):     DFT_RR3 over r3
//
// P1 is natural-order real space (the layout of every elementwise multislice step); P3 is the
// frequency domain with one frequency column kc per lane (coalesced stores).  The inverse transform
// (unnormalised, i.e. N^2 * ifft2) runs the adjoint passes in reverse order and goes back from P3
// to P1, so forward -> pointwise -> inverse never reorders data outside the two exchanges.
//
// Synchronisation contract: fwd() ends with an LDS read in P3 that inv() starts by overwriting from
// the same thread at the same addresses (and inv() ends / fwd() starts likewise in P1), so
// fwd/inv alternation needs no extra barrier.  Two fwd() (or two inv()) in a row need a
// __syncthreads() between them.
//
// Before each exchange, every register group's LDS writes are issued right after that group's last
// stage (a scheduling barrier keeps them there), so the LDS drains them while the next group is
// computed instead of after the whole pass.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

#include "twiddles.h"

namespace qem {

typedef float2 cf;

__device__ __forceinline__ cf mk(float x, float y) {
  cf r;
  r.x = x;
  r.y = y;
  return r;
}
__device__ __forceinline__ cf cadd(cf a, cf b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cf csub(cf a, cf b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cf cmul(cf a, cf b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
__device__ __forceinline__ cf cmulc(cf a, cf b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
// conj(a) * b
__device__ __forceinline__ cf cconjmul(cf a, cf b) { return mk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ cf cscale(cf a, float s) { return mk(a.x * s, a.y * s); }
__device__ __forceinline__ float cabs2(cf a) { return a.x * a.x + a.y * a.y; }

// Workgroup barrier that orders LDS only.  __syncthreads() is a workgroup-scope fence on all memory,
// which makes every wave wait for its outstanding global stores before the barrier; the FFT exchanges
// only hand data through LDS, so stores of stored waves / gradient patches issued before a transform
// can drain while it runs.  Code that hands data between threads through global memory must keep
// __syncthreads().
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// cos / sin of 2*pi*k/16.
constexpr double kCos16(int k) {
  constexpr double t[16] = {1.0,
                            0.92387953251128674,
                            0.70710678118654752,
                            0.38268343236508977,
                            0.0,
                            -0.38268343236508977,
                            -0.70710678118654752,
                            -0.92387953251128674,
                            -1.0,
                            -0.92387953251128674,
                            -0.70710678118654752,
                            -0.38268343236508977,
                            0.0,
                            0.38268343236508977,
                            0.70710678118654752,
                            0.92387953251128674};
  return t[k];
}
constexpr double kSin16(int k) { return kCos16((k + 12) & 15); }

// a * W_R^E with W_R = exp(-2 pi i / R) (forward) or its conjugate (INV).  R divides 16.
template <int R, int E, bool INV>
__device__ __forceinline__ cf twc(cf a) {
  constexpr int k = (((E % R) + R) % R) * (16 / R);
  if constexpr (k == 0) {
    return a;
  } else if constexpr (k == 8) {
    return mk(-a.x, -a.y);
  } else if constexpr (k == 4) {  // forward: -i
    return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
  } else if constexpr (k == 12) {  // forward: +i
    return INV ? mk(a.y, -a.x) : mk(-a.y, a.x);
  } else {
    constexpr float c = (float)kCos16(k);
    constexpr float s = INV ? (float)kSin16(k) : -(float)kSin16(k);
    return mk(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// In-register DFT of size R on v[0], v[S], ..., v[(R-1)S]; natural order in and out.
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<1, INV> {
  template <int S>
  __device__ __forceinline__ static void run(cf*) {}
};

template <bool INV>
struct Dft<2, INV> {
  template <int S>
  __device__ __forceinline__ static void run(cf* v) {
    cf a = v[0], b = v[S];
    v[0] = cadd(a, b);
    v[S] = csub(a, b);
  }
};

template <bool INV>
struct Dft<4, INV> {
  template <int S>
  __device__ __forceinline__ static void run(cf* v) {
    cf x0 = v[0], x1 = v[S], x2 = v[2 * S], x3 = v[3 * S];
    cf t0 = cadd(x0, x2), t1 = csub(x0, x2), t2 = cadd(x1, x3), t3 = csub(x1, x3);
    cf jt3 = INV ? mk(-t3.y, t3.x) : mk(t3.y, -t3.x);  // (+i or -i) * t3
    v[0] = cadd(t0, t2);
    v[2 * S] = csub(t0, t2);
    v[S] = cadd(t1, jt3);
    v[3 * S] = csub(t1, jt3);
  }
};

// Cooley-Tukey R = R1*R2: n = R2*n1 + n2, k = k1 + R1*k2.
template <int R1, int R2, bool INV, int S>
__device__ __forceinline__ void dft_ct(cf* v) {
  constexpr int R = R1 * R2;
  cf t[R];
  sfor<0, R2>([&](auto n2c) {
    constexpr int n2 = decltype(n2c)::value;
    cf a[R1];
    sfor<0, R1>([&](auto n1c) {
      constexpr int n1 = decltype(n1c)::value;
      a[n1] = v[(R2 * n1 + n2) * S];
    });
    Dft<R1, INV>::template run<1>(a);
    sfor<0, R1>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      t[k1 * R2 + n2] = twc<R, n2 * k1, INV>(a[k1]);
    });
  });
  sfor<0, R1>([&](auto k1c) {
    constexpr int k1 = decltype(k1c)::value;
    Dft<R2, INV>::template run<1>(&t[k1 * R2]);
    sfor<0, R2>([&](auto k2c) {
      constexpr int k2 = decltype(k2c)::value;
      v[(k1 + R1 * k2) * S] = t[k1 * R2 + k2];
    });
  });
}

template <bool INV>
struct Dft<8, INV> {
  template <int S>
  __device__ __forceinline__ static void run(cf* v) {
    dft_ct<2, 4, INV, S>(v);
  }
};

template <bool INV>
struct Dft<16, INV> {
  template <int S>
  __device__ __forceinline__ static void run(cf* v) {
    dft_ct<4, 4, INV, S>(v);
  }
};

template <int N>
struct Plan;
// Radices, plus padded LDS strides of the two exchanges (in complex64 units):
//   exchange 1 address = c0 + A1*kc0 + A2*rp + A2*NRP*k1
//   exchange 2 address = kc + B3*r3 + B3*RR3*k2 + B3*RR3*RR2*k1
// The strides were chosen by enumerating 64-bit LDS access patterns of both layouts touching each
// exchange and minimising bank collisions (16 lanes / 32 banks and 32 lanes / 64 banks); they
// are affine in the register index, so every LDS access is a per-thread base plus an immediate.
template <>
struct Plan<32> {
  static constexpr int RC1 = 8, RC2 = 4, RR1 = 2, RR2 = 4, RR3 = 4;
  static constexpr int A1 = 4, A2 = 34, B3 = 40;
};
template <>
struct Plan<64> {
  static constexpr int RC1 = 8, RC2 = 8, RR1 = 2, RR2 = 2, RR3 = 16;
  static constexpr int A1 = 9, A2 = 72, B3 = 72;
};
template <>
struct Plan<128> {
  static constexpr int RC1 = 16, RC2 = 8, RR1 = 2, RR2 = 4, RR3 = 16;
  static constexpr int A1 = 9, A2 = 152, B3 = 144;
};

// W_N^j from the W_256 table.
template <int N>
__device__ __forceinline__ cf twN(int j) {
  return kTw256[(j * (256 / N)) & 255];
}



template <int N>
struct Fft2 {
  static constexpr int RC1 = Plan<N>::RC1, RC2 = Plan<N>::RC2;
  static constexpr int RR1 = Plan<N>::RR1, RR2 = Plan<N>::RR2, RR3 = Plan<N>::RR3;
  static constexpr int NRP = N / RR1;
  static constexpr int PTS = RC1 * RR1;
  static constexpr int T = N * N / PTS;
  static constexpr int H3 = PTS / RR3;
  static_assert(RC1 * RC2 == N, "column radices");
  static_assert(RR1 * RR2 * RR3 == N, "row radices");
  static_assert(RC2 * RR2 == PTS, "pass-2 points");
  static_assert(PTS % RR3 == 0, "pass-3 points");
  static_assert(RR1 == 2, "first row stage is radix 2");
  static constexpr int A1 = Plan<N>::A1, A2 = Plan<N>::A2, A3 = Plan<N>::A2 * NRP;
  static constexpr int B3 = Plan<N>::B3, B2 = Plan<N>::B3 * RR3, B1 = Plan<N>::B3 * RR3 * RR2;
  static constexpr int LDS1 = (RC2 - 1) + A1 * (RC1 - 1) + A2 * (NRP - 1) + A3 * (RR1 - 1) + 1;
  static constexpr int LDS2 = (N - 1) + B3 * (RR3 - 1) + B2 * (RR2 - 1) + B1 * (RR1 - 1) + 1;
  static constexpr int LDS = LDS1 > LDS2 ? LDS1 : LDS2;  // complex64 slots of the exchange buffer
  static_assert(A1 >= RC2 && A2 >= RC2 + A1 * (RC1 - 1) && B3 >= N, "strides must not overlap");
  // Twiddle tables kept in LDS after the exchange buffer (all accesses are base + immediate):
  //   [c0][kc0] = W_N^(c0*kc0), rows padded to RC1+1 entries (bank-conflict free across c0)
  //   [r3][k2]  = W_N^(RR1*r3*k2)
  static constexpr int TWC = 0, TWC_ROW = RC1 + 1;
  static constexpr int TWR = TWC + RC2 * TWC_ROW;
  static constexpr int TWL = TWR + RR3 * RR2;
  static constexpr int LDS_TOTAL = LDS + TWL;  // complex64 slots a kernel reserves: exchange + twiddles

  // P1 role
  int rp, c0;
  // P2 role
  int k1b, r3b, kc0b;
  // P3 role
  int kc, q;
  int p1base;          // rp*N + c0
  const cf* twl;       // LDS twiddle tables
  int twc_base, twr_base;
  cf tw_rp;            // W_N^rp
  int e1p1, e1p2;      // exchange-1 base addresses of this thread in layouts P1 and P2
  int e2p2;            // exchange-2 base address in layout P2
  int e2p3[H3];        // exchange-2 base addresses in layout P3 (per register group)
  int p3k12[H3];       // k1 + RR1*k2 of each pass-3 register group
  int p3base[H3];      // p3k12*N + kc

  __device__ __forceinline__ explicit Fft2(int tid) {
    c0 = tid % RC2;
    rp = tid / RC2;
    kc0b = tid % RC1;
    r3b = (tid / RC1) % RR3;
    k1b = tid / (RC1 * RR3);
    kc = tid % N;
    q = tid / N;
    p1base = rp * N + c0;
    sfor<0, H3>([&](auto h) {
      const int j = q * H3 + h.value;
      p3k12[h.value] = (j % RR1) + RR1 * (j / RR1);
      p3base[h.value] = p3k12[h.value] * N + kc;
      e2p3[h.value] = kc + B2 * (j / RR1) + B1 * (j % RR1);
    });
    tw_rp = twN<N>(rp);
    twc_base = TWC + c0 * TWC_ROW;
    twr_base = TWR + r3b * RR2;
    e1p1 = c0 + A2 * rp;
    e1p2 = A1 * kc0b + A2 * r3b + A3 * k1b;
    e2p2 = kc0b + B3 * r3b + B1 * k1b;
  }

  // Fills the LDS twiddle tables at lds + LDS; the caller must __syncthreads() before fwd()/inv().
  __device__ __forceinline__ void init_twiddles(cf* lds, int tid) {
    cf* t = lds + LDS;
    for (int i = tid; i < TWL; i += T) {
      int j;
      if (i < TWR) {
        const int row = i / TWC_ROW, col = i - row * TWC_ROW;
        j = row * (col < RC1 ? col : 0);
      } else {
        const int r = (i - TWR) / RR2, k2 = (i - TWR) % RR2;
        j = RR1 * r * k2;
      }
      t[i] = twN<N>(j);
    }
    twl = t;
  }

  // Natural coordinates of register e in layout P1 (real space).
  __device__ __forceinline__ int p1_row(int e) const { return (e / RC1) * NRP + rp; }
  __device__ __forceinline__ int p1_col(int e) const { return (e % RC1) * RC2 + c0; }
  // Offset row*N + col of register e in an N x N row-major array (P1): thread base + constant.
  __device__ __forceinline__ int p1_off(int e) const { return p1base + (e / RC1) * NRP * N + (e % RC1) * RC2; }
  // Natural frequency row of register e in layout P3 (frequency column is kc).
  __device__ __forceinline__ int p3_row(int e) const { return p3k12[e / RR3] + RR1 * RR2 * (e % RR3); }
  // Offset of register e in the stored-wave layout (per-slice waves, gradient patches, checkpoints):
  // natural rows, each row's columns permuted so that registers 2k and 2k+1 of the RC2 threads that
  // share a row sit side by side, thread after thread (column c = r1 * RC2 + c0 at position
  // (r1 / 2) * 2 * RC2 + 2 * c0 + r1 % 2).  A 16-byte access of a register pair then covers whole
  // 128-byte lines across the lanes (RC2 = 8 at N = 128) instead of 64-byte pieces of natural rows.
  __device__ __forceinline__ int wv_off(int e) const {
    return rp * N + 2 * c0 + (e / RC1) * NRP * N + ((e % RC1) / 2) * 2 * RC2 + (e % 2);
  }
  // Pointer to register e of a stored-wave slot (wv_off) at a workgroup-uniform base: the base, a
  // 32-bit per-thread row offset (one register per row group e / RC1) and a constant below 1 KiB, so
  // that an access is a scalar-base load / store with an immediate offset and needs no 64-bit address
  // registers.
  template <typename T>
  __device__ __forceinline__ T* wv_ptr(T* base, int e) const {
    const unsigned row = (unsigned)(rp * N + 2 * c0 + (e / RC1) * NRP * N) * 8u;
    const unsigned col = (unsigned)(((e % RC1) / 2) * 2 * RC2 + (e % 2)) * 8u;
    return (T*)((char*)base + (size_t)row + (size_t)col);
  }
  // Pointer to register e of an N x N row-major array (P3) at a workgroup-uniform base, formed like
  // wv_ptr: the base plus a 32-bit per-thread offset.
  template <typename T>
  __device__ __forceinline__ T* p3_ptr(T* base, int e) const {
    return (T*)((char*)base + (size_t)((unsigned)p3base[e / RR3] * 8u) + (size_t)(RR1 * RR2 * N * (e % RR3) * 8));
  }
  // Offset kr*N + kc of register e in an N x N row-major array (P3).
  __device__ __forceinline__ int p3_off(int e) const { return p3base[e / RR3] + RR1 * RR2 * N * (e % RR3); }

  // Forward 2-D DFT (unnormalised), P1 -> P3.
  __device__ __forceinline__ void fwd(cf (&v)[PTS], cf* lds) const {
    fwd_groups(v, lds, [](auto) {}, [](auto) {});
  }

  // fwd() with hooks around each pass-3 register group h (registers h*RR3 .. h*RR3 + RR3 - 1, one
  // frequency column, rows p3_row): pre(h) runs before the group's transform and post(h) after it, so
  // that a pointwise step can issue its global loads for a group while that group is transformed.
  template <typename Pre, typename Post>
  __device__ __forceinline__ void fwd_groups(cf (&v)[PTS], cf* lds, Pre&& pre, Post&& post) const {
    // pass 1
    sfor<0, RC1>([&](auto c1) { Dft<RR1, false>::template run<RC1>(&v[c1.value]); });
    sfor<0, RC1>([&](auto c1) { v[RC1 + c1.value] = cmul(v[RC1 + c1.value], tw_rp); });
    sfor<0, RR1>([&](auto k1) {
      Dft<RC1, false>::template run<1>(&v[k1.value * RC1]);
      sfor<1, RC1>([&](auto kc0) { v[k1.value * RC1 + kc0.value] = cmul(v[k1.value * RC1 + kc0.value], twl[twc_base + kc0.value]); });
      sfor<0, RC1>([&](auto kc0) { lds[e1p1 + A3 * k1.value + A1 * kc0.value] = v[k1.value * RC1 + kc0.value]; });
      __builtin_amdgcn_sched_barrier(0);  // keep the group's exchange writes here
    });
    lds_barrier();
    sfor<0, RR2>([&](auto r2) {
      sfor<0, RC2>([&](auto c0i) {
        v[r2.value * RC2 + c0i.value] = lds[e1p2 + A2 * RR3 * r2.value + c0i.value];
      });
    });
    // pass 2
    sfor<0, RR2>([&](auto r2) { Dft<RC2, false>::template run<1>(&v[r2.value * RC2]); });
    lds_barrier();  // every thread's exchange-1 reads are done before any exchange-2 write
    // two column groups at a time: their writes at the same k2 are RC1 slots apart and pair up
    sfor<0, RC2 / 2>([&](auto g) {
      sfor<2 * g.value, 2 * g.value + 2>([&](auto kc1) {
        Dft<RR2, false>::template run<RC2>(&v[kc1.value]);
        sfor<1, RR2>([&](auto k2) { v[k2.value * RC2 + kc1.value] = cmul(v[k2.value * RC2 + kc1.value], twl[twr_base + k2.value]); });
      });
      sfor<0, RR2>([&](auto k2) {
        sfor<2 * g.value, 2 * g.value + 2>([&](auto kc1) { lds[e2p2 + B2 * k2.value + RC1 * kc1.value] = v[k2.value * RC2 + kc1.value]; });
      });
      __builtin_amdgcn_sched_barrier(0);  // keep the groups' exchange writes here
    });
    lds_barrier();
    sfor<0, H3>([&](auto h) {
      sfor<0, RR3>([&](auto r3) { v[h.value * RR3 + r3.value] = lds[e2p3[h.value] + B3 * r3.value]; });
    });
    // pass 3
    sfor<0, H3>([&](auto h) {
      pre(h);
      Dft<RR3, false>::template run<1>(&v[h.value * RR3]);
      post(h);
    });
  }

  // Inverse 2-D DFT without the 1/N^2 factor (= N^2 * ifft2), P3 -> P1.
  __device__ __forceinline__ void inv(cf (&v)[PTS], cf* lds) const {
    inv_hooked(v, lds, [] {});
  }

  // inv() with a hook run once the last exchange has been read, before the final register pass, so
  // that a caller can issue global loads (for the elementwise step after the transform) whose latency
  // that pass then hides.
  template <typename Pre1>
  __device__ __forceinline__ void inv_hooked(cf (&v)[PTS], cf* lds, Pre1&& pre1) const {
    sfor<0, H3>([&](auto h) {
      Dft<RR3, true>::template run<1>(&v[h.value * RR3]);
      sfor<0, RR3>([&](auto r3) { lds[e2p3[h.value] + B3 * r3.value] = v[h.value * RR3 + r3.value]; });
      __builtin_amdgcn_sched_barrier(0);  // keep the group's exchange writes here
    });
    lds_barrier();
    sfor<0, RR2>([&](auto k2) {
      sfor<0, RC2>([&](auto kc1) {
        v[k2.value * RC2 + kc1.value] = lds[e2p2 + B2 * k2.value + RC1 * kc1.value];
      });
    });
    sfor<1, RR2>([&](auto k2) {
      sfor<0, RC2>([&](auto kc1) { v[k2.value * RC2 + kc1.value] = cmulc(v[k2.value * RC2 + kc1.value], twl[twr_base + k2.value]); });
    });
    sfor<0, RC2>([&](auto kc1) { Dft<RR2, true>::template run<RC2>(&v[kc1.value]); });
    lds_barrier();  // every thread's exchange-2 reads are done before any exchange-1 write
    sfor<0, RR2>([&](auto r2) {
      Dft<RC2, true>::template run<1>(&v[r2.value * RC2]);
      sfor<0, RC2>([&](auto c0i) { lds[e1p2 + A2 * RR3 * r2.value + c0i.value] = v[r2.value * RC2 + c0i.value]; });
      __builtin_amdgcn_sched_barrier(0);  // keep the group's exchange writes here
    });
    lds_barrier();
    sfor<0, RR1>([&](auto k1) {
      sfor<0, RC1>([&](auto kc0) { v[k1.value * RC1 + kc0.value] = lds[e1p1 + A3 * k1.value + A1 * kc0.value]; });
    });
    pre1();
    sfor<1, RC1>([&](auto kc0) {
      sfor<0, RR1>([&](auto k1) { v[k1.value * RC1 + kc0.value] = cmulc(v[k1.value * RC1 + kc0.value], twl[twc_base + kc0.value]); });
    });
    sfor<0, RR1>([&](auto k1) { Dft<RC1, true>::template run<1>(&v[k1.value * RC1]); });
    sfor<0, RC1>([&](auto c1) { v[RC1 + c1.value] = cmulc(v[RC1 + c1.value], tw_rp); });
    sfor<0, RC1>([&](auto c1) { Dft<RR1, true>::template run<RC1>(&v[c1.value]); });
  }
};

}  // namespace qem
