// ptycho.hip — fused multislice mixed-probe ptychography forward / loss / adjoint for gfx950.
//
// One workgroup owns one (pattern b, probe mode m) wave for the whole multislice chain; the wave
// stays in registers (layout P1 / P3 of fft2.h) and only the FFT exchanges touch LDS.  HBM sees
// the object patches (L2-resident across modes), the stored per-slice incident waves (for the
// adjoint), the far-field wave and the per-pattern targets.
//
// Reference behaviour followed (quantem, src/quantem/diffractive_imaging/):
//   patch origin / fractional shift   dataset_models.py:164-167, 514-539, 1482-1499
//   object patch gather (periodic)    object_models.py:252-265
//   sub-pixel probe shift             ptycho_utils.py:128-139, 156-176; probe_models.py:718-720
//   multislice transmit + propagate   ptychography_base.py:934-947, 970-989
//   descan ramp on the exit wave      ptychography_base.py:903-905
//   far field, mode sum, fftshift     detector_models.py:22-30
//   loss                              ptychography_base.py:908-932
//   gradient                          torch autograd of the above (ptychography.py:327-329)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "common.h"
#include "fft2.h"
#include "ptycho_common.h"
#include "qem_ptycho.h"

namespace qem {

// ---------------------------------------------------------------------------------------------
// Probe spectra: pk[m] = FFT2(probe[m]).
// Separable propagators at N = 128.  The reference's propagator is exp(-i pi lambda dz (kr^2 + kc^2))
// times, for a tilt, one phase term per axis: a row factor times a column factor.  Per slice the
// workspace keeps a[r] = P[r][0] / N^2 and b[c] = P[0][c] / P[0][0] (kFacParts128 verdicts of 16 rows
// each say whether a[r] b[c] reproduces every entry within 1e-6 of the largest magnitude of the rows);
// the adjoint then reads both factors from LDS instead of N^2 entries from global memory per slice.
// The row factors are stored in the order of the adjoint's pass-3 registers: slot j * RR3 + i holds
// row j + RR1 * RR2 * i, so that register group h of lane-group q reads slots (q * H3 + h) * RR3 + i.
constexpr int kFacParts128 = 8;

template <int N>
__device__ __forceinline__ void prop_factors128(const Params& p, int blk) {
  using F = Fft2<N>;
  static_assert(F::T == 4 * N && N / kFacParts128 == 16, "four row groups of 16 rows per part");
  __shared__ double red[2][F::T / 64];
  const int s = blk / kFacParts128, part = blk - s * kFacParts128, t = threadIdx.x;
  const int c = t & (N - 1), g = t / N;
  const cf* P = p.prop + (size_t)s * N * N;
  cf* fa = p.pfac + (size_t)s * 2 * N;
  const cf p00 = P[0], p0c = P[c];
  const double d = (double)p00.x * p00.x + (double)p00.y * p00.y;
  const cf b_c = d > 0.0 ? mk((float)(((double)p0c.x * p00.x + (double)p0c.y * p00.y) / d),
                              (float)(((double)p0c.y * p00.x - (double)p0c.x * p00.y) / d))
                         : mk(0.0f, 0.0f);
  if (part == 0 && t < 2 * N) {
    if (t < N) {
      const int row = t / F::RR3 + F::RR1 * F::RR2 * (t % F::RR3);
      fa[t] = cscale(P[(size_t)row * N], 1.0f / (float)(N * N));
    } else {
      fa[t] = b_c;
    }
  }
  double dev = 0.0, mag = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = part * 16 + g + 4 * i;
    const cf pv = P[(size_t)r * N + c];
    const cf q = cmul(P[(size_t)r * N], b_c);
    const double ex = (double)q.x - (double)pv.x, ey = (double)q.y - (double)pv.y;
    const double e = ex * ex + ey * ey, m = (double)pv.x * pv.x + (double)pv.y * pv.y;
    // a NaN counts as an infinite deviation (or magnitude), and an infinity stays the maximum
    if (!(e <= dev)) dev = e == e ? e : INFINITY;
    if (!(m <= mag)) mag = m == m ? m : INFINITY;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double od = __shfl_down(dev, o, 64), om = __shfl_down(mag, o, 64);
    if (!(od <= dev)) dev = od;
    if (!(om <= mag)) mag = om;
  }
  if ((t & 63) == 0) {
    red[0][t >> 6] = dev;
    red[1][t >> 6] = mag;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < F::T / 64; ++w) {
      if (!(red[0][w] <= dev)) dev = red[0][w];
      if (!(red[1][w] <= mag)) mag = red[1][w];
    }
    // squared magnitudes: deviation <= 1e-6 * largest magnitude of the part, both finite
    p.pfac_ok[s * kFacParts128 + part] = (d > 0.0 && mag < INFINITY && dev <= 1e-12 * mag) ? 1 : 0;
  }
}

// Whether every propagator slice of the launch is separable (the verdicts of prop_factors128).
__device__ __forceinline__ bool all_separable128(const Params& p) {
  if (p.pfac == nullptr) return false;
  const int n = (p.S - 1) * kFacParts128;
  bool all = true;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + (int)(threadIdx.x & 63);
    all = all && __ballot(i < n && p.pfac_ok[i] == 0) == 0ull;
  }
  return all;
}

template <int N>
__global__ __launch_bounds__(Fft2<N>::T) void k_probe_spectrum(Params p) {
  __shared__ cf lds[Fft2<N>::LDS_TOTAL];
  using F = Fft2<N>;
  const int ramp_blocks = (p.B * 2 * N + F::T - 1) / F::T;
  const int fac_blocks = (N == 128 && p.pfac != nullptr) ? (p.S - 1) * kFacParts128 : 0;
  if constexpr (N == 128) {
    const int blk = (int)blockIdx.x - p.M - ramp_blocks;
    if (blk >= 0 && blk < fac_blocks) {
      // the blocks after the ramp table form and check the propagators' separable factors
      prop_factors128<N>(p, blk);
      return;
    }
  }
  if ((int)blockIdx.x >= p.M + ramp_blocks + fac_blocks) {
    // the last blocks scale the propagators by 1 / N^2 (exactly: a power of two) for the multislice
    // kernels, which then skip that multiply per element and slice
    const int n = (p.S - 1) * N * N;
    const float inv_n2 = 1.0f / (float)(N * N);
    const int first = p.M + ramp_blocks + fac_blocks;
    for (int idx = ((int)blockIdx.x - first) * F::T + (int)threadIdx.x; idx < n; idx += ((int)gridDim.x - first) * F::T)
      p.psc[idx] = cscale(p.prop[idx], inv_n2);
    return;
  }
  if ((int)blockIdx.x >= p.M) {
    // the blocks past the modes fill the shift-ramp table: ramps[b][0][k] = exp(-2 pi i f_k dr_b),
    // ramps[b][1][k] = exp(-2 pi i f_k dc_b), read by the forward prologue and the adjoint's last step
    const int idx = ((int)blockIdx.x - p.M) * F::T + (int)threadIdx.x;
    if (idx >= p.B * 2 * N) return;
    const int b = idx / (2 * N), r = idx - b * 2 * N, k = r & (N - 1);
    int r0, c0;
    float dr, dc;
    origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);
    p.ramps[idx] = expm2pi(fftfreq1<N>(k) * (r < N ? dr : dc));
    return;
  }
  F f(threadIdx.x);
  f.init_twiddles(lds, threadIdx.x);
  __syncthreads();
  const int m = blockIdx.x;
  cf v[F::PTS];
  const cf* src = p.probe + (size_t)m * N * N;
#pragma unroll
  for (int e = 0; e < F::PTS; ++e) v[e] = src[f.p1_off(e)];
  f.fwd(v, lds);
  cf* dst = p.pk + (size_t)m * N * N;
#pragma unroll
  for (int e = 0; e < F::PTS; ++e) dst[f.p3_off(e)] = v[e];
}

// Byte offsets of a thread's object pixels in one slice (layout P1: register e at patch row
// (e / RC1) * NRP + rp, column (e % RC1) * RC2 + c0, corner-centred).  The column of register
// e is qq + RC2 ((e % RC1 + RC1/2) % RC1) with qq the object column of patch column c0 - N/2, wrapped
// once (W >= N); only the RR1 row offsets and qq stay in registers.
template <int N>
struct ObjOffsets {
  using F = Fft2<N>;
  unsigned orow[F::RR1];
  int qq, W;
  __device__ __forceinline__ ObjOffsets(const F& f, int r0, int c0, int H, int W_) : W(W_) {
#pragma unroll
    for (int i = 0; i < F::RR1; ++i) {
      const int r = i * F::NRP + f.rp;
      orow[i] = (unsigned)(pmod(r0 + (r < N / 2 ? r : r - N), H) * W_) * 8u;
    }
    qq = pmod(c0 - N / 2 + f.c0, W_);
  }
  __device__ __forceinline__ unsigned at(int e) const {
    int x = qq + F::RC2 * ((e % F::RC1 + F::RC1 / 2) % F::RC1);
    x -= x >= W ? W : 0;
    return orow[e / F::RC1] + (unsigned)x * 8u;
  }
  // Keeps offsets derived from these values inside a slice loop from being hoisted out of it.
  __device__ __forceinline__ void opaque() {
#pragma unroll
    for (int i = 0; i < F::RR1; ++i) __asm__ volatile("" : "+v"(orow[i]));
    __asm__ volatile("" : "+v"(qq));
  }
};

// ---------------------------------------------------------------------------------------------
// Forward multislice for one (pattern, mode): psi_0 = ifft2(fft2(P_m) * ramp_b);
// for s: x_s = O_s[patch] * psi_s; psi_{s+1} = ifft2(fft2(x_s) * prop_s); X = fft2_ortho(x_{S-1}).
// ALL: every incident wave is kept (FWD_STORE_ALL, FWD_RECOMPUTE), so the stores are not under a
// branch; behind one, the wait-count analysis merged the two paths and made the transmission wait
// for the stores ahead of it.
// PG: also keep fft2 of every transmitted wave that is propagated (p.pspec), for the propagator gradient.
template <int N, bool DESCAN, bool ALL, bool PG = false>
__global__ __launch_bounds__(Fft2<N>::T) void k_forward(Params p) {
  __shared__ cf lds[Fft2<N>::LDS_TOTAL];
  using F = Fft2<N>;
  constexpr int PTS = F::PTS;
  F f(threadIdx.x);
  f.init_twiddles(lds, threadIdx.x);
  __syncthreads();
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);

  cf v[PTS];
  if (p.fwd_mode == FWD_RECOMPUTE) {
    // a segment of an earlier forward: start from its checkpoint wave
    const cf* src = p.ckpt + ((size_t)item * p.nck + p.s0 / p.ck) * N * N;
#pragma unroll
    for (int e = 0; e < PTS; e += 2) ld_stream2(src + f.wv_off(e), v[e], v[e + 1]);
    // through an empty asm, so that the values enter the slice loop as plain registers: as loads
    // still possibly in flight at the loop head they made every wave store of every slice wait
#pragma unroll
    for (int e = 0; e < PTS; ++e) asm volatile("" : "+v"(v[e].x), "+v"(v[e].y));
  } else {
    // shifted probe: spectrum * ramp / N^2, inverse transform (row and column ramps from the table)
    const cf* pk = p.pk + (size_t)m * N * N;
    const cf* rr = p.ramps + (size_t)b * 2 * N;
    const cf rc = rr[N + f.kc];
    const float inv_n2 = 1.0f / (float)(N * N);
#pragma unroll
    for (int e = 0; e < PTS; ++e) {
      const int kr = f.p3_row(e);
      v[e] = cscale(cmul(pk[kr * N + f.kc], cmul(rr[kr], rc)), inv_n2);
    }
    f.inv(v, lds);
  }

  ObjOffsets<N> oo(f, r0, c0, p.H, p.W);  // byte offsets of this thread's object pixels

  const int mode = p.fwd_mode;
  const bool far = mode != FWD_RECOMPUTE;
  const size_t HW = (size_t)p.H * p.W;
  // At N = 128 the object values of slice s + 1 are loaded inside the inverse transform of slice s's
  // propagation (after its last exchange, before its final register pass), which hides their L2
  // latency; at smaller N other resident waves hide it and the loads stay at the slice start.
  constexpr int OK = N >= 128 ? PTS : 0;  // object values loaded ahead
  // At N = 128 the propagator loads of the first QA registers (the first pass-3 register group) are
  // issued before the slice's wave stores: loads and stores share one vmcnt counter, so loads issued
  // after the stores (inside the transform) also waited for the stores.
  constexpr int QA = N >= 128 ? F::RR3 : 0;
  cf qa[QA > 0 ? QA : 1];
  cf o[PTS];
  oo.opaque();
#pragma unroll
  for (int e = 0; e < OK; ++e) o[e] = obj_at(p.obj + (size_t)p.s0 * HW, oo.at(e));
  for (int s = p.s0; s < p.s1; ++s) {
    const cf* O = p.obj + (size_t)s * HW;
    cf* dst = nullptr;
    if (ALL)  // FWD_STORE_ALL, FWD_RECOMPUTE
      dst = p.psi + ((size_t)item * p.wslots + (s - p.s0)) * N * N;
    else if (mode == FWD_CHECKPOINT && s % p.ck == 0)
      dst = p.ckpt + ((size_t)item * p.nck + s / p.ck) * N * N;
    const bool last_only = !far && s + 1 == p.s1;  // a recomputed segment needs only the incident waves
    // The object offsets are recomputed every slice: hoisted out of the loop they would hold ~46 VGPRs
    // through the transforms.  The object loads are issued before the wave stores, because loads and
    // stores share one vmcnt counter and a load issued after a store waits for that store as well.
    oo.opaque();
    if (!last_only) {
#pragma unroll
      for (int e = OK; e < PTS; ++e) o[e] = obj_at(O, oo.at(e));
    }
    if constexpr (QA > 0) {
      // the propagator values of the first QA registers are loaded before the wave stores; the last
      // slice (no propagation) reads the object's first N x N values instead, unused
      const cf* prq = s + 1 < p.s1 ? p.psc + (size_t)s * N * N : p.obj;
#pragma unroll
      for (int e = 0; e < QA; ++e) qa[e] = *f.p3_ptr(prq, e);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (ALL || dst != nullptr) {
#pragma unroll
      for (int e = 0; e < PTS; e += 2) st_stream2(dst + f.wv_off(e), v[e], v[e + 1]);
    }
    if (last_only) break;
#pragma unroll
    for (int e = 0; e < PTS; ++e) v[e] = cmul(o[e], v[e]);
    if (s + 1 < p.s1) {
      // propagation: the propagator values of each pass-3 register group are loaded before that
      // group's last transform stage, so that their L2 latency hides behind it
      const cf* pr = p.psc + (size_t)s * N * N;  // propagator / N^2
      cf* xs = PG ? p.pspec + ((size_t)item * (p.S - 1) + s) * N * N : nullptr;
      cf q[F::RR3];
      f.fwd_groups(
          v, lds,
          [&](auto h) {
            if constexpr (h.value * F::RR3 >= QA) {
#pragma unroll
              for (int i = 0; i < F::RR3; ++i) q[i] = pr[f.p3_off(h.value * F::RR3 + i)];
            }
          },
          [&](auto h) {
#pragma unroll
            for (int i = 0; i < F::RR3; ++i) {
              const int e = h.value * F::RR3 + i;
              if constexpr (PG) st_stream(xs + f.p3_off(e), v[e]);
              v[e] = cmul(v[e], e < QA ? qa[e < QA ? e : 0] : q[i]);
            }
          });
      const cf* On = O + HW;
      f.inv_hooked(v, lds, [&] {
        oo.opaque();
#pragma unroll
        for (int e = 0; e < OK; ++e) o[e] = obj_at(On, oo.at(e));
      });
    }
  }
  if (!far) return;
  if constexpr (DESCAN) {
    const float ddr = p.descan[2 * b], ddc = p.descan[2 * b + 1];
#pragma unroll
    for (int e = 0; e < PTS; ++e) {
      const cf ramp = cmul(expm2pi(fftfreq1<N>(f.p1_row(e)) * ddr), expm2pi(fftfreq1<N>(f.p1_col(e)) * ddc));
      v[e] = cmul(v[e], ramp);
    }
  }
  f.fwd(v, lds);
  cf* X = p.X + (size_t)item * N * N;
  const float inv_n = 1.0f / (float)N;
#pragma unroll
  for (int e = 0; e < PTS; ++e) X[f.p3_off(e)] = cscale(v[e], inv_n);
}

// ---------------------------------------------------------------------------------------------
// Loss: I = sum_m |X_m|^2 (unshifted k), compared with fftshifted targets; writes dL/dI and
// per-block partial sums.  256 threads per block, one frequency pixel per thread in each of the
// block's kLossTiles consecutive 256-pixel tiles (fewer partials for the single-block finalize);
// MC = min(M, 4) probe modes loaded together.
constexpr int kLossTiles = 8;

template <int N, int MC>
__global__ __launch_bounds__(256) void k_loss(Params p, int ntiles) {
  __shared__ double red[4];
  const bool amp = (p.loss_type == QEM_LOSS_L2_AMPLITUDE || p.loss_type == QEM_LOSS_L1_AMPLITUDE);
  const float scale = (p.loss_type == QEM_LOSS_POISSON) ? 1.0f / p.mean_I
                                                          : (float)p.P / ((float)p.B * p.mean_I);
  double acc = 0.0;
  // each thread sums its terms in tile order, then the block sums the threads
  for (int tile = blockIdx.x * kLossTiles; tile < ntiles && tile < (blockIdx.x + 1) * kLossTiles; ++tile) {
    const int idx = tile * 256 + threadIdx.x;
    const int b = idx / (N * N);
    const int k = idx - b * N * N;
    const int kr = k / N, kc = k - kr * N;
    float I = 0.0f;
    for (int m0 = 0; m0 < p.M; m0 += MC) {  // MC loads in flight; mode indices past M - 1 are not added
      cf x[MC];
#pragma unroll
      for (int j = 0; j < MC; ++j) x[j] = p.X[((size_t)b * p.M + (m0 + j < p.M ? m0 + j : p.M - 1)) * N * N + k];
#pragma unroll
      for (int j = 0; j < MC; ++j)
        if (m0 + j < p.M) I += cabs2(x[j]);
    }
    const int ks = ((kr + N / 2) & (N - 1)) * N + ((kc + N / 2) & (N - 1));
    if (p.intens_out != nullptr) p.intens_out[(size_t)b * N * N + ks] = I;
    const float t = p.targets[(size_t)b * N * N + ks];
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
    p.gI[(size_t)b * N * N + k] = scale * dI;
    acc += (double)term;
  }
  const double tot = block_sum<256>(acc, red);
  if (threadIdx.x == 0) p.loss_part[blockIdx.x] = tot;
}

// Fixed-order sum of the loss partials.
__global__ __launch_bounds__(256) void k_loss_finalize(Params p, int nparts) {
  __shared__ double red[4];
  double x = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) x += p.loss_part[i];
  const double tot = block_sum<256>(x, red);
  if (threadIdx.x == 0) {
    const double scale = (p.loss_type == QEM_LOSS_POISSON) ? 1.0 / (double)p.mean_I
                                                            : (double)p.P / ((double)p.B * (double)p.mean_I);
    p.loss_out[0] = (float)(tot * scale);
  }
}

// ---------------------------------------------------------------------------------------------
// Adjoint for one (pattern, mode).  g_X = 2 dL/dI X; g_x = ifft2_ortho^H(g_X); then for s = S-1..0:
//   dO_s[patch] (this mode) = conj(psi_s) g_x      -> stored over psi_s (natural order)
//   g_psi = conj(O_s) g_x;  g_x = ifft2(fft2(g_psi) conj(prop_{s-1}))   (s > 0)
// At s = 0: G = fft2(g_psi_0) and the probe-gradient spectrum conj(ramp_b) G is stored over X;
// the position gradient reduces Im(conj(G) P_m ramp_b) against the frequencies.
// PG: also the propagator-gradient contributions (over p.pspec, see Params).
// One (pattern, mode) item of the adjoint.  FAC (N = 128, no descan, no propagator gradient): the
// propagations apply the separable factors of prop_factors128, read from fab (4N slots: the factors of
// two propagations, double-buffered by slice parity, row factors then column factors).
template <int N, bool DESCAN, bool PG, bool FAC>
__device__ __forceinline__ void adjoint_item(const Params& p, cf* __restrict__ lds, double* __restrict__ red,
                                             cf* __restrict__ fab) {
  using F = Fft2<N>;
  constexpr int PTS = F::PTS;
  F f(threadIdx.x);
  f.init_twiddles(lds, threadIdx.x);
  __syncthreads();
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);
  const float dl = (p.flags & QEM_FLAG_ANALYTIC) ? 1.0f : p.dloss[0];
  const float inv_n = 1.0f / (float)N;
  const float inv_n2 = 1.0f / (float)(N * N);
  // With FAC the first 2N threads (four whole waves) copy the factors of propagator j into LDS: each
  // loads its value one slice ahead (fload, beside the next slice's wave prefetch) and writes it at
  // the start of the slice that uses it.  The load is branch-free (the other threads load a copy).
  cf fv = mk(0.0f, 0.0f);
  const bool fwriter = __builtin_amdgcn_readfirstlane((int)threadIdx.x) < 2 * N;
  auto fload = [&](int j) {
    if constexpr (FAC) fv = p.pfac[(size_t)(j > 0 ? j : 0) * 2 * N + (threadIdx.x & (2 * N - 1))];
  };

  cf v[PTS];
  cf* Xs = p.X + (size_t)item * N * N;
  const bool top = p.s1 == p.S;  // this launch starts at the far field
  if (!top) {
    // a lower segment: g_x of slice s1 left by the launch of the segment above (natural order)
#pragma unroll
    for (int e = 0; e < PTS; ++e) v[e] = Xs[f.p1_off(e)];
  } else if (p.flags & QEM_FLAG_ANALYTIC) {
    // the projection kernel left the Fourier-space residual X' - X over X: g = ifft2_ortho(residual)
#pragma unroll
    for (int e = 0; e < PTS; ++e) v[e] = cscale(Xs[f.p3_off(e)], inv_n);
  } else {
    const float* gI = p.gI + (size_t)b * N * N;
    const float sc = 2.0f * dl * inv_n;
#pragma unroll
    for (int e = 0; e < PTS; ++e) {
      const int k = f.p3_off(e);
      v[e] = cscale(Xs[k], sc * gI[k]);
    }
  }
  if (top) f.inv(v, lds);

  ObjOffsets<N> oo(f, r0, c0, p.H, p.W);  // byte offsets of this thread's object pixels

  const size_t HW = (size_t)p.H * p.W;
  const bool want_desc = DESCAN && (p.flags & QEM_FLAG_GRAD_DESCAN);
  // Descan (applied after the last slice in the forward): g_x <- conj(ramp) g_x, and its gradient
  // 2 pi sum f * Im(conj(g_x) x_out) with x_out = ramp * O_{S-1} psi_{S-1}.
  if (DESCAN && top) {
    const cf* O = p.obj + (size_t)(p.S - 1) * HW;
    const cf* ps = p.psi + ((size_t)item * p.wslots + (p.S - 1 - p.s0)) * N * N;
    const float ddr = p.descan[2 * b], ddc = p.descan[2 * b + 1];
    double ar = 0.0, ac = 0.0;
#pragma unroll
    for (int e = 0; e < PTS; ++e) {
      const int row = f.p1_row(e), col = f.p1_col(e);
      const float fr = fftfreq1<N>(row), fc = fftfreq1<N>(col);
      const cf ramp = cmul(expm2pi(fr * ddr), expm2pi(fc * ddc));
      if (want_desc) {
        const cf xo = cmul(cmul(obj_at(O, oo.at(e)), *f.wv_ptr(ps, e)), ramp);
        const float im = cconjmul(v[e], xo).y;
        ar += (double)(fr * im);
        ac += (double)(fc * im);
      }
      v[e] = cmulc(v[e], ramp);
    }
    if (want_desc) {
      const double tr = block_sum<F::T>(ar, red);
      const double tc = block_sum<F::T>(ac, red);
      if (threadIdx.x == 0) {
        p.gdesc_part[2 * item] = (float)(2.0 * M_PI * tr);
        p.gdesc_part[2 * item + 1] = (float)(2.0 * M_PI * tc);
      }
    }
  }
  // One slice back: object-gradient patch conj(psi_s) g_x (written over psi_s), then
  // g_psi = conj(O_s) g_x, in chunks of CH elements whose psi and object loads are issued before
  // their stores.  At N = 128 the next slice's psi values (all 32; 20 in the descan variant, whose
  // extra state would otherwise spill) are loaded during this slice's propagation, after the
  // propagator multiply (issued before it, they would hold up the propagator loads: one vmcnt), so
  // that their HBM latency hides behind the inverse transform; smaller N keep their occupancy instead.
  // With all of them prefetched (N = 128 without descan) the slice's object values are loaded ahead
  // too, inside that inverse transform (after its last exchange), and the slice step only stores.
  constexpr int CH = PTS < 16 ? PTS : 16;
  constexpr int PF = N >= 128 ? (DESCAN ? 20 : PTS) : 0;
  constexpr bool AHEAD = PF == PTS;
  cf pre[PF > 0 ? PF : 1];
  cf ob[AHEAD ? PTS : 1];
  // With every wave and object value prefetched (AHEAD), the slice step also issues the propagator
  // loads of the slice's propagation, before its patch stores: loads and stores share one vmcnt
  // counter, so propagator loads issued after the stores (inside the transform) waited for them.
  cf qa[AHEAD ? F::RR3 : 1];
  auto slice_step = [&](int s, bool have_pre) {
    const cf* O = p.obj + (size_t)s * HW;
    cf* ps = p.psi + ((size_t)item * p.wslots + (s - p.s0)) * N * N;
    oo.opaque();
    if constexpr (AHEAD) {
      // the patch values conj(psi_s) g_x go into pre, g_psi = conj(O_s) g_x into v
#pragma unroll
      for (int i = 0; i < PTS; ++i) {
        pre[i] = cconjmul(pre[i], v[i]);
        v[i] = cconjmul(ob[i], v[i]);
      }
      __builtin_amdgcn_sched_barrier(0);  // the object values are dead before the propagator loads
      // slice 0 has no propagation: its loads read the object's first N x N values, unused
      if constexpr (!FAC) {
        const cf* prq = s > 0 ? p.psc + (size_t)(s - 1) * N * N : p.obj;
#pragma unroll
        for (int i = 0; i < F::RR3; ++i) qa[i] = *f.p3_ptr(prq, i);
      }
      __builtin_amdgcn_sched_barrier(0);  // the first register group's loads are issued before the stores
#pragma unroll
      for (int i = 0; i < PTS; i += 2) st_stream2(f.wv_ptr(ps, i), pre[i], pre[i + 1]);
    } else {
#pragma unroll
      for (int e0 = 0; e0 < PTS; e0 += CH) {
        cf t[CH], o[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          if (e0 + i < PF && have_pre) {
            t[i] = pre[e0 + i];
          } else if (i % 2 == 0) {
            ld_stream2(f.wv_ptr(ps, e0 + i), t[i], t[i + 1]);
          }
          o[i] = obj_at(O, oo.at(e0 + i));
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          if (i % 2 == 0)
            st_stream2(f.wv_ptr(ps, e0 + i), cconjmul(t[i], v[e0 + i]), cconjmul(t[i + 1], v[e0 + i + 1]));
          v[e0 + i] = cconjmul(o[i], v[e0 + i]);
        }
      }
    }
  };
  if constexpr (AHEAD) {
    // the top slice's psi and object values, loaded as every later slice's are prefetched
    const cf* pt = p.psi + ((size_t)item * p.wslots + (p.s1 - 1 - p.s0)) * N * N;
#pragma unroll
    for (int i = 0; i < PF; i += 2) ld_stream2(f.wv_ptr(pt, i), pre[i], pre[i + 1]);
    oo.opaque();
#pragma unroll
    for (int i = 0; i < PTS; ++i) ob[i] = obj_at(p.obj + (size_t)(p.s1 - 1) * HW, oo.at(i));
  }
  fload(p.s1 - 2);
  for (int s = p.s1 - 1; s >= p.s0; --s) {
    // propagator s - 1's factors into LDS (this buffer's last reads, two slices up, are separated from
    // here by that slice's exchange barriers; the reads below follow the forward transform's barriers)
    if constexpr (FAC)
      if (fwriter) fab[((s - 1) & 1) * 2 * N + threadIdx.x] = fv;
    slice_step(s, s < p.s1 - 1);
    if (s > 0) {
      // g_x <- ifft2(fft2(g_psi) conj(prop_{s-1})), the adjoint of the forward propagation, with the
      // propagator loads of each pass-3 register group issued before that group's last stage
      const cf* __restrict__ pr = p.psc + (size_t)(s - 1) * N * N;  // propagator / N^2
      cf* xs = PG ? p.pspec + ((size_t)item * (p.S - 1) + (s - 1)) * N * N : nullptr;
      cf q[F::RR3];
      cf fb = mk(0.0f, 0.0f);  // column factor b[kc] (FAC)
      const cf* fa = FAC ? fab + ((s - 1) & 1) * 2 * N + f.q * F::H3 * F::RR3 : fab;  // this lane group's row factors
      cf xq[PG ? F::RR3 : 1];
      f.fwd_groups(
          v, lds,
          [&](auto h) {
            if constexpr (FAC) {
              if constexpr (h.value == 0) fb = fab[((s - 1) & 1) * 2 * N + N + f.kc];
            } else if constexpr (!AHEAD || h.value > 0) {
#pragma unroll
              for (int i = 0; i < F::RR3; ++i) q[i] = *f.p3_ptr(pr, h.value * F::RR3 + i);
            }
            if constexpr (PG) {
#pragma unroll
              for (int i = 0; i < F::RR3; ++i) xq[i] = ld_stream(f.p3_ptr(xs, h.value * F::RR3 + i));
            }
          },
          [&](auto h) {
#pragma unroll
            for (int i = 0; i < F::RR3; ++i) {
              const int e = h.value * F::RR3 + i;
              // dL/dP_{s-1} of this item: conj(fft2(x_{s-1})) fft2(g_psi_s) / N^2, over fft2(x_{s-1})
              if constexpr (PG) st_stream(f.p3_ptr(xs, e), cscale(cconjmul(xq[i], v[e]), inv_n2));
              if constexpr (FAC)
                v[e] = cmulc(cmulc(v[e], fa[h.value * F::RR3 + i]), fb);
              else
                v[e] = cmulc(v[e], AHEAD && h.value == 0 ? qa[i] : q[i]);
            }
          });
      if constexpr (PF > 0) {
        // Branch-free: at a segment bottom (s - 1 < s0) this reads the segment's first slot and the
        // values go unused.  Under a condition the compiler moved the transform's arithmetic below
        // the prefetch, behind a register reload that waits for every load in flight.
        const int sl = s - 1 >= p.s0 ? s - 1 - p.s0 : 0;
        const cf* pn = p.psi + ((size_t)item * p.wslots + sl) * N * N;
#pragma unroll
        for (int i = 0; i < PF; i += 2) ld_stream2(f.wv_ptr(pn, i), pre[i], pre[i + 1]);
      }
      fload(s - 2);
      const cf* On = p.obj + (size_t)(s - 1) * HW;
      f.inv_hooked(v, lds, [&] {
        if constexpr (AHEAD) {
          oo.opaque();
#pragma unroll
          for (int i = 0; i < PTS; ++i) ob[i] = obj_at(On, oo.at(i));
        }
      });
    }
  }
  if (p.s0 > 0) {
    // g_x of slice s0 - 1 for the segment below
#pragma unroll
    for (int e = 0; e < PTS; ++e) Xs[f.p1_off(e)] = v[e];
    return;
  }
  // G = fft2(g_psi_0) in natural order.  Over X goes the probe-gradient spectrum conj(ramp_b) G
  // (G itself for the analytical update: the reference sums g at the probe plane without the
  // shift's adjoint); the position-gradient partials are (2 pi / N^2) sum f Im(conj(G) P_m ramp_b).
  if (p.flags & QEM_FLAG_ANALYTIC) {
    f.fwd(v, lds);
#pragma unroll
    for (int e = 0; e < PTS; ++e) Xs[f.p3_off(e)] = v[e];
    return;
  }
  const cf* rr = p.ramps + (size_t)b * 2 * N;
  const cf* pk = p.pk + (size_t)m * N * N;
  const cf rc = rr[N + f.kc];
  const float fc = fftfreq1<N>(f.kc);
  double ar = 0.0, ac = 0.0;
  // per pass-3 register group: the probe-spectrum and ramp values are loaded before the group's last
  // transform stage and used after it (terms summed in register order, as one loop would)
  cf qp[F::RR3], qr[F::RR3];
  f.fwd_groups(
      v, lds,
      [&](auto h) {
#pragma unroll
        for (int i = 0; i < F::RR3; ++i) {
          const int e = h.value * F::RR3 + i;
          qp[i] = pk[f.p3_off(e)];
          qr[i] = rr[f.p3_row(e)];
        }
      },
      [&](auto h) {
#pragma unroll
        for (int i = 0; i < F::RR3; ++i) {
          const int e = h.value * F::RR3 + i;
          const int kr = f.p3_row(e), off = f.p3_off(e);
          const cf ramp = cmul(qr[i], rc);
          Xs[off] = cmulc(v[e], ramp);
          const float im = cconjmul(v[e], cmul(qp[i], ramp)).y;
          ar += (double)(fftfreq1<N>(kr) * im);
          ac += (double)(fc * im);
        }
      });
  if (p.flags & QEM_FLAG_GRAD_POSITIONS) {
    const double tr = block_sum<F::T>(ar, red);
    const double tc = block_sum<F::T>(ac, red);
    if (threadIdx.x == 0) {
      p.gpos_part[2 * item] = (float)(2.0 * M_PI * tr / (double)(N * N));
      p.gpos_part[2 * item + 1] = (float)(2.0 * M_PI * tc / (double)(N * N));
    }
  }
}

// The adjoint kernel.  At N = 128 without descan or propagator gradient it takes the separable-factor
// path when every propagator slice of the launch passed the factor check (a uniform branch: one
// launch, and the workgroup's LDS holds the factor tables as well).
template <int N, bool DESCAN, bool PG = false>
__global__ __launch_bounds__(Fft2<N>::T) void k_adjoint(Params p) {
  constexpr bool FACTORS = N == 128 && !DESCAN && !PG;
  __shared__ cf lds[Fft2<N>::LDS_TOTAL];
  __shared__ double red[Fft2<N>::T / 64];
  __shared__ cf fab[FACTORS ? 4 * N : 1];
  if constexpr (FACTORS) {
    if (all_separable128(p)) {
      adjoint_item<N, DESCAN, PG, true>(p, lds, red, fab);
      return;
    }
  }
  adjoint_item<N, DESCAN, PG, false>(p, lds, red, fab);
}

// ---------------------------------------------------------------------------------------------
// Column permutation of the stored-wave layout (Fft2::wv_off) for N <= 128; the N = 256 chain stores
// natural rows.  stored_pos: position of natural patch column c in a stored row; stored_col: inverse.
template <int N>
__device__ __forceinline__ int stored_pos(int c) {
  if constexpr (N <= 128) {
    constexpr int RC2 = Fft2<N>::RC2;
    const int c0 = c % RC2, r1 = c / RC2;
    return (r1 / 2) * 2 * RC2 + 2 * c0 + (r1 % 2);
  } else {
    return c;
  }
}
template <int N>
__device__ __forceinline__ int stored_col(int pos) {
  if constexpr (N <= 128) {
    constexpr int RC2 = Fft2<N>::RC2;
    const int rem = pos % (2 * RC2);
    return (2 * (pos / (2 * RC2)) + rem % 2) * RC2 + rem / 2;
  } else {
    return pos;
  }
}

// ---------------------------------------------------------------------------------------------
// Fourier projection residual of the analytical update (ptychography.py:356-372), over X:
//   one mode:  X' = A exp(i angle(X))             -> X' - X
//   M modes:   X'_m = (A / a) X_m, a = sqrt(sum_m |X_m + 1e-9|^2), A / a = 0 where a = 0
// with A the measured amplitude at the fftshifted pixel.
template <int N>
__global__ __launch_bounds__(256) void k_projection(Params p) {
  const int idx = blockIdx.x * 256 + threadIdx.x;  // over B*N*N
  const int b = idx / (N * N);
  const int k = idx - b * N * N;
  const int kr = k / N, kc = k - kr * N;
  const int ks = ((kr + N / 2) & (N - 1)) * N + ((kc + N / 2) & (N - 1));
  const float A = p.targets[(size_t)b * N * N + ks];
  cf* X = p.X + (size_t)b * p.M * N * N + k;
  if (p.M == 1) {
    const cf x = X[0];
    const float mag = sqrtf(cabs2(x));
    X[0] = mag > 0.0f ? mk(x.x * (A / mag) - x.x, x.y * (A / mag) - x.y) : mk(A - x.x, -x.y);
  } else {
    float s2 = 0.0f;
    for (int m = 0; m < p.M; ++m) {
      const cf x = X[(size_t)m * N * N];
      s2 += (x.x + 1e-9f) * (x.x + 1e-9f) + x.y * x.y;
    }
    const float a = sqrtf(s2);
    const float r = a > 0.0f ? A / a : 0.0f;
    for (int m = 0; m < p.M; ++m) {
      const cf x = X[(size_t)m * N * N];
      X[(size_t)m * N * N] = mk(x.x * (r - 1.0f), x.y * (r - 1.0f));
    }
  }
}

// Per-mode probe overlap sum_b |psi_s,m,b|^2 scattered into the object frame (the normalisation of
// ObjectPixelated.backward, object_models.py:704-706); same tiling as k_obj_reduce, one mode per
// blockIdx.z % M.  Output (S, M, H, W) float32.
template <int N>
__global__ __launch_bounds__(256) void k_overlap(Params p, float* out) {
  __shared__ int su0[256], sv0[256];
  const int s = blockIdx.z / p.M, m = blockIdx.z - (blockIdx.z / p.M) * p.M;
  const int X0 = blockIdx.x * 32, Y0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int x = X0 + tx;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int b0 = 0; b0 < p.B; b0 += 256) {
    const int nb = min(256, p.B - b0);
    if ((int)threadIdx.x < nb) {
      const int bb = b0 + threadIdx.x;
      const int r0 = (int)rintf(p.pos[2 * bb]), c0 = (int)rintf(p.pos[2 * bb + 1]);
      const int u0 = pmod(Y0 - r0 + N / 2, p.H), v0 = pmod(X0 - c0 + N / 2, p.W);
      const bool hit = (u0 < N || u0 + 31 >= p.H) && (v0 < N || v0 + 31 >= p.W);
      su0[threadIdx.x] = hit ? u0 : -1;
      sv0[threadIdx.x] = v0;
    }
    __syncthreads();
    for (int i = 0; i < nb; ++i) {
      const int u0 = su0[i];
      if (u0 < 0) continue;
      int ux = sv0[i] + tx;
      if (ux >= p.W) ux -= p.W;
      if (x >= p.W || ux >= N) continue;
      const int col = stored_pos<N>((ux + N / 2) & (N - 1));
      const cf* g = p.psi + (((size_t)(b0 + i) * p.M + m) * p.wslots + s) * N * N + col;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int dy = ty + 8 * j;
        int uy = u0 + dy;
        if (uy >= p.H) uy -= p.H;
        if (Y0 + dy < p.H && uy < N) acc[j] += cabs2(g[((uy + N / 2) & (N - 1)) * N]);
      }
    }
    __syncthreads();
  }
  if (x >= p.W) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int y = Y0 + ty + 8 * j;
    if (y < p.H) out[(((size_t)s * p.M + m) * p.H + y) * p.W + x] = acc[j];
  }
}

// Patch energy sum |O_s[patch b]|^2 (ProbePixelated.backward's normalisation, probe_models.py:775),
// one block per (s, b); output (S, B) float32.
template <int N>
__global__ __launch_bounds__(256) void k_patch_energy(Params p, float* out) {
  __shared__ double red[4];
  const int s = blockIdx.x / p.B, b = blockIdx.x - (blockIdx.x / p.B) * p.B;
  const int r0 = (int)rintf(p.pos[2 * b]), c0 = (int)rintf(p.pos[2 * b + 1]);
  const cf* O = p.obj + (size_t)s * p.H * p.W;
  double acc = 0.0;
  for (int idx = threadIdx.x; idx < N * N; idx += 256) {
    const int r = idx / N, c = idx - (idx / N) * N;
    const int y = pmod(r0 + (r < N / 2 ? r : r - N), p.H), xx = pmod(c0 + (c < N / 2 ? c : c - N), p.W);
    acc += (double)cabs2(O[(size_t)y * p.W + xx]);
  }
  const double t = block_sum<256>(acc, red);
  if (threadIdx.x == 0) out[(size_t)s * p.B + b] = (float)t;
}

// ---------------------------------------------------------------------------------------------
// Deterministic object-gradient scatter-add: each block owns a 32x32 tile of one slice and sums,
// in increasing (b, m) order, the per-mode patch gradients that land on it.  Patch row index of
// object row y for a patch with origin r0: u = (y - r0 + N/2) mod H (valid if u < N), stored row
// (u + N/2) mod N (fftfreq order).  The block first tabulates, for 256 patterns at a time, the
// tile origin in patch coordinates (or -1 when the patch misses the tile), then walks the table.
template <int N>
__global__ __launch_bounds__(256) void k_obj_reduce(Params p) {
  __shared__ int su0[256], sv0[256];
  const int s = p.s0 + blockIdx.z;  // object slice; its patches are in psi slot blockIdx.z
  const int X0 = blockIdx.x * 32, Y0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int x = X0 + tx;
  cf acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = mk(0.0f, 0.0f);
  const size_t slab = (size_t)p.wslots * N * N;  // stride between modes of one pattern
  for (int b0 = 0; b0 < p.B; b0 += 256) {
    const int nb = min(256, p.B - b0);
    if ((int)threadIdx.x < nb) {
      const int b = b0 + threadIdx.x;
      const int r0 = (int)rintf(p.pos[2 * b]), c0 = (int)rintf(p.pos[2 * b + 1]);
      const int u0 = pmod(Y0 - r0 + N / 2, p.H), v0 = pmod(X0 - c0 + N / 2, p.W);
      const bool hit = (u0 < N || u0 + 31 >= p.H) && (v0 < N || v0 + 31 >= p.W);
      su0[threadIdx.x] = hit ? u0 : -1;
      sv0[threadIdx.x] = v0;
    }
    __syncthreads();
    for (int i = 0; i < nb; ++i) {
      const int u0 = su0[i];
      if (u0 < 0) continue;
      int ux = sv0[i] + tx;
      if (ux >= p.W) ux -= p.W;
      const bool xok = x < p.W && ux < N;
      const int col = stored_pos<N>((ux + N / 2) & (N - 1));
      const cf* g = p.psi + ((size_t)(b0 + i) * p.M * p.wslots + blockIdx.z) * N * N + col;
      const cf* rowp[4];
      bool ok[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int dy = ty + 8 * j;
        int uy = u0 + dy;
        if (uy >= p.H) uy -= p.H;
        ok[j] = xok && (Y0 + dy < p.H) && uy < N;
        rowp[j] = g + ((uy + N / 2) & (N - 1)) * N;
      }
      // all loads of a pattern are issued before any add (masked lanes read a valid dummy address)
      for (int m0 = 0; m0 < p.M; m0 += 4) {
        cf t[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) {
            const bool use = ok[j] && (m0 + mm < p.M);
            t[j][mm] = ld_stream(use ? rowp[j] + (size_t)(m0 + mm) * slab : (const cf*)p.psi);
            if (!use) t[j][mm] = mk(0.0f, 0.0f);
          }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) {
            acc[j].x += t[j][mm].x;
            acc[j].y += t[j][mm].y;
          }
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

// Deterministic object-gradient scatter-add, one object row per block (the layout used when the row
// fits the LDS): the block lists, in batch order, the patterns whose patch covers the row; each of
// its N threads (one per stored position of a patch row) then reads that patch row of every mode (N * 8
// contiguous bytes per mode), sums the modes in order and adds the sum into a W-wide LDS row.  The
// columns of one patch are distinct (W >= N), those of consecutive patches may coincide, so a barrier
// separates the adds of consecutive patterns: every object pixel sums its patterns in increasing b.
// MC: probe modes whose loads are issued together (min(M, 4)).
template <int N, int MC>
__global__ __launch_bounds__(N) void k_obj_reduce_rows(Params p) {
  extern __shared__ cf acc_row[];  // p.W entries
  __shared__ int lb[N], lrow[N], lc0[N];
  __shared__ int wcount[(N + 63) / 64];
  constexpr int K = 4;  // patterns whose loads are in flight together
  const int t = threadIdx.x;
  const int y = blockIdx.x, slot = blockIdx.y, s = p.s0 + slot;
  const int ct = stored_col<N>(t);          // patch column at stored position t
  const int x_off = ct < N / 2 ? ct : ct - N;  // its natural column offset
  const size_t slab = (size_t)p.wslots * N * N;  // stride between modes of one pattern
  for (int x = t; x < p.W; x += N) acc_row[x] = mk(0.0f, 0.0f);
  for (int b0 = 0; b0 < p.B; b0 += N) {
    // compact the patterns of this chunk that cover row y, keeping batch order
    const int b = b0 + t;
    int u = N;
    int c0m = 0;
    if (b < p.B) {
      const int r0 = (int)rintf(p.pos[2 * b]), c0 = (int)rintf(p.pos[2 * b + 1]);
      u = pmod(y - r0 + N / 2, p.H);
      c0m = pmod(c0, p.W);
    }
    const bool hit = u < N;
    const unsigned long long bal = __ballot(hit);
    const int lane = t & 63, wave = t >> 6;
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();  // also orders the zeroing / the previous chunk's reads of the list
    int off = 0, cnt = 0;
    for (int w = 0; w < (N + 63) / 64; ++w) {
      off += w < wave ? wcount[w] : 0;
      cnt += wcount[w];
    }
    if (hit) {
      lb[off + pre] = b;
      lrow[off + pre] = (u + N / 2) & (N - 1);
      lc0[off + pre] = c0m;
    }
    __syncthreads();
    for (int i0 = 0; i0 < cnt; i0 += K) {
      cf val[K];
      int xo[K];
      size_t base[K];  // patch row of pattern i0 + k (of pattern i0 past the list's end: loaded, unused)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int i = i0 + k < cnt ? i0 + k : i0;
        base[k] = ((size_t)lb[i] * p.M * p.wslots + slot) * N * N + (size_t)lrow[i] * N + t;
        int x = lc0[i] + x_off;  // in [-N/2, W + N/2): one wrap each way
        x += x < 0 ? p.W : 0;
        x -= x >= p.W ? p.W : 0;
        xo[k] = x;
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
            if (m0 + j == 0) {
              val[k] = v[k][j];
            } else if (m0 + j < p.M) {
              val[k].x += v[k][j].x;
              val[k].y += v[k][j].y;
            }
          }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (i0 + k < cnt) {  // uniform over the block
          acc_row[xo[k]].x += val[k].x;
          acc_row[xo[k]].y += val[k].y;
          __syncthreads();
        }
      }
    }
    __syncthreads();  // the list is rewritten by the next chunk
  }
  cf* out = p.grad_obj + ((size_t)s * p.H + y) * p.W;
  for (int x = t; x < p.W; x += N) out[x] = acc_row[x];
}

// gsum[m][k] = sum_b spectra[b][m][k] in a fixed order: 128 frequencies per block, two adjacent ones per
// lane (one 16-byte load per pattern), 4 lanes per frequency pair each summing every 4th pattern,
// combined in lane order through LDS.
template <int N>
__global__ __launch_bounds__(256) void k_probe_reduce(Params p) {
  __shared__ float4 part[4][64];
  const int kk = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int idx = blockIdx.x * 128 + 2 * kk;  // over M*N*N (multiple of 128)
  const int m = idx / (N * N);
  const int k = idx - m * N * N;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4* src = reinterpret_cast<const float4*>(p.X + (size_t)m * N * N + k);
  const size_t stride = (size_t)p.M * N * N / 2;  // one pattern, in float4
#pragma unroll 8
  for (int b = g; b < p.B; b += 4) {
    const float4 t = src[(size_t)b * stride];
    acc.x += t.x;
    acc.y += t.y;
    acc.z += t.z;
    acc.w += t.w;
  }
  part[g][kk] = acc;
  __syncthreads();
  if (g == 0) {
    float4 t = part[0][kk];
    for (int i = 1; i < 4; ++i) {
      t.x += part[i][kk].x;
      t.y += part[i][kk].y;
      t.z += part[i][kk].z;
      t.w += part[i][kk].w;
    }
    reinterpret_cast<float4*>(p.gsum)[(size_t)idx / 2] = t;
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_prop_reduce(Params p, cf* grad_prop) {
  __shared__ cf part[4][64];
  const int kk = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int idx = blockIdx.x * 64 + kk;  // over (S-1)*N*N (a multiple of 64)
  const size_t slot = (size_t)(p.S - 1) * N * N;
  const int items = p.B * p.M;
  cf acc = mk(0.0f, 0.0f);
#pragma unroll 8
  for (int it = g; it < items; it += 4) {
    const cf t = ld_stream(p.pspec + (size_t)it * slot + idx);
    acc.x += t.x;
    acc.y += t.y;
  }
  part[g][kk] = acc;
  __syncthreads();
  if (g == 0) {
    cf t = part[0][kk];
    for (int i = 1; i < 4; ++i) {
      t.x += part[i][kk].x;
      t.y += part[i][kk].y;
    }
    grad_prop[idx] = t;
  }
}

// grad_probe[m] = ifft2(gsum[m]) (1/N^2 included).
template <int N>
__global__ __launch_bounds__(Fft2<N>::T) void k_probe_grad(Params p) {
  __shared__ cf lds[Fft2<N>::LDS_TOTAL];
  using F = Fft2<N>;
  F f(threadIdx.x);
  f.init_twiddles(lds, threadIdx.x);
  __syncthreads();
  const int m = blockIdx.x;
  cf v[F::PTS];
  const cf* src = p.gsum + (size_t)m * N * N;
#pragma unroll
  for (int e = 0; e < F::PTS; ++e) v[e] = src[f.p3_off(e)];
  f.inv(v, lds);
  const float inv_n2 = 1.0f / (float)(N * N);
  cf* dst = p.grad_probe + (size_t)m * N * N;
#pragma unroll
  for (int e = 0; e < F::PTS; ++e) dst[f.p1_off(e)] = cscale(v[e], inv_n2);
}

// grad[b][i] = sum_m part[(b*M + m)*2 + i]
__global__ __launch_bounds__(256) void k_pair_reduce(const float* part, float* out, int B, int M) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= 2 * B) return;
  const int b = idx >> 1, i = idx & 1;
  float acc = 0.0f;
  for (int m = 0; m < M; ++m) acc += part[((size_t)b * M + m) * 2 + i];
  out[idx] = acc;
}

// Batched 2-D FFT utility (torch default norms).
template <int N>
__global__ __launch_bounds__(Fft2<N>::T) void k_fft2(const cf* in, cf* out, int inverse) {
  __shared__ cf lds[Fft2<N>::LDS_TOTAL];
  using F = Fft2<N>;
  F f(threadIdx.x);
  f.init_twiddles(lds, threadIdx.x);
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * N * N;
  cf v[F::PTS];
  if (!inverse) {
#pragma unroll
    for (int e = 0; e < F::PTS; ++e) v[e] = in[base + f.p1_off(e)];
    f.fwd(v, lds);
#pragma unroll
    for (int e = 0; e < F::PTS; ++e) out[base + f.p3_off(e)] = v[e];
  } else {
#pragma unroll
    for (int e = 0; e < F::PTS; ++e) v[e] = in[base + f.p3_off(e)];
    f.inv(v, lds);
    const float s = 1.0f / (float)(N * N);
#pragma unroll
    for (int e = 0; e < F::PTS; ++e) out[base + f.p1_off(e)] = cscale(v[e], s);
  }
}

}  // namespace qem

// =============================================================================================
// C ABI
// =============================================================================================
namespace {
thread_local char g_err[512] = "";
}  // namespace

int qem::set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void qem::clear_error() { g_err[0] = 0; }

int qem::check_last_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return qem::set_error(QEM_ERR_LAUNCH, "%s launch failed: %s", what, hipGetErrorString(e));
  return QEM_OK;
}

namespace {

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wformat-security"
template <typename... A>
int fail(int code, const char* fmt, A... args) {
  return qem::set_error(code, fmt, args...);
}
#pragma clang diagnostic pop

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Optional event pairs recorded around the dominant kernels (k_forward, k_adjoint) for timing.
struct Profile {
  bool on = false;
  int cap = 0;
  int nf = 0, na = 0;
  hipEvent_t* fwd = nullptr;  // 2 * cap events
  hipEvent_t* adj = nullptr;
};
Profile g_prof;

void prof_mark(bool adjoint_kernel, int edge, hipStream_t st) {
  if (!g_prof.on) return;
  int& n = adjoint_kernel ? g_prof.na : g_prof.nf;
  hipEvent_t* ev = adjoint_kernel ? g_prof.adj : g_prof.fwd;
  if (n >= g_prof.cap) return;
  hipEventRecord(ev[2 * n + edge], st);
  if (edge == 1) ++n;
}

struct Layout {
  size_t psi, ckpt, scratch, X, gI, loss_part, pk, gsum, gpos, gdesc, ramps, pspec, pfac, pfac_ok, psc, total;
  bool pfactors;  // separable-propagator factors checked and kept (N = 128, 256, S > 1, no QEM_FLAG_FULL_PROPAGATORS)
  bool pgrad;  // propagator-gradient spectra kept (QEM_FLAG_GRAD_PROPAGATORS with stored waves)
  int nloss;
  int wslots;  // wave slots per item in psi
  int ck, nck; // checkpoint interval (0 = every incident wave is stored) and checkpoints per item
};

// Pattern columns (roi_cols, 0 = square) and whether the generic-size kernels run the descriptor.
int cols_of(const qem_ptycho_desc& d) { return d.roi_cols != 0 ? d.roi_cols : d.roi; }
bool fixed_size(const qem_ptycho_desc& d) {
  return cols_of(d) == d.roi && (d.roi == 32 || d.roi == 64 || d.roi == 128 || d.roi == 256);
}

Layout layout_of(const qem_ptycho_desc& d) {
  Layout L;
  const size_t N2 = (size_t)d.roi * cols_of(d);
  const bool two_pass = d.roi == 256 || !fixed_size(d);  // chains with a per-item row-transformed wave
  const size_t BM = (size_t)d.batch * d.num_modes;
  const int S = d.num_slices;
  const bool store = (d.flags & QEM_FLAG_STORE_WAVES) != 0;
  L.ck = (store && d.segment_slices >= 2 && d.segment_slices < S) ? d.segment_slices : 0;
  L.nck = L.ck ? (S + L.ck - 1) / L.ck : 0;
  // stored waves: all S slices, or one segment of ck slices (recomputed from a checkpoint)
  L.wslots = store ? (L.ck ? L.ck : S) : 0;
  size_t o = 0;
  L.psi = o;
  o += align_up(BM * L.wslots * N2 * 8);
  L.ckpt = o;
  o += align_up(BM * L.nck * N2 * 8);
  L.scratch = o;
  if (two_pass) o += align_up(BM * N2 * 8);  // the row-transformed wave of each item
  L.X = o;
  o += align_up(BM * N2 * 8);
  L.gI = o;
  o += align_up((size_t)d.batch * N2 * 4);
  L.nloss = (int)(((size_t)d.batch * N2 + 255) / 256);
  L.loss_part = o;
  o += align_up((size_t)L.nloss * 8);
  L.pk = o;
  o += align_up((size_t)d.num_modes * N2 * 8);
  L.gsum = o;
  o += align_up((size_t)d.num_modes * N2 * 8);
  L.gpos = o;
  o += align_up(BM * 2 * 4);
  L.gdesc = o;
  o += align_up(BM * 2 * 4);
  L.ramps = o;
  o += align_up((size_t)d.batch * (d.roi + cols_of(d)) * 8);
  L.pgrad = store && (d.flags & QEM_FLAG_GRAD_PROPAGATORS) && S > 1;
  L.pspec = o;
  if (L.pgrad) o += align_up(BM * (size_t)(S - 1) * N2 * 8);
  L.psc = o;
  if (fixed_size(d) && d.roi <= 128 && S > 1) o += align_up((size_t)(S - 1) * N2 * 8);
  L.pfactors = fixed_size(d) && (d.roi == 256 || d.roi == 128) && S > 1 && !(d.flags & QEM_FLAG_FULL_PROPAGATORS);
  L.pfac = o;
  if (L.pfactors) o += align_up((size_t)(S - 1) * 2 * d.roi * 8);
  L.pfac_ok = o;
  if (L.pfactors) o += align_up((size_t)(S - 1) * 8 * 4);  // eight verdicts per slice
  L.total = o;
  return L;
}

// The translation units that share qem::Params agree on its layout (see params_signature_local).
int check_params_layout() {
  static const bool same = qem::params_signature_local() == qem::params_signature_256() &&
                           qem::params_signature_local() == qem::params_signature_gen();
  if (!same)
    return fail(QEM_ERR_LAUNCH,
                "library build error: qem::Params has a different layout in ptycho.hip, ptycho256.hip and "
                "ptycho_gen.hip (translation units compiled from different headers); rebuild the library");
  return QEM_OK;
}

int check_desc(const qem_ptycho_desc* d) {
  if (int rc = check_params_layout()) return rc;
  if (d == nullptr) return fail(QEM_ERR_NULL, "desc is NULL");
  if (d->roi_cols < 0) return fail(QEM_ERR_ARG, "roi_cols must be >= 0 (got %d)", d->roi_cols);
  const int nc = cols_of(*d);
  if (!fixed_size(*d) && !qem::gen_supported(d->roi, nc))
    return fail(QEM_ERR_UNSUPPORTED,
                "pattern shape (%d, %d) not supported: square 32, 64, 128 or 256, or 2..512 per axis with "
                "prime factors 2, 3, 5 and 7 only",
                d->roi, nc);
  if (d->num_slices < 1) return fail(QEM_ERR_ARG, "num_slices must be >= 1 (got %d)", d->num_slices);
  if (d->num_modes < 1) return fail(QEM_ERR_ARG, "num_modes must be >= 1 (got %d)", d->num_modes);
  if (d->batch < 1) return fail(QEM_ERR_ARG, "batch must be >= 1 (got %d)", d->batch);
  if (d->obj_rows < d->roi || d->obj_cols < nc)
    return fail(QEM_ERR_ARG, "object (%d, %d) must be at least the pattern (%d, %d) in each dimension", d->obj_rows,
                d->obj_cols, d->roi, nc);
  if ((size_t)d->num_slices * d->obj_rows * d->obj_cols >= ((size_t)1 << 31) ||
      (size_t)d->obj_rows * d->obj_cols > ((size_t)1 << 28))
    return fail(QEM_ERR_ARG, "object has too many elements (at most 2^28 per slice, 2^31 in total)");
  if ((size_t)d->batch * d->num_modes * d->roi * nc >= ((size_t)1 << 31))
    return fail(QEM_ERR_ARG, "batch * num_modes * Nr * Nc must stay below 2^31 (batch %d, %d modes, pattern %d x %d)",
                d->batch, d->num_modes, d->roi, nc);
  if (d->num_gpts < 1) return fail(QEM_ERR_ARG, "num_gpts must be >= 1 (got %d)", d->num_gpts);
  if (d->segment_slices < 0) return fail(QEM_ERR_ARG, "segment_slices must be >= 0 (got %d)", d->segment_slices);
  if ((d->flags & QEM_FLAG_GRAD_PROPAGATORS) && d->segment_slices >= 2 && d->segment_slices < d->num_slices)
    return fail(QEM_ERR_UNSUPPORTED, "the propagator gradient needs every incident wave stored (segment_slices 0)");
  if (d->loss_type < QEM_LOSS_L2_AMPLITUDE || d->loss_type > QEM_LOSS_POISSON)
    return fail(QEM_ERR_ARG, "unknown loss_type %d", d->loss_type);
  if (!(d->mean_intensity > 0.0f) || !isfinite(d->mean_intensity))
    return fail(QEM_ERR_ARG, "mean_intensity must be finite and > 0");
  return QEM_OK;
}

int check_ptr(const void* ptr, const char* name, bool complex_) {
  if (ptr == nullptr) return fail(QEM_ERR_NULL, "%s is NULL", name);
  if (((uintptr_t)ptr) & (complex_ ? 7u : 3u)) return fail(QEM_ERR_ALIGN, "%s is misaligned", name);
  return QEM_OK;
}

int check_launch(const char* what) { return qem::check_last_launch(what); }

qem::Params make_params(const qem_ptycho_desc& d, const Layout& L, void* ws) {
  qem::Params p;
  memset(&p, 0, sizeof(p));
  p.S = d.num_slices;
  p.M = d.num_modes;
  p.B = d.batch;
  p.H = d.obj_rows;
  p.W = d.obj_cols;
  p.NR = d.roi;
  p.NC = cols_of(d);
  p.P = d.num_gpts;
  p.loss_type = d.loss_type;
  p.mean_I = d.mean_intensity;
  p.flags = d.flags;
  char* w = (char*)ws;
  p.psi = L.wslots ? (qem::cf*)(w + L.psi) : nullptr;
  p.ckpt = L.nck ? (qem::cf*)(w + L.ckpt) : nullptr;
  p.wslots = L.wslots;
  p.ck = L.ck;
  p.nck = L.nck;
  p.s0 = 0;
  p.s1 = d.num_slices;
  p.fwd_mode = !(d.flags & QEM_FLAG_STORE_WAVES) ? qem::FWD_NO_STORE
                                                 : (L.ck ? qem::FWD_CHECKPOINT : qem::FWD_STORE_ALL);
  p.scratch = (d.roi == 256 || !fixed_size(d)) ? (qem::cf*)(w + L.scratch) : nullptr;
  p.X = (qem::cf*)(w + L.X);
  p.gI = (float*)(w + L.gI);
  p.loss_part = (double*)(w + L.loss_part);
  p.pk = (qem::cf*)(w + L.pk);
  p.gsum = (qem::cf*)(w + L.gsum);
  p.ramps = (qem::cf*)(w + L.ramps);
  p.gpos_part = (float*)(w + L.gpos);
  p.gdesc_part = (float*)(w + L.gdesc);
  p.pspec = L.pgrad ? (qem::cf*)(w + L.pspec) : nullptr;
  p.pfac = L.pfactors ? (qem::cf*)(w + L.pfac) : nullptr;
  p.pfac_ok = L.pfactors ? (int*)(w + L.pfac_ok) : nullptr;
  p.psc = (fixed_size(d) && d.roi <= 128 && d.num_slices > 1) ? (qem::cf*)(w + L.psc) : nullptr;
  return p;
}

template <int N>
int launch_forward(const qem::Params& p, const Layout& L, hipStream_t st) {
  using F = qem::Fft2<N>;
  // M probe-spectrum blocks, then the blocks of the (B, 2, N) shift-ramp table
  const int ramp_blocks = (p.B * 2 * N + F::T - 1) / F::T;
  // and the blocks that scale the propagators (at most 256, grid-stride)
  const int n_prop = p.psc != nullptr ? (p.S - 1) * N * N : 0;
  const int prop_blocks = std::min(256, (n_prop + F::T - 1) / F::T);
  // and, at N = 128, the blocks that form and check the propagators' separable factors
  const int fac_blocks = (N == 128 && p.pfac != nullptr) ? (p.S - 1) * qem::kFacParts128 : 0;
  hipLaunchKernelGGL(qem::k_probe_spectrum<N>, dim3(p.M + ramp_blocks + fac_blocks + prop_blocks), dim3(F::T), 0,
                     st, p);
  if (int rc = check_launch("probe_spectrum")) return rc;
  prof_mark(false, 0, st);
  // every incident wave kept: the store-every-slice instantiation (no branch around the stores)
  const bool all = p.fwd_mode == qem::FWD_STORE_ALL;
  if (p.pspec != nullptr && p.descan != nullptr)
    hipLaunchKernelGGL((qem::k_forward<N, true, true, true>), dim3(p.B * p.M), dim3(F::T), 0, st, p);
  else if (p.pspec != nullptr)
    hipLaunchKernelGGL((qem::k_forward<N, false, true, true>), dim3(p.B * p.M), dim3(F::T), 0, st, p);
  else if (p.descan != nullptr && all)
    hipLaunchKernelGGL((qem::k_forward<N, true, true>), dim3(p.B * p.M), dim3(F::T), 0, st, p);
  else if (p.descan != nullptr)
    hipLaunchKernelGGL((qem::k_forward<N, true, false>), dim3(p.B * p.M), dim3(F::T), 0, st, p);
  else if (all)
    hipLaunchKernelGGL((qem::k_forward<N, false, true>), dim3(p.B * p.M), dim3(F::T), 0, st, p);
  else
    hipLaunchKernelGGL((qem::k_forward<N, false, false>), dim3(p.B * p.M), dim3(F::T), 0, st, p);
  prof_mark(false, 1, st);
  if (int rc = check_launch("forward")) return rc;
  const int nparts = (L.nloss + qem::kLossTiles - 1) / qem::kLossTiles;
  if (p.M >= 4)
    hipLaunchKernelGGL((qem::k_loss<N, 4>), dim3(nparts), dim3(256), 0, st, p, L.nloss);
  else if (p.M >= 2)
    hipLaunchKernelGGL((qem::k_loss<N, 2>), dim3(nparts), dim3(256), 0, st, p, L.nloss);
  else
    hipLaunchKernelGGL((qem::k_loss<N, 1>), dim3(nparts), dim3(256), 0, st, p, L.nloss);
  if (int rc = check_launch("loss")) return rc;
  hipLaunchKernelGGL(qem::k_loss_finalize, dim3(1), dim3(256), 0, st, p, nparts);
  return check_launch("loss_finalize");
}

int launch_forward256(const qem::Params& p, const Layout& L, hipStream_t st) {
  prof_mark(false, 0, st);
  if (int rc = qem::launch_chain256_forward(p, st)) return rc;
  prof_mark(false, 1, st);
  const int nparts = (L.nloss + qem::kLossTiles - 1) / qem::kLossTiles;
  if (p.M >= 4)
    hipLaunchKernelGGL((qem::k_loss<256, 4>), dim3(nparts), dim3(256), 0, st, p, L.nloss);
  else if (p.M >= 2)
    hipLaunchKernelGGL((qem::k_loss<256, 2>), dim3(nparts), dim3(256), 0, st, p, L.nloss);
  else
    hipLaunchKernelGGL((qem::k_loss<256, 1>), dim3(nparts), dim3(256), 0, st, p, L.nloss);
  if (int rc = check_launch("loss")) return rc;
  hipLaunchKernelGGL(qem::k_loss_finalize, dim3(1), dim3(256), 0, st, p, nparts);
  return check_launch("loss_finalize");
}

int launch_forward_gen(const qem::Params& p, const Layout& L, hipStream_t st) {
  prof_mark(false, 0, st);
  if (int rc = qem::launch_gen_forward(p, st)) return rc;
  prof_mark(false, 1, st);
  if (int rc = qem::launch_gen_loss(p, L.nloss, st)) return rc;
  hipLaunchKernelGGL(qem::k_loss_finalize, dim3(1), dim3(256), 0, st, p, qem::gen_loss_parts(L.nloss));
  return check_launch("loss_finalize");
}

int launch_reductions(const qem::Params& p, hipStream_t st);

// Slice segments of the adjoint, last first: [j ck, min(S, (j + 1) ck)) for j = nck-1 .. 0, or one
// segment [0, S) when every incident wave is stored.
template <typename Fn>
int for_each_segment(const qem::Params& p, Fn&& fn) {
  if (p.ck == 0) return fn(p);
  for (int j = p.nck - 1; j >= 0; --j) {
    qem::Params q = p;
    q.s0 = j * p.ck;
    q.s1 = (j + 1) * p.ck < p.S ? (j + 1) * p.ck : p.S;
    q.fwd_mode = qem::FWD_RECOMPUTE;
    if (int rc = fn(q)) return rc;
  }
  return QEM_OK;
}

template <int N>
int launch_obj_reduce(const qem::Params& q, hipStream_t st) {
  const size_t row_bytes = (size_t)q.W * sizeof(qem::cf);
  if (row_bytes <= 48 * 1024) {
    const dim3 g(q.H, q.s1 - q.s0);
    if (q.M >= 4)
      hipLaunchKernelGGL((qem::k_obj_reduce_rows<N, 4>), g, dim3(N), row_bytes, st, q);
    else if (q.M >= 2)
      hipLaunchKernelGGL((qem::k_obj_reduce_rows<N, 2>), g, dim3(N), row_bytes, st, q);
    else
      hipLaunchKernelGGL((qem::k_obj_reduce_rows<N, 1>), g, dim3(N), row_bytes, st, q);
  } else {
    dim3 g((q.W + 31) / 32, (q.H + 31) / 32, q.s1 - q.s0);
    hipLaunchKernelGGL(qem::k_obj_reduce<N>, g, dim3(256), 0, st, q);
  }
  return check_launch("obj_reduce");
}

int launch_adjoint256(const qem::Params& p, hipStream_t st) {
  // with checkpoints, the timed region covers the recomputation and the per-segment object reduce
  prof_mark(true, 0, st);
  int rc = for_each_segment(p, [&](const qem::Params& q) {
    if (q.fwd_mode == qem::FWD_RECOMPUTE)
      if (int r = qem::launch_forward256_kernel(q, st)) return r;
    if (int r = qem::launch_chain256_adjoint(q, st)) return r;
    return q.ck ? launch_obj_reduce<256>(q, st) : QEM_OK;
  });
  prof_mark(true, 1, st);
  if (rc) return rc;
  if (p.ck == 0)
    if (int r = launch_obj_reduce<256>(p, st)) return r;
  hipLaunchKernelGGL(qem::k_probe_reduce<256>, dim3(p.M * 256 * 256 / 128), dim3(256), 0, st, p);
  if (int rc = check_launch("probe_reduce")) return rc;
  if (int rc = qem::launch_probe_grad256(p, st)) return rc;
  return launch_reductions(p, st);
}

int launch_adjoint_gen(const qem::Params& p, hipStream_t st) {
  prof_mark(true, 0, st);
  int rc = for_each_segment(p, [&](const qem::Params& q) {
    if (q.fwd_mode == qem::FWD_RECOMPUTE)
      if (int r = qem::launch_gen_forward_kernel(q, st)) return r;
    if (int r = qem::launch_gen_adjoint(q, st)) return r;
    return q.ck ? qem::launch_gen_obj_reduce(q, st) : QEM_OK;
  });
  prof_mark(true, 1, st);
  if (rc) return rc;
  if (p.ck == 0)
    if (int r = qem::launch_gen_obj_reduce(p, st)) return r;
  if (int r = qem::launch_gen_probe_grad(p, st)) return r;
  return launch_reductions(p, st);
}

int launch_reductions(const qem::Params& p, hipStream_t st) {
  if (p.grad_pos != nullptr) {
    hipLaunchKernelGGL(qem::k_pair_reduce, dim3((2 * p.B + 255) / 256), dim3(256), 0, st, p.gpos_part, p.grad_pos,
                       p.B, p.M);
    if (int rc = check_launch("position_reduce")) return rc;
  }
  if (p.grad_descan != nullptr) {
    hipLaunchKernelGGL(qem::k_pair_reduce, dim3((2 * p.B + 255) / 256), dim3(256), 0, st, p.gdesc_part,
                       p.grad_descan, p.B, p.M);
    if (int rc = check_launch("descan_reduce")) return rc;
  }
  return QEM_OK;
}

// The adjoint without descan or propagator gradient (at N = 128 it picks the separable-factor path
// itself, by the factor verdicts).
template <int N>
int launch_adjoint_plain(const qem::Params& q, hipStream_t st) {
  using F = qem::Fft2<N>;
  hipLaunchKernelGGL((qem::k_adjoint<N, false>), dim3(q.B * q.M), dim3(F::T), 0, st, q);
  return check_launch("adjoint");
}

template <int N>
int launch_adjoint(const qem::Params& p, hipStream_t st) {
  using F = qem::Fft2<N>;
  // with checkpoints, the timed region covers the recomputation and the per-segment object reduce
  prof_mark(true, 0, st);
  int rc = for_each_segment(p, [&](const qem::Params& q) {
    if (q.fwd_mode == qem::FWD_RECOMPUTE) {
      hipLaunchKernelGGL((qem::k_forward<N, false, true>), dim3(q.B * q.M), dim3(F::T), 0, st, q);
      if (int r = check_launch("forward (recompute)")) return r;
    }
    if (q.pspec == nullptr && q.descan == nullptr) {
      if (int r = launch_adjoint_plain<N>(q, st)) return r;
    } else {
      if (q.pspec != nullptr && q.descan != nullptr)
        hipLaunchKernelGGL((qem::k_adjoint<N, true, true>), dim3(q.B * q.M), dim3(F::T), 0, st, q);
      else if (q.pspec != nullptr)
        hipLaunchKernelGGL((qem::k_adjoint<N, false, true>), dim3(q.B * q.M), dim3(F::T), 0, st, q);
      else
        hipLaunchKernelGGL((qem::k_adjoint<N, true>), dim3(q.B * q.M), dim3(F::T), 0, st, q);
      if (int r = check_launch("adjoint")) return r;
    }
    return q.ck ? launch_obj_reduce<N>(q, st) : QEM_OK;
  });
  prof_mark(true, 1, st);
  if (rc) return rc;
  if (p.ck == 0)
    if (int r = launch_obj_reduce<N>(p, st)) return r;
  hipLaunchKernelGGL(qem::k_probe_reduce<N>, dim3(p.M * N * N / 128), dim3(256), 0, st, p);
  if (int rc = check_launch("probe_reduce")) return rc;
  hipLaunchKernelGGL(qem::k_probe_grad<N>, dim3(p.M), dim3(F::T), 0, st, p);
  if (int rc = check_launch("probe_grad")) return rc;
  if (p.grad_pos != nullptr) {
    hipLaunchKernelGGL(qem::k_pair_reduce, dim3((2 * p.B + 255) / 256), dim3(256), 0, st, p.gpos_part, p.grad_pos,
                       p.B, p.M);
    if (int rc = check_launch("position_reduce")) return rc;
  }
  if (p.grad_descan != nullptr) {
    hipLaunchKernelGGL(qem::k_pair_reduce, dim3((2 * p.B + 255) / 256), dim3(256), 0, st, p.gdesc_part,
                       p.grad_descan, p.B, p.M);
    if (int rc = check_launch("descan_reduce")) return rc;
  }
  return QEM_OK;
}

template <int N>
int launch_projection_update(const qem::Params& p, const Layout& L, float* overlap, float* energy, hipStream_t st) {
  using F = qem::Fft2<N>;
  hipLaunchKernelGGL(qem::k_projection<N>, dim3(L.nloss), dim3(256), 0, st, p);
  if (int rc = check_launch("projection")) return rc;
  if (overlap != nullptr) {
    dim3 g((p.W + 31) / 32, (p.H + 31) / 32, p.S * p.M);
    hipLaunchKernelGGL(qem::k_overlap<N>, g, dim3(256), 0, st, p, overlap);
    if (int rc = check_launch("overlap")) return rc;
  }
  if (energy != nullptr) {
    hipLaunchKernelGGL(qem::k_patch_energy<N>, dim3(p.S * p.B), dim3(256), 0, st, p, energy);
    if (int rc = check_launch("patch_energy")) return rc;
  }
  if (int rc = launch_adjoint_plain<N>(p, st)) return rc;
  if (int rc = launch_obj_reduce<N>(p, st)) return rc;
  hipLaunchKernelGGL(qem::k_probe_reduce<N>, dim3(p.M * N * N / 128), dim3(256), 0, st, p);
  if (int rc = check_launch("probe_reduce")) return rc;
  hipLaunchKernelGGL(qem::k_probe_grad<N>, dim3(p.M), dim3(F::T), 0, st, p);
  return check_launch("probe_grad");
}

int launch_projection_update256(const qem::Params& p, const Layout& L, float* overlap, float* energy,
                                hipStream_t st) {
  hipLaunchKernelGGL(qem::k_projection<256>, dim3(L.nloss), dim3(256), 0, st, p);
  if (int rc = check_launch("projection")) return rc;
  if (overlap != nullptr) {
    dim3 g((p.W + 31) / 32, (p.H + 31) / 32, p.S * p.M);
    hipLaunchKernelGGL(qem::k_overlap<256>, g, dim3(256), 0, st, p, overlap);
    if (int rc = check_launch("overlap")) return rc;
  }
  if (energy != nullptr) {
    hipLaunchKernelGGL(qem::k_patch_energy<256>, dim3(p.S * p.B), dim3(256), 0, st, p, energy);
    if (int rc = check_launch("patch_energy")) return rc;
  }
  if (int rc = qem::launch_chain256_adjoint(p, st)) return rc;
  if (int rc = launch_obj_reduce<256>(p, st)) return rc;
  hipLaunchKernelGGL(qem::k_probe_reduce<256>, dim3(p.M * 256 * 256 / 128), dim3(256), 0, st, p);
  if (int rc = check_launch("probe_reduce")) return rc;
  return qem::launch_probe_grad256(p, st);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Workspace of the allocation-free forward (qem_ptycho_forward_loss_static): a file-scope device array,
// one per GPU the library runs on, held until the process ends.  In host code `scratch` names a
// host-side shadow of the array, not its device address: kernels get the address hipGetSymbolAddress
// returns, looked up once per device (with that device current) on the first call there.
static __device__ __align__(256) unsigned char scratch[QEM_STATIC_WORKSPACE_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's `scratch`.
int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(QEM_ERR_LAUNCH, "hipGetDevice failed: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(QEM_ERR_UNSUPPORTED, "device %d: the static workspace supports devices 0..%d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return fail(QEM_ERR_LAUNCH, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return QEM_OK;
}

int check_forward_ptrs(const qem_ptycho_desc& d, const void* obj, const void* probe, const float* positions,
                       const void* propagators, const float* targets, const float* mask, const float* descan,
                       float* loss, float* intensities);
int forward_loss_checked(const qem_ptycho_desc& d, const void* obj, const void* probe, const float* positions,
                         const void* propagators, const float* targets, const float* mask, const float* descan,
                         float* loss, float* intensities, void* workspace, size_t workspace_bytes, void* stream);

}  // namespace

extern "C" {

int32_t qem_abi_version(void) { return 104; }

int qem_profile_begin(int32_t max_launches) {
  qem::clear_error();
  if (g_prof.on) return fail(QEM_ERR_ARG, "profiling already active");
  if (max_launches < 1 || max_launches > 100000) return fail(QEM_ERR_ARG, "max_launches out of range");
  g_prof.cap = max_launches;
  g_prof.nf = g_prof.na = 0;
  g_prof.fwd = new hipEvent_t[2 * max_launches];
  g_prof.adj = new hipEvent_t[2 * max_launches];
  for (int i = 0; i < 2 * max_launches; ++i) {
    if (hipEventCreate(&g_prof.fwd[i]) != hipSuccess || hipEventCreate(&g_prof.adj[i]) != hipSuccess)
      return fail(QEM_ERR_LAUNCH, "hipEventCreate failed");
  }
  g_prof.on = true;
  return QEM_OK;
}

int qem_profile_end(float* fwd_ms, int32_t* n_fwd, float* adj_ms, int32_t* n_adj) {
  qem::clear_error();
  if (!g_prof.on) return fail(QEM_ERR_ARG, "profiling not active");
  float tf = 0.f, ta = 0.f;
  int rc = QEM_OK;
  for (int i = 0; i < g_prof.nf && rc == QEM_OK; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(g_prof.fwd[2 * i + 1]) != hipSuccess ||
        hipEventElapsedTime(&ms, g_prof.fwd[2 * i], g_prof.fwd[2 * i + 1]) != hipSuccess)
      rc = fail(QEM_ERR_LAUNCH, "event timing failed");
    tf += ms;
  }
  for (int i = 0; i < g_prof.na && rc == QEM_OK; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(g_prof.adj[2 * i + 1]) != hipSuccess ||
        hipEventElapsedTime(&ms, g_prof.adj[2 * i], g_prof.adj[2 * i + 1]) != hipSuccess)
      rc = fail(QEM_ERR_LAUNCH, "event timing failed");
    ta += ms;
  }
  if (fwd_ms) *fwd_ms = tf;
  if (adj_ms) *adj_ms = ta;
  if (n_fwd) *n_fwd = g_prof.nf;
  if (n_adj) *n_adj = g_prof.na;
  for (int i = 0; i < 2 * g_prof.cap; ++i) {
    hipEventDestroy(g_prof.fwd[i]);
    hipEventDestroy(g_prof.adj[i]);
  }
  delete[] g_prof.fwd;
  delete[] g_prof.adj;
  g_prof = Profile();
  return rc;
}

const char* qem_last_error(void) { return g_err; }

int qem_ptycho_workspace_bytes(const qem_ptycho_desc* desc, size_t* bytes) {
  qem::clear_error();
  if (int rc = check_desc(desc)) return rc;
  if (bytes == nullptr) return fail(QEM_ERR_NULL, "bytes is NULL");
  *bytes = layout_of(*desc).total;
  return QEM_OK;
}

int qem_ptycho_forward_loss(const qem_ptycho_desc* desc, const void* obj, const void* probe, const float* positions,
                            const void* propagators, const float* targets, const float* mask, const float* descan,
                            float* loss, float* intensities, void* workspace, size_t workspace_bytes, void* stream) {
  qem::clear_error();
  if (int rc = check_desc(desc)) return rc;
  return forward_loss_checked(*desc, obj, probe, positions, propagators, targets, mask, descan, loss, intensities,
                              workspace, workspace_bytes, stream);
}

int qem_ptycho_forward_loss_static(const qem_ptycho_desc* desc, const void* obj, const void* probe,
                                   const float* positions, const void* propagators, const float* targets,
                                   const float* mask, const float* descan, float* loss, float* intensities,
                                   void* stream) {
  qem::clear_error();
  if (int rc = check_desc(desc)) return rc;
  if (desc->flags & ~(uint32_t)QEM_FLAG_FULL_PROPAGATORS)
    return fail(QEM_ERR_ARG,
                "the static-workspace forward runs the no-gradient forward only (flags must be 0 or "
                "QEM_FLAG_FULL_PROPAGATORS)");
  const size_t need = layout_of(*desc).total;
  if (need > (size_t)QEM_STATIC_WORKSPACE_BYTES)
    return fail(QEM_ERR_WORKSPACE,
                "this call needs %zu bytes of workspace, more than the %zu of the static workspace; split the batch or "
                "use qem_ptycho_forward_loss with a workspace",
                need, (size_t)QEM_STATIC_WORKSPACE_BYTES);
  if (int rc = check_forward_ptrs(*desc, obj, probe, positions, propagators, targets, mask, descan, loss, intensities))
    return rc;
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  return forward_loss_checked(*desc, obj, probe, positions, propagators, targets, mask, descan, loss, intensities, ws,
                              (size_t)QEM_STATIC_WORKSPACE_BYTES, stream);
}

}  // extern "C"

namespace {

int check_forward_ptrs(const qem_ptycho_desc& d, const void* obj, const void* probe, const float* positions,
                       const void* propagators, const float* targets, const float* mask, const float* descan,
                       float* loss, float* intensities) {
  if (int rc = check_ptr(obj, "obj", true)) return rc;
  if (int rc = check_ptr(probe, "probe", true)) return rc;
  if (int rc = check_ptr(positions, "positions", false)) return rc;
  if (d.num_slices > 1)
    if (int rc = check_ptr(propagators, "propagators", true)) return rc;
  if (int rc = check_ptr(targets, "targets", false)) return rc;
  if (d.loss_type != QEM_LOSS_POISSON)
    if (int rc = check_ptr(mask, "mask", false)) return rc;
  if (descan != nullptr)
    if (int rc = check_ptr(descan, "descan", false)) return rc;
  if (int rc = check_ptr(loss, "loss", false)) return rc;
  if (intensities != nullptr)
    if (int rc = check_ptr(intensities, "intensities", false)) return rc;
  return QEM_OK;
}

int forward_loss_checked(const qem_ptycho_desc& d, const void* obj, const void* probe, const float* positions,
                         const void* propagators, const float* targets, const float* mask, const float* descan,
                         float* loss, float* intensities, void* workspace, size_t workspace_bytes, void* stream) {
  if (int rc = check_forward_ptrs(d, obj, probe, positions, propagators, targets, mask, descan, loss, intensities))
    return rc;
  if (int rc = check_ptr(workspace, "workspace", true)) return rc;
  const Layout L = layout_of(d);
  if (workspace_bytes < L.total)
    return fail(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", workspace_bytes, L.total);
  qem::Params p = make_params(d, L, workspace);
  p.obj = (const qem::cf*)obj;
  p.probe = (const qem::cf*)probe;
  p.pos = positions;
  p.prop = (const qem::cf*)propagators;
  p.targets = targets;
  p.mask = mask;
  p.descan = descan;
  p.loss_out = loss;
  p.intens_out = intensities;
  hipStream_t st = (hipStream_t)stream;
  if (!fixed_size(d)) return launch_forward_gen(p, L, st);
  switch (d.roi) {
    case 32: return launch_forward<32>(p, L, st);
    case 64: return launch_forward<64>(p, L, st);
    case 128: return launch_forward<128>(p, L, st);
    default: return launch_forward256(p, L, st);
  }
}

}  // namespace

extern "C" {

int qem_ptycho_adjoint(const qem_ptycho_desc* desc, const void* obj, const void* probe, const float* positions,
                       const void* propagators, const float* descan, const float* dloss, void* grad_obj,
                       void* grad_probe, float* grad_positions, float* grad_descan, void* workspace,
                       size_t workspace_bytes, void* stream) {
  qem::clear_error();
  if (int rc = check_desc(desc)) return rc;
  const qem_ptycho_desc& d = *desc;
  if (!(d.flags & QEM_FLAG_STORE_WAVES))
    return fail(QEM_ERR_ARG, "adjoint needs a forward pass run with QEM_FLAG_STORE_WAVES");
  if (int rc = check_ptr(obj, "obj", true)) return rc;
  if (int rc = check_ptr(probe, "probe", true)) return rc;
  if (int rc = check_ptr(positions, "positions", false)) return rc;
  if (d.num_slices > 1)
    if (int rc = check_ptr(propagators, "propagators", true)) return rc;
  if (descan != nullptr)
    if (int rc = check_ptr(descan, "descan", false)) return rc;
  if (int rc = check_ptr(dloss, "dloss", false)) return rc;
  if (int rc = check_ptr(grad_obj, "grad_obj", true)) return rc;
  if (int rc = check_ptr(grad_probe, "grad_probe", true)) return rc;
  if (grad_positions != nullptr) {
    if (!(d.flags & QEM_FLAG_GRAD_POSITIONS))
      return fail(QEM_ERR_ARG, "grad_positions given without QEM_FLAG_GRAD_POSITIONS");
    if (int rc = check_ptr(grad_positions, "grad_positions", false)) return rc;
  }
  if (grad_descan != nullptr) {
    if (!(d.flags & QEM_FLAG_GRAD_DESCAN) || descan == nullptr)
      return fail(QEM_ERR_ARG, "grad_descan needs QEM_FLAG_GRAD_DESCAN and a descan input");
    if (int rc = check_ptr(grad_descan, "grad_descan", false)) return rc;
  }
  if (int rc = check_ptr(workspace, "workspace", true)) return rc;
  const Layout L = layout_of(d);
  if (workspace_bytes < L.total)
    return fail(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", workspace_bytes, L.total);
  qem::Params p = make_params(d, L, workspace);
  p.obj = (const qem::cf*)obj;
  p.probe = (const qem::cf*)probe;
  p.pos = positions;
  p.prop = (const qem::cf*)propagators;
  p.descan = descan;
  p.dloss = dloss;
  p.grad_obj = (qem::cf*)grad_obj;
  p.grad_probe = (qem::cf*)grad_probe;
  p.grad_pos = grad_positions;
  p.grad_descan = grad_descan;
  hipStream_t st = (hipStream_t)stream;
  if (!fixed_size(d)) return launch_adjoint_gen(p, st);
  switch (d.roi) {
    case 32: return launch_adjoint<32>(p, st);
    case 64: return launch_adjoint<64>(p, st);
    case 128: return launch_adjoint<128>(p, st);
    default: return launch_adjoint256(p, st);
  }
}

int qem_ptycho_propagator_gradient(const qem_ptycho_desc* desc, void* grad_propagators, void* workspace,
                                   size_t workspace_bytes, void* stream) {
  qem::clear_error();
  if (int rc = check_desc(desc)) return rc;
  const qem_ptycho_desc& d = *desc;
  const Layout L = layout_of(d);
  if (d.num_slices < 2) return fail(QEM_ERR_ARG, "a single slice has no propagators");
  if (!L.pgrad)
    return fail(QEM_ERR_ARG, "the propagator gradient needs QEM_FLAG_STORE_WAVES | QEM_FLAG_GRAD_PROPAGATORS");
  if (int rc = check_ptr(grad_propagators, "grad_propagators", true)) return rc;
  if (int rc = check_ptr(workspace, "workspace", true)) return rc;
  if (workspace_bytes < L.total)
    return fail(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", workspace_bytes, L.total);
  const qem::Params p = make_params(d, L, workspace);
  hipStream_t st = (hipStream_t)stream;
  qem::cf* g = (qem::cf*)grad_propagators;
  if (!fixed_size(d)) return qem::launch_gen_prop_reduce(p, g, st);
  const int N = d.roi;
  const int blocks = (int)((size_t)(d.num_slices - 1) * N * N / 64);
  switch (N) {
    case 32: hipLaunchKernelGGL(qem::k_prop_reduce<32>, dim3(blocks), dim3(256), 0, st, p, g); break;
    case 64: hipLaunchKernelGGL(qem::k_prop_reduce<64>, dim3(blocks), dim3(256), 0, st, p, g); break;
    case 128: hipLaunchKernelGGL(qem::k_prop_reduce<128>, dim3(blocks), dim3(256), 0, st, p, g); break;
    default: hipLaunchKernelGGL(qem::k_prop_reduce<256>, dim3(blocks), dim3(256), 0, st, p, g); break;
  }
  return check_launch("propagator_gradient");
}

int qem_ptycho_projection_update(const qem_ptycho_desc* desc, const void* obj, const void* probe,
                                 const float* positions, const void* propagators, const float* targets,
                                 void* obj_update, void* probe_update, float* probe_overlap, float* patch_energy,
                                 void* workspace, size_t workspace_bytes, void* stream) {
  qem::clear_error();
  if (int rc = check_desc(desc)) return rc;
  qem_ptycho_desc d = *desc;
  if (!(d.flags & QEM_FLAG_STORE_WAVES))
    return fail(QEM_ERR_ARG, "projection update needs a forward pass run with QEM_FLAG_STORE_WAVES");
  if (layout_of(d).ck != 0)
    return fail(QEM_ERR_UNSUPPORTED, "projection update needs every incident wave stored (segment_slices 0)");
  if (!fixed_size(d))
    return fail(QEM_ERR_UNSUPPORTED, "projection update runs for square 32, 64, 128 and 256 patterns only (got %d x %d)",
                d.roi, cols_of(d));
  if (int rc = check_ptr(obj, "obj", true)) return rc;
  if (int rc = check_ptr(probe, "probe", true)) return rc;
  if (int rc = check_ptr(positions, "positions", false)) return rc;
  if (d.num_slices > 1)
    if (int rc = check_ptr(propagators, "propagators", true)) return rc;
  if (int rc = check_ptr(targets, "targets", false)) return rc;
  if (int rc = check_ptr(obj_update, "obj_update", true)) return rc;
  if (int rc = check_ptr(probe_update, "probe_update", true)) return rc;
  if (probe_overlap != nullptr)
    if (int rc = check_ptr(probe_overlap, "probe_overlap", false)) return rc;
  if (patch_energy != nullptr)
    if (int rc = check_ptr(patch_energy, "patch_energy", false)) return rc;
  if (int rc = check_ptr(workspace, "workspace", true)) return rc;
  const Layout L = layout_of(d);
  if (workspace_bytes < L.total)
    return fail(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", workspace_bytes, L.total);
  d.flags |= QEM_FLAG_ANALYTIC;
  d.flags &= ~(uint32_t)(QEM_FLAG_GRAD_POSITIONS | QEM_FLAG_GRAD_DESCAN);
  qem::Params p = make_params(d, L, workspace);
  p.obj = (const qem::cf*)obj;
  p.probe = (const qem::cf*)probe;
  p.pos = positions;
  p.prop = (const qem::cf*)propagators;
  p.targets = targets;
  p.grad_obj = (qem::cf*)obj_update;
  p.grad_probe = (qem::cf*)probe_update;
  hipStream_t st = (hipStream_t)stream;
  switch (d.roi) {
    case 32: return launch_projection_update<32>(p, L, probe_overlap, patch_energy, st);
    case 64: return launch_projection_update<64>(p, L, probe_overlap, patch_energy, st);
    case 128: return launch_projection_update<128>(p, L, probe_overlap, patch_energy, st);
    default: return launch_projection_update256(p, L, probe_overlap, patch_energy, st);
  }
}

int qem_fft2_c64(int32_t n, int32_t count, int32_t inverse, const void* in, void* out, void* stream) {
  qem::clear_error();
  if (n != 32 && n != 64 && n != 128) return fail(QEM_ERR_UNSUPPORTED, "fft size %d not supported", n);
  if (count < 0) return fail(QEM_ERR_ARG, "count must be >= 0");
  if (count == 0) return QEM_OK;
  if (int rc = check_ptr(in, "in", true)) return rc;
  if (int rc = check_ptr(out, "out", true)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const qem::cf* i = (const qem::cf*)in;
  qem::cf* o = (qem::cf*)out;
  switch (n) {
    case 32: hipLaunchKernelGGL(qem::k_fft2<32>, dim3(count), dim3(qem::Fft2<32>::T), 0, st, i, o, inverse); break;
    case 64: hipLaunchKernelGGL(qem::k_fft2<64>, dim3(count), dim3(qem::Fft2<64>::T), 0, st, i, o, inverse); break;
    default: hipLaunchKernelGGL(qem::k_fft2<128>, dim3(count), dim3(qem::Fft2<128>::T), 0, st, i, o, inverse); break;
  }
  return check_launch("fft2");
}

}  // extern "C"
