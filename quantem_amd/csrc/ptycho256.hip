// ptycho256.hip — multislice forward / adjoint for 256 x 256 patterns (two-pass transforms).
//
// A 256 x 256 complex64 wave (512 KiB) does not fit one workgroup's LDS, so each workgroup owns one
// (pattern, mode) item and streams every 2-D transform through two row passes: a pass reads 16 rows at
// a time, runs length-256 FFTs (16 x 16, one LDS exchange), and writes the result transposed; two
// passes give the 2-D transform in natural order.  The passes ping-pong between two per-item scratch
// buffers (L2 / Infinity-Cache resident in practice), and the multislice elementwise steps
// (transmission, propagator, descan, the adjoint's gradient patches) are fused into the passes' load
// (pre) and store (post) operations.  The chain follows the same reference lines as ptycho.hip.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "ptycho_common.h"

namespace qem {
namespace n256 {

constexpr int N = 256;
constexpr int N2 = N * N;
constexpr int LDX = 17;  // padded row stride (complex64) of the LDS buffers

struct Smem {
  cf ex[16 * 16 * LDX];  // per-row exchange of the 16 x 16 decomposition
  cf tile[N * LDX];      // transposed output staging (N frequencies x 16 rows)
  double red[4];
};

struct Ident {
  __device__ __forceinline__ cf operator()(int, int, int, cf x) const { return x; }
};
struct IdentPost {
  __device__ __forceinline__ cf operator()(int, int, cf x) const { return x; }
};

// out[k][r] = sum_c pre(r, c, i, src[r][c]) W^(+-c k) for every row r (W = exp(-2 pi i / 256), conjugated
// if INV), stored as post(k, r, value).  c = j + 16 i with j = lane % 16 and i the unrolled slot.
template <bool INV, typename Pre, typename Post>
__device__ __forceinline__ void pass(const cf* __restrict__ src, cf* __restrict__ dst, Smem& sm, Pre&& pre,
                                     Post&& post) {
  const int t = threadIdx.x, rr = t >> 4, j = t & 15;
  for (int r0 = 0; r0 < N; r0 += 16) {
    const int r = r0 + rr;
    cf v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = pre(r, j + 16 * i, i, src[r * N + j + 16 * i]);
    Dft<16, INV>::template run<1>(v);  // over i -> k1
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) {
      const cf w = kTw256[(j * k1) & 255];
      v[k1] = INV ? cmulc(v[k1], w) : cmul(v[k1], w);
    }
    __syncthreads();  // the previous chunk is done with sm
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) sm.ex[(rr * 16 + k1) * LDX + j] = v[k1];
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) v[jj] = sm.ex[(rr * 16 + j) * LDX + jj];  // this lane now owns k1 = j
    Dft<16, INV>::template run<1>(v);                                         // over jj -> k2; k = j + 16 k2
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) sm.tile[(j + 16 * k2) * LDX + rr] = v[k2];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 16; ++it) {
      const int k = (t >> 4) + 16 * it, rw = t & 15;
      dst[k * N + r0 + rw] = post(k, r0 + rw, sm.tile[k * LDX + rw]);
    }
  }
  __syncthreads();  // dst complete and visible to the whole workgroup
}

__device__ __forceinline__ cf ramp2(float fr, float fc, float dr, float dc) {
  return cmul(expm2pi(fr * dr), expm2pi(fc * dc));
}

// Object row / column byte offsets (periodic, fftfreq-ordered patch).
__device__ __forceinline__ unsigned obj_row_off(int r0, int r, int H, int W) {
  return (unsigned)(pmod(r0 + (r < N / 2 ? r : r - N), H) * W) * 8u;
}
__device__ __forceinline__ unsigned obj_col_off(int c0, int c, int W) {
  return (unsigned)pmod(c0 + (c < N / 2 ? c : c - N), W) * 8u;
}

__global__ __launch_bounds__(256) void k_spectrum256(Params p) {
  __shared__ Smem sm;
  const int m = blockIdx.x;
  cf* b1 = p.scratch + (size_t)(2 * m) * N2;
  pass<false>(p.probe + (size_t)m * N2, b1, sm, Ident(), IdentPost());
  pass<false>(b1, p.pk + (size_t)m * N2, sm, Ident(), IdentPost());
}

template <bool DESCAN>
__global__ __launch_bounds__(256) void k_forward256(Params p) {
  __shared__ Smem sm;
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);
  cf* b1 = p.scratch + (size_t)(2 * item) * N2;
  cf* b2 = b1 + N2;
  const bool store = (p.flags & QEM_FLAG_STORE_WAVES) != 0;
  auto slot = [&](int s) { return p.psi + ((size_t)item * (store ? p.S : 1) + (store ? s : 0)) * N2; };
  const float inv_n2 = 1.0f / (float)N2;
  const size_t HW = (size_t)p.H * p.W;
  unsigned ocol[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) ocol[i] = obj_col_off(c0, (threadIdx.x & 15) + 16 * i, p.W);

  // psi_0 = ifft2(fft2(P_m) * ramp_b)
  pass<true>(p.pk + (size_t)m * N2, b1, sm,
             [&](int kr, int kc, int, cf x) { return cscale(cmul(x, ramp2(fftfreq1<N>(kr), fftfreq1<N>(kc), dr, dc)), inv_n2); },
             IdentPost());
  pass<true>(b1, slot(0), sm, Ident(), IdentPost());

  for (int s = 0; s < p.S; ++s) {
    const cf* O = p.obj + (size_t)s * HW;
    const cf* A = slot(s);
    auto transmit = [&](int r, int c, int i, cf x) {
      cf y = cmul(obj_at(O, obj_row_off(r0, r, p.H, p.W) + ocol[i]), x);
      if constexpr (DESCAN) y = cmul(y, ramp2(fftfreq1<N>(r), fftfreq1<N>(c), p.descan[2 * b], p.descan[2 * b + 1]));
      return y;
    };
    if (s + 1 < p.S) {
      const cf* pr = p.prop + (size_t)s * N2;
      pass<false>(A, b1, sm, [&](int r, int c, int i, cf x) { return cmul(obj_at(O, obj_row_off(r0, r, p.H, p.W) + ocol[i]), x); },
                  IdentPost());
      pass<false>(b1, b2, sm, Ident(), [&](int kr, int kc, cf v) { return cscale(cmul(v, pr[kr * N + kc]), inv_n2); });
      pass<true>(b2, b1, sm, Ident(), IdentPost());
      pass<true>(b1, slot(s + 1), sm, Ident(), IdentPost());
    } else {
      const float inv_n = 1.0f / (float)N;
      pass<false>(A, b1, sm, transmit, IdentPost());
      pass<false>(b1, p.X + (size_t)item * N2, sm, Ident(), [&](int, int, cf v) { return cscale(v, inv_n); });
    }
  }
}

template <bool DESCAN>
__global__ __launch_bounds__(256) void k_adjoint256(Params p) {
  __shared__ Smem sm;
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);
  cf* b1 = p.scratch + (size_t)(2 * item) * N2;
  cf* b2 = b1 + N2;
  cf* Xs = p.X + (size_t)item * N2;
  const float* gI = p.gI + (size_t)b * N2;
  const float dl = p.dloss[0];
  const float inv_n2 = 1.0f / (float)N2;
  const size_t HW = (size_t)p.H * p.W;
  unsigned ocol[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) ocol[i] = obj_col_off(c0, (threadIdx.x & 15) + 16 * i, p.W);

  // g_x = (1/N) F^H (2 dL/dI X)
  const float sc = 2.0f * dl / (float)N;
  pass<true>(Xs, b1, sm, [&](int kr, int kc, int, cf x) { return cscale(x, sc * gI[kr * N + kc]); }, IdentPost());
  pass<true>(b1, b2, sm, Ident(), IdentPost());

  double dar = 0.0, dac = 0.0;
  for (int s = p.S - 1; s >= 0; --s) {
    const cf* O = p.obj + (size_t)s * HW;
    cf* A = p.psi + ((size_t)item * p.S + s) * N2;
    const bool last = (s == p.S - 1);
    // object-gradient patch conj(psi_s) g (over psi_s) and g_psi = conj(O_s) g, fused into the load
    pass<false>(b2, b1, sm,
                [&](int r, int c, int i, cf g) {
                  const cf o = obj_at(O, obj_row_off(r0, r, p.H, p.W) + ocol[i]);
                  const cf ps = A[r * N + c];
                  if constexpr (DESCAN) {
                    if (last) {
                      const float fr = fftfreq1<N>(r), fc = fftfreq1<N>(c);
                      const cf d = ramp2(fr, fc, p.descan[2 * b], p.descan[2 * b + 1]);
                      const float im = cconjmul(g, cmul(cmul(o, ps), d)).y;
                      dar += (double)(fr * im);
                      dac += (double)(fc * im);
                      g = cmulc(g, d);
                    }
                  }
                  A[r * N + c] = cconjmul(ps, g);
                  return cconjmul(o, g);
                },
                IdentPost());
    if (s > 0) {
      const cf* pr = p.prop + (size_t)(s - 1) * N2;
      pass<false>(b1, b2, sm, Ident(), [&](int kr, int kc, cf v) { return cscale(cmulc(v, pr[kr * N + kc]), inv_n2); });
      pass<true>(b2, b1, sm, Ident(), IdentPost());
      pass<true>(b1, b2, sm, Ident(), IdentPost());
    } else {
      pass<false>(b1, b2, sm, Ident(), IdentPost());  // G = fft2(g_psi_0) in b2
    }
  }
  if constexpr (DESCAN) {
    if (p.flags & QEM_FLAG_GRAD_DESCAN) {
      const double tr = block_sum<256>(dar, sm.red);
      const double tc = block_sum<256>(dac, sm.red);
      if (threadIdx.x == 0) {
        p.gdesc_part[2 * item] = (float)(2.0 * M_PI * tr);
        p.gdesc_part[2 * item + 1] = (float)(2.0 * M_PI * tc);
      }
    }
  }
  // probe-gradient spectrum conj(ramp) G over X, position-gradient partial sums
  const cf* pk = p.pk + (size_t)m * N2;
  double ar = 0.0, ac = 0.0;
  for (int idx = threadIdx.x; idx < N2; idx += 256) {
    const int kr = idx >> 8, kc = idx & (N - 1);
    const float fr = fftfreq1<N>(kr), fc = fftfreq1<N>(kc);
    const cf R = ramp2(fr, fc, dr, dc);
    const cf G = b2[idx];
    Xs[idx] = cmulc(G, R);
    const float im = cconjmul(G, cmul(pk[idx], R)).y;
    ar += (double)(fr * im);
    ac += (double)(fc * im);
  }
  if (p.flags & QEM_FLAG_GRAD_POSITIONS) {
    const double tr = block_sum<256>(ar, sm.red);
    const double tc = block_sum<256>(ac, sm.red);
    if (threadIdx.x == 0) {
      p.gpos_part[2 * item] = (float)(2.0 * M_PI * tr * (double)inv_n2);
      p.gpos_part[2 * item + 1] = (float)(2.0 * M_PI * tc * (double)inv_n2);
    }
  }
}

__global__ __launch_bounds__(256) void k_probe_grad256(Params p) {
  __shared__ Smem sm;
  const int m = blockIdx.x;
  cf* b1 = p.scratch + (size_t)(2 * m) * N2;
  const float inv_n2 = 1.0f / (float)N2;
  pass<true>(p.gsum + (size_t)m * N2, b1, sm, Ident(), IdentPost());
  pass<true>(b1, p.grad_probe + (size_t)m * N2, sm, Ident(), [&](int, int, cf v) { return cscale(v, inv_n2); });
}

}  // namespace n256

int launch_chain256_forward(const Params& p, hipStream_t st) {
  hipLaunchKernelGGL(n256::k_spectrum256, dim3(p.M), dim3(256), 0, st, p);
  if (int rc = check_last_launch("spectrum256")) return rc;
  if (p.descan != nullptr)
    hipLaunchKernelGGL((n256::k_forward256<true>), dim3(p.B * p.M), dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL((n256::k_forward256<false>), dim3(p.B * p.M), dim3(256), 0, st, p);
  return check_last_launch("forward256");
}

int launch_chain256_adjoint(const Params& p, hipStream_t st) {
  if (p.descan != nullptr)
    hipLaunchKernelGGL((n256::k_adjoint256<true>), dim3(p.B * p.M), dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL((n256::k_adjoint256<false>), dim3(p.B * p.M), dim3(256), 0, st, p);
  return check_last_launch("adjoint256");
}

int launch_probe_grad256(const Params& p, hipStream_t st) {
  hipLaunchKernelGGL(n256::k_probe_grad256, dim3(p.M), dim3(256), 0, st, p);
  return check_last_launch("probe_grad256");
}

}  // namespace qem
