// nn.hip — learned warm-start kernels for gfx950: per-pattern features and a 3x3 convolution as a
// bf16-in / fp32-accumulate implicit GEMM on the matrix cores (v_mfma_f32_16x16x16_bf16).
//
// MFMA 16x16x16 operand layout (wave64): lane l holds A[i = l % 16][k = 4*(l/16) + 0..3],
// B[k = 4*(l/16) + 0..3][j = l % 16], and D[i = 4*(l/16) + 0..3][j = l % 16].
//
// Convolution as GEMM: rows = output pixels (n, y, x), columns = output channels, reduction index
// k = (ky*3 + kx)*Cin + ci.  With Cin % 4 == 0 a lane's four consecutive k share one tap, so each A
// fragment is one 8-byte load of four channels of one input pixel (NHWC), or zero in the padding.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "common.h"
#include "qem_ptycho.h"
#include "qem_warmstart.h"

namespace qem {
namespace nn {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// ------------------------------------------------------------------------------------------------
// Features: one block per pattern, fixed-order double reductions.
__global__ __launch_bounds__(256) void k_features(int n, int amplitudes, const float* __restrict__ pat,
                                                  const float* __restrict__ edges, float* __restrict__ out) {
  __shared__ double red[4][9];
  const size_t base = (size_t)blockIdx.x * n * n;
  double s[9];
  for (int i = 0; i < 9; ++i) s[i] = 0.0;
  float e[7];
  for (int i = 0; i < 7; ++i) e[i] = edges[i];
  const float c = (float)(n / 2);
  for (int idx = threadIdx.x; idx < n * n; idx += 256) {
    float v = pat[base + idx];
    if (amplitudes) v = v * v;
    const float dr = (float)(idx / n) - c, dc = (float)(idx % n) - c;
    const float r = sqrtf(dr * dr + dc * dc);
    s[0] += v;
    s[1] += (double)(v * dr);
    s[2] += (double)(v * dc);
#pragma unroll
    for (int a = 0; a < 6; ++a)
      if (r >= e[a] && r < e[a + 1]) s[3 + a] += v;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    double x = s[i];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (l == 0) red[w][i] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[9];
    for (int i = 0; i < 9; ++i) t[i] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
    const double inv = t[0] > 0.0 ? 1.0 / t[0] : 0.0;
    float* o = out + (size_t)blockIdx.x * 8;
    o[0] = (float)(t[1] * inv);
    o[1] = (float)(t[2] * inv);
    for (int a = 0; a < 6; ++a) o[2 + a] = (float)(t[3 + a] * inv);
  }
}

// ------------------------------------------------------------------------------------------------
struct ConvArgs {
  int nb, h, w, cin, cout, relu, out_f32;
  const unsigned short* x;
  const unsigned short* wt;
  const float* bias;
  void* y;
};

// 256 threads = 4 waves; the block computes a 64 (pixels) x 64 (channels) tile, each wave 32 x 32
// as 2 x 2 MFMA 16x16 blocks.
__global__ __launch_bounds__(256) void k_conv3x3_fwd(ConvArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int M = a.nb * a.h * a.w;
  const int K = 9 * a.cin;
  const int m_base = blockIdx.x * 64 + (wave & 1) * 32;
  const int n_base = blockIdx.y * 64 + (wave >> 1) * 32;
  const int hw = a.h * a.w;
  int pn[2], py[2], px[2];
  bool pv[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int m = m_base + i * 16 + (lane & 15);
    pv[i] = m < M;
    const int mm = pv[i] ? m : 0;
    pn[i] = mm / hw;
    const int rem = mm - pn[i] * hw;
    py[i] = rem / a.w;
    px[i] = rem - py[i] * a.w;
  }
  int co[2];
  bool cv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    co[j] = n_base + j * 16 + (lane & 15);
    cv[j] = co[j] < a.cout;
  }
  const int kq = (lane >> 4) * 4;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const s4 zero = s4{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int k = k0 + kq;
    s4 af[2], bf[2];
    if (k < K) {
      const int tap = k / a.cin, ci = k - tap * a.cin;
      const int ky = tap / 3, kx = tap - ky * 3;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int yy = py[i] + ky - 1, xx = px[i] + kx - 1;
        const bool ok = pv[i] && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
        af[i] = ok ? *(const s4*)(a.x + (((size_t)pn[i] * a.h + yy) * a.w + xx) * a.cin + ci) : zero;
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = cv[j] ? *(const s4*)(a.wt + (size_t)co[j] * K + k) : zero;
    } else {
      af[0] = af[1] = bf[0] = bf[1] = zero;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bf[j], acc[i][j], 0, 0, 0);
  }
  // epilogue: lane holds rows 4*(lane/16) + r of each 16x16 block, column lane % 16
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (!cv[j]) continue;
    const float b = a.bias ? a.bias[co[j]] : 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_base + i * 16 + (lane >> 4) * 4 + r;
        if (m >= M) continue;
        float v = acc[i][j][r] + b;
        if (a.relu) v = fmaxf(v, 0.0f);
        const size_t o = (size_t)m * a.cout + co[j];
        if (a.out_f32)
          ((float*)a.y)[o] = v;
        else
          ((unsigned short*)a.y)[o] = f2bf(v);
      }
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Weight gradient: GEMM rows = cout, columns = K = 9*cin, reduction over pixels, split into chunks
// of kPixChunk pixels (grid.z); one wave per 32 x 32 tile.  Partials go to ws[split][cout][K].
constexpr int kPixChunk = 2048;

struct WgradArgs {
  int nb, h, w, cin, cout;
  const unsigned short* x;
  const unsigned short* dy;
  float* ws;
  float* dw;
  int splits;
};

__global__ __launch_bounds__(64) void k_conv3x3_wgrad(WgradArgs a) {
  const int lane = threadIdx.x;
  const int K = 9 * a.cin;
  const int P = a.nb * a.h * a.w;
  const int hw = a.h * a.w;
  const int r_base = blockIdx.x * 32, c_base = blockIdx.y * 32;
  const int p_begin = blockIdx.z * kPixChunk;
  const int p_end = min(P, p_begin + kPixChunk);
  // A = dy^T : A[co][p], lane: co = r_base + i*16 + lane%16, p = p0 + 4*(lane/16) + 0..3
  // B = im2col(x) : B[p][k], lane: k = c_base + j*16 + lane%16, p = p0 + 4*(lane/16) + 0..3
  int cor[2], kc[2], tapy[2], tapx[2], cic[2];
  bool rv[2], kv[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    cor[i] = r_base + i * 16 + (lane & 15);
    rv[i] = cor[i] < a.cout;
    kc[i] = c_base + i * 16 + (lane & 15);
    kv[i] = kc[i] < K;
    const int kk = kv[i] ? kc[i] : 0;
    const int tap = kk / a.cin;
    cic[i] = kk - tap * a.cin;
    tapy[i] = tap / 3 - 1;
    tapx[i] = tap % 3 - 1;
  }
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int p0 = p_begin; p0 < p_end; p0 += 16) {
    s4 af[2], bf[2];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int p = p0 + (lane >> 4) * 4 + q;
      const bool pok = p < p_end;
      const int pp = pok ? p : 0;
      const int n = pp / hw, rem = pp - n * hw, y = rem / a.w, x = rem - y * a.w;
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i][q] = (pok && rv[i]) ? (short)a.dy[(size_t)pp * a.cout + cor[i]] : (short)0;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int yy = y + tapy[j], xx = x + tapx[j];
        const bool ok = pok && kv[j] && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
        bf[j][q] = ok ? (short)a.x[(((size_t)n * a.h + yy) * a.w + xx) * a.cin + cic[j]] : (short)0;
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bf[j], acc[i][j], 0, 0, 0);
  }
  float* ws = a.ws + (size_t)blockIdx.z * a.cout * K;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = c_base + j * 16 + (lane & 15);
      if (col >= K) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = r_base + i * 16 + (lane >> 4) * 4 + r;
        if (row < a.cout) ws[(size_t)row * K + col] = acc[i][j][r];
      }
    }
}

__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* ws, float* dw, int count, int splits) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += ws[(size_t)k * count + i];
  dw[i] = s;
}

}  // namespace nn
}  // namespace qem

namespace {
int conv_shape_ok(int nb, int h, int w, int cin, int cout) {
  if (nb < 1 || h < 1 || w < 1 || cin < 1 || cout < 1)
    return qem::set_error(QEM_ERR_ARG, "conv3x3: sizes must be positive");
  if (cin % 4 != 0) return qem::set_error(QEM_ERR_ARG, "conv3x3: cin (%d) must be a multiple of 4", cin);
  if (cout % 16 != 0) return qem::set_error(QEM_ERR_ARG, "conv3x3: cout (%d) must be a multiple of 16", cout);
  if ((size_t)nb * h * w * (cin > cout ? cin : cout) >= ((size_t)1 << 31))
    return qem::set_error(QEM_ERR_ARG, "conv3x3: tensor too large");
  return QEM_OK;
}
bool aligned8(const void* p) { return ((uintptr_t)p & 7u) == 0; }
}  // namespace

extern "C" {

int qem_pattern_features(int32_t num_patterns, int32_t n, int32_t amplitudes, const float* patterns,
                         const float* edges7, float* out, void* stream) {
  qem::clear_error();
  if (num_patterns < 0 || n < 2) return qem::set_error(QEM_ERR_ARG, "features: bad sizes");
  if (num_patterns == 0) return QEM_OK;
  if (!patterns || !edges7 || !out) return qem::set_error(QEM_ERR_NULL, "features: NULL pointer");
  hipLaunchKernelGGL(qem::nn::k_features, dim3(num_patterns), dim3(256), 0, (hipStream_t)stream, n, amplitudes,
                     patterns, edges7, out);
  return qem::check_last_launch("features");
}

int qem_conv3x3_forward(int32_t nb, int32_t h, int32_t w, int32_t cin, int32_t cout, const void* x,
                        const void* weight, const float* bias, int32_t relu, int32_t out_f32, void* y,
                        void* stream) {
  qem::clear_error();
  if (int rc = conv_shape_ok(nb, h, w, cin, cout)) return rc;
  if (!x || !weight || !y) return qem::set_error(QEM_ERR_NULL, "conv3x3: NULL pointer");
  if (!aligned8(x) || !aligned8(weight)) return qem::set_error(QEM_ERR_ALIGN, "conv3x3: x/weight misaligned");
  qem::nn::ConvArgs a;
  a.nb = nb;
  a.h = h;
  a.w = w;
  a.cin = cin;
  a.cout = cout;
  a.relu = relu;
  a.out_f32 = out_f32;
  a.x = (const unsigned short*)x;
  a.wt = (const unsigned short*)weight;
  a.bias = bias;
  a.y = y;
  const int M = nb * h * w;
  dim3 grid((M + 63) / 64, (cout + 63) / 64);
  hipLaunchKernelGGL(qem::nn::k_conv3x3_fwd, grid, dim3(256), 0, (hipStream_t)stream, a);
  return qem::check_last_launch("conv3x3_forward");
}

int qem_conv3x3_wgrad_workspace_bytes(int32_t nb, int32_t h, int32_t w, int32_t cin, int32_t cout, size_t* bytes) {
  qem::clear_error();
  if (int rc = conv_shape_ok(nb, h, w, cin, cout)) return rc;
  if (!bytes) return qem::set_error(QEM_ERR_NULL, "bytes is NULL");
  const size_t P = (size_t)nb * h * w;
  const size_t splits = (P + qem::nn::kPixChunk - 1) / qem::nn::kPixChunk;
  *bytes = splits * (size_t)cout * 9 * cin * sizeof(float);
  return QEM_OK;
}

int qem_conv3x3_wgrad(int32_t nb, int32_t h, int32_t w, int32_t cin, int32_t cout, const void* x, const void* dy,
                      float* dw, void* workspace, size_t workspace_bytes, void* stream) {
  qem::clear_error();
  size_t need = 0;
  if (int rc = qem_conv3x3_wgrad_workspace_bytes(nb, h, w, cin, cout, &need)) return rc;
  if (!x || !dy || !dw || !workspace) return qem::set_error(QEM_ERR_NULL, "conv3x3_wgrad: NULL pointer");
  if (workspace_bytes < need)
    return qem::set_error(QEM_ERR_WORKSPACE, "conv3x3_wgrad: workspace has %zu bytes, needs %zu", workspace_bytes, need);
  qem::nn::WgradArgs a;
  a.nb = nb;
  a.h = h;
  a.w = w;
  a.cin = cin;
  a.cout = cout;
  a.x = (const unsigned short*)x;
  a.dy = (const unsigned short*)dy;
  a.ws = (float*)workspace;
  a.dw = dw;
  const int P = nb * h * w;
  a.splits = (P + qem::nn::kPixChunk - 1) / qem::nn::kPixChunk;
  const int K = 9 * cin;
  dim3 grid((cout + 31) / 32, (K + 31) / 32, a.splits);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(qem::nn::k_conv3x3_wgrad, grid, dim3(64), 0, st, a);
  if (int rc = qem::check_last_launch("conv3x3_wgrad")) return rc;
  const int count = cout * K;
  hipLaunchKernelGGL(qem::nn::k_wgrad_reduce, dim3((count + 255) / 256), dim3(256), 0, st, (const float*)workspace, dw,
                     count, a.splits);
  return qem::check_last_launch("conv3x3_wgrad_reduce");
}

}  // extern "C"
