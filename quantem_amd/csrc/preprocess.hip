// preprocess.hip — centre of mass and centring / normalisation of the measured diffraction data.
//
// Reference behaviour followed (quantem, src/quantem/diffractive_imaging/):
//   vectorised CoM            dataset_models.py:1063-1074
//   amplitude centring loop   dataset_models.py:1408-1440 (intensity clamp, sqrt, shift by -com_fit,
//                             clamp, fftshift)
//   Fourier shift             ptycho_utils.py:239-275 (shift_array, periodic, not bilinear), with the
//                             Fourier coordinates of make_Fourier_coords2D (ptycho_utils.py:303-321)
// One workgroup per pattern; n <= 128 transforms in registers (fft2.h), n = 256 through the row /
// column passes of ptycho256.hip.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "ptycho_common.h"
#include "qem_preprocess.h"
#include "qem_ptycho.h"

namespace qem {
namespace {

__global__ __launch_bounds__(256) void k_com(int n, const float* __restrict__ I, const float* __restrict__ mask,
                                              float* __restrict__ com) {
  __shared__ double red[4];
  const float* src = I + (size_t)blockIdx.x * n * n;
  double s = 0.0, sr = 0.0, sc = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += 256) {
    const int r = idx / n, c = idx - r * n;
    const double x = (double)src[idx] * (mask != nullptr ? (double)mask[idx] : 1.0);
    s += x;
    sr += x * r;
    sc += x * c;
  }
  const double ts = block_sum<256>(s, red);
  const double tr = block_sum<256>(sr, red);
  const double tc = block_sum<256>(sc, red);
  if (threadIdx.x == 0) {
    com[2 * blockIdx.x] = (float)(tr / ts);  // not finite for an empty pattern, as in the reference
    com[2 * blockIdx.x + 1] = (float)(tc / ts);
  }
}

template <int N>
__global__ __launch_bounds__(Fft2<N>::T) void k_center(const float* __restrict__ I, const float* __restrict__ com,
                                                         float* __restrict__ out, float* __restrict__ sums) {
  using F = Fft2<N>;
  __shared__ cf lds[F::LDS_TOTAL];
  __shared__ double red[F::T / 64];
  F f(threadIdx.x);
  f.init_twiddles(lds, threadIdx.x);
  __syncthreads();
  const int pat = blockIdx.x;
  const float* src = I + (size_t)pat * N * N;
  cf v[F::PTS];
  double si = 0.0, sa = 0.0;
#pragma unroll
  for (int e = 0; e < F::PTS; ++e) {
    const float x = fmaxf(src[f.p1_off(e)], 0.0f);
    const float a = sqrtf(x);
    si += (double)x;
    sa += (double)a;
    v[e] = mk(a, 0.0f);
  }
  f.fwd(v, lds);
  // shift by (dr, dc) = -com_fit: exp(-2 pi i (q_c dc + q_r dr)), q = fftfreq(N, 1)
  const float dr = -com[2 * pat], dc = -com[2 * pat + 1];
  const cf pc = expm2pi(fftfreq1<N>(f.kc) * dc);
#pragma unroll
  for (int e = 0; e < F::PTS; ++e) v[e] = cmul(v[e], cmul(expm2pi(fftfreq1<N>(f.p3_row(e)) * dr), pc));
  f.inv(v, lds);
  const float inv_n2 = 1.0f / (float)(N * N);
  float* dst = out + (size_t)pat * N * N;
#pragma unroll
  for (int e = 0; e < F::PTS; ++e) {
    const int row = (f.p1_row(e) + N / 2) & (N - 1), col = (f.p1_col(e) + N / 2) & (N - 1);
    dst[row * N + col] = fmaxf(v[e].x * inv_n2, 0.0f);
  }
  if (sums != nullptr) {
    const double ti = block_sum<F::T>(si, red);
    const double ta = block_sum<F::T>(sa, red);
    if (threadIdx.x == 0) {
      sums[2 * pat] = (float)ti;
      sums[2 * pat + 1] = (float)ta;
    }
  }
}

constexpr int kCenter256Blocks = 256;  // resident workgroups of the n = 256 kernel (one wave each)

int check_f32(const void* ptr, const char* name) {
  if (ptr == nullptr) return set_error(QEM_ERR_NULL, "%s is NULL", name);
  if (((uintptr_t)ptr) & 3u) return set_error(QEM_ERR_ALIGN, "%s is misaligned", name);
  return QEM_OK;
}

}  // namespace
}  // namespace qem

extern "C" {

int qem_pattern_com(int32_t num_patterns, int32_t n, const float* intensities, const float* mask, float* com,
                    void* stream) {
  qem::clear_error();
  if (num_patterns < 1) return qem::set_error(QEM_ERR_ARG, "num_patterns must be >= 1 (got %d)", num_patterns);
  if (n < 2 || n > 4096) return qem::set_error(QEM_ERR_ARG, "n must be in [2, 4096] (got %d)", n);
  if (int rc = qem::check_f32(intensities, "intensities")) return rc;
  if (mask != nullptr)
    if (int rc = qem::check_f32(mask, "mask")) return rc;
  if (int rc = qem::check_f32(com, "com")) return rc;
  hipLaunchKernelGGL(qem::k_com, dim3(num_patterns), dim3(256), 0, (hipStream_t)stream, (int)n, intensities, mask,
                     com);
  return qem::check_last_launch("pattern_com");
}

int qem_center_amplitudes_workspace_bytes(int32_t num_patterns, int32_t n, size_t* bytes) {
  qem::clear_error();
  if (bytes == nullptr) return qem::set_error(QEM_ERR_NULL, "bytes is NULL");
  if (num_patterns < 1) return qem::set_error(QEM_ERR_ARG, "num_patterns must be >= 1 (got %d)", num_patterns);
  if (n != 32 && n != 64 && n != 128 && n != 256)
    return qem::set_error(QEM_ERR_UNSUPPORTED, "n %d not supported (32, 64, 128 or 256)", n);
  const int blocks = num_patterns < qem::kCenter256Blocks ? num_patterns : qem::kCenter256Blocks;
  *bytes = n == 256 ? (size_t)blocks * 256 * 256 * sizeof(qem::cf) : 0;
  return QEM_OK;
}

int qem_center_amplitudes(int32_t num_patterns, int32_t n, const float* intensities, const float* com_fit,
                          float* centered, float* sums, void* workspace, size_t workspace_bytes, void* stream) {
  qem::clear_error();
  size_t need = 0;
  if (int rc = qem_center_amplitudes_workspace_bytes(num_patterns, n, &need)) return rc;
  if (int rc = qem::check_f32(intensities, "intensities")) return rc;
  if (int rc = qem::check_f32(com_fit, "com_fit")) return rc;
  if (int rc = qem::check_f32(centered, "centered")) return rc;
  if (sums != nullptr)
    if (int rc = qem::check_f32(sums, "sums")) return rc;
  hipStream_t st = (hipStream_t)stream;
  switch (n) {
    case 32:
      hipLaunchKernelGGL(qem::k_center<32>, dim3(num_patterns), dim3(qem::Fft2<32>::T), 0, st, intensities, com_fit,
                         centered, sums);
      break;
    case 64:
      hipLaunchKernelGGL(qem::k_center<64>, dim3(num_patterns), dim3(qem::Fft2<64>::T), 0, st, intensities, com_fit,
                         centered, sums);
      break;
    case 128:
      hipLaunchKernelGGL(qem::k_center<128>, dim3(num_patterns), dim3(qem::Fft2<128>::T), 0, st, intensities,
                         com_fit, centered, sums);
      break;
    default: {
      if (workspace == nullptr) return qem::set_error(QEM_ERR_NULL, "workspace is NULL");
      if (((uintptr_t)workspace) & 7u) return qem::set_error(QEM_ERR_ALIGN, "workspace is misaligned");
      if (workspace_bytes < need)
        return qem::set_error(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", workspace_bytes, need);
      const int blocks = num_patterns < qem::kCenter256Blocks ? num_patterns : qem::kCenter256Blocks;
      return qem::launch_center256(num_patterns, blocks, intensities, com_fit, centered, sums, (qem::cf*)workspace,
                                   st);
    }
  }
  return qem::check_last_launch("center_amplitudes");
}

}  // extern "C"
