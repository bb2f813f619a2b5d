// optim.hip — the per-batch optimizer step (object-gradient rescale + Adam) in one pass.
//
// torch.optim.Adam over the object and the probe moves each parameter, its gradient and both moment
// buffers through HBM in several kernels, plus a separate kernel for the reference's object-gradient
// rescale (ptychography.py:330-332).  Here one grid-stride kernel reads param, grad, exp_avg,
// exp_avg_sq once and writes param, exp_avg, exp_avg_sq once (28 bytes per float32 element),
// 4 elements per thread and iteration.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "common.h"
#include "qem_optim.h"
#include "qem_ptycho.h"

namespace qem {
namespace {

struct AdamArgs {
  float lr_bc1;    // lr / (1 - beta1^step)
  float sqrt_bc2;  // sqrt(1 - beta2^step)
  float one_m_b1, beta2, one_m_b2, eps, wd, gscale;
};

__device__ __forceinline__ float adam_one(float& p, float g, float& m, float& v, const AdamArgs& a) {
  g = g * a.gscale + a.wd * p;
  // the operation order of torch's exp_avg.lerp_, exp_avg_sq.mul_().addcmul_() and param.addcdiv_()
  m = m + a.one_m_b1 * (g - m);
  v = v * a.beta2 + a.one_m_b2 * (g * g);
  const float denom = sqrtf(v) / a.sqrt_bc2 + a.eps;
  p = p + (-a.lr_bc1) * (m / denom);
  return p;
}

__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ param, const float* __restrict__ grad,
                                               float* __restrict__ m1, float* __restrict__ m2, AdamArgs a) {
  const int64_t n4 = n / 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float4* P = (float4*)param;
  const float4* G = (const float4*)grad;
  float4* M = (float4*)m1;
  float4* V = (float4*)m2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    float4 p = P[i], m = M[i], v = V[i];
    const float4 g = G[i];
    adam_one(p.x, g.x, m.x, v.x, a);
    adam_one(p.y, g.y, m.y, v.y, a);
    adam_one(p.z, g.z, m.z, v.z, a);
    adam_one(p.w, g.w, m.w, v.w, a);
    P[i] = p;
    M[i] = m;
    V[i] = v;
  }
  // tail (n not a multiple of 4)
  const int64_t t = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) adam_one(param[t], grad[t], m1[t], m2[t], a);
}

}  // namespace
}  // namespace qem

extern "C" int qem_adam_step_f32(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                                 double lr, double beta1, double beta2, double eps, double weight_decay,
                                 double grad_scale, int32_t step, void* stream) {
  qem::clear_error();
  if (n < 0) return qem::set_error(QEM_ERR_ARG, "n must be >= 0 (got %lld)", (long long)n);
  if (step < 1) return qem::set_error(QEM_ERR_ARG, "step must be >= 1 (got %d)", step);
  if (!(lr >= 0.0) || !(eps >= 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) ||
      !isfinite(grad_scale) || !isfinite(weight_decay))
    return qem::set_error(QEM_ERR_ARG, "invalid Adam hyper-parameters");
  if (n == 0) return QEM_OK;
  const void* ptrs[4] = {param, grad, exp_avg, exp_avg_sq};
  const char* names[4] = {"param", "grad", "exp_avg", "exp_avg_sq"};
  for (int i = 0; i < 4; ++i) {
    if (ptrs[i] == nullptr) return qem::set_error(QEM_ERR_NULL, "%s is NULL", names[i]);
    if (((uintptr_t)ptrs[i]) & 15u) return qem::set_error(QEM_ERR_ALIGN, "%s is not 16-byte aligned", names[i]);
  }
  // scalars formed in double and rounded once, as torch passes Python floats to its float32 kernels
  const double bc1 = 1.0 - pow(beta1, (double)step);
  const double bc2 = 1.0 - pow(beta2, (double)step);
  qem::AdamArgs a;
  a.lr_bc1 = (float)(lr / bc1);
  a.sqrt_bc2 = (float)sqrt(bc2);
  a.one_m_b1 = (float)(1.0 - beta1);
  a.beta2 = (float)beta2;
  a.one_m_b2 = (float)(1.0 - beta2);
  a.eps = (float)eps;
  a.wd = (float)weight_decay;
  a.gscale = (float)grad_scale;
  const int64_t work = (n + 3) / 4;
  const int blocks = (int)(work < 256 * 2048 ? (work + 255) / 256 : 2048);
  hipLaunchKernelGGL(qem::k_adam, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, (hipStream_t)stream, n, param, grad,
                     exp_avg, exp_avg_sq, a);
  return qem::check_last_launch("adam");
}
