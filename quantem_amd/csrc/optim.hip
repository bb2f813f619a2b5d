// optim.hip — the per-batch optimizer step (object-gradient rescale + Adam, AdamW or SGD) in one pass.
//
// The reference optimises each model with torch.optim.Adam, AdamW or SGD (core/ml/optimizer_mixin.py:
// 117-122), which moves each parameter, its gradient and the optimizer state through HBM in several
// kernels, plus a separate kernel for the object-gradient rescale (ptychography.py:330-332).  Here one
// grid-stride kernel per parameter reads param, grad and the state once and writes param and state
// once (Adam / AdamW: 28 bytes per float32 element; SGD with momentum: 20), 4 elements per thread and
// iteration.
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
  float decay;     // AdamW: 1 - lr * weight_decay applied to the parameter first (wd = 0 then); Adam: 1
};

__device__ __forceinline__ float adam_one(float& p, float g, float& m, float& v, const AdamArgs& a) {
  p = p * a.decay;  // AdamW's decoupled decay (param.mul_(1 - lr * weight_decay)); exact for Adam (1)
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

// torch.optim.SGD: d = g + wd p; with momentum, buf = d on the first step, else
// buf = momentum buf + (1 - dampening) d, and d = buf (or d + momentum buf with Nesterov); p -= lr d.
struct SgdArgs {
  float lr, momentum, one_m_damp, wd, gscale;
  int has_momentum, nesterov, first;
};

__device__ __forceinline__ void sgd_one(float& p, float g, float* buf, const SgdArgs& a) {
  float d = g * a.gscale + a.wd * p;
  if (a.has_momentum) {
    const float b = a.first ? d : *buf * a.momentum + a.one_m_damp * d;
    *buf = b;
    d = a.nesterov ? d + a.momentum * b : b;
  }
  p = p + (-a.lr) * d;
}

__global__ __launch_bounds__(256) void k_sgd(int64_t n, float* __restrict__ param, const float* __restrict__ grad,
                                              float* __restrict__ buf, SgdArgs a) {
  const int64_t n4 = n / 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float4* P = (float4*)param;
  const float4* G = (const float4*)grad;
  float4* Bf = (float4*)buf;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    float4 p = P[i];
    const float4 g = G[i];
    float4 b;
    if (a.has_momentum && !a.first) b = Bf[i];
    sgd_one(p.x, g.x, &b.x, a);
    sgd_one(p.y, g.y, &b.y, a);
    sgd_one(p.z, g.z, &b.z, a);
    sgd_one(p.w, g.w, &b.w, a);
    P[i] = p;
    if (a.has_momentum) Bf[i] = b;
  }
  const int64_t t = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    float b = 0.0f;
    if (a.has_momentum && !a.first) b = buf[t];
    sgd_one(param[t], grad[t], &b, a);
    if (a.has_momentum) buf[t] = b;
  }
}

int check_vec_ptrs(const void* const* ptrs, const char* const* names, int k) {
  for (int i = 0; i < k; ++i) {
    if (ptrs[i] == nullptr) return qem::set_error(QEM_ERR_NULL, "%s is NULL", names[i]);
    if (((uintptr_t)ptrs[i]) & 15u) return qem::set_error(QEM_ERR_ALIGN, "%s is not 16-byte aligned", names[i]);
  }
  return QEM_OK;
}

int grid_of(int64_t n) {
  const int64_t work = (n + 3) / 4;
  const int blocks = (int)(work < 256 * 2048 ? (work + 255) / 256 : 2048);
  return blocks > 0 ? blocks : 1;
}

int adam_common(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq, double lr,
                double beta1, double beta2, double eps, double weight_decay, double grad_scale, int32_t step,
                bool decoupled, void* stream) {
  qem::clear_error();
  if (n < 0) return qem::set_error(QEM_ERR_ARG, "n must be >= 0 (got %lld)", (long long)n);
  if (step < 1) return qem::set_error(QEM_ERR_ARG, "step must be >= 1 (got %d)", step);
  if (!(lr >= 0.0) || !(eps >= 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) ||
      !isfinite(grad_scale) || !isfinite(weight_decay) || !(weight_decay >= 0.0))
    return qem::set_error(QEM_ERR_ARG, "invalid Adam hyper-parameters");
  if (n == 0) return QEM_OK;
  const void* ptrs[4] = {param, grad, exp_avg, exp_avg_sq};
  const char* names[4] = {"param", "grad", "exp_avg", "exp_avg_sq"};
  if (int rc = check_vec_ptrs(ptrs, names, 4)) return rc;
  // scalars formed in double and rounded once, as torch passes Python floats to its float32 kernels
  const double bc1 = 1.0 - pow(beta1, (double)step);
  const double bc2 = 1.0 - pow(beta2, (double)step);
  AdamArgs a;
  a.lr_bc1 = (float)(lr / bc1);
  a.sqrt_bc2 = (float)sqrt(bc2);
  a.one_m_b1 = (float)(1.0 - beta1);
  a.beta2 = (float)beta2;
  a.one_m_b2 = (float)(1.0 - beta2);
  a.eps = (float)eps;
  a.wd = decoupled ? 0.0f : (float)weight_decay;
  a.decay = decoupled ? (float)(1.0 - lr * weight_decay) : 1.0f;
  a.gscale = (float)grad_scale;
  hipLaunchKernelGGL(k_adam, dim3(grid_of(n)), dim3(256), 0, (hipStream_t)stream, n, param, grad, exp_avg,
                     exp_avg_sq, a);
  return qem::check_last_launch(decoupled ? "adamw" : "adam");
}

}  // namespace
}  // namespace qem

extern "C" int qem_adamw_step_f32(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                                  double lr, double beta1, double beta2, double eps, double weight_decay,
                                  double grad_scale, int32_t step, void* stream) {
  return qem::adam_common(n, param, grad, exp_avg, exp_avg_sq, lr, beta1, beta2, eps, weight_decay, grad_scale, step,
                          true, stream);
}

extern "C" int qem_sgd_step_f32(int64_t n, float* param, const float* grad, float* momentum_buffer, double lr,
                                double momentum, double dampening, double weight_decay, int32_t nesterov,
                                double grad_scale, int32_t first_step, void* stream) {
  qem::clear_error();
  if (n < 0) return qem::set_error(QEM_ERR_ARG, "n must be >= 0 (got %lld)", (long long)n);
  if (!(lr >= 0.0) || !(momentum >= 0.0) || !(weight_decay >= 0.0) || !isfinite(dampening) ||
      !isfinite(grad_scale) || !isfinite(momentum) || !isfinite(weight_decay))
    return qem::set_error(QEM_ERR_ARG, "invalid SGD hyper-parameters");
  if (nesterov && (momentum <= 0.0 || dampening != 0.0))
    return qem::set_error(QEM_ERR_ARG, "Nesterov momentum requires a momentum and zero dampening");
  if (n == 0) return QEM_OK;
  const bool mom = momentum != 0.0;
  const void* ptrs[3] = {param, grad, momentum_buffer};
  const char* names[3] = {"param", "grad", "momentum_buffer"};
  if (int rc = qem::check_vec_ptrs(ptrs, names, mom ? 3 : 2)) return rc;
  qem::SgdArgs a;
  a.lr = (float)lr;
  a.momentum = (float)momentum;
  a.one_m_damp = (float)(1.0 - dampening);
  a.wd = (float)weight_decay;
  a.gscale = (float)grad_scale;
  a.has_momentum = mom ? 1 : 0;
  a.nesterov = nesterov ? 1 : 0;
  a.first = first_step ? 1 : 0;
  hipLaunchKernelGGL(qem::k_sgd, dim3(qem::grid_of(n)), dim3(256), 0, (hipStream_t)stream, n, param, grad,
                     mom ? momentum_buffer : nullptr, a);
  return qem::check_last_launch("sgd");
}

extern "C" int qem_adam_step_f32(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                                 double lr, double beta1, double beta2, double eps, double weight_decay,
                                 double grad_scale, int32_t step, void* stream) {
  return qem::adam_common(n, param, grad, exp_avg, exp_avg_sq, lr, beta1, beta2, eps, weight_decay, grad_scale, step,
                          false, stream);
}
