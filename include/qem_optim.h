/*
 * qem_optim.h — C ABI of the optimizer step the reconstruction loop runs after every batch.
 *
 * The reference's models are optimised by torch.optim.Adam (optimizer type "adam", the default;
 * "adamw" and "sgd" select AdamW and SGD; quantem/core/ml/optimizer_mixin.py:84-126; object learning
 * rate 5e-3, object_models.py:44-48; probe 1e-3, probe_models.py:52-54), after the object gradient is
 * rescaled by upsample_factor**2 / 2 (quantem/diffractive_imaging/ptychography.py:330-332).  Each entry
 * point is that rescale and one optimizer step (Adam without amsgrad, AdamW, SGD) in one pass over a
 * float32 parameter vector; a complex64 parameter is passed as its interleaved float32 view, which is
 * how torch's optimizers update it.
 * Conventions as in qem_ptycho.h: device pointers, asynchronous on `stream`, no allocation.
 */
#ifndef QEM_OPTIM_H
#define QEM_OPTIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* In place, for i < n (torch.optim.Adam's update, torch/optim/adam.py, single-tensor path):
 *   g   = grad_scale * grad[i] + weight_decay * param[i]
 *   m   = exp_avg[i] + (1 - beta1) * (g - exp_avg[i])
 *   v   = beta2 * exp_avg_sq[i] + (1 - beta2) * g * g
 *   param[i] -= lr / (1 - beta1^step) * m / (sqrt(v) / sqrt(1 - beta2^step) + eps)
 * step is the 1-based count including this update.  The hyper-parameters are doubles (torch's Python
 * floats); each derived scalar (1 - beta1, beta2, 1 - beta2, lr / (1 - beta1^step),
 * sqrt(1 - beta2^step), eps) is formed in double and rounded to float32 once, as torch's kernels
 * receive them.  Returns QEM_ERR_ARG (1) for invalid scalars,
 * QEM_ERR_NULL (2) / QEM_ERR_ALIGN (3) for bad pointers, QEM_ERR_LAUNCH (5) on launch failure; the
 * message is in qem_last_error(). */
int qem_adam_step_f32(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                      double lr, double beta1, double beta2, double eps, double weight_decay, double grad_scale,
                      int32_t step, void* stream);

/* torch.optim.AdamW (decoupled weight decay; optimizer type "adamw", optimizer_mixin.py:119-120):
 *   param[i] *= 1 - lr * weight_decay      (the factor formed in double, rounded once)
 * then the update of qem_adam_step_f32 with weight_decay = 0.  Same arguments and error codes. */
int qem_adamw_step_f32(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                       double lr, double beta1, double beta2, double eps, double weight_decay, double grad_scale,
                       int32_t step, void* stream);

/* torch.optim.SGD (optimizer type "sgd", optimizer_mixin.py:121-122), single-tensor update:
 *   d = grad_scale * grad[i] + weight_decay * param[i]
 *   with momentum != 0:  b = first_step ? d : momentum * momentum_buffer[i] + (1 - dampening) * d,
 *                        momentum_buffer[i] = b,  d = nesterov ? d + momentum * b : b
 *   param[i] -= lr * d
 * momentum_buffer may be NULL when momentum == 0.  QEM_ERR_ARG for invalid scalars (Nesterov needs
 * momentum > 0 and dampening 0, as torch requires), QEM_ERR_NULL / QEM_ERR_ALIGN for bad pointers
 * (16-byte alignment), QEM_ERR_LAUNCH on launch failure. */
int qem_sgd_step_f32(int64_t n, float* param, const float* grad, float* momentum_buffer, double lr, double momentum,
                     double dampening, double weight_decay, int32_t nesterov, double grad_scale, int32_t first_step,
                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QEM_OPTIM_H */
