/*
 * qem_warmstart.h — C ABI of the learned warm-start kernels (gfx950).
 *
 * The reference has no warm-start network (its CNN2d / Conv2dBlock U-Net, core/ml/cnn.py:17-212 and
 * core/ml/blocks.py:53-210, is used only as a deep-image-prior re-parameterisation); these entry points
 * are new.  Inputs follow the reference's data conventions: diffraction patterns are fftshifted
 * (DC at (N/2, N/2), detector_models.py:30), centre of mass as in origin_models.py:95-126, annular
 * virtual detectors as in dataset4dstem.py:494-520.
 *
 * Activations are NHWC bfloat16 (raw 16-bit), weights (Cout, 3, 3, Cin) bfloat16, accumulation float32.
 * Same conventions as qem_ptycho.h: device pointers, asynchronous on `stream`, no allocation,
 * 0 = success, QEM_ERR_* otherwise, message in qem_last_error().
 */
#ifndef QEM_WARMSTART_H
#define QEM_WARMSTART_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-pattern features for P fftshifted N x N patterns (intensities, or amplitudes if
 * `amplitudes` != 0, squared on the fly):
 *   out[p][0..1] = centre of mass (row, col) relative to (N/2, N/2), in pixels
 *   out[p][2..7] = fraction of the total intensity in 6 annuli with radius edges edges[0..6] (pixels)
 * N is any size; patterns are float32 (P, N, N); out is float32 (P, 8). */
int qem_pattern_features(int32_t num_patterns, int32_t n, int32_t amplitudes, const float* patterns,
                         const float* edges7, float* out, void* stream);

/* 3x3 convolution, stride 1, zero padding 1 (Conv2dBlock's conv, blocks.py:53-140, with zero instead
 * of circular padding), as an implicit GEMM on MFMA (bf16 in, fp32 accumulate):
 *   y[n, y, x, co] = act(bias[co] + sum_{ky, kx, ci} w[co, ky, kx, ci] * x[n, y+ky-1, x+kx-1, ci])
 * act = ReLU if relu != 0.  y is float32 if out_f32 != 0, else bfloat16.  cin % 4 == 0, cout % 16 == 0.
 * bias may be NULL.  The input gradient of a convolution is this forward with the weights flipped
 * and transposed (w'[ci, ky, kx, co] = w[co, 2-ky, 2-kx, ci]). */
int qem_conv3x3_forward(int32_t nb, int32_t h, int32_t w, int32_t cin, int32_t cout, const void* x,
                        const void* weight, const float* bias, int32_t relu, int32_t out_f32, void* y,
                        void* stream);

/* Weight gradient dw[co, ky, kx, ci] = sum_{n, y, x} dy[n, y, x, co] * x[n, y+ky-1, x+kx-1, ci]
 * (float32 out), split over pixels and summed in a fixed order through `workspace`. */
int qem_conv3x3_wgrad_workspace_bytes(int32_t nb, int32_t h, int32_t w, int32_t cin, int32_t cout,
                                      size_t* bytes);
int qem_conv3x3_wgrad(int32_t nb, int32_t h, int32_t w, int32_t cin, int32_t cout, const void* x,
                      const void* dy, float* dw, void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QEM_WARMSTART_H */
