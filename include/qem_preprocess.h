/*
 * qem_preprocess.h — C ABI of the measured-data preprocessing that feeds the reconstruction:
 * per-pattern centre of mass and the centring / normalisation of the diffraction amplitudes.
 *
 * Replaces (quantem/diffractive_imaging/):
 *   the vectorised CoM of PtychographyDatasetRaster._set_intensities_com (dataset_models.py:1063-1074)
 *   the loop of _normalize_diffraction_intensities (dataset_models.py:1408-1440; no crop, Fourier
 *   shift) with shift_array (ptycho_utils.py:239-275, periodic, not bilinear)
 * The plane fit of the CoM over the scan (fit_origin, ptycho_utils.py:352-410) is a 3-parameter
 * least-squares problem that the host solves (quantem_amd/preprocess.py).
 * Conventions as in qem_ptycho.h: float32 device arrays, row-major, asynchronous on `stream`.
 */
#ifndef QEM_PREPROCESS_H
#define QEM_PREPROCESS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* com[p] = (sum_q I m q_r, sum_q I m q_c) / sum_q I m for the (P, n, n) intensities I (mask m (n, n)
 * or NULL = ones; q_r, q_c the pixel row / column indices), as float32 (P, 2).  n >= 2. */
int qem_pattern_com(int32_t num_patterns, int32_t n, const float* intensities, const float* mask, float* com,
                    void* stream);

/* Workspace bytes of qem_center_amplitudes (0 for n <= 128; a 512 KiB wave per resident workgroup for
 * n = 256). */
int qem_center_amplitudes_workspace_bytes(int32_t num_patterns, int32_t n, size_t* bytes);

/* For every pattern p (n in {32, 64, 128, 256}):
 *   a = sqrt(max(I, 0));  shifted = real(ifft2(fft2(a) exp(-2 pi i (q_r dr + q_c dc))))  with
 *   (dr, dc) = -com_fit[p], q = fftfreq(n, 1);  centered[p] = fftshift(max(shifted, 0));
 *   sums[p] = (sum max(I, 0), sum a) as float32 (the per-pattern terms of the mean intensity and
 *   mean amplitude).  centered (P, n, n) float32, sums (P, 2) float32 or NULL. */
int qem_center_amplitudes(int32_t num_patterns, int32_t n, const float* intensities, const float* com_fit,
                          float* centered, float* sums, void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QEM_PREPROCESS_H */
