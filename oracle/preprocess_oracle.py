"""CPU restatement of the measured-data preprocessing (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module; the product path runs the HIP kernels of
quantem_amd/csrc/preprocess.hip.  float64 numpy, following (quantem, src/quantem/diffractive_imaging/):
  pattern_com        dataset_models.py:1063-1074 (vectorised CoM, optional detector mask)
  fit_plane          ptycho_utils.py:327-328 (_plane) and 352-410 (fit_origin, "plane"): the linear
                     least-squares solution that curve_fit converges to for a linear model
  shift_array        ptycho_utils.py:239-275 (Fourier shift, periodic), coordinates of
                     make_Fourier_coords2D (ptycho_utils.py:303-321)
  center_amplitudes  dataset_models.py:1408-1440 (clamp, sqrt, shift by -com_fit, clamp, fftshift)
"""

import numpy as np


def pattern_com(intensities, mask=None):
    I = np.asarray(intensities, dtype=np.float64)
    if mask is not None:
        I = I * np.asarray(mask, dtype=np.float64)
    n_r, n_c = I.shape[-2:]
    kr = np.arange(n_r, dtype=np.float64)[:, None]
    kc = np.arange(n_c, dtype=np.float64)[None, :]
    s = I.sum(axis=(-2, -1))
    return np.stack([(I * kr).sum(axis=(-2, -1)) / s, (I * kc).sum(axis=(-2, -1)) / s], axis=-1)


def fit_plane(com, scan_shape):
    """com (P, 2) over a row-major (Rr, Rc) scan -> fitted (P, 2): mx * r + my * c + b per component."""
    com = np.asarray(com, dtype=np.float64)
    r, c = np.indices(scan_shape)
    A = np.stack([r.ravel(), c.ravel(), np.ones(r.size)], axis=1).astype(np.float64)
    out = np.empty_like(com)
    for k in range(2):
        y = com[:, k]
        ok = np.isfinite(y)
        coef, *_ = np.linalg.lstsq(A[ok], y[ok], rcond=None)
        out[:, k] = A @ coef
    return out


def shift_array(ar, rshift, cshift):
    n_r, n_c = ar.shape
    qr = np.fft.fftfreq(n_r, 1.0)[:, None]
    qc = np.fft.fftfreq(n_c, 1.0)[None, :]
    p = np.exp(-2j * np.pi * (cshift * qc + rshift * qr))
    return np.real(np.fft.ifft2(np.fft.fft2(ar) * p))


def center_amplitudes(intensities, com_fit):
    """(P, n, n) intensities, (P, 2) com_fit -> (centred amplitudes (P, n, n), sums (P, 2))."""
    I = np.maximum(np.asarray(intensities, dtype=np.float64), 0.0)
    out = np.empty_like(I)
    sums = np.empty((I.shape[0], 2))
    for p in range(I.shape[0]):
        a = np.sqrt(I[p])
        sums[p] = (I[p].sum(), a.sum())
        sh = shift_array(a, -com_fit[p, 0], -com_fit[p, 1])
        out[p] = np.fft.fftshift(np.maximum(sh, 0.0))
    return out, sums
