"""Measured-data preprocessing on the GPU: centre of mass, plane fit, centred amplitudes.

The quantities ``PtychographyDatasetRaster`` derives from the raw 4D-STEM intensities before
reconstruction (quantem ``dataset_models.py:1027-1112, 1372-1480``): the per-pattern centre of mass
(HIP kernel), its plane fit over the scan (a 3-parameter least-squares problem, on the host, as
``fit_origin`` with ``fit_function="plane"``, ``ptycho_utils.py:352-410``), and the centred, fftshifted
amplitudes that become the targets, with the mean diffraction intensity / amplitude and the descan
shifts (HIP kernel; the Fourier shift of ``shift_array``, ``ptycho_utils.py:239-275``).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat


def _check(t: torch.Tensor, name: str, ndim: int) -> torch.Tensor:
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on a GPU device, got {t.device}")
    if t.dim() != ndim:
        raise ValueError(f"{name} must have {ndim} dimensions, got {tuple(t.shape)}")
    return t.contiguous()


def _stream(dev) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def pattern_com(intensities: torch.Tensor, mask: torch.Tensor | None = None) -> torch.Tensor:
    """(P, n, n) intensities -> (P, 2) centre of mass (row, column) in pixels (dataset_models.py:1063-1074)."""
    I = _check(intensities, "intensities", 3)
    P, n, n2 = I.shape
    if n != n2:
        raise ValueError(f"patterns must be square, got {tuple(I.shape)}")
    m = None if mask is None else _check(mask, "mask", 2)
    out = torch.empty((P, 2), dtype=torch.float32, device=I.device)
    rc = nat.load().qem_pattern_com(P, n, I.data_ptr(), None if m is None else m.data_ptr(), out.data_ptr(),
                                    _stream(I.device))
    nat.check(rc, "qem_pattern_com")
    return out


def fit_plane(com: np.ndarray, scan_shape) -> np.ndarray:
    """Plane fit mx * r + my * c + b of each CoM component over the (Rr, Rc) scan grid (fit_origin)."""
    com = np.asarray(com, dtype=np.float64)
    r, c = np.indices(tuple(scan_shape))
    A = np.stack([r.ravel(), c.ravel(), np.ones(r.size)], axis=1)
    out = np.empty_like(com)
    for k in range(2):
        y = com[:, k]
        ok = np.isfinite(y)
        if ok.sum() < 3:
            raise ValueError("the plane fit needs at least 3 finite centre-of-mass values")
        coef, *_ = np.linalg.lstsq(A[ok], y[ok], rcond=None)
        out[:, k] = A @ coef
    return out


def center_amplitudes(intensities: torch.Tensor, com_fit: torch.Tensor) -> tuple:
    """(P, n, n) intensities, (P, 2) com_fit -> centred fftshifted amplitudes (P, n, n) and per-pattern
    (sum of clamped intensities, sum of amplitudes) (P, 2)."""
    I = _check(intensities, "intensities", 3)
    P, n, _ = I.shape
    cfit = _check(com_fit.to(device=I.device, dtype=torch.float32), "com_fit", 2)
    lib = nat.load()
    nb = ctypes.c_size_t(0)
    nat.check(lib.qem_center_amplitudes_workspace_bytes(P, n, ctypes.byref(nb)), "center workspace")
    ws = torch.empty(max(int(nb.value), 8), dtype=torch.uint8, device=I.device)
    out = torch.empty_like(I)
    sums = torch.empty((P, 2), dtype=torch.float32, device=I.device)
    rc = lib.qem_center_amplitudes(P, n, I.data_ptr(), cfit.data_ptr(), out.data_ptr(), sums.data_ptr(),
                                   ws.data_ptr(), ws.numel(), _stream(I.device))
    nat.check(rc, "qem_center_amplitudes")
    return out, sums


def preprocess(intensities: torch.Tensor, scan_shape, mask: torch.Tensor | None = None) -> dict:
    """Raw (Rr*Rc, n, n) intensities of a row-major scan -> the reconstruction's inputs: centred
    amplitudes (targets), mean diffraction intensity / amplitude, measured and fitted CoM, and the
    descan shifts n/2 - com_fit (dataset_models.py:1466-1469)."""
    com = pattern_com(intensities, mask)
    com_fit = fit_plane(com.cpu().numpy(), scan_shape)
    centered, sums = center_amplitudes(intensities, torch.from_numpy(com_fit).float())
    n = intensities.shape[-1]
    s = sums.double().sum(dim=0).cpu().numpy() / intensities.shape[0]
    return {"targets": centered, "mean_intensity": float(s[0]), "mean_amplitude": float(s[1]),
            "com_measured": com, "com_fit": com_fit, "descan_shifts": (n / 2.0 - com_fit).astype(np.float32)}
