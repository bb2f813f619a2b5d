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


def _rotate(com_r: np.ndarray, com_c: np.ndarray, angle: np.ndarray, transpose: bool) -> tuple:
    """Rotate CoM vectors by `angle` (radians, broadcastable), the transpose swapping the components first."""
    a_r, a_c = (com_c, com_r) if transpose else (com_r, com_c)
    ca, sa = np.cos(angle), np.sin(angle)
    return ca * a_r - sa * a_c, sa * a_r + ca * a_c


def _mean_abs_curl(rot_r: np.ndarray, rot_c: np.ndarray) -> np.ndarray:
    """Mean |d rot_c / d r - d rot_r / d c| over the interior of the scan (central differences)."""
    d_rc = rot_r[..., 1:-1, 2:] - rot_r[..., 1:-1, :-2]
    d_cr = rot_c[..., 2:, 1:-1] - rot_c[..., :-2, 1:-1]
    return np.mean(np.abs(d_cr - d_rc), axis=(-2, -1))


def com_rotation(com_measured: np.ndarray, com_fit: np.ndarray, scan_shape, reciprocal_sampling=(1.0, 1.0),
                 angles_deg: np.ndarray | None = None, force_rotation_deg: float | None = None,
                 force_transpose: bool | None = None) -> dict:
    """Relative rotation (and transpose) between the scan and the detector that makes the CoM field
    curl-free, as ``_set_com_relative_rotation`` (dataset_models.py:1113-1371, without the plots).

    com_measured, com_fit: (P, 2) over a row-major (Rr, Rc) scan.  The normalised CoM is
    nan_to_num(measured - fit) * reciprocal_sampling (dataset_models.py:824-828); the search minimises
    the mean absolute curl over `angles_deg` (default -89 .. 89 degrees in steps of 1) for both
    transpose options unless forced.  Returns the rotation (radians), the transpose flag, the curl
    curves and the rotated CoM (2, Rr, Rc)."""
    Rr, Rc = scan_shape
    d = np.nan_to_num(np.asarray(com_measured, dtype=np.float64) - np.asarray(com_fit, dtype=np.float64))
    com_r = d[:, 0].reshape(Rr, Rc) * reciprocal_sampling[0]
    com_c = d[:, 1].reshape(Rr, Rc) * reciprocal_sampling[1]
    angles_deg = np.arange(-89.0, 90.0, 1.0) if angles_deg is None else np.asarray(angles_deg, dtype=np.float64)
    curls = {}
    if force_rotation_deg is not None:
        best = np.deg2rad(force_rotation_deg)
        if force_transpose is None:
            c0 = float(_mean_abs_curl(*_rotate(com_r, com_c, best, False)))
            c1 = float(_mean_abs_curl(*_rotate(com_r, com_c, best, True)))
            transpose = c1 < c0
        else:
            transpose = bool(force_transpose)
    else:
        ang = np.deg2rad(angles_deg)[:, None, None]
        options = [bool(force_transpose)] if force_transpose is not None else [False, True]
        for t in options:
            curls[t] = _mean_abs_curl(*_rotate(com_r[None], com_c[None], ang, t))
        # smallest curl; the untransposed option wins ties
        t_best = min(options, key=lambda t: (float(curls[t].min()), t))
        transpose = t_best
        best = float(np.deg2rad(angles_deg[int(np.argmin(curls[t_best]))]))
    rot = np.stack(_rotate(com_r, com_c, best, transpose))
    return {"rotation_rad": float(best), "transpose": bool(transpose), "curl": curls, "com": rot}
