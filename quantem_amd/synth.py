"""Seeded synthetic multislice ptychography problems (objects, probes, scans, propagators, data).

Physics and conventions follow quantem:
  electron wavelength            core/utils/utils.py:97-104
  soft aperture, defocus chi     diffractive_imaging/complex_probe.py:51-86, 164-215 (C10 = -defocus,
                                 probe_models.py:178-180); Fourier probe normalised to unit L2 then
                                 inverse-FFT'd and renormalised (complex_probe.py:434-507)
  mode weights / intensity       probe_models.py:860-873 (sum |fft2_ortho(P)|^2 = mean intensity)
  Gram-Schmidt orthogonalisation probe_models.py:484-525
  Fresnel propagators            probe_models.py:376-427
  patch convention               corner-centred patches, fftfreq offsets (dataset_models.py:514-539)
Everything is generated in float64 numpy and cast once to complex64 / float32, so GPU and CPU
paths consume bit-identical inputs.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np


def electron_wavelength_angstrom(energy_ev: float) -> float:
    """Relativistic electron wavelength in Angstrom (same constants as the reference)."""
    m = 9.109383e-31
    e = 1.602177e-19
    c = 299792458.0
    h = 6.62607e-34
    return h / math.sqrt(2 * m * e * energy_ev) / math.sqrt(1 + e * energy_ev / 2 / m / c**2) * 1e10


def gaussian_atom_phase(shape, n_atoms: int, sigma: float, amp_range, rng: np.random.Generator) -> np.ndarray:
    """Sum of periodic Gaussian 'atoms' (peak U(amp_range) rad) at uniform random positions."""
    H, W = shape
    pos = rng.uniform(0.0, 1.0, size=(n_atoms, 2)) * np.array([H, W])
    amp = rng.uniform(amp_range[0], amp_range[1], size=n_atoms)
    phi = np.zeros((H, W), dtype=np.float64)
    rad = int(math.ceil(5 * sigma))
    dy = np.arange(-rad, rad + 1)
    for (py, px), a in zip(pos, amp):
        iy, ix = int(math.floor(py)), int(math.floor(px))
        gy = np.exp(-((iy + dy - py) ** 2) / (2 * sigma**2))
        gx = np.exp(-((ix + dy - px) ** 2) / (2 * sigma**2))
        rows = (iy + dy) % H
        cols = (ix + dy) % W
        phi[np.ix_(rows, cols)] += a * gy[:, None] * gx[None, :]
    return phi


def _shape2(n) -> tuple:
    """(rows, cols) of a pattern size given as one int (square) or a pair (roi_shape)."""
    return (int(n), int(n)) if np.ndim(n) == 0 else (int(n[0]), int(n[1]))


def fourier_probe(n, sampling: float, energy: float, semiangle_mrad: float, defocus: float,
                  soft_edges: bool = True) -> np.ndarray:
    """Aperture * exp(-i chi) in corner-centred Fourier space, unit L2 norm; n = N or (Nr, Nc)."""
    nr, nc = _shape2(n)
    lam = electron_wavelength_angstrom(energy)
    k_r, k_c = np.fft.fftfreq(nr, sampling), np.fft.fftfreq(nc, sampling)
    kx, ky = np.meshgrid(k_r, k_c, indexing="ij")
    kr = np.sqrt(kx**2 + ky**2)
    alpha = kr * lam
    phi = np.arctan2(ky, kx)
    cut = semiangle_mrad * 1e-3
    if soft_edges:
        # angular sampling (rad) along each axis (complex_probe.py:449, 51-86)
        da_r, da_c = (k_r[1] - k_r[0]) * lam, (k_c[1] - k_c[0]) * lam
        denom = np.sqrt((np.cos(phi) * da_r) ** 2 + (np.sin(phi) * da_c) ** 2)
        ap = np.clip((cut - alpha) / denom + 0.5, 0.0, 1.0)
    else:
        ap = (alpha <= cut).astype(np.float64)
    c10 = -defocus
    chi = 2 * np.pi / lam * 0.5 * alpha**2 * c10
    f = ap * np.exp(-1j * chi)
    return f / np.sqrt(np.sum(np.abs(f) ** 2))


def real_space_probe(n, sampling: float, energy: float, semiangle_mrad: float, defocus: float) -> np.ndarray:
    p = np.fft.ifft2(fourier_probe(n, sampling, energy, semiangle_mrad, defocus))
    return p / np.sqrt(np.sum(np.abs(p) ** 2))


def gram_schmidt(probes: np.ndarray) -> np.ndarray:
    """Orthogonalise modes keeping each original norm, sort by intensity (probe_models.py:484-525)."""
    norms = np.sqrt(np.sum(np.abs(probes) ** 2, axis=(-2, -1)))
    out = []
    for i in range(probes.shape[0]):
        p = probes[i].copy()
        for q in out:
            p = p - np.sum(np.conj(q) * p) * q
        out.append(p / max(np.sqrt(np.sum(np.abs(p) ** 2)), 1e-12))
    out = np.stack(out) * norms[:, None, None]
    order = np.argsort(-np.sum(np.abs(out) ** 2, axis=(-2, -1)), kind="stable")
    return out[order]


def mixed_probe(base: np.ndarray, weights: Sequence[float], total_intensity: float) -> np.ndarray:
    """Modes = base * low-order Hermite-Gaussian factors, orthogonalised, weighted, scaled.

    sum_m sum |fft2_ortho(P_m)|^2 = total_intensity and the per-mode power follows `weights`
    (the scaling of probe_models.py:860-873).
    """
    nr, nc = base.shape[-2:]
    M = len(weights)
    # corner-centred pixel offsets along rows (y) and columns (x)
    y, x = np.meshgrid(np.fft.fftfreq(nr, 1.0 / nr), np.fft.fftfreq(nc, 1.0 / nc), indexing="ij")
    w0 = max(np.sqrt(np.sum(np.abs(base) ** 2 * (x**2 + y**2)) / np.sum(np.abs(base) ** 2)), 1.0)
    u, v = x / w0, y / w0
    factors = [np.ones_like(u), u, v, u * v, u**2 - 0.5, v**2 - 0.5, u**3, v**3]
    modes = np.stack([base * factors[i % len(factors)] for i in range(M)]).astype(np.complex128)
    modes = gram_schmidt(modes)
    modes = modes / np.sqrt(np.sum(np.abs(modes) ** 2, axis=(-2, -1), keepdims=True))
    w = np.asarray(weights, dtype=np.float64)
    w = w / w.sum()
    return modes * np.sqrt(w * total_intensity)[:, None, None]


def raster_positions(ny: int, nx: int, step: float, jitter: float, rng: Optional[np.random.Generator],
                     origin=(0.0, 0.0)) -> np.ndarray:
    yy, xx = np.meshgrid(np.arange(ny) * step, np.arange(nx) * step, indexing="ij")
    pos = np.stack([yy.ravel(), xx.ravel()], axis=-1) + np.asarray(origin, dtype=np.float64)
    if jitter > 0 and rng is not None:
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    return pos


def fresnel_propagators(n, sampling: float, energy: float, thicknesses: Sequence[float],
                        tilt_mrad=(0.0, 0.0)) -> np.ndarray:
    """(S-1, Nr, Nc) propagators exp(-i pi lambda dz k^2) [* tilt ramps], fftfreq order (complex64);
    n = N or (Nr, Nc).  Evaluated the way probe_models.py:376-427 does: k^2 and the phase in complex64.
    """
    nr, nc = _shape2(n)
    if len(thicknesses) == 0:
        return np.zeros((0, nr, nc), dtype=np.complex64)
    lam = electron_wavelength_angstrom(energy)
    kr = np.fft.fftfreq(nr, sampling).astype(np.float32)
    kc = np.fft.fftfreq(nc, sampling).astype(np.float32)
    k2 = (kr[:, None] ** 2 + kc[None, :] ** 2).astype(np.complex64)
    dz = np.asarray(thicknesses, dtype=np.complex64)
    phase = (np.complex64(-1j * np.pi * lam) * dz)[:, None, None]
    prop = np.exp((phase * k2[None]).astype(np.complex64)).astype(np.complex64)
    tr, tc = tilt_mrad
    if tr != 0:
        prop = prop * np.exp(1j * (-2 * np.pi * dz.real[:, None, None] * np.tan(tr / 1e3)) * kr[None, :, None])
    if tc != 0:
        prop = prop * np.exp(1j * (-2 * np.pi * dz.real[:, None, None] * np.tan(tc / 1e3)) * kc[None, None, :])
    return prop.astype(np.complex64)


@dataclass
class ProblemSpec:
    """A seeded synthetic problem; defaults are the 8-slice 128x128 4-mode headline workload."""

    name: str = "multislice-8x384-128px-4modes-64x64scan"
    num_slices: int = 8
    obj_shape: tuple = (384, 384)
    roi: object = 128  # N (square) or the roi_shape (Nr, Nc)
    num_modes: int = 4
    mode_weights: tuple = (0.7, 0.15, 0.1, 0.05)
    scan: tuple = (64, 64)
    scan_step: float = 4.0
    jitter: float = 0.5
    slice_thickness: float = 2.0
    energy: float = 300e3
    sampling: float = 0.1
    semiangle_mrad: float = 20.0
    defocus: float = 100.0
    n_atoms: int = 300
    atom_sigma: float = 1.5
    atom_amp: tuple = (0.1, 0.5)
    dose: float = 1e5
    batch: int = 256
    extra: dict = field(default_factory=dict)


HEADLINE = ProblemSpec()
# deep-adjoint workload: 20 slices of 512x512 at 1 A, 256x256 patterns, 2 modes (the reference's default
# weights for two modes, probe_models.py:686-689), 48x48 raster; atom density as the headline slices
DEPTH = ProblemSpec(
    name="multislice-20x512-256px-2modes-48x48scan",
    num_slices=20, obj_shape=(512, 512), roi=256, num_modes=2, mode_weights=(0.98, 0.02), scan=(48, 48),
    slice_thickness=1.0, n_atoms=530, batch=256)
# large single-GPU reconstruction: 30 slices of 1024x1024 at 1.5 A, 256x256 patterns, 8 modes (the
# reference's default weights, probe_models.py:686-689), 128x128 raster (16384 patterns), batch 512
LARGE = ProblemSpec(
    name="multislice-30x1024-256px-8modes-128x128scan",
    num_slices=30, obj_shape=(1024, 1024), roi=256, num_modes=8, mode_weights=(0.86,) + (0.02,) * 7,
    scan=(128, 128), slice_thickness=1.5, n_atoms=2133, batch=512)
# the headline's slices, modes and scan with 128 x 120 patterns: the generic-size (mixed-radix) kernels
# on nearly the headline's work per pattern
GENERIC = ProblemSpec(name="multislice-8x384-128x120px-4modes-64x64scan", roi=(128, 120))
SINGLE_SLICE = ProblemSpec(
    name="single-slice-192px-64px-32x32scan",
    num_slices=1, obj_shape=(192, 192), roi=64, num_modes=1, mode_weights=(1.0,), scan=(32, 32),
    scan_step=4.0, jitter=0.5, dose=1e4, batch=64)

# named workloads of bench.py and the reconstruction command line
WORKLOADS = {
    "headline": HEADLINE,
    "depth20-256px": DEPTH,
    "large30-1024px": LARGE,
    "single-slice": SINGLE_SLICE,
    "generic-128x120px": GENERIC,
}


def scan_origin(spec: ProblemSpec) -> tuple:
    """Place the raster in the middle of the object (the reference pads the scan FOV)."""
    ext = (np.asarray(spec.scan) - 1) * spec.scan_step
    return tuple((np.asarray(spec.obj_shape) - ext) / 2.0)


def upsample_factor(scan: Sequence[int], scan_step_px: float, sampling: float) -> float:
    """The reference's ``upsample_factor`` (``dataset_models.py:451-456, 877-879``): the cropped object
    shape floor(fov / sampling), rounded up to even, over the scan grid, averaged over the two axes,
    with fov = scan step (Angstrom) * (gpts - 1).  Scan step given in object pixels here."""
    gpts = np.asarray(scan, dtype=np.float64)
    fov = (scan_step_px * sampling) * (gpts - 1)
    shp = np.floor(fov / sampling)
    shp += shp % 2
    return float((shp.astype("int") / gpts).mean())


def spec_upsample_factor(spec: "ProblemSpec") -> float:
    return upsample_factor(spec.scan, spec.scan_step, spec.sampling)


def make_problem(spec: ProblemSpec, seed: int = 0, with_data: bool = True) -> dict:
    """Object, probe, positions and propagators (and Poisson data computed by `intensity_fn`).

    Returns numpy arrays: obj (S,H,W) c64, probe (M,Nr,Nc) c64, positions (P,2) f32, propagators
    (S-1,Nr,Nc) c64, true_phase (S,H,W) f64.  Data are added by :func:`attach_data`.
    """
    rng = np.random.default_rng(seed)
    S = spec.num_slices
    phases = np.stack([gaussian_atom_phase(spec.obj_shape, spec.n_atoms, spec.atom_sigma, spec.atom_amp, rng)
                       for _ in range(S)])
    obj = np.exp(1j * phases).astype(np.complex64)
    base = real_space_probe(spec.roi, spec.sampling, spec.energy, spec.semiangle_mrad, spec.defocus)
    probe = mixed_probe(base, spec.mode_weights, spec.dose).astype(np.complex64)
    pos = raster_positions(spec.scan[0], spec.scan[1], spec.scan_step, spec.jitter, rng, scan_origin(spec))
    props = fresnel_propagators(spec.roi, spec.sampling, spec.energy, [spec.slice_thickness] * (S - 1))
    return {
        "obj": obj,
        "probe": probe,
        "positions": pos.astype(np.float32),
        "propagators": props,
        "true_phase": phases,
        "seed": seed,
    }


def poisson_amplitudes(intensities: np.ndarray, rng: np.random.Generator) -> tuple:
    """Poisson counts from model intensities (already scaled to the dose); returns (amplitudes, mean)."""
    counts = rng.poisson(np.clip(intensities.astype(np.float64), 0, None)).astype(np.float32)
    mean_intensity = float(counts.sum(axis=(-2, -1)).mean())
    return np.sqrt(counts).astype(np.float32), mean_intensity

