"""Writes reference_recipe_n64.npz: the reference test's simulation recipe as fixed vectors.

The recipe is quantem's tests/diffractive_imaging/test_ptychography.py:20-134 (N = 64, white-noise
phase object from default_rng(42) with the mean removed, corner-centred aperture probe with
Q_max = 0.5 1/A, Q_probe = Q_max/2, C10 = 50 A at 300 kV, 1 px raster, unnormalised fft2
intensities), re-typed here; the reference itself is not imported.  Expected gradients come from
the torch-autograd restatement (oracle/torch_ref.py, complex128), independent of the hand-written
FP64 adjoint that the tests check against them.  Eight scan positions keep the file small.

Run from the repository root: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import torch_ref  # noqa: E402
from quantem_amd.synth import electron_wavelength_angstrom  # noqa: E402

N = 64
Q_MAX = 0.5
Q_PROBE = Q_MAX / 2
ENERGY = 300e3
C10 = 50


def main():
    rng = np.random.default_rng(42)
    arr = rng.random((N, N))
    arr -= arr.mean()
    obj = np.exp(1.0j * arr.astype(np.float32)).astype(np.complex64)

    sampling = 1 / Q_MAX / 2
    recip = 2 * Q_MAX / N
    q = np.fft.fftfreq(N, sampling)
    qq = np.sqrt(q[:, None] ** 2 + q[None, :] ** 2)
    aperture = np.sqrt(np.clip((Q_PROBE - qq) / recip + 0.5, 0, 1))
    chi = qq**2 * electron_wavelength_angstrom(ENERGY) * np.pi * C10
    pf = aperture * np.exp(-1j * chi)
    pf /= np.sqrt(np.sum(np.abs(pf) ** 2))
    probe = (np.fft.ifft2(pf) * N).astype(np.complex64)

    all_pos = np.stack(np.meshgrid(np.arange(N, dtype=float), np.arange(N, dtype=float), indexing="ij"),
                       -1).reshape(-1, 2)
    sel = np.random.default_rng(7).choice(len(all_pos), 8, replace=False)
    pos = all_pos[sel].astype(np.float32)
    x0 = np.round(pos[:, 0]).astype(int)
    y0 = np.round(pos[:, 1]).astype(int)
    f = np.fft.fftfreq(N, d=1 / N).astype(int)
    row = (x0[:, None, None] + f[None, :, None]) % N
    col = (y0[:, None, None] + f[None, None, :]) % N
    exit_waves = obj.astype(np.complex128)[row, col] * probe.astype(np.complex128)
    intens = np.fft.fftshift(np.abs(np.fft.fft2(exit_waves)) ** 2, axes=(-2, -1))

    # targets: amplitudes of the recipe data x100 (the test scales by 100), in the ortho convention
    targets = np.sqrt(intens * 100 / (N * N)).astype(np.float32)
    mean_intensity = float(np.mean(np.sum(targets.astype(np.float64) ** 2, axis=(-2, -1))))
    num_gpts = N * N
    g = torch_ref.loss_and_grads(obj[None], probe[None], pos, None, targets, None, num_gpts, mean_intensity,
                                 grad_positions=False)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_recipe_n64.npz")
    np.savez_compressed(out, obj=obj, probe=probe, positions=pos, intensities_unnormalised=intens.astype(np.float32),
                        targets=targets, num_gpts=num_gpts, mean_intensity=mean_intensity, loss=g["loss"],
                        grad_obj=g["obj"][0].astype(np.complex64)[None], grad_probe=g["probe"].astype(np.complex64))
    print(out, os.path.getsize(out))


if __name__ == "__main__":
    main()
