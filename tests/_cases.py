"""Seeded test problems shared by the CPU and GPU tests (numpy in, numpy out)."""

from __future__ import annotations

import numpy as np

from quantem_amd import synth


def make_case(N=32, S=3, M=2, B=6, H=None, W=None, seed=0, descan=False, half_integer=True, sampling=None,
              dz=2.0, loss_type="l2_amplitude", weak=False):
    """Random multislice problem with realistic probe modes and a non-trivial loss."""
    rng = np.random.default_rng(seed)
    nr, nc = (N, N) if np.ndim(N) == 0 else (int(N[0]), int(N[1]))  # roi_shape
    H = H or nr + 24
    W = W or nc + 16
    phase = rng.uniform(-0.6, 0.6, (S, H, W))
    amp = rng.uniform(0.6, 1.0, (S, H, W)) if not weak else np.ones((S, H, W))
    obj = (amp * np.exp(1j * phase)).astype(np.complex64)
    samp = sampling or (2.0 / min(nr, nc))  # keeps the 20 mrad aperture inside the band for any N
    base = synth.real_space_probe(N, samp, 300e3, 20.0, 40.0)
    weights = [0.7, 0.15, 0.1, 0.05, 0.03, 0.02][:M]
    probe = synth.mixed_probe(base, weights, 1e3).astype(np.complex64)
    pos = rng.uniform(0, 1, (B, 2)) * np.array([H - 1, W - 1])
    if half_integer and B >= 3:
        pos[0] = [2.5, 3.5]           # round-half-even in both directions
        pos[1] = [H - 1, 0.0]         # clamped corner: patch wraps around the object
        pos[2] = [0.49, W - 1.51]
    pos = pos.astype(np.float32)
    props = synth.fresnel_propagators(N, samp, 300e3, [dz] * (S - 1))
    d = rng.uniform(-1.5, 1.5, (B, 2)).astype(np.float32) if descan else None
    mask = rng.uniform(0.5, 1.0, (nr, nc)).astype(np.float32)
    return {"obj": obj, "probe": probe, "positions": pos, "propagators": props, "descan": d, "mask": mask,
            "num_gpts": 4 * B, "rng": rng, "loss_type": loss_type}


def attach_targets(case, I):
    """Targets near the model so the loss is moderate (amplitudes or intensities per loss type)."""
    rng = case["rng"]
    noisy = I * rng.uniform(0.7, 1.3, I.shape)
    lt = case["loss_type"]
    tg = np.sqrt(noisy) if "amplitude" in lt else noisy
    case["targets"] = tg.astype(np.float32)
    case["mean_intensity"] = float(np.mean(np.sum(noisy, axis=(-2, -1))))
    return case


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
