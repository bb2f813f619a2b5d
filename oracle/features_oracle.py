"""CPU restatement of the warm-start pattern features (TEST INFRASTRUCTURE ONLY).

Centre of mass of an fftshifted pattern relative to (N/2, N/2) (the CoM of origin_models.py:95-126 /
dataset_models.py:1064-1074, in pixels) and the fraction of the total intensity inside 6 annuli
[edges[a], edges[a+1]) around the centre (annular virtual detectors, dataset4dstem.py:494-520).
"""

import numpy as np


def pattern_features(patterns: np.ndarray, edges: np.ndarray, amplitudes: bool = True) -> np.ndarray:
    p = np.asarray(patterns, dtype=np.float64)
    if amplitudes:
        p = p * p
    n = p.shape[-1]
    c = n // 2
    r, k = np.meshgrid(np.arange(n) - c, np.arange(n) - c, indexing="ij")
    rad = np.sqrt(r * r + k * k)
    tot = p.sum(axis=(-2, -1))
    out = np.zeros((p.shape[0], 8))
    out[:, 0] = (p * r).sum(axis=(-2, -1)) / tot
    out[:, 1] = (p * k).sum(axis=(-2, -1)) / tot
    for a in range(6):
        m = (rad >= edges[a]) & (rad < edges[a + 1])
        out[:, 2 + a] = (p * m).sum(axis=(-2, -1)) / tot
    return out
