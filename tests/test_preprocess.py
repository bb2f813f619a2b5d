"""Measured-data preprocessing: the numpy restatement pinned on known answers (CPU), the HIP kernels
against it (GPU)."""

import numpy as np
import pytest
import torch

from oracle import preprocess_oracle as pre
from quantem_amd import _native as nat
from quantem_amd import preprocess as qpre


def test_oracle_com_of_deltas_and_mask():
    I = np.zeros((3, 16, 16))
    I[0, 3, 11] = 2.0
    I[1, 8, 8] = 1.0
    I[1, 10, 12] = 1.0
    I[2, 5, 5] = 4.0
    I[2, 15, 0] = 4.0
    com = pre.pattern_com(I)
    assert com[0] == pytest.approx([3.0, 11.0])
    assert com[1] == pytest.approx([9.0, 10.0])
    mask = np.ones((16, 16))
    mask[15, 0] = 0.0
    assert pre.pattern_com(I, mask)[2] == pytest.approx([5.0, 5.0])


def test_oracle_plane_fit_recovers_a_plane():
    rng = np.random.default_rng(0)
    r, c = np.indices((7, 9))
    com = np.stack([0.3 * r.ravel() - 0.1 * c.ravel() + 31.0, -0.2 * r.ravel() + 0.05 * c.ravel() + 33.5], axis=1)
    assert np.allclose(pre.fit_plane(com, (7, 9)), com, atol=1e-10)
    noisy = com + rng.normal(0, 0.1, com.shape)
    noisy[4, 0] = np.nan  # non-finite CoM values are left out of the fit
    assert np.abs(pre.fit_plane(noisy, (7, 9)) - com).max() < 0.2
    assert np.allclose(qpre.fit_plane(noisy, (7, 9)), pre.fit_plane(noisy, (7, 9)))


def test_oracle_shift_and_centering():
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 1, (16, 16))
    assert np.allclose(pre.shift_array(a, 3.0, -5.0), np.roll(a, (3, -5), axis=(0, 1)))
    assert np.allclose(pre.shift_array(a, 0.0, 0.0), a)
    assert pre.shift_array(a, 0.37, -1.6).sum() == pytest.approx(a.sum())  # the DC term is unchanged
    I = np.zeros((1, 16, 16))
    I[0, 4, 13] = 9.0
    out, sums = pre.center_amplitudes(I, np.array([[4.0, 13.0]]))
    expect = np.zeros((16, 16))
    expect[8, 8] = 3.0  # the pattern's centre of mass moved to (N/2, N/2), amplitude sqrt(9)
    assert np.allclose(out[0], expect, atol=1e-12)
    assert sums[0] == pytest.approx([9.0, 3.0])


def test_preprocess_host_errors():
    lib = nat.load()
    assert lib.qem_pattern_com(0, 16, 8, None, 8, None) == nat.QEM_ERR_ARG
    assert lib.qem_pattern_com(2, 16, None, None, 8, None) == nat.QEM_ERR_NULL
    assert lib.qem_center_amplitudes(2, 48, 8, 8, 8, None, None, 0, None) == nat.QEM_ERR_UNSUPPORTED
    assert lib.qem_center_amplitudes(2, 256, 8, 8, 8, None, None, 0, None) == nat.QEM_ERR_NULL
    with pytest.raises(ValueError):
        qpre.fit_plane(np.full((4, 2), np.nan), (2, 2))


def _patterns(P, n, rng):
    # bright-field-like discs with sub-pixel centres and noise
    r, c = np.indices((n, n))
    cen = n / 2 + rng.uniform(-3, 3, (P, 2))
    rad = n / 6
    d = np.hypot(r[None] - cen[:, 0, None, None], c[None] - cen[:, 1, None, None])
    I = 1e3 * (d < rad) + rng.uniform(0, 20, (P, n, n)) - 2.0  # a few negative counts, as in raw data
    return I.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("n,P", [(32, 7), (64, 5), (128, 3), (256, 3)])
def test_preprocess_kernels_match_oracle(cuda, n, P):
    rng = np.random.default_rng(n)
    I = _patterns(P, n, rng)
    mask = rng.uniform(0.5, 1.0, (n, n)).astype(np.float32)
    It = torch.from_numpy(I).to(cuda)
    com = qpre.pattern_com(It, torch.from_numpy(mask).to(cuda)).cpu().numpy()
    ref_com = pre.pattern_com(I, mask)
    assert np.abs(com - ref_com).max() < 1e-4
    com_fit = ref_com + rng.uniform(-0.5, 0.5, ref_com.shape)  # arbitrary sub-pixel shifts
    out, sums = qpre.center_amplitudes(It, torch.from_numpy(com_fit).float())
    ref_out, ref_sums = pre.center_amplitudes(I, com_fit.astype(np.float32).astype(np.float64))
    rel = np.linalg.norm(out.cpu().numpy() - ref_out) / np.linalg.norm(ref_out)
    assert rel < 1e-5, rel
    assert np.allclose(sums.cpu().numpy(), ref_sums, rtol=1e-6)


@pytest.mark.gpu
def test_preprocess_end_to_end(cuda):
    rng = np.random.default_rng(3)
    I = _patterns(6 * 5, 64, rng)
    out = qpre.preprocess(torch.from_numpy(I).to(cuda), (6, 5))
    com = pre.pattern_com(I)
    fit = pre.fit_plane(com, (6, 5))
    assert np.allclose(out["com_fit"], fit, atol=1e-4)
    assert np.allclose(out["descan_shifts"], 32.0 - fit, atol=1e-4)
    assert out["mean_intensity"] == pytest.approx(np.maximum(I.astype(np.float64), 0).sum() / 30, rel=1e-6)


def test_com_rotation_undoes_a_known_rotation():
    """A curl-free CoM field (the gradient of a smooth potential) rotated by 23 degrees (and, in the
    second case, transposed) is found to need a -23 degree rotation (and the transpose)."""
    rr, cc = np.meshgrid(np.arange(24.0), np.arange(20.0), indexing="ij")
    phi = np.sin(rr / 4.0) * np.cos(cc / 5.0) + 0.02 * rr * cc
    g_r, g_c = np.gradient(phi)
    th = np.deg2rad(23.0)
    rot_r = np.cos(th) * g_r - np.sin(th) * g_c
    rot_c = np.sin(th) * g_r + np.cos(th) * g_c
    com_fit = np.zeros((24 * 20, 2))
    for transposed in (False, True):
        m_r, m_c = (rot_c, rot_r) if transposed else (rot_r, rot_c)  # a transpose swaps the detector axes
        meas = np.stack([m_r.ravel(), m_c.ravel()], axis=1)
        out = qpre.com_rotation(meas, com_fit, (24, 20))
        assert out["transpose"] == transposed
        # with the transpose the swap restores (rot_r, rot_c), so -23 degrees undoes the rotation in both cases
        assert np.rad2deg(out["rotation_rad"]) == pytest.approx(-23.0, abs=0.51)
        curl_best = min(float(v.min()) for v in out["curl"].values())
        assert curl_best < 1e-2 * max(float(v.max()) for v in out["curl"].values())
    forced = qpre.com_rotation(np.stack([rot_r.ravel(), rot_c.ravel()], 1), com_fit, (24, 20), force_rotation_deg=-23.0)
    assert forced["transpose"] is False
