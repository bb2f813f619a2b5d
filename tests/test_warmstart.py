"""Warm-start pieces: feature kernel vs the numpy restatement, MFMA conv vs a plain fp32 torch conv."""

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import features_oracle
from quantem_amd import warmstart as ws


def test_spec_window_and_edges():
    s = ws.WarmStartSpec()
    assert s.fov == 128
    w0, w1 = s.window
    assert w1 - w0 == 128 and 0 <= w0 and w1 <= s.obj_size
    first = s.origin
    assert w0 <= first - 1 and first + (s.scan - 1) * s.step <= w1  # scan positions inside the window
    e = ws.default_edges(10.0)
    assert e.numel() == 7 and torch.all(e[1:] > e[:-1])


def test_atom_phase_fft_is_periodic_sum_of_gaussians():
    rng = np.random.default_rng(0)
    ph = ws.atom_phase_fft(64, 1, 1.5, (0.4, 0.4), rng)
    assert ph.max() == pytest.approx(0.4, rel=0.15)  # one atom: peak ~ its amplitude (bilinear splat)
    assert ph.sum() == pytest.approx(0.4 * 2 * np.pi * 1.5**2, rel=1e-6)  # integral of the Gaussian


def test_features_oracle_on_known_patterns():
    """Pins the feature restatement: a delta's CoM is its offset from (N/2, N/2) and all of its
    intensity falls in the annulus holding its radius; a flat pattern has the grid's mean offset as
    CoM (-1/2 for even N) and annular fractions equal to pixel counts; `amplitudes` squares the input
    first."""
    n = 32
    edges = np.array([0.0, 2.0, 4.0, 6.0, 9.0, 12.0, 30.0])
    pats = np.zeros((3, n, n))
    pats[0, n // 2 + 3, n // 2 - 4] = 2.0   # radius 5 -> annulus [4, 6)
    pats[1, n // 2, n // 2 + 10] = 1.0      # radius 10 -> annulus [9, 12)
    pats[2] = 1.0
    f = features_oracle.pattern_features(pats, edges, amplitudes=False)
    assert f[0, :2] == pytest.approx([3.0, -4.0])
    assert f[0, 2:] == pytest.approx([0, 0, 1, 0, 0, 0])
    assert f[1, :2] == pytest.approx([0.0, 10.0])
    assert f[1, 2:] == pytest.approx([0, 0, 0, 0, 1, 0])
    r, k = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing="ij")
    rad = np.hypot(r, k)
    counts = [np.count_nonzero((rad >= edges[a]) & (rad < edges[a + 1])) for a in range(6)]
    assert f[2, :2] == pytest.approx([(np.arange(n) - n // 2).mean()] * 2)  # -0.5: even-size grid
    assert f[2, 2:] == pytest.approx(np.array(counts) / n**2)
    g = features_oracle.pattern_features(np.sqrt(pats), edges, amplitudes=True)
    assert g == pytest.approx(f)



@pytest.mark.gpu
def test_features_match_oracle(cuda):
    rng = np.random.default_rng(1)
    pats = rng.uniform(0, 3, (7, 64, 64)).astype(np.float32)
    edges = np.array([0, 3, 6, 9, 14, 20, 40], dtype=np.float32)
    got = ws.pattern_features(torch.from_numpy(pats).to(cuda), torch.from_numpy(edges), amplitudes=True).cpu().numpy()
    ref = features_oracle.pattern_features(pats, edges, amplitudes=True)
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-6)


def _bf(x):
    return x.to(torch.bfloat16).float()


@pytest.mark.gpu
@pytest.mark.parametrize("nb,h,w,cin,cout", [(2, 5, 7, 8, 16), (3, 16, 16, 32, 64), (1, 8, 8, 384, 128), (2, 32, 32, 16, 16)])
def test_conv3x3_forward_matches_torch(cuda, nb, h, w, cin, cout):
    g = torch.Generator(device=cuda).manual_seed(nb * 100 + cin)
    x = _bf(torch.randn((nb, cin, h, w), device=cuda, generator=g))
    wt = _bf(torch.randn((cout, cin, 3, 3), device=cuda, generator=g) * 0.1)
    b = torch.randn(cout, device=cuda, generator=g)
    ref = F.conv2d(x, wt, b, padding=1)
    y = ws.conv3x3_raw(ws._to_nhwc_bf16(x), wt.permute(0, 2, 3, 1).to(torch.bfloat16).contiguous(), b, True)
    err = (y - ref).norm() / ref.norm()
    assert err < 1e-5, float(err)


@pytest.mark.gpu
@pytest.mark.parametrize("nb,h,w,cin,cout", [(2, 6, 9, 16, 32), (4, 16, 16, 32, 16)])
def test_conv3x3_gradients_match_torch(cuda, nb, h, w, cin, cout):
    g = torch.Generator(device=cuda).manual_seed(7)
    x = _bf(torch.randn((nb, cin, h, w), device=cuda, generator=g))
    wt = _bf(torch.randn((cout, cin, 3, 3), device=cuda, generator=g) * 0.1)
    dy = _bf(torch.randn((nb, cout, h, w), device=cuda, generator=g))
    xr = x.clone().requires_grad_(True)
    wr = wt.clone().requires_grad_(True)
    F.conv2d(xr, wr, None, padding=1).backward(dy)
    xk = x.clone().requires_grad_(True)
    wk = wt.clone().requires_grad_(True)
    bk = torch.zeros(cout, device=cuda, requires_grad=True)
    ws.Conv3x3.apply(xk, wk, bk, True).backward(dy)
    assert (xk.grad - xr.grad).norm() / xr.grad.norm() < 1e-5
    assert (wk.grad - wr.grad).norm() / wr.grad.norm() < 1e-5
    assert torch.allclose(bk.grad, dy.sum(dim=(0, 2, 3)), rtol=1e-5, atol=1e-4)


@pytest.mark.gpu
def test_network_overfits_a_few_samples(cuda):
    torch.manual_seed(0)
    net = ws.WarmStartNet().to(cuda)
    feats = torch.randn((4, 8, 32, 32), device=cuda)
    target = torch.randn((4, 128, 128), device=cuda).cumsum(-1) * 0.05
    norm = ws.FeatureNorm(feats)
    hist = ws.train(net, feats, target, norm, epochs=40, batch=4, lr=2e-3)
    assert np.isfinite(hist).all()
    assert hist[-1] < 0.5 * hist[0]
