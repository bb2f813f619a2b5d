"""Warm-start pieces: feature kernel vs the numpy restatement, MFMA conv vs a float64 CPU conv, and the
learned warm start against the flat start."""

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
    # the padded cell grid covers the 256^2 object exactly: 16 + 32 + 16 cells of 4 px
    assert s.pad_cells == 16 and (2 * s.pad_cells + s.scan) * s.step == s.obj_size
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


def _cpu64(x):
    return x.detach().cpu().double()


@pytest.mark.gpu
@pytest.mark.parametrize("nb,h,w,cin,cout", [(2, 5, 7, 8, 16), (3, 16, 16, 32, 64), (1, 8, 8, 384, 128), (2, 32, 32, 16, 16)])
def test_conv3x3_forward_matches_cpu(cuda, nb, h, w, cin, cout):
    """MFMA conv (bf16 operands, fp32 accumulation, zero padding) against a float64 CPU conv of the same
    bf16-rounded operands."""
    g = torch.Generator(device=cuda).manual_seed(nb * 100 + cin)
    x = _bf(torch.randn((nb, cin, h, w), device=cuda, generator=g))
    wt = _bf(torch.randn((cout, cin, 3, 3), device=cuda, generator=g) * 0.1)
    b = torch.randn(cout, device=cuda, generator=g)
    ref = F.conv2d(_cpu64(x), _cpu64(wt), _cpu64(b), padding=1)
    y = ws.conv3x3_raw(ws._to_nhwc_bf16(x), wt.permute(0, 2, 3, 1).to(torch.bfloat16).contiguous(), b, True)
    err = (_cpu64(y) - ref).norm() / ref.norm()
    assert err < 1e-5, float(err)


@pytest.mark.gpu
@pytest.mark.parametrize("nb,h,w,cin,cout", [(2, 6, 9, 16, 32), (4, 16, 16, 32, 16)])
def test_conv3x3_gradients_match_cpu(cuda, nb, h, w, cin, cout):
    """Input and weight gradients of the MFMA conv against float64 CPU autograd of the same bf16-rounded
    operands (the incoming gradient rounded to bf16 as the kernels take it)."""
    g = torch.Generator(device=cuda).manual_seed(7)
    x = _bf(torch.randn((nb, cin, h, w), device=cuda, generator=g))
    wt = _bf(torch.randn((cout, cin, 3, 3), device=cuda, generator=g) * 0.1)
    dy = _bf(torch.randn((nb, cout, h, w), device=cuda, generator=g))
    xr = _cpu64(x).requires_grad_(True)
    wr = _cpu64(wt).requires_grad_(True)
    F.conv2d(xr, wr, None, padding=1).backward(_cpu64(dy))
    xk = x.clone().requires_grad_(True)
    wk = wt.clone().requires_grad_(True)
    bk = torch.zeros(cout, device=cuda, requires_grad=True)
    ws.Conv3x3.apply(xk, wk, bk, True).backward(dy)
    assert (_cpu64(xk.grad) - xr.grad).norm() / xr.grad.norm() < 1e-5
    assert (_cpu64(wk.grad) - wr.grad).norm() / wr.grad.norm() < 1e-5
    assert torch.allclose(_cpu64(bk.grad), _cpu64(dy).sum(dim=(0, 2, 3)), rtol=1e-5, atol=1e-4)


@pytest.mark.gpu
def test_network_overfits_a_few_samples(cuda):
    torch.manual_seed(0)
    net = ws.WarmStartNet(pad=16).to(cuda)
    feats = torch.randn((4, 8, 32, 32), device=cuda)
    # smooth random phase maps (bilinear upsampling of 16 x 16 noise) over the 256 x 256 object
    target = F.interpolate(torch.randn((4, 1, 16, 16), device=cuda), size=(256, 256), mode="bilinear",
                           align_corners=False)[:, 0]
    norm = ws.FeatureNorm(feats)
    hist = ws.train(net, feats, target, norm, epochs=150, batch=4, lr=2e-3)
    assert np.isfinite(hist).all()
    assert hist[-1] < 0.5 * hist[0]


@pytest.mark.gpu
def test_learned_warm_start_gains_over_flat(cuda):
    """The learned warm start of the 4-slice 256^2 configuration (128^2 patterns, 32 x 32 scan, 1e4 e-):
    pairs simulated by the multislice forward kernel, the U-Net trained briefly on them (every conv on
    the MFMA kernels), then on held-out objects the flat-start and the warm-started reconstruction
    (20 epochs each) are compared: the warm start begins with a lower projected-phase error and reaches
    the flat method's final error in fewer epochs and less wall-clock time (features and inference
    included)."""
    import time

    spec = ws.WarmStartSpec()
    sim = ws.PairSimulator(spec, cuda)
    t0 = time.perf_counter()
    feats, targets = ws.make_dataset(sim, range(600))
    norm = ws.FeatureNorm(feats)
    torch.manual_seed(0)
    net = ws.WarmStartNet(pad=spec.pad_cells).to(cuda)
    hist = ws.train(net, feats, targets, norm, epochs=12, batch=32, lr=1e-3)
    torch.cuda.synchronize()
    print(f"warm start: 600 pairs simulated and 12 epochs trained in {time.perf_counter() - t0:.1f} s, "
          f"loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < hist[0]
    assert net(norm(feats[:1])).shape == (1, 256, 256)
    ws.evaluate(net, norm, sim, 99999, epochs=20)  # untimed: one-time costs of the first runs
    ratios = []
    for k in range(4):
        r = ws.evaluate(net, norm, sim, 100000 + k, epochs=20)
        flat0, warm0 = r["flat"][0][1], r["warm"][0][1]
        ep_warm = next(i for i, (_, e) in enumerate(r["warm"]) if e <= r["target_nmse"]) \
            if any(e <= r["target_nmse"] for _, e in r["warm"]) else None
        print(f"  seed {100000 + k}: NMSE flat {flat0:.3f} -> {r['target_nmse']:.3f} in 20 epochs; warm start "
              f"{warm0:.3f}, reaches it after {ep_warm} epochs, time ratio {r['ratio']:.2f}", flush=True)
        assert warm0 < flat0
        assert ep_warm is not None and ep_warm < 20
        ratios.append(r["ratio"])
    assert float(np.median(ratios)) < 1.0
